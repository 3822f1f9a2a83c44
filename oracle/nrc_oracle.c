/*
 * nrc_oracle.c -- CPU restatement of VkNRC's neural-radiance-cache hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP kernels in
 * vknrc_amd/csrc and the CPU baseline that bench.py times beside them.  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it;
 * the product library never links or calls it.
 *
 * One source, two builds (oracle/Makefile):
 *   liboracle_f64.so  REAL = double  -- the parity oracle
 *   libcpu_f32.so     REAL = float   -- the FP32 CPU baseline ("port" of the
 *                                       reference's Eigen MLP, test/main.cpp:11-27)
 *
 * Every function cites the reference file:line whose behaviour it restates.
 * Pins (tests/test_oracle.py): closed-form KATs for every encoder primitive,
 * numpy matmul and torch float64 autograd fixtures (tests/golden) for the
 * forward and the weight gradient, central finite differences of the loss,
 * batch invariants (duplication, permutation, zero padding), and first-step
 * and recurrence KATs for Adam / EMA / the prepare scalars.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef REAL
#define REAL double
#endif

#define W_ 64           /* network width, src/VkNRCState.hpp:23 */
#define OUT_ 3          /* output width, src/VkNRCState.hpp:23 */
#define REC_ 14         /* floats per query record (UnpackedNRCInput, NRCRecord.glsl:40-45) */
#define TREC_ 17        /* floats per train record: query + target rgb */
#define BLK_ 64         /* samples per oracle block */

enum { ENC_TRI12 = 0, ENC_SINCOS6 = 1, ENC_UV_SINCOS16 = 2 };
enum { LOSS_REL_L2_LUM = 0, LOSS_L2 = 1 };

static int g_threads = 0;
void oracle_set_threads(int n) { g_threads = n; }
int oracle_get_threads(void) {
#ifdef _OPENMP
	return g_threads > 0 ? g_threads : omp_get_max_threads();
#else
	return 1;
#endif
}
int oracle_real_bytes(void) { return (int)sizeof(REAL); }

/* ---- binary16 round-to-nearest-even, written from the IEEE 754 definition ---- */
double oracle_round_f16(double x) {
	if (x != x) return x;
	double ax = fabs(x);
	if (ax == 0.0) return x;
	double q;
	if (ax < 6.103515625e-05) { /* below 2^-14: subnormal quantum 2^-24 */
		q = 5.9604644775390625e-08;
	} else {
		int e;
		frexp(ax, &e);             /* ax = m * 2^e, m in [0.5, 1) */
		q = ldexp(1.0, e - 11);    /* 11 significant bits */
	}
	double r = nearbyint(x / q) * q; /* default rounding mode: ties to even */
	if (fabs(r) > 65504.0) return x > 0 ? INFINITY : -INFINITY;
	return r;
}

/* ---- encoders (NRCRecord.glsl) ---- */

/* GLSL mod(a, b) = a - b * floor(a / b); _nrc_tri, NRCRecord.glsl:65-68 */
REAL oracle_tri(REAL x) {
	REAL a = x - (REAL)0.5;
	REAL m = a - (REAL)2 * floor(a / (REAL)2);
	return (REAL)2 * fabs(m - (REAL)1) - (REAL)1;
}

/* _quartic_cdf with inv_radius 4, NRCRecord.glsl:51-56 */
REAL oracle_quartic_cdf(REAL d) {
	REAL u = d * (REAL)4;
	REAL u2 = u * u;
	REAL u4 = u2 * u2;
	REAL c = (REAL)(15.0 / 16.0) * u * ((REAL)1 - (REAL)(2.0 / 3.0) * u2 + (REAL)(1.0 / 5.0) * u4) + (REAL)0.5;
	return c < 0 ? 0 : (c > 1 ? 1 : c);
}

/* NRCOneBlob4Encode, NRCRecord.glsl:59-63 */
void oracle_oneblob4(REAL x, REAL o[4]) {
	for (int k = 0; k < 4; ++k) {
		REAL l = (REAL)0.25 * (REAL)k, r = (REAL)0.25 * (REAL)(k + 1);
		o[k] = oracle_quartic_cdf(r - x) - oracle_quartic_cdf(l - x);
	}
}

/* NRCSphEncode, NRCRecord.glsl:47-49 */
void oracle_sph_encode(const REAL d[3], REAL o[2]) {
	REAL z = d[2] < -1 ? -1 : (d[2] > 1 ? 1 : d[2]);
	o[0] = (d[0] == 0 && d[1] == 0) ? (REAL)0.5 : (REAL)0.5 + atan2(d[1], d[0]) / ((REAL)2 * (REAL)M_PI);
	o[1] = acos(z) / (REAL)M_PI;
}

/* sin(pi x), cos(pi x) with exact reduction of x modulo 2 */
static REAL sinpi_(REAL x) { return sin((REAL)M_PI * fmod(x, (REAL)2)); }
static REAL cospi_(REAL x) { return cos((REAL)M_PI * fmod(x, (REAL)2)); }

/*
 * NRCInputEncode, NRCRecord.glsl:78-95: 64 features of one record.
 * ENC_TRI12 is the reference layout; ENC_SINCOS6 swaps each axis's 12 triangle
 * waves for sin(2^k pi p), k=0..5, then cos(2^k pi p), k=0..5; ENC_UV_SINCOS16
 * reads (u, v) = position.xy and emits sin/cos(2^k pi u), k=0..15, then the same
 * for v (image-learning workload).  Features 36..63 are shared by the first two.
 */
void oracle_encode(const float *rec, int enc, int round_f16, REAL out[W_]) {
	if (enc == ENC_UV_SINCOS16) {
		for (int a = 0; a < 2; ++a)
			for (int k = 0; k < 16; ++k) {
				REAL x = (REAL)rec[a] * (REAL)(1 << k);
				out[32 * a + k] = sinpi_(x);
				out[32 * a + 16 + k] = cospi_(x);
			}
	} else {
		for (int a = 0; a < 3; ++a) {
			for (int k = 0; k < 12; ++k) {
				if (enc == ENC_TRI12) {
					out[12 * a + k] = oracle_tri((REAL)rec[a] * (REAL)(1 << k));
				} else {
					REAL x = (REAL)rec[a] * (REAL)(1 << (k % 6));
					out[12 * a + k] = k < 6 ? sinpi_(x) : cospi_(x);
				}
			}
		}
		oracle_oneblob4((REAL)rec[3], out + 36);
		oracle_oneblob4((REAL)rec[4], out + 40);
		oracle_oneblob4((REAL)rec[5], out + 44);
		oracle_oneblob4((REAL)rec[6], out + 48);
		oracle_oneblob4((REAL)1 - exp(-(REAL)rec[7]), out + 52);
		for (int c = 0; c < 6; ++c) out[56 + c] = (REAL)rec[8 + c];
		out[62] = 1;
		out[63] = 1;
	}
	if (round_f16)
		for (int i = 0; i < W_; ++i) out[i] = (REAL)oracle_round_f16((double)out[i]);
}

void oracle_encode_batch(const float *recs, int64_t n, int64_t stride, int enc, int round_f16, REAL *out) {
#pragma omp parallel for num_threads(oracle_get_threads()) schedule(static)
	for (int64_t i = 0; i < n; ++i) oracle_encode(recs + i * stride, enc, round_f16, out + i * W_);
}

/* ---- MLP on a block of B samples, activations feature-major [64][BLK_] ---- */

/*
 * y = W x for a block: W row-major [out][in] (NN_nv.glsl:55-57,69-82;
 * test/main.cpp:15-23), x and y feature-major.  relu: max(., 0)
 * (NN_nv.glsl:121-126).
 */
static void block_layer(const REAL *Wm, int rows, const REAL *x, REAL *y, int nb, int relu) {
	for (int m = 0; m < rows; ++m) {
		REAL *yr = y + m * BLK_;
		for (int n = 0; n < nb; ++n) yr[n] = 0;
		for (int k = 0; k < W_; ++k) {
			const REAL w = Wm[m * W_ + k];
			const REAL *xr = x + k * BLK_;
#pragma omp simd
			for (int n = 0; n < nb; ++n) yr[n] += w * xr[n];
		}
		if (relu)
			for (int n = 0; n < nb; ++n) yr[n] = yr[n] > 0 ? yr[n] : 0;
	}
}

/* encode a block of records into a feature-major [64][BLK_] buffer */
static void block_encode(const float *recs, int64_t stride, int nb, int enc, int round_f16, REAL *x) {
	REAL f[W_];
	for (int n = 0; n < nb; ++n) {
		oracle_encode(recs + n * stride, enc, round_f16, f);
		for (int k = 0; k < W_; ++k) x[k * BLK_ + n] = f[k];
	}
}

/*
 * Forward of one block (nrc_inference.comp:40-48, nrc_gradient.comp:38-45):
 * acts[0] = x, acts[l+1] = relu(W_l acts[l]) for l < H, out = W_H acts[H].
 * acts holds (H+1) [64][BLK_] slabs, out a [3][BLK_] slab.
 */
static void block_forward(const REAL *Wt, int H, REAL *acts, REAL *out, int nb) {
	for (int l = 0; l < H; ++l)
		block_layer(Wt + (int64_t)l * W_ * W_, W_, acts + (int64_t)l * W_ * BLK_, acts + (int64_t)(l + 1) * W_ * BLK_, nb, 1);
	block_layer(Wt + (int64_t)H * W_ * W_, OUT_, acts + (int64_t)H * W_ * BLK_, out, nb, 0);
}

/*
 * Inference over n query records (nrc_inference.comp:30-48; test/main.cpp:11-27).
 * out: [n][3].  clamp0 != 0 applies max(o, 0) as nrc_inference.comp:48 does.
 */
void oracle_infer(const REAL *Wt, int H, const float *recs, int64_t stride, int64_t n, int enc, int round_f16,
                  int clamp0, REAL *out) {
	int64_t nblk = (n + BLK_ - 1) / BLK_;
#pragma omp parallel num_threads(oracle_get_threads())
	{
		REAL *acts = (REAL *)malloc(sizeof(REAL) * (size_t)(H + 1) * W_ * BLK_);
		REAL *o = (REAL *)malloc(sizeof(REAL) * OUT_ * BLK_);
#pragma omp for schedule(static)
		for (int64_t b = 0; b < nblk; ++b) {
			int64_t n0 = b * BLK_;
			int nb = (int)((n - n0) < BLK_ ? (n - n0) : BLK_);
			block_encode(recs + n0 * stride, stride, nb, enc, round_f16, acts);
			block_forward(Wt, H, acts, o, nb);
			for (int i = 0; i < nb; ++i)
				for (int c = 0; c < OUT_; ++c) {
					REAL v = o[c * BLK_ + i];
					out[(n0 + i) * OUT_ + c] = (clamp0 && !(v > 0)) ? 0 : v;
				}
		}
		free(acts);
		free(o);
	}
}

/* Same forward over pre-encoded features [n][64] (test/evaluate_NV.comp:16-31). */
void oracle_infer_encoded(const REAL *Wt, int H, const REAL *feat, int64_t n, REAL *out) {
	int64_t nblk = (n + BLK_ - 1) / BLK_;
#pragma omp parallel num_threads(oracle_get_threads())
	{
		REAL *acts = (REAL *)malloc(sizeof(REAL) * (size_t)(H + 1) * W_ * BLK_);
		REAL *o = (REAL *)malloc(sizeof(REAL) * OUT_ * BLK_);
#pragma omp for schedule(static)
		for (int64_t b = 0; b < nblk; ++b) {
			int64_t n0 = b * BLK_;
			int nb = (int)((n - n0) < BLK_ ? (n - n0) : BLK_);
			for (int i = 0; i < nb; ++i)
				for (int k = 0; k < W_; ++k) acts[k * BLK_ + i] = feat[(n0 + i) * W_ + k];
			block_forward(Wt, H, acts, o, nb);
			for (int i = 0; i < nb; ++i)
				for (int c = 0; c < OUT_; ++c) out[(n0 + i) * OUT_ + c] = o[c * BLK_ + i];
		}
		free(acts);
		free(o);
	}
}

/*
 * Per-sample loss gradient and loss value.
 * LOSS_REL_L2_LUM: NNLoadDA3_RelativeL2LuminanceLoss, NN_nv.glsl:178-196:
 *   lum = (0.299, 0.587, 0.114) . max(o, 0);  d = lum^2 + 0.01 (held constant)
 *   g = 2 s (o - t) / d;   loss = sum_c (o_c - t_c)^2 / d
 * LOSS_L2: NNLoadDA3_L2Loss, NN_nv.glsl:162-177:  g = 2 s (o - t);  loss = sum (o - t)^2
 */
void oracle_loss_grad(int loss, const REAL o[3], const REAL t[3], REAL s, REAL g[3], REAL *value) {
	REAL d = 1;
	if (loss == LOSS_REL_L2_LUM) {
		REAL lum = (REAL)0.299 * (o[0] > 0 ? o[0] : 0) + (REAL)0.587 * (o[1] > 0 ? o[1] : 0) +
		           (REAL)0.114 * (o[2] > 0 ? o[2] : 0);
		d = lum * lum + (REAL)0.01;
	}
	REAL v = 0;
	for (int c = 0; c < 3; ++c) {
		REAL e = o[c] - t[c];
		g[c] = (REAL)2 * s * e / d;
		v += e * e / d;
	}
	*value = v;
}

/* ReLU backward with the reference's NaN-mask semantics (NN_nv.glsl:198-220,
 * 240-242, 272-277): an inactive unit (act <= 0) or a NaN result gives 0,
 * +-Inf on an active unit passes through. */
static inline REAL relu_bwd(REAL v, REAL act) { return (act > 0 && v == v) ? v : 0; }

/*
 * Batch weight gradient over a block (nrc_gradient.comp:45-57, NN_nv.glsl:222-367):
 *   delta_H = g_o;  dW_H += delta_H a_H^T
 *   g_H = (W_H^T g_o) . 1[a_H > 0];  dW_{H-1} += g_H a_{H-1}^T;  ...
 *   g_l = (W_l^T g_{l+1}) . 1[a_l > 0] for l = H-1..1;  dW_0 += g_1 a_0^T
 * Accumulates into dW (sum over samples, not divided by the count).
 */
static void block_backward(const REAL *Wt, int H, const REAL *acts, const REAL *go, int nb, REAL *dW, REAL *g,
                           REAL *gn) {
	/* dW_H: 3 x 64 */
	{
		const REAL *a = acts + (int64_t)H * W_ * BLK_;
		REAL *d = dW + (int64_t)H * W_ * W_;
		for (int m = 0; m < OUT_; ++m)
			for (int k = 0; k < W_; ++k) {
				REAL s = 0;
#pragma omp simd reduction(+ : s)
				for (int n = 0; n < nb; ++n) s += go[m * BLK_ + n] * a[k * BLK_ + n];
				d[m * W_ + k] += s;
			}
	}
	/* g_H = W_H^T g_o masked by a_H */
	{
		const REAL *Wm = Wt + (int64_t)H * W_ * W_;
		const REAL *a = acts + (int64_t)H * W_ * BLK_;
		for (int k = 0; k < W_; ++k)
			for (int n = 0; n < nb; ++n) {
				REAL v = 0;
				for (int m = 0; m < OUT_; ++m) v += Wm[m * W_ + k] * go[m * BLK_ + n];
				g[k * BLK_ + n] = relu_bwd(v, a[k * BLK_ + n]);
			}
	}
	for (int l = H - 1; l >= 0; --l) {
		const REAL *a = acts + (int64_t)l * W_ * BLK_;
		REAL *d = dW + (int64_t)l * W_ * W_;
		for (int m = 0; m < W_; ++m)
			for (int k = 0; k < W_; ++k) {
				REAL s = 0;
#pragma omp simd reduction(+ : s)
				for (int n = 0; n < nb; ++n) s += g[m * BLK_ + n] * a[k * BLK_ + n];
				d[m * W_ + k] += s;
			}
		if (l == 0) break;
		const REAL *Wm = Wt + (int64_t)l * W_ * W_;
		for (int k = 0; k < W_; ++k) {
			REAL *gr = gn + k * BLK_;
			for (int n = 0; n < nb; ++n) gr[n] = 0;
		}
		for (int m = 0; m < W_; ++m)
			for (int k = 0; k < W_; ++k) {
				const REAL w = Wm[m * W_ + k];
				const REAL *gm = g + m * BLK_;
				REAL *gr = gn + k * BLK_;
#pragma omp simd
				for (int n = 0; n < nb; ++n) gr[n] += w * gm[n];
			}
		for (int k = 0; k < W_; ++k)
			for (int n = 0; n < nb; ++n) g[k * BLK_ + n] = relu_bwd(gn[k * BLK_ + n], a[k * BLK_ + n]);
	}
}

static int64_t weight_count(int H) { return (int64_t)H * W_ * W_ + OUT_ * W_; }

/*
 * Summed weight gradient of n train records (records: [n][17], query + target),
 * loss scale s.  dW_out receives the sum (length weight_count(H)); loss_sum the
 * summed per-sample loss value.
 */
void oracle_gradient(const REAL *Wt, int H, const float *recs, int64_t stride, int64_t n, int enc, int round_f16,
                     int loss, REAL s, REAL *dW_out, REAL *loss_sum) {
	int64_t WC = weight_count(H);
	int nt = oracle_get_threads();
	REAL *acc = (REAL *)calloc((size_t)nt * (size_t)WC, sizeof(REAL));
	REAL *lacc = (REAL *)calloc((size_t)nt, sizeof(REAL));
	int64_t nblk = (n + BLK_ - 1) / BLK_;
#pragma omp parallel num_threads(nt)
	{
		int tid = 0;
#ifdef _OPENMP
		tid = omp_get_thread_num();
#endif
		REAL *acts = (REAL *)malloc(sizeof(REAL) * (size_t)(H + 1) * W_ * BLK_);
		REAL *o = (REAL *)malloc(sizeof(REAL) * OUT_ * BLK_);
		REAL *go = (REAL *)malloc(sizeof(REAL) * OUT_ * BLK_);
		REAL *g = (REAL *)malloc(sizeof(REAL) * W_ * BLK_);
		REAL *gn = (REAL *)malloc(sizeof(REAL) * W_ * BLK_);
		REAL *myacc = acc + (size_t)tid * (size_t)WC;
#pragma omp for schedule(static)
		for (int64_t b = 0; b < nblk; ++b) {
			int64_t n0 = b * BLK_;
			int nb = (int)((n - n0) < BLK_ ? (n - n0) : BLK_);
			block_encode(recs + n0 * stride, stride, nb, enc, round_f16, acts);
			block_forward(Wt, H, acts, o, nb);
			for (int i = 0; i < nb; ++i) {
				const float *r = recs + (n0 + i) * stride;
				REAL oo[3] = {o[i], o[BLK_ + i], o[2 * BLK_ + i]};
				REAL tt[3] = {(REAL)r[14], (REAL)r[15], (REAL)r[16]};
				REAL gg[3], v;
				oracle_loss_grad(loss, oo, tt, s, gg, &v);
				for (int c = 0; c < 3; ++c) go[c * BLK_ + i] = gg[c];
				lacc[tid] += v;
			}
			block_backward(Wt, H, acts, go, nb, myacc, g, gn);
		}
		free(acts);
		free(o);
		free(go);
		free(g);
		free(gn);
	}
	for (int64_t w = 0; w < WC; ++w) {
		REAL v = 0;
		for (int t = 0; t < nt; ++t) v += acc[(size_t)t * (size_t)WC + w];
		dW_out[w] = v;
	}
	REAL ls = 0;
	for (int t = 0; t < nt; ++t) ls += lacc[t];
	*loss_sum = ls;
	free(acc);
	free(lacc);
}

/*
 * Gradient over pre-encoded features [n][64] and targets [n][3] with the L2 loss
 * (test/train_NV.comp:18-46 semantics; the target here is not rounded).
 */
void oracle_gradient_encoded(const REAL *Wt, int H, const REAL *feat, const REAL *tgt, int64_t n, int loss, REAL s,
                             REAL *dW_out, REAL *loss_sum) {
	int64_t WC = weight_count(H);
	memset(dW_out, 0, sizeof(REAL) * (size_t)WC);
	REAL *acts = (REAL *)malloc(sizeof(REAL) * (size_t)(H + 1) * W_ * BLK_);
	REAL *o = (REAL *)malloc(sizeof(REAL) * OUT_ * BLK_);
	REAL *go = (REAL *)malloc(sizeof(REAL) * OUT_ * BLK_);
	REAL *g = (REAL *)malloc(sizeof(REAL) * W_ * BLK_);
	REAL *gn = (REAL *)malloc(sizeof(REAL) * W_ * BLK_);
	REAL ls = 0;
	for (int64_t n0 = 0; n0 < n; n0 += BLK_) {
		int nb = (int)((n - n0) < BLK_ ? (n - n0) : BLK_);
		for (int i = 0; i < nb; ++i)
			for (int k = 0; k < W_; ++k) acts[k * BLK_ + i] = feat[(n0 + i) * W_ + k];
		block_forward(Wt, H, acts, o, nb);
		for (int i = 0; i < nb; ++i) {
			REAL oo[3] = {o[i], o[BLK_ + i], o[2 * BLK_ + i]};
			REAL gg[3], v;
			oracle_loss_grad(loss, oo, tgt + (n0 + i) * 3, s, gg, &v);
			for (int c = 0; c < 3; ++c) go[c * BLK_ + i] = gg[c];
			ls += v;
		}
		block_backward(Wt, H, acts, go, nb, dW_out, g, gn);
	}
	*loss_sum = ls;
	free(acts);
	free(o);
	free(go);
	free(g);
	free(gn);
}

/*
 * Optimizer scalars (OptimizerState, VkNRCState.cpp:25-28) and Adam/EMA
 * hyper-parameters (Constant.glsl:5-13, nrc_optimize.comp:10-11), plus the
 * dynamic loss-scale policy of the fused path.
 */
typedef struct {
	double t, beta1_t, beta2_t, alpha_t, alpha_t_1; /* running products, initial {0, 1, 1, 1, 0} */
	double loss_scale;                              /* current scale s */
	double clean_steps;                             /* steps since the last scale change */
	double skipped_steps;                           /* steps skipped on overflow */
} oracle_opt_state;

typedef struct {
	double lr, beta1, beta2, eps, ema_alpha;
	int dynamic_scale;       /* 0: per-element NaN/Inf -> 0 (reference); 1: skip step + halve on overflow */
	int growth_interval;     /* clean steps before the scale doubles */
} oracle_opt_params;

/*
 * One optimizer step over summed gradient G (length wc) for a batch of `count`
 * records.  Follows nrc_train_prepare.comp:16-29 (scalar recurrences, only when
 * count > 0) and nrc_optimize.comp:32-54 (per weight):
 *   g = G / count / s, NaN/Inf -> 0
 *   m = b1 m + (1-b1) g;  v = b2 v + (1-b2) g^2
 *   w -= lr (m / (1 - B1)) / (sqrt(v / (1 - B2)) + eps)
 *   ema = (1-a)/(1-A) w + a (1-A1) ema            (as written, nrc_optimize.comp:46-47)
 * Dynamic scale: any non-finite G skips the whole step and halves s; after
 * growth_interval clean steps s doubles.  Returns 1 when a step was applied.
 */
int oracle_adam_step(const oracle_opt_params *p, oracle_opt_state *st, const REAL *G, int64_t wc, int64_t count,
                     REAL *m, REAL *v, REAL *w, REAL *ema) {
	if (count <= 0) return 0;
	REAL s = (REAL)st->loss_scale;
	if (p->dynamic_scale) {
		int bad = 0;
		for (int64_t i = 0; i < wc; ++i)
			if (!isfinite(G[i])) { bad = 1; break; }
		if (bad) {
			st->loss_scale *= 0.5;
			st->clean_steps = 0;
			st->skipped_steps += 1;
			return 0;
		}
	}
	st->t += 1;
	st->beta1_t = (double)((REAL)st->beta1_t * (REAL)p->beta1);
	st->beta2_t = (double)((REAL)st->beta2_t * (REAL)p->beta2);
	st->alpha_t_1 = st->alpha_t;
	st->alpha_t = (double)((REAL)st->alpha_t * (REAL)p->ema_alpha);
	const REAL b1 = (REAL)p->beta1, b2 = (REAL)p->beta2, lr = (REAL)p->lr, eps = (REAL)p->eps, al = (REAL)p->ema_alpha;
	const REAL c1 = (REAL)1 - (REAL)st->beta1_t, c2 = (REAL)1 - (REAL)st->beta2_t;
	const REAL eta = (REAL)1 - (REAL)st->alpha_t, eta1 = (REAL)1 - (REAL)st->alpha_t_1;
	for (int64_t i = 0; i < wc; ++i) {
		REAL g = G[i] / (REAL)count / s;
		if (!isfinite(g)) g = 0;
		m[i] = b1 * m[i] + ((REAL)1 - b1) * g;
		v[i] = b2 * v[i] + ((REAL)1 - b2) * g * g;
		REAL mh = m[i] / c1, vh = v[i] / c2;
		w[i] -= lr * mh / (sqrt(vh) + eps);
		ema[i] = ((REAL)1 - al) / eta * w[i] + al * eta1 * ema[i];
	}
	if (p->dynamic_scale) {
		st->clean_steps += 1;
		if (p->growth_interval > 0 && st->clean_steps >= p->growth_interval) {
			st->loss_scale *= 2.0;
			st->clean_steps = 0;
		}
	}
	return 1;
}
