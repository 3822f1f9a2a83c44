// nrc_device.hpp -- device-side definitions shared by the NRC kernels (gfx950).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "nrc.h"

namespace nrc {

constexpr int kWidth = 64;
constexpr int kOut = 3;
constexpr int kQueryFloats = 14;
constexpr int kTrainFloats = 17;

__host__ __device__ constexpr int weight_count(int hidden) { return hidden * kWidth * kWidth + kOut * kWidth; }

// Fragment-major fp16 weight copies, the operand layout of the gradient kernel's
// v_mfma_f32_16x16x32_f16 A fragments: block b = 8 * layer + 2 * mt + c holds
// 64 lanes x 8 halves, lane 16g + r carrying row 16 mt + r at the columns
//   32c + 4g + j (j < 4),  32c + 16 + 4g + j - 4 (j >= 4)
// so each lane loads one fragment with a single 16-byte load.
//   forward copy : W_l for l < H (blocks 0 .. 8H-1), then the 3x64 output layer
//                  in blocks 8H, 8H+1 (rows 3..15 stay zero)
//   backward copy: W_l^T for l = 1 .. H-1 in blocks 8(l-1) ..; W_H^T (64 x 3)
//                  in the c = 0 blocks of layer slot H-1 (K columns 3..31 stay zero)
__host__ __device__ constexpr uint32_t fwd_frag_halves(int hidden) { return (uint32_t)(hidden * 8 + 2) * 512u; }
__host__ __device__ constexpr uint32_t bwd_frag_halves(int hidden) { return (uint32_t)hidden * 4096u; }

// Position of element (row, col) of the 64-column matrix in layer slot l.
__host__ __device__ inline uint32_t frag_pos(uint32_t l, uint32_t row, uint32_t col) {
	const uint32_t c = col >> 5, w = col & 31u, g = (w & 15u) >> 2, j = (w & 3u) + ((w >> 4) << 2);
	return ((((l * 4u + (row >> 4)) * 2u + c) * 64u) + g * 16u + (row & 15u)) * 8u + j;
}

// Scatters weight i (row-major layer order) into the fragment copies (wtf may be null).
template <typename T> __host__ __device__ inline void frag_scatter(uint32_t i, T h, T *wf, T *wtf) {
	const uint32_t l = i >> 12, row = (i >> 6) & 63u, col = i & 63u;
	wf[frag_pos(l, row, col)] = h;
	if (wtf && l >= 1) wtf[frag_pos(l - 1, col, row)] = h;
}

// Device-resident optimizer state, double-buffered by a parity the host flips
// on every launched train step: step p reads pub[p] and writes pub[p ^ 1], and
// its gradient kernels flag non-finite partials in overflow[p] (the optimizer
// clears overflow[p ^ 1] for the next step).  No block ever reads a buffer
// another block of the same launch writes, so no cross-block counter is needed.
struct DeviceState {
	nrc_optimizer_state pub[2];
	uint32_t overflow[2];
};

struct AdamParams {
	float lr, beta1, beta2, eps, ema_alpha;
	int32_t dynamic_scale;
	uint32_t growth_interval;
	int32_t use_ema;
};

__device__ __forceinline__ uint32_t effective_count(const uint32_t *d_count, uint32_t max_count, uint32_t clamp) {
	uint32_t c = d_count ? *d_count : max_count;
	c = c < max_count ? c : max_count;
	return c < clamp ? c : clamp;
}

// ---------------------------------------------------------------- encoders
// _nrc_tri, NRCRecord.glsl:65-68 with GLSL's floor-based mod(a, 2).
__device__ __forceinline__ float tri(float x) {
	float a = x - 0.5f;
	float m = a - 2.0f * floorf(a * 0.5f);
	return 2.0f * fabsf(m - 1.0f) - 1.0f;
}

// _quartic_cdf with inv_radius 4, NRCRecord.glsl:51-56.
__device__ __forceinline__ float quartic_cdf(float d) {
	float u = d * 4.0f;
	float u2 = u * u;
	float u4 = u2 * u2;
	float c = (15.0f / 16.0f) * u * (1.0f - (2.0f / 3.0f) * u2 + (1.0f / 5.0f) * u4) + 0.5f;
	return fminf(fmaxf(c, 0.0f), 1.0f);
}

// NRCOneBlob4Encode, NRCRecord.glsl:59-63.  Adjacent bins share a CDF value, so
// five evaluations give the four bins.
__device__ __forceinline__ void oneblob4(float x, float *o) {
	float c0 = quartic_cdf(0.0f - x), c1 = quartic_cdf(0.25f - x), c2 = quartic_cdf(0.5f - x),
	      c3 = quartic_cdf(0.75f - x), c4 = quartic_cdf(1.0f - x);
	o[0] = c1 - c0;
	o[1] = c2 - c1;
	o[2] = c3 - c2;
	o[3] = c4 - c3;
}

// sin(pi x) and cos(pi x).  With k = rint(x) and r = x - k in [-1/2, 1/2]
// (exact), sin(pi x) = (-1)^k sin(pi r) and cos(pi x) = (-1)^k cos(pi r).  The
// polynomials in z = r^2 are weighted least-squares fits on Chebyshev nodes of
// [0, 1/2] (max error 6e-9 for sin(pi r)/r, 2e-9 for cos(pi r) in exact
// arithmetic, about 2e-7 after FP32 evaluation).  Valid for |x| < 2^31.
__device__ __forceinline__ void sincospi_poly(float x, float *s, float *c) {
	const float k = rintf(x);
	const float r = x - k;
	const float z = r * r;
	float ps = fmaf(z, 0.07729381322860718f, -0.5980821251869202f);
	ps = fmaf(z, ps, 2.550037384033203f);
	ps = fmaf(z, ps, -5.167707443237305f);
	ps = fmaf(z, ps, 3.141592502593994f);
	float pc = fmaf(z, -0.02422487549483776f, 0.23481841385364532f);
	pc = fmaf(z, pc, -1.3351819515228271f);
	pc = fmaf(z, pc, 4.058705806732178f);
	pc = fmaf(z, pc, -4.934802055358887f);
	pc = fmaf(z, pc, 1.0f);
	const uint32_t sign = ((uint32_t)(int32_t)k & 1u) << 31;
	*s = __uint_as_float(__float_as_uint(r * ps) ^ sign);
	*c = __uint_as_float(__float_as_uint(pc) ^ sign);
}

// sin/cos(2^k pi x) for k = 0..N-1.  Every fourth octave is evaluated directly
// at the exact argument 2^k x; the three between use the
// double-angle recurrences sin 2a = 2 sin a cos a, cos 2a = 1 - 2 sin^2 a, whose
// error at most doubles per step (<= 8x the direct error before re-anchoring).
template <int N> __device__ __forceinline__ void sincos_octaves(float x, float *s, float *c) {
	float sv = 0.0f, cv = 1.0f;
#pragma unroll
	for (int k = 0; k < N; ++k) {
		if (k % 4 == 0) {
			sincospi_poly(x * (float)(1 << k), &sv, &cv);
		} else {
			float ns = 2.0f * sv * cv;
			float nc = 1.0f - 2.0f * sv * sv;
			sv = ns;
			cv = nc;
		}
		s[k] = sv;
		c[k] = cv;
	}
}

// Features of feature block q (16 features: 16q .. 16q+15) of one record.
// Layout (NRCInputEncode, NRCRecord.glsl:78-95), 64 features:
//   0..35  position: 12 per axis (TRI12: tri(2^k p), k=0..11;
//          SINCOS6: sin(2^k pi p), k=0..5, then cos(2^k pi p), k=0..5)
//   36..39 oneblob4(scattered_dir.u)   40..43 oneblob4(scattered_dir.v)
//   44..47 oneblob4(normal.u)          48..51 oneblob4(normal.v)
//   52..55 oneblob4(1 - exp(-roughness))
//   56..58 diffuse rgb   59..61 specular rgb   62, 63 = 1
// UV_SINCOS16: sin(2^k pi u) k=0..15 | cos(2^k pi u) | sin(2^k pi v) | cos(2^k pi v).
template <int ENC> __device__ __forceinline__ void axis12(float p, float *o) {
	if constexpr (ENC == NRC_ENCODING_TRI12) {
#pragma unroll
		for (int k = 0; k < 12; ++k) o[k] = tri(p * (float)(1 << k));
	} else {
		sincos_octaves<6>(p, o, o + 6);
	}
}

template <int ENC> __device__ __forceinline__ void encode_block(const float *rec, int q, float *o) {
	if constexpr (ENC == NRC_ENCODING_UV_SINCOS16) {
		float s[16], c[16];
		sincos_octaves<16>(rec[q >> 1], s, c);
#pragma unroll
		for (int k = 0; k < 16; ++k) o[k] = (q & 1) ? c[k] : s[k];
	} else {
		float t[12];
		if (q == 0) {
			axis12<ENC>(rec[0], t);
#pragma unroll
			for (int k = 0; k < 12; ++k) o[k] = t[k];
			axis12<ENC>(rec[1], t);
#pragma unroll
			for (int k = 0; k < 4; ++k) o[12 + k] = t[k];
		} else if (q == 1) {
			axis12<ENC>(rec[1], t);
#pragma unroll
			for (int k = 0; k < 8; ++k) o[k] = t[4 + k];
			axis12<ENC>(rec[2], t);
#pragma unroll
			for (int k = 0; k < 8; ++k) o[8 + k] = t[k];
		} else if (q == 2) {
			axis12<ENC>(rec[2], t);
#pragma unroll
			for (int k = 0; k < 4; ++k) o[k] = t[8 + k];
			oneblob4(rec[3], o + 4);
			oneblob4(rec[4], o + 8);
			oneblob4(rec[5], o + 12);
		} else {
			oneblob4(rec[6], o);
			oneblob4(1.0f - expf(-rec[7]), o + 4);
#pragma unroll
			for (int k = 0; k < 6; ++k) o[8 + k] = rec[8 + k];
			o[14] = 1.0f;
			o[15] = 1.0f;
		}
	}
}

// All 64 features of one record.
template <int ENC> __device__ __forceinline__ void encode_all(const float *rec, float *o) {
	if constexpr (ENC == NRC_ENCODING_UV_SINCOS16) {
		float s[16], c[16];
#pragma unroll
		for (int a = 0; a < 2; ++a) {
			sincos_octaves<16>(rec[a], s, c);
#pragma unroll
			for (int k = 0; k < 16; ++k) {
				o[32 * a + k] = s[k];
				o[32 * a + 16 + k] = c[k];
			}
		}
	} else {
		axis12<ENC>(rec[0], o);
		axis12<ENC>(rec[1], o + 12);
		axis12<ENC>(rec[2], o + 24);
		oneblob4(rec[3], o + 36);
		oneblob4(rec[4], o + 40);
		oneblob4(rec[5], o + 44);
		oneblob4(rec[6], o + 48);
		oneblob4(1.0f - expf(-rec[7]), o + 52);
#pragma unroll
		for (int k = 0; k < 6; ++k) o[56 + k] = rec[8 + k];
		o[62] = 1.0f;
		o[63] = 1.0f;
	}
}

// float -> int32 as v_cvt_i32_f32 converts: toward zero, saturating at the int32
// range, NaN -> 0 (a C++ cast of a NaN or of an out-of-range value is undefined).
__device__ __forceinline__ int cvt_i32_sat(float x) {
	int k;
	asm("v_cvt_i32_f32 %0, %1" : "=v"(k) : "v"(x));
	return k;
}

// b + f * p with the product and the sum rounded separately: the empty asm makes
// the product opaque, so the compiler cannot contract the pair into an FMA.
__device__ __forceinline__ float mul_add_rn(float b, float f, float p) {
	float m = f * p;
	asm volatile("" : "+v"(m));
	return b + m;
}

// NRCSphEncode, NRCRecord.glsl:47-49.
__device__ __forceinline__ void sph_encode(float x, float y, float z, float *uv) {
	constexpr float kPi = 3.14159265358979323846f;
	uv[0] = (x == 0.0f && y == 0.0f) ? 0.5f : 0.5f + atan2f(y, x) / (2.0f * kPi);
	uv[1] = acosf(fminf(fmaxf(z, -1.0f), 1.0f)) / kPi;
}

// ReLU backward with the reference's NaN-mask semantics (NN_nv.glsl:198-220,
// 240-242): inactive unit or NaN result -> 0; +-Inf on an active unit passes.
__device__ __forceinline__ float relu_bwd(float v, float act) { return (act > 0.0f && v == v) ? v : 0.0f; }

// Per-sample loss gradient and value, NN_nv.glsl:162-196 (scale s; the
// relative-L2 denominator is held constant).
template <int LOSS>
__device__ __forceinline__ float loss_grad(const float *o, const float *t, float s, float *g) {
	float d = 1.0f;
	if constexpr (LOSS == NRC_LOSS_RELATIVE_L2_LUMINANCE) {
		float lum = 0.299f * fmaxf(o[0], 0.0f) + 0.587f * fmaxf(o[1], 0.0f) + 0.114f * fmaxf(o[2], 0.0f);
		d = lum * lum + 0.01f;
	}
	float v = 0.0f;
#pragma unroll
	for (int c = 0; c < 3; ++c) {
		float e = o[c] - t[c];
		g[c] = 2.0f * s * e / d;
		v += e * e / d;
	}
	return v;
}

} // namespace nrc
