// nrc_fused.hip -- fully fused 64-wide MLP on gfx950 FP16 MFMA (FP32 accumulation).
//
// MFMA used: v_mfma_f32_16x16x32_f16.  Per instruction, lane l = 16*g + r holds
//   A: row r,    8 halves of K-group g      B: column r, 8 halves of K-group g
//   D: column r, rows 4g .. 4g+3 (FP32)
// A and B pair their halves by (g, j), so any bijection between (g, j) and the
// 32 K indices of an instruction gives the same product as long as A and B use
// the same one.  The kernels use
//   k(c, g, j) = 32c + 4g + j        (j < 4)
//              = 32c + 16 + 4g + j-4 (j >= 4)        for K-chunk c in {0, 1}
// because that is exactly where the D tiles of the previous layer leave the
// features: D of M-tile 2c holds rows 32c + 4g + i, D of M-tile 2c+1 holds rows
// 32c + 16 + 4g + i.  Layer l's output therefore becomes layer l+1's B operand
// in the same lanes, with a ReLU and an FP16 pack and no data movement.
//
// Weight layout in HBM (unchanged from the reference, NN_nv.glsl:55-57): the
// fp16 weights of layer l, row-major [out][in], at l*4096; the 3x64 output
// layer at H*4096 (padded to 16 rows with zeros inside the MFMA, NN_nv.glsl:77-82).
#include <cstdlib>

#include "nrc_internal.hpp"

namespace nrc {

namespace {

typedef _Float16 half8_t __attribute__((ext_vector_type(8)));
typedef _Float16 half4_t __attribute__((ext_vector_type(4)));
typedef float float4_t __attribute__((ext_vector_type(4)));

constexpr int kLdw = 72; // halves per LDS weight row (64 + 8 padding)
constexpr int kLdn = 72; // halves per LDS activation row (64 samples + 8 padding)

__device__ __forceinline__ float4_t mfma(half8_t a, half8_t b, float4_t c) {
	return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
}

typedef _Float16 half2_t __attribute__((ext_vector_type(2)));
typedef float float2_t __attribute__((ext_vector_type(2)));

// Two FP32 values -> packed FP16 (round to nearest even), then ReLU on the pair.
// max(cvt(x), 0) equals cvt(max(x, 0)) up to the sign of a zero.
__device__ __forceinline__ half2_t cvt_relu2(float a, float b) {
	const float2_t f = {a, b};
	return __builtin_elementwise_max(__builtin_convertvector(f, half2_t), (half2_t)0);
}

__device__ __forceinline__ half8_t pack_relu(float4_t lo, float4_t hi) {
	const half2_t a = cvt_relu2(lo[0], lo[1]), b = cvt_relu2(lo[2], lo[3]);
	const half2_t c = cvt_relu2(hi[0], hi[1]), d = cvt_relu2(hi[2], hi[3]);
	half8_t h;
	h[0] = a[0]; h[1] = a[1]; h[2] = b[0]; h[3] = b[1];
	h[4] = c[0]; h[5] = c[1]; h[6] = d[0]; h[7] = d[1];
	return h;
}

// Orders LDS traffic between lanes of one wavefront (LDS operations of a wave
// execute in order; this keeps the compiler from moving them across).  The
// fences are restricted to LDS so they wait on LDS counters only and leave the
// prefetched global loads in flight.
__device__ __forceinline__ void wave_lds_sync() {
	__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
	__builtin_amdgcn_wave_barrier();
	__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
}

// Row-major [rows][64] fp16 weights -> padded LDS rows.
__device__ __forceinline__ void stage_rows(_Float16 *sW, const _Float16 *__restrict__ w16, int rows) {
	const int chunks = rows * 8;
	for (int ci = threadIdx.x; ci < chunks; ci += blockDim.x) {
		const int row = ci >> 3, c8 = ci & 7;
		*reinterpret_cast<uint4 *>(sW + row * kLdw + c8 * 8) = reinterpret_cast<const uint4 *>(w16)[ci];
	}
}

// A fragment of row `row` for K-chunk c in lane group g (see the header).
__device__ __forceinline__ half8_t load_afrag(const _Float16 *sW, int row, int c, int g) {
	const half4_t lo = *reinterpret_cast<const half4_t *>(sW + row * kLdw + 32 * c + 4 * g);
	const half4_t hi = *reinterpret_cast<const half4_t *>(sW + row * kLdw + 32 * c + 16 + 4 * g);
	half8_t a;
	a[0] = lo[0]; a[1] = lo[1]; a[2] = lo[2]; a[3] = lo[3];
	a[4] = hi[0]; a[5] = hi[1]; a[6] = hi[2]; a[7] = hi[3];
	return a;
}

// --------------------------------------------------------------- encoders
// Emits the 16 feature quads (features 4p..4p+3) of one record; Emit(p, f0..f3).
template <int ENC, typename Emit> __device__ __forceinline__ void encode_quads(const float *rec, Emit &&emit) {
	if constexpr (ENC == NRC_ENCODING_UV_SINCOS16) {
#pragma unroll
		for (int a = 0; a < 2; ++a) {
#pragma unroll
			for (int k0 = 0; k0 < 16; k0 += 4) {
				float s[4], c[4];
				float sv, cv;
				sincospi_poly(rec[a] * (float)(1 << k0), &sv, &cv);
				s[0] = sv;
				c[0] = cv;
#pragma unroll
				for (int k = 1; k < 4; ++k) {
					const float ns = 2.0f * sv * cv, nc = 1.0f - 2.0f * sv * sv;
					sv = ns;
					cv = nc;
					s[k] = sv;
					c[k] = cv;
				}
				emit(8 * a + k0 / 4, s[0], s[1], s[2], s[3]);
				emit(8 * a + 4 + k0 / 4, c[0], c[1], c[2], c[3]);
			}
		}
	} else {
#pragma unroll
		for (int a = 0; a < 3; ++a) {
			float t[12];
			axis12<ENC>(rec[a], t);
			emit(3 * a + 0, t[0], t[1], t[2], t[3]);
			emit(3 * a + 1, t[4], t[5], t[6], t[7]);
			emit(3 * a + 2, t[8], t[9], t[10], t[11]);
		}
		float b[4];
		oneblob4(rec[3], b);
		emit(9, b[0], b[1], b[2], b[3]);
		oneblob4(rec[4], b);
		emit(10, b[0], b[1], b[2], b[3]);
		oneblob4(rec[5], b);
		emit(11, b[0], b[1], b[2], b[3]);
		oneblob4(rec[6], b);
		emit(12, b[0], b[1], b[2], b[3]);
		oneblob4(1.0f - expf(-rec[7]), b);
		emit(13, b[0], b[1], b[2], b[3]);
		emit(14, rec[8], rec[9], rec[10], rec[11]);
		emit(15, rec[12], rec[13], 1.0f, 1.0f);
	}
}

// ================================================================ inference
//
// Persistent: each wave loads the whole network's A fragments into VGPRs once
// (from the block's LDS copy), then streams 64-record tiles: lane i encodes
// record i straight into the B-fragment layout of its wave's LDS slot, and the
// 6 layers run as 4 N-tiles of 16 records, TILES at a time, activations in VGPRs.
template <int H, int ENC, bool F16IN, int TILES>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void fused_infer_kernel(
    const float *__restrict__ queries, const _Float16 *__restrict__ fin, const uint32_t *__restrict__ d_count,
    uint32_t max_count, const _Float16 *__restrict__ w16, float *__restrict__ out, int clamp0) {
	constexpr int kRows = H * kWidth + kOut;
	__shared__ __attribute__((aligned(16))) _Float16 sW[kRows * kLdw];
	__shared__ __attribute__((aligned(16))) half8_t sX[4][8 * 64];

	stage_rows(sW, w16, kRows);
	__syncthreads();

	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	const int g = lane >> 4, r = lane & 15;

	half8_t A[H][4][2];
#pragma unroll
	for (int l = 0; l < H; ++l)
#pragma unroll
		for (int mt = 0; mt < 4; ++mt)
#pragma unroll
			for (int c = 0; c < 2; ++c) A[l][mt][c] = load_afrag(sW, l * kWidth + mt * 16 + r, c, g);
	half8_t Ao[2];
#pragma unroll
	for (int c = 0; c < 2; ++c) {
		Ao[c] = load_afrag(sW, H * kWidth + (r < kOut ? r : 0), c, g);
		if (r >= kOut) Ao[c] = (half8_t)0;
	}

	const uint32_t count = effective_count(d_count, max_count, max_count);
	const uint32_t ntiles = (count + 63) / 64;
	const uint32_t nwaves = gridDim.x * 4;
	half8_t *sx = sX[wave];
	half4_t *sx4 = reinterpret_cast<half4_t *>(sx);
	const int et = lane >> 4, es = lane & 15; // this lane's record: tile et, column es

	// the record of this lane in tile t.  The index is clamped into [0, count)
	// instead of branching, so the load is unconditional (lanes past the count
	// encode zeros regardless); a branch here makes the wait-count pass drain
	// every outstanding load at the join.  count == 0 reads record 0, which
	// exists because max_count > 0.
	const uint32_t last = count ? count - 1 : 0;
	auto load_query = [&](uint32_t t, float *dst) __attribute__((always_inline)) {
		uint32_t m = t * 64 + lane;
		m = m < last ? m : last;
		const float2 *qp = reinterpret_cast<const float2 *>(queries + (size_t)m * kQueryFloats);
#pragma unroll
		for (int i = 0; i < 7; ++i) {
			const float2 v = qp[i];
			dst[2 * i] = v.x;
			dst[2 * i + 1] = v.y;
		}
	};
	// encode one record per lane straight into the B-fragment layout of this wave's slot
	auto encode_tile = [&](uint32_t base, const float *rec) __attribute__((always_inline)) {
		const uint32_t n = base + lane;
		auto emit = [&](int p, float f0, float f1, float f2, float f3) __attribute__((always_inline)) {
			const int c = p >> 3, gg = p & 3, hi = (p >> 2) & 1;
			half4_t h;
			h[0] = (_Float16)f0; h[1] = (_Float16)f1; h[2] = (_Float16)f2; h[3] = (_Float16)f3;
			sx4[((et * 2 + c) * 64 + gg * 16 + es) * 2 + hi] = h;
		};
		if constexpr (F16IN) {
			if (n < count) {
				const half4_t *src = reinterpret_cast<const half4_t *>(fin + (size_t)n * kWidth);
#pragma unroll
				for (int p = 0; p < 16; ++p) {
					const int c = p >> 3, gg = p & 3, hi = (p >> 2) & 1;
					sx4[((et * 2 + c) * 64 + gg * 16 + es) * 2 + hi] = src[p];
				}
			} else {
#pragma unroll
				for (int p = 0; p < 16; ++p) emit(p, 0.0f, 0.0f, 0.0f, 0.0f);
			}
		} else {
			if (n < count) {
				encode_quads<ENC>(rec, emit);
			} else {
#pragma unroll
				for (int p = 0; p < 16; ++p) emit(p, 0.0f, 0.0f, 0.0f, 0.0f);
			}
		}
		wave_lds_sync();
	};

	// the 6 layers over the 4 N-tiles of the staged tile, TILES at a time
	auto mlp_tile = [&](uint32_t base) __attribute__((always_inline)) {
#pragma unroll
		for (int tg = 0; tg < 4; tg += TILES) {
			half8_t B[TILES][2];
#pragma unroll
			for (int tt = 0; tt < TILES; ++tt)
#pragma unroll
				for (int c = 0; c < 2; ++c) B[tt][c] = sx[((tg + tt) * 2 + c) * 64 + lane];
#pragma unroll
			for (int l = 0; l < H; ++l) {
				float4_t acc[TILES][4];
#pragma unroll
				for (int tt = 0; tt < TILES; ++tt)
#pragma unroll
					for (int mt = 0; mt < 4; ++mt) acc[tt][mt] = (float4_t)0.0f;
#pragma unroll
				for (int c = 0; c < 2; ++c)
#pragma unroll
					for (int mt = 0; mt < 4; ++mt)
#pragma unroll
						for (int tt = 0; tt < TILES; ++tt) acc[tt][mt] = mfma(A[l][mt][c], B[tt][c], acc[tt][mt]);
#pragma unroll
				for (int tt = 0; tt < TILES; ++tt)
#pragma unroll
					for (int c = 0; c < 2; ++c) B[tt][c] = pack_relu(acc[tt][2 * c], acc[tt][2 * c + 1]);
			}
			float4_t o[TILES];
#pragma unroll
			for (int tt = 0; tt < TILES; ++tt) o[tt] = (float4_t)0.0f;
#pragma unroll
			for (int c = 0; c < 2; ++c)
#pragma unroll
				for (int tt = 0; tt < TILES; ++tt) o[tt] = mfma(Ao[c], B[tt][c], o[tt]);
			if (g == 0) {
#pragma unroll
				for (int tt = 0; tt < TILES; ++tt) {
					const uint32_t m = base + (tg + tt) * 16 + r;
					if (m < count) {
						float *dst = out + (size_t)m * kOut;
#pragma unroll
						for (int c = 0; c < kOut; ++c) dst[c] = clamp0 ? fmaxf(o[tt][c], 0.0f) : o[tt][c];
					}
				}
			}
		}
		wave_lds_sync();
	};

	uint32_t tile = blockIdx.x * 4 + wave;
	if constexpr (F16IN) {
		for (; tile < ntiles; tile += nwaves) {
			encode_tile(tile * 64, nullptr);
			mlp_tile(tile * 64);
		}
	} else {
		// two record buffers in ping-pong: the next tile's records are in flight
		// while this tile is encoded and run, and no register copy waits on them
		float ra[kQueryFloats], rb[kQueryFloats];
		load_query(tile, ra);
		for (; tile < ntiles; tile += 2 * nwaves) {
			load_query(tile + nwaves, rb);
			encode_tile(tile * 64, ra);
			mlp_tile(tile * 64);
			if (tile + nwaves >= ntiles) break;
			load_query(tile + 2 * nwaves, ra);
			encode_tile((tile + nwaves) * 64, rb);
			mlp_tile((tile + nwaves) * 64);
		}
	}
}

// ================================================================ gradient
//
// One workgroup = 4 waves; it walks `chunks` consecutive 64-record chunks (wave
// w owns N-tile w of each chunk) and accumulates its part of dW in VGPRs.  LDS
// holds the fp16 weights row-major and transposed (the transposed copy is
// written by the optimizer, so staging is a straight copy), every layer's
// activations feature-major [feature][record] -- the K = records operand
// layout of the dW MFMAs -- and the current layer's delta.
//   forward   : per wave, B operands in VGPRs, A operands from LDS
//   dW_l^T    : wave w owns output rows 16w..16w+15 of every layer; computing
//               the transposed tile (A = a_l, B = delta_l) leaves 4 consecutive
//               input columns per lane, so partials leave as float4 stores
//   delta_l-1 : per wave, (W_l^T delta_l) . 1[a_l > 0] with A from the
//               transposed copy and the ReLU mask from the saved B operands
// The workgroup writes its partial dW (FP32) and loss sum once; the optimizer
// kernel reduces the partials in a fixed order.
template <int H> struct GradSmem {
	_Float16 w[(H * kWidth + kOut) * kLdw];
	_Float16 wt[H * kWidth * kLdw];
	_Float16 act[(H + 1) * kWidth * kLdn];
	_Float16 delta[kWidth * kLdn];
	float tgt[64 * 3];
	float loss[64];
};

// feature index of element j of a B operand held by lane group g for K-chunk c
__device__ __forceinline__ int bfeat(int c, int g, int j) { return 32 * c + (j < 4 ? 4 * g + j : 16 + 4 * g + j - 4); }

template <int H, int ENC, int LOSS, bool F16IN>
__global__ __launch_bounds__(256) void fused_grad_kernel(const float *__restrict__ records,
                                                         const _Float16 *__restrict__ fin,
                                                         const float *__restrict__ ftgt,
                                                         const uint32_t *__restrict__ d_count, uint32_t max_count,
                                                         uint32_t clamp, uint32_t chunks,
                                                         const _Float16 *__restrict__ w16,
                                                         const _Float16 *__restrict__ w16t,
                                                         const DeviceState *__restrict__ st,
                                                         float *__restrict__ partials, uint32_t pstride,
                                                         uint32_t *__restrict__ overflow) {
	constexpr int WC = weight_count(H);
	extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
	GradSmem<H> &S = *reinterpret_cast<GradSmem<H> *>(smem_raw);

	const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
	const int g = lane >> 4, r = lane & 15;
	const uint32_t count = F16IN ? max_count : effective_count(d_count, max_count, clamp);
	const float scale = (F16IN || st == nullptr) ? 1.0f : st->pub.loss_scale;
	const int col = 16 * wave + r; // this lane's record column inside a chunk

	stage_rows(S.w, w16, H * kWidth + kOut);
	stage_rows(S.wt, w16t, H * kWidth);

	float4_t accW[H][4]; // dW_l^T tiles: rows = input columns 16nt + 4g + i, column = output row 16w + r
	float4_t accO = (float4_t)0.0f;
#pragma unroll
	for (int l = 0; l < H; ++l)
#pragma unroll
		for (int nt = 0; nt < 4; ++nt) accW[l][nt] = (float4_t)0.0f;
	float loss_sum = 0.0f;

	for (uint32_t ch = 0; ch < chunks; ++ch) {
		const uint32_t base = (blockIdx.x * chunks + ch) * 64u;
		if (base >= count) break; // uniform over the workgroup
		__syncthreads();          // previous chunk's LDS readers are done
		// ---- encode: wave w computes features 16w..16w+15 of all 64 records
		{
			const uint32_t n = base + lane;
			const bool valid = n < count;
			const uint32_t ns = valid ? n : 0u;
			_Float16 *a0 = S.act;
			if constexpr (F16IN) {
#pragma unroll
				for (int j = 0; j < 16; ++j)
					a0[(16 * wave + j) * kLdn + lane] = valid ? fin[(size_t)ns * kWidth + 16 * wave + j] : (_Float16)0.0f;
				if (wave == 0)
#pragma unroll
					for (int c = 0; c < 3; ++c) S.tgt[lane * 3 + c] = valid ? ftgt[(size_t)ns * 3 + c] : 0.0f;
			} else {
				float rec[kQueryFloats];
				const float *rp = records + (size_t)ns * kTrainFloats;
#pragma unroll
				for (int i = 0; i < kQueryFloats; ++i) rec[i] = valid ? rp[i] : 0.0f;
				float f[16];
				encode_block<ENC>(rec, wave, f);
#pragma unroll
				for (int j = 0; j < 16; ++j) a0[(16 * wave + j) * kLdn + lane] = valid ? (_Float16)f[j] : (_Float16)0.0f;
				if (wave == 0)
#pragma unroll
					for (int c = 0; c < 3; ++c) S.tgt[lane * 3 + c] = valid ? rp[14 + c] : 0.0f;
			}
		}
		__syncthreads();

		// ---- forward for this wave's 16 records
		half8_t Bs[H + 1][2]; // fp16 activations a_0..a_H as B operands
#pragma unroll
		for (int c = 0; c < 2; ++c)
#pragma unroll
			for (int j = 0; j < 8; ++j) Bs[0][c][j] = S.act[bfeat(c, g, j) * kLdn + col];
#pragma unroll
		for (int l = 0; l < H; ++l) {
			float4_t acc[4];
#pragma unroll
			for (int mt = 0; mt < 4; ++mt) acc[mt] = (float4_t)0.0f;
#pragma unroll
			for (int c = 0; c < 2; ++c)
#pragma unroll
				for (int mt = 0; mt < 4; ++mt)
					acc[mt] = mfma(load_afrag(S.w, l * kWidth + mt * 16 + r, c, g), Bs[l][c], acc[mt]);
#pragma unroll
			for (int c = 0; c < 2; ++c) Bs[l + 1][c] = pack_relu(acc[2 * c], acc[2 * c + 1]);
			_Float16 *al = S.act + (l + 1) * kWidth * kLdn;
#pragma unroll
			for (int c = 0; c < 2; ++c)
#pragma unroll
				for (int j = 0; j < 8; ++j) al[bfeat(c, g, j) * kLdn + col] = Bs[l + 1][c][j];
		}
		float4_t o = (float4_t)0.0f;
#pragma unroll
		for (int c = 0; c < 2; ++c) {
			half8_t a = load_afrag(S.w, H * kWidth + (r < kOut ? r : 0), c, g);
			if (r >= kOut) a = (half8_t)0;
			o = mfma(a, Bs[H][c], o);
		}

		// ---- loss gradient (lanes g == 0 hold the 3 outputs of record col)
		half8_t Bg = (half8_t)0; // g_o as the B operand of the output-layer backward (K rows 0..2)
		{
			float go[3] = {0.0f, 0.0f, 0.0f};
			if (g == 0) {
				const bool valid = base + col < count;
				const float oo[3] = {o[0], o[1], o[2]};
				const float tt[3] = {S.tgt[col * 3], S.tgt[col * 3 + 1], S.tgt[col * 3 + 2]};
				const float v = loss_grad<LOSS>(oo, tt, scale, go);
				if (!valid) go[0] = go[1] = go[2] = 0.0f;
				S.loss[col] = valid ? v : 0.0f;
				Bg[0] = (_Float16)go[0];
				Bg[1] = (_Float16)go[1];
				Bg[2] = (_Float16)go[2];
			}
			// delta rows 0..15 for record col: rows 0..2 = g_o, the rest 0 (padded output rows)
#pragma unroll
			for (int i = 0; i < 4; ++i) S.delta[(4 * g + i) * kLdn + col] = (g == 0 && i < 3) ? Bg[i] : (_Float16)0.0f;
		}
		__syncthreads();
		if (tid == 0)
			for (int i = 0; i < 64; ++i) loss_sum += S.loss[i];

		// ---- dW_H^T: rows = input columns 16w + 4g + i, column = output row r (< 3 valid)
#pragma unroll
		for (int c = 0; c < 2; ++c) {
			const half8_t a = *reinterpret_cast<const half8_t *>(S.act + (H * kWidth + 16 * wave + r) * kLdn + 32 * c + 8 * g);
			const half8_t b = *reinterpret_cast<const half8_t *>(S.delta + r * kLdn + 32 * c + 8 * g);
			accO = mfma(a, b, accO);
		}
		// ---- delta_{H-1} = (W_H^T g_o) . 1[a_H > 0] for this wave's records
		half8_t Bd[2];
		{
			float4_t acc[4];
#pragma unroll
			for (int mt = 0; mt < 4; ++mt) {
				half8_t a = (half8_t)0;
				if (g == 0) {
					const int m = 16 * mt + r;
					a[0] = S.w[(H * kWidth + 0) * kLdw + m];
					a[1] = S.w[(H * kWidth + 1) * kLdw + m];
					a[2] = S.w[(H * kWidth + 2) * kLdw + m];
				}
				acc[mt] = mfma(a, Bg, (float4_t)0.0f);
			}
#pragma unroll
			for (int c = 0; c < 2; ++c)
#pragma unroll
				for (int i = 0; i < 4; ++i) {
					Bd[c][i] = (_Float16)relu_bwd(acc[2 * c][i], (float)Bs[H][c][i]);
					Bd[c][4 + i] = (_Float16)relu_bwd(acc[2 * c + 1][i], (float)Bs[H][c][4 + i]);
				}
		}
		__syncthreads(); // every wave has finished reading the g_o rows of delta
#pragma unroll
		for (int c = 0; c < 2; ++c)
#pragma unroll
			for (int j = 0; j < 8; ++j) S.delta[bfeat(c, g, j) * kLdn + col] = Bd[c][j];
		__syncthreads();

		// ---- layers H-1 .. 0
#pragma unroll
		for (int l = H - 1; l >= 0; --l) {
			// dW_l^T tiles for output rows 16w .. 16w+15, all 64 input columns
#pragma unroll
			for (int c = 0; c < 2; ++c) {
				const half8_t b = *reinterpret_cast<const half8_t *>(S.delta + (16 * wave + r) * kLdn + 32 * c + 8 * g);
#pragma unroll
				for (int nt = 0; nt < 4; ++nt) {
					const half8_t a =
					    *reinterpret_cast<const half8_t *>(S.act + (l * kWidth + 16 * nt + r) * kLdn + 32 * c + 8 * g);
					accW[l][nt] = mfma(a, b, accW[l][nt]);
				}
			}
			if (l == 0) break;
			// delta_{l-1} = (W_l^T delta_l) . 1[a_l > 0]
			{
				float4_t acc[4];
#pragma unroll
				for (int mt = 0; mt < 4; ++mt) acc[mt] = (float4_t)0.0f;
#pragma unroll
				for (int c = 0; c < 2; ++c)
#pragma unroll
					for (int mt = 0; mt < 4; ++mt)
						acc[mt] = mfma(load_afrag(S.wt, l * kWidth + mt * 16 + r, c, g), Bd[c], acc[mt]);
#pragma unroll
				for (int c = 0; c < 2; ++c)
#pragma unroll
					for (int i = 0; i < 4; ++i) {
						Bd[c][i] = (_Float16)relu_bwd(acc[2 * c][i], (float)Bs[l][c][i]);
						Bd[c][4 + i] = (_Float16)relu_bwd(acc[2 * c + 1][i], (float)Bs[l][c][4 + i]);
					}
			}
			__syncthreads(); // dW_l reads of delta are done
#pragma unroll
			for (int c = 0; c < 2; ++c)
#pragma unroll
				for (int j = 0; j < 8; ++j) S.delta[bfeat(c, g, j) * kLdn + col] = Bd[c][j];
			__syncthreads();
		}
	}

	// ---- partial dW of this workgroup (float4 per lane and tile) and its loss sum
	float *part = partials + (size_t)blockIdx.x * pstride;
	bool bad = false;
#pragma unroll
	for (int l = 0; l < H; ++l)
#pragma unroll
		for (int nt = 0; nt < 4; ++nt) {
			const float4_t v4 = accW[l][nt];
			*reinterpret_cast<float4_t *>(part + l * kWidth * kWidth + (16 * wave + r) * kWidth + 16 * nt + 4 * g) = v4;
#pragma unroll
			for (int i = 0; i < 4; ++i) bad |= !isfinite(v4[i]);
		}
	if (r < kOut) {
		*reinterpret_cast<float4_t *>(part + H * kWidth * kWidth + r * kWidth + 16 * wave + 4 * g) = accO;
#pragma unroll
		for (int i = 0; i < 4; ++i) bad |= !isfinite(accO[i]);
	}
	if (tid == 0) part[WC] = loss_sum;
	if (bad && overflow) atomicOr(overflow, 1u);
}

// N-tiles each wave runs through the layers together (1 or 2); VKNRC_INFER_TILES overrides.
int infer_tiles() {
	static int t = 0;
	if (t == 0) {
		const char *e = getenv("VKNRC_INFER_TILES");
		t = (e && atoi(e) == 1) ? 1 : 2;
	}
	return t;
}

template <int H, int ENC, bool F16IN>
hipError_t infer_launch(const float *q, const _Float16 *fin, const uint32_t *dc, uint32_t mc, const _Float16 *w16,
                        float *out, int clamp0, hipStream_t s) {
	const uint32_t tiles = (mc + 63) / 64;
	uint32_t blocks = (tiles + 3) / 4;
	const uint32_t cap = (uint32_t)device_cu_count() * 2;
	if (blocks > cap) blocks = cap;
	if (blocks == 0) blocks = 1;
	if (infer_tiles() == 1)
		fused_infer_kernel<H, ENC, F16IN, 1><<<dim3(blocks), dim3(256), 0, s>>>(q, fin, dc, mc, w16, out, clamp0);
	else
		fused_infer_kernel<H, ENC, F16IN, 2><<<dim3(blocks), dim3(256), 0, s>>>(q, fin, dc, mc, w16, out, clamp0);
	return hipGetLastError();
}

template <int H, int ENC, int LOSS, bool F16IN>
hipError_t grad_launch(const float *rec, const _Float16 *fin, const float *ftgt, const uint32_t *dc, uint32_t mc,
                       uint32_t clamp, uint32_t chunks, const _Float16 *w16, const _Float16 *w16t,
                       const DeviceState *st, float *partials, uint32_t pstride, uint32_t *overflow, uint32_t nblocks,
                       hipStream_t s) {
	const size_t smem = sizeof(GradSmem<H>);
	auto kern = fused_grad_kernel<H, ENC, LOSS, F16IN>;
	static bool attr_set = false;
	if (!attr_set) {
		hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(kern),
		                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);
		if (e != hipSuccess) return e;
		attr_set = true;
	}
	kern<<<dim3(nblocks), dim3(256), smem, s>>>(rec, fin, ftgt, dc, mc, clamp, chunks, w16, w16t, st, partials,
	                                           pstride, overflow);
	return hipGetLastError();
}

template <int H, int ENC>
hipError_t grad_enc(const TrainArgs &a, uint32_t chunks, uint32_t nb, hipStream_t s) {
	if (a.loss == NRC_LOSS_L2)
		return grad_launch<H, ENC, NRC_LOSS_L2, false>(a.records, nullptr, nullptr, a.d_count, a.max_count, a.clamp,
		                                                chunks, a.w16, a.w16t, a.state, a.partials, a.pstride,
		                                                &a.state->overflow, nb, s);
	return grad_launch<H, ENC, NRC_LOSS_RELATIVE_L2_LUMINANCE, false>(a.records, nullptr, nullptr, a.d_count,
	                                                                  a.max_count, a.clamp, chunks, a.w16, a.w16t,
	                                                                  a.state, a.partials, a.pstride,
	                                                                  &a.state->overflow, nb, s);
}

template <int H> hipError_t grad_h(const TrainArgs &a, uint32_t chunks, uint32_t nb, hipStream_t s) {
	switch (a.encoding) {
	case NRC_ENCODING_TRI12: return grad_enc<H, NRC_ENCODING_TRI12>(a, chunks, nb, s);
	case NRC_ENCODING_SINCOS6: return grad_enc<H, NRC_ENCODING_SINCOS6>(a, chunks, nb, s);
	default: return grad_enc<H, NRC_ENCODING_UV_SINCOS16>(a, chunks, nb, s);
	}
}

template <int H>
hipError_t infer_h(int enc, const float *q, const uint32_t *dc, uint32_t mc, const _Float16 *w16, float *rgb,
                   hipStream_t s) {
	switch (enc) {
	case NRC_ENCODING_TRI12: return infer_launch<H, NRC_ENCODING_TRI12, false>(q, nullptr, dc, mc, w16, rgb, 1, s);
	case NRC_ENCODING_SINCOS6: return infer_launch<H, NRC_ENCODING_SINCOS6, false>(q, nullptr, dc, mc, w16, rgb, 1, s);
	default: return infer_launch<H, NRC_ENCODING_UV_SINCOS16, false>(q, nullptr, dc, mc, w16, rgb, 1, s);
	}
}

} // namespace

int device_cu_count() {
	static int cus = 0;
	if (cus == 0) {
		int dev = 0;
		(void)hipGetDevice(&dev);
		hipDeviceProp_t prop;
		if (hipGetDeviceProperties(&prop, dev) == hipSuccess && prop.multiProcessorCount > 0)
			cus = prop.multiProcessorCount;
		else
			cus = 256;
	}
	return cus;
}

hipError_t launch_fused_infer(int hidden, int encoding, const float *queries, const uint32_t *d_count,
                              uint32_t max_count, const _Float16 *w16, float *rgb, hipStream_t stream) {
	if (max_count == 0) return hipSuccess;
	if (hidden == 5) return infer_h<5>(encoding, queries, d_count, max_count, w16, rgb, stream);
	return infer_h<4>(encoding, queries, d_count, max_count, w16, rgb, stream);
}

hipError_t launch_fused_forward_f16(int hidden, const _Float16 *inputs, uint32_t n, const _Float16 *w16, float *out,
                                    hipStream_t stream) {
	if (n == 0) return hipSuccess;
	if (hidden == 5)
		return infer_launch<5, NRC_ENCODING_TRI12, true>(nullptr, inputs, nullptr, n, w16, out, 0, stream);
	return infer_launch<4, NRC_ENCODING_TRI12, true>(nullptr, inputs, nullptr, n, w16, out, 0, stream);
}

uint32_t fused_train_chunks(uint32_t n) {
	const uint32_t chunks64 = (n + kFusedTrainSamples - 1) / kFusedTrainSamples;
	const uint32_t target = (uint32_t)device_cu_count() / 2; // workgroups per step
	uint32_t c = target ? (chunks64 + target - 1) / target : 1;
	if (c < 1) c = 1;
	if (c > kFusedMaxChunks) c = kFusedMaxChunks;
	return c;
}

hipError_t launch_fused_gradient(const TrainArgs &a, uint32_t chunks, uint32_t *nparts, hipStream_t stream) {
	const uint32_t n = a.max_count < a.clamp ? a.max_count : a.clamp;
	const uint32_t chunks64 = (n + kFusedTrainSamples - 1) / kFusedTrainSamples;
	uint32_t nb = (chunks64 + chunks - 1) / chunks;
	if (nb == 0) nb = 1;
	*nparts = nb;
	if (a.hidden == 5) return grad_h<5>(a, chunks, nb, stream);
	return grad_h<4>(a, chunks, nb, stream);
}

hipError_t launch_fused_gradient_f16(int hidden, const _Float16 *inputs, const float *targets, uint32_t n,
                                     const _Float16 *w16, const _Float16 *w16t, float *partials, uint32_t pstride,
                                     uint32_t *nparts, hipStream_t stream) {
	uint32_t nb = (n + kFusedTrainSamples - 1) / kFusedTrainSamples;
	if (nb == 0) nb = 1;
	*nparts = nb;
	if (hidden == 5)
		return grad_launch<5, NRC_ENCODING_TRI12, NRC_LOSS_L2, true>(nullptr, inputs, targets, nullptr, n, n, 1, w16,
		                                                             w16t, nullptr, partials, pstride, nullptr, nb,
		                                                             stream);
	return grad_launch<4, NRC_ENCODING_TRI12, NRC_LOSS_L2, true>(nullptr, inputs, targets, nullptr, n, n, 1, w16, w16t,
	                                                             nullptr, partials, pstride, nullptr, nb, stream);
}

} // namespace nrc
