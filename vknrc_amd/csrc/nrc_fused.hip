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
#include <type_traits>

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

// Row-major [ROWS][64] fp16 weights -> padded LDS rows, by a 256-thread block.
// Loads go out four at a time before their LDS stores, so the block waits for
// one L2 round trip per four 16-byte chunks per thread instead of one per chunk.
template <int ROWS> __device__ __forceinline__ void stage_rows(_Float16 *sW, const _Float16 *__restrict__ w16) {
	constexpr int kChunks = ROWS * 8;
	const uint4 *src = reinterpret_cast<const uint4 *>(w16);
	auto put = [&](int ci, uint4 v) __attribute__((always_inline)) {
		*reinterpret_cast<uint4 *>(sW + (ci >> 3) * kLdw + (ci & 7) * 8) = v;
	};
#pragma unroll
	for (int b = 0; b < kChunks; b += 1024) {
		const int c0 = b + threadIdx.x, c1 = c0 + 256, c2 = c0 + 512, c3 = c0 + 768;
		uint4 v0 = {}, v1 = {}, v2 = {}, v3 = {};
		if (c0 < kChunks) v0 = src[c0];
		if (c1 < kChunks) v1 = src[c1];
		if (c2 < kChunks) v2 = src[c2];
		if (c3 < kChunks) v3 = src[c3];
		if (c0 < kChunks) put(c0, v0);
		if (c1 < kChunks) put(c1, v1);
		if (c2 < kChunks) put(c2, v2);
		if (c3 < kChunks) put(c3, v3);
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

	stage_rows<kRows>(sW, w16);
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
// One workgroup = 8 waves (2 per SIMD) walking `chunks` consecutive 128-record
// chunks; wave w owns N-tile w (records 16w..16w+15) of each chunk.  LDS holds
// the fp16 weights row-major and transposed (the optimizer writes the
// transposed copy, so staging is a straight copy) and, double-buffered, the
// current layer's activations and deltas feature-major [feature][record] -- the
// K = records operand layout of the dW MFMAs.  One barrier per layer.
//   forward   : per wave, B operands in VGPRs, A operands from LDS
//   dW_l^T    : wave w owns the 16x16 tiles (out rows 16(w&3).., in columns
//               16(2(w>>2)+j)..), K = the 128 records; the transposed tile
//               (A = a_l, B = delta_l) leaves 4 consecutive input columns per
//               lane, so partials leave as float4 stores
//   delta_l-1 : per wave, (W_l^T delta_l) . 1[a_l > 0] with A from the
//               transposed copy and the ReLU mask from the saved B operands
// The workgroup writes its partial dW (FP32) and loss sum once; the optimizer
// kernel reduces the partials in a fixed order.
constexpr int kGradWaves = 8;
constexpr int kGradRecords = 16 * kGradWaves; // records per chunk
constexpr int kGLdn = kGradRecords + 8;       // halves per LDS activation row

// Weights sit in LDS as unpadded 128-byte rows of 8 16-byte granules, granule j
// of row R stored at slot j ^ (R & 7) of that row.  The XOR spreads the 16 rows
// an A-fragment read touches over all banks, and keeps every granule a whole
// 16-byte slot, so the rows can be filled by direct global->LDS loads.
__device__ __forceinline__ int swz(int row, int col) {
	return row * 128 + ((((col >> 3) ^ (row & 7))) << 4) + ((col & 7) << 1);
}

template <int H> struct GradSmem {
	_Float16 w[(H * kWidth + kOut) * kWidth];
	_Float16 wt[H * kWidth * kWidth];
	_Float16 act[2][kWidth * kGLdn];   // a_l feature-major [feature][record], double-buffered
	_Float16 delta[2][kWidth * kGLdn]; // delta_l feature-major [feature][record], double-buffered
	float tgt[kGradRecords * 3];
	float loss[kGradRecords];
};

// Row-major [ROWS][64] fp16 weights -> the swizzled LDS rows, with direct
// global->LDS 16-byte loads (global_load_lds_dwordx4 puts lane i of a wave at
// the wave-uniform LDS base + 16 i).  Wave-instruction k fills slots
// 64k .. 64k+63; slot p holds granule (p & 7) ^ (row & 7) of row p >> 3.  All
// loads of the block are in flight together; the caller waits (vmcnt) and
// synchronises before reading.
template <int ROWS> __device__ __forceinline__ void stage_rows_lds(_Float16 *lds, const _Float16 *__restrict__ w16) {
	constexpr int kGran = ROWS * 8;
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	for (int k = wave; k * 64 < kGran; k += kGradWaves) {
		const int p = k * 64 + lane;
		if (p < kGran) {
			const int row = p >> 3, j = (p & 7) ^ (row & 7);
			__builtin_amdgcn_global_load_lds(
			    reinterpret_cast<const void *>(w16 + (row * 8 + j) * 8),
			    (__attribute__((address_space(3))) void *)(reinterpret_cast<char *>(lds) + k * 64 * 16), 16, 0, 0);
		}
	}
}

__device__ __forceinline__ half8_t load_afrag_g(const _Float16 *sW, int row, int c, int g) {
	const char *b = reinterpret_cast<const char *>(sW);
	const half4_t lo = *reinterpret_cast<const half4_t *>(b + swz(row, 32 * c + 4 * g));
	const half4_t hi = *reinterpret_cast<const half4_t *>(b + swz(row, 32 * c + 16 + 4 * g));
	half8_t a;
	a[0] = lo[0]; a[1] = lo[1]; a[2] = lo[2]; a[3] = lo[3];
	a[4] = hi[0]; a[5] = hi[1]; a[6] = hi[2]; a[7] = hi[3];
	return a;
}

// feature index of element j of a B operand held by lane group g for K-chunk c
__device__ __forceinline__ int bfeat(int c, int g, int j) { return 32 * c + (j < 4 ? 4 * g + j : 16 + 4 * g + j - 4); }

// writes this lane's 16 values of a B operand (record column col) feature-major
__device__ __forceinline__ void store_bcol(_Float16 *dst, const half8_t *b, int g, int col) {
#pragma unroll
	for (int c = 0; c < 2; ++c)
#pragma unroll
		for (int j = 0; j < 8; ++j) dst[bfeat(c, g, j) * kGLdn + col] = b[c][j];
}

template <int H, int ENC, int LOSS, bool F16IN, bool MULTI>
__global__ __launch_bounds__(512) void fused_grad_kernel(const float *__restrict__ records,
                                                         const _Float16 *__restrict__ fin,
                                                         const float *__restrict__ ftgt,
                                                         const uint32_t *__restrict__ d_count, uint32_t max_count,
                                                         uint32_t clamp, uint32_t chunks,
                                                         const _Float16 *__restrict__ w16,
                                                         const _Float16 *__restrict__ w16t,
                                                         const DeviceState *__restrict__ st, int parity,
                                                         float *__restrict__ partials, uint32_t pstride,
                                                         uint32_t *__restrict__ overflow) {
	constexpr int WC = weight_count(H);
	extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
	GradSmem<H> &S = *reinterpret_cast<GradSmem<H> *>(smem_raw);

	const int tid = threadIdx.x, wave = tid >> 6;
	const uint32_t count = F16IN ? max_count : effective_count(d_count, max_count, clamp);
	const float scale = (F16IN || st == nullptr) ? 1.0f : st->pub[parity].loss_scale;
	const int enc_q = wave & 3;                       // feature block this wave encodes
	const int mt_w = wave & 3, nt0 = 2 * (wave >> 2); // dW tiles of this wave: (mt_w, nt0), (mt_w, nt0 + 1)
	const uint32_t last = count ? count - 1 : 0;
	int lane = tid & 63;
	int g = lane >> 4, r = lane & 15;
	int col = 16 * wave + r;               // this lane's record column inside a chunk
	int enc_rec = 64 * (wave >> 2) + lane; // record this lane encodes

	// this lane's record to encode in chunk ch: the index is clamped into [0, count)
	// so the loads are unconditional; lanes past the count are zeroed at encode time
	float rec[kTrainFloats];
	auto load_rec = [&](uint32_t ch) __attribute__((always_inline)) {
		uint32_t n = (blockIdx.x * chunks + ch) * (uint32_t)kGradRecords + enc_rec;
		n = n < last ? n : last;
		if constexpr (!F16IN) {
			const float *rp = records + (size_t)n * kTrainFloats;
#pragma unroll
			for (int i = 0; i < kTrainFloats; ++i) rec[i] = rp[i];
		}
	};
	load_rec(0); // in flight while the weights are staged

	stage_rows_lds<H * kWidth + kOut>(S.w, w16);
	stage_rows_lds<H * kWidth>(S.wt, w16t);
	__builtin_amdgcn_s_waitcnt(0); // this wave's direct-to-LDS loads (and the record) have landed

	float4_t accW[MULTI ? H : 1][2]; // multi-chunk accumulators of this wave's two dW tiles per layer
	float4_t accO = (float4_t)0.0f;   // dW_H^T tile (waves 0..3)
#pragma unroll
	for (int l = 0; l < (MULTI ? H : 1); ++l) accW[l][0] = accW[l][1] = (float4_t)0.0f;
	float loss_sum = 0.0f;
	float *part = partials + (size_t)blockIdx.x * pstride;
	bool bad = false;
	auto store_tile = [&](int l, int nt, float4_t v4) __attribute__((always_inline)) {
		*reinterpret_cast<float4_t *>(part + l * kWidth * kWidth + (16 * mt_w + r) * kWidth + 16 * nt + 4 * g) = v4;
#pragma unroll
		for (int i = 0; i < 4; ++i) bad |= !isfinite(v4[i]);
	};

	for (uint32_t ch = 0; ch < chunks; ++ch) {
		const uint32_t base = (blockIdx.x * chunks + ch) * (uint32_t)kGradRecords;
		if (base >= count) { // uniform over the workgroup
			if constexpr (!MULTI) {
				// a chunk wholly past the count still owns its partial slot: zeros
#pragma unroll
				for (int l = 0; l < H; ++l) {
					store_tile(l, nt0, (float4_t)0.0f);
					store_tile(l, nt0 + 1, (float4_t)0.0f);
				}
			}
			break;
		}
		// recompute the lane-derived indices each chunk from an opaque copy of the
		// lane id, so the many LDS addresses built from them are not hoisted out of
		// the loop and kept live through the encoder
		asm volatile("" : "+v"(lane));
		g = lane >> 4;
		r = lane & 15;
		col = 16 * wave + r;
		enc_rec = 64 * (wave >> 2) + lane;
		__syncthreads(); // staging done / previous chunk's LDS readers done
		// ---- encode: wave w computes features 16(w&3).. of records 64(w>>2) + lane into act[0]
		{
			const uint32_t n = base + enc_rec;
			const bool valid = n < count;
			_Float16 *a0 = S.act[0];
			if constexpr (F16IN) {
				const uint32_t ns = n < last ? n : last;
#pragma unroll
				for (int j = 0; j < 16; ++j)
					a0[(16 * enc_q + j) * kGLdn + enc_rec] =
					    valid ? fin[(size_t)ns * kWidth + 16 * enc_q + j] : (_Float16)0.0f;
				if (enc_q == 0)
#pragma unroll
					for (int c = 0; c < 3; ++c) S.tgt[enc_rec * 3 + c] = valid ? ftgt[(size_t)ns * 3 + c] : 0.0f;
			} else {
				// one branch per feature block, each writing its own 16 features (no
				// 16-value merge across the branches)
				auto put = [&](auto qc) __attribute__((always_inline)) {
					constexpr int q = decltype(qc)::value;
					float f[16];
					encode_block<ENC>(rec, q, f);
#pragma unroll
					for (int j = 0; j < 16; ++j)
						a0[(16 * q + j) * kGLdn + enc_rec] = valid ? (_Float16)f[j] : (_Float16)0.0f;
				};
				switch (enc_q) {
				case 0: put(std::integral_constant<int, 0>{}); break;
				case 1: put(std::integral_constant<int, 1>{}); break;
				case 2: put(std::integral_constant<int, 2>{}); break;
				default: put(std::integral_constant<int, 3>{}); break;
				}
				if (enc_q == 0)
#pragma unroll
					for (int c = 0; c < 3; ++c) S.tgt[enc_rec * 3 + c] = valid ? rec[14 + c] : 0.0f;
			}
		}
		if (MULTI && ch + 1 < chunks) load_rec(ch + 1); // next chunk's records in flight during this one
		__syncthreads();

		// ---- forward for this wave's 16 records; activations stay in VGPRs as B operands
		half8_t Bs[H + 1][2];
#pragma unroll
		for (int c = 0; c < 2; ++c)
#pragma unroll
			for (int j = 0; j < 8; ++j) Bs[0][c][j] = S.act[0][bfeat(c, g, j) * kGLdn + col];
#pragma unroll
		for (int l = 0; l < H; ++l) {
			float4_t acc[4];
#pragma unroll
			for (int mt = 0; mt < 4; ++mt) acc[mt] = (float4_t)0.0f;
#pragma unroll
			for (int c = 0; c < 2; ++c)
#pragma unroll
				for (int mt = 0; mt < 4; ++mt)
					acc[mt] = mfma(load_afrag_g(S.w, l * kWidth + mt * 16 + r, c, g), Bs[l][c], acc[mt]);
#pragma unroll
			for (int c = 0; c < 2; ++c) Bs[l + 1][c] = pack_relu(acc[2 * c], acc[2 * c + 1]);
			__builtin_amdgcn_sched_barrier(0);
		}
		float4_t o = (float4_t)0.0f;
#pragma unroll
		for (int c = 0; c < 2; ++c) {
			half8_t a = load_afrag_g(S.w, H * kWidth + (r < kOut ? r : 0), c, g);
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
			// delta[1] rows 0..15 for record col: rows 0..2 = g_o, the rest 0 (padded output rows);
			// act[1] = a_H for the dW_H MFMAs
#pragma unroll
			for (int i = 0; i < 4; ++i) S.delta[1][(4 * g + i) * kGLdn + col] = (g == 0 && i < 3) ? Bg[i] : (_Float16)0.0f;
			store_bcol(S.act[1], Bs[H], g, col);
		}
		__syncthreads();
		if (tid == 0)
			for (int i = 0; i < kGradRecords; ++i) loss_sum += S.loss[i];

		// ---- dW_H^T (waves 0..3): rows = input columns 16w + 4g + i, column = output row r (< 3 valid)
		if (wave < 4) {
#pragma unroll
			for (int c = 0; c < kGradRecords / 32; ++c) {
				const half8_t a =
				    *reinterpret_cast<const half8_t *>(S.act[1] + (16 * wave + r) * kGLdn + 32 * c + 8 * g);
				const half8_t b = *reinterpret_cast<const half8_t *>(S.delta[1] + r * kGLdn + 32 * c + 8 * g);
				accO = mfma(a, b, accO);
			}
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
					const char *wb = reinterpret_cast<const char *>(S.w);
					a[0] = *reinterpret_cast<const _Float16 *>(wb + swz(H * kWidth + 0, m));
					a[1] = *reinterpret_cast<const _Float16 *>(wb + swz(H * kWidth + 1, m));
					a[2] = *reinterpret_cast<const _Float16 *>(wb + swz(H * kWidth + 2, m));
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
		// buffer 0 is free (its last readers passed the barrier above): delta_{H-1}, a_{H-1}
		store_bcol(S.delta[0], Bd, g, col);
		store_bcol(S.act[0], Bs[H - 1], g, col);
		__syncthreads();

		// ---- layers H-1 .. 0: buffer p holds delta_l and a_l
#pragma unroll
		for (int l = H - 1; l >= 0; --l) {
			const int p = (H - 1 - l) & 1;
			// this wave's two dW_l^T tiles over the 128 records
			{
				float4_t t0 = MULTI ? accW[MULTI ? l : 0][0] : (float4_t)0.0f;
				float4_t t1 = MULTI ? accW[MULTI ? l : 0][1] : (float4_t)0.0f;
#pragma unroll 2
				for (int c = 0; c < kGradRecords / 32; ++c) {
					const half8_t b =
					    *reinterpret_cast<const half8_t *>(S.delta[p] + (16 * mt_w + r) * kGLdn + 32 * c + 8 * g);
					const half8_t a0 =
					    *reinterpret_cast<const half8_t *>(S.act[p] + (16 * nt0 + r) * kGLdn + 32 * c + 8 * g);
					const half8_t a1 =
					    *reinterpret_cast<const half8_t *>(S.act[p] + (16 * nt0 + 16 + r) * kGLdn + 32 * c + 8 * g);
					t0 = mfma(a0, b, t0);
					t1 = mfma(a1, b, t1);
				}
				if constexpr (MULTI) {
					accW[MULTI ? l : 0][0] = t0;
					accW[MULTI ? l : 0][1] = t1;
				} else {
					store_tile(l, nt0, t0);
					store_tile(l, nt0 + 1, t1);
				}
			}
			if (l == 0) break;
			__builtin_amdgcn_sched_barrier(0);
			// delta_{l-1} = (W_l^T delta_l) . 1[a_l > 0]
			{
				float4_t acc[4];
#pragma unroll
				for (int mt = 0; mt < 4; ++mt) acc[mt] = (float4_t)0.0f;
#pragma unroll
				for (int c = 0; c < 2; ++c)
#pragma unroll
					for (int mt = 0; mt < 4; ++mt)
						acc[mt] = mfma(load_afrag_g(S.wt, l * kWidth + mt * 16 + r, c, g), Bd[c], acc[mt]);
#pragma unroll
				for (int c = 0; c < 2; ++c)
#pragma unroll
					for (int i = 0; i < 4; ++i) {
						Bd[c][i] = (_Float16)relu_bwd(acc[2 * c][i], (float)Bs[l][c][i]);
						Bd[c][4 + i] = (_Float16)relu_bwd(acc[2 * c + 1][i], (float)Bs[l][c][4 + i]);
					}
			}
			// the other buffer's last readers passed the previous barrier
			__builtin_amdgcn_sched_barrier(0);
			store_bcol(S.delta[p ^ 1], Bd, g, col);
			store_bcol(S.act[p ^ 1], Bs[l - 1], g, col);
			__syncthreads();
			__builtin_amdgcn_sched_barrier(0);
		}
	}

	// ---- partial dW of this workgroup (float4 per lane and tile) and its loss sum
	if constexpr (MULTI) {
#pragma unroll
		for (int l = 0; l < H; ++l) {
			store_tile(l, nt0, accW[MULTI ? l : 0][0]);
			store_tile(l, nt0 + 1, accW[MULTI ? l : 0][1]);
		}
	}
	if (wave < 4 && r < kOut) {
		*reinterpret_cast<float4_t *>(part + H * kWidth * kWidth + r * kWidth + 16 * wave + 4 * g) = accO;
#pragma unroll
		for (int i = 0; i < 4; ++i) bad |= !isfinite(accO[i]);
	}
	if (tid == 0) part[WC] = loss_sum;
	if (bad && overflow) atomicOr(overflow, 1u);
}

// ================================================ gradient, register-resident weights
//
// One workgroup = 4 waves (one per SIMD) and one 64-record chunk; wave w runs
// records 16w .. 16w+15 through the whole network with every A fragment it needs
// held in registers, loaded once with 16-byte loads from the fragment-major
// copies: W_l for the forward, W_l^T for the backward (312 registers for H = 5
// of the 512 a wave owns at one wave per SIMD).  Waves write their activations
// a_l and deltas feature-major into LDS as they produce them, so the forward and
// the backward chain run without any barrier between waves.  After the one
// barrier, wave w computes its 2x2 block of dW_l^T tiles (input tiles 2(w&1)..,
// output tiles 2(w>>1)..) for every layer from LDS, K = the 64 records, and
// stores them (float4 per lane) to its workgroup's partial-gradient row.
constexpr int kSmallLd = (int)kFusedSmallRecords + 8; // halves per LDS row

template <int H> struct SmallGradSmem {
	_Float16 act[H + 1][kWidth * kSmallLd]; // a_l, [feature][record]
	_Float16 dlt[H][kWidth * kSmallLd];     // delta_{l+1} = dL/dz at layer l's output, [unit][record]
	_Float16 go[16 * kSmallLd];             // output gradient rows 0..2; rows 3..15 stay zero
	float lsum[4];
};

// writes this lane's 16 values of a B operand (record column col) feature-major, row stride LD
template <int LD> __device__ __forceinline__ void store_bcol_ld(_Float16 *dst, const half8_t *b, int g, int col) {
#pragma unroll
	for (int c = 0; c < 2; ++c)
#pragma unroll
		for (int j = 0; j < 8; ++j) dst[bfeat(c, g, j) * LD + col] = b[c][j];
}

// delta = relu_bwd(acc) against the activations in B-operand layout, packed to fp16
__device__ __forceinline__ void relu_bwd_pack(const float4_t *acc, const half8_t *act, half8_t *d) {
#pragma unroll
	for (int c = 0; c < 2; ++c)
#pragma unroll
		for (int i = 0; i < 4; ++i) {
			d[c][i] = (_Float16)relu_bwd(acc[2 * c][i], (float)act[c][i]);
			d[c][4 + i] = (_Float16)relu_bwd(acc[2 * c + 1][i], (float)act[c][4 + i]);
		}
}

template <int H, int ENC, int LOSS, bool F16IN>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 1))) void fused_grad_small_kernel(
    const float *__restrict__ records, const _Float16 *__restrict__ fin, const float *__restrict__ ftgt,
    const uint32_t *__restrict__ d_count, uint32_t max_count, uint32_t clamp, const _Float16 *__restrict__ wf,
    const _Float16 *__restrict__ wtf, const DeviceState *__restrict__ st, int parity, float *__restrict__ partials,
    uint32_t pstride, uint32_t *__restrict__ overflow) {
	constexpr int WC = weight_count(H);
	constexpr int R = (int)kFusedSmallRecords;
	constexpr int LD = kSmallLd;
	extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
	SmallGradSmem<H> &S = *reinterpret_cast<SmallGradSmem<H> *>(smem_raw);

	const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
	const int g = lane >> 4, r = lane & 15;
	const int col = 16 * wave + r; // record column this lane runs through the network
	const uint32_t count = F16IN ? max_count : effective_count(d_count, max_count, clamp);
	const uint32_t base = blockIdx.x * (uint32_t)R;
	float *part = partials + (size_t)blockIdx.x * pstride;
	const int it0 = 2 * (wave & 1), ot0 = 2 * (wave >> 1); // this wave's dW tiles

	auto store_tile = [&](int l, int it, int ot, float4_t v) __attribute__((always_inline)) {
		*reinterpret_cast<float4_t *>(part + l * kWidth * kWidth + (16 * ot + r) * kWidth + 16 * it + 4 * g) = v;
	};
	if (base >= count) { // uniform: no valid record in this chunk, so its partial row is zero
#pragma unroll
		for (int l = 0; l < H; ++l)
#pragma unroll
			for (int i = 0; i < 2; ++i)
#pragma unroll
				for (int o = 0; o < 2; ++o) store_tile(l, it0 + i, ot0 + o, (float4_t)0.0f);
		if (r < kOut)
			*reinterpret_cast<float4_t *>(part + H * kWidth * kWidth + r * kWidth + 16 * wave + 4 * g) = (float4_t)0.0f;
		if (tid == 0) part[WC] = 0.0f;
		return;
	}
	const float scale = (F16IN || st == nullptr) ? 1.0f : st->pub[parity].loss_scale;
	const uint32_t last = count - 1;

	// ---- loads, in the order they are needed: the record this lane encodes
	// (base + lane), the targets of the record it runs (base + col), then every
	// weight fragment.  Indices are clamped so the loads are unconditional.
	const uint32_t n_enc = base + lane < last ? base + lane : last;
	const uint32_t n_run = base + col < last ? base + col : last;
	float q[kQueryFloats];
	uint4 fq[2];
	float tg[3];
	if constexpr (F16IN) {
		const uint4 *src = reinterpret_cast<const uint4 *>(fin + (size_t)n_enc * kWidth + 16 * wave);
		fq[0] = src[0];
		fq[1] = src[1];
#pragma unroll
		for (int c = 0; c < 3; ++c) tg[c] = ftgt[(size_t)n_run * 3 + c];
	} else {
		const float *rp = records + (size_t)n_enc * kTrainFloats;
#pragma unroll
		for (int i = 0; i < kQueryFloats; ++i) q[i] = rp[i];
		const float *tp = records + (size_t)n_run * kTrainFloats + kQueryFloats;
#pragma unroll
		for (int c = 0; c < 3; ++c) tg[c] = tp[c];
	}
	const half8_t *wfv = reinterpret_cast<const half8_t *>(wf);
	const half8_t *wtv = reinterpret_cast<const half8_t *>(wtf);
	half8_t Af[H][4][2], Ao[2], At[H - 1][4][2], Aot[4];
#pragma unroll
	for (int l = 0; l < H; ++l)
#pragma unroll
		for (int mt = 0; mt < 4; ++mt)
#pragma unroll
			for (int c = 0; c < 2; ++c) Af[l][mt][c] = wfv[((l * 4 + mt) * 2 + c) * 64 + lane];
#pragma unroll
	for (int c = 0; c < 2; ++c) Ao[c] = wfv[(H * 8 + c) * 64 + lane];
#pragma unroll
	for (int mt = 0; mt < 4; ++mt) Aot[mt] = wtv[((H - 1) * 4 + mt) * 2 * 64 + lane];
#pragma unroll
	for (int l = 0; l < H - 1; ++l)
#pragma unroll
		for (int mt = 0; mt < 4; ++mt)
#pragma unroll
			for (int c = 0; c < 2; ++c) At[l][mt][c] = wtv[((l * 4 + mt) * 2 + c) * 64 + lane];

	// rows 3..15 of the output-gradient operand stay zero
	for (int i = tid; i < 13 * 32; i += 256)
		reinterpret_cast<uint32_t *>(S.go + (3 + i / 32) * LD)[i % 32] = 0u;

	// ---- encode: wave w computes feature block w (features 16w ..) of record base + lane
	{
		_Float16 *a0 = S.act[0] + lane;
		const bool valid = base + lane < count;
		if constexpr (F16IN) {
#pragma unroll
			for (int h = 0; h < 2; ++h) {
				const half8_t v = __builtin_bit_cast(half8_t, fq[h]);
#pragma unroll
				for (int j = 0; j < 8; ++j) a0[(16 * wave + 8 * h + j) * LD] = valid ? v[j] : (_Float16)0.0f;
			}
		} else {
			auto put = [&](auto qc) __attribute__((always_inline)) {
				constexpr int qb = decltype(qc)::value;
				float f[16];
				encode_block<ENC>(q, qb, f);
#pragma unroll
				for (int j = 0; j < 16; ++j) a0[(16 * qb + j) * LD] = valid ? (_Float16)f[j] : (_Float16)0.0f;
			};
			switch (wave) {
			case 0: put(std::integral_constant<int, 0>{}); break;
			case 1: put(std::integral_constant<int, 1>{}); break;
			case 2: put(std::integral_constant<int, 2>{}); break;
			default: put(std::integral_constant<int, 3>{}); break;
			}
		}
	}
	__syncthreads();

	// ---- forward for this wave's 16 records; a_l kept as B operands and written feature-major
	half8_t Bs[H + 1][2];
#pragma unroll
	for (int c = 0; c < 2; ++c)
#pragma unroll
		for (int j = 0; j < 8; ++j) Bs[0][c][j] = S.act[0][bfeat(c, g, j) * LD + col];
#pragma unroll
	for (int l = 0; l < H; ++l) {
		float4_t acc[4];
#pragma unroll
		for (int mt = 0; mt < 4; ++mt) acc[mt] = (float4_t)0.0f;
#pragma unroll
		for (int c = 0; c < 2; ++c)
#pragma unroll
			for (int mt = 0; mt < 4; ++mt) acc[mt] = mfma(Af[l][mt][c], Bs[l][c], acc[mt]);
#pragma unroll
		for (int c = 0; c < 2; ++c) Bs[l + 1][c] = pack_relu(acc[2 * c], acc[2 * c + 1]);
		store_bcol_ld<LD>(S.act[l + 1], Bs[l + 1], g, col);
	}
	float4_t o = (float4_t)0.0f;
#pragma unroll
	for (int c = 0; c < 2; ++c) o = mfma(Ao[c], Bs[H][c], o);

	// ---- loss gradient: lanes g == 0 hold the 3 outputs of record col
	half8_t Bg = (half8_t)0;
	float lv = 0.0f;
	if (g == 0) {
		const bool valid = base + col < count;
		const float oo[3] = {o[0], o[1], o[2]};
		float go[3];
		const float v = loss_grad<LOSS>(oo, tg, scale, go);
		lv = valid ? v : 0.0f;
#pragma unroll
		for (int c = 0; c < kOut; ++c) {
			Bg[c] = valid ? (_Float16)go[c] : (_Float16)0.0f;
			S.go[c * LD + col] = Bg[c];
		}
	}
	// ---- backward chain: delta_H = (W_H^T g_o) . 1[a_H > 0], then delta_l for l = H-1 .. 1
	half8_t Bd[2];
	{
		float4_t acc[4];
#pragma unroll
		for (int mt = 0; mt < 4; ++mt) acc[mt] = mfma(Aot[mt], Bg, (float4_t)0.0f);
		relu_bwd_pack(acc, Bs[H], Bd);
		store_bcol_ld<LD>(S.dlt[H - 1], Bd, g, col);
	}
#pragma unroll
	for (int l = H - 1; l >= 1; --l) {
		float4_t acc[4];
#pragma unroll
		for (int mt = 0; mt < 4; ++mt) acc[mt] = (float4_t)0.0f;
#pragma unroll
		for (int c = 0; c < 2; ++c)
#pragma unroll
			for (int mt = 0; mt < 4; ++mt) acc[mt] = mfma(At[l - 1][mt][c], Bd[c], acc[mt]);
		relu_bwd_pack(acc, Bs[l], Bd);
		store_bcol_ld<LD>(S.dlt[l - 1], Bd, g, col);
	}
	// this wave's loss sum, in a fixed order
#pragma unroll
	for (int off = 32; off >= 1; off >>= 1) lv += __shfl_xor(lv, off);
	if (lane == 0) S.lsum[wave] = lv;
	__syncthreads();

	// ---- dW_l^T tiles of this wave over the chunk's records, straight to the partial row
	bool bad = false;
#pragma unroll
	for (int l = 0; l < H; ++l) {
		float4_t t[2][2];
#pragma unroll
		for (int i = 0; i < 2; ++i) t[i][0] = t[i][1] = (float4_t)0.0f;
#pragma unroll
		for (int kc = 0; kc < R / 32; ++kc) {
			half8_t a[2], b[2];
#pragma unroll
			for (int i = 0; i < 2; ++i) {
				a[i] = *reinterpret_cast<const half8_t *>(S.act[l] + (16 * (it0 + i) + r) * LD + 32 * kc + 8 * g);
				b[i] = *reinterpret_cast<const half8_t *>(S.dlt[l] + (16 * (ot0 + i) + r) * LD + 32 * kc + 8 * g);
			}
#pragma unroll
			for (int i = 0; i < 2; ++i)
#pragma unroll
				for (int o2 = 0; o2 < 2; ++o2) t[i][o2] = mfma(a[i], b[o2], t[i][o2]);
		}
#pragma unroll
		for (int i = 0; i < 2; ++i)
#pragma unroll
			for (int o2 = 0; o2 < 2; ++o2) {
				store_tile(l, it0 + i, ot0 + o2, t[i][o2]);
#pragma unroll
				for (int e = 0; e < 4; ++e) bad |= !isfinite(t[i][o2][e]);
			}
	}
	{ // dW_H^T: input tile `wave` against the output-gradient rows
		float4_t t = (float4_t)0.0f;
#pragma unroll
		for (int kc = 0; kc < R / 32; ++kc) {
			const half8_t a = *reinterpret_cast<const half8_t *>(S.act[H] + (16 * wave + r) * LD + 32 * kc + 8 * g);
			const half8_t b = *reinterpret_cast<const half8_t *>(S.go + r * LD + 32 * kc + 8 * g);
			t = mfma(a, b, t);
		}
		if (r < kOut) {
			*reinterpret_cast<float4_t *>(part + H * kWidth * kWidth + r * kWidth + 16 * wave + 4 * g) = t;
#pragma unroll
			for (int e = 0; e < 4; ++e) bad |= !isfinite(t[e]);
		}
	}
	if (tid == 0) part[WC] = (S.lsum[0] + S.lsum[1]) + (S.lsum[2] + S.lsum[3]);
	if (overflow && __any(bad) && lane == 0) atomicOr(overflow, 1u);
}

// N-tiles each wave runs through the layers together (1 or 2; both measure the
// same, and 1 needs no register spill); VKNRC_INFER_TILES overrides.
int infer_tiles() {
	static int t = 0;
	if (t == 0) {
		const char *e = getenv("VKNRC_INFER_TILES");
		t = (e && atoi(e) == 2) ? 2 : 1;
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
                       const DeviceState *st, int parity, float *partials, uint32_t pstride, uint32_t *overflow,
                       uint32_t nblocks, hipStream_t s) {
	const size_t smem = sizeof(GradSmem<H>);
	auto kern = chunks > 1 ? fused_grad_kernel<H, ENC, LOSS, F16IN, true> : fused_grad_kernel<H, ENC, LOSS, F16IN, false>;
	static bool attr_set[2] = {false, false};
	if (!attr_set[chunks > 1]) {
		hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(kern),
		                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);
		if (e != hipSuccess) return e;
		attr_set[chunks > 1] = true;
	}
	kern<<<dim3(nblocks), dim3(512), smem, s>>>(rec, fin, ftgt, dc, mc, clamp, chunks, w16, w16t, st, parity,
	                                           partials, pstride, overflow);
	return hipGetLastError();
}

template <int H, int ENC, int LOSS, bool F16IN>
hipError_t grad_small_launch(const float *rec, const _Float16 *fin, const float *ftgt, const uint32_t *dc,
                             uint32_t mc, uint32_t clamp, const _Float16 *wf, const _Float16 *wtf,
                             const DeviceState *st, int parity, float *partials, uint32_t pstride, uint32_t *overflow,
                             uint32_t nblocks, hipStream_t s) {
	const size_t smem = sizeof(SmallGradSmem<H>);
	auto kern = fused_grad_small_kernel<H, ENC, LOSS, F16IN>;
	static bool attr_set = false;
	if (!attr_set) {
		hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(kern),
		                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);
		if (e != hipSuccess) return e;
		attr_set = true;
	}
	kern<<<dim3(nblocks), dim3(256), smem, s>>>(rec, fin, ftgt, dc, mc, clamp, wf, wtf, st, parity, partials, pstride,
	                                           overflow);
	return hipGetLastError();
}

template <int H, int ENC, int LOSS>
hipError_t grad_loss(const TrainArgs &a, const FusedGradPlan &p, hipStream_t s) {
	uint32_t *ovf = &a.state->overflow[a.parity];
	if (p.records == kFusedSmallRecords)
		return grad_small_launch<H, ENC, LOSS, false>(a.records, nullptr, nullptr, a.d_count, a.max_count, a.clamp,
		                                              a.wf, a.wtf, a.state, a.parity, a.partials, a.pstride, ovf,
		                                              p.blocks, s);
	return grad_launch<H, ENC, LOSS, false>(a.records, nullptr, nullptr, a.d_count, a.max_count, a.clamp, p.chunks,
	                                        a.w16, a.w16t, a.state, a.parity, a.partials, a.pstride, ovf, p.blocks,
	                                        s);
}

template <int H, int ENC> hipError_t grad_enc(const TrainArgs &a, const FusedGradPlan &p, hipStream_t s) {
	if (a.loss == NRC_LOSS_L2) return grad_loss<H, ENC, NRC_LOSS_L2>(a, p, s);
	return grad_loss<H, ENC, NRC_LOSS_RELATIVE_L2_LUMINANCE>(a, p, s);
}

template <int H> hipError_t grad_h(const TrainArgs &a, const FusedGradPlan &p, hipStream_t s) {
	switch (a.encoding) {
	case NRC_ENCODING_TRI12: return grad_enc<H, NRC_ENCODING_TRI12>(a, p, s);
	case NRC_ENCODING_SINCOS6: return grad_enc<H, NRC_ENCODING_SINCOS6>(a, p, s);
	default: return grad_enc<H, NRC_ENCODING_UV_SINCOS16>(a, p, s);
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

// Multi-chunk kernel: chunks per workgroup so that about one workgroup runs per
// CU (VKNRC_TRAIN_CHUNKS overrides).
static uint32_t fused_train_chunks(uint32_t n) {
	static int forced = -1;
	if (forced < 0) {
		const char *e = getenv("VKNRC_TRAIN_CHUNKS");
		forced = e ? atoi(e) : 0;
	}
	if (forced > 0) return (uint32_t)(forced < (int)kFusedMaxChunks ? forced : (int)kFusedMaxChunks);
	const uint32_t nchunks = (n + kFusedTrainSamples - 1) / kFusedTrainSamples;
	const uint32_t target = (uint32_t)device_cu_count();
	uint32_t c = target ? (nchunks + target - 1) / target : 1;
	if (c < 1) c = 1;
	if (c > kFusedMaxChunks) c = kFusedMaxChunks;
	return c;
}

// Most workgroups of the single-chunk kernel before the multi-chunk kernel takes
// over (every workgroup writes a partial row); VKNRC_SMALL_GRAD_BLOCKS overrides.
static uint32_t small_grad_max_blocks() {
	static int v = -1;
	if (v < 0) {
		const char *e = getenv("VKNRC_SMALL_GRAD_BLOCKS");
		v = e ? atoi(e) : 2 * device_cu_count();
		if (v < 0) v = 0;
	}
	return (uint32_t)v;
}

FusedGradPlan plan_fused_gradient(uint32_t n) {
	FusedGradPlan p;
	const uint32_t small = (n + kFusedSmallRecords - 1) / kFusedSmallRecords;
	if (small <= small_grad_max_blocks()) {
		p.records = kFusedSmallRecords;
		p.chunks = 1;
		p.blocks = small ? small : 1;
		return p;
	}
	p.records = kFusedTrainSamples;
	p.chunks = fused_train_chunks(n);
	const uint32_t nchunks = (n + kFusedTrainSamples - 1) / kFusedTrainSamples;
	p.blocks = (nchunks + p.chunks - 1) / p.chunks;
	if (p.blocks == 0) p.blocks = 1;
	return p;
}

hipError_t launch_fused_gradient(const TrainArgs &a, const FusedGradPlan &plan, hipStream_t stream) {
	if (a.hidden == 5) return grad_h<5>(a, plan, stream);
	return grad_h<4>(a, plan, stream);
}

hipError_t launch_fused_gradient_f16(int hidden, const _Float16 *inputs, const float *targets, uint32_t n,
                                     const _Float16 *w16, const _Float16 *w16t, const _Float16 *wf,
                                     const _Float16 *wtf, float *partials, uint32_t pstride, uint32_t *nparts,
                                     hipStream_t stream) {
	const FusedGradPlan p = plan_fused_gradient(n);
	*nparts = p.blocks;
	if (p.records == kFusedSmallRecords) {
		if (hidden == 5)
			return grad_small_launch<5, NRC_ENCODING_TRI12, NRC_LOSS_L2, true>(
			    nullptr, inputs, targets, nullptr, n, n, wf, wtf, nullptr, 0, partials, pstride, nullptr, p.blocks,
			    stream);
		return grad_small_launch<4, NRC_ENCODING_TRI12, NRC_LOSS_L2, true>(
		    nullptr, inputs, targets, nullptr, n, n, wf, wtf, nullptr, 0, partials, pstride, nullptr, p.blocks, stream);
	}
	// one 128-record chunk per workgroup
	const uint32_t nb = (n + kFusedTrainSamples - 1) / kFusedTrainSamples;
	*nparts = nb;
	if (hidden == 5)
		return grad_launch<5, NRC_ENCODING_TRI12, NRC_LOSS_L2, true>(nullptr, inputs, targets, nullptr, n, n, 1, w16,
		                                                             w16t, nullptr, 0, partials, pstride, nullptr, nb,
		                                                             stream);
	return grad_launch<4, NRC_ENCODING_TRI12, NRC_LOSS_L2, true>(nullptr, inputs, targets, nullptr, n, n, 1, w16, w16t,
	                                                             nullptr, 0, partials, pstride, nullptr, nb, stream);
}

} // namespace nrc
