// nrc_fp8.hip -- the inference forward on FP8 (OCP e4m3) MFMA, v_mfma_f32_16x16x32_fp8_fp8.
//
// Scales are powers of two folded into the weights, so the kernel carries the
// activations pre-scaled and multiplies only at the output:
//   a_l is stored as C_l a_l (C_0 = 1: the encoded features lie in [-1, 1]),
//   layer l's weights as c_l W_l, so relu(c_l W_l C_l a_l) = C_{l+1} a_{l+1}
//   with C_{l+1} = c_l C_l, and the output as c_H W_H, unscaled by 1/(c_H C_H).
// c_l targets a largest scaled activation of 224 (half the e4m3 maximum 448)
// from the largest |a_{l+1}| of the cache's last 16 FP8 inferences (delayed
// scaling), and never scales c_l W_l past 448.  Values are converted with
// v_cvt_pk_fp8_f32 after one v_med3 that applies the ReLU and saturates at 448
// (the conversion itself turns values past the e4m3 range into NaN).
#include "nrc_mfma.hpp"

namespace nrc {

namespace {

constexpr float kFp8Max = 448.0f;
constexpr float kFp8Target = 224.0f;
constexpr uint32_t kFp8SampleStride = 8; // tiles whose activation maxima feed the scales

__device__ __forceinline__ float4_t mfma8(long a, long b, float4_t c) {
	return __builtin_amdgcn_mfma_f32_16x16x32_fp8_fp8(a, b, c, 0, 0, 0);
}

// four values -> four e4m3 bytes (first value in the low byte)
__device__ __forceinline__ uint32_t pack4_fp8(float a, float b, float c, float d) {
	int w = __builtin_amdgcn_cvt_pk_fp8_f32(a, b, 0, false);
	w = __builtin_amdgcn_cvt_pk_fp8_f32(c, d, w, true);
	return (uint32_t)w;
}

__device__ __forceinline__ float relu_sat(float x) { return __builtin_amdgcn_fmed3f(x, 0.0f, kFp8Max); }

// D tiles 2c (lo) and 2c+1 (hi) of a lane -> the next layer's B operand for K-chunk c
__device__ __forceinline__ long pack_relu_fp8(float4_t lo, float4_t hi) {
	const uint32_t a = pack4_fp8(relu_sat(lo[0]), relu_sat(lo[1]), relu_sat(lo[2]), relu_sat(lo[3]));
	const uint32_t b = pack4_fp8(relu_sat(hi[0]), relu_sat(hi[1]), relu_sat(hi[2]), relu_sat(hi[3]));
	return (long)((uint64_t)a | ((uint64_t)b << 32));
}

__device__ __forceinline__ float max4(float m, float4_t v) {
	return fmaxf(fmaxf(m, fmaxf(v[0], v[1])), fmaxf(v[2], v[3]));
}

// Same schedule as fused_infer_kernel (grouped, persistent, contiguous tile
// ranges per wave, encode into the wave's LDS slot, records of the next tile in
// flight), with 8-byte A and B fragments.  Every kFp8SampleStride-th tile of a
// cache also records the largest pre-saturation activation of each layer (a
// sample that does not depend on how the tiles are split among waves, so a
// cache gets the same scales alone and in a group); the wave's maxima go to its
// block's slot of the cache's FP8 state.
template <int H, int ENC, int TILES>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void fp8_infer_kernel(const InferGroup grp) {
	__shared__ __attribute__((aligned(16))) uint64_t sX[4][8 * 64];

	const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
	const int g = lane >> 4, r = lane & 15;
	const uint32_t nW = gridDim.x * 4, gw = blockIdx.x * 4 + wave;
	const uint32_t t0 = (uint32_t)((uint64_t)grp.total_tiles * gw / nW);
	const uint32_t t1 = (uint32_t)((uint64_t)grp.total_tiles * (gw + 1) / nW);
	uint64_t *sx = sX[wave];
	uint32_t *sx32 = reinterpret_cast<uint32_t *>(sx);
	const int et = lane >> 4, es = lane & 15; // this lane's record: tile et, column es

	for (uint32_t i = 0; i < grp.n_items; ++i) {
		const InferItem &it = grp.item[i];
		const uint32_t b0 = it.tile_begin, b1 = b0 + (it.max_count + 63) / 64;
		if (b1 <= t0) continue;
		if (b0 >= t1) break;
		const uint32_t count = __builtin_amdgcn_readfirstlane(effective_count(it.d_count, it.max_count, it.max_count));
		const uint32_t lt0 = (t0 > b0 ? t0 : b0) - b0;
		uint32_t lt1 = (t1 < b1 ? t1 : b1) - b0;
		const uint32_t ntiles = (count + 63) / 64;
		lt1 = lt1 < ntiles ? lt1 : ntiles;
		if (lt0 >= lt1) continue;

		const long *wv = reinterpret_cast<const long *>(it.w8);
		long A[H][4][2], Ao[2];
#pragma unroll
		for (int l = 0; l < H; ++l)
#pragma unroll
			for (int mt = 0; mt < 4; ++mt)
#pragma unroll
				for (int c = 0; c < 2; ++c) A[l][mt][c] = wv[((l * 4 + mt) * 2 + c) * 64 + lane];
#pragma unroll
		for (int c = 0; c < 2; ++c) Ao[c] = wv[(H * 8 + c) * 64 + lane]; // rows 3..15 are zero
		const float out_scale = it.f8->out_scale;

		const float *queries = it.queries;
		float *out = it.out;
		const uint32_t last = count - 1, tlast = lt1 - 1;
		auto load_query = [&](uint32_t t, float *dst) __attribute__((always_inline)) {
			t = t < tlast ? t : tlast;
			uint32_t m = t * 64 + lane;
			m = m < last ? m : last;
			const float2 *qp = reinterpret_cast<const float2 *>(queries + (size_t)m * kQueryFloats);
#pragma unroll
			for (int k = 0; k < 7; ++k) {
				const float2 v = qp[k];
				dst[2 * k] = v.x;
				dst[2 * k + 1] = v.y;
			}
		};
		auto encode_tile = [&](uint32_t base, const float *rec) __attribute__((always_inline)) {
			const uint32_t n = base + lane;
			auto emit = [&](int p, float f0, float f1, float f2, float f3) __attribute__((always_inline)) {
				const int c = p >> 3, gg = p & 3, hi = (p >> 2) & 1;
				sx32[((et * 2 + c) * 64 + gg * 16 + es) * 2 + hi] = pack4_fp8(f0, f1, f2, f3);
			};
			if (n < count) {
				encode_quads<ENC>(rec, emit);
			} else {
#pragma unroll
				for (int p = 0; p < 16; ++p) emit(p, 0.0f, 0.0f, 0.0f, 0.0f);
			}
			wave_lds_sync();
		};
		float amax[H]; // largest pre-saturation a_1 .. a_H of the recorded tile (scaled)
#pragma unroll
		for (int l = 0; l < H; ++l) amax[l] = 0.0f;
		auto mlp_tile = [&](uint32_t base, bool track) __attribute__((always_inline)) {
#pragma unroll
			for (int tg = 0; tg < 4; tg += TILES) {
				long B[TILES][2];
#pragma unroll
				for (int tt = 0; tt < TILES; ++tt)
#pragma unroll
					for (int c = 0; c < 2; ++c) B[tt][c] = (long)sx[((tg + tt) * 2 + c) * 64 + lane];
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
							for (int tt = 0; tt < TILES; ++tt) acc[tt][mt] = mfma8(A[l][mt][c], B[tt][c], acc[tt][mt]);
					if (track) { // uniform
#pragma unroll
						for (int tt = 0; tt < TILES; ++tt)
#pragma unroll
							for (int mt = 0; mt < 4; ++mt) amax[l] = max4(amax[l], acc[tt][mt]);
					}
#pragma unroll
					for (int tt = 0; tt < TILES; ++tt)
#pragma unroll
						for (int c = 0; c < 2; ++c) B[tt][c] = pack_relu_fp8(acc[tt][2 * c], acc[tt][2 * c + 1]);
				}
				float4_t o[TILES];
#pragma unroll
				for (int tt = 0; tt < TILES; ++tt) o[tt] = (float4_t)0.0f;
#pragma unroll
				for (int c = 0; c < 2; ++c)
#pragma unroll
					for (int tt = 0; tt < TILES; ++tt) o[tt] = mfma8(Ao[c], B[tt][c], o[tt]);
				if (g == 0) {
#pragma unroll
					for (int tt = 0; tt < TILES; ++tt) {
						const uint32_t m = base + (tg + tt) * 16 + r;
						if (m < count) {
							float *dst = out + (size_t)m * kOut;
#pragma unroll
							for (int c = 0; c < kOut; ++c) {
								const float y = o[tt][c] * out_scale;
								dst[c] = grp.clamp0 ? fmaxf(y, 0.0f) : y;
							}
						}
					}
				}
			}
			wave_lds_sync();
		};

		// ping-pong record buffers as in fused_infer_kernel; sampled tiles also
		// record the activation maxima
		float ra[kQueryFloats], rb[kQueryFloats];
		load_query(lt0, ra);
		for (uint32_t t = lt0; t < lt1; t += 2) {
			load_query(t + 1, rb);
			encode_tile(t * 64, ra);
			mlp_tile(t * 64, t % kFp8SampleStride == 0);
			if (t + 1 >= lt1) break;
			load_query(t + 2, ra);
			encode_tile((t + 1) * 64, rb);
			mlp_tile((t + 1) * 64, (t + 1) % kFp8SampleStride == 0);
		}
		// the wave's maxima into its block's slot (non-negative floats order as their bits)
#pragma unroll
		for (int l = 0; l < H; ++l) {
			float m = amax[l];
#pragma unroll
			for (int off = 32; off >= 1; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
			if (lane == 0 && m > 0.0f)
				atomicMax(reinterpret_cast<unsigned int *>(&it.f8->block_amax[blockIdx.x][l + 1]), __float_as_uint(m));
		}
	}
}

__device__ __forceinline__ float pow2_floor(float x) { return ldexpf(1.0f, ilogbf(x)); }

// One block per cache: fold the last FP8 inference's per-block maxima into the
// history and choose the layer scales.
template <int H> __global__ __launch_bounds__(256) void fp8_scale_kernel(const Fp8PrepGroup grp) {
	const Fp8PrepItem &it = grp.item[blockIdx.x];
	Fp8State *st = it.f8;
	const int t = threadIdx.x;
	__shared__ float red[kFp8Layers][256];

	// 1. the last inference's activation maxima (scaled by its C_l), cleared for the
	// next; every load is issued before the first store and the first maximum
	constexpr int kPer = kFp8MaxBlocks / 256;
	float m[kFp8Layers], v[kPer][kFp8Layers];
#pragma unroll
	for (int j = 0; j < kPer; ++j)
#pragma unroll
		for (int l = 0; l < kFp8Layers; ++l) v[j][l] = st->block_amax[t + 256 * j][l];
#pragma unroll
	for (int l = 0; l < kFp8Layers; ++l) {
		m[l] = v[0][l];
#pragma unroll
		for (int j = 1; j < kPer; ++j) m[l] = fmaxf(m[l], v[j][l]);
	}
#pragma unroll
	for (int j = 0; j < kPer; ++j)
#pragma unroll
		for (int l = 0; l < kFp8Layers; ++l) st->block_amax[t + 256 * j][l] = 0.0f;
#pragma unroll
	for (int l = 0; l < kFp8Layers; ++l) red[l][t] = m[l];
	__syncthreads();
	for (int s = 128; s > 0; s >>= 1) {
		if (t < s)
#pragma unroll
			for (int l = 0; l < kFp8Layers; ++l) red[l][t] = fmaxf(red[l][t], red[l][t + s]);
		__syncthreads();
	}
	__shared__ uint32_t nframes;
	if (t == 0) {
		uint32_t frames = st->frames;
		if (st->scale[0] > 0.0f) { // a previous call set the scales that inference used
			const uint32_t slot = frames % kFp8History;
			st->hist[slot][0] = 1.0f;
#pragma unroll
			for (int l = 1; l <= H; ++l) st->hist[slot][l] = red[l][0] / st->scale[l];
			st->frames = ++frames;
		}
		nframes = frames;
	}
	__syncthreads();
	// the history maximum of each layer, one thread per layer
	__shared__ float hmaxs[kFp8Layers];
	if (t < H) {
		const uint32_t nh = nframes < (uint32_t)kFp8History ? nframes : (uint32_t)kFp8History;
		float h = 0.0f;
		for (uint32_t k = 0; k < nh; ++k) h = fmaxf(h, st->hist[k][t + 1]);
		hmaxs[t] = nh == 0 ? 4.0f : h; // no history yet: a first guess, corrected after one inference
	}

	// 2. largest |W_l| per layer (all loads in flight at once)
	const float *w = it.use32;
	{
		constexpr int kRow = kWidth * kWidth / 256; // weights of one hidden layer per thread
		float wv[H][kRow];
#pragma unroll
		for (int l = 0; l < H; ++l)
#pragma unroll
			for (int j = 0; j < kRow; ++j) wv[l][j] = w[l * kWidth * kWidth + t + 256 * j];
		const float wo = t < kOut * kWidth ? w[H * kWidth * kWidth + t] : 0.0f;
#pragma unroll
		for (int l = 0; l < H; ++l) {
			float a = fabsf(wv[l][0]);
#pragma unroll
			for (int j = 1; j < kRow; ++j) a = fmaxf(a, fabsf(wv[l][j]));
			red[l][t] = a;
		}
		red[H][t] = fabsf(wo);
	}
	__syncthreads();
	for (int s = 128; s > 0; s >>= 1) {
		if (t < s)
#pragma unroll
			for (int l = 0; l <= H; ++l) red[l][t] = fmaxf(red[l][t], red[l][t + s]);
		__syncthreads();
	}

	// 3. the scales: c_l for l < H from the activation history, c_H from the weights
	if (t == 0) {
		float C = 1.0f;
		st->scale[0] = 1.0f;
#pragma unroll
		for (int l = 0; l < H; ++l) {
			const float hmax = hmaxs[l];
			const float wmax = red[l][0];
			float c = (hmax > 0.0f) ? pow2_floor(kFp8Target / (hmax * C)) : 1.0f;
			if (wmax > 0.0f && wmax * c > kFp8Max) c = pow2_floor(kFp8Max / wmax);
			st->wscale[l] = c;
			C *= c;
			st->scale[l + 1] = C;
		}
		const float wmax = red[H][0];
		const float cH = wmax > 0.0f ? pow2_floor(kFp8Max / wmax) : 1.0f;
		st->wscale[H] = cH;
		st->out_scale = 1.0f / (cH * C);
	}
}

// The fp8 fragment-major copy of c_l W_l, one weight per thread (grid.y = cache).
template <int H> __global__ __launch_bounds__(256) void fp8_quantize_kernel(const Fp8PrepGroup grp) {
	const Fp8PrepItem &it = grp.item[blockIdx.y];
	const uint32_t i = blockIdx.x * 256 + threadIdx.x;
	if (i >= (uint32_t)weight_count(H)) return;
	const uint32_t l = i >> 12, row = (i >> 6) & 63u, col = i & 63u;
	const float q = __builtin_amdgcn_fmed3f(it.use32[i] * it.f8->wscale[l], -kFp8Max, kFp8Max);
	it.w8[frag_pos(l, row, col)] = (uint8_t)(__builtin_amdgcn_cvt_pk_fp8_f32(q, 0.0f, 0, false) & 0xff);
}

template <int H, int ENC> hipError_t fp8_launch(InferGroup &grp, const Fp8PrepGroup &prep, hipStream_t s) {
	fp8_scale_kernel<H><<<dim3(prep.n_items), dim3(256), 0, s>>>(prep);
	hipError_t e = hipGetLastError();
	if (e != hipSuccess) return e;
	fp8_quantize_kernel<H><<<dim3((weight_count(H) + 255) / 256, prep.n_items), dim3(256), 0, s>>>(prep);
	e = hipGetLastError();
	if (e != hipSuccess) return e;
	uint32_t blocks = (grp.total_tiles + 3) / 4;
	uint32_t cap = (uint32_t)device_cu_count() * infer_wg_per_cu();
	if (cap > (uint32_t)kFp8MaxBlocks) cap = kFp8MaxBlocks;
	if (blocks > cap) blocks = cap;
	if (blocks == 0) blocks = 1;
	fp8_infer_kernel<H, ENC, 1><<<dim3(blocks), dim3(256), 0, s>>>(grp);
	return hipGetLastError();
}

template <int H> hipError_t fp8_h(int enc, InferGroup &grp, const Fp8PrepGroup &prep, hipStream_t s) {
	switch (enc) {
	case NRC_ENCODING_TRI12: return fp8_launch<H, NRC_ENCODING_TRI12>(grp, prep, s);
	case NRC_ENCODING_SINCOS6: return fp8_launch<H, NRC_ENCODING_SINCOS6>(grp, prep, s);
	default: return fp8_launch<H, NRC_ENCODING_UV_SINCOS16>(grp, prep, s);
	}
}

} // namespace

hipError_t launch_fp8_infer_group(int hidden, int encoding, InferGroup &grp, const float *const *use32,
                                  uint8_t *const *w8, hipStream_t stream) {
	uint32_t tiles = 0;
	Fp8PrepGroup prep{};
	prep.n_items = grp.n_items;
	for (uint32_t i = 0; i < grp.n_items; ++i) {
		grp.item[i].tile_begin = tiles;
		tiles += (grp.item[i].max_count + 63) / 64;
		prep.item[i] = Fp8PrepItem{grp.item[i].f8, use32[i], w8[i]};
	}
	grp.total_tiles = tiles;
	if (tiles == 0) return hipSuccess;
	if (hidden == 5) return fp8_h<5>(encoding, grp, prep, stream);
	return fp8_h<4>(encoding, grp, prep, stream);
}

} // namespace nrc
