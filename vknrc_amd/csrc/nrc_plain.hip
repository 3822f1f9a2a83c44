// nrc_plain.hip -- the plain method: FP32 encoding, weights and accumulation.
//
// One thread per record, all weights of the network staged in LDS as FP32, the
// 64-wide activation vector held in VGPRs.  The weight gradient is reduced over
// fixed 512-record chunks in a fixed order, so the result is deterministic.
// This path is the FP32 reference the fused FP16-MFMA path is measured against.
#include "nrc_internal.hpp"

namespace nrc {

namespace {

// Output units between scheduling barriers in the FP32 forward: groups of 4
// keep the H = 5 kernel within 256 VGPRs without spilling.
constexpr int kPlainSchedGroup = 4;

// x (64 input features) -> a_H in x, returns the 3 raw outputs in o.
// a_{l+1} = relu(W_l a_l) (NN_nv.glsl:99-127), o = W_H a_H (NN_nv.glsl:129-158).
template <int H> __device__ __forceinline__ void forward_f32(const float *sW, float *x, float *o) {
#pragma unroll 1
	for (int l = 0; l < H; ++l) {
		const float *Wl = sW + l * kWidth * kWidth;
		float y[kWidth];
#pragma unroll
		for (int m = 0; m < kWidth; ++m) {
			float acc = 0.0f;
#pragma unroll
			for (int k = 0; k < kWidth; k += 4) {
				const float4 w = *reinterpret_cast<const float4 *>(Wl + m * kWidth + k);
				acc = fmaf(w.x, x[k + 0], acc);
				acc = fmaf(w.y, x[k + 1], acc);
				acc = fmaf(w.z, x[k + 2], acc);
				acc = fmaf(w.w, x[k + 3], acc);
			}
			y[m] = fmaxf(acc, 0.0f);
			// bound how far the weight loads run ahead (and the registers they hold)
			if (m % kPlainSchedGroup == kPlainSchedGroup - 1) __builtin_amdgcn_sched_barrier(0);
		}
#pragma unroll
		for (int m = 0; m < kWidth; ++m) x[m] = y[m];
	}
	const float *Wo = sW + H * kWidth * kWidth;
#pragma unroll
	for (int c = 0; c < kOut; ++c) {
		float acc = 0.0f;
#pragma unroll
		for (int k = 0; k < kWidth; k += 4) {
			const float4 w = *reinterpret_cast<const float4 *>(Wo + c * kWidth + k);
			acc = fmaf(w.x, x[k + 0], acc);
			acc = fmaf(w.y, x[k + 1], acc);
			acc = fmaf(w.z, x[k + 2], acc);
			acc = fmaf(w.w, x[k + 3], acc);
		}
		o[c] = acc;
	}
}

// All weights into LDS by a BLOCK-thread block; loads go out four at a time
// before their LDS stores (one memory round trip per four float4 per thread).
template <int H, int BLOCK> __device__ __forceinline__ void stage_weights(float *sW, const float *__restrict__ w32) {
	constexpr int kVec = weight_count(H) / 4;
	const float4 *src = reinterpret_cast<const float4 *>(w32);
	float4 *dst = reinterpret_cast<float4 *>(sW);
#pragma unroll
	for (int b = 0; b < kVec; b += 4 * BLOCK) {
		const int j0 = b + threadIdx.x, j1 = j0 + BLOCK, j2 = j0 + 2 * BLOCK, j3 = j0 + 3 * BLOCK;
		float4 v0 = {}, v1 = {}, v2 = {}, v3 = {};
		if (j0 < kVec) v0 = src[j0];
		if (j1 < kVec) v1 = src[j1];
		if (j2 < kVec) v2 = src[j2];
		if (j3 < kVec) v3 = src[j3];
		if (j0 < kVec) dst[j0] = v0;
		if (j1 < kVec) dst[j1] = v1;
		if (j2 < kVec) dst[j2] = v2;
		if (j3 < kVec) dst[j3] = v3;
	}
	__syncthreads();
}

__device__ __forceinline__ void load_query(const float *__restrict__ q, float *rec) {
	const float2 *p = reinterpret_cast<const float2 *>(q);
#pragma unroll
	for (int i = 0; i < 7; ++i) {
		const float2 v = p[i];
		rec[2 * i] = v.x;
		rec[2 * i + 1] = v.y;
	}
}

// nrc_inference.comp:30-48 (without the scene unpack and the epilogue).
template <int H, int ENC>
__global__ __launch_bounds__(512) void plain_infer_kernel(const float *__restrict__ queries,
                                                          const uint32_t *__restrict__ d_count, uint32_t max_count,
                                                          const float *__restrict__ w32, float *__restrict__ rgb) {
	__shared__ __attribute__((aligned(16))) float sW[weight_count(H)];
	stage_weights<H, 512>(sW, w32);
	const uint32_t count = effective_count(d_count, max_count, max_count);
	const uint32_t n = blockIdx.x * blockDim.x + threadIdx.x;
	if (n >= count) return;
	float rec[kQueryFloats];
	load_query(queries + (size_t)n * kQueryFloats, rec);
	float x[kWidth], o[kOut];
	encode_all<ENC>(rec, x);
	forward_f32<H>(sW, x, o);
#pragma unroll
	for (int c = 0; c < kOut; ++c) rgb[(size_t)n * kOut + c] = fmaxf(o[c], 0.0f); // nrc_inference.comp:48
}

// Per record: encode, forward keeping a_0..a_H, loss gradient, then the
// backward recurrence delta_{l-1} = (W_l^T delta_l) . 1[a_l > 0]
// (nrc_gradient.comp:26-58, NN_nv.glsl:162-278).  Writes the activations and
// deltas feature-major so the chunk reduction reads them coalesced.
template <int H, int ENC, int LOSS>
__global__ __launch_bounds__(256) void plain_fwdbwd_kernel(const float *__restrict__ records,
                                                           const uint32_t *__restrict__ d_count, uint32_t max_count,
                                                           uint32_t clamp, const float *__restrict__ w32,
                                                           const DeviceState *__restrict__ st, int parity,
                                                           float *__restrict__ acts, float *__restrict__ deltas,
                                                           float *__restrict__ sample_loss, uint32_t ldn) {
	__shared__ __attribute__((aligned(16))) float sW[weight_count(H)];
	stage_weights<H, 256>(sW, w32);
	const uint32_t count = effective_count(d_count, max_count, clamp);
	const uint32_t n = blockIdx.x * blockDim.x + threadIdx.x;
	if (n >= count) return;
	const size_t slab = (size_t)kWidth * ldn;
	const float s = st->pub[parity].loss_scale;
	const float *r = records + (size_t)n * kTrainFloats;
	float rec[kQueryFloats];
#pragma unroll
	for (int i = 0; i < kQueryFloats; ++i) rec[i] = r[i];
	const float tgt[3] = {r[14], r[15], r[16]};

	float x[kWidth];
	encode_all<ENC>(rec, x);
#pragma unroll
	for (int k = 0; k < kWidth; ++k) acts[k * ldn + n] = x[k];
#pragma unroll 1
	for (int l = 0; l < H; ++l) {
		const float *Wl = sW + l * kWidth * kWidth;
		float y[kWidth];
#pragma unroll
		for (int m = 0; m < kWidth; ++m) {
			float acc = 0.0f;
#pragma unroll
			for (int k = 0; k < kWidth; k += 4) {
				const float4 w = *reinterpret_cast<const float4 *>(Wl + m * kWidth + k);
				acc = fmaf(w.x, x[k + 0], acc);
				acc = fmaf(w.y, x[k + 1], acc);
				acc = fmaf(w.z, x[k + 2], acc);
				acc = fmaf(w.w, x[k + 3], acc);
			}
			y[m] = fmaxf(acc, 0.0f);
		}
		float *al = acts + (size_t)(l + 1) * slab;
#pragma unroll
		for (int m = 0; m < kWidth; ++m) {
			x[m] = y[m];
			al[m * ldn + n] = y[m];
		}
	}
	const float *Wo = sW + H * kWidth * kWidth;
	float o[kOut];
#pragma unroll
	for (int c = 0; c < kOut; ++c) {
		float acc = 0.0f;
#pragma unroll
		for (int k = 0; k < kWidth; ++k) acc = fmaf(Wo[c * kWidth + k], x[k], acc);
		o[c] = acc;
	}
	float go[kOut];
	sample_loss[n] = loss_grad<LOSS>(o, tgt, s, go);
	float *dH = deltas + (size_t)H * slab;
#pragma unroll
	for (int c = 0; c < kOut; ++c) dH[c * ldn + n] = go[c];

	// delta_{H-1} = (W_H^T g_o) . 1[a_H > 0]   (NNBackwardDA3_ReLU, NN_nv.glsl:222-246)
	float g[kWidth];
#pragma unroll
	for (int k = 0; k < kWidth; ++k) {
		float v = Wo[k] * go[0];
		v = fmaf(Wo[kWidth + k], go[1], v);
		v = fmaf(Wo[2 * kWidth + k], go[2], v);
		g[k] = relu_bwd(v, x[k]);
	}
	{
		float *d = deltas + (size_t)(H - 1) * slab;
#pragma unroll
		for (int k = 0; k < kWidth; ++k) d[k * ldn + n] = g[k];
	}
	// delta_{l-1} = (W_l^T delta_l) . 1[a_l > 0], l = H-1..1  (NNBackwardDA64_ReLU, NN_nv.glsl:248-278)
#pragma unroll 1
	for (int l = H - 1; l >= 1; --l) {
		const float *Wl = sW + l * kWidth * kWidth;
		float gn[kWidth];
#pragma unroll
		for (int k = 0; k < kWidth; ++k) gn[k] = 0.0f;
#pragma unroll
		for (int m = 0; m < kWidth; ++m) {
#pragma unroll
			for (int k = 0; k < kWidth; k += 4) {
				const float4 w = *reinterpret_cast<const float4 *>(Wl + m * kWidth + k);
				gn[k + 0] = fmaf(w.x, g[m], gn[k + 0]);
				gn[k + 1] = fmaf(w.y, g[m], gn[k + 1]);
				gn[k + 2] = fmaf(w.z, g[m], gn[k + 2]);
				gn[k + 3] = fmaf(w.w, g[m], gn[k + 3]);
			}
		}
		const float *al = acts + (size_t)l * slab;
		float *d = deltas + (size_t)(l - 1) * slab;
#pragma unroll
		for (int k = 0; k < kWidth; ++k) {
			g[k] = relu_bwd(gn[k], al[k * ldn + n]);
			d[k * ldn + n] = g[k];
		}
	}
}

// Partial dW over one chunk of records: dW_l[m][k] = sum_n delta_l[m][n] a_l[k][n]
// (NNUpdateDW64 / NNUpdateDW3, NN_nv.glsl:280-367).  Grid: x over blocks of 256
// weights (one layer, four output rows each), y over chunks.  Also writes the
// chunk's loss sum after the weights and flags non-finite partials.
template <int H>
__global__ __launch_bounds__(256) void plain_dw_kernel(const uint32_t *__restrict__ d_count, uint32_t max_count,
                                                       uint32_t clamp, const float *__restrict__ acts,
                                                       const float *__restrict__ deltas,
                                                       const float *__restrict__ sample_loss, uint32_t ldn,
                                                       float *__restrict__ partials, uint32_t pstride,
                                                       uint32_t *__restrict__ overflow) {
	constexpr int WC = weight_count(H);
	__shared__ float sa[kWidth][kWidth + 1];
	__shared__ float sd[4][kWidth];
	__shared__ float sl[256];
	const uint32_t count = effective_count(d_count, max_count, clamp);
	const int t = threadIdx.x;
	const uint32_t w0 = blockIdx.x * 256u;
	const int l = (int)(w0 / (kWidth * kWidth));
	const int m0 = (int)((w0 % (kWidth * kWidth)) / kWidth);
	const int rows = (l == H) ? kOut : 4;
	const int mr = t / kWidth, k = t % kWidth;
	const size_t slab = (size_t)kWidth * ldn;
	const float *a = acts + (size_t)l * slab;
	const float *d = deltas + (size_t)l * slab;
	const uint32_t nb0 = blockIdx.y * kPlainChunk;
	uint32_t nend = nb0 + kPlainChunk;
	if (nend > count) nend = count;
	float acc = 0.0f;
	for (uint32_t nb = nb0; nb < nend; nb += kWidth) {
		// stage a_l[0..63][nb..nb+63] and delta_l[m0..m0+3][nb..nb+63]
#pragma unroll
		for (int i = 0; i < 16; ++i) {
			const int e = t + 256 * i;
			const int kk = e / kWidth, j = e % kWidth;
			sa[kk][j] = (nb + j < nend) ? a[kk * ldn + nb + j] : 0.0f;
		}
		{
			const int rr = t / kWidth, j = t % kWidth;
			sd[rr][j] = (rr < rows && nb + j < nend) ? d[(m0 + rr) * ldn + nb + j] : 0.0f;
		}
		__syncthreads();
#pragma unroll 8
		for (int j = 0; j < kWidth; ++j) acc = fmaf(sd[mr][j], sa[k][j], acc);
		__syncthreads();
	}
	const uint32_t w = w0 + (uint32_t)t;
	float *pp = partials + (size_t)blockIdx.y * pstride;
	if (w < (uint32_t)WC) {
		pp[w] = acc;
		if (!isfinite(acc)) atomicOr(overflow, 1u);
	}
	if (blockIdx.x == 0) {
		float v = 0.0f;
		for (uint32_t n = nb0 + t; n < nend; n += 256) v += sample_loss[n];
		sl[t] = v;
		__syncthreads();
		for (int s = 128; s > 0; s >>= 1) {
			if (t < s) sl[t] += sl[t + s];
			__syncthreads();
		}
		if (t == 0) pp[WC] = sl[0];
	}
}

template <int ENC>
__global__ __launch_bounds__(256) void encode_kernel(const float *__restrict__ queries, uint32_t n,
                                                     float *__restrict__ feat) {
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n) return;
	float rec[kQueryFloats];
	load_query(queries + (size_t)i * kQueryFloats, rec);
	float f[kWidth];
	encode_all<ENC>(rec, f);
#pragma unroll
	for (int k = 0; k < kWidth; k += 4)
		*reinterpret_cast<float4 *>(feat + (size_t)i * kWidth + k) = make_float4(f[k], f[k + 1], f[k + 2], f[k + 3]);
}

__global__ __launch_bounds__(256) void sph_kernel(const float *__restrict__ dirs, uint32_t n, float *__restrict__ uv) {
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n) return;
	float o[2];
	sph_encode(dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2], o);
	uv[2 * i] = o[0];
	uv[2 * i + 1] = o[1];
}

template <int H>
hipError_t plain_infer_h(int enc, const float *q, const uint32_t *dc, uint32_t mc, const float *w, float *rgb,
                         hipStream_t s) {
	const dim3 grid((mc + 511) / 512), block(512);
	switch (enc) {
	case NRC_ENCODING_TRI12: plain_infer_kernel<H, NRC_ENCODING_TRI12><<<grid, block, 0, s>>>(q, dc, mc, w, rgb); break;
	case NRC_ENCODING_SINCOS6: plain_infer_kernel<H, NRC_ENCODING_SINCOS6><<<grid, block, 0, s>>>(q, dc, mc, w, rgb); break;
	default: plain_infer_kernel<H, NRC_ENCODING_UV_SINCOS16><<<grid, block, 0, s>>>(q, dc, mc, w, rgb); break;
	}
	return hipGetLastError();
}

template <int H, int ENC, int LOSS> void plain_fwdbwd_launch(const TrainArgs &a, hipStream_t s) {
	const dim3 grid((a.max_count + 255) / 256), block(256);
	plain_fwdbwd_kernel<H, ENC, LOSS><<<grid, block, 0, s>>>(a.records, a.d_count, a.max_count, a.clamp, a.w32,
	                                                          a.state, a.parity, a.acts, a.deltas, a.sample_loss, a.ldn);
}

template <int H, int ENC> void plain_fwdbwd_loss(const TrainArgs &a, hipStream_t s) {
	if (a.loss == NRC_LOSS_L2) plain_fwdbwd_launch<H, ENC, NRC_LOSS_L2>(a, s);
	else plain_fwdbwd_launch<H, ENC, NRC_LOSS_RELATIVE_L2_LUMINANCE>(a, s);
}

template <int H> void plain_gradient_h(const TrainArgs &a, uint32_t nparts, hipStream_t s) {
	switch (a.encoding) {
	case NRC_ENCODING_TRI12: plain_fwdbwd_loss<H, NRC_ENCODING_TRI12>(a, s); break;
	case NRC_ENCODING_SINCOS6: plain_fwdbwd_loss<H, NRC_ENCODING_SINCOS6>(a, s); break;
	default: plain_fwdbwd_loss<H, NRC_ENCODING_UV_SINCOS16>(a, s); break;
	}
	const dim3 grid((weight_count(H) + 255) / 256, nparts), block(256);
	plain_dw_kernel<H><<<grid, block, 0, s>>>(a.d_count, a.max_count, a.clamp, a.acts, a.deltas, a.sample_loss, a.ldn,
	                                          a.partials, a.pstride, &a.state->overflow[a.parity]);
}

} // namespace

hipError_t launch_plain_infer(int hidden, int encoding, const float *queries, const uint32_t *d_count,
                              uint32_t max_count, const float *w32, float *rgb, hipStream_t stream) {
	if (max_count == 0) return hipSuccess;
	if (hidden == 5) return plain_infer_h<5>(encoding, queries, d_count, max_count, w32, rgb, stream);
	return plain_infer_h<4>(encoding, queries, d_count, max_count, w32, rgb, stream);
}

hipError_t launch_plain_gradient(const TrainArgs &a, uint32_t *nparts, hipStream_t stream) {
	const uint32_t n = a.max_count < a.clamp ? a.max_count : a.clamp;
	*nparts = (n + kPlainChunk - 1) / kPlainChunk;
	if (*nparts == 0) *nparts = 1;
	if (a.hidden == 5) plain_gradient_h<5>(a, *nparts, stream);
	else plain_gradient_h<4>(a, *nparts, stream);
	return hipGetLastError();
}

hipError_t launch_encode(int encoding, const float *queries, uint32_t n, float *features, hipStream_t stream) {
	if (n == 0) return hipSuccess;
	const dim3 grid((n + 255) / 256), block(256);
	switch (encoding) {
	case NRC_ENCODING_TRI12: encode_kernel<NRC_ENCODING_TRI12><<<grid, block, 0, stream>>>(queries, n, features); break;
	case NRC_ENCODING_SINCOS6: encode_kernel<NRC_ENCODING_SINCOS6><<<grid, block, 0, stream>>>(queries, n, features); break;
	default: encode_kernel<NRC_ENCODING_UV_SINCOS16><<<grid, block, 0, stream>>>(queries, n, features); break;
	}
	return hipGetLastError();
}

hipError_t launch_sph_encode(const float *dirs, uint32_t n, float *uv, hipStream_t stream) {
	if (n == 0) return hipSuccess;
	sph_kernel<<<dim3((n + 255) / 256), dim3(256), 0, stream>>>(dirs, n, uv);
	return hipGetLastError();
}

namespace {
// nrc_inference.comp:50-72: one thread per query.
__global__ __launch_bounds__(256) void composite_kernel(const CompositeArgs a) {
	const uint32_t i = blockIdx.x * 256u + threadIdx.x;
	const uint32_t count = effective_count(a.d_count, a.max_count, a.max_count);
	if (i >= count) return;
	const uint32_t e = a.dst[i];
	if (e == 0xFFFFFFFFu) return;
	const float p0 = a.rgb[(size_t)i * 3], p1 = a.rgb[(size_t)i * 3 + 1], p2 = a.rgb[(size_t)i * 3 + 2];
	const uint32_t v = e >> 1;
	// fields decoded as DecodeNRCEvalDstScreen/Train do (NRCRecord.glsl:24-33): the
	// top field keeps every remaining bit, so an over-wide y or r falls outside the
	// image or the batch and is skipped
	if ((e & 1u) == 0u) { // screen
		const uint32_t x = v & 0x7FFFu, y = v >> 15;
		if (a.bias_factor_r == nullptr || x >= a.width || y >= a.height) return;
		const size_t px = (size_t)y * a.pitch + x;
		float4 *bf = reinterpret_cast<float4 *>(a.bias_factor_r) + px;
		const float4 b = *bf;
		const float2 f = reinterpret_cast<const float2 *>(a.factor_gb)[px];
		*bf = make_float4(mul_add_rn(b.x, b.w, p0), mul_add_rn(b.y, f.x, p1), mul_add_rn(b.z, f.y, p2), 0.0f);
	} else { // train
		const uint32_t b = v & 3u, l = (v >> 2) & 0x3FFFu, r = v >> 16;
		if (a.batches == nullptr) return;
		const uint32_t rr = r < a.batch_stride ? r : a.batch_stride - 1u;
		for (uint32_t j = l; j <= rr && j < a.batch_stride; ++j) {
			const size_t row = (size_t)b * a.batch_stride + j;
			float *t = a.batches + row * kTrainFloats + kQueryFloats;
			const float *f = a.factors + row * 3;
			t[0] = mul_add_rn(t[0], f[0], p0);
			t[1] = mul_add_rn(t[1], f[1], p1);
			t[2] = mul_add_rn(t[2], f[2], p2);
		}
	}
}
} // namespace

hipError_t launch_composite(const CompositeArgs &a, hipStream_t stream) {
	if (a.max_count == 0) return hipSuccess;
	composite_kernel<<<dim3((a.max_count + 255) / 256), dim3(256), 0, stream>>>(a);
	return hipGetLastError();
}

} // namespace nrc
