// nrc_optim.hip -- partial-gradient reduction fused with Adam + EMA and the
// optimizer scalar recurrences (nrc_optimize.comp:32-54, nrc_train_prepare.comp:16-29).
//
// Each block owns 64 consecutive weights; its 256 threads sum the per-workgroup
// partial gradients as float4 columns in 16 interleaved slices (four loads in
// flight per thread) that are combined in a fixed order, so the reduced
// gradient is bit-identical from run to run.  Weight counts are multiples of 64
// and partial rows are 64-float aligned (the row stride holds the loss column and
// its 63 padding floats), so the float4 loads stay in bounds.  The last
// block to finish advances the scalar state (t, beta1^t, beta2^t, alpha^t,
// alpha^{t-1}, loss scale, loss), which every block read at its start.
#include "nrc_internal.hpp"

namespace nrc {

namespace {

__global__ __launch_bounds__(256) void adam_kernel(AdamArgs a) {
	const uint32_t count = effective_count(a.d_count, a.max_count, a.clamp);
	if (count == 0) return; // nrc_optimize.comp:33-34: an empty batch changes nothing
	const int t = threadIdx.x;
	DeviceState *st = a.state;
	const nrc_optimizer_state s0 = st->pub;
	const bool skip = a.p.dynamic_scale && (st->overflow != 0);

	// thread t sums partials q = slice, slice + 16, ... of weights w0 + 4*col4 .. +3
	const int col4 = t & 15, slice = t >> 4;
	const uint32_t w0 = blockIdx.x * 64u;
	__shared__ float4 red[16][16];
	float4 G4 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
	if (!skip) {
		const float4 *p = reinterpret_cast<const float4 *>(a.partials + w0) + col4;
		const size_t ps4 = a.pstride / 4;
		float4 acc[4];
#pragma unroll
		for (int u = 0; u < 4; ++u) acc[u] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
		uint32_t q = slice;
		for (; q + 48 < a.nparts; q += 64) {
#pragma unroll
			for (int u = 0; u < 4; ++u) {
				const float4 v = p[(q + 16 * u) * ps4];
				acc[u].x += v.x; acc[u].y += v.y; acc[u].z += v.z; acc[u].w += v.w;
			}
		}
		for (; q < a.nparts; q += 16) {
			const float4 v = p[q * ps4];
			acc[0].x += v.x; acc[0].y += v.y; acc[0].z += v.z; acc[0].w += v.w;
		}
		G4.x = (acc[0].x + acc[1].x) + (acc[2].x + acc[3].x);
		G4.y = (acc[0].y + acc[1].y) + (acc[2].y + acc[3].y);
		G4.z = (acc[0].z + acc[1].z) + (acc[2].z + acc[3].z);
		G4.w = (acc[0].w + acc[1].w) + (acc[2].w + acc[3].w);
	}
	red[slice][col4] = G4;
	__syncthreads();
	const uint32_t i = w0 + (uint32_t)t;
	if (blockIdx.x == gridDim.x - 1) {
		// the extra block: column weight_count of the partials holds each workgroup's loss sum
		if (t == 0 && !skip) {
			const float *rf = reinterpret_cast<const float *>(&red[0][0]);
			float L = 0.0f;
#pragma unroll
			for (int sl = 0; sl < 16; ++sl) L += rf[sl * 64];
			st->loss_sum = L;
		}
	} else if (t < 64 && !skip && i < a.weight_count) {
		const float *rf = reinterpret_cast<const float *>(&red[0][0]);
		float G = 0.0f;
#pragma unroll
		for (int sl = 0; sl < 16; ++sl) G += rf[sl * 64 + t];
		const float b1 = a.p.beta1, b2 = a.p.beta2, al = a.p.ema_alpha;
		const float B1 = s0.beta1_t * b1, B2 = s0.beta2_t * b2; // nrc_train_prepare.comp:23-27
		const float A1 = s0.alpha_t, A = s0.alpha_t * al;
		float g = G / (float)count / s0.loss_scale; // nrc_optimize.comp:36
		if (!isfinite(g)) g = 0.0f;                 // nrc_optimize.comp:37-38
		nrc_optimizer_entry e = a.entries[i];
		e.m = b1 * e.m + (1.0f - b1) * g;
		e.v = b2 * e.v + (1.0f - b2) * (g * g);
		const float mh = e.m / (1.0f - B1), vh = e.v / (1.0f - B2);
		e.weight -= a.p.lr * mh / (sqrtf(vh) + a.p.eps);
		const float eta = 1.0f - A, eta1 = 1.0f - A1; // nrc_optimize.comp:46-47, as written
		e.ema_weight = (1.0f - al) / eta * e.weight + al * eta1 * e.ema_weight;
		a.entries[i] = e;
		a.w32[i] = e.weight;
		const _Float16 h = (_Float16)e.weight;
		a.w16[i] = h;
		const uint32_t hid = (uint32_t)a.hidden * 4096u;
		a.w16t[i < hid ? ((i & ~4095u) | ((i & 63u) << 6) | ((i >> 6) & 63u)) : i] = h;
		if (a.write_use) { // nrc_optimize.comp:50-53
			const float u = a.p.use_ema ? e.ema_weight : e.weight;
			a.use32[i] = u;
			a.use16[i] = (_Float16)u;
		}
	}
	__syncthreads();
	if (t == 0) {
		__threadfence();
		const uint32_t done = atomicAdd(&st->done_blocks, 1u);
		if (done == gridDim.x - 1) {
			__threadfence();
			nrc_optimizer_state s = s0;
			if (skip) {
				s.loss_scale *= 0.5f;
				s.clean_steps = 0;
				s.skipped_steps += 1;
			} else {
				s.t += 1;
				s.beta1_t *= a.p.beta1;
				s.beta2_t *= a.p.beta2;
				s.alpha_t_1 = s.alpha_t;
				s.alpha_t *= a.p.ema_alpha;
				s.last_loss = *(volatile float *)&st->loss_sum / (float)count;
				s.last_count = count;
				if (a.p.dynamic_scale) {
					s.clean_steps += 1;
					if (a.p.growth_interval > 0 && s.clean_steps >= a.p.growth_interval) {
						s.loss_scale *= 2.0f;
						s.clean_steps = 0;
					}
				}
			}
			st->pub = s;
			st->overflow = 0;
			st->done_blocks = 0;
			__threadfence();
		}
	}
}

// out[i] = sum_q partials[q][i] in a fixed order (the gradient-only entry point).
__global__ __launch_bounds__(256) void reduce_kernel(const float *__restrict__ partials, uint32_t nparts,
                                                     uint32_t pstride, uint32_t n, float *__restrict__ out) {
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n) return;
	float s = 0.0f;
	for (uint32_t q = 0; q < nparts; ++q) s += partials[(size_t)q * pstride + i];
	out[i] = s;
}

__global__ __launch_bounds__(256) void transpose_hidden_kernel(const _Float16 *__restrict__ w16,
                                                                _Float16 *__restrict__ w16t, uint32_t n) {
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n) return;
	w16t[(i & ~4095u) | ((i & 63u) << 6) | ((i >> 6) & 63u)] = w16[i];
}

__global__ void init_state_kernel(DeviceState *st, float loss_scale) {
	nrc_optimizer_state s;
	s.t = 0;
	s.beta1_t = 1.0f;
	s.beta2_t = 1.0f;
	s.alpha_t = 1.0f;
	s.alpha_t_1 = 0.0f;
	s.loss_scale = loss_scale;
	s.clean_steps = 0;
	s.skipped_steps = 0;
	s.last_loss = 0.0f;
	s.last_count = 0;
	st->pub = s;
	st->done_blocks = 0;
	st->overflow = 0;
}

} // namespace

hipError_t launch_adam(const AdamArgs &a, hipStream_t stream) {
	// one block per 64 weights plus one for the loss column
	const dim3 grid(a.weight_count / 64 + 1), block(256);
	adam_kernel<<<grid, block, 0, stream>>>(a);
	return hipGetLastError();
}

hipError_t launch_reduce_partials(const float *partials, uint32_t nparts, uint32_t pstride, uint32_t n, float *out,
                                  hipStream_t stream) {
	reduce_kernel<<<dim3((n + 255) / 256), dim3(256), 0, stream>>>(partials, nparts, pstride, n, out);
	return hipGetLastError();
}

hipError_t launch_transpose_hidden(int hidden, const _Float16 *w16, _Float16 *w16t, hipStream_t stream) {
	const uint32_t n = (uint32_t)hidden * 4096u;
	transpose_hidden_kernel<<<dim3((n + 255) / 256), dim3(256), 0, stream>>>(w16, w16t, n);
	return hipGetLastError();
}

hipError_t launch_init_state(DeviceState *st, float loss_scale, hipStream_t stream) {
	init_state_kernel<<<1, 1, 0, stream>>>(st, loss_scale);
	return hipGetLastError();
}

} // namespace nrc
