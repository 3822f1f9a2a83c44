// nrc_optim.hip -- partial-gradient reduction fused with Adam + EMA and the
// optimizer scalar recurrences (nrc_optimize.comp:32-54, nrc_train_prepare.comp:16-29).
//
// Each block owns 64 consecutive weights; its 256 threads sum the per-workgroup
// partial gradients in four interleaved slices that are combined in a fixed
// order, so the reduced gradient is bit-identical from run to run.  The last
// block to finish advances the scalar state (t, beta1^t, beta2^t, alpha^t,
// alpha^{t-1}, loss scale, loss), which every block read at its start.
#include "nrc_internal.hpp"

namespace nrc {

namespace {

__global__ __launch_bounds__(256) void adam_kernel(AdamArgs a) {
	const uint32_t count = effective_count(a.d_count, a.max_count, a.clamp);
	if (count == 0) return; // nrc_optimize.comp:33-34: an empty batch changes nothing
	const int t = threadIdx.x;
	const int slice = t >> 6;
	const uint32_t i = blockIdx.x * 64u + (uint32_t)(t & 63);
	DeviceState *st = a.state;
	const nrc_optimizer_state s0 = st->pub;
	const bool skip = a.p.dynamic_scale && (st->overflow != 0);

	__shared__ float red[4][64];
	float G = 0.0f;
	if (!skip && i < a.weight_count) {
		const float *p = a.partials + i;
		for (uint32_t q = slice; q < a.nparts; q += 4) G += p[(size_t)q * a.pstride];
	}
	red[slice][t & 63] = G;
	__syncthreads();
	if (slice == 0 && !skip && i < a.weight_count) {
		G = ((red[0][t] + red[1][t]) + red[2][t]) + red[3][t];
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
		a.w16[i] = (_Float16)e.weight;
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
				float ls = 0.0f;
				for (uint32_t q = 0; q < a.nparts; ++q) ls += a.partials[(size_t)q * a.pstride + a.weight_count];
				s.last_loss = ls / (float)count;
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
	const dim3 grid((a.weight_count + 63) / 64), block(256);
	adam_kernel<<<grid, block, 0, stream>>>(a);
	return hipGetLastError();
}

hipError_t launch_reduce_partials(const float *partials, uint32_t nparts, uint32_t pstride, uint32_t n, float *out,
                                  hipStream_t stream) {
	reduce_kernel<<<dim3((n + 255) / 256), dim3(256), 0, stream>>>(partials, nparts, pstride, n, out);
	return hipGetLastError();
}

hipError_t launch_init_state(DeviceState *st, float loss_scale, hipStream_t stream) {
	init_state_kernel<<<1, 1, 0, stream>>>(st, loss_scale);
	return hipGetLastError();
}

} // namespace nrc
