// nrc_optim.hip -- partial-gradient reduction fused with Adam + EMA and the
// optimizer scalar recurrences (nrc_optimize.comp:32-54, nrc_train_prepare.comp:16-29).
//
// Each block owns W groups of 64 consecutive weights; its 256 threads sum the
// per-workgroup partial gradients as float4 columns in 16 interleaved slices
// (U loads in flight per group and thread) that are combined in a fixed order,
// so the reduced gradient is bit-identical from run to run.  U is the smallest
// power of two with 16 U >= the largest partial count of the launch (at most
// 16): with one round of loads every U gives the sums of U = 16, since the
// accumulators a smaller U leaves out would only add exact +0 to sums that
// start at +0.  W = 16 / U (at most 4) keeps 16 loads in flight a thread with
// fewer, fuller blocks when the rows are few.  Weight counts are multiples of 64
// and partial rows are 64-float aligned (the row stride holds the loss column and
// its 63 padding floats), so the float4 loads stay in bounds.  One extra block
// reduces the loss column.  The scalar state (t, beta1^t, beta2^t, alpha^t,
// alpha^{t-1}, loss scale, loss) is double-buffered by the step parity p: every
// block reads pub[p]; block 0 writes pub[p ^ 1] except the loss fields, which
// the loss block writes.
#include "nrc_internal.hpp"

namespace nrc {

namespace {

typedef float f4v __attribute__((ext_vector_type(4)));

// W: groups of 64 weights a block reduces (U W <= 16 loads a thread; each
// weight's sums are the same whichever block reduces it).
template <int U, int W> __global__ __launch_bounds__(256) void adam_kernel(const AdamGroup grp) {
	// this block's cache: the last item whose first block is at or below blockIdx.x
	// (block_begin does not decrease, so k counts the later items starting at or below
	// blockIdx.x; the unrolled count issues every table load at once)
	uint32_t k = 0;
#pragma unroll
	for (uint32_t j = 1; j < (uint32_t)kGroupMax; ++j)
		k += (j < grp.n_items && blockIdx.x >= grp.item[j].block_begin) ? 1u : 0u;
	const AdamArgs &a = grp.item[k];
	const uint32_t blk = blockIdx.x - a.block_begin;
	const int t = threadIdx.x;
	DeviceState *st = a.state;
	const int p = a.parity;
	// The state, count, overflow flag, partial rows and this thread's optimizer
	// entry are all loaded before the first decision that depends on one, so a
	// block waits on memory once for them: with few partial rows per cache the
	// kernel is latency-bound.  (The update below spells out its fused
	// multiply-adds, so where the entry is loaded does not change its rounding.)
	const nrc_optimizer_state s0 = st->pub[p];
	const uint32_t count = effective_count(a.d_count, a.max_count, a.clamp);
	const uint32_t ovf = st->overflow[p];
	const uint32_t wblocks = (a.weight_count + 64u * W - 1u) / (64u * W);
	const bool loss_block = blk == wblocks;
	// the block's columns: groups of 64 weights from w0 (the loss block: the loss
	// column and its padding, one group)
	const uint32_t w0 = loss_block ? a.weight_count : blk * 64u * W;
	const uint32_t ngroups = loss_block ? 1u : ((a.weight_count - w0) / 64u < W ? (a.weight_count - w0) / 64u : W);
	// thread t sums partials q = slice, slice + 16, ... of weights w0 + 64 gi + 4*col4 .. +3
	const int col4 = t & 15, slice = t >> 4;
	// and updates weight w0 + t
	const uint32_t i = w0 + (uint32_t)t;
	const bool updates = !loss_block && (uint32_t)t < 64u * ngroups;
	nrc_optimizer_entry e{};
	if (updates) e = a.entries[i];
	const float4 *src = reinterpret_cast<const float4 *>(a.partials + w0) + col4;
	const size_t ps4 = a.pstride / 4;
	float4 acc[W][U];
#pragma unroll
	for (int gi = 0; gi < W; ++gi)
#pragma unroll
		for (int u = 0; u < U; ++u) acc[gi][u] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
	// 16 U rows a round, all U W loads in flight before the first add (rows past
	// nparts load the last row and add as exact zeros, so the order of the sums is
	// that of the rows; groups past the last weight load group 0 and are unused)
	const uint32_t lastrow = a.nparts - 1;
	for (uint32_t q = slice; q < a.nparts; q += 16 * U) {
		float4 v[W][U];
#pragma unroll
		for (int gi = 0; gi < W; ++gi)
#pragma unroll
			for (int u = 0; u < U; ++u) {
				// non-temporal, as the gradient kernel stores the rows: nothing re-reads them
				const uint32_t row = q + 16 * u;
				const uint32_t col = (uint32_t)gi < ngroups ? 16u * gi : 0u;
				const f4v x = __builtin_nontemporal_load(
				    reinterpret_cast<const f4v *>(src + col + (size_t)(row < lastrow ? row : lastrow) * ps4));
				v[gi][u] = make_float4(x[0], x[1], x[2], x[3]);
			}
#pragma unroll
		for (int gi = 0; gi < W; ++gi)
#pragma unroll
			for (int u = 0; u < U; ++u) {
				const bool in = q + 16 * u < a.nparts;
				acc[gi][u].x += in ? v[gi][u].x : 0.0f;
				acc[gi][u].y += in ? v[gi][u].y : 0.0f;
				acc[gi][u].z += in ? v[gi][u].z : 0.0f;
				acc[gi][u].w += in ? v[gi][u].w : 0.0f;
			}
	}

	if (count == 0) { // nrc_optimize.comp:33-34: an empty batch changes nothing
		if (blk == 0 && t == 0) {
			st->pub[p ^ 1] = s0;
			st->overflow[p ^ 1] = 0;
		}
		return;
	}
	const bool skip = a.p.dynamic_scale && (ovf != 0);
	__shared__ float4 red[W][16][16];
#pragma unroll
	for (int gi = 0; gi < W; ++gi) {
		float4 G4 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
		if (!skip) {
			G4 = acc[gi][0];
#pragma unroll
			for (int u = 1; u < U; ++u) {
				G4.x += acc[gi][u].x; G4.y += acc[gi][u].y; G4.z += acc[gi][u].z; G4.w += acc[gi][u].w;
			}
		}
		red[gi][slice][col4] = G4;
	}
	__syncthreads();
	if (loss_block) {
		// column weight_count of the partials holds each workgroup's loss sum
		if (t == 0) {
			const float *rf = reinterpret_cast<const float *>(&red[0][0][0]);
			float L = 0.0f;
#pragma unroll
			for (int sl = 0; sl < 16; ++sl) L += rf[sl * 64];
			st->pub[p ^ 1].last_loss = skip ? s0.last_loss : L / (float)count;
			st->pub[p ^ 1].last_count = skip ? s0.last_count : count;
		}
		return;
	}
	const float b1 = a.p.beta1, b2 = a.p.beta2, al = a.p.ema_alpha;
	const float B1 = s0.beta1_t * b1, B2 = s0.beta2_t * b2; // nrc_train_prepare.comp:23-27
	const float A1 = s0.alpha_t, A = s0.alpha_t * al;
	if (updates && !skip) {
		const float *rf = reinterpret_cast<const float *>(&red[t >> 6][0][0]);
		float G = 0.0f;
#pragma unroll
		for (int sl = 0; sl < 16; ++sl) G += rf[sl * 64 + (t & 63)];
		float g = G / (float)count / s0.loss_scale; // nrc_optimize.comp:36
		if (!isfinite(g)) g = 0.0f;                 // nrc_optimize.comp:37-38
		// the products that are fused into the sums are written as fmaf; every
		// other product and quotient is rounded on its own
		e.m = fmaf(1.0f - b1, g, b1 * e.m);
		e.v = fmaf(1.0f - b2, g * g, b2 * e.v);
		const float mh = e.m / (1.0f - B1), vh = e.v / (1.0f - B2);
		e.weight -= a.p.lr * mh / (sqrtf(vh) + a.p.eps);
		const float eta = 1.0f - A, eta1 = 1.0f - A1; // nrc_optimize.comp:46-47, as written
		e.ema_weight = fmaf(al * eta1, e.ema_weight, (1.0f - al) / eta * e.weight);
		a.entries[i] = e;
		a.w32[i] = e.weight;
		const _Float16 h = (_Float16)e.weight;
		a.w16[i] = h;
		frag_scatter(i, h, a.wf, a.wtf);
		if (a.write_use) { // nrc_optimize.comp:50-53
			const float u = a.p.use_ema ? e.ema_weight : e.weight;
			a.use32[i] = u;
			const _Float16 uh = (_Float16)u;
			a.use16[i] = uh;
			frag_scatter(i, uh, a.usef, (_Float16 *)nullptr);
		}
	}
	if (blk == 0 && t == 0) {
		nrc_optimizer_state s = s0;
		if (skip) {
			s.loss_scale *= 0.5f;
			s.clean_steps = 0;
			s.skipped_steps += 1;
		} else {
			s.t += 1;
			s.beta1_t = B1;
			s.beta2_t = B2;
			s.alpha_t_1 = A1;
			s.alpha_t = A;
			if (a.p.dynamic_scale) {
				s.clean_steps += 1;
				if (a.p.growth_interval > 0 && s.clean_steps >= a.p.growth_interval) {
					s.loss_scale *= 2.0f;
					s.clean_steps = 0;
				}
			}
		}
		// every field but the loss pair, which the loss block writes
		nrc_optimizer_state *d = &st->pub[p ^ 1];
		d->t = s.t;
		d->beta1_t = s.beta1_t;
		d->beta2_t = s.beta2_t;
		d->alpha_t = s.alpha_t;
		d->alpha_t_1 = s.alpha_t_1;
		d->loss_scale = s.loss_scale;
		d->clean_steps = s.clean_steps;
		d->skipped_steps = s.skipped_steps;
		st->overflow[p ^ 1] = 0;
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

__global__ __launch_bounds__(256) void make_frags_kernel(const _Float16 *__restrict__ w16, _Float16 *__restrict__ wf,
                                                         _Float16 *__restrict__ wtf, uint32_t n) {
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n) return;
	frag_scatter(i, w16[i], wf, wtf);
}

__global__ __launch_bounds__(256) void zero_words_kernel(uint32_t *__restrict__ dst, size_t words) {
	for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < words; i += (size_t)gridDim.x * blockDim.x)
		dst[i] = 0u;
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
	st->pub[0] = s;
	st->pub[1] = s;
	st->overflow[0] = 0;
	st->overflow[1] = 0;
}

} // namespace

hipError_t launch_adam_group(AdamGroup &grp, hipStream_t stream) {
	// per cache: one block per W groups of 64 weights plus one for the loss column;
	// U loads a slice cover the largest partial count of the launch, and a block
	// takes W = 16 / U groups (at most 4), so a thread has at most 16 loads in flight
	uint32_t maxparts = 0;
	for (uint32_t i = 0; i < grp.n_items; ++i)
		if (grp.item[i].nparts > maxparts) maxparts = grp.item[i].nparts;
	const uint32_t w = maxparts <= 64 ? 4u : maxparts <= 128 ? 2u : 1u;
	uint32_t blocks = 0;
	for (uint32_t i = 0; i < grp.n_items; ++i) {
		grp.item[i].block_begin = blocks;
		blocks += (grp.item[i].weight_count + 64u * w - 1u) / (64u * w) + 1u;
	}
	if (blocks == 0) return hipSuccess;
	const dim3 g(blocks), b(256);
	if (maxparts <= 16) adam_kernel<1, 4><<<g, b, 0, stream>>>(grp);
	else if (maxparts <= 32) adam_kernel<2, 4><<<g, b, 0, stream>>>(grp);
	else if (maxparts <= 64) adam_kernel<4, 4><<<g, b, 0, stream>>>(grp);
	else if (maxparts <= 128) adam_kernel<8, 2><<<g, b, 0, stream>>>(grp);
	else adam_kernel<16, 1><<<g, b, 0, stream>>>(grp);
	return hipGetLastError();
}

hipError_t launch_adam(const AdamArgs &a, hipStream_t stream) {
	AdamGroup grp{};
	grp.n_items = 1;
	grp.item[0] = a;
	return launch_adam_group(grp, stream);
}

hipError_t launch_reduce_partials(const float *partials, uint32_t nparts, uint32_t pstride, uint32_t n, float *out,
                                  hipStream_t stream) {
	reduce_kernel<<<dim3((n + 255) / 256), dim3(256), 0, stream>>>(partials, nparts, pstride, n, out);
	return hipGetLastError();
}

hipError_t launch_make_frags(int hidden, const _Float16 *w16, _Float16 *wf, _Float16 *wtf, hipStream_t stream) {
	const uint32_t n = (uint32_t)weight_count(hidden);
	make_frags_kernel<<<dim3((n + 255) / 256), dim3(256), 0, stream>>>(w16, wf, wtf, n);
	return hipGetLastError();
}

hipError_t launch_zero_words(uint32_t *dst, size_t words, hipStream_t stream) {
	if (words == 0) return hipSuccess;
	size_t blocks = (words + 255) / 256;
	if (blocks > 1024) blocks = 1024;
	zero_words_kernel<<<dim3((uint32_t)blocks), dim3(256), 0, stream>>>(dst, words);
	return hipGetLastError();
}

hipError_t launch_init_state(DeviceState *st, float loss_scale, hipStream_t stream) {
	init_state_kernel<<<1, 1, 0, stream>>>(st, loss_scale);
	return hipGetLastError();
}

} // namespace nrc
