// nrc_internal.hpp -- host-side launchers shared between the HIP translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "nrc.h"
#include "nrc_device.hpp"

namespace nrc {

// Samples per chunk of the plain weight-gradient reduction.
constexpr uint32_t kPlainChunk = 512;
// Records per chunk of the multi-chunk fused gradient kernel (8 waves, one
// 16-record N-tile each).
constexpr uint32_t kFusedTrainSamples = 128;
// Most chunks one multi-chunk fused-gradient workgroup walks.
constexpr uint32_t kFusedMaxChunks = 16;
// Records per workgroup of the single-chunk fused gradient kernel (4 waves, one
// 16-record N-tile each, weights resident in registers).
constexpr uint32_t kFusedSmallRecords = 64;

// How a fused gradient launch covers n records: `blocks` workgroups, each
// writing one partial-gradient row, walking `chunks` chunks of `records`.
struct FusedGradPlan {
	uint32_t records, chunks, blocks;
};

struct TrainArgs {
	int hidden;
	int encoding;
	int loss;
	const float *records; // [max_count][17]
	const uint32_t *d_count;
	uint32_t max_count;
	uint32_t clamp; // train batch size
	const float *w32;
	const _Float16 *w16;
	const _Float16 *w16t; // hidden layers transposed, [l][in][out]
	const _Float16 *wf, *wtf; // fragment-major forward / backward copies
	DeviceState *state;
	int parity;      // state buffer this step reads
	float *partials; // [nparts][pstride]
	uint32_t pstride;
	// plain-path scratch, feature-major [(hidden+1)][64][ldn]
	float *acts, *deltas, *sample_loss;
	uint32_t ldn;
};

struct AdamArgs {
	const float *partials;
	uint32_t nparts, pstride;
	const uint32_t *d_count;
	uint32_t max_count, clamp;
	DeviceState *state;
	int parity;
	nrc_optimizer_entry *entries;
	_Float16 *w16, *w16t, *wf, *wtf, *use16;
	float *w32, *use32;
	int write_use;
	int hidden;
	uint32_t weight_count;
	AdamParams p;
};

// nrc_plain.hip
hipError_t launch_plain_infer(int hidden, int encoding, const float *queries, const uint32_t *d_count,
                              uint32_t max_count, const float *w32, float *rgb, hipStream_t stream);
// Forward + backward per record, then the chunked dW reduction; returns the number of partials written.
hipError_t launch_plain_gradient(const TrainArgs &a, uint32_t *nparts, hipStream_t stream);
hipError_t launch_encode(int encoding, const float *queries, uint32_t n, float *features, hipStream_t stream);
hipError_t launch_sph_encode(const float *dirs, uint32_t n, float *uv, hipStream_t stream);

// nrc_optim.hip
hipError_t launch_adam(const AdamArgs &a, hipStream_t stream);
hipError_t launch_reduce_partials(const float *partials, uint32_t nparts, uint32_t pstride, uint32_t n, float *out,
                                  hipStream_t stream);
hipError_t launch_init_state(DeviceState *st, float loss_scale, hipStream_t stream);
// w16t[l][k][m] = w16[l][m][k] for the hidden layers
hipError_t launch_transpose_hidden(int hidden, const _Float16 *w16, _Float16 *w16t, hipStream_t stream);
// Scatters row-major fp16 weights into the fragment-major copies (whose padding
// must already be zero).
hipError_t launch_make_frags(int hidden, const _Float16 *w16, _Float16 *wf, _Float16 *wtf, hipStream_t stream);

// nrc_fused.hip
hipError_t launch_fused_infer(int hidden, int encoding, const float *queries, const uint32_t *d_count,
                              uint32_t max_count, const _Float16 *w16, float *rgb, hipStream_t stream);
hipError_t launch_fused_forward_f16(int hidden, const _Float16 *inputs, uint32_t n, const _Float16 *w16, float *out,
                                    hipStream_t stream);
// Workgroups and chunking of the fused gradient for a step of n records.
FusedGradPlan plan_fused_gradient(uint32_t n);
hipError_t launch_fused_gradient(const TrainArgs &a, const FusedGradPlan &plan, hipStream_t stream);
// Gradient of the L2 loss over n fp16 input rows; wf/wtf are the fragment-major copies.
hipError_t launch_fused_gradient_f16(int hidden, const _Float16 *inputs, const float *targets, uint32_t n,
                                     const _Float16 *w16, const _Float16 *w16t, const _Float16 *wf,
                                     const _Float16 *wtf, float *partials, uint32_t pstride, uint32_t *nparts,
                                     hipStream_t stream);

int device_cu_count();

} // namespace nrc
