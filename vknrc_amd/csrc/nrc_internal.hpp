// nrc_internal.hpp -- host-side launchers shared between the HIP translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "nrc.h"
#include "nrc_device.hpp"

namespace nrc {

// Samples per chunk of the plain weight-gradient reduction.
constexpr uint32_t kPlainChunk = 512;
// Records per chunk of the fused gradient kernel (4 waves, one 16-record N-tile each).
constexpr uint32_t kFusedSmallRecords = 64;
// Most chunks one fused-gradient workgroup walks.
constexpr uint32_t kFusedMaxChunks = 16;

// Most caches one grouped launch takes (the tables travel as kernel arguments).
constexpr uint32_t kGroupMax = 16;

// FP8 inference state of one cache (device-resident).  Activations a_l are
// carried scaled by C_l (C_0 = 1, C_{l+1} = C_l c_l), the layer-l weights by c_l
// (powers of two); the output is unscaled by out_scale = 1 / (c_H C_H).
constexpr int kFp8History = 16;  // frames of activation maxima
constexpr int kFp8Layers = 6;    // a_0 .. a_H for H <= 5
constexpr int kFp8MaxBlocks = 1024;
struct Fp8State {
	float hist[kFp8History][kFp8Layers];     // largest |a_l| (unscaled) of past frames
	float block_amax[kFp8MaxBlocks][kFp8Layers]; // this frame: per inference block, scaled
	float scale[kFp8Layers];                  // C_l of the weights now in the fp8 copy
	float wscale[kFp8Layers];                 // c_l = C_{l+1} / C_l (c_H for the output layer)
	float out_scale;
	uint32_t frames; // frames recorded into hist
};

// One cache's share of a grouped fused inference.
struct InferItem {
	const float *queries;    // [max_count][14]
	const uint32_t *d_count; // device count or null (= max_count)
	float *out;              // [max_count][3]
	const _Float16 *wf;      // fragment-major use-weights (FP16 path)
	const uint8_t *w8;       // fragment-major fp8 use-weights (FP8 path)
	Fp8State *f8;            // FP8 path: scales in, activation maxima out
	uint32_t max_count;
	uint32_t tile_begin; // first 64-record tile of this cache in the group (set by the launcher)
};
// One cache's FP8 weight preparation.
struct Fp8PrepItem {
	Fp8State *f8;
	const float *use32; // FP32 use-weights, row-major
	uint8_t *w8;        // fragment-major e4m3 copy of the scaled use-weights
};
struct Fp8PrepGroup {
	Fp8PrepItem item[kGroupMax];
	uint32_t n_items;
};

struct InferGroup {
	InferItem item[kGroupMax];
	const _Float16 *fin; // pre-encoded fp16 inputs [n][64] (single-item test entry point only)
	uint32_t n_items, total_tiles;
	int clamp0; // max(pred, 0) as nrc_inference.comp:48
};

// One cache's share of a grouped fused gradient.
struct GradItem {
	const float *records; // [max_count][17]
	const uint32_t *d_count;
	const _Float16 *wf, *wtf; // fragment-major forward / backward weights
	const DeviceState *state; // loss scale source (null: scale 1)
	float *partials;          // this cache's partial rows [blocks][pstride]
	uint32_t *overflow;       // non-finite flag of this step (null: none)
	uint32_t max_count, clamp, chunks, blocks, pstride;
	uint32_t block_begin; // first workgroup of this cache in the group (set by the launcher)
	int parity;
};
struct GradGroup {
	GradItem item[kGroupMax];
	const _Float16 *fin; // pre-encoded fp16 inputs / targets (single-item test entry point only)
	const float *ftgt;
	uint32_t n_items, total_blocks;
};

// How a fused gradient launch covers n records: `blocks` workgroups, each
// writing one partial-gradient row, walking `chunks` 64-record chunks.
struct FusedGradPlan {
	uint32_t chunks, blocks;
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
	_Float16 *w16, *wf, *wtf, *use16, *usef;
	float *w32, *use32;
	int write_use;
	int hidden;
	uint32_t weight_count;
	uint32_t block_begin; // first block of this cache in a grouped launch (set by the launcher)
	AdamParams p;
};
struct AdamGroup {
	AdamArgs item[kGroupMax];
	uint32_t n_items;
};

// nrc_plain.hip
hipError_t launch_plain_infer(int hidden, int encoding, const float *queries, const uint32_t *d_count,
                              uint32_t max_count, const float *w32, float *rgb, hipStream_t stream);
// Forward + backward per record, then the chunked dW reduction; returns the number of partials written.
hipError_t launch_plain_gradient(const TrainArgs &a, uint32_t *nparts, hipStream_t stream);
hipError_t launch_encode(int encoding, const float *queries, uint32_t n, float *features, hipStream_t stream);
hipError_t launch_sph_encode(const float *dirs, uint32_t n, float *uv, hipStream_t stream);
// The inference epilogue (screen composite and training-target back-fill).
struct CompositeArgs {
	const float *rgb;
	const uint32_t *dst, *d_count;
	uint32_t max_count;
	float *bias_factor_r;
	const float *factor_gb;
	uint32_t width, height, pitch;
	float *batches; // nrc_train_record rows of kTrainFloats floats
	const float *factors;
	uint32_t batch_stride;
};
hipError_t launch_composite(const CompositeArgs &a, hipStream_t stream);

// nrc_optim.hip
hipError_t launch_adam(const AdamArgs &a, hipStream_t stream);
// One launch for every cache of the group; fills block_begin.
hipError_t launch_adam_group(AdamGroup &grp, hipStream_t stream);
hipError_t launch_reduce_partials(const float *partials, uint32_t nparts, uint32_t pstride, uint32_t n, float *out,
                                  hipStream_t stream);
hipError_t launch_init_state(DeviceState *st, float loss_scale, hipStream_t stream);
// Scatters row-major fp16 weights into the fragment-major copies (whose padding
// must already be zero); wtf may be null.
hipError_t launch_make_frags(int hidden, const _Float16 *w16, _Float16 *wf, _Float16 *wtf, hipStream_t stream);
// Zeroes `words` 32-bit words at dst with a kernel (no memset node, so the launch
// may be captured into a graph).
hipError_t launch_zero_words(uint32_t *dst, size_t words, hipStream_t stream);

// nrc_records.hip: the record front-end, the wire-record epilogue, the path
// suffix sums and the screen pass
hipError_t launch_pack_primitives(const nrc_scene &s, uint32_t n, nrc_scene_primitive *out, hipStream_t stream);
hipError_t launch_unpack_eval(const nrc_scene &s, const nrc_eval_record *rec, const uint32_t *d_count,
                              uint32_t max_count, float *queries, hipStream_t stream);
hipError_t launch_unpack_train(const nrc_scene &s, const nrc_wire_train_record *rec, const uint32_t *d_count,
                               uint32_t max_count, float *out, hipStream_t stream);
hipError_t launch_composite_records(const float *rgb, const nrc_eval_record *rec, const uint32_t *d_count,
                                    uint32_t max_count, float *bias_factor_r, const float *factor_gb, uint32_t width,
                                    uint32_t height, uint32_t pitch, nrc_wire_train_record *batches,
                                    uint32_t batch_stride, hipStream_t stream);
hipError_t launch_build_train_records(const nrc_traced_path *paths, uint32_t n_paths, nrc_wire_train_record *batches,
                                      uint32_t *batch_counts, uint32_t capacity, nrc_eval_record *evals,
                                      uint32_t *eval_count, uint32_t eval_capacity, hipStream_t stream);
hipError_t launch_screen(const float *color, float *accumulate_img, uint32_t width, uint32_t height, uint32_t pitch,
                         int accumulate, uint32_t count, float *out, hipStream_t stream);

// nrc_fused.hip
// All items share (hidden, encoding); fills tile_begin and total_tiles.
hipError_t launch_fused_infer_group(int hidden, int encoding, InferGroup &grp, hipStream_t stream);
// FP8 path: per item, folds the last inference's activation maxima into the
// history, picks the layer scales and requantizes the fp8 use-weights from
// use32; then the grouped FP8 inference.
hipError_t launch_fp8_infer_group(int hidden, int encoding, InferGroup &grp, const float *const *use32,
                                  uint8_t *const *w8, hipStream_t stream);
// Forward over n pre-encoded fp16 input rows with fragment-major weights wf (no clamp).
hipError_t launch_fused_forward_f16(int hidden, const _Float16 *inputs, uint32_t n, const _Float16 *wf, float *out,
                                    hipStream_t stream);
// Workgroups and chunking of the fused gradient for a step of n records: chunks
// 64-record chunks a workgroup (0: from n and the CU count).  The chunking fixes
// the summation order of dW, so a cache keeps one setting alone and in groups.
FusedGradPlan plan_fused_gradient(uint32_t n, uint32_t chunks = 0);
hipError_t launch_fused_gradient(const TrainArgs &a, const FusedGradPlan &plan, hipStream_t stream);
// One launch over a group sharing (hidden, encoding, loss); fills block_begin and
// total_blocks.
hipError_t launch_fused_gradient_group(int hidden, int encoding, int loss, GradGroup &grp, hipStream_t stream);
GradItem grad_item(const TrainArgs &a, const FusedGradPlan &plan);
// Gradient of the L2 loss over n fp16 input rows; wf/wtf are the fragment-major copies.
hipError_t launch_fused_gradient_f16(int hidden, const _Float16 *inputs, const float *targets, uint32_t n,
                                     const _Float16 *wf, const _Float16 *wtf, float *partials, uint32_t pstride,
                                     uint32_t *nparts, hipStream_t stream);

int device_cu_count();
// Persistent inference workgroups per CU (default 2, VKNRC_INFER_WG_PER_CU or
// set_infer_wg_per_cu override; results do not depend on it).
uint32_t infer_wg_per_cu();
void set_infer_wg_per_cu(uint32_t v);
// One-time launch set-up of the fused kernels a cache of this kind uses (kernel
// attributes, device properties), done at cache creation so that no launch
// performs it under stream capture.
hipError_t prepare_fused(int hidden, int encoding, int loss);

static_assert(sizeof(InferGroup) <= 4096 && sizeof(GradGroup) <= 4096 && sizeof(AdamGroup) <= 4096,
              "grouped launch tables travel as kernel arguments");

} // namespace nrc
