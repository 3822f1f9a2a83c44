/*
 * nrc.h -- C ABI of the MI355X neural-radiance-cache hot path (libvknrc.so).
 *
 * Drop-in boundary for VkNRC's NRC state object and NN passes: a cache owns the
 * network weights and optimizer state (VkNRCState), nrc_infer replaces the
 * NNDispatch<NNInference> pass, nrc_train replaces the NNTrain pass group
 * (zero gradient -> prepare -> gradient -> optimize).  Plain pointers, sizes and
 * status codes only; device pointers are HIP device memory, streams are
 * hipStream_t passed as void* (NULL = the null stream).  Calls on one cache are
 * ordered by the stream they are issued on; distinct caches may be driven from
 * distinct host threads.  No call aborts the process.
 */
#ifndef NRC_H
#define NRC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Network shape: 64 -> hidden x (64x64, ReLU) -> 3, no bias.
 * src/VkNRCState.hpp:23-25 (kNNHiddenLayers = 5, kNNWidth = 64, kNNOutWidth = 3). */
#define NRC_WIDTH 64
#define NRC_OUTPUT_WIDTH 3
#define NRC_MAX_HIDDEN_LAYERS 5
/* Constant.glsl:6-7: 4 train batches of at most 16384 records per frame. */
#define NRC_TRAIN_BATCH_COUNT 4
#define NRC_DEFAULT_TRAIN_BATCH_SIZE 16384
/* Largest train batch one nrc_train call accepts. */
#define NRC_MAX_TRAIN_BATCH_SIZE 65536
/* Most queries or records one call takes: max_count of nrc_infer and the record
 * and epilogue calls, max_queries of nrc_frame and the group calls, n of
 * nrc_encode, nrc_sph_encode and nrc_mlp_forward_f16, n_paths and eval_capacity
 * of nrc_build_train_records.  Counts are uint32 and the launches round them up to
 * whole tiles and blocks; larger counts give NRC_ERROR_CAPACITY. */
#define NRC_MAX_QUERIES 0x80000000u

typedef enum nrc_status {
	NRC_OK = 0,
	NRC_ERROR_INVALID_ARGUMENT = 1,
	NRC_ERROR_UNSUPPORTED = 2, /* NNInferenceShader.cpp:33-34 returns a null pipeline in this case */
	NRC_ERROR_OUT_OF_MEMORY = 3,
	NRC_ERROR_HIP = 4,
	NRC_ERROR_CAPACITY = 5
} nrc_status;

/* Kernel form.  PLAIN_F32 is the plain method: FP32 features, weights and
 * accumulation.  FUSED_F16 runs the layer GEMMs on FP16 MFMA with FP32
 * accumulation, FP16 activations and FP32 master weights.  FUSED_F8 runs the
 * inference forward on FP8 (OCP e4m3) MFMA with per-layer power-of-two scales
 * chosen from the largest activations of the cache's last 16 FP8 inferences;
 * its training steps are those of FUSED_F16 (FP16 backward, FP32 master
 * weights).  No counterpart in the reference, which runs FP16 throughout. */
typedef enum nrc_path { NRC_PATH_PLAIN_F32 = 0, NRC_PATH_FUSED_F16 = 1, NRC_PATH_FUSED_F8 = 2 } nrc_path;

/* Input encoding (NRCRecord.glsl:78-95).  TRI12 is the reference layout.
 * SINCOS6 replaces each position axis's 12 triangle waves by sin(2^k pi p),
 * k = 0..5, followed by cos(2^k pi p).  UV_SINCOS16 reads (u, v) from
 * position[0..1] and emits sin/cos(2^k pi u), k = 0..15, then the same for v. */
typedef enum nrc_encoding { NRC_ENCODING_TRI12 = 0, NRC_ENCODING_SINCOS6 = 1, NRC_ENCODING_UV_SINCOS16 = 2 } nrc_encoding;

/* NNLoadDA3_RelativeL2LuminanceLoss / NNLoadDA3_L2Loss, NN_nv.glsl:162-196. */
typedef enum nrc_loss { NRC_LOSS_RELATIVE_L2_LUMINANCE = 0, NRC_LOSS_L2 = 1 } nrc_loss;

/* VkNRCState.cpp:39-44 is Kaiming normal N(0, sqrt(2/64)) for every weight. */
typedef enum nrc_init { NRC_INIT_KAIMING_NORMAL = 0, NRC_INIT_XAVIER_UNIFORM = 1 } nrc_init;

/* One query: UnpackedNRCInput, NRCRecord.glsl:40-45 (14 floats, 56 bytes).
 * scattered_dir and normal are already sphere-encoded to [0,1]^2 (NRCSphEncode). */
typedef struct nrc_query {
	float position[3];
	float scattered_dir[2];
	float normal[2];
	float roughness;
	float diffuse[3];
	float specular[3];
} nrc_query;

/* One training record: the query plus its target radiance (68 bytes).  The
 * reference's NRCTrainRecord (NRCRecord.glsl:35-38) carries the target in its
 * bias field once the inference epilogue has back-filled it. */
typedef struct nrc_train_record {
	nrc_query query;
	float target[3];
} nrc_train_record;

/* OptimizerEntry, VkNRCState.cpp:29-31 / nrc_optimize.comp:19-22. */
typedef struct nrc_optimizer_entry {
	float m, v, weight, ema_weight;
} nrc_optimizer_entry;

/* OptimizerState, VkNRCState.cpp:25-28 / nrc_train_prepare.comp:11-14, extended
 * with the dynamic loss-scale state and the last step's loss. */
typedef struct nrc_optimizer_state {
	uint32_t t;
	float beta1_t, beta2_t, alpha_t, alpha_t_1;
	float loss_scale;
	uint32_t clean_steps;   /* steps since the loss scale last changed */
	uint32_t skipped_steps; /* steps skipped because the gradient overflowed */
	float last_loss;        /* mean per-record loss of the last non-empty step */
	uint32_t last_count;    /* records used by the last non-empty step */
} nrc_optimizer_state;

typedef struct nrc_cache_desc {
	uint32_t hidden_layers; /* 5 (reference) or 4 */
	nrc_encoding encoding;
	nrc_loss loss;
	nrc_init init;
	uint64_t seed;                 /* weight init seed (the reference draws std::random_device) */
	uint32_t train_batch_size;     /* per-step clamp, nrc_train_prepare.comp:17-19 (16384) */
	float learning_rate;           /* nrc_optimize.comp:10 (0.002) */
	float beta1, beta2, epsilon;   /* Constant.glsl:11-12, nrc_optimize.comp:11 (0.9, 0.999, 1e-8) */
	float ema_alpha;               /* Constant.glsl:13 (0.99) */
	float loss_scale;              /* Constant.glsl:8 LOSS_SCALE (1.0) or the dynamic start value */
	int32_t dynamic_loss_scale;    /* 1: skip the step and halve the scale on overflow */
	uint32_t loss_scale_growth_interval; /* clean steps before the scale doubles */
	int32_t use_ema_weights;       /* VkNRCState::SetUseEMAWeights, default off */
} nrc_cache_desc;

typedef struct nrc_cache nrc_cache;

/* Flags of nrc_train. */
#define NRC_TRAIN_WRITE_USE_WEIGHTS 1u /* -DWRITE_USE_WEIGHTS variant of nrc_optimize.comp:46-53 */

const char *nrc_status_string(nrc_status status);
/* Message of the last failed call on this host thread ("" when none). */
const char *nrc_last_error(void);

/* Fills the reference constants: 5 hidden layers, TRI12, relative L2,
 * Kaiming init, batch 16384, Adam 0.002/0.9/0.999/1e-8, EMA 0.99, scale 1. */
void nrc_cache_desc_default(nrc_cache_desc *desc);

/* Weight count: hidden * 64 * 64 + 3 * 64 (VkNRCState::GetWeightCount, VkNRCState.hpp:25,88);
 * 0 for an unsupported depth. */
uint32_t nrc_weight_count(uint32_t hidden_layers);

/* Host-only: the FP32 initial weights a cache created from desc with this seed
 * starts from (VkNRCState::initialize_weights, VkNRCState.cpp:39-44). */
nrc_status nrc_init_weights(const nrc_cache_desc *desc, uint64_t seed, float *host_weights);

/* VkNRCState ctor + ResetMLPBuffers (VkNRCState.hpp:42-45, VkNRCState.cpp:46-88):
 * allocates the cache on the current HIP device and initialises it (synchronous). */
nrc_status nrc_cache_create(const nrc_cache_desc *desc, nrc_cache **out_cache);
/* ResetMLPBuffers ("Re-Train", src/main.cpp:144-147) with a new seed (synchronous). */
nrc_status nrc_cache_reset(nrc_cache *cache, uint64_t seed);
void nrc_cache_destroy(nrc_cache *cache);
nrc_status nrc_cache_get_desc(const nrc_cache *cache, nrc_cache_desc *desc);
/* SetUseEMAWeights, VkNRCState.hpp:69. Takes effect at the next use-weights write. */
nrc_status nrc_set_use_ema(nrc_cache *cache, int32_t use_ema);

/*
 * NNDispatch<NNInference> (NNDispatch.hpp:13-57, NNInference.cpp:11-71,
 * nrc_inference.comp:30-48): rgb[i] = max(MLP(encode(queries[i])), 0) for
 * i < min(*d_count, max_count), using the cache's use-weights.  d_count is a
 * device pointer (NULL: count = max_count); d_rgb is [max_count][3] floats.
 */
nrc_status nrc_infer(nrc_cache *cache, const nrc_query *d_queries, const uint32_t *d_count, uint32_t max_count,
                     float *d_rgb, nrc_path path, void *stream);

/*
 * NNTrain (NNTrain.hpp:17-128, NNTrain.cpp:11-133): one optimizer step on the
 * first min(*d_count, max_count, train_batch_size) records.  An empty batch
 * changes nothing (nrc_optimize.comp:33-34).  max_count <= NRC_MAX_TRAIN_BATCH_SIZE.
 */
nrc_status nrc_train(nrc_cache *cache, const nrc_train_record *d_records, const uint32_t *d_count,
                     uint32_t max_count, nrc_path path, uint32_t flags, void *stream);

/*
 * One frame of NRCRenderGraph (NRCRenderGraph.cpp:46-80): inference with the
 * use-weights, then NRC_TRAIN_BATCH_COUNT chained train steps, the last of
 * which writes the use-weights.  d_batch_counts holds NRC_TRAIN_BATCH_COUNT
 * device counts (NULL: every batch uses max_batch_records); d_batches holds
 * NRC_TRAIN_BATCH_COUNT * max_batch_records records.
 */
nrc_status nrc_frame(nrc_cache *cache, const nrc_query *d_queries, const uint32_t *d_query_count,
                     uint32_t max_queries, float *d_rgb, const nrc_train_record *d_batches,
                     const uint32_t *d_batch_counts, uint32_t max_batch_records, nrc_path path, void *stream);

/* Synchronous state export/import (the buffers GetWeightBuffer,
 * GetUseWeightBuffer, GetOptimizerEntryBuffer, GetOptimizerStateBuffer expose,
 * VkNRCState.hpp:49-52).  Any pointer may be NULL to skip that part.  fp16
 * arrays are IEEE binary16 bit patterns.  nrc_set_state re-derives the fp16
 * weights (and the use-weights, as weight or EMA per use_ema) from the entries. */
nrc_status nrc_get_state(nrc_cache *cache, nrc_optimizer_entry *host_entries, nrc_optimizer_state *host_state,
                         uint16_t *host_weights_f16, uint16_t *host_use_weights_f16);
nrc_status nrc_set_state(nrc_cache *cache, const nrc_optimizer_entry *host_entries,
                         const nrc_optimizer_state *host_state);

/* Device bytes a cache with this description allocates (each allocation
 * rounded up to 4 KiB; the plain path's training scratch, allocated on its first
 * plain training step, counted when path is PLAIN_F32).  For admission control
 * by an HBM budget; no device call.  The reference sizes a cache's record
 * buffers with VkNRCState::GetEvalRecordBufferSize / GetBatchTrainRecordBufferSize
 * (VkNRCState.cpp:34-37) and has no admission control. */
nrc_status nrc_cache_footprint(const nrc_cache_desc *desc, nrc_path path, uint64_t *device_bytes);

/* Launch schedule of the cache's fused training steps: `chunks` 64-record chunks
 * per gradient workgroup, 1..16 (0, the default: chosen from the batch size so
 * that one cache's step fills the GPU).  The chunking fixes the summation order
 * of the batch-reduced dW, so a cache trains to the same bits alone and in any
 * group with one setting; the autotuner sets it per workload.  No counterpart in
 * the reference, which runs 128-invocation workgroups (nrc_gradient.comp:5-6) and
 * sums dW with float atomics (NN_nv.glsl:311-313, 361-363). */
nrc_status nrc_cache_set_grad_chunks(nrc_cache *cache, uint32_t chunks);
/* Persistent inference workgroups per CU, 1..8 (default 2), for every later
 * launch in the process.  Inference results do not depend on it. */
nrc_status nrc_set_infer_workgroups_per_cu(uint32_t workgroups);

/* The FP8 scales the cache's last FP8 inference ran with (synchronous):
 * layer_scales[l] = C_l for l = 0..hidden_layers (the activations a_l enter
 * layer l as C_l a_l in e4m3, layer l's weights as (C_{l+1} / C_l) W_l), and
 * out_scale = 1 / (c_H C_H) with c_H the output layer's weight scale;
 * frames = FP8 inferences recorded in the 16-frame activation history.
 * layer_scales needs hidden_layers + 1 entries; any pointer may be NULL.
 * No counterpart in the reference, which has no FP8 path. */
nrc_status nrc_get_fp8_scales(nrc_cache *cache, float *layer_scales, float *out_scale, uint32_t *frames);

/* NRCInputEncode (NRCRecord.glsl:78-95) alone: d_features is [n][64] FP32
 * (what the plain path feeds its first layer). */
nrc_status nrc_encode(nrc_encoding encoding, const nrc_query *d_queries, uint32_t n, float *d_features, void *stream);
/* NRCSphEncode (NRCRecord.glsl:47-49): d_dirs [n][3] -> d_uv [n][2]. */
nrc_status nrc_sph_encode(const float *d_dirs, uint32_t n, float *d_uv, void *stream);

/* ---- inference epilogue: screen composite and training-target back-fill
 * (nrc_inference.comp:50-72).  Every query carries a destination code
 * (NRCEvalRecord.dst, NRCRecord.glsl:12-33):
 *   NRC_DST_INVALID                    nothing;
 *   screen: ((x | y << 15) << 1)       pixel (x, y), x, y < 2^15;
 *   train:  ((b | l << 2 | r << 16) << 1) | 1   records l..r of batch b < 4. */
#define NRC_DST_INVALID 0xFFFFFFFFu
/* EncodeNRCEvalDstScreen / EncodeNRCEvalDstTrain (NRCRecord.glsl:19-22); the
 * fields are masked to their widths. */
uint32_t nrc_dst_screen(uint32_t x, uint32_t y);
uint32_t nrc_dst_train(uint32_t batch, uint32_t l, uint32_t r);

/* The screen of the composite: bias_factor_r is an RGBA32F image
 * [height][pitch] holding (bias rgb, factor r) and receives (color, 0);
 * factor_gb is an RG32F image [height][pitch] (uBias_FactorR / uFactorGB,
 * nrc_inference.comp:25-26).  Pitches are in pixels. */
typedef struct nrc_screen {
	float *d_bias_factor_r;
	const float *d_factor_gb;
	uint32_t width, height, pitch;
} nrc_screen;

/* For each query i < count (count = *d_count clamped to max_count, or
 * max_count when d_count is NULL), with pred = d_rgb[i] (the output of
 * nrc_infer, already clamped at 0):
 *   screen code: bias_factor_r[y][x] = (bias + (factor_r, factor_gb[y][x]) * pred, 0);
 *   train code:  for j = l .. r: d_batches[b * batch_stride + j].target
 *                += d_train_factors[(b * batch_stride + j) * 3 ..] * pred
 *                (the reference's NRCTrainRecord bias, which nrc_train reads as
 *                the target; its factor travels in d_train_factors).
 * Each product and sum is rounded separately (no FMA contraction).  Pixels
 * outside width x height and records j >= batch_stride are skipped, as a robust
 * image store or buffer write would drop them.  Destinations of different
 * queries must not overlap (true of the reference's path tracer, which gives
 * every eval record its own pixel or record range).  screen or d_batches may be
 * NULL when no code of that kind occurs. */
nrc_status nrc_composite(const float *d_rgb, const uint32_t *d_dst, const uint32_t *d_count, uint32_t max_count,
                         const nrc_screen *screen, nrc_train_record *d_batches, const float *d_train_factors,
                         uint32_t batch_stride, void *stream);

/* ---- record front-end: the renderer's wire records and the scene they point into.
 * The reference's path tracer writes 16-byte packed path vertices (PackedNRCInput,
 * NRCRecord.glsl:6-10; MakeNRCInput, path_tracer.comp:159-166) inside 20-byte eval
 * records (NRCEvalRecord, NRCRecord.glsl:12-17) and 40-byte train records
 * (NRCTrainRecord, NRCRecord.glsl:35-38); its shaders unpack them against the scene
 * (UnpackNRCInput, NRCRecord.glsl:98-125; Scene.glsl:8-69). */
typedef struct nrc_packed_input {
	uint32_t primitive_id;
	uint32_t flip_bit_instance_id; /* instance id | flip-normal bit << 31 */
	uint32_t barycentric_2x16u;    /* unorm16 (b1, b2); b0 = 1 - b1 - b2 */
	uint32_t scattered_dir_2x16u;  /* unorm16 sphere-encoded direction */
} nrc_packed_input;

typedef struct nrc_eval_record {
	uint32_t dst; /* NRC_DST_INVALID, nrc_dst_screen(x, y) or nrc_dst_train(b, l, r) */
	nrc_packed_input packed_input;
} nrc_eval_record;

/* bias: the radiance gathered after this vertex (the training target once the
 * inference epilogue has added factor * the cache's prediction at the path's end) */
typedef struct nrc_wire_train_record {
	float bias[3], factor[3];
	nrc_packed_input packed_input;
} nrc_wire_train_record;

/* Material (Scene.glsl:12-19, std430: 64 bytes); a texture id of 0xFFFFFFFF means
 * the constant colour. */
typedef struct nrc_material {
	float diffuse[3];
	uint32_t diffuse_texture_id;
	float specular[3];
	uint32_t specular_texture_id;
	float emission[3];
	uint32_t emission_texture_id;
	float metallic, roughness, ior, pad;
} nrc_material;

/* An RGBA8 sRGB texture (VkScene.cpp:193), sampled with linear filtering and
 * repeat addressing (NRCRenderGraph.cpp:132): each texel is decoded from sRGB
 * before the bilinear blend. */
typedef struct nrc_texture {
	const uint32_t *d_rgba8; /* [height][width], R in the low byte */
	uint32_t width, height;
} nrc_texture;

/* The scene buffers of Scene.glsl:26-47 (device pointers).  A vertex of primitive
 * p is d_vertices[d_vertex_indices[3p + k]] transformed by its instance's
 * d_transforms[i] (3 x 4 floats, row j giving world_j = row_j . (x, y, z, 1)). */
/* One primitive in one 64-byte record: its corners in object space (before the
 * instance transform), their texture coordinates and its material id, as the index
 * arrays of nrc_scene name them.  No counterpart in the reference, which gathers them
 * through the indices per record (Scene.glsl:50-62); see nrc_scene_pack_primitives. */
typedef struct nrc_scene_primitive {
	float corners[3][3];
	float texcoords[3][2];
	uint32_t material_id;
} nrc_scene_primitive;

typedef struct nrc_scene {
	const float *d_vertices;            /* [vertices][3] */
	const uint32_t *d_vertex_indices;   /* [primitives * 3] */
	const float *d_texcoords;           /* [texcoords][2] */
	const uint32_t *d_texcoord_indices; /* [primitives * 3] */
	const nrc_material *d_materials;
	const uint32_t *d_material_ids;     /* [primitives] */
	const float *d_transforms;          /* [instances][3][4] */
	const nrc_texture *d_textures;      /* device array of texture_count descriptors, or NULL */
	uint32_t texture_count;
	/* optional: the primitives packed by nrc_scene_pack_primitives (NULL: the unpack
	 * gathers through the index arrays); the unpacked records are the same bit for bit */
	const nrc_scene_primitive *d_primitives;
} nrc_scene;

/* Packs primitives [0, primitive_count) of the scene into d_out [primitive_count],
 * 64-byte aligned (device memory from the caller), once per scene: unpacking a record
 * then reads one 64-byte record for its primitive instead of about nine scattered
 * cache lines of indices, corners, texcoords and material ids.  primitive_count must
 * not exceed the primitives the index arrays hold, and every index must name an
 * existing vertex and texcoord (as the unpack itself assumes).  No counterpart in the
 * reference; for callers that unpack many records on one scene. */
nrc_status nrc_scene_pack_primitives(const nrc_scene *scene, uint32_t primitive_count, nrc_scene_primitive *d_out,
                                     void *stream);

/* UnpackNRCInput (NRCRecord.glsl:98-125) of the first min(*d_count, max_count)
 * eval records (d_count NULL: max_count) into d_queries [max_count]; the records'
 * dst codes stay in d_records for nrc_composite_records. */
nrc_status nrc_unpack_eval_records(const nrc_scene *scene, const nrc_eval_record *d_records, const uint32_t *d_count,
                                   uint32_t max_count, nrc_query *d_queries, void *stream);
/* The same for train records, into the records nrc_train reads: the unpacked
 * query and target = bias (nrc_gradient.comp:30-33). */
nrc_status nrc_unpack_train_records(const nrc_scene *scene, const nrc_wire_train_record *d_records,
                                    const uint32_t *d_count, uint32_t max_count, nrc_train_record *d_out,
                                    void *stream);
/* The inference epilogue over wire records (nrc_inference.comp:50-72): for each
 * eval record i < count, with pred = d_rgb[i] (the output of nrc_infer on the
 * unpacked queries): a screen code composites into the screen as nrc_composite;
 * a train code (b, l, r) sets bias = bias + factor * pred in records l..r of
 * d_batches + b * batch_stride (each product and sum rounded separately).  Records
 * j >= batch_stride and pixels outside the screen are skipped. */
nrc_status nrc_composite_records(const float *d_rgb, const nrc_eval_record *d_records, const uint32_t *d_count,
                                 uint32_t max_count, const nrc_screen *screen, nrc_wire_train_record *d_batches,
                                 uint32_t batch_stride, void *stream);

/* Path vertices per path (path_tracer.comp:11, 256-257). */
#define NRC_MAX_BOUNCE 8
/* One traced path as ExtendedPathTrace (path_tracer.comp:250-372) holds it before
 * its records are written: vertex i has its packed input, the light it sees
 * (emission, or the constant sky light at a miss) and the colour (BRDF weight) of
 * its scattering. */
typedef struct nrc_traced_path {
	uint32_t bounce;          /* vertices 0 .. bounce - 1 are valid, 1 <= bounce <= NRC_MAX_BOUNCE */
	uint32_t batch;           /* train batch 0 .. NRC_TRAIN_BATCH_COUNT - 1 */
	uint32_t light_terminate; /* the path ended at a light: no eval record for its end */
	uint32_t pad;
	float light[NRC_MAX_BOUNCE][3];
	float color[NRC_MAX_BOUNCE][3];
	nrc_packed_input input[NRC_MAX_BOUNCE];
} nrc_traced_path;

/* The tail of ExtendedPathTrace (path_tracer.comp:343-372) for n_paths paths:
 * each takes `bounce` slots of its batch by an atomic add on
 * d_batch_counts[batch]; if the first fits under batch_capacity, the suffix sums
 * light[i] += color[i] * light[i + 1], color[i] *= color[i + 1] (i = bounce - 2 .. 0)
 * give the bias and factor of its train records (those that fit), and, unless the
 * path ended at a light, an eval record appended at d_eval_count with the train
 * code (batch, first, last) and the packed input of vertex bounce - 1 asks the
 * inference for the radiance past the path's end.  Record order within a batch
 * follows the atomics; counts past the capacity keep growing, as the reference's
 * do (consumers clamp them). */
nrc_status nrc_build_train_records(const nrc_traced_path *d_paths, uint32_t n_paths, nrc_wire_train_record *d_batches,
                                   uint32_t *d_batch_counts, uint32_t batch_capacity, nrc_eval_record *d_eval_records,
                                   uint32_t *d_eval_count, uint32_t eval_capacity, void *stream);

/* The screen pass (screen.frag:10-31) over an RGBA32F colour image [height][pitch]:
 * when accumulate is set, colour = (colour + accumulated * count) / (count + 1)
 * (just colour when count == 0) and the RGBA32F accumulation image receives
 * (colour, 0); the output [height][pitch] RGBA32F pixel is
 * (pow(filmic(colour), 1/2.2), 1) with the Hejl 2015 filmic curve at white
 * point 3.2.  d_accumulate may be NULL when accumulate is 0. */
nrc_status nrc_screen_tonemap(const float *d_color, float *d_accumulate, uint32_t width, uint32_t height,
                              uint32_t pitch, int32_t accumulate, uint32_t accumulate_count, float *d_out,
                              void *stream);

/* The reference test kernels over pre-encoded inputs (test/evaluate_NV.comp:16-31,
 * test/train_NV.comp:18-46): fp16 weights [weight_count], fp16 inputs [n][64].
 * forward: d_out [n][3] FP32 raw outputs.  gradient: L2 loss against d_targets
 * [n][3] FP32, d_dw [weight_count] FP32 summed over the n samples (overwritten). */
nrc_status nrc_mlp_forward_f16(uint32_t hidden_layers, const uint16_t *d_weights_f16, const uint16_t *d_inputs_f16,
                               uint32_t n, float *d_out, void *stream);
nrc_status nrc_mlp_gradient_f16(uint32_t hidden_layers, const uint16_t *d_weights_f16, const uint16_t *d_inputs_f16,
                                const float *d_targets, uint32_t n, float *d_dw, void *stream);

/* Largest n nrc_mlp_gradient_f16_noalloc takes (its scratch is sized for it). */
#define NRC_NOALLOC_MAX_GRADIENT_RECORDS 16384u

/* nrc_mlp_gradient_f16 without any allocation: no counterpart in the reference.
 * For callers that cannot allocate device memory (code under HIP graph capture,
 * serving loops that forbid hipMalloc).  Same kernels and results (bit for bit) as
 * nrc_mlp_gradient_f16, with its scratch (the fragment-major weight copies and the
 * partial-gradient rows) in a fixed device array of 21.3 MB that each GPU the
 * library runs on holds until the process ends.  n <= NRC_NOALLOC_MAX_GRADIENT_RECORDS
 * (NRC_ERROR_CAPACITY beyond).  All calls on one GPU share that array, so calls
 * that overlap in time on one GPU (different streams or host threads without
 * ordering between them) race; calls ordered on one stream do not. */
nrc_status nrc_mlp_gradient_f16_noalloc(uint32_t hidden_layers, const uint16_t *d_weights_f16,
                                        const uint16_t *d_inputs_f16, const float *d_targets, uint32_t n,
                                        float *d_dw, void *stream);

/*
 * Many caches at once.  The reference drives one cache per renderer
 * (NRCRenderGraph.cpp:46-80); these run the same per-cache passes for a group
 * of independent caches in shared launches: every cache's result is the one
 * nrc_infer / nrc_train / nrc_frame would give it alone (bit-identical).
 * Caches in one call must be distinct and on the current device; they may
 * differ in hidden layers, encoding and loss (they are launched by kind).
 */
typedef struct nrc_group_item {
	nrc_cache *cache;
	const nrc_query *d_queries;          /* inference input, NULL with max_queries == 0 */
	const uint32_t *d_query_count;       /* device count or NULL (= max_queries) */
	uint32_t max_queries;
	float *d_rgb;                        /* [max_queries][3] */
	const nrc_train_record *d_batches;   /* NRC_TRAIN_BATCH_COUNT batches of max_batch_records */
	const uint32_t *d_batch_counts;      /* NRC_TRAIN_BATCH_COUNT device counts or NULL */
	uint32_t max_batch_records;          /* 0: no training for this cache */
} nrc_group_item;

/* nrc_infer for every item. */
nrc_status nrc_group_infer(const nrc_group_item *items, uint32_t n_items, nrc_path path, void *stream);
/* nrc_train on batch `batch` (0 .. NRC_TRAIN_BATCH_COUNT-1) of every item with training. */
nrc_status nrc_group_train(const nrc_group_item *items, uint32_t n_items, uint32_t batch, nrc_path path,
                           uint32_t flags, void *stream);
/* nrc_frame for every item: one grouped inference, then NRC_TRAIN_BATCH_COUNT grouped steps. */
nrc_status nrc_group_frame(const nrc_group_item *items, uint32_t n_items, nrc_path path, void *stream);

/*
 * A group frame captured once as a HIP graph and replayed per frame (device
 * counts may change between replays; buffers, max sizes and caches may not).
 * A frame takes an even number of steps, so every replay starts from the same
 * optimizer-state buffer; a replay after any other train call on one of the
 * caches is refused with NRC_ERROR_INVALID_ARGUMENT.
 */
typedef struct nrc_graph nrc_graph;
nrc_status nrc_group_frame_graph(const nrc_group_item *items, uint32_t n_items, nrc_path path, nrc_graph **out_graph);
nrc_status nrc_graph_launch(nrc_graph *graph, void *stream);
void nrc_graph_destroy(nrc_graph *graph);

#ifdef __cplusplus
}
#endif

#endif /* NRC_H */
