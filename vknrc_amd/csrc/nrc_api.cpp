// nrc_api.cpp -- the C ABI of include/nrc.h: the cache object (VkNRCState) and
// the NN passes (NNInference, NNTrain) as stream-ordered HIP launches.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "nrc_internal.hpp"

using nrc::DeviceState;

struct nrc_cache {
	nrc_cache_desc desc;
	int device = 0;
	uint32_t wc = 0;
	nrc_optimizer_entry *entries = nullptr;
	_Float16 *w16 = nullptr, *use16 = nullptr;
	_Float16 *wf = nullptr, *wtf = nullptr; // fragment-major copies for the fused gradient
	_Float16 *usef = nullptr;               // fragment-major use-weights for the fused inference
	uint32_t grad_chunks = 0;               // 64-record chunks per gradient workgroup (0: automatic)
	uint8_t *w8 = nullptr;                  // fragment-major e4m3 use-weights (FP8 inference)
	nrc::Fp8State *f8 = nullptr;            // FP8 scales and activation history
	float *w32 = nullptr, *use32 = nullptr;
	DeviceState *state = nullptr;
	float *partials = nullptr;
	uint32_t part_cap = 0, pstride = 0;
	int parity = 0; // state buffer the next train step reads (flipped per launched step)
	// plain-path scratch (allocated on the first plain train step)
	float *acts = nullptr, *deltas = nullptr, *sample_loss = nullptr;
	uint32_t ldn = 0;
};

namespace {

thread_local std::string g_error;

nrc_status fail(nrc_status s, const std::string &msg) {
	g_error = msg;
	return s;
}

nrc_status hip_fail(hipError_t e, const char *what) {
	return fail(e == hipErrorOutOfMemory ? NRC_ERROR_OUT_OF_MEMORY : NRC_ERROR_HIP,
	            std::string(what) + ": " + hipGetErrorString(e));
}

#define NRC_HIP(call, what)                                                                                            \
	do {                                                                                                               \
		hipError_t e_ = (call);                                                                                        \
		if (e_ != hipSuccess) return hip_fail(e_, what);                                                               \
	} while (0)

// A query or record count one call may take (NRC_MAX_QUERIES).
nrc_status check_count(uint32_t n) {
	return n > NRC_MAX_QUERIES ? fail(NRC_ERROR_CAPACITY, "count > NRC_MAX_QUERIES") : NRC_OK;
}

// SplitMix64 stream; uniform doubles in [0, 1) from the top 53 bits.
struct SplitMix64 {
	uint64_t s;
	uint64_t next() {
		uint64_t z = (s += 0x9E3779B97F4A7C15ull);
		z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
		z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
		return z ^ (z >> 31);
	}
	double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
	double normal() { // Box-Muller
		const double u1 = 1.0 - uniform(), u2 = uniform();
		return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586476925 * u2);
	}
};

bool valid_desc(const nrc_cache_desc *d, std::string &why) {
	if (!d) return why = "desc is NULL", false;
	if (d->hidden_layers != 4 && d->hidden_layers != 5) return why = "hidden_layers must be 4 or 5", false;
	if ((unsigned)d->encoding > NRC_ENCODING_UV_SINCOS16) return why = "unknown encoding", false;
	if ((unsigned)d->loss > NRC_LOSS_L2) return why = "unknown loss", false;
	if ((unsigned)d->init > NRC_INIT_XAVIER_UNIFORM) return why = "unknown init", false;
	if (d->train_batch_size == 0 || d->train_batch_size > NRC_MAX_TRAIN_BATCH_SIZE)
		return why = "train_batch_size must be in [1, NRC_MAX_TRAIN_BATCH_SIZE]", false;
	if (!(d->learning_rate > 0) || !(d->beta1 >= 0 && d->beta1 < 1) || !(d->beta2 >= 0 && d->beta2 < 1) ||
	    !(d->epsilon >= 0) || !(d->ema_alpha >= 0 && d->ema_alpha < 1))
		return why = "optimizer hyper-parameters out of range", false;
	if (!(d->loss_scale > 0) || !std::isfinite(d->loss_scale)) return why = "loss_scale must be positive", false;
	return true;
}

void host_init_weights(const nrc_cache_desc *d, uint64_t seed, float *w) {
	SplitMix64 rng{seed};
	const uint32_t H = d->hidden_layers, wc = nrc_weight_count(H);
	if (d->init == NRC_INIT_KAIMING_NORMAL) {
		// VkNRCState.cpp:39-44: N(0, sqrt(2 / 64)) for every weight, output layer included
		const double sigma = std::sqrt(2.0 / 64.0);
		for (uint32_t i = 0; i < wc; ++i) w[i] = (float)(sigma * rng.normal());
	} else {
		// Glorot uniform: U(-a, a), a = sqrt(6 / (fan_in + fan_out))
		for (uint32_t i = 0; i < wc; ++i) {
			const double fan_out = i < H * 4096u ? 64.0 : 3.0;
			const double a = std::sqrt(6.0 / (64.0 + fan_out));
			w[i] = (float)((2.0 * rng.uniform() - 1.0) * a);
		}
	}
}

// Uploads entries + the derived fp16/fp32 weight copies (ResetMLPBuffers, VkNRCState.cpp:46-88).
nrc_status upload_entries(nrc_cache *c, const nrc_optimizer_entry *entries) {
	const int H = (int)c->desc.hidden_layers;
	std::vector<_Float16> w16(c->wc), u16(c->wc);
	std::vector<_Float16> wf(nrc::fwd_frag_halves(H), (_Float16)0.0f), wtf(nrc::bwd_frag_halves(H), (_Float16)0.0f);
	std::vector<_Float16> usef(nrc::fwd_frag_halves(H), (_Float16)0.0f);
	std::vector<float> w32(c->wc), u32(c->wc);
	for (uint32_t i = 0; i < c->wc; ++i) {
		w32[i] = entries[i].weight;
		u32[i] = c->desc.use_ema_weights ? entries[i].ema_weight : entries[i].weight;
		w16[i] = (_Float16)w32[i];
		u16[i] = (_Float16)u32[i];
		nrc::frag_scatter(i, w16[i], wf.data(), wtf.data());
		nrc::frag_scatter(i, u16[i], usef.data(), (_Float16 *)nullptr);
	}
	NRC_HIP(hipSetDevice(c->device), "hipSetDevice");
	NRC_HIP(hipMemcpy(c->entries, entries, sizeof(nrc_optimizer_entry) * c->wc, hipMemcpyHostToDevice), "upload entries");
	NRC_HIP(hipMemcpy(c->w16, w16.data(), 2 * c->wc, hipMemcpyHostToDevice), "upload weights");
	NRC_HIP(hipMemcpy(c->wf, wf.data(), 2 * wf.size(), hipMemcpyHostToDevice), "upload weights");
	NRC_HIP(hipMemcpy(c->wtf, wtf.data(), 2 * wtf.size(), hipMemcpyHostToDevice), "upload weights");
	NRC_HIP(hipMemcpy(c->usef, usef.data(), 2 * usef.size(), hipMemcpyHostToDevice), "upload use weights");
	NRC_HIP(hipMemcpy(c->use16, u16.data(), 2 * c->wc, hipMemcpyHostToDevice), "upload use weights");
	NRC_HIP(hipMemcpy(c->w32, w32.data(), 4 * c->wc, hipMemcpyHostToDevice), "upload weights");
	NRC_HIP(hipMemcpy(c->use32, u32.data(), 4 * c->wc, hipMemcpyHostToDevice), "upload use weights");
	return NRC_OK;
}

nrc_status reset_cache(nrc_cache *c, uint64_t seed) {
	std::vector<float> w(c->wc);
	host_init_weights(&c->desc, seed, w.data());
	std::vector<nrc_optimizer_entry> e(c->wc);
	for (uint32_t i = 0; i < c->wc; ++i) e[i] = nrc_optimizer_entry{0.0f, 0.0f, w[i], w[i]};
	nrc_status s = upload_entries(c, e.data());
	if (s != NRC_OK) return s;
	NRC_HIP(nrc::launch_init_state(c->state, c->desc.loss_scale, nullptr), "init state");
	// FP8: no scales and no activation history yet; the padding of the fp8 copy stays zero
	NRC_HIP(hipMemset(c->f8, 0, sizeof(nrc::Fp8State)), "init fp8 state");
	NRC_HIP(hipMemset(c->w8, 0, (size_t)nrc::fwd_frag_halves((int)c->desc.hidden_layers)), "init fp8 weights");
	c->parity = 0;
	NRC_HIP(hipDeviceSynchronize(), "reset");
	c->desc.seed = seed;
	return NRC_OK;
}

void free_cache(nrc_cache *c) {
	if (!c) return;
	(void)hipSetDevice(c->device);
	void *ptrs[] = {c->entries, c->w16, c->wf, c->wtf, c->usef, c->w8, c->f8, c->use16, c->w32, c->use32, c->state, c->partials, c->acts, c->deltas, c->sample_loss};
	for (void *p : ptrs)
		if (p) (void)hipFree(p);
	delete c;
}

nrc::AdamParams adam_params(const nrc_cache *c) {
	nrc::AdamParams p;
	p.lr = c->desc.learning_rate;
	p.beta1 = c->desc.beta1;
	p.beta2 = c->desc.beta2;
	p.eps = c->desc.epsilon;
	p.ema_alpha = c->desc.ema_alpha;
	p.dynamic_scale = c->desc.dynamic_loss_scale;
	p.growth_interval = c->desc.loss_scale_growth_interval;
	p.use_ema = c->desc.use_ema_weights;
	return p;
}

// The plain path's scratch, allocated on its first use.  All three buffers are set
// together or not at all, so a failed attempt leaves the cache as it was.
nrc_status ensure_plain_scratch(nrc_cache *c) {
	if (c->acts) return NRC_OK;
	const uint32_t ldn = (c->desc.train_batch_size + 63u) & ~63u;
	const size_t slab = (size_t)(c->desc.hidden_layers + 1) * 64 * ldn;
	float *acts = nullptr, *deltas = nullptr, *loss = nullptr;
	hipError_t e = hipMalloc(&acts, slab * sizeof(float));
	if (e == hipSuccess) e = hipMalloc(&deltas, slab * sizeof(float));
	if (e == hipSuccess) e = hipMalloc(&loss, (size_t)ldn * sizeof(float));
	if (e != hipSuccess) {
		for (float *p : {acts, deltas, loss})
			if (p) (void)hipFree(p);
		return hip_fail(e, "plain scratch");
	}
	c->acts = acts;
	c->deltas = deltas;
	c->sample_loss = loss;
	c->ldn = ldn;
	return NRC_OK;
}


// Arguments of one training step on `c` (the plain-path scratch is added by the caller).
nrc::TrainArgs train_args(nrc_cache *c, const nrc_train_record *d_records, const uint32_t *d_count,
                          uint32_t max_count) {
	nrc::TrainArgs a{};
	a.hidden = (int)c->desc.hidden_layers;
	a.encoding = (int)c->desc.encoding;
	a.loss = (int)c->desc.loss;
	a.records = reinterpret_cast<const float *>(d_records);
	a.d_count = d_count;
	a.max_count = max_count;
	a.clamp = c->desc.train_batch_size;
	a.w32 = c->w32;
	a.wf = c->wf;
	a.wtf = c->wtf;
	a.state = c->state;
	a.parity = c->parity;
	a.partials = c->partials;
	a.pstride = c->pstride;
	return a;
}

nrc::AdamArgs adam_args(nrc_cache *c, const uint32_t *d_count, uint32_t max_count, uint32_t nparts, uint32_t flags) {
	nrc::AdamArgs o{};
	o.partials = c->partials;
	o.nparts = nparts;
	o.pstride = c->pstride;
	o.d_count = d_count;
	o.max_count = max_count;
	o.clamp = c->desc.train_batch_size;
	o.state = c->state;
	o.parity = c->parity;
	o.entries = c->entries;
	o.w16 = c->w16;
	o.wf = c->wf;
	o.wtf = c->wtf;
	o.use16 = c->use16;
	o.usef = c->usef;
	o.w32 = c->w32;
	o.use32 = c->use32;
	o.write_use = (flags & NRC_TRAIN_WRITE_USE_WEIGHTS) ? 1 : 0;
	o.weight_count = c->wc;
	o.hidden = (int)c->desc.hidden_layers;
	o.p = adam_params(c);
	return o;
}

// Argument checks of one training step, including the partial rows its plan
// writes, before anything is launched.
nrc_status check_train(nrc_cache *c, const nrc_train_record *d_records, uint32_t max_count, nrc_path path) {
	if (!d_records) return fail(NRC_ERROR_INVALID_ARGUMENT, "NULL records");
	if (max_count > NRC_MAX_TRAIN_BATCH_SIZE) return fail(NRC_ERROR_CAPACITY, "max_count > NRC_MAX_TRAIN_BATCH_SIZE");
	if (path != NRC_PATH_PLAIN_F32 && path != NRC_PATH_FUSED_F16 && path != NRC_PATH_FUSED_F8)
		return fail(NRC_ERROR_UNSUPPORTED, "unknown path");
	const uint32_t clamp = c->desc.train_batch_size;
	const uint32_t n = max_count < clamp ? max_count : clamp;
	const uint32_t need = path == NRC_PATH_PLAIN_F32 ? (n + nrc::kPlainChunk - 1) / nrc::kPlainChunk
	                                                 : nrc::plan_fused_gradient(n, c->grad_chunks).blocks;
	if (need > c->part_cap) return fail(NRC_ERROR_CAPACITY, "partial buffer too small");
	return NRC_OK;
}


// One training step on stream s (arguments already checked).
nrc_status train_step(nrc_cache *c, const nrc_train_record *d_records, const uint32_t *d_count, uint32_t max_count,
                      nrc_path path, uint32_t flags, hipStream_t s) {
	nrc::TrainArgs a = train_args(c, d_records, d_count, max_count);
	const uint32_t n = max_count < a.clamp ? max_count : a.clamp;
	uint32_t nparts = 0;
	hipError_t e;
	if (path == NRC_PATH_PLAIN_F32) {
		nrc_status st = ensure_plain_scratch(c);
		if (st != NRC_OK) return st;
		a.acts = c->acts;
		a.deltas = c->deltas;
		a.sample_loss = c->sample_loss;
		a.ldn = c->ldn;
		e = nrc::launch_plain_gradient(a, &nparts, s);
	} else {
		const nrc::FusedGradPlan plan = nrc::plan_fused_gradient(n, c->grad_chunks);
		e = nrc::launch_fused_gradient(a, plan, s);
		nparts = plan.blocks;
	}
	if (e != hipSuccess) return hip_fail(e, "nrc_train gradient launch");
	e = nrc::launch_adam(adam_args(c, d_count, max_count, nparts, flags), s);
	if (e != hipSuccess) return hip_fail(e, "nrc_train optimizer launch");
	c->parity ^= 1; // the optimizer wrote the other state buffer
	return NRC_OK;
}

} // namespace

extern "C" {

const char *nrc_status_string(nrc_status s) {
	switch (s) {
	case NRC_OK: return "ok";
	case NRC_ERROR_INVALID_ARGUMENT: return "invalid argument";
	case NRC_ERROR_UNSUPPORTED: return "unsupported";
	case NRC_ERROR_OUT_OF_MEMORY: return "out of memory";
	case NRC_ERROR_HIP: return "HIP error";
	case NRC_ERROR_CAPACITY: return "capacity exceeded";
	}
	return "unknown status";
}

const char *nrc_last_error(void) { return g_error.c_str(); }

void nrc_cache_desc_default(nrc_cache_desc *d) {
	if (!d) return;
	std::memset(d, 0, sizeof(*d));
	d->hidden_layers = 5;
	d->encoding = NRC_ENCODING_TRI12;
	d->loss = NRC_LOSS_RELATIVE_L2_LUMINANCE;
	d->init = NRC_INIT_KAIMING_NORMAL;
	d->seed = 0;
	d->train_batch_size = NRC_DEFAULT_TRAIN_BATCH_SIZE;
	d->learning_rate = 0.002f;
	d->beta1 = 0.9f;
	d->beta2 = 0.999f;
	d->epsilon = 1e-8f;
	d->ema_alpha = 0.99f;
	d->loss_scale = 1.0f;
	d->dynamic_loss_scale = 0;
	d->loss_scale_growth_interval = 2000;
	d->use_ema_weights = 0;
}

uint32_t nrc_weight_count(uint32_t hidden_layers) {
	return (hidden_layers == 4 || hidden_layers == 5) ? (uint32_t)nrc::weight_count((int)hidden_layers) : 0u;
}

nrc_status nrc_init_weights(const nrc_cache_desc *desc, uint64_t seed, float *host_weights) {
	std::string why;
	if (!valid_desc(desc, why)) return fail(NRC_ERROR_INVALID_ARGUMENT, why);
	if (!host_weights) return fail(NRC_ERROR_INVALID_ARGUMENT, "host_weights is NULL");
	host_init_weights(desc, seed, host_weights);
	return NRC_OK;
}

} // extern "C"

namespace {
// Device bytes of each allocation of a cache with this description (one table
// for nrc_cache_create and nrc_cache_footprint).
struct CacheSizes {
	size_t entries, w16, wf, wtf, usef, w8, f8, use16, w32, use32, state, partials, plain_scratch;
	explicit CacheSizes(const nrc_cache_desc &d) {
		const size_t wc = nrc_weight_count(d.hidden_layers);
		const int H = (int)d.hidden_layers;
		const size_t pstride = ((wc + 1) + 63u) & ~size_t(63);
		const size_t part_cap = (d.train_batch_size + nrc::kFusedSmallRecords - 1) / nrc::kFusedSmallRecords;
		entries = sizeof(nrc_optimizer_entry) * wc;
		w16 = use16 = 2 * wc;
		wf = usef = 2 * (size_t)nrc::fwd_frag_halves(H);
		wtf = 2 * (size_t)nrc::bwd_frag_halves(H);
		w8 = (size_t)nrc::fwd_frag_halves(H);
		f8 = sizeof(nrc::Fp8State);
		w32 = use32 = 4 * wc;
		state = sizeof(DeviceState);
		partials = sizeof(float) * pstride * part_cap;
		const size_t ldn = (d.train_batch_size + 63u) & ~size_t(63);
		plain_scratch = (2 * (size_t)(d.hidden_layers + 1) * 64 * ldn + ldn) * sizeof(float);
	}
	// what the allocator hands out: each allocation rounded up to 4 KiB
	static size_t page(size_t b) { return (b + 4095) & ~size_t(4095); }
	uint64_t total(bool plain) const {
		const size_t all[] = {entries, w16, wf, wtf, usef, w8, f8, use16, w32, use32, state, partials};
		uint64_t t = 0;
		for (size_t b : all) t += page(b);
		if (plain) t += 2 * page(plain_scratch / 2);
		return t;
	}
};
} // namespace

extern "C" {

nrc_status nrc_cache_footprint(const nrc_cache_desc *desc, nrc_path path, uint64_t *device_bytes) {
	std::string why;
	if (!device_bytes) return fail(NRC_ERROR_INVALID_ARGUMENT, "device_bytes is NULL");
	if (!valid_desc(desc, why)) return fail(NRC_ERROR_INVALID_ARGUMENT, why);
	*device_bytes = CacheSizes(*desc).total(path == NRC_PATH_PLAIN_F32);
	return NRC_OK;
}

nrc_status nrc_cache_create(const nrc_cache_desc *desc, nrc_cache **out) {
	std::string why;
	if (!out) return fail(NRC_ERROR_INVALID_ARGUMENT, "out_cache is NULL");
	*out = nullptr;
	if (!valid_desc(desc, why)) return fail(NRC_ERROR_INVALID_ARGUMENT, why);
	nrc_cache *c = new nrc_cache();
	c->desc = *desc;
	c->wc = nrc_weight_count(desc->hidden_layers);
	c->pstride = ((c->wc + 1) + 63u) & ~63u;
	// one partial row per 64 records covers every fused-gradient plan (and the plain path's 512)
	c->part_cap = (desc->train_batch_size + nrc::kFusedSmallRecords - 1) / nrc::kFusedSmallRecords;
	const CacheSizes sz(*desc);
	hipError_t e = hipGetDevice(&c->device);
	if (e == hipSuccess) e = hipMalloc(&c->entries, sz.entries);
	if (e == hipSuccess) e = hipMalloc(&c->w16, sz.w16);
	if (e == hipSuccess) e = hipMalloc(&c->wf, sz.wf);
	if (e == hipSuccess) e = hipMalloc(&c->wtf, sz.wtf);
	if (e == hipSuccess) e = hipMalloc(&c->usef, sz.usef);
	if (e == hipSuccess) e = hipMalloc(&c->w8, sz.w8);
	if (e == hipSuccess) e = hipMalloc(&c->f8, sz.f8);
	if (e == hipSuccess) e = hipMalloc(&c->use16, sz.use16);
	if (e == hipSuccess) e = hipMalloc(&c->w32, sz.w32);
	if (e == hipSuccess) e = hipMalloc(&c->use32, sz.use32);
	if (e == hipSuccess) e = hipMalloc(&c->state, sz.state);
	if (e == hipSuccess) e = hipMalloc(&c->partials, sz.partials);

	if (e != hipSuccess) {
		free_cache(c);
		return hip_fail(e, "nrc_cache_create");
	}
	nrc_status s = reset_cache(c, desc->seed);
	if (s == NRC_OK) {
		e = nrc::prepare_fused((int)desc->hidden_layers, (int)desc->encoding, (int)desc->loss);
		if (e != hipSuccess) s = hip_fail(e, "nrc_cache_create kernel set-up");
	}
	if (s != NRC_OK) {
		free_cache(c);
		return s;
	}
	*out = c;
	return NRC_OK;
}

nrc_status nrc_cache_reset(nrc_cache *c, uint64_t seed) {
	if (!c) return fail(NRC_ERROR_INVALID_ARGUMENT, "cache is NULL");
	return reset_cache(c, seed);
}

void nrc_cache_destroy(nrc_cache *c) { free_cache(c); }

nrc_status nrc_cache_get_desc(const nrc_cache *c, nrc_cache_desc *d) {
	if (!c || !d) return fail(NRC_ERROR_INVALID_ARGUMENT, "NULL argument");
	*d = c->desc;
	return NRC_OK;
}

nrc_status nrc_set_use_ema(nrc_cache *c, int32_t use_ema) {
	if (!c) return fail(NRC_ERROR_INVALID_ARGUMENT, "cache is NULL");
	c->desc.use_ema_weights = use_ema ? 1 : 0;
	return NRC_OK;
}

nrc_status nrc_infer(nrc_cache *c, const nrc_query *d_queries, const uint32_t *d_count, uint32_t max_count,
                     float *d_rgb, nrc_path path, void *stream) {
	if (!c) return fail(NRC_ERROR_INVALID_ARGUMENT, "cache is NULL");
	if (max_count == 0) return NRC_OK;
	if (!d_queries || !d_rgb) return fail(NRC_ERROR_INVALID_ARGUMENT, "NULL buffer");
	if (check_count(max_count) != NRC_OK) return NRC_ERROR_CAPACITY;
	const hipStream_t s = reinterpret_cast<hipStream_t>(stream);
	const int H = (int)c->desc.hidden_layers, enc = (int)c->desc.encoding;
	const float *q = reinterpret_cast<const float *>(d_queries);
	hipError_t e;
	switch (path) {
	case NRC_PATH_PLAIN_F32: e = nrc::launch_plain_infer(H, enc, q, d_count, max_count, c->use32, d_rgb, s); break;
	case NRC_PATH_FUSED_F16:
	case NRC_PATH_FUSED_F8: {
		nrc::InferGroup grp{};
		grp.n_items = 1;
		grp.clamp0 = 1;
		grp.item[0] = nrc::InferItem{q, d_count, d_rgb, c->usef, c->w8, c->f8, max_count, 0};
		if (path == NRC_PATH_FUSED_F16) {
			e = nrc::launch_fused_infer_group(H, enc, grp, s);
		} else {
			const float *u32[1] = {c->use32};
			uint8_t *w8[1] = {c->w8};
			e = nrc::launch_fp8_infer_group(H, enc, grp, u32, w8, s);
		}
		break;
	}
	default: return fail(NRC_ERROR_UNSUPPORTED, "unknown path");
	}
	if (e != hipSuccess) return hip_fail(e, "nrc_infer launch");
	return NRC_OK;
}

nrc_status nrc_cache_set_grad_chunks(nrc_cache *c, uint32_t chunks) {
	if (!c) return fail(NRC_ERROR_INVALID_ARGUMENT, "cache is NULL");
	if (chunks > (uint32_t)nrc::kFusedMaxChunks) return fail(NRC_ERROR_INVALID_ARGUMENT, "chunks > 16");
	c->grad_chunks = chunks;
	return NRC_OK;
}

nrc_status nrc_set_infer_workgroups_per_cu(uint32_t workgroups) {
	if (workgroups < 1 || workgroups > 8) return fail(NRC_ERROR_INVALID_ARGUMENT, "workgroups per CU outside 1..8");
	nrc::set_infer_wg_per_cu(workgroups);
	return NRC_OK;
}

nrc_status nrc_get_fp8_scales(nrc_cache *c, float *layer_scales, float *out_scale, uint32_t *frames) {
	if (!c) return fail(NRC_ERROR_INVALID_ARGUMENT, "cache is NULL");
	nrc::Fp8State st;
	NRC_HIP(hipSetDevice(c->device), "hipSetDevice");
	NRC_HIP(hipDeviceSynchronize(), "nrc_get_fp8_scales");
	NRC_HIP(hipMemcpy(&st, c->f8, sizeof(st), hipMemcpyDeviceToHost), "download fp8 state");
	if (layer_scales)
		for (uint32_t l = 0; l <= c->desc.hidden_layers; ++l) layer_scales[l] = st.scale[l];
	if (out_scale) *out_scale = st.out_scale;
	if (frames) *frames = st.frames;
	return NRC_OK;
}

nrc_status nrc_train(nrc_cache *c, const nrc_train_record *d_records, const uint32_t *d_count, uint32_t max_count,
                     nrc_path path, uint32_t flags, void *stream) {
	if (!c) return fail(NRC_ERROR_INVALID_ARGUMENT, "cache is NULL");
	if (max_count == 0) return NRC_OK;
	nrc_status st = check_train(c, d_records, max_count, path);
	if (st != NRC_OK) return st;
	return train_step(c, d_records, d_count, max_count, path, flags, reinterpret_cast<hipStream_t>(stream));
}

nrc_status nrc_frame(nrc_cache *c, const nrc_query *d_queries, const uint32_t *d_query_count, uint32_t max_queries,
                     float *d_rgb, const nrc_train_record *d_batches, const uint32_t *d_batch_counts,
                     uint32_t max_batch_records, nrc_path path, void *stream) {
	if (!c) return fail(NRC_ERROR_INVALID_ARGUMENT, "cache is NULL");
	if (max_queries && (!d_queries || !d_rgb)) return fail(NRC_ERROR_INVALID_ARGUMENT, "NULL buffer");
	if (check_count(max_queries) != NRC_OK) return NRC_ERROR_CAPACITY;
	nrc_status st = max_batch_records ? check_train(c, d_batches, max_batch_records, path) : NRC_OK;
	if (st != NRC_OK) return st;
	const hipStream_t s = reinterpret_cast<hipStream_t>(stream);
	auto batch = [&](uint32_t b) { return d_batches + (size_t)b * max_batch_records; };
	auto count = [&](uint32_t b) { return d_batch_counts ? d_batch_counts + b : nullptr; };
	auto flags = [](uint32_t b) { return b == NRC_TRAIN_BATCH_COUNT - 1 ? NRC_TRAIN_WRITE_USE_WEIGHTS : 0u; };
	st = nrc_infer(c, d_queries, d_query_count, max_queries, d_rgb, path, stream);
	for (uint32_t b = 0; st == NRC_OK && max_batch_records && b < NRC_TRAIN_BATCH_COUNT; ++b)
		st = train_step(c, batch(b), count(b), max_batch_records, path, flags(b), s);
	return st;
}

nrc_status nrc_get_state(nrc_cache *c, nrc_optimizer_entry *entries, nrc_optimizer_state *state, uint16_t *w16,
                         uint16_t *use16) {
	if (!c) return fail(NRC_ERROR_INVALID_ARGUMENT, "cache is NULL");
	NRC_HIP(hipSetDevice(c->device), "hipSetDevice");
	NRC_HIP(hipDeviceSynchronize(), "nrc_get_state");
	if (entries)
		NRC_HIP(hipMemcpy(entries, c->entries, sizeof(nrc_optimizer_entry) * c->wc, hipMemcpyDeviceToHost), "get entries");
	if (state)
		NRC_HIP(hipMemcpy(state, &c->state->pub[c->parity], sizeof(nrc_optimizer_state), hipMemcpyDeviceToHost),
		        "get state");
	if (w16) NRC_HIP(hipMemcpy(w16, c->w16, 2 * c->wc, hipMemcpyDeviceToHost), "get weights");
	if (use16) NRC_HIP(hipMemcpy(use16, c->use16, 2 * c->wc, hipMemcpyDeviceToHost), "get use weights");
	return NRC_OK;
}

nrc_status nrc_set_state(nrc_cache *c, const nrc_optimizer_entry *entries, const nrc_optimizer_state *state) {
	if (!c) return fail(NRC_ERROR_INVALID_ARGUMENT, "cache is NULL");
	NRC_HIP(hipSetDevice(c->device), "hipSetDevice");
	NRC_HIP(hipDeviceSynchronize(), "nrc_set_state");
	if (entries) {
		nrc_status s = upload_entries(c, entries);
		if (s != NRC_OK) return s;
	}
	if (state) {
		DeviceState ds{};
		ds.pub[0] = *state;
		ds.pub[1] = *state;
		NRC_HIP(hipMemcpy(c->state, &ds, sizeof(ds), hipMemcpyHostToDevice), "set state");
	}
	NRC_HIP(hipDeviceSynchronize(), "nrc_set_state");
	return NRC_OK;
}

nrc_status nrc_encode(nrc_encoding encoding, const nrc_query *d_queries, uint32_t n, float *d_features, void *stream) {
	if ((unsigned)encoding > NRC_ENCODING_UV_SINCOS16) return fail(NRC_ERROR_INVALID_ARGUMENT, "unknown encoding");
	if (n == 0) return NRC_OK;
	if (!d_queries || !d_features) return fail(NRC_ERROR_INVALID_ARGUMENT, "NULL buffer");
	if (check_count(n) != NRC_OK) return NRC_ERROR_CAPACITY;
	hipError_t e = nrc::launch_encode((int)encoding, reinterpret_cast<const float *>(d_queries), n, d_features,
	                                  reinterpret_cast<hipStream_t>(stream));
	if (e != hipSuccess) return hip_fail(e, "nrc_encode");
	return NRC_OK;
}

nrc_status nrc_sph_encode(const float *d_dirs, uint32_t n, float *d_uv, void *stream) {
	if (n == 0) return NRC_OK;
	if (!d_dirs || !d_uv) return fail(NRC_ERROR_INVALID_ARGUMENT, "NULL buffer");
	if (check_count(n) != NRC_OK) return NRC_ERROR_CAPACITY;
	hipError_t e = nrc::launch_sph_encode(d_dirs, n, d_uv, reinterpret_cast<hipStream_t>(stream));
	if (e != hipSuccess) return hip_fail(e, "nrc_sph_encode");
	return NRC_OK;
}

uint32_t nrc_dst_screen(uint32_t x, uint32_t y) { return ((x & 0x7FFFu) | ((y & 0x7FFFu) << 15)) << 1; }

uint32_t nrc_dst_train(uint32_t b, uint32_t l, uint32_t r) {
	return (((b & 3u) | ((l & 0x3FFFu) << 2) | ((r & 0x3FFFu) << 16)) << 1) | 1u;
}

nrc_status nrc_composite(const float *d_rgb, const uint32_t *d_dst, const uint32_t *d_count, uint32_t max_count,
                         const nrc_screen *screen, nrc_train_record *d_batches, const float *d_train_factors,
                         uint32_t batch_stride, void *stream) {
	if (max_count == 0) return NRC_OK;
	if (!d_rgb || !d_dst) return fail(NRC_ERROR_INVALID_ARGUMENT, "NULL buffer");
	if (check_count(max_count) != NRC_OK) return NRC_ERROR_CAPACITY;
	if (screen && (!screen->d_bias_factor_r || !screen->d_factor_gb || screen->pitch < screen->width))
		return fail(NRC_ERROR_INVALID_ARGUMENT, "screen images NULL or pitch < width");
	if (d_batches && (!d_train_factors || batch_stride == 0))
		return fail(NRC_ERROR_INVALID_ARGUMENT, "train factors NULL or batch_stride 0");
	nrc::CompositeArgs a{};
	a.rgb = d_rgb;
	a.dst = d_dst;
	a.d_count = d_count;
	a.max_count = max_count;
	if (screen) {
		a.bias_factor_r = screen->d_bias_factor_r;
		a.factor_gb = screen->d_factor_gb;
		a.width = screen->width;
		a.height = screen->height;
		a.pitch = screen->pitch;
	}
	a.batches = reinterpret_cast<float *>(d_batches);
	a.factors = d_train_factors;
	a.batch_stride = batch_stride;
	hipError_t e = nrc::launch_composite(a, reinterpret_cast<hipStream_t>(stream));
	if (e != hipSuccess) return hip_fail(e, "nrc_composite");
	return NRC_OK;
}

} // extern "C"

namespace {
static_assert(sizeof(nrc_scene_primitive) == 64, "a packed primitive is one 64-byte record");

nrc_status check_scene(const nrc_scene *s) {
	if (!s) return fail(NRC_ERROR_INVALID_ARGUMENT, "scene is NULL");
	if (!s->d_vertices || !s->d_vertex_indices || !s->d_texcoords || !s->d_texcoord_indices || !s->d_materials ||
	    !s->d_material_ids || !s->d_transforms)
		return fail(NRC_ERROR_INVALID_ARGUMENT, "scene buffer NULL");
	if (s->texture_count && !s->d_textures) return fail(NRC_ERROR_INVALID_ARGUMENT, "texture table NULL");
	if (reinterpret_cast<uintptr_t>(s->d_primitives) % 16u != 0u)
		return fail(NRC_ERROR_INVALID_ARGUMENT, "packed primitives not 16-byte aligned");
	return NRC_OK;
}
} // namespace

extern "C" {

nrc_status nrc_scene_pack_primitives(const nrc_scene *scene, uint32_t primitive_count, nrc_scene_primitive *d_out,
                                     void *stream) {
	if (primitive_count == 0) return NRC_OK;
	nrc_status st = check_scene(scene);
	if (st != NRC_OK) return st;
	if (!d_out) return fail(NRC_ERROR_INVALID_ARGUMENT, "NULL buffer");
	if (reinterpret_cast<uintptr_t>(d_out) % 64u != 0u) return fail(NRC_ERROR_INVALID_ARGUMENT, "d_out not 64-byte aligned");
	hipError_t e = nrc::launch_pack_primitives(*scene, primitive_count, d_out, reinterpret_cast<hipStream_t>(stream));
	if (e != hipSuccess) return hip_fail(e, "nrc_scene_pack_primitives");
	return NRC_OK;
}

nrc_status nrc_unpack_eval_records(const nrc_scene *scene, const nrc_eval_record *d_records, const uint32_t *d_count,
                                   uint32_t max_count, nrc_query *d_queries, void *stream) {
	if (max_count == 0) return NRC_OK;
	nrc_status st = check_scene(scene);
	if (st != NRC_OK) return st;
	if (!d_records || !d_queries) return fail(NRC_ERROR_INVALID_ARGUMENT, "NULL buffer");
	if (check_count(max_count) != NRC_OK) return NRC_ERROR_CAPACITY;
	hipError_t e = nrc::launch_unpack_eval(*scene, d_records, d_count, max_count, reinterpret_cast<float *>(d_queries),
	                                       reinterpret_cast<hipStream_t>(stream));
	if (e != hipSuccess) return hip_fail(e, "nrc_unpack_eval_records");
	return NRC_OK;
}

nrc_status nrc_unpack_train_records(const nrc_scene *scene, const nrc_wire_train_record *d_records,
                                    const uint32_t *d_count, uint32_t max_count, nrc_train_record *d_out,
                                    void *stream) {
	if (max_count == 0) return NRC_OK;
	nrc_status st = check_scene(scene);
	if (st != NRC_OK) return st;
	if (!d_records || !d_out) return fail(NRC_ERROR_INVALID_ARGUMENT, "NULL buffer");
	if (check_count(max_count) != NRC_OK) return NRC_ERROR_CAPACITY;
	hipError_t e = nrc::launch_unpack_train(*scene, d_records, d_count, max_count, reinterpret_cast<float *>(d_out),
	                                        reinterpret_cast<hipStream_t>(stream));
	if (e != hipSuccess) return hip_fail(e, "nrc_unpack_train_records");
	return NRC_OK;
}

nrc_status nrc_composite_records(const float *d_rgb, const nrc_eval_record *d_records, const uint32_t *d_count,
                                 uint32_t max_count, const nrc_screen *screen, nrc_wire_train_record *d_batches,
                                 uint32_t batch_stride, void *stream) {
	if (max_count == 0) return NRC_OK;
	if (!d_rgb || !d_records) return fail(NRC_ERROR_INVALID_ARGUMENT, "NULL buffer");
	if (check_count(max_count) != NRC_OK) return NRC_ERROR_CAPACITY;
	if (screen && (!screen->d_bias_factor_r || !screen->d_factor_gb || screen->pitch < screen->width))
		return fail(NRC_ERROR_INVALID_ARGUMENT, "screen images NULL or pitch < width");
	if (d_batches && batch_stride == 0) return fail(NRC_ERROR_INVALID_ARGUMENT, "batch_stride 0");
	hipError_t e = nrc::launch_composite_records(
	    d_rgb, d_records, d_count, max_count, screen ? screen->d_bias_factor_r : nullptr,
	    screen ? screen->d_factor_gb : nullptr, screen ? screen->width : 0u, screen ? screen->height : 0u,
	    screen ? screen->pitch : 0u, d_batches, batch_stride, reinterpret_cast<hipStream_t>(stream));
	if (e != hipSuccess) return hip_fail(e, "nrc_composite_records");
	return NRC_OK;
}

nrc_status nrc_build_train_records(const nrc_traced_path *d_paths, uint32_t n_paths, nrc_wire_train_record *d_batches,
                                   uint32_t *d_batch_counts, uint32_t batch_capacity, nrc_eval_record *d_eval_records,
                                   uint32_t *d_eval_count, uint32_t eval_capacity, void *stream) {
	if (n_paths == 0) return NRC_OK;
	if (!d_paths || !d_batches || !d_batch_counts || !d_eval_records || !d_eval_count)
		return fail(NRC_ERROR_INVALID_ARGUMENT, "NULL buffer");
	// the train code holds record indices in 14 bits (EncodeNRCEvalDstTrain, NRCRecord.glsl:20-22)
	if (batch_capacity == 0 || batch_capacity > (1u << 14))
		return fail(NRC_ERROR_INVALID_ARGUMENT, "batch_capacity must be in [1, 16384]");
	if (check_count(n_paths) != NRC_OK || check_count(eval_capacity) != NRC_OK) return NRC_ERROR_CAPACITY;
	hipError_t e = nrc::launch_build_train_records(d_paths, n_paths, d_batches, d_batch_counts, batch_capacity,
	                                               d_eval_records, d_eval_count, eval_capacity,
	                                               reinterpret_cast<hipStream_t>(stream));
	if (e != hipSuccess) return hip_fail(e, "nrc_build_train_records");
	return NRC_OK;
}

nrc_status nrc_screen_tonemap(const float *d_color, float *d_accumulate, uint32_t width, uint32_t height,
                              uint32_t pitch, int32_t accumulate, uint32_t accumulate_count, float *d_out,
                              void *stream) {
	if (width == 0 || height == 0) return NRC_OK;
	if (!d_color || !d_out || (accumulate && !d_accumulate)) return fail(NRC_ERROR_INVALID_ARGUMENT, "NULL buffer");
	if (pitch < width) return fail(NRC_ERROR_INVALID_ARGUMENT, "pitch < width");
	hipError_t e = nrc::launch_screen(d_color, d_accumulate, width, height, pitch, accumulate ? 1 : 0, accumulate_count,
	                                  d_out, reinterpret_cast<hipStream_t>(stream));
	if (e != hipSuccess) return hip_fail(e, "nrc_screen_tonemap");
	return NRC_OK;
}

nrc_status nrc_mlp_forward_f16(uint32_t hidden, const uint16_t *d_w, const uint16_t *d_in, uint32_t n, float *d_out,
                               void *stream) {
	if (hidden != 4 && hidden != 5) return fail(NRC_ERROR_INVALID_ARGUMENT, "hidden_layers must be 4 or 5");
	if (n == 0) return NRC_OK;
	if (!d_w || !d_in || !d_out) return fail(NRC_ERROR_INVALID_ARGUMENT, "NULL buffer");
	if (check_count(n) != NRC_OK) return NRC_ERROR_CAPACITY;
	const hipStream_t s = reinterpret_cast<hipStream_t>(stream);
	const size_t nf = nrc::fwd_frag_halves((int)hidden);
	_Float16 *wf = nullptr;
	NRC_HIP(hipMallocAsync(reinterpret_cast<void **>(&wf), 2 * nf, s), "forward scratch");
	hipError_t e = hipMemsetAsync(wf, 0, 2 * nf, s);
	if (e == hipSuccess)
		e = nrc::launch_make_frags((int)hidden, reinterpret_cast<const _Float16 *>(d_w), wf, nullptr, s);
	if (e == hipSuccess)
		e = nrc::launch_fused_forward_f16((int)hidden, reinterpret_cast<const _Float16 *>(d_in), n, wf, d_out, s);
	hipError_t e2 = hipFreeAsync(wf, s);
	if (e != hipSuccess) return hip_fail(e, "nrc_mlp_forward_f16");
	if (e2 != hipSuccess) return hip_fail(e2, "nrc_mlp_forward_f16 free");
	return NRC_OK;
}

nrc_status nrc_mlp_gradient_f16(uint32_t hidden, const uint16_t *d_w, const uint16_t *d_in, const float *d_targets,
                                uint32_t n, float *d_dw, void *stream) {
	if (hidden != 4 && hidden != 5) return fail(NRC_ERROR_INVALID_ARGUMENT, "hidden_layers must be 4 or 5");
	if (!d_w || !d_in || !d_targets || !d_dw) return fail(NRC_ERROR_INVALID_ARGUMENT, "NULL buffer");
	if (n > NRC_MAX_TRAIN_BATCH_SIZE) return fail(NRC_ERROR_CAPACITY, "n > NRC_MAX_TRAIN_BATCH_SIZE");
	const hipStream_t s = reinterpret_cast<hipStream_t>(stream);
	const uint32_t wc = nrc_weight_count(hidden);
	const uint32_t pstride = ((wc + 1) + 63u) & ~63u;
	// enough partial rows for any plan: one per 64 records
	const uint32_t nb = n == 0 ? 1 : (n + nrc::kFusedSmallRecords - 1) / nrc::kFusedSmallRecords;
	const size_t nf = nrc::fwd_frag_halves((int)hidden), nt = nrc::bwd_frag_halves((int)hidden);
	float *partials = nullptr;
	_Float16 *wf = nullptr, *wtf = nullptr;
	hipError_t e = hipMallocAsync(reinterpret_cast<void **>(&partials), sizeof(float) * (size_t)pstride * nb, s);
	if (e == hipSuccess) e = hipMallocAsync(reinterpret_cast<void **>(&wf), 2 * nf, s);
	if (e == hipSuccess) e = hipMallocAsync(reinterpret_cast<void **>(&wtf), 2 * nt, s);
	uint32_t nparts = 0;
	const _Float16 *w = reinterpret_cast<const _Float16 *>(d_w);
	if (e == hipSuccess) e = hipMemsetAsync(wf, 0, 2 * nf, s);
	if (e == hipSuccess) e = hipMemsetAsync(wtf, 0, 2 * nt, s);
	if (e == hipSuccess) e = nrc::launch_make_frags((int)hidden, w, wf, wtf, s);
	if (e == hipSuccess)
		e = nrc::launch_fused_gradient_f16((int)hidden, reinterpret_cast<const _Float16 *>(d_in), d_targets, n, wf, wtf,
		                                   partials, pstride, &nparts, s);
	if (e == hipSuccess) e = nrc::launch_reduce_partials(partials, nparts, pstride, wc, d_dw, s);
	// every scratch buffer that was allocated is freed, whatever failed
	hipError_t e2 = hipSuccess;
	for (void *p : {static_cast<void *>(partials), static_cast<void *>(wf), static_cast<void *>(wtf)})
		if (p) {
			const hipError_t f = hipFreeAsync(p, s);
			if (e2 == hipSuccess) e2 = f;
		}
	if (e != hipSuccess) return hip_fail(e, "nrc_mlp_gradient_f16");
	if (e2 != hipSuccess) return hip_fail(e2, "nrc_mlp_gradient_f16 free");
	return NRC_OK;
}

} // extern "C"

// ------------------------------------------------- gradient without allocation
namespace {

// Device scratch of nrc_mlp_gradient_f16_noalloc: the fragment-major weight
// copies (forward, then backward) and the partial-gradient rows of a batch of up
// to NRC_NOALLOC_MAX_GRADIENT_RECORDS records, one row per 64 records (the most
// rows any gradient plan writes), for the largest network (5 hidden layers).
constexpr size_t kNoallocWfBytes = 2 * (size_t)nrc::fwd_frag_halves(NRC_MAX_HIDDEN_LAYERS);
constexpr size_t kNoallocWtfBytes = 2 * (size_t)nrc::bwd_frag_halves(NRC_MAX_HIDDEN_LAYERS);
constexpr size_t kNoallocPstride = ((size_t)nrc::weight_count(NRC_MAX_HIDDEN_LAYERS) + 1 + 63) & ~size_t(63);
constexpr size_t kNoallocRows = NRC_NOALLOC_MAX_GRADIENT_RECORDS / nrc::kFusedSmallRecords;
constexpr size_t kNoallocBytes = kNoallocWfBytes + kNoallocWtfBytes + sizeof(float) * kNoallocPstride * kNoallocRows;
static_assert(kNoallocWfBytes % 256 == 0 && kNoallocWtfBytes % 256 == 0, "scratch parts stay 256-byte aligned");

// One copy per GPU, held until the process ends.  Host code never uses this name
// as an address: the device address of each GPU's copy comes from
// hipGetSymbolAddress (noalloc_scratch below).
static __device__ __align__(256) unsigned char scratch[kNoallocBytes];

constexpr int kNoallocMaxDevices = 64;
std::mutex g_noalloc_mu;
void *g_noalloc_ptr[kNoallocMaxDevices] = {};

// The device address of this GPU's scratch, looked up once per device with the
// device current (the first call on a device; later calls read the table).
nrc_status noalloc_scratch(unsigned char **out) {
	int dev = 0;
	NRC_HIP(hipGetDevice(&dev), "hipGetDevice");
	if (dev < 0 || dev >= kNoallocMaxDevices) return fail(NRC_ERROR_UNSUPPORTED, "device index beyond the scratch table");
	std::lock_guard<std::mutex> lock(g_noalloc_mu);
	if (!g_noalloc_ptr[dev]) {
		void *p = nullptr;
		NRC_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
		if (!p) return fail(NRC_ERROR_HIP, "hipGetSymbolAddress(scratch) returned NULL");
		g_noalloc_ptr[dev] = p;
	}
	*out = static_cast<unsigned char *>(g_noalloc_ptr[dev]);
	return NRC_OK;
}

} // namespace

extern "C" {

nrc_status nrc_mlp_gradient_f16_noalloc(uint32_t hidden, const uint16_t *d_w, const uint16_t *d_in,
                                        const float *d_targets, uint32_t n, float *d_dw, void *stream) {
	if (hidden != 4 && hidden != 5) return fail(NRC_ERROR_INVALID_ARGUMENT, "hidden_layers must be 4 or 5");
	if (!d_w || !d_in || !d_targets || !d_dw) return fail(NRC_ERROR_INVALID_ARGUMENT, "NULL buffer");
	if (n > NRC_NOALLOC_MAX_GRADIENT_RECORDS)
		return fail(NRC_ERROR_CAPACITY, "n > NRC_NOALLOC_MAX_GRADIENT_RECORDS (the scratch holds the partial rows of "
		                                "at most 16384 records)");
	unsigned char *base = nullptr;
	nrc_status st = noalloc_scratch(&base);
	if (st != NRC_OK) return st;
	const hipStream_t s = reinterpret_cast<hipStream_t>(stream);
	const uint32_t wc = nrc_weight_count(hidden);
	const uint32_t pstride = ((wc + 1) + 63u) & ~63u;
	const size_t nf = nrc::fwd_frag_halves((int)hidden), nt = nrc::bwd_frag_halves((int)hidden);
	_Float16 *wf = reinterpret_cast<_Float16 *>(base);
	_Float16 *wtf = reinterpret_cast<_Float16 *>(base + kNoallocWfBytes);
	float *partials = reinterpret_cast<float *>(base + kNoallocWfBytes + kNoallocWtfBytes);
	// the fragment copies keep the last call's values: their padding must start at zero
	hipError_t e = nrc::launch_zero_words(reinterpret_cast<uint32_t *>(wf), nf / 2, s);
	if (e == hipSuccess) e = nrc::launch_zero_words(reinterpret_cast<uint32_t *>(wtf), nt / 2, s);
	if (e == hipSuccess) e = nrc::launch_make_frags((int)hidden, reinterpret_cast<const _Float16 *>(d_w), wf, wtf, s);
	uint32_t nparts = 0;
	if (e == hipSuccess)
		e = nrc::launch_fused_gradient_f16((int)hidden, reinterpret_cast<const _Float16 *>(d_in), d_targets, n, wf, wtf,
		                                   partials, pstride, &nparts, s);
	if (e == hipSuccess && (size_t)nparts > kNoallocRows) return fail(NRC_ERROR_CAPACITY, "gradient plan exceeds the scratch");
	if (e == hipSuccess) e = nrc::launch_reduce_partials(partials, nparts, pstride, wc, d_dw, s);
	if (e != hipSuccess) return hip_fail(e, "nrc_mlp_gradient_f16_noalloc");
	return NRC_OK;
}

} // extern "C"

// ------------------------------------------------------------------ groups
struct nrc_graph {
	hipGraph_t graph = nullptr;
	hipGraphExec_t exec = nullptr;
	std::vector<std::pair<nrc_cache *, int>> parity; // each cache's step parity at capture
};

namespace {

nrc_status check_group(const nrc_group_item *items, uint32_t n, bool distinct) {
	if (n && !items) return fail(NRC_ERROR_INVALID_ARGUMENT, "items is NULL");
	int dev = 0;
	NRC_HIP(hipGetDevice(&dev), "hipGetDevice");
	for (uint32_t i = 0; i < n; ++i) {
		const nrc_cache *c = items[i].cache;
		if (!c) return fail(NRC_ERROR_INVALID_ARGUMENT, "group item has a NULL cache");
		if (c->device != dev) return fail(NRC_ERROR_INVALID_ARGUMENT, "group caches must be on the current device");
		if (distinct)
			for (uint32_t j = 0; j < i; ++j)
				if (items[j].cache == c) return fail(NRC_ERROR_INVALID_ARGUMENT, "a cache appears twice in one group");
	}
	return NRC_OK;
}

// Fused launches take items of one kind (same hidden layers, encoding and, for
// training, loss) and at most kGroupMax of them: `sel` lists the item indices of
// each launch.
std::vector<std::vector<uint32_t>> group_launches(const nrc_group_item *items, const std::vector<uint32_t> &use,
                                                  bool by_loss) {
	std::vector<std::vector<uint32_t>> out;
	std::vector<bool> taken(use.size(), false);
	for (size_t a = 0; a < use.size(); ++a) {
		if (taken[a]) continue;
		const nrc_cache_desc &da = items[use[a]].cache->desc;
		std::vector<uint32_t> cur;
		for (size_t b = a; b < use.size(); ++b) {
			const nrc_cache_desc &db = items[use[b]].cache->desc;
			if (taken[b] || db.hidden_layers != da.hidden_layers || db.encoding != da.encoding ||
			    (by_loss && db.loss != da.loss))
				continue;
			taken[b] = true;
			cur.push_back(use[b]);
			if (cur.size() == nrc::kGroupMax) {
				out.push_back(cur);
				cur.clear();
			}
		}
		if (!cur.empty()) out.push_back(cur);
	}
	return out;
}

// Argument checks of a grouped training step: distinct caches on the current
// device, and every item's step (records, sizes, partial rows) valid.
nrc_status check_group_train(const nrc_group_item *items, uint32_t n_items, nrc_path path) {
	nrc_status st = check_group(items, n_items, true);
	for (uint32_t i = 0; st == NRC_OK && i < n_items; ++i)
		if (items[i].max_batch_records)
			st = check_train(items[i].cache, items[i].d_batches, items[i].max_batch_records, path);
	return st;
}

// Argument checks of a grouped inference: distinct caches (the FP8 path updates
// each cache's activation history and scales) with their buffers.
nrc_status check_group_infer(const nrc_group_item *items, uint32_t n_items, nrc_path path) {
	nrc_status st = check_group(items, n_items, true);
	if (st != NRC_OK) return st;
	if (path != NRC_PATH_PLAIN_F32 && path != NRC_PATH_FUSED_F16 && path != NRC_PATH_FUSED_F8)
		return fail(NRC_ERROR_UNSUPPORTED, "unknown path");
	for (uint32_t i = 0; i < n_items; ++i) {
		if (items[i].max_queries && (!items[i].d_queries || !items[i].d_rgb))
			return fail(NRC_ERROR_INVALID_ARGUMENT, "NULL buffer");
		if (check_count(items[i].max_queries) != NRC_OK) return NRC_ERROR_CAPACITY;
	}
	return NRC_OK;
}

static nrc_status group_train_impl(const nrc_group_item *items, uint32_t n_items, uint32_t batch, nrc_path path,
                                   uint32_t flags, hipStream_t s) {
	if (batch >= NRC_TRAIN_BATCH_COUNT) return fail(NRC_ERROR_INVALID_ARGUMENT, "batch >= NRC_TRAIN_BATCH_COUNT");
	nrc_status st = check_group_train(items, n_items, path);
	if (st != NRC_OK) return st;
	std::vector<uint32_t> use;
	for (uint32_t i = 0; i < n_items; ++i)
		if (items[i].max_batch_records) use.push_back(i);
	auto records = [&](const nrc_group_item &it) { return it.d_batches + (size_t)batch * it.max_batch_records; };
	auto counts = [&](const nrc_group_item &it) { return it.d_batch_counts ? it.d_batch_counts + batch : nullptr; };
	if (path == NRC_PATH_PLAIN_F32) {
		for (uint32_t i : use) {
			const nrc_group_item &it = items[i];
			st = train_step(it.cache, records(it), counts(it), it.max_batch_records, path, flags, s);
			if (st != NRC_OK) return st;
		}
		return NRC_OK;
	}
	for (const auto &sel : group_launches(items, use, true)) {
		nrc::GradGroup grad{};
		nrc::AdamGroup adam{};
		for (uint32_t i : sel) {
			const nrc_group_item &it = items[i];
			nrc_cache *c = it.cache;
			const nrc::TrainArgs a = train_args(c, records(it), counts(it), it.max_batch_records);
			const uint32_t n = it.max_batch_records < a.clamp ? it.max_batch_records : a.clamp;
			const nrc::FusedGradPlan plan = nrc::plan_fused_gradient(n, c->grad_chunks);
			grad.item[grad.n_items++] = nrc::grad_item(a, plan);
			adam.item[adam.n_items++] = adam_args(c, counts(it), it.max_batch_records, plan.blocks, flags);
		}
		const nrc_cache_desc &d = items[sel[0]].cache->desc;
		hipError_t e = nrc::launch_fused_gradient_group((int)d.hidden_layers, (int)d.encoding, (int)d.loss, grad, s);
		if (e != hipSuccess) return hip_fail(e, "nrc_group_train gradient launch");
		e = nrc::launch_adam_group(adam, s);
		if (e != hipSuccess) return hip_fail(e, "nrc_group_train optimizer launch");
		for (uint32_t i : sel) items[i].cache->parity ^= 1;
	}
	return NRC_OK;
}

} // namespace

extern "C" {

nrc_status nrc_group_infer(const nrc_group_item *items, uint32_t n_items, nrc_path path, void *stream) {
	nrc_status st = check_group_infer(items, n_items, path);
	if (st != NRC_OK) return st;
	std::vector<uint32_t> use;
	for (uint32_t i = 0; i < n_items; ++i)
		if (items[i].max_queries) use.push_back(i);
	if (path == NRC_PATH_PLAIN_F32) {
		for (uint32_t i : use) {
			const nrc_group_item &it = items[i];
			st = nrc_infer(it.cache, it.d_queries, it.d_query_count, it.max_queries, it.d_rgb, path, stream);
			if (st != NRC_OK) return st;
		}
		return NRC_OK;
	}
	const hipStream_t s = reinterpret_cast<hipStream_t>(stream);
	for (const auto &sel : group_launches(items, use, false)) {
		nrc::InferGroup grp{};
		grp.clamp0 = 1;
		const float *u32[nrc::kGroupMax];
		uint8_t *w8[nrc::kGroupMax];
		for (uint32_t i : sel) {
			const nrc_group_item &it = items[i];
			u32[grp.n_items] = it.cache->use32;
			w8[grp.n_items] = it.cache->w8;
			grp.item[grp.n_items++] =
			    nrc::InferItem{reinterpret_cast<const float *>(it.d_queries), it.d_query_count, it.d_rgb,
			                   it.cache->usef, it.cache->w8, it.cache->f8, it.max_queries, 0};
		}
		const nrc_cache_desc &d = items[sel[0]].cache->desc;
		hipError_t e = path == NRC_PATH_FUSED_F8
		                   ? nrc::launch_fp8_infer_group((int)d.hidden_layers, (int)d.encoding, grp, u32, w8, s)
		                   : nrc::launch_fused_infer_group((int)d.hidden_layers, (int)d.encoding, grp, s);
		if (e != hipSuccess) return hip_fail(e, "nrc_group_infer launch");
	}
	return NRC_OK;
}

nrc_status nrc_group_train(const nrc_group_item *items, uint32_t n_items, uint32_t batch, nrc_path path,
                           uint32_t flags, void *stream) {
	return group_train_impl(items, n_items, batch, path, flags, reinterpret_cast<hipStream_t>(stream));
}

nrc_status nrc_group_frame(const nrc_group_item *items, uint32_t n_items, nrc_path path, void *stream) {
	// every check before the first launch, so a rejected frame launches nothing
	nrc_status st = check_group_infer(items, n_items, path);
	if (st == NRC_OK) st = check_group_train(items, n_items, path);
	if (st != NRC_OK) return st;
	st = nrc_group_infer(items, n_items, path, stream);
	for (uint32_t b = 0; st == NRC_OK && b < NRC_TRAIN_BATCH_COUNT; ++b)
		st = group_train_impl(items, n_items, b, path, b == NRC_TRAIN_BATCH_COUNT - 1 ? NRC_TRAIN_WRITE_USE_WEIGHTS : 0u,
		                      reinterpret_cast<hipStream_t>(stream));
	return st;
}

void nrc_graph_destroy(nrc_graph *g) {
	if (!g) return;
	if (g->exec) (void)hipGraphExecDestroy(g->exec);
	if (g->graph) (void)hipGraphDestroy(g->graph);
	delete g;
}

nrc_status nrc_group_frame_graph(const nrc_group_item *items, uint32_t n_items, nrc_path path, nrc_graph **out) {
	if (!out) return fail(NRC_ERROR_INVALID_ARGUMENT, "out_graph is NULL");
	*out = nullptr;
	nrc_status st = check_group(items, n_items, true);
	if (st != NRC_OK) return st;
	// nothing may allocate while the stream is captured
	for (uint32_t i = 0; i < n_items; ++i)
		if (path == NRC_PATH_PLAIN_F32 && items[i].max_batch_records) {
			st = ensure_plain_scratch(items[i].cache);
			if (st != NRC_OK) return st;
		}
	nrc_graph *g = new nrc_graph();
	for (uint32_t i = 0; i < n_items; ++i) g->parity.emplace_back(items[i].cache, items[i].cache->parity);
	hipStream_t s = nullptr;
	hipError_t e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
	if (e != hipSuccess) {
		delete g;
		return hip_fail(e, "nrc_group_frame_graph stream");
	}
	e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
	if (e == hipSuccess) {
		st = nrc_group_frame(items, n_items, path, s);
		hipError_t e2 = hipStreamEndCapture(s, &g->graph);
		if (e2 != hipSuccess && st == NRC_OK) st = hip_fail(e2, "nrc_group_frame_graph capture");
	} else {
		st = hip_fail(e, "nrc_group_frame_graph begin capture");
	}
	(void)hipStreamDestroy(s);
	// capturing ran no step: the parities are back where they were only if every
	// cache took an even number of steps, which a frame does
	for (auto &pc : g->parity) pc.first->parity = pc.second;
	if (st == NRC_OK) {
		e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
		if (e != hipSuccess) st = hip_fail(e, "nrc_group_frame_graph instantiate");
	}
	if (st != NRC_OK) {
		nrc_graph_destroy(g);
		return st;
	}
	*out = g;
	return NRC_OK;
}

nrc_status nrc_graph_launch(nrc_graph *g, void *stream) {
	if (!g || !g->exec) return fail(NRC_ERROR_INVALID_ARGUMENT, "graph is NULL");
	for (const auto &pc : g->parity)
		if (pc.first->parity != pc.second)
			return fail(NRC_ERROR_INVALID_ARGUMENT, "a cache of the graph took steps outside it since capture");
	NRC_HIP(hipGraphLaunch(g->exec, reinterpret_cast<hipStream_t>(stream)), "nrc_graph_launch");
	return NRC_OK;
}

} // extern "C"
