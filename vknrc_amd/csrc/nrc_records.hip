// nrc_records.hip -- the record front-end around the fused MLP: unpacking the
// renderer's wire records against the scene (UnpackNRCInput, NRCRecord.glsl:98-125),
// the inference epilogue over wire records (nrc_inference.comp:50-72), the suffix
// sums that turn a traced path into train records (path_tracer.comp:343-372) and
// the screen pass (screen.frag:10-31).
//
// All four are gathers and streams over HBM with a few dozen FLOPs a record, one
// thread per record (or pixel); the MLP kernels read the unpacked 56-byte queries.
#include "nrc_internal.hpp"

namespace nrc {

namespace {

constexpr uint32_t kNoTexture = 0xFFFFFFFFu;

// unpackUnorm2x16: c / 65535 per 16-bit half
__device__ __forceinline__ void unorm2x16(uint32_t v, float *o) {
	o[0] = (float)(v & 0xFFFFu) / 65535.0f;
	o[1] = (float)(v >> 16) / 65535.0f;
}

__device__ __forceinline__ float srgb_decode(uint32_t c8) {
	const float c = (float)c8 / 255.0f;
	return c <= 0.04045f ? c / 12.92f : powf((c + 0.055f) / 1.055f, 2.4f);
}

// The 256 sRGB decodes, built once per workgroup (one per thread of a 256-thread
// block) so that a texel channel costs one LDS read instead of a powf.
__device__ __forceinline__ void srgb_table_init(float *lut) {
	lut[threadIdx.x] = srgb_decode(threadIdx.x);
	__syncthreads();
}

// Texel coordinates of a repeat-addressed linear fetch along one axis of n texels:
// the coordinate's fraction (in [0, 1]) scaled to texels, so the two texel indices
// are in [-1, n] and wrap with one select each.  NaN or infinite coordinates give
// texel 0 (v_cvt of NaN is 0), so every index stays inside the image.
__device__ __forceinline__ void repeat_texels(float u, uint32_t n, uint32_t &i0, uint32_t &i1, float &a) {
	const float uf = u - floorf(u);
	const float x = uf * (float)n - 0.5f;
	const float x0 = floorf(x);
	a = x - x0;
	const int k = cvt_i32_sat(x0);
	i0 = k < 0 ? n - 1u : (uint32_t)k < n ? (uint32_t)k : 0u;
	i1 = (uint32_t)(k + 1) < n ? (uint32_t)(k + 1) : 0u;
}

// texture(sampler2D, uv).rgb with linear filtering and repeat addressing
// (NRCRenderGraph.cpp:132) of an RGBA8 sRGB image (VkScene.cpp:193)
__device__ void sample_texture(const nrc_texture &t, float u, float v, const float *lut, float *o) {
	uint32_t x0, x1, y0, y1;
	float ax, ay;
	repeat_texels(u, t.width, x0, x1, ax);
	repeat_texels(v, t.height, y0, y1, ay);
	const uint32_t p00 = t.d_rgba8[(size_t)y0 * t.width + x0], p01 = t.d_rgba8[(size_t)y0 * t.width + x1];
	const uint32_t p10 = t.d_rgba8[(size_t)y1 * t.width + x0], p11 = t.d_rgba8[(size_t)y1 * t.width + x1];
#pragma unroll
	for (int c = 0; c < 3; ++c) {
		const uint32_t sh = 8u * (uint32_t)c;
		const float c00 = lut[(p00 >> sh) & 255u], c01 = lut[(p01 >> sh) & 255u];
		const float c10 = lut[(p10 >> sh) & 255u], c11 = lut[(p11 >> sh) & 255u];
		const float top = c00 + ax * (c01 - c00), bot = c10 + ax * (c11 - c10);
		o[c] = top + ay * (bot - top);
	}
}

// GetSceneDiffuse / GetSceneSpecular (Scene.glsl:63-68); an id past the texture
// table, or an empty texture, takes the constant colour
__device__ __forceinline__ void material_color(const nrc_scene &s, const float *constant, uint32_t tex, float u, float v,
                                               const float *lut, float *o) {
	const nrc_texture *t = (tex == kNoTexture || s.d_textures == nullptr || tex >= s.texture_count) ? nullptr
	                                                                                               : s.d_textures + tex;
	if (t == nullptr || t->width == 0u || t->height == 0u || t->d_rgba8 == nullptr) {
		o[0] = constant[0];
		o[1] = constant[1];
		o[2] = constant[2];
	} else {
		sample_texture(*t, u, v, lut, o);
	}
}

// A primitive's corners (object space), their texture coordinates and its material:
// gathered through the index arrays (Scene.glsl:50-62), or read from the primitive's
// 64-byte packed record (nrc_scene_pack_primitives) as four 16-byte loads.
struct PrimitiveData {
	float c[3][3], t[3][2];
	uint32_t material;
};

template <bool PACKED> __device__ __forceinline__ PrimitiveData load_primitive(const nrc_scene &s, uint32_t prim) {
	PrimitiveData d;
	if constexpr (PACKED) {
		const float4 *r = reinterpret_cast<const float4 *>(s.d_primitives + prim);
		const float4 a = r[0], b = r[1], c = r[2], e = r[3];
		d.c[0][0] = a.x; d.c[0][1] = a.y; d.c[0][2] = a.z;
		d.c[1][0] = a.w; d.c[1][1] = b.x; d.c[1][2] = b.y;
		d.c[2][0] = b.z; d.c[2][1] = b.w; d.c[2][2] = c.x;
		d.t[0][0] = c.y; d.t[0][1] = c.z;
		d.t[1][0] = c.w; d.t[1][1] = e.x;
		d.t[2][0] = e.y; d.t[2][1] = e.z;
		d.material = __float_as_uint(e.w);
	} else {
#pragma unroll
		for (int k = 0; k < 3; ++k) {
			const uint32_t vi = s.d_vertex_indices[prim * 3u + k], ti = s.d_texcoord_indices[prim * 3u + k];
#pragma unroll
			for (int j = 0; j < 3; ++j) d.c[k][j] = s.d_vertices[(size_t)vi * 3 + j];
			d.t[k][0] = s.d_texcoords[(size_t)ti * 2];
			d.t[k][1] = s.d_texcoords[(size_t)ti * 2 + 1];
		}
		d.material = s.d_material_ids[prim];
	}
	return d;
}

// GetSceneVertex (Scene.glsl:50-53): (x, y, z, 1) times the instance's 3x4 rows
__device__ __forceinline__ void transform_vertex(const float *T, const float *v, float *o) {
#pragma unroll
	for (int j = 0; j < 3; ++j) o[j] = T[4 * j] * v[0] + T[4 * j + 1] * v[1] + T[4 * j + 2] * v[2] + T[4 * j + 3];
}

// UnpackNRCInput (NRCRecord.glsl:98-125) into the 14 floats of an nrc_query
template <bool PACKED>
__device__ void unpack_input(const nrc_scene &s, const nrc_packed_input &p, const float *lut, float *q) {
	const uint32_t prim = p.primitive_id, inst = p.flip_bit_instance_id & 0x7FFFFFFFu;
	const bool flip = (p.flip_bit_instance_id >> 31) != 0u;
	const PrimitiveData d = load_primitive<PACKED>(s, prim);
	const float *T = s.d_transforms + (size_t)inst * 12;
	float v0[3], v1[3], v2[3];
	transform_vertex(T, d.c[0], v0);
	transform_vertex(T, d.c[1], v1);
	transform_vertex(T, d.c[2], v2);
	const float e1[3] = {v1[0] - v0[0], v1[1] - v0[1], v1[2] - v0[2]};
	const float e2[3] = {v2[0] - v0[0], v2[1] - v0[1], v2[2] - v0[2]};
	float n[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
	const float inv = 1.0f / sqrtf(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
	const float sgn = flip ? -inv : inv;
	n[0] *= sgn;
	n[1] *= sgn;
	n[2] *= sgn;
	float b[3];
	unorm2x16(p.barycentric_2x16u, b + 1);
	b[0] = 1.0f - b[1] - b[2];
#pragma unroll
	for (int j = 0; j < 3; ++j) q[j] = v0[j] * b[0] + v1[j] * b[1] + v2[j] * b[2];
	unorm2x16(p.scattered_dir_2x16u, q + 3);
	sph_encode(n[0], n[1], n[2], q + 5);
	const nrc_material &m = s.d_materials[d.material];
	q[7] = m.roughness;
	const float u = d.t[0][0] * b[0] + d.t[1][0] * b[1] + d.t[2][0] * b[2];
	const float v = d.t[0][1] * b[0] + d.t[1][1] * b[1] + d.t[2][1] * b[2];
	material_color(s, m.diffuse, m.diffuse_texture_id, u, v, lut, q + 8);
	material_color(s, m.specular, m.specular_texture_id, u, v, lut, q + 11);
}

__device__ __forceinline__ nrc_packed_input load_packed(const uint32_t *w) {
	nrc_packed_input p;
	p.primitive_id = w[0];
	p.flip_bit_instance_id = w[1];
	p.barycentric_2x16u = w[2];
	p.scattered_dir_2x16u = w[3];
	return p;
}

template <bool PACKED>
__global__ __launch_bounds__(256) void unpack_eval_kernel(const nrc_scene s, const nrc_eval_record *__restrict__ rec,
                                                          const uint32_t *d_count, uint32_t max_count,
                                                          float *__restrict__ out) {
	__shared__ float lut[256];
	srgb_table_init(lut);
	const uint32_t i = blockIdx.x * 256u + threadIdx.x;
	if (i >= effective_count(d_count, max_count, max_count)) return;
	const uint32_t *w = reinterpret_cast<const uint32_t *>(rec + i);
	float q[kQueryFloats];
	unpack_input<PACKED>(s, load_packed(w + 1), lut, q);
	float2 *o = reinterpret_cast<float2 *>(out + (size_t)i * kQueryFloats);
#pragma unroll
	for (int k = 0; k < kQueryFloats / 2; ++k) o[k] = make_float2(q[2 * k], q[2 * k + 1]);
}

template <bool PACKED>
__global__ __launch_bounds__(256) void unpack_train_kernel(const nrc_scene s,
                                                           const nrc_wire_train_record *__restrict__ rec,
                                                           const uint32_t *d_count, uint32_t max_count,
                                                           float *__restrict__ out) {
	__shared__ float lut[256];
	srgb_table_init(lut);
	const uint32_t i = blockIdx.x * 256u + threadIdx.x;
	if (i >= effective_count(d_count, max_count, max_count)) return;
	const uint32_t *w = reinterpret_cast<const uint32_t *>(rec + i);
	float q[kTrainFloats];
	unpack_input<PACKED>(s, load_packed(w + 6), lut, q);
	// nrc_gradient.comp:30-33: the target is the record's bias
	q[14] = __uint_as_float(w[0]);
	q[15] = __uint_as_float(w[1]);
	q[16] = __uint_as_float(w[2]);
	float *o = out + (size_t)i * kTrainFloats;
#pragma unroll
	for (int k = 0; k < kTrainFloats; ++k) o[k] = q[k];
}

// One thread per primitive: the corners and texture coordinates its indices name and
// its material id, in the 64-byte record that load_primitive<true> reads (16-float
// order: corners 0..2 xyz, texcoords 0..2 uv, material id bits).
__global__ __launch_bounds__(256) void pack_primitives_kernel(const nrc_scene s, uint32_t n,
                                                              nrc_scene_primitive *__restrict__ out) {
	const uint32_t i = blockIdx.x * 256u + threadIdx.x;
	if (i >= n) return;
	const PrimitiveData d = load_primitive<false>(s, i);
	float4 *o = reinterpret_cast<float4 *>(out + i);
	o[0] = make_float4(d.c[0][0], d.c[0][1], d.c[0][2], d.c[1][0]);
	o[1] = make_float4(d.c[1][1], d.c[1][2], d.c[2][0], d.c[2][1]);
	o[2] = make_float4(d.c[2][2], d.t[0][0], d.t[0][1], d.t[1][0]);
	o[3] = make_float4(d.t[1][1], d.t[2][0], d.t[2][1], __uint_as_float(d.material));
}

// nrc_inference.comp:50-72 over wire records: the dst code comes with the eval
// record, the factor of a back-filled train record is in the record itself
__global__ __launch_bounds__(256) void composite_records_kernel(const float *__restrict__ rgb,
                                                                const nrc_eval_record *__restrict__ rec,
                                                                const uint32_t *d_count, uint32_t max_count,
                                                                float *bias_factor_r, const float *factor_gb,
                                                                uint32_t width, uint32_t height, uint32_t pitch,
                                                                nrc_wire_train_record *batches, uint32_t batch_stride) {
	const uint32_t i = blockIdx.x * 256u + threadIdx.x;
	if (i >= effective_count(d_count, max_count, max_count)) return;
	const uint32_t e = rec[i].dst;
	if (e == 0xFFFFFFFFu) return;
	const float p0 = rgb[(size_t)i * 3], p1 = rgb[(size_t)i * 3 + 1], p2 = rgb[(size_t)i * 3 + 2];
	const uint32_t v = e >> 1;
	if ((e & 1u) == 0u) { // screen (DecodeNRCEvalDstScreen, NRCRecord.glsl:24-27)
		const uint32_t x = v & 0x7FFFu, y = v >> 15;
		if (bias_factor_r == nullptr || x >= width || y >= height) return;
		const size_t px = (size_t)y * pitch + x;
		float4 *bf = reinterpret_cast<float4 *>(bias_factor_r) + px;
		const float4 b = *bf;
		const float2 f = reinterpret_cast<const float2 *>(factor_gb)[px];
		*bf = make_float4(mul_add_rn(b.x, b.w, p0), mul_add_rn(b.y, f.x, p1), mul_add_rn(b.z, f.y, p2), 0.0f);
	} else { // train (DecodeNRCEvalDstTrain, NRCRecord.glsl:28-33)
		const uint32_t b = v & 3u, l = (v >> 2) & 0x3FFFu, r = v >> 16;
		if (batches == nullptr) return;
		for (uint32_t j = l; j <= r && j < batch_stride; ++j) {
			nrc_wire_train_record &t = batches[(size_t)b * batch_stride + j];
			t.bias[0] = mul_add_rn(t.bias[0], t.factor[0], p0);
			t.bias[1] = mul_add_rn(t.bias[1], t.factor[1], p1);
			t.bias[2] = mul_add_rn(t.bias[2], t.factor[2], p2);
		}
	}
}

// path_tracer.comp:343-372: one thread per path
__global__ __launch_bounds__(256) void build_train_records_kernel(const nrc_traced_path *__restrict__ paths, uint32_t n_paths,
                                                                  nrc_wire_train_record *batches, uint32_t *batch_counts,
                                                                  uint32_t capacity, nrc_eval_record *evals,
                                                                  uint32_t *eval_count, uint32_t eval_capacity) {
	const uint32_t i = blockIdx.x * 256u + threadIdx.x;
	if (i >= n_paths) return;
	const nrc_traced_path &P = paths[i];
	const uint32_t bounce = P.bounce;
	if (bounce == 0u || bounce > (uint32_t)NRC_MAX_BOUNCE) return;
	const uint32_t batch = P.batch < (uint32_t)NRC_TRAIN_BATCH_COUNT ? P.batch : (uint32_t)NRC_TRAIN_BATCH_COUNT - 1u;
	const uint32_t train_id = atomicAdd(batch_counts + batch, bounce);
	if (train_id >= capacity) return;
	float L[NRC_MAX_BOUNCE][3], Cc[NRC_MAX_BOUNCE][3];
#pragma unroll
	for (int k = 0; k < NRC_MAX_BOUNCE; ++k)
#pragma unroll
		for (int c = 0; c < 3; ++c) {
			L[k][c] = (uint32_t)k < bounce ? P.light[k][c] : 0.0f;
			Cc[k][c] = (uint32_t)k < bounce ? P.color[k][c] : 1.0f;
		}
	// lights[i] += colors[i] * lights[i + 1]; colors[i] *= colors[i + 1] for i = bounce - 2 .. 0
#pragma unroll
	for (int k = NRC_MAX_BOUNCE - 2; k >= 0; --k)
		if ((uint32_t)k + 2u <= bounce)
#pragma unroll
			for (int c = 0; c < 3; ++c) {
				L[k][c] = mul_add_rn(L[k][c], Cc[k][c], L[k + 1][c]);
				Cc[k][c] = Cc[k][c] * Cc[k + 1][c];
			}
	const uint32_t left = capacity - train_id;
	const uint32_t train_count = bounce < left ? bounce : left;
	nrc_wire_train_record *dst = batches + (size_t)batch * capacity + train_id;
#pragma unroll
	for (int k = 0; k < NRC_MAX_BOUNCE; ++k)
		if ((uint32_t)k < train_count) {
			nrc_wire_train_record r;
			r.bias[0] = L[k][0];
			r.bias[1] = L[k][1];
			r.bias[2] = L[k][2];
			r.factor[0] = Cc[k][0];
			r.factor[1] = Cc[k][1];
			r.factor[2] = Cc[k][2];
			r.packed_input = P.input[k];
			dst[k] = r;
		}
	if (P.light_terminate == 0u) {
		const uint32_t eid = atomicAdd(eval_count, 1u);
		if (eid < eval_capacity) {
			nrc_eval_record er;
			// EncodeNRCEvalDstTrain (NRCRecord.glsl:20-22)
			er.dst = ((batch | (train_id << 2) | ((train_id + train_count - 1u) << 16)) << 1) | 1u;
			er.packed_input = P.input[bounce - 1u];
			evals[eid] = er;
		}
	}
}

// screen.frag:10-31
__device__ __forceinline__ float filmic_hejl(float h) {
	const float a = 1.425f * h + 0.05f;
	// the quotient through the hardware reciprocal (about 1 ulp), as a shader computes it
	return (h * a + 0.004f) * __builtin_amdgcn_rcpf(h * (a + 0.55f) + 0.0491f) - 0.0821f;
}

__global__ __launch_bounds__(256) void screen_kernel(const float *__restrict__ color, float *accumulate_img,
                                                     uint32_t width, uint32_t height, uint32_t pitch, int accumulate,
                                                     uint32_t count, float *__restrict__ out) {
	const uint32_t x = blockIdx.x * 16u + (threadIdx.x & 15u), y = blockIdx.y * 16u + (threadIdx.x >> 4);
	if (x >= width || y >= height) return;
	const size_t px = (size_t)y * pitch + x;
	const float4 c4 = reinterpret_cast<const float4 *>(color)[px];
	float c[3] = {c4.x, c4.y, c4.z};
	if (accumulate) {
		float4 *acc = reinterpret_cast<float4 *>(accumulate_img) + px;
		if (count != 0u) {
			const float4 a = *acc;
			// one reciprocal for the three channels (within 2 ulps of the division)
			const float n = (float)count, r = __builtin_amdgcn_rcpf((float)(count + 1u));
			c[0] = (c[0] + a.x * n) * r;
			c[1] = (c[1] + a.y * n) * r;
			c[2] = (c[2] + a.z * n) * r;
		}
		*acc = make_float4(c[0], c[1], c[2], 0.0f);
	}
	// ToneMapFilmic_Hejl2015(color, 3.2): the curve of each channel over the curve of the white point
	const float rfw = __builtin_amdgcn_rcpf(filmic_hejl(3.2f));
	float o[3];
#pragma unroll
	for (int k = 0; k < 3; ++k) {
		const float t = filmic_hejl(c[k]) * rfw;
		// GLSL pow(t, 1/2.2) is exp2(log2(t) / 2.2) (the precision GLSL specifies for it), here on the
		// hardware log2 and exp2; pow of the slightly negative values the curve gives near black is
		// undefined in GLSL, 0 here
		o[k] = t > 0.0f ? __builtin_amdgcn_exp2f(__builtin_amdgcn_logf(t) * (1.0f / 2.2f)) : 0.0f;
	}
	reinterpret_cast<float4 *>(out)[px] = make_float4(o[0], o[1], o[2], 1.0f);
}

} // namespace

hipError_t launch_pack_primitives(const nrc_scene &s, uint32_t n, nrc_scene_primitive *out, hipStream_t stream) {
	if (n == 0) return hipSuccess;
	pack_primitives_kernel<<<dim3((n + 255) / 256), dim3(256), 0, stream>>>(s, n, out);
	return hipGetLastError();
}

hipError_t launch_unpack_eval(const nrc_scene &s, const nrc_eval_record *rec, const uint32_t *d_count,
                              uint32_t max_count, float *queries, hipStream_t stream) {
	if (max_count == 0) return hipSuccess;
	if (s.d_primitives)
		unpack_eval_kernel<true><<<dim3((max_count + 255) / 256), dim3(256), 0, stream>>>(s, rec, d_count, max_count, queries);
	else
		unpack_eval_kernel<false><<<dim3((max_count + 255) / 256), dim3(256), 0, stream>>>(s, rec, d_count, max_count, queries);
	return hipGetLastError();
}

hipError_t launch_unpack_train(const nrc_scene &s, const nrc_wire_train_record *rec, const uint32_t *d_count,
                               uint32_t max_count, float *out, hipStream_t stream) {
	if (max_count == 0) return hipSuccess;
	if (s.d_primitives)
		unpack_train_kernel<true><<<dim3((max_count + 255) / 256), dim3(256), 0, stream>>>(s, rec, d_count, max_count, out);
	else
		unpack_train_kernel<false><<<dim3((max_count + 255) / 256), dim3(256), 0, stream>>>(s, rec, d_count, max_count, out);
	return hipGetLastError();
}

hipError_t launch_composite_records(const float *rgb, const nrc_eval_record *rec, const uint32_t *d_count,
                                    uint32_t max_count, float *bias_factor_r, const float *factor_gb, uint32_t width,
                                    uint32_t height, uint32_t pitch, nrc_wire_train_record *batches,
                                    uint32_t batch_stride, hipStream_t stream) {
	if (max_count == 0) return hipSuccess;
	composite_records_kernel<<<dim3((max_count + 255) / 256), dim3(256), 0, stream>>>(
	    rgb, rec, d_count, max_count, bias_factor_r, factor_gb, width, height, pitch, batches, batch_stride);
	return hipGetLastError();
}

hipError_t launch_build_train_records(const nrc_traced_path *paths, uint32_t n_paths, nrc_wire_train_record *batches,
                                      uint32_t *batch_counts, uint32_t capacity, nrc_eval_record *evals,
                                      uint32_t *eval_count, uint32_t eval_capacity, hipStream_t stream) {
	if (n_paths == 0) return hipSuccess;
	build_train_records_kernel<<<dim3((n_paths + 255) / 256), dim3(256), 0, stream>>>(
	    paths, n_paths, batches, batch_counts, capacity, evals, eval_count, eval_capacity);
	return hipGetLastError();
}

hipError_t launch_screen(const float *color, float *accumulate_img, uint32_t width, uint32_t height, uint32_t pitch,
                         int accumulate, uint32_t count, float *out, hipStream_t stream) {
	if (width == 0 || height == 0) return hipSuccess;
	screen_kernel<<<dim3((width + 15) / 16, (height + 15) / 16), dim3(256), 0, stream>>>(
	    color, accumulate_img, width, height, pitch, accumulate, count, out);
	return hipGetLastError();
}

} // namespace nrc
