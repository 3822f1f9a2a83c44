// nrc_mfma.hpp -- device helpers shared by the MFMA kernels (gfx950).
//
// MFMA used: v_mfma_f32_16x16x32_{f16,fp8_fp8}.  Per instruction, lane
// l = 16*g + r holds
//   A: row r,    8 elements of K-group g      B: column r, 8 elements of K-group g
//   D: column r, rows 4g .. 4g+3 (FP32)
// A and B pair their elements by (g, j), so any bijection between (g, j) and the
// 32 K indices of an instruction gives the same product as long as A and B use
// the same one.  The kernels use
//   k(c, g, j) = 32c + 4g + j        (j < 4)
//              = 32c + 16 + 4g + j-4 (j >= 4)        for K-chunk c in {0, 1}
// because that is exactly where the D tiles of the previous layer leave the
// features: D of M-tile 2c holds rows 32c + 4g + i, D of M-tile 2c+1 holds rows
// 32c + 16 + 4g + i.  Layer l's output therefore becomes layer l+1's B operand
// in the same lanes, with a ReLU and a pack and no data movement.
#pragma once

#include <cstdlib>
#include <type_traits>

#include "nrc_internal.hpp"

namespace nrc {
namespace {

typedef _Float16 half8_t __attribute__((ext_vector_type(8)));
typedef _Float16 half4_t __attribute__((ext_vector_type(4)));
typedef float float4_t __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float4_t mfma(half8_t a, half8_t b, float4_t c) {
	return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
}

typedef _Float16 half2_t __attribute__((ext_vector_type(2)));
typedef float float2_t __attribute__((ext_vector_type(2)));

// Two FP32 values -> packed FP16 (round to nearest even), then ReLU on the pair.
// max(cvt(x), 0) equals cvt(max(x, 0)) up to the sign of a zero.
__device__ __forceinline__ half2_t cvt_relu2(float a, float b) {
	const float2_t f = {a, b};
	return __builtin_elementwise_max(__builtin_convertvector(f, half2_t), (half2_t)0);
}

__device__ __forceinline__ half8_t pack_relu(float4_t lo, float4_t hi) {
	const half2_t a = cvt_relu2(lo[0], lo[1]), b = cvt_relu2(lo[2], lo[3]);
	const half2_t c = cvt_relu2(hi[0], hi[1]), d = cvt_relu2(hi[2], hi[3]);
	half8_t h;
	h[0] = a[0]; h[1] = a[1]; h[2] = b[0]; h[3] = b[1];
	h[4] = c[0]; h[5] = c[1]; h[6] = d[0]; h[7] = d[1];
	return h;
}

// Workgroup barrier that orders LDS traffic only: the fences are restricted to
// LDS, so the barrier waits on LDS counters and leaves global loads in flight
// (for data exchanged through LDS; global memory needs __syncthreads).
__device__ __forceinline__ void block_lds_sync() {
	__builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
	__builtin_amdgcn_s_barrier();
	__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Calls f(std::integral_constant<int, i>{}) for i = 0 .. N-1, unrolled at compile time.
template <typename F, int... I> __device__ __forceinline__ void static_for_seq(F &&f, std::integer_sequence<int, I...>) {
	(f(std::integral_constant<int, I>{}), ...);
}
template <int N, typename F> __device__ __forceinline__ void static_for(F &&f) {
	static_for_seq(f, std::make_integer_sequence<int, N>{});
}

// Orders LDS traffic between lanes of one wavefront (LDS operations of a wave
// execute in order; this keeps the compiler from moving them across).  The
// fences are restricted to LDS so they wait on LDS counters only and leave the
// prefetched global loads in flight.
__device__ __forceinline__ void wave_lds_sync() {
	__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
	__builtin_amdgcn_wave_barrier();
	__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
}

// --------------------------------------------------------------- encoders
// The inference kernels round the features to FP16 (or e4m3) right away, so they
// use shorter forms of two encoders than the FP32 restatements in nrc_device.hpp,
// equal in exact arithmetic and within a few FP32 ulps of them:
//   * one-blob (NRCOneBlob4Encode, NRCRecord.glsl:51-63): per knot k = 0..4,
//     c_k = clamp(0.5 + u (15/16 - 5/8 u^2 + 3/16 u^4), 0, 1) with u = k - 4x,
//     which is (15/16) u (1 - (2/3) u^2 + (1/5) u^4) + 0.5 in Horner form
//     (6 instructions a knot instead of 11);
//   * SINCOS6: sin/cos(2^k pi p) for k = 0..5 from one polynomial evaluation at
//     k = 0 and five double-angle steps (largest error 1.3e-5 at k = 5, a twentieth
//     of an FP16 half-ulp at 1);
//   * TRI12 (_nrc_tri, NRCRecord.glsl:65-68): with a/2 = 2^(k-1) p - 1/4, the GLSL
//     2 |mod(a, 2) - 1| - 1 is |4 fract(a/2) - 2| - 1, the same value (every step
//     scales by a power of two or rounds once in both) except where fract clamps
//     a result that rounds to 1 down to 1 - 2^-24.
__device__ __forceinline__ void oneblob4_knots(float x, float *o) {
	const float x4 = 4.0f * x;
	float c[5];
#pragma unroll
	for (int k = 0; k < 5; ++k) {
		const float u = (float)k - x4;
		const float u2 = u * u;
		float t = fmaf(u2, 0.1875f, -0.625f);
		t = fmaf(u2, t, 0.9375f);
		c[k] = fminf(fmaxf(fmaf(u, t, 0.5f), 0.0f), 1.0f);
	}
#pragma unroll
	for (int b = 0; b < 4; ++b) o[b] = c[b + 1] - c[b];
}

template <int ENC> __device__ __forceinline__ void axis12_infer(float p, float *o) {
	if constexpr (ENC == NRC_ENCODING_SINCOS6) {
		float sv, cv;
		sincospi_poly(p, &sv, &cv);
		o[0] = sv;
		o[6] = cv;
#pragma unroll
		for (int k = 1; k < 6; ++k) {
			const float ns = 2.0f * sv * cv, nc = 1.0f - 2.0f * sv * sv;
			sv = ns;
			cv = nc;
			o[k] = sv;
			o[6 + k] = cv;
		}
	} else {
#pragma unroll
		for (int k = 0; k < 12; ++k) {
			const float h = __builtin_amdgcn_fractf(fmaf(p, 0.5f * (float)(1 << k), -0.25f));
			o[k] = fabsf(fmaf(h, 4.0f, -2.0f)) - 1.0f;
		}
	}
}
// Emits the 16 feature quads (features 4p..4p+3) of one record; Emit(p, f0..f3).
template <int ENC, typename Emit> __device__ __forceinline__ void encode_quads(const float *rec, Emit &&emit) {
	if constexpr (ENC == NRC_ENCODING_UV_SINCOS16) {
#pragma unroll
		for (int a = 0; a < 2; ++a) {
#pragma unroll
			for (int k0 = 0; k0 < 16; k0 += 4) {
				float s[4], c[4];
				float sv, cv;
				sincospi_poly(rec[a] * (float)(1 << k0), &sv, &cv);
				s[0] = sv;
				c[0] = cv;
#pragma unroll
				for (int k = 1; k < 4; ++k) {
					const float ns = 2.0f * sv * cv, nc = 1.0f - 2.0f * sv * sv;
					sv = ns;
					cv = nc;
					s[k] = sv;
					c[k] = cv;
				}
				emit(8 * a + k0 / 4, s[0], s[1], s[2], s[3]);
				emit(8 * a + 4 + k0 / 4, c[0], c[1], c[2], c[3]);
			}
		}
	} else {
#pragma unroll
		for (int a = 0; a < 3; ++a) {
			float t[12];
			axis12_infer<ENC>(rec[a], t);
			emit(3 * a + 0, t[0], t[1], t[2], t[3]);
			emit(3 * a + 1, t[4], t[5], t[6], t[7]);
			emit(3 * a + 2, t[8], t[9], t[10], t[11]);
		}
		float b[4];
		oneblob4_knots(rec[3], b);
		emit(9, b[0], b[1], b[2], b[3]);
		oneblob4_knots(rec[4], b);
		emit(10, b[0], b[1], b[2], b[3]);
		oneblob4_knots(rec[5], b);
		emit(11, b[0], b[1], b[2], b[3]);
		oneblob4_knots(rec[6], b);
		emit(12, b[0], b[1], b[2], b[3]);
		oneblob4_knots(1.0f - expf(-rec[7]), b);
		emit(13, b[0], b[1], b[2], b[3]);
		emit(14, rec[8], rec[9], rec[10], rec[11]);
		emit(15, rec[12], rec[13], 1.0f, 1.0f);
	}
}


// The same quads in pieces, so that the encoding of one tile can be spread over
// the layers of another: piece P of encode_pieces<ENC>() emits
//   TRI12 / SINCOS6:  P = 0..2 the 12 features of position axis P (quads 3P..3P+2),
//                     P = 3..7 one one-blob quad each (quads 9..13), P = 8 quads 14, 15;
//   UV_SINCOS16:      P = 0..7 the sin and cos quads of octaves 4(P&3).. of axis P>>2.
template <int ENC> constexpr int encode_pieces() { return ENC == NRC_ENCODING_UV_SINCOS16 ? 8 : 9; }

template <int ENC, int P, typename Emit> __device__ __forceinline__ void encode_piece(const float *rec, Emit &&emit) {
	if constexpr (ENC == NRC_ENCODING_UV_SINCOS16) {
		constexpr int a = P >> 2, k0 = 4 * (P & 3);
		float s[4], c[4];
		float sv, cv;
		sincospi_poly(rec[a] * (float)(1 << k0), &sv, &cv);
		s[0] = sv;
		c[0] = cv;
#pragma unroll
		for (int k = 1; k < 4; ++k) {
			const float ns = 2.0f * sv * cv, nc = 1.0f - 2.0f * sv * sv;
			sv = ns;
			cv = nc;
			s[k] = sv;
			c[k] = cv;
		}
		emit(8 * a + k0 / 4, s[0], s[1], s[2], s[3]);
		emit(8 * a + 4 + k0 / 4, c[0], c[1], c[2], c[3]);
	} else if constexpr (P < 3) {
		float t[12];
		axis12_infer<ENC>(rec[P], t);
		emit(3 * P + 0, t[0], t[1], t[2], t[3]);
		emit(3 * P + 1, t[4], t[5], t[6], t[7]);
		emit(3 * P + 2, t[8], t[9], t[10], t[11]);
	} else if constexpr (P < 8) {
		float b[4];
		oneblob4_knots(P == 7 ? 1.0f - expf(-rec[7]) : rec[P], b);
		emit(6 + P, b[0], b[1], b[2], b[3]);
	} else {
		emit(14, rec[8], rec[9], rec[10], rec[11]);
		emit(15, rec[12], rec[13], 1.0f, 1.0f);
	}
}

// feature index of element j of a B operand held by lane group g for K-chunk c
__device__ __forceinline__ int bfeat(int c, int g, int j) { return 32 * c + (j < 4 ? 4 * g + j : 16 + 4 * g + j - 4); }

} // namespace
} // namespace nrc
