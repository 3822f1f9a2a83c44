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
//     (6 instructions a knot instead of 11; oneblob4_knots, which the FP8 kernel
//     uses); the FP16 kernels evaluate only the two unsaturated knots and place
//     the packed quad with byte permutes (oneblob4_packed);
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

// The same one-blob quad with two knots, packed to FP16.  With j = floor(4x) and
// f = 4x - j (exact), only knots j and j+1 have |u| < 1: c_j = C(-f), c_{j+1} =
// C(1 - f), every knot below j is 0 and every one above j+1 is 1.  So the quad holds
// c_j, c_{j+1} - c_j and 1 - c_{j+1} at positions j-1, j and j+1 and zeros elsewhere.
// The knot arguments equal the five-knot ones bit for bit (f and 1 - f are exact for
// 4x in [0, 4]), so the two forms agree except where the five-knot form evaluates a
// saturated knot (|u| just above 1) to one FP32 ulp short of 0 or 1: there it gives
// a feature of 6e-8 (the smallest FP16 subnormal) and this form the exact 0.  An
// exhaustive check over every float in [0, 1] finds 1,786 such inputs.
// The placement is two byte permutes of (c_j, c_{j+1} - c_j | 1 - c_{j+1}, 0) with
// selectors from an 8-entry table indexed by clamp(j, -2, 5) + 2 (selector byte 0x0C
// is a zero byte): j <= -2 (4x < -1) and j >= 5 (4x >= 5) give the all-zero quad,
// which is exact there, so no bound on x needs a register.
constexpr uint32_t kOneBlobSel[8][2] = {
    {0x0C0C0C0Cu, 0x0C0C0C0Cu}, // j <= -2: zeros
    {0x0C0C0504u, 0x0C0C0C0Cu}, // j = -1: (1 - c1, 0, 0, 0)
    {0x05040302u, 0x0C0C0C0Cu}, // j = 0:  (c1 - c0, 1 - c1, 0, 0)
    {0x03020100u, 0x0C0C0504u}, // j = 1:  (c0, c1 - c0, 1 - c1, 0)
    {0x01000C0Cu, 0x05040302u}, // j = 2:  (0, c0, c1 - c0, 1 - c1)
    {0x0C0C0C0Cu, 0x03020100u}, // j = 3:  (0, 0, c0, c1 - c0)
    {0x0C0C0C0Cu, 0x01000C0Cu}, // j = 4:  (0, 0, 0, c0)
    {0x0C0C0C0Cu, 0x0C0C0C0Cu}, // j >= 5: zeros
};

// writes the selector table into tab[0..7] (lanes 0..7 of the calling wave)
__device__ __forceinline__ void oneblob_sel_init(uint2 *tab, int lane) {
	if (lane < 8) tab[lane] = make_uint2(kOneBlobSel[lane][0], kOneBlobSel[lane][1]);
}

__device__ __forceinline__ float quartic_cdf(float u) {
	const float u2 = u * u;
	float t = fmaf(u2, 0.1875f, -0.625f);
	t = fmaf(u2, t, 0.9375f);
	return fminf(fmaxf(fmaf(u, t, 0.5f), 0.0f), 1.0f);
}

__device__ __forceinline__ uint32_t pack_half2_bits(float a, float b) {
	const float2_t f = {a, b};
	return __builtin_bit_cast(uint32_t, __builtin_convertvector(f, half2_t));
}

__device__ __forceinline__ uint2 oneblob4_packed(float x, const uint2 *sel) {
	const float x4 = 4.0f * x;
	const float j = __builtin_floorf(x4);
	const float f = x4 - j;
	const float c0 = quartic_cdf(-f), c1 = quartic_cdf(1.0f - f);
	const uint32_t lo = pack_half2_bits(c0, c1 - c0), hi = pack_half2_bits(1.0f - c1, 0.0f);
	// v_cvt_i32_f32 saturates at the int32 range and gives 0 for a NaN, so k is in [-2, 5]
	// for every input; the helper that states this conversion in asm (cvt_i32_sat) costs
	// this kernel spills and 1.6 % of its time, so the plain conversion stays here
	int k = (int)j;
	k = k < -2 ? -2 : (k > 5 ? 5 : k);
	const uint2 s = sel[k + 2];
	return make_uint2(__builtin_amdgcn_perm(hi, lo, s.x), __builtin_amdgcn_perm(hi, lo, s.y));
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
// PACKED: the one-blob quads come from oneblob4_packed through Emit.bits(p, halves)
// with the selector table Emit.sel (the FP16 kernels); otherwise from the five knots.
template <int ENC, bool PACKED = false, typename Emit> __device__ __forceinline__ void encode_quads(const float *rec, Emit &&emit) {
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
		if constexpr (PACKED) {
#pragma unroll
			for (int q = 0; q < 5; ++q) emit.bits(9 + q, oneblob4_packed(q == 4 ? 1.0f - expf(-rec[7]) : rec[3 + q], emit.sel));
		} else {
			float b[4];
#pragma unroll
			for (int q = 0; q < 5; ++q) {
				oneblob4_knots(q == 4 ? 1.0f - expf(-rec[7]) : rec[3 + q], b);
				emit(9 + q, b[0], b[1], b[2], b[3]);
			}
		}
		emit(14, rec[8], rec[9], rec[10], rec[11]);
		emit(15, rec[12], rec[13], 1.0f, 1.0f);
	}
}


// The same quads in pieces (the one-blob quads packed, as encode_quads<ENC, true>), so
// that the encoding of one tile can be spread over the layers of another: piece P of
// encode_pieces<ENC>() emits
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
		emit.bits(6 + P, oneblob4_packed(P == 7 ? 1.0f - expf(-rec[7]) : rec[P], emit.sel));
	} else {
		emit(14, rec[8], rec[9], rec[10], rec[11]);
		emit(15, rec[12], rec[13], 1.0f, 1.0f);
	}
}

// Writes feature quads of lane (et, es)'s record into a wave's LDS slot in the
// B-fragment layout: quad p holds features 4p..4p+3, K-chunk p >> 3, lane group
// p & 3, half (p >> 2) & 1 of the lane's 8 elements.
struct SlotEmit {
	half4_t *s4;
	int et, es;
	const uint2 *sel; // the one-blob selector table
	__device__ __forceinline__ int at(int p) const { return (((et * 2 + (p >> 3)) * 64 + (p & 3) * 16 + es) * 2 + ((p >> 2) & 1)); }
	__device__ __forceinline__ void operator()(int p, float f0, float f1, float f2, float f3) const {
		half4_t h;
		h[0] = (_Float16)f0; h[1] = (_Float16)f1; h[2] = (_Float16)f2; h[3] = (_Float16)f3;
		s4[at(p)] = h;
	}
	__device__ __forceinline__ void bits(int p, uint2 v) const { s4[at(p)] = __builtin_bit_cast(half4_t, v); }
};

// feature index of element j of a B operand held by lane group g for K-chunk c
__device__ __forceinline__ int bfeat(int c, int g, int j) { return 32 * c + (j < 4 ? 4 * g + j : 16 + 4 * g + j - 4); }

} // namespace
} // namespace nrc
