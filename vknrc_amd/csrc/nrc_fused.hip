// nrc_fused.hip -- fully fused 64-wide MLP on gfx950 FP16 MFMA (FP32 accumulation):
// the grouped inference and the grouped gradient kernels.  The MFMA operand
// layouts and the K permutation that chains the layers are described in
// nrc_mfma.hpp.
//
// Weight layout in HBM (unchanged from the reference, NN_nv.glsl:55-57): the
// fp16 weights of layer l, row-major [out][in], at l*4096; the 3x64 output
// layer at H*4096 (padded to 16 rows with zeros inside the MFMA, NN_nv.glsl:77-82).
// The kernels read the fragment-major copies of nrc_device.hpp.
#include <cstddef>
#include <cstdlib>
#include <type_traits>

#include "nrc_mfma.hpp"

#include <atomic>

namespace nrc {

namespace {

// ================================================================ inference
//
// Grouped and persistent.  The caches of a group lay their 64-record tiles end to
// end, and wave gw of W takes the contiguous tile range [T gw / W, T (gw + 1) / W).
// For each cache segment of its range a wave loads that cache's A fragments
// (the fragment-major use-weights, one 16-byte load per fragment and lane) into
// VGPRs, then streams the segment's tiles: lane i encodes record i of the tile
// straight into the B-fragment layout of its wave's LDS slot, and the 6 layers
// run as 4 N-tiles of 16 records, TILES at a time, activations in VGPRs, while
// the next tile's records are in flight.  Waves never wait on each other.
template <int H, int ENC, bool F16IN, int TILES>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void fused_infer_kernel(const InferGroup grp) {
	__shared__ __attribute__((aligned(16))) half8_t sX[4][8 * 64];
	__shared__ uint2 sSel[4][8]; // each wave's one-blob selector table

	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	oneblob_sel_init(sSel[wave], lane);
	wave_lds_sync();
	const int g = lane >> 4, r = lane & 15;
	const uint32_t nW = gridDim.x * 4, gw = blockIdx.x * 4 + wave;
	const uint32_t t0 = (uint32_t)((uint64_t)grp.total_tiles * gw / nW);
	const uint32_t t1 = (uint32_t)((uint64_t)grp.total_tiles * (gw + 1) / nW);
	half8_t *sx = sX[wave];
	half4_t *sx4 = reinterpret_cast<half4_t *>(sx);
	const int et = lane >> 4, es = lane & 15; // this lane's record: tile et, column es

	for (uint32_t i = 0; i < grp.n_items; ++i) {
		const InferItem &it = grp.item[i];
		const uint32_t b0 = it.tile_begin, b1 = b0 + (it.max_count + 63) / 64;
		if (b1 <= t0) continue;
		if (b0 >= t1) break;
		const uint32_t count = effective_count(it.d_count, it.max_count, it.max_count);
		const uint32_t lt0 = (t0 > b0 ? t0 : b0) - b0;
		uint32_t lt1 = (t1 < b1 ? t1 : b1) - b0;
		const uint32_t ntiles = (count + 63) / 64;
		lt1 = lt1 < ntiles ? lt1 : ntiles;
		if (lt0 >= lt1) continue;

		const half8_t *wv = reinterpret_cast<const half8_t *>(it.wf);
		half8_t A[H][4][2], Ao[2];
#pragma unroll
		for (int l = 0; l < H; ++l)
#pragma unroll
			for (int mt = 0; mt < 4; ++mt)
#pragma unroll
				for (int c = 0; c < 2; ++c) A[l][mt][c] = wv[((l * 4 + mt) * 2 + c) * 64 + lane];
#pragma unroll
		for (int c = 0; c < 2; ++c) Ao[c] = wv[(H * 8 + c) * 64 + lane]; // rows 3..15 are zero

		const float *queries = it.queries;
		float *out = it.out;
		// the record of this lane in local tile t.  The index is clamped into
		// [0, count) instead of branching, so the load is unconditional (lanes past
		// the count encode zeros regardless); a branch here makes the wait-count
		// pass drain every outstanding load at the join.
		// Prefetches past the wave's range re-read its last tile instead of the next
		// wave's first one.
		const uint32_t last = count - 1, tlast = lt1 - 1;
		auto load_query = [&](uint32_t t, float *dst) __attribute__((always_inline)) {
			t = t < tlast ? t : tlast;
			uint32_t m = t * 64 + lane;
			m = m < last ? m : last;
			const float2 *qp = reinterpret_cast<const float2 *>(queries + (size_t)m * kQueryFloats);
#pragma unroll
			for (int k = 0; k < 7; ++k) {
				const float2 v = qp[k];
				dst[2 * k] = v.x;
				dst[2 * k + 1] = v.y;
			}
		};
		// encode one record per lane straight into the B-fragment layout of this wave's slot
		auto encode_tile = [&](uint32_t base, const float *rec) __attribute__((always_inline)) {
			const uint32_t n = base + lane;
			const SlotEmit emit{sx4, et, es, sSel[wave]};
			if constexpr (F16IN) {
				if (n < count) {
					const half4_t *src = reinterpret_cast<const half4_t *>(grp.fin + (size_t)n * kWidth);
#pragma unroll
					for (int p = 0; p < 16; ++p) {
						const int c = p >> 3, gg = p & 3, hi = (p >> 2) & 1;
						sx4[((et * 2 + c) * 64 + gg * 16 + es) * 2 + hi] = src[p];
					}
				} else {
#pragma unroll
					for (int p = 0; p < 16; ++p) emit(p, 0.0f, 0.0f, 0.0f, 0.0f);
				}
			} else {
				if (n < count) {
					encode_quads<ENC, true>(rec, emit);
				} else {
#pragma unroll
					for (int p = 0; p < 16; ++p) emit(p, 0.0f, 0.0f, 0.0f, 0.0f);
				}
			}
			wave_lds_sync();
		};
		// the 6 layers over the 4 N-tiles of the staged tile, TILES at a time
		auto mlp_tile = [&](uint32_t base) __attribute__((always_inline)) {
#pragma unroll
			for (int tg = 0; tg < 4; tg += TILES) {
				half8_t B[TILES][2];
#pragma unroll
				for (int tt = 0; tt < TILES; ++tt)
#pragma unroll
					for (int c = 0; c < 2; ++c) B[tt][c] = sx[((tg + tt) * 2 + c) * 64 + lane];
#pragma unroll
				for (int l = 0; l < H; ++l) {
					float4_t acc[TILES][4];
#pragma unroll
					for (int tt = 0; tt < TILES; ++tt)
#pragma unroll
						for (int mt = 0; mt < 4; ++mt) acc[tt][mt] = (float4_t)0.0f;
#pragma unroll
					for (int c = 0; c < 2; ++c)
#pragma unroll
						for (int mt = 0; mt < 4; ++mt)
#pragma unroll
							for (int tt = 0; tt < TILES; ++tt) acc[tt][mt] = mfma(A[l][mt][c], B[tt][c], acc[tt][mt]);
#pragma unroll
					for (int tt = 0; tt < TILES; ++tt)
#pragma unroll
						for (int c = 0; c < 2; ++c) B[tt][c] = pack_relu(acc[tt][2 * c], acc[tt][2 * c + 1]);
				}
				float4_t o[TILES];
#pragma unroll
				for (int tt = 0; tt < TILES; ++tt) o[tt] = (float4_t)0.0f;
#pragma unroll
				for (int c = 0; c < 2; ++c)
#pragma unroll
					for (int tt = 0; tt < TILES; ++tt) o[tt] = mfma(Ao[c], B[tt][c], o[tt]);
				if (g == 0) {
#pragma unroll
					for (int tt = 0; tt < TILES; ++tt) {
						const uint32_t m = base + (tg + tt) * 16 + r;
						if (m < count) {
							float *dst = out + (size_t)m * kOut;
							// the clamp as the pipelined kernel applies it: an integer max of the bits
							// (sign bit set -> +0), so both kernels give the same bits
							const int lo_bits = grp.clamp0 ? 0 : INT_MIN;
#pragma unroll
							for (int c = 0; c < kOut; ++c) {
								const int b = __float_as_int(o[tt][c]);
								dst[c] = __int_as_float(b > lo_bits ? b : lo_bits);
							}
						}
					}
				}
			}
			wave_lds_sync();
		};

		if constexpr (F16IN) {
			for (uint32_t t = lt0; t < lt1; ++t) {
				encode_tile(t * 64, nullptr);
				mlp_tile(t * 64);
			}
		} else {
			// two record buffers in ping-pong: the next tile's records are in flight
			// while this tile is encoded and run, and no register copy waits on them
			float ra[kQueryFloats], rb[kQueryFloats];
			load_query(lt0, ra);
			for (uint32_t t = lt0; t < lt1; t += 2) {
				load_query(t + 1, rb);
				encode_tile(t * 64, ra);
				mlp_tile(t * 64);
				if (t + 1 >= lt1) break;
				load_query(t + 2, ra);
				encode_tile((t + 1) * 64, rb);
				mlp_tile((t + 1) * 64);
			}
		}
	}
}


// The same inference, software-pipelined: a wave encodes tile t+1 into one LDS
// slot while the layers of tile t run from the other, and nothing between two
// slot switches branches (records past the count encode the last record, and
// the outputs go out through buffer stores whose range check drops the lanes
// that hold no output and the records past the count), so the compiler can
// interleave the encoding's VALU work with the MFMA chains of the 4 N-tiles.
typedef unsigned int uint3_t __attribute__((ext_vector_type(3)));

template <int H, int ENC>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void fused_infer_pipe_kernel(const InferGroup grp) {
	// two separate arrays, so the compiler knows the encoding's stores into one
	// slot never alias the layers' loads from the other
	__shared__ __attribute__((aligned(16))) half8_t sX0[4][8 * 64];
	__shared__ __attribute__((aligned(16))) half8_t sX1[4][8 * 64];
	__shared__ uint2 sSel[4][8]; // each wave's one-blob selector table

	// the wave index and everything derived from it are wave-uniform: in SGPRs
	const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
	oneblob_sel_init(sSel[wave], lane);
	wave_lds_sync();
	const int g = lane >> 4, r = lane & 15;
	const uint32_t nW = gridDim.x * 4, gw = blockIdx.x * 4 + wave;
	const uint32_t t0 = (uint32_t)((uint64_t)grp.total_tiles * gw / nW);
	const uint32_t t1 = (uint32_t)((uint64_t)grp.total_tiles * (gw + 1) / nW);
	half8_t *slot0 = sX0[wave], *slot1 = sX1[wave];
	const int et = lane >> 4, es = lane & 15; // this lane's record: N-tile et, column es
	// the output clamp max(o, 0) as a signed integer max of the bits: every value with
	// the sign bit set (negative, -0, negative NaN) becomes +0, the rest pass (a
	// floating-point max would first canonicalize the MFMA result, an extra instruction
	// per value); without the clamp the floor INT_MIN passes every value
	const int lo_bits = grp.clamp0 ? 0 : INT_MIN;

	for (uint32_t i = 0; i < grp.n_items; ++i) {
		const InferItem &it = grp.item[i];
		const uint32_t b0 = it.tile_begin, b1 = b0 + (it.max_count + 63) / 64;
		if (b1 <= t0) continue;
		if (b0 >= t1) break;
		const uint32_t count = __builtin_amdgcn_readfirstlane(effective_count(it.d_count, it.max_count, it.max_count));
		const uint32_t lt0 = (t0 > b0 ? t0 : b0) - b0;
		uint32_t lt1 = (t1 < b1 ? t1 : b1) - b0;
		const uint32_t ntiles = (count + 63) / 64;
		lt1 = lt1 < ntiles ? lt1 : ntiles;
		if (lt0 >= lt1) continue;

		const half8_t *wv = reinterpret_cast<const half8_t *>(it.wf);
		half8_t A[H][4][2], Ao[2];
#pragma unroll
		for (int l = 0; l < H; ++l)
#pragma unroll
			for (int mt = 0; mt < 4; ++mt)
#pragma unroll
				for (int c = 0; c < 2; ++c) A[l][mt][c] = wv[((l * 4 + mt) * 2 + c) * 64 + lane];
#pragma unroll
		for (int c = 0; c < 2; ++c) Ao[c] = wv[(H * 8 + c) * 64 + lane]; // rows 3..15 are zero

		const float *queries = it.queries;
		float *out = it.out;
		const uint32_t last = count - 1, tlast = lt1 - 1;
		auto load_query = [&](uint32_t t, float *dst) __attribute__((always_inline)) {
			t = t < tlast ? t : tlast;
			uint32_t m = t * 64 + lane;
			m = m < last ? m : last;
			const float2 *qp = reinterpret_cast<const float2 *>(queries + (size_t)m * kQueryFloats);
#pragma unroll
			for (int k = 0; k < 7; ++k) {
				const float2 v = qp[k];
				dst[2 * k] = v.x;
				dst[2 * k + 1] = v.y;
			}
		};
		auto emitter = [&](half8_t *slot) __attribute__((always_inline)) {
			return SlotEmit{reinterpret_cast<half4_t *>(slot), et, es, sSel[wave]};
		};
		constexpr int NP = encode_pieces<ENC>();
		// piece k of the next tile's encoding goes after the MFMAs of step s = k * steps / NP
		auto piece_at = [&](auto step_c, const float *rec, half8_t *slot) __attribute__((always_inline)) {
			constexpr int step = decltype(step_c)::value;
			constexpr int steps = 4 * (H + 1);
			static_for<NP>([&](auto k_c) __attribute__((always_inline)) {
				constexpr int k = decltype(k_c)::value;
				if constexpr (k * steps / NP == step) encode_piece<ENC, k>(rec, emitter(slot));
			});
		};
		// the layers of the tile staged in `cur`, with the encoding of the next
		// tile (records `rec`) into `nxt` spread between them
		auto run_tile = [&](const half8_t *cur, uint32_t tile, const float *rec, half8_t *nxt) __attribute__((always_inline)) {
			const uint32_t left = count - tile * 64;
			const uint32_t nrec = left < 64u ? left : 64u;
			const __amdgpu_buffer_rsrc_t rs =
			    __builtin_amdgcn_make_buffer_rsrc(out + (size_t)tile * 64 * kOut, (short)0, (int)(nrec * kOut * 4), 0x00020000);
			static_for<4>([&](auto nt_c) __attribute__((always_inline)) {
				constexpr int nt = decltype(nt_c)::value;
				half8_t B[2];
#pragma unroll
				for (int c = 0; c < 2; ++c) B[c] = cur[(nt * 2 + c) * 64 + lane];
				static_for<H>([&](auto l_c) __attribute__((always_inline)) {
					constexpr int l = decltype(l_c)::value;
					float4_t acc[4];
#pragma unroll
					for (int mt = 0; mt < 4; ++mt) acc[mt] = (float4_t)0.0f;
#pragma unroll
					for (int c = 0; c < 2; ++c)
#pragma unroll
						for (int mt = 0; mt < 4; ++mt) acc[mt] = mfma(A[l][mt][c], B[c], acc[mt]);
					piece_at(std::integral_constant<int, nt * (H + 1) + l>{}, rec, nxt);
#pragma unroll
					for (int c = 0; c < 2; ++c) B[c] = pack_relu(acc[2 * c], acc[2 * c + 1]);
				});
				float4_t o = (float4_t)0.0f;
#pragma unroll
				for (int c = 0; c < 2; ++c) o = mfma(Ao[c], B[c], o);
				piece_at(std::integral_constant<int, nt * (H + 1) + H>{}, rec, nxt);
				uint3_t v;
#pragma unroll
				for (int e = 0; e < 3; ++e) {
					const int b = __float_as_int(o[e]);
					v[e] = (uint32_t)(b > lo_bits ? b : lo_bits);
				}
				const uint32_t off = g == 0 ? (uint32_t)((nt * 16 + r) * kOut * 4) : 0x7ffffff0u;
				__builtin_amdgcn_raw_buffer_store_b96(v, rs, off, 0, 0);
			});
		};

		// records: ra holds tile t (or t+2), rb tile t+1 (or t+3)
		float ra[kQueryFloats], rb[kQueryFloats];
		load_query(lt0, ra);
		load_query(lt0 + 1, rb);
		encode_quads<ENC, true>(ra, emitter(slot0));
		wave_lds_sync();
		for (uint32_t t = lt0; t < lt1; t += 2) {
			load_query(t + 2, ra);
			run_tile(slot0, t, rb, slot1); // encodes tile t+1 (a repeat of the last tile past the range)
			wave_lds_sync();
			if (t + 1 >= lt1) break;
			load_query(t + 3, rb);
			run_tile(slot1, t + 1, ra, slot0); // encodes tile t+2
			wave_lds_sync();
		}
	}
}

// ================================================================ gradient
//
// One workgroup = 4 waves (one per SIMD) walking `chunks` 64-record chunks;
// wave w runs records 16w .. 16w+15 of a chunk through the whole network.  The
// forward A fragments (W_l) sit in registers, 168 for H = 5, loaded once from the
// fragment-major copy (through LDS in a one-chunk workgroup); the backward ones
// (W_l^T) are copied once per workgroup into LDS with direct global->LDS loads.
// Waves write their activations a_l and deltas record-major into LDS as they
// produce them, so the forward and the backward chain run without any barrier
// between waves; the backward takes its ReLU masks from registers.  After the
// one barrier, wave w computes its 2x2 block of dW_l^T tiles (input tiles
// 2(w&1).., output tiles 2(w>>1)..) for every layer from LDS, K = the 64 records;
// a multi-chunk workgroup sums them in registers over its chunks, and the tiles
// go (float4 per lane, non-temporal) to the workgroup's partial-gradient row.
//
// LDS rows are records: a lane stores its B operand (its record's 16 features of
// the operand, two runs of 4 consecutive features a K-chunk) with plain 8-byte
// stores, and the dW operands, which need 8 consecutive records of one feature a
// lane, come back through transposing reads (ds_read_b64_tr_b16).  A row is 64
// halves with no padding; its 16 four-half chunks are permuted by a function of
// the record (row_swizzle) under which the 8-byte row stores of 16 consecutive
// records and the transposing reads of 4 records x 4 chunks hit distinct banks
// (with 32 banks and 16 lanes a pass, and with 64 banks and 32 lanes a pass).
constexpr int kRowLd = kWidth; // halves per record row
constexpr int kGoLd = 16;      // halves per output-gradient row (units 3..15 are zero)

template <int H> struct SmallGradSmem {
	_Float16 wt[H * 8 * 512];                         // backward A fragments, fragment-major as in HBM
	_Float16 act[H + 1][kFusedSmallRecords * kRowLd]; // a_l, [record][feature], chunks swizzled
	_Float16 dlt[H][kFusedSmallRecords * kRowLd];     // delta_{l+1} = dL/dz at layer l's output, [record][unit], swizzled
	_Float16 go[kFusedSmallRecords * kGoLd];          // output gradient, [record][unit]
	float lsum[4];
	float gmax[4]; // largest |output gradient| of each wave's records
};

// chunk ch (features 4ch .. 4ch+3) of record rec is chunk ch ^ row_swizzle(rec) of its
// row: bits 0, 1 of rec & 15 as they are, bit 2 = rec bit 1 ^ bit 2, bit 3 = rec bit 0 ^ bit 3
__device__ __forceinline__ int row_swizzle(int rec) {
	rec &= 15;
	return rec ^ ((rec & 2) << 1) ^ ((rec & 1) << 3);
}
// halves offset of features feat .. feat+3 (feat a multiple of 4) of record rec
__device__ __forceinline__ int row_at(int rec, int feat) { return rec * kRowLd + (((feat >> 2) ^ row_swizzle(rec)) << 2); }

// Output gradients at least this large (2^13) get a power-of-two scale that
// brings the workgroup's largest one below it, leaving 8x headroom under the
// FP16 maximum for the growth of the deltas through the layers.
constexpr float kDeltaLimit = 8192.0f;

// writes this lane's B operand (record col: features bfeat(c, g, 0..3) =
// 32c + 4g + 0..3 and bfeat(c, g, 4..7) = 32c + 16 + 4g + 0..3) into its row
__device__ __forceinline__ void store_brow(_Float16 *rows, const half8_t *b, int g, int col) {
#pragma unroll
	for (int c = 0; c < 2; ++c) {
		*reinterpret_cast<half4_t *>(rows + row_at(col, 32 * c + 4 * g)) = b[c].lo;
		*reinterpret_cast<half4_t *>(rows + row_at(col, 32 * c + 16 + 4 * g)) = b[c].hi;
	}
}

// reads back this lane's B operand (record col) written by store_brow
__device__ __forceinline__ void load_brow(const _Float16 *rows, half8_t *b, int g, int col) {
#pragma unroll
	for (int c = 0; c < 2; ++c) {
		const half4_t lo = *reinterpret_cast<const half4_t *>(rows + row_at(col, 32 * c + 4 * g));
		const half4_t hi = *reinterpret_cast<const half4_t *>(rows + row_at(col, 32 * c + 16 + 4 * g));
		b[c] = __builtin_shufflevector(lo, hi, 0, 1, 2, 3, 4, 5, 6, 7);
	}
}

// A K-operand of a dW MFMA (K = records): lane (g, r) gets X[record 32 kc + 8 g + k]
// [feature f0 + r], k = 0..7, from the swizzled record rows (SWZ) or from plain rows
// of stride LDR.  ds_read_b64_tr_b16 gives lane i of each 16-lane group element i % 4
// of the 4 halves that lane 4 k + i / 4 of the group addresses, as element k
// (k = 0..3); so lane j of group g addresses features f0 + 4 (j & 3) .. + 3 of record
// 32 kc + 8 g + (j >> 2), and again 4 records on for elements 4..7.
template <bool SWZ, int LDR>
__device__ __forceinline__ half8_t load_kop(const _Float16 *rows, int f0, int kc, int g, int r) {
	typedef short short4_t __attribute__((ext_vector_type(4)));
	typedef __attribute__((address_space(3))) short4_t lds_short4;
	const int rec = 32 * kc + 8 * g + (r >> 2), feat = f0 + 4 * (r & 3);
	const _Float16 *p0 = rows + (SWZ ? row_at(rec, feat) : rec * LDR + feat);
	const _Float16 *p1 = rows + (SWZ ? row_at(rec + 4, feat) : (rec + 4) * LDR + feat);
	const half4_t lo = __builtin_bit_cast(half4_t, __builtin_amdgcn_ds_read_tr16_b64_v4i16((lds_short4 *)p0));
	const half4_t hi = __builtin_bit_cast(half4_t, __builtin_amdgcn_ds_read_tr16_b64_v4i16((lds_short4 *)p1));
	return __builtin_shufflevector(lo, hi, 0, 1, 2, 3, 4, 5, 6, 7);
}

typedef short short2_t __attribute__((ext_vector_type(2)));

// The ReLU masks of a B operand of activations: all ones in each half whose fp16
// activation is > 0.  0 - a saturates, so +0 and -0 (0x8000) both give 0 and
// every positive value (finite or +Inf) a negative number; ReLU outputs are
// never NaN (the pack's max returns its non-NaN operand).
__device__ __forceinline__ uint4 act_masks(half8_t a) {
	const uint4 w = __builtin_bit_cast(uint4, a);
	auto m = [](uint32_t x) __attribute__((always_inline)) {
		const short2_t t = __builtin_elementwise_sub_sat((short2_t)0, __builtin_bit_cast(short2_t, x));
		return __builtin_bit_cast(uint32_t, (short2_t)(t >> 15));
	};
	return make_uint4(m(w.x), m(w.y), m(w.z), m(w.w));
}

// relu_bwd of two values packed to fp16: (fp16)x where the mask half is set and
// the value is not NaN, else +0 (fp16 keeps NaN a NaN and takes values beyond
// its range to +-Inf, so the NaN test is made on the converted halves: a half
// is NaN when its magnitude bits exceed 0x7C00, i.e. when 0x7C00 - magnitude,
// which stays within int16, is negative)
__device__ __forceinline__ uint32_t relu_bwd2(float x, float y, uint32_t mask) {
	const float2_t f = {x, y};
	const uint32_t h = __builtin_bit_cast(uint32_t, __builtin_convertvector(f, half2_t));
	short2_t mag = __builtin_bit_cast(short2_t, h & 0x7FFF7FFFu);
	asm("" : "+v"(mag)); // keeps the test in packed 16-bit arithmetic
	const short2_t nan = ((short2_t)0x7C00 - mag) >> 15; // all ones in a NaN half
	return h & mask & ~__builtin_bit_cast(uint32_t, nan);
}

// delta = relu_bwd(acc) against the ReLU masks of the activations (B-operand layout), packed to fp16
__device__ __forceinline__ void relu_bwd_pack(const float4_t *acc, const uint4 *mask, half8_t *d) {
#pragma unroll
	for (int c = 0; c < 2; ++c) {
		const uint4 o = make_uint4(relu_bwd2(acc[2 * c][0], acc[2 * c][1], mask[c].x),
		                           relu_bwd2(acc[2 * c][2], acc[2 * c][3], mask[c].y),
		                           relu_bwd2(acc[2 * c + 1][0], acc[2 * c + 1][1], mask[c].z),
		                           relu_bwd2(acc[2 * c + 1][2], acc[2 * c + 1][3], mask[c].w));
		d[c] = __builtin_bit_cast(half8_t, o);
	}
}

template <int H, int ENC, int LOSS, bool F16IN, bool MULTI>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 1))) void fused_grad_small_kernel(
    const GradGroup grp) {
	// this workgroup's cache: the last item whose first block is at or below blockIdx.x
	// (block_begin does not decrease, so k counts the later items starting at or below
	// blockIdx.x; the unrolled count issues every table load at once)
	uint32_t k = 0;
#pragma unroll
	for (uint32_t j = 1; j < (uint32_t)kGroupMax; ++j)
		k += (j < grp.n_items && blockIdx.x >= grp.item[j].block_begin) ? 1u : 0u;
	const GradItem &gi = grp.item[k];
	const uint32_t blk = blockIdx.x - gi.block_begin;
	const float *__restrict__ records = gi.records;
	const _Float16 *__restrict__ fin = grp.fin;
	const float *__restrict__ ftgt = grp.ftgt;
	const uint32_t *__restrict__ d_count = gi.d_count;
	const uint32_t max_count = gi.max_count, clamp = gi.clamp, chunks = gi.chunks, pstride = gi.pstride;
	const _Float16 *__restrict__ wf = gi.wf;
	const _Float16 *__restrict__ wtf = gi.wtf;
	const DeviceState *__restrict__ st = gi.state;
	const int parity = gi.parity;
	float *__restrict__ partials = gi.partials;
	uint32_t *__restrict__ overflow = gi.overflow;
	constexpr int WC = weight_count(H);
	constexpr int R = (int)kFusedSmallRecords;
	extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
	SmallGradSmem<H> &S = *reinterpret_cast<SmallGradSmem<H> *>(smem_raw);

	const int tid = threadIdx.x, wave = __builtin_amdgcn_readfirstlane(tid >> 6); // wave-uniform: in an SGPR
	int lane = tid & 63;
	int g = lane >> 4, r = lane & 15;
	int col = 16 * wave + r; // record column this lane runs through the network
	const uint32_t count = F16IN ? max_count : effective_count(d_count, max_count, clamp);
	const uint32_t nch = MULTI ? chunks : 1u;
	const uint32_t base0 = blk * nch * (uint32_t)R;
	float *part = partials + (size_t)blk * pstride;
	const int it0 = 2 * (wave & 1), ot0 = 2 * (wave >> 1); // this wave's dW tiles

	// The partial row is written with non-temporal stores and the optimizer reads it
	// with non-temporal loads: 21 MB a step that no kernel re-reads through L2.
	// Together they take 2.6 us off a 16,384-record step (either alone, nothing).
	bool bad = false;
	auto store_tile = [&](int l, int it, int ot, float4_t v) __attribute__((always_inline)) {
		__builtin_nontemporal_store(v, reinterpret_cast<float4_t *>(part + l * kWidth * kWidth + (16 * ot + r) * kWidth + 16 * it + 4 * g));
#pragma unroll
		for (int e = 0; e < 4; ++e) bad |= !isfinite(v[e]);
	};
	auto store_out_tile = [&](float4_t v) __attribute__((always_inline)) {
		if (r < kOut) {
			__builtin_nontemporal_store(v, reinterpret_cast<float4_t *>(part + H * kWidth * kWidth + r * kWidth + 16 * wave + 4 * g));
#pragma unroll
			for (int e = 0; e < 4; ++e) bad |= !isfinite(v[e]);
		}
	};
	if (base0 >= count) { // uniform: no valid record in this workgroup, so its partial row is zero
#pragma unroll
		for (int l = 0; l < H; ++l)
#pragma unroll
			for (int i = 0; i < 2; ++i)
#pragma unroll
				for (int o = 0; o < 2; ++o) store_tile(l, it0 + i, ot0 + o, (float4_t)0.0f);
		store_out_tile((float4_t)0.0f);
		if (tid == 0) part[WC] = 0.0f;
		return;
	}
	const float scale = (F16IN || st == nullptr) ? 1.0f : st->pub[parity].loss_scale;
	const uint32_t last = count - 1;

	// ---- the first chunk's records, then every weight fragment, in the order needed.
	// Indices are clamped so the loads are unconditional.
	float q[kQueryFloats];
	uint4 fq[2];
	float tg[3];
	// the inputs this lane encodes and the targets of the record it runs, loaded
	// separately: a multi-chunk workgroup loads the next chunk's inputs once the
	// current ones are encoded and its targets once the loss is taken, so both
	// loads are in flight during the rest of the chunk
	auto load_inputs = [&](uint32_t base) __attribute__((always_inline)) {
		const uint32_t n_enc = base + lane < last ? base + lane : last; // the record this lane encodes
		if constexpr (F16IN) {
			const uint4 *src = reinterpret_cast<const uint4 *>(fin + (size_t)n_enc * kWidth + 16 * wave);
			fq[0] = src[0];
			fq[1] = src[1];
		} else {
			const float *rp = records + (size_t)n_enc * kTrainFloats;
#pragma unroll
			for (int i = 0; i < kQueryFloats; ++i) q[i] = rp[i];
		}
	};
	auto load_targets = [&](uint32_t base) __attribute__((always_inline)) {
		const uint32_t n_run = base + col < last ? base + col : last; // the record this lane runs
		if constexpr (F16IN) {
#pragma unroll
			for (int c = 0; c < 3; ++c) tg[c] = ftgt[(size_t)n_run * 3 + c];
		} else {
			const float *tp = records + (size_t)n_run * kTrainFloats + kQueryFloats;
#pragma unroll
			for (int c = 0; c < 3; ++c) tg[c] = tp[c];
		}
	};
	load_inputs(base0);
	load_targets(base0);
	// the backward fragments go to LDS once per workgroup: a straight copy of 1 KB
	// blocks, wave-instruction k putting lane i at byte 1024 k + 16 i
	for (int k = wave; k < H * 8; k += 4)
		__builtin_amdgcn_global_load_lds(reinterpret_cast<const void *>(wtf + (k * 64 + lane) * 8),
		                                 (__attribute__((address_space(3))) void *)(S.wt + k * 512), 16, 0, 0);
	// The forward fragments stay in registers.  A one-chunk workgroup fetches them
	// once into the delta buffers and the output-gradient rows after them (first
	// written by the loss and the backward, past the barrier that every wave reaches
	// after its forward has consumed the fragments) and each wave reads them from
	// there after the first barrier, so the workgroup reads them from L2 once
	// rather than once a wave; a multi-chunk workgroup, which holds them across its
	// chunks, loads them straight into registers.
	static_assert(offsetof(SmallGradSmem<H>, go) == offsetof(SmallGradSmem<H>, dlt) + sizeof(S.dlt) &&
	                  sizeof(S.dlt) + sizeof(S.go) >= (H * 8 + 2) * 1024,
	              "the fragment staging area is the delta buffers and the output-gradient rows");
	_Float16 *const stage = &S.dlt[0][0];
	const half8_t *wfv = reinterpret_cast<const half8_t *>(wf);
	half8_t Af[H][4][2], Ao[2];
	auto fragments_from = [&](auto frag) __attribute__((always_inline)) {
#pragma unroll
		for (int l = 0; l < H; ++l)
#pragma unroll
			for (int mt = 0; mt < 4; ++mt)
#pragma unroll
				for (int c = 0; c < 2; ++c) Af[l][mt][c] = frag((l * 4 + mt) * 2 + c);
#pragma unroll
		for (int c = 0; c < 2; ++c) Ao[c] = frag(H * 8 + c);
	};
	if constexpr (MULTI) {
		fragments_from([&](int k) { return wfv[k * 64 + lane]; });
	} else {
		for (int k = wave; k < H * 8 + 2; k += 4)
			__builtin_amdgcn_global_load_lds(reinterpret_cast<const void *>(wf + (k * 64 + lane) * 8),
			                                 (__attribute__((address_space(3))) void *)(stage + k * 512), 16, 0, 0);
	}
	auto wt_frag = [&](int blk) __attribute__((always_inline)) {
		return *reinterpret_cast<const half8_t *>(S.wt + blk * 512 + lane * 8);
	};

	// the loss sums start at zero
	if (tid < 4) S.lsum[tid] = 0.0f;

	float4_t accW[MULTI ? H : 1][2][2], accO = (float4_t)0.0f; // dW tiles summed over the chunks
#pragma unroll
	for (int l = 0; l < (MULTI ? H : 1); ++l)
#pragma unroll
		for (int i = 0; i < 2; ++i) accW[l][i][0] = accW[l][i][1] = (float4_t)0.0f;

	for (uint32_t ch = 0; ch < nch; ++ch) {
		const uint32_t base = base0 + ch * (uint32_t)R;
		if (base >= count) break; // uniform
		if (ch > 0) block_lds_sync(); // the previous chunk's dW readers are done with the LDS buffers
		if constexpr (MULTI) {
			// recompute the lane-derived indices each chunk from an opaque copy of the
			// lane id, so the LDS addresses built from them are not kept live across chunks
			asm volatile("" : "+v"(lane));
			g = lane >> 4;
			r = lane & 15;
			col = 16 * wave + r;
		}
		// ---- encode: wave w computes feature block w (features 16w ..) of record base + lane
		{
			_Float16 *a0 = S.act[0];
			const bool valid = base + lane < count;
			if constexpr (F16IN) {
#pragma unroll
				for (int h = 0; h < 2; ++h) {
					const half8_t v = valid ? __builtin_bit_cast(half8_t, fq[h]) : (half8_t)0;
					*reinterpret_cast<half4_t *>(a0 + row_at(lane, 16 * wave + 8 * h)) = v.lo;
					*reinterpret_cast<half4_t *>(a0 + row_at(lane, 16 * wave + 8 * h + 4)) = v.hi;
				}
			} else {
				auto put = [&](auto qc) __attribute__((always_inline)) {
					constexpr int qb = decltype(qc)::value;
					float f[16];
					encode_block<ENC>(q, qb, f);
					half4_t h[4];
#pragma unroll
					for (int j = 0; j < 16; ++j) h[j >> 2][j & 3] = valid ? (_Float16)f[j] : (_Float16)0.0f;
#pragma unroll
					for (int k = 0; k < 4; ++k) *reinterpret_cast<half4_t *>(a0 + row_at(lane, 16 * qb + 4 * k)) = h[k];
				};
				switch (wave) {
				case 0: put(std::integral_constant<int, 0>{}); break;
				case 1: put(std::integral_constant<int, 1>{}); break;
				case 2: put(std::integral_constant<int, 2>{}); break;
				default: put(std::integral_constant<int, 3>{}); break;
				}
			}
		}
		if (ch == 0) __builtin_amdgcn_s_waitcnt(0); // this wave's fragment copies to LDS have landed
		block_lds_sync();
		if constexpr (!MULTI)
			fragments_from([&](int k) { return *reinterpret_cast<const half8_t *>(stage + k * 512 + lane * 8); });
		if constexpr (MULTI)
			if (ch + 1 < nch) load_inputs(base + R); // the next chunk's inputs (uniform branch)

		// ---- forward for this wave's 16 records; a_l kept as B operands and written to its record rows
		// (the activations are not kept in registers for the backward: its ReLU masks
		// come back from the LDS copy, which leaves room for the weight fragments)
		half8_t B[2];
		uint4 keep[H][2]; // ReLU masks of a_1 .. a_{H-1} in B layout, for the backward
		load_brow(S.act[0], B, g, col);
#pragma unroll
		for (int l = 0; l < H; ++l) {
			float4_t acc[4];
#pragma unroll
			for (int mt = 0; mt < 4; ++mt) acc[mt] = (float4_t)0.0f;
#pragma unroll
			for (int c = 0; c < 2; ++c)
#pragma unroll
				for (int mt = 0; mt < 4; ++mt) acc[mt] = mfma(Af[l][mt][c], B[c], acc[mt]);
#pragma unroll
			for (int c = 0; c < 2; ++c) B[c] = pack_relu(acc[2 * c], acc[2 * c + 1]);
			if (l + 1 < H) {
				keep[l + 1][0] = act_masks(B[0]);
				keep[l + 1][1] = act_masks(B[1]);
			}
			store_brow(S.act[l + 1], B, g, col);
		}
		float4_t o = (float4_t)0.0f;
#pragma unroll
		for (int c = 0; c < 2; ++c) o = mfma(Ao[c], B[c], o);

		// ---- loss gradient: lanes g == 0 hold the 3 outputs of record col
		float go[3] = {0.0f, 0.0f, 0.0f};
		float lv = 0.0f, gm = 0.0f;
		if (g == 0) {
			const bool valid = base + col < count;
			const float oo[3] = {o[0], o[1], o[2]};
			const float v = loss_grad<LOSS>(oo, tg, scale, go);
			lv = valid ? v : 0.0f;
#pragma unroll
			for (int c = 0; c < kOut; ++c) {
				go[c] = valid ? go[c] : 0.0f;
				gm = fmaxf(gm, fabsf(go[c]));
			}
		}
		if constexpr (MULTI)
			if (ch + 1 < nch) load_targets(base + R); // the next chunk's targets (uniform branch)
		// The deltas of the workgroup share one power-of-two scale 2^k: k = 0 unless
		// the largest output gradient reaches kDeltaLimit, where FP16 would overflow
		// and force a skipped step.  Scaling by 2^k is exact, and the dW tiles undo it
		// in FP32 when they are stored.
#pragma unroll
		for (int off = 32; off >= 1; off >>= 1) gm = fmaxf(gm, __shfl_xor(gm, off));
		if (lane == 0) S.gmax[wave] = gm;
		block_lds_sync();
		float dscale = 1.0f, undo = 1.0f;
		{
			const float wmax = fmaxf(fmaxf(S.gmax[0], S.gmax[1]), fmaxf(S.gmax[2], S.gmax[3]));
			if (wmax >= kDeltaLimit && wmax <= 3.0e38f) {
				const int k = 12 - ilogbf(wmax);
				dscale = ldexpf(1.0f, k);
				undo = ldexpf(1.0f, -k);
			}
		}
		half8_t Bg = (half8_t)0;
		if (g == 0) { // the record's whole row: units 0..2, then zeros
#pragma unroll
			for (int c = 0; c < kOut; ++c) Bg[c] = (_Float16)(go[c] * dscale);
			*reinterpret_cast<half8_t *>(S.go + col * kGoLd) = Bg;
			*reinterpret_cast<half8_t *>(S.go + col * kGoLd + 8) = (half8_t)0;
		}
		// A multi-chunk workgroup sums its dW tiles in the MFMA accumulators
		// themselves.  For a chunk whose deltas carry a scale 2^k, the sums are scaled
		// by 2^k before its dW and by 2^-k after (exact: powers of two), so the
		// scale costs nothing on the chunks that have none.
		auto scale_sums = [&](float s) __attribute__((always_inline)) {
			// kept behind its branch: the compiler would otherwise multiply on every
			// chunk, since x * 1 is x
			asm volatile("" ::: "memory");
#pragma unroll
			for (int l = 0; l < (MULTI ? H : 1); ++l)
#pragma unroll
				for (int i = 0; i < 2; ++i)
#pragma unroll
					for (int o2 = 0; o2 < 2; ++o2) accW[l][i][o2] *= s;
			accO *= s;
		};
		if constexpr (MULTI)
			if (undo != 1.0f) scale_sums(dscale); // uniform
		// dW_l^T tiles of this wave over the chunk's records (K = the 64 records, all
		// waves' a_l and delta): accumulated into accW by a multi-chunk workgroup,
		// stored at once by a one-chunk workgroup
		auto dw_layer = [&](int l) __attribute__((always_inline)) {
			float4_t t[2][2];
#pragma unroll
			for (int i = 0; i < 2; ++i)
#pragma unroll
				for (int o2 = 0; o2 < 2; ++o2) t[i][o2] = MULTI ? accW[MULTI ? l : 0][i][o2] : (float4_t)0.0f;
#pragma unroll
			for (int kc = 0; kc < R / 32; ++kc) {
				half8_t a[2], b[2];
#pragma unroll
				for (int i = 0; i < 2; ++i) {
					a[i] = load_kop<true, kRowLd>(S.act[l], 16 * (it0 + i), kc, g, r);
					b[i] = load_kop<true, kRowLd>(S.dlt[l], 16 * (ot0 + i), kc, g, r);
				}
#pragma unroll
				for (int i = 0; i < 2; ++i)
#pragma unroll
					for (int o2 = 0; o2 < 2; ++o2) t[i][o2] = mfma(a[i], b[o2], t[i][o2]);
			}
			if constexpr (!MULTI)
				if (undo != 1.0f) { // uniform; exact when taken
					asm volatile("" ::: "memory"); // keeps the multiplies behind the branch
#pragma unroll
					for (int i = 0; i < 2; ++i)
#pragma unroll
						for (int o2 = 0; o2 < 2; ++o2) t[i][o2] *= undo;
				}
#pragma unroll
			for (int i = 0; i < 2; ++i)
#pragma unroll
				for (int o2 = 0; o2 < 2; ++o2) {
					if constexpr (MULTI)
						accW[MULTI ? l : 0][i][o2] = t[i][o2];
					else
						store_tile(l, it0 + i, ot0 + o2, t[i][o2]);
				}
		};
		// dW_H^T: input tile `wave` against the output-gradient rows
		auto dw_output = [&]() __attribute__((always_inline)) {
			float4_t t = MULTI ? accO : (float4_t)0.0f;
#pragma unroll
			for (int kc = 0; kc < R / 32; ++kc) {
				const half8_t a = load_kop<true, kRowLd>(S.act[H], 16 * wave, kc, g, r);
				const half8_t b = load_kop<false, kGoLd>(S.go, 0, kc, g, r);
				t = mfma(a, b, t);
			}
			if constexpr (MULTI) {
				accO = t;
			} else {
				if (undo != 1.0f) {
					asm volatile("" ::: "memory");
					t *= undo;
				}
				accO += t;
			}
		};
		// this wave's loss sum, in a fixed order
#pragma unroll
		for (int off = 32; off >= 1; off >>= 1) lv += __shfl_xor(lv, off);
		if (lane == 0) S.lsum[wave] += lv;

		// ---- backward chain: delta_H = (W_H^T g_o) . 1[a_H > 0], then delta_l for l = H-1 .. 1.
		// A one-chunk workgroup takes each layer's dW as soon as every wave has stored
		// that layer's delta, so its partial-row stores drain while the backward runs
		// instead of all at the end; a multi-chunk workgroup, which stores once after
		// its last chunk, takes them all after the backward.
		half8_t Bd[2];
		{
			float4_t acc[4];
#pragma unroll
			for (int mt = 0; mt < 4; ++mt) acc[mt] = mfma(wt_frag(((H - 1) * 4 + mt) * 2), Bg, (float4_t)0.0f);
			const uint4 mH[2] = {act_masks(B[0]), act_masks(B[1])}; // B still holds a_H
			relu_bwd_pack(acc, mH, Bd);
			store_brow(S.dlt[H - 1], Bd, g, col);
		}
		if constexpr (!MULTI) {
			block_lds_sync(); // delta_H and the output gradients of every wave
			dw_layer(H - 1);
			dw_output();
		}
#pragma unroll
		for (int l = H - 1; l >= 1; --l) {
			// the layer's eight W^T fragments are read from LDS together, so the
			// MFMAs wait on one LDS round trip rather than on one each
			half8_t wb[4][2];
#pragma unroll
			for (int c = 0; c < 2; ++c)
#pragma unroll
				for (int mt = 0; mt < 4; ++mt) wb[mt][c] = wt_frag(((l - 1) * 4 + mt) * 2 + c);
			__builtin_amdgcn_sched_barrier(0);
			float4_t acc[4];
#pragma unroll
			for (int mt = 0; mt < 4; ++mt) acc[mt] = (float4_t)0.0f;
#pragma unroll
			for (int c = 0; c < 2; ++c)
#pragma unroll
				for (int mt = 0; mt < 4; ++mt) acc[mt] = mfma(wb[mt][c], Bd[c], acc[mt]);
			relu_bwd_pack(acc, keep[l], Bd);
			store_brow(S.dlt[l - 1], Bd, g, col);
			if constexpr (!MULTI) {
				block_lds_sync(); // delta_l of every wave
				dw_layer(l - 1);
			}
		}
		if constexpr (MULTI) {
			block_lds_sync(); // every wave's deltas (and loss sum)
#pragma unroll
			for (int l = 0; l < H; ++l) dw_layer(l);
			dw_output();
			if (undo != 1.0f) scale_sums(undo); // uniform
		}
	}
	if constexpr (!MULTI) block_lds_sync(); // every wave's loss sum

	// ---- the workgroup's partial row: dW tiles (float4 per lane) and the loss sum
	if constexpr (MULTI) {
#pragma unroll
		for (int l = 0; l < H; ++l)
#pragma unroll
			for (int i = 0; i < 2; ++i)
#pragma unroll
				for (int o2 = 0; o2 < 2; ++o2) store_tile(l, it0 + i, ot0 + o2, accW[MULTI ? l : 0][i][o2]);
	}
	store_out_tile(accO);
	if (tid == 0) part[WC] = (S.lsum[0] + S.lsum[1]) + (S.lsum[2] + S.lsum[3]);
	if (overflow && __any(bad) && lane == 0) atomicOr(overflow, 1u);
}

// N-tiles each wave runs through the layers together (1 or 2; both measure the
// same, and 1 needs no register spill); VKNRC_INFER_TILES overrides.
int infer_tiles() {
	static const int t = [] {
		const char *e = getenv("VKNRC_INFER_TILES");
		return (e && atoi(e) == 2) ? 2 : 1;
	}();
	return t;
}


// The software-pipelined inference kernel (VKNRC_INFER_PIPE=0 selects the other).
bool infer_pipelined() {
	static const bool p = [] {
		const char *e = getenv("VKNRC_INFER_PIPE");
		return !(e && atoi(e) == 0);
	}();
	return p;
}

template <int H, int ENC, bool F16IN> hipError_t infer_launch(const InferGroup &grp, hipStream_t s) {
	// persistent: a few workgroups per CU, each wave a contiguous tile range
	uint32_t blocks = (grp.total_tiles + 3) / 4;
	const uint32_t cap = (uint32_t)device_cu_count() * infer_wg_per_cu();
	if (blocks > cap) blocks = cap;
	if (blocks == 0) blocks = 1;
	if (!F16IN && infer_pipelined())
		fused_infer_pipe_kernel<H, ENC><<<dim3(blocks), dim3(256), 0, s>>>(grp);
	else if (infer_tiles() == 1)
		fused_infer_kernel<H, ENC, F16IN, 1><<<dim3(blocks), dim3(256), 0, s>>>(grp);
	else
		fused_infer_kernel<H, ENC, F16IN, 2><<<dim3(blocks), dim3(256), 0, s>>>(grp);
	return hipGetLastError();
}

// The dynamic LDS size of both variants, set once (before any stream capture; the
// initialiser of a function-local static runs once even under concurrent first calls).
template <int H, int ENC, int LOSS, bool F16IN> hipError_t grad_small_attributes() {
	static const hipError_t done = [] {
		const int smem = (int)sizeof(SmallGradSmem<H>);
		hipError_t e = hipFuncSetAttribute(
		    reinterpret_cast<const void *>(fused_grad_small_kernel<H, ENC, LOSS, F16IN, true>),
		    hipFuncAttributeMaxDynamicSharedMemorySize, smem);
		if (e == hipSuccess)
			e = hipFuncSetAttribute(reinterpret_cast<const void *>(fused_grad_small_kernel<H, ENC, LOSS, F16IN, false>),
			                        hipFuncAttributeMaxDynamicSharedMemorySize, smem);
		return e;
	}();
	return done;
}

template <int H, int ENC, int LOSS, bool F16IN> hipError_t grad_small_launch(const GradGroup &grp, hipStream_t s) {
	const size_t smem = sizeof(SmallGradSmem<H>);
	bool multi = false;
	for (uint32_t i = 0; i < grp.n_items; ++i) multi |= grp.item[i].chunks > 1;
	hipError_t e = grad_small_attributes<H, ENC, LOSS, F16IN>();
	if (e != hipSuccess) return e;
	if (multi)
		fused_grad_small_kernel<H, ENC, LOSS, F16IN, true><<<dim3(grp.total_blocks), dim3(256), smem, s>>>(grp);
	else
		fused_grad_small_kernel<H, ENC, LOSS, F16IN, false><<<dim3(grp.total_blocks), dim3(256), smem, s>>>(grp);
	return hipGetLastError();
}

template <int H, int ENC> hipError_t prepare_enc(int loss) {
	if (loss == NRC_LOSS_L2) return grad_small_attributes<H, ENC, NRC_LOSS_L2, false>();
	return grad_small_attributes<H, ENC, NRC_LOSS_RELATIVE_L2_LUMINANCE, false>();
}

template <int H, int ENC> hipError_t grad_group_enc(int loss, const GradGroup &g, hipStream_t s) {
	if (loss == NRC_LOSS_L2) return grad_small_launch<H, ENC, NRC_LOSS_L2, false>(g, s);
	return grad_small_launch<H, ENC, NRC_LOSS_RELATIVE_L2_LUMINANCE, false>(g, s);
}

template <int H> hipError_t grad_group_h(int encoding, int loss, const GradGroup &g, hipStream_t s) {
	switch (encoding) {
	case NRC_ENCODING_TRI12: return grad_group_enc<H, NRC_ENCODING_TRI12>(loss, g, s);
	case NRC_ENCODING_SINCOS6: return grad_group_enc<H, NRC_ENCODING_SINCOS6>(loss, g, s);
	default: return grad_group_enc<H, NRC_ENCODING_UV_SINCOS16>(loss, g, s);
	}
}

template <int H> hipError_t infer_h(int enc, const InferGroup &grp, hipStream_t s) {
	switch (enc) {
	case NRC_ENCODING_TRI12: return infer_launch<H, NRC_ENCODING_TRI12, false>(grp, s);
	case NRC_ENCODING_SINCOS6: return infer_launch<H, NRC_ENCODING_SINCOS6, false>(grp, s);
	default: return infer_launch<H, NRC_ENCODING_UV_SINCOS16, false>(grp, s);
	}
}

} // namespace

static std::atomic<uint32_t> g_infer_wg_per_cu{0};

uint32_t infer_wg_per_cu() {
	uint32_t v = g_infer_wg_per_cu.load(std::memory_order_relaxed);
	if (v == 0) {
		const char *e = getenv("VKNRC_INFER_WG_PER_CU");
		const int x = e ? atoi(e) : 2;
		v = x < 1 ? 1u : (uint32_t)x;
		g_infer_wg_per_cu.store(v, std::memory_order_relaxed);
	}
	return v;
}

void set_infer_wg_per_cu(uint32_t v) { g_infer_wg_per_cu.store(v, std::memory_order_relaxed); }

int device_cu_count() {
	static const int cus = [] {
		int dev = 0;
		(void)hipGetDevice(&dev);
		hipDeviceProp_t prop;
		if (hipGetDeviceProperties(&prop, dev) == hipSuccess && prop.multiProcessorCount > 0)
			return prop.multiProcessorCount;
		return 256;
	}();
	return cus;
}

hipError_t launch_fused_infer_group(int hidden, int encoding, InferGroup &grp, hipStream_t stream) {
	uint32_t tiles = 0;
	for (uint32_t i = 0; i < grp.n_items; ++i) {
		grp.item[i].tile_begin = tiles;
		tiles += (grp.item[i].max_count + 63) / 64;
	}
	grp.total_tiles = tiles;
	if (tiles == 0) return hipSuccess;
	if (hidden == 5) return infer_h<5>(encoding, grp, stream);
	return infer_h<4>(encoding, grp, stream);
}

hipError_t launch_fused_forward_f16(int hidden, const _Float16 *inputs, uint32_t n, const _Float16 *wf, float *out,
                                    hipStream_t stream) {
	if (n == 0) return hipSuccess;
	InferGroup grp{};
	grp.n_items = 1;
	grp.item[0].max_count = n;
	grp.item[0].wf = wf;
	grp.item[0].out = out;
	grp.item[0].tile_begin = 0;
	grp.total_tiles = (n + 63) / 64;
	grp.fin = inputs;
	grp.clamp0 = 0;
	if (hidden == 5) return infer_launch<5, NRC_ENCODING_TRI12, true>(grp, stream);
	return infer_launch<4, NRC_ENCODING_TRI12, true>(grp, stream);
}

static int env_int(const char *name, int dflt) {
	const char *e = getenv(name);
	return e ? atoi(e) : dflt;
}

// One 64-record chunk per workgroup up to about one workgroup per CU, then
// several chunks per workgroup (dW summed in registers) so that the partial rows
// stay about one per CU; VKNRC_GRAD_CHUNKS forces the chunk count.
FusedGradPlan plan_fused_gradient(uint32_t n, uint32_t chunks) {
	static const int forced = env_int("VKNRC_GRAD_CHUNKS", 0);
	FusedGradPlan p;
	const uint32_t nchunks = (n + kFusedSmallRecords - 1) / kFusedSmallRecords;
	const uint32_t cus = (uint32_t)device_cu_count();
	uint32_t c = chunks > 0 ? chunks : forced > 0 ? (uint32_t)forced : (nchunks + cus - 1) / cus;
	if (c < 1) c = 1;
	if (c > kFusedMaxChunks) c = kFusedMaxChunks;
	p.chunks = c;
	p.blocks = (nchunks + c - 1) / c;
	if (p.blocks == 0) p.blocks = 1;
	return p;
}

hipError_t prepare_fused(int hidden, int encoding, int loss) {
	(void)device_cu_count();
	switch (encoding) {
	case NRC_ENCODING_TRI12: return hidden == 5 ? prepare_enc<5, NRC_ENCODING_TRI12>(loss) : prepare_enc<4, NRC_ENCODING_TRI12>(loss);
	case NRC_ENCODING_SINCOS6:
		return hidden == 5 ? prepare_enc<5, NRC_ENCODING_SINCOS6>(loss) : prepare_enc<4, NRC_ENCODING_SINCOS6>(loss);
	default:
		return hidden == 5 ? prepare_enc<5, NRC_ENCODING_UV_SINCOS16>(loss)
		                   : prepare_enc<4, NRC_ENCODING_UV_SINCOS16>(loss);
	}
}

hipError_t launch_fused_gradient_group(int hidden, int encoding, int loss, GradGroup &grp, hipStream_t stream) {
	uint32_t blocks = 0;
	for (uint32_t i = 0; i < grp.n_items; ++i) {
		grp.item[i].block_begin = blocks;
		blocks += grp.item[i].blocks;
	}
	grp.total_blocks = blocks;
	if (blocks == 0) return hipSuccess;
	if (hidden == 5) return grad_group_h<5>(encoding, loss, grp, stream);
	return grad_group_h<4>(encoding, loss, grp, stream);
}

GradItem grad_item(const TrainArgs &a, const FusedGradPlan &plan) {
	GradItem it{};
	it.records = a.records;
	it.d_count = a.d_count;
	it.wf = a.wf;
	it.wtf = a.wtf;
	it.state = a.state;
	it.partials = a.partials;
	it.overflow = &a.state->overflow[a.parity];
	it.max_count = a.max_count;
	it.clamp = a.clamp;
	it.chunks = plan.chunks;
	it.blocks = plan.blocks;
	it.pstride = a.pstride;
	it.parity = a.parity;
	return it;
}

hipError_t launch_fused_gradient(const TrainArgs &a, const FusedGradPlan &plan, hipStream_t stream) {
	GradGroup grp{};
	grp.n_items = 1;
	grp.item[0] = grad_item(a, plan);
	return launch_fused_gradient_group(a.hidden, a.encoding, a.loss, grp, stream);
}

hipError_t launch_fused_gradient_f16(int hidden, const _Float16 *inputs, const float *targets, uint32_t n,
                                     const _Float16 *wf, const _Float16 *wtf, float *partials, uint32_t pstride,
                                     uint32_t *nparts, hipStream_t stream) {
	const FusedGradPlan p = plan_fused_gradient(n);
	*nparts = p.blocks;
	GradGroup grp{};
	grp.n_items = 1;
	grp.fin = inputs;
	grp.ftgt = targets;
	GradItem &it = grp.item[0];
	it.wf = wf;
	it.wtf = wtf;
	it.partials = partials;
	it.max_count = n;
	it.clamp = n;
	it.chunks = p.chunks;
	it.blocks = p.blocks;
	it.pstride = pstride;
	grp.total_blocks = p.blocks;
	if (hidden == 5) return grad_small_launch<5, NRC_ENCODING_TRI12, NRC_LOSS_L2, true>(grp, stream);
	return grad_small_launch<4, NRC_ENCODING_TRI12, NRC_LOSS_L2, true>(grp, stream);
}

} // namespace nrc
