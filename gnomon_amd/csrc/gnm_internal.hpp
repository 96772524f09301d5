// Internal declarations shared by the HIP kernels (gnm_kernels.hip) and the
// host-side plan/context code (gnm_api.hip).  Not part of the C ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cmath>

namespace gnm {

// People blocks of the split-BF16 adjoint (FP32 partial sums of
// integer-valued BF16 limbs stay exact: 8192 people * dosage 2 * |limb| 256
// = 2^22 < 2^24).
constexpr int kFlushRows = 8192;
// The split-int8 path has its own, longer scale blocks: an int32 sum of
// dosage (<= 2) x int8 limb (|.| <= 128) products stays exact up to 2^22
// rows; 65536 rows (sums < 2^24) keep the FP64 read-modify-write flushes of
// the partial sums rare.
constexpr int kI8FlushRows = 65536;
// Scale blocks of the split-BF16 scoring kernel: an FP32 sum of dosage (<= 2)
// x BF16 limb digit (|.| <= 255) products stays an exact integer below 2^24
// over 32768 rows (32768 * 510 = 16711680).  (The split-BF16 adjoint keeps
// kFlushRows-person blocks.)
constexpr int kBfFlushRows = 32768;
// fewest program rows per variant split of a scoring launch on the MFMA paths
constexpr uint64_t kMinSplitRows = 4096;
constexpr int flush_rows(bool int8) { return int8 ? kI8FlushRows : kBfFlushRows; }
constexpr int kLimbBits = 8;
constexpr int kFp32MaxClasses = 4;     // membership classes counted in-kernel by the FP32 kernel

// Signed base-256 digits of an integer with |Q| <= 2^(8L) - 2^(8(L-1)):
// low digits in [-128, 127], top digit in [-255, 255]; all exact in BF16.
__host__ __device__ inline void limb_digits(int64_t Q, uint32_t L, int32_t *out) {
    for (uint32_t j = L - 1; j > 0; --j) {
        const int64_t d = ((Q + 128) & 255) - 128;
        out[j] = (int32_t)d;
        Q = (Q - d) / 256;
    }
    out[0] = (int32_t)Q;
}

// Exponent marker of a (scale block, column) holding a non-finite value:
// its digits are zero and its limb scales NaN, so every score of that column
// comes out NaN (a NaN or infinite weight times any dosage, 0 included, is
// NaN in IEEE arithmetic).
constexpr int kBadExp = 0x40000000;

// max that keeps a NaN once seen (fmax drops it)
__host__ __device__ inline double nan_max(double a, double b) { return (b > a || b != b) ? b : a; }

// smallest e with qmax * 2^e >= m (0 for m == 0); kBadExp for m inf or NaN
__host__ __device__ inline int block_exponent_of(double m, double qmax) {
    if (!(m <= 1.7976931348623157e308)) return kBadExp;
    if (m == 0.0) return 0;
    int e = (int)ceil(log2(m / qmax));
    while (ldexp(qmax, e) < m) ++e;
    while (ldexp(qmax, e - 1) >= m) --e;
    return e;
}

// FP64 scale of limb digits at unit 2^(e + shift)
__host__ __device__ inline double limb_scale(int e, int shift) {
    return e == kBadExp ? __builtin_nan("") : ldexp(1.0, e + shift);
}

// |q| <= 256 is exact in BF16: the top half of its f32 pattern
__host__ __device__ inline uint16_t bf16_of_int(int32_t q) {
    return (uint16_t)(__builtin_bit_cast(uint32_t, (float)q) >> 16);
}

// Device-side view of the weight plan built by gnm_create.
struct DevicePlan {
    uint64_t num_prog_rows;    // R: reduction rows (one per variant, two for mixed-flip variants)
    const uint32_t *prog_var;  // [R] reconciled variant of each row
    const uint32_t *flipmask;  // [ceil(R/32)] bit j of word s: row 32s+j reads the A1 allele
    const uint32_t *prog_cls;  // [R] membership-class bitmask of count rows (FP32 kernel, <=32 classes)
    // FP32 path
    const float *w32;          // [Rpad][k32] effect-allele FP32 weights (k32 = fp32_weight_pitch(K))
    uint64_t k32;
    // split-BF16 path
    const uint32_t *bfrag;     // [R/32][tiles][64 lanes][4 dwords] B fragments (weights then counts)
    const double *scales;      // [R/flush_rows(int8)][ncol] power-of-two limb scales
    uint32_t limbs;            // 2 or 3 (BF16 limbs), 3 or 4 (int8 limbs)
    uint32_t int8;             // 1: int8 limbs, fragments [R/64][tiles][64 lanes][16 i8]
    const uint8_t *flipbyte;   // [Rpad] int8 path: 4 for rows that read the A1 allele, else 0
    uint32_t ncol;             // limbs * K (weight columns, column = limb*K + score)
    uint32_t wtiles;           // ceil(ncol/16)
    uint32_t ctiles;           // ceil(nclass/16)
    uint32_t nclass;           // membership classes (distinct CSR column supports)
    uint32_t vcount;           // 1: one class and no count tiles -- the MFMA kernels count missing
                               // calls on the vector ALUs (selector bit 7 marks count rows)
    const uint32_t *score_class; // [K]
};

// Kernel launchers (gnm_kernels.hip).  All are asynchronous on `stream`.
hipError_t launch_synth(uint64_t seed, uint32_t p_lo24, uint32_t p_hi24, uint32_t miss16,
                        uint64_t num_samples, uint64_t first_variant, uint64_t num_rows,
                        uint8_t *out, uint64_t row_stride, hipStream_t stream);
hipError_t launch_hbm_probe(int kind, const uint8_t *src, uint64_t src_bytes, uint8_t *dst, hipStream_t stream);
uint64_t hbm_probe_sink_bytes();   // output bytes the read probe writes (one dword per thread)
hipError_t launch_unpack(const uint8_t *rows, uint64_t stride, uint64_t num_rows,
                         uint64_t num_people, const uint8_t *flip, uint8_t *dosage,
                         uint32_t *missing, hipStream_t stream);
hipError_t launch_compact(const uint8_t *rows, uint64_t stride, uint64_t num_rows,
                          const uint32_t *out_to_fam, uint64_t num_people, uint8_t *out,
                          uint64_t out_stride, hipStream_t stream);
uint32_t fp32_splits(uint64_t num_people, uint64_t num_scores);
// row pitch (in floats) of the FP32 kernel's weight table for num_scores scores
uint64_t fp32_weight_pitch(uint64_t num_scores);
// Plain FP32 scoring (FP32 matrix cores).  identity: program row p0 + i is
// row i of the block (no per-row variant lookup).
hipError_t launch_score_fp32(const DevicePlan &plan, const uint8_t *rows, uint64_t stride,
                             uint64_t bpv, uint64_t v0, uint64_t p0, uint64_t p1,
                             uint64_t num_people, uint64_t num_scores, uint32_t splits,
                             bool count_in_kernel, bool identity, double *part, uint32_t *cpart,
                             hipStream_t stream);
// Split-int8 scoring (gnm_i8.hip); rows must be 4-byte aligned with a
// 4-byte multiple stride below 2^25 for the dword staging loads (else byte
// loads).  One launch covers every tile group of [tile_begin, tile_end)
// (groups of tiles_per_pass tiles: 0 = default, else 3, 4, 5, 6 or 8, or 18
// for two sets of 9-tile groups over the same people;
// batch_stages: 0 = default, else 2 or 4 stages whose loads are issued
// together).
hipError_t launch_score_i8(const DevicePlan &plan, const uint8_t *rows, uint64_t stride, uint64_t bpv, uint64_t v0,
                           uint64_t p0, uint64_t p1, uint64_t num_people, uint32_t splits, uint32_t tile_begin,
                           uint32_t tile_end, bool identity, double *part, uint32_t *cpart, hipStream_t stream,
                           uint32_t tiles_per_pass = 0, uint32_t batch_stages = 0);
// variant splits of the int8 kernel for num_people and ntiles tiles
uint32_t i8_splits(uint64_t num_people, uint32_t ntiles);
// cost-model units per wave and stage of a scoring pass over ntiles tiles at
// the default tiles per pass
uint64_t i8_default_cost(uint32_t ntiles);
uint64_t bf16_default_cost(uint32_t ntiles);
// One membership class: count its missing calls on the vector ALUs (about 30
// instructions per wave and stage in one tile group) rather than in a count
// tile, when the cost model says the extra tile costs more.
inline bool vcount_pays(bool int8, uint32_t wtiles) {
    const uint64_t valu = int8 ? i8_default_cost(wtiles) : bf16_default_cost(wtiles);
    const uint64_t ctile = int8 ? i8_default_cost(wtiles + 1) : bf16_default_cost(wtiles + 1);
    return valu + 120 < ctile;
}
// Split-BF16 scoring (gnm_bf16.hip): the same kernel with BF16 limbs (B
// fragments [64-row stage][tile][k-step h][lane][8 bf16], rows 16 g + 8 h + j
// of the stage for lane (n, g)), the same row and stride conditions, one
// launch over every tile group of [tile_begin, tile_end) (groups of
// tiles_per_pass tiles: 0 = default, else 3, 5, 6, 8 or 9).
hipError_t launch_score_bf16(const DevicePlan &plan, const uint8_t *rows, uint64_t stride, uint64_t bpv, uint64_t v0,
                             uint64_t p0, uint64_t p1, uint64_t num_people, uint32_t splits, uint32_t tile_begin,
                             uint32_t tile_end, bool identity, double *part, uint32_t *cpart, hipStream_t stream,
                             uint32_t tiles_per_pass = 0);
uint32_t bf16_splits(uint64_t num_people, uint32_t ntiles);
hipError_t launch_forward_frag_i8(const float *W, uint64_t V, uint64_t K, uint32_t L, const int *exps,
                                  uint32_t wtiles, uint32_t tiles, uint64_t nks, uint32_t *frag, double *scales,
                                  hipStream_t stream);
// split-int8 / split-BF16 adjoint (gnm_i8.hip): R fragments in super-step
// people order, then dW = G^T R through an FP64 workspace [V][Kp]; tiles =
// L*Kp/16 padded to 4, 8, 12 or 16; rows 4-byte aligned, stride a multiple
// of 4 below 2^25
hipError_t launch_adjoint_frag_i8(const float *R, uint64_t N, uint64_t K, uint64_t Kp, uint32_t L, const int *exps,
                                  uint32_t tiles, uint32_t tiles_used, uint64_t nks, uint32_t *frag, double *scales,
                                  hipStream_t stream);
hipError_t launch_adjoint_frag_bf(const float *R, uint64_t N, uint64_t K, uint64_t Kp, uint32_t L, const int *exps,
                                  uint32_t tiles, uint32_t tiles_used, uint64_t nks, uint32_t *frag, double *scales,
                                  hipStream_t stream);
hipError_t launch_adjoint_mx(bool bf, const uint8_t *rows, uint64_t stride, uint64_t V, uint64_t N,
                             const uint8_t *flip, const uint32_t *frag, const double *scales, uint32_t tiles,
                             uint64_t K, uint64_t Kp, uint32_t L, double *ws, float *dW, hipStream_t stream);
// per-row selector bytes of the scoring kernel: kFlipSelI8 (int8 path) or
// kFlipSelBf (BF16 path) for flipped rows
constexpr uint8_t kFlipSelI8 = 4, kFlipSelBf = 0x55;
hipError_t launch_flip_bytes(const uint8_t *flip, uint64_t V, uint64_t Rpad, uint8_t value, uint8_t *out,
                             hipStream_t stream);
constexpr int kI8KStep = 64;           // rows per k-step of the int8 kernel

// Device build of the weight plan (gnm_plan.hip).  wrow: f64 [R][K]
// (zeroed by the caller); rowcls: u32 [R][cw] class bits of count rows
// (zeroed; NULL when no count tiles).
hipError_t launch_plan_rows(const uint64_t *off, const uint32_t *cols, const float *w, const float *co,
                            const uint64_t *prog_off, const uint8_t *var_rows, uint64_t V, uint64_t K,
                            const uint32_t *score_class, uint32_t cw, double *wrow, uint32_t *rowcls,
                            hipStream_t stream);
// per (scale block of fr rows, score): exponent e with max|w| <= qmax * 2^e
// and the limb scales [block][L*K]
hipError_t launch_plan_exps(const double *wrow, uint64_t R, uint64_t K, uint64_t fr, double qmax, uint32_t L,
                            uint64_t ncol, uint64_t nblk, int *uexp, double *scales, hipStream_t stream);
// B fragments of nks 64-row stages (int8: one k-step each,
// [stage][tile][lane][16 i8]; BF16: two k-steps each,
// [stage][tile][k-step][lane][8 bf16])
hipError_t launch_plan_frag(bool int8, const double *wrow, uint64_t R, uint64_t K, const int *uexp, uint64_t fr,
                            uint32_t L, uint32_t ncol, uint32_t wtiles, uint32_t tiles, uint64_t nks,
                            const uint8_t *count_row, const uint32_t *rowcls, uint32_t cw, uint32_t nclass,
                            void *frag, hipStream_t stream);
// FP32-path weights w32[r][k] (row pitch k32) = W rounded to f32
hipError_t launch_plan_w32(const double *wrow, uint64_t R, uint64_t K, uint64_t k32, float *w32, hipStream_t stream);

// Sparse ingest (gnm_sparse.hip).  A sparse row goes as a bitmap of its
// non-zero 64-bit words (32 calls each) plus those words, in chunks of
// kSparseChunkWords words (4096 row bytes, 16384 people).
constexpr uint32_t kSparseChunkWords = 512;
constexpr uint64_t kSparseMaxSlotRows = 1u << 20;   // staged rows per slot
constexpr uint32_t kMapZero = 0xFFFFFFFFu;           // row map: an all-00 row
constexpr uint32_t kMapSparse = 0xFFFFFFFEu;         // row map: rebuilt from its words
// Non-zero words of one packed row (num_fam people; keep2: 2-bit mask per
// person, NULL = all) in row order; bitmap [nch][8] of the non-zero words;
// chunk_start[c] (c <= nch) = index of chunk c's first stored word.  out must
// hold max_words + 64 words.  -1 when there are more than max_words.
int64_t sparse_encode_row(const uint8_t *row, uint64_t num_fam, const uint64_t *keep2, uint64_t max_words,
                          uint32_t nch, uint64_t *out, uint64_t *bitmap, uint32_t *chunk_start);
// out row r (pitch bytes, a multiple of 64) = pack row map[r], zeros for
// kMapZero; rows marked kMapSparse are left alone
hipError_t launch_expand_rows(const uint8_t *pack, const uint32_t *map, uint64_t n, uint64_t pitch, uint8_t *out,
                              hipStream_t stream);
// rebuilds rows sv_row[s] of out from their records at word bmoff[s] of pack
// (bitmap [nch][8], then the stored words); cs: [nsv][nch+1] index of each
// chunk's first stored word within the record's words
hipError_t launch_sparse_rows(const uint64_t *pack, const uint64_t *bmoff, const uint32_t *sv_row, const uint32_t *cs,
                              uint64_t nsv, uint32_t nch, uint64_t pitch, uint8_t *out, hipStream_t stream);
// scores[p][k] += sum_s sum_l part[s][p][l*K + k]; counts[p][k] += sum_s cpart[s][p][class[k]]
hipError_t launch_reduce(const double *part, const uint32_t *cpart, uint32_t splits,
                         uint64_t num_people, uint64_t num_scores, uint32_t limbs,
                         uint32_t nclass, const uint32_t *score_class, double *scores,
                         uint32_t *counts, hipStream_t stream);
hipError_t launch_fill_f64(double *p, uint64_t n, double value, hipStream_t stream);
hipError_t launch_fill_u32(uint32_t *p, uint64_t n, uint32_t value, hipStream_t stream);

}  // namespace gnm
