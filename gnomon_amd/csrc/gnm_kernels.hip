// gfx950 kernels of the polygenic-scoring hot path.
//
// Genotype decode follows the PLINK SNP-major layout used by the reference
// (score/batch.rs:489-503, 797-844): sample j of a row sits in byte j>>2 at
// bits 2(j&3); codes 00 -> A2 dosage 0, 10 -> 1, 11 -> 2, 01 -> missing.
// Every scoring kernel works in effect-allele space: a row flagged "flip"
// reads the A1 allele, whose dosage is 2-d; the flip is the code map
// 00<->11 (01 and 10 fixed), i.e. swap the two bits of each field and invert.
// A missing genotype always contributes 0, which reproduces the reference's
// baseline-plus-missing-correction arithmetic (score/prepare.rs:197-208,
// score/batch.rs:372-385) without a separate correction contraction.
#include "gnm_internal.hpp"

#include <type_traits>
#include <utility>

namespace gnm {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));
typedef uint32_t u32x4_t __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Effect-allele flip of four packed genotypes: 00<->11, 01 and 10 unchanged.
__device__ __forceinline__ uint32_t flip_byte(uint32_t b) {
    return (~(((b & 0x55u) << 1) | ((b >> 1) & 0x55u))) & 0xFFu;
}

// ---------------------------------------------------------------------------
// Synthetic rows (counter-based hash, integer-only: identical to the CPU
// generator of the test oracle).
// ---------------------------------------------------------------------------
__global__ void k_synth(uint64_t seed, uint32_t p_lo24, uint32_t p_hi24, uint32_t miss16,
                        uint64_t num_samples, uint64_t first_variant, uint64_t num_rows,
                        uint64_t bpv, uint8_t *__restrict__ out, uint64_t row_stride) {
    const uint64_t b = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= bpv) return;
    for (uint64_t r = blockIdx.y; r < num_rows; r += gridDim.y) {
        const uint64_t v = first_variant + r;
        const uint64_t kv = mix64(seed ^ mix64(v + 0x632BE59BD9B4E019ull));
        const uint64_t u24 = mix64(kv ^ 0xA0761D6478BD642Full) >> 40;
        const uint32_t p24 = p_lo24 + (uint32_t)(((uint64_t)(p_hi24 - p_lo24) * u24) >> 24);
        const uint64_t q24 = (1ull << 24) - p24;
        const uint64_t t0 = q24 * q24;
        const uint64_t t1 = t0 + 2ull * (uint64_t)p24 * q24;
        uint32_t byte = 0;
#pragma unroll
        for (uint32_t j = 0; j < 4; ++j) {
            const uint64_t s = b * 4 + j;
            if (s < num_samples) {
                const uint64_t h = mix64(kv + (s + 1) * 0x9E3779B97F4A7C15ull);
                uint32_t code;
                if ((uint32_t)(h >> 48) < miss16) code = 1u;
                else {
                    const uint64_t u48 = h & 0xFFFFFFFFFFFFull;
                    code = (u48 < t0) ? 0u : (u48 < t1 ? 2u : 3u);
                }
                byte |= code << (2 * j);
            }
        }
        out[r * row_stride + b] = (uint8_t)byte;
    }
}

hipError_t launch_synth(uint64_t seed, uint32_t p_lo24, uint32_t p_hi24, uint32_t miss16,
                        uint64_t num_samples, uint64_t first_variant, uint64_t num_rows,
                        uint8_t *out, uint64_t row_stride, hipStream_t stream) {
    const uint64_t bpv = (num_samples + 3) / 4;
    if (num_rows == 0 || bpv == 0) return hipSuccess;
    dim3 grid((unsigned)((bpv + 255) / 256), (unsigned)(num_rows < 16384 ? num_rows : 16384));
    hipLaunchKernelGGL(k_synth, grid, dim3(256), 0, stream, seed, p_lo24, p_hi24, miss16,
                       num_samples, first_variant, num_rows, bpv, out, row_stride);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Standalone unpack: reference tile values (0,1,2 effect dosage, 3 missing)
// [row][person] plus per-person missing counts.
//
// Tile values are formed on 16 packed fields at once: with l, h the low and
// high bit of a code, value bit0 = l^h and bit1 = l (A2 dosage) or ~h (A1
// dosage, flipped row); the 2-bit fields are then spread to bytes.  Missing
// flags are summed in byte counters (at most 255 rows), drained into 16-bit
// pair counters, and added to the output with one atomic per person per row
// block.
//
// The stage writes four bytes per byte read, so its speed is set by how the
// stores coalesce.  Vector path: a wave owns 1024 packed bytes (4096 people)
// of each row; lane l decodes dwords l, l+64, l+128, l+192 of that span, so
// every load instruction reads 256 contiguous bytes and every store writes
// 1 KB contiguous (full 128-byte lines from one instruction).  Addresses are
// a per-row buffer resource (uniform base) plus a 32-bit lane offset; lanes
// past the row end read zeros and their stores are dropped by the buffer
// range check.  Stores are non-temporal: the tiles are not re-read by this
// kernel.
// ---------------------------------------------------------------------------
// rows loaded before any decode (a multiple of 8, dividing kUnpackRows):
// 100k people x 1M rows took 27.1 ms with 8 and 26.2 with 32; with the row
// blocks varying fastest over the grid, 16, 24 and 32 all ran 24.7-25.2 ms
constexpr int kUnpackBatch = 32;
constexpr int kUnpackBlock = 256;        // 4 waves
constexpr int kUnpackRows = 512;         // rows per workgroup (<= 65535 for the pair counters)
constexpr int kUnpackBytesRows = 2048;   // rows per workgroup of the byte-wise path
constexpr int kRsrcWord3 = 0x00020000;   // gfx9 raw buffer descriptor: 32-bit data format
constexpr int kStoreNT = 2;              // buffer store cache policy: non-temporal
static_assert(kUnpackBatch % 8 == 0 && kUnpackRows % kUnpackBatch == 0 && kUnpackBatch < 256,
              "whole flag words per batch, whole batches per workgroup, byte counters drained in time");

__device__ __forceinline__ uint32_t spread_fields(uint32_t byte4) {
    // four 2-bit fields of the low byte -> four bytes
    uint32_t z = (byte4 | (byte4 << 12)) & 0x000F000Fu;
    return (z | (z << 6)) & 0x03030303u;
}

// Tile values of one packed dword (16 people): o[k] byte t = person 4k+t.
// fm = 0 (A2 dosage) or ~0 (flipped row, A1 dosage), used as data so that a
// flag still in flight is only waited for where it is consumed.
__device__ __forceinline__ void decode4(uint32_t w, uint32_t fm, uint32_t (&o)[4], uint32_t (&cnt8)[4]) {
    const uint32_t L = w & 0x55555555u, H = (w >> 1) & 0x55555555u;
    const uint32_t B1 = ((L & ~fm) | (~H & fm)) & 0x55555555u;
    const uint32_t T = (L ^ H) | (B1 << 1);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const uint32_t z = spread_fields((T >> (8 * k)) & 0xFFu);
        o[k] = z;
        cnt8[k] += z & (z >> 1) & 0x01010101u;
    }
}

// Vector path.  Requires num_people % 16 == 0, 4-byte aligned rows and
// stride, 16-byte aligned dosage, 8-byte aligned flip (or NULL) and
// num_people < 2^32; num_dwords = num_people / 16.
__global__ __launch_bounds__(kUnpackBlock) void k_unpack(const uint8_t *__restrict__ rows, uint64_t stride,
                                                         uint64_t num_rows, uint64_t num_people,
                                                         uint64_t num_dwords, const uint8_t *__restrict__ flip,
                                                         uint8_t *__restrict__ dosage,
                                                         uint32_t *__restrict__ missing) {
    constexpr int R = kUnpackBatch, BS = kUnpackBlock;
    __shared__ uint32_t lcnt[32 * BS];   // 16-bit pair counters, [slot][thread]
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t wave = (uint64_t)blockIdx.y * (BS / 64) + (threadIdx.x >> 6);
    if (wave * 256 >= num_dwords) return;   // whole wave past the row end; no barriers below
    const uint32_t j0 = (uint32_t)(wave * 256) + lane;
    // Full per-lane byte offsets kept in VGPRs (opaque to constant folding),
    // so the range check sees the whole offset.
    uint32_t ioff[4], ooff[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        ioff[e] = 4u * j0 + 256u * e;
        ooff[e] = 16u * j0 + 1024u * e;
        asm volatile("" : "+v"(ioff[e]), "+v"(ooff[e]));
    }
    const uint32_t in_bytes = (uint32_t)(num_dwords * 4), out_bytes = (uint32_t)num_people;
    const uint64_t r0 = (uint64_t)blockIdx.x * kUnpackRows;
    const uint64_t r1 = r0 + kUnpackRows < num_rows ? r0 + kUnpackRows : num_rows;
    uint32_t cnt8[4][4];
#pragma unroll
    for (int e = 0; e < 4; ++e)
#pragma unroll
        for (int k = 0; k < 4; ++k) cnt8[e][k] = 0;
#pragma unroll
    for (int k = 0; k < 32; ++k) lcnt[k * BS + threadIdx.x] = 0;
    auto drain = [&]() __attribute__((always_inline)) {
#pragma unroll
        for (int e = 0; e < 4; ++e)
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                lcnt[(8 * e + 2 * k) * BS + threadIdx.x] += cnt8[e][k] & 0x00FF00FFu;
                lcnt[(8 * e + 2 * k + 1) * BS + threadIdx.x] += (cnt8[e][k] >> 8) & 0x00FF00FFu;
                cnt8[e][k] = 0;
            }
    };
    auto load = [&](uint32_t (&w)[4], uint64_t r) __attribute__((always_inline)) {
        const __amdgpu_buffer_rsrc_t ri =
            __builtin_amdgcn_make_buffer_rsrc((void *)(rows + r * stride), 0, in_bytes, kRsrcWord3);
#pragma unroll
        for (int e = 0; e < 4; ++e) w[e] = __builtin_amdgcn_raw_buffer_load_b32(ri, ioff[e], 0, 0);
    };
    auto row = [&](const uint32_t (&w)[4], uint32_t fm, uint64_t r) __attribute__((always_inline)) {
        const __amdgpu_buffer_rsrc_t ro =
            __builtin_amdgcn_make_buffer_rsrc(dosage + r * num_people, 0, out_bytes, kRsrcWord3);
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            uint32_t o[4];
            decode4(w[e], fm, o, cnt8[e]);
            __builtin_amdgcn_raw_buffer_store_b128(u32x4_t{o[0], o[1], o[2], o[3]}, ro, ooff[e], 0, kStoreNT);
        }
    };
    uint32_t since = 0;
    uint64_t r = r0;
    for (; r + R <= r1; r += R) {
        uint32_t w[R][4];
#pragma unroll
        for (int i = 0; i < R; ++i) load(w[i], r + i);
        // flip flags of the batch: uniform 8-byte loads (r % 8 == 0)
        uint64_t fw[R / 8];
#pragma unroll
        for (int q = 0; q < R / 8; ++q) fw[q] = flip ? *(const uint64_t *)(flip + r + 8 * q) : 0;
        __builtin_amdgcn_sched_barrier(0);   // all loads of the batch issue before its first store
#pragma unroll
        for (int i = 0; i < R; ++i) row(w[i], 0u - (uint32_t)(((fw[i / 8] >> (8 * (i % 8))) & 0xFFu) != 0), r + i);
        since += R;
        if (since + R > 255) {
            drain();
            since = 0;
        }
    }
    for (; r < r1; ++r) {   // fewer than R rows left
        uint32_t w[4];
        load(w, r);
        row(w, (flip && flip[r]) ? ~0u : 0u, r);
    }
    drain();
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const uint32_t j = j0 + 64u * e;
        if (j >= num_dwords) continue;
#pragma unroll
        for (int k = 0; k < 4; ++k)
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const uint32_t c = (lcnt[(8 * e + 2 * k + (t & 1)) * BS + threadIdx.x] >> (16 * (t >> 1))) & 0xFFFFu;
                if (c) atomicAdd(&missing[16u * j + 4 * k + t], c);
            }
    }
}

// Generic byte-wise path (any num_people, stride and alignment): a thread
// owns 16 packed bytes (64 people) of each row.
__device__ __forceinline__ void decode16(u32x4_t w, uint32_t fm, uint32_t (&out)[16], uint32_t (&cnt8)[16]) {
#pragma unroll
    for (int d = 0; d < 4; ++d) {
        uint32_t o[4], c[4] = {cnt8[4 * d], cnt8[4 * d + 1], cnt8[4 * d + 2], cnt8[4 * d + 3]};
        decode4(w[d], fm, o, c);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            out[4 * d + k] = o[k];
            cnt8[4 * d + k] = c[k];
        }
    }
}

// Byte counters (at most 255 rows) -> 16-bit pair counters: cnt16[2k] holds
// persons 4k (low half) and 4k+2, cnt16[2k+1] persons 4k+1 and 4k+3.
__device__ __forceinline__ void drain_counts(uint32_t (&cnt8)[16], uint32_t (&cnt16)[32]) {
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        cnt16[2 * k] += cnt8[k] & 0x00FF00FFu;
        cnt16[2 * k + 1] += (cnt8[k] >> 8) & 0x00FF00FFu;
        cnt8[k] = 0;
    }
}

__global__ __launch_bounds__(256) void k_unpack_bytes(const uint8_t *__restrict__ rows, uint64_t stride,
                                                      uint64_t num_rows, uint64_t num_people,
                                                      const uint8_t *__restrict__ flip,
                                                      uint8_t *__restrict__ dosage,
                                                      uint32_t *__restrict__ missing) {
    const uint64_t chunk = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t bpv = (num_people + 3) / 4;
    const uint64_t col = chunk * 16, p0 = col * 4;
    if (col >= bpv) return;
    const uint64_t r0 = (uint64_t)blockIdx.y * kUnpackBytesRows;
    const uint64_t r1 = r0 + kUnpackBytesRows < num_rows ? r0 + kUnpackBytesRows : num_rows;
    uint32_t cnt8[16], cnt16[32], out[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) cnt8[k] = 0;
#pragma unroll
    for (int k = 0; k < 32; ++k) cnt16[k] = 0;
    uint32_t since = 0;
    for (uint64_t r = r0; r < r1; ++r) {
        u32x4_t w;
#pragma unroll
        for (int d = 0; d < 4; ++d) {
            uint32_t x = 0;
#pragma unroll
            for (int bb = 0; bb < 4; ++bb) {
                const uint64_t o = col + 4 * d + bb;
                // past the row end: code 01 (missing), never stored or counted
                x |= (o < bpv ? (uint32_t)rows[r * stride + o] : 0x55u) << (8 * bb);
            }
            w[d] = x;
        }
        decode16(w, (flip && flip[r]) ? ~0u : 0u, out, cnt8);
        uint8_t *dst = dosage + r * num_people + p0;
#pragma unroll
        for (int k = 0; k < 16; ++k)
#pragma unroll
            for (int t = 0; t < 4; ++t)
                if (p0 + 4 * k + t < num_people) dst[4 * k + t] = (uint8_t)(out[k] >> (8 * t));
        if (++since == 255) {
            drain_counts(cnt8, cnt16);
            since = 0;
        }
    }
    drain_counts(cnt8, cnt16);
#pragma unroll
    for (int k = 0; k < 16; ++k)
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const uint32_t c = (cnt16[2 * k + (t & 1)] >> (16 * (t >> 1))) & 0xFFFFu;
            if (c && p0 + 4 * k + t < num_people) atomicAdd(&missing[p0 + 4 * k + t], c);
        }
}

static bool unpack_vector_path(const uint8_t *rows, uint64_t stride, uint64_t num_people, const uint8_t *flip,
                        const uint8_t *dosage) {
    const uintptr_t a = reinterpret_cast<uintptr_t>(rows), d = reinterpret_cast<uintptr_t>(dosage),
                    f = reinterpret_cast<uintptr_t>(flip);
    return num_people % 16 == 0 && num_people < (1ull << 32) && ((a | stride) & 3u) == 0 && (d & 15u) == 0 &&
           (f & 7u) == 0;
}

hipError_t launch_unpack(const uint8_t *rows, uint64_t stride, uint64_t num_rows,
                         uint64_t num_people, const uint8_t *flip, uint8_t *dosage,
                         uint32_t *missing, hipStream_t stream) {
    const uint64_t bpv = (num_people + 3) / 4;
    if (num_rows == 0 || bpv == 0) return hipSuccess;
    if (unpack_vector_path(rows, stride, num_people, flip, dosage)) {
        const uint64_t nd = num_people / 16;
        const uint64_t per_block = 256 * (kUnpackBlock / 64);
        if ((nd + per_block - 1) / per_block > 65535 || (num_rows + kUnpackRows - 1) / kUnpackRows >= (1ull << 31))
            return hipErrorInvalidValue;   // grid limits: 1.07e9 people, 2^40 rows
        // row blocks vary fastest over the grid: the workgroups running at
        // one time hold one span of many row blocks (at 100k x 1M 24.3-25.0
        // ms against 26.1-26.2 with spans varying fastest)
        hipLaunchKernelGGL(k_unpack, dim3((unsigned)((num_rows + kUnpackRows - 1) / kUnpackRows),
                                          (unsigned)((nd + per_block - 1) / per_block)),
                           dim3(kUnpackBlock), 0, stream, rows, stride, num_rows, num_people, nd, flip, dosage,
                           missing);
    } else {
        if ((num_rows + kUnpackBytesRows - 1) / kUnpackBytesRows > 65535 || ((bpv + 15) / 16 + 255) / 256 >= (1ull << 31))
            return hipErrorInvalidValue;   // grid limits: 134M rows on the byte path
        hipLaunchKernelGGL(k_unpack_bytes, dim3((unsigned)(((bpv + 15) / 16 + 255) / 256),
                                                (unsigned)((num_rows + kUnpackBytesRows - 1) / kUnpackBytesRows)),
                           dim3(256), 0, stream, rows, stride, num_rows, num_people, flip, dosage, missing);
    }
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Person-subset compaction (the --keep gather of output_idx_to_fam_idx,
// score/types.rs:151-158): output byte j of a row packs the codes of output
// people 4j..4j+3.
// ---------------------------------------------------------------------------
__global__ void k_compact(const uint8_t *__restrict__ rows, uint64_t stride, uint64_t num_rows,
                          const uint32_t *__restrict__ out_to_fam, uint64_t num_people,
                          uint8_t *__restrict__ out, uint64_t out_stride) {
    const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t obpv = (num_people + 3) / 4;
    if (j >= obpv) return;
    uint32_t fam[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) fam[t] = (j * 4 + t < num_people) ? out_to_fam[j * 4 + t] : 0xFFFFFFFFu;
    for (uint64_t r = blockIdx.y; r < num_rows; r += gridDim.y) {
        const uint8_t *row = rows + r * stride;
        uint32_t byte = 0;
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const uint32_t code = (fam[t] != 0xFFFFFFFFu) ? (row[fam[t] >> 2] >> (2 * (fam[t] & 3))) & 3u : 1u;
            byte |= code << (2 * t);
        }
        out[r * out_stride + j] = (uint8_t)byte;
    }
}

hipError_t launch_compact(const uint8_t *rows, uint64_t stride, uint64_t num_rows,
                          const uint32_t *out_to_fam, uint64_t num_people, uint8_t *out,
                          uint64_t out_stride, hipStream_t stream) {
    const uint64_t obpv = (num_people + 3) / 4;
    if (num_rows == 0 || obpv == 0) return hipSuccess;
    dim3 grid((unsigned)((obpv + 255) / 256), (unsigned)(num_rows < 8192 ? num_rows : 8192));
    hipLaunchKernelGGL(k_compact, grid, dim3(256), 0, stream, rows, stride, num_rows, out_to_fam,
                       num_people, out, out_stride);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Plain FP32 path on the FP32 matrix cores (v_mfma_f32_4x4x1_16b_f32).
//
// Arithmetic: FP32 weights, exact products (dosage 0/1/2 times an FP32 weight
// is exact in FP32), FP32 accumulation over blocks of kF32FlushRows program
// rows, each block sum then added in FP64 -- the arithmetic class of the
// reference kernel's FP32 sums over 256-variant mini-batches widened into
// FP64 accumulators (score/kernel.rs:124-176, score/batch.rs:34, 387-419).
//
// Layout: a thread owns one byte column of the rows (four people, one per
// "tile" t = the person's position in the byte); a workgroup of 4 waves reads
// 256 contiguous bytes of every row.  One MFMA multiplies 16 blocks of
// 4 x 1 by 1 x 4: the B operand is each lane's own dosage (the lane is one
// person of the block), the A operand holds the weights of 16 rows -- lane
// 4q + i carries score 4c + i of row q -- and the broadcast modifier
// (CBSZ 4, ABID q) hands row q's four weights to all 16 blocks.  Lane l then
// accumulates score 4c + i of its person in output register i.  One A
// register serves 16 rows, so the weights cost one load per 16 rows per
// four scores.
//
// Decode, per lane and row: the four 2-bit codes are spread into bytes
// (two shift-or steps), the row's flip selector is OR-ed in, and one
// v_perm_b32 maps (flip, code) to four dosage bytes, another to four missing
// flags; the dosage bytes convert to FP32 with v_cvt_f32_ubyte{0..3}.
// Missing calls are counted as packed bytes, at most kF32FlushRows per block.
// ---------------------------------------------------------------------------
constexpr int kF32Group = 16;                       // rows per A operand (ABID 0..15)
constexpr int kF32FlushGroups = 14;                 // 224-row FP32 blocks (byte counters stay below 256)
constexpr int kF32FlushRows = kF32Group * kF32FlushGroups;
constexpr uint32_t kLutA2 = 0x02010000u;            // code 00,01,10,11 -> A2 dosage 0,0(missing),1,2
constexpr uint32_t kLutA1 = 0x00010002u;            // the same codes read for the A1 allele: 2,0,1,0
constexpr uint32_t kLutMiss = 0x00000100u;          // code 01 -> 1
constexpr int nc1(int ncls) { return ncls > 0 ? ncls : 1; }

// Where the rows of a group of 16 program rows come from, and their weights
// and flags as loaded.  A group is multiplied one group after its loads
// were issued; row q's byte of the next group is loaded into the register
// of row q as soon as row q has been multiplied (one byte register per row,
// a 16-row prefetch distance).
template <int KC>
struct F32Src {
    uint32_t g0;                  // first row (offset in the split)
    __amdgpu_buffer_rsrc_t rr;    // IDENT: the group's first row (clamped)
    uint32_t last;                // IDENT: byte offset from it of the group's last row inside the split
    uint32_t vl;                  // not IDENT: lane q < 16 holds row q's block row index
    float w[KC / 4];              // lane 4q + i: score 4c + i of row q (rows past the split masked later)
    uint32_t fw0, fw1, fsh;       // flip words holding the group's rows, bit of row g0
    uint32_t pc;                  // lane q < 16: class bits of row q
};

template <int Q, int KC, int NCLS>
__device__ __forceinline__ void fp32_row(uint32_t b, uint32_t flips, uint32_t cnt0, const float (&w)[KC / 4],
                                         uint32_t g0, const uint32_t *__restrict__ pcls, uint32_t nrows,
                                         uint32_t vmask, f32x4_t (&acc)[4][KC / 4], uint32_t (&cnt4)[nc1(NCLS)]) {
    uint32_t y = b | (b << 6);
    y = y | (y << 12);
    const uint32_t fsel = (uint32_t)__builtin_amdgcn_sbfe((int)flips, Q, 1) & 0x04040404u;
    // (y & vmask) | fsel in one v_bitop3 (truth table 0xEA over y, vmask,
    // fsel); the code mask sits in a vector register, so the scalar flip
    // selector is the instruction's one scalar operand
    const uint32_t sel = __builtin_amdgcn_bitop3_b32(y, vmask, fsel, 0xEA);
    const uint32_t d4 = __builtin_amdgcn_perm(kLutA1, kLutA2, sel);
    const float d[4] = {(float)(d4 & 0xFFu), (float)((d4 >> 8) & 0xFFu), (float)((d4 >> 16) & 0xFFu),
                        (float)(d4 >> 24)};
#pragma unroll
    for (int c = 0; c < KC / 4; ++c)
#pragma unroll
        for (int t = 0; t < 4; ++t) acc[t][c] = __builtin_amdgcn_mfma_f32_4x4x1f32(w[c], d[t], acc[t][c], 4, Q, 0);
    if constexpr (NCLS == 1) {
        // y & vmask & cm (truth table 0x80): the codes of a count row, all
        // zero otherwise -- a selector of 0 picks byte 0 of the LUT, which is 0
        const uint32_t cm = (uint32_t)__builtin_amdgcn_sbfe((int)cnt0, Q, 1);
        cnt4[0] += __builtin_amdgcn_perm(kLutMiss, kLutMiss, __builtin_amdgcn_bitop3_b32(y, vmask, cm, 0x80));
    } else if constexpr (NCLS > 1) {
        const uint32_t o = g0 + Q;
        const uint32_t cls = o < nrows ? pcls[o] : 0u;
        const uint32_t m4 = __builtin_amdgcn_perm(kLutMiss, kLutMiss, sel);
#pragma unroll
        for (int c = 0; c < NCLS; ++c) cnt4[c] += ((cls >> c) & 1u) ? m4 : 0u;
    }
}

// Grid: one workgroup per (score pass, 1024-person block, variant split),
// numbered so that the passes and person blocks of one split are consecutive
// workgroups on one XCD (workgroup i runs on XCD i mod 8): the passes of a
// row strip then read it through one L2.  Row positions inside a split are
// 32-bit offsets from its first row (the scalar unit has no 64-bit ordered
// compare).  IDENT: the block's program rows are its variants in order, so
// row o of the split sits at rows + (base - p0 + o) * stride; otherwise the
// wave loads 16 rows' variant indices into lanes and reads them back one by
// one.  Rows past the split read its last row; they are multiplied with
// weight 0, no flip and no count.
// Held to 128 registers (four waves per SIMD; the FP64 sums and the
// accumulators then share the VGPRs, a few values outside the row loop are
// spilled): at 100k x 1M x 8 27.0 against 28.1 ms with three.  The
// multi-class count instances keep the compiler's choice (their count
// registers would spill inside the loop).
template <int KC, int NCLS, bool IDENT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(NCLS <= 1 || KC == 4 ? 4 : 1))) void k_score_fp32(DevicePlan plan, const uint8_t *__restrict__ rows,
                                                    uint64_t stride, uint64_t bpv, uint64_t v0, uint64_t p0,
                                                    uint64_t p1, uint64_t rps, uint32_t npass, uint32_t npb,
                                                    uint32_t nsplit, uint64_t num_people, uint64_t num_scores,
                                                    double *__restrict__ part, uint32_t *__restrict__ cpart) {
    const uint32_t total = npass * npb * nsplit;
    const uint32_t per_xcd = (total + 7) / 8;
    const uint32_t logical = (blockIdx.x % 8) * per_xcd + blockIdx.x / 8;
    if (logical >= total) return;
    const uint32_t pass = logical % npass;
    const uint32_t pb = (logical / npass) % npb;
    const uint32_t s = logical / (npass * npb);
    const int tid = threadIdx.x, lane = tid & 63;
    const uint32_t col = pb * 256u + (uint32_t)tid;
    const bool colok = col < bpv;
    // lanes past the row end read the last byte; their people are never written
    const uint32_t lcol = colok ? col : (uint32_t)(bpv - 1);
    const uint64_t sb = p0 + (uint64_t)s * rps;
    const uint64_t base = sb < p1 ? sb : p1;
    const uint32_t nrows = (uint32_t)(((sb + rps) < p1 ? (sb + rps) : p1) - base);
    const uint64_t kw = plan.k32;
    const uint64_t nfw = (plan.num_prog_rows + 31) / 32;   // flip words
    const uint32_t *__restrict__ pvar = plan.prog_var + base;
    const uint32_t *__restrict__ pcls = plan.prog_cls + base;
    const uint32_t *__restrict__ fmask = plan.flipmask;
    const float *__restrict__ wcol = plan.w32 + base * kw + (uint64_t)pass * KC + (lane & 3);
    const uint8_t *__restrict__ rows0 = rows + (IDENT ? (base - p0) * stride : 0);
    // the code mask as a per-lane value (never 1: col stays far below 2^32 - 16)
    const uint32_t vmask = 0x03030303u + (uint32_t)(col > 0xFFFFFFF0u);

    f32x4_t acc[4][KC / 4];
    double accd[4][KC];
    uint32_t cnt4[nc1(NCLS)], cntu[4][nc1(NCLS)];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
#pragma unroll
        for (int k = 0; k < KC; ++k) accd[t][k] = 0.0;
#pragma unroll
        for (int c = 0; c < nc1(NCLS); ++c) cntu[t][c] = 0;
    }

    // everything of group g0 but its row bytes
    auto source = [&](uint32_t g0, F32Src<KC> &x) {
        x.g0 = g0;
        const uint32_t ol = g0 + (uint32_t)(lane & 15);
        const uint32_t olc = ol < nrows ? ol : nrows - 1;
        if constexpr (IDENT) {
            const uint32_t gc = g0 < nrows ? g0 : nrows - 1;
            const uint32_t gl = nrows - 1 < g0 + (kF32Group - 1) ? nrows - 1 : g0 + (kF32Group - 1);
            x.rr = __builtin_amdgcn_make_buffer_rsrc((void *)(rows0 + (uint64_t)gc * stride), 0, 0x7FFFFFFF,
                                                     kRsrcWord3);
            x.last = (gl - gc) * (uint32_t)stride;
        } else {
            x.vl = pvar[olc] - (uint32_t)v0;
        }
        const uint32_t wo = g0 + (uint32_t)(lane >> 2);
        const uint32_t woc = wo < nrows ? wo : nrows - 1;
#pragma unroll
        for (int c = 0; c < KC / 4; ++c) x.w[c] = wcol[(uint64_t)woc * kw + 4 * c];
        const uint64_t r0 = base + (g0 < nrows ? g0 : nrows - 1);
        const uint64_t wi = r0 >> 5;
        x.fw0 = fmask[wi];
        x.fw1 = fmask[wi + 1 < nfw ? wi + 1 : wi];
        x.fsh = (uint32_t)(r0 & 31);
        if constexpr (NCLS == 1) x.pc = pass == 0 ? pcls[olc] : 0u;
    };
    auto load_byte = [&](const F32Src<KC> &x, int q) -> uint32_t {
        uint32_t v;
        if constexpr (IDENT) {
            const uint32_t so = (uint32_t)q * (uint32_t)stride;
            v = __builtin_amdgcn_raw_buffer_load_b8(x.rr, lcol, so < x.last ? so : x.last, 0);
        } else {
            const uint32_t var = (uint32_t)__builtin_amdgcn_readlane((int)x.vl, q);
            const __amdgpu_buffer_rsrc_t rr = __builtin_amdgcn_make_buffer_rsrc(
                (void *)(rows + (uint64_t)var * stride), 0, 0x7FFFFFFF, kRsrcWord3);
            v = __builtin_amdgcn_raw_buffer_load_b8(rr, lcol, 0, 0);
        }
        __builtin_assume(v <= 0xFFu);
        return v;
    };
    // multiply group cur (bytes in b), loading the next group's rows into b
    auto run_group = [&]<int... Qs>(std::integer_sequence<int, Qs...>, const F32Src<KC> &cur,
                                    const F32Src<KC> &nxt, uint32_t (&b)[kF32Group]) {
        const uint32_t left = cur.g0 < nrows ? nrows - cur.g0 : 0u;
        const uint32_t in = left >= 32u ? 0xFFFFFFFFu : ((1u << left) - 1u);
        const uint32_t flips = (uint32_t)((((uint64_t)cur.fw1 << 32) | cur.fw0) >> cur.fsh) & in;
        uint32_t cnt0 = 0;
        if constexpr (NCLS == 1)
            cnt0 = (uint32_t)__builtin_amdgcn_ballot_w64(lane < 16 && ((in >> (lane & 15)) & 1u) && (cur.pc & 1u));
        const bool wrow_in = (in >> (lane >> 2)) & 1u;   // lane 4q + i carries row q's weights
        float w[KC / 4];
#pragma unroll
        for (int c = 0; c < KC / 4; ++c) w[c] = wrow_in ? cur.w[c] : 0.0f;
        ((fp32_row<Qs, KC, NCLS>(b[Qs], flips, cnt0, w, cur.g0, pcls, nrows, vmask, acc, cnt4),
          b[Qs] = load_byte(nxt, Qs)), ...);
    };

    F32Src<KC> sa, sn;
    uint32_t b[kF32Group];
    if (nrows) {
        source(0, sa);
#pragma unroll
        for (int q = 0; q < kF32Group; ++q) b[q] = load_byte(sa, q);
    }
    // outer loop over FP32 blocks, inner loop over their groups: the FP64
    // flush sits after the inner loop, so the accumulators stay in
    // accumulation registers inside it
    for (uint32_t blk0 = 0; blk0 < nrows; blk0 += kF32FlushRows) {
#pragma unroll
        for (int t = 0; t < 4; ++t)
#pragma unroll
            for (int c = 0; c < KC / 4; ++c) acc[t][c] = f32x4_t{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int c = 0; c < nc1(NCLS); ++c) cnt4[c] = 0u;
        const uint32_t blk1 = nrows - blk0 > (uint32_t)kF32FlushRows ? blk0 + kF32FlushRows : nrows;
        for (uint32_t g0 = blk0; g0 < blk1; g0 += kF32Group) {
            source(g0 + kF32Group, sn);
            run_group(std::make_integer_sequence<int, kF32Group>{}, sa, sn, b);
            sa = sn;
        }
#pragma unroll
        for (int t = 0; t < 4; ++t) {
#pragma unroll
            for (int c = 0; c < KC / 4; ++c)
#pragma unroll
                for (int i = 0; i < 4; ++i) accd[t][4 * c + i] += (double)acc[t][c][i];
#pragma unroll
            for (int c = 0; c < nc1(NCLS); ++c) cntu[t][c] += (cnt4[c] >> (8 * t)) & 0xFFu;
        }
    }
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const uint64_t p = (uint64_t)col * 4 + t;
        if (!colok || p >= num_people) continue;
        double *dst = part + ((uint64_t)s * num_people + p) * num_scores + (uint64_t)pass * KC;
#pragma unroll
        for (int k = 0; k < KC; ++k)
            if ((uint64_t)pass * KC + k < num_scores) dst[k] = accd[t][k];
        if (NCLS > 0 && pass == 0) {
            uint32_t *cd = cpart + ((uint64_t)s * num_people + p) * plan.nclass;
#pragma unroll
            for (int c = 0; c < nc1(NCLS); ++c)
                if ((uint32_t)c < plan.nclass) cd[c] = cntu[t][c];
        }
    }
}

// scores per pass of the FP32 kernel for num_scores scores
static uint32_t fp32_pass_scores(uint64_t num_scores) { return num_scores <= 4 ? 4u : 8u; }

uint64_t fp32_weight_pitch(uint64_t num_scores) {
    const uint64_t kc = fp32_pass_scores(num_scores);
    return (num_scores + kc - 1) / kc * kc;
}

hipError_t launch_score_fp32(const DevicePlan &plan, const uint8_t *rows, uint64_t stride,
                             uint64_t bpv, uint64_t v0, uint64_t p0, uint64_t p1,
                             uint64_t num_people, uint64_t num_scores, uint32_t splits,
                             bool count_in_kernel, bool identity, double *part, uint32_t *cpart,
                             hipStream_t stream) {
    const uint64_t cols = (num_people + 3) / 4;
    const uint32_t kc = fp32_pass_scores(num_scores);
    if (plan.k32 != fp32_weight_pitch(num_scores)) return hipErrorInvalidValue;
    const uint32_t npass = (uint32_t)((num_scores + kc - 1) / kc);
    const uint32_t npb = (uint32_t)((cols + 255) / 256);
    const uint64_t rps = (p1 - p0 + splits - 1) / splits;
    const uint64_t total = (uint64_t)npass * npb * splits;
    // the row step of a 16-row group must fit the 32-bit scalar offset of a buffer load
    if (total >= (1ull << 31) || bpv == 0 || bpv > 0x7FFFFFFFull || stride > (0x7FFFFFFFull / kF32Group) ||
        rps >= (1ull << 32) || p1 < p0)
        return hipErrorInvalidValue;
    const int ncls = count_in_kernel ? (plan.nclass <= 1 ? 1 : kFp32MaxClasses) : 0;
    const dim3 grid((unsigned)((total + 7) / 8 * 8)), block(256);
#define GNM_F32(KC, NC, ID)                                                                                        \
    hipLaunchKernelGGL((k_score_fp32<KC, NC, ID>), grid, block, 0, stream, plan, rows, stride, bpv, v0, p0, p1,    \
                       rps, npass, npb, splits, num_people, num_scores, part, cpart)
#define GNM_F32_ID(KC, NC)                                                                                         \
    do {                                                                                                           \
        if (identity) GNM_F32(KC, NC, true);                                                                       \
        else GNM_F32(KC, NC, false);                                                                               \
    } while (0)
    if (kc == 4) {
        if (ncls == 1) GNM_F32_ID(4, 1);
        else if (ncls == 0) GNM_F32_ID(4, 0);
        else GNM_F32_ID(4, kFp32MaxClasses);
    } else {
        if (ncls == 1) GNM_F32_ID(8, 1);
        else if (ncls == 0) GNM_F32_ID(8, 0);
        else GNM_F32_ID(8, kFp32MaxClasses);
    }
#undef GNM_F32_ID
#undef GNM_F32
    return hipGetLastError();
}

static uint32_t splits_for(uint64_t wgs_per_split) {
    // enough workgroups to fill 256 CUs several times; a multiple of 8 so
    // every XCD owns whole splits
    uint64_t s = (3072 + wgs_per_split - 1) / (wgs_per_split ? wgs_per_split : 1);
    s = (s + 7) / 8 * 8;
    return (uint32_t)(s < 8 ? 8 : (s > 64 ? 64 : s));
}

uint32_t fp32_splits(uint64_t num_people, uint64_t num_scores) {
    const uint64_t pt = ((num_people + 3) / 4 + 255) / 256;
    const uint64_t passes = (num_scores + fp32_pass_scores(num_scores) - 1) / fp32_pass_scores(num_scores);
    return splits_for(pt * passes);
}

// ---------------------------------------------------------------------------
// Cross-split FP64 reduction into the persistent accumulators (fixed order,
// deterministic), and the per-score count gather from membership classes.
// ---------------------------------------------------------------------------
// One thread per (person, score); the sum runs split by split, limb by limb
// inside a split, the same order on every launch.  LC > 0: the limb count
// at compile time, and four splits' loads are issued before their adds (the
// order of the adds is unchanged), so more loads are in flight per thread.
template <int LC>
__global__ void k_reduce(const double *__restrict__ part, const uint32_t *__restrict__ cpart,
                         uint32_t splits, uint64_t num_people, uint64_t num_scores, uint32_t limbs,
                         uint32_t nclass, const uint32_t *__restrict__ score_class,
                         double *__restrict__ scores, uint32_t *__restrict__ counts) {
    const uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= num_people * num_scores) return;
    const uint64_t p = idx / num_scores, k = idx % num_scores;
    const uint32_t L = LC > 0 ? (uint32_t)LC : limbs;
    const uint64_t ncol = (uint64_t)L * num_scores;
    double acc = 0.0;
    uint32_t cnt = 0;
    const uint32_t cls = score_class[k];
    uint32_t s = 0;
    if constexpr (LC > 0) {
        for (; s + 4 <= splits; s += 4) {
            double v[4][LC];
            uint32_t c[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const double *ps = part + ((uint64_t)(s + u) * num_people + p) * ncol + k;
#pragma unroll
                for (int l = 0; l < LC; ++l) v[u][l] = ps[(uint64_t)l * num_scores];
                c[u] = cpart[((uint64_t)(s + u) * num_people + p) * nclass + cls];
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
#pragma unroll
                for (int l = 0; l < LC; ++l) acc += v[u][l];
                cnt += c[u];
            }
        }
    }
    for (; s < splits; ++s) {
        const double *ps = part + ((uint64_t)s * num_people + p) * ncol + k;
        for (uint32_t l = 0; l < L; ++l) acc += ps[(uint64_t)l * num_scores];
        cnt += cpart[((uint64_t)s * num_people + p) * nclass + cls];
    }
    scores[idx] += acc;
    counts[idx] += cnt;
}

hipError_t launch_reduce(const double *part, const uint32_t *cpart, uint32_t splits,
                         uint64_t num_people, uint64_t num_scores, uint32_t limbs,
                         uint32_t nclass, const uint32_t *score_class, double *scores,
                         uint32_t *counts, hipStream_t stream) {
    const uint64_t n = num_people * num_scores;
    if (n == 0) return hipSuccess;
    const dim3 grid((unsigned)((n + 255) / 256)), block(256);
#define GNM_REDUCE(LC)                                                                                       \
    hipLaunchKernelGGL((k_reduce<LC>), grid, block, 0, stream, part, cpart, splits, num_people, num_scores, \
                       limbs, nclass, score_class, scores, counts)
    if (limbs == 1) GNM_REDUCE(1);
    else if (limbs == 2) GNM_REDUCE(2);
    else if (limbs == 3) GNM_REDUCE(3);
    else if (limbs == 4) GNM_REDUCE(4);
    else GNM_REDUCE(0);
#undef GNM_REDUCE
    return hipGetLastError();
}

__global__ void k_fill_f64(double *p, uint64_t n, double v) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) p[i] = v;
}
__global__ void k_fill_u32(uint32_t *p, uint64_t n, uint32_t v) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) p[i] = v;
}

hipError_t launch_fill_f64(double *p, uint64_t n, double value, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    const uint64_t blocks = (n + 255) / 256 < 65536 ? (n + 255) / 256 : 65536;
    hipLaunchKernelGGL(k_fill_f64, dim3((unsigned)blocks), dim3(256), 0, stream, p, n, value);
    return hipGetLastError();
}
hipError_t launch_fill_u32(uint32_t *p, uint64_t n, uint32_t value, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    const uint64_t blocks = (n + 255) / 256 < 65536 ? (n + 255) / 256 : 65536;
    hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)blocks), dim3(256), 0, stream, p, n, value);
    return hipGetLastError();
}

// HBM ceiling probes: the access shapes of the unpack stage without its
// decode, so that its rate can be read against what the memory system
// delivers for the same shapes.  Grid-stride over 16-byte elements (n16 =
// bytes / 16 of the source) or dwords (n4), kProbeInFlight independent
// accesses per thread and iteration.
//   read:   16-byte non-temporal loads, folded into one dword per thread
//   write:  16-byte non-temporal buffer stores (the unpack stage's stores)
//   expand: every source dword becomes 16 bytes of output (1:4, the
//           unpack stage's ratio of packed bytes to tile bytes)
// Shape: a sweep over 16 GiB (grids of 1024 to 65536 workgroups of 256
// threads, one or four stores per iteration, plain or non-temporal) put
// writes at 4.4-5.8 TB/s and reads at 6.5-7.0 TB/s; 16384 workgroups with
// eight reads or four writes in flight gave the top of both ranges.
constexpr int kProbeGrid = 16384;
constexpr int kProbeInFlight = 8;
__global__ __launch_bounds__(256) void k_probe_read(const uint8_t *__restrict__ src, uint64_t n16,
                                                     uint32_t *__restrict__ sink) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x, nt = (uint64_t)gridDim.x * blockDim.x;
    u32x4_t acc = {0u, 0u, 0u, 0u};
    for (uint64_t i = t; i < n16; i += kProbeInFlight * nt) {
        u32x4_t v[kProbeInFlight];
#pragma unroll
        for (int u = 0; u < kProbeInFlight; ++u) {
            const uint64_t j = i + u * nt;
            v[u] = j < n16 ? __builtin_nontemporal_load((const u32x4_t *)(src + 16 * j)) : u32x4_t{0u, 0u, 0u, 0u};
        }
#pragma unroll
        for (int u = 0; u < kProbeInFlight; ++u) acc ^= v[u];
    }
    sink[t] = acc[0] ^ acc[1] ^ acc[2] ^ acc[3];
}
__global__ __launch_bounds__(256) void k_probe_write(uint8_t *__restrict__ dst, uint64_t n16, uint32_t value) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x, nt = (uint64_t)gridDim.x * blockDim.x;
    const u32x4_t v = {value, value ^ 1u, value ^ 2u, value ^ 3u};
    for (uint64_t i = t; i < n16; i += 4 * nt) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const uint64_t j = i + u * nt;
            if (j >= n16) break;
            const __amdgpu_buffer_rsrc_t r =
                __builtin_amdgcn_make_buffer_rsrc(dst + 16 * (j - j % 65536), 0, 0x7FFFFFFF, kRsrcWord3);
            __builtin_amdgcn_raw_buffer_store_b128(v, r, (uint32_t)(16 * (j % 65536)), 0, kStoreNT);
        }
    }
}
__global__ __launch_bounds__(256) void k_probe_expand(const uint8_t *__restrict__ src, uint64_t n4,
                                                       uint8_t *__restrict__ dst) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x, nt = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = t; i < n4; i += kProbeInFlight * nt) {
        uint32_t x[kProbeInFlight];
#pragma unroll
        for (int u = 0; u < kProbeInFlight; ++u) {
            const uint64_t j = i + u * nt;
            x[u] = j < n4 ? __builtin_nontemporal_load((const uint32_t *)(src + 4 * j)) : 0u;
        }
#pragma unroll
        for (int u = 0; u < kProbeInFlight; ++u) {
            const uint64_t j = i + u * nt;
            if (j >= n4) break;
            const u32x4_t o = {x[u] & 0x03030303u, (x[u] >> 2) & 0x03030303u, (x[u] >> 4) & 0x03030303u,
                               (x[u] >> 6) & 0x03030303u};
            const __amdgpu_buffer_rsrc_t r =
                __builtin_amdgcn_make_buffer_rsrc(dst + 16 * (j - j % 65536), 0, 0x7FFFFFFF, kRsrcWord3);
            __builtin_amdgcn_raw_buffer_store_b128(o, r, (uint32_t)(16 * (j % 65536)), 0, kStoreNT);
        }
    }
}

uint64_t hbm_probe_sink_bytes() { return (uint64_t)kProbeGrid * 256 * 4; }

hipError_t launch_hbm_probe(int kind, const uint8_t *src, uint64_t src_bytes, uint8_t *dst, hipStream_t stream) {
    const dim3 grid(kProbeGrid), block(256);   // 64 workgroups per CU, grid-stride
    if (kind == 0)
        hipLaunchKernelGGL(k_probe_read, grid, block, 0, stream, src, src_bytes / 16, (uint32_t *)dst);
    else if (kind == 1)
        hipLaunchKernelGGL(k_probe_write, grid, block, 0, stream, dst, src_bytes / 16, 0x5A5A5A5Au);
    else
        hipLaunchKernelGGL(k_probe_expand, grid, block, 0, stream, src, src_bytes / 4, dst);
    return hipGetLastError();
}

}  // namespace gnm
