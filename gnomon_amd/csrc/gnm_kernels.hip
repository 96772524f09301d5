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
constexpr int kUnpackBatch = 8;          // rows loaded before any decode
constexpr int kUnpackBlock = 256;        // 4 waves
constexpr int kUnpackRows = 512;         // rows per workgroup (<= 65535 for the pair counters)
constexpr int kUnpackBytesRows = 2048;   // rows per workgroup of the byte-wise path
constexpr int kRsrcWord3 = 0x00020000;   // gfx9 raw buffer descriptor: 32-bit data format
constexpr int kStoreNT = 2;              // buffer store cache policy: non-temporal

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
    const uint64_t wave = (uint64_t)blockIdx.x * (BS / 64) + (threadIdx.x >> 6);
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
    const uint64_t r0 = (uint64_t)blockIdx.y * kUnpackRows;
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
        // flip flags of the batch: one uniform 8-byte load (r % 8 == 0)
        const uint64_t fw = flip ? *(const uint64_t *)(flip + r) : 0;
        __builtin_amdgcn_sched_barrier(0);   // all loads of the batch issue before its first store
#pragma unroll
        for (int i = 0; i < R; ++i) row(w[i], 0u - (uint32_t)(((fw >> (8 * i)) & 0xFFu) != 0), r + i);
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
        hipLaunchKernelGGL(k_unpack, dim3((unsigned)((nd + per_block - 1) / per_block),
                                          (unsigned)((num_rows + kUnpackRows - 1) / kUnpackRows)),
                           dim3(kUnpackBlock), 0, stream, rows, stride, num_rows, num_people, nd, flip, dosage,
                           missing);
    } else {
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
// Plain FP32 path: a wavefront reduction with people on lanes (four per lane,
// one packed byte) and the row's weights broadcast from scalar registers.
// FP32 weights, exact products (dosage 0/1/2 times an FP32 weight is exact in
// FP64 as in FP32) and FP64 accumulation per genotype: v_fma_f64 runs at the
// same rate as unpacked v_fma_f32 on CDNA4, and it removes the FP32 partial-sum
// rounding of the reference kernel's 256-variant mini-batches
// (score/kernel.rs:124-176, score/batch.rs:34, 387-419), which alone exceeds
// 1e-6 relative at 1e5-1e6 variants.
// ---------------------------------------------------------------------------
template <int KC>
__global__ __launch_bounds__(256) void k_score_fp32(DevicePlan plan, const uint8_t *__restrict__ rows,
                                                    uint64_t stride, uint64_t bpv, uint64_t v0,
                                                    uint64_t p0, uint64_t p1, uint64_t split_base,
                                                    uint64_t rps, uint64_t num_people,
                                                    uint64_t num_scores, int count_in_kernel,
                                                    double *__restrict__ part,
                                                    uint32_t *__restrict__ cpart) {
    const uint32_t s = blockIdx.y;
    const uint32_t kc = blockIdx.z;
    const uint64_t col = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool colok = col < bpv;
    const uint64_t sb = split_base + (uint64_t)s * rps;
    const uint64_t r_begin = p0 > sb ? p0 : sb;
    const uint64_t se = sb + rps;
    const uint64_t r_end = p1 < se ? p1 : se;
    const bool do_count = count_in_kernel && kc == 0;
    const double *__restrict__ wbase = plan.w64 + (uint64_t)kc * KC;

    double acc[4][KC];
    uint32_t cnt[4][kFp32MaxClasses];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
#pragma unroll
        for (int k = 0; k < KC; ++k) acc[t][k] = 0.0;
#pragma unroll
        for (int c = 0; c < kFp32MaxClasses; ++c) cnt[t][c] = 0;
    }
    // eight rows per step, their loads issued together and one step ahead
    constexpr int kRB = 8;
    uint32_t nb[kRB];
    auto load_rows = [&](uint64_t r0, uint32_t *out) {
#pragma unroll
        for (int i = 0; i < kRB; ++i) {
            const uint64_t r = r0 + i;
            uint32_t b = 0u;  // rows past the split: code 00 unflipped, weight rows unused
            if (r < r_end) {
                const uint64_t var = (uint64_t)(plan.prog_var[r] - (uint32_t)v0);
                b = colok ? rows[var * stride + col] : 0x55u;
                if ((plan.flipmask[r >> 5] >> (r & 31)) & 1u) b = flip_byte(b);
            }
            out[i] = b;
        }
    };
    if (r_begin < r_end) load_rows(r_begin, nb);
    for (uint64_t r0 = r_begin; r0 < r_end; r0 += kRB) {
        uint32_t cb[kRB];
#pragma unroll
        for (int i = 0; i < kRB; ++i) cb[i] = nb[i];
        if (r0 + kRB < r_end) load_rows(r0 + kRB, nb);
#pragma unroll
        for (int i = 0; i < kRB; ++i) {
            const uint64_t r = r0 + i;
            if (r >= r_end) break;
            const uint32_t b = cb[i];
            const double *__restrict__ w = wbase + r * plan.k32;
            double wk[KC];
#pragma unroll
            for (int k = 0; k < KC; ++k) wk[k] = w[k];
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const uint32_t c = (b >> (2 * t)) & 3u;
                const double d = (double)((c >> 1) * ((c & 1u) + 1u));
#pragma unroll
                for (int k = 0; k < KC; ++k) acc[t][k] = __builtin_fma(d, wk[k], acc[t][k]);
            }
            if (do_count) {
                const uint32_t cls = plan.prog_cls[r];
                if (cls) {
#pragma unroll
                    for (int t = 0; t < 4; ++t) {
                        const uint32_t m = (((b >> (2 * t)) & 3u) == 1u) ? 1u : 0u;
#pragma unroll
                        for (int c = 0; c < kFp32MaxClasses; ++c) cnt[t][c] += m & (cls >> c);
                    }
                }
            }
        }
    }
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const uint64_t p = col * 4 + t;
        if (p >= num_people) continue;
        double *dst = part + ((uint64_t)s * num_people + p) * num_scores + (uint64_t)kc * KC;
#pragma unroll
        for (int k = 0; k < KC; ++k)
            if ((uint64_t)kc * KC + k < num_scores) dst[k] = acc[t][k];
        if (do_count) {
            uint32_t *cd = cpart + ((uint64_t)s * num_people + p) * plan.nclass;
#pragma unroll
            for (int c = 0; c < kFp32MaxClasses; ++c)
                if ((uint32_t)c < plan.nclass) cd[c] = cnt[t][c];
        }
    }
}

hipError_t launch_score_fp32(const DevicePlan &plan, const uint8_t *rows, uint64_t stride,
                             uint64_t bpv, uint64_t v0, uint64_t p0, uint64_t p1,
                             uint64_t num_people, uint64_t num_scores, uint32_t splits,
                             bool count_in_kernel, double *part, uint32_t *cpart, hipStream_t stream) {
    const uint64_t cols = (num_people + 3) / 4;
    const uint64_t split_base = p0;
    const uint64_t rows_per_split = (p1 - p0 + splits - 1) / splits;
    dim3 grid((unsigned)((cols + 255) / 256), splits,
              (unsigned)((num_scores + kFp32ScoreChunk - 1) / kFp32ScoreChunk));
    hipLaunchKernelGGL(k_score_fp32<kFp32ScoreChunk>, grid, dim3(256), 0, stream, plan, rows,
                       stride, bpv, v0, p0, p1, split_base, rows_per_split, num_people,
                       num_scores, count_in_kernel ? 1 : 0, part, cpart);
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
    const uint64_t chunks = (num_scores + kFp32ScoreChunk - 1) / kFp32ScoreChunk;
    return splits_for(pt * chunks);
}

// ---------------------------------------------------------------------------
// Cross-split FP64 reduction into the persistent accumulators (fixed order,
// deterministic), and the per-score count gather from membership classes.
// ---------------------------------------------------------------------------
__global__ void k_reduce(const double *__restrict__ part, const uint32_t *__restrict__ cpart,
                         uint32_t splits, uint64_t num_people, uint64_t num_scores, uint32_t limbs,
                         uint32_t nclass, const uint32_t *__restrict__ score_class,
                         double *__restrict__ scores, uint32_t *__restrict__ counts) {
    const uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= num_people * num_scores) return;
    const uint64_t p = idx / num_scores, k = idx % num_scores;
    const uint64_t ncol = (uint64_t)limbs * num_scores;
    double acc = 0.0;
    uint32_t cnt = 0;
    const uint32_t cls = score_class[k];
    for (uint32_t s = 0; s < splits; ++s) {
        const double *ps = part + ((uint64_t)s * num_people + p) * ncol + k;
        for (uint32_t l = 0; l < limbs; ++l) acc += ps[(uint64_t)l * num_scores];
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
    hipLaunchKernelGGL(k_reduce, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, part, cpart,
                       splits, num_people, num_scores, limbs, nclass, score_class, scores, counts);
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

}  // namespace gnm
