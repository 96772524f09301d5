// Sparse ingest of host rows: the GPU side of the reference CPU engine's
// density dispatch (assess_variant_density_for_dispatch and
// run_variant_major_path, score/batch.rs:583-760; the path decision of
// score/decide.rs:100-189).
//
// The reference's CPU engine skips the work of homozygous-A1 (00) calls of
// rare variants, 64-bit words at a time in its packed fast path
// (batch.rs:720-760).  On the device that work is not the cost (a packed row
// is read once at HBM speed whatever its content); the cost is moving host
// rows over PCIe.  So a host row whose 64-bit words (32 calls each) are
// mostly zero crosses PCIe as a bitmap of its non-zero words plus those
// words, and the device rebuilds the packed row (k_sparse_rows) before the
// dense kernels score it.  The rebuilt row equals the host row for every
// scored person, so the scores are bit-identical to sending every row
// packed.  The host encoder is branch-free per word and runs at copy speed.
#include "gnm_internal.hpp"

#include <algorithm>
#include <cstdlib>
#include <cstring>
#if !defined(__HIP_DEVICE_COMPILE__)
#include <immintrin.h>
#endif

namespace gnm {

// ---------------------------------------------------------------------------
// Host encoder
// ---------------------------------------------------------------------------
#if !defined(__HIP_DEVICE_COMPILE__)
namespace {

// 8-dword permutations that move the 64-bit lanes set in a 4-bit mask to the
// front, in order (AVX2 compress of non-zero words)
struct CompressTable {
    alignas(32) int32_t perm[16][8];
    CompressTable() {
        for (int m = 0; m < 16; ++m) {
            int k = 0;
            for (int j = 0; j < 4; ++j)
                if ((m >> j) & 1) {
                    perm[m][2 * k] = 2 * j;
                    perm[m][2 * k + 1] = 2 * j + 1;
                    ++k;
                }
            for (; k < 4; ++k) perm[m][2 * k] = perm[m][2 * k + 1] = 0;
        }
    }
};

// The same encoding four words at a time, branch-free: a zero test per
// 256-bit load, the non-zero words compressed to the front with one permute
// and stored whole (out holds max_words + 64), the count advancing by the
// number kept.
__attribute__((target("avx2,popcnt"))) int64_t encode_avx2(const uint8_t *row, uint64_t num_fam,
                                                            const uint64_t *keep2, uint64_t max_words,
                                                            uint32_t nch, uint64_t *out, uint64_t *bitmap,
                                                            uint32_t *chunk_start) {
    static const CompressTable table;
    const uint64_t full = num_fam / 32;
    const uint64_t rem = num_fam - 32 * full;
    const uint64_t words = full + (rem ? 1 : 0);
    uint64_t n = 0;
    for (uint32_t c = 0; c < nch; ++c) {
        chunk_start[c] = (uint32_t)n;
        for (uint64_t g = 0; g < kSparseChunkWords / 64; ++g) {
            const uint64_t w0 = (uint64_t)c * kSparseChunkWords + 64 * g;
            const uint64_t w1 = std::min(w0 + 64, full);
            uint64_t bm = 0, w = w0;
            for (; w + 4 <= w1; w += 4) {
                __m256i v = _mm256_loadu_si256((const __m256i *)(row + 8 * w));
                if (keep2) v = _mm256_and_si256(v, _mm256_loadu_si256((const __m256i *)(keep2 + w)));
                const __m256i z = _mm256_cmpeq_epi64(v, _mm256_setzero_si256());
                const unsigned nz = ~(unsigned)_mm256_movemask_pd(_mm256_castsi256_pd(z)) & 0xFu;
                const __m256i p = _mm256_load_si256((const __m256i *)table.perm[nz]);
                _mm256_storeu_si256((__m256i *)(out + n), _mm256_permutevar8x32_epi32(v, p));
                bm |= (uint64_t)nz << (w - w0);
                n += (uint64_t)__builtin_popcount(nz);
            }
            for (; w < w1; ++w) {
                uint64_t x;
                std::memcpy(&x, row + 8 * w, 8);
                if (keep2) x &= keep2[w];
                out[n] = x;
                bm |= (uint64_t)(x != 0) << (w - w0);
                n += (x != 0);
            }
            if (full < words && full >= w0 && full < w0 + 64) {   // last partial word: padding codes ignored
                uint64_t x = 0;
                std::memcpy(&x, row + 8 * full, (rem + 3) / 4);
                x &= (1ull << (2 * rem)) - 1;
                if (keep2) x &= keep2[full];
                out[n] = x;
                bm |= (uint64_t)(x != 0) << (full - w0);
                n += (x != 0);
            }
            bitmap[(uint64_t)c * (kSparseChunkWords / 64) + g] = bm;
            if (n > max_words) return -1;
        }
    }
    chunk_start[nch] = (uint32_t)n;
    return (int64_t)n;
}

}  // namespace
#endif

// Portable form, and the entry point: the AVX2 form where the host has it.
int64_t sparse_encode_row(const uint8_t *row, uint64_t num_fam, const uint64_t *keep2, uint64_t max_words,
                          uint32_t nch, uint64_t *out, uint64_t *bitmap, uint32_t *chunk_start) {
#if !defined(__HIP_DEVICE_COMPILE__)
    static const bool avx2 = __builtin_cpu_supports("avx2") && !std::getenv("GNM_SPARSE_SCALAR");
    if (avx2) return encode_avx2(row, num_fam, keep2, max_words, nch, out, bitmap, chunk_start);
#endif
    const uint64_t full = num_fam / 32;                  // words of 32 people
    const uint64_t rem = num_fam - 32 * full;            // people of the last partial word
    const uint64_t words = full + (rem ? 1 : 0);
    uint64_t n = 0;
    for (uint32_t c = 0; c < nch; ++c) {
        chunk_start[c] = (uint32_t)n;
        for (uint64_t g = 0; g < kSparseChunkWords / 64; ++g) {
            const uint64_t w0 = (uint64_t)c * kSparseChunkWords + 64 * g;
            const uint64_t w1 = std::min(w0 + 64, full);
            uint64_t bm = 0;
            for (uint64_t w = w0; w < w1; ++w) {
                uint64_t x;
                std::memcpy(&x, row + 8 * w, 8);
                if (keep2) x &= keep2[w];
                out[n] = x;                              // kept only when non-zero
                bm |= (uint64_t)(x != 0) << (w - w0);
                n += (x != 0);
            }
            if (full < words && full >= w0 && full < w0 + 64) {   // last partial word: padding codes ignored
                uint64_t x = 0;
                std::memcpy(&x, row + 8 * full, (rem + 3) / 4);
                x &= (1ull << (2 * rem)) - 1;
                if (keep2) x &= keep2[full];
                out[n] = x;
                bm |= (uint64_t)(x != 0) << (full - w0);
                n += (x != 0);
            }
            bitmap[(uint64_t)c * (kSparseChunkWords / 64) + g] = bm;
            if (n > max_words) return -1;            // out holds max_words + 64
        }
    }
    chunk_start[nch] = (uint32_t)n;
    return (int64_t)n;
}

// ---------------------------------------------------------------------------
// Device kernels
// ---------------------------------------------------------------------------

// out row r = pack row map[r] (dense), zeros (map[r] = kMapZero), or left
// to k_sparse_rows (map[r] = kMapSparse); 16-byte units
__global__ __launch_bounds__(256) void k_expand_rows(const uint4 *__restrict__ pack, const uint32_t *__restrict__ map,
                                                     uint64_t n, uint64_t pitch16, uint4 *__restrict__ out) {
    const uint64_t c = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (c >= pitch16) return;
    for (uint64_t r = blockIdx.y; r < n; r += gridDim.y) {
        const uint32_t src = map[r];
        if (src == kMapSparse) continue;
        uint4 v = make_uint4(0u, 0u, 0u, 0u);
        if (src != kMapZero) v = pack[(uint64_t)src * pitch16 + c];
        out[r * pitch16 + c] = v;
    }
}

// Workgroup (s, c) rebuilds chunk c (512 words, 4096 bytes) of sparse row s
// from its record at word bmoff[s] of the upload: nch*8 bitmap words, then
// the stored words.  Thread t writes words 2t and 2t+1, each the next stored
// word when its bitmap bit is set (its index: the chunk's first stored word
// plus the set bits before it), else zero.
__global__ __launch_bounds__(256) void k_sparse_rows(const uint64_t *__restrict__ pack,
                                                     const uint64_t *__restrict__ bmoff,
                                                     const uint32_t *__restrict__ sv_row,
                                                     const uint32_t *__restrict__ cs, uint32_t nch, uint64_t pitch,
                                                     uint8_t *__restrict__ out) {
    constexpr uint32_t kGroups = kSparseChunkWords / 64;
    const uint64_t s = blockIdx.x;
    uint8_t *row = out + (uint64_t)sv_row[s] * pitch;
    const uint64_t *rec = pack + bmoff[s];
    const uint64_t *words = rec + (uint64_t)nch * kGroups;
    const uint32_t t = threadIdx.x;
    const uint32_t g = (2 * t) / 64, bit = (2 * t) % 64;   // both words sit in bitmap word g
    for (uint32_t c = blockIdx.y; c < nch; c += gridDim.y) {
        const uint64_t b0 = (uint64_t)c * (kSparseChunkWords * 8) + 16 * t;   // byte of word 2t in the row
        if (b0 >= pitch) continue;
        const uint64_t *bm = rec + (uint64_t)c * kGroups;
        uint32_t before = 0;
        for (uint32_t q = 0; q < g; ++q) before += (uint32_t)__popcll(bm[q]);
        const uint64_t m = bm[g];
        before += (uint32_t)__popcll(m & ((1ull << bit) - 1));
        const uint64_t e = (uint64_t)cs[s * (nch + 1) + c] + before;
        const bool h0 = (m >> bit) & 1ull, h1 = (m >> (bit + 1)) & 1ull;
        const uint64_t x0 = h0 ? words[e] : 0ull;
        const uint64_t x1 = h1 ? words[e + (h0 ? 1 : 0)] : 0ull;
        *(uint4 *)(row + b0) = make_uint4((uint32_t)x0, (uint32_t)(x0 >> 32), (uint32_t)x1, (uint32_t)(x1 >> 32));
    }
}

hipError_t launch_expand_rows(const uint8_t *pack, const uint32_t *map, uint64_t n, uint64_t pitch, uint8_t *out,
                              hipStream_t stream) {
    if (n == 0) return hipSuccess;
    const uint64_t pitch16 = pitch / 16;
    dim3 grid((unsigned)((pitch16 + 255) / 256), (unsigned)(n < 8192 ? n : 8192));
    hipLaunchKernelGGL(k_expand_rows, grid, dim3(256), 0, stream, (const uint4 *)pack, map, n, pitch16, (uint4 *)out);
    return hipGetLastError();
}

hipError_t launch_sparse_rows(const uint64_t *pack, const uint64_t *bmoff, const uint32_t *sv_row, const uint32_t *cs,
                              uint64_t nsv, uint32_t nch, uint64_t pitch, uint8_t *out, hipStream_t stream) {
    if (nsv == 0) return hipSuccess;
    dim3 grid((unsigned)nsv, nch < 1024u ? nch : 1024u);
    hipLaunchKernelGGL(k_sparse_rows, grid, dim3(256), 0, stream, pack, bmoff, sv_row, cs, nch, pitch, out);
    return hipGetLastError();
}

}  // namespace gnm
