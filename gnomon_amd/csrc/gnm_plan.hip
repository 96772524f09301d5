// Device build of the weight plan of a scoring context: the GPU counterpart
// of the reference's densification of the CSR weights into weight tiles
// (scatter_batch_mats, score/cuda_backend.rs:141-172, run per mega-batch
// there), done once per context here.
//
// Inputs are the CSR (W_eff, corr, columns, row offsets) and the host's
// cheap O(V + nnz) index work: each variant's effect-allele program rows
// (var_prog_off, var_rows: bit 0 an A2 row, bit 1 an A1 row) and each
// score's membership class.  Steps:
//   k_plan_rows     effect-allele weight rows wrow[r][k] in f64 (A2 row:
//                   W_eff + corr/2, A1 row: corr/2) and the count classes of
//                   each variant's count row;
//   k_plan_exps     per (scale block, score) the power-of-two scale 2^e with
//                   max |w| <= qmax * 2^e, and the limb scales;
//   k_plan_frag_i8  split-int8 B fragments, [k-step][tile][lane][16 i8];
//   k_plan_frag_bf16 split-BF16 B fragments, [stage][tile][k-step][lane][8 bf16];
//   k_plan_w32      FP32-path weights (W rounded once to f32).
#include "gnm_internal.hpp"

#include <cmath>

namespace gnm {

__global__ __launch_bounds__(256) void k_plan_rows(const uint64_t *__restrict__ off, const uint32_t *__restrict__ cols,
                                                   const float *__restrict__ w, const float *__restrict__ co,
                                                   const uint64_t *__restrict__ prog_off,
                                                   const uint8_t *__restrict__ var_rows, uint64_t V, uint64_t K,
                                                   const uint32_t *__restrict__ score_class, uint32_t cw,
                                                   double *__restrict__ wrow, uint32_t *__restrict__ rowcls) {
    const uint64_t v = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (v >= V) return;
    const uint8_t m = var_rows[v];
    if (!m) return;
    const uint64_t r0 = prog_off[v];
    const bool has_d = m & 1, has_e = m & 2;
    const uint64_t rd = r0, re = has_d ? r0 + 1 : r0;
    for (uint64_t e = off[v]; e < off[v + 1]; ++e) {
        const uint32_t k = cols[e];
        const double we = w[e], c = co[e];
        if (has_d) wrow[rd * K + k] += we + 0.5 * c;
        if (has_e) wrow[re * K + k] += 0.5 * c;
        if (rowcls) {
            const uint32_t cl = score_class[k];
            rowcls[r0 * cw + cl / 32] |= 1u << (cl % 32);   // the first row of a variant is its count row
        }
    }
}

// grid (blocks, ceil(K/64)); thread (kk, q): score 64*blockIdx.y + kk, rows q, q+4, ...
__global__ __launch_bounds__(256) void k_plan_exps(const double *__restrict__ wrow, uint64_t R, uint64_t K,
                                                   uint64_t fr, double qmax, uint32_t L, uint64_t ncol,
                                                   int *__restrict__ uexp, double *__restrict__ scales) {
    __shared__ double red[4][64];
    const uint32_t kk = threadIdx.x % 64, q = threadIdx.x / 64;
    const uint64_t k = (uint64_t)blockIdx.y * 64 + kk, blk = blockIdx.x;
    const uint64_t r0 = blk * fr, r1 = r0 + fr < R ? r0 + fr : R;
    double m = 0.0;
    if (k < K)
        for (uint64_t r = r0 + q; r < r1; r += 4) m = nan_max(m, fabs(wrow[r * K + k]));
    red[q][kk] = m;
    __syncthreads();
    if (q != 0 || k >= K) return;
    m = nan_max(nan_max(red[0][kk], red[1][kk]), nan_max(red[2][kk], red[3][kk]));
    const int e = block_exponent_of(m, qmax);
    uexp[blk * K + k] = e;
    for (uint32_t j = 0; j < L; ++j) scales[blk * ncol + j * K + k] = limb_scale(e, 8 * (int)(L - 1 - j));
}

// one thread per (k-step, tile, lane): 16 bytes, rows 64ks + 16g + b of
// column 16*tile + n (n = lane % 16, g = lane / 16)
__global__ __launch_bounds__(256) void k_plan_frag_i8(const double *__restrict__ wrow, uint64_t R, uint64_t K,
                                                      const int *__restrict__ uexp, uint64_t fr, uint32_t L,
                                                      uint32_t ncol, uint32_t wtiles, uint32_t tiles, uint64_t nks,
                                                      const uint8_t *__restrict__ count_row,
                                                      const uint32_t *__restrict__ rowcls, uint32_t cw,
                                                      uint32_t nclass, uint4 *__restrict__ frag) {
    const uint64_t idx = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (idx >= nks * tiles * 64) return;
    const uint32_t lane = (uint32_t)(idx % 64), tile = (uint32_t)((idx / 64) % tiles);
    const uint64_t ks = idx / (64ull * tiles);
    const uint32_t n = lane % 16, g = lane / 16;
    const uint32_t col = tile * 16 + n;
    uint32_t word[4] = {0u, 0u, 0u, 0u};
    for (uint32_t b = 0; b < 16; ++b) {
        const uint64_t r = ks * kI8KStep + 16 * g + b;
        if (r >= R) break;
        uint32_t val = 0;
        if (tile < wtiles) {
            if (col < ncol) {
                const uint32_t limb = col / (uint32_t)K;
                const uint64_t k = col % K;
                const double w = wrow[r * K + k];
                const uint64_t blk = ks * kI8KStep / fr;
                const int e = uexp[blk * K + k];
                if (w != 0.0 && e != kBadExp) {
                    int32_t dig[4];
                    limb_digits((int64_t)rint(ldexp(w, -e)), L, dig);
                    val = (uint8_t)(int8_t)dig[limb];
                }
            }
        } else {
            const uint32_t cl = (tile - wtiles) * 16 + n;
            if (cl < nclass && count_row[r] && ((rowcls[r * cw + cl / 32] >> (cl % 32)) & 1u)) val = 1;
        }
        word[b / 4] |= val << (8 * (b % 4));
    }
    frag[idx] = make_uint4(word[0], word[1], word[2], word[3]);
}

// one thread per (64-row stage, tile, k-step h, lane): 8 bf16, rows
// 64 st + 16 g + 8 h + j (j = 0..7) of column 16*tile + n -- the row order in
// which the BF16 scoring kernel decodes a lane's 16 rows into two k-steps
__global__ __launch_bounds__(256) void k_plan_frag_bf16(const double *__restrict__ wrow, uint64_t R, uint64_t K,
                                                        const int *__restrict__ uexp, uint32_t L, uint32_t ncol,
                                                        uint32_t wtiles, uint32_t tiles, uint64_t nst,
                                                        const uint8_t *__restrict__ count_row,
                                                        const uint32_t *__restrict__ rowcls, uint32_t cw,
                                                        uint32_t nclass, uint4 *__restrict__ frag) {
    const uint64_t idx = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (idx >= nst * tiles * 128) return;
    const uint32_t lane = (uint32_t)(idx % 64), h = (uint32_t)((idx / 64) % 2), tile = (uint32_t)((idx / 128) % tiles);
    const uint64_t st = idx / (128ull * tiles);
    const uint32_t n = lane % 16, g = lane / 16;
    const uint32_t col = tile * 16 + n;
    uint32_t word[4] = {0u, 0u, 0u, 0u};
    for (uint32_t j = 0; j < 8; ++j) {
        const uint64_t r = st * 64 + 16 * g + 8 * h + j;
        if (r >= R) break;
        uint32_t val = 0;
        if (tile < wtiles) {
            if (col < ncol) {
                const uint32_t limb = col / (uint32_t)K;
                const uint64_t k = col % K;
                const double w = wrow[r * K + k];
                const uint64_t blk = r / kBfFlushRows;
                const int e = uexp[blk * K + k];
                if (w != 0.0 && e != kBadExp) {
                    int32_t dig[4];
                    limb_digits((int64_t)rint(ldexp(w, -e)), L, dig);
                    val = bf16_of_int(dig[limb]);
                }
            }
        } else {
            const uint32_t cl = (tile - wtiles) * 16 + n;
            if (cl < nclass && count_row[r] && ((rowcls[r * cw + cl / 32] >> (cl % 32)) & 1u)) val = 0x3F80u;
        }
        word[j / 2] |= val << (16 * (j % 2));
    }
    frag[idx] = make_uint4(word[0], word[1], word[2], word[3]);
}

__global__ __launch_bounds__(256) void k_plan_w32(const double *__restrict__ wrow, uint64_t R, uint64_t K, uint64_t k32,
                                                  float *__restrict__ w32) {
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= R * K) return;
    const uint64_t r = i / K, k = i % K;
    w32[r * k32 + k] = (float)wrow[i];
}

static unsigned blocks_for(uint64_t n) { return (unsigned)((n + 255) / 256); }

hipError_t launch_plan_rows(const uint64_t *off, const uint32_t *cols, const float *w, const float *co,
                            const uint64_t *prog_off, const uint8_t *var_rows, uint64_t V, uint64_t K,
                            const uint32_t *score_class, uint32_t cw, double *wrow, uint32_t *rowcls,
                            hipStream_t stream) {
    if (V == 0) return hipSuccess;
    hipLaunchKernelGGL(k_plan_rows, dim3(blocks_for(V)), dim3(256), 0, stream, off, cols, w, co, prog_off, var_rows, V,
                       K, score_class, cw, wrow, rowcls);
    return hipGetLastError();
}

hipError_t launch_plan_exps(const double *wrow, uint64_t R, uint64_t K, uint64_t fr, double qmax, uint32_t L,
                            uint64_t ncol, uint64_t nblk, int *uexp, double *scales, hipStream_t stream) {
    if (nblk == 0) return hipSuccess;
    hipLaunchKernelGGL(k_plan_exps, dim3((unsigned)nblk, (unsigned)((K + 63) / 64)), dim3(256), 0, stream, wrow, R, K,
                       fr, qmax, L, ncol, uexp, scales);
    return hipGetLastError();
}

hipError_t launch_plan_frag(bool int8, const double *wrow, uint64_t R, uint64_t K, const int *uexp, uint64_t fr,
                            uint32_t L, uint32_t ncol, uint32_t wtiles, uint32_t tiles, uint64_t nks,
                            const uint8_t *count_row, const uint32_t *rowcls, uint32_t cw, uint32_t nclass,
                            void *frag, hipStream_t stream) {
    // 16 bytes per (int8 k-step of 64 rows, tile, lane), or per (64-row
    // stage, tile, BF16 k-step, lane) with nks = stages
    const uint64_t n = nks * tiles * (int8 ? 64 : 128);
    if (n == 0) return hipSuccess;
    if (int8)
        hipLaunchKernelGGL(k_plan_frag_i8, dim3(blocks_for(n)), dim3(256), 0, stream, wrow, R, K, uexp, fr, L, ncol,
                           wtiles, tiles, nks, count_row, rowcls, cw, nclass, (uint4 *)frag);
    else
        hipLaunchKernelGGL(k_plan_frag_bf16, dim3(blocks_for(n)), dim3(256), 0, stream, wrow, R, K, uexp, L, ncol,
                           wtiles, tiles, nks, count_row, rowcls, cw, nclass, (uint4 *)frag);
    return hipGetLastError();
}

hipError_t launch_plan_w32(const double *wrow, uint64_t R, uint64_t K, uint64_t k32, float *w32, hipStream_t stream) {
    if (R * K == 0) return hipSuccess;
    hipLaunchKernelGGL(k_plan_w32, dim3(blocks_for(R * K)), dim3(256), 0, stream, wrow, R, K, k32, w32);
    return hipGetLastError();
}

}  // namespace gnm
