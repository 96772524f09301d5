// Device-side plan builders and the adjoint kernel of the differentiable
// path (dense weights on the device, e.g. a PyTorch-ROCm parameter).
//
// Forward:  S[p,k] = sum_v g_pv W[v,k]  -- the scoring kernels with an
//           identity variant program; the split-BF16 B fragments are built
//           here from device weights.
// Adjoint:  dW[v,k] = sum_p g_pv R[p,k] (R = dL/dS), recomputing g from the
//           packed .bed instead of storing dosages.  MFMA with variants on M,
//           scores (x limbs) on N and people on the reduction axis: a lane
//           reads two bytes of one row -- eight people of one variant -- so the
//           SNP-major layout needs no transpose here.
#include "gnm_dense.hpp"

namespace gnm {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));
typedef uint32_t u32x4_t __attribute__((ext_vector_type(4)));

union DFrag {
    bf16x8_t v;
    uint32_t u[4];
};

__device__ __forceinline__ uint32_t d_flip_word(uint32_t b) {
    return ~(((b & 0x55555555u) << 1) | ((b >> 1) & 0x55555555u));
}

// integer-valued BF16 limb j (of L) of w at power-of-two unit 2^e
__device__ __forceinline__ uint16_t limb_bf16(float w, int e, uint32_t L, uint32_t j) {
    if (w == 0.0f || e == kBadExp) return 0;
    int64_t Q = (int64_t)rint(ldexp((double)w, -e));
    int32_t dig = 0;
    for (uint32_t i = L - 1;; --i) {
        const int64_t d = ((Q + 128) & 255) - 128;
        if (i == j) {
            dig = (int32_t)(i == 0 ? Q : d);
            break;
        }
        Q = (Q - d) / 256;
        if (i == 0) break;
    }
    const float f = (float)dig;
    return (uint16_t)(__float_as_uint(f) >> 16);
}

// --------------------------------------------------------------------------
// exponent of the limb unit per (block of `block` rows, column):
// smallest e with max|src| <= qmax * 2^e
// --------------------------------------------------------------------------
__global__ void k_block_exponent(const float *__restrict__ src, uint64_t rows, uint64_t ld,
                                 uint64_t cols, uint64_t block, uint32_t L, int int8, int *__restrict__ out) {
    const uint64_t k = (uint64_t)blockIdx.x;
    const uint64_t blk = blockIdx.y;
    const uint64_t r0 = blk * block;
    const uint64_t r1 = r0 + block < rows ? r0 + block : rows;
    // a NaN or infinity anywhere in the block marks it (kBadExp)
    double m = 0.0;
    for (uint64_t r = r0 + threadIdx.x; r < r1; r += blockDim.x) m = nan_max(m, fabs((double)src[r * ld + k]));
    __shared__ double red[256];
    red[threadIdx.x] = m;
    __syncthreads();
    for (int sft = 128; sft > 0; sft >>= 1) {
        if ((int)threadIdx.x < sft) red[threadIdx.x] = nan_max(red[threadIdx.x], red[threadIdx.x + sft]);
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double qmax = int8 ? 127.0 * ldexp(1.0, 8 * (L - 1)) : ldexp(1.0, 8 * L) - ldexp(1.0, 8 * (L - 1));
        out[blk * cols + k] = block_exponent_of(red[0], qmax);
    }
}

hipError_t launch_block_exponent(const float *src, uint64_t rows, uint64_t ld, uint64_t cols,
                                 uint64_t block, uint32_t L, bool int8, int *out, hipStream_t stream) {
    const uint64_t nblk = (rows + block - 1) / block;
    if (nblk == 0 || cols == 0) return hipSuccess;
    hipLaunchKernelGGL(k_block_exponent, dim3((unsigned)cols, (unsigned)nblk), dim3(256), 0, stream, src,
                       rows, ld, cols, block, L, int8 ? 1 : 0, out);
    return hipGetLastError();
}

// --------------------------------------------------------------------------
// forward B fragments from dense W [V][K] in the split-BF16 scoring
// kernel's layout [64-row stage][tile][k-step h][lane][8 bf16] (lane (n, g)
// of k-step h: rows 64 st + 16 g + 8 h + j, column 16 tile + n); column
// c = limb*K + k; tile columns beyond ncol and rows beyond V are zero; the
// count tile (class 0) is 1.0 on every real row.  Scale blocks of
// kBfFlushRows rows.  One thread per (stage, tile, k-step, lane).
// --------------------------------------------------------------------------
__global__ void k_forward_frag(const float *__restrict__ W, uint64_t V, uint64_t K, uint32_t L,
                               const int *__restrict__ exps, uint32_t wtiles, uint32_t tiles,
                               uint64_t nst, uint16_t *__restrict__ frag, double *__restrict__ scales) {
    const uint64_t id = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= nst * tiles * 128) return;
    const uint32_t lane = id & 63;
    const uint32_t h = (uint32_t)((id >> 6) & 1);
    const uint32_t tile = (uint32_t)((id >> 7) % tiles);
    const uint64_t stg = (id >> 7) / tiles;
    const uint32_t n = lane & 15, g = lane >> 4;
    const uint64_t ncol = (uint64_t)L * K;
    uint16_t out[8];
    const uint64_t blk = stg * 64 / kBfFlushRows;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const uint64_t r = stg * 64 + g * 16 + h * 8 + j;
        uint16_t val = 0;
        if (r < V) {
            if (tile < wtiles) {
                const uint64_t c = (uint64_t)tile * 16 + n;
                if (c < ncol) {
                    const uint32_t limb = (uint32_t)(c / K);
                    const uint64_t k = c % K;
                    val = limb_bf16(W[r * K + k], exps[blk * K + k], L, limb);
                }
            } else if (tile == wtiles && n == 0) {
                val = 0x3F80;
            }
        }
        out[j] = val;
    }
    uint16_t *dst = frag + id * 8;
#pragma unroll
    for (int j = 0; j < 8; ++j) dst[j] = out[j];
    // scales: one thread per (block, column) -- the first stage of each block
    if (tile < wtiles && g == 0 && h == 0 && (stg * 64) % kBfFlushRows == 0) {
        const uint64_t c = (uint64_t)tile * 16 + n;
        if (c < ncol) {
            const uint32_t limb = (uint32_t)(c / K);
            const uint64_t k = c % K;
            scales[blk * ncol + c] = limb_scale(exps[blk * K + k], 8 * (int)(L - 1 - limb));
        }
    }
}

hipError_t launch_forward_frag(const float *W, uint64_t V, uint64_t K, uint32_t L, const int *exps,
                               uint32_t wtiles, uint32_t tiles, uint64_t nst, uint16_t *frag,
                               double *scales, hipStream_t stream) {
    const uint64_t n = nst * tiles * 128;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_forward_frag, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, W, V, K, L,
                       exps, wtiles, tiles, nst, frag, scales);
    return hipGetLastError();
}

// flips (one byte per variant) -> bit mask words; identity program; class bits
__global__ void k_dense_program(const uint8_t *__restrict__ flip, uint64_t V, uint64_t Rpad,
                                uint32_t *__restrict__ prog_var, uint32_t *__restrict__ flipmask,
                                uint32_t *__restrict__ prog_cls) {
    const uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r < Rpad) {
        prog_var[r] = (uint32_t)(r < V ? r : (V ? V - 1 : 0));
        prog_cls[r] = r < V ? 1u : 0u;
    }
    if ((r & 31) == 0 && r < Rpad) {
        uint32_t m = 0;
        if (flip)
            for (uint32_t j = 0; j < 32; ++j)
                if (r + j < V && flip[r + j]) m |= 1u << j;
        flipmask[r >> 5] = m;
    }
}

hipError_t launch_dense_program(const uint8_t *flip, uint64_t V, uint64_t Rpad, uint32_t *prog_var,
                                uint32_t *flipmask, uint32_t *prog_cls, hipStream_t stream) {
    if (Rpad == 0) return hipSuccess;
    hipLaunchKernelGGL(k_dense_program, dim3((unsigned)((Rpad + 255) / 256)), dim3(256), 0, stream, flip, V,
                       Rpad, prog_var, flipmask, prog_cls);
    return hipGetLastError();
}

// W [V][K] -> W32 [Rpad][k32] zero padded (FP32 kernel layout)
__global__ void k_pad_weights(const float *__restrict__ W, uint64_t V, uint64_t K, uint64_t Rpad,
                              uint64_t k32, float *__restrict__ out) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= Rpad * k32) return;
    const uint64_t r = i / k32, k = i % k32;
    out[i] = (r < V && k < K) ? W[r * K + k] : 0.0f;
}

hipError_t launch_pad_weights(const float *W, uint64_t V, uint64_t K, uint64_t Rpad, uint64_t k32, float *out,
                              hipStream_t stream) {
    const uint64_t n = Rpad * k32;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_pad_weights, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, W, V, K, Rpad,
                       k32, out);
    return hipGetLastError();
}

// counts [N][K] -> column 0 (the per-person missing count of a dense plan)
__global__ void k_take_col0(const uint32_t *__restrict__ counts, uint64_t N, uint64_t K, uint32_t *__restrict__ out) {
    const uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p < N) out[p] = counts[p * K];
}

hipError_t launch_take_col0(const uint32_t *counts, uint64_t N, uint64_t K, uint32_t *out, hipStream_t stream) {
    if (N == 0) return hipSuccess;
    hipLaunchKernelGGL(k_take_col0, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, stream, counts, N, K, out);
    return hipGetLastError();
}

// --------------------------------------------------------------------------
// Adjoint.
// People are consumed in chunks of 256 per iteration; inside a chunk, lane
// group g owns people [64g, 64g+64) (16 bytes of the row) and k-step s
// (0..7) takes bytes 2s, 2s+1 of that range.  Slot order inside a lane:
// VGPR q = (person 8s+q, person 8s+4+q) -- field q of byte 2s and of byte
// 2s+1 -- so slot j of lane group g is person 256*chunk + 64g + 8s + perm[j]
// with perm = {0,4,1,5,2,6,3,7}.  R fragments are built in the same order.
// --------------------------------------------------------------------------
__device__ __constant__ int kAdjPerm[8] = {0, 4, 1, 5, 2, 6, 3, 7};

__global__ void k_adjoint_frag(const float *__restrict__ R, uint64_t N, uint64_t K, uint64_t Kp, uint32_t L,
                               const int *__restrict__ exps, uint32_t tiles, uint64_t nks,
                               uint16_t *__restrict__ frag, double *__restrict__ scales) {
    const uint64_t id = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= nks * tiles * 64) return;
    const uint32_t lane = id & 63;
    const uint32_t tile = (uint32_t)((id >> 6) % tiles);
    const uint64_t ks = (id >> 6) / tiles;
    const uint32_t n = lane & 15, g = lane >> 4;
    const uint64_t c = (uint64_t)tile * 16 + n;
    const uint32_t limb = (uint32_t)(c / Kp);
    const uint64_t k = c % Kp;
    const uint64_t chunk = ks / 8, s8 = ks % 8;
    const uint64_t blk = chunk * 256 / kFlushRows;
    uint16_t *dst = frag + id * 8;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const uint64_t p = chunk * 256 + 64 * g + 8 * s8 + kAdjPerm[j];
        uint16_t val = 0;
        if (p < N && k < K && limb < L) val = limb_bf16(R[p * K + k], exps[blk * K + k], L, limb);
        dst[j] = val;
    }
    if (g == 0 && s8 == 0 && (chunk * 256) % kFlushRows == 0 && limb < L && k < K)
        scales[blk * (L * Kp) + c] = limb_scale(exps[blk * K + k], 8 * (int)(L - 1 - limb));
}

hipError_t launch_adjoint_frag(const float *R, uint64_t N, uint64_t K, uint64_t Kp, uint32_t L, const int *exps,
                               uint32_t tiles, uint64_t nks, uint16_t *frag, double *scales, hipStream_t stream) {
    const uint64_t n = nks * tiles * 64;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_adjoint_frag, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, R, N, K, Kp, L,
                       exps, tiles, nks, frag, scales);
    return hipGetLastError();
}

template <int TMAX>
__global__ __launch_bounds__(256, 2) void k_adjoint(const uint8_t *__restrict__ rows, uint64_t stride,
                                                    uint64_t V, uint64_t N, const uint8_t *__restrict__ flip,
                                                    const uint32_t *__restrict__ frag, const double *__restrict__ scales,
                                                    uint32_t tiles, uint64_t K, uint64_t Kp, uint32_t L,
                                                    float *__restrict__ dW) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int li = lane & 15, lg = lane >> 4;
    const uint64_t vt = (uint64_t)blockIdx.x * 4 + wave;  // 16-variant tile of this wave
    const uint64_t v = vt * 16 + li;                      // A row of this lane
    const bool vok = v < V;
    const uint64_t bpv = (N + 3) / 4;
    const uint8_t *__restrict__ row = rows + (vok ? v : 0) * stride;
    const bool fl = vok && flip && flip[v];
    // 128-bit row loads need 16-byte aligned rows (base and stride)
    const bool row16 = ((reinterpret_cast<uintptr_t>(rows) | stride) & 15u) == 0;
    const uint64_t nchunks = (N + 255) / 256;
    const uint64_t ncolp = (uint64_t)L * Kp;

    f32x4_t acc[TMAX];
    double acc64[TMAX][4];
#pragma unroll
    for (int j = 0; j < TMAX; ++j) {
        acc[j] = f32x4_t{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int q = 0; q < 4; ++q) acc64[j][q] = 0.0;
    }
    const uint64_t chunks_per_blk = kFlushRows / 256;
    for (uint64_t c0 = 0; c0 < nchunks; c0 += chunks_per_blk) {
        const uint64_t c1 = c0 + chunks_per_blk < nchunks ? c0 + chunks_per_blk : nchunks;
        const uint64_t blk = c0 / chunks_per_blk;
#pragma unroll 1
        for (uint64_t ch = c0; ch < c1; ++ch) {
            // 16 bytes of this lane's row: people 256ch + 64lg .. +64
            const uint64_t off = ch * 64 + 16 * lg;
            uint32_t w[4];
            if (vok && row16 && off + 16 <= bpv) {
                // whole 16 bytes inside the row: one 128-bit load
                const uint4 q = *(const uint4 *)(row + off);
                w[0] = q.x; w[1] = q.y; w[2] = q.z; w[3] = q.w;
                if (fl) {
#pragma unroll
                    for (int d = 0; d < 4; ++d) w[d] = d_flip_word(w[d]);
                }
            } else {
#pragma unroll
                for (int d = 0; d < 4; ++d) {
                    uint32_t x = 0x55555555u;  // missing: contributes nothing
                    if (vok) {
                        x = 0;
#pragma unroll
                        for (int bb = 0; bb < 4; ++bb) {
                            const uint64_t o = off + 4 * d + bb;
                            const uint32_t byte = o < bpv ? row[o] : 0x55u;
                            x |= byte << (8 * bb);
                        }
                        if (fl) x = d_flip_word(x);
                    }
                    w[d] = x;
                }
            }
            // a k-step's fragments load together, without branches (slots
            // past the last tile read the last tile again and their products
            // are never used); the next k-step's are in flight while this
            // one's MFMAs run
            auto load_b = [&](u32x4_t (&bv)[TMAX], uint64_t ks) __attribute__((always_inline)) {
                const u32x4_t *bf = (const u32x4_t *)frag + (ks * tiles) * 64 + lane;
#pragma unroll
                for (int j = 0; j < TMAX; ++j) bv[j] = bf[(uint64_t)((uint32_t)j < tiles ? j : tiles - 1) * 64];
            };
            // (prefetch only where registers allow: the 12-tile instance
            // would spill)
            constexpr bool kPrefetch = TMAX <= 4;
            u32x4_t bvb[kPrefetch ? 2 : 1][TMAX];
            if constexpr (kPrefetch) load_b(bvb[0], ch * 8);
#pragma unroll
            for (int s8 = 0; s8 < 8; ++s8) {
                if constexpr (kPrefetch) {
                    if (s8 < 7) load_b(bvb[(s8 + 1) & 1], ch * 8 + s8 + 1);
                } else {
                    load_b(bvb[0], ch * 8 + s8);
                }
                const uint32_t b0 = (w[s8 >> 1] >> (16 * (s8 & 1))) & 0xFFu;
                const uint32_t b1 = (w[s8 >> 1] >> (16 * (s8 & 1) + 8)) & 0xFFu;
                const uint32_t pp = (b0 | (b1 << 16)) * 0x0101u;
                DFrag A;
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const uint32_t x = ((pp >> (2 * q)) & 0x03030303u) | 0x04000400u;
                    A.u[q] = __builtin_amdgcn_perm(0x403F0000u, 0x00800000u, x);
                }
                const u32x4_t (&bv)[TMAX] = bvb[kPrefetch ? (s8 & 1) : 0];
#pragma unroll
                for (int j = 0; j < TMAX; ++j) {
                    if ((uint32_t)j < tiles) {
                        DFrag B;
                        B.u[0] = bv[j][0]; B.u[1] = bv[j][1]; B.u[2] = bv[j][2]; B.u[3] = bv[j][3];
                        acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A.v, B.v, acc[j], 0, 0, 0);
                    }
                }
            }
        }
#pragma unroll
        for (int j = 0; j < TMAX; ++j) {
            if ((uint32_t)j >= tiles) continue;
            const uint64_t c = (uint64_t)j * 16 + li;
            const double sc = c < ncolp ? scales[blk * ncolp + c] : 0.0;
#pragma unroll
            for (int q = 0; q < 4; ++q) acc64[j][q] += (double)acc[j][q] * sc;
            acc[j] = f32x4_t{0.f, 0.f, 0.f, 0.f};
        }
    }
    // output: lane holds column li of tile j, variant rows 4lg+q; limb l of
    // score k lives in tile (l*Kp + k)/16 at the same lane
    const uint32_t tiles_per_limb = (uint32_t)(Kp / 16);
#pragma unroll
    for (int j = 0; j < TMAX; ++j) {
        if ((uint32_t)j >= tiles_per_limb) continue;
        const uint64_t k = (uint64_t)j * 16 + li;
        if (k >= K) continue;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            double sum = 0.0;
#pragma unroll
            for (int j2 = 0; j2 < TMAX; ++j2)
                if ((uint32_t)j2 < tiles && (uint32_t)j2 % tiles_per_limb == (uint32_t)j) sum += acc64[j2][q];
            const uint64_t vv = vt * 16 + 4 * lg + q;
            if (vv < V) dW[vv * K + k] = (float)sum;
        }
    }
}

hipError_t launch_adjoint(const uint8_t *rows, uint64_t stride, uint64_t V, uint64_t N, const uint8_t *flip,
                          const uint32_t *frag, const double *scales, uint32_t tiles, uint64_t K, uint64_t Kp,
                          uint32_t L, float *dW, hipStream_t stream) {
    const uint64_t wgs = (V + 63) / 64;
    if (wgs == 0) return hipSuccess;
    if (tiles == 0) return hipErrorInvalidValue;   // the kernel reads tile min(j, tiles - 1)
    if (tiles <= 4) {
        hipLaunchKernelGGL(k_adjoint<4>, dim3((unsigned)wgs), dim3(256), 0, stream, rows, stride, V, N, flip, frag,
                           scales, tiles, K, Kp, L, dW);
    } else if (tiles <= 12) {
        hipLaunchKernelGGL(k_adjoint<12>, dim3((unsigned)wgs), dim3(256), 0, stream, rows, stride, V, N, flip, frag,
                           scales, tiles, K, Kp, L, dW);
    } else {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace gnm
