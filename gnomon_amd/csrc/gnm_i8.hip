// Split-int8 path: the launcher of the fused decode + contraction kernel
// (k_score_mx in gnm_score_mx.hpp with int8 limbs, v_mfma_i32_16x16x64_i8,
// exact int32 block sums over scale blocks of kI8FlushRows = 65536 program
// rows; L = 3 limbs represent 23 bits, L = 4 31 bits of the block maximum),
// the device builders of the dense (differentiable) path's int8 fragments,
// and the split-int8 adjoint.
#include "gnm_score_mx.hpp"

namespace gnm {

// Rough cost of one group pass per wave and stage, in cycles: decode VALU
// (~60 instructions) plus 4 MFMA of 16 cycles per tile slot.
static uint64_t i8_pass_cost(uint32_t ntiles, uint32_t tmax) {
    const uint64_t groups = (ntiles + tmax - 1) / tmax;
    return groups * (240u + 64u * tmax);
}

// default tiles per pass: the cheapest of 3, 4, 5, 6 and 8 by the cost model
uint32_t i8_default_tiles_per_pass(uint32_t ntiles) {
    uint32_t tmax = 3;
    for (uint32_t c : {4u, 5u, 6u, 8u})
        if (i8_pass_cost(ntiles, c) < i8_pass_cost(ntiles, tmax)) tmax = c;
    return tmax;
}

uint64_t i8_default_cost(uint32_t ntiles) {
    return i8_pass_cost(ntiles, i8_default_tiles_per_pass(ntiles));
}

// Variant splits of the int8 kernel: enough that all tile groups together
// launch about 12k workgroups (~48 per CU, so that three stay resident per
// CU through the tail and many row strips are read at once), a multiple of
// 8 (whole splits per XCD), at most 32: the FP64 partials and their
// reduction grow with the split count (at 50k people x 16 scores, 64 splits
// wrote and reduced 1.6 GB of partials; 32 splits ran the forward 9% faster,
// 16 no faster than 32).
uint32_t i8_splits(uint64_t num_people, uint32_t ntiles) {
    const uint32_t tmax = i8_default_tiles_per_pass(ntiles ? ntiles : 1);
    const uint64_t groups = ((ntiles ? ntiles : 1) + tmax - 1) / tmax;
    const uint64_t per = ((num_people + 255) / 256) * groups;
    uint64_t s = (12288 + per - 1) / per;
    s = (s + 7) / 8 * 8;
    return (uint32_t)(s < 8 ? 8 : (s > 32 ? 32 : s));
}

hipError_t launch_score_i8(const DevicePlan &plan, const uint8_t *rows, uint64_t stride, uint64_t bpv, uint64_t v0,
                           uint64_t p0, uint64_t p1, uint64_t num_people, uint32_t splits, uint32_t tile_begin,
                           uint32_t tile_end, bool identity, double *part, uint32_t *cpart, hipStream_t stream,
                           uint32_t tiles_per_pass, uint32_t batch_stages) {
    if (tile_end <= tile_begin) return hipSuccess;
    // dword row loads need 4-byte alignment; the fast path's buffer offsets
    // (up to 64 rows x stride) must stay below the descriptors' 2^31 - 1 range
    const bool vec = ((reinterpret_cast<uintptr_t>(rows) | stride) & 3u) == 0 && bpv >= 4 && stride < (1ull << 25);
    const uint64_t split_base = (p0 / kI8Stage) * kI8Stage;
    uint64_t rows_per_split = (p1 - split_base + splits - 1) / splits;
    rows_per_split = (rows_per_split + kI8Stage - 1) / kI8Stage * kI8Stage;
    const uint32_t ntiles = tile_end - tile_begin;
    // tiles per workgroup pass: every group computes tmax tile slots, so the
    // default takes the cheapest of 3, 4, 5, 6 and 8 by the cost model above;
    // 18 is two sets of 9-tile groups over the same people (one row read
    // for both)
    const uint32_t tp = tiles_per_pass ? tiles_per_pass : i8_default_tiles_per_pass(ntiles);
    if (tp != 3 && tp != 4 && tp != 5 && tp != 6 && tp != 8 && tp != 18) return hipErrorInvalidValue;
    // stages per batch by default: two for 3-tile groups (the smaller
    // register ring keeps four waves per SIMD), 4 and 8 and the two-set
    // shape; four for 6-tile groups (more loads in flight; 6% faster at
    // 100k x 1M x 64)
    const uint32_t depth = batch_stages ? batch_stages : (tp == 6 ? 4u : 2u);
    if (depth != 2 && depth != 4) return hipErrorInvalidValue;
    if (tp == 18 && depth != 2) return hipErrorInvalidValue;   // LDS: two-stage batches only
    const uint32_t ngroups = (ntiles + tp - 1) / tp;
    // 3-tile groups run 8 waves (512 people, 128 bytes of each row per
    // workgroup: whole 128-byte lines); larger groups 4 waves; the two-set
    // shape 2 x 4 waves over 256 people
    const uint32_t wv = (tp == 3 || tp == 18) ? 8u : 4u;
    const uint32_t people_per_wg = tp == 18 ? 256u : 64u * wv;
    const uint32_t pt = (uint32_t)((num_people + people_per_wg - 1) / people_per_wg);
    const uint64_t wgs = (uint64_t)pt * splits * ngroups;   /* splits is a multiple of 8 */
#define GNM_LAUNCH_I8(TM, WV, DD, VC, ID, SE)                                                                \
    hipLaunchKernelGGL((k_score_mx<TM, WV, DD, VC, ID, SE, false>), dim3((unsigned)wgs), dim3(64 * WV), 0, stream,     \
                       plan, rows, stride, bpv, v0, p0, p1, split_base, rows_per_split, splits, pt, num_people, \
                       tile_begin, tile_end, ngroups, part, cpart)
#define GNM_LAUNCH_I8_VI(TM, WV, DD, SE)                                                                     \
    do {                                                                                                     \
        if (vec) {                                                                                           \
            if (identity) GNM_LAUNCH_I8(TM, WV, DD, true, true, SE);                                         \
            else GNM_LAUNCH_I8(TM, WV, DD, true, false, SE);                                                 \
        } else {                                                                                             \
            if (identity) GNM_LAUNCH_I8(TM, WV, DD, false, true, SE);                                        \
            else GNM_LAUNCH_I8(TM, WV, DD, false, false, SE);                                                \
        }                                                                                                    \
    } while (0)
#define GNM_LAUNCH_I8_D(TM, WV)                                                                              \
    do {                                                                                                     \
        if (depth == 2) GNM_LAUNCH_I8_VI(TM, WV, 2, 1); else GNM_LAUNCH_I8_VI(TM, WV, 4, 1);                 \
    } while (0)
    if (tp == 3) GNM_LAUNCH_I8_D(3, 8);
    else if (tp == 4) GNM_LAUNCH_I8_D(4, 4);
    else if (tp == 5) GNM_LAUNCH_I8_D(5, 4);
    else if (tp == 6) GNM_LAUNCH_I8_D(6, 4);
    else if (tp == 8) GNM_LAUNCH_I8_D(8, 4);
    else GNM_LAUNCH_I8_VI(9, 8, 2, 2);
#undef GNM_LAUNCH_I8_D
#undef GNM_LAUNCH_I8_VI
#undef GNM_LAUNCH_I8
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Device-side builders of the dense (differentiable) path.
// ---------------------------------------------------------------------------

// signed base-256 digit j (of L) of the integer nearest w / 2^e, all digits in [-128, 127]
__device__ __forceinline__ int32_t limb_i8(float w, int e, uint32_t L, uint32_t j) {
    if (w == 0.0f || e == kBadExp) return 0;
    int64_t Q = (int64_t)rint(ldexp((double)w, -e));
    for (uint32_t i = L - 1; i > j; --i) {
        const int64_t d = ((Q + 128) & 255) - 128;
        Q = (Q - d) / 256;
    }
    return j == 0 ? (int32_t)Q : (int32_t)(((Q + 128) & 255) - 128);
}

// B fragments [k-step of 64 rows][tile][lane][16 i8] from dense W [V][K]:
// lane (n = lane%16, g = lane/16) byte b holds row 64ks + 16g + b of column
// tile*16 + n; column c = limb*K + k; the count tile (class 0) is 1 on every
// real row.  One thread per (k-step, tile, lane).
__global__ void k_forward_frag_i8(const float *__restrict__ W, uint64_t V, uint64_t K, uint32_t L,
                                  const int *__restrict__ exps, uint32_t wtiles, uint32_t tiles, uint64_t nks,
                                  uint32_t *__restrict__ frag, double *__restrict__ scales) {
    const uint64_t id = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= nks * tiles * 64) return;
    const uint32_t lane = id & 63;
    const uint32_t tile = (uint32_t)((id >> 6) % tiles);
    const uint64_t ks = (id >> 6) / tiles;
    const uint32_t n = lane & 15, g = lane >> 4;
    const uint64_t ncol = (uint64_t)L * K;
    const uint64_t blk = ks * kI8Stage / kI8FlushRows;
    uint32_t out[4] = {0, 0, 0, 0};
#pragma unroll
    for (int b = 0; b < 16; ++b) {
        const uint64_t r = ks * kI8Stage + g * 16 + b;
        int32_t val = 0;
        if (r < V) {
            if (tile < wtiles) {
                const uint64_t c = (uint64_t)tile * 16 + n;
                if (c < ncol) {
                    const uint32_t limb = (uint32_t)(c / K);
                    const uint64_t k = c % K;
                    val = limb_i8(W[r * K + k], exps[blk * K + k], L, limb);
                }
            } else if (tile == wtiles && n == 0) {
                val = 1;
            }
        }
        out[b >> 2] |= ((uint32_t)val & 0xFFu) << (8 * (b & 3));
    }
    uint32_t *dst = frag + id * 4;
#pragma unroll
    for (int q = 0; q < 4; ++q) dst[q] = out[q];
    if (tile < wtiles && g == 0 && (ks * kI8Stage) % kI8FlushRows == 0) {
        const uint64_t c = (uint64_t)tile * 16 + n;
        if (c < ncol) {
            const uint32_t limb = (uint32_t)(c / K);
            const uint64_t k = c % K;
            scales[blk * ncol + c] = limb_scale(exps[blk * K + k], 8 * (int)(L - 1 - limb));
        }
    }
}

hipError_t launch_forward_frag_i8(const float *W, uint64_t V, uint64_t K, uint32_t L, const int *exps,
                                  uint32_t wtiles, uint32_t tiles, uint64_t nks, uint32_t *frag, double *scales,
                                  hipStream_t stream) {
    const uint64_t n = nks * tiles * 64;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_forward_frag_i8, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, W, V, K, L, exps,
                       wtiles, tiles, nks, frag, scales);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Split-int8 adjoint dW = G^T R.  The reduction runs over people, which are
// contiguous inside a .bed row, so the MFMA A operand (16 variants x 64
// people) comes straight from each lane's own row: no transpose.  People are
// taken in super-steps of 256 (64 packed bytes per row = 4 k-steps); lane
// (i, g) loads bytes 64S + 16g .. +15 of its row with one 16-byte load, and
// dword d of that load is its part of k-step 4S + d.  In that k-step, A
// dword t byte b (K position 4t + b of the lane's 16) is person
// 256S + 64g + 16d + 4b + t; the B fragments of R follow the same order.
// R is split per (block of kI8FlushRows people, column) into L signed int8
// limbs like the forward weights; the int32 block sums are flushed, scaled,
// into an FP64 workspace [V][Kp].
// ---------------------------------------------------------------------------

// B fragments [ks][tile][lane][16 i8] of R [N][K]: column c = limb*Kp + k;
// tiles past L*Kp/16 and people past N are zero.  One thread per
// (k-step, tile, lane).
__global__ void k_adjoint_frag_i8(const float *__restrict__ R, uint64_t N, uint64_t K, uint64_t Kp, uint32_t L,
                                  const int *__restrict__ exps, uint32_t tiles, uint32_t tiles_used, uint64_t nks,
                                  uint32_t *__restrict__ frag, double *__restrict__ scales) {
    const uint64_t id = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= nks * tiles * 64) return;
    const uint32_t lane = id & 63;
    const uint32_t tile = (uint32_t)((id >> 6) % tiles);
    const uint64_t ks = (id >> 6) / tiles;
    const uint32_t n = lane & 15, g = lane >> 4;
    const uint64_t c = (uint64_t)tile * 16 + n;
    const uint32_t limb = (uint32_t)(c / Kp);
    const uint64_t k = c % Kp;
    const bool col_ok = tile < tiles_used && k < K;
    const uint64_t pbase = (ks >> 2) * 256 + 64 * g + 16 * (ks & 3);
    uint32_t out[4] = {0, 0, 0, 0};
#pragma unroll
    for (int p = 0; p < 16; ++p) {
        const int t = p >> 2, b = p & 3;
        const uint64_t person = pbase + 4 * b + t;
        int32_t val = 0;
        if (col_ok && person < N) {
            const uint64_t blk = person / kI8FlushRows;
            val = limb_i8(R[person * K + k], exps[blk * K + k], L, limb);
        }
        out[p >> 2] |= ((uint32_t)val & 0xFFu) << (8 * (p & 3));
    }
    uint32_t *dst = frag + id * 4;
#pragma unroll
    for (int q = 0; q < 4; ++q) dst[q] = out[q];
    if (col_ok && g == 0 && (ks * kI8Stage) % kI8FlushRows == 0) {
        const uint64_t blk = ks * kI8Stage / kI8FlushRows;
        scales[blk * (L * Kp) + c] = limb_scale(exps[blk * K + k], 8 * (int)(L - 1 - limb));
    }
}

hipError_t launch_adjoint_frag_i8(const float *R, uint64_t N, uint64_t K, uint64_t Kp, uint32_t L, const int *exps,
                                  uint32_t tiles, uint32_t tiles_used, uint64_t nks, uint32_t *frag, double *scales,
                                  hipStream_t stream) {
    const uint64_t n = nks * tiles * 64;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_adjoint_frag_i8, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, R, N, K, Kp, L,
                       exps, tiles, tiles_used, nks, frag, scales);
    return hipGetLastError();
}

// BF16 B fragments [ks][tile][lane][8 bf16] of R [N][K] for the split-BF16
// adjoint: column c = limb*Kp + k; k-step ks = 4 S + d covers, for lane
// group g, the 8 people 128 S + 32 g + 4 (2 d + h) + t (slot 2 t + h, h = 0, 1,
// t = 0..3), the order in which the adjoint kernel decodes a lane's row
// bytes 2d and 2d + 1 of its 8-byte share of super-step S.  One thread per
// (k-step, tile, lane).
__global__ void k_adjoint_frag_bf(const float *__restrict__ R, uint64_t N, uint64_t K, uint64_t Kp, uint32_t L,
                                  const int *__restrict__ exps, uint32_t tiles, uint32_t tiles_used, uint64_t nks,
                                  uint32_t *__restrict__ frag, double *__restrict__ scales) {
    const uint64_t id = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= nks * tiles * 64) return;
    const uint32_t lane = id & 63;
    const uint32_t tile = (uint32_t)((id >> 6) % tiles);
    const uint64_t ks = (id >> 6) / tiles;
    const uint32_t n = lane & 15, g = lane >> 4;
    const uint64_t c = (uint64_t)tile * 16 + n;
    const uint32_t limb = (uint32_t)(c / Kp);
    const uint64_t k = c % Kp;
    const bool col_ok = tile < tiles_used && k < K;
    const uint64_t S = ks >> 2, d = ks & 3;
    uint32_t out[4] = {0, 0, 0, 0};
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const uint64_t person = 128 * S + 32 * g + 4 * (2 * d + (j & 1)) + (j >> 1);
        uint32_t val = 0;
        if (col_ok && person < N) {
            const uint64_t blk = person / kBfFlushRows;
            const float w = R[person * K + k];
            const int e = exps[blk * K + k];
            if (w != 0.0f && e != kBadExp) {
                int32_t dig[4];
                limb_digits((int64_t)rint(ldexp((double)w, -e)), L, dig);
                val = bf16_of_int(dig[limb]);
            }
        }
        out[j >> 1] |= val << (16 * (j & 1));
    }
    uint32_t *dst = frag + id * 4;
#pragma unroll
    for (int q = 0; q < 4; ++q) dst[q] = out[q];
    if (col_ok && g == 0 && (ks * 32) % kBfFlushRows == 0) {
        const uint64_t blk = ks * 32 / kBfFlushRows;
        scales[blk * (L * Kp) + c] = limb_scale(exps[blk * K + k], 8 * (int)(L - 1 - limb));
    }
}

hipError_t launch_adjoint_frag_bf(const float *R, uint64_t N, uint64_t K, uint64_t Kp, uint32_t L, const int *exps,
                                  uint32_t tiles, uint32_t tiles_used, uint64_t nks, uint32_t *frag, double *scales,
                                  hipStream_t stream) {
    const uint64_t n = nks * tiles * 64;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_adjoint_frag_bf, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, R, N, K, Kp, L,
                       exps, tiles, tiles_used, nks, frag, scales);
    return hipGetLastError();
}

// The adjoint dW = G^T R, split-int8 (BF = false) or split-BF16 (BF = true).
// One workgroup = 4 waves x MT 16-variant tiles.  A super-step covers 256
// people (int8: each lane loads 16 bytes of its row, four k-steps of
// v_mfma_i32_16x16x64_i8) or 128 people (BF16: 8 bytes, four k-steps of
// v_mfma_f32_16x16x32_bf16, each decoding two of the lane's bytes into 8
// BF16 dosages).  Per super-step the workgroup's 256 threads load the B
// fragments of its 4 k-steps (16 B x TILES each) and store them to LDS
// (double-buffered, one barrier per super-step); each lane loads its MT
// rows' bytes.  Both are loaded one super-step ahead into registers.
// The 4-tile BF16 instance is held to 168 registers (three waves per SIMD):
// at 50k x 500k x 16 in bf16x4 3.16 against 3.43 ms with two; the int8
// instance ran slower that way (1.93 against 1.78 ms) and keeps two.
template <int TILES, int MT, bool BF>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(TILES == 4 && BF ? 3 : 1))) void k_adjoint_mx(const uint8_t *__restrict__ rows, uint64_t stride, uint64_t V,
                                                    uint64_t N, const uint8_t *__restrict__ flip,
                                                    const u32x4_t *__restrict__ frag,
                                                    const double *__restrict__ scales, uint64_t Kp, uint32_t L,
                                                    double *__restrict__ ws) {
    constexpr int kBPer = TILES;                  // B entries (16 B) per thread per super-step
    constexpr int kSSBytes = BF ? 32 : 64;        // row bytes per super-step
    constexpr int kLaneBytes = kSSBytes / 4;      // of which each lane group loads
    typedef std::conditional_t<BF, f32x4_t, i32x4_t> AccV;
    typedef std::conditional_t<BF, uint64_t, u32x4_t> RowV;
    __shared__ u32x4_t bt[2][4 * TILES * 64];     // [buf][(k-step d, tile j, lane)]
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int li = lane & 15, lg = lane >> 4;
    const uint64_t wv0 = ((uint64_t)blockIdx.x * 4 + wave) * (16 * MT);   // first variant of the wave
    const uint64_t bpv = (N + 3) / 4;
    const uint64_t nss = (bpv + kSSBytes - 1) / kSSBytes;                // super-steps
    const uint32_t tpl = (uint32_t)(Kp / 16);                            // tiles per limb
    // rows of the wave: one descriptor, lane offsets (16mt + i) * stride +
    // kLaneBytes g; rows past V read as zero (range check), bytes past the
    // row end belong to people whose R limbs are zero
    const uint64_t rem = wv0 < V ? (V - wv0) * stride : 0;
    const __amdgpu_buffer_rsrc_t rr = __builtin_amdgcn_make_buffer_rsrc(
        (void *)(rows + (wv0 < V ? wv0 : 0) * stride), 0, (int)(rem < 0x7FFFFFFFull ? rem : 0x7FFFFFFFull),
        kI8RsrcWord3);
    uint32_t voff[MT], F[MT];
#pragma unroll
    for (int mt = 0; mt < MT; ++mt) {
        const uint64_t v = wv0 + 16 * mt + li;
        voff[mt] = (uint32_t)((16 * mt + li) * stride + kLaneBytes * lg);
        const bool fl = v < V && flip && flip[v];
        // int8: flip selector bit; BF16: all-ones mask of the code swap
        F[mt] = fl ? (BF ? 0xFFFFFFFFu : 0x04040404u) : 0u;
    }
    const __amdgpu_buffer_rsrc_t rb = __builtin_amdgcn_make_buffer_rsrc(
        (void *)frag, 0, 0x7FFFFFFF, kI8RsrcWord3);
    auto load_ss = [&](uint64_t ss, RowV (&x)[MT], u32x4_t (&b)[kBPer]) __attribute__((always_inline)) {
        const uint32_t so = (uint32_t)(kSSBytes * ss);
#pragma unroll
        for (int mt = 0; mt < MT; ++mt) {
            if constexpr (BF)
                x[mt] = __builtin_bit_cast(uint64_t, __builtin_amdgcn_raw_buffer_load_b64(rr, voff[mt], so, 0));
            else
                x[mt] = __builtin_bit_cast(u32x4_t, __builtin_amdgcn_raw_buffer_load_b128(rr, voff[mt], so, 0));
        }
        const uint64_t e0 = ss * (4 * TILES * 64);   // first B entry of the super-step
#pragma unroll
        for (int m = 0; m < kBPer; ++m)
            b[m] = __builtin_bit_cast(u32x4_t, __builtin_amdgcn_raw_buffer_load_b128(
                                                   rb, (uint32_t)((e0 + tid + 256 * m) * 16), 0, 0));
    };
    AccV acc[MT][TILES];
    auto zero = [&]() __attribute__((always_inline)) {
        int z = 0;
        asm volatile("" : "+v"(z));
        const auto e = [&] {
            if constexpr (BF) return __builtin_bit_cast(float, z); else return z;
        }();
#pragma unroll
        for (int mt = 0; mt < MT; ++mt)
#pragma unroll
            for (int j = 0; j < TILES; ++j) acc[mt][j] = AccV{e, e, e, e};
    };
    zero();
    // block flush: acc[mt][j][q] is row 4g + q of tile mt (variant
    // wv0 + 16mt + 4g + q), column j*16 + i; limbs of score k = jj*16 + i sit
    // in tiles l*tpl + jj of the same lane
    auto flush = [&](uint64_t blk, bool first) __attribute__((always_inline)) {
        // opaque copies: what the flush derives from them is formed here,
        // not hoisted out of the loop into live registers
        uint32_t kp = (uint32_t)Kp, tp = tpl, ll = L, lane_i = (uint32_t)li;
        uint64_t v0 = wv0 + 4 * (uint64_t)lg;
        asm volatile("" : "+v"(kp), "+v"(tp), "+v"(ll), "+v"(lane_i), "+v"(v0));
        kp = __builtin_amdgcn_readfirstlane(kp);
        tp = __builtin_amdgcn_readfirstlane(tp);
        ll = __builtin_amdgcn_readfirstlane(ll);
#pragma unroll
        for (int mt = 0; mt < MT; ++mt) {
            // one tile row of sums at a time: kept in accumulation registers
            // up to here, so only TILES x 4 of them are converted at once
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int j = 0; j < TILES; ++j) asm volatile("" : "+a"(acc[mt][j]));
#pragma unroll 1
            for (uint32_t jj = 0; jj < tp; ++jj) {
                double sc[TILES];
#pragma unroll
                for (int j = 0; j < TILES; ++j) {
                    const uint32_t l = (uint32_t)j / tp;
                    sc[j] = ((uint32_t)j % tp == jj && l < ll)
                                ? scales[blk * (ll * (uint64_t)kp) + (uint64_t)j * 16 + lane_i] : 0.0;
                }
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const uint64_t v = v0 + 16 * mt + q;
                    double sum = 0.0;
#pragma unroll
                    for (int j = 0; j < TILES; ++j) {
                        const double a = acc[mt][j][q];
                        sum += a * sc[j];
                    }
                    if (v < V) {
                        double *d = ws + v * kp + (uint64_t)jj * 16 + lane_i;
                        *d = first ? sum : *d + sum;
                    }
                }
            }
            __builtin_amdgcn_sched_barrier(0);
        }
        zero();
    };

    RowV xa[MT];
    u32x4_t ba[kBPer];
    load_ss(0, xa, ba);
    // outer loop over person blocks, inner loop over the block's
    // super-steps: the flush sits after the inner loop on every path, so the
    // accumulators stay in accumulation registers
    constexpr uint64_t kSSPerBlk = BF ? (uint64_t)kBfFlushRows / 128 : (uint64_t)kI8FlushRows / 256;
    uint64_t ss = 0;
    bool first = true;
    do {   // nss > 0
        const uint64_t blk = ss / kSSPerBlk;
        const uint64_t ss_end = (blk + 1) * kSSPerBlk < nss ? (blk + 1) * kSSPerBlk : nss;
        do {
            const uint32_t buf = (uint32_t)(ss & 1);
            // this super-step's B fragments into LDS (the other half may still
            // be read by a wave finishing the previous super-step)
#pragma unroll
            for (int m = 0; m < kBPer; ++m) bt[buf][tid + 256 * m] = ba[m];
            RowV xc[MT];
#pragma unroll
            for (int mt = 0; mt < MT; ++mt) xc[mt] = xa[mt];
            lds_barrier_i8();
            // next super-step in flight during this one's MFMAs (clamped at
            // the end: the extra loads are never used)
            load_ss(ss + 1 < nss ? ss + 1 : ss, xa, ba);
#pragma unroll
            for (int d = 0; d < 4; ++d) {
                u32x4_t bv[TILES];
#pragma unroll
                for (int j = 0; j < TILES; ++j) bv[j] = bt[buf][(d * TILES + j) * 64 + lane];
#pragma unroll
                for (int mt = 0; mt < MT; ++mt) {
                    if constexpr (BF) {
                        // row bytes 2d, 2d+1 of the lane's 8: codes of A1
                        // rows swapped (00 <-> 11), the two bytes duplicated,
                        // then one v_perm_b32 per pair of BF16 dosages
                        // (slot 2t + h: person t of byte 2d + h)
                        uint32_t y = (uint32_t)(xc[mt] >> (32 * (d >> 1)));
                        const uint32_t u = ~(y ^ (y >> 1)) & (F[mt] & 0x55555555u);
                        y ^= u | (u << 1);
                        const uint32_t P = __builtin_amdgcn_perm(y, y, (d & 1) ? 0x03030202u : 0x01010000u);
                        BfFrag A;
#pragma unroll
                        for (int t = 0; t < 4; ++t) A.u[t] = bf_dose_pair(bf_sel_pair(P >> (2 * t)));
#pragma unroll
                        for (int j = 0; j < TILES; ++j) {
                            BfFrag B;
                            B.u[0] = bv[j][0]; B.u[1] = bv[j][1]; B.u[2] = bv[j][2]; B.u[3] = bv[j][3];
                            acc[mt][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A.v, B.v, acc[mt][j], 0, 0, 0);
                        }
                    } else {
                        i32x4_t A;
#pragma unroll
                        for (int t = 0; t < 4; ++t)
                            A[t] = (int)dose_i8(((xc[mt][d] >> (2 * t)) & 0x03030303u) | F[mt]);
#pragma unroll
                        for (int j = 0; j < TILES; ++j) {
                            const i32x4_t B = i32x4_t{(int)bv[j][0], (int)bv[j][1], (int)bv[j][2], (int)bv[j][3]};
                            acc[mt][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(A, B, acc[mt][j], 0, 0, 0);
                        }
                    }
                }
            }
            ++ss;
        } while (ss < ss_end);
        flush(blk, first);
        first = false;
    } while (ss < nss);
}

// dW [V][K] f32 from the FP64 workspace [V][Kp]
__global__ void k_adjoint_out(const double *__restrict__ ws, uint64_t V, uint64_t K, uint64_t Kp,
                              float *__restrict__ dW) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < V * K) dW[i] = (float)ws[(i / K) * Kp + i % K];
}

hipError_t launch_adjoint_mx(bool bf, const uint8_t *rows, uint64_t stride, uint64_t V, uint64_t N,
                             const uint8_t *flip, const uint32_t *frag, const double *scales, uint32_t tiles,
                             uint64_t K, uint64_t Kp, uint32_t L, double *ws, float *dW, hipStream_t stream) {
    if (V == 0) return hipSuccess;
    if ((reinterpret_cast<uintptr_t>(rows) | stride) & 3u) return hipErrorInvalidValue;
    if (stride >= (1ull << 25)) return hipErrorInvalidValue;
#define GNM_ADJ(T, M, B)                                                                                          \
    hipLaunchKernelGGL((k_adjoint_mx<T, M, B>), dim3((unsigned)((V + 64 * M - 1) / (64 * M))), dim3(256), 0,      \
                       stream, rows, stride, V, N, flip, (const u32x4_t *)frag, scales, Kp, L, ws)
#define GNM_ADJ_BF(T, M)                                                                                          \
    do {                                                                                                          \
        if (bf) GNM_ADJ(T, M, true); else GNM_ADJ(T, M, false);                                                    \
    } while (0)
    if (tiles == 4) GNM_ADJ_BF(4, 4);
    else if (tiles == 8) GNM_ADJ_BF(8, 2);
    else if (tiles == 12) GNM_ADJ_BF(12, 1);
    else if (tiles == 16) GNM_ADJ_BF(16, 1);
    else return hipErrorInvalidValue;
#undef GNM_ADJ_BF
#undef GNM_ADJ
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const uint64_t n = V * K;
    hipLaunchKernelGGL(k_adjoint_out, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, ws, V, K, Kp, dW);
    return hipGetLastError();
}

// per-row selector bytes (value for a flipped row, 0 otherwise) from a flip
// byte array
__global__ void k_flip_bytes(const uint8_t *__restrict__ flip, uint64_t V, uint64_t Rpad, uint8_t value,
                             uint8_t *__restrict__ out) {
    const uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r < Rpad) out[r] = (flip && r < V && flip[r]) ? value : (uint8_t)0;
}

hipError_t launch_flip_bytes(const uint8_t *flip, uint64_t V, uint64_t Rpad, uint8_t value, uint8_t *out,
                             hipStream_t stream) {
    if (Rpad == 0) return hipSuccess;
    hipLaunchKernelGGL(k_flip_bytes, dim3((unsigned)((Rpad + 255) / 256)), dim3(256), 0, stream, flip, V, Rpad,
                       value, out);
    return hipGetLastError();
}

}  // namespace gnm
