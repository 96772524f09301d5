// Fused decode + split-limb MFMA scoring kernel k_score_mx, shared by the
// split-int8 path (gnm_i8.hip, BF = false) and the split-BF16 path
// (gnm_bf16.hip, BF = true).
//
// Weights are written as L integer limbs (signed base-256 digits) times a
// power-of-two scale per (scale block of program rows, weight column):
//   * int8 limbs, every digit in [-128, 127], v_mfma_i32_16x16x64_i8 (64 rows
//     per MFMA); an int32 block sum is exact (65536 rows * 2 * 128 = 2^24);
//   * BF16 limbs, low digits in [-128, 127] and the top digit in
//     [-255, 255], all exact in BF16, v_mfma_f32_16x16x32_bf16 (32 rows per
//     MFMA); an FP32 block sum is an exact integer (32768 rows * 2 * 255 <
//     2^24).
// Dosages 0/1/2 are exact in both types, so every product and every partial
// sum inside a block is exact; blocks are flushed to FP64 with their scale.
//
// Data path, per workgroup of WAVES waves (64 people per wave, 16*WAVES
// packed bytes of every row) and per stage of 64 program rows (one int8
// k-step or two BF16 k-steps):
//   * each thread loads 4 rows x 4 packed bytes (one dword per row;
//     consecutive threads read consecutive dwords of a row), transposes the
//     4x4 bytes in registers and writes 4 dwords into a column-major LDS tile
//     [byte column][64 rows] (column pitch 80 bytes, row blocks swizzled per
//     column so writes and 128-bit reads avoid bank conflicts);
//   * one extra tile column holds a per-row selector byte for rows that read
//     the A1 allele (flipped): 4 on the int8 path, 0x55 on the BF16 path; 0
//     otherwise and for rows outside the split (whose data bytes are zero);
//     bit 7 marks count rows when missing calls are counted on the vector
//     ALUs instead of in count tiles (one membership class);
//   * lane (i, g) of wave w reads column 16w + i, rows 16g..16g+15 as one
//     128-bit LDS read: the packed codes of people 4(16w+i)+t, t = 0..3, for
//     16 variants, and decodes them with v_perm_b32 table lookups (see the
//     two compute steps);
//   * M-tile t of the wave holds people 64w + 4i + t; B fragments (16 bytes
//     per lane, the MFMA B operand, pre-arranged by the plan in the row order
//     of the decode) go from memory to LDS by buffer LDS DMA, without passing
//     through registers.
// Loads of a batch of D stages are issued together (buffer loads: one
// uniform descriptor per stage and a 32-bit lane offset, so no per-lane
// 64-bit addresses stay live), stored, then consumed behind LDS-only
// barriers.  The scale-block flush sits after the batch loop, so the
// accumulators carried by that loop stay in accumulation registers.
#pragma once

#include "gnm_internal.hpp"

#include <type_traits>
#include <utility>

namespace gnm {

typedef int i32x4_t __attribute__((ext_vector_type(4)));
typedef uint32_t u32x4_t __attribute__((ext_vector_type(4)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));

namespace {

constexpr int kI8Stage = 64;      // rows per stage = one k-step of v_mfma_i32_16x16x64_i8
constexpr int kColPitch = 80;     // LDS bytes per tile column (64 rows + 16 pad)
constexpr int kI8RsrcWord3 = 0x00020000;   // raw buffer descriptor (gfx9): 32-bit data format

// LDS dword of (tile column col, row quad m): the four 16-row blocks of a
// column are permuted by an XOR that depends on col / 16, so the transposed
// dword writes of a wave (16 columns x 4 row quads) fall on distinct banks,
// while the 128-bit reads of a 16-row block stay contiguous and aligned.
__device__ __forceinline__ int tile_dword(int col, int m) {
    return col * (kColPitch / 4) + (m ^ (4 * ((col >> 4) & 3)));
}

// LDS written and read back by the same wave: LDS instructions of one wave
// complete in order, so waiting for the writes is enough (no s_barrier).
__device__ __forceinline__ void lds_wave_fence() {
    asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
}

__device__ __forceinline__ void lds_barrier_i8() {
    asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

// Barrier after LDS DMA (global_load_lds): the DMA completes on the vector
// memory counter, so every wave waits for its own transfers first.
__device__ __forceinline__ void dma_lds_barrier() {
    // the same wait as an instruction the compiler sees (gfx9 encoding:
    // vmcnt 0, expcnt 7, lgkmcnt 0), so it knows the DMA has landed and
    // adds no waits for it to later LDS reads; the asm repeats it so that
    // no memory operation can slip between the wait and the barrier
    __builtin_amdgcn_s_waitcnt(0x0070);
    asm volatile("s_waitcnt vmcnt(0) lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

// 4x4 byte transpose: r[k] byte b -> c[b] byte k.
__device__ __forceinline__ void transpose4x4(const uint32_t (&r)[4], uint32_t (&c)[4]) {
    const uint32_t lo01 = __builtin_amdgcn_perm(r[1], r[0], 0x05010400u);
    const uint32_t hi01 = __builtin_amdgcn_perm(r[1], r[0], 0x07030602u);
    const uint32_t lo23 = __builtin_amdgcn_perm(r[3], r[2], 0x05010400u);
    const uint32_t hi23 = __builtin_amdgcn_perm(r[3], r[2], 0x07030602u);
    c[0] = __builtin_amdgcn_perm(lo23, lo01, 0x05040100u);
    c[1] = __builtin_amdgcn_perm(lo23, lo01, 0x07060302u);
    c[2] = __builtin_amdgcn_perm(hi23, hi01, 0x05040100u);
    c[3] = __builtin_amdgcn_perm(hi23, hi01, 0x07060302u);
}

// selector bytes (code | flip<<2) -> i8 dosage / missing indicator
__device__ __forceinline__ uint32_t dose_i8(uint32_t sel) {
    return __builtin_amdgcn_perm(0x00010002u, 0x02010000u, sel);
}
__device__ __forceinline__ uint32_t miss_i8(uint32_t sel) {
    return __builtin_amdgcn_perm(0x00000100u, 0x00000100u, sel);
}

// 16 bytes per lane from a buffer (lane offset voff) into LDS at
// dst + 16 * lane (buffer_load_dwordx4 ... lds); dst must be wave-uniform.
__device__ __forceinline__ void lds_dma_b128(__amdgpu_buffer_rsrc_t rsrc, void *dst, uint32_t voff) {
    __builtin_amdgcn_raw_ptr_buffer_load_lds(rsrc, (__attribute__((address_space(3))) void *)dst, 16, voff, 0, 0, 0);
}


// BF16 decode.  A 16-byte MFMA operand of 8 BF16 values as 4 dwords.
union BfFrag {
    bf16x8_t v;
    uint32_t u[4];
};

// selector [c0, c0+4, c1, c1+4] from the 2-bit codes in bits 0-1 of bytes 0
// and 2 of y (bytes 1 and 3 hold copies): (y & 0x03030303) | 0x04000400 as
// one v_and_or_b32
__device__ __forceinline__ uint32_t bf_sel_pair(uint32_t y) {
    uint32_t x;
    const uint32_t m = 0x03030303u, c = 0x04000400u;
    asm("v_and_or_b32 %0, %1, %2, %3" : "=v"(x) : "v"(y), "v"(m), "v"(c));
    return x;
}
// two BF16 dosages: low bytes {00 00 80 00}, high bytes {00 00 3F 40} for
// codes {00 01 10 11} give {0, 0, 1.0, 2.0}
__device__ __forceinline__ uint32_t bf_dose_pair(uint32_t x) {
    return __builtin_amdgcn_perm(0x403F0000u, 0x00800000u, x);
}
// two BF16 missing indicators: 1.0 for code 01
__device__ __forceinline__ uint32_t bf_miss_pair(uint32_t x) {
    return __builtin_amdgcn_perm(0x00003F00u, 0x00008000u, x);
}

}  // namespace

template <int TMAX, int WAVES, int D, bool VEC, bool IDENT, int SETS, bool BF>
// The 3-tile, 2-stage instances (int8 and BF16) are asked to fit four waves
// per SIMD (128 registers; a few values outside the MFMA blocks are spilled),
// the 2-stage 5-tile BF16 and 4- and 5-tile int8 instances three (168
// registers, three workgroups per CU; at 50k x 500k x 16 the 5-tile groups
// took 3.80 against 4.28 ms in bf16x4 and 2.33 against 2.72 ms in i8x4 with
// two), the other BF16 instances two waves per SIMD (256 registers).
// SETS > 1: the waves form SETS sets over the same people, each set with its
// own group of TMAX tiles, so one row read serves SETS groups (the first
// set loads and stores the rows for all).
__global__ __launch_bounds__(64 * WAVES) __attribute__((amdgpu_waves_per_eu(TMAX == 3 && D == 2 ? 4 : ((TMAX == 5 || (!BF && TMAX == 4)) && D == 2 ? 3 : (BF ? 2 : 1))))) void k_score_mx(DevicePlan plan, const uint8_t *__restrict__ rows,
                                                         uint64_t stride, uint64_t bpv, uint64_t v0, uint64_t p0,
                                                         uint64_t p1, uint64_t split_base, uint64_t rps,
                                                         uint32_t splits, uint32_t person_tiles,
                                                         uint64_t num_people, uint32_t tile_begin0,
                                                         uint32_t tile_end, uint32_t ngroups,
                                                         double *__restrict__ part,
                                                         uint32_t *__restrict__ cpart) {
    static_assert(WAVES % SETS == 0, "whole sets of waves");
    constexpr int PW = WAVES / SETS;                 // waves over distinct people
    constexpr int kBytes = 16 * PW;                  // packed bytes per row per workgroup
    constexpr int kCols = kBytes + 1;                // + the selector column
    constexpr int kTileWords = kCols * kColPitch / 4;
    constexpr int KS = BF ? 2 : 1;                   // MFMA k-steps per 64-row stage (BF16: 32 rows each)
    constexpr int kBT = SETS * TMAX;                 // B tiles per stage (all sets)
    constexpr int kBU = kBT * KS;                    // B fragments per stage
    constexpr int kBW = (kBU + WAVES - 1) / WAVES;   // B fragments each wave fetches per stage
    constexpr uint64_t kFlush = BF ? (uint64_t)kBfFlushRows : (uint64_t)kI8FlushRows;
    static_assert(kFlush % (2 * D * kI8Stage) == 0, "scale blocks must hold whole pairs of batches");
    // one LDS tile per stage of a batch: all D stages are stored, then one
    // barrier, then all D are consumed (two barriers per batch of D stages)
    __shared__ uint32_t tile[D][kTileWords];
    // B fragments are double-buffered in two LDS objects: the next batch's
    // DMA lands in one while this batch's MFMAs read the other.  Batches
    // alternate between them with a compile-time parity, so the compiler
    // sees that the reads never touch the object a DMA is writing and adds
    // no memory-counter wait before them.
    __shared__ u32x4_t btile0[D][kBU * 64];
    __shared__ u32x4_t btile1[D][kBU * 64];
    auto btile = [&](auto pc) __attribute__((always_inline)) -> u32x4_t(&)[D][kBU * 64] {
        if constexpr (decltype(pc)::value == 0) return btile0;
        else return btile1;
    };

    // Workgroups are dispatched round-robin over the 8 XCDs (L mod 8).  One
    // launch covers every tile group: the groups of one (split, person tile)
    // are consecutive slots of the same XCD, so they read the same .bed rows
    // through that XCD's L2 at about the same time, and every workgroup of a
    // split stays on one XCD (its B fragments are shared through that L2).
    const uint32_t L = blockIdx.x;
    const uint32_t xcd = L & 7u, slot = L >> 3;
    const uint32_t group = slot % ngroups, rest = slot / ngroups;
    const uint32_t s = xcd + 8u * (rest / person_tiles);
    const uint32_t ptile = rest % person_tiles;
    if (s >= splits) return;
    const uint32_t gtile0 = tile_begin0 + group * kBT;   // first tile of the workgroup's sets

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int pwave = wave % PW, set = wave / PW;
    // missing calls counted on the vector ALUs (one membership class, no
    // count tiles), by the first set of waves of the first tile group
    const bool vcount = plan.vcount != 0 && group == 0 && set == 0;
    uint32_t vcnt[4] = {0u, 0u, 0u, 0u};   // missing calls of people 4 li + t over the lane's rows
    // the first set loads and stores the rows of every stage
    const bool loader = SETS == 1 || wave < PW;
    const uint32_t tile_begin = gtile0 + (uint32_t)set * TMAX;
    const uint32_t ntile = tile_begin >= tile_end ? 0u
                           : (tile_end - tile_begin < (uint32_t)TMAX ? tile_end - tile_begin : (uint32_t)TMAX);
    const int li = lane & 15, lg = lane >> 4;
    const uint64_t col0 = (uint64_t)ptile * kBytes;
    const uint64_t pw0 = (uint64_t)ptile * (64 * PW) + (uint64_t)pwave * 64;
    const uint64_t sb = split_base + (uint64_t)s * rps;
    const uint64_t r_begin = p0 > sb ? p0 : sb;
    const uint64_t se = sb + rps;
    const uint64_t r_end = p1 < se ? p1 : se;
    const uint32_t wt = plan.wtiles;
    const uint32_t tiles_total = plan.wtiles + plan.ctiles;
    // weight tiles of this group (the first nw tiles); the others count missing calls
    const uint32_t nw = tile_begin < wt ? ((wt - tile_begin) < ntile ? (wt - tile_begin) : ntile) : 0;
    // load role: rows 4rg..4rg+3, packed bytes 4c4..4c4+3 of the workgroup's span
    // (consecutive threads take consecutive dwords of a row: each load
    // instruction reads 64 contiguous bytes of each of its rows)
    const int rg = tid / (4 * PW), c4 = tid % (4 * PW);
    const uint64_t ld_col = col0 + 4 * (uint64_t)c4;
    const bool ld_any = ld_col < bpv;
    // Every stage issues the same loads on every path (one dword per row):
    // past the row end the dword is read at the last full dword of the row
    // (bpv >= 4 on the vector path) and shifted / masked back at store time.
    const uint64_t ld_colc = ld_col + 4 <= bpv ? ld_col : bpv - 4;
    const uint32_t col_shift = ld_any ? (uint32_t)(ld_col - ld_colc) * 8u : 0u;
    const uint32_t col_mask = ld_col + 4 <= bpv ? 0xFFFFFFFFu
                              : (ld_any ? (1u << (8u * (uint32_t)(bpv - ld_col))) - 1u : 0u);
    const u32x4_t *__restrict__ bfrag = (const u32x4_t *)plan.bfrag;

    uint32_t ra[D][4];      // row dwords of each stage in flight
    uint32_t rs[D];         // selector bytes of rows 4rg..4rg+3 (threads with c4 == 0)
    uint32_t rk[D];         // keep mask bits 0..3 for rows 4rg+k inside the split
    bool rfull[D];          // whole stage inside the split (uniform): no masking
    // Fast path: every row of the stage is inside [r_begin, r_end) of an
    // identity block and the workgroup's byte span lies inside the row, so
    // the four row loads are one per-thread pointer plus uniform offsets.
    const bool span_full = VEC && col0 + kBytes <= bpv;
    // per-lane byte offset of the fast path's first row dword from the
    // stage's first row (the launcher keeps 64 rows x stride below 2^31)
    const uint32_t row_voff = (uint32_t)(4 * (uint64_t)rg * stride + ld_col);

    const uint64_t last_st = ((r_end - 1) / kI8Stage) * kI8Stage;
    // Stages are numbered from st0, the split's first row aligned down to a
    // pair of batches (scale blocks start at even batch indices, so a block
    // is a whole number of batch pairs); per-stage tests use these 32-bit indices (scalar compares, where
    // 64-bit row compares take vector instructions):
    //   nst      stages of the split
    //   fast_lo, fast_hi   stages wholly inside [r_begin, r_end) (fast path)
    //   last_i   the stage holding the split's last row
    const uint64_t st0 = (r_begin / (2 * D * kI8Stage)) * (2 * D * kI8Stage);
    const bool nonempty = r_begin < r_end;
    const uint32_t nst = nonempty ? (uint32_t)((r_end - st0 + kI8Stage - 1) / kI8Stage) : 0u;
    const uint32_t fast_lo = (uint32_t)((r_begin - st0 + kI8Stage - 1) / kI8Stage);
    const uint32_t fast_hi = (IDENT && span_full && nonempty) ? (uint32_t)((r_end - st0) / kI8Stage) : 0u;
    const uint32_t last_i = nonempty ? (uint32_t)((last_st - st0) / kI8Stage) : 0u;
    // the split's row range relative to st0, and its first row's variant
    // (non-identity rows)
    const uint32_t rb_rel = (uint32_t)(r_begin - st0), re_rel = nonempty ? (uint32_t)(r_end - st0) : rb_rel;
    const uint32_t var_begin = (!IDENT && nonempty) ? plan.prog_var[r_begin] : (uint32_t)v0;
    // B fragments: wave w fetches fragments u = w, w + WAVES, ... of the
    // sets' kBU fragments (tile u / KS, k-step u % KS) by LDS DMA (lane l's
    // 16 bytes land at btile(par)[d][64 u + l]); tiles past tile_end are fetched
    // clamped to the last tile and their products are never flushed
    uint32_t boff[kBW];     // byte offsets within a stage's fragment block
#pragma unroll
    for (int m = 0; m < kBW; ++m) {
        const uint32_t u = (uint32_t)(wave + m * WAVES);
        const uint32_t j = gtile0 + u / KS;
        const uint32_t jc = j < tile_end ? j : tile_end - 1;
        boff[m] = ((jc * KS + u % KS) * 64 + lane) * 16;
    }
    // fragment block of stage index si (clamped to last_i by the caller)
    const u32x4_t *__restrict__ bfrag_st0 = bfrag + (st0 / kI8Stage) * tiles_total * KS * 64;
    const uint32_t bstep = tiles_total * KS * 64;
    auto fetch_b = [&](auto par, int d, uint32_t si) __attribute__((always_inline)) {
        const __amdgpu_buffer_rsrc_t rb = __builtin_amdgcn_make_buffer_rsrc(
            (void *)(bfrag_st0 + (uint64_t)si * bstep), 0, 0x7FFFFFFF, kI8RsrcWord3);
#pragma unroll
        for (int m = 0; m < kBW; ++m) {
            const int u = wave + m * WAVES;
            if (kBW * WAVES == kBU || u < kBU)
                lds_dma_b128(rb, &btile(par)[d][u * 64], boff[m]);
        }
    };
    auto load_row = [&](uint64_t var, uint64_t lcol) __attribute__((always_inline)) -> uint32_t {
        const uint8_t *src = rows + var * stride;
        if (VEC) return *(const uint32_t *)(src + (lcol + 4 <= bpv ? lcol : bpv - 4));
        uint32_t x = 0;
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const uint64_t o = lcol + b;
            const uint32_t byte = src[o < bpv ? o : bpv - 1];
            x |= (o < bpv ? byte : 0u) << (8 * b);
        }
        return x;
    };
    auto issue = [&](int d, uint32_t si) __attribute__((always_inline)) {
        const uint64_t st_req = st0 + (uint64_t)si * kI8Stage;
        if (si >= fast_lo && si < fast_hi) {   // IDENT && span_full, stage inside the split
            // buffer loads: a uniform descriptor per stage, one 32-bit lane
            // offset, and the row step as the scalar offset
            const __amdgpu_buffer_rsrc_t rr = __builtin_amdgcn_make_buffer_rsrc(
                (void *)(rows + (st_req - p0) * stride), 0, 0x7FFFFFFF, kI8RsrcWord3);
#pragma unroll
            for (int k = 0; k < 4; ++k)
                ra[d][k] = __builtin_amdgcn_raw_buffer_load_b32(rr, row_voff, (uint32_t)(k * stride), 0);
            rk[d] = 0xFu;
            rfull[d] = true;
            const __amdgpu_buffer_rsrc_t rf = __builtin_amdgcn_make_buffer_rsrc(
                (void *)(plan.flipbyte + st_req), 0, 0x7FFFFFFF, kI8RsrcWord3);
            rs[d] = __builtin_amdgcn_raw_buffer_load_b32(rf, 4u * (uint32_t)rg, 0, 0);
            return;
        }
        // Slow path (split edges, partial spans, non-identity rows): its
        // per-lane inputs pass through an opaque copy, so what it derives
        // from them is computed here and not kept live across the loop.
        uint32_t rg4 = 4u * (uint32_t)rg;
        uint64_t lcol = ld_col;
        asm volatile("" : "+v"(rg4), "+v"(lcol));
        const uint64_t st = st_req < last_st ? st_req : last_st;
        const bool same = st == st_req;
        uint32_t keep = 0;
        if constexpr (IDENT) {
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const uint64_t r = st + rg4 + k;
                const bool ok = same && r >= r_begin && r < r_end && lcol < bpv;
                keep |= ok ? (1u << k) : 0u;
                const uint64_t rc = r < r_begin ? r_begin : (r >= r_end ? r_end - 1 : r);
                ra[d][k] = load_row(rc - p0, lcol);
            }
        } else {
            // the variants of the thread's four rows in one 16-byte load
            // (st + rg4 + 3 < Rpad: st is at most the split's last stage);
            // rows outside the split read the split's first variant instead,
            // a row of this block, and are masked at store time.  Row tests
            // in 32 bits, relative to the split's first stage.
            const u32x4_t pv = *(const u32x4_t *)(plan.prog_var + st + rg4);
            const uint32_t rel = (uint32_t)(st - st0) + rg4;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const bool ok = same && rel + k >= rb_rel && rel + k < re_rel && lcol < bpv;
                keep |= ok ? (1u << k) : 0u;
                ra[d][k] = load_row((uint64_t)((ok ? pv[k] : var_begin) - (uint32_t)v0), lcol);
            }
        }
        rk[d] = keep;
        rfull[d] = false;
        rs[d] = *(const uint32_t *)(plan.flipbyte + st + rg4);   // selector bytes of rows st+4rg..+3 (c4 == 0)
    };
    auto store = [&](int buf, int d) __attribute__((always_inline)) {
        uint32_t r[4], c[4];
        const uint32_t keep = rk[d];
        if (rfull[d]) {
#pragma unroll
            for (int k = 0; k < 4; ++k) r[k] = ra[d][k];
        } else {
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const uint32_t x = VEC ? (ra[d][k] >> col_shift) & col_mask : ra[d][k];
                r[k] = ((keep >> k) & 1u) ? x : 0u;
            }
        }
        transpose4x4(r, c);
        uint32_t *t = tile[buf];
#pragma unroll
        for (int b = 0; b < 4; ++b) t[tile_dword(4 * c4 + b, rg)] = c[b];
        if (c4 == 0) {
            // selector bytes, zero for rows outside the split
            uint32_t sel = rs[d];
            if (!rfull[d]) {
                const uint32_t km = ((keep & 1u) ? 0xFFu : 0u) | ((keep & 2u) ? 0xFF00u : 0u) |
                                    ((keep & 4u) ? 0xFF0000u : 0u) | ((keep & 8u) ? 0xFF000000u : 0u);
                sel &= km;
            }
            t[tile_dword(kBytes, rg)] = sel;
        }
    };

    // block sums [person tile t][tile j]: exact int32 (int8 limbs) or exact
    // FP32 integers (BF16 limbs)
    typedef std::conditional_t<BF, f32x4_t, i32x4_t> AccV;
    typedef AccV Acc[4][TMAX];
    // zero from an opaque copy of 0: the initialisation stays in the branch
    // that uses the set instead of being hoisted above the layout dispatch
    auto zero_acc = [&](Acc &acc) __attribute__((always_inline)) {
        int z = 0;
        asm volatile("" : "+v"(z));
        const auto e = [&] {
            if constexpr (BF) return __builtin_bit_cast(float, z); else return z;
        }();
#pragma unroll
        for (int t = 0; t < 4; ++t)
#pragma unroll
            for (int j = 0; j < TMAX; ++j) acc[t][j] = AccV{e, e, e, e};
    };

    // Block flush: acc[t][j][q] is the int32 block sum of person
    // pw0 + 16 lg + 4 q + t (MFMA output row 4 lg + q of M-tile t) and tile
    // column li.  The wave stages one tile through its own LDS buffer so
    // that each lane then owns one column for 16 people, and adds the values
    // to the split partials in FP64 (weight tiles, times the block scale) or
    // as counts.
    // the staging buffer shares the .bed tiles' LDS when they are large
    // enough (no tile is in use while a block is flushed)
    // Otherwise it is btile1 when that is large enough: a flush follows an
    // odd batch (or no batch), so btile1 has no DMA in flight, its last reads
    // retired at that batch's closing barrier, and the next DMA into it (the
    // next block's first batch) comes after the flush's closing barrier.
    constexpr bool kStageInTile = D * kTileWords >= WAVES * 64 * 16;
    constexpr bool kStageInB = !kStageInTile && D * kBU * 64 * 4 >= WAVES * 64 * 16;
    __shared__ int stage_own[(kStageInTile || kStageInB) ? 1 : WAVES * 64 * 16];
    auto flush = [&](Acc &acc, uint64_t blk, bool first, bool empty) __attribute__((always_inline)) {
        // opaque to the optimiser, so the addresses below are formed here
        // rather than hoisted out of the row loop into live registers
        uint64_t prow = (uint64_t)s * num_people + pw0;
        uint32_t ncol = plan.ncol, nclass = plan.nclass, tb = tile_begin, lane_c = (uint32_t)li;
        asm volatile("" : "+v"(prow), "+v"(ncol), "+v"(nclass), "+v"(tb), "+v"(lane_c));
        ncol = __builtin_amdgcn_readfirstlane(ncol);
        nclass = __builtin_amdgcn_readfirstlane(nclass);
        tb = __builtin_amdgcn_readfirstlane(tb);
        int *sbuf = (kStageInTile ? (int *)&tile[0][0] : (kStageInB ? (int *)&btile1[0][0] : stage_own)) +
                    wave * (64 * 16);
        auto flush_tile = [&](auto jc) __attribute__((always_inline)) {
            constexpr int j = decltype(jc)::value;
            __builtin_amdgcn_sched_barrier(0);
            // keeps this tile's sums in accumulation registers up to here,
            // so only 16 of them are copied out at a time
#pragma unroll
            for (int t = 0; t < 4; ++t) asm volatile("" : "+a"(acc[t][j]));
            lds_wave_fence();    // the previous tile's reads of sbuf are done
#pragma unroll
            for (int t = 0; t < 4; ++t) {
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    // the element is read into a scalar first: __builtin_bit_cast
                    // applied to a vector element subscript compiles to element 0
                    // for every q (seen with ROCm 7.2 clang)
                    if constexpr (BF) {
                        const float x = acc[t][j][q];
                        sbuf[(16 * lg + 4 * q + t) * 16 + li] = __float_as_int(x);
                    } else {
                        sbuf[(16 * lg + 4 * q + t) * 16 + li] = acc[t][j][q];
                    }
                }
                acc[t][j] = AccV{};
            }
            lds_wave_fence();
            const uint32_t tl = tb + j;
            const bool wtile = tl < wt;
            const uint32_t c = (wtile ? tl : tl - wt) * 16 + lane_c;
            const uint32_t nc = wtile ? ncol : nclass;
            if ((uint32_t)j < ntile && c < nc) {
                const double sc = (wtile && !empty) ? plan.scales[blk * ncol + c] : 0.0;
                // the first flush of a workgroup stores its partials; later
                // ones add to them with no-return device atomics (each
                // partial has one writer, so the sum is the same FP64 add a
                // read-modify-write would do, with no load latency and no
                // registers held for it)
#pragma unroll 1
                for (int i = 0; i < 16; ++i) {
                    const uint32_t pl = 4 * i + lg;  // person of the wave, 0..63
                    if (pw0 + pl >= num_people) continue;
                    const int vb = sbuf[pl * 16 + lane_c];
                    const double v = BF ? (double)__builtin_bit_cast(float, vb) : (double)vb;
                    const uint64_t row = prow + pl;
                    if (wtile) {
                        double *dd = part + row * ncol + c;
                        if (first) *dd = v * sc;
                        else unsafeAtomicAdd(dd, v * sc);
                    } else {
                        uint32_t *dd = cpart + row * nclass + c;
                        if (first) *dd = (uint32_t)v;
                        else atomicAdd(dd, (uint32_t)v);
                    }
                }
            }
            __builtin_amdgcn_sched_barrier(0);
        };
        [&]<int... J>(std::integer_sequence<int, J...>) __attribute__((always_inline)) {
            (flush_tile(std::integral_constant<int, J>{}), ...);
        }(std::make_integer_sequence<int, TMAX>{});
        if (vcount) {
            // the four lanes of a byte column (lg = 0..3) hold the counts of
            // its people over different rows: summed through LDS, one person
            // per lane
            lds_wave_fence();
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                sbuf[lg * 64 + li * 4 + t] = (int)vcnt[t];
                vcnt[t] = 0u;
            }
            lds_wave_fence();
            const uint32_t pl = (uint32_t)lane;
            if (pw0 + pl < num_people) {
                const uint32_t x = (uint32_t)(sbuf[pl] + sbuf[64 + pl] + sbuf[128 + pl] + sbuf[192 + pl]);
                uint32_t *dd = cpart + (prow + pl) * nclass;
                if (first) *dd = x;
                else atomicAdd(dd, x);
            }
        }
        // staging inside the .bed tiles: no wave may store the next batch's
        // tiles while another still reads its staged sums
        if constexpr (kStageInTile || kStageInB) lds_barrier_i8();
    };

    // Missing calls (code 01) of the lane's four people over its 16 rows of
    // the stage: per dword of 4 rows the missing bits sit at 2t + 8b
    // (person t, row b); two dwords are interleaved (the second shifted to
    // the odd bits) so one popcount per person covers 8 rows.  Rows that are
    // not count rows (the second row of a mixed-flip variant) are masked by
    // the selector's bit 7; identity blocks have none.
    auto count_missing = [&](const u32x4_t &X, const u32x4_t &F) __attribute__((always_inline)) {
        uint32_t m[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const uint32_t x = X[q];
            uint32_t mm = x & ~(x >> 1) & 0x55555555u;
            if constexpr (!IDENT) {
                // 0xFF in each byte whose selector has bit 7 set
                const uint32_t cm = __builtin_amdgcn_perm(F[q], F[q] << 8, 0x0B090A08u);
                mm &= cm;
            }
            m[q] = mm;
        }
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const uint32_t comb = m[2 * h] | (m[2 * h + 1] << 1);
#pragma unroll
            for (int t = 0; t < 4; ++t) vcnt[t] += (uint32_t)__builtin_popcount(comb & (0x03030303u << (2 * t)));
        }
    };

    // BF16 stage: the lane's 16 rows of its byte column (X, rows 16 lg + 4 q
    // + b in byte b of X[q]) feed two v_mfma_f32_16x16x32_bf16 k-steps, rows
    // 16 lg + 0..7 (h = 0) and 16 lg + 8..15 (h = 1); the plan writes the B
    // fragments in the same row order.  A1 rows (selector byte 0x55) have
    // their codes 00 and 11 swapped first, so one table serves every row:
    // a row pair's bytes are duplicated ([b0, b0, b1, b1], shared by the four
    // person tiles), each person's 2-bit codes are isolated into selectors
    // [c0, c0 + 4, c1, c1 + 4] and one v_perm_b32 maps them to two BF16
    // dosages (low bytes 00 00 80 00 and high bytes 00 00 3F 40 for codes
    // 00 01 10 11: 0, 0, 1.0, 2.0), another to the missing indicator.
    auto compute_bf = [&](Acc &acc, auto nwc, auto par, int buf, const u32x4_t &X,
                          const u32x4_t &F) __attribute__((always_inline)) {
        if constexpr (BF) {
        constexpr int NW = decltype(nwc)::value;
        uint32_t P[8];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const uint32_t x = X[q];
            const uint32_t u = ~(x ^ (x >> 1)) & (F[q] & 0x55555555u);
            const uint32_t y = x ^ (u | (u << 1));
            P[2 * q] = __builtin_amdgcn_perm(y, y, 0x01010000u);
            P[2 * q + 1] = __builtin_amdgcn_perm(y, y, 0x03030202u);
        }
        // B fragments of the first k-step; those of the second are read
        // into the same registers right after each one's last MFMA of the
        // first k-step, so their LDS latency runs under the remaining MFMAs
        u32x4_t bv[TMAX];
#pragma unroll
        for (int j = 0; j < TMAX; ++j) bv[j] = btile(par)[buf][((set * TMAX + j) * 2) * 64 + lane];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
#pragma unroll
            for (int tt = 0; tt < 4; ++tt) {
                BfFrag A, M;
                uint32_t s[4];
#pragma unroll
                for (int m = 0; m < 4; ++m) s[m] = bf_sel_pair(P[4 * h + m] >> (2 * tt));
                if constexpr (NW != 0) {
#pragma unroll
                    for (int m = 0; m < 4; ++m) A.u[m] = bf_dose_pair(s[m]);
                }
                auto mfma = [&](int j, const BfFrag &op) __attribute__((always_inline)) {
                    BfFrag B;
                    B.u[0] = bv[j][0]; B.u[1] = bv[j][1]; B.u[2] = bv[j][2]; B.u[3] = bv[j][3];
                    acc[tt][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(op.v, B.v, acc[tt][j], 0, 0, 0);
                };
                if constexpr (NW >= 0) {
                    if constexpr (NW < TMAX) {
#pragma unroll
                        for (int m = 0; m < 4; ++m) M.u[m] = bf_miss_pair(s[m]);
                    }
                    __builtin_amdgcn_s_setprio(1);
#pragma unroll
                    for (int j = 0; j < TMAX; ++j) {
                        mfma(j, j < NW ? A : M);
                        if (h == 0 && tt == 3) bv[j] = btile(par)[buf][((set * TMAX + j) * 2 + 1) * 64 + lane];
                    }
                    __builtin_amdgcn_s_setprio(0);
                } else {
#pragma unroll
                    for (int m = 0; m < 4; ++m) A.u[m] = bf_dose_pair(s[m]);
#pragma unroll
                    for (int j = 0; j < TMAX; ++j)
                        if ((uint32_t)j < nw) mfma(j, A);
                    if (nw < ntile) {
#pragma unroll
                        for (int m = 0; m < 4; ++m) M.u[m] = bf_miss_pair(s[m]);
#pragma unroll
                        for (int j = 0; j < TMAX; ++j)
                            if ((uint32_t)j >= nw && (uint32_t)j < ntile) mfma(j, M);
                    }
                    if (h == 0 && tt == 3) {
#pragma unroll
                        for (int j = 0; j < TMAX; ++j) bv[j] = btile(par)[buf][((set * TMAX + j) * 2 + 1) * 64 + lane];
                    }
                }
            }
        }
        }
    };

    // compute(nwc, par, buf): one stage.  With a compile-time layout
    // (nwc >= 0: the first nwc tiles are weight tiles, the others count
    // tiles) the stage is straight-line code, so the scheduler can overlap
    // one stage's decode with the previous stage's MFMAs; tile slots past
    // ntile are computed on a duplicated fragment and never flushed.  With
    // nwc < 0 the layout is read at run time (per-tile branches).
    auto compute = [&](Acc &acc, auto nwc, auto par, int buf) __attribute__((always_inline)) {
        constexpr int NW = decltype(nwc)::value;
        const uint32_t *t = tile[buf];
        const u32x4_t X = *(const u32x4_t *)&t[tile_dword(pwave * 16 + li, 4 * lg)];
        const u32x4_t F = *(const u32x4_t *)&t[tile_dword(kBytes, 4 * lg)];
        if (vcount) count_missing(X, F);
        if constexpr (BF) {
            compute_bf(acc, nwc, par, buf, X, F);
        } else {
        u32x4_t bv[TMAX];
#pragma unroll
        for (int j = 0; j < TMAX; ++j) bv[j] = btile(par)[buf][(set * TMAX + j) * 64 + lane];
        // person tile by person tile: only one tile's selectors and operands
        // are live at a time
#pragma unroll
        for (int tt = 0; tt < 4; ++tt) {
            uint32_t sel[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) sel[q] = ((X[q] >> (2 * tt)) & 0x03030303u) | (F[q] & 0x07070707u);
            i32x4_t A, M;
            if constexpr (NW != 0) {
#pragma unroll
                for (int q = 0; q < 4; ++q) A[q] = (int)dose_i8(sel[q]);
            }
            auto mfma = [&](int j, const i32x4_t &op) __attribute__((always_inline)) {
                const i32x4_t B = i32x4_t{(int)bv[j][0], (int)bv[j][1], (int)bv[j][2], (int)bv[j][3]};
                acc[tt][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(op, B, acc[tt][j], 0, 0, 0);
            };
            if constexpr (NW >= 0) {
                if constexpr (NW < TMAX) {
#pragma unroll
                    for (int q = 0; q < 4; ++q) M[q] = (int)miss_i8(sel[q]);
                }
                // a wave issuing its MFMAs goes ahead of waves decoding, so
                // the matrix pipe is fed while they run (3% at 64 scores)
                __builtin_amdgcn_s_setprio(1);
#pragma unroll
                for (int j = 0; j < TMAX; ++j) mfma(j, j < NW ? A : M);
                __builtin_amdgcn_s_setprio(0);
            } else {
#pragma unroll
                for (int j = 0; j < TMAX; ++j)
                    if ((uint32_t)j < nw) mfma(j, A);
                if (nw < ntile) {
#pragma unroll
                    for (int q = 0; q < 4; ++q) M[q] = (int)miss_i8(sel[q]);
#pragma unroll
                    for (int j = 0; j < TMAX; ++j)
                        if ((uint32_t)j >= nw && (uint32_t)j < ntile) mfma(j, M);
                }
            }
        }
        }
    };

    if (r_begin >= r_end) {
        Acc acc;
        zero_acc(acc);
        flush(acc, 0, true, true);
        return;
    }
    // Stages are processed in batches of D, software-pipelined: a batch's
    // row loads and B-fragment DMA are issued before the previous batch is
    // computed, so their latency overlaps the MFMAs; a batch is stored to
    // LDS only once its loads have landed.
    auto issue_batch = [&](uint32_t i) __attribute__((always_inline)) {
        // stages past the split are loaded clamped and never stored
        if (!loader) return;
        [&]<int... I>(std::integer_sequence<int, I...>) __attribute__((always_inline)) {
            (issue(I, i + I), ...);
        }(std::make_integer_sequence<int, D>{});
    };
    auto fetch_batch = [&](auto par, uint32_t i) __attribute__((always_inline)) {
        [&]<int... I>(std::integer_sequence<int, I...>) __attribute__((always_inline)) {
            (fetch_b(par, I, i + I < last_i ? i + I : last_i), ...);
        }(std::make_integer_sequence<int, D>{});
    };
    auto main_loop = [&](auto nwc) __attribute__((always_inline)) {
        Acc acc;   // one accumulator set per copy of the loop
        zero_acc(acc);
        using P0 = std::integral_constant<int, 0>;
        using P1 = std::integral_constant<int, 1>;
        uint32_t i0 = 0;
        issue_batch(i0);
        fetch_batch(P0{}, i0);
        // One batch whose B fragments sit in btile(par): store its rows,
        // then (after the barrier that also retires the previous batch's
        // reads of the other buffer) issue the next batch's row loads and
        // fragment DMA before computing this one, so both have a whole
        // batch of MFMAs to land.  Batches past the split's last stage
        // (the second of a pair at the split's end) store and compute
        // nothing.
        auto batch = [&](auto par) __attribute__((always_inline)) {
            constexpr int NP = 1 - decltype(par)::value;
            if (loader) {
                [&]<int... I>(std::integer_sequence<int, I...>) __attribute__((always_inline)) {
                    ((i0 + I < nst ? (store(I, I), 0) : 0), ...);
                }(std::make_integer_sequence<int, D>{});
            }
            dma_lds_barrier();   // this batch's B fragments landed and its tiles are stored
            issue_batch(i0 + D);
            fetch_batch(std::integral_constant<int, NP>{}, i0 + D);
            [&]<int... I>(std::integer_sequence<int, I...>) __attribute__((always_inline)) {
                ((i0 + I < nst ? (compute(acc, nwc, par, I), 0) : 0), ...);
            }(std::make_integer_sequence<int, D>{});
            lds_barrier_i8();   // every wave is done with the tiles before the next batch stores
            i0 += D;
        };
        // Outer loop over scale blocks, inner loop over pairs of batches of
        // the block (even batch indices use btile0, odd ones btile1; a block
        // starts at an even index): the flush sits after the inner loop on
        // every path, so the accumulators carried by the inner loop are only
        // ever MFMA results and stay in accumulation registers.
        bool first = true;
        do {   // nst > 0
            const uint64_t blk = (st0 + (uint64_t)i0 * kI8Stage) / kFlush;
            const uint64_t blk_end = ((blk + 1) * kFlush - st0) / kI8Stage;   // a multiple of 2 D
            const uint32_t i_end = blk_end < nst ? (uint32_t)blk_end : nst;
            do {   // i0 < i_end on entry: the block holds stage i0
                batch(P0{});
                batch(P1{});
            } while (i0 < i_end);
            flush(acc, blk, first, false);
            first = false;
        } while (i0 < nst);
    };
    // One copy of the loop per group layout on the common path (dword row
    // loads, identity rows): nw is fixed per workgroup, so the choice is made
    // once, and no accumulator value is merged across the copies.
    if constexpr (VEC && IDENT) {
        [&]<int... N>(std::integer_sequence<int, N...>) __attribute__((always_inline)) {
            ((nw == (uint32_t)N ? (main_loop(std::integral_constant<int, N>{}), 0) : 0), ...);
        }(std::make_integer_sequence<int, TMAX + 1>{});
    } else {
        main_loop(std::integral_constant<int, -1>{});
    }
    // the last, unused batch's DMA must land before the workgroup's LDS is released
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
}

}  // namespace gnm
