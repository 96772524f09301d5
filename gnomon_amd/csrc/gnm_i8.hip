// Split-int8 scoring kernel: the fused decode + contraction of
// gnm_kernels.hip with weights written as signed 8-bit integer limbs and
// v_mfma_i32_16x16x64_i8 (64 variants per MFMA, exact int32 accumulation).
//
// Per (scale block of kI8FlushRows = 65536 program rows, weight column) the
// limb unit is a power of two 2^e with max|w| <= 127 * 256^(L-1) * 2^e;
// w / 2^e is rounded to an integer Q and written in L signed base-256
// digits, every digit in [-128, 127].  Dosages are 0/1/2, so each product is
// an exact small integer and an int32 block sum is exact (65536 rows * 2 *
// 128 = 2^24 < 2^31); blocks are flushed to FP64 with their scale.  L = 3
// represents 23 bits, L = 4 31 bits of the block maximum.
//
// Data path, per workgroup of WAVES waves (64 people per wave, 16*WAVES
// packed bytes of every row: 8 waves for 3-tile groups, so a workgroup reads
// whole 128-byte lines of each row, else 4) and per stage of 64 program rows
// (one k-step):
//   * each thread loads 4 rows x 4 packed bytes (one dword per row;
//     consecutive threads read consecutive dwords of a row), transposes the
//     4x4 bytes in registers and writes 4 dwords into a column-major LDS tile
//     [byte column][64 rows] (column pitch 80 bytes, row blocks swizzled per
//     column so writes and 128-bit reads avoid bank conflicts);
//   * one extra tile column holds a per-row selector byte: 4 for rows that
//     read the A1 allele (flipped), 0 otherwise, and 0 for rows outside the
//     split (whose data bytes are zero);
//   * lane (i, g) of wave w reads column 16w + i, rows 16g..16g+15 as one
//     128-bit LDS read: the packed codes of people 4(16w+i)+t, t = 0..3, for
//     16 variants.  Person t's codes are isolated per byte and OR-ed with the
//     row selector byte; one v_perm_b32 then maps (flip, code) to the i8
//     dosage, another to the missing indicator:
//        code 00 01 10 11  ->  dose 0 0 1 2 (A2 row) or 2 0 1 0 (A1 row),
//                               missing 0 1 0 0;
//   * M-tile t of the wave holds people 64w + 4i + t; B fragments (16 i8 per
//     lane, the MFMA B operand) go from memory to LDS by buffer LDS DMA,
//     without passing through registers.
// Loads of a batch of D stages are issued together (buffer loads: one
// uniform descriptor per stage and a 32-bit lane offset, so no per-lane
// 64-bit addresses stay live), stored, then consumed behind LDS-only
// barriers.  The scale-block flush sits after the batch loop, so the
// accumulators carried by that loop stay in accumulation registers.
#include "gnm_internal.hpp"

#include <utility>

namespace gnm {

typedef int i32x4_t __attribute__((ext_vector_type(4)));
typedef uint32_t u32x4_t __attribute__((ext_vector_type(4)));

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

}  // namespace

template <int TMAX, int WAVES, int D, bool VEC, bool IDENT, int SETS>
// The 3-tile, 2-stage instance is asked to fit four waves per SIMD (128
// registers: a few values move to accumulation registers, none to memory).
// SETS > 1: the waves form SETS sets over the same people, each set with its
// own group of TMAX tiles, so one row read serves SETS groups (the first
// set loads and stores the rows for all).
__global__ __launch_bounds__(64 * WAVES) __attribute__((amdgpu_waves_per_eu(TMAX == 3 && D == 2 ? 4 : 1))) void k_score_i8(DevicePlan plan, const uint8_t *__restrict__ rows,
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
    constexpr int kBT = SETS * TMAX;                 // B tiles per stage (all sets)
    constexpr int kBW = (kBT + WAVES - 1) / WAVES;   // B tiles each wave fetches per stage
    static_assert(kI8FlushRows % (D * kI8Stage) == 0, "scale blocks must hold whole batches");
    // one LDS tile per stage of a batch: all D stages are stored, then one
    // barrier, then all D are consumed (two barriers per batch of D stages)
    __shared__ uint32_t tile[D][kTileWords];
    // B fragments are double-buffered: the next batch's DMA lands in one
    // half while this batch's MFMAs read the other
    __shared__ u32x4_t btile[2][D][kBT * 64];

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
    // batch; per-stage tests use these 32-bit indices (scalar compares, where
    // 64-bit row compares take vector instructions):
    //   nst      stages of the split
    //   fast_lo, fast_hi   stages wholly inside [r_begin, r_end) (fast path)
    //   last_i   the stage holding the split's last row
    const uint64_t st0 = (r_begin / (D * kI8Stage)) * (D * kI8Stage);
    const bool nonempty = r_begin < r_end;
    const uint32_t nst = nonempty ? (uint32_t)((r_end - st0 + kI8Stage - 1) / kI8Stage) : 0u;
    const uint32_t fast_lo = (uint32_t)((r_begin - st0 + kI8Stage - 1) / kI8Stage);
    const uint32_t fast_hi = (IDENT && span_full && nonempty) ? (uint32_t)((r_end - st0) / kI8Stage) : 0u;
    const uint32_t last_i = nonempty ? (uint32_t)((last_st - st0) / kI8Stage) : 0u;
    // B fragments: wave w fetches tiles w, w + WAVES, ... of the sets' kBT
    // tiles by LDS DMA (lane l's 16 bytes land at btile[d][64 j + l]); tiles
    // past tile_end are fetched clamped to the last tile and their products
    // are never flushed
    uint32_t boff[kBW];     // byte offsets within a stage's fragment block
#pragma unroll
    for (int m = 0; m < kBW; ++m) {
        const uint32_t j = gtile0 + (uint32_t)(wave + m * WAVES);
        const uint32_t jc = j < tile_end ? j : tile_end - 1;
        boff[m] = (jc * 64 + lane) * 16;
    }
    // fragment block of stage index si (clamped to last_i by the caller)
    const u32x4_t *__restrict__ bfrag_st0 = bfrag + (st0 / kI8Stage) * tiles_total * 64;
    const uint32_t bstep = tiles_total * 64;
    auto fetch_b = [&](uint32_t par, int d, uint32_t si) __attribute__((always_inline)) {
        const __amdgpu_buffer_rsrc_t rb = __builtin_amdgcn_make_buffer_rsrc(
            (void *)(bfrag_st0 + (uint64_t)si * bstep), 0, 0x7FFFFFFF, kI8RsrcWord3);
#pragma unroll
        for (int m = 0; m < kBW; ++m) {
            const int j = wave + m * WAVES;
            if (kBW * WAVES == kBT || j < kBT)
                lds_dma_b128(rb, &btile[par][d][j * 64], boff[m]);
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
    auto issue = [&](uint32_t par, int d, uint32_t si) __attribute__((always_inline)) {
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
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const uint64_t r = st + rg4 + k;
            const bool ok = same && r >= r_begin && r < r_end && lcol < bpv;
            keep |= ok ? (1u << k) : 0u;
            const uint64_t rc = r < r_begin ? r_begin : (r >= r_end ? r_end - 1 : r);
            const uint64_t var = IDENT ? (rc - p0) : (uint64_t)(plan.prog_var[rc] - (uint32_t)v0);
            ra[d][k] = load_row(var, lcol);
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

    typedef i32x4_t Acc[4][TMAX];   // int32 block sums [person tile t][tile j]
    // zero from an opaque copy of 0: the initialisation stays in the branch
    // that uses the set instead of being hoisted above the layout dispatch
    auto zero_acc = [&](Acc &acc) __attribute__((always_inline)) {
        int z = 0;
        asm volatile("" : "+v"(z));
#pragma unroll
        for (int t = 0; t < 4; ++t)
#pragma unroll
            for (int j = 0; j < TMAX; ++j) acc[t][j] = i32x4_t{z, z, z, z};
    };

    // Block flush: acc[t][j][q] is the int32 block sum of person
    // pw0 + 16 lg + 4 q + t (MFMA output row 4 lg + q of M-tile t) and tile
    // column li.  The wave stages one tile through its own LDS buffer so
    // that each lane then owns one column for 16 people, and adds the values
    // to the split partials in FP64 (weight tiles, times the block scale) or
    // as counts.
    // the staging buffer shares the .bed tiles' LDS when they are large
    // enough (no tile is in use while a block is flushed)
    constexpr bool kStageInTile = D * kTileWords >= WAVES * 64 * 16;
    __shared__ int stage_own[kStageInTile ? 1 : WAVES * 64 * 16];
    auto flush = [&](Acc &acc, uint64_t blk, bool first, bool empty) __attribute__((always_inline)) {
        // opaque to the optimiser, so the addresses below are formed here
        // rather than hoisted out of the row loop into live registers
        uint64_t prow = (uint64_t)s * num_people + pw0;
        uint32_t ncol = plan.ncol, nclass = plan.nclass, tb = tile_begin, lane_c = (uint32_t)li;
        asm volatile("" : "+v"(prow), "+v"(ncol), "+v"(nclass), "+v"(tb), "+v"(lane_c));
        ncol = __builtin_amdgcn_readfirstlane(ncol);
        nclass = __builtin_amdgcn_readfirstlane(nclass);
        tb = __builtin_amdgcn_readfirstlane(tb);
        int *sbuf = (kStageInTile ? (int *)&tile[0][0] : stage_own) + wave * (64 * 16);
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
                for (int q = 0; q < 4; ++q) sbuf[(16 * lg + 4 * q + t) * 16 + li] = acc[t][j][q];
                acc[t][j] = i32x4_t{0, 0, 0, 0};
            }
            lds_wave_fence();
            const uint32_t tl = tb + j;
            const bool wtile = tl < wt;
            const uint32_t c = (wtile ? tl : tl - wt) * 16 + lane_c;
            const uint32_t nc = wtile ? ncol : nclass;
            if ((uint32_t)j < ntile && c < nc) {
                const double sc = (wtile && !empty) ? plan.scales[blk * ncol + c] : 0.0;
#pragma unroll 1
                for (int i = 0; i < 16; ++i) {
                    const uint32_t pl = 4 * i + lg;  // person of the wave, 0..63
                    if (pw0 + pl >= num_people) continue;
                    const int v = sbuf[pl * 16 + lane_c];
                    const uint64_t row = prow + pl;
                    if (wtile) {
                        double *dd = part + row * ncol + c;
                        const double x = (double)v * sc;
                        *dd = first ? x : *dd + x;
                    } else {
                        uint32_t *dd = cpart + row * nclass + c;
                        const uint32_t x = (uint32_t)v;
                        *dd = first ? x : *dd + x;
                    }
                }
            }
            __builtin_amdgcn_sched_barrier(0);
        };
        [&]<int... J>(std::integer_sequence<int, J...>) __attribute__((always_inline)) {
            (flush_tile(std::integral_constant<int, J>{}), ...);
        }(std::make_integer_sequence<int, TMAX>{});
        // staging inside the .bed tiles: no wave may store the next batch's
        // tiles while another still reads its staged sums
        if constexpr (kStageInTile) lds_barrier_i8();
    };

    // compute(nwc, par, buf): one stage.  With a compile-time layout
    // (nwc >= 0: the first nwc tiles are weight tiles, the others count
    // tiles) the stage is straight-line code, so the scheduler can overlap
    // one stage's decode with the previous stage's MFMAs; tile slots past
    // ntile are computed on a duplicated fragment and never flushed.  With
    // nwc < 0 the layout is read at run time (per-tile branches).
    auto compute = [&](Acc &acc, auto nwc, uint32_t par, int buf) __attribute__((always_inline)) {
        constexpr int NW = decltype(nwc)::value;
        const uint32_t *t = tile[buf];
        const u32x4_t X = *(const u32x4_t *)&t[tile_dword(pwave * 16 + li, 4 * lg)];
        const u32x4_t F = *(const u32x4_t *)&t[tile_dword(kBytes, 4 * lg)];
        u32x4_t bv[TMAX];
#pragma unroll
        for (int j = 0; j < TMAX; ++j) bv[j] = btile[par][buf][(set * TMAX + j) * 64 + lane];
        // person tile by person tile: only one tile's selectors and operands
        // are live at a time
#pragma unroll
        for (int tt = 0; tt < 4; ++tt) {
            uint32_t sel[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) sel[q] = ((X[q] >> (2 * tt)) & 0x03030303u) | F[q];
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
    auto issue_batch = [&](uint32_t par, uint32_t i) __attribute__((always_inline)) {
        // stages past the split are loaded clamped and never stored
        if (!loader) return;
        [&]<int... I>(std::integer_sequence<int, I...>) __attribute__((always_inline)) {
            (issue(par, I, i + I), ...);
        }(std::make_integer_sequence<int, D>{});
    };
    auto fetch_batch = [&](uint32_t par, uint32_t i) __attribute__((always_inline)) {
        [&]<int... I>(std::integer_sequence<int, I...>) __attribute__((always_inline)) {
            (fetch_b(par, I, i + I < last_i ? i + I : last_i), ...);
        }(std::make_integer_sequence<int, D>{});
    };
    auto main_loop = [&](auto nwc) __attribute__((always_inline)) {
        Acc acc;   // one accumulator set per copy of the loop
        zero_acc(acc);
        uint32_t par = 0;
        uint32_t i0 = 0;
        issue_batch(par, i0);
        fetch_batch(par, i0);
        // Outer loop over scale blocks, inner loop over batches of the block:
        // the flush sits after the inner loop on every path, so the
        // accumulators carried by the inner loop are only ever MFMA results and
        // stay in accumulation registers.
        bool first = true;
        do {   // nst > 0
            const uint64_t blk = (st0 + (uint64_t)i0 * kI8Stage) / kI8FlushRows;
            const uint64_t blk_end = ((blk + 1) * kI8FlushRows - st0) / kI8Stage;   // a multiple of D
            const uint32_t i_end = blk_end < nst ? (uint32_t)blk_end : nst;
            do {   // i0 < i_end on entry: the block holds stage i0
                if (loader) {
                    [&]<int... I>(std::integer_sequence<int, I...>) __attribute__((always_inline)) {
                        ((i0 + I < nst ? (store(I, I), 0) : 0), ...);
                    }(std::make_integer_sequence<int, D>{});
                }
                dma_lds_barrier();   // this batch's B fragments landed and its tiles are stored
                issue_batch(par ^ 1u, i0 + D);
                [&]<int... I>(std::integer_sequence<int, I...>) __attribute__((always_inline)) {
                    ((i0 + I < nst ? (compute(acc, nwc, par, I), 0) : 0), ...);
                }(std::make_integer_sequence<int, D>{});
                // after this wave's last read of the batch's B fragments: the
                // DMA into the other half never overlaps an LDS read in flight
                fetch_batch(par ^ 1u, i0 + D);
                lds_barrier_i8();   // every wave is done with the tiles before the next batch stores
                par ^= 1u;
                i0 += D;
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
    hipLaunchKernelGGL((k_score_i8<TM, WV, DD, VC, ID, SE>), dim3((unsigned)wgs), dim3(64 * WV), 0, stream,     \
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

// One workgroup = 4 waves x MT 16-variant tiles.  Per super-step the
// workgroup's 256 threads load the B fragments of 4 k-steps (16 B x TILES
// each) and store them to LDS (double-buffered, one barrier per
// super-step); each lane loads its MT rows' 16 bytes.  Both are loaded one
// super-step ahead into registers.
template <int TILES, int MT>
__global__ __launch_bounds__(256) void k_adjoint_i8(const uint8_t *__restrict__ rows, uint64_t stride, uint64_t V,
                                                    uint64_t N, const uint8_t *__restrict__ flip,
                                                    const u32x4_t *__restrict__ frag,
                                                    const double *__restrict__ scales, uint64_t Kp, uint32_t L,
                                                    double *__restrict__ ws) {
    constexpr int kBPer = TILES;                  // B entries (16 B) per thread per super-step
    __shared__ u32x4_t bt[2][4 * TILES * 64];     // [buf][(k-step d, tile j, lane)]
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int li = lane & 15, lg = lane >> 4;
    const uint64_t wv0 = ((uint64_t)blockIdx.x * 4 + wave) * (16 * MT);   // first variant of the wave
    const uint64_t bpv = (N + 3) / 4;
    const uint64_t nss = (bpv + 63) / 64;                                // super-steps
    const uint32_t tpl = (uint32_t)(Kp / 16);                            // tiles per limb
    // rows of the wave: one descriptor, lane offsets (16mt + i) * stride + 16g;
    // rows past V read as zero (range check), bytes past the row end belong
    // to people whose R limbs are zero
    const uint64_t rem = wv0 < V ? (V - wv0) * stride : 0;
    const __amdgpu_buffer_rsrc_t rr = __builtin_amdgcn_make_buffer_rsrc(
        (void *)(rows + (wv0 < V ? wv0 : 0) * stride), 0, (int)(rem < 0x7FFFFFFFull ? rem : 0x7FFFFFFFull),
        kI8RsrcWord3);
    uint32_t voff[MT], F[MT];
#pragma unroll
    for (int mt = 0; mt < MT; ++mt) {
        const uint64_t v = wv0 + 16 * mt + li;
        voff[mt] = (uint32_t)((16 * mt + li) * stride + 16 * lg);
        F[mt] = (v < V && flip && flip[v]) ? 0x04040404u : 0u;
    }
    const __amdgpu_buffer_rsrc_t rb = __builtin_amdgcn_make_buffer_rsrc(
        (void *)frag, 0, 0x7FFFFFFF, kI8RsrcWord3);
    auto load_ss = [&](uint64_t ss, u32x4_t (&x)[MT], u32x4_t (&b)[kBPer]) __attribute__((always_inline)) {
        const uint32_t so = (uint32_t)(64 * ss);
#pragma unroll
        for (int mt = 0; mt < MT; ++mt)
            x[mt] = __builtin_bit_cast(u32x4_t, __builtin_amdgcn_raw_buffer_load_b128(rr, voff[mt], so, 0));
        const uint64_t e0 = ss * (4 * TILES * 64);   // first B entry of the super-step
#pragma unroll
        for (int m = 0; m < kBPer; ++m)
            b[m] = __builtin_bit_cast(u32x4_t, __builtin_amdgcn_raw_buffer_load_b128(
                                                   rb, (uint32_t)((e0 + tid + 256 * m) * 16), 0, 0));
    };
    i32x4_t acc[MT][TILES];
    auto zero = [&]() __attribute__((always_inline)) {
        int z = 0;
        asm volatile("" : "+v"(z));
#pragma unroll
        for (int mt = 0; mt < MT; ++mt)
#pragma unroll
            for (int j = 0; j < TILES; ++j) acc[mt][j] = i32x4_t{z, z, z, z};
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
                    for (int j = 0; j < TILES; ++j) sum += (double)acc[mt][j][q] * sc[j];
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

    u32x4_t xa[MT], ba[kBPer];
    load_ss(0, xa, ba);
    // outer loop over person blocks, inner loop over the block's
    // super-steps: the flush sits after the inner loop on every path, so the
    // accumulators stay in accumulation registers
    constexpr uint64_t kSSPerBlk = kI8FlushRows / 256;
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
            u32x4_t xc[MT];
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
                    i32x4_t A;
#pragma unroll
                    for (int t = 0; t < 4; ++t) A[t] = (int)dose_i8(((xc[mt][d] >> (2 * t)) & 0x03030303u) | F[mt]);
#pragma unroll
                    for (int j = 0; j < TILES; ++j) {
                        const i32x4_t B = i32x4_t{(int)bv[j][0], (int)bv[j][1], (int)bv[j][2], (int)bv[j][3]};
                        acc[mt][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(A, B, acc[mt][j], 0, 0, 0);
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

hipError_t launch_adjoint_i8(const uint8_t *rows, uint64_t stride, uint64_t V, uint64_t N, const uint8_t *flip,
                             const uint32_t *frag, const double *scales, uint32_t tiles, uint64_t K, uint64_t Kp,
                             uint32_t L, double *ws, float *dW, hipStream_t stream) {
    if (V == 0) return hipSuccess;
    if ((reinterpret_cast<uintptr_t>(rows) | stride) & 3u) return hipErrorInvalidValue;
    if (stride >= (1ull << 25)) return hipErrorInvalidValue;
#define GNM_ADJ(T, M)                                                                                              \
    hipLaunchKernelGGL((k_adjoint_i8<T, M>), dim3((unsigned)((V + 64 * M - 1) / (64 * M))), dim3(256), 0, stream, \
                       rows, stride, V, N, flip, (const u32x4_t *)frag, scales, Kp, L, ws)
    if (tiles == 4) GNM_ADJ(4, 4);
    else if (tiles == 8) GNM_ADJ(8, 2);
    else if (tiles == 12) GNM_ADJ(12, 1);
    else if (tiles == 16) GNM_ADJ(16, 1);
    else return hipErrorInvalidValue;
#undef GNM_ADJ
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const uint64_t n = V * K;
    hipLaunchKernelGGL(k_adjoint_out, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, ws, V, K, Kp, dW);
    return hipGetLastError();
}

// per-row selector bytes (4 = flipped row) from a flip byte array
__global__ void k_flip_bytes(const uint8_t *__restrict__ flip, uint64_t V, uint64_t Rpad, uint8_t *__restrict__ out) {
    const uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r < Rpad) out[r] = (flip && r < V && flip[r]) ? 4u : 0u;
}

hipError_t launch_flip_bytes(const uint8_t *flip, uint64_t V, uint64_t Rpad, uint8_t *out, hipStream_t stream) {
    if (Rpad == 0) return hipSuccess;
    hipLaunchKernelGGL(k_flip_bytes, dim3((unsigned)((Rpad + 255) / 256)), dim3(256), 0, stream, flip, V, Rpad, out);
    return hipGetLastError();
}

}  // namespace gnm
