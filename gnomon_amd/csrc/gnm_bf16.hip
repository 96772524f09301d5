// Split-BF16 path: the launcher of the fused decode + contraction kernel
// (k_score_mx in gnm_score_mx.hpp with BF16 limbs: v_mfma_f32_16x16x32_bf16,
// two k-steps per 64-row stage, exact FP32 block sums over scale blocks of
// kBfFlushRows = 32768 program rows, FP64 across blocks).  L = 4 limbs
// represent 32 bits of each block maximum, L = 3 24 bits, L = 2 16 bits.
#include "gnm_score_mx.hpp"

namespace gnm {

// Rough cost of one group pass per wave and stage, in cycles: decode VALU
// (about 150 instructions of 4 cycles: code flip, row-pair duplication,
// selectors, dosage and missing lookups, shared by the group's tiles) plus
// 8 MFMA of 16 cycles per tile slot (two k-steps x four person tiles).
static uint64_t bf16_pass_cost(uint32_t ntiles, uint32_t tmax) {
    const uint64_t groups = (ntiles + tmax - 1) / tmax;
    return groups * (600u + 128u * tmax);
}

// default tiles per pass: the cheapest of 3, 5, 6 and 8 by the cost model.
// 9-tile groups stay an explicit choice only: their 144 accumulators and
// the decode do not fit 256 registers (over a thousand spilled values), and
// at 100k x 1M x 64 in bf16x4 two 9-tile groups took 145 ms against 40 ms
// for two 8-tile groups (6-tile groups 51, 5-tile 59, 3-tile 61 ms).
static uint32_t bf16_default_tiles_per_pass(uint32_t ntiles) {
    uint32_t tmax = 3;
    for (uint32_t c : {5u, 6u, 8u})
        if (bf16_pass_cost(ntiles, c) < bf16_pass_cost(ntiles, tmax)) tmax = c;
    return tmax;
}

uint64_t bf16_default_cost(uint32_t ntiles) {
    return bf16_pass_cost(ntiles, bf16_default_tiles_per_pass(ntiles));
}

// 3- and 9-tile groups run 8 waves (512 people, whole 128-byte lines of
// each row; the 9-tile group's 18 KB of B fragments per stage,
// double-buffered, leave room for one workgroup per CU, so 8 waves keep two
// per SIMD); 5-, 6- and 8-tile groups run 4 waves, two workgroups per CU
// (the 8-tile group's flush staged in its idle B buffer: 74 KB of LDS).  At
// 100k x 1M x 64 in bf16x4 the 4-wave 8-tile group took 39.4 ms against
// 41.0 ms with 8 waves.
static uint32_t bf16_waves(uint32_t tmax) { return (tmax == 3 || tmax == 9) ? 8u : 4u; }

// Variant splits: about 12k workgroups over all tile groups, a multiple of 8
// (whole splits per XCD), 8..32 (the FP64 partials grow with the count).
uint32_t bf16_splits(uint64_t num_people, uint32_t ntiles) {
    const uint32_t nt = ntiles ? ntiles : 1;
    const uint32_t tmax = bf16_default_tiles_per_pass(nt);
    const uint64_t groups = (nt + tmax - 1) / tmax;
    const uint64_t ppw = 64ull * bf16_waves(tmax);
    const uint64_t per = ((num_people + ppw - 1) / ppw) * groups;
    uint64_t s = (12288 + per - 1) / per;
    s = (s + 7) / 8 * 8;
    return (uint32_t)(s < 8 ? 8 : (s > 32 ? 32 : s));
}

hipError_t launch_score_bf16(const DevicePlan &plan, const uint8_t *rows, uint64_t stride, uint64_t bpv, uint64_t v0,
                             uint64_t p0, uint64_t p1, uint64_t num_people, uint32_t splits, uint32_t tile_begin,
                             uint32_t tile_end, bool identity, double *part, uint32_t *cpart, hipStream_t stream,
                             uint32_t tiles_per_pass) {
    if (tile_end <= tile_begin) return hipSuccess;
    // dword row loads need 4-byte alignment; the fast path's buffer offsets
    // (up to 64 rows x stride) must stay below the descriptors' 2^31 - 1 range
    const bool vec = ((reinterpret_cast<uintptr_t>(rows) | stride) & 3u) == 0 && bpv >= 4 && stride < (1ull << 25);
    const uint64_t split_base = (p0 / kI8Stage) * kI8Stage;
    uint64_t rows_per_split = (p1 - split_base + splits - 1) / splits;
    rows_per_split = (rows_per_split + kI8Stage - 1) / kI8Stage * kI8Stage;
    const uint32_t ntiles = tile_end - tile_begin;
    const uint32_t tp = tiles_per_pass ? tiles_per_pass : bf16_default_tiles_per_pass(ntiles);
    if (tp != 3 && tp != 5 && tp != 6 && tp != 8 && tp != 9) return hipErrorInvalidValue;
    const uint32_t ngroups = (ntiles + tp - 1) / tp;
    const uint32_t wv = bf16_waves(tp);
    const uint32_t people_per_wg = 64u * wv;
    const uint32_t pt = (uint32_t)((num_people + people_per_wg - 1) / people_per_wg);
    const uint64_t wgs = (uint64_t)pt * splits * ngroups;   /* splits is a multiple of 8 */
#define GNM_LAUNCH_BF(TM, WV, VC, ID)                                                                        \
    hipLaunchKernelGGL((k_score_mx<TM, WV, 2, VC, ID, 1, true>), dim3((unsigned)wgs), dim3(64 * WV), 0, stream, \
                       plan, rows, stride, bpv, v0, p0, p1, split_base, rows_per_split, splits, pt, num_people, \
                       tile_begin, tile_end, ngroups, part, cpart)
#define GNM_LAUNCH_BF_VI(TM, WV)                                                                             \
    do {                                                                                                     \
        if (vec) {                                                                                           \
            if (identity) GNM_LAUNCH_BF(TM, WV, true, true);                                                 \
            else GNM_LAUNCH_BF(TM, WV, true, false);                                                         \
        } else {                                                                                             \
            if (identity) GNM_LAUNCH_BF(TM, WV, false, true);                                                \
            else GNM_LAUNCH_BF(TM, WV, false, false);                                                        \
        }                                                                                                    \
    } while (0)
    if (tp == 3) GNM_LAUNCH_BF_VI(3, 8);
    else if (tp == 5) GNM_LAUNCH_BF_VI(5, 4);
    else if (tp == 6) GNM_LAUNCH_BF_VI(6, 4);
    else if (tp == 8) GNM_LAUNCH_BF_VI(8, 4);
    else GNM_LAUNCH_BF_VI(9, 8);
#undef GNM_LAUNCH_BF_VI
#undef GNM_LAUNCH_BF
    return hipGetLastError();
}

}  // namespace gnm
