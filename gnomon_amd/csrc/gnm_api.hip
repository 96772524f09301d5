// Host side of the C ABI declared in include/gnomon_hip.h.
//
// gnm_create turns the reference's CSR weight matrix (PreparationResult,
// score/types.rs:166-365) into a device plan:
//   * effect-allele rows: a variant whose CSR entries are all unflipped
//     (corr == 0) becomes one A2-dosage row with weights W_eff; one whose
//     entries are all flipped becomes one A1-dosage row with weights corr/2
//     (the original effect weight, score/prepare.rs:197-208); a variant with
//     both kinds gets both rows.  In effect-allele space a missing call
//     contributes 0, which equals baseline + W_eff*d - corr-on-missing.
//   * membership classes: scores whose CSR columns hold the same variant set
//     share one missing-count accumulator.
//   * FP32 weights (plain path) or split-BF16 integer limbs with per
//     (scale block, column) power-of-two scales, pre-arranged in the lane
//     order of v_mfma_f32_16x16x32_bf16 (MFMA path).
// Rows then stream through gnm_score_rows (pinned staging, asynchronous
// copies) or arrive device-resident; partial sums are reduced in FP64 into
// persistent device accumulators.
#include "../../include/gnomon_hip.h"
#include "gnm_internal.hpp"
#include "gnm_dense.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <map>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <tuple>
#include <vector>

using namespace gnm;

namespace {

thread_local std::string g_last_error;

gnm_status fail(gnm_status st, const std::string &msg) {
    g_last_error = msg;
    return st;
}

#define GNM_HIP(call)                                                                  \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess) {                                                        \
            return fail(e_ == hipErrorOutOfMemory ? GNM_ERR_OUT_OF_MEMORY : GNM_ERR_HIP, \
                        std::string(#call) + ": " + hipGetErrorString(e_));            \
        }                                                                              \
    } while (0)

uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

bool env_hides_devices() {
    for (const char *name : {"HIP_VISIBLE_DEVICES", "ROCR_VISIBLE_DEVICES", "CUDA_VISIBLE_DEVICES"}) {
        const char *v = std::getenv(name);
        if (!v) continue;
        std::string s(v);
        while (!s.empty() && (s.back() == ' ' || s.back() == '\t')) s.pop_back();
        while (!s.empty() && (s.front() == ' ' || s.front() == '\t')) s.erase(s.begin());
        std::string low = s;
        for (auto &ch : low) ch = (char)std::tolower((unsigned char)ch);
        if (s.empty() || s == "-1" || low == "none") return true;
    }
    return false;
}

struct DevBuf {
    void *ptr = nullptr;
    size_t bytes = 0;
};

}  // namespace

struct gnm_ctx {
    // problem
    uint64_t N = 0, Nfam = 0, bpv = 0, K = 0, V = 0;
    int precision = GNM_PRECISION_BF16X3;
    int device = 0;
    bool identity = true;
    // host copies needed after create
    std::vector<uint64_t> var_prog_off;   // [V+1]
    std::vector<uint64_t> non_single;     // [V+1] variants before v with other than one program row
    std::vector<uint64_t> row_offsets;    // [V+1] CSR
    std::vector<uint32_t> cols;           // [nnz]
    std::vector<float> corr;              // [nnz]
    std::vector<uint32_t> score_class;    // [K]
    uint64_t R = 0;                       // program rows
    uint32_t nclass = 0;
    uint32_t limbs = 1;
    bool fp32_count_in_kernel = true;
    // device
    DevicePlan plan{};
    std::vector<DevBuf> allocations;
    uint32_t *d_score_class = nullptr;
    uint32_t *d_out_to_fam = nullptr;
    double *d_scores = nullptr;
    uint32_t *d_counts = nullptr;
    double *d_part = nullptr;
    uint32_t *d_cpart = nullptr;
    uint32_t splits = 8;
    uint8_t *d_compact = nullptr;
    uint64_t compact_rows = 0;     // capacity in rows
    uint64_t cbpv = 0;             // compact bytes per variant
    uint8_t *d_stage[2] = {nullptr, nullptr};
    uint8_t *h_stage[2] = {nullptr, nullptr};
    hipEvent_t stage_free[2] = {nullptr, nullptr};    // compute finished with the slot
    hipEvent_t stage_copied[2] = {nullptr, nullptr};  // the slot's rows are on the device
    hipStream_t copy_stream = nullptr;                // host-to-device copies of the staging slots
    bool staging_ready = false;    // slots, events and copy stream made
    uint64_t stage_rows = 0;       // capacity of each staging slot in rows
    uint64_t stage_pitch = 0;      // staging row pitch (bytes_per_variant rounded up to 64)
    hipStream_t stream = nullptr;
    hipEvent_t ev_main0 = nullptr, ev_main1 = nullptr, ev_all0 = nullptr, ev_all1 = nullptr;
    bool timing_valid = false;
    uint64_t completed = 0;
    uint32_t tiles_per_pass = 0;   // int8 kernel launch shape (0 = default), gnm_set_launch_config
    uint32_t batch_stages = 0;
    // sparse ingest of host rows (gnm_set_sparse_ingest)
    double sparse_fraction = 0.5;         // a row with at most this fraction of non-zero words goes sparse
    bool sparse_ready = false;            // upload buffers allocated
    uint32_t sparse_nch = 0;              // 512-word chunks per staged row (pitch / 4096, rounded up)
    std::vector<uint64_t> keep2;          // [ceil(Nfam/32)] 2-bit masks of the scored people (subsets only)
    uint8_t *d_pack[2] = {nullptr, nullptr};      // uploaded dense rows + entries
    uint32_t *h_map[2] = {nullptr, nullptr}, *d_map[2] = {nullptr, nullptr};     // [stage_rows]
    uint32_t *h_svrow[2] = {nullptr, nullptr}, *d_svrow[2] = {nullptr, nullptr}; // [stage_rows] sparse rows
    uint64_t *h_bmoff[2] = {nullptr, nullptr}, *d_bmoff[2] = {nullptr, nullptr}; // [stage_rows] their records
    uint32_t *h_cs[2] = {nullptr, nullptr}, *d_cs[2] = {nullptr, nullptr};       // [stage_rows][nch+1]
    uint64_t st_dense = 0, st_sparse = 0, st_words = 0, st_bytes = 0;            // gnm_ingest_stats
    uint32_t sparse_rest = 0;             // slots left to send packed without classifying

    template <class T>
    gnm_status alloc(T **out, size_t count) {
        void *p = nullptr;
        const size_t bytes = std::max<size_t>(count * sizeof(T), 256);
        hipError_t e = hipMalloc(&p, bytes);
        if (e != hipSuccess)
            return fail(e == hipErrorOutOfMemory ? GNM_ERR_OUT_OF_MEMORY : GNM_ERR_HIP,
                        "hipMalloc(" + std::to_string(bytes) + "): " + hipGetErrorString(e));
        allocations.push_back({p, bytes});
        *out = (T *)p;
        return GNM_OK;
    }
    template <class T>
    gnm_status upload(T **out, const std::vector<T> &host) {
        gnm_status st = alloc(out, host.size());
        if (st != GNM_OK) return st;
        if (!host.empty()) GNM_HIP(hipMemcpy(*out, host.data(), host.size() * sizeof(T), hipMemcpyHostToDevice));
        return GNM_OK;
    }
};

namespace {

gnm_status validate(const gnm_problem *p) {
    if (!p) return fail(GNM_ERR_INVALID_ARG, "problem is NULL");
    if (p->num_people == 0 || p->num_scores == 0)
        return fail(GNM_ERR_INVALID_ARG, "num_people and num_scores must be positive");
    if (p->total_people_in_fam < p->num_people)
        return fail(GNM_ERR_INVALID_ARG, "total_people_in_fam < num_people");
    if (p->bytes_per_variant != (p->total_people_in_fam + 3) / 4)
        return fail(GNM_ERR_INVALID_ARG, "bytes_per_variant must equal ceil(total_people_in_fam/4)");
    if (p->num_variants > 0xFFFFFFF0ull) return fail(GNM_ERR_INVALID_ARG, "too many variants");
    if (!p->sparse_row_offsets) return fail(GNM_ERR_INVALID_ARG, "sparse_row_offsets is NULL");
    if (p->sparse_row_offsets[0] != 0 || p->sparse_row_offsets[p->num_variants] != p->nnz)
        return fail(GNM_ERR_INVALID_ARG, "sparse_row_offsets must start at 0 and end at nnz");
    if (p->nnz && (!p->sparse_weights || !p->sparse_missing_corrections || !p->sparse_score_columns))
        return fail(GNM_ERR_INVALID_ARG, "CSR arrays are NULL");
    for (uint64_t v = 0; v < p->num_variants; ++v)
        if (p->sparse_row_offsets[v + 1] < p->sparse_row_offsets[v])
            return fail(GNM_ERR_INVALID_ARG, "sparse_row_offsets must be non-decreasing");
    for (uint64_t e = 0; e < p->nnz; ++e) {
        if (p->sparse_score_columns[e] >= p->num_scores)
            return fail(GNM_ERR_INVALID_ARG, "CSR score column out of range");
        // non-finite weights are refused, as the reference refuses its split
        // path for them (score/cuda_backend.rs:1178-1181)
        if (!std::isfinite(p->sparse_weights[e]) || !std::isfinite(p->sparse_missing_corrections[e]))
            return fail(GNM_ERR_INVALID_ARG, "CSR weight or missing correction is not finite (entry " +
                                                 std::to_string(e) + ")");
    }
    if (!p->output_idx_to_fam_idx && p->num_people != p->total_people_in_fam)
        return fail(GNM_ERR_INVALID_ARG, "output_idx_to_fam_idx is required when num_people < total_people_in_fam");
    if (p->output_idx_to_fam_idx)
        for (uint64_t i = 0; i < p->num_people; ++i)
            if (p->output_idx_to_fam_idx[i] >= p->total_people_in_fam)
                return fail(GNM_ERR_INVALID_ARG, "output_idx_to_fam_idx entry out of range");
    if (p->precision < GNM_PRECISION_FP32 || p->precision > GNM_PRECISION_BF16X4)
        return fail(GNM_ERR_INVALID_ARG, "unknown precision");
    for (int i = 0; i < 4; ++i)
        if (p->reserved[i]) return fail(GNM_ERR_INVALID_ARG, "reserved fields must be zero");
    return GNM_OK;
}

inline bool is_int8(int precision) {
    return precision == GNM_PRECISION_I8X3 || precision == GNM_PRECISION_I8X4;
}

inline uint32_t limbs_of(int precision) {
    switch (precision) {
        case GNM_PRECISION_BF16X2: return 2;
        case GNM_PRECISION_BF16X3: return 3;
        case GNM_PRECISION_BF16X4: return 4;
        case GNM_PRECISION_I8X3: return 3;
        case GNM_PRECISION_I8X4: return 4;
        default: return 1;
    }
}

// largest |Q| the limb digits represent: BF16 limbs allow a top digit up to
// 255, int8 limbs keep every digit in [-128, 127]
inline double limb_qmax(uint32_t L, bool int8) {
    return int8 ? 127.0 * std::ldexp(1.0, 8 * (int)(L - 1)) : std::ldexp(1.0, 8 * (int)L) - std::ldexp(1.0, 8 * (int)(L - 1));
}


template <class F>
void parallel_for(uint64_t n, F &&fn, uint64_t min_items = 64) {
    unsigned T = std::max(1u, std::min(32u, std::thread::hardware_concurrency()));
    if (n < min_items || T == 1) {
        fn(0, n);
        return;
    }
    std::vector<std::thread> th;
    // an exception in a worker is carried to the caller (and from there to
    // the entry point's catch-all) instead of terminating the process
    std::vector<std::exception_ptr> err(T);
    uint64_t chunk = (n + T - 1) / T;
    for (unsigned t = 0; t < T; ++t) {
        uint64_t a = t * chunk, b = std::min(n, a + chunk);
        if (a >= b) break;
        th.emplace_back([&fn, &err, t, a, b] {
            try {
                fn(a, b);
            } catch (...) {
                err[t] = std::current_exception();
            }
        });
    }
    for (auto &x : th) x.join();
    for (auto &e : err)
        if (e) std::rethrow_exception(e);
}

// Device buffers needed only while gnm_create builds the plan.
struct BuildScratch {
    std::vector<void *> ptrs;
    ~BuildScratch() {
        for (void *q : ptrs) (void)hipFree(q);
    }
    template <class T>
    gnm_status get(T **out, size_t count) {
        void *q = nullptr;
        const hipError_t e = hipMalloc(&q, std::max<size_t>(count * sizeof(T), 256));
        if (e != hipSuccess)
            return fail(e == hipErrorOutOfMemory ? GNM_ERR_OUT_OF_MEMORY : GNM_ERR_HIP,
                        std::string("hipMalloc (plan build): ") + hipGetErrorString(e));
        ptrs.push_back(q);
        *out = (T *)q;
        return GNM_OK;
    }
    template <class T>
    gnm_status put(T **out, const T *host, size_t count) {
        gnm_status st = get(out, count);
        if (st == GNM_OK && count) {
            const hipError_t e = hipMemcpy(*out, host, count * sizeof(T), hipMemcpyHostToDevice);
            if (e != hipSuccess) return fail(GNM_ERR_HIP, std::string("hipMemcpy (plan build): ") + hipGetErrorString(e));
        }
        return st;
    }
};

gnm_status build_context(const gnm_problem *p, gnm_ctx *c) {
    c->N = p->num_people;
    c->Nfam = p->total_people_in_fam;
    c->bpv = p->bytes_per_variant;
    c->K = p->num_scores;
    c->V = p->num_variants;
    c->precision = p->precision;   // AUTO is resolved by the caller (create_impl)
    c->device = p->device < 0 ? 0 : p->device;
    c->limbs = limbs_of(c->precision);
    const bool int8 = is_int8(c->precision);
    const uint64_t V = c->V, K = c->K, nnz = p->nnz;

    c->identity = true;
    if (p->output_idx_to_fam_idx) {
        if (c->N != c->Nfam) c->identity = false;
        for (uint64_t i = 0; i < c->N && c->identity; ++i)
            if (p->output_idx_to_fam_idx[i] != i) c->identity = false;
    }

    c->row_offsets.assign(p->sparse_row_offsets, p->sparse_row_offsets + V + 1);
    c->cols.assign(p->sparse_score_columns, p->sparse_score_columns + nnz);
    c->corr.assign(p->sparse_missing_corrections, p->sparse_missing_corrections + nnz);
    if (!c->identity) {   // the sparse ingest sends the calls of scored people only
        c->keep2.assign((c->Nfam + 31) / 32, 0);
        for (uint64_t i = 0; i < c->N; ++i) {
            const uint32_t f = p->output_idx_to_fam_idx[i];
            c->keep2[f / 32] |= 3ull << (2 * (f % 32));
        }
    }

    // ---- membership classes (identical CSR column supports share a count)
    std::vector<uint64_t> h1(K, 0), h2(K, 0), cnt(K, 0);
    for (uint64_t v = 0; v < V; ++v) {
        const uint64_t a = mix64(v ^ 0x5851F42D4C957F2Dull), b = mix64(v + 0x14057B7EF767814Full);
        for (uint64_t e = c->row_offsets[v]; e < c->row_offsets[v + 1]; ++e) {
            const uint32_t k = c->cols[e];
            h1[k] += a;
            h2[k] ^= b + (h1[k] << 7);  // order-dependent mix is fine: variants arrive ascending
            cnt[k] += 1;
        }
    }
    std::map<std::tuple<uint64_t, uint64_t, uint64_t>, uint32_t> cls_of;
    c->score_class.resize(K);
    for (uint64_t k = 0; k < K; ++k) {
        auto key = std::make_tuple(cnt[k], h1[k], h2[k]);
        auto it = cls_of.find(key);
        if (it == cls_of.end()) it = cls_of.emplace(key, (uint32_t)cls_of.size()).first;
        c->score_class[k] = it->second;
    }
    c->nclass = (uint32_t)cls_of.size();

    // ---- effect-allele program rows
    c->var_prog_off.assign(V + 1, 0);
    c->non_single.assign(V + 1, 0);
    std::vector<uint8_t> var_rows(V, 0);  // bit0: A2 row, bit1: A1 row
    for (uint64_t v = 0; v < V; ++v) {
        bool need_d = false, need_e = false;
        for (uint64_t e = c->row_offsets[v]; e < c->row_offsets[v + 1]; ++e) {
            const double we = p->sparse_weights[e], co = p->sparse_missing_corrections[e];
            if (co != 0.0) need_e = true;
            if (we + 0.5 * co != 0.0) need_d = true;
        }
        // a variant without CSR entries contributes nothing: no program row
        const bool empty = c->row_offsets[v] == c->row_offsets[v + 1];
        uint8_t m = empty ? 0 : (need_e ? (need_d ? 3 : 2) : 1);
        var_rows[v] = m;
        c->var_prog_off[v + 1] = c->var_prog_off[v] + (m == 3 ? 2 : (m ? 1 : 0));
        c->non_single[v + 1] = c->non_single[v] + ((m == 1 || m == 2) ? 0 : 1);
    }
    const uint64_t R = c->var_prog_off[V];
    c->R = R;
    // padded to whole 64-row LDS stages of the MFMA kernel
    const uint64_t Rpad = std::max<uint64_t>((R + 63) / 64 * 64, 64);
    std::vector<uint32_t> prog_var(Rpad, 0), flipmask(Rpad / 32, 0), prog_cls(Rpad, 0);
    std::vector<uint8_t> prog_count_row(Rpad, 0);
    for (uint64_t v = 0; v < V; ++v) {
        uint64_t r = c->var_prog_off[v];
        if (var_rows[v] & 1) {
            prog_var[r] = (uint32_t)v;
            prog_count_row[r] = 1;
            ++r;
        }
        if (var_rows[v] & 2) {
            prog_var[r] = (uint32_t)v;
            flipmask[r >> 5] |= 1u << (r & 31);
            if (!(var_rows[v] & 1)) prog_count_row[r] = 1;
        }
    }
    // fill padding rows with the last real variant (never read: masked by range)
    for (uint64_t r = R; r < Rpad; ++r) prog_var[r] = R ? prog_var[R - 1] : 0;
    for (uint64_t v = 0; v < V; ++v) {
        if (!var_rows[v]) continue;
        const uint64_t r = c->var_prog_off[v];
        uint32_t m = 0;
        for (uint64_t e = c->row_offsets[v]; e < c->row_offsets[v + 1]; ++e) {
            const uint32_t cl = c->score_class[c->cols[e]];
            if (cl < 32) m |= 1u << cl;
        }
        prog_cls[r] = m;  // first row of the variant is its count row
    }

    // ---- device
    GNM_HIP(hipSetDevice(c->device));
    GNM_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    gnm_status st;
    uint32_t *d_prog_var, *d_flipmask, *d_prog_cls;
    if ((st = c->upload(&d_prog_var, prog_var)) != GNM_OK) return st;
    if ((st = c->upload(&d_flipmask, flipmask)) != GNM_OK) return st;
    if ((st = c->upload(&d_prog_cls, prog_cls)) != GNM_OK) return st;
    if ((st = c->upload(&c->d_score_class, c->score_class)) != GNM_OK) return st;
    DevicePlan &pl = c->plan;
    pl.num_prog_rows = R;
    pl.prog_var = d_prog_var;
    pl.flipmask = d_flipmask;
    pl.prog_cls = d_prog_cls;
    pl.nclass = c->nclass;
    pl.score_class = c->d_score_class;
    pl.limbs = c->limbs;
    pl.int8 = int8 ? 1u : 0u;
    {
        // selector byte of each program row for the MFMA kernels' decode
        // (bit 7: a count row, when the MFMA kernels count missing calls
        // themselves: one membership class)
        std::vector<uint8_t> flipbyte(Rpad, 0);
        const uint8_t sel = int8 ? kFlipSelI8 : kFlipSelBf;
        const bool vc = c->precision != GNM_PRECISION_FP32 && c->nclass == 1 &&
                        vcount_pays(int8, (uint32_t)((c->limbs * K + 15) / 16));
        for (uint64_t r = 0; r < R; ++r) {
            if ((flipmask[r >> 5] >> (r & 31)) & 1u) flipbyte[r] = sel;
            if (vc && prog_count_row[r]) flipbyte[r] |= 0x80u;
        }
        uint8_t *d_flipbyte;
        if ((st = c->upload(&d_flipbyte, flipbyte)) != GNM_OK) return st;
        pl.flipbyte = d_flipbyte;
    }

    const bool mfma_weights = c->precision != GNM_PRECISION_FP32;
    c->fp32_count_in_kernel = c->nclass <= (uint32_t)kFp32MaxClasses;
    const bool mfma_counts = mfma_weights || !c->fp32_count_in_kernel;
    const uint32_t L = c->limbs;
    pl.ncol = mfma_weights ? (uint32_t)(L * K) : 0;
    pl.wtiles = (pl.ncol + 15) / 16;
    // one membership class: the MFMA kernels count missing calls on the
    // vector ALUs (no count tile) where the cost model says a count tile
    // costs more (it does when it would need a tile group of its own, not
    // when it fills a slot the weight tiles leave free)
    pl.vcount = (mfma_weights && c->nclass == 1 && vcount_pays(int8, pl.wtiles)) ? 1u : 0u;
    pl.ctiles = (mfma_counts && !pl.vcount) ? (c->nclass + 15) / 16 : 0;
    const uint32_t tiles = pl.wtiles + pl.ctiles;

    // ---- the weight plan, built on the device from the CSR (gnm_plan.hip):
    // f64 effect-allele rows, then scales and B fragments or FP32 weights
    {
        BuildScratch sc;
        uint64_t *d_off, *d_prog_off;
        uint32_t *d_cols, *d_rowcls = nullptr;
        float *d_w, *d_co;
        uint8_t *d_var_rows, *d_count_row = nullptr;
        double *d_wrow;
        if ((st = sc.put(&d_off, c->row_offsets.data(), V + 1)) != GNM_OK) return st;
        if ((st = sc.put(&d_cols, c->cols.data(), nnz)) != GNM_OK) return st;
        if ((st = sc.put(&d_w, p->sparse_weights, nnz)) != GNM_OK) return st;
        if ((st = sc.put(&d_co, p->sparse_missing_corrections, nnz)) != GNM_OK) return st;
        if ((st = sc.put(&d_prog_off, c->var_prog_off.data(), V + 1)) != GNM_OK) return st;
        if ((st = sc.put(&d_var_rows, var_rows.data(), V)) != GNM_OK) return st;
        if ((st = sc.get(&d_wrow, Rpad * K)) != GNM_OK) return st;
        GNM_HIP(launch_fill_f64(d_wrow, Rpad * K, 0.0, c->stream));
        const uint32_t cw = (c->nclass + 31) / 32;
        if (pl.ctiles) {
            if ((st = sc.get(&d_rowcls, std::max<uint64_t>(R, 1) * cw)) != GNM_OK) return st;
            GNM_HIP(launch_fill_u32(d_rowcls, std::max<uint64_t>(R, 1) * cw, 0u, c->stream));
            if ((st = sc.put(&d_count_row, prog_count_row.data(), Rpad)) != GNM_OK) return st;
        }
        GNM_HIP(launch_plan_rows(d_off, d_cols, d_w, d_co, d_prog_off, d_var_rows, V, K, c->d_score_class, cw, d_wrow,
                                 d_rowcls, c->stream));
        if (!mfma_weights) {
            const uint64_t k32 = fp32_weight_pitch(K);
            float *d_w32;
            if ((st = c->alloc(&d_w32, Rpad * k32)) != GNM_OK) return st;
            GNM_HIP(launch_fill_u32((uint32_t *)d_w32, Rpad * k32, 0u, c->stream));
            GNM_HIP(launch_plan_w32(d_wrow, R, K, k32, d_w32, c->stream));
            pl.w32 = d_w32;
            pl.k32 = k32;
        }
        if (tiles) {
            const uint64_t nst = Rpad / kI8KStep;   // 64-row stages
            const uint64_t fr = (uint64_t)flush_rows(int8);   // rows per scale block
            const uint64_t nblk = (Rpad + fr - 1) / fr;
            int *d_uexp = nullptr;
            if (mfma_weights) {
                double *d_scales;
                if ((st = sc.get(&d_uexp, nblk * K)) != GNM_OK) return st;
                if ((st = c->alloc(&d_scales, nblk * pl.ncol)) != GNM_OK) return st;
                GNM_HIP(launch_plan_exps(d_wrow, R, K, fr, limb_qmax(L, int8), L, pl.ncol, nblk, d_uexp, d_scales,
                                         c->stream));
                pl.scales = d_scales;
            }
            // 16 bytes per (stage, tile, lane) of 16 int8 values, or per
            // (stage, tile, k-step, lane) of 8 BF16 values
            uint4 *d_frag;
            if ((st = c->alloc(&d_frag, nst * tiles * (int8 ? 64 : 128))) != GNM_OK) return st;
            GNM_HIP(launch_plan_frag(int8, d_wrow, R, K, d_uexp, fr, L, pl.ncol, pl.wtiles, tiles, nst, d_count_row,
                                     d_rowcls, cw, c->nclass, d_frag, c->stream));
            pl.bfrag = (const uint32_t *)d_frag;
        }
        GNM_HIP(hipStreamSynchronize(c->stream));
    }

    // ---- accumulators and partials
    const uint64_t cells = c->N * K;
    if ((st = c->alloc(&c->d_scores, cells)) != GNM_OK) return st;
    if ((st = c->alloc(&c->d_counts, cells)) != GNM_OK) return st;
    GNM_HIP(launch_fill_f64(c->d_scores, cells, 0.0, c->stream));
    GNM_HIP(launch_fill_u32(c->d_counts, cells, 0u, c->stream));
    c->splits = mfma_weights ? (int8 ? i8_splits(c->N, tiles) : bf16_splits(c->N, tiles)) : fp32_splits(c->N, K);
    if (!mfma_weights && !c->fp32_count_in_kernel)
        c->splits = std::max(c->splits, bf16_splits(c->N, tiles));
    const uint64_t ncol_part = mfma_weights ? (uint64_t)L * K : K;
    if ((st = c->alloc(&c->d_part, (uint64_t)c->splits * c->N * ncol_part)) != GNM_OK) return st;
    if ((st = c->alloc(&c->d_cpart, (uint64_t)c->splits * c->N * c->nclass)) != GNM_OK) return st;
    if (!c->identity) {
        std::vector<uint32_t> o2f(p->output_idx_to_fam_idx, p->output_idx_to_fam_idx + c->N);
        if ((st = c->upload(&c->d_out_to_fam, o2f)) != GNM_OK) return st;
        c->cbpv = ((c->N + 3) / 4 + 63) / 64 * 64;  // padded: 128-bit staging loads
    }
    GNM_HIP(hipEventCreate(&c->ev_main0));
    GNM_HIP(hipEventCreate(&c->ev_main1));
    GNM_HIP(hipEventCreate(&c->ev_all0));
    GNM_HIP(hipEventCreate(&c->ev_all1));
    GNM_HIP(hipStreamSynchronize(c->stream));
    return GNM_OK;
}

void destroy_context(gnm_ctx *c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->copy_stream) (void)hipStreamSynchronize(c->copy_stream);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (int i = 0; i < 2; ++i) {
        if (c->stage_free[i]) (void)hipEventDestroy(c->stage_free[i]);
        if (c->stage_copied[i]) (void)hipEventDestroy(c->stage_copied[i]);
        if (c->h_stage[i]) (void)hipHostFree(c->h_stage[i]);
        if (c->h_map[i]) (void)hipHostFree(c->h_map[i]);
        if (c->h_svrow[i]) (void)hipHostFree(c->h_svrow[i]);
        if (c->h_bmoff[i]) (void)hipHostFree(c->h_bmoff[i]);
        if (c->h_cs[i]) (void)hipHostFree(c->h_cs[i]);
    }
    if (c->copy_stream) (void)hipStreamDestroy(c->copy_stream);
    for (hipEvent_t e : {c->ev_main0, c->ev_main1, c->ev_all0, c->ev_all1})
        if (e) (void)hipEventDestroy(e);
    for (auto it = c->allocations.rbegin(); it != c->allocations.rend(); ++it) (void)hipFree(it->ptr);
    c->allocations.clear();
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

// Sum of missing corrections of variants [from, V) per score: the part of
// the reference baseline (score/prepare.rs:632-650) not yet accounted for.
std::vector<double> baseline_tail(const gnm_ctx *c, uint64_t from) {
    std::vector<double> b(c->K, 0.0);
    for (uint64_t v = from; v < c->V; ++v)
        for (uint64_t e = c->row_offsets[v]; e < c->row_offsets[v + 1]; ++e) b[c->cols[e]] += (double)c->corr[e];
    return b;
}

gnm_status score_device_rows(gnm_ctx *c, const uint8_t *d_rows, uint64_t stride, uint64_t v0, uint64_t n) {
    if (n == 0) return GNM_OK;
    const uint8_t *rows = d_rows;
    uint64_t rstride = stride, bpv = c->bpv;
    GNM_HIP(hipEventRecord(c->ev_all0, c->stream));
    if (!c->identity) {
        if (c->compact_rows < n) {
            // grow the compaction buffer (rare: sized by the largest block seen)
            if (c->d_compact) {
                GNM_HIP(hipStreamSynchronize(c->stream));
                for (auto it = c->allocations.begin(); it != c->allocations.end(); ++it)
                    if (it->ptr == c->d_compact) {
                        (void)hipFree(it->ptr);
                        c->allocations.erase(it);
                        break;
                    }
                c->d_compact = nullptr;
            }
            gnm_status st = c->alloc(&c->d_compact, n * c->cbpv);
            if (st == GNM_OK) GNM_HIP(launch_fill_u32((uint32_t *)c->d_compact, n * c->cbpv / 4, 0u, c->stream));
            if (st != GNM_OK) return st;
            c->compact_rows = n;
        }
        GNM_HIP(launch_compact(d_rows, stride, n, c->d_out_to_fam, c->N, c->d_compact, c->cbpv, c->stream));
        rows = c->d_compact;
        rstride = c->cbpv;
        bpv = (c->N + 3) / 4;
    }
    const uint64_t p0 = c->var_prog_off[v0], p1 = c->var_prog_off[v0 + n];
    // variant splits of this launch: the context's count, but on the MFMA
    // paths no more than one per kMinSplitRows program rows of the block (at
    // least 8, a multiple of 8), so that a block of few rows does not launch
    // thousands of workgroups of a stage or two each
    uint32_t S = c->splits;
    if (c->precision != GNM_PRECISION_FP32) {
        const uint64_t want = ((p1 - p0 + kMinSplitRows - 1) / kMinSplitRows + 7) / 8 * 8;
        S = (uint32_t)std::min<uint64_t>(S, std::max<uint64_t>(want, 8));
    }
    if (p1 == p0) {   // no variant of the block is in any score: nothing to add
        GNM_HIP(hipEventRecord(c->ev_main0, c->stream));
        GNM_HIP(hipEventRecord(c->ev_main1, c->stream));
        GNM_HIP(hipEventRecord(c->ev_all1, c->stream));
        c->timing_valid = true;
        c->completed = v0 + n;
        return GNM_OK;
    }
    // program row p0 + i is variant v0 + i only when every variant of the
    // block has exactly one row (an empty variant and a mixed-flip one
    // would leave the row count equal to n but shift the rows)
    const bool identity = c->non_single[v0 + n] == c->non_single[v0];
    GNM_HIP(hipEventRecord(c->ev_main0, c->stream));
    if (c->precision == GNM_PRECISION_FP32) {
        GNM_HIP(launch_score_fp32(c->plan, rows, rstride, bpv, v0, p0, p1, c->N, c->K, S,
                                  c->fp32_count_in_kernel, identity, c->d_part, c->d_cpart, c->stream));
        if (!c->fp32_count_in_kernel)
            GNM_HIP(launch_score_bf16(c->plan, rows, rstride, bpv, v0, p0, p1, c->N, S, 0, c->plan.ctiles, identity,
                                      c->d_part, c->d_cpart, c->stream));
    } else if (c->plan.int8) {
        GNM_HIP(launch_score_i8(c->plan, rows, rstride, bpv, v0, p0, p1, c->N, S, 0,
                                c->plan.wtiles + c->plan.ctiles, identity, c->d_part, c->d_cpart, c->stream,
                                c->tiles_per_pass, c->batch_stages));
    } else {
        GNM_HIP(launch_score_bf16(c->plan, rows, rstride, bpv, v0, p0, p1, c->N, S, 0,
                                  c->plan.wtiles + c->plan.ctiles, identity, c->d_part, c->d_cpart, c->stream,
                                  c->tiles_per_pass));
    }
    GNM_HIP(hipEventRecord(c->ev_main1, c->stream));
    GNM_HIP(launch_reduce(c->d_part, c->d_cpart, S, c->N, c->K,
                          c->precision == GNM_PRECISION_FP32 ? 1 : c->limbs, c->nclass,
                          c->d_score_class, c->d_scores, c->d_counts, c->stream));
    GNM_HIP(hipEventRecord(c->ev_all1, c->stream));
    c->timing_valid = true;
    c->completed = v0 + n;
    return GNM_OK;
}

// Upload buffers of the sparse ingest, allocated on the first
// gnm_score_rows call that uses it.
gnm_status sparse_init(gnm_ctx *c) {
    c->sparse_nch = (uint32_t)((c->stage_pitch + kSparseChunkWords * 8 - 1) / (kSparseChunkWords * 8));
    const uint64_t ncs = c->stage_rows * (c->sparse_nch + 1);
    gnm_status st;
    // each resource is made once: a call after a failed allocation resumes
    for (int i = 0; i < 2; ++i) {
        if (!c->d_pack[i] && (st = c->alloc(&c->d_pack[i], c->stage_rows * c->stage_pitch)) != GNM_OK) return st;
        if (!c->d_map[i] && (st = c->alloc(&c->d_map[i], c->stage_rows)) != GNM_OK) return st;
        if (!c->d_svrow[i] && (st = c->alloc(&c->d_svrow[i], c->stage_rows)) != GNM_OK) return st;
        if (!c->d_bmoff[i] && (st = c->alloc(&c->d_bmoff[i], c->stage_rows)) != GNM_OK) return st;
        if (!c->d_cs[i] && (st = c->alloc(&c->d_cs[i], ncs)) != GNM_OK) return st;
        if (!c->h_map[i])
            GNM_HIP(hipHostMalloc((void **)&c->h_map[i], c->stage_rows * sizeof(uint32_t), hipHostMallocDefault));
        if (!c->h_svrow[i])
            GNM_HIP(hipHostMalloc((void **)&c->h_svrow[i], c->stage_rows * sizeof(uint32_t), hipHostMallocDefault));
        if (!c->h_bmoff[i])
            GNM_HIP(hipHostMalloc((void **)&c->h_bmoff[i], c->stage_rows * sizeof(uint64_t), hipHostMallocDefault));
        if (!c->h_cs[i]) GNM_HIP(hipHostMalloc((void **)&c->h_cs[i], ncs * sizeof(uint32_t), hipHostMallocDefault));
    }
    c->sparse_ready = true;
    return GNM_OK;
}

gnm_status check_block(gnm_ctx *c, uint64_t first_variant, uint64_t num_rows, uint64_t row_stride) {
    if (!c) return fail(GNM_ERR_INVALID_ARG, "ctx is NULL");
    if (first_variant > c->V || num_rows > c->V - first_variant)
        return fail(GNM_ERR_INVALID_ARG, "row block exceeds num_variants");
    // blocks arrive in ascending order without gaps or repeats (the
    // producer's order, score/io.rs:156-321): a block starts where the last
    // one ended (after a seed, at the seeded count), or at variant 0 once
    // every variant has been scored (a further pass accumulates on top)
    if (first_variant != c->completed && !(first_variant == 0 && c->completed == c->V))
        return fail(GNM_ERR_INVALID_ARG, "row block starts at variant " + std::to_string(first_variant) +
                                             " but the next variant to score is " + std::to_string(c->completed));
    if (row_stride < c->bpv) return fail(GNM_ERR_INVALID_ARG, "row_stride < bytes_per_variant");
    return GNM_OK;
}

// Known-answer check of the split-int8 path on one device, the counterpart
// of the reference's split-path self-test calibrate_fp16
// (score/cuda_backend.rs:2424-2494), with its inputs: 64 people x 64 variants
// x 64 scores, dosage (64 v + p) mod 3, weight ((7 i mod 101) - 50) / 13 for
// i = 64 v + k.  Passes when the counts are exact (all zero) and every score
// is within the limb bound sum_v d_pv * max_v |w_vk| / qmax of the exact sum.
bool split_self_test(int device) {
    constexpr uint64_t D = 64;
    std::vector<float> w(D * D), co(D * D, 0.0f);
    std::vector<uint32_t> cols(D * D);
    std::vector<uint64_t> off(D + 1);
    for (uint64_t v = 0; v < D; ++v) {
        off[v] = v * D;
        for (uint64_t k = 0; k < D; ++k) {
            const int64_t i = (int64_t)(v * D + k);
            w[v * D + k] = (float)((i * 7 % 101) - 50) / 13.0f;
            cols[v * D + k] = (uint32_t)k;
        }
    }
    off[D] = D * D;
    // packed rows: code 00 -> 0, 10 -> 1, 11 -> 2
    std::vector<uint8_t> rows(D * (D / 4), 0);
    for (uint64_t v = 0; v < D; ++v)
        for (uint64_t p = 0; p < D; ++p) {
            const uint32_t d = (uint32_t)((v * D + p) % 3), code = d == 0 ? 0u : (d == 1 ? 2u : 3u);
            rows[v * (D / 4) + p / 4] |= (uint8_t)(code << (2 * (p % 4)));
        }
    gnm_problem pr{};
    pr.num_people = pr.total_people_in_fam = D;
    pr.bytes_per_variant = D / 4;
    pr.num_scores = pr.num_variants = D;
    pr.nnz = D * D;
    pr.sparse_weights = w.data();
    pr.sparse_missing_corrections = co.data();
    pr.sparse_score_columns = cols.data();
    pr.sparse_row_offsets = off.data();
    pr.precision = GNM_PRECISION_I8X4;
    pr.device = device;
    gnm_ctx *c = new gnm_ctx();
    uint8_t *d_rows = nullptr;
    bool ok = build_context(&pr, c) == GNM_OK && hipMalloc(&d_rows, rows.size()) == hipSuccess &&
              hipMemcpy(d_rows, rows.data(), rows.size(), hipMemcpyHostToDevice) == hipSuccess &&
              score_device_rows(c, d_rows, D / 4, 0, D) == GNM_OK && hipStreamSynchronize(c->stream) == hipSuccess;
    std::vector<double> S(D * D);
    std::vector<uint32_t> C(D * D);
    ok = ok && hipMemcpy(S.data(), c->d_scores, S.size() * sizeof(double), hipMemcpyDeviceToHost) == hipSuccess &&
         hipMemcpy(C.data(), c->d_counts, C.size() * sizeof(uint32_t), hipMemcpyDeviceToHost) == hipSuccess;
    if (d_rows) (void)hipFree(d_rows);
    destroy_context(c);
    if (!ok) return false;
    const double qmax = limb_qmax(4, true);
    for (uint64_t k = 0; k < D; ++k) {
        double wmax = 0.0;
        for (uint64_t v = 0; v < D; ++v) wmax = std::max(wmax, std::fabs((double)w[v * D + k]));
        for (uint64_t p = 0; p < D; ++p) {
            double exact = 0.0, dsum = 0.0;
            for (uint64_t v = 0; v < D; ++v) {
                const double d = (double)((v * D + p) % 3);
                exact += d * (double)w[v * D + k];
                dsum += d;
            }
            if (C[p * D + k] != 0 || !(std::fabs(S[p * D + k] - exact) <= dsum * wmax / qmax * (1.0 + 1e-9)))
                return false;
        }
    }
    return true;
}

std::mutex g_auto_mu;
std::map<int, int> g_auto;   // device -> precision AUTO resolves to

// AUTO: split-int8 (I8X4) once its known-answer check has passed on the
// device, else FP32, decided once per device and process (the reference
// keeps exact FP32 when its self-test fails, score/cuda_backend.rs:1183-1207).
int resolve_auto(int device) {
    std::lock_guard<std::mutex> lk(g_auto_mu);
    auto it = g_auto.find(device);
    if (it != g_auto.end()) return it->second;
    int prec = GNM_PRECISION_FP32;
    try {
        if (hipSetDevice(device) == hipSuccess && split_self_test(device)) prec = GNM_PRECISION_I8X4;
    } catch (...) {
    }
    g_auto.emplace(device, prec);
    return prec;
}

}  // namespace

extern "C" {

const char *gnm_last_error(void) { return g_last_error.c_str(); }
int gnm_abi_version(void) { return GNM_ABI_VERSION; }

int gnm_size_gate(uint64_t num_people, uint64_t num_scores, uint64_t num_variants) {
    const unsigned __int128 cells = (unsigned __int128)num_people * num_scores;
    const unsigned __int128 work = cells * num_variants;
    return (cells >= 100000u || work >= 100000000u) ? 1 : 0;
}

int gnm_device_count(void) {
    try {
        if (env_hides_devices()) return 0;
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess) return 0;
        int ok = 0;
        for (int i = 0; i < n; ++i) {
            hipDeviceProp_t prop;
            if (hipGetDeviceProperties(&prop, i) != hipSuccess) continue;
            if (std::strncmp(prop.gcnArchName, "gfx950", 6) == 0) ++ok;
        }
        return ok;
    } catch (...) {
        return 0;
    }
}

static gnm_status create_impl(const gnm_problem *problem, gnm_ctx **out_ctx, bool gate) {
    try {
        if (!out_ctx) return fail(GNM_ERR_INVALID_ARG, "out_ctx is NULL");
        *out_ctx = nullptr;
        gnm_status st = validate(problem);
        if (st != GNM_OK) return st;
        if (gate && !gnm_size_gate(problem->num_people, problem->num_scores, problem->num_variants))
            return fail(GNM_DECLINED, "problem size below the device threshold");
        if (gnm_device_count() == 0) return fail(GNM_ERR_NO_DEVICE, "no visible gfx950 device");
        gnm_problem resolved = *problem;
        if (resolved.precision == GNM_PRECISION_AUTO) resolved.precision = resolve_auto(resolved.device < 0 ? 0 : resolved.device);
        gnm_ctx *c = new gnm_ctx();
        st = build_context(&resolved, c);
        if (st != GNM_OK) {
            std::string msg = g_last_error;
            destroy_context(c);
            g_last_error = msg;
            return st;
        }
        *out_ctx = c;
        return GNM_OK;
    } catch (const std::bad_alloc &) {
        return fail(GNM_ERR_OUT_OF_MEMORY, "host allocation failed while building the plan");
    } catch (const std::exception &e) {
        return fail(GNM_ERR_INTERNAL, e.what());
    } catch (...) {
        return fail(GNM_ERR_INTERNAL, "unknown exception");
    }
}

gnm_status gnm_create(const gnm_problem *problem, gnm_ctx **out_ctx) { return create_impl(problem, out_ctx, true); }
gnm_status gnm_create_forced(const gnm_problem *problem, gnm_ctx **out_ctx) { return create_impl(problem, out_ctx, false); }

gnm_status gnm_score_rows_device(gnm_ctx *ctx, const uint8_t *d_rows, uint64_t row_stride,
                                 uint64_t first_variant, uint64_t num_rows) {
    try {
        gnm_status st = check_block(ctx, first_variant, num_rows, row_stride);
        if (st != GNM_OK) return st;
        if (num_rows && !d_rows) return fail(GNM_ERR_INVALID_ARG, "d_rows is NULL");
        GNM_HIP(hipSetDevice(ctx->device));
        return score_device_rows(ctx, d_rows, row_stride, first_variant, num_rows);
    } catch (...) {
        return fail(GNM_ERR_INTERNAL, "exception in gnm_score_rows_device");
    }
}

gnm_status gnm_score_rows(gnm_ctx *ctx, const uint8_t *rows, uint64_t row_stride,
                          uint64_t first_variant, uint64_t num_rows) {
    try {
        gnm_status st = check_block(ctx, first_variant, num_rows, row_stride);
        if (st != GNM_OK) return st;
        if (num_rows && !rows) return fail(GNM_ERR_INVALID_ARG, "rows is NULL");
        GNM_HIP(hipSetDevice(ctx->device));
        gnm_ctx *c = ctx;
        const uint64_t bpv = c->bpv;
        if (!c->staging_ready) {
            // two pinned slots of ~256 MiB (at least one row) -- the reference's
            // PIPELINE_SLOTS = 2 double buffer (cuda_backend.rs:42, 1137-1147).
            // Copies run on their own stream (default priority), so the copy of
            // one slot overlaps the scoring of the other.  Each resource is
            // made once: a call after a failed allocation resumes here.
            c->stage_pitch = (bpv + 63) / 64 * 64;  // 128-bit aligned rows for the staging loads
            c->stage_rows = std::min<uint64_t>(kSparseMaxSlotRows, std::max<uint64_t>(1, (256ull << 20) / c->stage_pitch));
            if (!c->copy_stream) GNM_HIP(hipStreamCreateWithFlags(&c->copy_stream, hipStreamNonBlocking));
            for (int i = 0; i < 2; ++i) {
                if (!c->h_stage[i]) {
                    GNM_HIP(hipHostMalloc((void **)&c->h_stage[i], c->stage_rows * c->stage_pitch, hipHostMallocDefault));
                    std::memset(c->h_stage[i], 0, c->stage_rows * c->stage_pitch);
                }
                if (!c->d_stage[i]) {
                    gnm_status s2 = c->alloc(&c->d_stage[i], c->stage_rows * c->stage_pitch);
                    if (s2 != GNM_OK) return s2;
                }
                if (!c->stage_free[i]) {
                    GNM_HIP(hipEventCreateWithFlags(&c->stage_free[i], hipEventDisableTiming));
                    GNM_HIP(hipEventRecord(c->stage_free[i], c->stream));
                }
                if (!c->stage_copied[i]) GNM_HIP(hipEventCreateWithFlags(&c->stage_copied[i], hipEventDisableTiming));
            }
            c->staging_ready = true;
        }
        const bool use_sparse = c->sparse_fraction > 0.0;
        if (use_sparse && !c->sparse_ready) {
            st = sparse_init(c);
            if (st != GNM_OK) return st;
        }
        const uint32_t nch = c->sparse_nch;
        const uint64_t nbm = (uint64_t)nch * (kSparseChunkWords / 64);   // bitmap words per sparse row
        // a sparse row (bitmap + stored words) must take less than a packed row
        const uint64_t row_words = (c->Nfam + 31) / 32;
        const uint64_t cap = c->stage_pitch > nbm * 8 ? (c->stage_pitch - nbm * 8) / 8 : 0;
        const uint64_t max_words = std::min<uint64_t>((uint64_t)(c->sparse_fraction * (double)row_words), cap);
        const bool sparse_on = use_sparse && max_words > 0;
        const uint64_t *keep2 = c->identity ? nullptr : c->keep2.data();
        struct TaskOut {
            std::vector<uint32_t> svrow;   // row of each sparse row of the task
            std::vector<uint64_t> bmoff;   // word offset of its record (bitmap, then stored words) in the slot
            std::vector<uint32_t> cs;      // [nch+1] per sparse row, relative to its first stored word
            uint64_t lo = 0, hi = 0;       // task bytes [0, lo): dense rows; [hi, end): sparse records
            uint64_t ndense = 0, nwords = 0;
        };
        uint64_t done = 0;
        int slot = 0;
        while (done < num_rows) {
            const uint64_t n = std::min(c->stage_rows, num_rows - done);
            // the slot's previous rows are scored (and so copied): refill it
            GNM_HIP(hipEventSynchronize(c->stage_free[slot]));
            const uint8_t *src = rows + done * row_stride;
            const uint64_t pitch = c->stage_pitch;
            const uint64_t vblk = first_variant + done;
            uint8_t *dst = c->h_stage[slot];
            // host work split over threads in tasks of ~8 MiB of rows; task t
            // owns the slot bytes of its rows
            const uint64_t rows_per_task = std::max<uint64_t>(1, (8ull << 20) / pitch);
            const uint64_t ntask = (n + rows_per_task - 1) / rows_per_task;
            uint64_t nd = n, ne = 0, nsv = 0;
            std::vector<TaskOut> tout;
            // after a slot in which every row went packed, the next 7 slots
            // skip the classification (dense regions of the genome); every
            // eighth slot looks again
            const bool classify = sparse_on && c->sparse_rest == 0;
            if (!classify && c->sparse_rest) --c->sparse_rest;
            if (classify) {
                // one pass: a task fills its bytes with its dense rows from the
                // bottom and the records of its sparse rows from the top (rows
                // of variants in no score are not sent at all)
                uint32_t *map = c->h_map[slot];
                tout.resize(ntask);
                parallel_for(ntask, [&](uint64_t a, uint64_t b) {
                    std::vector<uint64_t> tmp(max_words + 64), tbm(nbm);
                    std::vector<uint32_t> cst(nch + 1);
                    for (uint64_t t = a; t < b; ++t) {
                        TaskOut &o = tout[t];
                        const uint64_t r0 = t * rows_per_task, r1 = std::min(n, r0 + rows_per_task);
                        uint8_t *reg = dst + r0 * pitch;
                        uint64_t lo = 0, hi = (r1 - r0) * pitch;   // each row takes at most pitch bytes
                        for (uint64_t r = r0; r < r1; ++r) {
                            const uint64_t v = vblk + r;
                            const uint8_t *row = src + r * row_stride;
                            map[r] = kMapZero;
                            if (c->row_offsets[v] == c->row_offsets[v + 1]) continue;
                            const int64_t k = sparse_encode_row(row, c->Nfam, keep2, max_words, nch, tmp.data(),
                                                                tbm.data(), cst.data());
                            if (k < 0) {
                                std::memcpy(reg + lo, row, bpv);
                                std::memset(reg + lo + bpv, 0, pitch - bpv);
                                map[r] = (uint32_t)(r0 + o.ndense++);   // row slot in the upload
                                lo += pitch;
                            } else if (k > 0) {
                                hi -= (nbm + (uint64_t)k) * 8;
                                std::memcpy(reg + hi, tbm.data(), nbm * 8);
                                std::memcpy(reg + hi + nbm * 8, tmp.data(), (uint64_t)k * 8);
                                map[r] = kMapSparse;
                                o.svrow.push_back((uint32_t)r);
                                o.bmoff.push_back((r0 * pitch + hi) / 8);
                                o.cs.insert(o.cs.end(), cst.begin(), cst.end());
                                o.nwords += (uint64_t)k;
                            }
                        }
                        o.lo = lo;
                        o.hi = hi;
                    }
                }, 2);
                nd = 0;
                for (const TaskOut &o : tout) {
                    nd += o.ndense;
                    ne += o.nwords;
                    nsv += o.svrow.size();
                }
                if (nd == n) c->sparse_rest = 7;
            }
            if (nd < n) {
                uint64_t bytes = 0, s0 = 0;
                uint32_t *svrow = c->h_svrow[slot], *cs = c->h_cs[slot];
                uint64_t *bmoff = c->h_bmoff[slot];
                for (uint64_t t = 0; t < ntask; ++t) {
                    const TaskOut &o = tout[t];
                    const uint64_t base = t * rows_per_task * pitch;
                    const uint64_t end = (std::min(n, (t + 1) * rows_per_task) - t * rows_per_task) * pitch;
                    if (o.lo)
                        GNM_HIP(hipMemcpyAsync(c->d_pack[slot] + base, dst + base, o.lo, hipMemcpyHostToDevice,
                                               c->copy_stream));
                    if (o.hi < end)
                        GNM_HIP(hipMemcpyAsync(c->d_pack[slot] + base + o.hi, dst + base + o.hi, end - o.hi,
                                               hipMemcpyHostToDevice, c->copy_stream));
                    bytes += o.lo + (end - o.hi);
                    std::memcpy(svrow + s0, o.svrow.data(), o.svrow.size() * sizeof(uint32_t));
                    std::memcpy(bmoff + s0, o.bmoff.data(), o.bmoff.size() * sizeof(uint64_t));
                    std::memcpy(cs + s0 * (nch + 1), o.cs.data(), o.cs.size() * sizeof(uint32_t));
                    s0 += o.svrow.size();
                }
                GNM_HIP(hipMemcpyAsync(c->d_map[slot], c->h_map[slot], n * sizeof(uint32_t), hipMemcpyHostToDevice,
                                       c->copy_stream));
                if (nsv) {
                    GNM_HIP(hipMemcpyAsync(c->d_svrow[slot], svrow, nsv * sizeof(uint32_t), hipMemcpyHostToDevice,
                                           c->copy_stream));
                    GNM_HIP(hipMemcpyAsync(c->d_bmoff[slot], bmoff, nsv * sizeof(uint64_t), hipMemcpyHostToDevice,
                                           c->copy_stream));
                    GNM_HIP(hipMemcpyAsync(c->d_cs[slot], cs, nsv * (nch + 1) * sizeof(uint32_t),
                                           hipMemcpyHostToDevice, c->copy_stream));
                }
                GNM_HIP(hipEventRecord(c->stage_copied[slot], c->copy_stream));
                GNM_HIP(hipStreamWaitEvent(c->stream, c->stage_copied[slot], 0));
                GNM_HIP(launch_expand_rows(c->d_pack[slot], c->d_map[slot], n, pitch, c->d_stage[slot], c->stream));
                GNM_HIP(launch_sparse_rows((const uint64_t *)c->d_pack[slot], c->d_bmoff[slot], c->d_svrow[slot],
                                           c->d_cs[slot], nsv, nch, pitch, c->d_stage[slot], c->stream));
                st = score_device_rows(c, c->d_stage[slot], pitch, vblk, n);
                if (st != GNM_OK) return st;
                c->st_bytes += bytes + n * sizeof(uint32_t) + nsv * (nch + 4) * sizeof(uint32_t);   // + metadata
            } else {
                // every row dense: the rows as they are (a classifying pass
                // already wrote them in order)
                if (!classify) parallel_for(ntask, [&](uint64_t a, uint64_t b) {
                    const uint64_t r0 = a * rows_per_task, r1 = std::min(n, b * rows_per_task);
                    if (row_stride == pitch) {
                        std::memcpy(dst + r0 * pitch, src + r0 * pitch, (r1 - r0) * pitch);
                    } else {
                        for (uint64_t r = r0; r < r1; ++r) std::memcpy(dst + r * pitch, src + r * row_stride, bpv);
                    }
                }, 2);
                GNM_HIP(hipMemcpyAsync(c->d_stage[slot], dst, n * pitch, hipMemcpyHostToDevice, c->copy_stream));
                GNM_HIP(hipEventRecord(c->stage_copied[slot], c->copy_stream));
                GNM_HIP(hipStreamWaitEvent(c->stream, c->stage_copied[slot], 0));
                st = score_device_rows(c, c->d_stage[slot], pitch, vblk, n);
                if (st != GNM_OK) return st;
                c->st_bytes += n * pitch;
            }
            c->st_dense += nd;
            c->st_sparse += nsv;
            c->st_words += ne;
            GNM_HIP(hipEventRecord(c->stage_free[slot], c->stream));
            done += n;
            slot ^= 1;
        }
        // the caller may reuse `rows` now: every byte has been copied to pinned staging
        return GNM_OK;
    } catch (const std::bad_alloc &) {
        return fail(GNM_ERR_OUT_OF_MEMORY, "host allocation failed in gnm_score_rows");
    } catch (...) {
        return fail(GNM_ERR_INTERNAL, "exception in gnm_score_rows");
    }
}

gnm_status gnm_read_accumulators(gnm_ctx *ctx, double *scores, uint32_t *counts, uint64_t *completed_variants) {
    try {
        if (!ctx || !scores || !counts) return fail(GNM_ERR_INVALID_ARG, "NULL argument");
        GNM_HIP(hipSetDevice(ctx->device));
        GNM_HIP(hipStreamSynchronize(ctx->stream));
        const uint64_t cells = ctx->N * ctx->K;
        GNM_HIP(hipMemcpy(scores, ctx->d_scores, cells * sizeof(double), hipMemcpyDeviceToHost));
        GNM_HIP(hipMemcpy(counts, ctx->d_counts, cells * sizeof(uint32_t), hipMemcpyDeviceToHost));
        const std::vector<double> tail = baseline_tail(ctx, ctx->completed);
        for (uint64_t p = 0; p < ctx->N; ++p)
            for (uint64_t k = 0; k < ctx->K; ++k) scores[p * ctx->K + k] += tail[k];
        if (completed_variants) *completed_variants = ctx->completed;
        return GNM_OK;
    } catch (...) {
        return fail(GNM_ERR_INTERNAL, "exception in gnm_read_accumulators");
    }
}

gnm_status gnm_seed_accumulators(gnm_ctx *ctx, const double *scores, const uint32_t *counts, uint64_t completed_variants) {
    try {
        if (!ctx || !scores || !counts) return fail(GNM_ERR_INVALID_ARG, "NULL argument");
        if (completed_variants > ctx->V) return fail(GNM_ERR_INVALID_ARG, "completed_variants > num_variants");
        GNM_HIP(hipSetDevice(ctx->device));
        GNM_HIP(hipStreamSynchronize(ctx->stream));
        const uint64_t cells = ctx->N * ctx->K;
        const std::vector<double> tail = baseline_tail(ctx, completed_variants);
        std::vector<double> s(scores, scores + cells);
        for (uint64_t p = 0; p < ctx->N; ++p)
            for (uint64_t k = 0; k < ctx->K; ++k) s[p * ctx->K + k] -= tail[k];
        GNM_HIP(hipMemcpy(ctx->d_scores, s.data(), cells * sizeof(double), hipMemcpyHostToDevice));
        GNM_HIP(hipMemcpy(ctx->d_counts, counts, cells * sizeof(uint32_t), hipMemcpyHostToDevice));
        ctx->completed = completed_variants;
        return GNM_OK;
    } catch (...) {
        return fail(GNM_ERR_INTERNAL, "exception in gnm_seed_accumulators");
    }
}

gnm_status gnm_finish(gnm_ctx *ctx, double *scores, uint32_t *counts) {
    uint64_t done = 0;
    gnm_status st = gnm_read_accumulators(ctx, scores, counts, &done);
    if (st != GNM_OK) return st;
    if (done != ctx->V) return fail(GNM_ERR_INVALID_ARG, "gnm_finish before all variants were scored");
    return GNM_OK;
}

void gnm_destroy(gnm_ctx *ctx) {
    try {
        destroy_context(ctx);
    } catch (...) {
    }
}

gnm_status gnm_run(const gnm_problem *problem, const uint8_t *rows, uint64_t row_stride, double *scores,
                   uint32_t *counts) {
    gnm_ctx *c = nullptr;
    gnm_status st = gnm_create(problem, &c);
    if (st != GNM_OK) return st;
    st = gnm_score_rows(c, rows, row_stride, 0, problem->num_variants);
    if (st == GNM_OK) st = gnm_finish(c, scores, counts);
    std::string msg = g_last_error;
    gnm_destroy(c);
    g_last_error = msg;
    return st;
}

gnm_status gnm_last_timing(gnm_ctx *ctx, float *main_kernel_ms, float *total_ms) {
    try {
        if (!ctx || !ctx->timing_valid) return fail(GNM_ERR_INVALID_ARG, "no timed call on this context");
        GNM_HIP(hipSetDevice(ctx->device));
        GNM_HIP(hipEventSynchronize(ctx->ev_all1));
        float a = 0.f, b = 0.f;
        GNM_HIP(hipEventElapsedTime(&a, ctx->ev_main0, ctx->ev_main1));
        GNM_HIP(hipEventElapsedTime(&b, ctx->ev_all0, ctx->ev_all1));
        if (main_kernel_ms) *main_kernel_ms = a;
        if (total_ms) *total_ms = b;
        return GNM_OK;
    } catch (...) {
        return fail(GNM_ERR_INTERNAL, "exception in gnm_last_timing");
    }
}

gnm_status gnm_unpack_device(const uint8_t *d_rows, uint64_t bytes_per_variant, uint64_t num_variants,
                             uint64_t num_people, const uint8_t *d_flip, uint8_t *d_dosage,
                             uint32_t *d_missing_per_person, void *stream) {
    try {
        if (num_variants && (!d_rows || !d_dosage || !d_missing_per_person))
            return fail(GNM_ERR_INVALID_ARG, "NULL device pointer");
        if (bytes_per_variant < (num_people + 3) / 4) return fail(GNM_ERR_INVALID_ARG, "bytes_per_variant too small");
        // the launch grid holds at most 65535 spans of 16384 people
        if (num_people > 65535ull * 16384ull)
            return fail(GNM_ERR_INVALID_ARG, "num_people exceeds the unpack limit of 1073725440");
        GNM_HIP(launch_unpack(d_rows, bytes_per_variant, num_variants, num_people, d_flip, d_dosage,
                              d_missing_per_person, (hipStream_t)stream));
        return GNM_OK;
    } catch (...) {
        return fail(GNM_ERR_INTERNAL, "exception in gnm_unpack_device");
    }
}

gnm_status gnm_hbm_probe(int kind, const void *d_src, uint64_t src_bytes, void *d_dst, uint64_t dst_bytes,
                         void *stream) {
    try {
        const uintptr_t a = reinterpret_cast<uintptr_t>(d_src), b = reinterpret_cast<uintptr_t>(d_dst);
        if (kind < 0 || kind > 2) return fail(GNM_ERR_INVALID_ARG, "probe kind must be 0 (read), 1 (write) or 2 (expand)");
        if (src_bytes == 0 || src_bytes % 16) return fail(GNM_ERR_INVALID_ARG, "probe size must be a positive multiple of 16");
        if (!d_dst || (b & 15u)) return fail(GNM_ERR_INVALID_ARG, "probe output must be a 16-byte aligned device pointer");
        if (kind != 1 && (!d_src || (a & 15u))) return fail(GNM_ERR_INVALID_ARG, "probe input must be a 16-byte aligned device pointer");
        const uint64_t need = kind == 0 ? hbm_probe_sink_bytes() : (kind == 1 ? src_bytes : 4 * src_bytes);
        if (dst_bytes < need) return fail(GNM_ERR_INVALID_ARG, "probe output buffer too small");
        GNM_HIP(launch_hbm_probe(kind, (const uint8_t *)d_src, src_bytes, (uint8_t *)d_dst, (hipStream_t)stream));
        return GNM_OK;
    } catch (...) {
        return fail(GNM_ERR_INTERNAL, "exception in gnm_hbm_probe");
    }
}

gnm_status gnm_set_launch_config(gnm_ctx *ctx, uint32_t tiles_per_pass, uint32_t batch_stages) {
    if (!ctx) return fail(GNM_ERR_INVALID_ARG, "ctx is NULL");
    if (!is_int8(ctx->precision)) {
        // split-BF16 kernel: tile groups of 3, 5, 6, 8 or 9, two-stage batches
        if (tiles_per_pass != 0 && tiles_per_pass != 3 && tiles_per_pass != 5 && tiles_per_pass != 6 &&
            tiles_per_pass != 8 && tiles_per_pass != 9)
            return fail(GNM_ERR_INVALID_ARG, "tiles_per_pass must be 0, 3, 5, 6, 8 or 9 on the split-BF16 path");
        if (batch_stages != 0 && batch_stages != 2)
            return fail(GNM_ERR_INVALID_ARG, "batch_stages must be 0 or 2 on the split-BF16 path");
        ctx->tiles_per_pass = tiles_per_pass;
        ctx->batch_stages = batch_stages;
        return GNM_OK;
    }
    if (tiles_per_pass != 0 && tiles_per_pass != 3 && tiles_per_pass != 4 && tiles_per_pass != 5 &&
        tiles_per_pass != 6 && tiles_per_pass != 8 && tiles_per_pass != 18)
        return fail(GNM_ERR_INVALID_ARG, "tiles_per_pass must be 0, 3, 4, 5, 6, 8 or 18");
    if (tiles_per_pass == 18 && batch_stages == 4)
        return fail(GNM_ERR_INVALID_ARG, "tiles_per_pass 18 runs two-stage batches only");
    if (batch_stages != 0 && batch_stages != 2 && batch_stages != 4)
        return fail(GNM_ERR_INVALID_ARG, "batch_stages must be 0, 2 or 4");
    ctx->tiles_per_pass = tiles_per_pass;
    ctx->batch_stages = batch_stages;
    return GNM_OK;
}

gnm_status gnm_context_shape(const gnm_ctx *ctx, uint64_t *program_rows, uint32_t *mfma_tiles) {
    if (!ctx || !program_rows || !mfma_tiles) return fail(GNM_ERR_INVALID_ARG, "NULL argument");
    *program_rows = ctx->R;
    *mfma_tiles = ctx->precision == GNM_PRECISION_FP32 ? 0u : ctx->plan.wtiles + ctx->plan.ctiles;
    return GNM_OK;
}

gnm_status gnm_set_sparse_ingest(gnm_ctx *ctx, double max_word_fraction) {
    if (!ctx) return fail(GNM_ERR_INVALID_ARG, "ctx is NULL");
    if (!(max_word_fraction >= 0.0 && max_word_fraction <= 0.5))
        return fail(GNM_ERR_INVALID_ARG, "max_word_fraction must lie in [0, 1/2]");
    ctx->sparse_fraction = max_word_fraction;
    ctx->sparse_rest = 0;
    return GNM_OK;
}

gnm_status gnm_ingest_stats(gnm_ctx *ctx, uint64_t *dense_rows, uint64_t *sparse_rows, uint64_t *stored_words,
                            uint64_t *bytes_to_device) {
    if (!ctx || !dense_rows || !sparse_rows || !stored_words || !bytes_to_device)
        return fail(GNM_ERR_INVALID_ARG, "NULL argument");
    *dense_rows = ctx->st_dense;
    *sparse_rows = ctx->st_sparse;
    *stored_words = ctx->st_words;
    *bytes_to_device = ctx->st_bytes;
    return GNM_OK;
}

int64_t gnm_sparse_encode_row(const uint8_t *row, uint64_t num_fam, const uint64_t *keep2, uint64_t max_words,
                              uint64_t *words, uint64_t *bitmap, uint32_t *chunk_start) {
    if ((!row && num_fam) || !words || !bitmap || !chunk_start) return -2;
    const uint32_t nch = (uint32_t)((num_fam + 32ull * kSparseChunkWords - 1) / (32ull * kSparseChunkWords));
    return sparse_encode_row(row, num_fam, keep2, max_words, nch, words, bitmap, chunk_start);
}

gnm_status gnm_synth_rows_device(uint64_t seed, uint32_t p_lo24, uint32_t p_hi24, uint32_t miss16,
                                 uint64_t num_samples, uint64_t first_variant, uint64_t num_rows,
                                 uint8_t *d_out, uint64_t row_stride, void *stream) {
    try {
        if (num_rows && !d_out) return fail(GNM_ERR_INVALID_ARG, "d_out is NULL");
        if (p_hi24 < p_lo24 || p_hi24 > (1u << 24)) return fail(GNM_ERR_INVALID_ARG, "bad frequency range");
        if (miss16 > 65536) return fail(GNM_ERR_INVALID_ARG, "bad missing rate");
        if (row_stride < (num_samples + 3) / 4) return fail(GNM_ERR_INVALID_ARG, "row_stride too small");
        GNM_HIP(launch_synth(seed, p_lo24, p_hi24, miss16, num_samples, first_variant, num_rows, d_out,
                             row_stride, (hipStream_t)stream));
        return GNM_OK;
    } catch (...) {
        return fail(GNM_ERR_INTERNAL, "exception in gnm_synth_rows_device");
    }
}

}  // extern "C"

namespace {

// Grow-only device scratch of the dense entry points, one set per device.
// Calls on one device are expected to be serialised by the caller (they run
// on the caller's stream); growing waits for that stream before freeing.
struct DenseWorkspace {
    std::vector<std::pair<void *, size_t>> bufs;
    std::mutex mu;
};
std::mutex g_ws_mu;
std::map<int, DenseWorkspace *> g_ws;

gnm_status ws_get(DenseWorkspace *ws, size_t slot, size_t bytes, hipStream_t stream, void **out) {
    if (ws->bufs.size() <= slot) ws->bufs.resize(slot + 1, {nullptr, 0});
    auto &b = ws->bufs[slot];
    bytes = std::max<size_t>(bytes, 256);
    if (b.second < bytes) {
        if (b.first) {
            GNM_HIP(hipStreamSynchronize(stream));
            GNM_HIP(hipFree(b.first));
            b = {nullptr, 0};
        }
        void *p = nullptr;
        hipError_t e = hipMalloc(&p, bytes);
        if (e != hipSuccess)
            return fail(e == hipErrorOutOfMemory ? GNM_ERR_OUT_OF_MEMORY : GNM_ERR_HIP,
                        "hipMalloc(" + std::to_string(bytes) + "): " + hipGetErrorString(e));
        b = {p, bytes};
    }
    *out = b.first;
    return GNM_OK;
}

DenseWorkspace *workspace_for_current_device() {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return nullptr;
    std::lock_guard<std::mutex> lk(g_ws_mu);
    auto it = g_ws.find(dev);
    if (it == g_ws.end()) it = g_ws.emplace(dev, new DenseWorkspace()).first;
    return it->second;
}

template <class T>
gnm_status ws_typed(DenseWorkspace *ws, size_t slot, size_t count, hipStream_t st, T **out) {
    void *p = nullptr;
    gnm_status s = ws_get(ws, slot, count * sizeof(T), st, &p);
    *out = (T *)p;
    return s;
}

// Buffers of the split-int8 / split-BF16 adjoint, in elements: block
// exponents (int), R limb fragments (u32), limb scales (f64); the FP64
// accumulator is [V][Kp].  R is split per block of kI8FlushRows (int8) or
// kBfFlushRows (BF16) people.
struct AdjointI8Shape {
    bool bf;
    uint32_t L, used, tiles;
    uint64_t Kp, nks, nblk, exps, frag, scales;
};

gnm_status adjoint_i8_shape(uint64_t N, uint64_t K, int32_t precision, bool aligned, AdjointI8Shape *sh) {
    sh->bf = !is_int8(precision);
    sh->L = limbs_of(precision);
    sh->Kp = (K + 15) / 16 * 16;
    sh->used = (uint32_t)(sh->L * (sh->Kp / 16));
    sh->tiles = (sh->used + 3) / 4 * 4;
    if (sh->tiles > 16)
        return fail(GNM_ERR_UNSUPPORTED, sh->bf ? "the split-BF16 adjoint supports up to 16 limb tiles per call "
                                                  "(64 scores in bf16x4)"
                                                : "the int8 adjoint supports up to 64 scores per call");
    if (!aligned)
        return fail(GNM_ERR_UNSUPPORTED, "the split-limb adjoint needs 4-byte aligned rows with a stride that is "
                                         "a multiple of 4 (below 2^25)");
    // super-steps of 256 people (int8) or 128 people (BF16), 4 k-steps each
    const uint64_t nss = ((N + 3) / 4 + (sh->bf ? 31 : 63)) / (sh->bf ? 32 : 64);
    const uint64_t fr = sh->bf ? (uint64_t)kBfFlushRows : (uint64_t)kI8FlushRows;
    sh->nks = nss * 4;
    sh->nblk = (N + fr - 1) / fr;
    sh->exps = sh->nblk * K;
    sh->frag = sh->nks * sh->tiles * 64 * 4;
    sh->scales = sh->nblk * sh->L * sh->Kp;
    return GNM_OK;
}

// Every buffer is fully written before it is read: scales are cleared here,
// exponents and fragments are written whole, and the adjoint kernel's first
// flush per variant overwrites the accumulator.
gnm_status adjoint_i8_launch(const AdjointI8Shape &sh, const uint8_t *d_rows, uint64_t stride, uint64_t V,
                             uint64_t N, const uint8_t *d_flip, const float *d_grad_scores, uint64_t K, int *exps,
                             uint32_t *frag, double *scales, double *acc, float *d_grad_weights, hipStream_t st) {
    GNM_HIP(launch_fill_f64(scales, sh.scales, 0.0, st));
    GNM_HIP(launch_block_exponent(d_grad_scores, N, K, K, sh.bf ? kBfFlushRows : kI8FlushRows, sh.L, !sh.bf, exps,
                                  st));
    if (sh.bf)
        GNM_HIP(launch_adjoint_frag_bf(d_grad_scores, N, K, sh.Kp, sh.L, exps, sh.tiles, sh.used, sh.nks, frag,
                                       scales, st));
    else
        GNM_HIP(launch_adjoint_frag_i8(d_grad_scores, N, K, sh.Kp, sh.L, exps, sh.tiles, sh.used, sh.nks, frag,
                                       scales, st));
    GNM_HIP(launch_adjoint_mx(sh.bf, d_rows, stride, V, N, d_flip, frag, scales, sh.tiles, K, sh.Kp, sh.L, acc,
                              d_grad_weights, st));
    return GNM_OK;
}

}  // namespace

extern "C" {

gnm_status gnm_dense_forward_device(const uint8_t *d_rows, uint64_t bytes_per_variant, uint64_t num_variants,
                                    uint64_t num_people, const uint8_t *d_flip, const float *d_weights,
                                    uint64_t num_scores, int32_t precision, double *d_scores,
                                    uint32_t *d_missing_per_person, void *stream_) {
    try {
        const uint64_t V = num_variants, N = num_people, K = num_scores;
        if (N == 0 || K == 0) return fail(GNM_ERR_INVALID_ARG, "num_people and num_scores must be positive");
        if (V && (!d_rows || !d_weights)) return fail(GNM_ERR_INVALID_ARG, "NULL device pointer");
        if (!d_scores || !d_missing_per_person) return fail(GNM_ERR_INVALID_ARG, "NULL output pointer");
        if (bytes_per_variant < (N + 3) / 4) return fail(GNM_ERR_INVALID_ARG, "bytes_per_variant too small");
        if (precision < GNM_PRECISION_FP32 || precision > GNM_PRECISION_BF16X4)
            return fail(GNM_ERR_INVALID_ARG, "unknown precision");
        if (precision == GNM_PRECISION_AUTO) {
            int dev = 0;
            GNM_HIP(hipGetDevice(&dev));
            precision = resolve_auto(dev);
        }
        hipStream_t st = (hipStream_t)stream_;
        DenseWorkspace *ws = workspace_for_current_device();
        if (!ws) return fail(GNM_ERR_NO_DEVICE, "no current HIP device");
        std::lock_guard<std::mutex> lk(ws->mu);
        const uint32_t L = limbs_of(precision);
        const bool int8 = is_int8(precision);
        const uint64_t Rpad = std::max<uint64_t>((V + 63) / 64 * 64, 64);
        const uint32_t Lw = L == 1 ? 0 : L;
        const uint32_t ntl = (uint32_t)((Lw * K + 15) / 16);   // weight tiles; missing calls counted on the VALU
        const uint32_t S = precision == GNM_PRECISION_FP32 ? fp32_splits(N, K)
                                                           : (int8 ? i8_splits(N, ntl) : bf16_splits(N, ntl));
        gnm_status s;
        DevicePlan pl{};
        uint32_t *prog_var, *flipmask, *prog_cls, *score_class, *cnt_tmp, *cpart;
        double *part;
        if ((s = ws_typed(ws, 0, Rpad, st, &prog_var)) != GNM_OK) return s;
        if ((s = ws_typed(ws, 1, Rpad / 32, st, &flipmask)) != GNM_OK) return s;
        if ((s = ws_typed(ws, 2, Rpad, st, &prog_cls)) != GNM_OK) return s;
        if ((s = ws_typed(ws, 3, K, st, &score_class)) != GNM_OK) return s;
        GNM_HIP(launch_fill_u32(score_class, K, 0u, st));
        GNM_HIP(launch_dense_program(d_flip, V, Rpad, prog_var, flipmask, prog_cls, st));
        pl.num_prog_rows = V;
        pl.prog_var = prog_var;
        pl.flipmask = flipmask;
        pl.prog_cls = prog_cls;
        pl.nclass = 1;
        pl.score_class = score_class;
        pl.limbs = L;
        const uint64_t ncol_part = L == 1 ? K : (uint64_t)L * K;
        if ((s = ws_typed(ws, 4, (uint64_t)S * N * ncol_part, st, &part)) != GNM_OK) return s;
        if ((s = ws_typed(ws, 5, (uint64_t)S * N, st, &cpart)) != GNM_OK) return s;
        if ((s = ws_typed(ws, 6, N * K, st, &cnt_tmp)) != GNM_OK) return s;
        GNM_HIP(launch_fill_f64(d_scores, N * K, 0.0, st));
        GNM_HIP(launch_fill_u32(cnt_tmp, N * K, 0u, st));
        if (precision == GNM_PRECISION_FP32) {
            const uint64_t k32 = fp32_weight_pitch(K);
            float *w32;
            if ((s = ws_typed(ws, 7, Rpad * k32, st, &w32)) != GNM_OK) return s;
            GNM_HIP(launch_pad_weights(d_weights, V, K, Rpad, k32, w32, st));
            pl.w32 = w32;
            pl.k32 = k32;
            GNM_HIP(launch_score_fp32(pl, d_rows, bytes_per_variant, (N + 3) / 4, 0, 0, V, N, K, S, true, true,
                                      part, cpart, st));
        } else {
            pl.ncol = (uint32_t)(L * K);
            pl.wtiles = (pl.ncol + 15) / 16;
            // one class, every row a count row: counted on the vector ALUs or
            // in one count tile, whichever the cost model prefers
            pl.vcount = vcount_pays(int8, pl.wtiles) ? 1u : 0u;
            pl.ctiles = pl.vcount ? 0u : 1u;
            const uint32_t tiles = pl.wtiles + pl.ctiles;
            const uint64_t nst = Rpad / kI8KStep;   // 64-row stages
            const uint64_t fr = (uint64_t)flush_rows(int8);
            const uint64_t nblk = (Rpad + fr - 1) / fr;
            int *exps;
            uint16_t *frag;   // 16 bytes per (stage, tile, lane) int8, per (stage, tile, k-step, lane) BF16
            double *scales;
            uint8_t *flipbyte;
            if ((s = ws_typed(ws, 8, nblk * K, st, &exps)) != GNM_OK) return s;
            if ((s = ws_typed(ws, 9, nst * tiles * (int8 ? 64 : 128) * 8, st, &frag)) != GNM_OK) return s;
            if ((s = ws_typed(ws, 10, nblk * pl.ncol, st, &scales)) != GNM_OK) return s;
            if ((s = ws_typed(ws, 14, Rpad, st, &flipbyte)) != GNM_OK) return s;
            GNM_HIP(launch_block_exponent(d_weights, V, K, K, fr, L, int8, exps, st));
            GNM_HIP(launch_flip_bytes(d_flip, V, Rpad, int8 ? kFlipSelI8 : kFlipSelBf, flipbyte, st));
            pl.scales = scales;
            pl.bfrag = (const uint32_t *)frag;
            pl.flipbyte = flipbyte;
            if (int8) {
                GNM_HIP(launch_forward_frag_i8(d_weights, V, K, L, exps, pl.wtiles, tiles, nst, (uint32_t *)frag,
                                               scales, st));
                pl.int8 = 1;
                GNM_HIP(launch_score_i8(pl, d_rows, bytes_per_variant, (N + 3) / 4, 0, 0, V, N, S, 0, tiles, true,
                                        part, cpart, st));
            } else {
                GNM_HIP(launch_forward_frag(d_weights, V, K, L, exps, pl.wtiles, tiles, nst, frag, scales, st));
                GNM_HIP(launch_score_bf16(pl, d_rows, bytes_per_variant, (N + 3) / 4, 0, 0, V, N, S, 0, tiles,
                                          true, part, cpart, st));
            }
        }
        GNM_HIP(launch_reduce(part, cpart, S, N, K, L, 1, score_class, d_scores, cnt_tmp, st));
        GNM_HIP(launch_take_col0(cnt_tmp, N, K, d_missing_per_person, st));
        return GNM_OK;
    } catch (...) {
        return fail(GNM_ERR_INTERNAL, "exception in gnm_dense_forward_device");
    }
}

gnm_status gnm_dense_adjoint_device(const uint8_t *d_rows, uint64_t bytes_per_variant, uint64_t num_variants,
                                    uint64_t num_people, const uint8_t *d_flip, const float *d_grad_scores,
                                    uint64_t num_scores, int32_t precision, float *d_grad_weights,
                                    void *stream_) {
    try {
        const uint64_t V = num_variants, N = num_people, K = num_scores;
        if (N == 0 || K == 0) return fail(GNM_ERR_INVALID_ARG, "num_people and num_scores must be positive");
        if (V && (!d_rows || !d_grad_weights)) return fail(GNM_ERR_INVALID_ARG, "NULL device pointer");
        if (!d_grad_scores) return fail(GNM_ERR_INVALID_ARG, "NULL grad_scores");
        if (bytes_per_variant < (N + 3) / 4) return fail(GNM_ERR_INVALID_ARG, "bytes_per_variant too small");
        if (precision < GNM_PRECISION_FP32 || precision > GNM_PRECISION_BF16X4)
            return fail(GNM_ERR_INVALID_ARG, "unknown precision");
        const bool aligned = ((reinterpret_cast<uintptr_t>(d_rows) | bytes_per_variant) & 3u) == 0 &&
                             bytes_per_variant < (1ull << 25);
        if (precision == GNM_PRECISION_AUTO) {
            int dev = 0;
            GNM_HIP(hipGetDevice(&dev));
            precision = (aligned && resolve_auto(dev) == GNM_PRECISION_I8X4) ? GNM_PRECISION_I8X4
                                                                             : GNM_PRECISION_BF16X3;
        }
        // split-int8, and split-BF16 on aligned rows: the shared adjoint
        // kernel; split-BF16 on unaligned rows: the byte-load kernel below
        const bool bf = precision == GNM_PRECISION_BF16X2 || precision == GNM_PRECISION_BF16X3 ||
                        precision == GNM_PRECISION_BF16X4;
        if (is_int8(precision) || (bf && aligned && limbs_of(precision) * ((K + 15) / 16) <= 16)) {
            AdjointI8Shape sh;
            gnm_status s = adjoint_i8_shape(N, K, precision, aligned, &sh);
            if (s != GNM_OK) return s;
            hipStream_t st = (hipStream_t)stream_;
            DenseWorkspace *ws = workspace_for_current_device();
            if (!ws) return fail(GNM_ERR_NO_DEVICE, "no current HIP device");
            std::lock_guard<std::mutex> lk(ws->mu);
            int *exps;
            uint32_t *frag;
            double *scales, *acc;
            if ((s = ws_typed(ws, 11, sh.exps, st, &exps)) != GNM_OK) return s;
            if ((s = ws_typed(ws, 12, sh.frag, st, &frag)) != GNM_OK) return s;
            if ((s = ws_typed(ws, 13, sh.scales, st, &scales)) != GNM_OK) return s;
            if ((s = ws_typed(ws, 15, std::max<uint64_t>(V, 1) * sh.Kp, st, &acc)) != GNM_OK) return s;
            return adjoint_i8_launch(sh, d_rows, bytes_per_variant, V, N, d_flip, d_grad_scores, K, exps, frag,
                                     scales, acc, d_grad_weights, st);
        }
        if (precision != GNM_PRECISION_BF16X2 && precision != GNM_PRECISION_BF16X3 && precision != GNM_PRECISION_BF16X4)
            return fail(GNM_ERR_UNSUPPORTED, "the adjoint runs in split-BF16 (bf16x2, bf16x3, bf16x4) or "
                                             "split-int8 (i8x3, i8x4)");
        const uint32_t L = limbs_of(precision);
        const uint64_t Kp = (K + 15) / 16 * 16;
        const uint32_t tiles = (uint32_t)(L * (Kp / 16));
        if (tiles > 12)
            return fail(GNM_ERR_UNSUPPORTED, "the split-BF16 adjoint supports up to 12 limb tiles per call "
                                             "(64 scores in bf16x3, 48 in bf16x4)");
        hipStream_t st = (hipStream_t)stream_;
        DenseWorkspace *ws = workspace_for_current_device();
        if (!ws) return fail(GNM_ERR_NO_DEVICE, "no current HIP device");
        std::lock_guard<std::mutex> lk(ws->mu);
        const uint64_t nchunks = (N + 255) / 256;
        const uint64_t nks = nchunks * 8;
        const uint64_t nblk = (nchunks * 256 + kFlushRows - 1) / kFlushRows;
        gnm_status s;
        int *exps;
        uint16_t *frag;
        double *scales;
        if ((s = ws_typed(ws, 11, nblk * K, st, &exps)) != GNM_OK) return s;
        if ((s = ws_typed(ws, 12, nks * tiles * 64 * 8, st, &frag)) != GNM_OK) return s;
        if ((s = ws_typed(ws, 13, nblk * L * Kp, st, &scales)) != GNM_OK) return s;
        GNM_HIP(launch_fill_f64(scales, nblk * L * Kp, 0.0, st));
        GNM_HIP(launch_block_exponent(d_grad_scores, N, K, K, kFlushRows, L, false, exps, st));
        GNM_HIP(launch_adjoint_frag(d_grad_scores, N, K, Kp, L, exps, tiles, nks, frag, scales, st));
        GNM_HIP(launch_adjoint(d_rows, bytes_per_variant, V, N, d_flip, (const uint32_t *)frag, scales, tiles, K, Kp,
                               L, d_grad_weights, st));
        return GNM_OK;
    } catch (...) {
        return fail(GNM_ERR_INTERNAL, "exception in gnm_dense_adjoint_device");
    }
}

}  // extern "C"

namespace {

// Fixed device scratch of gnm_dense_adjoint_device_noalloc.  Every GPU the
// library runs on holds its own copy of this array from the first use of this
// file's code on that GPU until the process ends.  Host code never passes the
// array's name to a kernel (on the host it is only a stand-in for the symbol):
// its device address is looked up once per device with hipGetSymbolAddress
// and kept in g_scratch_ptr.
constexpr size_t kNoAllocScratchBytes = 96ull << 20;
constexpr int kNoAllocMaxDevices = 64;
}  // namespace

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
std::mutex g_scratch_mu;
void *g_scratch_ptr[kNoAllocMaxDevices] = {};

gnm_status scratch_for_current_device(void **out) {
    int dev = 0;
    GNM_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kNoAllocMaxDevices) return fail(GNM_ERR_UNSUPPORTED, "device ordinal above 63");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(GNM_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return GNM_OK;
}

uint64_t align256(uint64_t b) { return (b + 255) / 256 * 256; }

}  // namespace

extern "C" {

gnm_status gnm_dense_adjoint_device_noalloc(const uint8_t *d_rows, uint64_t bytes_per_variant,
                                            uint64_t num_variants, uint64_t num_people, const uint8_t *d_flip,
                                            const float *d_grad_scores, uint64_t num_scores, int32_t precision,
                                            float *d_grad_weights, void *stream_) {
    try {
        const uint64_t V = num_variants, N = num_people, K = num_scores;
        if (N == 0 || K == 0) return fail(GNM_ERR_INVALID_ARG, "num_people and num_scores must be positive");
        if (V && (!d_rows || !d_grad_weights)) return fail(GNM_ERR_INVALID_ARG, "NULL device pointer");
        if (!d_grad_scores) return fail(GNM_ERR_INVALID_ARG, "NULL grad_scores");
        if (bytes_per_variant < (N + 3) / 4) return fail(GNM_ERR_INVALID_ARG, "bytes_per_variant too small");
        if (!is_int8(precision))
            return fail(GNM_ERR_UNSUPPORTED, "the allocation-free adjoint runs in split-int8 (i8x3, i8x4) only");
        const bool aligned = ((reinterpret_cast<uintptr_t>(d_rows) | bytes_per_variant) & 3u) == 0 &&
                             bytes_per_variant < (1ull << 25);
        AdjointI8Shape sh;
        gnm_status s = adjoint_i8_shape(N, K, precision, aligned, &sh);
        if (s != GNM_OK) return s;
        const uint64_t b_exps = align256(sh.exps * sizeof(int));
        const uint64_t b_frag = align256(sh.frag * sizeof(uint32_t));
        const uint64_t b_scales = align256(sh.scales * sizeof(double));
        const uint64_t b_acc = align256(std::max<uint64_t>(V, 1) * sh.Kp * sizeof(double));
        const uint64_t need = b_exps + b_frag + b_scales + b_acc;
        if (need > kNoAllocScratchBytes)
            return fail(GNM_ERR_INVALID_ARG,
                        "problem needs " + std::to_string(need) + " bytes of scratch; the allocation-free adjoint "
                        "has " + std::to_string(kNoAllocScratchBytes) + " (about 8*V*round16(K) + "
                        "16*N*round4(limbs*round16(K)/16) bytes are needed)");
        if (V == 0) return GNM_OK;
        void *base = nullptr;
        if ((s = scratch_for_current_device(&base)) != GNM_OK) return s;
        unsigned char *p = (unsigned char *)base;
        int *exps = (int *)p;
        uint32_t *frag = (uint32_t *)(p + b_exps);
        double *scales = (double *)(p + b_exps + b_frag);
        double *acc = (double *)(p + b_exps + b_frag + b_scales);
        return adjoint_i8_launch(sh, d_rows, bytes_per_variant, V, N, d_flip, d_grad_scores, K, exps, frag, scales,
                                 acc, d_grad_weights, (hipStream_t)stream_);
    } catch (...) {
        return fail(GNM_ERR_INTERNAL, "exception in gnm_dense_adjoint_device_noalloc");
    }
}

}  // extern "C"
