/*
 * gnomon_oracle.c -- CPU oracle for the polygenic-scoring hot path.
 *
 * TEST INFRASTRUCTURE ONLY (see gnomon_oracle.h): loaded by tests/, by
 * __graft_entry__.smoke() and by the cpu_baseline leg of bench.py, never by
 * the product library.
 *
 * Build: oracle/Makefile (gcc -O3 -march=native -ffp-contract=off -pthread).
 * -ffp-contract=off keeps every f32/f64 operation exactly as written; the one
 * fused multiply-add the reference kernel performs is written as fmaf().
 */
#define _GNU_SOURCE
#include "gnomon_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdatomic.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

/* ======================================================================= */
/* Synthetic genotypes: integer-only so the HIP generator matches bit for   */
/* bit.  Not a reference function: HWE draws as configured for the         */
/* benchmark workloads (A2 frequency p, P(d=0,1,2) = (1-p)^2, 2p(1-p), p^2, */
/* independent missingness).                                               */
/* ======================================================================= */

uint64_t orc_mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static inline uint64_t synth_variant_key(uint64_t seed, uint64_t v) {
    return orc_mix64(seed ^ orc_mix64(v + 0x632BE59BD9B4E019ull));
}

static inline uint32_t synth_p24(uint64_t kv, uint32_t p_lo24, uint32_t p_hi24) {
    uint64_t u24 = orc_mix64(kv ^ 0xA0761D6478BD642Full) >> 40; /* 24-bit uniform */
    return p_lo24 + (uint32_t)(((uint64_t)(p_hi24 - p_lo24) * u24) >> 24);
}

/* PLINK code of one genotype: 0b00 hom A1 (d=0), 0b10 het (d=1),
 * 0b11 hom A2 (d=2), 0b01 missing. */
static inline uint32_t synth_code(uint64_t kv, uint64_t s, uint32_t p24, uint32_t miss16) {
    uint64_t r = orc_mix64(kv + (s + 1) * 0x9E3779B97F4A7C15ull);
    if ((uint32_t)(r >> 48) < miss16) return 1u;
    uint64_t u48 = r & 0xFFFFFFFFFFFFull;
    uint64_t q24 = (1ull << 24) - p24;
    uint64_t t0 = q24 * q24;
    uint64_t t1 = t0 + 2ull * (uint64_t)p24 * q24;
    if (u48 < t0) return 0u;
    if (u48 < t1) return 2u;
    return 3u;
}

void orc_synth_rows(uint64_t seed, uint32_t p_lo24, uint32_t p_hi24, uint32_t miss16,
                    uint64_t num_samples, uint64_t first_variant, uint64_t num_rows,
                    uint8_t *out) {
    uint64_t bpv = (num_samples + 3) / 4;
    for (uint64_t r = 0; r < num_rows; ++r) {
        uint64_t v = first_variant + r;
        uint64_t kv = synth_variant_key(seed, v);
        uint32_t p24 = synth_p24(kv, p_lo24, p_hi24);
        uint8_t *row = out + r * bpv;
        for (uint64_t b = 0; b < bpv; ++b) {
            uint32_t byte = 0;
            for (uint32_t j = 0; j < 4; ++j) {
                uint64_t s = b * 4 + j;
                if (s < num_samples) byte |= synth_code(kv, s, p24, miss16) << (2 * j);
            }
            row[b] = (uint8_t)byte;
        }
    }
}

/* ======================================================================= */
/* Decode.  Sample j of a row sits in byte j>>2 at bits 2(j&3)             */
/* (score/batch.rs:797-844).  Tile value per score/batch.rs:489-503:        */
/* dosage = hi*(lo+1), missing (0b01) -> sentinel 3.                       */
/* ======================================================================= */

static inline uint32_t code_at(const uint8_t *row, uint64_t fam) {
    return (row[fam >> 2] >> (2 * (fam & 3))) & 3u;
}

static inline uint8_t tile_value(uint32_t code) {
    uint32_t lo = code & 1u, hi = (code >> 1) & 1u;
    uint32_t dosage = hi * (lo + 1u);
    uint32_t is_missing = (1u - hi) * lo;
    return (uint8_t)(dosage + 3u * is_missing);
}

void orc_decode_rows(const uint8_t *rows, uint64_t bpv, uint64_t num_rows,
                     const uint32_t *out_to_fam, uint64_t num_people, uint8_t *out) {
    for (uint64_t r = 0; r < num_rows; ++r) {
        const uint8_t *row = rows + r * bpv;
        for (uint64_t p = 0; p < num_people; ++p) {
            uint64_t fam = out_to_fam ? out_to_fam[p] : p;
            out[r * num_people + p] = tile_value(code_at(row, fam));
        }
    }
}

/* ======================================================================= */
/* FP64 closed form (Neumaier-compensated).                                */
/* Per CSR entry (v,k): contribution corr + w_eff*d for d in {0,1,2};      */
/* a missing genotype contributes 0 (the reference adds the baseline       */
/* sum(corr) and subtracts corr per missing call: score/prepare.rs:197-208, */
/* 632-650; score/batch.rs:372-385; score/pipeline.rs:1739-1775).           */
/* ======================================================================= */

static inline void neumaier_add(double *s, double *c, double x) {
    double t = *s + x;
    if (fabs(*s) >= fabs(x))
        *c += (*s - t) + x;
    else
        *c += (x - t) + *s;
    *s = t;
}

void orc_score_fp64(const uint8_t *rows, uint64_t bpv, uint64_t num_variants,
                    const uint32_t *out_to_fam, uint64_t num_people, uint64_t num_scores,
                    const float *w_eff, const float *corr, const uint32_t *cols,
                    const uint64_t *row_offsets, double *scores, uint32_t *counts) {
    size_t cells = (size_t)num_people * num_scores;
    double *sum = calloc(cells, sizeof(double));
    double *comp = calloc(cells, sizeof(double));
    memset(counts, 0, cells * sizeof(uint32_t));
    for (uint64_t v = 0; v < num_variants; ++v) {
        const uint8_t *row = rows + v * bpv;
        uint64_t e0 = row_offsets[v], e1 = row_offsets[v + 1];
        if (e0 == e1) continue;
        for (uint64_t p = 0; p < num_people; ++p) {
            uint64_t fam = out_to_fam ? out_to_fam[p] : p;
            uint8_t t = tile_value(code_at(row, fam));
            size_t base = (size_t)p * num_scores;
            for (uint64_t e = e0; e < e1; ++e) {
                size_t cell = base + cols[e];
                if (t == 3) {
                    counts[cell] += 1;
                } else {
                    double x = (double)corr[e] + (double)w_eff[e] * (double)t;
                    neumaier_add(&sum[cell], &comp[cell], x);
                }
            }
        }
    }
    for (size_t i = 0; i < cells; ++i) scores[i] = sum[i] + comp[i];
    free(sum);
    free(comp);
}

/* ======================================================================= */
/* Reference CPU engine restatement.                                       */
/* ======================================================================= */

enum { PATH_PIVOT = 0, PATH_NOPIVOT = 1 };

/* Same thresholds and outcomes as score/decide.rs:107-189
 * (decide_path_with_freq); the strategy is fixed to the frequency-aware
 * tree (score/pipeline.rs:534). */
int orc_decide_path(float n, float k, float sf, float f) {
    const int tiny_subset = (n <= 10.0000f) && (sf <= 0.2750f);
    if (n <= 20000.002f) {
        if (f <= 0.0894f) {
            if (f <= 0.0045f) return tiny_subset ? PATH_PIVOT : PATH_NOPIVOT;
            if (n <= 7071.067f) return tiny_subset ? PATH_PIVOT : PATH_NOPIVOT;
            if (sf <= 0.7500f) return PATH_NOPIVOT;
            return (k <= 15.8114f) ? PATH_PIVOT : PATH_NOPIVOT;
        }
        if (n <= 7071.067f) {
            if (k <= 15.8114f) {
                if (sf <= 0.7500f) return tiny_subset ? PATH_PIVOT : PATH_NOPIVOT;
                return (n <= 2236.0677f) ? PATH_NOPIVOT : PATH_PIVOT;
            }
            return tiny_subset ? PATH_PIVOT : PATH_NOPIVOT;
        }
        if (sf <= 0.7500f) {
            if (k <= 15.8114f) return (sf <= 0.2750f) ? PATH_NOPIVOT : PATH_PIVOT;
            return PATH_NOPIVOT;
        }
        return PATH_PIVOT;
    }
    if (sf <= 0.2750f) return PATH_NOPIVOT;
    return (f <= 0.0045f) ? PATH_NOPIVOT : PATH_PIVOT;
}

/* score/batch.rs:583-614: popcount of the packed row over 8-byte words,
 * early exit to 1.0 once above the highest dispatch threshold. */
float orc_variant_density(const uint8_t *row, uint64_t bpv, uint64_t total_people) {
    if (total_people == 0) return 0.0f;
    uint64_t cutoff = (uint64_t)((float)total_people * 0.0894f);
    uint64_t bits = 0, w = 0;
    for (; w + 8 <= bpv; w += 8) {
        uint64_t x;
        memcpy(&x, row + w, 8);
        bits += (uint64_t)__builtin_popcountll(x);
        if (bits > cutoff) return 1.0f;
    }
    for (; w < bpv; ++w) bits += (uint64_t)__builtin_popcount(row[w]);
    return (float)bits / (float)total_people;
}

typedef struct {
    const uint8_t *rows;
    uint64_t bpv, num_people, num_scores, stride;
    const uint32_t *out_to_fam;
    const float *w_eff, *corr;
    const uint32_t *cols;
    const uint64_t *row_offsets;
    /* work list: items < n_batches are dense batches, the rest sparse variants */
    const uint64_t *dense_list;
    uint64_t n_dense, n_batches;
    const uint64_t *sparse_list;
    uint64_t n_sparse;
    atomic_uint_fast64_t next;
    double **part_scores;
    uint32_t **part_counts;
} refcpu_job;

enum { DENSE_BATCH = 256, PERSON_BLOCK = 4096, SCORE_CHUNK = 64 };

/* score/batch.rs:720-844 (all people) and :622-711 (subset): per genotype,
 * f64 arithmetic directly into the thread partial. */
static void refcpu_sparse_variant(const refcpu_job *J, uint64_t v, double *S, uint32_t *C) {
    const uint8_t *row = J->rows + v * J->bpv;
    uint64_t e0 = J->row_offsets[v], e1 = J->row_offsets[v + 1];
    for (uint64_t p = 0; p < J->num_people; ++p) {
        uint64_t fam = J->out_to_fam ? J->out_to_fam[p] : p;
        uint32_t g = code_at(row, fam);
        if (g == 0) continue;
        size_t base = (size_t)p * J->num_scores;
        if (g == 1) {
            for (uint64_t e = e0; e < e1; ++e) {
                C[base + J->cols[e]] += 1;
                S[base + J->cols[e]] -= (double)J->corr[e];
            }
        } else {
            double dosage = (g == 2) ? 1.0 : 2.0;
            for (uint64_t e = e0; e < e1; ++e) S[base + J->cols[e]] += (double)J->w_eff[e] * dosage;
        }
    }
}

/* score/batch.rs:48-422 + score/kernel.rs:124-238 for one dense batch of up
 * to 256 variants: person-major tile, g1/g2/missing lists, f32 kernel per
 * 64-score chunk (adds for dosage 1, fma(w,2,acc) for dosage 2), f64 widen. */
static void refcpu_dense_batch(const refcpu_job *J, const uint64_t *vars, uint64_t nb,
                               double *S, uint32_t *C, uint8_t *tile, float *Wb, float *Cb) {
    const uint64_t K = J->num_scores, stride = J->stride;
    memset(Wb, 0, nb * stride * sizeof(float));
    memset(Cb, 0, nb * stride * sizeof(float));
    for (uint64_t i = 0; i < nb; ++i) {
        uint64_t v = vars[i];
        for (uint64_t e = J->row_offsets[v]; e < J->row_offsets[v + 1]; ++e) {
            Wb[i * stride + J->cols[e]] = J->w_eff[e];
            Cb[i * stride + J->cols[e]] = J->corr[e];
        }
    }
    uint16_t g1[DENSE_BATCH], g2[DENSE_BATCH], gm[DENSE_BATCH];
    float acc[SCORE_CHUNK];
    for (uint64_t pb = 0; pb < J->num_people; pb += PERSON_BLOCK) {
        uint64_t np = J->num_people - pb < PERSON_BLOCK ? J->num_people - pb : PERSON_BLOCK;
        /* pivot: tile[person][variant] */
        for (uint64_t i = 0; i < nb; ++i) {
            const uint8_t *row = J->rows + vars[i] * J->bpv;
            for (uint64_t q = 0; q < np; ++q) {
                uint64_t p = pb + q;
                uint64_t fam = J->out_to_fam ? J->out_to_fam[p] : p;
                tile[q * nb + i] = tile_value(code_at(row, fam));
            }
        }
        for (uint64_t q = 0; q < np; ++q) {
            const uint8_t *drow = tile + q * nb;
            int n1 = 0, n2 = 0, nm = 0;
            for (uint64_t i = 0; i < nb; ++i) {
                uint8_t d = drow[i];
                if (d == 1) g1[n1++] = (uint16_t)i;
                else if (d == 2) g2[n2++] = (uint16_t)i;
                else if (d == 3) gm[nm++] = (uint16_t)i;
            }
            size_t base = (size_t)(pb + q) * K;
            for (int m = 0; m < nm; ++m) {
                uint64_t v = vars[gm[m]];
                for (uint64_t e = J->row_offsets[v]; e < J->row_offsets[v + 1]; ++e) {
                    uint32_t col = J->cols[e];
                    C[base + col] += 1;
                    S[base + col] -= (double)Cb[(uint64_t)gm[m] * stride + col];
                }
            }
            for (uint64_t c0 = 0; c0 < K; c0 += SCORE_CHUNK) {
                uint64_t clen = K - c0 < SCORE_CHUNK ? K - c0 : SCORE_CHUNK;
                uint64_t lanes = (clen + 7) / 8 * 8;
                for (uint64_t j = 0; j < lanes; ++j) acc[j] = 0.0f;
                for (int a = 0; a < n1; ++a) {
                    const float *w = Wb + (uint64_t)g1[a] * stride + c0;
                    for (uint64_t j = 0; j < lanes; ++j) acc[j] = acc[j] + w[j];
                }
                for (int a = 0; a < n2; ++a) {
                    const float *w = Wb + (uint64_t)g2[a] * stride + c0;
                    for (uint64_t j = 0; j < lanes; ++j) acc[j] = fmaf(w[j], 2.0f, acc[j]);
                }
                for (uint64_t j = 0; j < clen; ++j) S[base + c0 + j] += (double)acc[j];
            }
        }
    }
}

static void *refcpu_worker(void *arg) {
    refcpu_job *J = ((refcpu_job **)arg)[0];
    uint64_t tid = (uint64_t)(uintptr_t)((void **)arg)[1];
    double *S = J->part_scores[tid];
    uint32_t *C = J->part_counts[tid];
    uint8_t *tile = malloc((size_t)PERSON_BLOCK * DENSE_BATCH);
    float *Wb = malloc((size_t)DENSE_BATCH * J->stride * sizeof(float));
    float *Cb = malloc((size_t)DENSE_BATCH * J->stride * sizeof(float));
    uint64_t total = J->n_batches + J->n_sparse;
    for (;;) {
        uint64_t it = atomic_fetch_add(&J->next, 1);
        if (it >= total) break;
        if (it < J->n_batches) {
            uint64_t s = it * DENSE_BATCH;
            uint64_t nb = J->n_dense - s < DENSE_BATCH ? J->n_dense - s : DENSE_BATCH;
            refcpu_dense_batch(J, J->dense_list + s, nb, S, C, tile, Wb, Cb);
        } else {
            refcpu_sparse_variant(J, J->sparse_list[it - J->n_batches], S, C);
        }
    }
    free(tile);
    free(Wb);
    free(Cb);
    return NULL;
}

int orc_refcpu_run(const uint8_t *rows, uint64_t bpv, uint64_t num_variants,
                   const uint32_t *out_to_fam, uint64_t num_people,
                   uint64_t total_people_in_fam, uint64_t num_scores,
                   const float *w_eff, const float *corr, const uint32_t *cols,
                   const uint64_t *row_offsets, const double *baseline,
                   int threads, int force_path, double *scores, uint32_t *counts) {
    if (threads <= 0) threads = (int)sysconf(_SC_NPROCESSORS_ONLN);
    if (threads < 1) threads = 1;
    refcpu_job J;
    memset(&J, 0, sizeof(J));
    J.rows = rows;
    J.bpv = bpv;
    J.num_people = num_people;
    J.num_scores = num_scores;
    J.stride = (num_scores + 7) / 8 * 8;
    J.out_to_fam = out_to_fam;
    J.w_eff = w_eff;
    J.corr = corr;
    J.cols = cols;
    J.row_offsets = row_offsets;
    uint64_t *dense = malloc((num_variants + 1) * sizeof(uint64_t));
    uint64_t *sparse = malloc((num_variants + 1) * sizeof(uint64_t));
    uint64_t nd = 0, ns = 0;
    /* producer-side routing (score/pipeline.rs:697-707) */
    float n_cohort = (float)total_people_in_fam;
    float subset_frac = (float)num_people / (float)total_people_in_fam;
    for (uint64_t v = 0; v < num_variants; ++v) {
        int path = force_path;
        if (path < 0) {
            float f = orc_variant_density(rows + v * bpv, bpv, total_people_in_fam);
            path = orc_decide_path(n_cohort, (float)num_scores, subset_frac, f);
        }
        if (path == PATH_PIVOT) dense[nd++] = v;
        else sparse[ns++] = v;
    }
    J.dense_list = dense;
    J.n_dense = nd;
    J.n_batches = (nd + DENSE_BATCH - 1) / DENSE_BATCH;
    J.sparse_list = sparse;
    J.n_sparse = ns;
    atomic_init(&J.next, 0);
    size_t cells = (size_t)num_people * num_scores;
    J.part_scores = malloc(threads * sizeof(double *));
    J.part_counts = malloc(threads * sizeof(uint32_t *));
    pthread_t *tids = malloc(threads * sizeof(pthread_t));
    void ***args = malloc(threads * sizeof(void **));
    for (int t = 0; t < threads; ++t) {
        J.part_scores[t] = calloc(cells, sizeof(double));
        J.part_counts[t] = calloc(cells, sizeof(uint32_t));
        args[t] = malloc(2 * sizeof(void *));
        args[t][0] = &J;
        args[t][1] = (void *)(uintptr_t)t;
        pthread_create(&tids[t], NULL, refcpu_worker, args[t]);
    }
    for (int t = 0; t < threads; ++t) pthread_join(tids[t], NULL);
    /* score/pipeline.rs:1739-1775, 771-796: baseline per person, then the
     * per-thread partials (the reference reduces them pairwise in a
     * scheduling-dependent order; here in thread order). */
    for (uint64_t p = 0; p < num_people; ++p)
        for (uint64_t k = 0; k < num_scores; ++k) scores[p * num_scores + k] = baseline[k];
    memset(counts, 0, cells * sizeof(uint32_t));
    for (int t = 0; t < threads; ++t) {
        for (size_t i = 0; i < cells; ++i) {
            scores[i] += J.part_scores[t][i];
            counts[i] += J.part_counts[t][i];
        }
        free(J.part_scores[t]);
        free(J.part_counts[t]);
        free(args[t]);
    }
    free(args);
    free(tids);
    free(J.part_scores);
    free(J.part_counts);
    free(dense);
    free(sparse);
    return threads;
}

/* ======================================================================= */
/* Effect-allele space forward and adjoint (differentiation path; not a    */
/* reference feature).  g = d (flip 0) or 2-d (flip 1); missing -> 0.      */
/* ======================================================================= */

static inline double effect_dosage(uint32_t code, int flip) {
    uint8_t t = tile_value(code);
    if (t == 3) return 0.0;
    return flip ? 2.0 - (double)t : (double)t;
}

void orc_dense_forward_fp64(const uint8_t *rows, uint64_t bpv, uint64_t num_variants,
                            uint64_t num_people, const uint8_t *flip, const float *W,
                            uint64_t num_scores, double *scores, uint32_t *missing) {
    size_t cells = (size_t)num_people * num_scores;
    double *comp = calloc(cells, sizeof(double));
    memset(scores, 0, cells * sizeof(double));
    memset(missing, 0, num_people * sizeof(uint32_t));
    for (uint64_t v = 0; v < num_variants; ++v) {
        const uint8_t *row = rows + v * bpv;
        for (uint64_t p = 0; p < num_people; ++p) {
            uint32_t code = code_at(row, p);
            if (code == 1) {
                missing[p] += 1;
                continue;
            }
            double g = effect_dosage(code, flip ? flip[v] : 0);
            if (g == 0.0) continue;
            for (uint64_t k = 0; k < num_scores; ++k)
                neumaier_add(&scores[p * num_scores + k], &comp[p * num_scores + k],
                             g * (double)W[v * num_scores + k]);
        }
    }
    for (size_t i = 0; i < cells; ++i) scores[i] += comp[i];
    free(comp);
}

void orc_adjoint_fp64(const uint8_t *rows, uint64_t bpv, uint64_t num_variants,
                      uint64_t num_people, const uint8_t *flip, const double *R,
                      uint64_t num_scores, double *out) {
    double *s = malloc(num_scores * sizeof(double));
    double *c = malloc(num_scores * sizeof(double));
    for (uint64_t v = 0; v < num_variants; ++v) {
        const uint8_t *row = rows + v * bpv;
        for (uint64_t k = 0; k < num_scores; ++k) s[k] = c[k] = 0.0;
        for (uint64_t p = 0; p < num_people; ++p) {
            double g = effect_dosage(code_at(row, p), flip ? flip[v] : 0);
            if (g == 0.0) continue;
            for (uint64_t k = 0; k < num_scores; ++k)
                neumaier_add(&s[k], &c[k], g * R[p * num_scores + k]);
        }
        for (uint64_t k = 0; k < num_scores; ++k) out[v * num_scores + k] = s[k] + c[k];
    }
    free(s);
    free(c);
}
