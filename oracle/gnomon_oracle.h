/*
 * gnomon_oracle.h -- CPU oracle for the polygenic-scoring hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * (or the timed CPU baseline), never as the thing measured or shipped.  The
 * product library (gnomon_amd/csrc) never links or calls it.
 *
 * Contents (each function cites the reference code it restates):
 *   - counter-based synthetic genotype generator (integer-only, so the HIP
 *     generator in the product reproduces it bit for bit);
 *   - PLINK .bed decode (score/batch.rs:489-503, 812-844);
 *   - FP64 closed-form scorer with compensated summation (the "truth");
 *   - a restatement of the reference CPU engine (score/kernel.rs,
 *     score/batch.rs, score/decide.rs, score/pipeline.rs aggregation), used
 *     for arithmetic-class parity and as the timed CPU baseline;
 *   - FP64 adjoint G^T R for the differentiation path.
 *
 * Parity pinning: see tests/test_oracle_golden.py.  Every function here is
 * checked against the reference's own golden vectors and/or closed forms.
 */
#ifndef GNOMON_ORACLE_H
#define GNOMON_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- synthetic data ---------------------------------------------------- */
uint64_t orc_mix64(uint64_t z);
/* Packed SNP-major rows for variants [first_variant, first_variant+num_rows).
 * p_lo24/p_hi24: A2 allele frequency range in units of 2^-24; miss16: missing
 * probability in units of 2^-16.  Padding bits of the last byte are zero. */
void orc_synth_rows(uint64_t seed, uint32_t p_lo24, uint32_t p_hi24, uint32_t miss16,
                    uint64_t num_samples, uint64_t first_variant, uint64_t num_rows,
                    uint8_t *out);

/* ---- decode ------------------------------------------------------------ */
/* Per (row, person) A2 dosage tile value: 0,1,2 or 3 (= missing sentinel),
 * gathering fam index out_to_fam[p] (NULL = identity).  out is rows x people. */
void orc_decode_rows(const uint8_t *rows, uint64_t bpv, uint64_t num_rows,
                     const uint32_t *out_to_fam, uint64_t num_people, uint8_t *out);

/* ---- FP64 closed-form scorer ------------------------------------------ */
/* S[p,k] = sum over CSR entries (v,k) with d_pv != missing of corr + w*d,
 * miss[p,k] = #{CSR entries (v,k) with d_pv missing}.  Compensated sums. */
void orc_score_fp64(const uint8_t *rows, uint64_t bpv, uint64_t num_variants,
                    const uint32_t *out_to_fam, uint64_t num_people, uint64_t num_scores,
                    const float *w_eff, const float *corr, const uint32_t *cols,
                    const uint64_t *row_offsets, double *scores, uint32_t *counts);

/* ---- reference CPU engine restatement ---------------------------------- */
/* Dispatch tree restated from score/decide.rs:107-189 (0 = pivot/dense,
 * 1 = no-pivot/sparse). */
int orc_decide_path(float n_cohort, float k_scores, float subset_frac, float freq);
/* score/batch.rs:583-614 */
float orc_variant_density(const uint8_t *row, uint64_t bpv, uint64_t total_people);
/* Full engine: density dispatch, 256-variant dense batches with f32 kernel
 * accumulation, f64 sparse path, baseline + per-thread partials.  threads<=0
 * means hardware concurrency.  force_path: -1 dispatch, 0 all dense, 1 all
 * sparse.  Returns the number of threads used. */
int orc_refcpu_run(const uint8_t *rows, uint64_t bpv, uint64_t num_variants,
                   const uint32_t *out_to_fam, uint64_t num_people,
                   uint64_t total_people_in_fam, uint64_t num_scores,
                   const float *w_eff, const float *corr, const uint32_t *cols,
                   const uint64_t *row_offsets, const double *baseline,
                   int threads, int force_path, double *scores, uint32_t *counts);

/* ---- adjoint ------------------------------------------------------------ */
/* out[v,k] = sum_p g_pv * R[p,k] with g the effect-allele dosage under
 * flip[v] (0: A2 dosage, 1: 2 - A2 dosage), missing -> 0.  Compensated. */
void orc_adjoint_fp64(const uint8_t *rows, uint64_t bpv, uint64_t num_variants,
                      uint64_t num_people, const uint8_t *flip, const double *R,
                      uint64_t num_scores, double *out);
/* Dense forward in effect-allele space: S[p,k] = sum_v g_pv(flip[v]) W[v,k]
 * (compensated f64), miss[p] = number of missing genotypes of person p. */
void orc_dense_forward_fp64(const uint8_t *rows, uint64_t bpv, uint64_t num_variants,
                            uint64_t num_people, const uint8_t *flip, const float *W,
                            uint64_t num_scores, double *scores, uint32_t *missing);

#ifdef __cplusplus
}
#endif
#endif
