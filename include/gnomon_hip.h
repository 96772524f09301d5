/*
 * gnomon_hip.h -- C ABI of the MI355X (gfx950) polygenic-scoring backend.
 *
 * This library is a drop-in for what sits behind gnomon's GPU entry point
 *   score/cuda_backend.rs:604  try_run_cuda(&PipelineContext)
 *       -> Result<Option<(Vec<f64>, Vec<u32>)>, PipelineError>
 * called from score/pipeline.rs:480.  Every entry point below names the
 * reference code it replaces.  No C++ exception, abort or allocation crosses
 * the ABI; all pointers are plain host or device addresses and sizes.
 *
 * Semantics (identical to the reference):
 *   scores[p*K + k] = baseline[k] + sum_v W_eff[v,k]*d_pv - sum_{v: missing} corr[v,k]
 *                   = sum over CSR entries (v,k) with d_pv not missing of corr + W_eff*d
 *   counts[p*K + k] = number of CSR entries (v,k) whose genotype for p is missing
 * with d the A2-allele dosage decoded from the SNP-major PLINK .bed row
 * (sample j: byte j>>2, bits 2(j&3); 00->0, 10->1, 11->2, 01->missing).
 * Output layout is [person][score], persons in output order
 * (output_idx_to_fam_idx), as score/cuda_backend.rs:1674-1679, 2063.
 */
#ifndef GNOMON_HIP_H
#define GNOMON_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GNM_ABI_VERSION 1

/* Replaces PipelineError (shared/pipeline_error.rs:5-10) plus the Ok(None)
 * fallback of try_run_cuda (cuda_backend.rs:608-634). */
typedef enum gnm_status {
    GNM_OK = 0,
    GNM_DECLINED = 1,          /* below the device size gate: caller runs the CPU path */
    GNM_ERR_NO_DEVICE = 2,     /* no usable gfx950 device, or devices hidden by env */
    GNM_ERR_INVALID_ARG = 3,
    GNM_ERR_OUT_OF_MEMORY = 4,
    GNM_ERR_HIP = 5,           /* a HIP runtime call failed after start */
    GNM_ERR_INTERNAL = 6,
    GNM_ERR_UNSUPPORTED = 7
} gnm_status;

/* Arithmetic of the genotype x weight contraction.
 *   FP32:   plain FP32-weight wavefront reduction on the vector ALUs: FP32
 *           weights, exact dosage x weight products, FP64 accumulation per
 *           genotype (no split, no matrix cores).  Replaces the reference's
 *           FP32 kernels (score/kernel.rs:124-176, cuda_backend.rs:2554-2584),
 *           whose FP32 partial sums it makes exact.
 *   BF16X2/BF16X3/BF16X4: split-BF16 MFMA.  Each weight column is written
 *           as 2, 3 or 4 integer-valued BF16 limbs (signed base-256 digits)
 *           times a power-of-two scale per (variant block, column); dosages
 *           0/1/2 are exact in BF16, so every product and every FP32 partial
 *           sum inside a variant block is exact; blocks are combined in FP64.
 *           BF16X4 represents 32 bits of each block maximum, so every FP32
 *           weight within 2^8 of its block maximum is represented exactly
 *           and the rest to 2^-32 of that maximum; BF16X3 24 bits, BF16X2 16.
 *           Replaces the split-FP16 tensor-core path
 *           (cuda_backend.rs:2360-2415).
 *   I8X3/I8X4: split-int8 MFMA (v_mfma_i32_16x16x64_i8).  Each weight column
 *           is written as 3 or 4 signed 8-bit integer limbs times a power-of-two
 *           scale per (variant block, column); products and int32 block sums
 *           are exact; blocks are combined in FP64.  I8X4 represents 31 bits
 *           of each block maximum, I8X3 23 bits.  Same role as the split paths
 *           above (no reference counterpart for the int8 limbs themselves).
 *   AUTO:   I8X4 once a known-answer check of that path has passed on the
 *           device (the calibrate_fp16 inputs of cuda_backend.rs:2424-2494,
 *           checked against exact host sums; run once per device and
 *           process), else FP32 -- the reference likewise enables its split
 *           path only after its self-test (cuda_backend.rs:1176-1196).
 *   Weights and corrections must be finite (GNM_ERR_INVALID_ARG otherwise),
 *   as the reference refuses its split path for non-finite weights
 *   (cuda_backend.rs:1178-1181). */
typedef enum gnm_precision {
    GNM_PRECISION_FP32 = 0,
    GNM_PRECISION_BF16X2 = 1,
    GNM_PRECISION_BF16X3 = 2,
    GNM_PRECISION_AUTO = 3,
    GNM_PRECISION_I8X3 = 4,
    GNM_PRECISION_I8X4 = 5,
    GNM_PRECISION_BF16X4 = 6
} gnm_precision;

/* The slice of PreparationResult (score/types.rs:166-365) the hot path reads. */
typedef struct gnm_problem {
    uint64_t num_people;            /* num_people_to_score (N) */
    uint64_t total_people_in_fam;   /* .fam rows (N_fam) */
    uint64_t bytes_per_variant;     /* ceil(N_fam / 4) (prepare.rs:811) */
    uint64_t num_scores;            /* score_names.len() (K) */
    uint64_t num_variants;          /* num_reconciled_variants (V, CSR rows) */
    uint64_t nnz;                   /* sparse_row_offsets[V] */
    const float *sparse_weights;              /* [nnz] W_eff, A2-dosage space */
    const float *sparse_missing_corrections;  /* [nnz] corr */
    const uint32_t *sparse_score_columns;     /* [nnz] column < K */
    const uint64_t *sparse_row_offsets;       /* [V+1] */
    const uint32_t *output_idx_to_fam_idx;    /* [N]; NULL = identity */
    int32_t precision;              /* gnm_precision */
    int32_t device;                 /* HIP ordinal; <0 = 0 */
    uint64_t reserved[4];           /* must be zero */
} gnm_problem;

typedef struct gnm_ctx gnm_ctx;

/* Thread-local description of the last failure on this thread. */
const char *gnm_last_error(void);
int gnm_abi_version(void);

/* Size gate of try_run_cuda (cuda_backend.rs:38-39, 608-617): 1 when
 * N*K >= 1e5 or N*K*V >= 1e8, else 0. */
int gnm_size_gate(uint64_t num_people, uint64_t num_scores, uint64_t num_variants);

/* Device probe, replaces cuda_driver_likely_available (cuda_backend.rs:675-695):
 * 0 when HIP_VISIBLE_DEVICES / ROCR_VISIBLE_DEVICES is "", "-1" or "none",
 * or when no gfx950 device is present; else the number of devices. */
int gnm_device_count(void);

/* Replaces CudaRuntime::new (cuda_backend.rs:994-1235): validates the CSR,
 * builds the device weight plan (flip-aware effect-allele rows, split limbs,
 * membership classes), allocates accumulators.  GNM_DECLINED when below the
 * size gate (no context is created). */
gnm_status gnm_create(const gnm_problem *problem, gnm_ctx **out_ctx);
/* Same, skipping the size gate (tests and explicit GPU requests). */
gnm_status gnm_create_forced(const gnm_problem *problem, gnm_ctx **out_ctx);

/* Replaces the mega-batch body of process_dense_stream_cuda /
 * run_pending_compute_cuda (cuda_backend.rs:1827-1978, 2067-2333).
 * rows: num_rows packed .bed rows (bytes_per_variant each, row stride given)
 * for reconciled variants [first_variant, first_variant + num_rows), in
 * ascending order, on the host.  Blocks follow each other without gaps or
 * repeats: first_variant must be the number of variants completed so far
 * (after gnm_seed_accumulators, the seeded count), or 0 once all
 * num_variants are done (a further pass then accumulates on top);
 * GNM_ERR_INVALID_ARG otherwise.  The call copies the rows through pinned
 * staging and returns once they have been consumed (the caller may reuse
 * the buffer); compute stays asynchronous. */
gnm_status gnm_score_rows(gnm_ctx *ctx, const uint8_t *rows, uint64_t row_stride,
                          uint64_t first_variant, uint64_t num_rows);
/* Same with rows already resident in device memory (device-resident .bed). */
gnm_status gnm_score_rows_device(gnm_ctx *ctx, const uint8_t *d_rows, uint64_t row_stride,
                                 uint64_t first_variant, uint64_t num_rows);

/* Sparse ingest of gnm_score_rows, the device counterpart of the CPU
 * engine's density dispatch (assess_variant_density_for_dispatch and
 * run_variant_major_path, score/batch.rs:583-760; path choice
 * score/decide.rs:100-189).  A host row in which at most max_word_fraction
 * of the 64-bit words (32 calls each) hold a non-zero call of a scored
 * person crosses PCIe as a bitmap of those words plus the words, and the
 * device rebuilds the packed row before scoring it; rows of variants in no
 * score are not sent.  Results are bit-identical to sending every row
 * packed.  max_word_fraction in [0, 1/2]; 0 sends every row packed.
 * Default 1/2.  After a staging slot in which every row went packed, the
 * next 7 slots are sent packed without looking (dense regions cost no
 * classification); every eighth slot looks again.  Device-resident rows (gnm_score_rows_device) are read as
 * they are. */
gnm_status gnm_set_sparse_ingest(gnm_ctx *ctx, double max_word_fraction);
/* Totals since gnm_create of the rows gnm_score_rows sent packed, the rows
 * sent sparse, their stored words, and the bytes copied to the device. */
gnm_status gnm_ingest_stats(gnm_ctx *ctx, uint64_t *dense_rows, uint64_t *sparse_rows, uint64_t *stored_words,
                            uint64_t *bytes_to_device);
/* The host encoder of the sparse ingest (host only, no device use): the
 * non-zero 64-bit words of one packed row of num_fam people (keep2: 2-bit
 * mask per person, 32 people per word, NULL = all; padding calls of the
 * last byte are dropped) in row order, with a bitmap of them, 8 words of
 * bitmap per chunk of 512 row words, and chunk_start[c] = index of the first
 * stored word of chunk c, for c = 0 .. ceil(num_fam / 16384).  Returns the
 * number of stored words, -1 when there are more than max_words (the row
 * then goes packed), -2 for NULL buffers.  words must hold max_words + 64
 * values. */
int64_t gnm_sparse_encode_row(const uint8_t *row, uint64_t num_fam, const uint64_t *keep2, uint64_t max_words,
                              uint64_t *words, uint64_t *bitmap, uint32_t *chunk_start);

/* Checkpoint contract (score/checkpoint.rs:293-339, cuda_backend.rs:1969-2011):
 * exact f64 sums (baseline included, as the reference stores them) and u32
 * counts for variants [0, completed_variants). */
gnm_status gnm_read_accumulators(gnm_ctx *ctx, double *scores, uint32_t *counts,
                                 uint64_t *completed_variants);
/* Resume (pipeline.rs:1777-1799, cuda_backend.rs:1680-1694): seed the
 * accumulators from a checkpoint taken after completed_variants variants. */
gnm_status gnm_seed_accumulators(gnm_ctx *ctx, const double *scores, const uint32_t *counts,
                                 uint64_t completed_variants);

/* Final result: waits for all work and copies N*K f64 scores (baseline
 * included) and u32 missing counts to host buffers owned by the caller. */
gnm_status gnm_finish(gnm_ctx *ctx, double *scores, uint32_t *counts);

/* Synchronises and frees every device resource in a fixed order
 * (cuda_backend.rs:2014-2058).  NULL is accepted. */
void gnm_destroy(gnm_ctx *ctx);

/* One-shot form of try_run_cuda over a host .bed image: rows for all V
 * reconciled variants (row_stride apart).  GNM_DECLINED = caller falls back. */
gnm_status gnm_run(const gnm_problem *problem, const uint8_t *rows, uint64_t row_stride,
                   double *scores, uint32_t *counts);

/* ---- device-pointer entry points (PyTorch-ROCm autograd path, tests) ------
 * All pointers are device addresses; stream is a hipStream_t (NULL = default).
 * The .bed block is num_variants rows of bytes_per_variant, persons
 * [0, num_people) of each row (identity person map). */

/* Standalone unpack: dosage[v*num_people + p] = reference tile value
 * (score/batch.rs:489-503): effect-allele dosage 0/1/2 (2-d when flip[v]),
 * 3 = missing; missing_per_person[p] += number of missing calls in the block
 * (accumulated, caller zeroes).  flip may be NULL.  num_people at most
 * 1073725440.  Replaces unpack_plink (cuda_backend.rs:70-125) as a separate
 * stage. */
gnm_status gnm_unpack_device(const uint8_t *d_rows, uint64_t bytes_per_variant,
                             uint64_t num_variants, uint64_t num_people, const uint8_t *d_flip,
                             uint8_t *d_dosage, uint32_t *d_missing_per_person, void *stream);

/* Shape of the context's scoring launch, for roofline accounting (no
 * reference counterpart): program rows (variants with CSR entries; a variant
 * whose cells mix flipped and unflipped effect alleles counts twice) and the
 * 16-column MFMA tiles each row is multiplied by (weight limbs plus missing
 * counts; 0 on the FP32 path).  One pass over all rows executes
 * 2 * num_people * program_rows * 16 * mfma_tiles matrix operations. */
gnm_status gnm_context_shape(const gnm_ctx *ctx, uint64_t *program_rows, uint32_t *mfma_tiles);

/* HBM ceiling probes for the unpack stage's roofline (no reference
 * counterpart: measurement only).  kind 0 reads src_bytes from d_src (output:
 * one dword per probe thread, d_dst of at least 16 MiB); kind 1 writes
 * src_bytes to d_dst with the unpack stage's non-temporal 16-byte stores
 * (d_src unused); kind 2 expands every source dword into 16 output bytes
 * (reads src_bytes, writes 4 * src_bytes), the unpack stage's 1:4 ratio.
 * src_bytes: a positive multiple of 16; pointers 16-byte aligned. */
gnm_status gnm_hbm_probe(int kind, const void *d_src, uint64_t src_bytes, void *d_dst, uint64_t dst_bytes,
                         void *stream);

/* Dense effect-allele forward, S[p,k] = sum_v g_pv W[v,k] with
 * g = d (flip 0) or 2-d (flip 1), missing -> 0; missing_per_person[p] =
 * number of missing calls.  W is [num_variants][num_scores] f32, S is
 * [num_people][num_scores] f64 (overwritten).  A NaN or infinity in W
 * makes the affected columns of S NaN.  New (no reference counterpart): the
 * forward of the differentiable path. */
gnm_status gnm_dense_forward_device(const uint8_t *d_rows, uint64_t bytes_per_variant,
                                    uint64_t num_variants, uint64_t num_people,
                                    const uint8_t *d_flip, const float *d_weights,
                                    uint64_t num_scores, int32_t precision, double *d_scores,
                                    uint32_t *d_missing_per_person, void *stream);

/* Adjoint dL/dW = G^T dL/dS, recomputing G from the .bed (no stored
 * dosages).  grad_scores [num_people][num_scores] f32, grad_weights
 * [num_variants][num_scores] f32 (overwritten).  precision: BF16X2/BF16X3/
 * BF16X4 split dL/dS into BF16 limbs per block of 8192 people (up to 12 limb
 * tiles: 64 scores in BF16X3, 48 in BF16X4); I8X3/I8X4 into int8 limbs per
 * block of 65536 people (rows 4-byte aligned, bytes_per_variant a multiple
 * of 4 below 2^25; up to 64 scores); AUTO = I8X4 where the rows allow it and
 * the split self-test passed, else BF16X3.  A NaN or infinity in
 * grad_scores makes the affected columns of grad_weights NaN.  New (no
 * reference counterpart). */
gnm_status gnm_dense_adjoint_device(const uint8_t *d_rows, uint64_t bytes_per_variant,
                                    uint64_t num_variants, uint64_t num_people,
                                    const uint8_t *d_flip, const float *d_grad_scores,
                                    uint64_t num_scores, int32_t precision,
                                    float *d_grad_weights, void *stream);

/* gnm_dense_adjoint_device without any device allocation, for callers that
 * cannot allocate (code under HIP graph capture, tight serving or training
 * loops, harnesses that forbid hipMalloc).  No counterpart in the reference.
 * Split-int8 only (I8X3 / I8X4; other precisions GNM_ERR_UNSUPPORTED), same
 * operands, kernels and results as gnm_dense_adjoint_device at that
 * precision.  Its scratch (block exponents, R limb fragments, limb scales and
 * an FP64 [V][round16(K)] accumulator) lives in a fixed 96 MiB device array
 * per GPU, held from the library's first use on that GPU until the process
 * ends; problems needing more (about 8*V*round16(K) bytes plus 16 bytes per
 * person per 16-column limb tile) get GNM_ERR_INVALID_ARG before any launch.
 * Every call on one GPU shares that array: calls that overlap on one GPU
 * (different streams or threads without ordering) race. */
gnm_status gnm_dense_adjoint_device_noalloc(const uint8_t *d_rows, uint64_t bytes_per_variant,
                                            uint64_t num_variants, uint64_t num_people,
                                            const uint8_t *d_flip, const float *d_grad_scores,
                                            uint64_t num_scores, int32_t precision,
                                            float *d_grad_weights, void *stream);

/* Synthetic SNP-major rows on the device, bit-identical to the test oracle's
 * generator (variant v, sample s keyed counter-based hash; HWE with A2
 * frequency in [p_lo24, p_hi24] * 2^-24; missing with probability
 * miss16 * 2^-16).  Benchmark/test utility, not a reference function. */
gnm_status gnm_synth_rows_device(uint64_t seed, uint32_t p_lo24, uint32_t p_hi24,
                                 uint32_t miss16, uint64_t num_samples, uint64_t first_variant,
                                 uint64_t num_rows, uint8_t *d_out, uint64_t row_stride,
                                 void *stream);

/* Launch shape of the split-int8 scoring kernel on this context (tuning; no
 * reference counterpart, used by the project's autotuner):
 *   tiles_per_pass: 16-column weight/count tiles each workgroup accumulates
 *                   (0 = default, 3, 4, 5, 6 or 8; 18 = two sets of waves with a
 *                   9-tile group each over the same people, one row read for
 *                   both, two-stage batches only; all tile groups of a block
 *                   still run in one launch);
 *   batch_stages:   64-row stages whose loads are issued together (0 =
 *                   default, 2 or 4).
 * Results are identical for every shape (exact integer block sums, the same
 * FP64 order). */
gnm_status gnm_set_launch_config(gnm_ctx *ctx, uint32_t tiles_per_pass, uint32_t batch_stages);

/* Kernel timing of the last gnm_score_rows[_device] call on this context,
 * measured with HIP events on the library's compute stream: milliseconds of
 * the main contraction kernel, and of the whole call (all kernels). */
gnm_status gnm_last_timing(gnm_ctx *ctx, float *main_kernel_ms, float *total_ms);

#ifdef __cplusplus
}
#endif
#endif
