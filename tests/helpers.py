"""Shared helpers of the test-suite (bed packing, tolerance rules)."""
import numpy as np

# PLINK 2-bit codes of the A2 dosage (score/tests/gpu_tests.rs:621-628)
CODE_OF_DOSAGE = {0: 0b00, 1: 0b10, 2: 0b11, -1: 0b01}


def pack_bed(dosage: np.ndarray) -> np.ndarray:
    """dosage [V, N] in {0,1,2,-1(missing)} -> packed SNP-major rows [V, ceil(N/4)]."""
    d = np.asarray(dosage)
    V, N = d.shape
    code = np.select([d == 0, d == 1, d == 2], [0, 2, 3], default=1).astype(np.uint8)
    bpv = (N + 3) // 4
    padded = np.zeros((V, bpv * 4), np.uint8)
    padded[:, :N] = code
    q = padded.reshape(V, bpv, 4)
    return (q[..., 0] | (q[..., 1] << 2) | (q[..., 2] << 4) | (q[..., 3] << 6)).astype(np.uint8)


def reference_agree(a, b, tol=1e-6):
    """GPU/CPU agreement rule of score/tests/gpu_tests.rs:134-137:
    |a-b| <= tol or |a-b| / max(|a|, |b|, 1) <= tol."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    d = np.abs(a - b)
    return (d <= tol) | (d / np.maximum(np.maximum(np.abs(a), np.abs(b)), 1.0) <= tol)


def normwise_rel(a, ref):
    """max |a - ref| / max |ref| per score column."""
    a = np.asarray(a, np.float64)
    ref = np.asarray(ref, np.float64)
    den = np.maximum(np.abs(ref).max(axis=0), 1e-300)
    return (np.abs(a - ref).max(axis=0) / den)


def gpu_tests_genotype(person, variant):
    """score/tests/gpu_tests.rs:611-619."""
    key = (person * 131 + variant * 17 + 23) % 31
    return np.select([key <= 1, key <= 9, key <= 20], [-1, 0, 1], default=2)


def per_sample_rel(S, S64, floor=1e-3):
    """Per score: max over people with |S64| >= floor * max|S64| of
    |S - S64| / |S64| (0 for a score whose oracle column is all zero)."""
    S = np.asarray(S, np.float64)
    S64 = np.asarray(S64, np.float64)
    out = np.zeros(S64.shape[1])
    for k in range(S64.shape[1]):
        ref = S64[:, k]
        m = np.abs(ref).max()
        if m == 0:
            continue
        sel = np.abs(ref) >= floor * m
        out[k] = (np.abs(S[sel, k] - ref[sel]) / np.abs(ref[sel])).max()
    return out


def oracle_fp64_parallel(orc, rows, csr, num_people, threads=0):
    """The FP64 oracle over slices of people in a thread pool."""
    return orc.score_fp64_parallel(rows, csr, num_people, threads)


# FP32 path: products are exact and summed in FP32 over blocks of at most
# F32_BLOCK_ROWS program rows, each block sum then added in FP64.  A block of
# n terms takes n - 1 rounded FP32 additions, so by the recursive-summation
# bound |block - exact| <= gamma_{n-1} * sum |terms| with
# gamma_m = m u / (1 - m u), u = 2^-24 (Higham, Accuracy and Stability of
# Numerical Algorithms, 2nd ed., sec. 4.2); the FP64 steps add at most
# ~1e-15 of the same sum.
F32_BLOCK_ROWS = 224


def fp32_sum_bound(abs_terms_sum):
    """Error bound of the FP32 path for cells whose sum of |dosage x weight|
    terms is abs_terms_sum."""
    m = F32_BLOCK_ROWS - 1
    u = 2.0 ** -24
    return (m * u / (1 - m * u) + 1e-14) * np.asarray(abs_terms_sum, np.float64)


def abs_terms(orc, rows, csr, num_people, out_to_fam=None):
    """Per (person, score): sum over the score's variants of |effect dosage x
    effect weight|, with the oracle's FP64 scorer on a CSR of magnitudes.  A
    cell's A2-allele weight u = W_eff + corr/2 and A1-allele weight f =
    corr/2 contribute |u| d + |f| (2 - d) for a non-missing call, i.e.
    corr' + W'·d with corr' = 2|f| and W' = |u| - |f|."""
    from gnomon_amd.prep import ScoreCsr

    u = csr.weights.astype(np.float64) + csr.corrections.astype(np.float64) / 2
    f = csr.corrections.astype(np.float64) / 2
    acsr = ScoreCsr(weights=(np.abs(u) - np.abs(f)).astype(np.float32),
                    corrections=(2 * np.abs(f)).astype(np.float32), columns=csr.columns,
                    row_offsets=csr.row_offsets, num_scores=csr.num_scores,
                    baseline=np.zeros(csr.num_scores), score_variant_counts=csr.score_variant_counts)
    T, _ = orc.score_fp64(rows, acsr, num_people, out_to_fam)
    return T


def avg_agree(S, S2, C, csr, tol=1e-6):
    """The reference's GPU/CPU agreement rule (score/tests/gpu_tests.rs:134-137)
    on what it compares: the .sscore AVG values (score/main.rs:1501-1527)."""
    from gnomon_amd.prep import sscore_columns

    a, _ = sscore_columns(np.asarray(S, np.float64), C, csr.score_variant_counts)
    b, _ = sscore_columns(np.asarray(S2, np.float64), C, csr.score_variant_counts)
    return reference_agree(a, b, tol)
