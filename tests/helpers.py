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
