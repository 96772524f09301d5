"""Weight canonicalisation into the reference's CSR form.

This mirrors the parts of ``score/prepare.rs`` that fix the numbers the hot
path consumes (the parsing and merge-join around them are out of scope):

* ``apply_simple_score_assignment`` (prepare.rs:197-208): every matched score
  row is written in A2-dosage space -- unflipped: ``w_eff += w``; flipped
  (effect allele == BIM A1, prepare.rs:557-558): ``w_eff -= w`` and
  ``corr += 2w``, all in f32; duplicate rows of one (variant, score)
  accumulate in input order (prepare.rs:562-572);
* CSR emission, baseline and counts (prepare.rs:632-650): one CSR entry per
  present (variant, score) cell -- zero weights included --, baseline
  ``sum(corr)`` per score in f64, ``score_variant_counts`` per score.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass
class ScoreCsr:
    weights: np.ndarray        # f32 [nnz]  W_eff
    corrections: np.ndarray    # f32 [nnz]  corr
    columns: np.ndarray        # u32 [nnz]
    row_offsets: np.ndarray    # u64 [V+1]
    num_scores: int
    baseline: np.ndarray       # f64 [K]  sum of corr per score
    score_variant_counts: np.ndarray  # u32 [K]

    @property
    def num_variants(self) -> int:
        return int(self.row_offsets.shape[0] - 1)

    @property
    def nnz(self) -> int:
        return int(self.row_offsets[-1])


def csr_from_records(num_variants: int, num_scores: int, records) -> ScoreCsr:
    """Build the CSR from (variant, score, weight, flipped) records.

    Records of the same (variant, score) cell are aggregated in the order
    given, with the f32 arithmetic of apply_simple_score_assignment.
    """
    cells: dict[tuple[int, int], list] = {}
    for v, k, w, flipped in records:
        if not (0 <= v < num_variants and 0 <= k < num_scores):
            raise ValueError("record out of range")
        e = cells.setdefault((int(v), int(k)), [np.float32(0.0), np.float32(0.0)])
        w = np.float32(w)
        if flipped:
            e[0] = np.float32(e[0] - w)
            e[1] = np.float32(e[1] + np.float32(2.0) * w)
        else:
            e[0] = np.float32(e[0] + w)
    keys = sorted(cells)
    nnz = len(keys)
    weights = np.zeros(nnz, np.float32)
    corr = np.zeros(nnz, np.float32)
    cols = np.zeros(nnz, np.uint32)
    offs = np.zeros(num_variants + 1, np.uint64)
    for i, (v, k) in enumerate(keys):
        weights[i], corr[i] = cells[(v, k)]
        cols[i] = k
        offs[v + 1] += 1
    offs = np.cumsum(offs).astype(np.uint64)
    return _finish(weights, corr, cols, offs, num_scores)


def csr_from_dense(weights: np.ndarray, flipped: np.ndarray, present: np.ndarray | None = None) -> ScoreCsr:
    """Vectorised form for one record per cell.

    weights: f32 [V, K] effect-allele weights; flipped: bool [V] or [V, K]
    (effect allele is BIM A1); present: bool [V, K] cells that exist in the
    score files (default: all; a present zero weight still gets an entry).
    """
    W = np.ascontiguousarray(weights, dtype=np.float32)
    V, K = W.shape
    F = np.asarray(flipped, dtype=bool)
    if F.ndim == 1:
        F = np.broadcast_to(F[:, None], (V, K))
    P = np.ones((V, K), bool) if present is None else np.asarray(present, dtype=bool)
    vv, kk = np.nonzero(P)
    w = W[vv, kk]
    f = F[vv, kk]
    zero = np.float32(0.0)
    w_eff = np.where(f, zero - w, zero + w).astype(np.float32)
    corr = np.where(f, zero + np.float32(2.0) * w, zero).astype(np.float32)
    offs = np.zeros(V + 1, np.uint64)
    np.add.at(offs, vv + 1, 1)
    offs = np.cumsum(offs).astype(np.uint64)
    return _finish(w_eff, corr, kk.astype(np.uint32), offs, K)


def _finish(weights, corr, cols, offs, K) -> ScoreCsr:
    baseline = np.zeros(K, np.float64)
    np.add.at(baseline, cols.astype(np.int64), corr.astype(np.float64))
    counts = np.bincount(cols.astype(np.int64), minlength=K).astype(np.uint32)
    return ScoreCsr(weights=np.ascontiguousarray(weights), corrections=np.ascontiguousarray(corr),
                    columns=np.ascontiguousarray(cols), row_offsets=np.ascontiguousarray(offs),
                    num_scores=int(K), baseline=baseline, score_variant_counts=counts)


def sscore_columns(scores: np.ndarray, counts: np.ndarray, score_variant_counts: np.ndarray):
    """AVG and MISSING_PCT of the .sscore output (score/main.rs:1501-1527).

    AVG = sum / (count_k - missing) in f64 (0 when no variant is used);
    MISSING_PCT = missing / count_k * 100 in f32.
    """
    n = score_variant_counts.astype(np.int64)[None, :]
    used = np.maximum(n - counts.astype(np.int64), 0)
    avg = np.where(used > 0, scores / np.maximum(used, 1), 0.0)
    pct = np.where(n > 0, (counts.astype(np.float32) / np.maximum(n, 1).astype(np.float32)) * np.float32(100.0),
                   np.float32(0.0)).astype(np.float32)
    return avg, pct
