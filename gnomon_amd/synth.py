"""Synthetic workloads of the benchmark configurations.

Genotypes come from a counter-based hash keyed by (seed, variant, sample), so
the device generator (``gnm_synth_rows_device``) and the CPU generator used by
the tests produce identical bytes without shipping a 25 GB .bed.  HWE draws
with A2 frequency p per variant, P(d = 0, 1, 2) = ((1-p)^2, 2p(1-p), p^2), and
independent missingness.  Weights are drawn on the host with numpy.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field

import numpy as np

from ._lib import check, lib


def p24(x: float) -> int:
    return int(round(x * (1 << 24)))


def m16(x: float) -> int:
    return int(round(x * (1 << 16)))


@dataclass
class Workload:
    name: str
    num_people: int
    num_variants: int
    num_scores: int
    p_lo: float
    p_hi: float
    missing: float
    weights: str            # "normal", "laplace"
    weight_scale: float     # sigma (normal) or b (laplace)
    zero_fraction: float = 0.0
    flip_fraction: float = 0.5
    seed: int = 20240611
    files: int = 0          # >0: grouped files with log-uniform variant counts
    file_min: int = 0
    extra: dict = field(default_factory=dict)

    @property
    def bytes_per_variant(self) -> int:
        return (self.num_people + 3) // 4

    def scaled(self, num_people=None, num_variants=None, num_scores=None, files=None):
        d = dict(self.__dict__)
        if num_people is not None:
            d["num_people"] = num_people
        if num_variants is not None:
            d["num_variants"] = num_variants
        if num_scores is not None:
            d["num_scores"] = num_scores
        if files is not None:
            d["files"] = files
        return Workload(**d)


# The five configurations of the benchmark definition.  "N(0, 1e-4)" is read
# as variance 1e-4 (sigma 0.01), matching the small plain case.
WORKLOADS = {
    "plain_small": Workload("plain_small", 2_000, 50_000, 1, 0.05, 0.5, 0.01, "normal", 0.01),
    "plain_fp32_biobank": Workload("plain_fp32_biobank", 100_000, 1_000_000, 8, 0.01, 0.5, 0.02,
                                   "normal", 0.01),
    "precision_64": Workload("precision_64", 100_000, 1_000_000, 64, 0.01, 0.5, 0.02, "laplace",
                             1e-3, zero_fraction=0.3),
    "batching_200": Workload("batching_200", 100_000, 1_200_000, 200, 0.01, 0.5, 0.02, "normal",
                             0.01, files=200, file_min=1_000),
    "adjoint_16": Workload("adjoint_16", 50_000, 500_000, 16, 0.01, 0.5, 0.02, "laplace", 1e-3),
    # sequencing-style rows for the sparse ingest of host rows: rare variants
    # (A2 frequency ~U(1e-4, 0.02), 0.1% missing: half the rows sparse) and
    # very rare ones (~U(1e-5, 0.002), 0.05% missing: every row sparse)
    "wgs_rare": Workload("wgs_rare", 100_000, 1_000_000, 8, 1e-4, 0.02, 0.001, "normal", 0.01),
    "wgs_ultrarare": Workload("wgs_ultrarare", 100_000, 1_000_000, 8, 1e-5, 0.002, 0.0005, "normal", 0.01),
}


def draw_weights(w: Workload, rng: np.random.Generator | None = None):
    """Returns (weights f32 [V,K], flipped bool [V], present bool [V,K] or None)."""
    rng = rng or np.random.default_rng(w.seed + 1)
    V, K = w.num_variants, w.num_scores
    if w.weights == "normal":
        W = rng.normal(0.0, w.weight_scale, size=(V, K)).astype(np.float32)
    elif w.weights == "laplace":
        W = rng.laplace(0.0, w.weight_scale, size=(V, K)).astype(np.float32)
    else:
        raise ValueError(w.weights)
    if w.zero_fraction:
        W[rng.random((V, K)) < w.zero_fraction] = np.float32(0.0)
    flipped = rng.random(V) < w.flip_fraction
    present = None
    if w.files:
        lo, hi = np.log(max(1, w.file_min)), np.log(V)
        sizes = np.exp(rng.uniform(lo, hi, size=K)).astype(np.int64).clip(1, V)
        present = np.zeros((V, K), dtype=bool)
        for k in range(K):
            present[rng.choice(V, size=int(sizes[k]), replace=False), k] = True
    return W, flipped, present


def synth_rows_device(w: Workload, d_out_ptr: int, row_stride: int, first_variant: int, num_rows: int,
                      stream=None) -> None:
    """Generate packed rows [first_variant, first_variant+num_rows) into device memory."""
    check(lib().gnm_synth_rows_device(w.seed, p24(w.p_lo), p24(w.p_hi), m16(w.missing), w.num_people,
                                      first_variant, num_rows, ctypes.c_void_p(d_out_ptr), row_stride,
                                      None if stream is None else ctypes.c_void_p(stream)))
