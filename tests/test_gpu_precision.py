"""Per-sample precision of the split paths at the shape of the precision
configuration: 64 scores, weights ~ Laplace(0, 1e-3) with 30% exact zeros
(kept present), HWE genotypes with A2 frequency ~ U(0.01, 0.5), 2% missing,
the effect allele A1 for half the variants.  Scaled to 2,000 people x 70,000
variants so the FP64 oracle finishes in seconds.

Metric, per score: the largest |S - S64| / |S64| over the people with
|S64| >= 1e-3 max|S64| (S64 = the FP64 oracle).  Asserted <= 1e-6 for
bf16x4 (four 8-bit BF16 limbs), i8x4 (four int8 limbs) and AUTO; measured and
printed for the shorter splits bf16x3, bf16x2 and i8x3, whose 16-24 bits of
the block maximum do not reach it for small-|S| people.  Missing counts are
exact on every path.  The plain FP32 path (FP32 sums over blocks of program
rows, combined in FP64, the reference kernel's arithmetic class) is checked
against the restated reference CPU engine and the FP64 oracle under the
reference's rule (score/tests/gpu_tests.rs:134-137) at 1e-6, and against the
FP32 summation bound.
"""
import numpy as np
import pytest

from gnomon_amd import synth
from gnomon_amd.pipeline import ScoringContext
from gnomon_amd.prep import csr_from_dense
from tests.helpers import fp32_sum_bound, normwise_rel, oracle_fp64_parallel, per_sample_rel, reference_agree

pytestmark = pytest.mark.gpu

N, V = 2000, 70_000


def _flush(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def c3(orc):
    w = synth.WORKLOADS["precision_64"].scaled(num_people=N, num_variants=V)
    W, flipped, present = synth.draw_weights(w)
    csr = csr_from_dense(W, flipped, present)
    rows = orc.synth_rows(w.seed, N, 0, V, synth.p24(w.p_lo), synth.p24(w.p_hi), synth.m16(w.missing))
    S64, C64 = oracle_fp64_parallel(orc, rows, csr, N)
    return csr, rows, S64, C64


def _score(csr, rows, precision):
    with ScoringContext(csr, N, precision=precision, force=True) as ctx:
        ctx.score_rows(rows, 0)
        return ctx.finish()


@pytest.mark.parametrize("precision", ["bf16x4", "i8x4", "auto"])
def test_split_path_per_sample_error_within_1e6(c3, precision):
    csr, rows, S64, C64 = c3
    _flush(f"precision C3-shape N={N} V={V} K=64 [{precision}]")
    S, C = _score(csr, rows, precision)
    assert np.array_equal(C, C64), "missing counts differ from the oracle"
    rel = per_sample_rel(S, S64)
    _flush(f"[{precision}] per-sample max rel {rel.max():.3e}, normwise {normwise_rel(S, S64).max():.3e}")
    assert rel.max() <= 1e-6, f"worst score {int(rel.argmax())}: {rel.max():.3e}"


@pytest.mark.parametrize("precision,normwise", [("bf16x3", 1e-6), ("bf16x2", 1e-3), ("i8x3", 1e-5)])
def test_short_split_paths_measured(c3, precision, normwise):
    """Shorter splits: counts exact, normwise error within the representation
    of 24, 16 and 23 bits of the block maximum; per-sample error printed."""
    csr, rows, S64, C64 = c3
    _flush(f"precision C3-shape [{precision}]")
    S, C = _score(csr, rows, precision)
    assert np.array_equal(C, C64)
    nw = normwise_rel(S, S64).max()
    _flush(f"[{precision}] per-sample max rel {per_sample_rel(S, S64).max():.3e}, normwise {nw:.3e}")
    assert nw <= normwise


def test_fp32_path_matches_reference_engine(c3, orc):
    """Plain FP32 against the restated reference CPU engine (its f32
    mini-batch arithmetic) under the reference's agreement rule at 1e-6,
    and against the FP64 oracle."""
    csr, rows, S64, C64 = c3
    _flush("precision C3-shape [fp32] vs reference engine")
    S, C = _score(csr, rows, "fp32")
    assert np.array_equal(C, C64)
    Sr, Cr, _ = orc.refcpu_run(rows, csr, N, threads=0)
    assert np.array_equal(Cr, C64)
    ok = reference_agree(S, Sr, 1e-6)
    _flush(f"[fp32] vs engine max|d| {np.abs(S - Sr).max():.3e}; vs FP64 normwise {normwise_rel(S, S64).max():.3e},"
           f" per-sample {per_sample_rel(S, S64).max():.3e}")
    assert ok.all(), f"{(~ok).sum()} cells outside the reference rule"
    assert reference_agree(S, S64, 1e-6).all()
    # FP32 sums over blocks of program rows: every |dosage x weight| term is
    # at most 2 |w|, so per score the error is within the FP32 block bound of
    # 2 sum_v |w_vk| for every person
    absw = np.zeros(S64.shape[1])
    np.add.at(absw, csr.columns, 2.0 * np.abs(csr.weights.astype(np.float64)) +
              np.abs(csr.corrections.astype(np.float64)))
    assert np.all(np.abs(S - S64) <= fp32_sum_bound(absw)[None, :])
