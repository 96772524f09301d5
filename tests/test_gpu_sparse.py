"""Sparse ingest of host rows (gnm_set_sparse_ingest) on the device, through
the C ABI.

Rows whose 64-bit words are mostly zero cross PCIe as a bitmap plus their
non-zero words and are rebuilt on the device before the dense kernels score
them.  The scores and missing counts must be bit-identical to sending every
row packed, and pass the precision's rule against the FP64 oracle.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from gnomon_amd._lib import GnmError
from gnomon_amd.pipeline import ScoringContext
from gnomon_amd.prep import csr_from_dense
from tests.helpers import abs_terms, avg_agree, fp32_sum_bound, normwise_rel, reference_agree

pytestmark = pytest.mark.gpu
TOL = {"fp32": None, "bf16x3": 1e-6, "i8x4": 1e-6}


def _flush(msg):
    print(msg, flush=True)


def _check(S, C, S64, C64, precision, fp32=None):
    """fp32: (orc, rows, csr, N, out_to_fam) -- the FP32 path is held to the
    FP32 summation bound and the reference's rule on AVG values."""
    assert np.array_equal(C, C64), "missing counts differ"
    if precision == "fp32":
        orc, rows, csr, n, o2f = fp32
        assert np.all(np.abs(S - S64) <= fp32_sum_bound(abs_terms(orc, rows, csr, n, o2f)))
        assert avg_agree(S, S64, C, csr, 1e-6).all()
        return
    ok = reference_agree(S, S64, 1e-6)
    assert ok.all(), f"{(~ok).sum()} cells outside the 1e-6 rule; worst {np.abs(S - S64).max()}"
    if TOL[precision]:
        assert normwise_rel(S, S64).max() <= TOL[precision]


def _rare_problem(orc, seed, N, V, K, nfam=None, p_hi=0.03):
    """HWE rows with A2 frequency ~U(1e-4, p_hi) and 0.2% missing calls (sparse
    and dense rows), per-cell flips and presence, 5% of variants in no score."""
    nf = nfam or N
    rows = orc.synth_rows(seed, nf, 0, V, int(1e-4 * 2**24), int(p_hi * 2**24), 131)
    rng = np.random.default_rng(seed)
    W = rng.laplace(0, 1e-3, (V, K)).astype(np.float32)
    W[rng.random((V, K)) < 0.2] = 0.0
    flip = rng.random((V, K)) < 0.5
    present = rng.random((V, K)) < 0.9
    present[rng.random(V) < 0.05] = False
    return rows, csr_from_dense(W, flip, present)


def _run(csr, rows, N, precision, fraction, nfam=None, o2f=None, blocks=1):
    with ScoringContext(csr, N, nfam, o2f, precision=precision, force=True) as ctx:
        ctx.set_sparse_ingest(fraction)
        edges = np.linspace(0, csr.num_variants, blocks + 1).astype(int)
        for a, b in zip(edges[:-1], edges[1:]):
            ctx.score_rows(rows[a:b], int(a))
        S, C = ctx.finish()
        return S, C, ctx.ingest_stats()


@pytest.mark.parametrize("precision", ["fp32", "bf16x3", "i8x4"])
def test_sparse_ingest_bit_identical_to_packed(orc, precision):
    _flush(f"sparse ingest {precision}")
    N, V, K = 40001, 3000, 6      # three 16384-person chunks, the last one partial
    rows, csr = _rare_problem(orc, 41, N, V, K)
    S64, C64 = orc.score_fp64(rows, csr, N)
    S, C, st = _run(csr, rows, N, precision, 0.5, blocks=3)
    assert st["sparse_rows"] > V // 4 and st["dense_rows"] > 0, st
    assert st["bytes_to_device"] < V * rows.shape[1]
    _check(S, C, S64, C64, precision, (orc, rows, csr, N, None))
    Sd, Cd, std = _run(csr, rows, N, precision, 0.0, blocks=3)
    assert std["sparse_rows"] == 0 and std["dense_rows"] == V
    assert np.array_equal(Sd, S) and np.array_equal(Cd, C)


def test_every_row_sparse_and_empty_rows(orc):
    """Every row under the threshold, some all-00, some variants in no score."""
    _flush("sparse all rows")
    N, V, K = 2048, 3000, 3
    rows, csr = _rare_problem(orc, 43, N, V, K, p_hi=0.003)
    rows[::7] = 0
    S64, C64 = orc.score_fp64(rows, csr, N)
    S, C, st = _run(csr, rows, N, "i8x4", 0.5)
    assert st["dense_rows"] == 0 and st["sparse_rows"] > 0
    _check(S, C, S64, C64, "i8x4")
    Sd, Cd, _ = _run(csr, rows, N, "i8x4", 0.0)
    assert np.array_equal(Sd, S) and np.array_equal(Cd, C)


def test_sparse_ingest_person_subset(orc):
    _flush("sparse subset")
    rng = np.random.default_rng(45)
    Nfam, V, K = 20001, 2000, 4
    o2f = rng.permutation(Nfam)[:9000].astype(np.uint32)
    rows, csr = _rare_problem(orc, 45, 9000, V, K, nfam=Nfam)
    S64, C64 = orc.score_fp64(rows, csr, 9000, o2f)
    for precision in ("fp32", "i8x4"):
        S, C, st = _run(csr, rows, 9000, precision, 0.5, nfam=Nfam, o2f=o2f, blocks=2)
        assert st["sparse_rows"] > 0
        _check(S, C, S64, C64, precision, (orc, rows, csr, 9000, o2f))
        Sd, Cd, _ = _run(csr, rows, 9000, precision, 0.0, nfam=Nfam, o2f=o2f, blocks=2)
        assert np.array_equal(Sd, S) and np.array_equal(Cd, C)


def test_sparse_ingest_checkpoint_resume(orc):
    """Accumulators read mid-stream seed a fresh context that finishes the rest."""
    _flush("sparse checkpoint")
    N, V, K = 2500, 5000, 5
    rows, csr = _rare_problem(orc, 47, N, V, K)
    S64, C64 = orc.score_fp64(rows, csr, N)
    cut = 2100
    with ScoringContext(csr, N, precision="i8x4", force=True) as ctx:
        ctx.score_rows(rows[:cut], 0)
        Sm, Cm, done = ctx.read_accumulators()
    assert done == cut
    with ScoringContext(csr, N, precision="i8x4", force=True) as ctx2:
        ctx2.seed_accumulators(Sm, Cm, cut)
        ctx2.score_rows(rows[cut:], cut)
        S, C = ctx2.finish()
    _check(S, C, S64, C64, "i8x4")


def test_sparse_ingest_rejects_bad_threshold(orc):
    N, V, K = 512, 300, 2
    rows, csr = _rare_problem(orc, 49, N, V, K)
    with ScoringContext(csr, N, precision="i8x4", force=True) as ctx:
        for bad in (-0.1, 0.6, float("nan")):
            with pytest.raises(GnmError):
                ctx.set_sparse_ingest(bad)
