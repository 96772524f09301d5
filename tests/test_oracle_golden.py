"""Pins the CPU oracle against the reference's golden vectors and closed forms.

Each check is chosen so that a plausible mistake in the oracle (wrong code
table, swapped people/variants, dropped missing correction, wrong flip sign,
wrong dispatch threshold) fails it.
"""
import json
import os

import numpy as np
import pytest

from gnomon_amd.prep import csr_from_dense, csr_from_records, sscore_columns
from tests.helpers import gpu_tests_genotype, pack_bed, reference_agree

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "golden.json")))
LD = os.path.join(os.path.dirname(__file__), "golden", "ld.bed")


def _single_variant_csr(w_eff, corr, K=1):
    from gnomon_amd.prep import ScoreCsr
    return ScoreCsr(weights=np.full(K, w_eff, np.float32), corrections=np.full(K, corr, np.float32),
                    columns=np.arange(K, dtype=np.uint32), row_offsets=np.array([0, K], np.uint64),
                    num_scores=K, baseline=np.full(K, corr, np.float64),
                    score_variant_counts=np.ones(K, np.uint32))


def test_variant_major_golden_byte(orc):
    g = GOLD["variant_major_missing_correction"]
    rows = np.array([[g["byte"]]], np.uint8)
    csr = _single_variant_csr(g["w_eff"], g["corr"])
    S, C = orc.score_fp64(rows, csr, g["num_people"])
    # the oracle returns final scores (baseline included): partial + baseline
    assert np.allclose(S[:, 0] - g["corr"], g["partial_scores"], atol=1e-12)
    assert C[:, 0].tolist() == g["missing"]
    # the restated engine through both of its paths
    for path in (0, 1):
        S2, C2, _ = orc.refcpu_run(rows, csr, g["num_people"], threads=1, force_path=path)
        assert np.allclose(S2[:, 0] - g["corr"], g["partial_scores"], atol=1e-12)
        assert C2[:, 0].tolist() == g["missing"]


def test_word_skip_and_tail_golden(orc):
    g = GOLD["word_skip_and_tail"]
    N = g["num_people"]
    row = np.zeros((1, (N + 3) // 4), np.uint8)
    for person, code in g["codes"]:
        row[0, person // 4] |= code << ((person % 4) * 2)
    csr = _single_variant_csr(g["w_eff"], g["corr"])
    for S, C in (orc.score_fp64(row, csr, N), orc.refcpu_run(row, csr, N, threads=2, force_path=1)[:2],
                 orc.refcpu_run(row, csr, N, threads=1, force_path=0)[:2]):
        partial = S[:, 0] - g["corr"]
        for p, val in g["partial_scores"].items():
            assert partial[int(p)] == val
        assert int(np.count_nonzero(partial)) == g["nonzero_scores"]
        assert int(C.sum()) == g["missing_total"] and C[99, 0] == 1


def test_many_scores_golden(orc):
    g = GOLD["many_scores"]
    K = g["num_scores"]
    rows = pack_bed(np.array([g["dosages"]]))
    csr = _single_variant_csr(1.0, 0.0, K)
    for S, C in (orc.score_fp64(rows, csr, g["num_people"]),
                 orc.refcpu_run(rows, csr, g["num_people"], threads=1, force_path=0)[:2]):
        for p, expect in enumerate(g["scores"]):
            assert np.all(S[p] == expect)
        assert C.sum() == 0


def test_aggregation_golden():
    g = GOLD["aggregation"]
    tol = g["tolerance"]
    for key in ("same_orientation", "flipped"):
        c = csr_from_records(1, 1, [(0, 0, w, f) for w, f in g[key]["rows"]])
        assert abs(float(c.weights[0]) - g[key]["w_eff"]) < tol
        assert abs(float(c.corrections[0]) - g[key]["corr"]) < tol
    rows = g["mixed_rows"]
    c = csr_from_records(1, 1, [(0, 0, w, f) for w, f in rows])
    for d in (0.0, 1.0, 2.0):
        row_by_row = sum(np.float32(w) * ((2.0 - d) if f else d) for w, f in rows)
        agg = float(c.corrections[0]) + float(c.weights[0]) * d
        assert abs(row_by_row - agg) < tol
    # one CSR entry per present cell, zero weights included (prepare.rs:1579-1609)
    c = csr_from_dense(np.zeros((3, 2), np.float32), np.array([0, 1, 0], bool))
    assert c.nnz == 6 and c.score_variant_counts.tolist() == [3, 3]


def test_special_case_known_answer(orc):
    g = GOLD["special_case"]
    d = np.array(g["a2_dosages"])[:, None]  # 50 variants x 1 person
    rows = pack_bed(d)
    csr = csr_from_dense(np.full((50, 1), g["weight"], np.float32), np.ones(50, bool))
    S, C = orc.score_fp64(rows, csr, 1)
    assert abs(S[0, 0] - g["sum"]) < 1e-6
    avg, pct = sscore_columns(S, C, csr.score_variant_counts)
    assert abs(avg[0, 0] - g["avg"]) < 1e-7 and pct[0, 0] == 0.0


def test_decode_gpu_tests_generator(orc):
    V, N = 37, 53
    truth = gpu_tests_genotype(np.arange(N)[None, :], np.arange(V)[:, None])
    rows = pack_bed(truth)
    tiles = orc.decode_rows(rows, N).astype(np.int64)
    expect = np.where(truth < 0, 3, truth)
    assert np.array_equal(tiles, expect)
    # person subset / reorder gather
    o2f = np.array([52, 0, 7, 7, 31], np.uint32)
    t2 = orc.decode_rows(rows, len(o2f), o2f)
    assert np.array_equal(t2, expect[:, o2f])


def test_decode_ld_fixture_brute_force(orc):
    raw = np.fromfile(LD, np.uint8)
    assert raw[:3].tolist() == [0x6C, 0x1B, 0x01]  # SNP-major magic (shared/files.rs:373-400)
    N, V = 200, 1500
    rows = raw[3:].reshape(V, N // 4)
    tiles = orc.decode_rows(rows, N)
    # brute force from the PLINK spec, one genotype at a time
    rng = np.random.default_rng(0)
    for _ in range(2000):
        v, p = int(rng.integers(V)), int(rng.integers(N))
        code = (int(rows[v, p // 4]) >> (2 * (p % 4))) & 3
        assert tiles[v, p] == {0: 0, 1: 3, 2: 1, 3: 2}[code]


def test_fp64_oracle_closed_form_and_invariants(orc):
    rng = np.random.default_rng(5)
    V, N, K = 300, 77, 5
    dos = rng.choice([0, 1, 2, -1], size=(V, N), p=[0.4, 0.35, 0.2, 0.05])
    W = rng.normal(0, 1, (V, K)).astype(np.float32)
    flip = rng.random(V) < 0.5
    csr = csr_from_dense(W, flip)
    rows = pack_bed(dos)
    S, C = orc.score_fp64(rows, csr, N)
    eff = np.where(dos < 0, 0, np.where(flip[:, None], 2 - dos, dos)).astype(np.float64)
    assert np.allclose(S, eff.T @ W.astype(np.float64), rtol=1e-12, atol=1e-12)
    assert np.array_equal(C, np.repeat((dos < 0).sum(0)[:, None], K, 1))
    # flip symmetry: w * g_f(d) = 2w - w * g_{not f}(d) on non-missing calls, so
    # scoring -W with every flip inverted gives S - 2 * sum_{non-missing} w
    csr2 = csr_from_dense(-W, ~flip)
    S2, _ = orc.score_fp64(rows, csr2, N)
    nonmiss = (dos >= 0).astype(np.float64)
    assert np.allclose(S2, S - nonmiss.T @ (2.0 * W.astype(np.float64)), rtol=1e-10, atol=1e-10)
    # all missing -> 0
    S3, C3 = orc.score_fp64(pack_bed(np.full((V, N), -1)), csr, N)
    assert np.all(S3 == 0.0) and np.all(C3 == V)


def test_refcpu_matches_fp64_within_reference_rule(orc):
    rng = np.random.default_rng(11)
    V, N, K = 700, 300, 70
    dos = rng.choice([0, 1, 2, -1], size=(V, N), p=[0.5, 0.3, 0.18, 0.02])
    W = rng.normal(0, 0.01, (V, K)).astype(np.float32)
    present = rng.random((V, K)) < 0.8
    csr = csr_from_dense(W, rng.random(V) < 0.5, present)
    rows = pack_bed(dos)
    S64, C64 = orc.score_fp64(rows, csr, N)
    for path in (-1, 0, 1):
        S, C, used = orc.refcpu_run(rows, csr, N, threads=3, force_path=path)
        assert used == 3
        assert np.array_equal(C, C64)
        assert reference_agree(S, S64, 1e-6).all()


def test_decide_tree_spot_values(orc):
    # score/decide.rs:107-189 at points on both sides of each threshold
    P, NP = 0, 1
    assert orc.decide_path(100000, 8, 1.0, 0.2) == P
    assert orc.decide_path(100000, 8, 1.0, 0.004) == NP
    assert orc.decide_path(100000, 8, 0.2, 0.2) == NP
    assert orc.decide_path(2000, 1, 1.0, 0.5) == NP
    assert orc.decide_path(5000, 1, 1.0, 0.5) == P
    assert orc.decide_path(5000, 20, 1.0, 0.5) == NP
    assert orc.decide_path(10, 1, 0.2, 0.001) == P
    assert orc.decide_path(15000, 8, 0.5, 0.5) == P
    assert orc.decide_path(15000, 8, 0.2, 0.5) == NP
    assert orc.decide_path(15000, 8, 1.0, 0.05) == P
    assert orc.decide_path(15000, 20, 1.0, 0.05) == NP
    assert orc.decide_path(15000, 20, 1.0, 0.5) == P
    # density (batch.rs:583-614): early exit to 1.0, else set bits / people
    assert orc.variant_density(np.zeros(25, np.uint8), 100) == 0.0
    assert orc.variant_density(np.array([0b11] + [0] * 24, np.uint8), 100) == pytest.approx(0.02)
    assert orc.variant_density(np.full(25, 0xFF, np.uint8), 100) == 1.0


def test_synth_generator_statistics(orc):
    rows = orc.synth_rows(7, 4001, 0, 400, int(0.05 * 2**24), int(0.5 * 2**24), int(0.01 * 2**16))
    assert rows.shape == (400, 1001)
    tiles = orc.decode_rows(rows, 4001)
    miss = (tiles == 3).mean()
    assert 0.008 < miss < 0.012
    # padding bits of the last byte are zero
    assert np.all((rows[:, -1] >> 2) == 0)
    # HWE: per-variant het fraction close to 2p(1-p) with p from the A2 count
    d = np.where(tiles == 3, 0, tiles).astype(np.float64)
    nm = (tiles != 3).sum(1)
    p = d.sum(1) / (2 * nm)
    het = (tiles == 1).sum(1) / nm
    assert np.mean(np.abs(het - 2 * p * (1 - p))) < 0.02
    # same key, same bytes; sub-range consistency
    again = orc.synth_rows(7, 4001, 100, 50, int(0.05 * 2**24), int(0.5 * 2**24), int(0.01 * 2**16))
    assert np.array_equal(again, rows[100:150])


def test_adjoint_and_forward_fp64(orc):
    rng = np.random.default_rng(3)
    V, N, K = 60, 41, 3
    dos = rng.choice([0, 1, 2, -1], size=(V, N))
    flip = (rng.random(V) < 0.5).astype(np.uint8)
    W = rng.normal(size=(V, K)).astype(np.float32)
    rows = pack_bed(dos)
    S, M = orc.dense_forward_fp64(rows, N, flip, W)
    G = np.where(dos < 0, 0, np.where(flip[:, None] == 1, 2 - dos, dos)).astype(np.float64)
    assert np.allclose(S, G.T @ W.astype(np.float64), rtol=1e-12, atol=1e-12)
    assert np.array_equal(M, (dos < 0).sum(0))
    R = rng.normal(size=(N, K))
    A = orc.adjoint_fp64(rows, N, flip, R)
    assert np.allclose(A, G @ R, rtol=1e-12, atol=1e-12)
    # adjoint identity <G W, R> = <W, G^T R>
    assert np.isclose((S * R).sum(), (W.astype(np.float64) * A).sum(), rtol=1e-10)
