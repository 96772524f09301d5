"""GPU parity of the HIP path against the CPU oracle, through the C ABI.

Integer work (decoded dosages, missing counts) must be bit-exact.  Scores:
  * fp32   -- FP32 sums of exact products over blocks of program rows, added
              in FP64 (the reference kernel's arithmetic class): within the
              FP32 summation bound of the FP64 oracle cell by cell, and the
              reference's own GPU/CPU rule (score/tests/gpu_tests.rs:134-137,
              |a-b| <= 1e-6 or |a-b|/max(|a|,|b|,1) <= 1e-6) on the .sscore
              AVG values it compares, against FP64 and the reference engine;
  * bf16x4, bf16x3, i8x4 -- that rule on the raw sums against FP64, and
              normwise max|S-S64|/max|S64| <= 1e-6 per score;
  * bf16x2 -- normwise <= 1e-4 (16-bit limbs; reported, not a production mode).
"""
import ctypes
import json
import os

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from gnomon_amd import _lib
from gnomon_amd.pipeline import ScoringContext, try_run_hip
from gnomon_amd.prep import ScoreCsr, csr_from_dense
from tests.helpers import (abs_terms, avg_agree, fp32_sum_bound, gpu_tests_genotype, normwise_rel, pack_bed,
                           reference_agree)

pytestmark = pytest.mark.gpu
GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "golden.json")))
LD = os.path.join(os.path.dirname(__file__), "golden", "ld.bed")
TOL = {"fp32": None, "bf16x4": 1e-6, "bf16x3": 1e-6, "bf16x2": 1e-4, "i8x4": 1e-6, "i8x3": 1e-5}


def _flush(msg):
    print(msg, flush=True)


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _fp32_terms(orc, precision, rows, csr, N, o2f=None):
    """For the FP32 path: per (person, score) the sum of |dosage x weight|
    terms its FP32 block sums add (the summation error bound scales with
    it); None for the other paths."""
    return (abs_terms(orc, rows, csr, N, o2f), csr) if precision == "fp32" else None


def _check_scores(S, C, S64, C64, precision, fp32=None):
    assert np.array_equal(C, C64), "missing counts differ"
    if precision == "fp32":
        # FP32 block sums (the reference kernel's arithmetic class): within
        # the FP32 summation bound, and the reference's own GPU/CPU rule on
        # the .sscore AVG values
        terms, csr = fp32
        err = np.abs(S - S64)
        assert np.all(err <= fp32_sum_bound(terms)), f"worst excess {(err - fp32_sum_bound(terms)).max()}"
        assert avg_agree(S, S64, C, csr, 1e-6).all()
        return
    if precision in ("bf16x2", "i8x3"):
        assert normwise_rel(S, S64).max() <= TOL[precision]
        return
    ok = reference_agree(S, S64, 1e-6)
    assert ok.all(), f"{(~ok).sum()} cells outside the 1e-6 rule; worst {np.abs(S - S64).max()}"
    if precision in ("bf16x4", "bf16x3", "i8x4"):
        assert normwise_rel(S, S64).max() <= TOL[precision]


def _score(csr, rows, N, precision, nfam=None, o2f=None, blocks=1):
    with ScoringContext(csr, N, nfam, o2f, precision=precision, force=True) as ctx:
        V = csr.num_variants
        edges = np.linspace(0, V, blocks + 1).astype(int)
        for a, b in zip(edges[:-1], edges[1:]):
            ctx.score_rows(rows[a:b], int(a))
        return ctx.finish()


def test_device_visible():
    assert _lib.lib().gnm_device_count() >= 1


def test_synth_device_matches_cpu_generator(orc):
    _flush("synth N=1001 V=300")
    N, V, stride = 1001, 300, 260
    d = torch.zeros(V * stride, dtype=torch.uint8, device="cuda")
    _lib.check(_lib.lib().gnm_synth_rows_device(99, 838861, 8388608, 1311, N, 5, V,
                                                ctypes.c_void_p(d.data_ptr()), stride, None))
    torch.cuda.synchronize()
    got = d.cpu().numpy().reshape(V, stride)[:, :(N + 3) // 4]
    want = orc.synth_rows(99, N, 5, V, 838861, 8388608, 1311)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("N,V,with_flip", [(64, 517, True), (1001, 517, True), (4099, 517, True),
                                           (4112, 517, True), (4112, 517, False), (128, 5000, True)])
def test_unpack_bit_exact(orc, N, V, with_flip):
    """Vector path (N a multiple of 16; 4112 people = one full wave span plus
    16 people), byte path (N odd), several row blocks (V > 512), row counts
    that are not a multiple of the 32-row batch, no flip array."""
    _flush(f"unpack N={N} V={V} flip={with_flip}")
    rng = np.random.default_rng(N)
    dos = rng.choice([0, 1, 2, -1], size=(V, N), p=[0.45, 0.3, 0.2, 0.05])
    dos[3] = -1          # all missing
    dos[4] = 2           # all hom A2
    rows = pack_bed(dos)
    flip = (rng.random(V) < 0.5).astype(np.uint8) if with_flip else np.zeros(V, np.uint8)
    d_rows, d_flip = _dev(rows), _dev(flip)
    d_out = torch.zeros(V * N, dtype=torch.uint8, device="cuda")
    d_miss = torch.zeros(N, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    _lib.check(_lib.lib().gnm_unpack_device(ctypes.c_void_p(d_rows.data_ptr()), rows.shape[1], V, N,
                                            ctypes.c_void_p(d_flip.data_ptr() if with_flip else None),
                                            ctypes.c_void_p(d_out.data_ptr()),
                                            ctypes.c_void_p(d_miss.data_ptr()), None))
    torch.cuda.synchronize()
    tiles = orc.decode_rows(rows, N)
    flipped = np.array([2, 1, 0, 3], np.uint8)[tiles]
    want = np.where(flip[:, None] == 1, flipped, tiles)
    assert np.array_equal(d_out.cpu().numpy().reshape(V, N), want)
    assert np.array_equal(d_miss.cpu().numpy().astype(np.int64), (tiles == 3).sum(0))


def test_unpack_ld_fixture(orc):
    raw = np.fromfile(LD, np.uint8)
    N, V = 200, 1500
    rows = raw[3:].reshape(V, N // 4)
    d_rows = _dev(rows)
    d_out = torch.zeros(V * N, dtype=torch.uint8, device="cuda")
    d_miss = torch.zeros(N, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    _lib.check(_lib.lib().gnm_unpack_device(ctypes.c_void_p(d_rows.data_ptr()), N // 4, V, N, None,
                                            ctypes.c_void_p(d_out.data_ptr()),
                                            ctypes.c_void_p(d_miss.data_ptr()), None))
    torch.cuda.synchronize()
    assert np.array_equal(d_out.cpu().numpy().reshape(V, N), orc.decode_rows(rows, N))


@pytest.mark.parametrize("precision", ["fp32", "bf16x4", "bf16x3", "bf16x2", "i8x4", "i8x3"])
def test_plain_small_config(orc, precision):
    """2,000 people x 50,000 variants x 1 score, HWE p~U(0.05,0.5), 1% missing,
    w ~ N(0, 0.01^2), effect allele A1 for a random half."""
    _flush(f"plain_small {precision}")
    N, V = 2000, 50_000
    rows = orc.synth_rows(1, N, 0, V, int(0.05 * 2**24), int(0.5 * 2**24), 655)
    rng = np.random.default_rng(2)
    csr = csr_from_dense(rng.normal(0, 0.01, (V, 1)).astype(np.float32), rng.random(V) < 0.5)
    S64, C64 = orc.score_fp64(rows, csr, N)
    S, C = _score(csr, rows, N, precision, blocks=3)
    _check_scores(S, C, S64, C64, precision, _fp32_terms(orc, precision, rows, csr, N))


@pytest.mark.parametrize("precision", ["fp32", "bf16x3", "i8x4"])
def test_many_scores_classes_and_mixed_flips(orc, precision):
    """K=70 with per-cell presence (membership classes > 4) and per-cell flips
    (variants needing both an A2 and an A1 row); N not a multiple of 4."""
    _flush(f"classes {precision}")
    rng = np.random.default_rng(7)
    N, V, K = 1234, 3000, 70
    dos = rng.choice([0, 1, 2, -1], size=(V, N), p=[0.5, 0.3, 0.17, 0.03])
    rows = pack_bed(dos)
    W = rng.laplace(0, 1e-3, (V, K)).astype(np.float32)
    W[rng.random((V, K)) < 0.3] = 0.0
    flip = rng.random((V, K)) < 0.5
    present = rng.random((V, K)) < 0.85
    present[:, :3] = True   # a few identical classes
    csr = csr_from_dense(W, flip, present)
    S64, C64 = orc.score_fp64(rows, csr, N)
    S, C = _score(csr, rows, N, precision, blocks=2)
    _check_scores(S, C, S64, C64, precision, _fp32_terms(orc, precision, rows, csr, N))
    # the restated reference engine agrees with both under the reference rule
    Sr, Cr, _ = orc.refcpu_run(rows, csr, N, threads=4)
    assert np.array_equal(Cr, C)
    assert avg_agree(S, Sr, C, csr, 1e-6).all()


@pytest.mark.parametrize("precision", ["fp32", "bf16x3", "i8x4"])
def test_person_subset_gather(orc, precision):
    _flush(f"subset {precision}")
    rng = np.random.default_rng(9)
    Nfam, V, K = 1001, 2500, 4
    dos = rng.choice([0, 1, 2, -1], size=(V, Nfam), p=[0.4, 0.35, 0.2, 0.05])
    rows = pack_bed(dos)
    o2f = rng.permutation(Nfam)[:600].astype(np.uint32)
    csr = csr_from_dense(rng.normal(0, 0.01, (V, K)).astype(np.float32), rng.random(V) < 0.5)
    S64, C64 = orc.score_fp64(rows, csr, len(o2f), o2f)
    S, C = _score(csr, rows, len(o2f), precision, nfam=Nfam, o2f=o2f)
    _check_scores(S, C, S64, C64, precision, _fp32_terms(orc, precision, rows, csr, len(o2f), o2f))


def _single(w_eff, corr, K=1):
    return ScoreCsr(weights=np.full(K, w_eff, np.float32), corrections=np.full(K, corr, np.float32),
                    columns=np.arange(K, dtype=np.uint32), row_offsets=np.array([0, K], np.uint64),
                    num_scores=K, baseline=np.full(K, corr, np.float64),
                    score_variant_counts=np.ones(K, np.uint32))


@pytest.mark.parametrize("precision", ["fp32", "bf16x3", "i8x4"])
def test_reference_golden_vectors_on_device(precision):
    _flush(f"golden {precision}")
    g = GOLD["variant_major_missing_correction"]
    S, C = _score(_single(g["w_eff"], g["corr"]), np.array([[g["byte"]]], np.uint8), 4, precision)
    assert np.allclose(S[:, 0] - g["corr"], g["partial_scores"], atol=1e-12)
    assert C[:, 0].tolist() == g["missing"]
    g = GOLD["word_skip_and_tail"]
    row = np.zeros((1, 26), np.uint8)
    for person, code in g["codes"]:
        row[0, person // 4] |= code << ((person % 4) * 2)
    S, C = _score(_single(g["w_eff"], g["corr"]), row, 101, precision)
    partial = S[:, 0] - g["corr"]
    for p, val in g["partial_scores"].items():
        assert partial[int(p)] == val
    assert np.count_nonzero(partial) == g["nonzero_scores"] and C.sum() == 1 and C[99, 0] == 1
    g = GOLD["many_scores"]
    S, C = _score(_single(1.0, 0.0, g["num_scores"]), pack_bed(np.array([g["dosages"]])), 3, precision)
    for p, expect in enumerate(g["scores"]):
        assert np.all(S[p] == expect)
    g = GOLD["special_case"]
    csr = csr_from_dense(np.full((50, 1), g["weight"], np.float32), np.ones(50, bool))
    S, C = _score(csr, pack_bed(np.array(g["a2_dosages"])[:, None]), 1, precision)
    assert abs(S[0, 0] - g["sum"]) < 1e-6 and C[0, 0] == 0


@pytest.mark.parametrize("precision", ["fp32", "bf16x4", "bf16x3", "i8x4"])
def test_split_self_test_known_answer(precision):
    """Inputs of the reference's split self-test (cuda_backend.rs:2432-2438):
    dosage[p][v] = (64p+v) % 3, weight[v][k] = ((7(64v+k)) mod 101 - 50)/13.
    The exact answer is an integer sum divided by 13."""
    n = 64
    i = np.arange(n * n)
    dos = (i % 3).reshape(n, n)                     # [people][variants]
    num = ((i * 7) % 101 - 50).reshape(n, n)        # [variants][scores], weight = num / 13
    W = (num / 13.0).astype(np.float32)
    csr = csr_from_dense(W, np.zeros(n, bool))
    S, C = _score(csr, pack_bed(dos.T), n, precision)
    exact = (dos.astype(np.int64) @ num.astype(np.int64)) / 13.0
    W64 = W.astype(np.float64)
    exact_f32w = dos.astype(np.float64) @ W64
    assert np.all(C == 0)
    if precision == "fp32":
        # FP32 sums of exact products over 64 variants
        assert np.all(np.abs(S - exact_f32w) <= fp32_sum_bound(dos.astype(np.float64) @ np.abs(W64)))
    else:
        # fixed-point limbs: the unit u = 2^e is the smallest power of two with
        # max|w| <= qmax u, so u < 2 max|w| / qmax and |w - Q u| <= u/2 <
        # max|w| / qmax per weight (qmax = 2^24 - 2^16 for three BF16 limbs,
        # 127 * 256^3 for four int8 limbs); |dS| <= that times sum_v d
        qmax = {"bf16x4": 2.0**32 - 2.0**24, "bf16x3": 2.0**24 - 2.0**16, "i8x4": 127.0 * 256.0**3}[precision]
        bound = np.abs(W64).max() / qmax * dos.sum(1, keepdims=True)
        assert np.all(np.abs(S - exact_f32w) <= bound + 1e-12)
        assert normwise_rel(S, exact_f32w).max() <= 1e-6
    assert np.abs(S - exact).max() < 1e-4   # f32 weights vs the rational thirteenths


def test_streaming_checkpoint_contract(orc):
    """After a block, read_accumulators returns what the reference checkpoints
    (baseline of all variants + contributions of the completed ones); a new
    context seeded from it finishes with the full result."""
    _flush("checkpoint")
    rng = np.random.default_rng(4)
    N, V, K = 777, 4000, 3
    dos = rng.choice([0, 1, 2, -1], size=(V, N), p=[0.4, 0.35, 0.2, 0.05])
    rows = pack_bed(dos)
    csr = csr_from_dense(rng.normal(0, 0.01, (V, K)).astype(np.float32), rng.random(V) < 0.5)
    S64, C64 = orc.score_fp64(rows, csr, N)
    cut = 1500
    with ScoringContext(csr, N, precision="bf16x3", force=True) as ctx:
        ctx.score_rows(rows[:cut], 0)
        Sck, Cck, done = ctx.read_accumulators()
    assert done == cut
    head = ScoreCsr(weights=csr.weights[:int(csr.row_offsets[cut])],
                    corrections=csr.corrections[:int(csr.row_offsets[cut])],
                    columns=csr.columns[:int(csr.row_offsets[cut])],
                    row_offsets=csr.row_offsets[:cut + 1], num_scores=K,
                    baseline=csr.baseline, score_variant_counts=csr.score_variant_counts)
    Sh, Ch = orc.score_fp64(rows[:cut], head, N)
    tail_base = np.zeros(K)
    np.add.at(tail_base, csr.columns[int(csr.row_offsets[cut]):].astype(int),
              csr.corrections[int(csr.row_offsets[cut]):].astype(np.float64))
    assert np.array_equal(Cck, Ch)
    assert reference_agree(Sck, Sh + tail_base, 1e-6).all()
    with ScoringContext(csr, N, precision="bf16x3", force=True) as ctx2:
        ctx2.seed_accumulators(Sck, Cck, cut)
        ctx2.score_rows(rows[cut:], cut)
        S, C = ctx2.finish()
    _check_scores(S, C, S64, C64, "bf16x3")


def test_try_run_hip_drop_in(orc):
    _flush("try_run_hip")
    rng = np.random.default_rng(12)
    N, V, K = 2000, 50_000, 1
    rows = orc.synth_rows(3, N, 0, V, int(0.05 * 2**24), int(0.5 * 2**24), 655)
    csr = csr_from_dense(rng.normal(0, 0.01, (V, K)).astype(np.float32), rng.random(V) < 0.5)
    out = try_run_hip(csr, rows, N)
    assert out is not None
    S64, C64 = orc.score_fp64(rows, csr, N)
    _check_scores(out[0], out[1], S64, C64, "bf16x3")
    # below the gate the backend declines
    assert try_run_hip(csr_from_dense(np.ones((10, 1), np.float32), np.zeros(10, bool)),
                       rows[:10], N) is None


def test_device_resident_rows_match_host_rows(orc):
    _flush("device-resident")
    rng = np.random.default_rng(21)
    N, V, K = 3000, 20_000, 8
    stride = 768
    d = torch.zeros(V * stride, dtype=torch.uint8, device="cuda")
    _lib.check(_lib.lib().gnm_synth_rows_device(5, 167772, 8388608, 1311, N, 0, V,
                                                ctypes.c_void_p(d.data_ptr()), stride, None))
    torch.cuda.synchronize()
    rows = orc.synth_rows(5, N, 0, V, 167772, 8388608, 1311)
    csr = csr_from_dense(rng.normal(0, 0.01, (V, K)).astype(np.float32), rng.random(V) < 0.5)
    S64, C64 = orc.score_fp64(rows, csr, N)
    for precision in ("fp32", "bf16x3", "i8x4"):
        with ScoringContext(csr, N, precision=precision, force=True) as ctx:
            ctx.score_rows_device(d.data_ptr(), stride, 0, V)
            S, C = ctx.finish()
            main_ms, total_ms = ctx.last_timing()
            assert 0 < main_ms <= total_ms
        _check_scores(S, C, S64, C64, precision, _fp32_terms(orc, precision, rows, csr, N))


@pytest.mark.parametrize("precision", ["i8x4", "i8x3"])
def test_int8_scale_blocks_and_launch_shapes(orc, precision):
    """150,000 variants = two full int8 scale blocks of 65,536 rows and a
    partial one; 700 people (ragged, one split per 2,344 rows, so splits start
    and end inside blocks); 24 scores with per-cell presence (6 weight tiles +
    a count tile: groups that mix weight and count tiles, count-only groups
    and unused tile slots, depending on the shape).  Every launch shape gives
    bit-identical results (exact integer block sums, the same FP64 order),
    and they match the FP64 oracle."""
    _flush(f"int8 blocks {precision}")
    rng = np.random.default_rng(31)
    N, V, K = 700, 150_000, 24
    rows = orc.synth_rows(8, N, 3, V, int(0.02 * 2**24), int(0.5 * 2**24), 1311)
    W = rng.normal(0, 0.01, (V, K)).astype(np.float32)
    W[60_000:70_000] *= 1e-3            # one block with much smaller weights
    present = rng.random((V, K)) < 0.9
    present[:, :4] = True
    csr = csr_from_dense(W, rng.random((V, K)) < 0.5, present)
    S64, C64 = orc.score_fp64(rows, csr, N)
    first = None
    for tp, bs in [(0, 0), (3, 2), (3, 4), (4, 4), (5, 2), (5, 4), (6, 2), (6, 4), (8, 2), (8, 4), (18, 2)]:
        _flush(f"  tiles_per_pass={tp} batch_stages={bs}")
        with ScoringContext(csr, N, precision=precision, force=True) as ctx:
            ctx.set_launch_config(tp, bs)
            for a, b in [(0, 40_000), (40_000, 131_072), (131_072, V)]:
                ctx.score_rows(rows[a:b], a)
            S, C = ctx.finish()
        if first is None:
            first = (S, C)
            _check_scores(S, C, S64, C64, precision, _fp32_terms(orc, precision, rows, csr, N))
        else:
            assert np.array_equal(C, first[1])
            assert np.array_equal(S, first[0]), f"shape {tp}x{bs} differs"


def test_launch_config_rejects_unknown_shapes(orc):
    """Tiles per pass outside {0, 3, 4, 5, 6, 8, 18}, batch depths outside
    {0, 2, 4}, and the two-set shape with four-stage batches are refused
    with an invalid-argument error; the context still scores afterwards."""
    from gnomon_amd._lib import GnmError
    rng = np.random.default_rng(41)
    N, V, K = 300, 2_000, 3
    rows = orc.synth_rows(8, N, 3, V, int(0.02 * 2**24), int(0.5 * 2**24), 1409)
    csr = csr_from_dense(rng.normal(0, 0.01, (V, K)).astype(np.float32), rng.random((V, K)) < 0.5,
                         np.ones((V, K), bool))
    S64, C64 = orc.score_fp64(rows, csr, N)
    with ScoringContext(csr, N, precision="i8x4", force=True) as ctx:
        for tp, bs in [(1, 0), (2, 2), (7, 2), (9, 2), (16, 2), (18, 4), (5, 3), (3, 8)]:
            with pytest.raises(GnmError):
                ctx.set_launch_config(tp, bs)
        ctx.set_launch_config(5, 2)
        ctx.score_rows(rows, 0)
        S, C = ctx.finish()
    _check_scores(S, C, S64, C64, "i8x4")


def test_int8_two_set_shape_many_scores(orc):
    """The two-set shape (two 9-tile groups over the same people, one row read
    for both) with 40 scores: 10 weight tiles and 3 count tiles, so the second
    set mixes weight and count tiles; 1,300 people (a ragged last person
    tile), two scale blocks.  Bit-identical to the default shape and within
    the i8x4 rule of the FP64 oracle."""
    _flush("int8 two-set shape")
    rng = np.random.default_rng(37)
    N, V, K = 1300, 70_000, 40
    rows = orc.synth_rows(9, N, 0, V, int(0.02 * 2**24), int(0.5 * 2**24), 1311)
    W = rng.laplace(0, 1e-3, (V, K)).astype(np.float32)
    W[rng.random((V, K)) < 0.3] = 0.0
    csr = csr_from_dense(W, rng.random((V, K)) < 0.5, rng.random((V, K)) < 0.85)
    S64, C64 = orc.score_fp64(rows, csr, N)
    out = {}
    for tp in (0, 18):
        with ScoringContext(csr, N, precision="i8x4", force=True) as ctx:
            ctx.set_launch_config(tp, 2 if tp else 0)
            ctx.score_rows(rows, 0)
            out[tp] = ctx.finish()
    _check_scores(*out[0], S64, C64, "i8x4")
    assert np.array_equal(out[18][1], out[0][1])
    assert np.array_equal(out[18][0], out[0][0])


@pytest.mark.parametrize("precision", ["bf16x4", "bf16x3"])
def test_bf16_scale_blocks_and_launch_shapes(orc, precision):
    """100,000 variants = three full split-BF16 scale blocks of 32,768 rows
    and a partial one; 700 people (ragged); 24 scores with per-cell presence
    (6 or 5 weight tiles and 2 count tiles: mixed, count-only and partly
    unused groups depending on the shape); rows in three blocks that start
    inside scale blocks.  Every tile-group size (3, 5, 6, 8, 9) gives
    bit-identical results (exact FP32 block sums, the same FP64 order), and
    they match the FP64 oracle."""
    _flush(f"bf16 blocks {precision}")
    rng = np.random.default_rng(33)
    N, V, K = 700, 100_000, 24
    rows = orc.synth_rows(8, N, 3, V, int(0.02 * 2**24), int(0.5 * 2**24), 1311)
    W = rng.normal(0, 0.01, (V, K)).astype(np.float32)
    W[40_000:50_000] *= 1e-3            # part of one block with much smaller weights
    present = rng.random((V, K)) < 0.9
    present[:, :4] = True
    csr = csr_from_dense(W, rng.random((V, K)) < 0.5, present)
    S64, C64 = orc.score_fp64(rows, csr, N)
    first = None
    for tp in (0, 3, 5, 6, 8, 9):
        _flush(f"  tiles_per_pass={tp}")
        with ScoringContext(csr, N, precision=precision, force=True) as ctx:
            ctx.set_launch_config(tp, 0)
            for a, b in [(0, 30_000), (30_000, 70_001), (70_001, V)]:
                ctx.score_rows(rows[a:b], a)
            S, C = ctx.finish()
        if first is None:
            first = (S, C)
            if precision == "bf16x4":
                _check_scores(S, C, S64, C64, precision, _fp32_terms(orc, precision, rows, csr, N))
            else:
                # 24 bits of a 32768-row block maximum: the block with 1000x
                # smaller weights keeps ~14 bits of them, so only the
                # normwise bound of the representation holds
                assert np.array_equal(C, C64)
                assert normwise_rel(S, S64).max() <= 1e-6
        else:
            assert np.array_equal(C, first[1])
            assert np.array_equal(S, first[0]), f"shape {tp} differs"


def test_bf16_launch_config_rejects_int8_shapes(orc):
    """The split-BF16 path takes tile groups of 3, 5, 6, 8 or 9 and two-stage
    batches; other shapes are refused and the context still scores."""
    from gnomon_amd._lib import GnmError
    rng = np.random.default_rng(43)
    N, V, K = 300, 2_000, 3
    rows = orc.synth_rows(8, N, 3, V, int(0.02 * 2**24), int(0.5 * 2**24), 1409)
    csr = csr_from_dense(rng.normal(0, 0.01, (V, K)).astype(np.float32), rng.random((V, K)) < 0.5,
                         np.ones((V, K), bool))
    S64, C64 = orc.score_fp64(rows, csr, N)
    with ScoringContext(csr, N, precision="bf16x4", force=True) as ctx:
        for tp, bs in [(4, 0), (7, 0), (18, 2), (3, 4), (2, 0)]:
            with pytest.raises(GnmError):
                ctx.set_launch_config(tp, bs)
        ctx.set_launch_config(5, 2)
        ctx.score_rows(rows, 0)
        S, C = ctx.finish()
    _check_scores(S, C, S64, C64, "bf16x4")


@pytest.mark.parametrize("precision", ["fp32", "bf16x3", "bf16x4", "i8x4"])
@pytest.mark.parametrize("N,V,K", [(1, 65, 1), (3, 7, 2), (5, 130, 3), (33, 1, 2), (63, 200, 5)])
def test_tiny_and_ragged_shapes(orc, precision, N, V, K):
    """Single person, fewer than 4 people (one partial byte), one variant,
    person counts straddling 32 and 64; every genotype code present."""
    _flush(f"tiny {precision} N={N} V={V} K={K}")
    rng = np.random.default_rng(N * 1000 + V + K)
    dos = rng.choice([0, 1, 2, -1], size=(V, N), p=[0.4, 0.3, 0.2, 0.1])
    rows = pack_bed(dos)
    W = rng.normal(0, 0.05, (V, K)).astype(np.float32)
    csr = csr_from_dense(W, rng.random((V, K)) < 0.5, rng.random((V, K)) < 0.8)
    S64, C64 = orc.score_fp64(rows, csr, N)
    S, C = _score(csr, rows, N, precision)
    _check_scores(S, C, S64, C64, precision, _fp32_terms(orc, precision, rows, csr, N))


@pytest.mark.parametrize("precision", ["fp32", "bf16x3", "i8x4"])
def test_degenerate_inputs(orc, precision):
    """Variants in no score, a score with no variants, an all-missing row, an
    all-00 row, a score whose weights are all zero (its missing counts still
    count), weights spanning 1e-30 .. 1e+3, an empty row block."""
    _flush(f"degenerate {precision}")
    rng = np.random.default_rng(77)
    N, V, K = 301, 3000, 6
    dos = rng.choice([0, 1, 2, -1], size=(V, N), p=[0.45, 0.3, 0.2, 0.05])
    dos[10] = -1
    dos[11] = 0
    rows = pack_bed(dos)
    W = rng.normal(0, 0.01, (V, K)).astype(np.float32)
    W[:, 2] = 0.0                                        # all-zero score
    W[100:200, 3] = np.float32(1e-30)                    # tiny weights
    W[200:210, 4] = np.float32(1e3)                      # large weights
    present = rng.random((V, K)) < 0.9
    present[:, 5] = False                                # a score with no variants
    present[500:600] = False                             # variants in no score
    csr = csr_from_dense(W, rng.random((V, K)) < 0.5, present)
    S64, C64 = orc.score_fp64(rows, csr, N)
    with ScoringContext(csr, N, precision=precision, force=True) as ctx:
        ctx.score_rows(rows[:0], 0)                      # empty block: no-op
        ctx.score_rows(rows[:1500], 0)
        ctx.score_rows(rows[1500:1500], 1500)
        ctx.score_rows(rows[1500:], 1500)
        S, C = ctx.finish()
    assert np.array_equal(C, C64)
    assert (C[:, 5] == 0).all() and (S[:, 5] == 0).all()
    assert (C[:, 2] > 0).any()
    if precision == "fp32":
        _check_scores(S, C, S64, C64, precision, _fp32_terms(orc, precision, rows, csr, N))
    else:
        # normwise per score, skipping the empty score
        assert normwise_rel(S[:, :5], S64[:, :5]).max() <= TOL[precision]


def test_forty_people_hundred_scores_partial_overlap(orc):
    """40 people x 100 scores with partial overlap (each score covers a random
    40-90% of 30,000 variants) and per-cell effect-allele flips: the shape of
    the reference's forty_genomes_hundred_scores test (score/tests/
    gpu_tests.rs:322-411), above the device size gate through N*K*V."""
    _flush("forty x hundred")
    rng = np.random.default_rng(101)
    N, V, K = 40, 30_000, 100
    rows = orc.synth_rows(12, N, 0, V, int(0.05 * 2**24), int(0.5 * 2**24), 1311)
    W = rng.normal(0, 0.02, (V, K)).astype(np.float32)
    cover = rng.uniform(0.4, 0.9, K)
    present = rng.random((V, K)) < cover[None, :]
    csr = csr_from_dense(W, rng.random((V, K)) < 0.5, present)
    assert _lib.lib().gnm_size_gate(N, K, V) == 1
    S64, C64 = orc.score_fp64(rows, csr, N)
    Sr, Cr, _ = orc.refcpu_run(rows, csr, N, threads=4)
    for precision in ("fp32", "i8x4"):
        S, C = _score(csr, rows, N, precision, blocks=2)
        _check_scores(S, C, S64, C64, precision, _fp32_terms(orc, precision, rows, csr, N))
        assert np.array_equal(C, Cr)
        # the reference's GPU/CPU rule on the AVG values it compares
        assert avg_agree(S, Sr, C, csr, 1e-6).all()
        assert avg_agree(S, S64, C, csr, 1e-6).all()


@pytest.mark.parametrize("precision", ["fp32", "bf16x3", "i8x4"])
def test_empty_and_mixed_flip_variants_in_one_block(orc, precision):
    """A variant in no score (no program row) and a variant whose cells mix
    effect alleles (two program rows) in the same block: the row count equals
    the variant count, but row i is not variant i."""
    _flush(f"empty+mixed {precision}")
    rng = np.random.default_rng(5)
    N, V, K = 200, 64, 2
    dos = rng.choice([0, 1, 2, -1], size=(V, N), p=[0.4, 0.3, 0.2, 0.1])
    rows = pack_bed(dos)
    W = rng.normal(0, 0.05, (V, K)).astype(np.float32)
    flip = np.zeros((V, K), bool)
    flip[9, 0] = True                     # variant 9: score 0 reads A1, score 1 reads A2
    present = np.ones((V, K), bool)
    present[5] = False                    # variant 5 is in no score
    csr = csr_from_dense(W, flip, present)
    S64, C64 = orc.score_fp64(rows, csr, N)
    S, C = _score(csr, rows, N, precision)
    _check_scores(S, C, S64, C64, precision, _fp32_terms(orc, precision, rows, csr, N))


@pytest.mark.parametrize("precision", ["fp32", "bf16x3", "i8x4"])
def test_device_rows_odd_stride_unaligned_base_and_subset(orc, precision):
    """Device-resident rows at an odd stride (bytes_per_variant itself) and
    from an odd base address take the byte-load kernel variants; the same
    rows with a person subset go through the on-device compaction."""
    _flush(f"device odd stride {precision}")
    rng = np.random.default_rng(61)
    Nfam, V, K = 1001, 4000, 5
    dos = rng.choice([0, 1, 2, -1], size=(V, Nfam), p=[0.4, 0.3, 0.25, 0.05])
    rows = pack_bed(dos)                              # stride = bpv = 251
    W = rng.normal(0, 0.02, (V, K)).astype(np.float32)
    csr = csr_from_dense(W, rng.random((V, K)) < 0.5, rng.random((V, K)) < 0.9)
    buf = torch.zeros(rows.size + 1, dtype=torch.uint8, device="cuda")
    buf[1:] = torch.from_numpy(rows.reshape(-1)).cuda()     # rows start at an odd address
    base = buf.data_ptr() + 1
    torch.cuda.synchronize()
    S64, C64 = orc.score_fp64(rows, csr, Nfam)
    with ScoringContext(csr, Nfam, precision=precision, force=True) as ctx:
        ctx.score_rows_device(base, rows.shape[1], 0, 1700)
        ctx.score_rows_device(base + 1700 * rows.shape[1], rows.shape[1], 1700, V - 1700)
        S, C = ctx.finish()
    _check_scores(S, C, S64, C64, precision, _fp32_terms(orc, precision, rows, csr, Nfam))
    o2f = rng.permutation(Nfam)[:777].astype(np.uint32)
    S64s, C64s = orc.score_fp64(rows, csr, len(o2f), o2f)
    with ScoringContext(csr, len(o2f), Nfam, o2f, precision=precision, force=True) as ctx:
        ctx.score_rows_device(base, rows.shape[1], 0, V)
        S, C = ctx.finish()
    _check_scores(S, C, S64s, C64s, precision, _fp32_terms(orc, precision, rows, csr, len(o2f), o2f))
