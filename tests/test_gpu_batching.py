"""Many score files of uneven sizes in one grouped context, the shape of the
batching configuration scaled down: 3,000 people x 150,000 variants, 60
files whose variant counts are log-uniform in [1,000, 150,000], each drawn
as a random subset, weights ~ N(0, 0.01^2), the effect allele A1 for half
the variants.  All files are scored as one K-wide problem, as the reference
merges score files (score/prepare.rs:1947, score/cuda_backend.rs:2126).

Every file is compared with the FP64 oracle (counts bit-exact, scores under
the reference's rule and the per-sample 1e-6 bound), and every launch shape
the autotuner tries gives bit-identical results.
"""
import numpy as np
import pytest

from gnomon_amd import autotune, synth
from gnomon_amd.pipeline import ScoringContext
from gnomon_amd.prep import csr_from_dense
from tests.helpers import normwise_rel, per_sample_rel, reference_agree

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

N, V, FILES = 3000, 150_000, 60


@pytest.fixture(scope="module")
def c4(orc):
    w = synth.WORKLOADS["batching_200"].scaled(num_people=N, num_variants=V, files=FILES)
    W, flipped, present = synth.draw_weights(w)
    csr = csr_from_dense(W, flipped, present)
    rows = orc.synth_rows(w.seed, N, 0, V, synth.p24(w.p_lo), synth.p24(w.p_hi), synth.m16(w.missing))
    S64, C64 = orc.score_fp64_parallel(rows, csr, N)
    sizes = present.sum(0)
    assert sizes.min() < 5_000 and sizes.max() > 50_000      # uneven, as in the configuration
    return w, csr, rows, S64, C64


@pytest.mark.parametrize("precision", ["i8x4", "bf16x4"])
def test_grouped_files_match_oracle_and_every_shape_is_identical(c4, precision):
    w, csr, rows, S64, C64 = c4
    print(f"batching {FILES} files N={N} V={V} [{precision}]", flush=True)
    stride = (w.bytes_per_variant + 255) // 256 * 256
    d = torch.zeros(V, stride, dtype=torch.uint8, device="cuda")
    d[:, :rows.shape[1]] = torch.from_numpy(rows).cuda()
    torch.cuda.synchronize()
    with ScoringContext(csr, N, precision=precision, force=True) as ctx:
        S0, C0, _ = ctx.read_accumulators()
        tuned = autotune.tune(ctx, d.data_ptr(), stride, 40_000, reps=1)
        ctx.score_rows_device(d.data_ptr(), stride, 0, V)
        S, C, done = ctx.read_accumulators()
        assert done == V
        assert np.array_equal(C, C64), "missing counts differ from the oracle"
        assert reference_agree(S, S64, 1e-6).all()
        assert normwise_rel(S, S64).max() <= 1e-6
        assert per_sample_rel(S, S64).max() <= 1e-6
        for tp, bs in autotune.candidates_for(precision):
            ctx.set_launch_config(tp, bs)
            ctx.seed_accumulators(S0, C0, 0)
            ctx.score_rows_device(d.data_ptr(), stride, 0, V)
            S2, C2, _ = ctx.read_accumulators()
            assert np.array_equal(C2, C), f"counts differ at shape {tp}x{bs}"
            assert np.array_equal(S2, S), f"scores differ at shape {tp}x{bs}"
    print(f"[{precision}] autotuned shape {tuned['choice']}", flush=True)


def test_row_blocks_out_of_order_are_refused(orc):
    """A block that leaves a gap, repeats rows or starts before the last one
    ended is refused; starting over at variant 0 after a full pass is allowed."""
    from gnomon_amd._lib import GnmError
    rng = np.random.default_rng(5)
    n, v = 200, 3000
    rows = orc.synth_rows(3, n, 0, v, int(0.05 * 2**24), int(0.5 * 2**24), 655)
    csr = csr_from_dense(rng.normal(0, 0.01, (v, 2)).astype(np.float32), rng.random(v) < 0.5)
    S64, C64 = orc.score_fp64(rows, csr, n)
    with ScoringContext(csr, n, precision="i8x4", force=True) as ctx:
        ctx.score_rows(rows[:1000], 0)
        with pytest.raises(GnmError):
            ctx.score_rows(rows[1500:2000], 1500)      # gap
        with pytest.raises(GnmError):
            ctx.score_rows(rows[500:1000], 500)        # repeat
        with pytest.raises(GnmError):
            ctx.score_rows(rows[:1000], 0)             # restart before the pass is complete
        ctx.score_rows(rows[1000:], 1000)
        S, C = ctx.finish()
        assert np.array_equal(C, C64)
        assert reference_agree(S, S64, 1e-6).all()
        ctx.score_rows(rows, 0)                        # a second full pass accumulates
        S2, C2 = ctx.finish()
        assert np.array_equal(C2, 2 * C64)
