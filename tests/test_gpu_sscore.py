"""End to end on the GPU: scores from the HIP path written as a .sscore file
(score/main.rs:1389-1548), read back with the reference's reader (restated
in tests/test_sscore.py from score/tests/gpu_tests.rs:1202-1260), and
compared with the FP64 oracle's AVG and MISSING_PCT under the reference's
GPU/CPU agreement rule (gpu_tests.rs:134-162): |a - b| <= 1e-6 or
|a - b| / max(|a|, |b|, 1) <= 1e-6 for AVG, and MISSING_PCT within 1e-6.
"""
import numpy as np
import pytest

from gnomon_amd.pipeline import ScoringContext
from gnomon_amd.prep import csr_from_dense, sscore_columns
from gnomon_amd.sscore import write_sscore
from tests.helpers import reference_agree
from tests.test_sscore import _parse_sscore_table

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", ["bf16x4", "i8x4", "fp32"])
def test_sscore_from_device_scores(tmp_path, orc, precision):
    print(f"sscore end to end [{precision}]", flush=True)
    rng = np.random.default_rng(71)
    N, V, K = 1500, 20_000, 5
    rows = orc.synth_rows(5, N, 0, V, int(0.02 * 2**24), int(0.5 * 2**24), 2621)   # 4% missing
    W = rng.laplace(0, 1e-3, (V, K)).astype(np.float32)
    W[rng.random((V, K)) < 0.3] = 0.0
    present = rng.random((V, K)) < 0.8
    present[:, 0] = True
    csr = csr_from_dense(W, rng.random((V, K)) < 0.5, present)
    with ScoringContext(csr, N, precision=precision, force=True) as ctx:
        ctx.score_rows(rows, 0)
        S, C = ctx.finish()
    S64, C64 = orc.score_fp64_parallel(rows, csr, N)
    assert np.array_equal(C, C64)
    names = [f"PGS{k:06d}" for k in range(K)]
    iids = [f"id{p}" for p in range(N)]
    path = tmp_path / f"run_{precision}.sscore"
    write_sscore(path, iids, names, csr.score_variant_counts, S, C)
    got_names, table = _parse_sscore_table(path.read_text())
    assert got_names == names and list(table) == iids
    avg64, pct64 = sscore_columns(S64, C64, np.asarray(csr.score_variant_counts))
    avg = np.array([[table[i][k][0] for k in range(K)] for i in iids])
    pct = np.array([[table[i][k][1] for k in range(K)] for i in iids])
    assert reference_agree(avg, avg64, 1e-6).all()
    assert np.abs(pct - pct64.astype(np.float64)).max() <= 1e-6
