"""The .sscore writer (score/main.rs:1386-1548) on the CPU.

Number formatting follows the ryu crate's Buffer::format (ryu 1.0.23 per the
reference's Cargo.lock).  It is pinned by the layout examples of each branch
of its algorithm, by round trips (the written text parses back to the exact
value, and no shorter text does), and by the reference's own reader
(parse_sscore_table, score/tests/gpu_tests.rs:1202-1260), restated here.
"""
import os

import numpy as np
import pytest

from gnomon_amd.prep import sscore_columns
from gnomon_amd.sscore import format_float, sscore_lines, write_sscore

# (value, f64 text, f32 text), one per layout branch and edge
KNOWN = [
    (0.0, "0.0", "0.0"),
    (-0.0, "-0.0", "-0.0"),
    (1.0, "1.0", "1.0"),
    (100.0, "100.0", "100.0"),
    (12.34, "12.34", "12.34"),
    (-12.34, "-12.34", "-12.34"),
    (0.001234, "0.001234", "0.001234"),
    (1.234e-5, "0.00001234", "0.00001234"),   # kk = -4: decimal for both types
    (1.5e-6, "1.5e-6", "0.0000015"),          # kk = -5: f64 turns scientific, f32 not yet
    (1.5e-7, "1.5e-7", "1.5e-7"),             # kk = -6: scientific for both
    (1e-7, "1e-7", "1e-7"),
    (12340000000.0, "12340000000.0", "12340000000.0"),
    (1e12, "1000000000000.0", "1000000000000.0"),   # kk = 13: the f32 limit
    (1e13, "10000000000000.0", "1e13"),       # kk = 14 > 13: f32 scientific
    (1e15, "1000000000000000.0", "1e15"),     # kk = 16: the f64 limit
    (1e16, "1e16", "1e16"),                   # kk = 17 > 16: f64 scientific
    (1.5e16, "1.5e16", "1.5e16"),
    (1.234e33, "1.234e33", "1.234e33"),
    (2.5, "2.5", "2.5"),
    (float("nan"), "NaN", "NaN"),
    (float("inf"), "inf", "inf"),
    (float("-inf"), "-inf", "-inf"),
]


@pytest.mark.parametrize("x,f64,f32", KNOWN)
def test_layout_known_answers(x, f64, f32):
    assert format_float(x, 64) == f64
    assert format_float(x, 32) == f32


def test_f32_shortest_digits():
    # 0.1 as f32 is 0.100000001490116...: the f32 text is "0.1", the f64 text of
    # the same (widened) value needs 17 digits
    x = float(np.float32(0.1))
    assert format_float(x, 32) == "0.1"
    assert format_float(x, 64) == "0.10000000149011612"
    assert format_float(1.0 / 3.0, 64) == "0.3333333333333333"
    assert format_float(np.float32(1.0) / np.float32(3.0), 32) == "0.33333334"


def _shorter_exists(x, text, cast):
    """Whether a text with fewer significant digits parses to the same value."""
    mant = text.split("e")[0].replace("-", "").replace(".", "").strip("0")
    n = len(mant)
    if n <= 1:
        return False
    for d in range(1, n):
        t = np.format_float_scientific(x, precision=d - 1, unique=False)
        if cast(float(t)) == x:
            return True
    return False


def test_round_trip_and_shortest():
    rng = np.random.default_rng(5)
    vals = np.concatenate([rng.normal(0, 1, 400), rng.laplace(0, 1e-3, 400),
                           10.0 ** rng.uniform(-12, 20, 400) * rng.choice([-1, 1], 400)])
    for x in vals:
        t = format_float(float(x), 64)
        assert float(t) == float(x)
        assert not _shorter_exists(float(x), t, float)
        x32 = np.float32(x)
        t32 = format_float(float(x32), 32)
        assert np.float32(float(t32)) == x32
        assert not _shorter_exists(x32, t32, np.float32)


def _parse_sscore_table(text):
    """Restatement of the reference's reader (gpu_tests.rs:1202-1260)."""
    lines = [ln for ln in text.split("\n") if not ln.startswith("#REGION")]
    header = lines[0].split("\t")
    assert header[0] == "#IID" and (len(header) - 1) % 2 == 0
    names = []
    for i in range(1, len(header), 2):
        assert header[i].endswith("_AVG")
        name = header[i][:-4]
        assert header[i + 1] == f"{name}_MISSING_PCT"
        names.append(name)
    rows = {}
    for ln in lines[1:]:
        if not ln.strip() or ln.startswith("#"):
            continue
        cols = ln.split("\t")
        assert len(cols) == len(header)
        rows[cols[0]] = [(float(cols[1 + 2 * k]), float(cols[2 + 2 * k])) for k in range(len(names))]
    return names, rows


def test_file_layout_and_values(tmp_path):
    rng = np.random.default_rng(9)
    N, K = 7, 3
    names = ["PGS_A", "PGS_B", "PGS_C"]
    counts_k = np.array([10, 0, 250], np.uint32)          # a score with no variants
    sums = rng.normal(0, 1, (N, K))
    miss = rng.integers(0, 11, (N, K)).astype(np.uint32)
    miss[:, 1] = 0
    miss[2, 0] = 10                                        # every variant missing: AVG 0
    iids = [f"s{i}" for i in range(N)]
    regions = {"PGS_C": "chr1:100-200"}
    path = tmp_path / "out.sscore"
    write_sscore(path, iids, names, counts_k, sums, miss, regions)
    text = path.read_text()
    assert text.startswith("#REGION\tSCORE\tINTERVAL\n#REGION\tPGS_C\tchr1:100-200\n#IID\tPGS_A_AVG\t")
    assert text.endswith("\n") and "\r" not in text
    assert [f for f in os.listdir(tmp_path) if f.endswith(".partial")] == []   # temp file renamed away
    got_names, rows = _parse_sscore_table(text)
    assert got_names == names and list(rows) == iids
    avg, pct = sscore_columns(sums, miss, counts_k)
    for p, iid in enumerate(iids):
        for k in range(K):
            a, m = rows[iid][k]
            assert a == avg[p, k]                         # f64 text round-trips exactly
            assert np.float32(m) == pct[p, k]             # f32 text round-trips exactly
    assert rows["s2"][0][0] == 0.0 and rows["s0"][1] == (0.0, 0.0)
    assert list(sscore_lines(iids, names, counts_k, sums, miss))[0].startswith("#IID")


def test_rejects_mismatched_rows():
    with pytest.raises(ValueError):
        list(sscore_lines(["a", "b"], ["S"], np.array([1]), np.zeros((3, 1)), np.zeros((3, 1), np.uint32)))
