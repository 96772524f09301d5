"""The host encoder of the sparse ingest (gnm_sparse_encode_row), on the CPU.

A sparse row is a bitmap of its non-zero 64-bit words (32 calls each) plus
those words.  The checks: rebuilding the row from bitmap and words gives back
the row with the padding calls of the last byte and the calls of unscored
people cleared (the device rebuild does the same, and the dense kernels read
only scored people); the calls read back through the oracle's decode (pinned
against the reference's ld.bed fixture) match the original row's for every
scored person; the word budget and the chunk starts.
"""
import ctypes

import numpy as np
import pytest

from gnomon_amd._lib import lib
from tests.helpers import pack_bed

CHUNK_WORDS = 512


def _encode(row, num_fam, max_words, keep2=None):
    row = np.ascontiguousarray(row, np.uint8)
    nch = (num_fam + 32 * CHUNK_WORDS - 1) // (32 * CHUNK_WORDS)
    words = np.zeros(max_words + 64, np.uint64)
    bitmap = np.zeros(nch * 8, np.uint64)
    cs = np.zeros(nch + 1, np.uint32)
    k = None if keep2 is None else np.ascontiguousarray(keep2, np.uint64)
    n = lib().gnm_sparse_encode_row(row.ctypes.data_as(ctypes.c_void_p), num_fam,
                                    None if k is None else k.ctypes.data_as(ctypes.c_void_p), max_words,
                                    words.ctypes.data_as(ctypes.c_void_p), bitmap.ctypes.data_as(ctypes.c_void_p),
                                    cs.ctypes.data_as(ctypes.c_void_p))
    return n, words[:max(n, 0)], bitmap, cs


def _rebuild(words, bitmap, nbytes):
    """Packed row from bitmap + stored words (the device rebuild, restated)."""
    nw = len(bitmap) * 64
    bits = ((bitmap[:, None] >> np.arange(64, dtype=np.uint64)) & np.uint64(1)).astype(bool).reshape(-1)
    full = np.zeros(nw, np.uint64)
    full[bits] = words
    return full.view(np.uint8)[:nbytes]


def _keep2(nfam, kept):
    k = np.zeros((nfam + 31) // 32, np.uint64)
    for f in kept:
        k[f // 32] |= np.uint64(3) << np.uint64(2 * (f % 32))
    return k


@pytest.mark.parametrize("num_fam", [1, 3, 31, 32, 33, 64, 1000, 4099, 16384, 16417, 40000])
@pytest.mark.parametrize("density", [0.0, 0.002, 0.03, 0.3])
def test_rebuild_gives_back_the_row(orc, num_fam, density):
    rng = np.random.default_rng(num_fam * 7 + int(density * 1e4))
    dos = np.where(rng.random(num_fam) < density, rng.choice([1, 2, -1], num_fam), 0)
    row = pack_bed(dos[None, :])[0]
    nw = (num_fam + 31) // 32
    n, words, bitmap, cs = _encode(row, num_fam, nw)
    assert n == sum(bin(int(b)).count("1") for b in bitmap)
    assert (words != 0).all()
    back = _rebuild(words, bitmap, len(row))
    assert np.array_equal(back, row)          # pack_bed leaves the padding calls at 00
    assert np.array_equal(orc.decode_rows(back[None, :], num_fam), orc.decode_rows(row[None, :], num_fam))
    # chunk starts count the stored words of the chunks before
    per_chunk = [sum(bin(int(b)).count("1") for b in bitmap[8 * c:8 * c + 8]) for c in range(len(cs) - 1)]
    assert list(cs) == list(np.concatenate([[0], np.cumsum(per_chunk)]))
    # the word budget: one word fewer sends the row packed
    if n > 0:
        assert _encode(row, num_fam, n - 1)[0] == -1
        assert _encode(row, num_fam, n)[0] == n


def test_padding_calls_of_the_last_byte_are_dropped():
    # 5 people: byte 1 holds person 4 in its low bits and three padding calls
    row = np.array([0b00000000, 0b11111100], np.uint8)
    n, words, bitmap, _ = _encode(row, 5, 4)
    assert n == 0 and bitmap[0] == 0
    row[1] = 0b11111110   # person 4 het, padding still set
    n, words, bitmap, _ = _encode(row, 5, 4)
    assert n == 1 and int(words[0]) == 0b10 << 8 and int(bitmap[0]) == 1


def test_unscored_people_are_cleared(orc):
    rng = np.random.default_rng(3)
    nfam = 20000
    dos = np.where(rng.random(nfam) < 0.01, rng.choice([1, 2, -1], nfam), 0)
    row = pack_bed(dos[None, :])[0]
    kept = np.sort(rng.permutation(nfam)[:9000])
    n, words, bitmap, _ = _encode(row, nfam, (nfam + 31) // 32, _keep2(nfam, kept))
    back = _rebuild(words, bitmap, len(row))
    got = orc.decode_rows(back[None, :], nfam)[0]
    want = orc.decode_rows(row[None, :], nfam)[0]
    mask = np.zeros(nfam, bool)
    mask[kept] = True
    assert np.array_equal(got[mask], want[mask])
    assert (got[~mask] == 0).all()
    # fewer stored words than without the mask
    assert n < _encode(row, nfam, (nfam + 31) // 32)[0]


def test_null_buffers():
    assert lib().gnm_sparse_encode_row(None, 10, None, 4, None, None, None) == -2


def test_vector_and_portable_encoders_agree(tmp_path):
    """The AVX2 encoder (used where the host has AVX2) and the portable one
    (forced with GNM_SPARSE_SCALAR=1, read once per process) give the same
    words, bitmaps and chunk starts, with and without a person mask."""
    import json
    import os
    import subprocess
    import sys

    script = r'''
import json, sys
import numpy as np
sys.path.insert(0, sys.argv[1])
from tests.test_sparse_encode import _encode, _keep2
from tests.helpers import pack_bed
out = []
rng = np.random.default_rng(11)
for nfam, d in [(5, 0.5), (130, 0.05), (16417, 0.003), (40000, 0.02), (33000, 0.2)]:
    dos = np.where(rng.random(nfam) < d, rng.choice([1, 2, -1], nfam), 0)
    row = pack_bed(dos[None, :])[0]
    kept = np.sort(rng.permutation(nfam)[: nfam // 2])
    for keep in (None, _keep2(nfam, kept)):
        n, w, b, c = _encode(row, nfam, (nfam + 31) // 32, keep)
        out.append([int(n), [int(x) for x in w], [int(x) for x in b], [int(x) for x in c]])
        out.append(int(_encode(row, nfam, max(int(n) - 1, 0), keep)[0]))
print(json.dumps(out))
'''
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    res = {}
    for scalar in (False, True):
        env = dict(os.environ)
        env.pop("GNM_SPARSE_SCALAR", None)
        if scalar:
            env["GNM_SPARSE_SCALAR"] = "1"
        p = subprocess.run([sys.executable, "-c", script, root], env=env, capture_output=True, text=True,
                           timeout=120, cwd=root)
        assert p.returncode == 0, p.stderr
        res[scalar] = json.loads(p.stdout.strip().splitlines()[-1])
    assert res[False] == res[True]
