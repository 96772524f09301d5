"""The allocation-free adjoint (gnm_dense_adjoint_device_noalloc) against the
allocating one (gnm_dense_adjoint_device) through the C ABI: the same kernels
on the same operands, so dW must be bit-identical.  Two calls in a row, the
second smaller and at another precision and score count, show that what the
first call left in the shared scratch does not leak into the second.  Size
and precision refusals are host-side checks that return before any launch."""
import ctypes

import numpy as np
import pytest
import torch

from gnomon_amd import _lib
from helpers import pack_bed

pytestmark = pytest.mark.gpu


def _flush(msg):
    print(f"\n[case] {msg}", flush=True)


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _operands(N, V, K, seed):
    rng = np.random.default_rng(seed)
    dos = rng.choice([0, 1, 2, -1], size=(V, N), p=[0.45, 0.35, 0.17, 0.03])
    rows = pack_bed(dos)
    rows = np.pad(rows, ((0, 0), (0, -rows.shape[1] % 4)))   # the int8 adjoint reads dwords
    flip = (rng.random(V) < 0.5).astype(np.uint8)
    R = rng.normal(0, 1, (N, K)).astype(np.float32)
    R[:, 1] *= 1e-6                                          # per-column scales
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    return dev(rows), dev(flip), dev(R)


def _adjoint(entry, rows, flip, R, V, precision):
    N, K = R.shape
    dW = torch.full((V, K), float("nan"), dtype=torch.float32, device=R.device)
    st = getattr(_lib.lib(), entry)(_ptr(rows), rows.shape[1], V, N, _ptr(flip), _ptr(R), K,
                                    _lib.PRECISION[precision], _ptr(dW), _stream())
    _lib.check(st)
    torch.cuda.synchronize()
    return dW.cpu().numpy()


def test_noalloc_adjoint_matches_allocating_entry_twice():
    cases = [(70_000, 3000, 40, "i8x4", 21),   # two int8 person blocks, 12 limb tiles
             (517, 1100, 16, "i8x3", 22)]      # then smaller: fewer people, variants, tiles
    for N, V, K, precision, seed in cases:
        _flush(f"noalloc adjoint {precision} N={N} V={V} K={K}")
        rows, flip, R = _operands(N, V, K, seed)
        ref = _adjoint("gnm_dense_adjoint_device", rows, flip, R, V, precision)
        got = _adjoint("gnm_dense_adjoint_device_noalloc", rows, flip, R, V, precision)
        assert np.isfinite(ref).all()
        assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))


def test_noalloc_adjoint_refuses_before_launch():
    _flush("noalloc adjoint refusals")
    rows, flip, R = _operands(64, 8, 16, 23)
    L = _lib.lib()
    dW = torch.zeros((8, 16), dtype=torch.float32, device="cuda")
    args = (_ptr(rows), rows.shape[1])
    # 2M variants x 16 scores need a 256 MB FP64 accumulator: above the fixed scratch
    st = L.gnm_dense_adjoint_device_noalloc(*args, 2_000_000, 64, _ptr(flip), _ptr(R), 16,
                                            _lib.PRECISION["i8x4"], _ptr(dW), _stream())
    assert st == _lib.GNM_ERR_INVALID_ARG and "scratch" in _lib.last_error()
    st = L.gnm_dense_adjoint_device_noalloc(*args, 8, 64, _ptr(flip), _ptr(R), 16,
                                            _lib.PRECISION["bf16x3"], _ptr(dW), _stream())
    assert st == _lib.GNM_ERR_UNSUPPORTED
    torch.cuda.synchronize()
    assert not dW.any()
