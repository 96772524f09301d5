"""ctypes wrapper over the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()``
and by the ``cpu_baseline`` leg of ``bench.py`` -- never by the product package
``gnomon_amd``.  Every function restates a reference function (cited in
``gnomon_oracle.c``) or is a closed form used to check the HIP path.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_u8p = ctypes.POINTER(ctypes.c_uint8)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_u64p = ctypes.POINTER(ctypes.c_uint64)
_f32p = ctypes.POINTER(ctypes.c_float)
_f64p = ctypes.POINTER(ctypes.c_double)
_u64 = ctypes.c_uint64


def build() -> None:
    """Compile liboracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-B", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_mix64.restype = ctypes.c_uint64
        L.orc_mix64.argtypes = [ctypes.c_uint64]
        L.orc_synth_rows.restype = None
        L.orc_synth_rows.argtypes = [_u64, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                     _u64, _u64, _u64, _u8p]
        L.orc_decode_rows.restype = None
        L.orc_decode_rows.argtypes = [_u8p, _u64, _u64, _u32p, _u64, _u8p]
        L.orc_score_fp64.restype = None
        L.orc_score_fp64.argtypes = [_u8p, _u64, _u64, _u32p, _u64, _u64,
                                     _f32p, _f32p, _u32p, _u64p, _f64p, _u32p]
        L.orc_decide_path.restype = ctypes.c_int
        L.orc_decide_path.argtypes = [ctypes.c_float] * 4
        L.orc_variant_density.restype = ctypes.c_float
        L.orc_variant_density.argtypes = [_u8p, _u64, _u64]
        L.orc_refcpu_run.restype = ctypes.c_int
        L.orc_refcpu_run.argtypes = [_u8p, _u64, _u64, _u32p, _u64, _u64, _u64,
                                     _f32p, _f32p, _u32p, _u64p, _f64p,
                                     ctypes.c_int, ctypes.c_int, _f64p, _u32p]
        L.orc_adjoint_fp64.restype = None
        L.orc_adjoint_fp64.argtypes = [_u8p, _u64, _u64, _u64, _u8p, _f64p, _u64, _f64p]
        L.orc_dense_forward_fp64.restype = None
        L.orc_dense_forward_fp64.argtypes = [_u8p, _u64, _u64, _u64, _u8p, _f32p, _u64,
                                             _f64p, _u32p]
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(t) if a is not None else None


def mix64(z: int) -> int:
    return int(lib().orc_mix64(z))


def synth_rows(seed, num_samples, first_variant, num_rows, p_lo24, p_hi24, miss16):
    bpv = (num_samples + 3) // 4
    out = np.zeros((num_rows, bpv), dtype=np.uint8)
    lib().orc_synth_rows(seed, p_lo24, p_hi24, miss16, num_samples, first_variant, num_rows,
                         _p(out, _u8p))
    return out


def decode_rows(rows, num_people, out_to_fam=None):
    rows = np.ascontiguousarray(rows, dtype=np.uint8)
    o2f = None if out_to_fam is None else np.ascontiguousarray(out_to_fam, dtype=np.uint32)
    out = np.zeros((rows.shape[0], num_people), dtype=np.uint8)
    lib().orc_decode_rows(_p(rows, _u8p), rows.shape[1], rows.shape[0], _p(o2f, _u32p),
                          num_people, _p(out, _u8p))
    return out


def score_fp64(rows, csr, num_people, out_to_fam=None):
    """FP64 closed-form scores (baseline included) and missing counts."""
    rows = np.ascontiguousarray(rows, dtype=np.uint8)
    o2f = None if out_to_fam is None else np.ascontiguousarray(out_to_fam, dtype=np.uint32)
    K = csr.num_scores
    S = np.zeros((num_people, K), dtype=np.float64)
    C = np.zeros((num_people, K), dtype=np.uint32)
    lib().orc_score_fp64(_p(rows, _u8p), rows.shape[1], csr.num_variants, _p(o2f, _u32p),
                         num_people, K, _p(csr.weights, _f32p), _p(csr.corrections, _f32p),
                         _p(csr.columns, _u32p), _p(csr.row_offsets, _u64p),
                         _p(S, _f64p), _p(C, _u32p))
    return S, C


def score_fp64_parallel(rows, csr, num_people, threads=0):
    """score_fp64 over slices of people in a thread pool (the C entry point
    runs without the GIL); identical results to one score_fp64 call."""
    import os
    from concurrent.futures import ThreadPoolExecutor

    threads = threads or min(16, os.cpu_count() or 1)
    edges = np.linspace(0, num_people, threads + 1).astype(np.int64)
    parts = [(int(a), int(b)) for a, b in zip(edges[:-1], edges[1:]) if b > a]

    def run(ab):
        a, b = ab
        return score_fp64(rows, csr, b - a, np.arange(a, b, dtype=np.uint32))

    with ThreadPoolExecutor(len(parts)) as ex:
        res = list(ex.map(run, parts))
    return np.concatenate([r[0] for r in res]), np.concatenate([r[1] for r in res])


def refcpu_run(rows, csr, num_people, total_people_in_fam=None, out_to_fam=None,
               threads=0, force_path=-1):
    """Restated reference CPU engine. Returns (scores, counts, threads_used)."""
    rows = np.ascontiguousarray(rows, dtype=np.uint8)
    o2f = None if out_to_fam is None else np.ascontiguousarray(out_to_fam, dtype=np.uint32)
    if total_people_in_fam is None:
        total_people_in_fam = num_people
    K = csr.num_scores
    S = np.zeros((num_people, K), dtype=np.float64)
    C = np.zeros((num_people, K), dtype=np.uint32)
    base = np.ascontiguousarray(csr.baseline, dtype=np.float64)
    used = lib().orc_refcpu_run(_p(rows, _u8p), rows.shape[1], csr.num_variants,
                                _p(o2f, _u32p), num_people, total_people_in_fam, K,
                                _p(csr.weights, _f32p), _p(csr.corrections, _f32p),
                                _p(csr.columns, _u32p), _p(csr.row_offsets, _u64p),
                                _p(base, _f64p), threads, force_path, _p(S, _f64p), _p(C, _u32p))
    return S, C, used


def decide_path(n_cohort, k_scores, subset_frac, freq):
    return int(lib().orc_decide_path(n_cohort, k_scores, subset_frac, freq))


def variant_density(row, total_people):
    row = np.ascontiguousarray(row, dtype=np.uint8)
    return float(lib().orc_variant_density(_p(row, _u8p), row.shape[0], total_people))


def dense_forward_fp64(rows, num_people, flip, W):
    rows = np.ascontiguousarray(rows, dtype=np.uint8)
    W = np.ascontiguousarray(W, dtype=np.float32)
    fl = None if flip is None else np.ascontiguousarray(flip, dtype=np.uint8)
    V, K = W.shape
    S = np.zeros((num_people, K), dtype=np.float64)
    M = np.zeros(num_people, dtype=np.uint32)
    lib().orc_dense_forward_fp64(_p(rows, _u8p), rows.shape[1], V, num_people, _p(fl, _u8p),
                                 _p(W, _f32p), K, _p(S, _f64p), _p(M, _u32p))
    return S, M


def adjoint_fp64(rows, num_people, flip, R):
    rows = np.ascontiguousarray(rows, dtype=np.uint8)
    R = np.ascontiguousarray(R, dtype=np.float64)
    fl = None if flip is None else np.ascontiguousarray(flip, dtype=np.uint8)
    V = rows.shape[0]
    K = R.shape[1]
    out = np.zeros((V, K), dtype=np.float64)
    lib().orc_adjoint_fp64(_p(rows, _u8p), rows.shape[1], V, num_people, _p(fl, _u8p),
                           _p(R, _f64p), K, _p(out, _f64p))
    return out
