"""Host-side mirror of the reference's GPU entry point over the C ABI.

``try_run_hip`` keeps the contract of ``cuda_backend::try_run_cuda``
(score/cuda_backend.rs:604-656, called from score/pipeline.rs:480):

* ``None``  -- the backend declines (below the size gate, no device, or the
  device plan cannot be built): the caller runs its CPU pipeline;
* ``(scores, counts)`` -- f64 [N, K] sums (baseline included) and u32 [N, K]
  missing counts, [person][score] layout;
* an exception (``GnmError``) for a failure after computation started
  (the reference's ``Err(PipelineError::Compute)``).

``ScoringContext`` exposes the streaming form (blocks of rows, checkpoint
read/seed) used by the streaming ingest of score/cuda_backend.rs:1827-1978.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import GnmError, Problem, check, lib
from .prep import ScoreCsr


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


class ScoringContext:
    """Owns one ``gnm_ctx``.  Arrays passed in are kept alive for its lifetime."""

    def __init__(self, csr: ScoreCsr, num_people: int, total_people_in_fam: int | None = None,
                 out_to_fam=None, precision: str = "auto", device: int = 0, force: bool = False):
        L = lib()
        self.csr = csr
        self.precision = precision
        self.N = int(num_people)
        self.Nfam = int(total_people_in_fam if total_people_in_fam is not None else num_people)
        self.K = csr.num_scores
        self.V = csr.num_variants
        self.bpv = (self.Nfam + 3) // 4
        self._keep = [csr.weights, csr.corrections, csr.columns, csr.row_offsets]
        o2f = None
        if out_to_fam is not None:
            o2f = np.ascontiguousarray(out_to_fam, dtype=np.uint32)
            self._keep.append(o2f)
        pr = Problem()
        pr.num_people = self.N
        pr.total_people_in_fam = self.Nfam
        pr.bytes_per_variant = self.bpv
        pr.num_scores = self.K
        pr.num_variants = self.V
        pr.nnz = csr.nnz
        pr.sparse_weights = _ptr(csr.weights)
        pr.sparse_missing_corrections = _ptr(csr.corrections)
        pr.sparse_score_columns = _ptr(csr.columns)
        pr.sparse_row_offsets = _ptr(csr.row_offsets)
        pr.output_idx_to_fam_idx = _ptr(o2f)
        pr.precision = _lib.PRECISION[precision]
        pr.device = device
        self._problem = pr
        self._ctx = ctypes.c_void_p()
        create = L.gnm_create_forced if force else L.gnm_create
        st = create(ctypes.byref(pr), ctypes.byref(self._ctx))
        if st != _lib.GNM_OK:
            raise GnmError(st, _lib.last_error())

    def launch_shape(self):
        """(program rows, MFMA tiles per row) of the scoring launch; one pass
        executes 2 * N * rows * 16 * tiles matrix operations (tiles 0: FP32 path)."""
        rows, tiles = ctypes.c_uint64(), ctypes.c_uint32()
        check(lib().gnm_context_shape(self._ctx, ctypes.byref(rows), ctypes.byref(tiles)))
        return int(rows.value), int(tiles.value)

    # -- rows ---------------------------------------------------------------
    def score_rows(self, rows: np.ndarray, first_variant: int) -> None:
        rows = np.ascontiguousarray(rows, dtype=np.uint8)
        if rows.ndim != 2 or rows.shape[1] < self.bpv:
            raise ValueError("rows must be [n, >= bytes_per_variant] uint8")
        check(lib().gnm_score_rows(self._ctx, _ptr(rows), rows.shape[1], first_variant, rows.shape[0]))

    def score_rows_device(self, d_rows_ptr: int, row_stride: int, first_variant: int, num_rows: int) -> None:
        check(lib().gnm_score_rows_device(self._ctx, ctypes.c_void_p(d_rows_ptr), row_stride,
                                          first_variant, num_rows))

    def set_launch_config(self, tiles_per_pass: int = 0, batch_stages: int = 0) -> None:
        """Launch shape of the split-int8 kernel (results do not depend on it)."""
        check(lib().gnm_set_launch_config(self._ctx, tiles_per_pass, batch_stages))

    def set_sparse_ingest(self, max_word_fraction: float) -> None:
        """Rows of score_rows in which at most this fraction of the 64-bit
        words hold a non-zero call cross PCIe as a bitmap plus those words
        (0 = every row packed; default 1/2).  Results do not change."""
        check(lib().gnm_set_sparse_ingest(self._ctx, float(max_word_fraction)))

    def ingest_stats(self) -> dict:
        """Totals of score_rows since creation: rows sent packed, rows sent
        sparse, their stored words, bytes copied to the device."""
        v = [ctypes.c_uint64() for _ in range(4)]
        check(lib().gnm_ingest_stats(self._ctx, *(ctypes.byref(x) for x in v)))
        return dict(zip(("dense_rows", "sparse_rows", "stored_words", "bytes_to_device"), (int(x.value) for x in v)))

    def last_timing(self):
        a, b = ctypes.c_float(), ctypes.c_float()
        check(lib().gnm_last_timing(self._ctx, ctypes.byref(a), ctypes.byref(b)))
        return float(a.value), float(b.value)

    # -- accumulators -------------------------------------------------------
    def read_accumulators(self):
        S = np.zeros((self.N, self.K), np.float64)
        C = np.zeros((self.N, self.K), np.uint32)
        done = ctypes.c_uint64()
        check(lib().gnm_read_accumulators(self._ctx, _ptr(S), _ptr(C), ctypes.byref(done)))
        return S, C, int(done.value)

    def seed_accumulators(self, scores, counts, completed_variants: int) -> None:
        S = np.ascontiguousarray(scores, dtype=np.float64)
        C = np.ascontiguousarray(counts, dtype=np.uint32)
        check(lib().gnm_seed_accumulators(self._ctx, _ptr(S), _ptr(C), completed_variants))

    def finish(self):
        S = np.zeros((self.N, self.K), np.float64)
        C = np.zeros((self.N, self.K), np.uint32)
        check(lib().gnm_finish(self._ctx, _ptr(S), _ptr(C)))
        return S, C

    def close(self) -> None:
        if self._ctx:
            lib().gnm_destroy(self._ctx)
            self._ctx = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def try_run_hip(csr: ScoreCsr, rows: np.ndarray, num_people: int, total_people_in_fam: int | None = None,
                out_to_fam=None, precision: str = "auto", device: int = 0):
    """Drop-in for try_run_cuda over an in-memory .bed image (rows [V, bpv])."""
    L = lib()
    if not L.gnm_size_gate(num_people, csr.num_scores, csr.num_variants):
        return None  # "Backend: CPU fallback (problem size below threshold)"
    if L.gnm_device_count() == 0:
        return None  # "Backend: CPU fallback (no device detected)"
    try:
        ctx = ScoringContext(csr, num_people, total_people_in_fam, out_to_fam, precision, device)
    except GnmError as e:
        if e.status in (_lib.GNM_DECLINED, _lib.GNM_ERR_NO_DEVICE, _lib.GNM_ERR_OUT_OF_MEMORY,
                        _lib.GNM_ERR_UNSUPPORTED):
            return None  # init failure degrades to the CPU path (cuda_backend.rs:624-634)
        raise
    with ctx:
        ctx.score_rows(rows, 0)
        return ctx.finish()
