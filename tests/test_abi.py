"""C ABI checks that need no GPU: the library loads, exports every entry
point declared in include/gnomon_hip.h, and rejects bad input before it
touches a device."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

from gnomon_amd import _lib
from gnomon_amd.prep import csr_from_dense

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "gnomon_hip.h")).read()
    return sorted(set(re.findall(r"\b(gnm_[a-z0-9_]+)\s*\(", text)))


def test_header_and_binding_agree():
    assert _declared() == sorted(_lib.EXPORTS)


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    for name in _declared():
        assert hasattr(L, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (gnm_\w+)", out))
    assert set(_declared()) <= exported
    assert L.gnm_abi_version() == 1


def test_size_gate_matches_reference():
    L = _lib.lib()
    # cuda_backend.rs:38-39, 608-617: N*K >= 1e5 or N*K*V >= 1e8
    assert L.gnm_size_gate(2000, 1, 50000) == 1          # 1e8 exactly
    assert L.gnm_size_gate(2000, 1, 49999) == 0
    assert L.gnm_size_gate(100000, 1, 1) == 1
    assert L.gnm_size_gate(99999, 1, 1000) == 0
    assert L.gnm_size_gate(5, 3, 10) == 0


def test_devices_hidden_by_env(monkeypatch):
    L = _lib.lib()
    for val in ("", "-1", "none", " NONE "):
        monkeypatch.setenv("HIP_VISIBLE_DEVICES", val)
        assert L.gnm_device_count() == 0


def _problem(csr, N, **kw):
    pr = _lib.Problem()
    pr.num_people = N
    pr.total_people_in_fam = kw.get("nfam", N)
    pr.bytes_per_variant = kw.get("bpv", (pr.total_people_in_fam + 3) // 4)
    pr.num_scores = csr.num_scores
    pr.num_variants = csr.num_variants
    pr.nnz = csr.nnz
    pr.sparse_weights = csr.weights.ctypes.data
    pr.sparse_missing_corrections = csr.corrections.ctypes.data
    pr.sparse_score_columns = csr.columns.ctypes.data
    pr.sparse_row_offsets = csr.row_offsets.ctypes.data
    pr.precision = kw.get("precision", 2)
    return pr


def test_invalid_arguments_rejected_without_device():
    L = _lib.lib()
    csr = csr_from_dense(np.ones((4, 2), np.float32), np.zeros(4, bool))
    ctx = ctypes.c_void_p()
    assert L.gnm_create_forced(None, ctypes.byref(ctx)) == _lib.GNM_ERR_INVALID_ARG
    pr = _problem(csr, 10, bpv=2)  # wrong bytes_per_variant
    assert L.gnm_create_forced(ctypes.byref(pr), ctypes.byref(ctx)) == _lib.GNM_ERR_INVALID_ARG
    assert "bytes_per_variant" in _lib.last_error()
    bad = csr_from_dense(np.ones((4, 2), np.float32), np.zeros(4, bool))
    bad.columns[1] = 7
    pr = _problem(bad, 10)
    assert L.gnm_create_forced(ctypes.byref(pr), ctypes.byref(ctx)) == _lib.GNM_ERR_INVALID_ARG
    pr = _problem(csr, 10, precision=9)
    assert L.gnm_create_forced(ctypes.byref(pr), ctypes.byref(ctx)) == _lib.GNM_ERR_INVALID_ARG
    pr = _problem(csr, 10, nfam=12)  # subset without a person map
    assert L.gnm_create_forced(ctypes.byref(pr), ctypes.byref(ctx)) == _lib.GNM_ERR_INVALID_ARG
    assert not ctx.value


def test_unpack_rejects_more_people_than_its_grid_holds_without_device():
    L = _lib.lib()
    buf = ctypes.c_void_p(1 << 20)   # never dereferenced: the call fails its checks first
    N = 65535 * 16384 + 16
    assert L.gnm_unpack_device(buf, (N + 3) // 4, 1, N, None, buf, buf, None) == _lib.GNM_ERR_INVALID_ARG
    assert "limit" in _lib.last_error()


def test_hbm_probe_rejects_bad_arguments_without_device():
    L = _lib.lib()
    buf = ctypes.c_void_p(1 << 20)   # never dereferenced: every call below fails its checks first
    assert L.gnm_hbm_probe(3, buf, 4096, buf, 1 << 30, None) == _lib.GNM_ERR_INVALID_ARG
    assert "kind" in _lib.last_error()
    assert L.gnm_hbm_probe(0, buf, 4100, buf, 1 << 30, None) == _lib.GNM_ERR_INVALID_ARG
    assert L.gnm_hbm_probe(0, buf, 0, buf, 1 << 30, None) == _lib.GNM_ERR_INVALID_ARG
    assert L.gnm_hbm_probe(2, buf, 4096, ctypes.c_void_p((1 << 20) + 4), 1 << 30, None) == _lib.GNM_ERR_INVALID_ARG
    assert L.gnm_hbm_probe(2, buf, 4096, buf, 4 * 4096 - 16, None) == _lib.GNM_ERR_INVALID_ARG
    assert "too small" in _lib.last_error()
    assert L.gnm_hbm_probe(0, buf, 4096, buf, (16 << 20) - 16, None) == _lib.GNM_ERR_INVALID_ARG   # sink < 16 MiB
    assert L.gnm_hbm_probe(0, None, 4096, buf, 1 << 30, None) == _lib.GNM_ERR_INVALID_ARG


def test_small_problem_declines_and_no_device_is_reported(monkeypatch):
    L = _lib.lib()
    csr = csr_from_dense(np.ones((4, 2), np.float32), np.zeros(4, bool))
    ctx = ctypes.c_void_p()
    pr = _problem(csr, 10)
    assert L.gnm_create(ctypes.byref(pr), ctypes.byref(ctx)) == _lib.GNM_DECLINED
    monkeypatch.setenv("HIP_VISIBLE_DEVICES", "-1")
    assert L.gnm_create_forced(ctypes.byref(pr), ctypes.byref(ctx)) == _lib.GNM_ERR_NO_DEVICE
    from gnomon_amd import try_run_hip
    assert try_run_hip(csr, np.zeros((4, 3), np.uint8), 10) is None
    L.gnm_destroy(None)


def test_oracle_is_not_imported_by_the_product():
    for dirpath, _, files in os.walk(os.path.join(ROOT, "gnomon_amd")):
        for f in files:
            if f.endswith((".py", ".hip", ".hpp", ".cpp", ".h")):
                text = open(os.path.join(dirpath, f), errors="replace").read()
                assert "oracle" not in re.sub(r"#.*|//.*", "", text).replace("the oracle", ""), f


@pytest.mark.parametrize("field,value", [("weights", np.nan), ("weights", np.inf), ("corrections", -np.inf)])
def test_non_finite_weights_rejected_without_device(field, value):
    """Non-finite CSR weights or corrections are refused before any device
    work, as the reference refuses its split path for them
    (score/cuda_backend.rs:1178-1181)."""
    L = _lib.lib()
    csr = csr_from_dense(np.ones((4, 2), np.float32), np.array([False, True, False, True]))
    getattr(csr, field)[3] = value
    ctx = ctypes.c_void_p()
    for prec in (0, 3, 5, 6):
        pr = _problem(csr, 10, precision=prec)
        assert L.gnm_create_forced(ctypes.byref(pr), ctypes.byref(ctx)) == _lib.GNM_ERR_INVALID_ARG
        assert "not finite" in _lib.last_error()
    assert not ctx.value


def test_precision_codes():
    """bf16x4 is code 6; codes past it are refused."""
    assert _lib.PRECISION["bf16x4"] == 6
    L = _lib.lib()
    csr = csr_from_dense(np.ones((4, 2), np.float32), np.zeros(4, bool))
    ctx = ctypes.c_void_p()
    pr = _problem(csr, 10, precision=7)
    assert L.gnm_create_forced(ctypes.byref(pr), ctypes.byref(ctx)) == _lib.GNM_ERR_INVALID_ARG
    assert "unknown precision" in _lib.last_error()


def _rust_string_list(text, const_name):
    m = re.search(const_name + r"\s*:\s*\[&str;\s*(\d+)\]\s*=\s*\[(.*?)\];", text, re.S)
    assert m, const_name
    items = re.findall(r'"([^"]+)"', m.group(2))
    assert len(items) == int(m.group(1))
    return items


def test_rust_build_lists_every_source():
    """score/hip_build.rs and INTEGRATION.md compile what gnomon_amd/build.py
    compiles, with the same flags, so the Rust build links with no undefined
    symbols."""
    from gnomon_amd import build

    rs = open(os.path.join(ROOT, "score", "hip_build.rs")).read()
    assert _rust_string_list(rs, "SOURCES") == list(build.SOURCE_NAMES)
    assert _rust_string_list(rs, "COMPILE_FLAGS") == build.HIPCC_FLAGS + ["-c"]
    assert _rust_string_list(rs, "LINK_FLAGS") == build.LINK_FLAGS
    doc = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    listed = re.search(r"in link order:\s*```text\s*\n(.*?)\n```", doc, re.S)
    assert listed and listed.group(1).split() == list(build.SOURCE_NAMES)
    on_disk = sorted(f for f in os.listdir(os.path.join(ROOT, "gnomon_amd", "csrc")) if f.endswith(".hip"))
    assert sorted(build.SOURCE_NAMES) == on_disk


def test_rust_shim_declares_header_functions():
    """Every function score/hip_backend.rs binds is declared in the header and
    exported by the library; GnmProblem has gnm_problem's fields in order."""
    rs = open(os.path.join(ROOT, "score", "hip_backend.rs")).read()
    block = re.search(r'extern "C"\s*\{(.*?)\n\}', rs, re.S).group(1)
    bound = re.findall(r"fn (gnm_\w+)\s*\(", block)
    assert bound and set(bound) <= set(_declared())
    called = set(re.findall(r"\b(gnm_\w+)\s*\(", rs.replace(block, "")))
    assert called <= set(bound), called - set(bound)
    hdr = open(os.path.join(ROOT, "include", "gnomon_hip.h")).read()
    c_struct = re.search(r"typedef struct gnm_problem \{(.*?)\} gnm_problem;", hdr, re.S).group(1)
    c_fields = re.findall(r"\b(\w+)(?:\[\d+\])?;", re.sub(r"/\*.*?\*/", "", c_struct, flags=re.S))
    r_struct = re.search(r"struct GnmProblem \{(.*?)\}", rs, re.S).group(1)
    r_fields = re.findall(r"(\w+)\s*:", r_struct)
    assert r_fields == c_fields
    assert [f for f, _ in _lib.Problem._fields_] == c_fields
