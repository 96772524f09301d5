"""Differentiable polygenic score on PyTorch-ROCm.

    S = PolygenicScore.apply(W, bed, flip, num_people, precision)

S[p, k] = sum_v g_pv W[v, k] with g the effect-allele dosage decoded from the
packed SNP-major .bed rows on the device (A2 dosage d, or 2 - d for a flipped
variant; a missing call contributes 0).  The forward runs the scoring kernels
(gnm_dense_forward_device); the backward is the adjoint kernel
(gnm_dense_adjoint_device), dL/dW = G^T dL/dS, which re-decodes G from the
.bed instead of storing dosages.  Both run on torch's current stream.

This path is new (the reference has no gradient); its forward semantics are
the dense form of the reference score with the baseline term dropped
(score/prepare.rs:197-208: flipped weights become -w plus a 2w baseline,
i.e. w times the A1 dosage 2 - d).
"""
from __future__ import annotations

import ctypes

import torch

from . import _lib

_PREC = {k: _lib.PRECISION[k] for k in ("fp32", "bf16x2", "bf16x3", "bf16x4", "i8x3", "i8x4", "auto")}


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _check_bed(bed, num_people, V):
    if bed.device.type != "cuda" or bed.dtype != torch.uint8 or bed.dim() != 2 or not bed.is_contiguous():
        raise ValueError("bed must be a contiguous uint8 CUDA tensor [variants, row_stride]")
    if bed.shape[0] != V:
        raise ValueError(f"bed has {bed.shape[0]} rows, weights have {V}")
    if bed.shape[1] < (num_people + 3) // 4:
        raise ValueError("bed rows are shorter than ceil(num_people / 4) bytes")


def _check_operand(name, t, bed, rows, cols=None):
    """t must be a 2-D (or, with cols None, 1-D) tensor on bed's device with
    the given shape; checked before any kernel is launched."""
    if not isinstance(t, torch.Tensor) or t.device != bed.device:
        raise ValueError(f"{name} must be a tensor on {bed.device}")
    shape = (rows,) if cols is None else (rows, cols)
    if tuple(t.shape) != shape:
        raise ValueError(f"{name} must have shape {shape}, got {tuple(t.shape)}")


def _check_flip(flip, bed, V):
    if flip is None:
        return
    _check_operand("flip", flip, bed, V)
    if flip.dtype not in (torch.uint8, torch.bool) or not flip.is_contiguous():
        raise ValueError("flip must be a contiguous uint8 or bool tensor (1 = effect allele is A1)")


def _check_precision(precision):
    if precision not in _PREC:
        raise ValueError(f"unknown precision {precision!r}; one of {sorted(_PREC)}")


def dense_forward(W, bed, flip, num_people, precision="auto"):
    """Scores [N, K] float64 and per-person missing counts [N] int32."""
    _check_precision(precision)
    if not isinstance(W, torch.Tensor) or W.dim() != 2:
        raise ValueError("W must be a 2-D tensor [variants, scores]")
    V, K = W.shape
    _check_bed(bed, num_people, V)
    _check_operand("W", W, bed, V, K)
    _check_flip(flip, bed, V)
    W = W.detach().to(torch.float32).contiguous()
    S = torch.empty(num_people, K, dtype=torch.float64, device=W.device)
    miss = torch.empty(num_people, dtype=torch.int32, device=W.device)
    _lib.check(_lib.lib().gnm_dense_forward_device(
        _ptr(bed), bed.shape[1], V, num_people, _ptr(flip), _ptr(W), K, _PREC[precision], _ptr(S), _ptr(miss),
        _stream()))
    return S, miss


def dense_adjoint(R, bed, flip, num_variants, precision="auto"):
    """dW [V, K] float32 = G^T R for R [N, K]."""
    _check_precision(precision)
    if not isinstance(R, torch.Tensor) or R.dim() != 2:
        raise ValueError("R must be a 2-D tensor [people, scores]")
    N, K = R.shape
    _check_bed(bed, N, num_variants)
    _check_operand("R", R, bed, N, K)
    _check_flip(flip, bed, num_variants)
    R = R.detach().to(torch.float32).contiguous()
    dW = torch.empty(num_variants, K, dtype=torch.float32, device=R.device)
    _lib.check(_lib.lib().gnm_dense_adjoint_device(
        _ptr(bed), bed.shape[1], num_variants, N, _ptr(flip), _ptr(R), K,
        _PREC[_adjoint_precision(precision, bed)], _ptr(dW), _stream()))
    return dW


def _adjoint_precision(precision, bed):
    """The adjoint's arithmetic for a forward precision: the same split when
    the rows are 4-byte aligned with a row stride that is a multiple of 4
    (what the split-limb adjoint kernel's row loads need); on other rows the
    int8 paths fall back to bf16x3 and bf16x4 to bf16x3 (the byte-load BF16
    kernel holds at most 12 limb tiles); "auto" is resolved by the library
    (split-int8 where the rows allow it and its self-test passed)."""
    if precision == "auto":
        return "auto"
    aligned = bed.data_ptr() % 4 == 0 and bed.shape[1] % 4 == 0
    if precision in ("i8x4", "i8x3", "bf16x4"):
        return precision if aligned else "bf16x3"
    return "bf16x2" if precision == "bf16x2" else "bf16x3"


class PolygenicScore(torch.autograd.Function):
    """S = G W on the device; backward re-decodes G from the .bed.

    W: [V, K] float32 CUDA tensor; bed: uint8 CUDA tensor [V, row_stride]
    (packed SNP-major rows); flip: uint8 CUDA tensor [V] (1 = effect allele
    is A1) or None.  Returns S [N, K] float64.  precision selects the forward
    arithmetic ("auto", "bf16x4", "bf16x3", "bf16x2", "i8x4", "i8x3" or
    "fp32"; "auto" is split-int8 after the library's self-test); the adjoint
    runs split-int8 with the int8 forwards, else split-BF16 (bf16x3 unless
    "bf16x2" is asked for).
    """

    @staticmethod
    def forward(ctx, W, bed, flip, num_people, precision="auto"):
        S, miss = dense_forward(W, bed, flip, num_people, precision)
        ctx.save_for_backward(bed, flip) if flip is not None else ctx.save_for_backward(bed)
        ctx.has_flip = flip is not None
        ctx.precision = precision
        ctx.V = W.shape[0]
        ctx.w_dtype = W.dtype
        return S

    @staticmethod
    def backward(ctx, grad_S):
        saved = ctx.saved_tensors
        bed, flip = (saved[0], saved[1]) if ctx.has_flip else (saved[0], None)
        dW = dense_adjoint(grad_S, bed, flip, ctx.V, ctx.precision)
        return dW.to(ctx.w_dtype), None, None, None, None


def polygenic_score(W, bed, flip, num_people, precision="auto"):
    return PolygenicScore.apply(W, bed, flip, num_people, precision)
