"""Differentiable path on the GPU: dense forward and adjoint against the FP64
oracle, and the torch.autograd.Function against central finite differences
and the FP64 CPU gradient.

Error bounds follow from the split-limb construction: a value x in a
(block, column) cell with maximum magnitude m is represented as an integer
multiple of the smallest power of two u with m <= qmax u (qmax = 2^(8L) -
2^(8(L-1)) for BF16 limbs, 127 * 256^(L-1) for int8 limbs), so u < 2 m / qmax
and |x - x_hat| <= u / 2 < m / qmax.  Dosages are exact, and
products and partial sums inside a block are exact (integers below 2^24), so
the contraction error is the sum of those representation errors weighted by
the dosages.
"""
import ctypes

import numpy as np
import pytest
import torch

from gnomon_amd import _lib
from gnomon_amd import autograd as ag
from helpers import fp32_sum_bound, pack_bed

pytestmark = pytest.mark.gpu

BLOCK = 8192   # rows (forward: variants; adjoint: people) per scale block
I8_BLOCK = 65536   # variants per scale block of the split-int8 forward
BF_ADJ_BLOCK = 32768   # people per scale block of the split-limb BF16 adjoint
LIMBS = {"bf16x4": 4, "bf16x3": 3, "bf16x2": 2, "i8x4": 4, "i8x3": 3}


def _qmax(precision):
    L = LIMBS[precision]
    return 127.0 * 256.0 ** (L - 1) if precision.startswith("i8") else 2.0 ** (8 * L) - 2.0 ** (8 * (L - 1))


def _flush(msg):
    print(f"\n[case] {msg}", flush=True)


def _data(N, V, K, seed, zero_frac=0.3):
    rng = np.random.default_rng(seed)
    dos = rng.choice([0, 1, 2, -1], size=(V, N), p=[0.45, 0.35, 0.17, 0.03])
    rows = pack_bed(dos)
    flip = (rng.random(V) < 0.5).astype(np.uint8)
    W = rng.laplace(0, 1e-3, (V, K)).astype(np.float32)
    W[rng.random((V, K)) < zero_frac] = 0.0
    return dos, rows, flip, W


def _effect_dosage(dos, flip):
    """[V, N] effect-allele dosage, missing -> 0 (float64)."""
    g = np.where(dos < 0, 0, dos).astype(np.float64)
    g = np.where((flip[:, None] == 1) & (dos >= 0), 2.0 - g, g)
    return g


def _limb_bound(G, X, precision, block=BLOCK):
    """max error of sum over rows r of G[r, :] * X[r, k] when X is limb-split
    per `block` rows and column: sum_blk (sum_{r in blk} G[r, i]) * m_blk,k / qmax."""
    d = _qmax(precision)
    R = X.shape[0]
    out = 0.0
    for b0 in range(0, R, block):
        m = np.abs(X[b0:b0 + block]).max(0)          # [K]
        gsum = G[b0:b0 + block].sum(0)               # [I]
        out = out + np.outer(gsum, m) / d
    return out


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


@pytest.mark.parametrize("precision", ["fp32", "bf16x4", "bf16x3", "bf16x2", "i8x4", "i8x3"])
def test_dense_forward_matches_fp64(orc, precision):
    N, V, K = 3001, 9000, 16          # two variant scale blocks, ragged N and V
    _flush(f"dense forward {precision} N={N} V={V} K={K}")
    dos, rows, flip, W = _data(N, V, K, 11)
    S64, M64 = orc.dense_forward_fp64(rows, N, flip, W)
    S, miss = ag.dense_forward(_dev(W), _dev(rows), _dev(flip), N, precision)
    torch.cuda.synchronize()
    S, miss = S.cpu().numpy(), miss.cpu().numpy()
    assert np.array_equal(miss.astype(np.int64), M64.astype(np.int64))
    err = np.abs(S - S64)
    if precision == "fp32":
        G = _effect_dosage(dos, flip)                 # [V, N]
        assert np.all(err <= fp32_sum_bound(G.T @ np.abs(W.astype(np.float64)))), err.max()
        assert err.max() / np.abs(S64).max() <= 1e-6
    else:
        G = _effect_dosage(dos, flip)                 # [V, N]
        bound = _limb_bound(G, W.astype(np.float64), precision,
                            I8_BLOCK if precision.startswith("i8") else BLOCK)
        assert np.all(err <= bound + 1e-12 * np.abs(S64).max()), (err - bound).max()
        if precision in ("bf16x4", "bf16x3", "i8x4"):
            assert err.max() / np.abs(S64).max() <= 1e-6


@pytest.mark.parametrize("precision,K,N,V,pad", [
    ("bf16x3", 16, 9000, 4100, False), ("bf16x3", 40, 9000, 4100, False), ("bf16x2", 16, 9000, 4100, False),
    ("bf16x4", 16, 9000, 4100, True), ("bf16x3", 40, 9000, 4100, True), ("bf16x4", 64, 9000, 300, True),
    ("bf16x4", 16, 70_000, 300, True), ("bf16x2", 8, 9000, 300, True),
    ("i8x4", 16, 9000, 4100, True), ("i8x4", 40, 9000, 4100, True), ("i8x3", 16, 9000, 4100, True),
    ("i8x4", 64, 9000, 300, True), ("i8x4", 16, 70_000, 300, True)])
def test_adjoint_matches_fp64(orc, precision, K, N, V, pad):
    """Rows padded to a multiple of 4 bytes run the split-limb adjoint kernel
    (int8 and BF16); 2250-byte rows run the byte-load BF16 kernel.  N = 9000:
    a last super-step that runs past the row end, a partial last workgroup,
    two person blocks of the byte-load kernel; N = 70,000: two person blocks
    of the split-limb kernel (an FP64 flush across blocks); K = 64: 16
    tiles."""
    _flush(f"adjoint {precision} N={N} V={V} K={K} pad={pad}")
    dos, rows, flip, _ = _data(N, V, 1, 12)
    rng = np.random.default_rng(13)
    R = rng.normal(0, 1, (N, K)).astype(np.float32)
    R[:, 3] *= 1e-6                    # a small column: scales are per column
    dW64 = orc.adjoint_fp64(rows, N, flip, R.astype(np.float64))
    if pad:
        # the split-limb adjoint reads 4-byte aligned rows: pad each row with zero bytes
        rows = np.pad(rows, ((0, 0), (0, -rows.shape[1] % 4)))
        assert ag._adjoint_precision(precision, _dev(rows)) == precision
    dW = ag.dense_adjoint(_dev(R), _dev(rows), _dev(flip), V, precision)
    torch.cuda.synchronize()
    dW = dW.cpu().numpy().astype(np.float64)
    G = _effect_dosage(dos, flip).T                   # [N, V]
    block = I8_BLOCK if precision.startswith("i8") else (BF_ADJ_BLOCK if pad else BLOCK)
    bound = _limb_bound(G, R.astype(np.float64), precision, block)
    err = np.abs(dW - dW64)
    assert np.all(err <= bound + 2.0 ** -23 * np.abs(dW64) + 1e-30), (err - bound).max()


def test_int8_adjoint_falls_back_on_unaligned_rows():
    bed = torch.zeros(10, 2250, dtype=torch.uint8, device="cuda")   # 2250-byte rows
    assert ag._adjoint_precision("i8x4", bed) == "bf16x3"


def test_adjoint_rejects_too_many_scores():
    # 100 scores in bf16x3 are 21 limb tiles: above the 16 of the split-limb
    # kernel and the 12 of the byte-load kernel
    N, V, K = 256, 64, 100
    rows = np.zeros((V, N // 4), np.uint8)
    R = torch.zeros(N, K, dtype=torch.float32, device="cuda")
    with pytest.raises(_lib.GnmError):
        ag.dense_adjoint(R, _dev(rows), None, V, "bf16x3")


def _ridge_problem(N, V, K, seed):
    dos, rows, flip, W = _data(N, V, K, seed, zero_frac=0.0)
    rng = np.random.default_rng(seed + 1)
    G = _effect_dosage(dos, flip)                     # [V, N]
    w_star = rng.normal(0, 0.02, V)
    y = G.T @ w_star + rng.normal(0, 1, N)
    return dos, rows, flip, W, G, y


@pytest.mark.parametrize("precision", ["bf16x3", "i8x4"])
def test_autograd_matches_finite_differences_and_fp64_gradient(orc, precision):
    """Ridge-penalised MSE L(W) = sum_k mean_p (S[p,k] - y[p])^2 + lam ||W||^2
    with S from the split forward of `precision`; gradient through the
    adjoint of the same arithmetic (500-byte rows: the int8 adjoint runs)."""
    N, V, K, lam = 2000, 3000, 4, 1e-3
    _flush(f"autograd FD {precision} N={N} V={V} K={K}")
    dos, rows, flip, W0, G, y = _ridge_problem(N, V, K, 21)
    d_rows, d_flip = _dev(rows), _dev(flip)
    y_t = torch.from_numpy(y).cuda()

    W = torch.from_numpy(W0).cuda().requires_grad_(True)
    S = ag.polygenic_score(W, d_rows, d_flip, N, precision)
    loss = ((S - y_t[:, None]) ** 2).mean(0).sum() + lam * (W.double() ** 2).sum()
    loss.backward()
    g_auto = W.grad.detach().cpu().numpy().astype(np.float64)

    # FP64 CPU gradient: dL/dW = G (2/N)(S64 - y) + 2 lam W
    S64, _ = orc.dense_forward_fp64(rows, N, flip, W0)
    R64 = 2.0 / N * (S64 - y[:, None])
    g64 = orc.adjoint_fp64(rows, N, flip, R64) + 2 * lam * W0.astype(np.float64)
    scale = np.abs(g64).max()
    assert np.abs(g_auto - g64).max() <= 1e-5 * scale

    # central finite differences of the FP64 loss on 100 random weights, the
    # loss evaluated with the FP64 oracle's forward
    def loss_at(Wnp):
        Sx, _ = orc.dense_forward_fp64(rows, N, flip, Wnp)
        return float(((Sx - y[:, None]) ** 2).mean(0).sum() + lam * (Wnp.astype(np.float64) ** 2).sum())

    rng = np.random.default_rng(5)
    picks = rng.choice(V * K, 100, replace=False)
    h = 2.0 ** -12                       # exact in f32 next to weights ~1e-3
    fd = np.empty(100)
    for i, idx in enumerate(picks):
        v, k = divmod(int(idx), K)
        Wp, Wm = W0.copy(), W0.copy()
        Wp[v, k] += h
        Wm[v, k] -= h
        hp = float(Wp[v, k]) - float(W0[v, k])
        hm = float(W0[v, k]) - float(Wm[v, k])
        fd[i] = (loss_at(Wp) - loss_at(Wm)) / (hp + hm)
    ga = g_auto.reshape(-1)[picks]
    g6 = g64.reshape(-1)[picks]
    assert np.abs(fd - g6).max() <= 1e-7 * scale       # the FD itself is sound
    assert np.abs(ga - fd).max() <= 1e-5 * scale


def test_gradient_steps_reduce_the_loss():
    """20 plain gradient steps on the ridge problem with step 1/lambda_max of
    the Hessian (2/N) G G^T + 2 lam I: the loss decreases at every step."""
    N, V, K, lam = 2000, 3000, 2, 1e-3
    _flush("autograd 20 steps")
    _, rows, flip, W0, G, y = _ridge_problem(N, V, K, 31)
    lam_max = 2.0 / N * np.linalg.norm(G, 2) ** 2 + 2 * lam
    d_rows, d_flip = _dev(rows), _dev(flip)
    y_t = torch.from_numpy(y).cuda()
    W = torch.zeros(V, K, dtype=torch.float32, device="cuda", requires_grad=True)
    opt = torch.optim.SGD([W], lr=1.0 / lam_max)
    losses = []
    for _ in range(20):
        opt.zero_grad()
        S = ag.polygenic_score(W, d_rows, d_flip, N, "bf16x3")
        loss = ((S - y_t[:, None]) ** 2).mean(0).sum() + lam * (W.double() ** 2).sum()
        loss.backward()
        opt.step()
        losses.append(float(loss.detach()))
    assert all(np.isfinite(losses))
    assert all(b <= a * (1 + 1e-9) for a, b in zip(losses, losses[1:]))
    assert losses[-1] < losses[0]


@pytest.mark.parametrize("precision", ["fp32", "bf16x3", "i8x4"])
@pytest.mark.parametrize("N,V,K", [(1, 65, 1), (5, 1, 3), (33, 63, 2), (130, 200, 17)])
def test_dense_forward_and_adjoint_tiny_shapes(orc, precision, N, V, K):
    """One person, one variant, fewer than a k-step of rows, ragged bytes,
    17 scores (a partial tile): forward within the limb bound and exact
    missing counts; the adjoint within its bound (the int8 adjoint on rows
    padded to 4 bytes)."""
    _flush(f"tiny dense {precision} N={N} V={V} K={K}")
    dos, rows, flip, W = _data(N, V, K, 100 + N + V + K)
    S64, M64 = orc.dense_forward_fp64(rows, N, flip, W)
    S, miss = ag.dense_forward(_dev(W), _dev(rows), _dev(flip), N, precision)
    torch.cuda.synchronize()
    S, miss = S.cpu().numpy(), miss.cpu().numpy()
    assert np.array_equal(miss.astype(np.int64), M64.astype(np.int64))
    scale = max(np.abs(S64).max(), 1e-30)
    if precision == "fp32":
        G = _effect_dosage(dos, flip)
        assert np.all(np.abs(S - S64) <= fp32_sum_bound(G.T @ np.abs(W.astype(np.float64))) + 1e-300)
    else:
        G = _effect_dosage(dos, flip)
        bound = _limb_bound(G, W.astype(np.float64), precision,
                            I8_BLOCK if precision.startswith("i8") else BLOCK)
        assert np.all(np.abs(S - S64) <= bound + 1e-12 * scale)
    if precision == "fp32":
        return   # the adjoint has split paths only
    rng = np.random.default_rng(7)
    R = rng.normal(0, 1, (N, K)).astype(np.float32)
    dW64 = orc.adjoint_fp64(rows, N, flip, R.astype(np.float64))
    arows = np.pad(rows, ((0, 0), (0, -rows.shape[1] % 4))) if precision.startswith("i8") else rows
    dW = ag.dense_adjoint(_dev(R), _dev(arows), _dev(flip), V, precision)
    torch.cuda.synchronize()
    dW = dW.cpu().numpy().astype(np.float64)
    Gt = _effect_dosage(dos, flip).T
    bound = _limb_bound(Gt, R.astype(np.float64), precision, I8_BLOCK if precision.startswith("i8") else BLOCK)
    assert np.all(np.abs(dW - dW64) <= bound + 2.0 ** -23 * np.abs(dW64) + 1e-30)


def test_operands_checked_before_launch():
    """Wrong device, shape or dtype of W, R or flip raise ValueError before
    any kernel runs."""
    _flush("operand checks")
    N, V, K = 64, 32, 4
    _, rows, flip, W = _data(N, V, K, 3)
    bed, fl = _dev(rows), _dev(flip)
    with pytest.raises(ValueError):
        ag.dense_forward(torch.from_numpy(W), bed, fl, N)                 # W on the host
    with pytest.raises(ValueError):
        ag.dense_forward(_dev(W), bed, fl.to(torch.int64), N)             # wide flip
    with pytest.raises(ValueError):
        ag.dense_forward(_dev(W), bed, fl[:-1].contiguous(), N)           # short flip
    with pytest.raises(ValueError):
        ag.dense_forward(_dev(W), bed, fl, N, "bf16x9")
    R = torch.zeros(N, K, device="cuda")
    with pytest.raises(ValueError):
        ag.dense_adjoint(R.cpu(), bed, fl, V)
    with pytest.raises(ValueError):
        ag.dense_adjoint(torch.zeros(N + 8, K, device="cuda"), bed, fl, V)   # more people than the rows hold
    with pytest.raises(ValueError):
        ag.dense_adjoint(R, bed, fl.cpu(), V)
    torch.cuda.synchronize()
    S, _ = ag.dense_forward(_dev(W), bed, fl, N, "i8x4")                # still usable after
    assert torch.isfinite(S).all()


@pytest.mark.parametrize("precision", ["bf16x4", "i8x4", "fp32"])
def test_non_finite_weight_gives_nan_column(precision):
    """A NaN or infinite weight makes its whole score column non-finite: the
    split paths give NaN for every person (the block scale is NaN), the FP32
    path IEEE's NaN or inf per person (inf * 0 = NaN); the other columns are
    unchanged."""
    N, V, K = 300, 700, 3
    _flush(f"non-finite weight {precision}")
    _, rows, flip, W = _data(N, V, K, 5)
    S0, _ = ag.dense_forward(_dev(W), _dev(rows), _dev(flip), N, precision)
    for bad in (np.nan, np.inf):
        Wb = W.copy()
        Wb[123, 1] = bad
        S, _ = ag.dense_forward(_dev(Wb), _dev(rows), _dev(flip), N, precision)
        torch.cuda.synchronize()
        S, s0 = S.cpu().numpy(), S0.cpu().numpy()
        assert not np.isfinite(S[:, 1]).any()
        assert np.array_equal(S[:, [0, 2]], s0[:, [0, 2]])


@pytest.mark.parametrize("precision", ["i8x4", "bf16x3"])
def test_nan_gradient_reaches_dw(precision):
    """A NaN in dL/dS makes the matching column of dL/dW NaN."""
    N, V, K = 300, 700, 3
    _flush(f"NaN gradient {precision}")
    _, rows, flip, _ = _data(N, V, K, 6)
    R = torch.randn(N, K, device="cuda")
    R[17, 2] = float("nan")
    dW = ag.dense_adjoint(R, _dev(rows), _dev(flip), V, precision)
    torch.cuda.synchronize()
    dW = dW.cpu().numpy()
    assert np.isnan(dW[:, 2]).all()
    assert np.isfinite(dW[:, :2]).all()
