"""Learned candidate proposer: the MLP kernels against the FP64 CPU MLP.

The oracle's orc_mlp_fwd_bwd is pinned on the CPU by central finite
differences and by closed forms (test_oracle.py).  Here the GPU forward and
backward (BF16 MFMA products, FP32 accumulation) are compared with it in two
ways: against the oracle run with the same BF16 rounding points ("mirror"),
which leaves only FP32-vs-FP64 accumulation differences and rare BF16 rounding
flips, and against the unrounded FP64 network, which bounds the effect of BF16
training arithmetic.
"""
import numpy as np
import pytest

from oracle import oracle as orc

torch = pytest.importorskip("torch")
import kmeans_amd as km  # noqa: E402

pytestmark = pytest.mark.gpu


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _bf16_bits(a):
    """float32 values that are BF16-representable -> uint16 bit patterns (as int16 storage)."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)
    assert np.all((u & 0xFFFF) == 0)
    return (u >> 16).astype(np.uint16).view(np.int16)


def _params(din, h, g, rng):
    parts = []
    for fan_out, fan_in in ((h, din), (h, 1), (h, h), (h, 1), (g, h), (g, 1)):
        sc = 1.0 / np.sqrt(max(fan_in, 1)) if fan_in > 1 else 0.1
        parts.append((rng.standard_normal(fan_out * fan_in) * sc).astype(np.float32))
    return np.concatenate(parts)


def _blocks(din, h, g):
    sizes = [h * din, h, h * h, h, g * h, g]
    off = np.concatenate([[0], np.cumsum(sizes)])
    names = ["W1", "b1", "W2", "b2", "W3", "b3"]
    return [(names[i], off[i], off[i + 1]) for i in range(6)]


def _gpu_fwd_bwd(din, h, g, X, t, P):
    B = X.shape[0]
    assert km.mlp_param_count(din, h, g) == P.shape[0]
    Xd = _dev(_bf16_bits(X))
    td = _dev(t.astype(np.int32))
    Pd = _dev(P)
    Gd = torch.full_like(Pd, float("nan"))
    loss = torch.empty(1, dtype=torch.float32, device="cuda")
    wsb = km.mlp_workspace_size(din, h, g, B)
    ws = torch.empty(wsb, dtype=torch.uint8, device="cuda")
    km.mlp_forward_backward(din, h, g, B, Xd.data_ptr(), td.data_ptr(), Pd.data_ptr(),
                            Gd.data_ptr(), loss.data_ptr(), ws.data_ptr(), wsb)
    torch.cuda.synchronize()
    return float(loss.item()), Gd.cpu().numpy()


@pytest.mark.parametrize("din,h,g,B,seed", [(16, 16, 16, 5, 1), (96, 128, 128, 1000, 2),
                                            (96, 256, 128, 777, 3), (128, 64, 32, 4096, 4)])
def test_mlp_gradients_match_fp64_oracle(din, h, g, B, seed):
    rng = np.random.default_rng(seed)
    X = orc.bf16_round(rng.standard_normal((B, din)).astype(np.float32) * 2.0)
    t = rng.integers(0, g, B).astype(np.int32)
    P = _params(din, h, g, rng)
    loss_gpu, G_gpu = _gpu_fwd_bwd(din, h, g, X, t, P)
    assert np.all(np.isfinite(G_gpu))
    loss_m, G_m = orc.mlp_fwd_bwd(din, h, g, X, t, P.astype(np.float64), mirror=True)
    loss_x, G_x = orc.mlp_fwd_bwd(din, h, g, X, t, P.astype(np.float64), mirror=False)
    # loss: FP32 accumulation of BF16 products vs FP64 of the same rounded operands
    assert abs(loss_gpu - loss_m) <= 2e-3 * max(1.0, abs(loss_m)), (loss_gpu, loss_m)
    assert abs(loss_gpu - loss_x) <= 3e-2 * max(1.0, abs(loss_x)), (loss_gpu, loss_x)
    for name, a, b in _blocks(din, h, g):
        gm, gx, gg = G_m[a:b], G_x[a:b], G_gpu[a:b].astype(np.float64)
        nm = np.linalg.norm(gm)
        assert nm > 0, name
        # same rounding points: per-block relative Frobenius error within 1e-2
        assert np.linalg.norm(gg - gm) <= 1e-2 * nm, (name, np.linalg.norm(gg - gm) / nm)
        # against the unrounded network: BF16 training arithmetic stays within 10 %
        assert np.linalg.norm(gg - gx) <= 1e-1 * np.linalg.norm(gx), (
            name, np.linalg.norm(gg - gx) / np.linalg.norm(gx))

