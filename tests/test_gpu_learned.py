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



# ---------------------------------------------------------------- engine level
def _c5_like(n=120_000, d=32, k=1024, seed=3):
    """Gaussian mixture shaped like the 20M x 96 x 16384 workload, scaled down:
    k means uniform in [-6, 6]^d, sigma uniform in [0.3, 1.5]; seeded sampled
    initial centres."""
    X, _, _, _ = orc.mixture(seed, n, d, k, -6.0, 6.0, 0.3, 1.5)
    C0 = X[orc.sample_init(seed, n, k)].copy()
    return X, C0


def _two_level(n=100_000, d=32, k=1024, seed=5):
    """Centres in 16 well separated clusters of 64: the group bound's case."""
    rng = np.random.default_rng(seed)
    sup = rng.uniform(-60, 60, (16, d))
    means = (np.repeat(sup, k // 16, axis=0) + rng.normal(0, 1.2, (k, d))).astype(np.float32)
    z = rng.integers(0, k, n)
    X = (means[z] + rng.normal(0, 0.6, (n, d))).astype(np.float32)
    C0 = X[orc.sample_init(seed, n, k)].copy()
    return X, C0


_ORACLE_CACHE = {}


def _oracle_run(key, X, C0, iters):
    if key not in _ORACLE_CACHE:
        _ORACLE_CACHE[key] = orc.lloyd(X, C0, iters)
    return _ORACLE_CACHE[key]


def _run(X, C0, iters, **kw):
    e = km.KMeans(X.shape[0], X.shape[1], C0.shape[0], **kw)
    e.set_data(X)
    e.set_centers(C0)
    e.run(iters)
    out = e.labels(), e.centers(), e.stats()
    e.close()
    return out


def _check_counts(st, m, iters, train_iters):
    learned_iters = iters - train_iters
    assert st["proposed"] == m * learned_iters
    assert st["certified"] + st["certified_group"] + st["fallback"] == st["proposed"]


@pytest.mark.parametrize("proposer,precision", [("mlp", "fp32"), ("centroid", "fp32"),
                                                ("mlp", "bf16")])
def test_learned_lloyd_equals_oracle(proposer, precision):
    X, C0 = _c5_like()
    iters = 7
    ref = _oracle_run("c5like", X, C0, iters)
    lab, C, st = _run(X, C0, iters, precision=precision, proposer=proposer, groups=32,
                      top_groups=4, train_iters=3, hidden=128)
    print(f"\n{proposer}/{precision}: {st}")
    # every iteration's labels equal the plain FP32 labels, so the whole
    # trajectory (labels and canonical centres) is bit-identical
    assert np.array_equal(lab, ref["labels"]), int((lab != ref["labels"]).sum())
    assert np.array_equal(C, ref["centers"])
    _check_counts(st, X.shape[0], iters, 3)
    assert st["certified"] > 0
    # far fewer exact scores than a full search of the proposer iterations
    assert st["candidate_pairs"] < 0.5 * st["proposed"] * C0.shape[0]
    if proposer == "mlp":
        assert st["train_loss_last"] < st["train_loss_first"]


def test_learned_group_bound_certifies_separated_groups():
    X, C0 = _two_level()
    iters = 6
    ref = _oracle_run("twolevel", X, C0, iters)
    lab, C, st = _run(X, C0, iters, proposer="centroid", groups=16, top_groups=2, train_iters=2)
    print(f"\ntwo-level: {st}")
    assert np.array_equal(lab, ref["labels"])
    assert np.array_equal(C, ref["centers"])
    _check_counts(st, X.shape[0], iters, 2)
    assert st["certified_group"] > 0


def test_learned_mlp_gain_over_centroid_proposer():
    """What the trained network gains over the method without it.  Against
    the nearest-centroid proposer (same groups, same number of proposed
    groups): the MLP names the group of the final label for more points, more
    points are certified and fewer are searched in full.  Against the full
    search: a fraction of the exact scores, with the same labels (the parity
    tests above).  Data: 200,000 points, D = 64, K = 2048 means uniform in
    [-6, 6]^64, sigma uniform in [0.3, 1.5], 32 groups, top 4, BF16 engine."""
    X, _, _, _ = orc.mixture(3, 200_000, 64, 2048, -6.0, 6.0, 0.3, 1.5)
    C0 = X[orc.sample_init(3, X.shape[0], 2048)].copy()
    res = {}
    for prop in ("mlp", "centroid"):
        _, _, st = _run(X, C0, 6, precision="bf16", proposer=prop, groups=32, top_groups=4,
                        train_iters=3, hidden=128)
        res[prop] = st
    m, c = res["mlp"], res["centroid"]
    n = m["proposed"]
    print("\ngain:", {p: {f: s[f] / s["proposed"] for f in ("proposal_hits", "certified", "fallback")}
                      for p, s in res.items()})
    assert c["proposed"] == n
    assert m["proposal_hits"] >= c["proposal_hits"] + 0.02 * n
    assert m["certified"] > c["certified"]
    assert m["fallback"] < c["fallback"]
    assert m["candidate_pairs"] < 0.25 * n * 2048
