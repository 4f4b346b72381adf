"""Parity at the measured workloads' own shapes, on slices of their own
generators (kmeans_amd.datagen, the same counter-based streams as the bench):

  N=10M x D=128 x K=1024 (seed 1, sampled initial centres), BF16 and FP16
      engines on a 200,000-point slice against the oracle's Lloyd loop;
  N=2e8 x D=64 x K=4096 (seed 2, Zipf(1.1) weights), FP16 engine streamed
      from host memory in KM_BLOCK-row tiles (two full tiles and a ragged
      one) against the resident FP16 run and the resident FP32 run;
  N=2e7 x D=96 x K=16384 (seed 3) with the learned proposer at its own
      configuration (128 groups, MLP 96->256->256->128, top 4 groups, 3 plain
      iterations, BF16 engine) on a 200,000-point slice against the oracle.

Bar: labels, centres and compactness bit for bit."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
import kmeans_amd as km  # noqa: E402
from kmeans_amd import datagen  # noqa: E402


def run(X, C0, iters, **kw):
    e = km.KMeans(X.shape[0], X.shape[1], C0.shape[0], **kw)
    e.set_data(X)
    e.set_centers(C0)
    e.run(iters)
    out = e.labels(), e.centers(), e.stats()
    e.close()
    return out


@pytest.mark.parametrize("prec", ["bf16", "fp16"])
def test_c2_slice_equals_oracle(orc, prec):
    print(f"c2 slice n=200000 d=128 k=1024 {prec}", flush=True)
    Xd, C0, spec = datagen.generate_on_gpu("c2", n=200_000, torch=torch)
    X = Xd.cpu().numpy()
    assert spec["seed"] == 1 and spec["init"] == "sample"
    assert np.array_equal(C0, X[spec["init_idx"]])
    ref = orc.lloyd(X, C0, 3)
    lab, C, st = run(X, C0, 3, precision=prec)
    assert np.array_equal(lab, ref["labels"])
    assert np.array_equal(C, ref["centers"])
    assert st["compactness"] == ref["comp_hist"][-1]


def test_c3_streamed_fp16_tiles_equal_resident():
    n = 2 * km.KM_BLOCK + 1_000_003  # two full tiles and a ragged one
    print(f"c3 slice n={n} d=64 k=4096 fp16 streamed", flush=True)
    Xd, C0, spec = datagen.generate_on_gpu("c3", n=n, torch=torch)
    assert spec["zipf"] == 1.1 and spec["d"] == 64 and spec["k"] == 4096
    X = Xd.cpu().numpy()
    del Xd
    torch.cuda.empty_cache()
    iters = 2
    res = run(X, C0, iters, precision="fp16")
    st_ = run(X, C0, iters, precision="fp16", tile_rows=km.KM_BLOCK)
    f32 = run(X, C0, iters, precision="fp32")
    for a, b in ((st_, res), (res, f32)):
        assert np.array_equal(a[0], b[0]), int((a[0] != b[0]).sum())
        assert np.array_equal(a[1], b[1])
        assert a[2]["compactness"] == b[2]["compactness"]
    assert st_[2]["iterations"] == iters


def test_c5_shape_learned_equals_oracle(orc):
    print("c5 slice n=200000 d=96 k=16384 bf16 mlp 128 groups hidden 256 top 4", flush=True)
    Xd, C0, spec = datagen.generate_on_gpu("c5", n=200_000, torch=torch)
    X = Xd.cpu().numpy()
    del Xd
    assert spec["seed"] == 3 and spec["d"] == 96 and spec["k"] == 16384
    iters, train = 5, 3
    ref = orc.lloyd(X, C0, iters)
    lab, C, st = run(X, C0, iters, precision="bf16", proposer="mlp", groups=128, top_groups=4,
                     train_iters=train, hidden=256)
    print({f: st[f] for f in ("proposed", "proposal_hits", "certified", "certified_group",
                              "fallback", "candidate_pairs")})
    assert np.array_equal(lab, ref["labels"]), int((lab != ref["labels"]).sum())
    assert np.array_equal(C, ref["centers"])
    assert st["compactness"] == ref["comp_hist"][-1]
    m = X.shape[0]
    assert st["proposed"] == m * (iters - train)
    assert st["certified"] + st["certified_group"] + st["fallback"] == st["proposed"]
    assert st["certified"] > 0
    assert st["train_loss_last"] < st["train_loss_first"]
