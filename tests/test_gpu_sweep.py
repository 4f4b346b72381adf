"""Seeded random shapes through every engine precision, against the oracle
bit for bit (labels and centres after three iterations): odd and tiny D, K
not a multiple of any tile, N from a few hundred to tens of thousands."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
import kmeans_amd as km  # noqa: E402


def shapes(seed, count):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        n = int(rng.integers(300, 30_000))
        d = int(rng.choice([1, 2, 3, 5, 8, 13, 16, 31, 48, 64, 97, 128, 160]))
        k = int(rng.integers(1, min(n, 400) + 1))
        out.append((n, d, k))
    return out


@pytest.mark.parametrize("case", range(16))
def test_random_shape_all_precisions(orc, case):
    n, d, k = shapes(2024, 16)[case]
    prec = ("fp32", "bf16", "fp16")[case % 3]
    print(f"sweep case {case}: n={n} d={d} k={k} {prec}", flush=True)
    X, _, _, _ = orc.mixture(500 + case, n, d, max(k, 2), -8, 8, 0.5, 2.0)
    C0 = X[np.linspace(0, n - 1, k).astype(int)].copy()
    ref = orc.lloyd(X, C0, 3)
    e = km.KMeans(n, d, k, precision=prec)
    e.set_data(X)
    e.set_centers(C0)
    e.run(3)
    assert np.array_equal(e.labels(), ref["labels"])
    assert np.array_equal(e.centers(), ref["centers"])
    assert e.stats()["compactness"] == ref["comp_hist"][-1]
    e.close()


@pytest.mark.parametrize("case", range(5))
def test_random_shape_double(orc, case):
    n, d, k = shapes(77, 5)[case]
    rng = np.random.default_rng(case)
    X = rng.standard_normal((n, d)) * rng.uniform(0.5, 4.0, d) + rng.uniform(-3, 3, d)
    C0 = X[:k].copy()
    st, res = km.compute_kmeans_d(X, k, -1.0, 3, result=C0.T.copy(), init_from_result=True)
    ref = orc.lloyd64(X, C0, 3)
    assert np.array_equal(res, ref["centers"].T)
