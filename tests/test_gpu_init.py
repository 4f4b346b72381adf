"""GPU k-means++ seeding against the CPU oracle's canonical picks (bit-exact
indices: FP64 distances and the same 256-way sum tree), through the C-ABI."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
import kmeans_amd as km  # noqa: E402


@pytest.mark.parametrize("n,d,k,seed", [(200, 5, 3, 1), (300, 8, 4, 2), (1000, 3, 5, 3),
                                        (70_000, 17, 50, 4), (250_000, 64, 200, 5)])
def test_plus_plus_picks_equal_oracle(orc, n, d, k, seed):
    print(f"kmeans++ n={n} d={d} k={k}", flush=True)
    X, _, _, _ = orc.mixture(70 + d, n, d, max(k // 2, 2), -5, 5, 1.0, 1.0)
    idx = orc.kmeanspp(X, k, seed)
    e = km.KMeans(n, d, k)
    e.set_data(X)
    e.init_plus_plus(seed)
    C = e.centers()
    e.close()
    assert np.array_equal(C, X[idx])


def test_plus_plus_three_level_tree(orc):
    # 17M points: the sum tree has three levels (66,407 / 260 / 2 sums)
    n, d, k = 17_000_000, 2, 3
    rng = np.random.default_rng(0)
    X = rng.standard_normal((n, d)).astype(np.float32)
    idx = orc.kmeanspp(X, k, 11)
    e = km.KMeans(n, d, k)
    e.set_data(X)
    e.init_plus_plus(11)
    C = e.centers()
    e.close()
    assert np.array_equal(C, X[idx])


def test_plus_plus_identical_points(orc):
    X = np.ones((3000, 4), np.float32)
    X[5] = 2.0  # one distinct point, the rest coincide
    idx = orc.kmeanspp(X, 4, 9)
    e = km.KMeans(3000, 4, 4)
    e.set_data(X)
    e.init_plus_plus(9)
    assert np.array_equal(e.centers(), X[idx])
    e.close()


def test_compute_kmeans_with_plus_plus_init(orc):
    X, _, _, _ = orc.mixture(80, 20_000, 6, 12, -8, 8, 1.0, 1.0)
    idx = orc.kmeanspp(X, 12, 42)
    ref = orc.lloyd(X, X[idx], 5)
    res = np.zeros((6, 12), np.float32)
    st = km.lib().computeKmeansF(km._np_ptr(X), 20_000, 6, 12, -1.0, 5, 0, 2, 0, 0,
                                 km._np_ptr(res), 1.0)
    assert st == km.KMEANS_NONCONVERGENCE
    assert np.array_equal(res, ref["centers"].T)
    X64 = X.astype(np.float64) + 1e-6
    idx64 = orc.kmeanspp(X64, 12, 42)
    ref64 = orc.lloyd64(X64, X64[idx64], 5)
    res64 = np.zeros((6, 12), np.float64)
    st = km.lib().computeKmeansD(km._np_ptr(X64), 20_000, 6, 12, -1.0, 5, 0, 2, 0, 0,
                                 km._np_ptr(res64), 1.0)
    assert st == km.KMEANS_NONCONVERGENCE
    assert np.array_equal(res64, ref64["centers"].T)


def test_plus_plus_needs_resident_data():
    e = km.KMeans(2 * km.KM_BLOCK, 2, 3, tile_rows=km.KM_BLOCK)
    e.set_data(np.zeros((2 * km.KM_BLOCK, 2), np.float32))
    with pytest.raises(km.KmeansError):
        e.init_plus_plus(1)
    e.close()


@pytest.mark.parametrize("n,d,k", [(64, 3, 64), (257, 1, 5), (5000, 2, 1)])
def test_plus_plus_edge_shapes(orc, n, d, k):
    # k == m (every point is picked once), one dimension with 257 points (a
    # second chunk of one point), a single centre
    rng = np.random.default_rng(n)
    X = rng.standard_normal((n, d)).astype(np.float32)
    idx = orc.kmeanspp(X, k, 3)
    if k == n:
        assert sorted(idx.tolist()) == list(range(n))
    e = km.KMeans(n, d, k)
    e.set_data(X)
    e.init_plus_plus(3)
    assert np.array_equal(e.centers(), X[idx])
    e.close()
