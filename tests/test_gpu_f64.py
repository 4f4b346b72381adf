"""GPU parity of the double-precision engine (computeKmeansD and the ...D
kernel names) and of rowTransformMinimumF/D against the CPU oracle, through
the C-ABI.

Bar: bit-exact centres, labels, norms, counts and compactness (the GPU runs
the same double fma chains and the same canonical FP64 summation order)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
import kmeans_amd as km  # noqa: E402


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def mixture64(orc, seed, n, d, k, lo=-6.0, hi=6.0, s=1.0):
    """Double points with bits below float precision (not float-valued)."""
    X, _, _, _ = orc.mixture(seed, n, d, k, lo, hi, s, s)
    rng = np.random.default_rng(seed)
    return X.astype(np.float64) + rng.uniform(-1e-4, 1e-4, X.shape)


@pytest.mark.parametrize("n,d,k,iters", [(20_000, 17, 37, 5), (50_000, 128, 300, 4),
                                         (9_000, 3, 2, 6), (30_000, 64, 1024, 3),
                                         (4_000, 200, 65, 3)])
def test_compute_kmeans_d_bit_exact(orc, n, d, k, iters):
    X = mixture64(orc, 31 + d, n, d, k)
    C0 = X[7:7 + k].copy()
    st, res = km.compute_kmeans_d(X, k, -1.0, iters, result=C0.T.copy(), init_from_result=True)
    ref = orc.lloyd64(X, C0, iters)
    assert st == km.KMEANS_NONCONVERGENCE  # criterion < 0 is never met
    assert res.dtype == np.float64
    assert np.array_equal(res, ref["centers"].T)


def test_compute_kmeans_d_criterion_and_sampled_init(orc):
    X = mixture64(orc, 40, 30_000, 8, 10, -10, 10, 0.7)
    idx = orc.sample_init(42, X.shape[0], 10)
    # 1e-5 stops after 5 iterations; 1e-7 is still unmet after 60
    for crit, status in ((1e-5, km.KMEANS_SUCCESS), (1e-7, km.KMEANS_NONCONVERGENCE)):
        ref = orc.lloyd64(X, X[idx], 60, criterion=crit)
        st, res = km.compute_kmeans_d(X, 10, crit, 60, init_algorithm=3)
        assert st == status == (0 if ref["status"] == 0 else km.KMEANS_NONCONVERGENCE)
        assert np.array_equal(res, ref["centers"].T)


def test_compute_kmeans_d_mini_batch(orc):
    X = mixture64(orc, 41, 10_000, 5, 8)
    C = X[:8].copy()
    st, res = km.compute_kmeans_d(X, 8, -1.0, 5, result=C.T.copy(), init_from_result=True,
                                  mini_batch_fraction=0.3)
    # windows of 3000 rows: [0,3000) [3000,6000) [6000,9000) [9000,10000) [0,3000)
    for b0, b1 in [(0, 3000), (3000, 6000), (6000, 9000), (9000, 10000), (0, 3000)]:
        lab = orc.assign64(X[b0:b1], C)
        C = orc.update64(X[b0:b1], lab, C)[0]
    assert np.array_equal(res, C.T)


def test_compute_kmeans_d_empty_cluster_keeps_centre(orc):
    X = mixture64(orc, 42, 5_000, 4, 3)
    C0 = np.concatenate([X[:3], np.full((1, 4), 1e6)])  # centre 3 attracts nothing
    st, res = km.compute_kmeans_d(X, 4, -1.0, 3, result=C0.T.copy(), init_from_result=True)
    ref = orc.lloyd64(X, C0, 3)
    assert np.array_equal(res, ref["centers"].T)
    assert np.all(res[:, 3] == 1e6)


def test_compute_kmeans_d_rejects_bad_shapes(orc):
    X = np.zeros((10, 3))
    st, _ = km.compute_kmeans_d(X, 11, -1.0, 3)
    assert st == km.KMEANS_ERROR_INVALID_ARGUMENT


def test_legacy_double_kernel_names(orc):
    X = mixture64(orc, 43, 6000, 20, 70, -5, 5)
    C = X[:70].copy()
    n, d = X.shape
    k = 70
    L = km.lib()
    Xd, B = dev(X), dev(C.T.copy())  # D x K as the reference passes it
    ncg = torch.empty(k, dtype=torch.float64, device="cuda")
    assert L.colNormalizeD(d, k, km._vp(B.data_ptr()), km._vp(ncg.data_ptr())) == 0
    rng = torch.empty(n, dtype=torch.float64, device="cuda")
    assert L.rowNormalizeD(0, n, d, km._vp(Xd.data_ptr()), km._vp(rng.data_ptr())) == 0
    lab = torch.empty(n, dtype=torch.int32, device="cuda")
    for m0 in (0, 2500):  # two tiles
        m = 2500 if m0 == 0 else n - 2500
        assert L.ClosestCentersD(m0, m, d, km._vp(Xd.data_ptr() + 8 * m0 * d), k,
                                 km._vp(B.data_ptr()), None, None, km._vp(ncg.data_ptr()), 0,
                                 None, None, km._vp(lab.data_ptr())) == 0
    torch.cuda.synchronize()
    assert np.array_equal(ncg.cpu().numpy(), orc.col_norms64(C))
    assert np.array_equal(rng.cpu().numpy(), orc.row_norms64(X))
    ref = orc.assign64(X, C)
    assert np.array_equal(lab.cpu().numpy(), ref)
    # two tiles of the update, then lastTile
    large = torch.zeros(4 * k * d, dtype=torch.float64, device="cuda")
    clarge = torch.zeros(4 * k, dtype=torch.int32, device="cuda")
    cen = torch.zeros((k, d), dtype=torch.float64, device="cuda")
    cnt = torch.empty(k, dtype=torch.int32, device="cuda")
    assert L.ClusterCentersD(0, n, d, k, True, km._vp(Xd.data_ptr()), km._vp(lab.data_ptr()),
                             km._vp(large.data_ptr()), km._vp(clarge.data_ptr()),
                             km._vp(cen.data_ptr()), km._vp(cnt.data_ptr())) == 0
    rC, rcnt, empty, _ = orc.update64(X, ref, C)
    assert np.array_equal(cnt.cpu().numpy(), rcnt)
    assert np.array_equal(cen.cpu().numpy()[rcnt > 0], rC[rcnt > 0])
    comp = torch.zeros(4, dtype=torch.float64, device="cuda")
    assert L.CompactnessD(0, n, True, d, k, km._vp(Xd.data_ptr()), km._vp(lab.data_ptr()),
                          km._vp(cen.data_ptr()), km._vp(comp.data_ptr())) == 0
    assert comp[0].item() == orc.compactness64(X, ref, cen.cpu().numpy())


def test_closest_centers_unaligned_centre_pointer(orc):
    # B one element into its allocation (not 16/32-byte aligned): padded copy path
    for dt, esz in ((np.float32, 4), (np.float64, 8)):
        X, _, _, _ = orc.mixture(44, 3000, 16, 256, -5, 5, 1.0, 1.0)
        X = X.astype(dt)
        C = X[:256].copy()
        L = km.lib()
        buf = dev(np.concatenate([np.zeros(1, dt), C.T.copy().ravel()]))
        Xd = dev(X)
        f = L.ClosestCentersF if dt == np.float32 else L.ClosestCentersD
        nc = dev(orc.col_norms(C) if dt == np.float32 else orc.col_norms64(C))
        lab = torch.empty(3000, dtype=torch.int32, device="cuda")
        assert f(0, 3000, 16, km._vp(Xd.data_ptr()), 256, km._vp(buf.data_ptr() + esz), None, None,
                 km._vp(nc.data_ptr()), 0, None, None, km._vp(lab.data_ptr())) == 0
        torch.cuda.synchronize()
        ref = orc.assign(X, C) if dt == np.float32 else orc.assign64(X, C)
        assert np.array_equal(lab.cpu().numpy(), ref)


@pytest.mark.parametrize("dt", [np.float32, np.float64])
@pytest.mark.parametrize("m,n", [(3000, 64), (777, 1000), (5, 3), (2000, 4097)])
def test_row_transform_minimum(orc, dt, m, n):
    rng = np.random.default_rng(m + n)
    X = rng.standard_normal((m, 24)).astype(dt)
    C = rng.standard_normal((n, 24)).astype(dt)
    C[n // 2] = C[0]  # an exact tie: the lower index wins
    P = (X @ C.T).astype(dt)
    P[1, :] = -np.inf  # every value +inf: -1
    P[2, n - 1] = np.nan
    rn = (X * X).sum(1).astype(dt)
    cn = (C * C).sum(1).astype(dt)
    m0 = 11
    ref = orc.row_transform_min(P, rn, cn, m0=m0)
    out = dev(np.full(m0 + m, -7, np.int32))
    Pd, rnd, cnd = dev(P), dev(rn), dev(cn)
    f = km.lib().rowTransformMinimumF if dt == np.float32 else km.lib().rowTransformMinimumD
    assert f(m0, m, n, km._vp(rnd.data_ptr()), km._vp(cnd.data_ptr()), km._vp(Pd.data_ptr()),
             km._vp(out.data_ptr())) == 0
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    assert np.array_equal(got, ref)
    assert got[m0 + 1] == -1 and np.all(got[:m0] == -7)


@pytest.mark.parametrize("n,d,k", [(500, 1, 3), (400, 6, 1), (64, 5, 64), (1, 3, 1)])
def test_compute_kmeans_d_edge_shapes(orc, n, d, k):
    # one dimension, one centre, every point its own centre, a single point
    rng = np.random.default_rng(n + d + k)
    X = rng.standard_normal((n, d)) * 3
    C0 = X[:k].copy()
    st, res = km.compute_kmeans_d(X, k, -1.0, 3, result=C0.T.copy(), init_from_result=True)
    ref = orc.lloyd64(X, C0, 3)
    assert np.array_equal(res, ref["centers"].T)
