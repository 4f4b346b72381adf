"""GPU checks of the library-GEMM assignment (the reference's doFullSGEMM = 1
path: rocBLAS SGEMM / DGEMM into an HBM product, then the row argmin).

The GEMM's summation order is the library's, so labels are not bit-exact to
the canonical fma chain.  Bar: every label's canonical score lies within the
GEMM's error bound of the best canonical score (2 gamma_D sum_d |x_d c_d| plus
the rounding of the norm sum), equality with the oracle's label wherever the
best two canonical scores are farther apart than that bound, and, on data
with no such near-ties, centres bit-identical to the fused path's (the update
is canonical, so equal labels give equal centres)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
import kmeans_amd as km  # noqa: E402


def check_labels(orc, X, C, lab, tol_scale=1.0):
    """Labels from the GEMM path against the canonical FP32 scores."""
    X64, C64 = X.astype(np.float64), C.astype(np.float64)
    s = X64 @ C64.T - 0.5 * (C64 * C64).sum(1)[None]
    n, d = X.shape
    eps = 2.0 ** -24
    gam = (d + 2) * eps / (1 - (d + 2) * eps)
    # error of |x|^2 + |c|^2 - 2 x.c evaluated in FP32 (GEMM product, fma-chain
    # norms, the FP32 sum and the final rounding), halved into score units
    nrm = (X64 * X64).sum(1)[:, None] + (C64 * C64).sum(1)[None]
    bound = tol_scale * (gam * (np.abs(X64) @ np.abs(C64).T) + (gam + 4 * eps) * nrm)
    best = s.max(1)
    chosen = s[np.arange(n), lab]
    assert np.all(best - chosen <= 2 * bound.max(1) + 1e-12)
    ref = orc.assign(X, C)
    srt = np.sort(s, 1)
    clear = (srt[:, -1] - srt[:, -2]) > 4 * bound.max(1)
    assert clear.mean() > 0.95
    assert np.array_equal(lab[clear], ref[clear])


@pytest.mark.parametrize("n,d,k", [(20_000, 32, 64), (30_000, 128, 300), (4_000, 17, 5)])
def test_gemm_assign_labels_within_bound(orc, n, d, k):
    X, _, _, _ = orc.mixture(50 + d, n, d, k, -5, 5, 1.0, 2.0)
    C = X[:k].copy()
    e = km.KMeans(n, d, k, full_gemm=True)
    e.set_data(X)
    e.set_centers(C)
    e.run(1)
    check_labels(orc, X, C, e.labels())
    e.close()


def test_gemm_assign_many_product_tiles(orc):
    # k = 16384: 32768 rows per product tile, so 100000 points take 4 GEMMs
    X, _, _, _ = orc.mixture(60, 100_000, 16, 64, -5, 5, 1.0, 1.0)
    rng = np.random.default_rng(0)
    C = (rng.standard_normal((16384, 16)) * 4).astype(np.float32)
    e = km.KMeans(100_000, 16, 16384, full_gemm=True)
    e.set_data(X)
    e.set_centers(C)
    e.run(1)
    lab = e.labels()
    e.close()
    idx = rng.choice(100_000, 4000, replace=False)  # the oracle on a sample of rows
    check_labels(orc, X[idx], C, lab[idx])


def test_gemm_resident_row_norms_equal_streamed(orc):
    # resident data: the row norms are computed once per data set; streamed
    # tiles compute theirs per tile; same kernel, so identical results
    # (streamed tiles are whole canonical blocks: two tiles here, one short)
    n, d, k = km.KM_BLOCK + 50_000, 8, 64
    X, _, _, _ = orc.mixture(61, n, d, k, -6, 6, 0.7, 1.5)
    C0 = X[:k].copy()
    out = {}
    for tr in (0, km.KM_BLOCK):
        e = km.KMeans(n, d, k, full_gemm=True, tile_rows=tr)
        e.set_data(X)
        e.set_centers(C0)
        e.run(3)
        out[tr] = (e.labels(), e.centers())
        e.close()
    assert np.array_equal(out[0][0], out[km.KM_BLOCK][0])
    assert np.array_equal(out[0][1], out[km.KM_BLOCK][1])


def test_compute_kmeans_f_full_gemm_equals_fused_on_separated_data(orc):
    # centres start near the component means: no near-ties in any iteration
    X, _, means, _ = orc.mixture(61, 40_000, 24, 16, -20, 20, 0.5, 0.5)
    C0 = (means + 0.05).astype(np.float32)
    ref = orc.lloyd(X, C0, 6)
    for full in (False, True):
        st, res = km.compute_kmeans_f(X, 16, -1.0, 6, result=C0.T.copy(), init_from_result=True,
                                      full_gemm=full)
        assert st == km.KMEANS_NONCONVERGENCE
        assert np.array_equal(res, ref["centers"].T)


def test_compute_kmeans_d_full_gemm_equals_fused_on_separated_data(orc):
    X, _, means, _ = orc.mixture(62, 40_000, 24, 16, -20, 20, 0.5, 0.5)
    X = X.astype(np.float64) + 1e-6
    C0 = means.astype(np.float64) + 0.05
    ref = orc.lloyd64(X, C0, 6)
    st, res = km.compute_kmeans_d(X, 16, -1.0, 6, result=C0.T.copy(), init_from_result=True,
                                  full_gemm=True)
    assert np.array_equal(res, ref["centers"].T)


def test_full_gemm_rejects_reduced_precision():
    with pytest.raises(km.KmeansError):
        km.KMeans(1000, 8, 4, precision="bf16", full_gemm=True)


def test_gemm_assign_single_centre_and_dimension(orc):
    # K = 1: every label 0; D = 1 with two centres: the argmin on a line
    X = np.linspace(-5, 5, 1001, dtype=np.float32)[:, None]
    e = km.KMeans(1001, 1, 1, full_gemm=True)
    e.set_data(X)
    e.set_centers(np.array([[0.5]], np.float32))
    e.run(1)
    assert np.all(e.labels() == 0)
    e.close()
    e = km.KMeans(1001, 1, 2, full_gemm=True)
    e.set_data(X)
    e.set_centers(np.array([[-2.0], [3.0]], np.float32))
    e.run(1)
    lab = e.labels()
    e.close()
    check_labels(orc, X, np.array([[-2.0], [3.0]], np.float32), lab)
