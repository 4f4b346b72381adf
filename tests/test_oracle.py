"""Pins for the CPU oracle (no GPU).  The reference ships no golden vectors
or known-answer tests for this path, so each oracle function is pinned here by
exact rational arithmetic, brute force in FP64, closed forms, hand-computed
known answers or an independent library (scikit-learn).  Each check is chosen
so that a plausible slip (wrong sign on the norm term, wrong tie rule, a
dropped chunk level, old instead of new centres) fails it."""
from fractions import Fraction
import math

import numpy as np
import pytest


# ------------------------------------------------------------------ helpers
def rnd32(q):
    """Round an exact rational to the nearest float32 (ties to even)."""
    q = Fraction(q)
    if q == 0:
        return 0.0
    s = -1 if q < 0 else 1
    a = abs(q)
    e = a.numerator.bit_length() - a.denominator.bit_length()
    if Fraction(2) ** e > a:
        e -= 1
    m = a / Fraction(2) ** (e - 23)  # in [2^23, 2^24)
    r = round(m)  # Fraction.__round__ rounds half to even
    return float(s * r * Fraction(2) ** (e - 23))


def exact_chain(x, c, ncj):
    s = rnd32(Fraction(-0.5) * Fraction(float(np.float32(ncj))))
    for a, b in zip(x, c):
        s = rnd32(Fraction(float(a)) * Fraction(float(b)) + Fraction(s))
    return s


def exact_norm(c):
    s = 0.0
    for a in c:
        s = rnd32(Fraction(float(a)) ** 2 + Fraction(s))
    return s


# ------------------------------------------------------------------ scores
def test_norm_and_score_are_single_rounding_fma_chains(orc):
    rng = np.random.default_rng(7)
    for d in (1, 3, 17, 40):
        for _ in range(20):
            x = rng.standard_normal(d).astype(np.float32) * 3
            c = rng.standard_normal(d).astype(np.float32) * 3
            nc = orc.col_norms(c[None, :])[0]
            assert float(nc) == exact_norm(c)
            assert float(orc.row_norms(c[None, :])[0]) == exact_norm(c)
            assert float(orc.score(x, c, nc)) == exact_chain(x, c, nc)


def test_row_norms_fma_chain_known_answer(orc):
    # each row against the chain computed in exact rationals with one
    # float32 rounding per step: the same squares in both orders, an exact
    # case (3-4-5) and squares below the smallest subnormal (they vanish)
    rows = np.array([[1.0 + 2.0 ** -12, 1.0], [1.0, 1.0 + 2.0 ** -12], [3.0, 4.0],
                     [2.0 ** -70, 2.0 ** -70]], np.float32)
    got = orc.row_norms(rows)
    for r, g in zip(rows, got):
        assert float(g) == exact_norm(r)
    assert float(got[2]) == 25.0


def test_fma_chain_differs_from_mul_add(orc):
    # a case where rounding the product first changes the result: the chain
    # must keep the product unrounded (fma), as fma.hcu:13 does.  With
    # ncj = 2 the chain starts at -1 and adds (1+2^-12)^2 = 1 + 2^-11 + 2^-24.
    x = np.array([1.0 + 2.0 ** -12], np.float32)
    s = orc.score(x, x, 2.0)
    assert s == exact_chain(x, x, 2.0) == 2.0 ** -11 + 2.0 ** -24
    twostep = float(np.float32(x[0] * x[0])) - 1.0
    assert twostep == 2.0 ** -11 and s != twostep


def test_assign_matches_fp64_bruteforce(orc):
    X, z, means, _ = orc.mixture(11, 3000, 24, 40, -10, 10, 0.5, 0.5)
    C = means + 0.01
    lab, best, second = orc.assign(X, C, with_scores=True)
    D2 = ((X.astype(np.float64)[:, None, :] - C.astype(np.float64)[None]) ** 2).sum(-1)
    ref = D2.argmin(1)
    srt = np.sort(D2, 1)
    clear = (srt[:, 1] - srt[:, 0]) > 1e-2  # far from any FP32 near-tie
    assert clear.mean() > 0.99
    assert np.array_equal(lab[clear], ref[clear])
    assert np.all(best >= second)


def test_assign_uses_norm_term_with_correct_sign(orc):
    # x.c alone would pick the long centre; the distance picks the short one.
    X = np.array([[1.0, 0.0]], np.float32)
    C = np.array([[10.0, 0.0], [1.0, 0.1]], np.float32)
    assert orc.assign(X, C)[0] == 1


def test_assign_tie_goes_to_lowest_index(orc):
    rng = np.random.default_rng(3)
    X = rng.standard_normal((500, 8)).astype(np.float32)
    base = rng.standard_normal((6, 8)).astype(np.float32)
    C = np.concatenate([base, base[[2, 0, 5]]])  # centres 6,7,8 duplicate 2,0,5
    lab = orc.assign(X, C)
    assert lab.max() < 6
    lab2 = orc.assign(X, C[::-1].copy())  # duplicates now come first at 0,1,2
    inv = {8 - j: j for j in range(9)}
    expect = np.array([min(inv[l], {2: 8 - 6, 0: 8 - 7, 5: 8 - 8}.get(l, 99)) for l in lab])
    assert np.array_equal(lab2, expect)


def test_assign_k_not_multiple_of_block(orc):
    X, _, means, _ = orc.mixture(5, 700, 9, 300, -20, 20, 0.3, 0.3)
    lab = orc.assign(X, means)
    D2 = ((X.astype(np.float64)[:, None, :] - means.astype(np.float64)[None]) ** 2).sum(-1)
    assert (lab == D2.argmin(1)).mean() > 0.999


# ------------------------------------------------------------------ canonical sum
def test_csum_short_is_left_to_right(orc):
    rng = np.random.default_rng(1)
    for n in (0, 1, 2, 255, 256):
        v = rng.standard_normal(n) * 10.0 ** rng.integers(-8, 8, n)
        s = 0.0
        for t in v:
            s += float(t)
        assert orc.csum(v) == s


def test_csum_known_answer_pins_chunking(orc):
    # 2^53 followed by ones: a left-to-right sum loses every +1 (ties to even),
    # fsum keeps all, CSUM keeps exactly the ones outside the first chunk.
    v = np.ones(512)
    v[0] = 2.0 ** 53
    assert orc.csum(v) == 2.0 ** 53 + 256
    assert math.fsum(v) == 2.0 ** 53 + 511
    v = np.ones(256 * 256 + 256)
    v[0] = 2.0 ** 53
    # level 1: chunk0 = 2^53, 256 other chunks of 256; level 2: group0 =
    # 2^53 + 255*256, group1 = 256; total 2^53 + 65536.
    assert orc.csum(v) == 2.0 ** 53 + 65536


def test_csum_accuracy(orc):
    rng = np.random.default_rng(2)
    v = rng.standard_normal(100_000)
    assert abs(orc.csum(v) - math.fsum(v)) < 1e-11 * np.abs(v).sum()


# ------------------------------------------------------------------ update
def test_update_matches_fp64_means(orc):
    X, z, means, _ = orc.mixture(4, 20_000, 16, 32, -5, 5, 1.0, 1.0)
    C0 = X[:32].copy()
    lab = orc.assign(X, C0)
    C1, cnt, S, empty = orc.update(X, lab, C0)
    M, cnt64 = orc.means_fp64(X, lab, 32)
    assert np.array_equal(cnt, np.bincount(lab, minlength=32))
    assert np.array_equal(cnt, cnt64)
    ok = cnt > 0
    assert np.allclose(C1[ok], M[ok], rtol=1e-6, atol=1e-6)
    assert np.array_equal(C1[ok], M[ok].astype(np.float32)) or np.abs(C1[ok] - M[ok]).max() < 1e-5
    assert np.array_equal(C1[~ok], C0[~ok])
    assert empty == (~ok).sum()


def test_update_empty_cluster_keeps_previous_centre(orc):
    X = np.array([[0.0, 0.0], [0.1, 0.0], [5.0, 5.0]], np.float32)
    C0 = np.array([[0.0, 0.0], [5.0, 5.0], [100.0, 100.0]], np.float32)
    lab = orc.assign(X, C0)
    assert list(lab) == [0, 0, 1]
    C1, cnt, S, empty = orc.update(X, lab, C0)
    assert empty == 1 and list(cnt) == [2, 1, 0]
    assert np.array_equal(C1[2], C0[2])
    assert C1[0, 0] == np.float32(0.05) and C1[1, 0] == 5.0


def test_update_block_structure_known_answer(orc):
    # one centre, d=1: x0 = 2^53 then ones across two KM_BLOCK blocks.
    n = orc.KM_BLOCK + 1000
    X = np.ones((n, 1), np.float32)
    X[0, 0] = 2.0 ** 53
    lab = np.zeros(n, np.int32)
    C1, cnt, S, _ = orc.update(X, lab, np.zeros((1, 1), np.float32))
    # block 0: chunk0 loses its 255 ones, every other one is kept; block 1 exact
    assert S[0, 0] == 2.0 ** 53 + (orc.KM_BLOCK - 256) + 1000
    assert cnt[0] == n


def q_order_rows(d=2):
    """One centre, 513 members: row 0 = (2^30, 0, ...), 255 rows (0, 8, ...),
    256 rows (0, 1, ...), one row (0, 12, ...)."""
    X = np.zeros((513, d), np.float32)
    X[0, 0] = 2.0 ** 30
    X[1:256, 1] = 8.0
    X[256:512, 1] = 1.0
    X[512, 1] = 12.0
    return X


Q_ORDER_ANSWER = 2.0 ** 60 + 16896.0


def test_moments_squared_norm_order_known_answer(orc):
    # The squared-norm moment sums each KM_CHUNK piece of members dimension by
    # dimension, then combines the pieces as CSUM does:
    #   piece 0 = fl(2^60 + 255 * 64) = 2^60 + 16384 (ulp 256 at 2^60),
    #   piece 1 = 256, piece 2 = 144,
    #   Q = fl(fl(2^60 + 16384 + 256) + 144) = fl(2^60 + 16784) = 2^60 + 16896.
    # Summing each point's squared norm first would give 2^60 + 512, and
    # per-dimension totals over all members 2^60 + 16640: both are excluded.
    for d in (2, 16):
        X = q_order_rows(d)
        lab = np.zeros(513, np.int32)
        _, Q, cnt = orc.moments(X, lab, 1)
        assert cnt[0] == 513
        assert Q[0] == Q_ORDER_ANSWER
        _, Q64, _ = orc.moments64(X.astype(np.float64), lab, 1)
        assert Q64[0] == Q_ORDER_ANSWER
        # against zero centres the compactness is the moment itself
        assert orc.compactness(X, lab, np.zeros((1, d), np.float32)) == Q_ORDER_ANSWER


# ------------------------------------------------------------------ compactness
def test_compactness_fp64_and_identity(orc):
    X, z, means, _ = orc.mixture(9, 30_000, 20, 16, -6, 6, 0.8, 1.2)
    C0 = X[:16].copy()
    lab = orc.assign(X, C0)
    C1, cnt, S, _ = orc.update(X, lab, C0)
    comp = orc.compactness(X, lab, C1)
    ref = orc.compactness_fp64(X, lab, C1)
    assert abs(comp - ref) <= 1e-5 * ref
    # sum ||x - mu||^2 = sum ||x||^2 - sum_j n_j ||mu_j||^2 at the exact means
    M, cnt64 = orc.means_fp64(X, lab, 16)
    X64 = X.astype(np.float64)
    lhs = ((X64 - M[lab]) ** 2).sum()
    rhs = (X64 ** 2).sum() - (cnt64[:, None] * M ** 2).sum()
    assert abs(lhs - rhs) < 1e-9 * lhs
    assert abs(comp - lhs) <= 1e-5 * lhs
    # the moments the formula uses
    S, Q, cnt = orc.moments(X, lab, 16)
    assert np.array_equal(cnt, cnt64)
    assert np.allclose(Q, np.bincount(lab, (X64 ** 2).sum(1), minlength=16), rtol=1e-13)
    assert np.allclose(S, np.stack([np.bincount(lab, X64[:, q], minlength=16) for q in range(20)], 1),
                       rtol=1e-12, atol=1e-9)
    _, _, _, _, comp2 = orc.update(X, lab, C0, with_comp=True)
    assert comp2 == comp


def test_compactness_uses_given_centres(orc):
    X = np.array([[0.0], [2.0]], np.float32)
    lab = np.array([0, 0], np.int32)
    assert orc.compactness(X, lab, np.array([[1.0]], np.float32)) == 2.0
    assert orc.compactness(X, lab, np.array([[0.0]], np.float32)) == 4.0


# ------------------------------------------------------------------ Lloyd loop
def test_lloyd_1d_two_clusters_by_hand(orc):
    X = np.array([[0.0], [1.0], [2.0], [10.0], [11.0]], np.float32)
    r = orc.lloyd(X, np.array([[0.0], [1.0]], np.float32), 10, criterion=0.0)
    # it1: {0} {1,2,10,11} -> c=(0, 6); comp = 0 + 25+16+16+25 = 82
    # it2: {0,1,2} {10,11} -> c=(1, 10.5); comp = 1+0+1+0.25+0.25 = 2.5
    # it3: unchanged -> comp 2.5, relErr 0 -> stop
    assert list(r["comp_hist"]) == [82.0, 2.5, 2.5]
    assert r["iters"] == 3 and r["status"] == 0
    assert np.array_equal(r["centers"], np.array([[1.0], [10.5]], np.float32))
    assert list(r["labels"]) == [0, 0, 0, 1, 1]


def test_lloyd_nonconvergence_status(orc):
    X, _, _, _ = orc.mixture(1, 5000, 4, 8, -3, 3, 1.0, 1.0)
    r = orc.lloyd(X, X[:8], 2, criterion=1e-12)
    assert r["iters"] == 2 and r["status"] == 13


def test_lloyd_closed_form_tiny_sigma(orc):
    # sigma -> 0: one step from one point per component lands on the means.
    X, z, means, _ = orc.mixture(21, 4000, 6, 5, -10, 10, 1e-6, 1e-6)
    first = [int(np.nonzero(z == j)[0][0]) for j in range(5)]
    r = orc.lloyd(X, X[first], 1)
    assert np.array_equal(r["labels"], z)
    assert np.allclose(r["centers"], means, atol=1e-5)


def test_lloyd_matches_sklearn(orc):
    from sklearn.cluster import KMeans
    X, z, means, _ = orc.mixture(0, 20_000, 8, 12, -10, 10, 1.0, 1.0)
    C0 = X[:12].copy()
    r = orc.lloyd(X, C0, 15)
    km = KMeans(n_clusters=12, init=C0.astype(np.float64), n_init=1, max_iter=15, tol=0.0,
                algorithm="lloyd").fit(X.astype(np.float64))
    # sklearn's labels_ come from a final assignment against the final centres
    assert np.array_equal(orc.assign(X, r["centers"]), km.labels_)
    assert np.abs(r["centers"] - km.cluster_centers_).max() < 1e-5


# ------------------------------------------------------------------ generators
def test_generator_is_order_free_and_matches_spec(orc):
    X, z, means, sig = orc.mixture(1, 50_000, 4, 8, -5, 5, 0.5, 2.0)
    X2, z2, _, _ = orc.mixture(1, 1000, 4, 8, -5, 5, 0.5, 2.0, i0=20_000)
    assert np.array_equal(X[20_000:21_000], X2) and np.array_equal(z[20_000:21_000], z2)
    assert means.min() >= -5 and means.max() <= 5 and sig.min() >= 0.5 and sig.max() <= 2.0
    for j in range(8):
        r = X[z == j] - means[j]
        assert abs(r.std() - sig[j]) < 0.05 * sig[j]
        assert np.abs(r.mean(0)).max() < 0.1
    assert abs(np.bincount(z).std() / np.bincount(z).mean()) < 0.05


def test_generator_zipf_weights(orc):
    X, z, _, _ = orc.mixture(2, 200_000, 2, 64, -8, 8, 1.0, 1.0, zipf_s=1.1)
    cnt = np.bincount(z, minlength=64).astype(float)
    w = np.arange(1, 65) ** -1.1
    w /= w.sum()
    assert np.allclose(cnt / cnt.sum(), w, atol=3e-3)


def test_sample_init_distinct(orc):
    idx = orc.sample_init(1, 10_000_000, 1024)
    assert len(set(idx.tolist())) == 1024 and idx.min() >= 0 and idx.max() < 10_000_000
    assert np.array_equal(idx, orc.sample_init(1, 10_000_000, 1024))


# ------------------------------------------------------------------ MLP proposer
def test_mlp_oracle_gradients_by_finite_differences(orc):
    rng = np.random.default_rng(4)
    din, h, g, B = 16, 24, 7, 40
    n = h * din + h + h * h + h + g * h + g
    P = rng.standard_normal(n) * 0.3
    X = rng.standard_normal((B, din)).astype(np.float32)
    t = rng.integers(0, g, B)
    loss, G = orc.mlp_fwd_bwd(din, h, g, X, t, P)
    assert np.isfinite(loss) and loss > 0
    for i in rng.choice(n, 40, replace=False):
        e = np.zeros(n)
        e[i] = 1e-6
        lp, _ = orc.mlp_fwd_bwd(din, h, g, X, t, P + e)
        lm, _ = orc.mlp_fwd_bwd(din, h, g, X, t, P - e)
        fd = (lp - lm) / 2e-6
        assert abs(fd - G[i]) <= 1e-6 + 1e-4 * abs(G[i]), (i, fd, G[i])


def test_mlp_oracle_loss_closed_form(orc):
    # zero weights: uniform softmax, loss = log(g), db3 = mean(1/g - onehot)
    din, h, g, B = 16, 8, 5, 10
    n = h * din + h + h * h + h + g * h + g
    X = np.ones((B, din), np.float32)
    t = np.arange(B) % g
    loss, G = orc.mlp_fwd_bwd(din, h, g, X, t, np.zeros(n))
    assert abs(loss - np.log(g)) < 1e-12
    gb3 = G[-g:]
    expect = np.full(g, 1.0 / g) - np.bincount(t, minlength=g) / B
    assert np.allclose(gb3, expect, atol=1e-15)


# ------------------------------------------------------------------ double precision
def exact_chain64(x, c, ncj):
    """fma chain in double by exact rationals: float(Fraction) rounds to nearest even."""
    s = float(Fraction(-0.5) * Fraction(float(ncj)))
    for a, b in zip(x, c):
        s = float(Fraction(float(a)) * Fraction(float(b)) + Fraction(s))
    return s


def test_f64_norm_and_score_are_single_rounding_fma_chains(orc):
    rng = np.random.default_rng(8)
    for d in (1, 3, 17, 40):
        for _ in range(20):
            x = rng.standard_normal(d) * 3
            c = rng.standard_normal(d) * 3
            nc = orc.col_norms64(c[None, :])[0]
            ref = 0.0
            for a in c:
                ref = float(Fraction(float(a)) ** 2 + Fraction(ref))
            assert nc == ref
            assert orc.row_norms64(c[None, :])[0] == ref
            assert orc.score64(x, c, nc) == exact_chain64(x, c, nc)


def test_f64_fma_chain_differs_from_mul_add(orc):
    # (1 + 2^-30)^2 = 1 + 2^-29 + 2^-60: the 2^-60 survives only without a
    # rounding of the product (score = -1 + x*x with ncj = 2)
    x = np.array([1.0 + 2.0 ** -30])
    s = orc.score64(x, x, 2.0)
    assert s == 2.0 ** -29 + 2.0 ** -60
    assert (x[0] * x[0]) - 1.0 == 2.0 ** -29


def test_f64_assign_matches_bruteforce_and_ties(orc):
    rng = np.random.default_rng(9)
    X = rng.standard_normal((2000, 13)) * 4
    C = rng.standard_normal((37, 13)) * 4
    lab = orc.assign64(X, C)
    D2 = ((X[:, None, :] - C[None]) ** 2).sum(-1)
    srt = np.sort(D2, 1)
    clear = (srt[:, 1] - srt[:, 0]) > 1e-9
    assert clear.all()
    assert np.array_equal(lab, D2.argmin(1))
    # duplicate centres: the lower index wins
    C2 = np.concatenate([C[:5], C[:5]])
    lab2 = orc.assign64(X, C2)
    assert lab2.max() < 5


def test_f64_moments_equal_fp32_moments_on_float_data(orc):
    # on float-valued points the double engine's sums and squared norms are the
    # FP32 engine's bit for bit (same canonical order; the squares are exact),
    # and rounding its centres to float gives the FP32 centres
    X, _, _, _ = orc.mixture(21, 20_000, 9, 12, -5, 5, 1.0, 1.0)
    C = X[:12].copy()
    lab = orc.assign(X, C)
    S32, Q32, n32 = orc.moments(X, lab, 12)
    S64, Q64, n64 = orc.moments64(X.astype(np.float64), lab, 12)
    assert np.array_equal(S32, S64) and np.array_equal(Q32, Q64) and np.array_equal(n32, n64)
    C32 = orc.update(X, lab, C)[0]
    C64, cnt, empty, comp = orc.update64(X.astype(np.float64), lab, C.astype(np.float64))
    assert empty == 0
    assert np.array_equal(C64.astype(np.float32), C32)


def test_f64_update_and_compactness_vs_numpy(orc):
    rng = np.random.default_rng(10)
    X = rng.standard_normal((30_000, 7)) * 2 + 1
    lab = rng.integers(0, 9, 30_000).astype(np.int32)
    lab[lab == 4] = 5  # centre 4 empty
    C0 = rng.standard_normal((9, 7))
    C, cnt, empty, comp = orc.update64(X, lab, C0)
    assert empty == 1 and cnt[4] == 0
    assert np.array_equal(C[4], C0[4])
    for j in range(9):
        if cnt[j]:
            np.testing.assert_allclose(C[j], X[lab == j].mean(0), rtol=1e-13, atol=1e-14)
    direct = ((X - C[lab]) ** 2).sum()
    assert abs(comp - direct) <= 1e-10 * direct
    assert orc.compactness64(X, lab, C) == comp


def test_f64_lloyd_closed_form_and_criterion(orc):
    # well separated, tiny spread: one step lands on the component means
    rng = np.random.default_rng(11)
    means = rng.uniform(-50, 50, (6, 5))
    z = np.arange(6000) % 6
    X = means[z] + rng.standard_normal((6000, 5)) * 1e-3
    r = orc.lloyd64(X, X[:6], 3)
    for j in range(6):
        np.testing.assert_allclose(r["centers"][j], X[z == j].mean(0), rtol=0, atol=1e-12)
    assert r["status"] == 13
    r2 = orc.lloyd64(X, X[:6], 50, criterion=1e-9)
    assert r2["status"] == 0 and r2["iters"] < 50


def test_row_transform_min_known_answers(orc):
    # one row, n = 4: values rn + cn_j - 2 P_j = [5, 3, 3, 7] -> index 1 (tie to the lower)
    for dt in (np.float32, np.float64):
        P = np.array([[0.0, 1.0, 0.5, -1.0]], dt)
        rn = np.array([1.0], dt)
        cn = np.array([4.0, 4.0, 3.0, 4.0], dt)
        out = orc.row_transform_min(P, rn, cn, m0=2)
        assert list(out) == [-7, -7, 1]
        # nothing below FLT_MAX (for double too): -1 (P = -inf gives +inf, NaN never compares)
        big = np.array([[-np.inf, np.nan]], dt)
        assert orc.row_transform_min(big, np.zeros(1, dt), np.zeros(2, dt))[0] == -1
        if dt == np.float64:
            P2 = np.array([[0.0, 0.0]], dt)
            assert orc.row_transform_min(P2, np.zeros(1, dt), np.array([1e39, 2e39]))[0] == -1


def test_row_transform_min_matches_nearest_centre(orc):
    # with P = X C^T and the squared norms, the minimiser is the nearest centre
    X, z, means, _ = orc.mixture(22, 3000, 16, 30, -10, 10, 0.5, 0.5)
    C = means + 0.01
    X64, C64 = X.astype(np.float64), C.astype(np.float64)
    D2 = ((X64[:, None, :] - C64[None]) ** 2).sum(-1)
    for dt in (np.float32, np.float64):
        P = (X.astype(dt) @ C.astype(dt).T).astype(dt)
        rn = (X.astype(dt) ** 2).sum(1).astype(dt)
        cn = (C.astype(dt) ** 2).sum(1).astype(dt)
        out = orc.row_transform_min(P, rn, cn)
        srt = np.sort(D2, 1)
        clear = (srt[:, 1] - srt[:, 0]) > 1e-1
        assert clear.mean() > 0.99
        assert np.array_equal(out[clear], D2.argmin(1)[clear])


# ------------------------------------------------------------------ k-means++ seeding
def test_kmeanspp_one_pick_per_separated_component(orc):
    X, z, _, _ = orc.mixture(5, 50_000, 8, 20, -50, 50, 0.1, 0.1)
    for X_ in (X, X.astype(np.float64)):
        idx = orc.kmeanspp(X_, 20, 1)
        assert len(set(idx.tolist())) == 20
        assert len(set(z[idx].tolist())) == 20


def test_kmeanspp_first_pick_seed_and_type(orc):
    X, _, _, _ = orc.mixture(6, 3000, 5, 7, -5, 5, 1.0, 1.0)
    for seed in (0, 1, 99):
        idx = orc.kmeanspp(X, 7, seed)
        assert idx[0] == orc.lib().orc_rng_u64(seed, 6, 0) % 3000
        assert np.array_equal(idx, orc.kmeanspp(X, 7, seed))
        # on float-valued data the double variant draws the same points
        assert np.array_equal(idx, orc.kmeanspp(X.astype(np.float64), 7, seed))
    assert not np.array_equal(orc.kmeanspp(X, 7, 0), orc.kmeanspp(X, 7, 1))


def test_kmeanspp_draw_is_proportional_to_squared_distance(orc):
    # 1-D points; over many seeds the second pick must follow D(x) / sum D for
    # the first pick (not sqrt(D), not uniform): chi-square against expectation
    X = np.array([[0.0], [1.0], [3.0], [3.5], [7.0], [10.0]], np.float32)
    obs = np.zeros(6)
    exp = np.zeros(6)
    for seed in range(6000):
        i0, i1 = orc.kmeanspp(X, 2, seed)
        D = (X[:, 0].astype(np.float64) - X[i0, 0]) ** 2
        exp += D / D.sum()
        obs[i1] += 1
    chi2 = ((obs - exp) ** 2 / exp).sum()
    assert chi2 < 25.0  # 5 degrees of freedom: p ~ 1e-4
    # the alternatives are far off
    alt = np.zeros(6)
    for seed in range(6000):
        i0 = orc.kmeanspp(X, 1, seed)[0]
        D = np.abs(X[:, 0].astype(np.float64) - X[i0, 0])
        alt += D / D.sum()
    assert ((obs - alt) ** 2 / alt).sum() > 100.0


def test_kmeanspp_tree_descent_and_zero_total(orc):
    # 70,000 points (a two-level tree) all at the origin but one: step 1 must
    # find the only point with a positive distance
    X = np.zeros((70_000, 3), np.float32)
    X[61_234] = [5.0, -1.0, 2.0]
    for seed in range(5):
        idx = orc.kmeanspp(X, 2, seed)
        if idx[0] != 61_234:
            assert idx[1] == 61_234
    # all points identical: total 0, the draw is rng_u64(seed, 6, t) % m
    Y = np.ones((1000, 4), np.float32)
    idx = orc.kmeanspp(Y, 3, 7)
    assert idx[1] == orc.lib().orc_rng_u64(7, 6, 1) % 1000
    assert idx[2] == orc.lib().orc_rng_u64(7, 6, 2) % 1000
