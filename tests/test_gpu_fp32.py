"""GPU parity of the plain FP32 path against the CPU oracle, through the C-ABI.

Bar: bit-exact labels, scores, centres, counts and compactness (the GPU runs
the same fma chains and the same canonical FP64 summation order)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
import kmeans_amd as km  # noqa: E402


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def gpu_assign(X, C):
    k, d = C.shape
    Xd, Cd = dev(X), dev(C)
    ld = km.centers_ld(k)
    Ct = torch.zeros((d, ld), dtype=torch.float32, device="cuda")
    nc = torch.empty(k, dtype=torch.float32, device="cuda")
    lab = torch.empty(X.shape[0], dtype=torch.int32, device="cuda")
    best = torch.empty(X.shape[0], dtype=torch.float32, device="cuda")
    km.transpose_centers(k, d, Cd.data_ptr(), Ct.data_ptr())
    km.col_norms(k, d, Cd.data_ptr(), nc.data_ptr())
    km.assign_f32(X.shape[0], d, Xd.data_ptr(), k, Ct.data_ptr(), nc.data_ptr(), lab.data_ptr(),
                  best.data_ptr())
    torch.cuda.synchronize()
    return lab.cpu().numpy(), best.cpu().numpy(), nc.cpu().numpy()


@pytest.mark.parametrize("n,d,k", [(1000, 3, 2), (777, 16, 5), (4096, 17, 64), (3000, 32, 65),
                                   (2500, 128, 300), (5000, 128, 1024), (513, 200, 33)])
def test_assign_bit_exact(orc, n, d, k):
    print(f"assign n={n} d={d} k={k}", flush=True)
    X, _, means, _ = orc.mixture(100 + d, n, d, k, -5, 5, 0.5, 2.0)
    C = X[np.arange(k) * (n // k)].copy()
    lab, best, nc = gpu_assign(X, C)
    rl, rb, _ = orc.assign(X, C, with_scores=True)
    assert np.array_equal(nc, orc.col_norms(C))
    assert np.array_equal(lab, rl)
    assert np.array_equal(best, rb)


@pytest.mark.parametrize("case", ["wide", "subnormal", "cancel"])
def test_assign_adversarial_rounding(orc, case):
    """The matrix-core chain must round exactly like the fmaf chain where the
    orders and roundings of alternative evaluations differ most: values over
    ~100 binades, subnormal products and partial sums, and near-cancelling
    terms.  Labels and best scores bit for bit against the oracle."""
    rng = np.random.default_rng({"wide": 1, "subnormal": 2, "cancel": 3}[case])
    n, d, k = 3000, 37, 70
    if case == "wide":
        X = rng.standard_normal((n, d)) * np.exp2(rng.integers(-50, 50, (n, d)))
        C = rng.standard_normal((k, d)) * np.exp2(rng.integers(-50, 50, (k, d)))
    elif case == "subnormal":
        X = rng.standard_normal((n, d)) * 2.0 ** -70
        C = rng.standard_normal((k, d)) * 2.0 ** -62
    else:
        X = rng.standard_normal((n, d)) * 1e3
        C = rng.standard_normal((k, d))
        C[:, 1::2] = C[:, 0::2][:, : d // 2] * -1.0000001  # pairs that nearly cancel
        X[:, 1::2] = X[:, 0::2][:, : d // 2]
    X, C = X.astype(np.float32), C.astype(np.float32)
    lab, best, nc = gpu_assign(X, C)
    rl, rb, _ = orc.assign(X, C, with_scores=True)
    assert np.array_equal(nc, orc.col_norms(C))
    assert np.array_equal(lab, rl)
    assert np.array_equal(best.view(np.uint32), rb.view(np.uint32))


def test_assign_ties_lowest_index(orc):
    rng = np.random.default_rng(5)
    X = rng.standard_normal((2048, 24)).astype(np.float32)
    base = rng.standard_normal((300, 24)).astype(np.float32)
    C = np.concatenate([base, base[::-1]])  # every centre duplicated, 600 total
    lab, _, _ = gpu_assign(X, C)
    assert np.array_equal(lab, orc.assign(X, C))
    assert lab.max() < 300


def test_assign_rows_indirect(orc):
    X, _, _, _ = orc.mixture(3, 4000, 40, 100, -4, 4, 1.0, 1.0)
    C = X[:100].copy()
    k, d = C.shape
    rows = np.sort(np.random.default_rng(0).choice(4000, 999, replace=False)).astype(np.int32)
    Xd, Cd, rd = dev(X), dev(C), dev(rows)
    Ct = torch.zeros((d, km.centers_ld(k)), dtype=torch.float32, device="cuda")
    nc = torch.empty(k, dtype=torch.float32, device="cuda")
    lab = torch.full((4000,), -7, dtype=torch.int32, device="cuda")
    mdev = dev(np.array([999], np.int32))
    km.transpose_centers(k, d, Cd.data_ptr(), Ct.data_ptr())
    km.col_norms(k, d, Cd.data_ptr(), nc.data_ptr())
    km.assign_f32(4000, d, Xd.data_ptr(), k, Ct.data_ptr(), nc.data_ptr(), lab.data_ptr(),
                  rows_ptr=rd.data_ptr(), m_dev_ptr=mdev.data_ptr())
    out = lab.cpu().numpy()
    ref = orc.assign(X, C)
    assert np.array_equal(out[rows], ref[rows])
    mask = np.ones(4000, bool)
    mask[rows] = False
    assert np.all(out[mask] == -7)


def gpu_update(X, lab, C, misalign=False):
    n, d = X.shape
    k = C.shape[0]
    Xd, ld, Cd = dev(X), dev(lab), dev(C)
    if misalign:  # the same rows at an address 4 bytes past a 16-byte boundary
        buf = torch.empty(n * d + 1, dtype=torch.float32, device="cuda")
        buf[1:] = Xd.reshape(-1)
        Xd = buf[1:]
    Cn = torch.empty_like(Cd)
    cnt = torch.empty(k, dtype=torch.int64, device="cuda")
    emp = torch.zeros(1, dtype=torch.int32, device="cuda")
    nc = torch.empty(k, dtype=torch.float32, device="cuda")
    wsb = km.update_workspace_size(n, d, k)
    ws = torch.empty(wsb, dtype=torch.uint8, device="cuda")
    km.update(n, d, k, Xd.data_ptr(), ld.data_ptr(), Cd.data_ptr(), Cn.data_ptr(), ws.data_ptr(),
              wsb, nc_ptr=nc.data_ptr(), counts_ptr=cnt.data_ptr(), empty_ptr=emp.data_ptr())
    out = torch.empty(1, dtype=torch.float64, device="cuda")
    km.compactness(n, d, k, Xd.data_ptr(), ld.data_ptr(), Cn.data_ptr(), out.data_ptr(),
                   ws.data_ptr(), wsb)
    torch.cuda.synchronize()
    return (Cn.cpu().numpy(), cnt.cpu().numpy(), int(emp.item()), nc.cpu().numpy(),
            float(out.item()))


@pytest.mark.parametrize("n,d,k", [(5000, 1, 4), (5000, 3, 7), (20000, 16, 64), (30000, 33, 100),
                                   (12345, 128, 1024), (4000, 300, 20),
                                   # four dimensions per thread (d % 4 == 0, 16 <= d <= 256):
                                   # unstaged member indices (d = 20), staged (96, 256)
                                   (50000, 20, 37), (40000, 96, 300), (8000, 256, 16),
                                   # k > 2048: the larger scatter sub-block
                                   (60000, 8, 2500)])
def test_update_and_compactness_bit_exact(orc, n, d, k):
    print(f"update n={n} d={d} k={k}", flush=True)
    X, _, _, _ = orc.mixture(7 + k, n, d, k, -6, 6, 0.5, 1.5)
    C0 = X[:k].copy()
    lab = orc.assign(X, C0)
    C1, cnt, emp, nc, comp = gpu_update(X, lab, C0)
    rC, rcnt, _, rempty = orc.update(X, lab, C0)
    assert np.array_equal(cnt, rcnt)
    assert emp == rempty
    assert np.array_equal(C1, rC)
    assert np.array_equal(nc, orc.col_norms(rC))
    assert comp == orc.compactness(X, lab, rC)


def test_update_skewed_and_empty(orc):
    # one giant cluster (many chunks), several empty ones
    rng = np.random.default_rng(9)
    n, d, k = 300_000, 8, 50
    X = rng.standard_normal((n, d)).astype(np.float32)
    lab = np.where(rng.random(n) < 0.9, 3, rng.integers(10, 20, n)).astype(np.int32)
    C0 = rng.standard_normal((k, d)).astype(np.float32)
    C1, cnt, emp, nc, comp = gpu_update(X, lab, C0)
    rC, rcnt, _, rempty = orc.update(X, lab, C0)
    assert emp == rempty == k - 11
    assert np.array_equal(C1, rC) and np.array_equal(cnt, rcnt)
    assert comp == orc.compactness(X, lab, rC)


def test_update_misaligned_rows(orc):
    # float4 row reads need 16-byte aligned rows; a misaligned borrowed pointer
    # takes the one-dimension-per-thread kernel and gives the same sums
    X, _, _, _ = orc.mixture(3, 20000, 32, 50, -6, 6, 0.5, 1.5)
    lab = orc.assign(X, X[:50].copy())
    C0 = X[:50].copy()
    a = gpu_update(X, lab, C0)
    b = gpu_update(X, lab, C0, misalign=True)
    rC, rcnt, _, _ = orc.update(X, lab, C0)
    assert np.array_equal(a[0], rC) and np.array_equal(b[0], rC)
    assert a[4] == b[4] == orc.compactness(X, lab, rC)


@pytest.mark.parametrize("d", [16, 64])
def test_update_skewed_wide(orc, d):
    # a giant cluster (hundreds of full 256-member chunks plus a remainder) with
    # the four-dimensions-per-thread chunk sums, and empty clusters
    rng = np.random.default_rng(10 + d)
    n, k = 200_003, 40
    X = (rng.standard_normal((n, d)) * 3).astype(np.float32)
    lab = np.where(rng.random(n) < 0.8, 7, rng.integers(20, 30, n)).astype(np.int32)
    C0 = rng.standard_normal((k, d)).astype(np.float32)
    C1, cnt, emp, nc, comp = gpu_update(X, lab, C0)
    rC, rcnt, _, rempty = orc.update(X, lab, C0)
    assert emp == rempty == k - 11
    assert np.array_equal(C1, rC) and np.array_equal(cnt, rcnt)
    assert np.array_equal(nc, orc.col_norms(rC))
    assert comp == orc.compactness(X, lab, rC)


@pytest.mark.parametrize("d", [2, 16])
def test_update_squared_norm_order_known_answer(orc, d):
    # the oracle's pinned moment order (per KM_CHUNK piece, dimension by
    # dimension): against zero centres the compactness is the moment itself,
    # 2^60 + 16896; d = 2 takes the one-dimension-per-thread chunk sums, d = 16
    # the four-dimensions-per-thread ones
    X = np.zeros((513, d), np.float32)
    X[0, 0] = 2.0 ** 30
    X[1:256, 1] = 8.0
    X[256:512, 1] = 1.0
    X[512, 1] = 12.0
    lab = np.zeros(513, np.int32)
    n, k = 513, 1
    Xd, ld, Cz = dev(X), dev(lab), dev(np.zeros((1, d), np.float32))
    wsb = km.update_workspace_size(n, d, k)
    ws = torch.empty(wsb, dtype=torch.uint8, device="cuda")
    out = torch.empty(1, dtype=torch.float64, device="cuda")
    km.compactness(n, d, k, Xd.data_ptr(), ld.data_ptr(), Cz.data_ptr(), out.data_ptr(),
                   ws.data_ptr(), wsb)
    torch.cuda.synchronize()
    assert float(out.item()) == 2.0 ** 60 + 16896.0
    C1, cnt, emp, nc, comp = gpu_update(X, lab, np.zeros((1, d), np.float32))
    rC, _, _, _ = orc.update(X, lab, np.zeros((1, d), np.float32))
    assert np.array_equal(C1, rC) and comp == orc.compactness(X, lab, rC)


def test_update_spans_two_blocks(orc):
    # the oracle's block known answer: 2^53 then ones over KM_BLOCK + 1000 rows
    n = km.KM_BLOCK + 1000
    X = np.ones((n, 1), np.float32)
    X[0, 0] = 2.0 ** 53
    lab = np.zeros(n, np.int32)
    C0 = np.zeros((1, 1), np.float32)
    C1, cnt, emp, nc, comp = gpu_update(X, lab, C0)
    rC, rcnt, S, _ = orc.update(X, lab, C0)
    assert S[0, 0] == 2.0 ** 53 + (km.KM_BLOCK - 256) + 1000
    assert np.array_equal(C1, rC) and cnt[0] == n
    assert comp == orc.compactness(X, lab, rC)


@pytest.mark.parametrize("k", [101, 1024])
def test_update_staged_scatter_odd_and_full_k(orc, k):
    # a block of 37+ sub-blocks takes the LDS-staged scatter (k <= 1024):
    # odd k (the int arrays before its 64-bit masks are padded) and the
    # largest k; labels skewed so that some labels fill whole tiles
    n, d = 600_000, 5
    rng = np.random.default_rng(29)
    X = rng.standard_normal((n, d)).astype(np.float32)
    lab = np.where(rng.random(n) < 0.3, 0, rng.integers(0, k, n)).astype(np.int32)
    lab[: 20_000] = k - 1
    C0 = rng.standard_normal((k, d)).astype(np.float32)
    C1, cnt, emp, nc, comp = gpu_update(X, lab, C0)
    rC, rcnt, _, _ = orc.update(X, lab, C0)
    assert np.array_equal(C1, rC)
    assert np.array_equal(cnt, rcnt)
    assert comp == orc.compactness(X, lab, rC)


def test_engine_lloyd_config1_bit_exact(orc):
    X, C0 = orc.config1()
    ref = orc.lloyd(X, C0, 20)
    e = km.KMeans(X.shape[0], X.shape[1], 64)
    e.set_data(X)
    e.set_centers(C0)
    e.run(20)
    st = e.stats()
    assert st["iterations"] == 20
    assert np.array_equal(e.labels(), ref["labels"])
    assert np.array_equal(e.centers(), ref["centers"])
    assert st["compactness"] == ref["comp_hist"][-1]
    e.close()


def test_engine_update_overlaps_four_blocks(orc):
    # resident points over four canonical blocks (the last one short): the
    # blocks' sorts and chunk sums run on two side streams, the additions in
    # block order; labels, centres and compactness bit for bit
    n, d, k = 3 * km.KM_BLOCK + 4321, 4, 50
    X, _, _, _ = orc.mixture(18, n, d, k, -10, 10, 1.0, 1.0)
    C0 = X[:k].copy()
    ref = orc.lloyd(X, C0, 2)
    e = km.KMeans(n, d, k)
    e.set_data(X)
    e.set_centers(C0)
    e.run(2)
    assert np.array_equal(e.labels(), ref["labels"])
    assert np.array_equal(e.centers(), ref["centers"])
    assert e.stats()["compactness"] == ref["comp_hist"][-1]
    e.close()


@pytest.mark.parametrize("prec", ["fp32", "bf16"])
def test_update_reuses_unchanged_chunks(orc, prec):
    # every full update after the first reuses the chunk partials of the
    # centres' unchanged member prefixes (three canonical blocks, the last
    # short); labels, centres and compactness after every run length
    n, d, k = 2 * km.KM_BLOCK + 12_345, 8, 40
    X, _, _, _ = orc.mixture(21, n, d, k, -10, 10, 1.0, 1.0)
    C0 = X[:k].copy()
    e = km.KMeans(n, d, k, precision=prec)
    e.set_data(X)
    e.set_centers(C0)
    done = 0
    for steps in (1, 1, 2, 3):
        e.run(steps)
        done += steps
        ref = orc.lloyd(X, C0, done)
        assert np.array_equal(e.labels(), ref["labels"]), done
        assert np.array_equal(e.centers(), ref["centers"]), done
        assert e.stats()["compactness"] == ref["comp_hist"][-1], done
    e.close()


def test_update_reuse_not_across_data_sets(orc):
    # new points of the same shape: nothing of the previous data's partials
    # may be reused (one and two canonical blocks)
    for n in (150_000, km.KM_BLOCK + 1000):
        d, k = 8, 30
        X1, _, _, _ = orc.mixture(22, n, d, k, -10, 10, 1.0, 1.0)
        X2, _, _, _ = orc.mixture(23, n, d, k, -10, 10, 1.0, 1.0)
        e = km.KMeans(n, d, k)
        e.set_data(X1)
        e.set_centers(X1[:k].copy())
        e.run(3)
        e.set_data(X2)
        e.set_centers(X2[:k].copy())
        e.run(3)
        ref = orc.lloyd(X2, X2[:k].copy(), 3)
        assert np.array_equal(e.labels(), ref["labels"]), n
        assert np.array_equal(e.centers(), ref["centers"]), n
        e.close()


def test_update_reuse_after_mini_batch_windows(orc):
    # mini-batch windows update partial row ranges; the full updates after
    # them must not reuse partials from before them: the result equals a fresh
    # engine started from the centres the windows left
    n, d, k = 300_000, 8, 30
    X, _, _, _ = orc.mixture(24, n, d, k, -10, 10, 1.0, 1.0)
    C0 = X[:k].copy()
    e = km.KMeans(n, d, k)
    e.set_data(X)
    e.set_centers(C0)
    e.run(2)
    e.set_mini_batch(100_000)
    e.run(2)
    C_mb = e.centers()
    e.set_mini_batch(0)
    e.run(2)
    f = km.KMeans(n, d, k)
    f.set_data(X)
    f.set_centers(C_mb)
    f.run(2)
    assert np.array_equal(e.labels(), f.labels())
    assert np.array_equal(e.centers(), f.centers())
    ref = orc.lloyd(X, C_mb, 2)
    assert np.array_equal(e.centers(), ref["centers"])
    e.close()
    f.close()


@pytest.mark.parametrize("prec", ["fp32", "bf16"])
def test_engine_criterion_stops_like_reference(orc, prec):
    # the convergence test runs on the device: the iterations enqueued past
    # the converged one must leave its labels, centres and compactness as
    # they are, and the statistics must count iterations up to it
    X, _, _, _ = orc.mixture(12, 30_000, 8, 10, -10, 10, 0.7, 0.7)
    C0 = X[:10].copy()
    ref = orc.lloyd(X, C0, 50, criterion=1e-6)
    e = km.KMeans(X.shape[0], 8, 10, precision=prec, timing=True)
    e.set_data(X)
    e.set_centers(C0)
    e.run(50, 1e-6)
    st = e.stats()
    assert st["iterations"] == ref["iters"] < 50
    assert st["rel_err"] == ref["rel_err"]
    assert st["compactness"] == ref["comp_hist"][-1]
    assert np.array_equal(e.centers(), ref["centers"])
    assert np.array_equal(e.labels(), ref["labels"])
    # a second run from the converged state: its first iteration compares
    # against the converged compactness (the oracle's fresh run against 0,
    # so it always continues); from the second iteration on both compare
    # consecutive compactness values
    e.run(50, 1e-6)
    ref3 = orc.lloyd(X, ref["centers"], 50, criterion=1e-6)
    c1 = ref3["comp_hist"][0]
    first_stops = abs(c1 - ref["comp_hist"][-1]) / abs(c1) <= 1e-6
    n2 = 1 if first_stops else ref3["iters"]
    assert e.stats()["iterations"] == ref["iters"] + n2
    if not first_stops:
        assert np.array_equal(e.labels(), ref3["labels"])
        assert np.array_equal(e.centers(), ref3["centers"])
    e.close()


@pytest.mark.parametrize("n,d,k", [(100_000, 32, 64), (3001, 7, 5), (40_000, 128, 300)])
def test_graph_iterations_equal_eager(orc, n, d, k):
    # FP32 iterations over one canonical block with timing off run as graph
    # launches (one graph per reuse state and parity): every run length, a
    # device-side convergence run and new centres between runs give the
    # oracle's results, and the eagerly launched engine (timing on) the same
    print(f"graph iterations n={n} d={d} k={k}", flush=True)
    X, _, _, _ = orc.mixture(30 + d, n, d, k, -8, 8, 1.0, 1.5)
    C0 = X[:: n // k][:k].copy()
    g = km.KMeans(n, d, k)
    e = km.KMeans(n, d, k, timing=True)
    for eng in (g, e):
        eng.set_data(X)
        eng.set_centers(C0)
    done = 0
    for steps in (1, 2, 3, 1):
        for eng in (g, e):
            eng.run(steps)
        done += steps
        ref = orc.lloyd(X, C0, done)
        for eng in (g, e):
            assert np.array_equal(eng.labels(), ref["labels"]), done
            assert np.array_equal(eng.centers(), ref["centers"]), done
            assert eng.stats()["compactness"] == ref["comp_hist"][-1], done
    assert g.stats()["graph_launches"] == done
    assert e.stats()["graph_launches"] == 0
    # new centres, then a run with the device-side convergence test
    C1 = X[1 :: n // k][:k].copy()
    ref = orc.lloyd(X, C1, 60, criterion=1e-5)
    for eng in (g, e):
        eng.set_centers(C1)
        eng.run(60, 1e-5)
        st = eng.stats()
        assert st["iterations"] == done + ref["iters"]
        assert st["compactness"] == ref["comp_hist"][-1]
        assert st["rel_err"] == ref["rel_err"]
        assert np.array_equal(eng.labels(), ref["labels"])
        assert np.array_equal(eng.centers(), ref["centers"])
    assert g.stats()["graph_launches"] > done
    g.close()
    e.close()


def test_graph_iterations_after_steps_and_new_points(orc):
    # the step calls (assign, update) between graph runs and new points of
    # the same shape: the graph runs continue from the state they left
    n, d, k = 20_000, 16, 24
    X, _, _, _ = orc.mixture(41, n, d, k, -8, 8, 1.0, 1.0)
    X2, _, _, _ = orc.mixture(42, n, d, k, -8, 8, 1.0, 1.0)
    C0 = X[:k].copy()
    g = km.KMeans(n, d, k)
    g.set_data(X)
    g.set_centers(C0)
    g.run(2)
    assert km.lib().kmeansAssign(g._h) == 0
    assert km.lib().kmeansUpdateCenters(g._h) == 0
    g.run(2)
    ref = orc.lloyd(X, C0, 5)
    assert np.array_equal(g.labels(), ref["labels"])
    assert np.array_equal(g.centers(), ref["centers"])
    g.set_data(X2)
    g.set_centers(X2[:k].copy())
    g.run(3)
    ref = orc.lloyd(X2, X2[:k].copy(), 3)
    assert np.array_equal(g.labels(), ref["labels"])
    assert np.array_equal(g.centers(), ref["centers"])
    assert g.stats()["graph_launches"] == 7
    g.close()


@pytest.mark.parametrize("crit", [-1.0, 0.0])
def test_engine_counts_empty_clusters_per_iteration(orc, crit):
    # centre 4 attracts no point in any iteration: one empty cluster per
    # iteration, with a fixed count and with the device-side convergence test
    X, _, _, _ = orc.mixture(19, 8_000, 5, 4, -6, 6, 0.5, 0.5)
    C0 = np.concatenate([X[:4], np.full((1, 5), 1e6, np.float32)])
    e = km.KMeans(X.shape[0], 5, 5)
    e.set_data(X)
    e.set_centers(C0)
    e.run(9, crit)
    st = e.stats()
    ref = orc.lloyd(X, C0, 9, criterion=crit)
    assert st["iterations"] == ref["iters"]
    assert st["empty_clusters"] == st["iterations"]
    e.close()


def test_engine_device_data_borrowed(orc):
    X, _, _, _ = orc.mixture(13, 10_000, 12, 16, -5, 5, 1.0, 1.0)
    Xd = dev(X)
    torch.cuda.synchronize()
    e = km.KMeans(10_000, 12, 16)
    e.set_data(device_ptr=Xd.data_ptr())
    e.set_centers(X[:16])
    e.run(3)
    ref = orc.lloyd(X, X[:16], 3)
    assert np.array_equal(e.labels(), ref["labels"])
    assert np.array_equal(e.centers(), ref["centers"])


def test_compute_kmeans_f_with_given_init(orc):
    X, _, _, _ = orc.mixture(14, 8000, 6, 12, -8, 8, 1.0, 1.0)
    C0 = X[100:112].copy()
    st, res = km.compute_kmeans_f(X, 12, -1.0, 7, result=C0.T.copy(), init_from_result=True)
    ref = orc.lloyd(X, C0, 7)
    assert st == km.KMEANS_NONCONVERGENCE  # criterion < 0 can never be met
    assert np.array_equal(res, ref["centers"].T)


@pytest.mark.parametrize("alg", [0, 7])
def test_compute_kmeans_starts_from_the_reference_rows(orc, alg):
    # initAlgorithm other than 1, 2, 3: the rows the reference's own shuffle
    # and srand(42) pick select (the reference ignores initAlgorithm)
    X, _, _, _ = orc.mixture(18, 20_000, 6, 12, -8, 8, 0.8, 0.8)
    rows = km.reference_init_rows(X.shape[0], 12)
    st, res = km.compute_kmeans_f(X, 12, -1.0, 0, init_algorithm=alg)
    assert np.array_equal(res, X[rows].T)
    ref = orc.lloyd(X, X[rows].copy(), 5)
    st, res = km.compute_kmeans_f(X, 12, -1.0, 5, init_algorithm=alg)
    assert st == km.KMEANS_NONCONVERGENCE
    assert np.array_equal(res, ref["centers"].T)
    X64 = X.astype(np.float64)
    ref64 = orc.lloyd64(X64, X64[rows].copy(), 5)
    st, res64 = km.compute_kmeans_d(X64, 12, -1.0, 5, init_algorithm=alg)
    assert np.array_equal(res64, ref64["centers"].T)


def test_compute_kmeans_f_sampled_init_converges(orc):
    X, z, means, _ = orc.mixture(15, 20_000, 4, 5, -20, 20, 0.5, 0.5)
    st, res = km.compute_kmeans_f(X, 5, 1e-4, 100)
    assert st == km.KMEANS_SUCCESS
    assert res.shape == (4, 5)


def test_mini_batch_matches_oracle(orc):
    X, _, _, _ = orc.mixture(16, 10_000, 5, 8, -6, 6, 1.0, 1.0)
    C = X[:8].copy()
    e = km.KMeans(10_000, 5, 8)
    e.set_data(X)
    e.set_centers(C)
    e.set_mini_batch(3000)
    e.run(5)
    # batches: [0,3000) [3000,6000) [6000,9000) [9000,10000) [0,3000)
    for b0, b1 in [(0, 3000), (3000, 6000), (6000, 9000), (9000, 10000), (0, 3000)]:
        lab = orc.assign(X[b0:b1], C)
        C, _, _, _ = orc.update(X[b0:b1], lab, C)
    assert np.array_equal(e.centers(), C)


def test_legacy_kernel_names(orc):
    X, _, _, _ = orc.mixture(17, 6000, 20, 70, -5, 5, 1.0, 1.0)
    C = X[:70].copy()
    n, d = X.shape
    k = 70
    L = km.lib()
    Xd, B = dev(X), dev(C.T.copy())  # D x K as the reference passes it
    ncg = torch.empty(k, dtype=torch.float32, device="cuda")
    assert L.colNormalizeF(d, k, km._vp(B.data_ptr()), km._vp(ncg.data_ptr())) == 0
    # rowNormalizeF of two tiles into one output at offsets m0 (Normalize.cu:28)
    rn = torch.full((n,), -1.0, dtype=torch.float32, device="cuda")
    for m0, m in ((0, 3001), (3001, n - 3001)):
        assert L.rowNormalizeF(m0, m, d, km._vp(Xd.data_ptr() + 4 * m0 * d),
                               km._vp(rn.data_ptr())) == 0
    lab = torch.empty(n, dtype=torch.int32, device="cuda")
    for m0 in (0, 2500):  # two tiles
        m = 2500 if m0 == 0 else n - 2500
        assert L.ClosestCentersF(m0, m, d, km._vp(Xd.data_ptr() + 4 * m0 * d), k,
                                 km._vp(B.data_ptr()), None, None, km._vp(ncg.data_ptr()), 0,
                                 None, None, km._vp(lab.data_ptr())) == 0
    torch.cuda.synchronize()
    assert np.array_equal(ncg.cpu().numpy(), orc.col_norms(C))
    assert np.array_equal(rn.cpu().numpy(), orc.row_norms(X))
    ref = orc.assign(X, C)
    assert np.array_equal(lab.cpu().numpy(), ref)
    large =torch.zeros(4 * k * d, dtype=torch.float32, device="cuda")
    clarge = torch.zeros(4 * k, dtype=torch.int32, device="cuda")
    cen = torch.empty((k, d), dtype=torch.float32, device="cuda")
    cnt = torch.empty(k, dtype=torch.int32, device="cuda")
    assert L.ClusterCentersF(0, n, d, k, True, km._vp(Xd.data_ptr()), km._vp(lab.data_ptr()),
                             km._vp(large.data_ptr()), km._vp(clarge.data_ptr()),
                             km._vp(cen.data_ptr()), km._vp(cnt.data_ptr())) == 0
    rC, rcnt, _, _ = orc.update(X, ref, C)
    # empty clusters keep the value already in `centers`; none here
    assert np.array_equal(cnt.cpu().numpy(), rcnt)
    assert np.array_equal(cen.cpu().numpy()[rcnt > 0], rC[rcnt > 0])
    comp = torch.zeros(8, dtype=torch.float32, device="cuda")
    assert L.CompactnessF(0, n, True, d, k, km._vp(Xd.data_ptr()), km._vp(lab.data_ptr()),
                          km._vp(cen.data_ptr()), km._vp(comp.data_ptr())) == 0
    assert comp[0].item() == np.float32(orc.compactness(X, ref, cen.cpu().numpy()))


def test_generator_matches_cpu(orc):
    n, d, k = 5000, 16, 32
    X, z, means, sig = orc.mixture(1, n, d, k, -5, 5, 0.5, 2.0, i0=123)
    cdf = np.empty(k, np.float64)
    orc.lib().orc_gen_cdf(k, 0.0, cdf.ctypes.data)
    md, sd, cd = dev(means), dev(sig), dev(cdf)
    Xg = torch.empty((n, d), dtype=torch.float32, device="cuda")
    zg = torch.empty(n, dtype=torch.int32, device="cuda")
    km.generate_mixture(1, 123, n, d, k, md.data_ptr(), sd.data_ptr(), cd.data_ptr(),
                        Xg.data_ptr(), zg.data_ptr())
    torch.cuda.synchronize()
    assert np.array_equal(zg.cpu().numpy(), z)
    assert np.allclose(Xg.cpu().numpy(), X, rtol=1e-6, atol=1e-6)


@pytest.mark.parametrize("prec", ["fp32", "bf16", "fp16"])
def test_streamed_tiles_equal_resident(orc, prec):
    # points in host memory, two KM_BLOCK tiles plus a ragged one through HBM
    n, d, k = 2 * km.KM_BLOCK + 12345, 4, 16
    X, _, _, _ = orc.mixture(51, n, d, k, -20, 20, 1.0, 1.0)
    C0 = X[:k].copy()
    res = km.KMeans(n, d, k, precision=prec)
    res.set_data(X)
    res.set_centers(C0)
    res.run(3)
    st = km.KMeans(n, d, k, precision=prec, tile_rows=km.KM_BLOCK)
    st.set_data(X)  # pageable NumPy memory: staged through pinned buffers
    st.set_centers(C0)
    st.run(3)
    assert np.array_equal(st.labels(), res.labels())
    assert np.array_equal(st.centers(), res.centers())
    assert st.stats()["compactness"] == res.stats()["compactness"]
    # pinned host memory: direct asynchronous copies
    Xp = torch.from_numpy(X).pin_memory()
    st2 = km.KMeans(n, d, k, precision=prec, tile_rows=km.KM_BLOCK)
    st2.set_data(host_ptr=Xp.data_ptr())
    st2.set_centers(C0)
    st2.run(3)
    assert np.array_equal(st2.centers(), res.centers())
    if prec == "fp32":
        ref = orc.lloyd(X, C0, 3)
        assert np.array_equal(st.centers(), ref["centers"])
        assert st.stats()["compactness"] == ref["comp_hist"][-1]


def test_streamed_from_a_mapped_file(orc, tmp_path):
    # the reference's input format (headerless float32 rows) memory-mapped and
    # streamed through HBM tile by tile; equal to the resident run
    n, d, k = 2 * km.KM_BLOCK + 777, 3, 8
    X, _, _, _ = orc.mixture(52, n, d, k, -20, 20, 1.0, 1.0)
    path = tmp_path / "points.f32"
    X.tofile(path)
    M = km.load_raw_f32(str(path), d)
    st = km.KMeans(n, d, k, tile_rows=km.KM_BLOCK)
    st.set_data(M)
    st.set_centers(X[:k])
    st.run(2)
    ref = orc.lloyd(X, X[:k].copy(), 2)
    assert np.array_equal(st.centers(), ref["centers"])
    assert np.array_equal(st.labels(), ref["labels"])
    st.close()


def test_streamed_rejects_unaligned_tiles(orc):
    X = np.zeros((1000, 2), np.float32)
    e = km.KMeans(1000, 2, 4, tile_rows=500)
    with pytest.raises(km.KmeansError):
        e.set_data(X)
