"""BF16/FP16 MFMA assignment: labels must equal the plain FP32 labels (and so
the oracle's) on every input, including adversarial near-ties; the certified
fraction must be high on clustered data (otherwise the speed-up is fake)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
import kmeans_amd as km  # noqa: E402


def run_engine(X, C0, iters, precision):
    e = km.KMeans(X.shape[0], X.shape[1], C0.shape[0], precision=precision)
    e.set_data(X)
    e.set_centers(C0)
    e.run(iters)
    out = (e.labels(), e.centers(), e.stats())
    e.close()
    return out


@pytest.mark.parametrize("prec", ["bf16", "fp16"])
@pytest.mark.parametrize("n,d,k", [(3000, 3, 1), (3000, 3, 2), (3000, 5, 3), (5000, 16, 64),
                                   (4000, 17, 65), (6000, 32, 100), (5000, 64, 256),
                                   (5000, 96, 300), (20000, 128, 1024), (3000, 200, 40)])
def test_mfma_labels_equal_oracle(orc, prec, n, d, k):
    print(f"mfma {prec} n={n} d={d} k={k}", flush=True)
    X, _, _, _ = orc.mixture(200 + d + k, n, d, max(k, 2), -5, 5, 0.5, 2.0)
    C0 = X[np.arange(k) * (n // k)].copy()
    lab, C, st = run_engine(X, C0, 3, prec)
    ref = orc.lloyd(X, C0, 3)
    assert np.array_equal(lab, ref["labels"])
    assert np.array_equal(C, ref["centers"])
    assert st["compactness"] == ref["comp_hist"][-1]


@pytest.mark.parametrize("prec", ["bf16", "fp16"])
def test_mfma_certifies_most_points(orc, prec):
    X, z, means, _ = orc.mixture(31, 50_000, 128, 256, -5, 5, 0.5, 2.0)
    C0 = X[:256].copy()
    lab, C, st = run_engine(X, C0, 4, prec)
    ref = orc.lloyd(X, C0, 4)
    assert np.array_equal(lab, ref["labels"])
    frac = st["reresolved"] / (4 * 50_000)
    print(f"{prec}: re-resolved fraction {frac:.4f}")
    assert frac < 0.2


@pytest.mark.parametrize("prec", ["bf16", "fp16"])
def test_mfma_adversarial_near_ties(orc, prec):
    rng = np.random.default_rng(77)
    d, k = 48, 40
    base = (rng.standard_normal((k // 2, d)) * 4).astype(np.float32)
    # pairs of centres closer than the BF16 resolution, and exact duplicates
    C0 = np.concatenate([base, base + (rng.standard_normal((k // 2, d)) * 1e-3).astype(np.float32)])
    C0[5] = C0[25]
    C0[7] = C0[30]
    # points on (or next to) bisectors of the close pairs
    mids = 0.5 * (C0[: k // 2] + C0[k // 2:])
    X = np.concatenate([
        np.repeat(mids, 100, axis=0) + (rng.standard_normal((100 * (k // 2), d)) * 1e-4).astype(np.float32),
        (rng.standard_normal((3000, d)) * 4).astype(np.float32),
    ]).astype(np.float32)
    nc = orc.col_norms(C0)
    ref = orc.assign(X, C0, nc)
    e = km.KMeans(X.shape[0], d, k, precision=prec)
    e.set_data(X)
    e.set_centers(C0)
    e.run(1)
    # after one iteration labels are those of the first assignment
    assert np.array_equal(e.labels(), ref)
    st = e.stats()
    assert st["reresolved"] > 0  # the near-ties must have gone to the FP32 path
    e.close()


@pytest.mark.parametrize("prec", ["bf16", "fp16"])
def test_mfma_fp32_search_split_over_centres_ties(orc, prec):
    # every centre four times (indices j, j + 100, j + 200, j + 300): no
    # reduced-precision certificate holds, so every point goes to the FP32
    # search, whose centre slices each see a different copy; the label must
    # be the lowest index of the best score, as in the single pass
    print(f"split ties {prec}", flush=True)
    rng = np.random.default_rng(11)
    d, kb = 24, 100
    base = rng.standard_normal((kb, d)).astype(np.float32) * 3
    C0 = np.concatenate([base] * 4)
    X = (base[rng.integers(0, kb, 5000)] + rng.standard_normal((5000, d)) * 0.5).astype(np.float32)
    ref = orc.assign(X, C0)
    assert ref.max() < kb
    e = km.KMeans(X.shape[0], d, C0.shape[0], precision=prec)
    e.set_data(X)
    e.set_centers(C0)
    e.run(1)
    assert np.array_equal(e.labels(), ref)
    st = e.stats()
    assert st["reresolved"] == X.shape[0]
    e.close()


@pytest.mark.parametrize("prec", ["fp32", "bf16", "fp16"])
def test_split_search_signed_zero_tie(orc, prec):
    # x = (-1, -0, -0, -0): the zero centre 0 scores -0 (every product is -0)
    # and the six centres at distance |x| in the next centre slices score +0.
    # -0 and +0 are equal scores, so the label is the lowest index, 0.  The
    # reduced-precision certificates cannot separate seven equal scores, so
    # these points reach the FP32 search split over 64-centre slices.
    print(f"signed-zero tie {prec}", flush=True)
    rng = np.random.default_rng(5)
    d, k = 4, 256
    C0 = rng.uniform(20, 60, (k, d)).astype(np.float32)
    C0[0] = 0.0
    ties = {64: (-2, 0, 0, 0), 65: (-1, 1, 0, 0), 128: (-1, -1, 0, 0), 129: (-1, 0, 1, 0),
            192: (-1, 0, -1, 0), 193: (-1, 0, 0, 1)}
    for j, c in ties.items():
        C0[j] = c
    x = np.array([-1.0, -0.0, -0.0, -0.0], np.float32)
    far = rng.integers(1, 64, 448)
    X = np.concatenate([np.tile(x, (64, 1)),
                        C0[far] + rng.standard_normal((448, d)).astype(np.float32) * 0.1])
    X = X.astype(np.float32)
    ref, best, _ = orc.assign(X, C0, with_scores=True)
    assert np.all(ref[:64] == 0) and np.all(best[:64] == 0.0)
    assert np.signbit(best[0])  # the zero centre's chain really is -0
    e = km.KMeans(X.shape[0], d, k, precision=prec)
    e.set_data(X)
    e.set_centers(C0)
    e.run(1)
    lab = e.labels()
    st = e.stats()
    e.close()
    assert np.array_equal(lab, ref)
    if prec != "fp32":
        assert st["reresolved"] >= 64  # the tie points took the split FP32 search


@pytest.mark.parametrize("prec", ["bf16", "fp16"])
def test_mfma_deep_pass_twins_in_one_tile_row(orc, prec):
    # centres j and j + 128 are near-twins in the same row of two different
    # 32-centre tiles, so their reduced-precision scores (split BF16 in the
    # deep pass) agree to within the packed low bits; the deep pass must
    # either tell them apart by their exact FP32 chains or send the point on
    print(f"deep twins {prec}", flush=True)
    rng = np.random.default_rng(23)
    d, kb = 64, 128
    base = (rng.standard_normal((kb, d)) * 3).astype(np.float32)
    twin = base + (rng.standard_normal((kb, d)) * 2e-5).astype(np.float32)
    C0 = np.concatenate([base, twin]).astype(np.float32)
    X = (base[rng.integers(0, kb, 20000)] + rng.standard_normal((20000, d)) * 0.4).astype(np.float32)
    ref = orc.assign(X, C0)
    e = km.KMeans(X.shape[0], d, C0.shape[0], precision=prec)
    e.set_data(X)
    e.set_centers(C0)
    e.run(1)
    assert np.array_equal(e.labels(), ref)
    assert e.stats()["deep"] > 0
    e.close()


def test_mfma_large_magnitudes_fp16_scaling(orc):
    # values far outside FP16 range must be scaled, not overflow
    X, _, _, _ = orc.mixture(41, 4000, 24, 20, -5, 5, 1.0, 1.0)
    X = X * np.float32(1e6)
    C0 = X[:20].copy()
    lab, C, st = run_engine(X, C0, 2, "fp16")
    ref = orc.lloyd(X, C0, 2)
    assert np.array_equal(lab, ref["labels"]) and np.array_equal(C, ref["centers"])


def test_mfma_tiny_magnitudes(orc):
    X, _, _, _ = orc.mixture(42, 4000, 24, 20, -5, 5, 1.0, 1.0)
    X = X * np.float32(1e-15)
    C0 = X[:20].copy()
    for prec in ("bf16", "fp16"):
        lab, C, st = run_engine(X, C0, 2, prec)
        ref = orc.lloyd(X, C0, 2)
        assert np.array_equal(lab, ref["labels"]) and np.array_equal(C, ref["centers"])


@pytest.mark.parametrize("prec", ["bf16", "fp16"])
def test_mfma_more_elements_than_a_dispatch_dimension(orc, prec):
    """68M x 64 points: the reduced-precision copy has more than 2^32 elements,
    more than one dispatch's 32-bit grid size can address; the rows past that
    mark must be converted and labelled like the first ones."""
    from kmeans_amd import datagen
    n = 68_000_000
    X, C0, spec = datagen.generate_on_gpu("c3", n=n, torch=torch)
    d, k = spec["d"], spec["k"]
    assert n * 64 > 2 ** 32
    labs = {}
    for p in ("fp32", prec):
        e = km.KMeans(n, d, k, precision=p)
        e.set_data(device_ptr=X.data_ptr())
        e.set_centers(C0)
        e.run(1)
        labs[p] = e.labels()
        e.close()
    assert np.array_equal(labs["fp32"], labs[prec])
    rows = np.sort(np.random.default_rng(1).choice(np.arange(2 ** 32 // 64, n), 2000, replace=False))
    Xs = X[torch.from_numpy(rows).cuda()].cpu().numpy()
    assert np.array_equal(labs[prec][rows], orc.assign(Xs, C0))
    del X
    torch.cuda.empty_cache()


@pytest.mark.parametrize("prec", ["bf16", "fp16"])
def test_mfma_misaligned_borrowed_rows(orc, prec):
    # a borrowed device pointer 4 bytes past a 16-byte boundary: the kernels
    # that read FP32 rows as 16-byte pieces must fall back to scalar reads
    print(f"misaligned rows {prec}", flush=True)
    X, _, _, _ = orc.mixture(61, 6000, 64, 40, -5, 5, 0.5, 2.0)
    C0 = X[:40].copy()
    buf = torch.empty(X.size + 1, dtype=torch.float32, device="cuda")
    buf[1:] = torch.from_numpy(X.reshape(-1)).cuda()
    Xd = buf[1:]
    assert Xd.data_ptr() % 16 == 4
    torch.cuda.synchronize()
    e = km.KMeans(X.shape[0], 64, 40, precision=prec)
    e.set_data(device_ptr=Xd.data_ptr())
    e.set_centers(C0)
    e.run(3)
    ref = orc.lloyd(X, C0, 3)
    assert np.array_equal(e.labels(), ref["labels"])
    assert np.array_equal(e.centers(), ref["centers"])
    e.close()


def test_batch_rejects_misaligned_borrowed_rows():
    n = np.array([1000, 1200])
    X = torch.zeros(int(n.sum()) * 16 + 1, dtype=torch.float32, device="cuda")
    b = km.KMeansBatch(n, 16, 8)
    with pytest.raises(km.KmeansError) as err:
        b.set_data(device_ptr=X[1:].data_ptr())
    assert err.value.code == km.KMEANS_ERROR_INVALID_ARGUMENT
    b.close()


@pytest.mark.parametrize("prec", ["bf16", "fp16"])
@pytest.mark.parametrize("sig_lo,sig_hi,init", [(0.2, 0.3, "means"), (0.3, 2.0, "means"),
                                                (0.3, 2.0, "sample")])
def test_mfma_ordered_assign_tile_cut(orc, prec, sig_lo, sig_hi, init):
    # tight, well-separated clusters in label order: workgroups stop after a
    # few centre tiles once the triangle bound clears every remaining tile;
    # with a spread of widths some workgroups stop and others do not.  Labels
    # and centres must stay those of the oracle on every iteration.
    print(f"tile cut {prec} sigma [{sig_lo}, {sig_hi}] init {init}", flush=True)
    n, d, k = 400_000, 64, 256
    X, _, means, _ = orc.mixture(71, n, d, k, -5, 5, sig_lo, sig_hi)
    C0 = means.copy() if init == "means" else X[orc.sample_init(71, n, k)].copy()
    e = km.KMeans(n, d, k, precision=prec)
    e.set_data(X)
    e.set_centers(C0)
    e.run(1)  # the first assign has no label order yet
    s1 = e.stats()
    for _ in range(4):
        # setting the (same) centres again drops the drift bounds, so every
        # point is scored in the label order and the cut decides the tiles
        e.set_centers(e.centers())
        e.run(1)
    lab, C, st = e.labels(), e.centers(), e.stats()
    e.close()
    ref = orc.lloyd(X, C0, 5)
    assert np.array_equal(lab, ref["labels"])
    assert np.array_equal(C, ref["centers"])
    assert st["compactness"] == ref["comp_hist"][-1]
    t, tt = st["tiles"] - s1["tiles"], st["tiles_total"] - s1["tiles_total"]
    print(f"ordered iterations: tiles computed {t} of {tt}")
    assert s1["tiles"] == s1["tiles_total"] > 0 and 0 < t <= tt
    if sig_hi <= 0.3:  # most workgroups hold one tight cluster: the cut must happen
        assert t < 0.75 * tt


@pytest.mark.parametrize("prec", ["bf16", "fp16"])
def test_mfma_drift_bounds_keep_labels(orc, prec):
    # late iterations: centres barely move, so most points keep their labels
    # on the drift bounds alone; every run length must match the oracle
    print(f"drift bounds {prec}", flush=True)
    n, d, k = 300_000, 48, 200
    X, _, means, _ = orc.mixture(91, n, d, k, -5, 5, 0.4, 1.2)
    C0 = X[orc.sample_init(91, n, k)].copy()
    e = km.KMeans(n, d, k, precision=prec)
    e.set_data(X)
    e.set_centers(C0)
    done = 0
    for steps in (1, 3, 4, 4):
        e.run(steps)
        done += steps
        ref = orc.lloyd(X, C0, done)
        assert np.array_equal(e.labels(), ref["labels"]), done
        assert np.array_equal(e.centers(), ref["centers"]), done
    st = e.stats()
    print(f"kept by bounds {st['bounded']}, decided between two {st['paired']} of {done * n}")
    assert st["bounded"] > 2 * n
    assert st["paired"] > 0
    # centres set from outside: the bounds must not be used for them
    C1 = means.copy()
    e.set_centers(C1)
    e.run(2)
    ref = orc.lloyd(X, C1, 2)
    assert np.array_equal(e.labels(), ref["labels"])
    assert np.array_equal(e.centers(), ref["centers"])
    e.close()


def test_mfma_drift_bounds_across_a_converged_run(orc):
    # a device-side convergence stop skips the later assigns and their bound
    # updates; a run after it must still match the oracle's continuation
    n, d, k = 200_000, 32, 100
    X, _, _, _ = orc.mixture(93, n, d, k, -5, 5, 0.5, 1.0)
    C0 = X[orc.sample_init(93, n, k)].copy()
    e = km.KMeans(n, d, k, precision="bf16")
    e.set_data(X)
    e.set_centers(C0)
    e.run(40, criterion=1e-4)
    it = e.stats()["iterations"]
    ref = orc.lloyd(X, C0, 40, criterion=1e-4)
    assert it == ref["iters"]
    assert np.array_equal(e.labels(), ref["labels"])
    C_conv = e.centers()
    assert np.array_equal(C_conv, ref["centers"])
    e.run(3)
    ref2 = orc.lloyd(X, C_conv, 3)
    assert np.array_equal(e.labels(), ref2["labels"])
    assert np.array_equal(e.centers(), ref2["centers"])
    e.close()


def pair_fixed_point_data(m=1 << 10, d=16, seed=7):
    """Two mirrored pairs of clusters at a Lloyd fixed point.  In each pair,
    cluster a: 3m points at x0 = -1, m at the origin (exact ties between a
    and b, which go to a, the lower index) and m at x0 = -2^-12; cluster b:
    4m points at x0 = +3/4 and m at x0 = +2^-12 (near-ties).  The other
    coordinates are +-pairs (multiples of 1/64) within every group, so their
    sums are exactly zero, and both means sit at x0 = -+3072.25 / 5120.  The
    second pair is the first shifted by 8 in x1 (small enough that the FP32
    chains still resolve the near-ties)."""
    rng = np.random.default_rng(seed)
    eps = 2.0 ** -12

    def group(cnt, x0):
        h = rng.integers(1, 64, size=(cnt // 2, d - 1)).astype(np.float32) / 64.0
        G = np.zeros((cnt, d), np.float32)
        G[:, 0] = x0
        G[:, 1:] = np.concatenate([h, -h])
        return G

    def pair(base):
        P = np.concatenate([group(3 * m, -1.0), group(m, 0.0), group(m, -eps),
                            group(4 * m, 0.75), group(m, eps)])
        P[:, 1] += base
        return P

    X = np.concatenate([pair(0.0), pair(8.0)])
    X = np.ascontiguousarray(X[rng.permutation(X.shape[0])])
    c = np.float32(3072.25 / 5120.0)
    C0 = np.zeros((4, d), np.float32)
    C0[:, 0] = [-c, c, -c, c]
    C0[2:, 1] = 8.0
    return X, C0


@pytest.mark.parametrize("prec", ["bf16", "fp16"])
def test_mfma_pair_decisions_at_a_fixed_point(orc, prec):
    # The origin points are exact ties between a and b and the points at
    # x0 = +-2^-12 near-ties; every other centre is >= 8 away.  The run starts
    # with both centres of each pair shifted by 1/4 in x1, which keeps every
    # tie and every label, so the first update moves them onto the fixed
    # point: the second assign decides these points by the two FP32 chains in
    # the drift-bound filter (the ties to the lower index), and later ones,
    # with the centres unmoved, keep them.
    X, C0 = pair_fixed_point_data()
    n = X.shape[0]
    Cs = C0.copy()
    Cs[:, 1] += 0.25
    ref = orc.lloyd(X, Cs, 6)
    assert np.array_equal(ref["centers"], C0)  # one update reaches the fixed point
    assert np.array_equal(orc.lloyd(X, Cs, 1)["centers"], C0)
    e = km.KMeans(n, X.shape[1], 4, precision=prec)
    e.set_data(X)
    e.set_centers(Cs)
    done = 0
    for steps in (1, 1, 2, 2):
        e.run(steps)
        done += steps
        assert np.array_equal(e.labels(), ref["labels"]), done
        assert np.array_equal(e.centers(), C0), done
    st = e.stats()
    print(f"{prec}: paired {st['paired']} bounded {st['bounded']} of {done * n}")
    assert st["paired"] > 0
    assert st["bounded"] > 0
    e.close()
