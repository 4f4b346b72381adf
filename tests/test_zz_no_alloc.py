"""kmUpdateNoAlloc (the update with its workspace in the library's fixed
per-device scratch array) against kmUpdate with a caller workspace, through
the C-ABI: centres, transposed copy, norms, counts and the empty count must
be identical bit for bit, on two calls in a row (the second smaller, so the
scratch holds the first call's contents), and both must equal the oracle.
Shapes beyond the scratch are rejected before any device work."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
import kmeans_amd as km  # noqa: E402


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def outputs(n, d, k):
    ld = km.centers_ld(k)
    return dict(C=torch.empty((k, d), dtype=torch.float32, device="cuda"),
                Ct=torch.full((d, ld), 7.0, dtype=torch.float32, device="cuda"),
                nc=torch.empty(k, dtype=torch.float32, device="cuda"),
                cnt=torch.empty(k, dtype=torch.int64, device="cuda"),
                emp=torch.full((1,), 99, dtype=torch.int32, device="cuda"))


def host(o):
    torch.cuda.synchronize()
    return {key: v.cpu().numpy() for key, v in o.items()}


def run_both(orc, n, d, k, seed, empty_centre):
    X, _, _, _ = orc.mixture(seed, n, d, k, -6, 6, 0.5, 1.5)
    C0 = X[:k].copy()
    lab = orc.assign(X, C0)
    if empty_centre:  # no member for centre 3: it keeps its value
        lab[lab == 3] = 4
    Xd, ld, Cd = dev(X), dev(lab), dev(C0)
    a, b = outputs(n, d, k), outputs(n, d, k)
    wsb = km.update_workspace_size(n, d, k)
    assert wsb <= km.KM_NOALLOC_SCRATCH_BYTES
    ws = torch.empty(wsb, dtype=torch.uint8, device="cuda")
    km.update(n, d, k, Xd.data_ptr(), ld.data_ptr(), Cd.data_ptr(), a["C"].data_ptr(),
              ws.data_ptr(), wsb, Ct_ptr=a["Ct"].data_ptr(), nc_ptr=a["nc"].data_ptr(),
              counts_ptr=a["cnt"].data_ptr(), empty_ptr=a["emp"].data_ptr())
    km.update_no_alloc(n, d, k, Xd.data_ptr(), ld.data_ptr(), Cd.data_ptr(), b["C"].data_ptr(),
                       Ct_ptr=b["Ct"].data_ptr(), nc_ptr=b["nc"].data_ptr(),
                       counts_ptr=b["cnt"].data_ptr(), empty_ptr=b["emp"].data_ptr())
    ha, hb = host(a), host(b)
    for key in ha:
        assert np.array_equal(ha[key], hb[key]), key
    rC, rcnt, _, rempty = orc.update(X, lab, C0)
    assert np.array_equal(hb["C"], rC)
    assert np.array_equal(hb["cnt"], rcnt)
    assert int(hb["emp"][0]) == rempty == (1 if empty_centre else 0)
    assert np.array_equal(hb["nc"], orc.col_norms(rC))
    assert np.array_equal(hb["Ct"][:, :k], rC.T)


def test_no_alloc_update_equals_workspace_update_twice(orc):
    print("no-alloc update: n=300000 d=64 k=200, then n=40000 d=16 k=37", flush=True)
    run_both(orc, 300_000, 64, 200, 41, empty_centre=False)
    run_both(orc, 40_000, 16, 37, 42, empty_centre=True)


def test_no_alloc_update_rejects_shapes_beyond_its_scratch():
    n, d, k = 4_000_000, 128, 1024
    assert km.update_workspace_size(n, d, k) > km.KM_NOALLOC_SCRATCH_BYTES
    dummy = torch.empty(16, dtype=torch.float32, device="cuda")
    p = dummy.data_ptr()
    with pytest.raises(km.KmeansError) as err:
        km.update_no_alloc(n, d, k, p, p, p, p)
    assert err.value.code == 8  # ERROR_INVALID_ARGUMENT
