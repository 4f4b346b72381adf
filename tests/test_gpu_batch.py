"""Batched independent clusterings against the per-instance oracle.

Every instance of a batch must follow exactly the single-problem canonical
Lloyd trajectory: labels, centres and per-iteration compactness bit-identical
to oracle.lloyd on that instance's points alone, whatever the work-item tile
size, and across replays of the captured graph.
"""
import numpy as np
import pytest

from oracle import oracle as orc

torch = pytest.importorskip("torch")
import kmeans_amd as km  # noqa: E402
from kmeans_amd import datagen  # noqa: E402

pytestmark = pytest.mark.gpu


def _instances(nb, d, k, median, lo, hi, seed=7):
    n, seeds = datagen.c4_instances(nb, 1.0, seed, median, lo, hi)
    Xs = []
    for i in range(nb):
        X, _, _, _ = orc.mixture(int(seeds[i]), int(n[i]), d, k, -4.0, 4.0, 0.7, 0.7)
        Xs.append(X)
    return n, Xs


def _oracle(Xs, k, iters):
    return [orc.lloyd(X, X[:k].copy(), iters) for X in Xs]


def _check(b, n, Xs, ref, iters_last):
    lab = b.labels()
    C = b.centers()
    comp = b.compactness()
    off = np.concatenate([[0], np.cumsum(n)])
    for i, r in enumerate(ref):
        assert np.array_equal(lab[off[i]:off[i + 1]], r["labels"]), i
        assert np.array_equal(C[i], r["centers"]), i
        assert np.array_equal(comp[:, i], r["comp_hist"][-iters_last:]), i


@pytest.mark.parametrize("precision,d,k,median,lo,hi", [
    ("fp32", 16, 32, 3000, 300, 20000), ("fp32", 5, 7, 400, 40, 3000), ("fp32", 20, 64, 1500, 200, 9000),
    ("bf16", 16, 32, 3000, 300, 20000), ("bf16", 5, 7, 400, 40, 3000), ("bf16", 20, 64, 1500, 200, 9000),
    ("bf16", 16, 256, 3000, 300, 20000), ("fp32", 8, 400, 3000, 1000, 9000)])
def test_batch_equals_per_instance_oracle(precision, d, k, median, lo, hi):
    nb, iters = 24, 6
    n, Xs = _instances(nb, d, k, median, lo, hi)
    ref = _oracle(Xs, k, iters)
    b = km.KMeansBatch(n, d, k, precision=precision)
    b.set_data(np.concatenate(Xs))
    b.init_first_k()
    b.run(iters)
    st = b.stats()
    print(f"\n{precision} d={d} k={k}: {st}")
    assert st["tile_rows"] in (256, 512, 1024, 2048, 4096)
    # per iteration: counter reset, (BF16: drift-bound classification,)
    # assign, sort, chunk sums, finalize, parity flip
    assert st["graph_nodes"] == (7 if precision == "bf16" else 6) * iters
    _check(b, n, Xs, ref, iters)
    b.close()


@pytest.mark.parametrize("precision", ["fp32", "bf16"])
def test_batch_graph_replays_and_tile_sizes(precision):
    """Three replays of one captured 3-iteration graph equal 9 eager oracle
    iterations; every explicit tile size gives the same result."""
    nb, d, k = 16, 16, 32
    n, Xs = _instances(nb, d, k, 2500, 300, 12000, seed=9)
    ref = _oracle(Xs, k, 9)
    for tile in (0, 256, 2048):
        b = km.KMeansBatch(n, d, k, precision=precision)
        b.set_data(np.concatenate(Xs))
        b.init_first_k()
        if tile:
            b.set_tile_rows(tile)
        for rep in range(3):
            b.run(3)
            torch.cuda.synchronize()  # host work between replays
        st = b.stats()
        assert st["iterations"] == 9 and st["last_run_iters"] == 3
        if tile:
            assert st["tile_rows"] == tile
        _check(b, n, Xs, ref, 3)
        b.close()


def test_batch_explicit_centres_and_device_data():
    nb, d, k = 8, 16, 16
    n, Xs = _instances(nb, d, k, 1000, 200, 4000, seed=11)
    rng = np.random.default_rng(0)
    C0 = rng.uniform(-4, 4, (nb, k, d)).astype(np.float32)
    ref = [orc.lloyd(X, C0[i], 4) for i, X in enumerate(Xs)]
    Xd = torch.from_numpy(np.concatenate(Xs)).cuda()
    Cd = torch.from_numpy(C0).cuda()
    b = km.KMeansBatch(n, d, k)
    b.set_data(device_ptr=Xd.data_ptr())
    b.set_centers(device_ptr=Cd.data_ptr())
    b.run(4)
    _check(b, n, Xs, ref, 4)
    b.close()


def test_batch_drift_bounds_keep_labels():
    """BF16 batches keep the points whose drift bounds separate their label
    (or decide them between two candidates) without scoring them; every run
    length matches the per-instance oracle, and centres set from outside
    afterwards drop the bounds."""
    nb, d, k = 24, 16, 64
    n, Xs = _instances(nb, d, k, 4000, 500, 20000, seed=13)
    b = km.KMeansBatch(n, d, k, precision="bf16")
    b.set_data(np.concatenate(Xs))
    b.init_first_k()
    b.set_tile_rows(1024)  # no tuning launches in the counts
    done = 0
    for steps in (2, 3, 5):
        b.run(steps)
        done += steps
        _check(b, n, Xs, _oracle(Xs, k, done), steps)
    st = b.stats()
    print(f"kept {st['kept']} paired {st['paired']} scored {st['scored']} of {done * int(n.sum())}")
    assert st["kept"] > 2 * int(n.sum())
    assert st["scored"] < (done - 1) * int(n.sum())
    C1 = b.centers()
    C1[:, 0, :] += 0.5  # from outside: the bounds no longer hold
    b.set_centers(C1)
    b.run(3)
    ref = [orc.lloyd(X, C1[i], 3) for i, X in enumerate(Xs)]
    _check(b, n, Xs, ref, 3)
    b.close()


def test_batch_drift_bounds_across_tile_sizes():
    """The open lists are per work item: changing the tile size between runs
    (more, smaller items: the per-item counts are reallocated) keeps every
    instance on the oracle's trajectory, with the bounds carried over."""
    nb, d, k = 16, 16, 48
    n, Xs = _instances(nb, d, k, 3000, 400, 12000, seed=17)
    b = km.KMeansBatch(n, d, k, precision="bf16")
    b.set_data(np.concatenate(Xs))
    b.init_first_k()
    done = 0
    for rows, steps in ((4096, 3), (256, 3), (1024, 2), (2, 1)):
        b.set_tile_rows(rows)
        b.run(steps)
        done += steps
        _check(b, n, Xs, _oracle(Xs, k, done), steps)
    assert b.stats()["kept"] > 0
    b.close()


@pytest.mark.parametrize("precision", ["fp32", "bf16"])
def test_batch_new_data_drops_reuse(precision):
    """The update reuses chunk partials of unchanged clusters across
    iterations and graph replays; new points of the same shape must start
    afresh (their sums and labels equal the oracle's on the new points)."""
    nb, d, k = 12, 16, 32
    n, Xs1 = _instances(nb, d, k, 2000, 300, 8000, seed=17)
    Xs2 = [orc.mixture(500 + i, int(n[i]), d, k, -4.0, 4.0, 0.7, 0.7)[0] for i in range(nb)]
    b = km.KMeansBatch(n, d, k, precision=precision)
    b.set_data(np.concatenate(Xs1))
    b.init_first_k()
    b.run(4)
    _check(b, n, Xs1, _oracle(Xs1, k, 4), 4)
    b.set_data(np.concatenate(Xs2))
    b.init_first_k()
    b.run(3)
    b.run(2)
    _check(b, n, Xs2, _oracle(Xs2, k, 5), 2)
    b.close()
