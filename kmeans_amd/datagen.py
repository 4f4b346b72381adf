"""Synthetic Gaussian-mixture workloads, generated on the GPU.

The points come from kmGenerateMixture (a counter-based kernel: any row range
can be made independently).  The small per-component tables (means, sigmas,
cumulative weights) and the initial-centre indices are computed here on the
host with the same counter-based generator (SplitMix64 finaliser of (seed,
stream, index); Steele, Lea, Flood, OOPSLA 2014).

Workloads (sizes of the K-means configurations this engine is measured on):
  c1  N=100,000     D=32  K=64    means U[-10,10], sigma 1,          seed 0, C0 = first K points
  c2  N=10,000,000  D=128 K=1024  means U[-5,5],   sigma U[0.5,2.0], seed 1, C0 = K sampled points
  c3  N=200,000,000 D=64  K=4096  means U[-8,8],   sigma 1, Zipf(1.1) weights, seed 2
  c5  N=20,000,000  D=96  K=16384 means U[-6,6],   sigma U[0.3,1.5], seed 3
"""
import math

import numpy as np

M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def _mix64(z):
    with np.errstate(over="ignore"):
        z = np.asarray(z, dtype=np.uint64)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def rng_u64(seed, stream, i):
    with np.errstate(over="ignore"):
        k = _mix64(np.uint64(seed) * np.uint64(0x9E3779B97F4A7C15)
                   + np.uint64(stream) * np.uint64(0xD1B54A32D192ED03)
                   + np.uint64(0x632BE59BD9B4E019))
        return _mix64(k + np.asarray(i, dtype=np.uint64) * np.uint64(0x9E3779B97F4A7C15))


def rng_unif(seed, stream, i):
    u = rng_u64(seed, stream, i)
    return ((u >> np.uint64(11)).astype(np.float64) + 0.5) * (1.0 / 9007199254740992.0)


def mixture_tables(seed, d, ncomp, mean_lo, mean_hi, sig_lo, sig_hi, zipf_s=0.0):
    means = (mean_lo + (mean_hi - mean_lo) * rng_unif(seed, 0, np.arange(ncomp * d))).astype(np.float32)
    sig = (sig_lo + (sig_hi - sig_lo) * rng_unif(seed, 1, np.arange(ncomp))).astype(np.float32)
    # math.pow is the C library pow, the same routine the oracle's table uses
    w = [math.pow(j + 1, -zipf_s) for j in range(ncomp)] if zipf_s > 0 else [1.0] * ncomp
    tot = 0.0
    for v in w:  # same left-to-right order as the oracle's table
        tot += v
    cdf = np.empty(ncomp)
    acc = 0.0
    for j, v in enumerate(w):
        acc += v / tot
        cdf[j] = acc
    cdf[-1] = 1.0
    return means.reshape(ncomp, d), sig, cdf


def sample_init(seed, n, k):
    """k distinct indices in [0, n) by Floyd's algorithm (stream 4)."""
    idx = []
    seen = set()
    r_all = rng_u64(seed, 4, np.arange(k)).tolist()
    for t, j in enumerate(range(n - k, n)):
        r = int(r_all[t] % (j + 1))
        v = j if r in seen else r
        seen.add(v)
        idx.append(v)
    return np.array(idx, np.int32)


WORKLOADS = {
    "c1": dict(n=100_000, d=32, k=64, seed=0, mlo=-10.0, mhi=10.0, slo=1.0, shi=1.0, zipf=0.0, init="first"),
    "c2": dict(n=10_000_000, d=128, k=1024, seed=1, mlo=-5.0, mhi=5.0, slo=0.5, shi=2.0, zipf=0.0, init="sample"),
    "c3": dict(n=200_000_000, d=64, k=4096, seed=2, mlo=-8.0, mhi=8.0, slo=1.0, shi=1.0, zipf=1.1, init="sample"),
    "c5": dict(n=20_000_000, d=96, k=16384, seed=3, mlo=-6.0, mhi=6.0, slo=0.3, shi=1.5, zipf=0.0, init="sample"),
}


def generate_on_gpu(name, n=None, torch=None, i0=0):
    """Returns (X device tensor n x d, C0 host k x d, spec)."""
    import kmeans_amd as km
    spec = dict(WORKLOADS[name])
    if n is not None:
        spec["n"] = n
    n, d, k, seed = spec["n"], spec["d"], spec["k"], spec["seed"]
    means, sig, cdf = mixture_tables(seed, d, k, spec["mlo"], spec["mhi"], spec["slo"], spec["shi"],
                                     spec["zipf"])
    dm = torch.from_numpy(means).cuda()
    ds = torch.from_numpy(sig).cuda()
    dc = torch.from_numpy(cdf).cuda()
    X = torch.empty((n, d), dtype=torch.float32, device="cuda")
    km.generate_mixture(seed, i0, n, d, k, dm.data_ptr(), ds.data_ptr(), dc.data_ptr(), X.data_ptr())
    idx = np.arange(k, dtype=np.int32) if spec["init"] == "first" else sample_init(seed, n, k)
    di = torch.from_numpy(idx).cuda()
    C0 = torch.empty((k, d), dtype=torch.float32, device="cuda")
    km.gather_rows(k, d, X.data_ptr(), di.data_ptr(), C0.data_ptr())
    torch.cuda.synchronize()
    spec["init_idx"] = idx
    return X, C0.cpu().numpy(), spec


def c4_instances(nb=2048, sigma=1.0, seed=4, median=20000, lo=1000, hi=200000):
    """Batched workload sizes and seeds: N_i = clip(round(median * exp(sigma * Z_i)), lo, hi)
    with Z_i standard normal (Box-Muller on the counter-based generator, stream 10);
    instance seed = generator output (stream 11) masked to 31 bits."""
    i = np.arange(nb, dtype=np.uint64)
    u1 = rng_unif(seed, 10, 2 * i)
    u2 = rng_unif(seed, 10, 2 * i + 1)
    z = np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)
    n = np.clip(np.rint(median * np.exp(sigma * z)), lo, hi).astype(np.int64)
    seeds = (rng_u64(seed, 11, i) & np.uint64(0x7FFFFFFF)).astype(np.int64)
    return n, seeds


C4 = dict(nb=2048, d=16, k=256, mlo=-4.0, mhi=4.0, sig=0.7, iters=15)


def generate_c4_on_gpu(torch, nb=None, sigma=1.0):
    """All instances of the batched workload concatenated on the GPU: each
    instance is its own Gaussian mixture (k means U[-4,4]^16, sigma 0.7, equal
    weights) from its own seed.  Returns (X device tensor, n per instance, seeds)."""
    import kmeans_amd as km
    nb = C4["nb"] if nb is None else nb
    d, k = C4["d"], C4["k"]
    n, seeds = c4_instances(nb, sigma)
    X = torch.empty((int(n.sum()), d), dtype=torch.float32, device="cuda")
    off = 0
    for i in range(nb):
        means, sig, cdf = mixture_tables(int(seeds[i]), d, k, C4["mlo"], C4["mhi"], C4["sig"], C4["sig"])
        dm = torch.from_numpy(means).cuda()
        ds = torch.from_numpy(sig).cuda()
        dc = torch.from_numpy(cdf).cuda()
        km.generate_mixture(int(seeds[i]), 0, int(n[i]), d, k, dm.data_ptr(), ds.data_ptr(),
                            dc.data_ptr(), X[off:].data_ptr())
        off += int(n[i])
    torch.cuda.synchronize()
    return X, n, seeds
