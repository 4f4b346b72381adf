"""The host-side tables of kmeans_amd.datagen equal the oracle's generator
(same counter-based streams), so GPU-generated workloads match the CPU ones."""
import numpy as np

from kmeans_amd import datagen


def test_tables_match_oracle(orc):
    for seed, d, k, z in [(1, 16, 32, 0.0), (2, 8, 100, 1.1)]:
        X, zz, means, sig = orc.mixture(seed, 10, d, k, -5, 5, 0.5, 2.0, zipf_s=z)
        m2, s2, cdf2 = datagen.mixture_tables(seed, d, k, -5, 5, 0.5, 2.0, z)
        assert np.array_equal(means, m2) and np.array_equal(sig, s2)
        cdf = np.empty(k, np.float64)
        orc.lib().orc_gen_cdf(k, z, cdf.ctypes.data)
        assert np.array_equal(cdf, cdf2)


def test_sample_init_matches_oracle(orc):
    assert np.array_equal(datagen.sample_init(1, 10_000_000, 1024), orc.sample_init(1, 10_000_000, 1024))
    assert np.array_equal(datagen.sample_init(3, 5000, 5000), orc.sample_init(3, 5000, 5000))


def test_rng_matches_oracle(orc):
    for i in (0, 1, 12345, 2**40 + 7):
        assert int(datagen.rng_u64(9, 3, i)) == orc.lib().orc_rng_u64(9, 3, i)
