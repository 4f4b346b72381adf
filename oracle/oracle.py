"""ctypes wrapper over oracle/libkmoracle.so -- TEST INFRASTRUCTURE ONLY.

The CPU oracle restates the reference's Lloyd hot path (see the header of
kmeans_oracle.c for the file:line of each function it follows).  Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg import this module;
the product package kmeans_amd never does.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "libkmoracle.so")
KM_CHUNK = 256
KM_BLOCK = 1 << 23

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", _HERE, "-B"], check=True)


_march = "x86-64-v3"


def use_native():
    """Build and load the oracle with -march=native for this host (the CPU
    baseline's build; the portable x86-64-v3 build stays for the tests).
    Returns False, keeping the portable build, if compiling fails or the
    library is already loaded."""
    global _LIB, _march
    if _lib is not None:
        return False
    native = os.path.join(_HERE, "libkmoracle_native.so")
    r = subprocess.run(["make", "-s", "-C", _HERE, "-B", "libkmoracle_native.so"],
                       capture_output=True)
    if r.returncode != 0 or not os.path.exists(native):
        return False
    _LIB, _march = native, "native"
    return True


def march():
    return _march


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = ctypes.CDLL(_LIB)
        P = ctypes.c_void_p
        i64, i32, u64, f32, f64 = ctypes.c_long, ctypes.c_int, ctypes.c_uint64, ctypes.c_float, ctypes.c_double
        sig = {
            "orc_rng_u64": (u64, [u64, u64, u64]),
            "orc_rng_unif": (f64, [u64, u64, u64]),
            "orc_gen_means": (None, [u64, i32, i32, f32, f32, f32, f32, P, P]),
            "orc_gen_cdf": (None, [i32, f64, P]),
            "orc_gen_points": (None, [u64, i64, i64, i32, i32, P, P, P, P, P]),
            "orc_sample_init": (i32, [u64, i64, i32, P]),
            "orc_col_norms": (None, [P, i32, i32, P]),
            "orc_row_norms": (None, [P, i64, i32, P]),
            "orc_assign": (None, [P, i64, i32, P, i32, P, P, P, P]),
            "orc_score": (f32, [P, P, i32, f32]),
            "orc_csum": (f64, [P, i64]),
            "orc_update": (i32, [P, i64, i32, P, i32, P, P, P, P, P, P]),
            "orc_compactness": (f64, [P, i64, i32, P, P, i32]),
            "orc_moments": (None, [P, i64, i32, P, i32, P, P, P]),
            "orc_lloyd": (i32, [P, i64, i32, i32, P, i32, f64, P, P, P, P]),
            "orc_means_fp64": (None, [P, i64, i32, P, i32, P, P]),
            "orc_col_norms_f64": (None, [P, i32, i32, P]),
            "orc_row_norms_f64": (None, [P, i64, i32, P]),
            "orc_score_f64": (f64, [P, P, i32, f64]),
            "orc_assign_f64": (None, [P, i64, i32, P, i32, P, P]),
            "orc_moments_f64": (None, [P, i64, i32, P, i32, P, P, P]),
            "orc_update_f64": (i32, [P, i64, i32, P, i32, P, P, P, P]),
            "orc_compactness_f64": (f64, [P, i64, i32, P, P, i32]),
            "orc_lloyd_f64": (i32, [P, i64, i32, i32, P, i32, f64, P, P, P, P]),
            "orc_row_transform_min_f": (None, [i64, i64, i32, P, P, P, P]),
            "orc_row_transform_min_d": (None, [i64, i64, i32, P, P, P, P]),
            "orc_kmeanspp": (i32, [P, i64, i32, i32, u64, P]),
            "orc_kmeanspp_f64": (i32, [P, i64, i32, i32, u64, P]),
            "orc_compactness_fp64": (f64, [P, i64, i32, P, P]),
            "orc_mlp_fwd_bwd": (f64, [i32, i32, i32, i32, P, P, P, P, i32]),
            "orc_num_threads": (i32, []),
            "orc_set_threads": (None, [i32]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


# ---------------------------------------------------------------- generators
def mixture(seed, n, d, k, mean_lo, mean_hi, sig_lo, sig_hi, zipf_s=0.0, i0=0, ncomp=None):
    """Rows [i0, i0+n) of a Gaussian mixture (counter-based, order-free).

    Returns (X n x d float32, z component ids, means, sigmas)."""
    ncomp = k if ncomp is None else ncomp
    means = np.empty((ncomp, d), np.float32)
    sig = np.empty(ncomp, np.float32)
    lib().orc_gen_means(seed, d, ncomp, mean_lo, mean_hi, sig_lo, sig_hi, _p(means), _p(sig))
    cdf = np.empty(ncomp, np.float64)
    lib().orc_gen_cdf(ncomp, float(zipf_s), _p(cdf))
    X = np.empty((n, d), np.float32)
    z = np.empty(n, np.int32)
    lib().orc_gen_points(seed, i0, n, d, ncomp, _p(means), _p(sig), _p(cdf), _p(X), _p(z))
    return X, z, means, sig


def sample_init(seed, n, k):
    idx = np.empty(k, np.int32)
    if lib().orc_sample_init(seed, n, k, _p(idx)) != 0:
        raise ValueError("k > n")
    return idx


# ---------------------------------------------------------------- hot path
def col_norms(C):
    C = _f32(C)
    nc = np.empty(C.shape[0], np.float32)
    lib().orc_col_norms(_p(C), C.shape[0], C.shape[1], _p(nc))
    return nc


def assign(X, C, nc=None, with_scores=False):
    X, C = _f32(X), _f32(C)
    n, d = X.shape
    k = C.shape[0]
    nc = col_norms(C) if nc is None else _f32(nc)
    lab = np.empty(n, np.int32)
    best = np.empty(n, np.float32) if with_scores else None
    second = np.empty(n, np.float32) if with_scores else None
    lib().orc_assign(_p(X), n, d, _p(C), k, _p(nc), _p(lab), _p(best), _p(second))
    return (lab, best, second) if with_scores else lab


def score(x, c, ncj):
    x, c = _f32(x), _f32(c)
    return lib().orc_score(_p(x), _p(c), x.shape[0], float(ncj))


def csum(v):
    v = np.ascontiguousarray(v, dtype=np.float64)
    return lib().orc_csum(_p(v), v.shape[0])


def update(X, labels, C_old, with_comp=False):
    """Canonical update.  Returns (C_new, counts, S, empty[, comp vs C_new])."""
    X, C_old, labels = _f32(X), _f32(C_old), _i32(labels)
    n, d = X.shape
    k = C_old.shape[0]
    C_new = np.empty_like(C_old)
    counts = np.empty(k, np.int64)
    S = np.empty((k, d), np.float64)
    Q = np.empty(k, np.float64)
    comp = ctypes.c_double(0)
    empty = lib().orc_update(_p(X), n, d, _p(labels), k, _p(C_old), _p(C_new), _p(counts), _p(S),
                             _p(Q), ctypes.byref(comp))
    if with_comp:
        return C_new, counts, S, empty, comp.value
    return C_new, counts, S, empty


def moments(X, labels, k):
    X, labels = _f32(X), _i32(labels)
    n, d = X.shape
    S = np.empty((k, d), np.float64)
    Q = np.empty(k, np.float64)
    cnt = np.empty(k, np.int64)
    lib().orc_moments(_p(X), n, d, _p(labels), k, _p(S), _p(Q), _p(cnt))
    return S, Q, cnt


def compactness(X, labels, C):
    X, C, labels = _f32(X), _f32(C), _i32(labels)
    n, d = X.shape
    return lib().orc_compactness(_p(X), n, d, _p(labels), _p(C), C.shape[0])


def lloyd(X, C0, max_iters, criterion=-1.0):
    """Full canonical Lloyd loop.  Returns dict(labels, centers, comp_hist, iters, rel_err, status)."""
    X = _f32(X)
    C = _f32(C0).copy()
    n, d = X.shape
    k = C.shape[0]
    lab = np.empty(n, np.int32)
    hist = np.zeros(max(max_iters, 1), np.float64)
    it = ctypes.c_int(0)
    rel = ctypes.c_double(0)
    st = lib().orc_lloyd(_p(X), n, d, k, _p(C), max_iters, float(criterion), _p(lab), _p(hist),
                         ctypes.byref(it), ctypes.byref(rel))
    return dict(labels=lab, centers=C, comp_hist=hist[: it.value], iters=it.value,
                rel_err=rel.value, status=st)


def means_fp64(X, labels, k):
    X, labels = _f32(X), _i32(labels)
    n, d = X.shape
    M = np.empty((k, d), np.float64)
    cnt = np.empty(k, np.int64)
    lib().orc_means_fp64(_p(X), n, d, _p(labels), k, _p(M), _p(cnt))
    return M, cnt


def compactness_fp64(X, labels, C):
    X, C, labels = _f32(X), _f32(C), _i32(labels)
    return lib().orc_compactness_fp64(_p(X), X.shape[0], X.shape[1], _p(labels), _p(C))


# ---------------------------------------------------------------- double precision
def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def row_norms(X):
    """rowNormalizeF: the fma chain of each row's squares (float32)."""
    X = _f32(X)
    out = np.empty(X.shape[0], np.float32)
    lib().orc_row_norms(_p(X), X.shape[0], X.shape[1], _p(out))
    return out


def col_norms64(C):
    C = _f64(C)
    nc = np.empty(C.shape[0], np.float64)
    lib().orc_col_norms_f64(_p(C), C.shape[0], C.shape[1], _p(nc))
    return nc


def row_norms64(X):
    X = _f64(X)
    out = np.empty(X.shape[0], np.float64)
    lib().orc_row_norms_f64(_p(X), X.shape[0], X.shape[1], _p(out))
    return out


def score64(x, c, ncj):
    x, c = _f64(x), _f64(c)
    return lib().orc_score_f64(_p(x), _p(c), x.shape[0], float(ncj))


def assign64(X, C, nc=None):
    X, C = _f64(X), _f64(C)
    nc = col_norms64(C) if nc is None else _f64(nc)
    lab = np.empty(X.shape[0], np.int32)
    lib().orc_assign_f64(_p(X), X.shape[0], X.shape[1], _p(C), C.shape[0], _p(nc), _p(lab))
    return lab


def moments64(X, labels, k):
    X, labels = _f64(X), _i32(labels)
    n, d = X.shape
    S = np.empty((k, d), np.float64)
    Q = np.empty(k, np.float64)
    cnt = np.empty(k, np.int64)
    lib().orc_moments_f64(_p(X), n, d, _p(labels), k, _p(S), _p(Q), _p(cnt))
    return S, Q, cnt


def update64(X, labels, C_old):
    """Double-precision update.  Returns (C_new, counts, empty, comp vs C_new)."""
    X, C_old, labels = _f64(X), _f64(C_old), _i32(labels)
    n, d = X.shape
    k = C_old.shape[0]
    C_new = np.empty_like(C_old)
    counts = np.empty(k, np.int64)
    comp = ctypes.c_double(0)
    empty = lib().orc_update_f64(_p(X), n, d, _p(labels), k, _p(C_old), _p(C_new), _p(counts),
                                 ctypes.byref(comp))
    return C_new, counts, empty, comp.value


def compactness64(X, labels, C):
    X, C, labels = _f64(X), _f64(C), _i32(labels)
    return lib().orc_compactness_f64(_p(X), X.shape[0], X.shape[1], _p(labels), _p(C), C.shape[0])


def lloyd64(X, C0, max_iters, criterion=-1.0):
    """Double-precision Lloyd loop; same result dict as lloyd()."""
    X = _f64(X)
    C = _f64(C0).copy()
    n, d = X.shape
    k = C.shape[0]
    lab = np.empty(n, np.int32)
    hist = np.zeros(max(max_iters, 1), np.float64)
    it = ctypes.c_int(0)
    rel = ctypes.c_double(0)
    st = lib().orc_lloyd_f64(_p(X), n, d, k, _p(C), max_iters, float(criterion), _p(lab), _p(hist),
                             ctypes.byref(it), ctypes.byref(rel))
    return dict(labels=lab, centers=C, comp_hist=hist[: it.value], iters=it.value,
                rel_err=rel.value, status=st)


def row_transform_min(P, rn, cn, m0=0, out=None):
    """rowTransformMinimum{F,D} on a host m x n product (dtype picks the variant)."""
    f = P.dtype == np.float32
    cv = _f32 if f else _f64
    P, rn, cn = cv(P), cv(rn), cv(cn)
    m, n = P.shape
    if out is None:
        out = np.full(m0 + m, -7, np.int32)
    fn = lib().orc_row_transform_min_f if f else lib().orc_row_transform_min_d
    fn(m0, m, n, _p(rn), _p(cn), _p(P), _p(out))
    return out


def kmeanspp(X, k, seed):
    """k-means++ seeding: the k chosen row indices in pick order (dtype picks the variant)."""
    f64 = np.asarray(X).dtype == np.float64
    X = _f64(X) if f64 else _f32(X)
    idx = np.empty(k, np.int32)
    fn = lib().orc_kmeanspp_f64 if f64 else lib().orc_kmeanspp
    assert fn(_p(X), X.shape[0], X.shape[1], k, seed, _p(idx)) == 0
    return idx


def num_threads():
    return lib().orc_num_threads()


def set_threads(t):
    lib().orc_set_threads(int(t))


# ---------------------------------------------------------------- workloads
def config1():
    """N=100,000, D=32, K=64 mixture: means U[-10,10], sigma 1, seed 0; C0 = first K points."""
    X, z, means, sig = mixture(0, 100_000, 32, 64, -10.0, 10.0, 1.0, 1.0)
    return X, X[:64].copy()


def mlp_fwd_bwd(din, h, g, X, t, P, mirror=False):
    """FP64 MLP loss and gradients (flat parameter layout of the GPU kernels)."""
    X = _f32(X)
    t = _i32(t)
    P = np.ascontiguousarray(P, dtype=np.float64)
    G = np.empty_like(P)
    loss = lib().orc_mlp_fwd_bwd(din, h, g, X.shape[0], _p(X), _p(t), _p(P), _p(G), int(mirror))
    return loss, G


def bf16_round(a):
    """Round float32 values to the nearest BF16 (ties to even), kept as float32."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    return u.view(np.float32)
