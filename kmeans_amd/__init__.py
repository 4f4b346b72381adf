"""kmeans_amd -- Python mirror of the MI355X K-means engine's C-ABI.

The product is the shared library libkmeans_amd.so (hand-written HIP for
gfx950 behind the extern "C" boundary in include/kmeans_amd.h).  This module
binds it with ctypes so Python callers, the tests and bench.py reach the same
entry points a C or C++ caller links against:

  compute_kmeans_f(...)        computeKmeansF (reference main/cpp/kmeans.hpp:102-107)
  compute_kmeans_d(...)        computeKmeansD (kmeans.hpp:125-130), double precision
  load_raw_f32(path, n)        the reference's headerless float32 input files, memory-mapped
  KMeans(m, n, k, ...)         the engine handle (kmeans<T> of kmeans.hpp:174-658)
  col_norms / assign_f32 /     kernel API, device pointers (KmeansCudaKernels.h:
  update / compactness          238-474 replaced with stream + workspace arguments)

There is no CPU fallback: if the library is missing or no GPU is present the
calls raise.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libkmeans_amd.so")

KMEANS_SUCCESS = 0
KMEANS_NONCONVERGENCE = 13
KMEANS_ERROR_INVALID_ARGUMENT = 14
PREC = {"fp32": 0, "bf16": 1, "fp16": 2}
PROPOSER = {None: 0, "none": 0, "mlp": 1, "centroid": 2}
KM_CHUNK = 256
KM_BLOCK = 1 << 23
KM_NOALLOC_SCRATCH_BYTES = 32 << 20  # kmUpdateNoAlloc's per-device scratch array

ERRORS = {
    1: "KMEANS_ERROR_BUILDDATA", 2: "KMEANS_ERROR_INITIALIZE", 3: "KMEANS_ERROR_CLOSEST_CENTERS",
    4: "KMEANS_ERROR_CLUSTER_CENTERS", 5: "KMEANS_ERROR_COMPACTNESS", 6: "KMEANS_ERROR_TRANSPOSE",
    7: "KMEANS_ERROR_RESET", 8: "KMEANS_ERROR_COMPUTE", 9: "KMEANS_ERROR_GET_CENTERS",
    10: "KMEANS_ERROR_TILING", 11: "KMEANS_ERROR_COPY_TILE", 12: "KMEANS_ERROR_CONSTRUCT_MINIBATCH",
    13: "KMEANS_NONCONVERGENCE", 14: "KMEANS_ERROR_INVALID_ARGUMENT",
}


class KmeansError(RuntimeError):
    def __init__(self, code, where=""):
        self.code = code
        super().__init__(f"{where}: {ERRORS.get(code, code)}")


class Options(ctypes.Structure):
    _fields_ = [("precision", ctypes.c_int), ("device", ctypes.c_int), ("stream", ctypes.c_void_p),
                ("tile_rows", ctypes.c_long), ("timing", ctypes.c_int), ("proposer", ctypes.c_int),
                ("groups", ctypes.c_int), ("top_groups", ctypes.c_int), ("train_iters", ctypes.c_int),
                ("hidden", ctypes.c_int), ("full_gemm", ctypes.c_int)]


class Stats(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int), ("compactness", ctypes.c_double),
                ("rel_err", ctypes.c_double), ("empty_clusters", ctypes.c_longlong),
                ("reresolved", ctypes.c_longlong), ("rescored", ctypes.c_longlong),
                ("ms_assign", ctypes.c_double),
                ("ms_reresolve", ctypes.c_double), ("ms_update", ctypes.c_double),
                ("ms_compactness", ctypes.c_double), ("ms_total", ctypes.c_double),
                ("proposed", ctypes.c_longlong), ("certified", ctypes.c_longlong),
                ("certified_group", ctypes.c_longlong), ("fallback", ctypes.c_longlong),
                ("candidate_pairs", ctypes.c_longlong), ("proposal_hits", ctypes.c_longlong),
                ("ms_train", ctypes.c_double), ("train_loss_first", ctypes.c_double),
                ("train_loss_last", ctypes.c_double), ("deep", ctypes.c_longlong),
                ("tiles", ctypes.c_longlong), ("tiles_total", ctypes.c_longlong),
                ("bounded", ctypes.c_longlong), ("paired", ctypes.c_longlong),
                ("scored", ctypes.c_longlong), ("assign_launches", ctypes.c_longlong),
                ("ms_assign_kernel", ctypes.c_double), ("graph_launches", ctypes.c_longlong)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


# Every symbol declared in include/kmeans_amd.h with its ctypes signature.
_P, _I, _L, _F, _D, _U64, _SZ = (ctypes.c_void_p, ctypes.c_int, ctypes.c_long, ctypes.c_float,
                                 ctypes.c_double, ctypes.c_uint64, ctypes.c_size_t)
SIGNATURES = {
    "kmeansGetErrorString": (ctypes.c_char_p, [_I]),
    "computeKmeansF": (_I, [_P, _I, _I, _I, _F, _I, _I, _I, _I, _I, _P, _F]),
    "computeKmeansD": (_I, [_P, _I, _I, _I, _D, _I, _I, _I, _I, _I, _P, _D]),
    "kmeansCreate": (_I, [_P, _L, _I, _I, _P]),
    "kmeansDestroy": (None, [_P]),
    "kmeansSetData": (_I, [_P, _P, _I]),
    "kmeansSetCenters": (_I, [_P, _P]),
    "kmeansInitPlusPlus": (_I, [_P, _U64]),
    "kmeansInitialize": (_I, [_P, _I, _U64, _P]),
    "kmeansAssign": (_I, [_P]),
    "kmeansUpdateCenters": (_I, [_P]),
    "kmeansCompactness": (_I, [_P, _P]),
    "kmeansRun": (_I, [_P, _I, _D]),
    "kmeansGetCenters": (_I, [_P, _P]),
    "kmeansGetLabels": (_I, [_P, _P]),
    "kmeansGetStats": (_I, [_P, _P]),
    "kmeansSetMiniBatch": (_I, [_P, _L]),
    "kmeansReferenceInitRows": (_I, [_L, _I, _P]),
    "kmeansDeviceCenters": (_P, [_P]),
    "kmeansDeviceLabels": (_P, [_P]),
    "kmeansSynchronize": (_I, [_P]),
    "kmeansCreateD": (_I, [_P, _L, _I, _I, _P]),
    "kmeansDestroyD": (None, [_P]),
    "kmeansSetDataD": (_I, [_P, _P, _I]),
    "kmeansSetCentersD": (_I, [_P, _P]),
    "kmeansInitializeD": (_I, [_P, _I, _U64, _P]),
    "kmeansAssignD": (_I, [_P]),
    "kmeansUpdateCentersD": (_I, [_P]),
    "kmeansCompactnessD": (_I, [_P, _P]),
    "kmeansRunD": (_I, [_P, _I, _D]),
    "kmeansSetMiniBatchD": (_I, [_P, _L]),
    "kmeansGetCentersD": (_I, [_P, _P]),
    "kmeansGetLabelsD": (_I, [_P, _P]),
    "kmeansGetStatsD": (_I, [_P, _P]),
    "kmeansBatchCreate": (_I, [_P, _I, _P, _I, _I, _P]),
    "kmeansBatchDestroy": (None, [_P]),
    "kmeansBatchSetData": (_I, [_P, _P, _I]),
    "kmeansBatchSetCenters": (_I, [_P, _P, _I]),
    "kmeansBatchInitFirstK": (_I, [_P]),
    "kmeansBatchSetTileRows": (_I, [_P, _I]),
    "kmeansBatchRun": (_I, [_P, _I]),
    "kmeansBatchGetCenters": (_I, [_P, _P]),
    "kmeansBatchGetLabels": (_I, [_P, _P]),
    "kmeansBatchGetCompactness": (_I, [_P, _P]),
    "kmeansBatchGetStats": (_I, [_P, _P]),
    "kmeansBatchDeviceCenters": (_P, [_P]),
    "kmeansBatchDeviceLabels": (_P, [_P]),
    "kmColNorms": (_I, [_P, _I, _I, _P, _P]),
    "kmCentersLd": (_I, [_I]),
    "kmTransposeCenters": (_I, [_P, _I, _I, _P, _P]),
    "kmAssignF32": (_I, [_P, _L, _P, _I, _P, _P, _I, _P, _P, _P, _P]),
    "kmUpdateWorkspaceSize": (_SZ, [_L, _I, _I]),
    "kmUpdate": (_I, [_P, _L, _I, _I, _P, _P, _P, _P, _P, _P, _P, _P, _P, _SZ]),
    "kmUpdateNoAlloc": (_I, [_P, _L, _I, _I, _P, _P, _P, _P, _P, _P, _P, _P]),
    "kmCompactness": (_I, [_P, _L, _I, _I, _P, _P, _P, _P, _P, _SZ]),
    "kmRowNorms": (_I, [_P, _L, _L, _I, _P, _P]),
    "kmGenerateMixture": (_I, [_P, _U64, _L, _L, _I, _I, _P, _P, _P, _P, _P]),
    "kmGatherRows": (_I, [_P, _L, _I, _P, _P, _P]),
    "kmMlpParamCount": (_SZ, [_I, _I, _I]),
    "kmMlpWorkspaceSize": (_SZ, [_I, _I, _I, _I]),
    "kmMlpForwardBackward": (_I, [_P, _I, _I, _I, _I, _P, _P, _P, _P, _P, _P, _SZ]),
    "ClosestCentersF": (_I, [_I, _I, _I, _P, _I, _P, _P, _P, _P, _I, _P, _P, _P]),
    "ClosestCentersD": (_I, [_I, _I, _I, _P, _I, _P, _P, _P, _P, _I, _P, _P, _P]),
    "rowTransformMinimumF": (_I, [_I, _I, _I, _P, _P, _P, _P]),
    "rowTransformMinimumD": (_I, [_I, _I, _I, _P, _P, _P, _P]),
    "ClusterCentersF": (_I, [_I, _I, _I, _I, ctypes.c_bool, _P, _P, _P, _P, _P, _P]),
    "ClusterCentersD": (_I, [_I, _I, _I, _I, ctypes.c_bool, _P, _P, _P, _P, _P, _P]),
    "CompactnessF": (_I, [_I, _I, ctypes.c_bool, _I, _I, _P, _P, _P, _P]),
    "CompactnessD": (_I, [_I, _I, ctypes.c_bool, _I, _I, _P, _P, _P, _P]),
    "rowNormalizeF": (_I, [_I, _I, _I, _P, _P]),
    "rowNormalizeD": (_I, [_I, _I, _I, _P, _P]),
    "colNormalizeF": (_I, [_I, _I, _P, _P]),
    "colNormalizeD": (_I, [_I, _I, _P, _P]),
}

_lib = None


def lib():
    """Load libkmeans_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: run __graft_entry__.build() first")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _np_ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _check(code, where):
    if code != 0:
        raise KmeansError(code, where)


# ------------------------------------------------------------------ data files
def load_raw_f32(path, n, m=None):
    """Map a headerless row-major float32 file of m x n points, the reference's
    input format (test/cpp/KmeansTests.cpp:84-87 reads m*n floats from it),
    without reading it into memory.  m=None takes every whole row in the file.
    The returned array can go to KMeans.set_data directly: resident engines
    copy it once, engines with tile_rows stream it through pinned staging
    buffers tile by tile."""
    size = os.path.getsize(path)
    if size % 4:
        raise ValueError(f"{path}: {size} bytes is not a whole number of float32 values")
    rows = size // (4 * n)
    if m is None:
        m = rows
    if m <= 0 or m > rows:
        raise ValueError(f"{path}: holds {size // 4} floats, {m} x {n} requested")
    return np.memmap(path, dtype=np.float32, mode="r", shape=(m, n))


def load_npy(path):
    """Map a NumPy .npy file of m x n float32 points (C order) without
    reading it into memory, like load_raw_f32; other dtypes, ranks or Fortran
    order are refused rather than converted (a conversion would read and copy
    the whole file).  Object arrays are never unpickled."""
    a = np.load(path, mmap_mode="r", allow_pickle=False)
    if a.dtype != np.float32 or a.ndim != 2:
        raise ValueError(f"{path}: {a.dtype} array of rank {a.ndim}, need 2-D float32")
    if not a.flags.c_contiguous:
        raise ValueError(f"{path}: Fortran-ordered array, need row-major points")
    return a


# ------------------------------------------------------------------ host API
def _init_alg(init_algorithm, init_from_result):
    if init_algorithm is not None:
        return int(init_algorithm)
    return 1 if init_from_result else 0


def reference_init_rows(m, k):
    """The k rows the reference's initialize picks as initial centres
    (kmeansReferenceInitRows), as indices into the unshuffled rows."""
    out = np.empty(k, np.int32)
    _check(lib().kmeansReferenceInitRows(int(m), int(k), _np_ptr(out)), "kmeansReferenceInitRows")
    return out


def compute_kmeans_f(data, k, criterion, max_iters, result=None, init_from_result=False,
                     mini_batch_fraction=1.0, full_gemm=False, init_algorithm=None):
    """computeKmeansF: returns (status, centers as D x K).  full_gemm: doFullSGEMM = 1.
    init_algorithm (default: 1 with init_from_result, else 0): 0 the reference's
    own initial rows, 1 the centres in result, 2 k-means++, 3 seeded sampling."""
    data = np.ascontiguousarray(data, dtype=np.float32)
    m, n = data.shape
    res = np.zeros((n, k), np.float32) if result is None else np.ascontiguousarray(result, np.float32)
    st = lib().computeKmeansF(_np_ptr(data), m, n, k, criterion, max_iters, 0,
                              _init_alg(init_algorithm, init_from_result), 1 if full_gemm else 0, 0,
                              _np_ptr(res),
                              mini_batch_fraction)
    return st, res


def compute_kmeans_d(data, k, criterion, max_iters, result=None, init_from_result=False,
                     mini_batch_fraction=1.0, full_gemm=False, init_algorithm=None):
    """computeKmeansD: returns (status, centers as D x K, float64).  full_gemm: doFullSGEMM = 1.
    init_algorithm as for compute_kmeans_f."""
    data = np.ascontiguousarray(data, dtype=np.float64)
    m, n = data.shape
    res = np.zeros((n, k), np.float64) if result is None else np.ascontiguousarray(result, np.float64)
    st = lib().computeKmeansD(_np_ptr(data), m, n, k, criterion, max_iters, 0,
                              _init_alg(init_algorithm, init_from_result), 1 if full_gemm else 0, 0,
                              _np_ptr(res),
                              mini_batch_fraction)
    return st, res


class KMeans:
    """The engine handle: kmeansCreate / SetData / SetCenters / Run / Get*."""

    def __init__(self, m, n, k, precision="fp32", device=0, stream=None, timing=False, tile_rows=0,
                 proposer=None, groups=0, top_groups=0, train_iters=0, hidden=0, full_gemm=False):
        """tile_rows > 0 (a multiple of KM_BLOCK) with host data: stream the points
        through HBM in tiles of that many rows instead of keeping them resident.
        proposer "mlp" | "centroid": after train_iters plain iterations, score each
        point only against the members of top_groups proposed centre groups and
        certify the label (uncertified points are searched in full).
        full_gemm: assign through a library GEMM (rocBLAS SGEMM) and a row argmin,
        the reference's doFullSGEMM = 1 path (fp32 only)."""
        self.m, self.n, self.k = int(m), int(n), int(k)
        opt = Options(PREC[precision], device, stream, int(tile_rows), 1 if timing else 0,
                      PROPOSER[proposer], int(groups), int(top_groups), int(train_iters), int(hidden),
                      1 if full_gemm else 0)
        h = ctypes.c_void_p()
        _check(lib().kmeansCreate(ctypes.byref(h), self.m, self.n, self.k, ctypes.byref(opt)),
               "kmeansCreate")
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().kmeansDestroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def set_data(self, data=None, device_ptr=None, host_ptr=None):
        """data: NumPy array (copied to the GPU, or streamed when tile_rows is set);
        device_ptr: device address (borrowed); host_ptr: host address (borrowed,
        e.g. a pinned buffer, used in streamed mode)."""
        if device_ptr is not None:
            _check(lib().kmeansSetData(self._h, ctypes.c_void_p(device_ptr), 1), "kmeansSetData")
        elif host_ptr is not None:
            _check(lib().kmeansSetData(self._h, ctypes.c_void_p(host_ptr), 0), "kmeansSetData")
        else:
            self._host = np.ascontiguousarray(data, dtype=np.float32)
            assert self._host.shape == (self.m, self.n)
            _check(lib().kmeansSetData(self._h, _np_ptr(self._host), 0), "kmeansSetData")

    def set_centers(self, C):
        C = np.ascontiguousarray(C, dtype=np.float32)
        assert C.shape == (self.k, self.n)
        _check(lib().kmeansSetCenters(self._h, _np_ptr(C)), "kmeansSetCenters")

    def init_plus_plus(self, seed=42):
        """k-means++ seeding of the centres on the GPU (kmeansInitPlusPlus)."""
        _check(lib().kmeansInitPlusPlus(self._h, int(seed)), "kmeansInitPlusPlus")

    def set_mini_batch(self, rows):
        _check(lib().kmeansSetMiniBatch(self._h, int(rows)), "kmeansSetMiniBatch")

    def run(self, max_iters, criterion=-1.0):
        _check(lib().kmeansRun(self._h, int(max_iters), float(criterion)), "kmeansRun")

    def centers(self):
        out = np.empty((self.k, self.n), np.float32)
        _check(lib().kmeansGetCenters(self._h, _np_ptr(out)), "kmeansGetCenters")
        return out

    def labels(self):
        out = np.empty(self.m, np.int32)
        _check(lib().kmeansGetLabels(self._h, _np_ptr(out)), "kmeansGetLabels")
        return out

    def stats(self):
        s = Stats()
        _check(lib().kmeansGetStats(self._h, ctypes.byref(s)), "kmeansGetStats")
        return s.as_dict()

    def synchronize(self):
        _check(lib().kmeansSynchronize(self._h), "kmeansSynchronize")

    def device_labels_ptr(self):
        return lib().kmeansDeviceLabels(self._h)

    def device_centers_ptr(self):
        return lib().kmeansDeviceCenters(self._h)


class BatchStats(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int), ("last_run_iters", ctypes.c_int),
                ("tile_rows", ctypes.c_int), ("work_items", ctypes.c_int),
                ("graph_nodes", ctypes.c_int), ("tuned_from_cache", ctypes.c_int),
                ("empty_clusters", ctypes.c_longlong), ("ms_graph", ctypes.c_double),
                ("ms_capture", ctypes.c_double), ("tune_ms", ctypes.c_float * 5),
                ("exact_searched", ctypes.c_longlong), ("kept", ctypes.c_longlong),
                ("paired", ctypes.c_longlong), ("scored", ctypes.c_longlong),
                ("ms_assign", ctypes.c_double), ("ms_update", ctypes.c_double)]

    def as_dict(self):
        d = {f: getattr(self, f) for f, _ in self._fields_}
        d["tune_ms"] = list(self.tune_ms)
        return d


class KMeansBatch:
    """Many independent clusterings of uneven size (kmeansBatch*): instance i
    owns n[i] consecutive rows of the data and k centres."""

    def __init__(self, n, d, k, precision="fp32", device=0, stream=None, timing=False):
        self.n = np.ascontiguousarray(n, dtype=np.int64)
        self.nb, self.d, self.k = len(self.n), int(d), int(k)
        self.total = int(self.n.sum())
        opt = Options(PREC[precision], device, stream, 0, 1 if timing else 0)
        h = ctypes.c_void_p()
        _check(lib().kmeansBatchCreate(ctypes.byref(h), self.nb, _np_ptr(self.n), self.d, self.k,
                                       ctypes.byref(opt)), "kmeansBatchCreate")
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().kmeansBatchDestroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def set_data(self, X=None, device_ptr=None):
        if device_ptr is not None:
            _check(lib().kmeansBatchSetData(self._h, ctypes.c_void_p(device_ptr), 1), "kmeansBatchSetData")
        else:
            X = np.ascontiguousarray(X, dtype=np.float32)
            assert X.shape == (self.total, self.d)
            _check(lib().kmeansBatchSetData(self._h, _np_ptr(X), 0), "kmeansBatchSetData")

    def set_centers(self, C=None, device_ptr=None):
        if device_ptr is not None:
            _check(lib().kmeansBatchSetCenters(self._h, ctypes.c_void_p(device_ptr), 1),
                   "kmeansBatchSetCenters")
        else:
            C = np.ascontiguousarray(C, dtype=np.float32)
            assert C.shape == (self.nb, self.k, self.d)
            _check(lib().kmeansBatchSetCenters(self._h, _np_ptr(C), 0), "kmeansBatchSetCenters")

    def init_first_k(self):
        _check(lib().kmeansBatchInitFirstK(self._h), "kmeansBatchInitFirstK")

    def set_tile_rows(self, rows):
        _check(lib().kmeansBatchSetTileRows(self._h, int(rows)), "kmeansBatchSetTileRows")

    def run(self, iters):
        _check(lib().kmeansBatchRun(self._h, int(iters)), "kmeansBatchRun")

    def centers(self):
        out = np.empty((self.nb, self.k, self.d), np.float32)
        _check(lib().kmeansBatchGetCenters(self._h, _np_ptr(out)), "kmeansBatchGetCenters")
        return out

    def labels(self):
        out = np.empty(self.total, np.int32)
        _check(lib().kmeansBatchGetLabels(self._h, _np_ptr(out)), "kmeansBatchGetLabels")
        return out

    def compactness(self):
        it = self.stats()["last_run_iters"]
        out = np.empty((it, self.nb), np.float64)
        _check(lib().kmeansBatchGetCompactness(self._h, _np_ptr(out)), "kmeansBatchGetCompactness")
        return out

    def stats(self):
        s = BatchStats()
        _check(lib().kmeansBatchGetStats(self._h, ctypes.byref(s)), "kmeansBatchGetStats")
        return s.as_dict()


# ------------------------------------------------------------------ kernel API
def _vp(x):
    return None if x is None else ctypes.c_void_p(int(x))


def col_norms(k, d, C_ptr, nc_ptr, stream=None):
    _check(lib().kmColNorms(_vp(stream), k, d, _vp(C_ptr), _vp(nc_ptr)), "kmColNorms")


def centers_ld(k):
    return lib().kmCentersLd(k)


def transpose_centers(k, d, C_ptr, Ct_ptr, stream=None):
    _check(lib().kmTransposeCenters(_vp(stream), k, d, _vp(C_ptr), _vp(Ct_ptr)), "kmTransposeCenters")


def assign_f32(m, d, X_ptr, k, Ct_ptr, nc_ptr, labels_ptr, best_ptr=None, rows_ptr=None,
               m_dev_ptr=None, stream=None):
    _check(lib().kmAssignF32(_vp(stream), m, _vp(m_dev_ptr), d, _vp(X_ptr), _vp(rows_ptr), k,
                             _vp(Ct_ptr), _vp(nc_ptr), _vp(labels_ptr), _vp(best_ptr)), "kmAssignF32")


def update_workspace_size(n, d, k):
    return lib().kmUpdateWorkspaceSize(n, d, k)


def update(n, d, k, X_ptr, labels_ptr, C_old_ptr, C_new_ptr, ws_ptr, ws_bytes, Ct_ptr=None,
           nc_ptr=None, counts_ptr=None, empty_ptr=None, stream=None):
    _check(lib().kmUpdate(_vp(stream), n, d, k, _vp(X_ptr), _vp(labels_ptr), _vp(C_old_ptr),
                          _vp(C_new_ptr), _vp(Ct_ptr), _vp(nc_ptr), _vp(counts_ptr), _vp(empty_ptr),
                          _vp(ws_ptr), ws_bytes), "kmUpdate")


def update_no_alloc(n, d, k, X_ptr, labels_ptr, C_old_ptr, C_new_ptr, Ct_ptr=None, nc_ptr=None,
                    counts_ptr=None, empty_ptr=None, stream=None):
    """kmUpdateNoAlloc: kmUpdate with its workspace in the library's fixed
    per-device scratch array (no allocation; shapes beyond it are rejected)."""
    _check(lib().kmUpdateNoAlloc(_vp(stream), n, d, k, _vp(X_ptr), _vp(labels_ptr),
                                   _vp(C_old_ptr), _vp(C_new_ptr), _vp(Ct_ptr), _vp(nc_ptr),
                                   _vp(counts_ptr), _vp(empty_ptr)), "kmUpdateNoAlloc")


def compactness(n, d, k, X_ptr, labels_ptr, C_ptr, out_ptr, ws_ptr, ws_bytes, stream=None):
    _check(lib().kmCompactness(_vp(stream), n, d, k, _vp(X_ptr), _vp(labels_ptr), _vp(C_ptr),
                               _vp(out_ptr), _vp(ws_ptr), ws_bytes), "kmCompactness")


def row_norms(m0, m, d, X_ptr, out_ptr, stream=None):
    _check(lib().kmRowNorms(_vp(stream), m0, m, d, _vp(X_ptr), _vp(out_ptr)), "kmRowNorms")


def generate_mixture(seed, i0, n, d, ncomp, means_ptr, sigmas_ptr, cdf_ptr, X_ptr, z_ptr=None,
                     stream=None):
    _check(lib().kmGenerateMixture(_vp(stream), seed, i0, n, d, ncomp, _vp(means_ptr),
                                   _vp(sigmas_ptr), _vp(cdf_ptr), _vp(X_ptr), _vp(z_ptr)),
           "kmGenerateMixture")


def gather_rows(n, d, X_ptr, idx_ptr, out_ptr, stream=None):
    _check(lib().kmGatherRows(_vp(stream), n, d, _vp(X_ptr), _vp(idx_ptr), _vp(out_ptr)),
           "kmGatherRows")


def mlp_param_count(din, hidden, groups):
    return lib().kmMlpParamCount(din, hidden, groups)


def mlp_workspace_size(din, hidden, groups, batch):
    return lib().kmMlpWorkspaceSize(din, hidden, groups, batch)


def mlp_forward_backward(din, hidden, groups, batch, X_ptr, t_ptr, P_ptr, G_ptr, loss_ptr, ws_ptr,
                         ws_bytes, stream=None):
    _check(lib().kmMlpForwardBackward(_vp(stream), din, hidden, groups, batch, _vp(X_ptr), _vp(t_ptr),
                                      _vp(P_ptr), _vp(G_ptr), _vp(loss_ptr), _vp(ws_ptr), ws_bytes),
           "kmMlpForwardBackward")
