"""C-ABI checks that need no GPU: the library loads, exports every function
include/kmeans_amd.h declares, and host-side argument checks reject bad
shapes before touching a device."""
import ctypes
import os
import re

import numpy as np
import pytest

import kmeans_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    txt = open(os.path.join(ROOT, "include", "kmeans_amd.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = re.findall(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(", txt)
    skip = {"if", "sizeof", "defined"}
    return sorted({n for n in names if n not in skip and not n.isupper()})


def test_library_exports_every_declared_symbol():
    L = kmeans_amd.lib()
    names = declared_functions()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    # and the Python mirror binds each of them with a signature
    assert set(names) <= set(kmeans_amd.SIGNATURES), set(names) - set(kmeans_amd.SIGNATURES)


def test_error_strings_match_reference_table():
    L = kmeans_amd.lib()
    assert L.kmeansGetErrorString(0) == b"no error"
    assert L.kmeansGetErrorString(3) == b"error in computing the closest centers"
    assert L.kmeansGetErrorString(13) == b"algorithm did not converge"


def test_create_rejects_bad_shapes_without_device():
    L = kmeans_amd.lib()
    h = ctypes.c_void_p()
    for m, n, k in [(10, 4, 11), (0, 4, 1), (10, 0, 2), (10, 4, 0)]:
        assert L.kmeansCreate(ctypes.byref(h), m, n, k, None) == kmeans_amd.KMEANS_ERROR_INVALID_ARGUMENT
    opt = kmeans_amd.Options(7, 0, None, 0, 0)  # unknown precision
    assert L.kmeansCreate(ctypes.byref(h), 100, 4, 2, ctypes.byref(opt)) == 14


def test_kernel_api_rejects_bad_arguments():
    L = kmeans_amd.lib()
    assert L.kmColNorms(None, 0, 4, None, None) == 8
    assert L.kmAssignF32(None, 10, None, 0, None, None, 4, None, None, None, None) == 8
    assert L.kmUpdate(None, 10, 4, 2, None, None, None, None, None, None, None, None, None, 0) == 8
    assert L.kmCompactness(None, 10, 4, 2, None, None, None, None, None, 0) == 8
    assert L.rowTransformMinimumF(0, 10, 4, None, None, None, None) == 8
    assert L.rowTransformMinimumD(-1, 10, 4, None, None, None, None) == 8
    assert L.ClosestCentersD(0, 10, 0, None, 4, None, None, None, None, 0, None, None, None) == 8


def test_compute_kmeans_d_rejects_bad_shapes_without_device():
    X = np.zeros((10, 3))
    for k in (11, 0):
        st, _ = kmeans_amd.compute_kmeans_d(X, k, -1.0, 3) if k else (
            kmeans_amd.lib().computeKmeansD(kmeans_amd._np_ptr(X), 10, 3, 0, -1.0, 3, 0, 0, 0, 0,
                                            kmeans_amd._np_ptr(np.zeros(3)), 1.0), None)
        assert st == kmeans_amd.KMEANS_ERROR_INVALID_ARGUMENT


def test_mlp_api_rejects_bad_arguments():
    L = kmeans_amd.lib()
    assert L.kmMlpParamCount(96, 256, 128) == 256 * 96 + 256 + 256 * 256 + 256 + 128 * 256 + 128
    wsb = L.kmMlpWorkspaceSize(96, 128, 128, 8)
    args = (1, 1, 1, 1, 1, 1)
    assert L.kmMlpForwardBackward(None, 96, 100, 128, 8, *args, wsb) == 8  # hidden % 16
    assert L.kmMlpForwardBackward(None, 96, 128, 120, 8, *args, wsb) == 8  # groups % 16
    assert L.kmMlpForwardBackward(None, 90, 128, 128, 8, *args, wsb) == 8  # din % 16
    assert L.kmMlpForwardBackward(None, 96, 128, 128, 8, *args, 16) == 8   # workspace too small
    assert L.kmMlpForwardBackward(None, 96, 128, 128, 8, None, 1, 1, 1, 1, 1, wsb) == 8


def test_workspace_size_and_layout():
    L = kmeans_amd.lib()
    assert L.kmCentersLd(1) == 256 and L.kmCentersLd(256) == 256 and L.kmCentersLd(257) == 512
    small = L.kmUpdateWorkspaceSize(1000, 8, 4)
    big = L.kmUpdateWorkspaceSize(10_000_000, 128, 1024)
    assert 0 < small < big
    # bounded by one canonical block of points, not by N
    assert L.kmUpdateWorkspaceSize(200_000_000, 64, 4096) < 2 * L.kmUpdateWorkspaceSize(kmeans_amd.KM_BLOCK, 64, 4096)


def test_no_cpu_fallback_in_package():
    src = open(os.path.join(ROOT, "kmeans_amd", "__init__.py")).read()
    assert "oracle" not in src.split('"""', 2)[2]


def test_create_rejects_bad_proposer_options_without_device():
    L = kmeans_amd.lib()
    h = ctypes.c_void_p()
    bad = [dict(proposer=3), dict(proposer=1, groups=20), dict(proposer=1, hidden=100),
           dict(proposer=2, groups=2000), dict(proposer=2, top_groups=9),
           dict(proposer=2, groups=4, top_groups=5), dict(proposer=1, tile_rows=1 << 23)]
    for kw in bad:
        opt = kmeans_amd.Options()
        for f, v in kw.items():
            setattr(opt, f, v)
        assert L.kmeansCreate(ctypes.byref(h), 100_000, 8, 1024, ctypes.byref(opt)) == 14, kw


def test_batch_create_rejects_bad_shapes_without_device():
    L = kmeans_amd.lib()
    h = ctypes.c_void_p()
    n = np.array([500, 800, 1200], np.int64)
    p = n.ctypes.data_as(ctypes.c_void_p)
    assert L.kmeansBatchCreate(ctypes.byref(h), 3, p, 16, 600, None) == 14   # k > n[0]
    assert L.kmeansBatchCreate(ctypes.byref(h), 3, p, 65, 16, None) == 14    # d > 64
    assert L.kmeansBatchCreate(ctypes.byref(h), 3, p, 16, 2000, None) == 14  # k > 1024
    assert L.kmeansBatchCreate(ctypes.byref(h), 0, p, 16, 16, None) == 14
    opt = kmeans_amd.Options(2, 0, None, 0, 0)  # FP16: not offered
    assert L.kmeansBatchCreate(ctypes.byref(h), 3, p, 16, 16, ctypes.byref(opt)) == 14
    opt = kmeans_amd.Options(1, 0, None, 0, 0)  # BF16: d <= 32 only
    assert L.kmeansBatchCreate(ctypes.byref(h), 3, p, 33, 16, ctypes.byref(opt)) == 14


def test_c4_instance_sizes():
    from kmeans_amd import datagen
    n, seeds = datagen.c4_instances()
    assert len(n) == 2048 and n.min() >= 1000 and n.max() <= 200000
    # lognormal with median 20,000 and sigma 1: sum about 2048 * 20000 * exp(1/2)
    assert 5.5e7 < n.sum() < 7.5e7
    assert abs(np.median(n) / 20000 - 1) < 0.1
    assert len(set(seeds.tolist())) == 2048


def test_load_raw_f32_maps_reference_files(tmp_path):
    X = np.arange(7 * 5, dtype=np.float32).reshape(7, 5)
    path = tmp_path / "points.bin"
    X.tofile(path)
    M = kmeans_amd.load_raw_f32(str(path), 5)
    assert isinstance(M, np.memmap) and M.shape == (7, 5)
    assert np.array_equal(M, X)
    assert np.array_equal(kmeans_amd.load_raw_f32(str(path), 5, m=3), X[:3])
    with pytest.raises(ValueError):
        kmeans_amd.load_raw_f32(str(path), 5, m=8)  # fewer floats than m*n, as fread would find
    (tmp_path / "odd.bin").write_bytes(b"\0" * 6)
    with pytest.raises(ValueError):
        kmeans_amd.load_raw_f32(str(tmp_path / "odd.bin"), 1)


def test_create_rejects_reduced_precision_above_256_dimensions():
    L = kmeans_amd.lib()
    h = ctypes.c_void_p()
    opt = kmeans_amd.Options(1, 0, None, 0, 0)  # BF16
    assert L.kmeansCreate(ctypes.byref(h), 1000, 257, 4, ctypes.byref(opt)) == 14


def test_cpp_class_header_compiles():
    src = '#include "kmeans_amd.hpp"\nint main() { kmeans_amd::kmeans<float> *a = nullptr; ' \
          'kmeans_amd::kmeans<double> *b = nullptr; return a == nullptr && b == nullptr ? 0 : 1; }\n'
    import subprocess
    r = subprocess.run(["g++", "-std=c++14", "-Wall", "-Werror", "-fsyntax-only", "-I",
                        os.path.join(ROOT, "include"), "-x", "c++", "-"], input=src, text=True,
                       capture_output=True)
    assert r.returncode == 0, r.stderr


def test_reference_init_rows_follow_the_reference_steps():
    """kmeansReferenceInitRows against a program that runs the reference's
    own steps (kmeans.cpp:553-556 shuffle, :876-883 pick with the global
    srand/rand), for small and large m, k close to m (where the unsorted
    binary_search lets duplicates through) and k = m."""
    import subprocess
    helper = os.path.join(ROOT, "tests", "cpp", "reference_init")
    assert os.path.exists(helper), "build the library first (it builds the helper too)"
    for m, k in [(10, 3), (1000, 8), (64, 60), (5, 5), (100_000, 1024)]:
        out = subprocess.run([helper, str(m), str(k)], capture_output=True, text=True,
                             check=True).stdout.split()
        assert kmeans_amd.reference_init_rows(m, k).tolist() == [int(v) for v in out], (m, k)
    # known index sequences (this image's libstdc++ and glibc)
    assert kmeans_amd.reference_init_rows(10, 3).tolist() == [4, 2, 7]
    assert kmeans_amd.reference_init_rows(1000, 8).tolist() == [946, 542, 598, 509, 471, 220,
                                                                299, 710]
    rows = kmeans_amd.reference_init_rows(64, 60)
    assert len(set(rows.tolist())) < 60  # the reference's dedup misses some repeats


def test_reference_init_rows_leave_the_callers_rand_alone():
    libc = ctypes.CDLL("libc.so.6")
    libc.srand(42)
    first = [libc.rand() for _ in range(3)]
    assert first == [71876166, 708592740, 1483128881]  # glibc's srand(42) stream
    libc.srand(7)
    a = libc.rand()
    libc.srand(7)
    kmeans_amd.reference_init_rows(1000, 50)
    assert libc.rand() == a
    L = kmeans_amd.lib()
    out = np.zeros(4, np.int32)
    assert L.kmeansReferenceInitRows(3, 4, out.ctypes.data_as(ctypes.c_void_p)) == 14


def test_load_npy_maps_float32_points(tmp_path):
    X = np.arange(24, dtype=np.float32).reshape(6, 4)
    np.save(tmp_path / "x.npy", X)
    M = kmeans_amd.load_npy(str(tmp_path / "x.npy"))
    assert isinstance(M, np.memmap) and np.array_equal(M, X)
    np.save(tmp_path / "d.npy", X.astype(np.float64))
    with pytest.raises(ValueError):
        kmeans_amd.load_npy(str(tmp_path / "d.npy"))
    np.save(tmp_path / "f.npy", np.asfortranarray(X))
    with pytest.raises(ValueError):
        kmeans_amd.load_npy(str(tmp_path / "f.npy"))
    np.save(tmp_path / "o.npy", np.array([{"a": 1}], dtype=object), allow_pickle=True)
    with pytest.raises(ValueError):
        kmeans_amd.load_npy(str(tmp_path / "o.npy"))
