"""The reference-style C++ class (include/kmeans_amd.hpp) on the GPU: a
compiled program (tests/cpp/class_api.cpp, built with the library) runs the
loop through compute(), then step by step through closestCenters /
clusterCenters / compactness, in float and in double; results against the
oracle bit for bit."""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "class_api_test")


@pytest.mark.parametrize("n,d,k,iters", [(20_000, 16, 24, 4), (3000, 3, 5, 6)])
def test_class_api_matches_oracle(orc, tmp_path, n, d, k, iters):
    assert os.path.exists(BIN), "build the library first (it builds the test program too)"
    X, _, _, _ = orc.mixture(90 + d, n, d, k, -6, 6, 1.0, 1.0)
    src = tmp_path / "points.f32"
    out = tmp_path / "out.bin"
    X.tofile(src)
    r = subprocess.run([BIN, str(src), str(n), str(d), str(k), str(iters), str(out)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, (r.returncode, r.stdout, r.stderr)
    raw = out.read_bytes()
    off = 0
    C = np.frombuffer(raw, np.float32, d * k, off).reshape(d, k); off += 4 * d * k
    lab = np.frombuffer(raw, np.int32, n, off); off += 4 * n
    comp = np.frombuffer(raw, np.float64, 1, off)[0]; off += 8
    Cd = np.frombuffer(raw, np.float64, d * k, off).reshape(d, k); off += 8 * d * k
    labd = np.frombuffer(raw, np.int32, n, off); off += 4 * n
    compd = np.frombuffer(raw, np.float64, 1, off)[0]; off += 8
    fmatch, fconv, dmatch, dconv, cmatch = np.frombuffer(raw, np.int32, 5, off)
    ref = orc.lloyd(X, X[:k].copy(), iters)
    assert np.array_equal(C, ref["centers"].T)
    assert np.array_equal(lab, ref["labels"])
    assert comp == ref["comp_hist"][-1]
    assert fmatch == 1  # float step-by-step equals compute()
    assert fconv == 13  # criterion < 0 is never met
    assert cmatch == 1  # computeKmeansF with the default miniBatchFraction
    ref64 = orc.lloyd64(X.astype(np.float64), X[:k].astype(np.float64), iters)
    assert np.array_equal(Cd, ref64["centers"].T)
    assert dmatch == 1  # double step-by-step equals compute()
    assert dconv == 13
    assert np.array_equal(labd, ref64["labels"])
    assert compd == ref64["comp_hist"][-1]
