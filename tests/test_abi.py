"""CPU tests of the C ABI boundary: libplanc_hip.so loads and exports every
function include/planc_hip.h declares (no GPU call is made), and the
Philox4x64 generator of planc_amd/csrc/philox.h reproduces numpy's stream."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "planc_hip.h")


def declared_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(planc_\w+)\s*\(", src, flags=re.M)
    return sorted(set(names))


def test_header_declares_entry_points():
    names = declared_functions()
    for n in ["planc_create", "planc_nmf_run", "planc_ntf_run", "planc_mttkrp", "planc_gemm_tn"]:
        assert n in names


def test_library_exports_every_declared_symbol():
    from planc_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.fail("libplanc_hip.so not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, f"declared but not exported: {missing}"
    # the Python binding declares exactly the header's functions
    assert sorted(_lib.SIGNATURES) == declared_functions()


def test_ntf_tree_split_rule():
    """planc_ntf_tree_split (host arithmetic, no GPU): the s minimising the
    larger block product, nmodes/2 on ties, against a brute force."""
    from planc_amd import _lib
    lib = _lib.load()

    def brute(dims):
        N = len(dims)
        best = None
        for s in range(N - 1):
            v = max(int(np.prod(dims[:s + 1])), int(np.prod(dims[s + 1:])))
            if best is None or v < best[0] or (v == best[0] and s == N // 2):
                best = (v, s)
        return best[1]

    rng = np.random.default_rng(3)
    cases = [(128, 128, 128, 128), (256, 256, 256), (10, 1000, 10, 10), (5, 7), (2, 3, 5, 7, 11, 13, 17)]
    cases += [tuple(int(x) for x in rng.integers(1, 60, size=rng.integers(2, 8))) for _ in range(200)]
    for dims in cases:
        arr = (ctypes.c_long * len(dims))(*dims)
        assert lib.planc_ntf_tree_split(len(dims), arr) == brute(dims), dims
    assert lib.planc_ntf_tree_split(1, (ctypes.c_long * 1)(5)) == -1
    assert lib.planc_ntf_tree_split(3, (ctypes.c_long * 3)(4, 0, 2)) == -1
    assert _lib.SIGNATURES["planc_ntf_tree_split"][0] is ctypes.c_int


def test_version_string_without_gpu():
    from planc_amd import _lib
    lib = _lib.load()
    assert b"gfx950" in lib.planc_version()


def test_philox_matches_numpy(tmp_path):
    src = tmp_path / "t.cpp"
    src.write_text('#include "philox.h"\n#include <cstdio>\nint main(){\n'
                   ' for (int e = 0; e < 9; ++e) printf("%llu\\n", (unsigned long long)'
                   'planc_rng::word(12345, 7, e));\n'
                   ' printf("%.17g\\n", planc_rng::normal(99, 3, 5));\n}\n')
    exe = tmp_path / "t"
    subprocess.run(["g++", "-O2", "-I", os.path.join(ROOT, "planc_amd", "csrc"), str(src), "-o",
                    str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()
    got = [int(x) for x in out[:9]]
    ref = np.random.Philox(key=[12345, 0], counter=[0, 7, 0, 0]).random_raw(9)
    assert got == [int(x) for x in ref]
    from planc_amd import datagen as D
    z = D.normal(99, 3, (6,))
    assert abs(float(out[9]) - z[5]) < 1e-14
