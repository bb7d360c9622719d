"""The C ABI used from C: tests/host/c_client.c, compiled as C99 with
-pedantic -Werror against include/tbh.h and linked to libtbh.so, writes a
triangle's MIS network as one tbh_slice by hand.  Its host-only calls
(complexity, peak memory, plan) run here and equal the oracle's and the
closed form; with a GPU it contracts the slice in both modes and both
element types to the triangle's MIS, 1.  The client is built as a shared
library (its main renamed) and called in this process, so that no program
is started from a process that uses the GPU."""
import ctypes as C
import math
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.join(ROOT, "tensorbranching.jl_amd", "lib")
# the slice c_client.c writes: vertex leaves 0..2, edge leaves (0,1), (1,2),
# (0,2); tree nodes 6 = (0, 3), 7 = (1, 6), 8 = (4, 7), 9 = (2, 8), 10 = (5, 9)
IX_PTR = [0, 1, 2, 3, 5, 7, 9]
IX = [0, 1, 2, 0, 1, 1, 2, 0, 2]
TREE = [0, 3, 1, 6, 4, 7, 2, 8, 5, 9]


@pytest.fixture(scope="module")
def client(tmp_path_factory):
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("gcc not available")
    if not os.path.exists(os.path.join(LIBDIR, "libtbh.so")):
        pytest.skip("libtbh.so is not built")
    so = str(tmp_path_factory.mktemp("c_client") / "libc_client.so")
    cmd = [gcc, "-std=c99", "-Wall", "-Wextra", "-Werror", "-pedantic", "-fPIC", "-shared",
           "-Dmain=tbh_c_client_main", "-I", os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "host", "c_client.c"),
           "-L", LIBDIR, "-ltbh", "-Wl,-rpath," + LIBDIR, "-o", so]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
    lib = C.CDLL(so)
    lib.tbh_c_client_main.restype = C.c_int
    lib.tbh_c_client_main.argtypes = [C.c_int, C.POINTER(C.c_char_p)]
    return lib


def _run(lib, capfd, *args):
    argv = [b"c_client"] + [a.encode() for a in args]
    arr = (C.c_char_p * len(argv))(*argv)
    capfd.readouterr()
    rc = lib.tbh_c_client_main(len(argv), arr)
    out, err = capfd.readouterr()
    assert rc == 0, (out, err)
    return dict(kv.split("=") for kv in out.split())


def test_c_client_host_calls(tb, orc, client, capfd):
    got = _run(client, capfd)
    code = tb.CompressedEinsum.from_arrays(np.array(IX_PTR, np.int32), np.array(IX, np.int32),
                                           np.array(TREE, np.int32))
    tc, sc, rwc = orc.complexity(code)
    # closed form: the five nodes span {0,1}, {0,1}, {0,1,2}, {0,2}, {0,2}
    assert abs(tc - math.log2(4 + 4 + 8 + 4 + 4)) < 1e-12 and sc == 2
    assert abs(float(got["tc"]) - tc) < 1e-6
    assert abs(float(got["sc"]) - sc) < 1e-6
    assert abs(float(got["rwc"]) - rwc) < 1e-6
    assert abs(float(got["peak"]) - orc.peak_memory(code)) < 1e-6
    g = tb.SimpleGraph(3, [(0, 1), (1, 2), (0, 2)])
    br = tb.SlicedBranch(tb.MISProblem(g), code, 0)
    assert int(got["steps"]) == len(tb.plan_steps(br, np.float32, "batched")) > 0
    assert int(got["malformed"]) == -1  # TBH_EINVAL


@pytest.mark.gpu
def test_c_client_contracts_on_gpu(client, capfd):
    got = _run(client, capfd, "--gpu")
    assert float(got["batched"]) == 1.0 and float(got["plain"]) == 1.0
    assert int(got["int32"]) == 1
