"""The host planner (csrc/tbh_plan.cpp) built with AddressSanitizer and
UndefinedBehaviorSanitizer (g++, host code only) plans slices of several
graph families in both lowerings and under tuning variants without a
report."""
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _write_slices(path, branches):
    items = [b for b in branches if b.code is not None]
    with open(path, "wb") as f:
        f.write(np.int32(len(items)).tobytes())
        for b in items:
            c = b.code
            tree = np.asarray(c.tree, np.int32)
            f.write(np.asarray([b.p.g.nv, c.nleaves, len(c.ix_ptr), len(c.ix), len(tree)],
                               np.int32).tobytes())
            for a in (c.ix_ptr, c.ix, tree):
                f.write(np.ascontiguousarray(a, np.int32).tobytes())
    return len(items)


def test_planner_clean_under_sanitizers(tb, tmp_path):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    rocm_inc = "/opt/rocm/include"
    if not os.path.isdir(rocm_inc):
        pytest.skip("HIP headers not available")
    exe = tmp_path / "plan_check"
    cmd = [gxx, "-O1", "-g", "-std=c++17", "-fsanitize=address,undefined",
           "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-D__HIP_PLATFORM_AMD__",
           "-I", os.path.join(ROOT, "include"),
           "-I", os.path.join(ROOT, "tensorbranching.jl_amd", "csrc"), "-I", rocm_inc,
           os.path.join(ROOT, "tests", "host", "plan_check.cpp"),
           os.path.join(ROOT, "tensorbranching.jl_amd", "csrc", "tbh_plan.cpp"), "-o", str(exe)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
    branches = []
    for n, sc, seed in ((100, 10, 0), (160, 18, 5), (200, 22, 3)):
        g = tb.random_regular_graph(n, 3, seed=seed)
        branches += tb.slice_graph(g, None, tb.ContractionTreeSlicer(sc_target=sc, order_trials=10),
                                   seed=seed)
    g = tb.random_ksg(12, 12, 0.8, seed=1)
    branches += tb.slice_graph(g, None, tb.ContractionTreeSlicer(sc_target=12, order_trials=10),
                               seed=1)
    for n in (1, 2, 7):  # tiny graphs: single leaves, one edge
        g = tb.SimpleGraph(n, [(i, i + 1) for i in range(n - 1)])
        branches.append(tb.SlicedBranch(tb.MISProblem(g), tb.initialize_code(g, 1), 0))
    data = tmp_path / "slices.bin"
    count = _write_slices(str(data), branches)
    assert count > 100
    # the ASan runtime need not come first among preloaded libraries
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1:verify_asan_link_order=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([str(exe), str(data)], capture_output=True, text=True, env=env,
                       timeout=600)
    assert r.returncode == 0, (r.stdout[-2000:], r.stderr[-4000:])
    assert "planned %d slice variants" % (9 * count) in r.stdout
