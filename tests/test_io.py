"""The on-disk slice format (src/io.jl:1-120): round trips as the
reference's own test does (test/slice.jl:22-35: write a slice directory,
load_all_finished, the maximum over the loaded slices is the MIS), and files
written in the Julia packages' formats by hand.  The reference holds no
sample files, so the formats follow Graphs.jl's LG writer and OMEinsum's
writejson schema."""
import json
import os

import numpy as np
import pytest


def _same(a, b):
    assert a.p.g.nv == b.p.g.nv and np.array_equal(a.p.g.edges, b.p.g.edges)
    if a.p.weights is None:
        assert b.p.weights is None
    else:
        assert a.p.weights.dtype == b.p.weights.dtype
        assert np.array_equal(a.p.weights, b.p.weights)
    assert a.r == b.r
    if a.code is None:
        assert b.code is None
    else:
        assert a.code.ixs == b.code.ixs
        assert np.array_equal(a.code.tree, b.code.tree)


def test_bfs_rw_round_trip_unit_weights(tb, orc, tmp_path):
    g = tb.random_regular_graph(100, 3, seed=0)
    br = tb.slice_graph(g, None, tb.ContractionTreeSlicer(sc_target=10, order_trials=20))
    ids = tb.save_slices(str(tmp_path), br)
    assert len(ids) == len(br)
    assert os.path.exists(tmp_path / f"graph_{ids[0]}.dot")
    back = tb.load_all_finished(str(tmp_path))
    assert len(back) == len(br)
    for a, b in zip(br, back):
        _same(a, b)
    assert int(orc.contract_slices(back, np.float32).max()) == orc.mis_exact(g)


def test_round_trip_float32_weights_bit_exact(tb, orc, tmp_path):
    """Non-integer Float32 weights (as test/slice.jl:28) survive the text
    format bit for bit, so the per-slice optima are identical."""
    g = tb.random_regular_graph(60, 3, seed=4)
    w = (1.0 + np.random.default_rng(4).random(g.nv)).astype(np.float32)
    br = tb.slice_graph(g, w, tb.ContractionTreeSlicer(sc_target=8, order_trials=20))
    tb.save_slices(str(tmp_path), br)
    back = tb.load_all_finished(str(tmp_path))
    for a, b in zip(br, back):
        _same(a, b)
    assert np.array_equal(orc.contract_slices(br, np.float32), orc.contract_slices(back, np.float32))


def test_float32_weights_written_as_julia_string(tb, tmp_path):
    """save_weights writes what string(w) gives in Julia (src/io.jl:17): plain
    decimals that parse(Float32, line) (src/io.jl:40) accepts, with no `f0`
    suffix, each the shortest text that reads back to the same Float32."""
    w = np.array([1.5, 2.0, 0.1, 1e-7, 3e7, 100.0], np.float32)
    f = tmp_path / "weights_1.txt"
    tb.save_weights(str(f), w)
    lines = f.read_text().split()
    assert lines[0] == "Float32"
    assert lines[1:4] == ["1.5", "2.0", "0.1"]
    for t, x in zip(lines[1:], w):
        assert "f" not in t
        assert np.float32(float(t)) == x
    assert np.array_equal(tb.load_weights(str(f)), w)
    # the reader also takes Julia's Float32 literal form (repr / show)
    f.write_text("Float32\n1.5f0\n2.0f-3\n")
    assert tb.load_weights(str(f)).tolist() == [np.float32(1.5), np.float32(2e-3)]


def test_empty_graph_slice_round_trip(tb, tmp_path):
    empty = tb.SlicedBranch(tb.MISProblem(tb.SimpleGraph(0)), None, 7)
    tb.save_slices(str(tmp_path), [empty])
    (back,) = tb.load_all_finished(str(tmp_path))
    _same(empty, back)
    assert (tmp_path / "eincode_1.json").read_text() == "nothing"


def test_julia_written_files(tb, orc, tmp_path):
    """Files as Graphs.savegraph, OMEinsum writejson, save_weights and CSV.jl
    write them: a path 1-2-3 with weights [1.5, 2, 1.25] (Float32 printed
    by string(w), without a type suffix), r = 0.5."""
    (tmp_path / "graph_7.dot").write_text(
        "# written by Graphs.jl\n3,2,u,graph,Int64,simplegraph\n1,2\n2,3\n")
    (tmp_path / "weights_7.txt").write_text("Float32\n1.5\n2.0\n1.25\n")
    code = {"label-type": "Int64", "inputs": [[1], [2], [3], [1, 2], [2, 3]], "output": [],
            "tree": {"isleaf": False, "eins": {"ixs": [[1, 2], [2, 3]], "iy": []},
                     "args": [
                         {"isleaf": False, "eins": {"ixs": [[1], [1, 2]], "iy": [1, 2]},
                          "args": [
                              {"isleaf": False, "eins": {"ixs": [[1], [2]], "iy": [1, 2]},
                               "args": [{"isleaf": True, "tensorindex": 1},
                                        {"isleaf": True, "tensorindex": 2}]},
                              {"isleaf": True, "tensorindex": 4}]},
                         {"isleaf": False, "eins": {"ixs": [[3], [2, 3]], "iy": [2, 3]},
                          "args": [{"isleaf": True, "tensorindex": 3},
                                   {"isleaf": True, "tensorindex": 5}]}]}}
    # the second child's iy keeps 2 and 3? 3 occurs only in leaves 3 and 5,
    # both inside it, so the rule drops 3: fix the file to what the rule gives
    code["tree"]["args"][1]["eins"]["iy"] = [2]
    code["tree"]["args"][0]["eins"]["iy"] = [2]
    (tmp_path / "eincode_7.json").write_text(json.dumps(code))
    (tmp_path / "slices.csv").write_text("id,sc,tc,r\n7,2.0,3.0,0.5\n")
    (br,) = tb.load_all_finished(str(tmp_path))
    assert br.p.g.nv == 3 and br.p.g.edges.tolist() == [[0, 1], [1, 2]]
    assert br.p.weights.dtype == np.float32 and br.p.weights.tolist() == [1.5, 2.0, 1.25]
    assert br.code.ixs == [[0], [1], [2], [0, 1], [1, 2]]
    assert br.r == 0.5
    # MIS of the weighted path: {1, 3} = 2.75, {2} = 2; plus r
    assert float(orc.contract_slices([br], np.float32)[0]) == 2.75 + 0.5


def test_unit_weight_and_int_files(tb, tmp_path):
    p = tmp_path / "w.txt"
    p.write_text("UnitWeight\n12\n")
    w = tb.load_weights(str(p))
    assert isinstance(w, tb.UnitWeight) and w.n == 12
    p.write_text("Int32\n3\n-4\n5\n")
    assert tb.load_weights(str(p)).tolist() == [3, -4, 5]
    p.write_text("Complex64\n1\n")
    with pytest.raises(ValueError):
        tb.load_weights(str(p))


def test_malformed_codes_are_rejected(tb, tmp_path):
    p = tmp_path / "c.json"
    base = {"label-type": "Int64", "inputs": [[1], [1]], "output": [],
            "tree": {"isleaf": False, "eins": {"ixs": [[1], [1]], "iy": []},
                     "args": [{"isleaf": True, "tensorindex": 1},
                              {"isleaf": True, "tensorindex": 2}]}}
    p.write_text(json.dumps(base))
    assert tb.load_code(str(p)).ixs == [[0], [0]]
    bad = json.loads(json.dumps(base))
    bad["tree"]["eins"]["iy"] = [1]  # the root keeps no label
    p.write_text(json.dumps(bad))
    with pytest.raises(ValueError):
        tb.load_code(str(p))
    bad = json.loads(json.dumps(base))
    bad["tree"]["args"].append({"isleaf": True, "tensorindex": 2})  # not binary
    p.write_text(json.dumps(bad))
    with pytest.raises(ValueError):
        tb.load_code(str(p))
    bad = json.loads(json.dumps(base))
    bad["tree"]["args"][1]["tensorindex"] = 1  # leaf used twice
    p.write_text(json.dumps(bad))
    with pytest.raises(ValueError):
        tb.load_code(str(p))


def test_deep_tree_round_trip(tb, tmp_path):
    """A chain of 3000 leaves (deeper than Python's default recursion
    limit) is written and read back."""
    n = 3000
    ixs = [[i, i + 1] for i in range(n)]
    ct = 0
    for i in range(1, n):
        ct = tb.ContractionTree(ct, i)
    code = tb.CompressedEinsum(ixs, [], ct)
    p = tmp_path / "deep.json"
    tb.save_code(str(p), code)
    back = tb.load_code(str(p))
    assert back.ixs == ixs and np.array_equal(back.tree, code.tree)
