"""Seeded random parity cases: graph family (random 3-, 4-, 5-regular,
Erdos-Renyi, King's subgraph), size, weight type (unit, integer Float32,
Int32), slicing target (up to six below the graph's unsliced sc) and a
tuning variant (defaults, every eligible EW step direct, none direct, one
stream, no chain re-association, tiled EW prefetched into registers, persistent
grids, 128 x 128 GEMM tiles only, wider tall-skinny GEMM tiles, chain rotations
under the rank caps alone, tiled claims of one tile) drawn per case.  Each case's per-slice
optima must equal the oracle's, bit for bit, in both executors.  Cases whose
unsliced sc exceeds 28, whose slices exceed 2^28 work or 4000 in number are
skipped (deterministically, from the seed), so the oracle stays within
seconds."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GROUPS, PER_GROUP = 6, 15


def _case(tb, rng):
    fam = str(rng.choice(["reg3", "reg4", "reg5", "er", "ksg"]))
    seed = int(rng.integers(1 << 30))
    if fam.startswith("reg"):
        d = int(fam[3])
        n = int(rng.integers(20, 160))
        if d % 2:
            n &= ~1
        g = tb.random_regular_graph(n, d, seed=seed)
    elif fam == "er":
        g = tb.erdos_renyi(int(rng.integers(20, 200)), float(rng.uniform(0.01, 0.06)), seed=seed)
    else:
        g = tb.random_ksg(int(rng.integers(6, 24)), int(rng.integers(6, 24)),
                          float(rng.uniform(0.5, 1.0)), seed=seed)
    wt = str(rng.choice(["none", "f32", "i32"]))
    dtype = np.int32 if wt == "i32" else np.float32
    w = None if wt == "none" else rng.integers(1, 101, g.nv).astype(dtype)
    variant = int(rng.integers(len(VARIANTS)))
    drop = int(rng.integers(0, 7))
    nat = tb.sc(tb.SlicedBranch(tb.MISProblem(g), tb.initialize_code(g, order_trials=5, seed=seed), 0))
    desc = f"{fam} nv={g.nv} seed={seed} weights={wt} variant={variant}"
    if nat > 28:
        return None, desc
    sc = int(max(5, nat - drop))
    try:
        br = tb.slice_graph(g, w, tb.ContractionTreeSlicer(sc_target=sc, order_trials=10), seed=seed)
    except RuntimeError:
        return None, desc
    if len(br) > 4000 or max((tb.complexity(b).tc for b in br if b.code is not None),
                             default=0) > 28:
        return None, desc
    return (br, dtype, variant), f"{desc} sc_target={sc} slices={len(br)}"


VARIANTS = [{}, {"direct_max_reduce": 12}, {"direct_max_reduce": -1}, {"concurrent": 0},
            {"chain_big": 30}, {"ew_prefetch": 1}, {"ew_grid_mult": 1, "gemm_grid_mult": 1},
            {"gemm_tiles": 0}, {"gemm_tiles": 2}, {"chain_cost": 0}, {"ew_chunk": 1}]


@pytest.mark.parametrize("group", range(GROUPS))
def test_random_cases_bit_exact(tb, orc, backend, group):
    rng = np.random.default_rng(1000 + group)
    t0 = backend.tuning()
    ran = 0
    try:
        for _ in range(PER_GROUP):
            case, desc = _case(tb, rng)
            if case is None:
                continue
            br, dtype, variant = case
            ref = orc.contract_slices(br, dtype, nthreads=16)
            backend.set_tuning(**dict(t0, **VARIANTS[variant]))
            for mode in ("batched", "plain"):
                got = backend.contract_slices(br, dtype, mode)
                assert got.dtype == dtype, desc
                assert np.array_equal(got, ref), f"{mode}: {desc}"
            backend.set_tuning(**t0)
            ran += 1
    finally:
        backend.set_tuning(**t0)
    assert ran >= PER_GROUP // 2
