"""Max-plus GEMM (tbh_maxplus_gemm, the TropicalGEMM mul! seam) on the GPU
against the oracle, bit-exact."""
import numpy as np
import pytest

from conftest import rand_tropical

pytestmark = pytest.mark.gpu

SHAPES = [(1, 1, 1), (3, 5, 7), (16, 16, 16), (64, 64, 64), (100, 37, 129), (128, 128, 16),
          (256, 64, 512), (16, 4096, 16), (4096, 16, 64), (512, 512, 256), (130, 257, 3),
          (2048, 256, 256),
          # tall-skinny tiles (256 x 64, 256 x 32): ragged M, N, K, and K not a
          # multiple of 4 (scalar loads)
          (300, 33, 20), (257, 64, 128), (1000, 32, 17), (4100, 7, 9), (512, 65, 64)]


@pytest.mark.parametrize("dtype", [np.float32, np.int32])
@pytest.mark.parametrize("shape", SHAPES)
def test_gemm_matches_oracle(backend, orc, dtype, shape):
    M, N, K = shape
    rng = np.random.default_rng(M * 7919 + N * 31 + K)
    A = rand_tropical(rng, (M, K), dtype)
    B = rand_tropical(rng, (K, N), dtype)
    got = backend.maxplus_gemm(A, B)
    ref = orc.maxplus_gemm(A, B)
    assert got.dtype == ref.dtype
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("dtype", [np.float32, np.int32])
def test_gemm_batched_skinny(backend, orc, dtype):
    rng = np.random.default_rng(6)
    A = rand_tropical(rng, (768, 40, 3), dtype)
    B = rand_tropical(rng, (40, 48, 3), dtype)
    assert np.array_equal(backend.maxplus_gemm(A, B), orc.maxplus_gemm(A, B))


@pytest.mark.parametrize("dtype", [np.float32, np.int32])
def test_gemm_batched(backend, orc, dtype):
    rng = np.random.default_rng(5)
    A = rand_tropical(rng, (64, 32, 6), dtype)
    B = rand_tropical(rng, (32, 96, 6), dtype)
    assert np.array_equal(backend.maxplus_gemm(A, B), orc.maxplus_gemm(A, B))


def test_gemm_large_identity(backend):
    """A (x) I = A at M = N = K = 4096 (the tropical identity has 0 on the
    diagonal and -inf elsewhere)."""
    n = 4096
    rng = np.random.default_rng(11)
    A = rand_tropical(rng, (n, n), np.float32)
    I = np.full((n, n), -np.inf, np.float32, order="F")
    np.fill_diagonal(I, 0)
    assert np.array_equal(backend.maxplus_gemm(A, I), A)
    assert np.array_equal(backend.maxplus_gemm(I, A), A)


def test_gemm_large_rank_one(backend):
    """A[m,k] = u[m] + v[k] gives C[m,n] = u[m] + max_k(v[k] + B[k,n])."""
    rng = np.random.default_rng(12)
    M, N, K = 4096, 2048, 1024
    u = rng.integers(0, 1000, M).astype(np.float32)
    v = rng.integers(0, 1000, K).astype(np.float32)
    B = rand_tropical(rng, (K, N), np.float32, neg_frac=0.0)
    A = np.asfortranarray(u[:, None] + v[None, :])
    C = backend.maxplus_gemm(A, B)
    expect = u[:, None] + (v[:, None] + B).max(axis=0)[None, :]
    assert np.array_equal(C, expect)


def test_valu_peak_bounds_the_gemm(backend):
    """The measured VALU ceiling of the max-plus mix is positive and the
    register-blocked GEMM stays below it (it issues the same mix plus
    loads)."""
    import time
    for dt in (np.float32, np.int32):
        peak = backend.valu_peak(dt)
        assert 5e12 < peak < 4e14, peak
    n = 2048
    rng = np.random.default_rng(1)
    A = np.asfortranarray(rng.integers(0, 1024, (n, n)).astype(np.float32))
    B = np.asfortranarray(rng.integers(0, 1024, (n, n)).astype(np.float32))
    backend.maxplus_gemm(A, B)
    backend.set_profiling(True)
    backend.reset_stats()
    backend.maxplus_gemm(A, B)
    ms = backend.stats()["kind_ms"][2]
    backend.set_profiling(False)
    assert 2 * n ** 3 / (ms * 1e-3) < 1.1 * backend.valu_peak(np.float32)


@pytest.mark.parametrize("dtype", [np.float32, np.int32])
def test_gemm_padded_c_keeps_gaps(backend, orc, dtype):
    """Host operands with ldc = M + 3 and a batch stride of C with a gap
    after every entry: the M x N blocks hold A (x) B, and every entry of C's
    span outside them (padding rows, gaps) comes back unchanged."""
    import ctypes as C
    from tensorbranching_jl_amd import _lib
    M, N, K, nb = 37, 20, 29, 3
    ldc, sC = M + 3, (M + 3) * N + 11
    rng = np.random.default_rng(11)
    A = rand_tropical(rng, (M, K, nb), dtype)
    B = rand_tropical(rng, (K, N, nb), dtype)
    sentinel = dtype(12345)
    span = (nb - 1) * sC + (N - 1) * ldc + M
    Cbuf = np.full(span + 5, sentinel, dtype)
    code = 0 if dtype == np.float32 else 1
    _lib.check(_lib.tbh().tbh_maxplus_gemm(backend.handle, code, M, N, K, nb, A.ctypes.data, M,
                                           M * K, B.ctypes.data, K, K * N, Cbuf.ctypes.data, ldc,
                                           sC, _lib.TBH_HOST))
    ref = orc.maxplus_gemm(A, B)
    inside = np.zeros(Cbuf.shape, bool)
    for t in range(nb):
        for n in range(N):
            o = t * sC + n * ldc
            assert np.array_equal(Cbuf[o:o + M], ref[:, n, t])
            inside[o:o + M] = True
    assert np.all(Cbuf[~inside] == sentinel)
