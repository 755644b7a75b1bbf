"""sdx_csr_from_stencil on device arrays against the oracle's CSR of the same
seven diagonals: random 7-point operators with structural zeros (boundary
couplings, identity rows, dropped couplings), a unit diagonal given as a null
pointer, a too-short output (pattern only), and bad arguments.  Device
memory comes from the library's own sdx_dev_* entry points."""
import ctypes

import numpy as np
import pytest

import semisim_oracle as O

pytestmark = pytest.mark.gpu

SHAPES = [(5, 4, 3), (64, 9, 7), (97, 33, 5), (1, 1, 1), (3000, 2, 1)]


def _stencil(shape, seed, unit_diag=False):
    rng = np.random.default_rng(seed)
    Nx, Ny, Nz = shape
    N = Nx * Ny * Nz
    A = -rng.uniform(0.1, 1.0, (7, N))
    A[0] = 1.0 if unit_diag else rng.uniform(4.0, 5.0, N)
    A[1:][rng.random((6, N)) < 0.2] = 0.0                     # dropped couplings
    ident = rng.random(N) < 0.05                              # identity rows
    A[1:, ident] = 0.0
    A[0, ident] = 1.0
    return A


class Dev:
    """Device buffers through the library (freed on exit)."""

    def __init__(self, hip):
        self.hip, self.bufs = hip, []

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = self.hip.sdx_dev_alloc(max(a.nbytes, 8))
        assert p
        self.bufs.append(p)
        if a.nbytes:
            assert self.hip.sdx_dev_copy(p, a.ctypes.data, a.nbytes, 1) == 1
        return p

    def get(self, p, n, dtype):
        out = np.zeros(n, dtype)
        if out.nbytes:
            assert self.hip.sdx_dev_copy(out.ctypes.data, p, out.nbytes, 2) == 1
        return out

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        for p in self.bufs:
            self.hip.sdx_dev_free(p)


def _run(hip, A, shape, cap=None, unit_diag=False, fn=None):
    fn = fn or hip.sdx_csr_from_stencil
    N = A.shape[1]
    cap = 7 * N if cap is None else cap
    with Dev(hip) as d:
        D = [d.put(a) for a in A]
        if unit_diag:
            D[0] = None
        rp = d.put(np.zeros(N + 1, np.int32))
        col = d.put(np.zeros(max(cap, 1), np.int32))
        val = d.put(np.zeros(max(cap, 1), np.float64))
        ptrs = (ctypes.c_void_p * 7)(*D)
        nnz = fn(*shape, ptrs, rp, col, val, cap)
        return nnz, d.get(rp, N + 1, np.int32), d.get(col, max(cap, 1), np.int32), d.get(val, max(cap, 1), np.float64)


@pytest.mark.parametrize("shape", SHAPES)
def test_csr_matches_oracle(hip, shape):
    print(f"\n[gpu csr] shape {shape}", flush=True)
    A = _stencil(shape, sum(shape))
    nnz, rp, col, val = _run(hip, A, shape)
    rp_ref, col_ref, val_ref = O.stencil_csr(A, shape)
    assert nnz == len(col_ref)
    np.testing.assert_array_equal(rp, rp_ref)
    np.testing.assert_array_equal(col[:nnz], col_ref)
    np.testing.assert_array_equal(val[:nnz], val_ref)       # copied, not computed: bit-exact


def test_csr_unit_diagonal_and_short_output(hip):
    shape = (40, 12, 6)
    print(f"\n[gpu csr] unit diagonal, shape {shape}", flush=True)
    A = _stencil(shape, 7, unit_diag=True)
    rp_ref, col_ref, val_ref = O.stencil_csr(A, shape)
    nnz, rp, col, val = _run(hip, A, shape, unit_diag=True)      # c null: unit diagonal
    assert nnz == len(col_ref)
    np.testing.assert_array_equal(rp, rp_ref)
    np.testing.assert_array_equal(col[:nnz], col_ref)
    np.testing.assert_array_equal(val[:nnz], val_ref)
    # too short an output: the row pointers (the pattern) are still complete
    nnz2, rp2, _, _ = _run(hip, A, shape, cap=10)
    assert nnz2 == nnz
    np.testing.assert_array_equal(rp2, rp_ref)


def test_csr_rejects_bad_arguments(hip):
    print("\n[gpu csr] bad arguments", flush=True)
    ptrs = (ctypes.c_void_p * 7)()
    assert hip.sdx_csr_from_stencil(0, 4, 4, ptrs, None, None, None, 0) == -1
    assert hip.sdx_csr_from_stencil(4, 4, 4, None, None, None, None, 0) == -1
    assert hip.sdx_csr_from_stencil(4, 4, 4, ptrs, None, None, None, 0) == -1
