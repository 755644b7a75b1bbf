"""sdx_csr_from_stencil on device arrays against the oracle's CSR of the same
seven diagonals: random 7-point operators with structural zeros (boundary
couplings, identity rows, dropped couplings), a unit diagonal given as a null
pointer, a too-short output (pattern only), and bad arguments."""
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


def _run(hip, A, shape, cap=None):
    import torch
    dev = torch.device("cuda:0")
    N = A.shape[1]
    D = [torch.from_numpy(np.ascontiguousarray(a)).to(dev) for a in A]
    cap = 7 * N if cap is None else cap
    rp = torch.zeros(N + 1, dtype=torch.int32, device=dev)
    col = torch.zeros(max(cap, 1), dtype=torch.int32, device=dev)
    val = torch.zeros(max(cap, 1), dtype=torch.float64, device=dev)
    ptrs = (ctypes.c_void_p * 7)(*[d.data_ptr() for d in D])
    torch.cuda.synchronize()
    nnz = hip.sdx_csr_from_stencil(*shape, ptrs, ctypes.c_void_p(rp.data_ptr()), ctypes.c_void_p(col.data_ptr()),
                                   ctypes.c_void_p(val.data_ptr()), cap)
    return nnz, rp.cpu().numpy(), col.cpu().numpy(), val.cpu().numpy()


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
    import torch
    A = _stencil(shape, 7, unit_diag=True)
    rp_ref, col_ref, val_ref = O.stencil_csr(A, shape)
    dev = torch.device("cuda:0")
    N = A.shape[1]
    D = [torch.from_numpy(np.ascontiguousarray(a)).to(dev) for a in A]
    rp = torch.zeros(N + 1, dtype=torch.int32, device=dev)
    col = torch.zeros(7 * N, dtype=torch.int32, device=dev)
    val = torch.zeros(7 * N, dtype=torch.float64, device=dev)
    ptrs = (ctypes.c_void_p * 7)(None, *[d.data_ptr() for d in D[1:]])      # c null: unit diagonal
    torch.cuda.synchronize()
    nnz = hip.sdx_csr_from_stencil(*shape, ptrs, ctypes.c_void_p(rp.data_ptr()), ctypes.c_void_p(col.data_ptr()),
                                   ctypes.c_void_p(val.data_ptr()), 7 * N)
    assert nnz == len(col_ref)
    np.testing.assert_array_equal(rp.cpu().numpy(), rp_ref)
    np.testing.assert_array_equal(col.cpu().numpy()[:nnz], col_ref)
    # too short an output: the row pointers (the pattern) are still complete
    nnz2, rp2, _, _ = _run(hip, A, shape, cap=10)
    assert nnz2 == nnz
    np.testing.assert_array_equal(rp2, rp_ref)


def test_csr_rejects_bad_arguments(hip):
    print("\n[gpu csr] bad arguments", flush=True)
    ptrs = (ctypes.c_void_p * 7)()
    assert hip.sdx_csr_from_stencil(0, 4, 4, ptrs, None, None, None, 0) == -1
    assert hip.sdx_csr_from_stencil(4, 4, 4, None, None, None, None, 0) == -1
