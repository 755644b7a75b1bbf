"""sdx_csr_from_stencil_noalloc, the allocation-free variant of
sdx_csr_from_stencil (its scan partials in a fixed device array of the
library instead of a per-call allocation), against the allocating entry
point: two calls in a row, the second on a smaller operator, must give the
same row pointers, columns and values bit for bit; rows beyond the fixed
scratch are rejected before any device work."""
import ctypes

import numpy as np
import pytest

from test_gpu_csr import Dev, _run, _stencil

pytestmark = pytest.mark.gpu


def test_noalloc_matches_the_allocating_export(hip):
    for shape, seed, unit in (((97, 33, 9), 11, False), ((40, 12, 5), 12, True)):     # the second call is smaller
        print(f"\n[gpu no-alloc] shape {shape} unit diagonal {unit}", flush=True)
        A = _stencil(shape, seed, unit_diag=unit)
        want = _run(hip, A, shape, unit_diag=unit)
        got = _run(hip, A, shape, unit_diag=unit, fn=hip.sdx_csr_from_stencil_noalloc)
        assert want[0] > 0 and got[0] == want[0]
        nnz = want[0]
        np.testing.assert_array_equal(got[1], want[1])
        np.testing.assert_array_equal(got[2][:nnz], want[2][:nnz])
        np.testing.assert_array_equal(got[3][:nnz], want[3][:nnz])


def test_noalloc_rejects_more_rows_than_its_scratch(hip):
    print("\n[gpu no-alloc] oversize", flush=True)
    with Dev(hip) as d:
        small = d.put(np.zeros(8))
        ptrs = (ctypes.c_void_p * 7)(*([small] * 7))
        rp = d.put(np.zeros(8, np.int32))
        # 2^26 + 1 rows: rejected on the host, nothing is launched
        assert hip.sdx_csr_from_stencil_noalloc((1 << 26) + 1, 1, 1, ptrs, rp, None, None, 0) == -4
