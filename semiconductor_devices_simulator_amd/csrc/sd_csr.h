// sd_csr.h - CSR export of a 7-point stencil operator (k_csr.hip).
#pragma once
#include <hip/hip_runtime.h>

#define SD_CSR_EARG (-1LL)        // bad shape or pointer
#define SD_CSR_ERANGE (-2LL)      // more nonzeros than an int row pointer holds
#define SD_CSR_EDEVICE (-3LL)     // a HIP call failed
#define SD_CSR_ESIZE (-4LL)       // more rows than the allocation-free variant's fixed scratch holds

// ints of scratch csr_from_stencil needs for N rows
int csr_scratch_words(long N);
// CSR of the operator with diagonals diag[0..6] = c, xm, xp, ym, yp, zm, zp
// (device arrays of N = Nx Ny Nz doubles each, natural node order; diag[0]
// null: unit diagonal) into the device arrays rowptr (N + 1), col and val
// (cap entries): every nonzero entry, columns ascending within a row.
// scratch: csr_scratch_words(N) ints of device memory, 8-byte aligned.
// Returns the number of nonzeros (entries beyond cap are not written) after
// synchronising stream s, or SD_CSR_EDEVICE.
long long csr_from_stencil(hipStream_t s, int Nx, int Ny, int Nz, const double *const *diag, int *rowptr,
                           int *col, double *val, long long cap, int *scratch);
