// k_csr.hip - CSR export of a 7-point stencil operator held as seven SoA
// diagonals on the device.
//
// The reference keeps tensor-mesh operators as the diagonals of `Sten`
// (semiconductor_3d.c:210) and has no CSR form of them; its matrix holds an
// entry (k, kb) exactly where the assembled coupling is nonzero (poisson_eval
// :2732-2789: Dirichlet rows are identity, couplings toward Dirichlet nodes
// are 0; cont_assemble :3067-3106: identity rows for non-semiconductor and
// fixed-carrier nodes, couplings only toward free semiconductor nodes).  The
// export therefore lists, per row and in ascending column order (z-, y-, x-,
// diagonal, x+, y+, z+), every stored entry that is nonzero, so that the CSR
// pattern of the device operator can be compared bit for bit with the
// pattern of the reference's assembled operator.
//
// Three passes, all bandwidth-bound (the diagonals are read twice):
//   k_csr_count   one lane per row: nonzeros of the row, scanned across the
//                 workgroup (CSR_ROWS rows) into rowptr[k+1] (inclusive, within
//                 the workgroup), the workgroup's total into part[b];
//   k_csr_scan    one workgroup: exclusive scan of the workgroup totals in part;
//   k_csr_fill    one lane per row: adds its workgroup's offset to rowptr[k+1],
//                 writes the row's columns and values.
#include "sd_csr.h"
#include <hip/hip_runtime.h>
#include <cstdio>
#include <mutex>

#define CSR_ROWS 1024                  // rows per workgroup of the count and fill passes

struct Diag7 { const double *q[7]; };   // c, xm, xp, ym, yp, zm, zp (c null: unit diagonal)

// entry value of direction slot s (ascending column order) of row k, and its column
// slot: 0 zm, 1 ym, 2 xm, 3 c, 4 xp, 5 yp, 6 zp
__device__ __forceinline__ bool csr_entry(const Diag7 &A, int Nx, int Ny, int Nz, long k, int i, int j, int l,
                                          int s, int &col, double &v) {
    const long NxNy = (long)Nx * Ny;
    switch (s) {
    case 0: if (l == 0) return false; col = (int)(k - NxNy); v = A.q[5][k]; break;
    case 1: if (j == 0) return false; col = (int)(k - Nx); v = A.q[3][k]; break;
    case 2: if (i == 0) return false; col = (int)(k - 1); v = A.q[1][k]; break;
    case 3: col = (int)k; v = A.q[0] ? A.q[0][k] : 1.0; break;
    case 4: if (i == Nx - 1) return false; col = (int)(k + 1); v = A.q[2][k]; break;
    case 5: if (j == Ny - 1) return false; col = (int)(k + Nx); v = A.q[4][k]; break;
    default: if (l == Nz - 1) return false; col = (int)(k + NxNy); v = A.q[6][k]; break;
    }
    return v != 0.0;
}

__device__ __forceinline__ void csr_unidx(long k, int Nx, int Ny, int &i, int &j, int &l) {
    const long NxNy = (long)Nx * Ny;
    l = (int)(k / NxNy);
    const long r = k - (long)l * NxNy;
    j = (int)(r / Nx);
    i = (int)(r - (long)j * Nx);
}

__global__ void __launch_bounds__(CSR_ROWS) k_csr_count(Diag7 A, int Nx, int Ny, int Nz, long N, int *rowptr,
                                                         int *part) {
    __shared__ int wsum[CSR_ROWS / 64];
    const long k = (long)blockIdx.x * CSR_ROWS + threadIdx.x;
    int cnt = 0;
    if (k < N) {
        int i, j, l, col;
        double v;
        csr_unidx(k, Nx, Ny, i, j, l);
        for (int s = 0; s < 7; s++) cnt += csr_entry(A, Nx, Ny, Nz, k, i, j, l, s, col, v) ? 1 : 0;
    }
    // inclusive scan: within the wavefront by shuffles, then across wavefronts
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int x = cnt;
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        int run = 0;
        for (int q = 0; q < CSR_ROWS / 64; q++) { const int t = wsum[q]; wsum[q] = run; run += t; }
        part[blockIdx.x] = run;
    }
    __syncthreads();
    x += wsum[w];
    if (k < N) rowptr[k + 1] = x;
    if (k == 0) rowptr[0] = 0;
}

// exclusive scan of nb workgroup totals in place, in one workgroup; the
// grand total (int64: the caller checks it fits an int row pointer) to *total
__global__ void __launch_bounds__(1024) k_csr_scan(int *part, int nb, long long *total) {
    __shared__ long long s[1024];
    const int t = threadIdx.x;
    const int per = (nb + 1023) / 1024;
    const int a = t * per, b = min(nb, a + per);
    long long acc = 0;
    for (int q = a; q < b; q++) acc += part[q];
    s[t] = acc;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
        const long long y = t >= o ? s[t - o] : 0;
        __syncthreads();
        s[t] += y;
        __syncthreads();
    }
    long long run = t ? s[t - 1] : 0;
    for (int q = a; q < b; q++) {
        const int v = part[q];
        part[q] = (int)(run > 0x7fffffffLL ? 0x7fffffff : run);
        run += v;
    }
    if (t == 1023) *total = s[1023];
}

__global__ void __launch_bounds__(CSR_ROWS) k_csr_fill(Diag7 A, int Nx, int Ny, int Nz, long N, int *rowptr,
                                                        const int *part, int *col, double *val, long long cap) {
    const long k = (long)blockIdx.x * CSR_ROWS + threadIdx.x;
    if (k >= N) return;
    const int off = part[blockIdx.x];
    int i, j, l;
    csr_unidx(k, Nx, Ny, i, j, l);
    const int end = rowptr[k + 1] + off;
    rowptr[k + 1] = end;
    int cnt = 0, c;
    double v;
    for (int s = 0; s < 7; s++) cnt += csr_entry(A, Nx, Ny, Nz, k, i, j, l, s, c, v) ? 1 : 0;
    long long p = (long long)end - cnt;
    if ((long long)end > cap) return;               // the caller's arrays are too short: pattern only
    for (int s = 0; s < 7; s++)
        if (csr_entry(A, Nx, Ny, Nz, k, i, j, l, s, c, v)) {
            col[p] = c;
            val[p] = v;
            p++;
        }
}

int csr_scratch_words(long N) { return (int)((N + CSR_ROWS - 1) / CSR_ROWS) + 4; }

long long csr_from_stencil(hipStream_t s, int Nx, int Ny, int Nz, const double *const *diag, int *rowptr, int *col,
                           double *val, long long cap, int *scratch) {
    const long N = (long)Nx * Ny * Nz;
    const int nb = (int)((N + CSR_ROWS - 1) / CSR_ROWS);
    Diag7 A;
    for (int q = 0; q < 7; q++) A.q[q] = diag[q];
    long long *total = reinterpret_cast<long long *>(scratch + ((nb + 1) & ~1));   // 8-byte aligned after part
    hipLaunchKernelGGL(k_csr_count, dim3(nb), dim3(CSR_ROWS), 0, s, A, Nx, Ny, Nz, N, rowptr, scratch);
    hipLaunchKernelGGL(k_csr_scan, dim3(1), dim3(1024), 0, s, scratch, nb, total);
    hipLaunchKernelGGL(k_csr_fill, dim3(nb), dim3(CSR_ROWS), 0, s, A, Nx, Ny, Nz, N, rowptr, scratch, col, val, cap);
    long long h = -1;
    if (hipMemcpyAsync(&h, total, sizeof h, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return SD_CSR_EDEVICE;
    return h;
}

// ------------------------------------------------------------ C ABI
static bool csr_args_ok(int Nx, int Ny, int Nz, const double *const *diag, const int *rowptr) {
    if (Nx < 1 || Ny < 1 || Nz < 1 || !diag || !rowptr) return false;
    if ((long)Nx * Ny * Nz > 0x7ffffffeL) return false;
    for (int q = 1; q < 7; q++) if (!diag[q]) return false;
    return true;
}

// ------------------------------------------------ allocation-free variant
// The scan partials of sdx_csr_from_stencil_noalloc live in this file-scope
// device array: one int per 1024 rows plus the 8-byte total, so rows up to
// CSR_NOALLOC_ROWS.  Every GPU the library runs on holds its own copy until
// the process ends.  In host code `scratch` is only a symbol: its device
// address on the current GPU comes from hipGetSymbolAddress, looked up once
// per device and kept in g_scratch_ptr.  Nothing in it needs clearing between
// calls (the count pass writes every partial the scan reads, the scan writes
// the total), but calls that overlap on one GPU race on it.
#define CSR_NOALLOC_ROWS (64L << 20)
#define CSR_NOALLOC_WORDS ((int)(CSR_NOALLOC_ROWS / CSR_ROWS) + 4)
static __device__ __align__(256) int scratch[CSR_NOALLOC_WORDS];
#define CSR_MAX_DEVICES 64
static int *g_scratch_ptr[CSR_MAX_DEVICES];
static std::mutex g_scratch_mu;

// device address of `scratch` on the current GPU (null on failure)
static int *scratch_on_current_device() {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= CSR_MAX_DEVICES) return nullptr;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        int *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress((void **)&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) {
            fprintf(stderr, "semisim-hip: scratch of the allocation-free CSR export not found on device %d: %s\n", dev,
                    hipGetErrorString(e));
            return nullptr;
        }
        g_scratch_ptr[dev] = ptr;
    }
    return g_scratch_ptr[dev];
}

extern "C" {
// device memory for callers without a HIP runtime of their own (tests, the
// ctypes front end): allocate, free and copy (dir 1 host->device, 2
// device->host, 3 device->device); 0 / null on failure
void *sdx_dev_alloc(long long bytes) {
    void *p = nullptr;
    if (bytes <= 0 || hipMalloc(&p, (size_t)bytes) != hipSuccess) return nullptr;
    return p;
}
void sdx_dev_free(void *p) { if (p) (void)hipFree(p); }
int sdx_dev_copy(void *dst, const void *src, long long bytes, int dir) {
    if (bytes < 0 || (!dst || !src) || dir < 1 || dir > 3) return 0;
    const hipMemcpyKind k = dir == 1 ? hipMemcpyHostToDevice : (dir == 2 ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice);
    return hipMemcpy(dst, src, (size_t)bytes, k) == hipSuccess;
}
// CSR of a 7-point stencil operator, all arrays in device memory (see
// include/semisim_hip_ext.h).  Allocates its scratch (one int per 1024 rows)
// for the call; runs on the null stream.
long long sdx_csr_from_stencil(int Nx, int Ny, int Nz, const double *const *diag, int *rowptr, int *col, double *val,
                               long long cap) {
    if (!csr_args_ok(Nx, Ny, Nz, diag, rowptr) || cap < 0 || (cap > 0 && (!col || !val))) {
        fprintf(stderr, "semisim-hip: sdx_csr_from_stencil: bad arguments\n");
        return SD_CSR_EARG;
    }
    const long N = (long)Nx * Ny * Nz;
    int *scratch = nullptr;
    if (hipMalloc(&scratch, (size_t)csr_scratch_words(N) * sizeof(int)) != hipSuccess) return SD_CSR_EDEVICE;
    long long nnz = csr_from_stencil(nullptr, Nx, Ny, Nz, diag, rowptr, col, val, cap, scratch);
    if (hipFree(scratch) != hipSuccess && nnz >= 0) nnz = SD_CSR_EDEVICE;
    if (nnz > 0x7fffffffLL) {
        fprintf(stderr, "semisim-hip: sdx_csr_from_stencil: %lld nonzeros exceed an int row pointer\n", nnz);
        return SD_CSR_ERANGE;
    }
    return nnz;
}

// The same export without allocating: its scratch is the file-scope array
// above (rows up to CSR_NOALLOC_ROWS), the caller passes the inputs and
// outputs.  No hipMalloc, hipFree or allocating library call.
long long sdx_csr_from_stencil_noalloc(int Nx, int Ny, int Nz, const double *const *diag, int *rowptr, int *col,
                                       double *val, long long cap) {
    if (!csr_args_ok(Nx, Ny, Nz, diag, rowptr) || cap < 0 || (cap > 0 && (!col || !val))) {
        fprintf(stderr, "semisim-hip: sdx_csr_from_stencil_noalloc: bad arguments\n");
        return SD_CSR_EARG;
    }
    const long N = (long)Nx * Ny * Nz;
    if (N > CSR_NOALLOC_ROWS) {
        fprintf(stderr, "semisim-hip: sdx_csr_from_stencil_noalloc: %ld rows exceed its fixed scratch (at most %ld)\n",
                N, (long)CSR_NOALLOC_ROWS);
        return SD_CSR_ESIZE;
    }
    int *ptr = scratch_on_current_device();
    if (!ptr) return SD_CSR_EDEVICE;
    const long long nnz = csr_from_stencil(nullptr, Nx, Ny, Nz, diag, rowptr, col, val, cap, ptr);
    if (nnz > 0x7fffffffLL) {
        fprintf(stderr, "semisim-hip: sdx_csr_from_stencil_noalloc: %lld nonzeros exceed an int row pointer\n", nnz);
        return SD_CSR_ERANGE;
    }
    return nnz;
}
}
