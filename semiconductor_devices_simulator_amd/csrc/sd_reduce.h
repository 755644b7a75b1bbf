// sd_reduce.h - deterministic device reductions shared by the kernel files.
#pragma once
#include "sd_linalg.h"

// Last-block-done reduction: block partials -> rb.part, the last block to
// arrive sums them in index order (bit-reproducible for a fixed grid).
template <int NV, bool MAXR>
__device__ __forceinline__ void red_finish(double (&acc)[NV], RedBuf rb, int slot0) {
    __shared__ double lds[NV][SD_BLOCK / 64];
    __shared__ int amlast;
    int w = threadIdx.x >> 6, ln = threadIdx.x & 63;
    for (int v = 0; v < NV; v++) {
        double x = MAXR ? sd_wave_max(acc[v]) : sd_wave_sum(acc[v]);
        if (ln == 0) lds[v][w] = x;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int v = 0; v < NV; v++) {
            double s = lds[v][0];
            for (int q = 1; q < SD_BLOCK / 64; q++) s = MAXR ? sd_nanmax(s, lds[v][q]) : s + lds[v][q];
            rb.part[v * SD_RED_MAXB + blockIdx.x] = s;
        }
        __threadfence();
        unsigned t = __hip_atomic_fetch_add(rb.cnt, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        amlast = (t == gridDim.x - 1);
    }
    __syncthreads();
    if (!amlast) return;
    __threadfence();
    for (int v = 0; v < NV; v++) {
        double s = 0.0;
        for (int b = threadIdx.x; b < (int)gridDim.x; b += SD_BLOCK) {
            double x = __hip_atomic_load(&rb.part[v * SD_RED_MAXB + b], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            s = MAXR ? sd_nanmax(s, x) : s + x;
        }
        s = MAXR ? sd_wave_max(s) : sd_wave_sum(s);
        __syncthreads();
        if (ln == 0) lds[v][w] = s;
        __syncthreads();
        if (threadIdx.x == 0) {
            double r = lds[v][0];
            for (int q = 1; q < SD_BLOCK / 64; q++) r = MAXR ? sd_nanmax(r, lds[v][q]) : r + lds[v][q];
            rb.scal[slot0 + v] = r;
        }
    }
    if (threadIdx.x == 0) __hip_atomic_store(rb.cnt, 0u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

static inline int red_grid(int n) {
    int nb = sd_nblocks(n);
    return nb < 1 ? 1 : (nb > SD_RED_MAXB ? SD_RED_MAXB : nb);
}
