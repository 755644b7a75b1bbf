// sd_reduce.h - deterministic device reductions shared by the kernel files.
#pragma once
#include "sd_linalg.h"

// Two-stage deterministic reduction: every block writes its partial (NV
// values) to rb.part; a one-workgroup finalize kernel (red_final) sums them
// in index order and can apply a scalar update of the Krylov coefficients.
// No device-scope fences or atomics: the kernel boundary orders the stages.
template <int NV, bool MAXR>
__device__ __forceinline__ void red_partial(double (&acc)[NV], RedBuf rb) {
    __shared__ double lds[NV][SD_BLOCK / 64];
    int w = threadIdx.x >> 6, ln = threadIdx.x & 63;
    for (int v = 0; v < NV; v++) {
        double x = MAXR ? sd_wave_max(acc[v]) : sd_wave_sum(acc[v]);
        if (ln == 0) lds[v][w] = x;
    }
    __syncthreads();
    if (threadIdx.x < NV) {
        int v = threadIdx.x;
        double s = lds[v][0];
        for (int q = 1; q < SD_BLOCK / 64; q++) s = MAXR ? sd_nanmax(s, lds[v][q]) : s + lds[v][q];
        rb.part[v * SD_RED_MAXB + blockIdx.x] = s;
    }
}
// finalize nv values over nb blocks into rb.scal[slot0..]; then scalar op `post` (SOP_*, -1 none)
void red_final(hipStream_t s, RedBuf rb, int nv, int nb, int slot0, int is_max, int post = -1);

static inline int red_grid(int n) {
    int nb = sd_nblocks(n);
    return nb < 1 ? 1 : (nb > SD_RED_MAXB ? SD_RED_MAXB : nb);
}
