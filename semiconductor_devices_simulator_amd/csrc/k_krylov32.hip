// k_krylov32.hip - kernels of the mixed-precision method's inner Krylov
// solves.
//
// The FP32 method solves each linear system as FP64 defect correction around
// an FP32 Krylov solve: the outer residual b - A x is formed in FP64 with the
// FP64 operator, the correction d of A d = r is found by PCG or BiCGSTAB whose
// vectors are FP32 and live in the level-0 storage of the FP32 multigrid
// (parity-split x order, the order the V-cycle reads and writes), and x += d
// is applied in FP64.  Every dot product is accumulated in FP64 and reduced
// by the deterministic two-stage reduction; the Krylov scalars stay FP64 in
// the device slot array.  All kernels here are HBM-bound streams.
#include "sd_reduce.h"

// storage position of x index i in a row of a level-0 array
__device__ __forceinline__ int k32_xpos(const L0Geo &g, int i) {
    return g.px ? (i & 1) * ((g.Nx + 1) >> 1) + (i >> 1) : i;
}
// index of a neighbour load: the neighbour's, or the node's own where there
// is none, so that every load of a row is issued before the first use (the
// empty asm keeps the compiler from folding the clamped load into the node's
// own load and moving the neighbour's into a branch)
__device__ __forceinline__ int k32_nb(bool has, int nb, int self) {
    int v = has ? nb : self;
    asm("" : "+v"(v));
    return v;
}
// x index of storage position p of a row
__device__ __forceinline__ int k32_xidx(const L0Geo &g, int p) {
    if (!g.px) return p;
    const int h = (g.Nx + 1) >> 1;
    return p < h ? 2 * p : 2 * (p - h) + 1;
}

// y = A x on the level-0 storage with the FP32 operator (unit diagonal not
// read when UD); slot <- y.(u0, or y when u0 is null), slot+1 <- y.u1 (NV 2).
// Loop over storage positions; NodeWalk's i is the position within the row.
// Two nodes per iteration, both rows' loads issued before either is summed;
// the operator is streamed with non-temporal loads (read once per product,
// so it need not displace the multigrid's data from the caches).
template <bool UD, int NV>
__global__ void __launch_bounds__(SD_BLOCK) k32_spmv_dots(L0Geo g, Sten7<float> A, const float *__restrict__ x,
                                                          float *__restrict__ y, const float *__restrict__ u0,
                                                          const float *__restrict__ u1, RedBuf rb) {
    const int Nx = g.Nx, Ny = g.Ny, Nz = g.Nz, N = Nx * Ny * Nz, NxNy = Nx * Ny;
    double acc[NV] = {};
    int k0, st, k1;
    sd_node_loop(N, k0, st, k1);
    auto ld = [](const float *p, int q) -> float { return __builtin_nontemporal_load(p + q); };
    auto row = [&](int q, const NodeWalk &w) -> float {
        const int i = k32_xidx(g, w.i), r0 = q - w.i, j = w.j, l = w.l;
        const bool h0 = i > 0, h1 = i < Nx - 1, h2 = j > 0, h3 = j < Ny - 1, h4 = l > 0, h5 = l < Nz - 1;
        const float xq = x[q];
        const float a0 = ld(A.xm, q), a1 = ld(A.xp, q), a2 = ld(A.ym, q), a3 = ld(A.yp, q);
        const float a4 = ld(A.zm, q), a5 = ld(A.zp, q);
        const float x0 = x[k32_nb(h0, r0 + k32_xpos(g, i - 1), q)], x1 = x[k32_nb(h1, r0 + k32_xpos(g, i + 1), q)];
        const float x2 = x[k32_nb(h2, q - Nx, q)], x3 = x[k32_nb(h3, q + Nx, q)];
        const float x4 = x[k32_nb(h4, q - NxNy, q)], x5 = x[k32_nb(h5, q + NxNy, q)];
        float s = UD ? xq : ld(A.c, q) * xq;
        s = h0 ? s + a0 * x0 : s;
        s = h1 ? s + a1 * x1 : s;
        s = h2 ? s + a2 * x2 : s;
        s = h3 ? s + a3 * x3 : s;
        s = h4 ? s + a4 * x4 : s;
        s = h5 ? s + a5 * x5 : s;
        return s;
    };
    NodeWalk nw(k0, st, Nx, Ny);
    for (int q = k0; q < k1; q += 2 * st) {
        const int qb = q + st;
        NodeWalk nb = nw;
        nb.next();
        const bool hb = qb < k1;
        const int qc = hb ? qb : q;
        const float ua = u0 ? u0[q] : 0.f, ub = u0 ? u0[qc] : 0.f;
        const float va = NV == 2 ? u1[q] : 0.f, vb = NV == 2 ? u1[qc] : 0.f;
        const float sa = row(q, nw);
        const float sb = hb ? row(qb, nb) : 0.f;
        y[q] = sa;
        acc[0] += (double)sa * (u0 ? (double)ua : (double)sa);
        if constexpr (NV == 2) acc[1] += (double)sa * (double)va;
        if (hb) {
            y[qb] = sb;
            acc[0] += (double)sb * (u0 ? (double)ub : (double)sb);
            if constexpr (NV == 2) acc[1] += (double)sb * (double)vb;
        }
        nb.next();
        nw = nb;
    }
    red_partial<NV, false>(acc, rb);
}
void k32_spmv_dots_launch(hipStream_t s, const L0Geo &g, Sten7<float> A, bool unit, const float *x, float *y,
                          const float *u0, const float *u1, int nv, RedBuf rb, int slot, int post) {
    const int n = g.Nx * g.Ny * g.Nz, nb = red_grid(n);
    if (nv == 2) {
        if (unit) k32_spmv_dots<true, 2><<<nb, SD_BLOCK, 0, s>>>(g, A, x, y, u0, u1, rb);
        else k32_spmv_dots<false, 2><<<nb, SD_BLOCK, 0, s>>>(g, A, x, y, u0, u1, rb);
    } else {
        if (unit) k32_spmv_dots<true, 1><<<nb, SD_BLOCK, 0, s>>>(g, A, x, y, u0, u1, rb);
        else k32_spmv_dots<false, 1><<<nb, SD_BLOCK, 0, s>>>(g, A, x, y, u0, u1, rb);
    }
    red_final(s, rb, nv, nb, slot, 0, post);
}

// Outer residual: r = b - A x in FP64 (natural order, FP64 operator), r.r
// accumulated in FP64, r stored as FP32 at its level-0 position.  With rh:
// the inner BiCGSTAB's start as well (rh = r, p = v = d = 0).
template <bool UD>
__global__ void __launch_bounds__(SD_BLOCK) k32_resid(L0Geo g, Sten7<double> A, const double *__restrict__ b,
                                                      const double *__restrict__ x, float *__restrict__ r,
                                                      float *__restrict__ rh, float *__restrict__ p,
                                                      float *__restrict__ v, float *__restrict__ d, RedBuf rb) {
    const int Nx = g.Nx, Ny = g.Ny, Nz = g.Nz, N = Nx * Ny * Nz, NxNy = Nx * Ny;
    double acc[1] = {0.0};
    int k0, st, k1;
    sd_node_loop(N, k0, st, k1);
    NodeWalk nw(k0, st, Nx, Ny);
    for (int k = k0; k < k1; k += st, nw.next()) {
        const int i = nw.i, j = nw.j, l = nw.l;
        double s = UD ? x[k] : A.c[k] * x[k];
        if (i > 0) s += A.xm[k] * x[k - 1];
        if (i < Nx - 1) s += A.xp[k] * x[k + 1];
        if (j > 0) s += A.ym[k] * x[k - Nx];
        if (j < Ny - 1) s += A.yp[k] * x[k + Nx];
        if (l > 0) s += A.zm[k] * x[k - NxNy];
        if (l < Nz - 1) s += A.zp[k] * x[k + NxNy];
        const double y = b[k] - s;
        acc[0] += y * y;
        const int q = k - i + k32_xpos(g, i);
        const float yf = (float)y;
        r[q] = yf;
        if (rh) { rh[q] = yf; p[q] = 0.f; v[q] = 0.f; d[q] = 0.f; }
    }
    red_partial<1, false>(acc, rb);
}
void k32_resid_launch(hipStream_t s, const L0Geo &g, Sten7<double> A, bool unit, const double *b, const double *x,
                      float *r, float *rh, float *p, float *v, float *d, RedBuf rb, int slot) {
    const int n = g.Nx * g.Ny * g.Nz, nb = red_grid(n);
    if (unit) k32_resid<true><<<nb, SD_BLOCK, 0, s>>>(g, A, b, x, r, rh, p, v, d, rb);
    else k32_resid<false><<<nb, SD_BLOCK, 0, s>>>(g, A, b, x, r, rh, p, v, d, rb);
    red_final(s, rb, 1, nb, slot, 0);
}

// x += d: the FP32 correction at its level-0 position added in FP64
__global__ void k32_add_out(L0Geo g, double *__restrict__ x, const float *__restrict__ d) {
    const int n = g.Nx * g.Ny * g.Nz;
    const int k = blockIdx.x * SD_BLOCK + threadIdx.x;
    if (k >= n) return;
    const int i = k % g.Nx;
    x[k] += (double)d[k - i + k32_xpos(g, i)];
}
void k32_add_out_launch(hipStream_t s, const L0Geo &g, double *x, const float *d) {
    k32_add_out<<<sd_nblocks((long)g.Nx * g.Ny * g.Nz), SD_BLOCK, 0, s>>>(g, x, d);
}

// slot <- a.b (FP32 vectors, FP64 accumulation)
__global__ void __launch_bounds__(SD_BLOCK) k32_dot(const float *__restrict__ a, const float *__restrict__ b, int n,
                                                    RedBuf rb) {
    double acc[1] = {0.0};
    for (int k = blockIdx.x * SD_BLOCK + threadIdx.x; k < n; k += gridDim.x * SD_BLOCK)
        acc[0] += (double)a[k] * (double)b[k];
    red_partial<1, false>(acc, rb);
}
void k32_dot_launch(hipStream_t s, const float *a, const float *b, int n, RedBuf rb, int slot, int post) {
    const int nb = red_grid(n);
    k32_dot<<<nb, SD_BLOCK, 0, s>>>(a, b, n, rb);
    red_final(s, rb, 1, nb, slot, 0, post);
}

__global__ void k32_fill(float *a, float v, int n) {
    const int k = blockIdx.x * SD_BLOCK + threadIdx.x;
    if (k < n) a[k] = v;
}
void k32_fill_launch(hipStream_t s, float *a, float v, int n) { k32_fill<<<sd_nblocks(n), SD_BLOCK, 0, s>>>(a, v, n); }
__global__ void k32_copy(float *__restrict__ d, const float *__restrict__ a, int n) {
    const int k = blockIdx.x * SD_BLOCK + threadIdx.x;
    if (k < n) d[k] = a[k];
}
void k32_copy_launch(hipStream_t s, float *d, const float *a, int n) {
    k32_copy<<<sd_nblocks(n), SD_BLOCK, 0, s>>>(d, a, n);
}

// ------------------------------------------------------------------ PCG
// d += al p, r -= al q (skipped when p.q <= 0, as in the FP64 solver)
__global__ void k32_pcg_xr(float *__restrict__ d, float *__restrict__ r, const float *__restrict__ p,
                           const float *__restrict__ q, int n, const double *sc) {
    const int k = blockIdx.x * SD_BLOCK + threadIdx.x;
    if (k >= n || !(sc[S_PQ] > 0.)) return;
    const double al = sc[S_AL];
    d[k] = (float)((double)d[k] + al * (double)p[k]);
    r[k] = (float)((double)r[k] - al * (double)q[k]);
}
void k32_pcg_xr_launch(hipStream_t s, float *d, float *r, const float *p, const float *q, int n, RedBuf rb) {
    k32_pcg_xr<<<sd_nblocks(n), SD_BLOCK, 0, s>>>(d, r, p, q, n, rb.scal);
}
// p = z + be p
__global__ void k32_pcg_p(float *__restrict__ p, const float *__restrict__ z, int n, const double *sc) {
    const int k = blockIdx.x * SD_BLOCK + threadIdx.x;
    if (k >= n) return;
    p[k] = (float)((double)z[k] + sc[S_BE] * (double)p[k]);
}
void k32_pcg_p_launch(hipStream_t s, float *p, const float *z, int n, RedBuf rb) {
    k32_pcg_p<<<sd_nblocks(n), SD_BLOCK, 0, s>>>(p, z, n, rb.scal);
}

// ------------------------------------------------------------- BiCGSTAB
// p = r + be (p - om v)
__global__ void k32_bi_p(float *__restrict__ p, const float *__restrict__ r, const float *__restrict__ v, int n,
                         const double *sc) {
    const int k = blockIdx.x * SD_BLOCK + threadIdx.x;
    if (k >= n) return;
    const double be = sc[S_BE], om = sc[S_OM];
    p[k] = (float)((double)r[k] + be * ((double)p[k] - om * (double)v[k]));
}
void k32_bi_p_launch(hipStream_t s, float *p, const float *r, const float *v, int n, RedBuf rb) {
    k32_bi_p<<<sd_nblocks(n), SD_BLOCK, 0, s>>>(p, r, v, n, rb.scal);
}
// r -= al v (r becomes s); s.s -> slot
__global__ void __launch_bounds__(SD_BLOCK) k32_bi_s(float *__restrict__ r, const float *__restrict__ v, int n,
                                                     RedBuf rb) {
    const double al = rb.scal[S_AL];
    double acc[1] = {0.0};
    for (int k = blockIdx.x * SD_BLOCK + threadIdx.x; k < n; k += gridDim.x * SD_BLOCK) {
        const float y = (float)((double)r[k] - al * (double)v[k]);
        r[k] = y;
        acc[0] += (double)y * (double)y;
    }
    red_partial<1, false>(acc, rb);
}
void k32_bi_s_launch(hipStream_t s, float *r, const float *v, int n, RedBuf rb, int slot) {
    const int nb = red_grid(n);
    k32_bi_s<<<nb, SD_BLOCK, 0, s>>>(r, v, n, rb);
    red_final(s, rb, 1, nb, slot, 0);
}
// d += al ph
__global__ void k32_bi_x1(float *__restrict__ d, const float *__restrict__ ph, int n, const double *sc) {
    const int k = blockIdx.x * SD_BLOCK + threadIdx.x;
    if (k < n) d[k] = (float)((double)d[k] + sc[S_AL] * (double)ph[k]);
}
void k32_bi_x1_launch(hipStream_t s, float *d, const float *ph, int n, RedBuf rb) {
    k32_bi_x1<<<sd_nblocks(n), SD_BLOCK, 0, s>>>(d, ph, n, rb.scal);
}
// d += al ph + om sh; r -= om t; r.r and rh.r, then the next beta
__global__ void __launch_bounds__(SD_BLOCK) k32_bi_xr2(float *__restrict__ d, float *__restrict__ r,
                                                       const float *__restrict__ ph, const float *__restrict__ sh,
                                                       const float *__restrict__ t, const float *__restrict__ rh,
                                                       int n, RedBuf rb) {
    const double al = rb.scal[S_AL], om = rb.scal[S_OM];
    double acc[2] = {0.0, 0.0};
    for (int k = blockIdx.x * SD_BLOCK + threadIdx.x; k < n; k += gridDim.x * SD_BLOCK) {
        d[k] = (float)((double)d[k] + al * (double)ph[k] + om * (double)sh[k]);
        const float y = (float)((double)r[k] - om * (double)t[k]);
        r[k] = y;
        acc[0] += (double)y * (double)y;
        acc[1] += (double)rh[k] * (double)y;
    }
    red_partial<2, false>(acc, rb);
}
void k32_bi_xr2_launch(hipStream_t s, float *d, float *r, const float *ph, const float *sh, const float *t,
                       const float *rh, int n, RedBuf rb) {
    const int nb = red_grid(n);
    k32_bi_xr2<<<nb, SD_BLOCK, 0, s>>>(d, r, ph, sh, t, rh, n, rb);
    red_final(s, rb, 2, nb, S_TMP2, 0, SOP_BI_NEXT_BE);
}
// shadow restart: rh = r, p = v = 0
__global__ void k32_bi_restart(float *rh, float *p, float *v, const float *r, int n) {
    const int k = blockIdx.x * SD_BLOCK + threadIdx.x;
    if (k < n) { rh[k] = r[k]; p[k] = 0.f; v[k] = 0.f; }
}
void k32_bi_restart_launch(hipStream_t s, float *rh, float *p, float *v, const float *r, int n) {
    k32_bi_restart<<<sd_nblocks(n), SD_BLOCK, 0, s>>>(rh, p, v, r, n);
}
