// k_linalg.hip - Krylov and multigrid kernels of the MI355X drift-diffusion
// core (gfx950).  Every kernel is HBM-bandwidth bound: 7-point stencils as
// SoA diagonals, one node per lane, coalesced along x; reductions are
// deterministic (wavefront shuffle, LDS, fixed-order partial sums).
//
// Reference counterparts (semiconductor_3d.c): matvec_L :1975-1993,
// galerkin :2003-2027, restrict_sum/prolong_add :2028-2048, dense LU
// :2050-2081, mg_weights :2085-2117, vcycle :2151-2165, dot :2174-2179,
// pcg :2182-2210, bicgstab_core :2234-2320.  The chunked D-ILU smoother of the
// reference is replaced by red-black Gauss-Seidel (symmetric ordering in the
// V-cycle), which changes iteration counts but not the solution.
#include <type_traits>
#include "sd_reduce.h"

#define PH_NONE (-1e300)

// ------------------------------------------------------------ reductions

__global__ void __launch_bounds__(SD_BLOCK) k_dot(const double *a, const double *b, int n, RedBuf rb, int slot) {
    double acc[1] = {0.0};
    for (int k = blockIdx.x * SD_BLOCK + threadIdx.x; k < n; k += gridDim.x * SD_BLOCK) acc[0] += b ? a[k] * b[k] : a[k];
    red_partial<1, false>(acc, rb);
}
__global__ void __launch_bounds__(SD_BLOCK) k_dot2(const double *a, const double *b, const double *c, int n,
                                                   RedBuf rb, int slot) {
    double acc[2] = {0.0, 0.0};
    for (int k = blockIdx.x * SD_BLOCK + threadIdx.x; k < n; k += gridDim.x * SD_BLOCK) {
        double x = a[k];
        acc[0] += x * b[k];
        acc[1] += x * c[k];
    }
    red_partial<2, false>(acc, rb);
}
__global__ void __launch_bounds__(SD_BLOCK) k_maxabs(const double *a, const double *b, int n, RedBuf rb, int slot) {
    double acc[1] = {0.0};
    for (int k = blockIdx.x * SD_BLOCK + threadIdx.x; k < n; k += gridDim.x * SD_BLOCK) {
        double x = b ? fabs(a[k] - b[k]) : fabs(a[k]);
        acc[0] = sd_nanmax(acc[0], x);
    }
    red_partial<1, true>(acc, rb);
}

void sd_dot(hipStream_t s, const double *a, const double *b, int n, RedBuf rb, int slot, int post) {
    int nb = red_grid(n);
    k_dot<<<nb, SD_BLOCK, 0, s>>>(a, b, n, rb, slot);
    red_final(s, rb, 1, nb, slot, 0, post);
}
void sd_dot2(hipStream_t s, const double *a, const double *b, const double *c, int n, RedBuf rb, int sab, int sac,
             int post) {
    // slots must be adjacent: sac == sab + 1
    (void)sac;
    int nb = red_grid(n);
    k_dot2<<<nb, SD_BLOCK, 0, s>>>(a, b, c, n, rb, sab);
    red_final(s, rb, 2, nb, sab, 0, post);
}
void sd_maxabs(hipStream_t s, const double *a, int n, RedBuf rb, int slot) {
    int nb = red_grid(n);
    k_maxabs<<<nb, SD_BLOCK, 0, s>>>(a, nullptr, n, rb, slot);
    red_final(s, rb, 1, nb, slot, 1);
}
void sd_maxabsdiff(hipStream_t s, const double *a, const double *b, int n, RedBuf rb, int slot) {
    int nb = red_grid(n);
    k_maxabs<<<nb, SD_BLOCK, 0, s>>>(a, b, n, rb, slot);
    red_final(s, rb, 1, nb, slot, 1);
}

int g_spmv_nt = 1;           // 1: the Krylov SpMV streams the operator with non-temporal loads

// ------------------------------------------------ level-0 staging
// storage position of natural node k in the staged level-0 arrays
__device__ __forceinline__ int st_pos(const Stage0 &st, int k) {
    if (!st.px) return k;
    const int i = k % st.Nx;
    return k - i + (i & 1) * ((st.Nx + 1) >> 1) + (i >> 1);
}
__device__ __forceinline__ void st_put(const Stage0 &st, int k, double v) {
    const int q = st_pos(st, k);
    if (st.f32) ((float *)st.b)[q] = (float)v;
    else ((double *)st.b)[q] = v;
}
// the staged solution's storage type is a template argument, so that no
// load is issued for the type not in use
template <class TS> __device__ __forceinline__ double st_get(const Stage0 &st, int q) {
    return (double)((const TS *)st.x)[q];
}
template <class TS>
__global__ void __launch_bounds__(SD_BLOCK) k_dot_stage(const double *r, double *z, Stage0 st, int n, RedBuf rb) {
    double acc[1] = {0.0};
    for (int k = blockIdx.x * SD_BLOCK + threadIdx.x; k < n; k += gridDim.x * SD_BLOCK) {
        const double zk = st_get<TS>(st, st_pos(st, k));
        if (z) z[k] = zk;
        acc[0] += r[k] * zk;
    }
    red_partial<1, false>(acc, rb);
}
void sd_dot_stage(hipStream_t s, const double *r, double *z, Stage0 st, int n, RedBuf rb, int slot, int post) {
    int nb = red_grid(n);
    if (st.f32) k_dot_stage<float><<<nb, SD_BLOCK, 0, s>>>(r, z, st, n, rb);
    else k_dot_stage<double><<<nb, SD_BLOCK, 0, s>>>(r, z, st, n, rb);
    red_final(s, rb, 1, nb, slot, 0, post);
}

// -------------------------------------------------------- vector kernels
__global__ void k_fill(double *a, double v, int n) {
    int k = blockIdx.x * SD_BLOCK + threadIdx.x;
    if (k < n) a[k] = v;
}
__global__ void k_copy(double *d, const double *s, int n) {
    int k = blockIdx.x * SD_BLOCK + threadIdx.x;
    if (k < n) d[k] = s[k];
}
__global__ void k_axpy(double *y, const double *x, double a, int n) {
    int k = blockIdx.x * SD_BLOCK + threadIdx.x;
    if (k < n) y[k] += a * x[k];
}
void sd_fill(hipStream_t s, double *a, double v, int n) { k_fill<<<sd_nblocks(n), SD_BLOCK, 0, s>>>(a, v, n); }
void sd_copy(hipStream_t s, double *d, const double *src, int n) {
    hipMemcpyAsync(d, src, (size_t)n * 8, hipMemcpyDeviceToDevice, s);
}
void sd_axpy(hipStream_t s, double *y, const double *x, double a, int n) {
    k_axpy<<<sd_nblocks(n), SD_BLOCK, 0, s>>>(y, x, a, n);
}

// -------------------------------------------------------------- stencil
// s + a xv (nb_add) or s - a xv (nb_sub) where the neighbour exists, s
// otherwise.  The callers load a and xv whether or not the neighbour exists
// (its index clamped to the node itself), so that no load sits inside a
// branch: a load whose value is used in a conditional block makes the
// compiler wait for it inside the block, and the loads of a row then
// complete one after another instead of together.  The select leaves the
// arithmetic (one fused multiply-add where the neighbour exists) unchanged.
// Index of such a neighbour load: the neighbour's, or the node's own where
// there is none.  The empty asm hides the value from the compiler, which
// would otherwise fold the clamped load into the node's own load of the same
// array and move the neighbour's back into a branch.
__device__ __forceinline__ int nb_idx(bool has, int nb, int self) {
    int v = has ? nb : self;
    asm("" : "+v"(v));
    return v;
}
template <class V> __device__ __forceinline__ V nb_add(bool has, V s, V a, V xv) { return has ? s + a * xv : s; }
template <class V> __device__ __forceinline__ V nb_sub(bool has, V s, V a, V xv) { return has ? s - a * xv : s; }
template <class T, class V>
__device__ __forceinline__ V sten_apply(const Sten7<T> &A, const V *x, int k, int i, int j, int l,
                                        int Nx, int Ny, int Nz, bool unitdiag) {
    int NxNy = Nx * Ny;
    const bool h0 = i > 0, h1 = i < Nx - 1, h2 = j > 0, h3 = j < Ny - 1, h4 = l > 0, h5 = l < Nz - 1;
    V s = unitdiag ? (V)x[k] : (V)A.c[k] * x[k];
    s = nb_add<V>(h0, s, (V)A.xm[k], x[nb_idx(h0, k - 1, k)]);
    s = nb_add<V>(h1, s, (V)A.xp[k], x[nb_idx(h1, k + 1, k)]);
    s = nb_add<V>(h2, s, (V)A.ym[k], x[nb_idx(h2, k - Nx, k)]);
    s = nb_add<V>(h3, s, (V)A.yp[k], x[nb_idx(h3, k + Nx, k)]);
    s = nb_add<V>(h4, s, (V)A.zm[k], x[nb_idx(h4, k - NxNy, k)]);
    s = nb_add<V>(h5, s, (V)A.zp[k], x[nb_idx(h5, k + NxNy, k)]);
    return s;
}

template <class T>
__global__ void __launch_bounds__(SD_BLOCK) k_spmv(int Nx, int Ny, int Nz, Sten7<T> A, const T *x, const T *b, T *y) {
    int N = Nx * Ny * Nz;
    int k = sd_tile() * SD_BLOCK + threadIdx.x;
    if (k >= N) return;
    int NxNy = Nx * Ny, l = k / NxNy, r = k - l * NxNy, j = r / Nx, i = r - j * Nx;
    T s = sten_apply<T, T>(A, x, k, i, j, l, Nx, Ny, Nz, false);
    y[k] = b ? b[k] - s : s;
}
void sd_spmv(hipStream_t s, const Geo &g, Sten7<double> A, const double *x, const double *b, double *y) {
    k_spmv<double><<<sd_nblocks(g.N), SD_BLOCK, 0, s>>>(g.Nx, g.Ny, g.Nz, A, x, b, y);
}
template <class T> void sd_spmv_T(hipStream_t s, const Geo &g, Sten7<T> A, const T *x, const T *b, T *y) {
    k_spmv<T><<<sd_nblocks(g.N), SD_BLOCK, 0, s>>>(g.Nx, g.Ny, g.Nz, A, x, b, y);
}
template void sd_spmv_T<double>(hipStream_t, const Geo &, Sten7<double>, const double *, const double *, double *);
template void sd_spmv_T<float>(hipStream_t, const Geo &, Sten7<float>, const float *, const float *, float *);

// y = A x fused with the dots y.(u0 or y) and y.u1 (Krylov: q = A p with p.q;
// v = A ph with rh.v; t = A sh with t.t and t.r)
// TS != void: x is the staged level-0 solution `in` (storage order of the
// level, type TS), copied to xw in FP64 natural order on the way
// NT: the operator is read with non-temporal loads (streamed once per
// product, so it need not displace the multigrid data from the caches)
// UD (A.c == nullptr): unit diagonal (not read); a template argument, so that
// no branch separates the node's own load from its use
template <class TS, bool NT, bool UD>
__global__ void __launch_bounds__(SD_BLOCK) k_spmv_dots(int Nx, int Ny, int Nz, Sten7<double> A, const double *x,
                                                        double *y, const double *u0, const double *u1, int nv,
                                                        RedBuf rb, Stage0 in, double *xw) {
    int N = Nx * Ny * Nz, NxNy = Nx * Ny;
    double acc[2] = {0.0, 0.0};
    int k0, st, k1;
    sd_node_loop(N, k0, st, k1);
    auto ld = [](const double *p, int k) -> double {
        if constexpr (NT) return __builtin_nontemporal_load(p + k);
        else return p[k];
    };
    // (A x)_k, and x_k itself
    auto row = [&](int k, int i, int j, int l, double &xk) -> double {
        if constexpr (!std::is_void_v<TS>) {
            // row start is the same in both orders; x positions within the row differ
            const int h = (Nx + 1) >> 1, r0 = k - i;
            auto xp = [&](int ii) { return in.px ? (ii & 1) * h + (ii >> 1) : ii; };
            const int q = r0 + xp(i);
            xk = st_get<TS>(in, q);
            const bool h0 = i > 0, h1 = i < Nx - 1, h2 = j > 0, h3 = j < Ny - 1, h4 = l > 0, h5 = l < Nz - 1;
            double s = UD ? xk : ld(A.c, k) * xk;
            s = nb_add<double>(h0, s, ld(A.xm, k), st_get<TS>(in, nb_idx(h0, r0 + xp(i - 1), q)));
            s = nb_add<double>(h1, s, ld(A.xp, k), st_get<TS>(in, nb_idx(h1, r0 + xp(i + 1), q)));
            s = nb_add<double>(h2, s, ld(A.ym, k), st_get<TS>(in, nb_idx(h2, q - Nx, q)));
            s = nb_add<double>(h3, s, ld(A.yp, k), st_get<TS>(in, nb_idx(h3, q + Nx, q)));
            s = nb_add<double>(h4, s, ld(A.zm, k), st_get<TS>(in, nb_idx(h4, q - NxNy, q)));
            s = nb_add<double>(h5, s, ld(A.zp, k), st_get<TS>(in, nb_idx(h5, q + NxNy, q)));
            return s;
        } else {
            xk = 0.0;
            const bool h0 = i > 0, h1 = i < Nx - 1, h2 = j > 0, h3 = j < Ny - 1, h4 = l > 0, h5 = l < Nz - 1;
            double s = UD ? x[k] : ld(A.c, k) * x[k];
            s = nb_add<double>(h0, s, ld(A.xm, k), x[nb_idx(h0, k - 1, k)]);
            s = nb_add<double>(h1, s, ld(A.xp, k), x[nb_idx(h1, k + 1, k)]);
            s = nb_add<double>(h2, s, ld(A.ym, k), x[nb_idx(h2, k - Nx, k)]);
            s = nb_add<double>(h3, s, ld(A.yp, k), x[nb_idx(h3, k + Nx, k)]);
            s = nb_add<double>(h4, s, ld(A.zm, k), x[nb_idx(h4, k - NxNy, k)]);
            s = nb_add<double>(h5, s, ld(A.zp, k), x[nb_idx(h5, k + NxNy, k)]);
            return s;
        }
    };
    // two nodes per iteration, both computed before either is stored, so
    // that the loads of both are in flight together
    NodeWalk nw(k0, st, Nx, Ny);
    for (int k = k0; k < k1; k += 2 * st) {
        const int kb = k + st;
        NodeWalk nb = nw;
        nb.next();
        double xa, xb = 0.0, sb = 0.0;
        const bool hb = kb < k1;
        // the dot vectors' values are loaded with the rows' operands
        const int kc = hb ? kb : k;
        const double u0a = u0 ? u0[k] : 0.0, u0b = u0 ? u0[kc] : 0.0;
        const double va = nv > 1 ? u1[k] : 0.0, v1b = nv > 1 ? u1[kc] : 0.0;
        const double sa = row(k, nw.i, nw.j, nw.l, xa);
        if (hb) sb = row(kb, nb.i, nb.j, nb.l, xb);
        const double ua = u0 ? u0a : sa;
        double ub = 0.0, vb = 0.0;
        if (hb) { ub = u0 ? u0b : sb; vb = v1b; }
        y[k] = sa;
        if constexpr (!std::is_void_v<TS>) { if (xw) xw[k] = xa; }
        acc[0] += sa * ua;
        if (nv > 1) acc[1] += sa * va;
        if (hb) {
            y[kb] = sb;
            if constexpr (!std::is_void_v<TS>) { if (xw) xw[kb] = xb; }
            acc[0] += sb * ub;
            if (nv > 1) acc[1] += sb * vb;
        }
        nb.next();
        nw = nb;
    }
    red_partial<2, false>(acc, rb);
}
void sd_spmv_dots(hipStream_t s, const Geo &g, Sten7<double> A, const double *x, double *y, const double *u0,
                  const double *u1, int nv, RedBuf rb, int slot, int post, Stage0 in) {
    int nb = red_grid(g.N);
    double *xw = const_cast<double *>(x);
    auto go = [&](auto nt, auto ud) {
        constexpr bool NT = decltype(nt)::value, UD = decltype(ud)::value;
        if (in.x && in.f32)
            k_spmv_dots<float, NT, UD><<<nb, SD_BLOCK, 0, s>>>(g.Nx, g.Ny, g.Nz, A, x, y, u0, u1, nv, rb, in, xw);
        else if (in.x)
            k_spmv_dots<double, NT, UD><<<nb, SD_BLOCK, 0, s>>>(g.Nx, g.Ny, g.Nz, A, x, y, u0, u1, nv, rb, in, xw);
        else
            k_spmv_dots<void, NT, UD><<<nb, SD_BLOCK, 0, s>>>(g.Nx, g.Ny, g.Nz, A, x, y, u0, u1, nv, rb, in, nullptr);
    };
    const bool ud = A.c == nullptr;
    if (g_spmv_nt) { if (ud) go(std::true_type{}, std::true_type{}); else go(std::true_type{}, std::false_type{}); }
    else { if (ud) go(std::false_type{}, std::true_type{}); else go(std::false_type{}, std::false_type{}); }
    red_final(s, rb, nv, nb, slot, 0, post);
}

void sd_set_tile_mode_linalg(int m) {
    if (hipMemcpyToSymbol(HIP_SYMBOL(c_tile_mode), &m, sizeof(int)) != hipSuccess)
        fprintf(stderr, "semisim-hip: setting the tile order failed\n");
}

// rounding floor of a continuity solve, semiconductor_3d.c:2243-2265
__global__ void __launch_bounds__(SD_BLOCK) k_floor(int Nx, int Ny, int Nz, Sten7<double> A, const double *x,
                                                    const double *b, RedBuf rb, int slot) {
    int N = Nx * Ny * Nz, NxNy = Nx * Ny;
    double acc[1] = {0.0};
    int k0, st, k1;
    sd_node_loop(N, k0, st, k1);
    NodeWalk nw(k0, st, Nx, Ny);
    for (int k = k0; k < k1; k += st, nw.next()) {
        const int i = nw.i, j = nw.j, l = nw.l;
        double s = fabs(A.c[k] * x[k]) + fabs(b[k]);
        if (i > 0) s += fabs(A.xm[k] * x[k - 1]);
        if (i < Nx - 1) s += fabs(A.xp[k] * x[k + 1]);
        if (j > 0) s += fabs(A.ym[k] * x[k - Nx]);
        if (j < Ny - 1) s += fabs(A.yp[k] * x[k + Nx]);
        if (l > 0) s += fabs(A.zm[k] * x[k - NxNy]);
        if (l < Nz - 1) s += fabs(A.zp[k] * x[k + NxNy]);
        acc[0] += s * s;
    }
    red_partial<1, false>(acc, rb);
}
void sd_floor(hipStream_t s, const Geo &g, Sten7<double> A, const double *x, const double *b, RedBuf rb, int slot) {
    int nb = red_grid(g.N);
    k_floor<<<nb, SD_BLOCK, 0, s>>>(g.Nx, g.Ny, g.Nz, A, x, b, rb, slot);
    red_final(s, rb, 1, nb, slot, 0);
}

// ------------------------------------------------------- Krylov updates
__device__ __forceinline__ void scalar_op(int op, double *sc) {
    switch (op) {
    case SOP_PCG_AL: sc[S_AL] = (sc[S_PQ] > 0.) ? sc[S_RZ] / sc[S_PQ] : 0.0; break;
    case SOP_PCG_BE: sc[S_BE] = sc[S_TMP0] / sc[S_RZ]; sc[S_RZ] = sc[S_TMP0]; break;
    case SOP_BI_BE: sc[S_BE] = (sc[S_RHO1] / sc[S_RHO]) * (sc[S_AL] / sc[S_OM]); break;
    case SOP_BI_AL: sc[S_AL] = (fabs(sc[S_RV]) < 1e-300) ? 0.0 : sc[S_RHO1] / sc[S_RV]; break;
    case SOP_BI_OM: sc[S_OM] = sc[S_TT] > 0. ? sc[S_TS] / sc[S_TT] : 0.; break;
    case SOP_BI_NEXT: sc[S_RHO] = sc[S_RHO1]; break;
    case SOP_BI_NEXT_BE:       // r.r and rh.r of the new r, then the next beta
        sc[S_RR] = sc[S_TMP2];
        sc[S_RHO] = sc[S_RHO1];
        sc[S_RHO1] = sc[S_TMP3];
        sc[S_BE] = (sc[S_RHO1] / sc[S_RHO]) * (sc[S_AL] / sc[S_OM]);
        break;
    case SOP_BI_RESTART: sc[S_RHO] = sc[S_AL] = sc[S_OM] = 1.0; break;
    }
}
__global__ void k_scalar_op(int op, double *sc) {
    if (threadIdx.x == 0) scalar_op(op, sc);
}
void sd_scalar_op(hipStream_t s, int op, RedBuf rb) { k_scalar_op<<<1, 64, 0, s>>>(op, rb.scal); }

__global__ void __launch_bounds__(SD_BLOCK) k_red_final(RedBuf rb, int nv, int nb, int slot0, int is_max, int post) {
    __shared__ double lds[SD_BLOCK / 64];
    int w = threadIdx.x >> 6, ln = threadIdx.x & 63;
    for (int v = 0; v < nv; v++) {
        double s = 0.0;
        for (int b = threadIdx.x; b < nb; b += SD_BLOCK) {
            double x = rb.part[v * SD_RED_MAXB + b];
            s = is_max ? sd_nanmax(s, x) : s + x;
        }
        s = is_max ? sd_wave_max(s) : sd_wave_sum(s);
        if (ln == 0) lds[w] = s;
        __syncthreads();
        if (threadIdx.x == 0) {
            double r = lds[0];
            for (int q = 1; q < SD_BLOCK / 64; q++) r = is_max ? sd_nanmax(r, lds[q]) : r + lds[q];
            rb.scal[slot0 + v] = r;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0 && post >= 0) scalar_op(post, rb.scal);
}
void red_final(hipStream_t s, RedBuf rb, int nv, int nb, int slot0, int is_max, int post) {
    k_red_final<<<1, SD_BLOCK, 0, s>>>(rb, nv, nb, slot0, is_max, post);
}

__global__ void k_pcg_xr(double *x, double *r, const double *p, const double *q, int n, const double *sc,
                         Stage0 st) {
    int k = blockIdx.x * SD_BLOCK + threadIdx.x;
    if (k >= n || !(sc[S_PQ] > 0.)) return;
    double al = sc[S_AL];
    x[k] += al * p[k];
    const double rk = r[k] - al * q[k];
    r[k] = rk;
    if (st.b) st_put(st, k, rk);
}
__global__ void k_pcg_p(double *p, const double *z, int n, const double *sc) {
    int k = blockIdx.x * SD_BLOCK + threadIdx.x;
    if (k >= n) return;
    p[k] = z[k] + sc[S_BE] * p[k];
}
// z read where the V-cycle left it (the staged level-0 solution)
template <class TS>
__global__ void k_pcg_p_st(double *p, Stage0 st, int n, const double *sc) {
    int k = blockIdx.x * SD_BLOCK + threadIdx.x;
    if (k >= n) return;
    p[k] = st_get<TS>(st, st_pos(st, k)) + sc[S_BE] * p[k];
}
void sd_pcg_update_p_st(hipStream_t s, double *p, Stage0 st, int n, RedBuf rb) {
    if (st.f32) k_pcg_p_st<float><<<sd_nblocks(n), SD_BLOCK, 0, s>>>(p, st, n, rb.scal);
    else k_pcg_p_st<double><<<sd_nblocks(n), SD_BLOCK, 0, s>>>(p, st, n, rb.scal);
}
void sd_pcg_update_xr(hipStream_t s, double *x, double *r, const double *p, const double *q, int n, RedBuf rb,
                      Stage0 st) {
    k_pcg_xr<<<sd_nblocks(n), SD_BLOCK, 0, s>>>(x, r, p, q, n, rb.scal, st);
}
void sd_pcg_update_p(hipStream_t s, double *p, const double *z, int n, RedBuf rb) {
    k_pcg_p<<<sd_nblocks(n), SD_BLOCK, 0, s>>>(p, z, n, rb.scal);
}

__global__ void k_bi_p(double *p, const double *r, const double *v, int n, const double *sc, Stage0 st) {
    int k = blockIdx.x * SD_BLOCK + threadIdx.x;
    if (k >= n) return;
    double be = sc[S_BE], om = sc[S_OM];
    const double pk = r[k] + be * (p[k] - om * v[k]);
    p[k] = pk;
    if (st.b) st_put(st, k, pk);
}
__global__ void __launch_bounds__(SD_BLOCK) k_bi_r1(double *r, const double *v, int n, RedBuf rb, Stage0 st) {
    double al = rb.scal[S_AL];
    double acc[1] = {0.0};
    for (int k = blockIdx.x * SD_BLOCK + threadIdx.x; k < n; k += gridDim.x * SD_BLOCK) {
        double x = r[k] - al * v[k];
        r[k] = x;
        if (st.b) st_put(st, k, x);
        acc[0] += x * x;
    }
    red_partial<1, false>(acc, rb);
}
__global__ void k_bi_x1(double *x, const double *ph, int n, const double *sc) {
    int k = blockIdx.x * SD_BLOCK + threadIdx.x;
    if (k < n) x[k] += sc[S_AL] * ph[k];
}
template <class TS>
__global__ void k_bi_x1_st(double *x, Stage0 st, int n, const double *sc) {
    int k = blockIdx.x * SD_BLOCK + threadIdx.x;
    if (k < n) x[k] += sc[S_AL] * st_get<TS>(st, st_pos(st, k));
}
// x += al ph + om sh, r -= om t with ph and sh read where the two V-cycles
// left them (staged level-0 solutions st.x and st.x2): the same values as
// their FP64 copies, so the same iterates
template <int NV, class TS>
__global__ void __launch_bounds__(SD_BLOCK) k_bi_xr2_st(double *x, double *r, Stage0 st, const double *t,
                                                        const double *rh, int n, RedBuf rb) {
    double al = rb.scal[S_AL], om = rb.scal[S_OM];
    double acc[NV] = {};
    Stage0 s2 = st;
    s2.x = st.x2;
    for (int k = blockIdx.x * SD_BLOCK + threadIdx.x; k < n; k += gridDim.x * SD_BLOCK) {
        const int q = st_pos(st, k);
        x[k] += al * st_get<TS>(st, q) + om * st_get<TS>(s2, q);
        double y = r[k] - om * t[k];
        r[k] = y;
        acc[0] += y * y;
        if constexpr (NV == 2) acc[1] += rh[k] * y;
    }
    red_partial<NV, false>(acc, rb);
}
// x += al ph + om sh, r -= om t; r.r and (with rh) the next iteration's rh.r
template <int NV>
__global__ void __launch_bounds__(SD_BLOCK) k_bi_xr2(double *x, double *r, const double *ph, const double *sh,
                                                     const double *t, const double *rh, int n, RedBuf rb) {
    double al = rb.scal[S_AL], om = rb.scal[S_OM];
    double acc[NV] = {};
    for (int k = blockIdx.x * SD_BLOCK + threadIdx.x; k < n; k += gridDim.x * SD_BLOCK) {
        x[k] += al * ph[k] + om * sh[k];
        double y = r[k] - om * t[k];
        r[k] = y;
        acc[0] += y * y;
        if constexpr (NV == 2) acc[1] += rh[k] * y;
    }
    red_partial<NV, false>(acc, rb);
}
__global__ void __launch_bounds__(SD_BLOCK) k_bi_init(double *r, double *rh, double *p, double *v, const double *b,
                                                      const double *t, int n, RedBuf rb) {
    double acc[1] = {0.0};
    for (int k = blockIdx.x * SD_BLOCK + threadIdx.x; k < n; k += gridDim.x * SD_BLOCK) {
        double y = b[k] - t[k];
        r[k] = y; rh[k] = y; p[k] = 0.; v[k] = 0.;
        acc[0] += y * y;
    }
    red_partial<1, false>(acc, rb);
}
__global__ void k_bi_restart(double *rh, double *p, double *v, const double *r, int n) {
    int k = blockIdx.x * SD_BLOCK + threadIdx.x;
    if (k < n) { rh[k] = r[k]; p[k] = 0.; v[k] = 0.; }
}
void sd_bi_p(hipStream_t s, double *p, const double *r, const double *v, int n, RedBuf rb, Stage0 st) {
    k_bi_p<<<sd_nblocks(n), SD_BLOCK, 0, s>>>(p, r, v, n, rb.scal, st);
}
void sd_bi_r1(hipStream_t s, double *r, const double *v, int n, RedBuf rb, Stage0 st) {
    int nb = red_grid(n);
    k_bi_r1<<<nb, SD_BLOCK, 0, s>>>(r, v, n, rb, st);
    red_final(s, rb, 1, nb, S_SS, 0);
}
void sd_bi_x1(hipStream_t s, double *x, const double *ph, int n, RedBuf rb) {
    k_bi_x1<<<sd_nblocks(n), SD_BLOCK, 0, s>>>(x, ph, n, rb.scal);
}
void sd_bi_x1_st(hipStream_t s, double *x, Stage0 st, int n, RedBuf rb) {
    if (st.f32) k_bi_x1_st<float><<<sd_nblocks(n), SD_BLOCK, 0, s>>>(x, st, n, rb.scal);
    else k_bi_x1_st<double><<<sd_nblocks(n), SD_BLOCK, 0, s>>>(x, st, n, rb.scal);
}
void sd_bi_xr2_st(hipStream_t s, double *x, double *r, Stage0 st, const double *t, const double *rh, int n,
                  RedBuf rb) {
    int nb = red_grid(n);
    if (rh) {
        if (st.f32) k_bi_xr2_st<2, float><<<nb, SD_BLOCK, 0, s>>>(x, r, st, t, rh, n, rb);
        else k_bi_xr2_st<2, double><<<nb, SD_BLOCK, 0, s>>>(x, r, st, t, rh, n, rb);
        red_final(s, rb, 2, nb, S_TMP2, 0, SOP_BI_NEXT_BE);
    } else {
        if (st.f32) k_bi_xr2_st<1, float><<<nb, SD_BLOCK, 0, s>>>(x, r, st, t, rh, n, rb);
        else k_bi_xr2_st<1, double><<<nb, SD_BLOCK, 0, s>>>(x, r, st, t, rh, n, rb);
        red_final(s, rb, 1, nb, S_RR, 0, SOP_BI_NEXT);
    }
}
void sd_bi_xr2(hipStream_t s, double *x, double *r, const double *ph, const double *sh, const double *t,
               const double *rh, int n, RedBuf rb) {
    int nb = red_grid(n);
    if (rh) {
        k_bi_xr2<2><<<nb, SD_BLOCK, 0, s>>>(x, r, ph, sh, t, rh, n, rb);
        red_final(s, rb, 2, nb, S_TMP2, 0, SOP_BI_NEXT_BE);
    } else {
        k_bi_xr2<1><<<nb, SD_BLOCK, 0, s>>>(x, r, ph, sh, t, rh, n, rb);
        red_final(s, rb, 1, nb, S_RR, 0, SOP_BI_NEXT);
    }
}
void sd_bi_init(hipStream_t s, double *r, double *rh, double *p, double *v, const double *b, const double *t, int n,
                RedBuf rb) {
    int nb = red_grid(n);
    k_bi_init<<<nb, SD_BLOCK, 0, s>>>(r, rh, p, v, b, t, n, rb);
    red_final(s, rb, 1, nb, S_RR, 0);
}
void sd_bi_restart(hipStream_t s, double *rh, double *p, double *v, const double *r, int n) {
    k_bi_restart<<<sd_nblocks(n), SD_BLOCK, 0, s>>>(rh, p, v, r, n);
}

// ------------------------------------------------------------ multigrid
// aggregate of coarse node (I,J,L) on the fine level: 2x2x2 where the fine
// level is coarsened along that axis (semiconductor_3d.c:1995-1998)
struct Agg { int i0, i1, j0, j1, l0, l1; };
__device__ __forceinline__ Agg agg_range(int fNx, int fNy, int fNz, int cx, int cy, int cz, int I, int J, int L) {
    Agg a;
    a.i0 = cx ? 2 * I : I; a.i1 = cx ? 2 * I + 1 : I; if (a.i1 > fNx - 1) a.i1 = fNx - 1;
    a.j0 = cy ? 2 * J : J; a.j1 = cy ? 2 * J + 1 : J; if (a.j1 > fNy - 1) a.j1 = fNy - 1;
    a.l0 = cz ? 2 * L : L; a.l1 = cz ? 2 * L + 1 : L; if (a.l1 > fNz - 1) a.l1 = fNz - 1;
    return a;
}

// Slotboom weights of level L into level L+1 (semiconductor_3d.c:2085-2117)
template <class T>
__global__ void __launch_bounds__(SD_BLOCK) k_mg_weights(MGLevel<T> f, MGLevel<T> c, int L, DevView d, int qsign,
                                                         double VT) {
    int K = blockIdx.x * SD_BLOCK + threadIdx.x;
    if (K >= c.N) return;
    int cNxNy = c.Nx * c.Ny, LL = K / cNxNy, rr = K - LL * cNxNy, J = rr / c.Nx, I = rr - J * c.Nx;
    Agg a = agg_range(f.Nx, f.Ny, f.Nz, f.cx, f.cy, f.cz, I, J, LL);
    double pv[8];
    int kk[8], m = 0;
    double mx = PH_NONE;
    for (int l = a.l0; l <= a.l1; l++)
        for (int j = a.j0; j <= a.j1; j++)
            for (int i = a.i0; i <= a.i1; i++) {
                const int k = mg_idx(f, i, j, l), kn = i + f.Nx * (j + f.Ny * l);
                double p;
                if (L == 0) {
                    if (d.kind[kn] != K_SEMI || d.con_mask[kn]) p = PH_NONE;
                    else p = qsign > 0 ? d.phin[kn] / VT : -d.phip[kn] / VT;
                } else p = f.ph[k];
                if (p > mx) mx = p;
                pv[m] = p; kk[m] = k; m++;
            }
    c.ph[mg_idx(c, I, J, LL)] = mx;
    for (int q = 0; q < m; q++) {
        double e = (pv[q] <= 0.5 * PH_NONE) ? -800. : pv[q] - mx;
        f.pw[kk[q]] = (T)(e < -700. ? 0. : exp(e));
    }
}
template <class T>
void mg_weights_level(hipStream_t s, const MGLevel<T> &f, MGLevel<T> &c, int L, const DevView &d, int qsign,
                      double VT) {
    k_mg_weights<T><<<sd_nblocks(c.N), SD_BLOCK, 0, s>>>(f, c, L, d, qsign, VT);
}

template <class T> __global__ void k_fill_T(T *a, T v, int n) {
    int k = blockIdx.x * SD_BLOCK + threadIdx.x;
    if (k < n) a[k] = v;
}
template <class T> void mg_set_pw_one(hipStream_t s, MGLevel<T> &v) {
    k_fill_T<T><<<sd_nblocks(v.N), SD_BLOCK, 0, s>>>(v.pw, (T)1, v.N);
}

// Galerkin collapse P^T W A P of the 7-point stencil (semiconductor_3d.c:
// 2003-2027).  W = row scale of the fine level; the coarse row is stored
// scaled to unit diagonal with its FP64 scale in c.rs.
template <class T>
__global__ void __launch_bounds__(SD_BLOCK) k_galerkin(MGLevel<T> f, MGLevel<T> c) {
    int K = blockIdx.x * SD_BLOCK + threadIdx.x;
    if (K >= c.N) return;
    int cNxNy = c.Nx * c.Ny, LL = K / cNxNy, rr = K - LL * cNxNy, J = rr / c.Nx, I = rr - J * c.Nx;
    Agg a = agg_range(f.Nx, f.Ny, f.Nz, f.cx, f.cy, f.cz, I, J, LL);
    const int NxNy = f.Nx * f.Ny;
    double cc = 0, xm = 0, xp = 0, ym = 0, yp = 0, zm = 0, zp = 0;
    for (int l = a.l0; l <= a.l1; l++)
        for (int j = a.j0; j <= a.j1; j++)
            for (int i = a.i0; i <= a.i1; i++) {
                const int k = mg_idx(f, i, j, l), kx = k - mg_xpos(f, i);
                double wk = f.rs ? f.rs[k] : 1.0, t;
                if (wk == 0.0) continue;
                double ac = f.A.c ? (double)f.A.c[k] : 1.0;
                cc += wk * ac * (double)f.pw[k];
                if (i > 0) { t = wk * (double)f.A.xm[k] * (double)f.pw[kx + mg_xpos(f, i - 1)]; if (i - 1 >= a.i0) cc += t; else xm += t; }
                if (i < f.Nx - 1) { t = wk * (double)f.A.xp[k] * (double)f.pw[kx + mg_xpos(f, i + 1)]; if (i + 1 <= a.i1) cc += t; else xp += t; }
                if (j > 0) { t = wk * (double)f.A.ym[k] * (double)f.pw[k - f.Nx]; if (j - 1 >= a.j0) cc += t; else ym += t; }
                if (j < f.Ny - 1) { t = wk * (double)f.A.yp[k] * (double)f.pw[k + f.Nx]; if (j + 1 <= a.j1) cc += t; else yp += t; }
                if (l > 0) { t = wk * (double)f.A.zm[k] * (double)f.pw[k - NxNy]; if (l - 1 >= a.l0) cc += t; else zm += t; }
                if (l < f.Nz - 1) { t = wk * (double)f.A.zp[k] * (double)f.pw[k + NxNy]; if (l + 1 <= a.l1) cc += t; else zp += t; }
            }
    // an aggregate of identity rows only carries no equation: identity row.
    // A zero pivot is replaced by 1 as the reference smoother does (:1939-1940).
    double sc = cc;
    if (cc == 0. && xm == 0. && xp == 0. && ym == 0. && yp == 0. && zm == 0. && zp == 0.) sc = 1.0;
    if (!(sc != 0.0) || !isfinite(sc)) sc = 1.0;
    double is = 1.0 / sc;
    const int Kc = mg_idx(c, I, J, LL);
    c.rs[Kc] = sc;
    c.A.xm[Kc] = (T)(xm * is); c.A.xp[Kc] = (T)(xp * is);
    c.A.ym[Kc] = (T)(ym * is); c.A.yp[Kc] = (T)(yp * is);
    c.A.zm[Kc] = (T)(zm * is); c.A.zp[Kc] = (T)(zp * is);
}
template <class T> void mg_galerkin(hipStream_t s, const MGLevel<T> &f, MGLevel<T> &c) {
    k_galerkin<T><<<sd_nblocks(c.N), SD_BLOCK, 0, s>>>(f, c);
}

// Red-black Gauss-Seidel half sweep on the nodes of one colour
// ((i+j+l) & 1 == color).  zero != 0: the sweep starts from x = 0 (every node
// is written: the colour's nodes get b/d, the others 0).  pro: the neighbours
// carry the coarse correction x += pw * x_parent, applied on the fly.
template <class T>
__device__ __forceinline__ void rb_zero_node(const MGLevel<T> &v, const T *b, T *x, int color, int k) {
    const int Nx = v.Nx, NxNy = Nx * v.Ny;
    int l = k / NxNy, r = k - l * NxNy, j = r / Nx, i = r - j * Nx;
    if (((i + j + l) & 1) != color) { x[k] = (T)0; return; }
    T d = v.A.c ? v.A.c[k] : (T)1;
    if (!(d != (T)0)) d = (T)1;
    x[k] = b[k] / d;
}
// t indexes the colour's nodes: ceil(Nx/2) per x-line
template <class T>
__device__ __forceinline__ void rb_node(const MGLevel<T> &v, const T *b, T *x, int color, int t,
                                        const MGLevel<T> &c, int pro) {
    const int Nx = v.Nx, Ny = v.Ny, Nz = v.Nz, NxNy = Nx * Ny;
    int Hx = (Nx + 1) >> 1;
    int line = t / Hx, i2 = t - line * Hx, l = line / Ny, j = line - l * Ny;
    int i = 2 * i2 + ((j + l + color) & 1);
    if (i >= Nx) return;
    int k = i + Nx * (j + Ny * l);
    const bool h0 = i > 0, h1 = i < Nx - 1, h2 = j > 0, h3 = j < Ny - 1, h4 = l > 0, h5 = l < Nz - 1;
    T d = v.A.c ? v.A.c[k] : (T)1;          // loaded with the neighbours
    T s = b[k];
    if (pro) {
        auto xc = [&](int ii, int jj, int ll, int kk) -> T {
            int I = v.cx ? ii >> 1 : ii, J = v.cy ? jj >> 1 : jj, L = v.cz ? ll >> 1 : ll;
            return x[kk] + v.pw[kk] * c.x[I + c.Nx * (J + c.Ny * L)];
        };
        s = nb_sub<T>(h0, s, v.A.xm[k], xc(i - h0, j, l, k - h0));
        s = nb_sub<T>(h1, s, v.A.xp[k], xc(i + h1, j, l, k + h1));
        s = nb_sub<T>(h2, s, v.A.ym[k], xc(i, j - h2, l, k - h2 * Nx));
        s = nb_sub<T>(h3, s, v.A.yp[k], xc(i, j + h3, l, k + h3 * Nx));
        s = nb_sub<T>(h4, s, v.A.zm[k], xc(i, j, l - h4, k - h4 * NxNy));
        s = nb_sub<T>(h5, s, v.A.zp[k], xc(i, j, l + h5, k + h5 * NxNy));
    } else {
        s = nb_sub<T>(h0, s, v.A.xm[k], x[nb_idx(h0, k - 1, k)]);
        s = nb_sub<T>(h1, s, v.A.xp[k], x[nb_idx(h1, k + 1, k)]);
        s = nb_sub<T>(h2, s, v.A.ym[k], x[nb_idx(h2, k - Nx, k)]);
        s = nb_sub<T>(h3, s, v.A.yp[k], x[nb_idx(h3, k + Nx, k)]);
        s = nb_sub<T>(h4, s, v.A.zm[k], x[nb_idx(h4, k - NxNy, k)]);
        s = nb_sub<T>(h5, s, v.A.zp[k], x[nb_idx(h5, k + NxNy, k)]);
    }
    if (!(d != (T)0)) d = (T)1;
    x[k] = s / d;
}
// Both half sweeps of a red-black sweep from x = 0 in one pass: the first
// colour's nodes get b/d, and a node of the other colour reads its
// neighbours' values as b/d directly, so every node is written once.  The
// operations and their order are those of the two half sweeps, so x is the
// same to the last bit.
template <class T>
__device__ __forceinline__ void rb_zero2_node(const MGLevel<T> &v, const T *b, T *x, int color, int k) {
    const int Nx = v.Nx, Ny = v.Ny, Nz = v.Nz, NxNy = Nx * Ny;
    int l = k / NxNy, r = k - l * NxNy, j = r / Nx, i = r - j * Nx;
    auto dg = [&](int m) -> T {
        T d = v.A.c ? v.A.c[m] : (T)1;
        return d != (T)0 ? d : (T)1;
    };
    if (((i + j + l) & 1) == color) { x[k] = b[k] / dg(k); return; }
    const bool h0 = i > 0, h1 = i < Nx - 1, h2 = j > 0, h3 = j < Ny - 1, h4 = l > 0, h5 = l < Nz - 1;
    const int n0 = nb_idx(h0, k - 1, k), n1 = nb_idx(h1, k + 1, k), n2 = nb_idx(h2, k - Nx, k), n3 = nb_idx(h3, k + Nx, k),
              n4 = nb_idx(h4, k - NxNy, k), n5 = nb_idx(h5, k + NxNy, k);
    T s = b[k];
    s = nb_sub<T>(h0, s, v.A.xm[k], b[n0] / dg(n0));
    s = nb_sub<T>(h1, s, v.A.xp[k], b[n1] / dg(n1));
    s = nb_sub<T>(h2, s, v.A.ym[k], b[n2] / dg(n2));
    s = nb_sub<T>(h3, s, v.A.yp[k], b[n3] / dg(n3));
    s = nb_sub<T>(h4, s, v.A.zm[k], b[n4] / dg(n4));
    s = nb_sub<T>(h5, s, v.A.zp[k], b[n5] / dg(n5));
    x[k] = s / dg(k);
}
// zero: 1 the colour's half sweep from x = 0, 2 the whole sweep from x = 0
// (`color` first)
template <class T>
__global__ void __launch_bounds__(SD_BLOCK) k_rbgs(MGLevel<T> v, const T *b, T *x, int color, int zero,
                                                 MGLevel<T> c, int pro) {
    int t = sd_tile() * SD_BLOCK + threadIdx.x;
    if (zero == 2) {
        if (t < v.N) rb_zero2_node(v, b, x, color, t);
        return;
    }
    if (zero) {
        if (t < v.N) rb_zero_node(v, b, x, color, t);
        return;
    }
    if (t < ((v.Nx + 1) >> 1) * v.Ny * v.Nz) rb_node(v, b, x, color, t, c, pro);
}
int g_rb_fuse = 1;            // 1: the pre-smoothing red-black sweep from x = 0 runs as one launch

// Both half sweeps of the post-smoothing red-black sweep in one pass.  A
// colour-1 node computes its new value as the first half sweep does (its
// neighbours from x plus the coarse correction pw * x_parent); a colour-0
// node recomputes the new values of its colour-1 neighbours the same way and
// then its own from them.  The result goes to a second array, since the
// colour-0 values of x are read by other nodes' recomputations.
template <class T>
__global__ void __launch_bounds__(SD_BLOCK) k_rbgs_post2(MGLevel<T> v, const T *b, const T *x, T *xo, MGLevel<T> c) {
    const int k = sd_tile() * SD_BLOCK + threadIdx.x;
    if (k >= v.N) return;
    const int Nx = v.Nx, Ny = v.Ny, Nz = v.Nz, NxNy = Nx * Ny;
    const int l = k / NxNy, r = k - l * NxNy, j = r / Nx, i = r - j * Nx;
    auto dg = [&](int m) -> T {
        T d = v.A.c ? v.A.c[m] : (T)1;
        return d != (T)0 ? d : (T)1;
    };
    // x of a colour-0 node with the coarse correction
    auto xcor = [&](int ii, int jj, int ll, int kk) -> T {
        int I = v.cx ? ii >> 1 : ii, J = v.cy ? jj >> 1 : jj, L = v.cz ? ll >> 1 : ll;
        return x[kk] + v.pw[kk] * c.x[I + c.Nx * (J + c.Ny * L)];
    };
    // new value of colour-1 node (ii, jj, ll) = kk
    auto new1 = [&](int ii, int jj, int ll, int kk) -> T {
        T s = b[kk];
        if (ii > 0) s -= v.A.xm[kk] * xcor(ii - 1, jj, ll, kk - 1);
        if (ii < Nx - 1) s -= v.A.xp[kk] * xcor(ii + 1, jj, ll, kk + 1);
        if (jj > 0) s -= v.A.ym[kk] * xcor(ii, jj - 1, ll, kk - Nx);
        if (jj < Ny - 1) s -= v.A.yp[kk] * xcor(ii, jj + 1, ll, kk + Nx);
        if (ll > 0) s -= v.A.zm[kk] * xcor(ii, jj, ll - 1, kk - NxNy);
        if (ll < Nz - 1) s -= v.A.zp[kk] * xcor(ii, jj, ll + 1, kk + NxNy);
        return s / dg(kk);
    };
    if ((i + j + l) & 1) { xo[k] = new1(i, j, l, k); return; }
    T s = b[k];
    if (i > 0) s -= v.A.xm[k] * new1(i - 1, j, l, k - 1);
    if (i < Nx - 1) s -= v.A.xp[k] * new1(i + 1, j, l, k + 1);
    if (j > 0) s -= v.A.ym[k] * new1(i, j - 1, l, k - Nx);
    if (j < Ny - 1) s -= v.A.yp[k] * new1(i, j + 1, l, k + Nx);
    if (l > 0) s -= v.A.zm[k] * new1(i, j, l - 1, k - NxNy);
    if (l < Nz - 1) s -= v.A.zp[k] * new1(i, j, l + 1, k + NxNy);
    xo[k] = s / dg(k);
}
int g_rb_fuse_post = 0;       // 1: the post-smoothing red-black sweep of a coarse level runs as one launch (measured slower: off)
template <class T> void mg_rbgs_post2(hipStream_t s, const MGLevel<T> &v, const T *b, const T *x, T *xo,
                                      const MGLevel<T> &c) {
    k_rbgs_post2<T><<<sd_nblocks(v.N), SD_BLOCK, 0, s>>>(v, b, x, xo, c);
}
template <class T> void mg_rbgs(hipStream_t s, const MGLevel<T> &v, const T *b, T *x, int color, int zero,
                                 const MGLevel<T> *c) {
    MGLevel<T> cc = c ? *c : v;
    if (zero) k_rbgs<T><<<sd_nblocks(v.N), SD_BLOCK, 0, s>>>(v, b, x, color, zero, cc, 0);
    else {
        int n = ((v.Nx + 1) >> 1) * v.Ny * v.Nz;
        k_rbgs<T><<<sd_nblocks(n), SD_BLOCK, 0, s>>>(v, b, x, color, 0, cc, c != nullptr);
    }
}

// Zebra line Gauss-Seidel: every line along axis `ax` whose two other indices
// have parity `color` is solved exactly (tridiagonal in-line couplings, the
// off-line neighbours taken from x).  The LU factors of every line depend only
// on the operator, so mg_line_factor computes them once per multigrid setup:
// for line element q, iw = 1/w_q, ai = a_q/w_q, cp = c_q/w_q with
// w_q = d_q - a_q cp_{q-1}.  A sweep is then a short recurrence per line:
//   y_q = r_q iw_q - ai_q y_{q-1};  x_q = y_q - cp_q x_{q+1}.
// One workgroup owns NL lines of one colour: all lanes compute the line
// right-hand sides r_q = b - (off-line couplings) x with coalesced loads and
// stage them with the factors in LDS, NL lanes run the recurrences, all lanes
// store x.  zero: the sweep starts from x = 0 (off-line terms vanish).  pro:
// the off-line neighbours carry the coarse correction x + pw * x_parent (first
// post-smoothing sweep).
#define LINE_MAXE 4
#define LINE_MAXEQ (SD_BLOCK * LINE_MAXE)
struct LineGeom { int n, Ntd, s_ax, s_td, s_3; };
__device__ __forceinline__ LineGeom line_geom(int Nx, int Ny, int Nz, int ax) {
    LineGeom g;
    if (ax == 0) { g.n = Nx; g.Ntd = Ny; g.s_ax = 1; g.s_td = Nx; g.s_3 = Nx * Ny; }
    else if (ax == 1) { g.n = Ny; g.Ntd = Nx; g.s_ax = Nx; g.s_td = 1; g.s_3 = Nx * Ny; }
    else { g.n = Nz; g.Ntd = Nx; g.s_ax = Nx * Ny; g.s_td = 1; g.s_3 = Nx; }
    return g;
}
// One lane per line (both colours), sequential over the line.  The
// coefficients of the next LF_CHUNK elements are loaded while the current
// chunk is factored, so a line waits for memory about once (the first chunk)
// rather than once per chunk; 64-lane workgroups spread the few thousand
// lines over all CUs.  The FP64 hierarchy factors in FP64 with IEEE division;
// the FP32 hierarchy in FP32 with the hardware reciprocal refined by one
// Newton step (the factors are stored in FP32, and BF16 on level 0, either
// way): the dependent chain along a line is then a few single-precision
// operations per element instead of an FP64 division.
#define LF_CHUNK 16
template <class T> __device__ __forceinline__ T lf_recip(T w) {
    if constexpr (std::is_same_v<T, float>) {
        float r = __builtin_amdgcn_rcpf(w);
        return r * (2.0f - w * r);
    } else {
        return 1.0 / w;
    }
}
template <class T>
__global__ void __launch_bounds__(64) k_line_factor(MGLevel<T> v, int ax) {
    LineGeom g = line_geom(v.Nx, v.Ny, v.Nz, ax);
    int N3 = ax == 2 ? v.Ny : v.Nz;
    int t = blockIdx.x * 64 + threadIdx.x;
    if (t >= g.Ntd * N3) return;
    int td = t % g.Ntd, p3 = t / g.Ntd;
    const T *am = ax == 0 ? v.A.xm : (ax == 1 ? v.A.ym : v.A.zm);
    const T *ap = ax == 0 ? v.A.xp : (ax == 1 ? v.A.yp : v.A.zp);
    T *ai = v.lu[ax][0], *iw = v.lu[ax][1], *cp = v.lu[ax][2];
    auto node = [&](int q) {
        return ax == 0 ? mg_idx(v, q, td, p3) : (ax == 1 ? mg_idx(v, td, q, p3) : mg_idx(v, td, p3, q));
    };
    T dv[LF_CHUNK], av[LF_CHUNK], cv[LF_CHUNK];
    int kk[LF_CHUNK];
    auto load = [&](int q0, T (&d)[LF_CHUNK], T (&a)[LF_CHUNK], T (&c)[LF_CHUNK], int (&kq)[LF_CHUNK]) {
#pragma unroll
        for (int u = 0; u < LF_CHUNK; u++) {
            const int q = q0 + u < g.n ? q0 + u : g.n - 1;
            const int k = node(q);
            kq[u] = k;
            d[u] = v.A.c ? v.A.c[k] : (T)1;
            a[u] = am[k];
            c[u] = ap[k];
        }
    };
    load(0, dv, av, cv, kk);
    T cprev = 0;
    for (int q0 = 0; q0 < g.n; q0 += LF_CHUNK) {
        T dn[LF_CHUNK], an[LF_CHUNK], cn[LF_CHUNK];
        int kn[LF_CHUNK];
        const bool more = q0 + LF_CHUNK < g.n;
        if (more) load(q0 + LF_CHUNK, dn, an, cn, kn);
#pragma unroll
        for (int u = 0; u < LF_CHUNK; u++) {
            const int q = q0 + u;
            if (q >= g.n) break;
            T dg = dv[u];
            if (!(dg != (T)0)) dg = 1;
            const T a = q > 0 ? av[u] : (T)0, c = q < g.n - 1 ? cv[u] : (T)0;
            T w = dg - a * cprev;
            if (!(w != (T)0)) w = dg;
            const T iwq = lf_recip<T>(w);
            cprev = c * iwq;
            const int k = kk[u];
            ai[k] = a * iwq; iw[k] = iwq; cp[k] = cprev;
        }
        if (more) {
#pragma unroll
            for (int u = 0; u < LF_CHUNK; u++) { dv[u] = dn[u]; av[u] = an[u]; cv[u] = cn[u]; kk[u] = kn[u]; }
        }
    }
}
template <class T> void mg_line_factor(hipStream_t s, const MGLevel<T> &v, int ax) {
    int N3 = ax == 2 ? v.Ny : v.Nz, Ntd = ax == 0 ? v.Ny : v.Nx;
    k_line_factor<T><<<sd_nblocks(Ntd * N3, 64), 64, 0, s>>>(v, ax);
}

template <class T>
__global__ void __launch_bounds__(SD_BLOCK) k_line_gs(MGLevel<T> v, const T *__restrict__ b, T *__restrict__ x, int ax,
                                                      int color, int zero, MGLevel<T> c, int pro, int NL) {
    __shared__ T sr[LINE_MAXEQ], sa[LINE_MAXEQ], sw[LINE_MAXEQ], sp[LINE_MAXEQ], sy[LINE_MAXEQ];
    const int Nx = v.Nx, Ny = v.Ny, Nz = v.Nz, NxNy = Nx * Ny;
    const LineGeom g = line_geom(Nx, Ny, Nz, ax);
    const int n = g.n;
    const int ntile = (g.Ntd + 2 * NL - 1) / (2 * NL);
    const int p3 = blockIdx.x / ntile, t0 = (blockIdx.x - p3 * ntile) * 2 * NL;
    const int sh = (t0 + p3 + color) & 1;
    const int neq = NL * n;
    const T *__restrict__ AI = v.lu[ax][0];
    const T *__restrict__ IW = v.lu[ax][1];
    const T *__restrict__ CP = v.lu[ax][2];
    // phase 1: every lane gathers the right-hand sides of up to LINE_MAXE line
    // elements; the loop is unrolled and predicated so that all loads issue
    // before any of them is consumed
#pragma unroll
    for (int m = 0; m < LINE_MAXE; m++) {
        const int e = threadIdx.x + SD_BLOCK * m;
        const int q = e / NL, L = e - q * NL, td = t0 + 2 * L + sh;
        const bool ok = e < neq && td < g.Ntd;
        T r = 0, ai = 0, iw = 1, cp = 0;
        if (ok) {
            const int k = td * g.s_td + q * g.s_ax + p3 * g.s_3;
            int i, j, l;
            if (ax == 0) { i = q; j = td; l = p3; } else if (ax == 1) { i = td; j = q; l = p3; } else { i = td; j = p3; l = q; }
            ai = AI[k]; iw = IW[k]; cp = CP[k];
            T s = b[k];
            if (!zero) {
                auto xv = [&](int ii, int jj, int ll, int kk) -> T {
                    if (!pro) return x[kk];
                    int I = v.cx ? ii >> 1 : ii, J = v.cy ? jj >> 1 : jj, LL = v.cz ? ll >> 1 : ll;
                    return x[kk] + v.pw[kk] * c.x[I + c.Nx * (J + c.Ny * LL)];
                };
                if (ax != 0) {
                    if (i > 0) s -= v.A.xm[k] * xv(i - 1, j, l, k - 1);
                    if (i < Nx - 1) s -= v.A.xp[k] * xv(i + 1, j, l, k + 1);
                }
                if (ax != 1) {
                    if (j > 0) s -= v.A.ym[k] * xv(i, j - 1, l, k - Nx);
                    if (j < Ny - 1) s -= v.A.yp[k] * xv(i, j + 1, l, k + Nx);
                }
                if (ax != 2) {
                    if (l > 0) s -= v.A.zm[k] * xv(i, j, l - 1, k - NxNy);
                    if (l < Nz - 1) s -= v.A.zp[k] * xv(i, j, l + 1, k + NxNy);
                }
            }
            r = s;
        }
        if (e < neq) { sr[e] = r * iw; sa[e] = ai; sp[e] = cp; }
    }
    __syncthreads();
    // phase 2: one lane per line runs the two recurrences (separate LDS
    // arrays in and out, so the next elements' loads can run ahead)
    if (threadIdx.x < NL) {
        const int L = threadIdx.x;
        T y = 0;
#pragma unroll 8
        for (int q = 0; q < n; q++) {
            const int e = q * NL + L;
            y = sr[e] - sa[e] * y;
            sy[e] = y;
        }
        T xq = 0;
#pragma unroll 8
        for (int q = n - 1; q >= 0; q--) {
            const int e = q * NL + L;
            xq = sy[e] - sp[e] * xq;
            sw[e] = xq;
        }
    }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < LINE_MAXE; m++) {
        const int e = threadIdx.x + SD_BLOCK * m;
        const int q = e / NL, L = e - q * NL, td = t0 + 2 * L + sh;
        if (e < neq && td < g.Ntd) x[td * g.s_td + q * g.s_ax + p3 * g.s_3] = sw[e];
    }
}

// The same zebra sweep with the two recurrences as wavefront scans, for lines
// of up to 64*P elements.  A wavefront owns S slots of 64*P elements; each
// slot holds LW = floor(64 P / n) whole lines packed end to end, P consecutive
// line elements per lane.  The forward recurrence
// y_e = r_e iw_e - ai_e y_{e-1} composes the affine maps y -> -ai_e y + r_e iw_e:
// every lane composes its P maps, a 6-step shuffle scan composes the lanes'
// maps across the wavefront, and every lane replays its P elements from the
// value that enters it.  The back substitution x_e = y_e - cp_e x_{e+1} runs
// the same way from the other end.  ai = 0 at the first element of a line and
// cp = 0 at its last cut the chain between packed lines, so the scan needs no
// segment flags.  The right-hand sides are gathered as in k_line_gs and staged
// in LDS: along the line for x-lines (contiguous), across the workgroup's
// lines otherwise, where S > 1 widens the contiguous run each load covers.
#define LS_WAVES 8
#define LS_THREADS (64 * LS_WAVES)
// Wavefront scans of affine maps y -> A y + B.  Lane i holds the composition
// of its own elements' maps; the scan leaves in lane i the composition of
// lanes 0..i (forward, the later map applied last) or of lanes i..63
// (backward).  FP32: four DPP row shifts give the scan within each row of 16
// lanes, then the row totals, read with readlane, carry it across rows.
// FP64: shuffles, log2(64) steps.
#define DPP_ROW_SHL(n) (0x100 + (n))
#define DPP_ROW_SHR(n) (0x110 + (n))
#define DPP_WAVE_SHL1 0x130
#define DPP_WAVE_SHR1 0x138
// lane move; lanes without a source lane get `old`
template <int CTRL> __device__ __forceinline__ float dpp_mov(float old, float src) {
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(old), __float_as_int(src), CTRL, 0xf, 0xf, false));
}
__device__ __forceinline__ float lane_f(float v, int l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
template <int CTRL> __device__ __forceinline__ void scan_step(float &A, float &B) {
    const float Ap = dpp_mov<CTRL>(1.0f, A), Bp = dpp_mov<CTRL>(0.0f, B);
    B = B + A * Bp;
    A = A * Ap;
}
template <class T> __device__ __forceinline__ void wave_scan_fwd(T &A, T &B, int lane) {
    if constexpr (std::is_same_v<T, float>) {
        scan_step<DPP_ROW_SHR(1)>(A, B);
        scan_step<DPP_ROW_SHR(2)>(A, B);
        scan_step<DPP_ROW_SHR(4)>(A, B);
        scan_step<DPP_ROW_SHR(8)>(A, B);
        // rows 0..2 end at lanes 15, 31, 47; the map entering row r composes them
        const float a0 = lane_f(A, 15), b0 = lane_f(B, 15), a1 = lane_f(A, 31), b1 = lane_f(B, 31);
        const float a2 = lane_f(A, 47), b2 = lane_f(B, 47);
        const float e2a = a1 * a0, e2b = b1 + a1 * b0;
        const float e3a = a2 * e2a, e3b = b2 + a2 * e2b;
        const int r = lane >> 4;
        if (r > 0) {
            const float ea = r == 1 ? a0 : (r == 2 ? e2a : e3a), eb = r == 1 ? b0 : (r == 2 ? e2b : e3b);
            B = B + A * eb;
            A = A * ea;
        }
    } else {
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            T Ap = __shfl_up(A, d, 64), Bp = __shfl_up(B, d, 64);
            if (lane >= d) { B = B + A * Bp; A = A * Ap; }
        }
    }
}
template <class T> __device__ __forceinline__ void wave_scan_bwd(T &A, T &B, int lane) {
    if constexpr (std::is_same_v<T, float>) {
        scan_step<DPP_ROW_SHL(1)>(A, B);
        scan_step<DPP_ROW_SHL(2)>(A, B);
        scan_step<DPP_ROW_SHL(4)>(A, B);
        scan_step<DPP_ROW_SHL(8)>(A, B);
        // rows 1..3 start at lanes 16, 32, 48
        const float a1 = lane_f(A, 16), b1 = lane_f(B, 16), a2 = lane_f(A, 32), b2 = lane_f(B, 32);
        const float a3 = lane_f(A, 48), b3 = lane_f(B, 48);
        const float e1a = a2 * a3, e1b = b2 + a2 * b3;
        const float e0a = a1 * e1a, e0b = b1 + a1 * e1b;
        const int r = lane >> 4;
        if (r < 3) {
            const float ea = r == 2 ? a3 : (r == 1 ? e1a : e0a), eb = r == 2 ? b3 : (r == 1 ? e1b : e0b);
            B = B + A * eb;
            A = A * ea;
        }
    } else {
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            T Al = __shfl_down(A, d, 64), Bl = __shfl_down(B, d, 64);
            if (lane + d < 64) { B = B + A * Bl; A = A * Al; }
        }
    }
}
// B of lane i - 1 (forward) or i + 1 (backward); 0 past the ends
template <class T> __device__ __forceinline__ T wave_prev(T B, int lane) {
    if constexpr (std::is_same_v<T, float>) return dpp_mov<DPP_WAVE_SHR1>(0.0f, B);
    else { T y = __shfl_up(B, 1, 64); return lane == 0 ? (T)0 : y; }
}
template <class T> __device__ __forceinline__ T wave_next(T B, int lane) {
    if constexpr (std::is_same_v<T, float>) return dpp_mov<DPP_WAVE_SHL1>(0.0f, B);
    else { T y = __shfl_down(B, 1, 64); return lane == 63 ? (T)0 : y; }
}

// Element idx of an n-element array through a raw buffer load or store: a
// 32-bit byte offset from a descriptor held in scalar registers, instead of a
// 64-bit address per lane, and a read past n bytes returns 0 (a write past it
// is dropped) instead of leaving the array.  Callers keep n * sizeof(U) below
// 2^32 (the launcher checks the level size).  Descriptor word 3 is the GFX9
// raw-buffer form (32-bit data format, no swizzle, stride 0).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void *base, unsigned bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), (short)0, (int)bytes, 0x00020000);
}
template <class U> __device__ __forceinline__ U ld_buf(const U *base, unsigned n, unsigned idx) {
    const __amdgpu_buffer_rsrc_t r = buf_rsrc(base, n * (unsigned)sizeof(U));
    const int off = (int)(idx * (unsigned)sizeof(U));
    if constexpr (sizeof(U) == 2) return __builtin_bit_cast(U, __builtin_amdgcn_raw_buffer_load_b16(r, off, 0, 0));
    else if constexpr (sizeof(U) == 4) return __builtin_bit_cast(U, __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
    else if constexpr (sizeof(U) == 8) return __builtin_bit_cast(U, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
    else {
        static_assert(sizeof(U) == 16, "2, 4, 8 or 16-byte elements");
        return __builtin_bit_cast(U, __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0));
    }
}
template <class U> __device__ __forceinline__ void st_buf(U *base, unsigned n, unsigned idx, U val) {
    const __amdgpu_buffer_rsrc_t r = buf_rsrc(base, n * (unsigned)sizeof(U));
    const int off = (int)(idx * (unsigned)sizeof(U));
    if constexpr (sizeof(U) == 4) __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, val), r, off, 0, 0);
    else {
        static_assert(sizeof(U) == 8, "4 or 8-byte elements");
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(__attribute__((ext_vector_type(2))) unsigned, val), r, off, 0, 0);
    }
}
// The wavefront scans of one sweep: wavefront w solves its S slots of 64 P
// elements in LDS (right-hand side r iw in s_r, ai in s_a, cp in s_c); the
// solution replaces s_r.
template <class T, int P, int S>
__device__ __forceinline__ void line_scan_slots(T *s_r, const T *s_a, const T *s_c) {
    constexpr int SP = 64 * P + 1;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int si = 0; si < S; si++) {
        T *R = s_r + (w * S + si) * SP;
        const T *Aa = s_a + (w * S + si) * SP, *Cc = s_c + (w * S + si) * SP;
        T ra[P], aa[P], ca[P];
#pragma unroll
        for (int p = 0; p < P; p++) { ra[p] = R[lane * P + p]; aa[p] = Aa[lane * P + p]; ca[p] = Cc[lane * P + p]; }
        T A = 1, B = 0;
#pragma unroll
        for (int p = 0; p < P; p++) { A = -aa[p] * A; B = ra[p] - aa[p] * B; }
        wave_scan_fwd<T>(A, B, lane);
        T y = wave_prev<T>(B, lane);
#pragma unroll
        for (int p = 0; p < P; p++) { y = ra[p] - aa[p] * y; ra[p] = y; }
        A = 1; B = 0;
#pragma unroll
        for (int p = P - 1; p >= 0; p--) { A = -ca[p] * A; B = ra[p] - ca[p] * B; }
        wave_scan_bwd<T>(A, B, lane);
        T xq = wave_next<T>(B, lane);
#pragma unroll
        for (int p = P - 1; p >= 0; p--) { xq = ra[p] - ca[p] * xq; ra[p] = xq; }
#pragma unroll
        for (int p = 0; p < P; p++) R[lane * P + p] = ra[p];
    }
}
// MD: 0 the sweep starts from x = 0 (no neighbour reads), 1 plain sweep, 2
// first post-smoothing sweep (adds the coarse correction to the neighbours it
// reads).  A template argument, so that the plain sweep holds no registers for
// the other two.
// H16: the off-line couplings and line factors come from the level's BF16
// copies (o16, lu16), converted to T on load.
template <class T, int P, int S, int MD, bool H16>
__global__ void __launch_bounds__(LS_THREADS) k_line_scan(MGLevel<T> v, const T *__restrict__ b, T *__restrict__ x,
                                                          int ax, int color, MGLevel<T> c, int LW, int xcd) {
    constexpr bool zero = MD == 0, pro = MD == 2;
    constexpr int SL = 64 * P;                       // elements per slot
    constexpr int SP = SL + 1;                       // slot pitch in LDS: odd, so that the gathers
                                                     // across lines (consecutive lanes, one slot
                                                     // apart) fall in different banks
    __shared__ T s_r[LS_WAVES * S * SP], s_a[LS_WAVES * S * SP], s_c[LS_WAVES * S * SP];
    const int Nx = v.Nx, Ny = v.Ny, Nz = v.Nz, NxNy = Nx * Ny;
    const LineGeom g = line_geom(Nx, Ny, Nz, ax);
    const int n = g.n;
    const int NLW = LS_WAVES * S * LW;               // lines per workgroup
    // the colour's lines of plane p3 are every other td: td = 2 h + sh.
    // Workgroup b runs on XCD b % 8; with xcd set, each XCD takes a
    // contiguous range of tiles, so that neighbouring tiles (which share
    // cache lines) and neighbouring planes meet in the same L2.
    const int ntile = (((g.Ntd + 1) >> 1) + NLW - 1) / NLW;
    const int bid = xcd ? (blockIdx.x & 7) * (gridDim.x >> 3) + (blockIdx.x >> 3) : blockIdx.x;
    const int p3 = bid / ntile, h0 = (bid - p3 * ntile) * NLW, sh = (p3 + color) & 1;
    const int E = NLW * n, used = LW * n, pad = SL - used;
    const T *__restrict__ AI = v.lu[ax][0];
    const T *__restrict__ IW = v.lu[ax][1];
    const T *__restrict__ CP = v.lu[ax][2];
    const __bf16 *__restrict__ AI16 = v.lu16[ax][0];
    const __bf16 *__restrict__ IW16 = v.lu16[ax][1];
    const __bf16 *__restrict__ CP16 = v.lu16[ax][2];
    // slot elements past the packed lines carry the identity
    for (int t = threadIdx.x; t < LS_WAVES * S * pad; t += LS_THREADS) {
        int q = (t / pad) * SP + used + t % pad;
        s_r[q] = 0; s_a[q] = 0; s_c[q] = 0;
    }
    // Gather, first pass: every load of this thread's M elements is issued
    // before any of them is used.  Elements outside the level and missing
    // neighbours read the element itself (clamped index) and are masked.
    // The two off-line axes in the order the terms are subtracted: for x-lines
    // y then z, for y-lines x then z, for z-lines x then y.
    constexpr int M = P * S;
    const int s1 = ax == 0 ? Nx : 1, s2 = ax == 2 ? Nx : NxNy;
    const T *o0 = ax == 0 ? v.A.ym : v.A.xm, *o1 = ax == 0 ? v.A.yp : v.A.xp;
    const T *o2 = ax == 2 ? v.A.ym : v.A.zm, *o3 = ax == 2 ? v.A.yp : v.A.zp;
    const __bf16 *h0p = ax == 0 ? v.o16[2] : v.o16[0], *h1p = ax == 0 ? v.o16[3] : v.o16[1];
    const __bf16 *h2p = ax == 2 ? v.o16[2] : v.o16[4], *h3p = ax == 2 ? v.o16[3] : v.o16[5];
    T vb[M], vai[M], viw[M], vcp[M], cf[M][4], xn[M][4], pwn[pro ? M : 1][4], xcn[pro ? M : 1][4];
    int msk[M];
#pragma unroll
    for (int m = 0; m < M; m++) {
        const int e = threadIdx.x + LS_THREADS * m;
        int L, q;
        if (ax == 0) { L = e / n; q = e - L * n; } else { q = e / NLW; L = e - q * NLW; }
        const int td = 2 * (h0 + L) + sh;
        const bool ok = e < E && td < g.Ntd;
        int i, j, l;
        if (ax == 0) { i = q; j = td; l = p3; } else if (ax == 1) { i = td; j = q; l = p3; } else { i = td; j = p3; l = q; }
        const int k = ok ? mg_idx(v, i, j, l) : 0;
        const int u1 = ax == 0 ? j : i, n1 = ax == 0 ? Ny : Nx;
        const int u2 = ax == 2 ? j : l, n2 = ax == 2 ? Ny : Nz;
        const bool hm1 = ok && u1 > 0, hp1 = ok && u1 < n1 - 1, hm2 = ok && u2 > 0, hp2 = ok && u2 < n2 - 1;
        // off-line neighbours: along x through the storage order, else strided
        const int kx = k - mg_xpos(v, i);
        const int m1 = ax == 0 ? k - s1 : kx + mg_xpos(v, i - 1), p1 = ax == 0 ? k + s1 : kx + mg_xpos(v, i + 1);
        const int nk[4] = {hm1 ? m1 : k, hp1 ? p1 : k, hm2 ? k - s2 : k, hp2 ? k + s2 : k};
        msk[m] = (int)ok | ((int)hm1 << 1) | ((int)hp1 << 2) | ((int)hm2 << 3) | ((int)hp2 << 4);
        vb[m] = b[k];
        if constexpr (H16) {
            vai[m] = (T)AI16[k]; viw[m] = (T)IW16[k]; vcp[m] = (T)CP16[k];
        } else {
            vai[m] = AI[k]; viw[m] = IW[k]; vcp[m] = CP[k];
        }
        if constexpr (!zero) {
            if constexpr (H16) {
                cf[m][0] = (T)h0p[k]; cf[m][1] = (T)h1p[k]; cf[m][2] = (T)h2p[k]; cf[m][3] = (T)h3p[k];
            } else {
                cf[m][0] = o0[k]; cf[m][1] = o1[k]; cf[m][2] = o2[k]; cf[m][3] = o3[k];
            }
#pragma unroll
            for (int t = 0; t < 4; t++) xn[m][t] = x[nk[t]];
        }
        if constexpr (pro) {
            // the neighbours' coarse parents (first post-smoothing sweep)
#pragma unroll
            for (int t = 0; t < 4; t++) {
                const int d = (t & 1) ? 1 : -1, on = (msk[m] >> (t + 1)) & 1;
                int ii = i, jj = j, ll = l;
                if (t < 2) { if (ax == 0) jj += on * d; else ii += on * d; }
                else { if (ax == 2) jj += on * d; else ll += on * d; }
                const int I = v.cx ? ii >> 1 : ii, J = v.cy ? jj >> 1 : jj, LL = v.cz ? ll >> 1 : ll;
                pwn[m][t] = v.pw[nk[t]];
                xcn[m][t] = c.x[ok ? mg_idx(c, I, J, LL) : 0];
            }
        }
    }
    // second pass: right-hand sides into LDS
#pragma unroll
    for (int m = 0; m < M; m++) {
        const int e = threadIdx.x + LS_THREADS * m;
        if (e >= E) break;
        int L, q;
        if (ax == 0) { L = e / n; q = e - L * n; } else { q = e / NLW; L = e - q * NLW; }
        T r = 0, ai = 0, iw = 0, cp = 0;
        if (msk[m] & 1) {
            T sr = vb[m];
            if constexpr (!zero) {
#pragma unroll
                for (int t = 0; t < 4; t++) {
                    if (!((msk[m] >> (t + 1)) & 1)) continue;
                    if constexpr (pro) sr -= cf[m][t] * (xn[m][t] + pwn[m][t] * xcn[m][t]);
                    else sr -= cf[m][t] * xn[m][t];
                }
            }
            r = sr; ai = vai[m]; iw = viw[m]; cp = vcp[m];
        }
        const int q0 = (L / LW) * SP + (L % LW) * n + q;
        s_r[q0] = r * iw; s_a[q0] = ai; s_c[q0] = cp;
    }
    __syncthreads();
    line_scan_slots<T, P, S>(s_r, s_a, s_c);
    __syncthreads();
#pragma unroll
    for (int m = 0; m < P * S; m++) {
        const int e = threadIdx.x + LS_THREADS * m;
        if (e >= E) break;
        int L, q;
        if (ax == 0) { L = e / n; q = e - L * n; } else { q = e / NLW; L = e - q * NLW; }
        const int td = 2 * (h0 + L) + sh;
        if (td < g.Ntd) {
            const int k = ax == 0 ? mg_idx(v, q, td, p3) : (ax == 1 ? mg_idx(v, td, q, p3) : mg_idx(v, td, p3, q));
            x[k] = s_r[(L / LW) * SP + (L % LW) * n + q];
        }
    }
}
// The same sweep for one whole line per slot (LW = 1: lines of 64 (P-1) + 1
// to 64 P elements), with the axis a template argument: lines per workgroup
// is then a compile-time power of two, the indices of every element are
// computed once and kept for the three phases, and the x positions of a
// line and of its neighbours across x follow from the line's parity without
// division.  HM: coefficients in T (0), in separate BF16 arrays (1), or in
// one packed 16-byte BF16 record per node (2).  The choice is a template
// argument: a run-time branch between the layouts joins after each node's
// loads, so the compiler waits for every load in flight there and a lane's
// gathers no longer overlap.
template <class T, int P, int S, int MD, int HM, int AX>
__global__ void __launch_bounds__(LS_THREADS) k_line_scan1(MGLevel<T> v, const T *__restrict__ b, T *__restrict__ x,
                                                           int color, MGLevel<T> c, int xcd) {
    constexpr bool zero = MD == 0, pro = MD == 2, H16 = HM != 0;
    constexpr int SL = 64 * P, SP = SL + 1, NLW = LS_WAVES * S, M = P * S, NT = LS_THREADS;
    __shared__ T s_r[NLW * SP], s_a[NLW * SP], s_c[NLW * SP];
    const int Nx = v.Nx, Ny = v.Ny, Nz = v.Nz, NxNy = Nx * Ny;
    const unsigned NU = (unsigned)v.N;
    const int n = AX == 0 ? Nx : (AX == 1 ? Ny : Nz);        // line length
    const int Ntd = AX == 0 ? Ny : Nx;                       // lines of both colours per plane
    const int ntile = (((Ntd + 1) >> 1) + NLW - 1) / NLW;
    const int bid = xcd ? (blockIdx.x & 7) * (gridDim.x >> 3) + (blockIdx.x >> 3) : blockIdx.x;
    const int p3 = bid / ntile, h0 = (bid - p3 * ntile) * NLW, sh = (p3 + color) & 1;
    const int E = NLW * n, pad = SL - n, hx = (Nx + 1) >> 1;
    // LDS position of element q of line L of the tile
    auto lds = [&](int L, int q) { return L * SP + q; };
    for (int t = threadIdx.x; t < NLW * pad; t += NT) {
        const int q = lds(t / pad, n + t % pad);
        s_r[q] = 0; s_a[q] = 0; s_c[q] = 0;
    }
    constexpr int a_lu = AX;
    const T *o0 = AX == 0 ? v.A.ym : v.A.xm, *o1 = AX == 0 ? v.A.yp : v.A.xp;
    const T *o2 = AX == 2 ? v.A.ym : v.A.zm, *o3 = AX == 2 ? v.A.yp : v.A.zp;
    const __bf16 *h0p = AX == 0 ? v.o16[2] : v.o16[0], *h1p = AX == 0 ? v.o16[3] : v.o16[1];
    const __bf16 *h2p = AX == 2 ? v.o16[2] : v.o16[4], *h3p = AX == 2 ? v.o16[3] : v.o16[5];
    const float rn = 1.0f / (float)n;
    // line L of the tile and position q along it of element e
    auto split = [&](int e, int &L, int &q) {
        if constexpr (AX == 0) {
            // e / n exactly: (e + 1/2) / n is at least 1/(2n) away from an integer
            L = (int)(((float)e + 0.5f) * rn);
            q = e - L * n;
        } else {
            q = e / NLW;
            L = e % NLW;
        }
    };
    unsigned kk[M];
    int msk[M];
    T vb[M], vai[M], viw[M], vcp[M], cf[M][4], xn[M][4], pwn[pro ? M : 1][4], xcn[pro ? M : 1][4];
#pragma unroll
    for (int m = 0; m < M; m++) {
        const int e = threadIdx.x + NT * m;
        int L, q;
        split(e, L, q);
        const int a = h0 + L, td = 2 * a + sh;
        const bool ok = e < E && td < Ntd;
        int i, j, l;
        if constexpr (AX == 0) { i = q; j = td; l = p3; }
        else if constexpr (AX == 1) { i = td; j = q; l = p3; }
        else { i = td; j = p3; l = q; }
        const int row = Nx * (j + Ny * l);
        int k, km1, kp1, km2, kp2;
        bool hm1, hp1, hm2, hp2;
        if constexpr (AX == 0) {
            k = row + (v.px ? (q & 1) * hx + (q >> 1) : q);
            hm1 = j > 0; hp1 = j < Ny - 1; km1 = k - Nx; kp1 = k + Nx;
            hm2 = l > 0; hp2 = l < Nz - 1; km2 = k - NxNy; kp2 = k + NxNy;
        } else {
            // x = td = 2 a + sh; x - 1 and x + 1 have the other parity
            const int xi = v.px ? sh * hx + a : td;
            const int xm = v.px ? (1 - sh) * hx + a - 1 + sh : td - 1;
            const int xp = v.px ? (1 - sh) * hx + a + sh : td + 1;
            k = row + xi; km1 = row + xm; kp1 = row + xp;
            hm1 = td > 0; hp1 = td < Nx - 1;
            if constexpr (AX == 1) { hm2 = l > 0; hp2 = l < Nz - 1; km2 = k - NxNy; kp2 = k + NxNy; }
            else { hm2 = j > 0; hp2 = j < Ny - 1; km2 = k - Nx; kp2 = k + Nx; }
        }
        hm1 = hm1 && ok; hp1 = hp1 && ok; hm2 = hm2 && ok; hp2 = hp2 && ok;
        if (!ok) k = 0;
        const unsigned nk[4] = {(unsigned)(hm1 ? km1 : k), (unsigned)(hp1 ? kp1 : k), (unsigned)(hm2 ? km2 : k),
                                (unsigned)(hp2 ? kp2 : k)};
        msk[m] = (int)ok | ((int)hm1 << 1) | ((int)hp1 << 2) | ((int)hm2 << 3) | ((int)hp2 << 4);
        kk[m] = (unsigned)k;
        const unsigned ku = (unsigned)k;
        vb[m] = ld_buf(b, NU, ku);
        if constexpr (HM == 2) {
            // the node's seven BF16 coefficients in one 16-byte load (a
            // BF16 value is the upper half of the FP32 word)
            const uint4 r = ld_buf(v.rec16[a_lu], NU, ku);
            vai[m] = __uint_as_float(r.x << 16); viw[m] = __uint_as_float(r.x & 0xffff0000u);
            vcp[m] = __uint_as_float(r.y << 16);
            if constexpr (!zero) {
                cf[m][0] = __uint_as_float(r.y & 0xffff0000u); cf[m][1] = __uint_as_float(r.z << 16);
                cf[m][2] = __uint_as_float(r.z & 0xffff0000u); cf[m][3] = __uint_as_float(r.w << 16);
            }
        } else if constexpr (HM == 1) {
            vai[m] = (T)ld_buf(v.lu16[a_lu][0], NU, ku); viw[m] = (T)ld_buf(v.lu16[a_lu][1], NU, ku);
            vcp[m] = (T)ld_buf(v.lu16[a_lu][2], NU, ku);
            if constexpr (!zero) {
                cf[m][0] = (T)ld_buf(h0p, NU, ku); cf[m][1] = (T)ld_buf(h1p, NU, ku);
                cf[m][2] = (T)ld_buf(h2p, NU, ku); cf[m][3] = (T)ld_buf(h3p, NU, ku);
            }
        } else {
            vai[m] = ld_buf(v.lu[a_lu][0], NU, ku); viw[m] = ld_buf(v.lu[a_lu][1], NU, ku); vcp[m] = ld_buf(v.lu[a_lu][2], NU, ku);
        }
        if constexpr (!zero) {
            if constexpr (!H16) {
                cf[m][0] = ld_buf(o0, NU, ku); cf[m][1] = ld_buf(o1, NU, ku); cf[m][2] = ld_buf(o2, NU, ku); cf[m][3] = ld_buf(o3, NU, ku);
            }
            // a neighbour outside the mesh reads as 0 (its index is the
            // node's own, whose x the sweep has not necessarily written yet)
            const bool hv[4] = {hm1, hp1, hm2, hp2};
#pragma unroll
            for (int t = 0; t < 4; t++) {
                const T xv = ld_buf(x, NU, nk[t]);
                xn[m][t] = hv[t] ? xv : (T)0;
            }
        }
        if constexpr (pro) {
            // the neighbours' coarse parents (first post-smoothing sweep)
#pragma unroll
            for (int t = 0; t < 4; t++) {
                const int d = (t & 1) ? 1 : -1, on = (msk[m] >> (t + 1)) & 1;
                int ii = i, jj = j, ll = l;
                if (t < 2) { if (AX == 0) jj += on * d; else ii += on * d; }
                else { if (AX == 2) jj += on * d; else ll += on * d; }
                const int I = v.cx ? ii >> 1 : ii, J = v.cy ? jj >> 1 : jj, LL = v.cz ? ll >> 1 : ll;
                const T pv = ld_buf(v.pw, NU, nk[t]);
                const T cv = ld_buf(c.x, (unsigned)c.N, (unsigned)(ok ? mg_idx(c, I, J, LL) : 0));
                pwn[m][t] = on ? pv : (T)0;
                xcn[m][t] = on ? cv : (T)0;
            }
        }
    }
    // right-hand sides into LDS
#pragma unroll
    for (int m = 0; m < M; m++) {
        const int e = threadIdx.x + NT * m;
        if (e >= E) break;
        int L, q;
        split(e, L, q);
        T r = 0, ai = 0, iw = 0, cp = 0;
        if (msk[m] & 1) {
            T sr = vb[m];
            if constexpr (!zero) {
#pragma unroll
                for (int t = 0; t < 4; t++) {
                    // the packed records hold 0 for a neighbour outside the
                    // mesh, and its x, weight and coarse value were read as 0:
                    // nothing to skip
                    if constexpr (HM != 2)
                        if (!((msk[m] >> (t + 1)) & 1)) continue;
                    if constexpr (pro) sr -= cf[m][t] * (xn[m][t] + pwn[m][t] * xcn[m][t]);
                    else sr -= cf[m][t] * xn[m][t];
                }
            }
            r = sr; ai = vai[m]; iw = viw[m]; cp = vcp[m];
        }
        const int q0 = lds(L, q);
        s_r[q0] = r * iw; s_a[q0] = ai; s_c[q0] = cp;
    }
    __syncthreads();
    line_scan_slots<T, P, S>(s_r, s_a, s_c);
    __syncthreads();
#pragma unroll
    for (int m = 0; m < M; m++) {
        if (!(msk[m] & 1)) continue;
        int L, q;
        split(threadIdx.x + NT * m, L, q);
        st_buf(x, NU, kk[m], s_r[lds(L, q)]);
    }
}
// slots per wavefront allowed by the 64 KB of LDS a workgroup may hold
template <class T> constexpr int ls_smax(int P) {
    return 3 * LS_WAVES * (64 * P + 1) * 4 * (int)sizeof(T) <= 65536 ? 4
         : (3 * LS_WAVES * (64 * P + 1) * 2 * (int)sizeof(T) <= 65536 ? 2 : 1);
}
template <class T, int P, int MD, bool H>
static void ls_launch_h(hipStream_t s, int S, int nb, const MGLevel<T> &v, const T *b, T *x, int ax, int color,
                        const MGLevel<T> &cc, int LW, int xcd) {
    constexpr int SM = ls_smax<T>(P);
    if (S >= 4 && SM >= 4)
        k_line_scan<T, P, (SM >= 4 ? 4 : 1), MD, H><<<nb, LS_THREADS, 0, s>>>(v, b, x, ax, color, cc, LW, xcd);
    else if (S >= 2 && SM >= 2)
        k_line_scan<T, P, (SM >= 2 ? 2 : 1), MD, H><<<nb, LS_THREADS, 0, s>>>(v, b, x, ax, color, cc, LW, xcd);
    else
        k_line_scan<T, P, 1, MD, H><<<nb, LS_THREADS, 0, s>>>(v, b, x, ax, color, cc, LW, xcd);
}
template <class T, int P, int MD>
static void ls_launch_md(hipStream_t s, int S, int nb, const MGLevel<T> &v, const T *b, T *x, int ax, int color,
                         const MGLevel<T> &cc, int LW, int xcd) {
    // BF16 coefficients exist only for FP32 levels (and lines up to 128 long)
    if constexpr (std::is_same_v<T, float> && P <= 2) {
        if (v.o16[0] && v.lu16[ax][0]) {
            ls_launch_h<T, P, MD, true>(s, S, nb, v, b, x, ax, color, cc, LW, xcd);
            return;
        }
    }
    ls_launch_h<T, P, MD, false>(s, S, nb, v, b, x, ax, color, cc, LW, xcd);
}
int g_line_scan1 = 1;         // 1: axis-specialised kernel for one line per slot
// launch of k_line_scan1 (one line per slot): slots per wavefront S, sweep
// kind MD, coefficient layout H, axis from the runtime value
template <class T, int P, int MD, int H, int AX>
static void ls1_launch_ax(hipStream_t s, int S, int nb, const MGLevel<T> &v, const T *b, T *x, int color,
                          const MGLevel<T> &cc, int xcd) {
    constexpr int SM = ls_smax<T>(P);
    if (S >= 4 && SM >= 4)
        k_line_scan1<T, P, (SM >= 4 ? 4 : 1), MD, H, AX><<<nb, LS_THREADS, 0, s>>>(v, b, x, color, cc, xcd);
    else if (S >= 2 && SM >= 2)
        k_line_scan1<T, P, (SM >= 2 ? 2 : 1), MD, H, AX><<<nb, LS_THREADS, 0, s>>>(v, b, x, color, cc, xcd);
    else
        k_line_scan1<T, P, 1, MD, H, AX><<<nb, LS_THREADS, 0, s>>>(v, b, x, color, cc, xcd);
}
template <class T, int P, int MD, int H>
static void ls1_launch_h(hipStream_t s, int S, int nb, const MGLevel<T> &v, const T *b, T *x, int ax, int color,
                         const MGLevel<T> &cc, int xcd) {
    if (ax == 0) ls1_launch_ax<T, P, MD, H, 0>(s, S, nb, v, b, x, color, cc, xcd);
    else if (ax == 1) ls1_launch_ax<T, P, MD, H, 1>(s, S, nb, v, b, x, color, cc, xcd);
    else ls1_launch_ax<T, P, MD, H, 2>(s, S, nb, v, b, x, color, cc, xcd);
}
template <class T, int P, int MD>
static void ls1_launch_md(hipStream_t s, int S, int nb, const MGLevel<T> &v, const T *b, T *x, int ax, int color,
                          const MGLevel<T> &cc, int xcd) {
    if constexpr (std::is_same_v<T, float> && P <= 2) {
        if (v.rec16[ax]) {
            ls1_launch_h<T, P, MD, 2>(s, S, nb, v, b, x, ax, color, cc, xcd);
            return;
        }
        if (v.o16[0] && v.lu16[ax][0]) {
            ls1_launch_h<T, P, MD, 1>(s, S, nb, v, b, x, ax, color, cc, xcd);
            return;
        }
    }
    ls1_launch_h<T, P, MD, 0>(s, S, nb, v, b, x, ax, color, cc, xcd);
}
template <class T, int P>
static void ls_launch(hipStream_t s, int S, int nb, const MGLevel<T> &v, const T *b, T *x, int ax, int color, int zero,
                      const MGLevel<T> &cc, int pro, int LW, int xcd) {
    // k_line_scan1 addresses its arrays with 32-bit byte offsets (16-byte records)
    if (LW == 1 && g_line_scan1 && (size_t)v.N * 16 < ((size_t)1 << 32)) {
        if (zero) ls1_launch_md<T, P, 0>(s, S, nb, v, b, x, ax, color, cc, xcd);
        else if (pro) ls1_launch_md<T, P, 2>(s, S, nb, v, b, x, ax, color, cc, xcd);
        else ls1_launch_md<T, P, 1>(s, S, nb, v, b, x, ax, color, cc, xcd);
        return;
    }
    if (zero) ls_launch_md<T, P, 0>(s, S, nb, v, b, x, ax, color, cc, LW, xcd);
    else if (pro) ls_launch_md<T, P, 2>(s, S, nb, v, b, x, ax, color, cc, LW, xcd);
    else ls_launch_md<T, P, 1>(s, S, nb, v, b, x, ax, color, cc, LW, xcd);
}

int g_line_nl_max = 8;       // lines per workgroup cap of k_line_gs (tuning)
int g_line_scan = 1;         // 1: scan kernel for lines of up to 256 elements
int g_line_slots = 4;        // most line slots per wavefront of the scan kernel (1, 2, 4)
int g_line_minwg = 512;      // fewest workgroups a wider slot choice may leave
int g_line_xcd = 1;          // XCD-contiguous tile order
template <class T>
int mg_line(hipStream_t s, const MGLevel<T> &v, const T *b, T *x, int ax, int color, int zero, const MGLevel<T> *c) {
    int n = ax == 0 ? v.Nx : (ax == 1 ? v.Ny : v.Nz);
    int Ntd = ax == 0 ? v.Ny : v.Nx, N3 = ax == 2 ? v.Ny : v.Nz;
    if (!v.lu[ax][0]) return 0;
    const int P = (n + 63) / 64;
    if (g_line_scan && P <= 4) {
        const int LW = 64 * P / n;
        const int lpp = (Ntd + 1) / 2, npl = N3;      // lines per plane and colour, planes
        // lines across the workgroup (y- and z-lines): widen the contiguous
        // runs with more slots per wavefront while at least g_line_minwg
        // workgroups remain
        int S = 1;
        if (ax != 0)
            for (int sc = g_line_slots; sc > 1; sc >>= 1) {
                int nlw = LS_WAVES * sc * LW;
                if ((long)((lpp + nlw - 1) / nlw) * npl >= g_line_minwg) { S = sc; break; }
            }
        // at most the slots the LDS allows (ls_launch makes the same choice)
        const int SM = P == 1 ? ls_smax<T>(1) : (P == 2 ? ls_smax<T>(2) : (P == 3 ? ls_smax<T>(3) : ls_smax<T>(4)));
        int Srun = S;
        while (Srun > SM) Srun >>= 1;
        const int NLWr = LS_WAVES * Srun * LW, nb = (lpp + NLWr - 1) / NLWr * npl;
        if (nb == 0) return 1;
        const int xcd = g_line_xcd && (nb & 7) == 0;      // the remap is a bijection for nb % 8 == 0
        MGLevel<T> cc = c ? *c : v;
        const int pro = c != nullptr;
        switch (P) {
        case 1: ls_launch<T, 1>(s, Srun, nb, v, b, x, ax, color, zero, cc, pro, LW, xcd); break;
        case 2: ls_launch<T, 2>(s, Srun, nb, v, b, x, ax, color, zero, cc, pro, LW, xcd); break;
        case 3: ls_launch<T, 3>(s, Srun, nb, v, b, x, ax, color, zero, cc, pro, LW, xcd); break;
        default: ls_launch<T, 4>(s, Srun, nb, v, b, x, ax, color, zero, cc, pro, LW, xcd); break;
        }
        return 1;
    }
    if (n > LINE_MAXEQ) return 0;
    int NL = LINE_MAXEQ / n;
    if (NL > g_line_nl_max) NL = g_line_nl_max;
    int half = (Ntd + 1) / 2;
    if (NL > half) NL = half;
    if (NL < 1) NL = 1;
    int ntile = (Ntd + 2 * NL - 1) / (2 * NL);
    MGLevel<T> cc = c ? *c : v;
    k_line_gs<T><<<ntile * N3, SD_BLOCK, 0, s>>>(v, b, x, ax, color, zero, cc, c != nullptr, NL);
    return 1;
}

// Residual of the fine level summed over each aggregate with the fine row
// scale (restriction R = P^T W, semiconductor_3d.c:2028-2037), divided by the
// coarse row scale (the coarse rows are stored with unit diagonal).  Rows
// solved exactly by the last smoothing sweep have no residual and are
// skipped: skip 1 = point black nodes, 2+a = colour-1 lines along axis a.
// One lane per fine node: the eight lanes of an aggregate are adjacent (lane
// bits 0, 1, 2 = x, y, z offset in the aggregate), each evaluates its node's
// weighted residual, and three xor shuffles sum the aggregate, so every lane
// issues its loads at once.  t indexes lanes 8 K + sub; every lane of the
// wavefront must call it (the shuffles), lanes past 8 c.N contribute nothing.
// weighted residual of fine node (2I+di, 2J+dj, 2L+dl) of coarse node K
// (0 outside the aggregate or on a skipped row)
// NT: the operator and row scales are read with non-temporal loads (level 0
// with BF16 smoother coefficients, where nothing else reads them in the V-cycle)
// UD: unit diagonal (f.A.c not read), known at compile time
template <class T, bool NT = false, bool UD = false>
__device__ __forceinline__ double agg_resid(const MGLevel<T> &f, const MGLevel<T> &c, const T *b, const T *x, int skip,
                                            int I, int J, int LL, int di, int dj, int dl) {
    const int i = f.cx ? 2 * I + di : I, j = f.cy ? 2 * J + dj : J, l = f.cz ? 2 * LL + dl : LL;
    bool ok = (f.cx || !di) && (f.cy || !dj) && (f.cz || !dl) && i < f.Nx && j < f.Ny && l < f.Nz;
    if (ok && skip == 1 && ((i + j + l) & 1)) ok = false;
    if (ok && skip >= 2) {
        const int ax = skip - 2, par = ax == 0 ? (j + l) : (ax == 1 ? (i + l) : (i + j));
        if (par & 1) ok = false;
    }
    if (!ok) return 0.0;
    auto ld = [](const auto *p, int k) {
        if constexpr (NT) return __builtin_nontemporal_load(p + k);
        else return p[k];
    };
    const int k = mg_idx(f, i, j, l), kx = k - mg_xpos(f, i);
    const double w = f.rs ? ld(f.rs, k) : 1.0;
    const int NxNy = f.Nx * f.Ny;
    const bool h0 = i > 0, h1 = i < f.Nx - 1, h2 = j > 0, h3 = j < f.Ny - 1, h4 = l > 0, h5 = l < f.Nz - 1;
    T s = UD ? x[k] : (f.A.c ? ld(f.A.c, k) * x[k] : x[k]);
    s = nb_add<T>(h0, s, ld(f.A.xm, k), x[nb_idx(h0, kx + mg_xpos(f, i - 1), k)]);
    s = nb_add<T>(h1, s, ld(f.A.xp, k), x[nb_idx(h1, kx + mg_xpos(f, i + 1), k)]);
    s = nb_add<T>(h2, s, ld(f.A.ym, k), x[nb_idx(h2, k - f.Nx, k)]);
    s = nb_add<T>(h3, s, ld(f.A.yp, k), x[nb_idx(h3, k + f.Nx, k)]);
    s = nb_add<T>(h4, s, ld(f.A.zm, k), x[nb_idx(h4, k - NxNy, k)]);
    s = nb_add<T>(h5, s, ld(f.A.zp, k), x[nb_idx(h5, k + NxNy, k)]);
    T r = b[k] - s;
    return w == 0.0 ? 0.0 : w * (double)r;
}
template <class T>
__device__ __forceinline__ double agg_resid(const MGLevel<T> &f, const MGLevel<T> &c, const T *b, const T *x, int skip,
                                            int K, int di, int dj, int dl) {
    const int cNxNy = c.Nx * c.Ny, LL = K / cNxNy, rr = K - LL * cNxNy, J = rr / c.Nx, I = rr - J * c.Nx;
    return agg_resid(f, c, b, x, skip, I, J, LL, di, dj, dl);
}
// storage position of coarse node K (natural numbering)
template <class T> __device__ __forceinline__ int mg_kpos(const MGLevel<T> &c, int K) {
    const int cNxNy = c.Nx * c.Ny, LL = K / cNxNy, rr = K - LL * cNxNy, J = rr / c.Nx, I = rr - J * c.Nx;
    return mg_idx(c, I, J, LL);
}
template <class T>
__device__ __forceinline__ void restrict8(const MGLevel<T> &f, const MGLevel<T> &c, const T *b, const T *x, int skip,
                                          int t) {
    const int K = t >> 3, sub = t & 7;
    double v = K < c.N ? agg_resid(f, c, b, x, skip, K, sub & 1, (sub >> 1) & 1, sub >> 2) : 0.0;
    v += __shfl_xor(v, 1, 64);
    v += __shfl_xor(v, 2, 64);
    v += __shfl_xor(v, 4, 64);
    if (sub == 0 && K < c.N) { const int Kc = mg_kpos(c, K); c.b[Kc] = (T)(v / c.rs[Kc]); }
}
// Level-wide restriction: a workgroup of four wavefronts covers 32
// consecutive coarse nodes; wavefront w takes the fine rows (dj, dl) =
// (w & 1, w >> 1) and lane pairs the two x offsets, so each load of a
// wavefront spans 64 consecutive fine nodes of one row.  Pair sums by
// shuffle, then the four rows through LDS.
template <class T>
__global__ void __launch_bounds__(256) k_resid_restrict(MGLevel<T> f, MGLevel<T> c, const T *b, const T *x, int skip,
                                                        double alpha) {
    __shared__ double part[4][32];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, tb = sd_tile();
    const int K = tb * 32 + (lane >> 1);
    double v = K < c.N ? agg_resid(f, c, b, x, skip, K, lane & 1, w & 1, w >> 1) : 0.0;
    v += __shfl_xor(v, 1, 64);
    if (!(lane & 1)) part[w][lane >> 1] = v;
    __syncthreads();
    if (threadIdx.x < 32) {
        const int K2 = tb * 32 + threadIdx.x;
        if (K2 < c.N) {
            double sum = part[0][threadIdx.x] + part[1][threadIdx.x] + part[2][threadIdx.x] + part[3][threadIdx.x];
            const int Kc = mg_kpos(c, K2);
            c.b[Kc] = (T)(alpha * sum / c.rs[Kc]);
        }
    }
}
// The same restriction when the last pre-smoothing sweep solved half of every
// aggregate exactly (skip != 0) and all three axes are coarsened: only the 4
// fine nodes of an aggregate whose residual is not 0 get a lane.  Wavefront w
// takes the w-th of these offsets (di, dj, dl) for 64 consecutive coarse
// nodes, so every lane works and each load spans 64 consecutive fine nodes of
// one parity (contiguous in the parity-split order).  The four partial sums
// meet in LDS and are added in offset order.
__device__ __forceinline__ void act_offset(int skip, int w, int &di, int &dj, int &dl) {
    const int lo = w & 1, hi = w >> 1;
    switch (skip) {
    case 1: di = lo ^ hi; dj = lo; dl = hi; break;    // (i + j + l) even
    case 2: di = lo; dj = hi; dl = hi; break;         // x lines: (j + l) even
    case 3: di = hi; dj = lo; dl = hi; break;         // y lines: (i + l) even
    default: di = hi; dj = hi; dl = lo; break;        // z lines: (i + j) even
    }
}
template <class T, bool NT, bool UD>
__global__ void __launch_bounds__(256) k_resid_restrict_act(MGLevel<T> f, MGLevel<T> c, const T *b, const T *x,
                                                            int skip, double alpha) {
    __shared__ double part[4][64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, tb = sd_tile();
    const int K = tb * 64 + lane;
    int di, dj, dl;
    act_offset(skip, w, di, dj, dl);
    // coarse coordinates: the workgroup's first node by division, then a
    // carry along x and y for the lane
    const int cNxNy = c.Nx * c.Ny, K0 = tb * 64, L0 = K0 / cNxNy, r0 = K0 - L0 * cNxNy, J0 = r0 / c.Nx;
    int I = r0 - J0 * c.Nx + lane, J = J0, LL = L0;
    while (I >= c.Nx) {
        I -= c.Nx;
        if (++J == c.Ny) { J = 0; LL++; }
    }
    part[w][lane] = K < c.N ? agg_resid<T, NT, UD>(f, c, b, x, skip, I, J, LL, di, dj, dl) : 0.0;
    __syncthreads();
    if (threadIdx.x < 64 && K < c.N) {
        const double sum = part[0][lane] + part[1][lane] + part[2][lane] + part[3][lane];
        const int Kc = mg_idx(c, I, J, LL);
        c.b[Kc] = (T)(alpha * sum / c.rs[Kc]);
    }
}
int g_restrict_act = 1;      // 1: active-lane restriction where it applies
template <class T> void mg_resid_restrict(hipStream_t s, const MGLevel<T> &f, const MGLevel<T> &c, const T *b, const T *x,
                                          int skip, double alpha) {
    if (g_restrict_act && skip >= 1 && skip <= 4 && f.cx && f.cy && f.cz) {
        const int nb = (c.N + 63) / 64;
        const bool nt = f.o16[0] && g_spmv_nt, ud = f.A.c == nullptr;
        if (nt && ud) k_resid_restrict_act<T, true, true><<<nb, 256, 0, s>>>(f, c, b, x, skip, alpha);
        else if (nt) k_resid_restrict_act<T, true, false><<<nb, 256, 0, s>>>(f, c, b, x, skip, alpha);
        else if (ud) k_resid_restrict_act<T, false, true><<<nb, 256, 0, s>>>(f, c, b, x, skip, alpha);
        else k_resid_restrict_act<T, false, false><<<nb, 256, 0, s>>>(f, c, b, x, skip, alpha);
    }
    else
        k_resid_restrict<T><<<(c.N + 31) / 32, 256, 0, s>>>(f, c, b, x, skip, alpha);
}

// The coarse end of the V-cycle in one workgroup: every level from lv[0]
// (whose b holds the restricted residual) to the coarsest lv[nt-1] (dense
// inverse), with nsw red-black Gauss-Seidel sweeps before and after each
// coarse correction, the same arithmetic as the per-level kernels, separated
// by workgroup barriers instead of kernel launches.  Result in lv[0].x.
#define TAIL_BLOCK 1024
template <class T>
__global__ void __launch_bounds__(TAIL_BLOCK) k_tail(const MGLevel<T> *lv, int nt, const double *inv, int nsw) {
    const int tid = threadIdx.x;
    for (int L = 0; L < nt - 1; L++) {
        const MGLevel<T> v = lv[L];
        const int half = ((v.Nx + 1) >> 1) * v.Ny * v.Nz;
        for (int sw = 0; sw < nsw; sw++) {
            if (sw == 0) { for (int k = tid; k < v.N; k += TAIL_BLOCK) rb_zero_node(v, v.b, v.x, 0, k); }
            else { for (int t = tid; t < half; t += TAIL_BLOCK) rb_node(v, v.b, v.x, 0, t, v, 0); }
            __syncthreads();
            for (int t = tid; t < half; t += TAIL_BLOCK) rb_node(v, v.b, v.x, 1, t, v, 0);
            __syncthreads();
        }
        const MGLevel<T> c = lv[L + 1];
        for (int t0 = 0; t0 < 8 * c.N; t0 += TAIL_BLOCK) restrict8(v, c, v.b, v.x, 1, t0 + tid);
        __syncthreads();
    }
    {
        const MGLevel<T> z = lv[nt - 1];
        const int n = z.N;
        for (int r = tid; r < n; r += TAIL_BLOCK) {
            double s = 0.0;
            for (int q = 0; q < n; q++) s += inv[r * n + q] * (double)z.b[q];
            z.x[r] = (T)s;
        }
        __syncthreads();
    }
    for (int L = nt - 2; L >= 0; L--) {
        const MGLevel<T> v = lv[L], c = lv[L + 1];
        const int half = ((v.Nx + 1) >> 1) * v.Ny * v.Nz;
        for (int sw = 0; sw < nsw; sw++) {
            for (int t = tid; t < half; t += TAIL_BLOCK) rb_node(v, v.b, v.x, 1, t, c, sw == 0);
            __syncthreads();
            for (int t = tid; t < half; t += TAIL_BLOCK) rb_node(v, v.b, v.x, 0, t, v, 0);
            __syncthreads();
        }
    }
}
// The same coarse end with every level's x and b held in LDS and each
// thread's row of every level (node t of level L) in registers, loaded once:
// after the first loads each phase is LDS work between barriers.  The
// prolongation is applied explicitly before the post-smoothing sweeps (the
// same values the fused form reads).  Up to TL_MAXT levels, the first of at
// most TAIL_BLOCK nodes, TL_MAXN nodes in all, the coarsest of at most 64.
#define TL_MAXT 4
#define TL_MAXN 2048
int g_tail_lds = 1;          // 1: LDS-resident tail kernel where the levels fit
template <class T>
__global__ void __launch_bounds__(TAIL_BLOCK) k_tail_lds(const MGLevel<T> *lv, int nt, const double *inv, int nsw) {
    __shared__ T sx[TL_MAXN], sb[TL_MAXN];
    __shared__ double sr[TAIL_BLOCK];
    const int t = threadIdx.x;
    int off[TL_MAXT + 1];
    off[0] = 0;
    for (int L = 0; L < nt; L++) off[L + 1] = off[L] + lv[L].N;
    T a[TL_MAXT][6], pw[TL_MAXT];
    double rs[TL_MAXT];
#pragma unroll
    for (int L = 0; L < TL_MAXT; L++) {
        pw[L] = 0; rs[L] = 1.0;
        for (int q = 0; q < 6; q++) a[L][q] = 0;
        if (L < nt && t < lv[L].N) {
            const MGLevel<T> &v = lv[L];
            if (L < nt - 1) {
                a[L][0] = v.A.xm[t]; a[L][1] = v.A.xp[t]; a[L][2] = v.A.ym[t];
                a[L][3] = v.A.yp[t]; a[L][4] = v.A.zm[t]; a[L][5] = v.A.zp[t];
            }
            pw[L] = v.pw[t];
            rs[L] = v.rs ? v.rs[t] : 1.0;
        }
    }
    if (t < lv[0].N) sb[t] = lv[0].b[t];
    __syncthreads();
    // x of node t of level L from its neighbours in LDS, colour `color`
    auto sweep = [&](int L, int color, int zero) {
        const MGLevel<T> &v = lv[L];
        if (t < v.N) {
            const int Nx = v.Nx, NxNy = v.Nx * v.Ny;
            const int l = t / NxNy, r = t - l * NxNy, j = r / Nx, i = r - j * Nx;
            T *x = sx + off[L];
            if (((i + j + l) & 1) == color) {
                T s = sb[off[L] + t];
                if (!zero) {
                    const bool h0 = i > 0, h1 = i < Nx - 1, h2 = j > 0, h3 = j < v.Ny - 1, h4 = l > 0,
                               h5 = l < v.Nz - 1;
                    s = nb_sub<T>(h0, s, a[L][0], x[nb_idx(h0, t - 1, t)]);
                    s = nb_sub<T>(h1, s, a[L][1], x[nb_idx(h1, t + 1, t)]);
                    s = nb_sub<T>(h2, s, a[L][2], x[nb_idx(h2, t - Nx, t)]);
                    s = nb_sub<T>(h3, s, a[L][3], x[nb_idx(h3, t + Nx, t)]);
                    s = nb_sub<T>(h4, s, a[L][4], x[nb_idx(h4, t - NxNy, t)]);
                    s = nb_sub<T>(h5, s, a[L][5], x[nb_idx(h5, t + NxNy, t)]);
                }
                x[t] = s;
            } else if (zero) {
                x[t] = 0;
            }
        }
        __syncthreads();
    };
    for (int L = 0; L < nt - 1; L++) {
        const MGLevel<T> &v = lv[L], &c = lv[L + 1];
        for (int sw = 0; sw < nsw; sw++) {
            sweep(L, 0, sw == 0);
            sweep(L, 1, 0);
        }
        // weighted residual of the rows the last sweep did not solve (colour 0)
        if (t < v.N) {
            const int Nx = v.Nx, NxNy = v.Nx * v.Ny;
            const int l = t / NxNy, r = t - l * NxNy, j = r / Nx, i = r - j * Nx;
            double w = 0.0;
            if (!((i + j + l) & 1)) {
                const T *x = sx + off[L];
                const bool h0 = i > 0, h1 = i < Nx - 1, h2 = j > 0, h3 = j < v.Ny - 1, h4 = l > 0, h5 = l < v.Nz - 1;
                T s = sb[off[L] + t] - x[t];
                s = nb_sub<T>(h0, s, a[L][0], x[nb_idx(h0, t - 1, t)]);
                s = nb_sub<T>(h1, s, a[L][1], x[nb_idx(h1, t + 1, t)]);
                s = nb_sub<T>(h2, s, a[L][2], x[nb_idx(h2, t - Nx, t)]);
                s = nb_sub<T>(h3, s, a[L][3], x[nb_idx(h3, t + Nx, t)]);
                s = nb_sub<T>(h4, s, a[L][4], x[nb_idx(h4, t - NxNy, t)]);
                s = nb_sub<T>(h5, s, a[L][5], x[nb_idx(h5, t + NxNy, t)]);
                w = rs[L] * (double)s;
            }
            sr[t] = w;
        }
        __syncthreads();
        if (t < c.N) {
            const int cNxNy = c.Nx * c.Ny, LL = t / cNxNy, rr = t - LL * cNxNy, J = rr / c.Nx, I = rr - J * c.Nx;
            Agg g = agg_range(v.Nx, v.Ny, v.Nz, v.cx, v.cy, v.cz, I, J, LL);
            double acc = 0.0;
            for (int l = g.l0; l <= g.l1; l++)
                for (int j = g.j0; j <= g.j1; j++)
                    for (int i = g.i0; i <= g.i1; i++) acc += sr[i + v.Nx * (j + v.Ny * l)];
            sb[off[L + 1] + t] = (T)(acc / rs[L + 1]);
        }
        __syncthreads();
    }
    {
        const MGLevel<T> &z = lv[nt - 1];
        const int n = z.N;
        if (t < n) {
            double s = 0.0;
            for (int q = 0; q < n; q++) s += inv[t * n + q] * (double)sb[off[nt - 1] + q];
            sx[off[nt - 1] + t] = (T)s;
        }
        __syncthreads();
    }
    for (int L = nt - 2; L >= 0; L--) {
        const MGLevel<T> &v = lv[L], &c = lv[L + 1];
        if (t < v.N) {
            const int Nx = v.Nx, NxNy = v.Nx * v.Ny;
            const int l = t / NxNy, r = t - l * NxNy, j = r / Nx, i = r - j * Nx;
            const int I = v.cx ? i >> 1 : i, J = v.cy ? j >> 1 : j, LL = v.cz ? l >> 1 : l;
            sx[off[L] + t] += pw[L] * sx[off[L + 1] + I + c.Nx * (J + c.Ny * LL)];
        }
        __syncthreads();
        for (int sw = 0; sw < nsw; sw++) {
            sweep(L, 1, 0);
            sweep(L, 0, 0);
        }
    }
    if (t < lv[0].N) lv[0].x[t] = sx[t];
}
template <class T> void mg_tail(hipStream_t s, const MGLevel<T> *d_levels, const MGLevel<T> *h_levels, int nt,
                                const double *inv, int nsw) {
    int tot = 0;
    for (int L = 0; L < nt; L++) tot += h_levels[L].N;
    bool lds = g_tail_lds && nt <= TL_MAXT && nt >= 2 && tot <= TL_MAXN && h_levels[0].N <= TAIL_BLOCK &&
               h_levels[nt - 1].N <= 64;
    for (int L = 0; L + 1 < nt && lds; L++) lds = h_levels[L].A.c == nullptr;     // unit-diagonal levels
    if (lds) k_tail_lds<T><<<1, TAIL_BLOCK, 0, s>>>(d_levels, nt, inv, nsw);
    else k_tail<T><<<1, TAIL_BLOCK, 0, s>>>(d_levels, nt, inv, nsw);
}

// x += pw * x_parent (semiconductor_3d.c:2038-2048, MG_ALPHA = 1)
template <class T>
__global__ void __launch_bounds__(SD_BLOCK) k_prolong(MGLevel<T> f, MGLevel<T> c, T *x) {
    int k = blockIdx.x * SD_BLOCK + threadIdx.x;
    if (k >= f.N) return;
    int NxNy = f.Nx * f.Ny, l = k / NxNy, r = k - l * NxNy, j = r / f.Nx, i = r - j * f.Nx;
    int I = f.cx ? i >> 1 : i, J = f.cy ? j >> 1 : j, L = f.cz ? l >> 1 : l;
    const int kf = mg_idx(f, i, j, l);
    x[kf] += f.pw[kf] * c.x[mg_idx(c, I, J, L)];
}
template <class T> void mg_prolong(hipStream_t s, const MGLevel<T> &f, const MGLevel<T> &c, T *x) {
    k_prolong<T><<<sd_nblocks(f.N), SD_BLOCK, 0, s>>>(f, c, x);
}

// Coarsest level (<= 64 unknowns): dense LU with partial pivoting in LDS
// (semiconductor_3d.c:2050-2081), then the explicit inverse, one column per
// lane.  A zero pivot is replaced by 1e-300 as in the reference.
// NM: the coarsest level's size rounded up to 16, 32 or 64.  Every loop over
// the rows runs to NM (rows past n do nothing), so the solution column y of
// each lane is indexed by constants after unrolling and stays in registers.
template <class T, int NM>
__global__ void __launch_bounds__(64) k_coarse_factor(MGLevel<T> v, double *inv) {
    __shared__ double M[64][65];
    __shared__ int piv[64];
    const int n = v.N, t = threadIdx.x, NxNy = v.Nx * v.Ny;
    for (int r = 0; r < n; r++) if (t < n) M[r][t] = 0.0;
    __syncthreads();
    if (t < n) {
        int k = t, l = k / NxNy, rr = k - l * NxNy, j = rr / v.Nx, i = rr - j * v.Nx;
        M[k][k] = v.A.c ? (double)v.A.c[k] : 1.0;
        if (i > 0) M[k][k - 1] = v.A.xm[k];
        if (i < v.Nx - 1) M[k][k + 1] = v.A.xp[k];
        if (j > 0) M[k][k - v.Nx] = v.A.ym[k];
        if (j < v.Ny - 1) M[k][k + v.Nx] = v.A.yp[k];
        if (l > 0) M[k][k - NxNy] = v.A.zm[k];
        if (l < v.Nz - 1) M[k][k + NxNy] = v.A.zp[k];
    }
    __syncthreads();
    for (int c = 0; c < n; c++) {
        if (t == 0) {
            int p = c; double mx = fabs(M[c][c]);
            for (int r = c + 1; r < n; r++) if (fabs(M[r][c]) > mx) { mx = fabs(M[r][c]); p = r; }
            piv[c] = p;
        }
        __syncthreads();
        int p = piv[c];
        if (p != c && t < n) { double tt = M[c][t]; M[c][t] = M[p][t]; M[p][t] = tt; }
        __syncthreads();
        if (t == 0 && !(fabs(M[c][c]) > 1e-300)) M[c][c] = 1e-300;
        __syncthreads();
        double dg = M[c][c];
        if (t > c && t < n) {
            double fct = M[t][c] / dg;
            M[t][c] = fct;
            if (fct != 0.) for (int q = c + 1; q < n; q++) M[t][q] -= fct * M[c][q];
        }
        __syncthreads();
    }
    // column t of the inverse: solve LU y = P e_t
    if (t < n) {
        double y[NM];
#pragma unroll
        for (int r = 0; r < NM; r++) y[r] = (r == t) ? 1.0 : 0.0;
#pragma unroll
        for (int c = 0; c < NM; c++) {
            if (c < n) {
                const int p = piv[c];
                double yc = y[c], yp = 0.0;
#pragma unroll
                for (int r = c + 1; r < NM; r++) if (r == p) yp = y[r];
                if (p != c) {
#pragma unroll
                    for (int r = c + 1; r < NM; r++) if (r == p) y[r] = yc;
                    y[c] = yp;
                }
            }
        }
#pragma unroll
        for (int r = 0; r < NM; r++) {
            if (r < n) {
                double s = y[r];
#pragma unroll
                for (int q = 0; q < r; q++) s -= M[r][q] * y[q];
                y[r] = s;
            }
        }
#pragma unroll
        for (int r = NM - 1; r >= 0; r--) {
            if (r < n) {
                double s = y[r];
#pragma unroll
                for (int q = r + 1; q < NM; q++) if (q < n) s -= M[r][q] * y[q];
                y[r] = s / M[r][r];
            }
        }
#pragma unroll
        for (int r = 0; r < NM; r++) if (r < n) inv[r * n + t] = y[r];
    }
}
template <class T> void mg_coarse_factor(hipStream_t s, const MGLevel<T> &v, double *inv) {
    if (v.N <= 16) k_coarse_factor<T, 16><<<1, 64, 0, s>>>(v, inv);
    else if (v.N <= 32) k_coarse_factor<T, 32><<<1, 64, 0, s>>>(v, inv);
    else k_coarse_factor<T, 64><<<1, 64, 0, s>>>(v, inv);
}
template <class T>
__global__ void __launch_bounds__(64) k_coarse_solve(int n, const double *inv, const T *b, T *x) {
    int r = threadIdx.x;
    if (r >= n) return;
    double s = 0.0;
    for (int c = 0; c < n; c++) s += inv[r * n + c] * (double)b[c];
    x[r] = (T)s;
}
template <class T> void mg_coarse_solve(hipStream_t s, const MGLevel<T> &v, const double *inv, const T *b, T *x) {
    k_coarse_solve<T><<<1, 64, 0, s>>>(v.N, inv, b, x);
}

__global__ void k_copy_bw(const double *a, double *b, long n) {
    long k = (long)blockIdx.x * SD_BLOCK + threadIdx.x;
    if (k < n) b[k] = a[k];
}
void sd_copy_bw(hipStream_t s, const double *a, double *b, long n) {
    k_copy_bw<<<(int)((n + SD_BLOCK - 1) / SD_BLOCK), SD_BLOCK, 0, s>>>(a, b, n);
}
// natural-order FP64 node data into (and out of) the level's storage order
// and precision; one thread per x-row pair of positions keeps both the
// natural reads and the split writes contiguous
template <class T> __global__ void k_vec_in(const double *r, MGLevel<T> v, T *b) {
    int k = blockIdx.x * SD_BLOCK + threadIdx.x;
    if (k >= v.N) return;
    const int i = k % v.Nx;
    b[k - i + mg_xpos(v, i)] = (T)r[k];
}
template <class T> __global__ void k_vec_out(const T *x, MGLevel<T> v, double *z) {
    int k = blockIdx.x * SD_BLOCK + threadIdx.x;
    if (k >= v.N) return;
    const int i = k % v.Nx;
    z[k] = (double)x[k - i + mg_xpos(v, i)];
}
__global__ void k_o16(MGLevel<float> v) {
    int k = blockIdx.x * SD_BLOCK + threadIdx.x;
    if (k >= v.N) return;
    const float *src[6] = {v.A.xm, v.A.xp, v.A.ym, v.A.yp, v.A.zm, v.A.zp};
#pragma unroll
    for (int q = 0; q < 6; q++) v.o16[q][k] = (__bf16)src[q][k];
}
__global__ void k_lu16(MGLevel<float> v, int ax) {
    int k = blockIdx.x * SD_BLOCK + threadIdx.x;
    if (k >= v.N) return;
#pragma unroll
    for (int q = 0; q < 3; q++) v.lu16[ax][q][k] = (__bf16)v.lu[ax][q][k];
}
// one 16-byte record per node for the sweeps along axis ax: the line factors
// ai, iw, cp and the off-line couplings in the order k_line_scan1 subtracts
// them (x-lines: ym, yp, zm, zp; y-lines: xm, xp, zm, zp; z-lines: xm, xp, ym,
// yp), each rounded to BF16 exactly as k_o16 / k_lu16 round them
__device__ __forceinline__ unsigned bf16_bits(float f) {
    const __bf16 h = (__bf16)f;
    return (unsigned)__builtin_bit_cast(unsigned short, h);
}
__global__ void k_rec16(MGLevel<float> v, int ax, uint4 *rec) {
    int k = blockIdx.x * SD_BLOCK + threadIdx.x;
    if (k >= v.N) return;
    const float *o0 = ax == 0 ? v.A.ym : v.A.xm, *o1 = ax == 0 ? v.A.yp : v.A.xp;
    const float *o2 = ax == 2 ? v.A.ym : v.A.zm, *o3 = ax == 2 ? v.A.yp : v.A.zp;
    // node (i, j, l) at storage position k; a coupling to a neighbour outside
    // the mesh is stored as 0, so that the sweep subtracts it unconditionally
    const int row = k / v.Nx, p = k - row * v.Nx, j = row % v.Ny, l = row / v.Ny, hx = (v.Nx + 1) >> 1;
    const int i = v.px ? (p < hx ? 2 * p : 2 * (p - hx) + 1) : p;
    const bool hx0 = i > 0, hx1 = i < v.Nx - 1, hy0 = j > 0, hy1 = j < v.Ny - 1, hz0 = l > 0, hz1 = l < v.Nz - 1;
    const bool h0 = ax == 0 ? hy0 : hx0, h1 = ax == 0 ? hy1 : hx1;
    const bool h2 = ax == 2 ? hy0 : hz0, h3 = ax == 2 ? hy1 : hz1;
    uint4 r;
    r.x = bf16_bits(v.lu[ax][0][k]) | (bf16_bits(v.lu[ax][1][k]) << 16);
    r.y = bf16_bits(v.lu[ax][2][k]) | (bf16_bits(h0 ? o0[k] : 0.0f) << 16);
    r.z = bf16_bits(h1 ? o1[k] : 0.0f) | (bf16_bits(h2 ? o2[k] : 0.0f) << 16);
    r.w = bf16_bits(h3 ? o3[k] : 0.0f);
    rec[k] = r;
}
void mg_rec16(hipStream_t s, const MGLevel<float> &v, int ax, uint4 *rec) {
    k_rec16<<<sd_nblocks(v.N), SD_BLOCK, 0, s>>>(v, ax, rec);
}
void mg_o16(hipStream_t s, const MGLevel<float> &v) { k_o16<<<sd_nblocks(v.N), SD_BLOCK, 0, s>>>(v); }
void mg_lu16(hipStream_t s, const MGLevel<float> &v, int ax) {
    k_lu16<<<sd_nblocks(v.N), SD_BLOCK, 0, s>>>(v, ax);
}
template <class T> __global__ void k_sten_in(Sten7<double> A, MGLevel<T> v, Sten7<T> B) {
    int k = blockIdx.x * SD_BLOCK + threadIdx.x;
    if (k >= v.N) return;
    const int i = k % v.Nx, q = k - i + mg_xpos(v, i);
    B.c[q] = (T)A.c[k]; B.xm[q] = (T)A.xm[k]; B.xp[q] = (T)A.xp[k];
    B.ym[q] = (T)A.ym[k]; B.yp[q] = (T)A.yp[k]; B.zm[q] = (T)A.zm[k]; B.zp[q] = (T)A.zp[k];
}
template <class T> __global__ void k_rs_in(const double *w, MGLevel<T> v, double *rs) {
    int k = blockIdx.x * SD_BLOCK + threadIdx.x;
    if (k >= v.N) return;
    const int i = k % v.Nx;
    rs[k - i + mg_xpos(v, i)] = w[k];
}
template <class T> void mg_vec_in(hipStream_t s, const double *r, const MGLevel<T> &v, T *b) {
    k_vec_in<T><<<sd_nblocks(v.N), SD_BLOCK, 0, s>>>(r, v, b);
}
template <class T> void mg_vec_out(hipStream_t s, const T *x, const MGLevel<T> &v, double *z) {
    k_vec_out<T><<<sd_nblocks(v.N), SD_BLOCK, 0, s>>>(x, v, z);
}
template <class T> void mg_sten_in(hipStream_t s, Sten7<double> A, const MGLevel<T> &v, Sten7<T> B) {
    k_sten_in<T><<<sd_nblocks(v.N), SD_BLOCK, 0, s>>>(A, v, B);
}
template <class T> void mg_rs_in(hipStream_t s, const double *w, const MGLevel<T> &v, double *rs) {
    k_rs_in<T><<<sd_nblocks(v.N), SD_BLOCK, 0, s>>>(w, v, rs);
}

#define INST(T)                                                                                                   \
    template void mg_weights_level<T>(hipStream_t, const MGLevel<T> &, MGLevel<T> &, int, const DevView &, int, double); \
    template void mg_galerkin<T>(hipStream_t, const MGLevel<T> &, MGLevel<T> &);                                  \
    template void mg_rbgs<T>(hipStream_t, const MGLevel<T> &, const T *, T *, int, int, const MGLevel<T> *);                         \
    template void mg_rbgs_post2<T>(hipStream_t, const MGLevel<T> &, const T *, const T *, T *, const MGLevel<T> &);                 \
    template void mg_resid_restrict<T>(hipStream_t, const MGLevel<T> &, const MGLevel<T> &, const T *, const T *, int, double); \
    template int mg_line<T>(hipStream_t, const MGLevel<T> &, const T *, T *, int, int, int, const MGLevel<T> *); \
    template void mg_line_factor<T>(hipStream_t, const MGLevel<T> &, int); \
    template void mg_tail<T>(hipStream_t, const MGLevel<T> *, const MGLevel<T> *, int, const double *, int); \
    template void mg_prolong<T>(hipStream_t, const MGLevel<T> &, const MGLevel<T> &, T *);                       \
    template void mg_coarse_factor<T>(hipStream_t, const MGLevel<T> &, double *);                                \
    template void mg_vec_in<T>(hipStream_t, const double *, const MGLevel<T> &, T *);                             \
    template void mg_vec_out<T>(hipStream_t, const T *, const MGLevel<T> &, double *);                            \
    template void mg_sten_in<T>(hipStream_t, Sten7<double>, const MGLevel<T> &, Sten7<T>);                       \
    template void mg_rs_in<T>(hipStream_t, const double *, const MGLevel<T> &, double *);                        \
    template void mg_coarse_solve<T>(hipStream_t, const MGLevel<T> &, const double *, const T *, T *);           \
    template void mg_set_pw_one<T>(hipStream_t, MGLevel<T> &);
INST(double)
INST(float)
