// k_graph.hip - kernels of the graph (general box-method mesh) path (gfx950).
//
// The same finite-volume equations as the tensor kernels (k_physics.hip),
// with neighbours and geometry from the CSR graph: one lane per node walks
// its row of directed edges (length h, dual face area ar).  Reference
// counterparts (semiconductor_3d.c): poisson_core_g :4181-4213, face_flux_g
// :4215-4226, upd_mob_ref_g :4240-4292, mob_display_g :4294-4311,
// cont_assemble_g :4313-4346, float_fix_g majority test :4355-4356,
// terminal_flux_g :4412-4424, g_node_J/upd_J_g :4428-4463; graph linear
// algebra :2532-2705 (gmatvec_L, gfloor, g_galerkin, g_restrict, g_prolong,
// g_weights).  The reference's chunked D-ILU(0) smoother is replaced by
// multicolour Gauss-Seidel (one launch per colour, every lane of a colour
// independent), and its dense LU of the coarsest level by an explicit
// inverse computed in one workgroup.
#include "sd_graph.h"
#include "sd_reduce.h"

#define PH_NONE (-1e300)

// ---------------------------------------------------------------- Poisson
__global__ void __launch_bounds__(SD_BLOCK) k_gpoisson(DevView d, GraphDev g, const double *phi, int build,
                                                       double *F, GOp A, double *sc, RedBuf rb) {
    double acc[1] = {0.0};
    for (int k = blockIdx.x * SD_BLOCK + threadIdx.x; k < g.N; k += gridDim.x * SD_BLOCK) {
        const int e0 = g.rp[k], e1 = g.rp[k + 1];
        if (d.con_mask[k]) {
            F[k] = 0.0;
            if (build) {
                A.dg[k] = 1.0;
                for (int e = e0; e < e1; e++) A.va[e] = 0.0;
                sc[k] = 1.0;
            }
            continue;
        }
        const double ek = d.eps[k], pk = phi[k];
        double f = 0.0, dg = 0.0;
        for (int e = e0; e < e1; e++) {
            const int kb = g.cj[e];
            const double eb = d.eps[kb], cf = (2. * ek * eb / (ek + eb)) * g.ar[e] / g.h[e];
            f += cf * (phi[kb] - pk);
            dg += cf;
            if (build) A.va[e] = d.con_mask[kb] ? 0.0 : -cf;
        }
        const double vol = g.vol[k];
        if (d.kind[k] == K_SEMI) {
            const double VT = d.mats[d.mat[k]].VT;
            const double n = sd_nK(d, k, pk, d.phin[k], VT), p = sd_pK(d, k, pk, d.phip[k], VT);
            f += vol * SD_Q * (p - n + d.Nd[k] - d.Na[k] + d.Nf[k]);
            dg += vol * SD_Q * (n + p) / VT;
        } else f += vol * SD_Q * d.Nf[k];
        const int gm = d.gate_mask[k];
        if (gm) {                                       // oxide (Robin) gate
            const ConDev &c = d.cons[gm - 1];
            const double ra = d.robA[k];
            f += ra * (c.Cox * (c.VGp - pk) + c.qox);
            dg += c.Cox * ra;
        }
        F[k] = f;
        double s;
        if (build) { A.dg[k] = dg; sc[k] = dg; s = dg; }
        else s = sc[k];
        const double r = f / s;
        acc[0] += r * r;
    }
    red_partial<1, false>(acc, rb);
}
void gph_poisson(hipStream_t s, const DevView &d, const GraphDev &g, const double *phi, int build, double *F,
                 GOp A, double *sc, RedBuf rb, int slot) {
    const int nb = red_grid(g.N);
    k_gpoisson<<<nb, SD_BLOCK, 0, s>>>(d, g, phi, build, F, A, sc, rb);
    red_final(s, rb, 1, nb, slot, 0);
}

// ------------------------------------------------------- edge mobilities
// mobility anisotropy of a hexagonal crystal projected on the edge k -> kp:
// 1 + (a - 1) t_c^2, t_c the edge's direction cosine with the c-axis
__device__ __forceinline__ double ani_edge(const DevView &d, const GraphDev &g, int k, int kp, double a) {
    if (d.caxis < 0 || !(a > 0.) || a == 1.) return 1.;
    double L2 = 0., dc = 0.;
    for (int q = 0; q < 3; q++) {
        const double v = g.pos[3 * kp + q] - g.pos[3 * k + q];
        L2 += v * v;
        if (q == d.caxis) dc = v;
    }
    return L2 > 0. ? 1. + (a - 1.) * dc * dc / L2 : 1.;
}
// Least-squares gradient of psi over the semiconductor neighbours of a
// semiconductor node near an interface (qd < 0.5 m; 0 elsewhere): g = M^-1
// sum_e (psi_b - psi_k) dr_e / |dr_e|^2 (g_grad :2678-2690, upd_mob_ref_g
// :4250-4256)
__global__ void k_ggrad(DevView d, GraphDev g) {
    const int k = blockIdx.x * SD_BLOCK + threadIdx.x;
    if (k >= g.N) return;
    double r[3] = {0., 0., 0.};
    if (d.kind[k] == K_SEMI && d.qd[k] < 0.5) {
        for (int e = g.rp[k]; e < g.rp[k + 1]; e++) {
            const int b = g.cj[e];
            if (d.kind[b] != K_SEMI) continue;
            const double dx = g.pos[3 * b] - g.pos[3 * k], dy = g.pos[3 * b + 1] - g.pos[3 * k + 1],
                         dz = g.pos[3 * b + 2] - g.pos[3 * k + 2];
            const double w = (d.psi[b] - d.psi[k]) / (dx * dx + dy * dy + dz * dz + 1e-40);
            r[0] += w * dx; r[1] += w * dy; r[2] += w * dz;
        }
        const double *m = g.lsq + 6 * (size_t)k;
        g.grad[3 * k] = m[0] * r[0] + m[1] * r[1] + m[2] * r[2];
        g.grad[3 * k + 1] = m[1] * r[0] + m[3] * r[1] + m[4] * r[2];
        g.grad[3 * k + 2] = m[2] * r[0] + m[4] * r[1] + m[5] * r[2];
    } else {
        g.grad[3 * k] = g.grad[3 * k + 1] = g.grad[3 * k + 2] = 0.;
    }
}
// every undirected edge once, by its lower node
__global__ void k_gupd_mob(DevView d, GraphDev g, const double *qn, const double *qp, double *gmu) {
    const int k = blockIdx.x * SD_BLOCK + threadIdx.x;
    if (k >= g.N) return;
    const int sat = (d.models & 4) != 0;
    for (int e = g.rp[k]; e < g.rp[k + 1]; e++) {
        const int kp = g.cj[e];
        if (kp < k) continue;
        const int ue = g.ue[e];
        if (d.kind[k] != K_SEMI || d.kind[kp] != K_SEMI) { gmu[ue] = 0.0; gmu[g.nue + ue] = 0.0; continue; }
        const MatDev &m = d.mats[d.mat[k]];
        double mn = 0.5 * (d.mu_n[k] + d.mu_n[kp]), mp = 0.5 * (d.mu_p[k] + d.mu_p[kp]);
        mn *= ani_edge(d, g, k, kp, m.ani_n);
        mp *= ani_edge(d, g, k, kp, m.ani_p);
        if (d.qd && m.lb_on) {          // Lombardi: the field normal to the edge from the node gradients
            const double Dw = exp(-0.5 * (d.qd[k] + d.qd[kp]) / SD_LB_LCRIT);
            if (Dw > 1e-8) {
                const double h = g.h[e];
                double t[3];
                for (int q = 0; q < 3; q++) t[q] = (g.pos[3 * kp + q] - g.pos[3 * k + q]) / h;
                const double *ga = g.grad + 3 * (size_t)k, *gb = g.grad + 3 * (size_t)kp;
                const double pa = ga[0] * t[0] + ga[1] * t[1] + ga[2] * t[2];
                const double pb = gb[0] * t[0] + gb[1] * t[1] + gb[2] * t[2];
                const double Fa = ga[0] * ga[0] + ga[1] * ga[1] + ga[2] * ga[2] - pa * pa;
                const double Fb = gb[0] * gb[0] + gb[1] * gb[1] + gb[2] * gb[2] - pb * pb;
                const double F = sqrt(fmax(0.5 * (Fa + Fb), 0.));
                const double Nt = 0.5 * (d.Nd[k] + d.Na[k] + d.Nd[kp] + d.Na[kp]);
                if (m.lb_on & 1) mn = sd_lombardi(m, mn, F, Nt, Dw, 0);
                if (m.lb_on & 2) mp = sd_lombardi(m, mp, F, Nt, Dw, 1);
            }
        }
        if (sat) {                      // velocity saturation; no metal-contact column on this path (weight 1)
            const double h = g.h[e], dps = d.psi[kp] - d.psi[k];
            const double Fn = sd_heat_field(qn[kp] - qn[k], dps, h), Fp = sd_heat_field(qp[kp] - qp[k], dps, h);
            const double sn = sd_mob_hf(mn, Fn, m.vsat_n, m.beta_n), sp = sd_mob_hf(mp, Fp, m.vsat_p, m.beta_p);
            mn += 1.0 * (sn - mn);
            mp += 1.0 * (sp - mp);
        }
        gmu[ue] = mn;
        gmu[g.nue + ue] = mp;
    }
}
void gph_upd_mob(hipStream_t s, const DevView &d, const GraphDev &g, const double *qn, const double *qp,
                 double *gmu) {
    if (d.qd) k_ggrad<<<sd_nblocks(g.N), SD_BLOCK, 0, s>>>(d, g);
    k_gupd_mob<<<sd_nblocks(g.N), SD_BLOCK, 0, s>>>(d, g, qn, qp, gmu);
}

// ------------------------------------------- Scharfetter-Gummel continuity
// Rows scaled to unit diagonal as on the tensor path: unknown u = c/s, row
// scale 1/(diag s), dsc = diag s for the multigrid restriction.
__global__ void k_gcont(DevView d, GraphDev g, int hole, const double *gmu, GOp A, double *rhs, double *sc,
                        double *dsc) {
    const int k = blockIdx.x * SD_BLOCK + threadIdx.x;
    if (k >= g.N) return;
    const double *car = hole ? d.p : d.n;
    const int e0 = g.rp[k], e1 = g.rp[k + 1];
    if (d.kind[k] != K_SEMI || d.con_mask[k]) {
        A.dg[k] = 1.0;
        for (int e = e0; e < e1; e++) A.va[e] = 0.0;
        const bool semi = d.kind[k] == K_SEMI;
        rhs[k] = semi ? 1.0 : 0.0;
        sc[k] = semi ? car[k] : 0.0;
        dsc[k] = 0.0;
        return;
    }
    const MatDev &ma = d.mats[d.mat[k]];
    const double VT = ma.VT, sk = fmax(car[k], SD_NFLOOR);
    const double *mu = gmu + (size_t)hole * g.nue;
    double dg = 0.0, b = 0.0;
    for (int e = e0; e < e1; e++) {
        const int kb = g.cj[e];
        A.va[e] = 0.0;
        if (d.kind[kb] != K_SEMI) continue;
        const double dl = hole ? -(d.lnvK[kb] - d.lnvK[k]) : (d.lncK[kb] - d.lncK[k]);
        const double D = (d.psi[kb] - d.psi[k]) / VT + dl;
        const double gg = g.ar[e] * mu[g.ue[e]] * VT / g.h[e];
        double bp, bm;
        sd_bern2(D, bp, bm);
        const double self = hole ? gg * bp : gg * bm;
        const double nb = hole ? gg * bm : gg * bp;
        dg += self;
        if (d.con_mask[kb]) b += nb * car[kb];
        else A.va[e] = -nb * fmax(car[kb], SD_NFLOOR);
    }
    double R, dRn, dRp;
    sd_recomb(ma, d.niK[k], sd_tauf(d, k, ma), d.n[k], d.p[k], R, dRn, dRp);
    double dR = hole ? dRp : dRn;
    if (dR < 0.) dR = 0.;
    const double vol = g.vol[k];
    dg += vol * dR;
    b += vol * (dR * car[k] - R);
    if (d.tk) { dg += d.tk[k]; b += d.tg[k]; }
    if (d.gii) b += d.gii[k];
    const double w = 1.0 / (dg * sk);
    A.dg[k] = 1.0;
    for (int e = e0; e < e1; e++) A.va[e] *= w;
    rhs[k] = b * w;
    sc[k] = sk;
    dsc[k] = dg * sk;
}
void gph_cont(hipStream_t s, const DevView &d, const GraphDev &g, int hole, const double *gmu, GOp A, double *rhs,
              double *sc, double *dsc) {
    k_gcont<<<sd_nblocks(g.N), SD_BLOCK, 0, s>>>(d, g, hole, gmu, A, rhs, sc, dsc);
}

// --------------------------------------------------------- edge fluxes
// electron and hole particle fluxes a -> b along edge e (and their
// magnitudes), face_flux_g :4215-4226
__device__ __forceinline__ void gflux(const DevView &d, const GraphDev &g, const double *gmu, int ka, int kb, int e,
                                      double &Jn, double &Jp, double &Sn, double &Sp) {
    const double VT = d.mats[d.mat[ka]].VT, dps = (d.psi[kb] - d.psi[ka]) / VT;
    const double Dn = dps + (d.lncK[kb] - d.lncK[ka]), Dp = dps - (d.lnvK[kb] - d.lnvK[ka]);
    const int ue = g.ue[e];
    const double gn = SD_Q * g.ar[e] * gmu[ue] * VT / g.h[e], gp = SD_Q * g.ar[e] * gmu[g.nue + ue] * VT / g.h[e];
    double bnp, bnm, bpp, bpm;
    sd_bern2(Dn, bnp, bnm);
    sd_bern2(Dp, bpp, bpm);
    const double a1 = gn * d.n[kb] * bnp, a2 = gn * d.n[ka] * bnm;
    const double b1 = gp * d.p[ka] * bpp, b2 = gp * d.p[kb] * bpm;
    Jn = a1 - a2; Jp = b1 - b2; Sn = a1 + a2; Sp = b1 + b2;
}
__global__ void __launch_bounds__(SD_BLOCK) k_gcurrents(DevView d, GraphDev g, const double *gmu, const int *ck,
                                                        const int *ce, const int *off, RedBuf rb) {
    __shared__ double lds[3][SD_BLOCK / 64];
    const int c = blockIdx.x;
    double a0 = 0., a1 = 0., a2 = 0.;
    for (int q = off[c] + threadIdx.x; q < off[c + 1]; q += SD_BLOCK) {
        double Jn, Jp, Sn, Sp;
        const int k = ck[q], e = ce[q];
        gflux(d, g, gmu, k, g.cj[e], e, Jn, Jp, Sn, Sp);
        a0 += Jn + Jp; a1 += Sn + Sp; a2 += Jn;
    }
    a0 = sd_wave_sum(a0); a1 = sd_wave_sum(a1); a2 = sd_wave_sum(a2);
    const int w = threadIdx.x >> 6, ln = threadIdx.x & 63;
    if (ln == 0) { lds[0][w] = a0; lds[1][w] = a1; lds[2][w] = a2; }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int v = 0; v < 3; v++) {
            double s = lds[v][0];
            for (int q = 1; q < SD_BLOCK / 64; q++) s += lds[v][q];
            rb.scal[S_CUR + 3 * c + v] = s;
        }
    }
}
void gph_currents(hipStream_t s, const DevView &d, const GraphDev &g, const double *gmu, const int *ck,
                  const int *ce, const int *off, int ncon, RedBuf rb) {
    if (ncon > 0) k_gcurrents<<<ncon, SD_BLOCK, 0, s>>>(d, g, gmu, ck, ce, off, rb);
}

// majority test of float_fix_g (:4355-4356); the blocks are traced on the host
__global__ void k_gmaj_flags(DevView d, int hole, const double *u, const double *s, unsigned char *flag) {
    const int k = blockIdx.x * SD_BLOCK + threadIdx.x;
    if (k >= d.g.N) return;
    const double *oth = hole ? d.n : d.p;
    const double c = s[k] * u[k], D = 0.5 * fabs(d.Nd[k] - d.Na[k]);
    flag[k] = d.kind[k] == K_SEMI && !d.con_mask[k] && c >= oth[k] && c >= D && s[k] >= D && c > 10. * d.niK[k];
}
void gph_maj_flags(hipStream_t s, const DevView &d, int hole, const double *u, const double *sc,
                   unsigned char *flag) {
    k_gmaj_flags<<<sd_nblocks(d.g.N), SD_BLOCK, 0, s>>>(d, hole, u, sc, flag);
}

// ----------------------------------------------------- display quantities
// node current density by least squares over the node's edges: J = (sum_e
// A_e t t^T)^-1 sum_e A_e t j_e, t the unit edge vector, j_e the flux
// density (g_node_J :4428-4452); which: 0 total, 1 electrons, 2 holes
__device__ void gnode_J(const DevView &d, const GraphDev &g, const double *gmu, int k, int which, double J[3]) {
    double M[6] = {0, 0, 0, 0, 0, 0}, r[3] = {0, 0, 0};
    J[0] = J[1] = J[2] = 0.;
    if (d.kind[k] != K_SEMI) return;
    for (int e = g.rp[k]; e < g.rp[k + 1]; e++) {
        const int kb = g.cj[e];
        if (d.kind[kb] != K_SEMI) continue;
        const double h = g.h[e], ar = g.ar[e];
        double t[3], Jn, Jp, Sn, Sp;
        for (int q = 0; q < 3; q++) t[q] = (g.pos[3 * kb + q] - g.pos[3 * k + q]) / h;
        gflux(d, g, gmu, k, kb, e, Jn, Jp, Sn, Sp);
        const double j = (which == 1 ? Jn : (which == 2 ? Jp : Jn + Jp)) / (ar > 0. ? ar : 1e-300);
        const double w = ar;
        M[0] += w * t[0] * t[0]; M[1] += w * t[0] * t[1]; M[2] += w * t[0] * t[2];
        M[3] += w * t[1] * t[1]; M[4] += w * t[1] * t[2]; M[5] += w * t[2] * t[2];
        r[0] += w * t[0] * j; r[1] += w * t[1] * j; r[2] += w * t[2] * j;
    }
    const double tr = M[0] + M[3] + M[5], eps = 1e-6 * tr + 1e-300;
    M[0] += eps; M[3] += eps; M[5] += eps;
    const double a = M[0], b = M[1], c = M[2], dd = M[3], e = M[4], f = M[5];
    const double Ai = dd * f - e * e, Bi = -(b * f - c * e), Ci = b * e - c * dd;
    const double Di = a * f - c * c, Ei = -(a * e - b * c), Fi = a * dd - b * b;
    const double det = a * Ai + b * Bi + c * Ci;
    if (!(fabs(det) > 0.)) return;
    J[0] = (Ai * r[0] + Bi * r[1] + Ci * r[2]) / det;
    J[1] = (Bi * r[0] + Di * r[1] + Ei * r[2]) / det;
    J[2] = (Ci * r[0] + Ei * r[1] + Fi * r[2]) / det;
}
__global__ void k_gdisplay(DevView d, GraphDev g, const double *gmu, double *Jx, double *Jy, double *Jz, double *Rt,
                           double *mun, double *mup) {
    const int k = blockIdx.x * SD_BLOCK + threadIdx.x;
    if (k >= g.N) return;
    if (d.kind[k] != K_SEMI) {
        Jx[k] = Jy[k] = Jz[k] = 0.; Rt[k] = 0.;
        mun[k] = d.mu_n[k]; mup[k] = d.mu_p[k];
        return;
    }
    double J[3];
    gnode_J(d, g, gmu, k, 0, J);
    Jx[k] = J[0]; Jy[k] = J[1]; Jz[k] = J[2];
    const MatDev &m = d.mats[d.mat[k]];
    double R, a, b;
    sd_recomb(m, d.niK[k], sd_tauf(d, k, m), d.n[k], d.p[k], R, a, b);
    Rt[k] = R;
    double sn = 0., sp = 0.;
    int c = 0;
    for (int e = g.rp[k]; e < g.rp[k + 1]; e++) {
        const int kb = g.cj[e];
        if (d.kind[kb] != K_SEMI) continue;
        const MatDev &ml = d.mats[d.mat[k < kb ? k : kb]];       // shown across the c-axis
        sn += gmu[g.ue[e]] / ani_edge(d, g, k, kb, ml.ani_n);
        sp += gmu[g.nue + g.ue[e]] / ani_edge(d, g, k, kb, ml.ani_p);
        c++;
    }
    if (c) { mun[k] = sn / c; mup[k] = sp / c; }
    else { mun[k] = d.mu_n[k]; mup[k] = d.mu_p[k]; }
}
void gph_display(hipStream_t s, const DevView &d, const GraphDev &g, const double *gmu, double *Jx, double *Jy,
                 double *Jz, double *R, double *mun, double *mup) {
    k_gdisplay<<<sd_nblocks(g.N), SD_BLOCK, 0, s>>>(d, g, gmu, Jx, Jy, Jz, R, mun, mup);
}
__global__ void k_gnode_J(DevView d, GraphDev g, const double *gmu, int ax, int carrier, double *out) {
    const int k = blockIdx.x * SD_BLOCK + threadIdx.x;
    if (k >= g.N) return;
    double J[3];
    gnode_J(d, g, gmu, k, carrier, J);
    out[k] = J[ax];
}
void gph_node_J(hipStream_t s, const DevView &d, const GraphDev &g, const double *gmu, int ax, int carrier,
                double *out) {
    k_gnode_J<<<sd_nblocks(g.N), SD_BLOCK, 0, s>>>(d, g, gmu, ax, carrier, out);
}

// ---------------------------------------------------------- linear algebra
// Rows walked by LPR lanes together: lane j of the row's group takes entries
// j, j + LPR, ..., so that the group's loads of a row's (contiguous) columns
// and values are coalesced; a shuffle sum closes the row.  LPR is chosen per
// level near its mean row length (gla_lanes).
template <int LPR> __device__ __forceinline__ double grow_part(const GOp &A, const double *x, int k, int j) {
    double s = 0.;
    for (int e = A.rp[k] + j; e < A.rp[k + 1]; e += LPR) s += A.va[e] * x[A.cj[e]];
#pragma unroll
    for (int o = LPR / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, LPR);
    return s;
}
int gla_lanes(int n, int nnz) {
    const double m = n > 0 ? (double)nnz / n : 1.;
    return m <= 4.5 ? 4 : (m <= 9. ? 8 : (m <= 18. ? 16 : 32));
}
__device__ __forceinline__ double grow(const GOp &A, const double *x, int k) {
    double s = A.dg ? A.dg[k] * x[k] : x[k];
    for (int e = A.rp[k]; e < A.rp[k + 1]; e++) s += A.va[e] * x[A.cj[e]];
    return s;
}
__global__ void k_gspmv(GOp A, const double *x, const double *b, double *y) {
    const int k = blockIdx.x * SD_BLOCK + threadIdx.x;
    if (k >= A.N) return;
    const double s = grow(A, x, k);
    y[k] = b ? b[k] - s : s;
}
void gla_spmv(hipStream_t s, GOp A, const double *x, const double *b, double *y) {
    k_gspmv<<<sd_nblocks(A.N), SD_BLOCK, 0, s>>>(A, x, b, y);
}
// y = A x fused with y.(u0 or y) and y.u1, as k_spmv_dots
__global__ void __launch_bounds__(SD_BLOCK) k_gspmv_dots(GOp A, const double *x, double *y, const double *u0,
                                                         const double *u1, int nv, RedBuf rb) {
    double acc[2] = {0.0, 0.0};
    for (int k = blockIdx.x * SD_BLOCK + threadIdx.x; k < A.N; k += gridDim.x * SD_BLOCK) {
        const double s = grow(A, x, k);
        y[k] = s;
        acc[0] += s * (u0 ? u0[k] : s);
        if (nv > 1) acc[1] += s * u1[k];
    }
    red_partial<2, false>(acc, rb);
}
template <int LPR>
__global__ void __launch_bounds__(SD_BLOCK) k_gspmv_dots_v(GOp A, const double *x, double *y, const double *u0,
                                                           const double *u1, int nv, RedBuf rb) {
    double acc[2] = {0.0, 0.0};
    const int j = threadIdx.x % LPR, rows = SD_BLOCK / LPR;
    for (int k0 = blockIdx.x * rows; k0 < A.N; k0 += gridDim.x * rows) {
        const int k = k0 + (int)threadIdx.x / LPR;
        if (k >= A.N) continue;                         // whole row groups leave together
        double s = grow_part<LPR>(A, x, k, j);
        if (j == 0) {
            s += A.dg ? A.dg[k] * x[k] : x[k];
            y[k] = s;
            acc[0] += s * (u0 ? u0[k] : s);
            if (nv > 1) acc[1] += s * u1[k];
        }
    }
    red_partial<2, false>(acc, rb);
}
void gla_spmv_dots(hipStream_t s, GOp A, const double *x, double *y, const double *u0, const double *u1, int nv,
                   RedBuf rb, int slot, int post, int lanes) {
    if (lanes <= 1) {
        const int nb = red_grid(A.N);
        k_gspmv_dots<<<nb, SD_BLOCK, 0, s>>>(A, x, y, u0, u1, nv, rb);
        red_final(s, rb, nv, nb, slot, 0, post);
        return;
    }
    const int nb = red_grid((int)(((long)A.N * lanes + SD_BLOCK - 1) / SD_BLOCK * SD_BLOCK));
    switch (lanes) {
    case 4: k_gspmv_dots_v<4><<<nb, SD_BLOCK, 0, s>>>(A, x, y, u0, u1, nv, rb); break;
    case 8: k_gspmv_dots_v<8><<<nb, SD_BLOCK, 0, s>>>(A, x, y, u0, u1, nv, rb); break;
    case 16: k_gspmv_dots_v<16><<<nb, SD_BLOCK, 0, s>>>(A, x, y, u0, u1, nv, rb); break;
    default: k_gspmv_dots_v<32><<<nb, SD_BLOCK, 0, s>>>(A, x, y, u0, u1, nv, rb); break;
    }
    red_final(s, rb, nv, nb, slot, 0, post);
}
// rounding floor of a continuity solve (gfloor :2556-2566)
__global__ void __launch_bounds__(SD_BLOCK) k_gfloor(GOp A, const double *x, const double *b, RedBuf rb) {
    double acc[1] = {0.0};
    for (int k = blockIdx.x * SD_BLOCK + threadIdx.x; k < A.N; k += gridDim.x * SD_BLOCK) {
        double s = fabs((A.dg ? A.dg[k] : 1.0) * x[k]) + fabs(b[k]);
        for (int e = A.rp[k]; e < A.rp[k + 1]; e++) s += fabs(A.va[e] * x[A.cj[e]]);
        acc[0] += s * s;
    }
    red_partial<1, false>(acc, rb);
}
void gla_floor(hipStream_t s, GOp A, const double *x, const double *b, RedBuf rb, int slot) {
    const int nb = red_grid(A.N);
    k_gfloor<<<nb, SD_BLOCK, 0, s>>>(A, x, b, rb);
    red_final(s, rb, 1, nb, slot, 0);
}

// Slotboom prolongation weights of the members of every aggregate: the
// members' reference potential relative to the aggregate's largest, 0 for
// nodes whose carrier density is fixed (g_weights :2617-2639, ph0 :2085-2091)
__global__ void k_gweights(DevView d, int L, int nc, GLev f, GLev c, int qsign, double VT) {
    const int I = blockIdx.x * SD_BLOCK + threadIdx.x;
    if (I >= nc) return;
    auto ph = [&](int i) -> double {
        if (L > 0) return f.ph[i];
        if (d.kind[i] != K_SEMI || d.con_mask[i]) return PH_NONE;
        return qsign > 0 ? d.phin[i] / VT : -d.phip[i] / VT;
    };
    double mx = PH_NONE;
    for (int q = f.mrp[I]; q < f.mrp[I + 1]; q++) { const double p = ph(f.mem[q]); if (p > mx) mx = p; }
    c.ph[I] = mx;
    for (int q = f.mrp[I]; q < f.mrp[I + 1]; q++) {
        const int i = f.mem[q];
        const double p = ph(i);
        const double e = (p <= 0.5 * PH_NONE) ? -800. : p - mx;
        f.pw[i] = e < -700. ? 0. : exp(e);
    }
}
void gla_weights(hipStream_t s, const DevView &d, int L, int nc, const GLev &f, const GLev &c, int qsign, double VT) {
    k_gweights<<<sd_nblocks(nc), SD_BLOCK, 0, s>>>(d, L, nc, f, c, qsign, VT);
}
__global__ void k_gfill(double *a, double v, int n) {
    const int k = blockIdx.x * SD_BLOCK + threadIdx.x;
    if (k < n) a[k] = v;
}
void gla_pw_one(hipStream_t s, int n, double *pw) { k_gfill<<<sd_nblocks(n), SD_BLOCK, 0, s>>>(pw, 1.0, n); }

// Petrov-Galerkin product of one coarse row: sum over the aggregate's members
// i of w_i (row i of A) pw, with the entries of edges inside the aggregate
// added to the diagonal; a row of zeros keeps an identity row (g_galerkin
// :2569-2589).  One lane per coarse row, members and edges in stored order.
__global__ void k_ggalerkin(int nc, GLev f, GLev c, const double *w) {
    const int I = blockIdx.x * SD_BLOCK + threadIdx.x;
    if (I >= nc) return;
    const int s0 = c.A.rp[I], s1 = c.A.rp[I + 1];
    for (int s = s0; s < s1; s++) c.A.va[s] = 0.;
    double cc = 0.;
    for (int q = f.mrp[I]; q < f.mrp[I + 1]; q++) {
        const int i = f.mem[q];
        const double wi = w ? w[i] : 1.;
        cc += wi * (f.A.dg ? f.A.dg[i] : 1.0) * f.pw[i];
        for (int e = f.A.rp[i]; e < f.A.rp[i + 1]; e++) {
            const double a = wi * f.A.va[e] * f.pw[f.A.cj[e]];
            const int m = f.emap[e];
            if (m < 0) cc += a;
            else c.A.va[m] += a;
        }
    }
    bool zero = cc == 0.;
    for (int s = s0; s < s1 && zero; s++) if (c.A.va[s] != 0.) zero = false;
    c.A.dg[I] = zero ? 1. : cc;
}
void gla_galerkin(hipStream_t s, int nc, const GLev &f, const GLev &c, const double *w) {
    k_ggalerkin<<<sd_nblocks(nc), SD_BLOCK, 0, s>>>(nc, f, c, w);
}

// one colour of a Gauss-Seidel sweep: the nodes perm[n0 .. n1) share no edge
__global__ void k_ggs(GOp A, const double *b, double *x, const int *perm, int n0, int n1) {
    const int t = n0 + blockIdx.x * SD_BLOCK + threadIdx.x;
    if (t >= n1) return;
    const int k = perm[t];
    double s = b[k];
    for (int e = A.rp[k]; e < A.rp[k + 1]; e++) s -= A.va[e] * x[A.cj[e]];
    x[k] = A.dg ? s / A.dg[k] : s;
}
template <int LPR> __global__ void k_ggs_v(GOp A, const double *b, double *x, const int *perm, int n0, int n1) {
    const int t = n0 + (blockIdx.x * SD_BLOCK + (int)threadIdx.x) / LPR, j = threadIdx.x % LPR;
    if (t >= n1) return;                                // whole row groups leave together
    const int k = perm[t];
    const double s = grow_part<LPR>(A, x, k, j);
    if (j == 0) x[k] = A.dg ? (b[k] - s) / A.dg[k] : b[k] - s;
}
void gla_gs(hipStream_t s, GOp A, const double *b, double *x, const int *perm, int n0, int n1, int lanes) {
    if (n1 <= n0) return;
    const int nb = sd_nblocks((long)(n1 - n0) * (lanes > 1 ? lanes : 1));
    switch (lanes) {
    case 4: k_ggs_v<4><<<nb, SD_BLOCK, 0, s>>>(A, b, x, perm, n0, n1); break;
    case 8: k_ggs_v<8><<<nb, SD_BLOCK, 0, s>>>(A, b, x, perm, n0, n1); break;
    case 16: k_ggs_v<16><<<nb, SD_BLOCK, 0, s>>>(A, b, x, perm, n0, n1); break;
    case 32: k_ggs_v<32><<<nb, SD_BLOCK, 0, s>>>(A, b, x, perm, n0, n1); break;
    default: k_ggs<<<nb, SD_BLOCK, 0, s>>>(A, b, x, perm, n0, n1);
    }
}
// residual b - A x of the members of every aggregate, summed with the
// restriction weights into the coarse right-hand side (g_restrict :2590-2598
// after gmatvec_L); the residual itself is not stored
__global__ void k_gresid_restrict(int nc, GLev f, const double *b, const double *x, const double *w, double *bc) {
    const int I = blockIdx.x * SD_BLOCK + threadIdx.x;
    if (I >= nc) return;
    double s = 0.;
    for (int q = f.mrp[I]; q < f.mrp[I + 1]; q++) {
        const int i = f.mem[q];
        const double r = b[i] - grow(f.A, x, i);
        s += w ? w[i] * r : r;
    }
    bc[I] = s;
}
// One l1-Jacobi sweep (groups of LPR lanes per row): xo = xi + (b - A xi) /
// (|dg| + sum |off-diagonal|); xi == nullptr: from xi = 0.  The l1 diagonal
// makes the sweep convergent for the symmetric positive definite Poisson
// operator, and equal pre- and post-smoothing keeps the V-cycle symmetric.
template <int LPR> __global__ void k_gjacobi(GOp A, const double *b, const double *xi, double *xo) {
    const int k = (blockIdx.x * SD_BLOCK + (int)threadIdx.x) / LPR, j = threadIdx.x % LPR;
    if (k >= A.N) return;
    double s = 0., a = 0.;
    for (int e = A.rp[k] + j; e < A.rp[k + 1]; e += LPR) {
        const double v = A.va[e];
        if (xi) s += v * xi[A.cj[e]];
        a += fabs(v);
    }
#pragma unroll
    for (int o = LPR / 2; o > 0; o >>= 1) { s += __shfl_xor(s, o, LPR); a += __shfl_xor(a, o, LPR); }
    if (j == 0) {
        const double d = A.dg ? A.dg[k] : 1.0, x0 = xi ? xi[k] : 0.;
        xo[k] = x0 + (b[k] - d * x0 - s) / (fabs(d) + a);
    }
}
void gla_jacobi(hipStream_t s, GOp A, const double *b, const double *xi, double *xo, int lanes) {
    const int nb = sd_nblocks((long)A.N * lanes);
    switch (lanes) {
    case 4: k_gjacobi<4><<<nb, SD_BLOCK, 0, s>>>(A, b, xi, xo); break;
    case 8: k_gjacobi<8><<<nb, SD_BLOCK, 0, s>>>(A, b, xi, xo); break;
    case 16: k_gjacobi<16><<<nb, SD_BLOCK, 0, s>>>(A, b, xi, xo); break;
    default: k_gjacobi<32><<<nb, SD_BLOCK, 0, s>>>(A, b, xi, xo); break;
    }
}

// residual of every node (one lane per row), weighted, stored in member order
// of the aggregates: r[q] for q = position of node mem[q]
template <int LPR>
__global__ void k_gresid_mem(int n, GLev f, const double *b, const double *x, const double *w, double *r) {
    const int q = (blockIdx.x * SD_BLOCK + (int)threadIdx.x) / LPR, j = threadIdx.x % LPR;
    if (q >= n) return;
    const int i = f.mem[q];
    const double s = grow_part<LPR>(f.A, x, i, j);
    if (j == 0) {
        const double v = b[i] - (f.A.dg ? f.A.dg[i] * x[i] : x[i]) - s;
        r[q] = w ? w[i] * v : v;
    }
}
// the coarse right-hand side: each aggregate's members are contiguous in r
__global__ void k_gsum_mem(int nc, GLev f, const double *r, double *bc) {
    const int I = blockIdx.x * SD_BLOCK + threadIdx.x;
    if (I >= nc) return;
    double s = 0.;
    for (int q = f.mrp[I]; q < f.mrp[I + 1]; q++) s += r[q];
    bc[I] = s;
}
void gla_resid_restrict(hipStream_t s, int nc, const GLev &f, const double *b, const double *x, const double *w,
                        double *bc, double *r, int lanes) {
    if (!r) {
        k_gresid_restrict<<<sd_nblocks(nc), SD_BLOCK, 0, s>>>(nc, f, b, x, w, bc);
        return;
    }
    const int n = f.A.N, nb = sd_nblocks((long)n * lanes);
    switch (lanes) {
    case 4: k_gresid_mem<4><<<nb, SD_BLOCK, 0, s>>>(n, f, b, x, w, r); break;
    case 8: k_gresid_mem<8><<<nb, SD_BLOCK, 0, s>>>(n, f, b, x, w, r); break;
    case 16: k_gresid_mem<16><<<nb, SD_BLOCK, 0, s>>>(n, f, b, x, w, r); break;
    default: k_gresid_mem<32><<<nb, SD_BLOCK, 0, s>>>(n, f, b, x, w, r); break;
    }
    k_gsum_mem<<<sd_nblocks(nc), SD_BLOCK, 0, s>>>(nc, f, r, bc);
}
__global__ void k_gprolong(int n, GLev f, const double *xc, double *x, double alpha) {
    const int i = blockIdx.x * SD_BLOCK + threadIdx.x;
    if (i < n) x[i] += alpha * f.pw[i] * xc[f.agg[i]];
}
void gla_prolong(hipStream_t s, int n, const GLev &f, const double *xc, double *x, double alpha) {
    k_gprolong<<<sd_nblocks(n), SD_BLOCK, 0, s>>>(n, f, xc, x, alpha);
}

// Explicit inverse of the coarsest operator (n <= a few hundred unknowns):
// Gauss-Jordan elimination with partial pivoting on [A | I] (n x 2n, row
// major, in `work`) by one workgroup; barriers order the steps.  A pivot
// below 1e-300 is replaced by it (the reference's dense LU does the same).
#define GDENSE_T 1024
#define GDENSE_LDS_N 80            // [A | I] in LDS up to 80 unknowns (100 KB of the 160 KB of a CU)
__global__ void __launch_bounds__(GDENSE_T) k_gdense_inverse(GOp A, double *Mg, double *inv) {
    const int n = A.N, w = 2 * n, t = threadIdx.x;
    __shared__ int piv;
    __shared__ double Ml[2 * GDENSE_LDS_N * GDENSE_LDS_N];
    double *M = n <= GDENSE_LDS_N ? Ml : Mg;
    for (int q = t; q < n * w; q += GDENSE_T) M[q] = 0.;
    __syncthreads();
    for (int r = t; r < n; r += GDENSE_T) {
        M[(size_t)r * w + r] = A.dg ? A.dg[r] : 1.0;
        for (int e = A.rp[r]; e < A.rp[r + 1]; e++) M[(size_t)r * w + A.cj[e]] += A.va[e];
        M[(size_t)r * w + n + r] = 1.;
    }
    __syncthreads();
    for (int c = 0; c < n; c++) {
        if (t == 0) {
            int p = c;
            double mx = fabs(M[(size_t)c * w + c]);
            for (int r = c + 1; r < n; r++) {
                const double v = fabs(M[(size_t)r * w + c]);
                if (v > mx) { mx = v; p = r; }
            }
            piv = p;
        }
        __syncthreads();
        const int p = piv;
        if (p != c)
            for (int q = t; q < w; q += GDENSE_T) {
                const double a = M[(size_t)c * w + q];
                M[(size_t)c * w + q] = M[(size_t)p * w + q];
                M[(size_t)p * w + q] = a;
            }
        __syncthreads();
        double dgv = M[(size_t)c * w + c];
        if (!(fabs(dgv) > 1e-300)) dgv = 1e-300;
        __syncthreads();
        for (int q = t; q < w; q += GDENSE_T) M[(size_t)c * w + q] = q == c ? 1.0 : M[(size_t)c * w + q] / dgv;
        __syncthreads();
        // eliminate column c from every other row
        for (int idx = t; idx < n * w; idx += GDENSE_T) {
            const int r = idx / w, q = idx - r * w;
            if (r == c || q == c) continue;
            const double f = M[(size_t)r * w + c];
            if (f != 0.) M[(size_t)r * w + q] -= f * M[(size_t)c * w + q];
        }
        __syncthreads();
        for (int r = t; r < n; r += GDENSE_T) if (r != c) M[(size_t)r * w + c] = 0.;
        __syncthreads();
    }
    for (int q = t; q < n * n; q += GDENSE_T) {
        const int r = q / n, c = q - r * n;
        inv[q] = M[(size_t)r * w + n + c];
    }
}
void gla_dense_inverse(hipStream_t s, GOp A, double *work, double *inv) {
    k_gdense_inverse<<<1, GDENSE_T, 0, s>>>(A, work, inv);
}
__global__ void k_gdense_apply(int n, const double *inv, const double *b, double *x) {
    const int r = blockIdx.x * SD_BLOCK + threadIdx.x;
    if (r >= n) return;
    double s = 0.;
    for (int c = 0; c < n; c++) s += inv[(size_t)r * n + c] * b[c];
    x[r] = s;
}
void gla_dense_apply(hipStream_t s, int n, const double *inv, const double *b, double *x) {
    k_gdense_apply<<<sd_nblocks(n), SD_BLOCK, 0, s>>>(n, inv, b, x);
}

// The small levels of a V-cycle in one workgroup: per level, Gauss-Seidel
// colour by colour from x = 0, the residual restricted to the next level;
// the coarsest solved; then back up, prolongation and the backward sweep.
// Barriers separate every dependent step; the levels' data stay in L2.
#define GTAIL_T 1024
#define GTAIL_LPR 16
__global__ void __launch_bounds__(GTAIL_T) k_gtail(GTail T, int L0, const double *b0, double *x0) {
    const int t = threadIdx.x;
    auto B = [&](int L) -> const double * { return L == L0 ? b0 : T.lev[L].b; };
    auto X = [&](int L) -> double * { return L == L0 ? x0 : T.lev[L].x; };
    // rows by groups of GTAIL_LPR lanes (the coarse rows hold tens of entries)
    const int j = t % GTAIL_LPR, grp = t / GTAIL_LPR;
    constexpr int NG = GTAIL_T / GTAIL_LPR;
    auto sweep = [&](int L, bool fwd) {
        const GLev &v = T.lev[L];
        const double *b = B(L);
        double *x = X(L);
        for (int q = 0; q < v.ncol; q++) {
            const int c = fwd ? q : v.ncol - 1 - q;
            for (int i0 = v.coff[c]; i0 < v.coff[c + 1]; i0 += NG) {
                const int i = i0 + grp;
                if (i < v.coff[c + 1]) {
                    const int k = v.perm[i];
                    const double sum = grow_part<GTAIL_LPR>(v.A, x, k, j);
                    if (j == 0) x[k] = v.A.dg ? (b[k] - sum) / v.A.dg[k] : b[k] - sum;
                }
            }
            __syncthreads();
        }
    };
    const int last = T.nl - 1;
    for (int L = L0; L < last; L++) {
        const GLev &v = T.lev[L];
        double *x = X(L);
        for (int k = t; k < v.A.N; k += GTAIL_T) x[k] = 0.;
        __syncthreads();
        sweep(L, true);
        const int nc = T.lev[L + 1].A.N;
        const double *b = B(L);
        // residuals of the members (groups of lanes per row) into the next
        // level's x as scratch in member order, then the per-aggregate sums
        double *rq = T.lev[L + 1].x;
        for (int q0 = 0; q0 < v.A.N; q0 += NG) {
            const int q = q0 + grp;
            if (q < v.A.N) {
                const int i = v.mem[q];
                const double sum = grow_part<GTAIL_LPR>(v.A, x, i, j);
                if (j == 0) rq[q] = b[i] - (v.A.dg ? v.A.dg[i] * x[i] : x[i]) - sum;
            }
        }
        __syncthreads();
        for (int I = t; I < nc; I += GTAIL_T) {
            double sum = 0.;
            for (int q = v.mrp[I]; q < v.mrp[I + 1]; q++) sum += rq[q];
            T.lev[L + 1].b[I] = sum;
        }
        __syncthreads();
    }
    {
        const GLev &v = T.lev[last];
        const double *b = B(last);
        double *x = X(last);
        const int n = v.A.N;
        if (T.dense) {
            for (int r = t; r < n; r += GTAIL_T) {
                double sum = 0.;
                for (int c = 0; c < n; c++) sum += T.inv[(size_t)r * n + c] * b[c];
                x[r] = sum;
            }
            __syncthreads();
        } else {
            for (int k = t; k < n; k += GTAIL_T) x[k] = 0.;
            __syncthreads();
            for (int it = 0; it < 4; it++) { sweep(last, true); sweep(last, false); }
        }
    }
    for (int L = last - 1; L >= L0; L--) {
        const GLev &v = T.lev[L];
        double *x = X(L);
        const double *xc = X(L + 1);
        for (int i = t; i < v.A.N; i += GTAIL_T) x[i] += v.pw[i] * xc[v.agg[i]];
        __syncthreads();
        sweep(L, false);
    }
}
void gla_tail(hipStream_t s, const GTail &T, int L0, const double *b0, double *x0) {
    k_gtail<<<1, GTAIL_T, 0, s>>>(T, L0, b0, x0);
}
