// k_physics.hip - box-method drift-diffusion physics on the device (gfx950).
//
// One lane per mesh node, SoA FP64 node data, 7-point neighbourhoods read
// straight from HBM/L2 (x neighbours share cache lines, y/z neighbours are
// L2 hits for the planes in flight).  The physics follows the reference core
// (semiconductor_3d.c): poisson_eval :2732-2789, rebuild_np :2826-2834,
// face_flux :2843-2854, upd_mob_ref :2962-3006, cont_assemble :3067-3106,
// cont_update :3112-3146, terminal_currents :3598-3623, upd_J :3627-3658,
// apply_bc/neutral_psi/init_eq :1782-1845, aa_step :3709-3757, gummel's
// neutral-region predictor :3853-3859, float_fix majority test :3168-3169.
#include "sd_physics.h"
#include "sd_reduce.h"

#define IDX3(i, j, l) ((i) + g.Nx * ((j) + g.Ny * (l)))

__device__ __forceinline__ double neutral_psi(const DevView &d, int k, double V, double VT) {
    double D = d.Nd[k] - d.Na[k];
    return V + VT * (asinh(D / (2.0 * d.niK[k])) - 0.5 * (d.lncK[k] - d.lnvK[k]));
}

void ph_set_tile_mode(int m) {
    if (hipMemcpyToSymbol(HIP_SYMBOL(c_tile_mode), &m, sizeof(int)) != hipSuccess)
        fprintf(stderr, "semisim-hip: setting the tile order failed\n");
}

// ---------------------------------------------------------------- Poisson
__global__ void __launch_bounds__(SD_BLOCK) k_poisson(DevView d, const double *phi, int build, double *F,
                                                      Sten7<double> A, double *sc, RedBuf rb, int slot) {
    const Geo g = d.g;
    double acc[1] = {0.0};
    int k0, st, k1;
    sd_node_loop(g.N, k0, st, k1);
    NodeWalk nw(k0, st, g.Nx, g.Ny);
    for (int k = k0; k < k1; k += st, nw.next()) {
        if (d.con_mask[k]) {
            F[k] = 0.0;
            if (build) {
                A.c[k] = 1.0; A.xm[k] = A.xp[k] = A.ym[k] = A.yp[k] = A.zm[k] = A.zp[k] = 0.0;
                sc[k] = 1.0;
            }
            continue;
        }
        const int i = nw.i, j = nw.j, l = nw.l;
        // every neighbour's operands loaded before the first is used
        Nbr6 nb;
        sd_nbr6(g, i, j, l, k, nb);
        double ebv[6], phv[6];
        int cmv[6];
#pragma unroll
        for (int dir = 0; dir < 6; dir++) {
            const int q = nb.kb[dir];
            ebv[dir] = d.eps[q]; phv[dir] = phi[q]; cmv[dir] = d.con_mask[q];
        }
        double ek = d.eps[k], pk = phi[k], f = 0.0, dg = 0.0, off[6];
#pragma unroll
        for (int dir = 0; dir < 6; dir++) {
            off[dir] = 0.0;
            if (!nb.has[dir]) continue;
            double eb = ebv[dir], cf = (2. * ek * eb / (ek + eb)) * nb.ar[dir] / nb.h[dir];
            f += cf * (phv[dir] - pk);
            dg += cf;
            if (!cmv[dir]) off[dir] = -cf;
        }
        double vol = g.Lx[i] * g.Ly[j] * g.Lz[l];
        if (d.kind[k] == K_SEMI) {
            double VT = d.mats[d.mat[k]].VT;
            double n = sd_nK(d, k, pk, d.phin[k], VT), p = sd_pK(d, k, pk, d.phip[k], VT);
            f += vol * SD_Q * (p - n + d.Nd[k] - d.Na[k] + d.Nf[k]);
            dg += vol * SD_Q * (n + p) / VT;
        } else f += vol * SD_Q * d.Nf[k];
        int gm = d.gate_mask[k];
        if (gm) {                                       // oxide (Robin) gate
            const ConDev &c = d.cons[gm - 1];
            double ra = d.robA[k];
            f += ra * (c.Cox * (c.VGp - pk) + c.qox);
            dg += c.Cox * ra;
        }
        F[k] = f;
        double s;
        if (build) {
            A.c[k] = dg; A.xm[k] = off[0]; A.xp[k] = off[1]; A.ym[k] = off[2];
            A.yp[k] = off[3]; A.zm[k] = off[4]; A.zp[k] = off[5];
            sc[k] = dg;
            s = dg;
        } else s = sc[k];
        double r = f / s;
        acc[0] += r * r;
    }
    red_partial<1, false>(acc, rb);
}
void ph_poisson(hipStream_t s, const DevView &d, const double *phi, int build, double *F, Sten7<double> A,
                double *sc, RedBuf rb, int slot) {
    int nb = red_grid(d.g.N);
    k_poisson<<<nb, SD_BLOCK, 0, s>>>(d, phi, build, F, A, sc, rb, slot);
    red_final(s, rb, 1, nb, slot, 0);
}

__global__ void k_update_psi(int N, const double *phi, const double *dx, double lam, double *out) {
    int k = blockIdx.x * SD_BLOCK + threadIdx.x;
    if (k < N) out[k] = phi[k] + lam * dx[k];
}
void ph_update_psi(hipStream_t s, int N, double *phi, const double *dx, double lam, double *out) {
    k_update_psi<<<sd_nblocks(N), SD_BLOCK, 0, s>>>(N, phi, dx, lam, out);
}

// --------------------------------------------------------------- carriers
__global__ void k_rebuild_np(DevView d) {
    int k = blockIdx.x * SD_BLOCK + threadIdx.x;
    if (k >= d.g.N) return;
    if (d.kind[k] != K_SEMI) { d.n[k] = d.p[k] = 0.0; return; }
    double VT = d.mats[d.mat[k]].VT, ps = d.psi[k];
    d.n[k] = sd_nK(d, k, ps, d.phin[k], VT);
    d.p[k] = sd_pK(d, k, ps, d.phip[k], VT);
}
void ph_rebuild_np(hipStream_t s, const DevView &d) { k_rebuild_np<<<sd_nblocks(d.g.N), SD_BLOCK, 0, s>>>(d); }

// ------------------------------------------------------- face mobilities
// potential gradient along ax inside the semiconductor (one-sided at an
// interface), semiconductor_3d.c:2909-2917
__device__ __forceinline__ double grad_semi(const DevView &d, int i, int j, int l, int k, int ax) {
    const Geo &g = d.g;
    int km = 0, kp = 0; double hm = 1., hp = 1., ar;
    bool a = sd_nbr(g, i, j, l, k, 2 * ax, km, hm, ar) && d.kind[km] == K_SEMI;
    bool b = sd_nbr(g, i, j, l, k, 2 * ax + 1, kp, hp, ar) && d.kind[kp] == K_SEMI;
    if (a && b) return (d.psi[kp] - d.psi[km]) / (hm + hp);
    if (b) return (d.psi[kp] - d.psi[k]) / hp;
    if (a) return (d.psi[k] - d.psi[km]) / hm;
    return 0.;
}
// squared field normal to an edge along ax: the gradients along the two
// other axes (fnorm2, semiconductor_3d.c:2918-2921)
__device__ __forceinline__ double fnorm2(const DevView &d, int i, int j, int l, int k, int ax) {
    const double g1 = grad_semi(d, i, j, l, k, (ax + 1) % 3), g2 = grad_semi(d, i, j, l, k, (ax + 2) % 3);
    return g1 * g1 + g2 * g2;
}
// one Lombardi evaluation for a material (api3_lombardi :4894-4897)
__global__ void k_lombardi_eval(MatDev m, double mub, double Fv, double Nt, double Dw, int hole, double *out) {
    if (threadIdx.x == 0) out[0] = sd_lombardi(m, mub, Fv, Nt, Dw, hole);
}
void ph_lombardi_eval(hipStream_t s, const MatDev &m, double mub, double Fv, double Nt, double Dw, int hole,
                      double *out) {
    k_lombardi_eval<<<1, 64, 0, s>>>(m, mub, Fv, Nt, Dw, hole, out);
}
__global__ void k_upd_mob(DevView d, const double *qn, const double *qp) {
    const Geo g = d.g;
    int k = sd_tile() * SD_BLOCK + threadIdx.x;
    if (k >= g.N) return;
    int i, j, l;
    sd_unidx(g, k, i, j, l);
    const int sat = (d.models & 4) != 0;
    const size_t N = g.N;
    // the three upper neighbours' operands loaded before the first is used
    Nbr6 ng;
    sd_nbr6(g, i, j, l, k, ng);
    int kdv[3];
    double mnv[3], mpv[3], psv[3], qnv[3], qpv[3];
#pragma unroll
    for (int ax = 0; ax < 3; ax++) {
        const int q = ng.kb[2 * ax + 1];
        kdv[ax] = d.kind[q]; mnv[ax] = d.mu_n[q]; mpv[ax] = d.mu_p[q];
        if (sat) { psv[ax] = d.psi[q]; qnv[ax] = qn[q]; qpv[ax] = qp[q]; }
        else psv[ax] = qnv[ax] = qpv[ax] = 0.0;
    }
    const bool semi = d.kind[k] == K_SEMI;
    for (int ax = 0; ax < 3; ax++) {
        const int kp = ng.kb[2 * ax + 1];
        const double h = ng.h[2 * ax + 1];
        bool ok = semi && ng.has[2 * ax + 1] && kdv[ax] == K_SEMI;
        if (!ok) { d.fmu[ax * N + k] = 0.0; d.fmu[(3 + ax) * N + k] = 0.0; continue; }
        const MatDev &m = d.mats[d.mat[k]];
        double mn = 0.5 * (d.mu_n[k] + mnv[ax]), mp = 0.5 * (d.mu_p[k] + mpv[ax]);
        if (ax == d.caxis) {
            if (m.ani_n > 0.) mn *= m.ani_n;
            if (m.ani_p > 0.) mp *= m.ani_p;
        }
        if (d.qd && m.lb_on) {            // Lombardi surface mobility near an interface
            const double Dw = exp(-0.5 * (d.qd[k] + d.qd[kp]) / SD_LB_LCRIT);
            if (Dw > 1e-8) {
                const int ip = i + (ax == 0), jp = j + (ax == 1), lp = l + (ax == 2);
                const double F = sqrt(0.5 * (fnorm2(d, i, j, l, k, ax) + fnorm2(d, ip, jp, lp, kp, ax)));
                const double Nt = 0.5 * (d.Nd[k] + d.Na[k] + d.Nd[kp] + d.Na[kp]);
                if (m.lb_on & 1) mn = sd_lombardi(m, mn, F, Nt, Dw, 0);
                if (m.lb_on & 2) mp = sd_lombardi(m, mp, F, Nt, Dw, 1);
            }
        }
        if (sat) {
            double wv = 1.0;
            double dps = psv[ax] - d.psi[k];
            double Fn = sd_heat_field(qnv[ax] - qn[k], dps, h), Fp = sd_heat_field(qpv[ax] - qp[k], dps, h);
            double sn = sd_mob_hf(mn, Fn, m.vsat_n, m.beta_n), sp = sd_mob_hf(mp, Fp, m.vsat_p, m.beta_p);
            mn += wv * (sn - mn);
            mp += wv * (sp - mp);
        }
        d.fmu[ax * N + k] = mn;
        d.fmu[(3 + ax) * N + k] = mp;
    }
}
void ph_upd_mob(hipStream_t s, const DevView &d, const double *qn, const double *qp) {
    k_upd_mob<<<sd_nblocks(d.g.N), SD_BLOCK, 0, s>>>(d, qn, qp);
}

// ---------------------------------------------- Scharfetter-Gummel continuity
// Scaled rows: unknown u = c/s (s = current carrier density), row scale
// 1/(diag*s) so the diagonal is 1; dsc = diag*s restores the conservation
// form for the multigrid restriction.
__global__ void k_cont(DevView d, int hole, Sten7<double> A, double *rhs, double *sc, double *dsc) {
    const Geo g = d.g;
    int k = sd_tile() * SD_BLOCK + threadIdx.x;
    if (k >= g.N) return;
    const double *car = hole ? d.p : d.n;
    if (d.kind[k] != K_SEMI || d.con_mask[k]) {
        A.c[k] = 1.0; A.xm[k] = A.xp[k] = A.ym[k] = A.yp[k] = A.zm[k] = A.zp[k] = 0.0;
        bool semi = d.kind[k] == K_SEMI;
        rhs[k] = semi ? 1.0 : 0.0;
        sc[k] = semi ? car[k] : 0.0;
        dsc[k] = 0.0;
        return;
    }
    int i, j, l;
    sd_unidx(g, k, i, j, l);
    const MatDev &ma = d.mats[d.mat[k]];
    const double VT = ma.VT, sk = fmax(car[k], SD_NFLOOR), pk = d.psi[k];
    const double lk = hole ? d.lnvK[k] : d.lncK[k];
    const size_t N = g.N;
    const double *fm = d.fmu + (size_t)(hole ? 3 : 0) * N;
    const double *lnd = hole ? d.lnvK : d.lncK;
    // every neighbour's operands loaded before the first is used
    Nbr6 ng;
    sd_nbr6(g, i, j, l, k, ng);
    double lv[6], psv[6], fmv[6], carv[6];
    int kdv[6], cmv[6];
#pragma unroll
    for (int dir = 0; dir < 6; dir++) {
        const int q = ng.kb[dir], kl = (dir & 1) ? k : q;
        kdv[dir] = d.kind[q]; lv[dir] = lnd[q]; psv[dir] = d.psi[q];
        fmv[dir] = fm[(size_t)(dir >> 1) * N + kl]; cmv[dir] = d.con_mask[q]; carv[dir] = car[q];
    }
    double dg = 0.0, b = 0.0, off[6];
#pragma unroll
    for (int dir = 0; dir < 6; dir++) {
        off[dir] = 0.0;
        if (!ng.has[dir] || kdv[dir] != K_SEMI) continue;
        double dl = lv[dir] - lk;
        double D = (psv[dir] - pk) / VT + (hole ? -dl : dl);
        double gg = ng.ar[dir] * fmv[dir] * VT / ng.h[dir];
        double bp, bm;
        sd_bern2(D, bp, bm);
        double self = hole ? gg * bp : gg * bm;
        double nb = hole ? gg * bm : gg * bp;
        dg += self;
        if (cmv[dir]) b += nb * carv[dir];
        else off[dir] = -nb * fmax(carv[dir], SD_NFLOOR);
    }
    double R, dRn, dRp;
    sd_recomb(ma, d.niK[k], sd_tauf(d, k, ma), d.n[k], d.p[k], R, dRn, dRp);
    double dR = hole ? dRp : dRn;
    if (dR < 0.) dR = 0.;
    double vol = g.Lx[i] * g.Ly[j] * g.Lz[l];
    dg += vol * dR;
    b += vol * (dR * car[k] - R);
    if (d.tk) { dg += d.tk[k]; b += d.tg[k]; }     // tunnelling loss linear in the carrier, plus its source
    if (d.gii) b += d.gii[k];
    double w = 1.0 / (dg * sk);
    A.c[k] = 1.0;
    A.xm[k] = off[0] * w; A.xp[k] = off[1] * w; A.ym[k] = off[2] * w;
    A.yp[k] = off[3] * w; A.zm[k] = off[4] * w; A.zp[k] = off[5] * w;
    rhs[k] = b * w;
    sc[k] = sk;
    dsc[k] = dg * sk;
}
// The same assembly over 32x4x2-node tiles whose neighbour operands (kind,
// contact mask, psi, ln Nc or ln Nv, carrier density) are staged in LDS with a
// one-node halo: a workgroup reads each value of its 34x6x4 halo tile once
// from memory instead of every node reading its six neighbours through the
// caches.  Same operations in the same order as k_cont: identical rows.
#define CT_X 32
#define CT_Y 4
#define CT_Z 2
#define CT_H ((CT_X + 2) * (CT_Y + 2) * (CT_Z + 2))
__global__ void __launch_bounds__(CT_X * CT_Y * CT_Z) k_cont_lds(DevView d, int hole, Sten7<double> A, double *rhs,
                                                                 double *sc, double *dsc) {
    __shared__ double s_ps[CT_H], s_ln[CT_H], s_car[CT_H];
    __shared__ int s_cm[CT_H];
    __shared__ unsigned char s_kd[CT_H];
    const Geo g = d.g;
    const int nbx = (g.Nx + CT_X - 1) / CT_X, nby = (g.Ny + CT_Y - 1) / CT_Y;
    const int b = blockIdx.x, bx = b % nbx, by = (b / nbx) % nby, bz = b / (nbx * nby);
    const int i0 = bx * CT_X - 1, j0 = by * CT_Y - 1, l0 = bz * CT_Z - 1;
    const double *car = hole ? d.p : d.n;
    const double *lnd = hole ? d.lnvK : d.lncK;
    for (int t = threadIdx.x; t < CT_H; t += CT_X * CT_Y * CT_Z) {
        const int hx = t % (CT_X + 2), r = t / (CT_X + 2), hy = r % (CT_Y + 2), hz = r / (CT_Y + 2);
        const int i = i0 + hx, j = j0 + hy, l = l0 + hz;
        if (i >= 0 && i < g.Nx && j >= 0 && j < g.Ny && l >= 0 && l < g.Nz) {
            const int q = i + g.Nx * (j + g.Ny * l);
            s_kd[t] = d.kind[q]; s_cm[t] = d.con_mask[q];
            s_ps[t] = d.psi[q]; s_ln[t] = lnd[q]; s_car[t] = car[q];
        } else {
            s_kd[t] = 255; s_cm[t] = 0; s_ps[t] = s_ln[t] = s_car[t] = 0.0;
        }
    }
    __syncthreads();
    const int tx = threadIdx.x % CT_X, ty = (threadIdx.x / CT_X) % CT_Y, tz = threadIdx.x / (CT_X * CT_Y);
    const int i = i0 + 1 + tx, j = j0 + 1 + ty, l = l0 + 1 + tz;
    if (i >= g.Nx || j >= g.Ny || l >= g.Nz) return;
    const int k = i + g.Nx * (j + g.Ny * l);
    const int c0 = (tx + 1) + (CT_X + 2) * ((ty + 1) + (CT_Y + 2) * (tz + 1));
    if (s_kd[c0] != K_SEMI || s_cm[c0]) {
        A.c[k] = 1.0; A.xm[k] = A.xp[k] = A.ym[k] = A.yp[k] = A.zm[k] = A.zp[k] = 0.0;
        bool semi = s_kd[c0] == K_SEMI;
        rhs[k] = semi ? 1.0 : 0.0;
        sc[k] = semi ? s_car[c0] : 0.0;
        dsc[k] = 0.0;
        return;
    }
    const MatDev &ma = d.mats[d.mat[k]];
    const double VT = ma.VT, sk = fmax(s_car[c0], SD_NFLOOR), pk = s_ps[c0];
    const double lk = s_ln[c0];
    const size_t N = g.N;
    const double *fm = d.fmu + (size_t)(hole ? 3 : 0) * N;
    Nbr6 ng;
    sd_nbr6(g, i, j, l, k, ng);
    const int hs[6] = {-1, 1, -(CT_X + 2), CT_X + 2, -(CT_X + 2) * (CT_Y + 2), (CT_X + 2) * (CT_Y + 2)};
    double fmv[6];
#pragma unroll
    for (int dir = 0; dir < 6; dir++) {
        const int kl = (dir & 1) ? k : ng.kb[dir];
        fmv[dir] = fm[(size_t)(dir >> 1) * N + kl];
    }
    double dg = 0.0, bb = 0.0, off[6];
#pragma unroll
    for (int dir = 0; dir < 6; dir++) {
        off[dir] = 0.0;
        const int h = c0 + hs[dir];
        if (!ng.has[dir] || s_kd[h] != K_SEMI) continue;
        double dl = s_ln[h] - lk;
        double D = (s_ps[h] - pk) / VT + (hole ? -dl : dl);
        double gg = ng.ar[dir] * fmv[dir] * VT / ng.h[dir];
        double bp, bm;
        sd_bern2(D, bp, bm);
        double self = hole ? gg * bp : gg * bm;
        double nb = hole ? gg * bm : gg * bp;
        dg += self;
        if (s_cm[h]) bb += nb * s_car[h];
        else off[dir] = -nb * fmax(s_car[h], SD_NFLOOR);
    }
    double R, dRn, dRp;
    sd_recomb(ma, d.niK[k], sd_tauf(d, k, ma), d.n[k], d.p[k], R, dRn, dRp);
    double dR = hole ? dRp : dRn;
    if (dR < 0.) dR = 0.;
    double vol = g.Lx[i] * g.Ly[j] * g.Lz[l];
    dg += vol * dR;
    bb += vol * (dR * s_car[c0] - R);
    if (d.tk) { dg += d.tk[k]; bb += d.tg[k]; }
    if (d.gii) bb += d.gii[k];
    double w = 1.0 / (dg * sk);
    A.c[k] = 1.0;
    A.xm[k] = off[0] * w; A.xp[k] = off[1] * w; A.ym[k] = off[2] * w;
    A.yp[k] = off[3] * w; A.zm[k] = off[4] * w; A.zp[k] = off[5] * w;
    rhs[k] = bb * w;
    sc[k] = sk;
    dsc[k] = dg * sk;
}
int g_cont_lds = 1;          // 1: the LDS-staged continuity assembly
void ph_cont(hipStream_t s, const DevView &d, int hole, Sten7<double> A, double *rhs, double *sc, double *dsc) {
    if (g_cont_lds) {
        const Geo &g = d.g;
        const int nb = ((g.Nx + CT_X - 1) / CT_X) * ((g.Ny + CT_Y - 1) / CT_Y) * ((g.Nz + CT_Z - 1) / CT_Z);
        k_cont_lds<<<nb, CT_X * CT_Y * CT_Z, 0, s>>>(d, hole, A, rhs, sc, dsc);
    } else {
        k_cont<<<sd_nblocks(d.g.N), SD_BLOCK, 0, s>>>(d, hole, A, rhs, sc, dsc);
    }
}

// c = s*u -> quasi-Fermi potential clamped to [lo, hi]; weighted max change
__global__ void __launch_bounds__(SD_BLOCK) k_cont_update(DevView d, int hole, const double *sol, const double *sc,
                                                          double *dq, double lo, double hi, RedBuf rb, int slot) {
    double acc[1] = {0.0};
    for (int k = blockIdx.x * SD_BLOCK + threadIdx.x; k < d.g.N; k += gridDim.x * SD_BLOCK) {
        if (d.kind[k] != K_SEMI || d.con_mask[k]) { dq[k] = 0.0; continue; }
        double VT = d.mats[d.mat[k]].VT, ps = d.psi[k];
        double u = sol[k];
        if (!(u > 1e-30)) u = 1e-30; else if (u > 1e30) u = 1e30;
        double c = sc[k] * u, q;
        if (hole) q = ps + VT * (log(c) - d.lnvK[k]);
        else q = ps - VT * (log(c) - d.lncK[k]);
        if (q < lo) q = lo; else if (q > hi) q = hi;
        double *qf = hole ? d.phip : d.phin;
        double ch = q - qf[k];
        qf[k] = q;
        dq[k] = ch;
        double cn = hole ? sd_pK(d, k, ps, q, VT) : sd_nK(d, k, ps, q, VT);
        if (hole) d.p[k] = cn; else d.n[k] = cn;
        double scale = d.n[k] + d.p[k] + fabs(d.Nd[k] - d.Na[k]);
        double r = scale > 0. ? cn / scale : 0.;
        if (r > 1e-6) {
            double w = r >= 1e-3 ? 1. : r * 1e3, a = w * fabs(ch);
            acc[0] = sd_nanmax(acc[0], a);
        }
    }
    red_partial<1, true>(acc, rb);
}
void ph_cont_update(hipStream_t s, const DevView &d, int hole, const double *sol, const double *sc, double *dq,
                    double lo, double hi, RedBuf rb, int slot) {
    int nb = red_grid(d.g.N);
    k_cont_update<<<nb, SD_BLOCK, 0, s>>>(d, hole, sol, sc, dq, lo, hi, rb, slot);
    red_final(s, rb, 1, nb, slot, 1);
}

__global__ void k_clamp_qf(DevView d, double lo, double hi) {
    int k = blockIdx.x * SD_BLOCK + threadIdx.x;
    if (k >= d.g.N) return;
    double a = d.phin[k], b = d.phip[k];
    if (a < lo) a = lo; else if (a > hi) a = hi;
    if (b < lo) b = lo; else if (b > hi) b = hi;
    d.phin[k] = a;
    d.phip[k] = b;
}
void ph_clamp_qf(hipStream_t s, const DevView &d, double lo, double hi) {
    k_clamp_qf<<<sd_nblocks(d.g.N), SD_BLOCK, 0, s>>>(d, lo, hi);
}

// Dirichlet values of contact nodes at the applied voltages
__global__ void k_apply_bc(DevView d, int have_solution) {
    int k = blockIdx.x * SD_BLOCK + threadIdx.x;
    if (k >= d.g.N) return;
    int c = d.con_mask[k];
    if (!c) return;
    const ConDev &cn = d.cons[c - 1];
    double V = cn.V;
    if (d.kind[k] == K_METAL) {
        double pm = cn.phi_m > 0. ? cn.phi_m : d.Cref;
        d.psi[k] = V + d.Cref - pm;
        d.phin[k] = d.phip[k] = V;
        d.n[k] = d.p[k] = 0.0;
        return;
    }
    double VT = d.mats[d.mat[k]].VT;
    if (cn.bc == 1) {
        double pm = cn.phi_m > 0. ? cn.phi_m : d.Cref;
        d.psi[k] = V + d.Cref - pm;
        if (cn.te && have_solution) {
            d.n[k] = sd_nK(d, k, d.psi[k], d.phin[k], VT);
            d.p[k] = sd_pK(d, k, d.psi[k], d.phip[k], VT);
            return;
        }
    } else d.psi[k] = neutral_psi(d, k, V, VT);
    d.phin[k] = d.phip[k] = V;
    d.n[k] = sd_nK(d, k, d.psi[k], V, VT);
    d.p[k] = sd_pK(d, k, d.psi[k], V, VT);
}
void ph_apply_bc(hipStream_t s, const DevView &d, int have_solution) {
    k_apply_bc<<<sd_nblocks(d.g.N), SD_BLOCK, 0, s>>>(d, have_solution);
}

__global__ void k_init_eq(DevView d) {
    int k = blockIdx.x * SD_BLOCK + threadIdx.x;
    if (k >= d.g.N) return;
    d.phin[k] = d.phip[k] = 0.0;
    if (d.kind[k] == K_SEMI) {
        double VT = d.mats[d.mat[k]].VT;
        double ps = neutral_psi(d, k, 0.0, VT);
        d.psi[k] = ps;
        d.n[k] = sd_nK(d, k, ps, 0.0, VT);
        d.p[k] = sd_pK(d, k, ps, 0.0, VT);
    } else {
        d.psi[k] = 0.0;
        d.n[k] = d.p[k] = 0.0;
    }
}
void ph_init_eq(hipStream_t s, const DevView &d) { k_init_eq<<<sd_nblocks(d.g.N), SD_BLOCK, 0, s>>>(d); }

// quasi-neutral nodes follow their quasi-Fermi change (gummel :3853-3859)
__global__ void k_neutral_pred(DevView d, const double *dqn, const double *dqp) {
    int k = blockIdx.x * SD_BLOCK + threadIdx.x;
    if (k >= d.g.N) return;
    if (d.kind[k] != K_SEMI || d.con_mask[k]) return;
    double n = d.n[k], p = d.p[k], D = fabs(d.Nd[k] - d.Na[k]);
    double den = fmax(n + p, D);
    if (!(den > 0.)) return;
    d.psi[k] += (n * dqn[k] + p * dqp[k]) / den;
}
void ph_neutral_pred(hipStream_t s, const DevView &d, const double *dqn, const double *dqp) {
    k_neutral_pred<<<sd_nblocks(d.g.N), SD_BLOCK, 0, s>>>(d, dqn, dqp);
}

// ---------------------------------------------------- SG edge flux (A)
__device__ __forceinline__ void face_flux(const DevView &d, int ka, int kb, double h, double ar, double &Jn,
                                          double &Jp, double &Sn, double &Sp) {
    const double VT = d.mats[d.mat[ka]].VT;
    double dps = (d.psi[kb] - d.psi[ka]) / VT;
    double Dn = dps + (d.lncK[kb] - d.lncK[ka]), Dp = dps - (d.lnvK[kb] - d.lnvK[ka]);
    int df = ka < kb ? kb - ka : ka - kb;
    int ax = df == 1 ? 0 : (df == d.g.Nx ? 1 : 2);
    int kl = ka < kb ? ka : kb;
    size_t N = d.g.N;
    double gn = SD_Q * ar * d.fmu[ax * N + kl] * VT / h;
    double gp = SD_Q * ar * d.fmu[(3 + ax) * N + kl] * VT / h;
    double bpn, bmn, bpp, bmp;
    sd_bern2(Dn, bpn, bmn);
    sd_bern2(Dp, bpp, bmp);
    double a1 = gn * d.n[kb] * bpn, a2 = gn * d.n[ka] * bmn;
    double b1 = gp * d.p[ka] * bpp, b2 = gp * d.p[kb] * bmp;
    Jn = a1 - a2; Jp = b1 - b2; Sn = a1 + a2; Sp = b1 + b2;
}

// ------------------------------------------------------- gate tunnelling
// Tsu-Esaki current with WKB transmission through the layer stack of each
// path, semiconductor_3d.c:3230-3390.  The flux of a carrier between the
// semiconductor and the gate metal is
//     F = 4 pi m_s q^2 kT / h^3  Int T(E) ln(1 + exp((E_F - E)/kT)) dE
// (1/m^2 s), integrated by Simpson's rule over 64 intervals from the
// effective band edge up to min(max(E_F - E0, 0) + 25 kT, 3 eV) (holes
// mirrored).  The supply of carriers pressed against the interface by the
// field follows from the contact theorem summed over the path's column;
// carriers pushed away use the local density extrapolated to the interface.
#define SD_HBAR    1.054571817e-34
#define SD_M0      9.1093837015e-31
#define SD_HPLANCK 6.62607015e-34
#define SD_TUN_NE   64
#define SD_TUN_EMAX 3.0
__device__ __forceinline__ double tun_softplus(double x) { return x > 35. ? x + log1p(exp(-x)) : log1p(exp(x)); }
// 2 Int kappa dx over a layer of length L whose barrier above the carrier
// energy runs linearly from u0 to u1 (eV); kappa = sqrt(2 m q u) / hbar
__device__ __forceinline__ double tun_wkb(double m, double L, double u0, double u1) {
    double a = u0 > 0. ? u0 : 0., b = u1 > 0. ? u1 : 0.;
    if (a == 0. && b == 0.) return 0.;
    double c = 2. * sqrt(2. * m * SD_M0 * SD_Q) / SD_HBAR, du = u1 - u0, I;
    if (fabs(du) < 1e-9 * (a + b)) I = L * sqrt(0.5 * (a + b));
    else I = L * (2. / 3.) * (b * sqrt(b) - a * sqrt(a)) / du;
    return c * I;
}
struct TunStack {
    int nl, mi[SD_TUN_MAXL];
    double t[SD_TUN_MAXL], p0[SD_TUN_MAXL], dp[SD_TUN_MAXL];
    double D, psis;
};
// potentials along the stack between the node at psik and the metal at
// potential psim (= V + Cref - work function): D = (psim - psik) / sum t/eps,
// psi linear in every layer (tun_stack :3300-3312)
__device__ void tun_stack(const MatDev *mats, TunStack &P, int nl, const int *mi, const double *t, double psik,
                          double psim) {
    P.nl = nl;
    double R = 0.;
    for (int i = 0; i < nl; i++) { P.mi[i] = mi[i]; P.t[i] = t[i]; R += t[i] / (mats[mi[i]].eps_r * SD_EPS0); }
    double ps = psik;
    P.D = R > 0. ? (psim - psik) / R : 0.;
    P.psis = psik;
    for (int i = 0; i < nl; i++) {
        P.p0[i] = ps; P.dp[i] = P.D * t[i] / (mats[mi[i]].eps_r * SD_EPS0); ps += P.dp[i];
        if (i == 0 && !mats[mi[0]].insulator) P.psis = ps;
    }
}
// WKB transmission at energy E (eV) through the insulating layers
__device__ double tun_T(const MatDev *mats, const TunStack &P, double E, int hole) {
    double s = 0.;
    for (int i = 0; i < P.nl; i++) {
        const MatDev &m = mats[P.mi[i]];
        if (!m.insulator) continue;
        double mt = hole ? m.mt_h : m.mt_e;
        if (!(mt > 0.)) mt = 0.5;
        double ea = -P.p0[i] - m.chi, eb = -(P.p0[i] + P.dp[i]) - m.chi;     // Ec of the layer at both ends
        double u0, u1;
        if (hole) { u0 = E - (ea - m.Eg); u1 = E - (eb - m.Eg); }
        else { u0 = ea - E; u1 = eb - E; }
        s += tun_wkb(mt, P.t[i], u0, u1);
        if (s > 745.) return 0.;
    }
    return exp(-s);
}
// Int T(E) ln(1 + exp(+-(EF - E)/kT)) dE: the 65 Simpson nodes spread over
// the lanes of one wavefront (lane 0 also takes the last); the sum is valid
// in lane 0
__device__ double tun_int_wave(const MatDev *mats, const TunStack &P, double E0, double EF, double kT, int hole,
                               int lane) {
    double sg = hole ? -1. : 1.;
    double top = sg * (EF - E0);
    if (top < 0.) top = 0.;
    double span = top + 25. * kT;
    if (span > SD_TUN_EMAX) span = SD_TUN_EMAX;
    double h = span / SD_TUN_NE, s = 0.;
    for (int q = lane; q <= SD_TUN_NE; q += 64) {
        double E = E0 + sg * q * h;
        double f = tun_T(mats, P, E, hole) * tun_softplus(sg * (EF - E) / kT);
        s += f * (q == 0 || q == SD_TUN_NE ? 1. : (q & 1 ? 4. : 2.));
    }
    s = sd_wave_sum(s);
    return s * h / 3.;
}
// the four fluxes (1/m^2 s): electrons s->g, g->s, holes s->g, g->s, for
// effective band edges Ecp, Evp and Fermi levels EFn, EFp (semiconductor),
// EFm (metal), in eV; valid in lane 0
__device__ void tun_flux_wave(const MatDev *mats, const MatDev &ms, const TunStack &P, double Ecp, double Evp,
                              double EFn, double EFp, double EFm, double *F, int lane) {
    double kT = ms.VT, h3 = SD_HPLANCK * SD_HPLANCK * SD_HPLANCK;
    double ce = 4. * M_PI * ms.ms_e * SD_M0 * SD_Q * SD_Q * kT / h3;
    double chh = 4. * M_PI * ms.ms_h * SD_M0 * SD_Q * SD_Q * kT / h3;
    F[0] = ce * tun_int_wave(mats, P, Ecp, EFn, kT, 0, lane);
    F[1] = ce * tun_int_wave(mats, P, Ecp, EFm, kT, 0, lane);
    F[2] = chh * tun_int_wave(mats, P, Evp, EFp, kT, 1, lane);
    F[3] = chh * tun_int_wave(mats, P, Evp, EFm, kT, 1, lane);
}
// escape rates (1/s) of path p at the current densities (tun_escape :3382-3389)
__device__ __forceinline__ void tun_escape(const DevView &d, const TunDev &T, int p, double &se, double &sh) {
    double a = 0., b = 0.;
    for (int q = T.pc0[p]; q < T.pc0[p] + T.pcn[p]; q++) {
        int kq = T.ck[q];
        a += T.cK[2 * q] * d.n[kq]; b += T.cK[2 * q + 1] * d.p[kq];
    }
    se = a; sh = b;
}
// one wavefront per path (tun_path :3326-3366): effective supply densities
// from the column, the fluxes, and the linearised exchange: cK (escape per
// unit density of each column node), pG (injection from the metal)
__global__ void __launch_bounds__(SD_BLOCK) k_tun_path(DevView d, TunDev T) {
    const int lane = threadIdx.x & 63;
    const int p = blockIdx.x * (SD_BLOCK / 64) + (threadIdx.x >> 6);
    if (p >= T.np) return;                        // uniform over the wavefront
    const int k = T.pk[p], c = T.pc[p];
    const MatDev &ms = d.mats[d.mat[k]];
    const double VT = ms.VT, A = T.pA[p];
    TunStack P;
    tun_stack(d.mats, P, T.pnl[p], T.pm + (size_t)p * SD_TUN_MAXL, T.pt + (size_t)p * SD_TUN_MAXL, d.psi[k],
              d.cons[c].VGp);
    const int q0 = T.pc0[p], nq = T.pcn[p];
    const double Fs = P.D / (ms.eps_r * SD_EPS0);  // surface field, + pushes electrons to the wall
    double neff[2];
    for (int hole = 0; hole < 2; hole++) {
        const double sg = hole ? -1. : 1., Fsc = sg * Fs;
        const double *car = hole ? d.p : d.n;
        double *K = T.cK + 2 * q0 + hole;        // column weights, written by lane 0
        if (Fsc > 0.) {                           // pressed against the wall: contact theorem
            double P0 = 0., Fl = Fsc;
            for (int q = 0; q < nq; q++) {
                int kq = T.ck[q0 + q];
                double Fr = (q < nq - 1 && T.ch[q0 + q] > 0.)
                                ? sg * (d.psi[kq] - d.psi[T.ck[q0 + q + 1]]) / T.ch[q0 + q] : Fl;
                double Fm = 0.5 * (Fl + Fr);
                if (Fm < 0.) Fm = 0.;
                double w = Fm * T.cl[q0 + q] / VT;
                if (lane == 0) K[2 * q] = q == nq - 1 ? w + 1. : w;     // + kT n(z1) at the column's end
                P0 += w * car[kq];
                Fl = Fr;
            }
            P0 += car[T.ck[q0 + nq - 1]];
            neff[hole] = P0;
        } else {                                  // pushed away: local density at the interface
            const double *qc = hole ? T.qcp : T.qcn;
            double qv = qc ? qc[k] : 0.;
            neff[hole] = car[k] * exp(-qv + sg * (P.psis - d.psi[k]) / VT);
            if (lane == 0)
                for (int q = 0; q < nq; q++) K[2 * q] = q == 0 ? neff[hole] / fmax(car[k], SD_NFLOOR) : 0.;
        }
        if (!(neff[hole] > 1e-100)) neff[hole] = 1e-100;
    }
    const double EFn = -d.phin[k] - d.Cref, EFp = -d.phip[k] - d.Cref, EFm = -d.cons[c].V - d.Cref;
    const double Ecp = EFn - VT * log(neff[0] / ms.Nc), Evp = EFp + VT * log(neff[1] / ms.Nv);
    double F[4];
    tun_flux_wave(d.mats, ms, P, Ecp, Evp, EFn, EFp, EFm, F, lane);
    if (lane == 0) {
        double be = A * F[0] / neff[0], bh = A * F[2] / neff[1];
        for (int q = q0; q < q0 + nq; q++) { T.cK[2 * q] *= be; T.cK[2 * q + 1] *= bh; }
        T.pG[2 * p] = A * F[1]; T.pG[2 * p + 1] = A * F[3];
    }
}
// fluxes through a given stack from a semiconductor surface at psi with
// quasi-Fermi potentials phin, phip, local supply (api3_tun_eval :4855-4867)
struct TunEvalArgs {
    int smat, nl, mi[SD_TUN_MAXL];
    double t[SD_TUN_MAXL];
    double psi, phin, phip, Vg, pm, Cref;
};
__global__ void __launch_bounds__(64) k_tun_eval(const MatDev *mats, TunEvalArgs a, double *F) {
    const int lane = threadIdx.x & 63;
    const MatDev &ms = mats[a.smat];
    double pm = a.pm > 0. ? a.pm : a.Cref;
    TunStack P;
    tun_stack(mats, P, a.nl, a.mi, a.t, a.psi, a.Vg + a.Cref - pm);
    double Ec = -P.psis - ms.chi, Ev = Ec - ms.Eg, Fl[4];
    tun_flux_wave(mats, ms, P, Ec, Ev, -a.phin - a.Cref, -a.phip - a.Cref, -a.Vg - a.Cref, Fl, lane);
    if (lane == 0) for (int i = 0; i < 4; i++) F[i] = Fl[i];
}
void ph_tun_eval(hipStream_t s, const MatDev *mats, int smat, int nl, const int *mi, const double *t, double psi,
                 double phin, double phip, double Vg, double pm, double Cref, double *F) {
    TunEvalArgs a{};
    a.smat = smat; a.nl = nl;
    for (int i = 0; i < nl; i++) { a.mi[i] = mi[i]; a.t[i] = t[i]; }
    a.psi = psi; a.phin = phin; a.phip = phip; a.Vg = Vg; a.pm = pm; a.Cref = Cref;
    k_tun_eval<<<1, 64, 0, s>>>(mats, a, F);
}
void ph_tun_update(hipStream_t s, const DevView &d, const TunDev &T) {
    if (T.np > 0) k_tun_path<<<sd_nblocks(T.np, SD_BLOCK / 64), SD_BLOCK, 0, s>>>(d, T);
}
// sink tk and source tg of one carrier at every node a column touches, summed
// in the reference's order (tun_scatter :3373-3380); other nodes stay 0
__global__ void k_tun_gather(TunDev T, int hole, double *tk, double *tg) {
    int t = blockIdx.x * SD_BLOCK + threadIdx.x;
    if (t >= T.ntn) return;
    double a = 0., b = 0.;
    for (int e = T.offK[t]; e < T.offK[t + 1]; e++) a += T.cK[2 * T.entK[e] + hole];
    for (int e = T.offG[t]; e < T.offG[t + 1]; e++) b += T.pG[2 * T.entG[e] + hole];
    tk[T.tn[t]] = a; tg[T.tn[t]] = b;
}
void ph_tun_gather(hipStream_t s, const TunDev &T, int hole, double *tk, double *tg) {
    if (T.ntn > 0) k_tun_gather<<<sd_nblocks(T.ntn), SD_BLOCK, 0, s>>>(T, hole, tk, tg);
}
// gate tunnelling current density (A/m^2, into the device through the gate)
// at the interface nodes, for display (upd_J :3648-3657); out is 0 elsewhere
__global__ void k_tun_J(DevView d, TunDev T, double *out) {
    int t = blockIdx.x * SD_BLOCK + threadIdx.x;
    if (t >= T.ntn) return;
    double J = 0., A = 0.;
    for (int e = T.offG[t]; e < T.offG[t + 1]; e++) {
        int p = T.entG[e];
        double se, sh;
        tun_escape(d, T, p, se, sh);
        J += SD_Q * ((se - T.pG[2 * p]) - (sh - T.pG[2 * p + 1]));
        A += T.pA[p];
    }
    out[T.tn[t]] = A > 0. ? J / A : 0.;
}
void ph_tun_J(hipStream_t s, const DevView &d, const TunDev &T, double *out) {
    if (T.ntn > 0) k_tun_J<<<sd_nblocks(T.ntn), SD_BLOCK, 0, s>>>(d, T, out);
}

// terminal currents: each contact's edges (and tunnelling paths) split over
// `split` workgroups, each writing its three partial sums; a one-workgroup
// pass then adds every contact's partials in order (fixed-order sums)
__global__ void __launch_bounds__(SD_BLOCK) k_currents(DevView d, CurEdges e, RedBuf rb, int split) {
    __shared__ double lds[3][SD_BLOCK / 64];
    const int c = blockIdx.x / split, part = blockIdx.x % split, stride = split * SD_BLOCK;
    double a0 = 0., a1 = 0., a2 = 0.;
    for (int q = e.off[c] + part * SD_BLOCK + threadIdx.x; q < e.off[c + 1]; q += stride) {
        double Jn, Jp, Sn, Sp;
        face_flux(d, e.k[q], e.kb[q], e.h[q], e.ar[q], Jn, Jp, Sn, Sp);
        a0 += Jn + Jp; a1 += Sn + Sp; a2 += Jn;
    }
    // gate tunnelling (:3614-3621): electrons leaving to the gate are current
    // into the device through it, holes leaving are current out of it
    if (e.poff)
        for (int q = e.poff[c] + part * SD_BLOCK + threadIdx.x; q < e.poff[c + 1]; q += stride) {
            const int p = e.pidx[q];
            double se, sh;
            tun_escape(d, e.tun, p, se, sh);
            const double G0 = e.tun.pG[2 * p], G1 = e.tun.pG[2 * p + 1];
            double Ie = SD_Q * (se - G0), Ih = -SD_Q * (sh - G1);
            a0 += Ie + Ih; a1 += SD_Q * (se + G0 + sh + G1); a2 += Ie;
        }
    a0 = sd_wave_sum(a0); a1 = sd_wave_sum(a1); a2 = sd_wave_sum(a2);
    int w = threadIdx.x >> 6, ln = threadIdx.x & 63;
    if (ln == 0) { lds[0][w] = a0; lds[1][w] = a1; lds[2][w] = a2; }
    __syncthreads();
    if (threadIdx.x < 3) {
        const int v = threadIdx.x;
        double s = lds[v][0];
        for (int q = 1; q < SD_BLOCK / 64; q++) s += lds[v][q];
        rb.part[v * SD_RED_MAXB + blockIdx.x] = s;
    }
}
__global__ void k_currents_sum(RedBuf rb, int ncon, int split) {
    const int t = threadIdx.x;
    if (t >= 3 * ncon) return;
    const int c = t / 3, v = t % 3;
    double s = 0.;
    for (int q = 0; q < split; q++) s += rb.part[v * SD_RED_MAXB + c * split + q];
    rb.scal[S_CUR + 3 * c + v] = s;
}
void ph_currents(hipStream_t s, const DevView &d, CurEdges e, RedBuf rb) {
    if (e.ncon <= 0) return;
    int split = SD_RED_MAXB / e.ncon;
    if (split > 32) split = 32;
    if (split < 1) split = 1;
    k_currents<<<e.ncon * split, SD_BLOCK, 0, s>>>(d, e, rb, split);
    k_currents_sum<<<1, (3 * e.ncon + 63) / 64 * 64, 0, s>>>(rb, e.ncon, split);
}

__global__ void k_extrap(int N, double *d0, double *d1, double *d2, const double *a0, const double *a1,
                         const double *a2, const double *b0, const double *b1, const double *b2, double r) {
    int k = blockIdx.x * SD_BLOCK + threadIdx.x;
    if (k >= N) return;
    double x1 = a0[k], x2 = b0[k]; d0[k] = x1 + r * (x1 - x2);
    x1 = a1[k]; x2 = b1[k]; d1[k] = x1 + r * (x1 - x2);
    x1 = a2[k]; x2 = b2[k]; d2[k] = x1 + r * (x1 - x2);
}
void ph_extrap(hipStream_t s, int N, double *dst0, double *dst1, double *dst2, const double *a0, const double *a1,
               const double *a2, const double *b0, const double *b1, const double *b2, double r) {
    k_extrap<<<sd_nblocks(N), SD_BLOCK, 0, s>>>(N, dst0, dst1, dst2, a0, a1, a2, b0, b1, b2, r);
}
// quadratic extrapolation through three states a (step 0), b (step -1), c
// (step -2) of uniform step, to step r: a + r (a - b) + r (r + 1) / 2 (a - 2 b + c)
struct Ptr3 { double *p[3]; };
struct CPtr3 { const double *p[3]; };
__global__ void k_extrap2v(int N, Ptr3 d, CPtr3 a, CPtr3 b, CPtr3 c, double r) {
    int k = blockIdx.x * SD_BLOCK + threadIdx.x;
    if (k >= N) return;
    const double q = 0.5 * r * (r + 1.0);
#pragma unroll
    for (int f = 0; f < 3; f++) {
        const double x0 = a.p[f][k], x1 = b.p[f][k], x2 = c.p[f][k];
        d.p[f][k] = x0 + r * (x0 - x1) + q * (x0 - 2.0 * x1 + x2);
    }
}
void ph_extrap2(hipStream_t s, int N, double *const d[3], const double *const a[3], const double *const b[3],
                const double *const c[3], double r) {
    Ptr3 D{{d[0], d[1], d[2]}};
    CPtr3 A{{a[0], a[1], a[2]}}, B{{b[0], b[1], b[2]}}, C{{c[0], c[1], c[2]}};
    k_extrap2v<<<sd_nblocks(N), SD_BLOCK, 0, s>>>(N, D, A, B, C, r);
}

// --------------------------------------------- floating majority regions
__global__ void k_maj_flags(DevView d, int hole, const double *u, const double *s, unsigned char *flag) {
    const Geo g = d.g;
    int k = blockIdx.x * SD_BLOCK + threadIdx.x;
    if (k >= g.N) return;
    const double *oth = hole ? d.n : d.p;
    auto maj = [&](int q) -> bool {
        double c = s[q] * u[q], D = 0.5 * fabs(d.Nd[q] - d.Na[q]);
        return d.kind[q] == K_SEMI && !d.con_mask[q] && c >= oth[q] && c >= D && s[q] >= D && c > 10. * d.niK[q];
    };
    if (!maj(k)) { flag[k] = 0; return; }
    int i, j, l;
    sd_unidx(g, k, i, j, l);
    unsigned char f = 1;
    for (int dir = 0; dir < 6; dir++) {
        int kb; double h, ar;
        if (!sd_nbr(g, i, j, l, k, dir, kb, h, ar) || d.kind[kb] != K_SEMI) continue;
        if (d.con_mask[kb]) f = 2;
    }
    flag[k] = f;
}
void ph_maj_flags(hipStream_t s, const DevView &d, int hole, const double *u, const double *sc, unsigned char *flag) {
    k_maj_flags<<<sd_nblocks(d.g.N), SD_BLOCK, 0, s>>>(d, hole, u, sc, flag);
}
// anchoring travels along whole lines of one axis per launch, one wavefront
// per line: 64 consecutive elements per step, whose flags become two ballot
// masks (majority M: flag != 0, anchored A: flag == 2).  Within a step the
// anchored set spreads through runs of M by doubling shifts (after shift s,
// bit p holds whether an anchored element lies within the 2s elements of its
// run ending at p); the run's state carries into the next step.  A forward
// pass, then a backward pass over the flags it left: every majority run of
// the line that holds an anchored element ends anchored, as a sequential walk
// each way leaves it.
__device__ __forceinline__ unsigned long long maj_spread(unsigned long long A, unsigned long long M, bool fwd) {
    unsigned long long P = M;
#pragma unroll
    for (int s = 1; s < 64; s <<= 1) {
        if (fwd) { A |= (A << s) & P; P &= P << s; }
        else { A |= (A >> s) & P; P &= P >> s; }
    }
    return A;
}
#define MP_LINES 4
__global__ void __launch_bounds__(64 * MP_LINES) k_maj_prop(Geo g, unsigned char *flag, int axis, int *changed) {
    const int lane = threadIdx.x & 63, t = blockIdx.x * MP_LINES + (threadIdx.x >> 6);
    int n, stride, k0;
    if (axis == 0) {
        if (t >= g.Ny * g.Nz) return;
        n = g.Nx; stride = 1; k0 = t * g.Nx;
    } else if (axis == 1) {
        if (t >= g.Nx * g.Nz) return;
        int i = t % g.Nx, l = t / g.Nx;
        n = g.Ny; stride = g.Nx; k0 = i + g.Nx * g.Ny * l;
    } else {
        if (t >= g.Nx * g.Ny) return;
        n = g.Nz; stride = g.Nx * g.Ny; k0 = t;
    }
    // (every lane of a wavefront has the same t, so whole wavefronts return)
    const int nch = (n + 63) >> 6;
    bool ch = false;
    for (int dir = 0; dir < 2; dir++) {
        const bool fwd = dir == 0;
        bool carry = false;
        for (int c = 0; c < nch; c++) {
            const int cc = fwd ? c : nch - 1 - c, q = (cc << 6) + lane;
            const bool ok = q < n;
            const size_t k = (size_t)k0 + (size_t)q * stride;
            const unsigned char f = ok ? flag[k] : 0;
            const unsigned long long M = __ballot(f != 0), A0 = __ballot(f == 2);
            const unsigned long long in = fwd ? 1ull : (1ull << 63);
            const unsigned long long A = maj_spread(A0 | (carry ? (M & in) : 0ull), M, fwd);
            if (((A & ~A0) >> lane) & 1ull) { flag[k] = 2; ch = true; }
            carry = ((fwd ? A >> 63 : A) & 1ull) != 0;
        }
    }
    if (__ballot(ch) && lane == 0) *changed = 1;
}
void ph_maj_propagate(hipStream_t s, const DevView &d, unsigned char *flag, int axis, int *changed) {
    const Geo &g = d.g;
    int n = axis == 0 ? g.Ny * g.Nz : (axis == 1 ? g.Nx * g.Nz : g.Nx * g.Ny);
    k_maj_prop<<<(n + MP_LINES - 1) / MP_LINES, 64 * MP_LINES, 0, s>>>(g, flag, axis, changed);
}
// indices of the majority nodes whose block reached no contact (flag 1);
// the order of the list is arbitrary, the host sorts it
__global__ void k_maj_list(int N, const unsigned char *flag, int *cnt, int *list, int cap) {
    int k = blockIdx.x * SD_BLOCK + threadIdx.x;
    if (k >= N || flag[k] != 1) return;
    int pos = atomicAdd(cnt, 1);
    if (pos < cap) list[pos] = k;
}
void ph_maj_list(hipStream_t s, int N, const unsigned char *flag, int *cnt, int *list, int cap) {
    k_maj_list<<<sd_nblocks(N), SD_BLOCK, 0, s>>>(N, flag, cnt, list, cap);
}
// per listed node the values the block balance reads: u, s, the other
// carrier, psi, then per direction psi of the neighbour, the face mobility
// and the neighbour's carrier (its contact value on a contact node, else s*u)
__global__ void k_maj_gather(DevView d, int hole, const double *u, const double *s, const int *list, int m,
                             double *out) {
    const Geo g = d.g;
    int t = blockIdx.x * SD_BLOCK + threadIdx.x;
    if (t >= m) return;
    int k = list[t], i, j, l;
    sd_unidx(g, k, i, j, l);
    const double *oth = hole ? d.n : d.p, *car = hole ? d.p : d.n;
    double *o = out + (size_t)t * SD_MAJ_REC;
    o[0] = u[k]; o[1] = s[k]; o[2] = oth[k]; o[3] = d.psi[k];
    for (int dir = 0; dir < 6; dir++) {
        int kb; double h, ar;
        double *q = o + 4 + 3 * dir;
        if (!sd_nbr(g, i, j, l, k, dir, kb, h, ar)) { q[0] = q[1] = q[2] = 0.; continue; }
        int kl = k < kb ? k : kb;
        q[0] = d.psi[kb];
        q[1] = d.fmu[(size_t)((hole ? 3 : 0) + dir / 2) * g.N + kl];
        q[2] = d.con_mask[kb] ? car[kb] : s[kb] * u[kb];
    }
    o[22] = d.gii ? d.gii[k] : 0.;
    o[23] = d.tk ? d.tk[k] : 0.;
    o[24] = d.tg ? d.tg[k] : 0.;
}
void ph_maj_gather(hipStream_t s, const DevView &d, int hole, const double *u, const double *sc, const int *list,
                   int m, double *out) {
    k_maj_gather<<<sd_nblocks(m), SD_BLOCK, 0, s>>>(d, hole, u, sc, list, m, out);
}
__global__ void k_scale_list(double *u, const int *list, const double *lam, int m) {
    int t = blockIdx.x * SD_BLOCK + threadIdx.x;
    if (t < m) u[list[t]] *= lam[t];
}
void ph_scale_list(hipStream_t s, double *u, const int *list, const double *lam, int m) {
    k_scale_list<<<sd_nblocks(m), SD_BLOCK, 0, s>>>(u, list, lam, m);
}
__global__ void __launch_bounds__(SD_BLOCK) k_maj_count(int N, const unsigned char *flag, RedBuf rb, int slot) {
    double acc[1] = {0.0};
    for (int k = blockIdx.x * SD_BLOCK + threadIdx.x; k < N; k += gridDim.x * SD_BLOCK) acc[0] += flag[k] == 1;
    red_partial<1, false>(acc, rb);
}
void ph_maj_count(hipStream_t s, int N, const unsigned char *flag, RedBuf rb, int slot) {
    int nb = red_grid(N);
    k_maj_count<<<nb, SD_BLOCK, 0, s>>>(N, flag, rb, slot);
    red_final(s, rb, 1, nb, slot, 0);
}

// ------------------------------------------------------------- Anderson
__device__ __forceinline__ bool aa_sig(const DevView &d, int k, int hole) {
    if (d.kind[k] != K_SEMI || d.con_mask[k]) return false;
    double c = hole ? d.p[k] : d.n[k];
    return c > 1e-6 * (d.n[k] + d.p[k] + fabs(d.Nd[k] - d.Na[k]));
}
__global__ void k_aa_diff(DevView d, const double *x0, const double *x1, const double *x2, const double *g0,
                          const double *g1, const double *g2, double *f, double *gp, float *dF, float *dG,
                          int have_prev) {
    int N = d.g.N;
    int t = blockIdx.x * SD_BLOCK + threadIdx.x;
    if (t >= 3 * N) return;
    int b = t / N, k = t - b * N;
    const double *xb = b == 0 ? x0 : (b == 1 ? x1 : x2);
    const double *gb = b == 0 ? g0 : (b == 1 ? g1 : g2);
    double gk = gb[k];
    double fk = (b == 0 || aa_sig(d, k, b - 1)) ? gk - xb[k] : 0.0;
    if (have_prev) {
        dF[t] = (float)(fk - f[t]);
        dG[t] = (float)(gk - gp[t]);
    }
    f[t] = fk;
    gp[t] = gk;
}
void ph_aa_diff(hipStream_t s, const DevView &d, const double *x0, const double *x1, const double *x2,
                const double *g0, const double *g1, const double *g2, double *f, double *gp, float *dF, float *dG,
                int have_prev) {
    k_aa_diff<<<sd_nblocks(3L * d.g.N), SD_BLOCK, 0, s>>>(d, x0, x1, x2, g0, g1, g2, f, gp, dF, dG, have_prev);
}
struct VecList { float *v[12]; };
// dots of up to 12 float vectors with one vector (float y or double yd)
__global__ void __launch_bounds__(SD_BLOCK) k_aa_dots(int L, VecList vl, int nv, const float *y, const double *yd,
                                                      RedBuf rb, int slot0) {
    double acc[12];
#pragma unroll
    for (int q = 0; q < 12; q++) acc[q] = 0.0;
    for (int k = blockIdx.x * SD_BLOCK + threadIdx.x; k < L; k += gridDim.x * SD_BLOCK) {
        double yk = y ? (double)y[k] : yd[k];
#pragma unroll
        for (int q = 0; q < 12; q++)
            if (q < nv) acc[q] += (double)vl.v[q][k] * yk;
    }
    red_partial<12, false>(acc, rb);
}
void ph_aa_dots(hipStream_t s, int L, float *const *vecs, int nv, const float *y, const double *yd, RedBuf rb,
                int slot0) {
    VecList vl;
    for (int q = 0; q < 12; q++) vl.v[q] = q < nv ? vecs[q] : nullptr;
    int nb = red_grid(L);
    k_aa_dots<<<nb, SD_BLOCK, 0, s>>>(L, vl, nv, y, yd, rb, slot0);
    red_final(s, rb, nv, nb, slot0, 0);
}
struct GamList { float *v[12]; double gam[12]; };
__global__ void k_aa_mix(DevView d, double *g0, double *g1, double *g2, GamList gl, int m) {
    int N = d.g.N;
    int t = blockIdx.x * SD_BLOCK + threadIdx.x;
    if (t >= 3 * N) return;
    int b = t / N, k = t - b * N;
    if (b == 0 ? d.con_mask[k] != 0 : d.con_mask[k] != 0) return;     // psi and carriers fixed on contacts
    if (b > 0 && !aa_sig(d, k, b - 1)) return;
    double *gb = b == 0 ? g0 : (b == 1 ? g1 : g2);
    double s = gb[k];
    for (int i = 0; i < m; i++) s -= gl.gam[i] * (double)gl.v[i][t];
    gb[k] = s;
}
void ph_aa_mix(hipStream_t s, const DevView &d, double *g0, double *g1, double *g2, float *const *dG, const double *gam,
               int m) {
    GamList gl;
    for (int q = 0; q < 12; q++) { gl.v[q] = q < m ? dG[q] : nullptr; gl.gam[q] = q < m ? gam[q] : 0.0; }
    k_aa_mix<<<sd_nblocks(3L * d.g.N), SD_BLOCK, 0, s>>>(d, g0, g1, g2, gl, m);
}

// ------------------------------------------------------ display fields
__global__ void k_upd_J(DevView d, double *Jx, double *Jy, double *Jz, double *Rt) {
    const Geo g = d.g;
    int k = blockIdx.x * SD_BLOCK + threadIdx.x;
    if (k >= g.N) return;
    if (d.kind[k] != K_SEMI) { Jx[k] = Jy[k] = Jz[k] = 0.; Rt[k] = 0.; return; }
    int i, j, l;
    sd_unidx(g, k, i, j, l);
    double *Jo[3] = {Jx, Jy, Jz};
    for (int ax = 0; ax < 3; ax++) {
        double s = 0.; int cnt = 0, kb; double h, ar, Jn, Jp, Sn, Sp;
        if (sd_nbr(g, i, j, l, k, 2 * ax, kb, h, ar) && d.kind[kb] == K_SEMI) {
            face_flux(d, kb, k, h, ar, Jn, Jp, Sn, Sp); s += (Jn + Jp) / ar; cnt++;
        }
        if (sd_nbr(g, i, j, l, k, 2 * ax + 1, kb, h, ar) && d.kind[kb] == K_SEMI) {
            face_flux(d, k, kb, h, ar, Jn, Jp, Sn, Sp); s += (Jn + Jp) / ar; cnt++;
        }
        Jo[ax][k] = cnt ? s / cnt : 0.;
    }
    const MatDev &m = d.mats[d.mat[k]];
    double R, a, b;
    sd_recomb(m, d.niK[k], sd_tauf(d, k, m), d.n[k], d.p[k], R, a, b);
    Rt[k] = R;
}
void ph_upd_J(hipStream_t s, const DevView &d, double *Jx, double *Jy, double *Jz, double *R) {
    k_upd_J<<<sd_nblocks(d.g.N), SD_BLOCK, 0, s>>>(d, Jx, Jy, Jz, R);
}
__global__ void k_mob_display(DevView d, double *mun, double *mup) {
    const Geo g = d.g;
    int k = blockIdx.x * SD_BLOCK + threadIdx.x;
    if (k >= g.N) return;
    if (d.kind[k] != K_SEMI) { mun[k] = d.mu_n[k]; mup[k] = d.mu_p[k]; return; }
    int i, j, l;
    sd_unidx(g, k, i, j, l);
    size_t N = g.N;
    double sn = 0., sp = 0.; int c = 0;
    for (int dir = 0; dir < 6; dir++) {
        int kb; double h, ar;
        if (!sd_nbr(g, i, j, l, k, dir, kb, h, ar) || d.kind[kb] != K_SEMI) continue;
        int kl = k < kb ? k : kb, ax = dir / 2;
        double fn = d.fmu[ax * N + kl], fp = d.fmu[(3 + ax) * N + kl];
        if (ax == d.caxis) {
            const MatDev &m = d.mats[d.mat[kl]];
            if (m.ani_n > 0.) fn /= m.ani_n;
            if (m.ani_p > 0.) fp /= m.ani_p;
        }
        sn += fn; sp += fp; c++;
    }
    if (c) { mun[k] = sn / c; mup[k] = sp / c; }
    else { mun[k] = d.mu_n[k]; mup[k] = d.mu_p[k]; }
}
void ph_mob_display(hipStream_t s, const DevView &d, double *mun, double *mup) {
    k_mob_display<<<sd_nblocks(d.g.N), SD_BLOCK, 0, s>>>(d, mun, mup);
}
__global__ void k_node_Jc(DevView d, int ax, int hole, double *out) {
    const Geo g = d.g;
    int k = blockIdx.x * SD_BLOCK + threadIdx.x;
    if (k >= g.N) return;
    if (d.kind[k] != K_SEMI) { out[k] = 0.; return; }
    int i, j, l;
    sd_unidx(g, k, i, j, l);
    double s = 0.; int cnt = 0, kb; double h, ar, Jn, Jp, Sn, Sp;
    if (sd_nbr(g, i, j, l, k, 2 * ax, kb, h, ar) && d.kind[kb] == K_SEMI) {
        face_flux(d, kb, k, h, ar, Jn, Jp, Sn, Sp); s += (hole ? Jp : Jn) / ar; cnt++;
    }
    if (sd_nbr(g, i, j, l, k, 2 * ax + 1, kb, h, ar) && d.kind[kb] == K_SEMI) {
        face_flux(d, k, kb, h, ar, Jn, Jp, Sn, Sp); s += (hole ? Jp : Jn) / ar; cnt++;
    }
    out[k] = cnt ? s / cnt : 0.;
}
void ph_node_Jc(hipStream_t s, const DevView &d, int ax, int hole, double *out) {
    k_node_Jc<<<sd_nblocks(d.g.N), SD_BLOCK, 0, s>>>(d, ax, hole, out);
}
__global__ void k_face_J(DevView d, int ax, int carrier, double *out) {
    const Geo g = d.g;
    int k = blockIdx.x * SD_BLOCK + threadIdx.x;
    if (k >= g.N) return;
    int i, j, l, kb;
    double h, ar, Jn, Jp, Sn, Sp;
    sd_unidx(g, k, i, j, l);
    out[k] = 0.;
    if (d.kind[k] != K_SEMI || !sd_nbr(g, i, j, l, k, 2 * ax + 1, kb, h, ar) || d.kind[kb] != K_SEMI) return;
    face_flux(d, k, kb, h, ar, Jn, Jp, Sn, Sp);
    out[k] = (carrier ? Jn : Jn + Jp) / ar;
}
void ph_face_J(hipStream_t s, const DevView &d, int ax, int carrier, double *out) {
    k_face_J<<<sd_nblocks(d.g.N), SD_BLOCK, 0, s>>>(d, ax, carrier, out);
}

// ------------------------------------------------ impact ionisation (bit 8)
__device__ __forceinline__ double ii_alpha(const MatDev &m, double F, int c) {
    if (!(F > 1e5)) return 0.;
    int r = F >= m.ii_E0[c];
    double x = m.ii_g * m.ii_b[c][r] / F;
    if (m.ii_m[c] != 1.) x = pow(x, m.ii_m[c]);
    if (x > 700.) return 0.;
    return m.ii_g * m.ii_a[c][r] * exp(-x);
}
// G = (alpha_n |J_n| + alpha_p |J_p|)/q * vol, semiconductor_3d.c:4486-4523
__global__ void k_ii_update(DevView d) {
    const Geo g = d.g;
    int k = blockIdx.x * SD_BLOCK + threadIdx.x;
    if (k >= g.N) return;
    d.gii[k] = 0.;
    if (d.kind[k] != K_SEMI || d.con_mask[k]) return;
    const MatDev &m = d.mats[d.mat[k]];
    if (!m.ii_on) return;
    int i, j, l;
    sd_unidx(g, k, i, j, l);
    double Jn[3], Jp[3], E[3];
    for (int ax = 0; ax < 3; ax++) {
        double sn = 0., sp = 0., h, ar, a, b, c2, e; int cnt = 0, kb;
        if (sd_nbr(g, i, j, l, k, 2 * ax, kb, h, ar) && d.kind[kb] == K_SEMI) {
            face_flux(d, kb, k, h, ar, a, b, c2, e); sn += a / ar; sp += b / ar; cnt++;
        }
        if (sd_nbr(g, i, j, l, k, 2 * ax + 1, kb, h, ar) && d.kind[kb] == K_SEMI) {
            face_flux(d, k, kb, h, ar, a, b, c2, e); sn += a / ar; sp += b / ar; cnt++;
        }
        Jn[ax] = cnt ? sn / cnt : 0.;
        Jp[ax] = cnt ? sp / cnt : 0.;
        E[ax] = -grad_semi(d, i, j, l, k, ax);
    }
    double vol = g.Lx[i] * g.Ly[j] * g.Lz[l], G = 0.;
    for (int c = 0; c < 2; c++) {
        const double *J = c ? Jp : Jn;
        double jj = sqrt(J[0] * J[0] + J[1] * J[1] + J[2] * J[2]);
        if (!(jj > 0.)) continue;
        double F = fabs(E[0] * J[0] + E[1] * J[1] + E[2] * J[2]) / jj;
        G += ii_alpha(m, F, c) * jj;
    }
    d.gii[k] = G / SD_Q * vol;
}
void ph_ii_update(hipStream_t s, const DevView &d) { k_ii_update<<<sd_nblocks(d.g.N), SD_BLOCK, 0, s>>>(d); }
