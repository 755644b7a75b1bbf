// engine.cpp - solver drivers of the MI355X drift-diffusion core.
//
// Control flow follows the reference (semiconductor_3d.c): prepare :1638-1775,
// pcg :2182-2210, bicgstab :2216-2320, poisson_newton :2791-2823, solve_cont
// :3580-3594, float_fix :3159-3228, terminal_currents :3598-3623, aa_step
// :3709-3757, gummel :3781-3930, region_shift :3956-3987, equilibrium
// :3989-4004, cont_path :4018-4129, run_continuation :4135-4173.  All vector
// work runs on the device (k_physics.hip, k_linalg.hip); the host keeps the
// scalar decisions, reading a handful of reduction results per iteration.
#include "engine.h"
#include "sd_csr.h"
#include <type_traits>
#include "sd_nn.h"
#include <chrono>
#include <cstdlib>
#include <algorithm>
#include <queue>
#include <unordered_map>

Engine G;
volatile int g_stop = 0;
volatile double g_prog[8];

static int g_err = 0;
#define HC(x)                                                                                   \
    do {                                                                                        \
        hipError_t e_ = (x);                                                                    \
        if (e_ != hipSuccess) {                                                                 \
            fprintf(stderr, "semisim-hip: %s failed: %s (%s:%d)\n", #x, hipGetErrorString(e_),  \
                    __FILE__, __LINE__);                                                        \
            g_err = 1;                                                                          \
        }                                                                                       \
    } while (0)

int eng_ok() { return !g_err; }

static double wtime() {
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

// ------------------------------------------------------------ memory
template <class T> static T *dalloc(size_t n) {
    void *p = nullptr;
    if (n == 0) n = 1;
    if (hipMalloc(&p, n * sizeof(T)) != hipSuccess) { g_err = 1; return nullptr; }
    G.dev_allocs.push_back(p);
    return (T *)p;
}

// buffers whose size prepare derives from the device (contact edges,
// tunnelling paths): released by the next prepare, so re-preparing the same
// grid (new models, gate stacks or contacts) does not grow device memory
template <class T> static T *palloc(size_t n) {
    void *p = nullptr;
    if (n == 0) n = 1;
    if (hipMalloc(&p, n * sizeof(T)) != hipSuccess) { g_err = 1; return nullptr; }
    G.prep_allocs.push_back(p);
    return (T *)p;
}
void *eng_dalloc_bytes(size_t bytes) { return dalloc<unsigned char>(bytes); }
void *eng_palloc_bytes(size_t bytes) { return palloc<unsigned char>(bytes); }
void eng_set_error() { g_err = 1; }
static void prep_free() {
    if (G.st) hipStreamSynchronize(G.st);
    for (void *p : G.prep_allocs) hipFree(p);
    G.prep_allocs.clear();
    G.tun = TunDev{};
    G.d_tp_off = nullptr;
    G.d_tp_idx = nullptr;
    G.d_ek = G.d_ekb = G.d_eoff = nullptr;
    G.d_eh = G.d_ear = nullptr;
    G.d_gck = G.d_gce = G.d_gcoff = nullptr;
}

void eng_sync() { HC(hipStreamSynchronize(G.st)); }
void eng_download(double *dst, const double *src, size_t n) {
    HC(hipMemcpyAsync(dst, src, n * 8, hipMemcpyDeviceToHost, G.st));
    eng_sync();
}
void eng_upload(double *dst, const double *src, size_t n) {
    HC(hipMemcpyAsync(dst, src, n * 8, hipMemcpyHostToDevice, G.st));
    eng_sync();
}
// read scalar slots [a, b) into the pinned mirror (one copy, one sync)
static void rd(int a, int b) {
    HC(hipMemcpyAsync(G.h_scal + a, G.rb.scal + a, (size_t)(b - a) * 8, hipMemcpyDeviceToHost, G.st));
    eng_sync();
}
static double S(int slot) { return G.h_scal[slot]; }

// Captured preconditioner launch sequences (see precond): they bake in the
// hierarchy's pointers, shapes, smoother choices and the coarse-correction
// scale of the operator kind, so every change of those drops them.

template <class T> static void hier_free(Hier<T> &H) {
    H.lev.clear(); H.inv = nullptr; H.alloc = false;
    H.A0 = Sten7<T>{}; H.rs0 = nullptr; H.b0 = H.x0 = H.x0b = nullptr; H.c16 = nullptr; H.r16 = nullptr;
    H.r16b = nullptr;
}
// Storage order of each level: the line-smoothed levels above the coarse tail
// are stored parity-split along x (MGLevel::px) when their lines fit the scan
// kernel, so that a zebra sweep across x reads and writes its colour's lines
// contiguously.  The coarse levels (red-black sweeps, tail kernel, coarsest
// solve) keep the natural order.
template <class T> static void hier_layout(Hier<T> &H) {
    const int nl = (int)H.lev.size();
    for (int L = 0; L < nl; L++) {
        MGLevel<T> &v = H.lev[L];
        int px = G.mg_perm && G.n_axes > 0 && L < G.line_levels && L < nl - 1 && (H.tail < 0 || L < H.tail);
        for (int a = 0; a < G.n_axes && px; a++) {
            int len = G.mg_axes[a] == 0 ? v.Nx : (G.mg_axes[a] == 1 ? v.Ny : v.Nz);
            if (len > 256) px = 0;
        }
        v.px = px;
    }
}

void eng_free() {
    if (G.st) hipStreamSynchronize(G.st);
    prep_free();
    for (void *p : G.dev_allocs) hipFree(p);
    G.dev_allocs.clear();
    if (G.d_mrec) hipFree(G.d_mrec);
    G.d_mrec = nullptr;
    G.mrec_cap = 0;
    hier_free(G.H64);
    hier_free(G.H32);
    G.d_b32 = G.d_x32 = G.d_x32b = nullptr;
    G.d_k32 = nullptr;
    G.d_ws[0] = G.d_ws[1] = G.d_ws[2] = nullptr;
    G.d_e3[0] = G.d_e3[1] = G.d_e3[2] = nullptr;
    G.e3_ok = 0;
    G.ws_have = 0;
    G.allocated = 0;
    G.have_solution = 0;
    G.graph = 0;
    G.nE = G.nUE = 0;
    G.grp.clear(); G.gcol.clear(); G.grev.clear(); G.gue.clear();
    G.gh.clear(); G.ga.clear(); G.gvol.clear(); G.gpos.clear(); G.gwall.clear();
    G.gd = GraphDev{};
    G.d_gval = G.d_gmu = nullptr;
    gr_free();
}

// Multigrid hierarchy shape: 2x2x2 aggregation of every axis longer than 2
// until N <= 64 (semiconductor_3d.c:1894-1921).
// the coarse end (levels of at most G.tail_n nodes, not level 0) runs as one
// single-workgroup kernel; its level descriptors are copied to the device
template <class T> static void hier_tail(Hier<T> &H) {
    H.tail = -1;
    int nl = (int)H.lev.size();
    for (int L = 1; L < nl; L++) if (H.lev[L].N <= G.tail_n) { H.tail = L; break; }
    if (H.tail == nl - 1) H.tail = -1;            // a lone coarsest level: the dense solve kernel
    if (H.tail > 0 && H.d_tail)
        HC(hipMemcpy(H.d_tail, &H.lev[H.tail], sizeof(MGLevel<T>) * (nl - H.tail), hipMemcpyHostToDevice));
}
template <class T> static int hier_alloc(Hier<T> &H, bool level0_own) {
    H.lev.clear();
    MGLevel<T> f{};
    f.Nx = G.Nx; f.Ny = G.Ny; f.Nz = G.Nz; f.N = G.N;
    H.lev.push_back(f);
    while ((int)H.lev.size() < 24) {
        MGLevel<T> &a = H.lev.back();
        if (a.N <= 64) break;
        int cx = a.Nx > 2, cy = a.Ny > 2, cz = a.Nz > 2;
        if (!cx && !cy && !cz) break;
        a.cx = cx; a.cy = cy; a.cz = cz;
        MGLevel<T> c{};
        c.Nx = cx ? (a.Nx + 1) / 2 : a.Nx;
        c.Ny = cy ? (a.Ny + 1) / 2 : a.Ny;
        c.Nz = cz ? (a.Nz + 1) / 2 : a.Nz;
        c.N = c.Nx * c.Ny * c.Nz;
        H.lev.push_back(c);
    }
    for (size_t L = 0; L < H.lev.size(); L++) {
        MGLevel<T> &v = H.lev[L];
        size_t n = v.N;
        if (L == 0) {
            if (level0_own) {
                v.A.c = dalloc<T>(n); v.A.xm = dalloc<T>(n); v.A.xp = dalloc<T>(n); v.A.ym = dalloc<T>(n);
                v.A.yp = dalloc<T>(n); v.A.zm = dalloc<T>(n); v.A.zp = dalloc<T>(n);
            }
        } else {
            v.A.c = nullptr;
            v.A.xm = dalloc<T>(n); v.A.xp = dalloc<T>(n); v.A.ym = dalloc<T>(n);
            v.A.yp = dalloc<T>(n); v.A.zm = dalloc<T>(n); v.A.zp = dalloc<T>(n);
            v.rs = dalloc<double>(n);
            v.ph = dalloc<double>(n);
            v.x = dalloc<T>(n);
            v.xs = dalloc<T>(n);
            v.b = dalloc<T>(n);
        }
        v.pw = dalloc<T>(n);
        for (int a = 0; a < 3; a++)
            for (int q = 0; q < 3; q++) v.lu[a][q] = dalloc<T>(n);
    }
    {
        const size_t n = H.lev[0].N;
        if (level0_own) H.A0 = H.lev[0].A;
        else {
            H.A0.c = dalloc<T>(n); H.A0.xm = dalloc<T>(n); H.A0.xp = dalloc<T>(n); H.A0.ym = dalloc<T>(n);
            H.A0.yp = dalloc<T>(n); H.A0.zm = dalloc<T>(n); H.A0.zp = dalloc<T>(n);
            H.b0 = dalloc<T>(n); H.x0 = dalloc<T>(n); H.x0b = dalloc<T>(n);
        }
        H.rs0 = dalloc<double>(n);
    }
    int nz = H.lev.back().N;
    H.inv = dalloc<double>((size_t)nz * nz);
    H.d_tail = dalloc<MGLevel<T>>(H.lev.size());
    hier_tail(H);
    H.alloc = !g_err;
    H.pw_one = false;
    return H.alloc;
}

int eng_alloc(int Nx, int Ny, int Nz, int graph) {
    eng_free();
    g_err = 0;
    G.graph = graph;
    if (!G.st) {
        // The solvers read a few device scalars back after most reductions,
        // so the host's wait latency is on the critical path: wait by
        // spinning unless SEMISIM_SYNC asks for yield, blocking or auto.
        // Applies only before the process's HIP context exists.
        const char *e = getenv("SEMISIM_SYNC");
        unsigned fl = hipDeviceScheduleSpin;
        if (e && !strcmp(e, "yield")) fl = hipDeviceScheduleYield;
        if (e && !strcmp(e, "blocking")) fl = hipDeviceScheduleBlockingSync;
        if (e && !strcmp(e, "auto")) fl = hipDeviceScheduleAuto;
        (void)hipSetDeviceFlags(fl);
        (void)hipGetLastError();
        if (hipStreamCreateWithFlags(&G.st, hipStreamNonBlocking) != hipSuccess) return 0;
    }
    size_t N = (size_t)Nx * Ny * Nz;
    if (N > 2000000000ULL) return 0;
    G.Nx = Nx; G.Ny = Ny; G.Nz = Nz; G.N = (int)N;
    G.d_geo = dalloc<double>(2 * (Nx + Ny + Nz));
    G.d_kind = dalloc<unsigned char>(N); G.d_mat = dalloc<unsigned char>(N); G.d_flag = dalloc<unsigned char>(N);
    G.d_con_mask = dalloc<int>(N); G.d_gate_mask = dalloc<int>(N); G.d_changed = dalloc<int>(4);
    G.d_mlist = dalloc<int>(N);
    double **nd[] = {&G.d_Nd, &G.d_Na, &G.d_Nf, &G.d_eps, &G.d_bgn, &G.d_robA, &G.d_lncK, &G.d_lnvK, &G.d_niK,
                     &G.d_psi, &G.d_phin, &G.d_phip, &G.d_n, &G.d_p, &G.d_mun, &G.d_mup, &G.d_gii,
                     &G.A.c, &G.A.xm, &G.A.xp, &G.A.ym, &G.A.yp, &G.A.zm, &G.A.zp,
                     &G.d_rhs, &G.d_sc, &G.d_dsc, &G.d_sol, &G.d_xsave, &G.d_tr, &G.d_Ft,
                     &G.d_xk, &G.d_dqn, &G.d_dqp, &G.d_qn0, &G.d_qp0,
                     &G.d_Jx, &G.d_Jy, &G.d_Jz, &G.d_Rt, &G.d_mudn, &G.d_mudp, &G.d_qd,
                     &G.d_tk, &G.d_tg, &G.d_qcn, &G.d_qcp};
    for (double **p : nd) *p = dalloc<double>(N);
    for (int q = 0; q < 7; q++) G.d_wk[q] = dalloc<double>(N);
    for (int a = 0; a < 3; a++) { G.d_h1[a] = dalloc<double>(N); G.d_h2[a] = dalloc<double>(N); G.d_e2[a] = dalloc<double>(N); }
    G.d_fmu = dalloc<double>(6 * N);
    G.d_aaf = dalloc<double>(3 * N);
    G.d_aag = dalloc<double>(3 * N);
    for (int q = 0; q < G.aa_m; q++) { G.d_aadF[q] = dalloc<float>(3 * N); G.d_aadG[q] = dalloc<float>(3 * N); }
    G.d_mats = dalloc<MatDev>(SD_NM_MAX);
    G.d_cons = dalloc<ConDev>(SD_NC_MAX);
    G.rb.part = dalloc<double>((size_t)SD_RED_MAXV * SD_RED_MAXB);
    G.rb.cnt = dalloc<unsigned>(4);
    G.rb.scal = dalloc<double>(S_NSLOT);
    if (!G.h_scal) {
        if (hipHostMalloc((void **)&G.h_scal, S_NSLOT * 8, hipHostMallocDefault) != hipSuccess) g_err = 1;
    }
    if (g_err) { eng_free(); return 0; }
    HC(hipMemsetAsync(G.rb.cnt, 0, 4 * sizeof(unsigned), G.st));
    HC(hipMemsetAsync(G.rb.scal, 0, S_NSLOT * 8, G.st));
    HC(hipMemsetAsync(G.d_gii, 0, N * 8, G.st));
    for (double *p : {G.d_psi, G.d_phin, G.d_phip, G.d_n, G.d_p, G.d_mun, G.d_mup, G.d_Jx, G.d_Jy, G.d_Jz, G.d_Rt,
                      G.d_mudn, G.d_mudp})
        HC(hipMemsetAsync(p, 0, N * 8, G.st));
    HC(hipMemsetAsync(G.d_fmu, 0, 6 * N * 8, G.st));
    // the tensor multigrid (a graph device has its own, built by eng_set_graph)
    if (!graph) {
        if (!hier_alloc(G.H64, false)) { eng_free(); return 0; }
        G.H64.lev[0].A = G.A;
    }
    eng_sync();
    if (g_err) { eng_free(); return 0; }
    G.allocated = 1;
    G.have_solution = 0;
    G.dirty = 1;
    return 1;
}

static int ensure_h32() {
    if (G.H32.alloc) return 1;
    if (!hier_alloc(G.H32, true)) return 0;
    G.d_b32 = dalloc<float>(G.N);
    G.d_x32 = dalloc<float>(G.N);
    G.d_x32b = dalloc<float>(G.N);
    G.d_k32 = dalloc<float>((size_t)8 * G.N);
    return !g_err;
}

DevView eng_view() {
    DevView d;
    d.g.Nx = G.Nx; d.g.Ny = G.Ny; d.g.Nz = G.Nz; d.g.N = G.N;
    d.g.xs = G.d_geo; d.g.ys = G.d_geo + G.Nx; d.g.zs = G.d_geo + G.Nx + G.Ny;
    int o = G.Nx + G.Ny + G.Nz;
    d.g.Lx = G.d_geo + o; d.g.Ly = G.d_geo + o + G.Nx; d.g.Lz = G.d_geo + o + G.Nx + G.Ny;
    d.kind = G.d_kind; d.con_mask = G.d_con_mask; d.gate_mask = G.d_gate_mask; d.mat = G.d_mat;
    d.Nd = G.d_Nd; d.Na = G.d_Na; d.Nf = G.d_Nf; d.eps = G.d_eps; d.bgn = G.d_bgn; d.robA = G.d_robA;
    d.lncK = G.d_lncK; d.lnvK = G.d_lnvK; d.niK = G.d_niK;
    d.psi = G.d_psi; d.phin = G.d_phin; d.phip = G.d_phip; d.n = G.d_n; d.p = G.d_p;
    d.mu_n = G.d_mun; d.mu_p = G.d_mup; d.fmu = G.d_fmu;
    d.gii = (G.models & 256) ? G.d_gii : nullptr;
    d.qd = (G.models & 16) ? G.d_qd : nullptr;
    d.tk = G.tun.np ? G.d_tk : nullptr;
    d.tg = G.tun.np ? G.d_tg : nullptr;
    d.mats = G.d_mats; d.cons = G.d_cons;
    d.models = G.models; d.caxis = G.caxis; d.ref_mat = G.ref_mat; d.Cref = G.Cref;
    return d;
}

// ------------------------------------------------------------ prepare
static void mat_prepare() {
    G.ref_mat = 0;
    for (int i = 0; i < G.n_mats; i++) if (!G.mats[i].insulator) { G.ref_mat = i; break; }
    const HostMat &r = G.mats[G.ref_mat];
    G.Cref = r.insulator ? 4.6 : r.chi + 0.5 * r.Eg + 0.5 * r.VT * log(r.Nc / r.Nv);
    for (int i = 0; i < G.n_mats; i++) {
        HostMat &m = G.mats[i];
        if (m.insulator) { m.lnc = m.lnv = -1e300; continue; }
        m.lnc = log(m.Nc) + (m.chi - G.Cref) / m.VT;
        m.lnv = log(m.Nv) + (G.Cref - m.chi - m.Eg) / m.VT;
    }
}

// Tunnelling fluxes through a stack (api3_tun_eval :4855-4867), evaluated on
// the device by the kernels of the solve; F: e s->g, e g->s, h s->g, h g->s
// (1/m^2 s).  Returns the net current density into the semiconductor (A/m^2).
double eng_tun_eval(int smat, int nl, const int *mat, const double *t, double psi, double phin, double phip,
                    double Vg, double pm, double *F) {
    mat_prepare();
    MatDev md[SD_NM_MAX] = {};
    for (int i = 0; i < G.n_mats; i++) md[i] = host_mat_dev(G.mats[i]);
    MatDev *d_md = nullptr;
    double *d_F = nullptr, h[4] = {NAN, NAN, NAN, NAN};
    hipStream_t s = G.st;
    if (hipMalloc(&d_md, sizeof md) == hipSuccess && hipMalloc(&d_F, 4 * sizeof(double)) == hipSuccess &&
        hipMemcpyAsync(d_md, md, sizeof md, hipMemcpyHostToDevice, s) == hipSuccess) {
        ph_tun_eval(s, d_md, smat, nl, mat, t, psi, phin, phip, Vg, pm, G.Cref, d_F);
        if (hipMemcpyAsync(h, d_F, sizeof h, hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess)
            for (double &v : h) v = NAN;
    }
    if (d_md) hipFree(d_md);
    if (d_F) hipFree(d_F);
    if (F) for (int i = 0; i < 4; i++) F[i] = h[i];
    return SD_Q * ((h[0] - h[1]) - (h[2] - h[3]));
}

// Lombardi edge mobility (m^2/Vs) of material mi by the device function of
// the mobility kernel (api3_lombardi :4894-4897)
double eng_lombardi(int mi, double mub, double Fv, double Nt, double Dw, int hole) {
    double *d_o = nullptr, h = NAN;
    hipStream_t s = G.st;
    if (hipMalloc(&d_o, sizeof(double)) == hipSuccess) {
        ph_lombardi_eval(s, host_mat_dev(G.mats[mi]), mub, Fv, Nt, Dw, hole, d_o);
        if (hipMemcpyAsync(&h, d_o, sizeof h, hipMemcpyDeviceToHost, s) != hipSuccess || hipStreamSynchronize(s) != hipSuccess)
            h = NAN;
        hipFree(d_o);
    }
    return h;
}
// MLDA density factor of material mi at the cell range [a, b] from a wall
// (api3_mlda_factor :4869-4872): the host function prepare uses
static double mlda_factor(const HostMat &m, double a, double b, int hole);
double eng_mlda_factor(int mi, double a, double b, int hole) { return mlda_factor(G.mats[mi], a, b, hole); }

// nodes of a face contact (face 0..5) or of a box (face 6), v6 index
// convention, clipped to the mesh (semiconductor_3d.c:1047-1065); a contact
// given as a node list (a shaped contact, api3_set_con_nodes) visits its
// listed nodes instead, with their listed areas (FOR_CON_ANY :1422-1426)
template <class F> static void for_con_nodes(const HostCon &cn, F body) {
    const int cid = (int)(&cn - G.cons);
    if (cid >= 0 && cid < SD_NC_MAX && !G.con_nodes[cid].empty()) {
        const std::vector<int> &nl = G.con_nodes[cid];
        for (size_t q = 0; q < nl.size(); q++) {
            const int k = nl[q];
            int i, j, l;
            host_unidx(k, i, j, l);
            body(i, j, l, k, G.con_areas[cid][q]);
        }
        return;
    }
    if (G.graph) return;                 // a contact of a graph device has no nodes but its list
    int ia = 0, ib = G.Nx - 1, ja = 0, jb = G.Ny - 1, la = 0, lb = G.Nz - 1;
    switch (cn.face) {
    case 0: ia = ib = 0; ja = cn.i0; jb = cn.i1; la = cn.j0; lb = cn.j1; break;
    case 1: ia = ib = G.Nx - 1; ja = cn.i0; jb = cn.i1; la = cn.j0; lb = cn.j1; break;
    case 2: ja = jb = 0; ia = cn.i0; ib = cn.i1; la = cn.j0; lb = cn.j1; break;
    case 3: ja = jb = G.Ny - 1; ia = cn.i0; ib = cn.i1; la = cn.j0; lb = cn.j1; break;
    case 4: la = lb = 0; ia = cn.i0; ib = cn.i1; ja = cn.j0; jb = cn.j1; break;
    case 5: la = lb = G.Nz - 1; ia = cn.i0; ib = cn.i1; ja = cn.j0; jb = cn.j1; break;
    default: ia = cn.i0; ib = cn.i1; ja = cn.j0; jb = cn.j1; la = cn.k0; lb = cn.k1;
    }
    if (ia < 0) ia = 0; if (ja < 0) ja = 0; if (la < 0) la = 0;
    if (ib > G.Nx - 1) ib = G.Nx - 1; if (jb > G.Ny - 1) jb = G.Ny - 1; if (lb > G.Nz - 1) lb = G.Nz - 1;
    for (int l = la; l <= lb; l++)
        for (int j = ja; j <= jb; j++)
            for (int i = ia; i <= ib; i++) {
                int k = i + G.Nx * (j + G.Ny * l);
                double area = cn.face <= 1 ? G.Ly[j] * G.Lz[l] : (cn.face <= 3 ? G.Lx[i] * G.Lz[l] : (cn.face <= 5 ? G.Lx[i] * G.Ly[j] : 0.));
                body(i, j, l, k, area);
            }
}

bool host_nbr(int i, int j, int l, int k, int dir, int &kb, double &h, double &ar) {
    switch (dir) {
    case 0: if (i == 0) return false; kb = k - 1; h = G.xs[i] - G.xs[i - 1]; ar = G.Ly[j] * G.Lz[l]; return true;
    case 1: if (i == G.Nx - 1) return false; kb = k + 1; h = G.xs[i + 1] - G.xs[i]; ar = G.Ly[j] * G.Lz[l]; return true;
    case 2: if (j == 0) return false; kb = k - G.Nx; h = G.ys[j] - G.ys[j - 1]; ar = G.Lx[i] * G.Lz[l]; return true;
    case 3: if (j == G.Ny - 1) return false; kb = k + G.Nx; h = G.ys[j + 1] - G.ys[j]; ar = G.Lx[i] * G.Lz[l]; return true;
    case 4: if (l == 0) return false; kb = k - G.Nx * G.Ny; h = G.zs[l] - G.zs[l - 1]; ar = G.Lx[i] * G.Ly[j]; return true;
    default: if (l == G.Nz - 1) return false; kb = k + G.Nx * G.Ny; h = G.zs[l + 1] - G.zs[l]; ar = G.Lx[i] * G.Ly[j]; return true;
    }
}
void host_unidx(int k, int &i, int &j, int &l) {
    int NxNy = G.Nx * G.Ny;
    l = k / NxNy; int r = k - l * NxNy; j = r / G.Nx; i = r - j * G.Nx;
}

// Distance of every semiconductor node to the nearest insulator interface or
// Robin gate face along the mesh lines (1 m where there is none), as the
// reference's wall scan computes it in prepare (walls_tensor,
// semiconductor_3d.c:1100-1135; the MLDA factors it also forms belong to
// out-of-scope physics).  Used by the Lombardi surface mobility (prepare) and
// by the readback.
// MLDA: fraction of the bulk density kept in the cell [a, b] from a wall
// (Gaussian population average, thermal length lam), mlda_pop
// semiconductor_3d.c:1073-1081; series for small cells, no cancellation
static double mlda_pop(double a, double b, double lam) {
    const double u = a / lam, v = b / lam;
    if (v < 0.05) {
        const double d = v - u;
        if (d < 1e-12) return u * u;
        return ((v * v * v - u * u * u) / 3. - (pow(v, 5) - pow(u, 5)) / 10. + (pow(v, 7) - pow(u, 7)) / 42.) / d;
    }
    if (v - u < 1e-9 * v) return 1. - exp(-u * u);
    return 1. - 0.5 * sqrt(M_PI) * (erf(v) - erf(u)) / (v - u);
}
// cell-averaged MLDA factor of carrier `hole` for a cell [a, b] from the
// wall, two valley masses weighted, mlda_factor :1082-1089
static double mlda_factor(const HostMat &m, double a, double b, int hole) {
    const double kT = SD_KB * m.T, m1 = hole ? m.mq_h1 : m.mq_e1, m2 = hole ? m.mq_h2 : m.mq_e2;
    const double w = hole ? m.wq_h1 : m.wq_e1;
    if (!(m1 > 0.)) return 1.;
    if (a < 0.) a = 0.;
    if (b < a) b = a;
    const double hb = 1.054571817e-34, m0 = 9.1093837015e-31;
    double f = w * mlda_pop(a, b, hb / sqrt(2. * m1 * m0 * kT));
    if (w < 1. && m2 > 0.) f += (1. - w) * mlda_pop(a, b, hb / sqrt(2. * m2 * m0 * kT));
    return f;
}
// With qcn/qcp given (MLDA on): also the sum over the walls within 40 nm of
// ln(cell-averaged factor) per carrier, clamped at -60.
void eng_walls(std::vector<double> &qd, std::vector<double> *qcn, std::vector<double> *qcp) {
    const int Nx = G.Nx, Ny = G.Ny, Nz = G.Nz;
    qd.assign((size_t)G.N, 1.0);
    if (qcn) { qcn->assign((size_t)G.N, 0.); qcp->assign((size_t)G.N, 0.); }
    for (int ax = 0; ax < 3; ax++) {
        const int n = ax == 0 ? Nx : (ax == 1 ? Ny : Nz);
        const std::vector<double> &c = ax == 0 ? G.xs : (ax == 1 ? G.ys : G.zs);
        const size_t st = ax == 0 ? 1 : (ax == 1 ? (size_t)Nx : (size_t)Nx * Ny);
        const int n1 = ax == 0 ? Ny : Nx, n2 = ax == 2 ? Ny : Nz;
        for (int b = 0; b < n2; b++)
            for (int a = 0; a < n1; a++) {
                const size_t k0 = ax == 0 ? (size_t)Nx * (a + (size_t)Ny * b)
                                          : (ax == 1 ? a + (size_t)Nx * Ny * b : a + (size_t)Nx * b);
                for (int side = 0; side < 2; side++) {      // wall on the low (0) or high (1) side
                    double w = 0.;
                    bool have = false;
                    for (int t = 0; t < n; t++) {
                        const int q = side ? n - 1 - t : t;
                        const size_t k = k0 + (size_t)q * st;
                        if (G.kind[k] != K_SEMI) { have = false; continue; }
                        const int qp = side ? q + 1 : q - 1;
                        if (qp < 0 || qp >= n) {
                            const int g = G.gate_mask[k];
                            have = false;
                            if (g) {
                                const int f = G.cons[g - 1].face;
                                if (f / 2 == ax && f % 2 == side) { w = c[q]; have = true; }
                            }
                        } else {
                            const size_t kp = k0 + (size_t)qp * st;
                            if (G.kind[kp] == K_INS) { w = 0.5 * (c[q] + c[qp]); have = true; }
                            else if (G.kind[kp] != K_SEMI) have = false;
                        }
                        if (!have) continue;
                        const double lo = q > 0 ? 0.5 * (c[q - 1] + c[q]) : c[q];
                        const double hi = q < n - 1 ? 0.5 * (c[q] + c[q + 1]) : c[q];
                        double za = side ? w - hi : lo - w;
                        const double zb = side ? w - lo : hi - w, zn = fabs(c[q] - w);
                        if (za < 0.) za = 0.;
                        if (zn < qd[k]) qd[k] = zn;
                        if (qcn && za < 40e-9) {
                            const HostMat &m = G.mats[G.mat8[k]];
                            (*qcn)[k] += log(fmax(mlda_factor(m, za, zb, 0), 1e-30));
                            (*qcp)[k] += log(fmax(mlda_factor(m, za, zb, 1), 1e-30));
                        }
                    }
                }
            }
    }
    if (qcn)
        for (int k = 0; k < G.N; k++) {
            if ((*qcn)[k] < -60.) (*qcn)[k] = -60.;
            if ((*qcp)[k] < -60.) (*qcp)[k] = -60.;
        }
}
void eng_wall_distance(std::vector<double> &qd) { eng_walls(qd, nullptr, nullptr); }

// ------------------------------------------------------ gate tunnelling
// Tunnelling paths (tun_build :1213-1247, tun_push :1172-1210), in the
// reference's order: first the paths through meshed insulator layers (every
// semiconductor node and direction, ascending, whose straight march through
// insulator nodes, at most TUN_LMAX of it and TUN_MAXL material runs, ends on
// the metal node of a gate: a semiconductor half cell, then the insulator
// runs midpoint to midpoint, the last to the metal), then one path per node of
// a Robin gate with a layer stack.  Each has its supply column: the
// semiconductor nodes inward from the interface over 5 nm, stopping at
// contacts and non-semiconductors.
#define TUN_COL 5e-9
#define TUN_LMAX 30e-9
#define TUN_MAXL 6
static int tun_build() {
    G.tp_k.clear(); G.tp_c.clear(); G.tp_c0.clear(); G.tp_cn.clear(); G.col_k.clear();
    G.tp_A.clear(); G.col_l.clear(); G.col_h.clear();
    G.tp_nl.clear(); G.tp_m.clear(); G.tp_t.clear();
    for (int c = 0; c < G.n_cons; c++) G.cons[c].ntp = 0;
    if (!(G.models & 32)) return 1;
    // one path from node k (toward the metal along dir) with its layers, and
    // its supply column inward (dir ^ 1)
    auto push = [&](int k, double A, int c, int dir, int nl, const int *m, const double *t) {
        G.tp_k.push_back(k); G.tp_c.push_back(c); G.tp_A.push_back(A);
        G.tp_nl.push_back(nl);
        for (int i = 0; i < TUN_MAXL; i++) { G.tp_m.push_back(i < nl ? m[i] : 0); G.tp_t.push_back(i < nl ? t[i] : 0.); }
        G.tp_c0.push_back((int)G.col_k.size());
        const int in = dir ^ 1, ax = dir / 2;
        int cnt = 0, cur = k;
        double depth = 0.;
        for (;;) {
            int ii, jj, ll;
            host_unidx(cur, ii, jj, ll);
            double L = ax == 0 ? G.Lx[ii] : (ax == 1 ? G.Ly[jj] : G.Lz[ll]);
            int kn; double h, ar;
            bool more = host_nbr(ii, jj, ll, cur, in, kn, h, ar) && G.kind[kn] == K_SEMI && !G.con_mask[kn];
            G.col_k.push_back(cur); G.col_l.push_back(L); G.col_h.push_back(more ? h : 0.);
            cnt++; depth += L;
            if (!more || depth >= TUN_COL) break;
            cur = kn;
        }
        G.tp_cn.push_back(cnt);
        G.cons[c].ntp++;
    };
    for (int k = 0; k < G.N; k++) {
        if (G.kind[k] != K_SEMI || G.con_mask[k]) continue;
        int i, j, l;
        host_unidx(k, i, j, l);
        for (int dir = 0; dir < 6; dir++) {
            int kb; double h, ar;
            if (!host_nbr(i, j, l, k, dir, kb, h, ar) || G.kind[kb] != K_INS || G.con_mask[kb]) continue;
            int m[TUN_MAXL]; double t[TUN_MAXL];
            int nl = 1, ok = 0, cur = kb;
            m[0] = G.mat_idx[k]; t[0] = 0.5 * h;              // semiconductor half cell
            double tins = 0., hin = h;
            for (int guard = 0; guard < 100000; guard++) {
                int ci, cj, cl, kn; double h2, ar2;
                host_unidx(cur, ci, cj, cl);
                if (!host_nbr(ci, cj, cl, cur, dir, kn, h2, ar2)) break;                // domain edge
                const int mi = G.mat_idx[cur];
                const bool end = G.kind[kn] == K_METAL && G.con_mask[kn] && G.cons[G.con_mask[kn] - 1].bc == 2;
                const double seg = 0.5 * hin + (end ? h2 : 0.5 * h2);   // midpoint to midpoint (or to the metal)
                if (nl > 1 && m[nl - 1] == mi) t[nl - 1] += seg;
                else { if (nl >= TUN_MAXL) break; m[nl] = mi; t[nl] = seg; nl++; }
                tins += seg;
                if (tins > TUN_LMAX) break;
                if (end) { ok = G.con_mask[kn]; break; }
                if (G.kind[kn] != K_INS || G.con_mask[kn]) break;
                cur = kn; hin = h2;
            }
            if (ok) push(k, ar, ok - 1, dir, nl, m, t);
        }
    }
    for (int c = 0; c < G.n_cons; c++) {
        HostCon &cn = G.cons[c];
        if (cn.bc != 2 || cn.face > 5 || cn.snl < 1) continue;
        for_con_nodes(cn, [&](int, int, int, int k, double area) {
            if (G.gate_mask[k] != c + 1 || G.con_mask[k]) return;
            push(k, area, c, cn.face, cn.snl, cn.smat, cn.st);
        });
    }
    if (G.verbose && !G.tp_k.empty())
        fprintf(stderr, "semisim-hip: %zu tunnelling paths, %zu column nodes\n", G.tp_k.size(), G.col_k.size());
    return 1;
}
template <class T> static T *dput(const std::vector<T> &v) {
    T *p = palloc<T>(v.size());
    if (p && !v.empty()) HC(hipMemcpyAsync(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, G.st));
    return p;
}
// device copies of the paths (reallocated per prepare, like the current
// edges), the per-node gather lists, and the MLDA terms the supply reads
static void tun_upload() {
    TunDev &T = G.tun;
    T = TunDev{};
    const int np = (int)G.tp_k.size(), N = G.N;
    if (np == 0) return;
    static_assert(SD_TUN_MAXL == TUN_MAXL, "one layer limit");
    // the paths of contact c are pidx[poff[c]..poff[c+1]) (meshed paths come
    // in node order, so a contact's paths need not be adjacent)
    std::vector<int> poff(G.n_cons + 1, 0), pidx(np);
    for (int p = 0; p < np; p++) poff[G.tp_c[p] + 1]++;
    for (int c = 0; c < G.n_cons; c++) poff[c + 1] += poff[c];
    {
        std::vector<int> fill(poff.begin(), poff.end() - 1);
        for (int p = 0; p < np; p++) pidx[fill[G.tp_c[p]]++] = p;
    }
    // touched nodes (ascending) with their column entries and path starts,
    // both in ascending (= the reference's summation) order
    std::vector<int> tn(G.col_k);
    std::sort(tn.begin(), tn.end());
    tn.erase(std::unique(tn.begin(), tn.end()), tn.end());
    const int ntn = (int)tn.size();
    std::unordered_map<int, int> at;
    at.reserve((size_t)ntn * 2);
    for (int t = 0; t < ntn; t++) at[tn[t]] = t;
    std::vector<std::vector<int>> eK(ntn), eG(ntn);
    for (int q = 0; q < (int)G.col_k.size(); q++) eK[at[G.col_k[q]]].push_back(q);
    for (int p = 0; p < np; p++) eG[at[G.tp_k[p]]].push_back(p);
    std::vector<int> offK(ntn + 1, 0), offG(ntn + 1, 0), entK, entG;
    for (int t = 0; t < ntn; t++) {
        entK.insert(entK.end(), eK[t].begin(), eK[t].end()); offK[t + 1] = (int)entK.size();
        entG.insert(entG.end(), eG[t].begin(), eG[t].end()); offG[t + 1] = (int)entG.size();
    }
    T.np = np; T.ntn = ntn;
    T.pk = dput(G.tp_k); T.pc = dput(G.tp_c); T.pnl = dput(G.tp_nl); T.pc0 = dput(G.tp_c0); T.pcn = dput(G.tp_cn);
    T.pm = dput(G.tp_m); T.pA = dput(G.tp_A); T.pt = dput(G.tp_t);
    T.ck = dput(G.col_k); T.cl = dput(G.col_l); T.ch = dput(G.col_h);
    T.cK = palloc<double>(2 * G.col_k.size()); T.pG = palloc<double>(2 * (size_t)np);
    T.tn = dput(tn); T.offK = dput(offK); T.entK = dput(entK); T.offG = dput(offG); T.entG = dput(entG);
    G.d_tp_off = dput(poff);
    G.d_tp_idx = dput(pidx);
    if (G.models & 8) {
        HC(hipMemcpyAsync(G.d_qcn, G.qcn.data(), (size_t)N * 8, hipMemcpyHostToDevice, G.st));
        HC(hipMemcpyAsync(G.d_qcp, G.qcp.data(), (size_t)N * 8, hipMemcpyHostToDevice, G.st));
        T.qcn = G.d_qcn; T.qcp = G.d_qcp;
    }
    HC(hipMemsetAsync(T.cK, 0, 2 * G.col_k.size() * 8, G.st));
    HC(hipMemsetAsync(T.pG, 0, 2 * (size_t)np * 8, G.st));
    HC(hipMemsetAsync(G.d_tk, 0, (size_t)N * 8, G.st));
    HC(hipMemsetAsync(G.d_tg, 0, (size_t)N * 8, G.st));
    if (g_err) G.tun = TunDev{};
}

int eng_prepare() {
    if (!G.allocated || G.n_mats < 1) return 0;
    {   // physics outside this core is rejected where it would take effect, never approximated
        const char *why = nullptr;
        for (int i = 0; i < G.n_mats && !why; i++)
            if ((G.models & 16) && G.mats[i].id == 4) why = "Coulomb surface mobility of 4H-SiC (models bit 4)";
        for (int c = 0; c < G.n_cons && !why; c++)
            if ((G.models & (64 | 128)) && G.cons[c].bc == 1 && G.cons[c].face <= 5)
                why = "metal-contact emission on a Schottky face contact (models bits 6-7)";
        for (int i = 0; i < G.n_mats && !why; i++)
            if ((G.models & 512) && G.mats[i].id == 4) why = "incomplete ionisation (models bit 9)";
        if (G.graph && !why) {
            if (G.models & 32) why = "gate tunnelling on a graph mesh (models bit 5)";
            else if (G.models & 256) why = "impact ionisation on a graph mesh (models bit 8)";
        }
        if (why) { fprintf(stderr, "semisim-hip: %s is outside this core: not solved\n", why); return 0; }
    }
    if (!G.dirty) return 1;
    prep_free();                       // the previous prepare's edge and path buffers
    const int N = G.N;
    mat_prepare();
    G.kind.assign(N, 0); G.mat8.assign(N, 0); G.eps.assign(N, 0.); G.bgn.assign(N, 0.);
    G.con_mask.assign(N, 0); G.gate_mask.assign(N, 0); G.robA.assign(N, 0.);
    for (int k = 0; k < N; k++) {
        int mi = G.mat_idx[k];
        if (mi < 0 || mi >= G.n_mats) mi = G.mat_idx[k] = 0;
        const HostMat &m = G.mats[mi];
        G.mat8[k] = (unsigned char)mi;
        G.kind[k] = m.insulator ? K_INS : K_SEMI;
        G.eps[k] = m.eps_r * SD_EPS0;
        G.bgn[k] = (!m.insulator && (G.models & 1)) ? host_bgn_half(m, G.Nd[k] + G.Na[k]) : 0.;
    }
    // contacts; later contacts override earlier ones on shared nodes
    int nrob = 0;
    for (int c = 0; c < G.n_cons; c++) G.cons[c].te = 0;
    for (int c = 0; c < G.n_cons; c++) {
        HostCon &cn = G.cons[c];
        cn.A = 0.;
        for_con_nodes(cn, [&](int, int, int, int k, double area) {
            if (cn.face <= 5) cn.A += area;
            if (cn.bc == 2 && cn.face <= 5 && G.kind[k] == K_SEMI) { G.gate_mask[k] = c + 1; nrob++; }
            else if (cn.bc == 2 || G.kind[k] != K_SEMI) { G.kind[k] = K_METAL; G.con_mask[k] = c + 1; G.gate_mask[k] = 0; }
            else { G.con_mask[k] = c + 1; G.gate_mask[k] = 0; }
        });
    }
    G.n_rob = 0;
    if (nrob > 0) {
        for (int c = 0; c < G.n_cons; c++) {
            HostCon &cn = G.cons[c];
            if (cn.bc != 2 || cn.face > 5) continue;
            for_con_nodes(cn, [&](int, int, int, int k, double area) {
                if (G.gate_mask[k] == c + 1 && !G.con_mask[k] && G.n_rob < nrob) { G.robA[k] = area; G.n_rob++; }
            });
        }
    }
    for (int k = 0; k < N; k++) if (G.gate_mask[k] && (G.con_mask[k] || G.robA[k] == 0.)) G.gate_mask[k] = 0;
    if (!G.graph && !tun_build()) return 0;
    // box contacts: area = interface with non-contact semiconductor (graph:
    // gr_prepare_edges below)
    for (int c = 0; c < G.n_cons && !G.graph; c++) {
        HostCon &cn = G.cons[c];
        if (cn.face <= 5) continue;
        double Ab = 0.;
        for_con_nodes(cn, [&](int i, int j, int l, int k, double) {
            if (G.con_mask[k] != c + 1) return;
            for (int dir = 0; dir < 6; dir++) {
                int kb; double h, ar;
                if (host_nbr(i, j, l, k, dir, kb, h, ar) && G.kind[kb] == K_SEMI && G.con_mask[kb] != c + 1) Ab += ar;
            }
        });
        cn.A = Ab;
    }
    // node band data and low-field mobility (constant for the device); with
    // MLDA (models bit 3) the confinement factors qcn, qcp of the effective
    // densities of states near interfaces (lncK, lnvK, niK :736-738)
    if (G.graph) {
        gr_walls(G.qd, G.qcn, G.qcp, (G.models & 8) != 0, mlda_factor);
    } else if (G.models & 8) {
        std::vector<double> qd;
        eng_walls(qd, &G.qcn, &G.qcp);
    } else {
        G.qcn.assign(N, 0.); G.qcp.assign(N, 0.);
    }
    G.lncK.assign(N, -1e300); G.lnvK.assign(N, -1e300); G.niK.assign(N, 0.); G.mu0n.assign(N, 0.); G.mu0p.assign(N, 0.);
    for (int k = 0; k < N; k++) {
        const HostMat &m = G.mats[G.mat8[k]];
        if (m.insulator) { G.lncK[k] = G.lnvK[k] = m.lnc; continue; }
        G.lncK[k] = m.lnc + G.bgn[k] + G.qcn[k];
        G.lnvK[k] = m.lnv + G.bgn[k] + G.qcp[k];
        G.niK[k] = m.ni * exp(G.bgn[k] + 0.5 * (G.qcn[k] + G.qcp[k]));
        if (G.kind[k] == K_SEMI) {
            double Nt = G.Nd[k] + G.Na[k];
            G.mu0n[k] = host_mob_low(m, Nt, 0);
            G.mu0p[k] = host_mob_low(m, Nt, 1);
        }
    }
    // terminal-current edges, grouped by contact in node order
    std::vector<int> ek, ekb, eoff(G.n_cons + 1, 0);
    std::vector<double> eh, ear;
    std::vector<std::vector<int>> per(G.n_cons);
    for (int k = 0; k < N && !G.graph; k++) {
        int c = G.con_mask[k];
        if (!c || G.kind[k] != K_SEMI || G.cons[c - 1].te) continue;
        per[c - 1].push_back(k);
    }
    for (int c = 0; c < G.n_cons; c++) {
        eoff[c] = (int)ek.size();
        for (int k : per[c]) {
            int i, j, l;
            host_unidx(k, i, j, l);
            for (int dir = 0; dir < 6; dir++) {
                int kb; double h, ar;
                if (!host_nbr(i, j, l, k, dir, kb, h, ar) || G.kind[kb] != K_SEMI || G.con_mask[kb] == c + 1) continue;
                ek.push_back(k); ekb.push_back(kb); eh.push_back(h); ear.push_back(ar);
            }
        }
    }
    eoff[G.n_cons] = (int)ek.size();
    G.n_edges = (int)ek.size();
    // line-smoother axes: the two axes with the largest geometric coupling
    // sum_edges area/h (weaker of the two first); point red-black GS if asked
    // for (SEMISIM_MG_SMOOTH=rb) or if a line is longer than the kernel takes
    if (!G.graph) {
        double sx = 0., sy = 0., sz = 0., SLx = 0., SLy = 0., SLz = 0.;
        for (double v : G.Lx) SLx += v;
        for (double v : G.Ly) SLy += v;
        for (double v : G.Lz) SLz += v;
        for (int i = 0; i + 1 < G.Nx; i++) sx += 1. / (G.xs[i + 1] - G.xs[i]);
        for (int j = 0; j + 1 < G.Ny; j++) sy += 1. / (G.ys[j + 1] - G.ys[j]);
        for (int l = 0; l + 1 < G.Nz; l++) sz += 1. / (G.zs[l + 1] - G.zs[l]);
        double w[3] = {sx * SLy * SLz, sy * SLx * SLz, sz * SLx * SLy};
        int ord[3] = {0, 1, 2};
        std::sort(ord, ord + 3, [&](int a, int b) { return w[a] < w[b]; });
        const char *mode = getenv("SEMISIM_MG_SMOOTH");
        int want = 2;
        if (mode && !strcmp(mode, "rb")) want = 0;
        else if (mode && !strcmp(mode, "lines3")) want = 3;
        else if (mode && !strcmp(mode, "lines1")) want = 1;
        int maxlen = std::max(G.Nx, std::max(G.Ny, G.Nz));
        if (maxlen > 1024) want = 0;
        G.n_axes = want;
        for (int a = 0; a < want; a++) G.mg_axes[a] = ord[3 - want + a];
    }
    // upload
    std::vector<double> geo;
    geo.insert(geo.end(), G.xs.begin(), G.xs.end()); geo.insert(geo.end(), G.ys.begin(), G.ys.end());
    geo.insert(geo.end(), G.zs.begin(), G.zs.end()); geo.insert(geo.end(), G.Lx.begin(), G.Lx.end());
    geo.insert(geo.end(), G.Ly.begin(), G.Ly.end()); geo.insert(geo.end(), G.Lz.begin(), G.Lz.end());
    hipStream_t s = G.st;
    HC(hipMemcpyAsync(G.d_geo, geo.data(), geo.size() * 8, hipMemcpyHostToDevice, s));
    HC(hipMemcpyAsync(G.d_kind, G.kind.data(), N, hipMemcpyHostToDevice, s));
    HC(hipMemcpyAsync(G.d_mat, G.mat8.data(), N, hipMemcpyHostToDevice, s));
    HC(hipMemcpyAsync(G.d_con_mask, G.con_mask.data(), N * 4, hipMemcpyHostToDevice, s));
    HC(hipMemcpyAsync(G.d_gate_mask, G.gate_mask.data(), N * 4, hipMemcpyHostToDevice, s));
    const std::vector<double> *hv[] = {&G.Nd, &G.Na, &G.Nf, &G.eps, &G.bgn, &G.robA, &G.lncK, &G.lnvK, &G.niK, &G.mu0n, &G.mu0p};
    double *dv[] = {G.d_Nd, G.d_Na, G.d_Nf, G.d_eps, G.d_bgn, G.d_robA, G.d_lncK, G.d_lnvK, G.d_niK, G.d_mun, G.d_mup};
    for (int q = 0; q < 11; q++) HC(hipMemcpyAsync(dv[q], hv[q]->data(), (size_t)N * 8, hipMemcpyHostToDevice, s));
    if (G.models & 16) {                  // Lombardi: distance of every node to the nearest interface
        if (!G.graph) eng_wall_distance(G.qd);
        HC(hipMemcpyAsync(G.d_qd, G.qd.data(), (size_t)N * 8, hipMemcpyHostToDevice, s));
    }
    MatDev md[SD_NM_MAX];
    for (int i = 0; i < G.n_mats; i++) md[i] = host_mat_dev(G.mats[i]);
    HC(hipMemcpyAsync(G.d_mats, md, sizeof(MatDev) * G.n_mats, hipMemcpyHostToDevice, s));
    // edge arrays are reallocated per prepare (small; the previous prepare's
    // were released at its start)
    G.d_ek = palloc<int>(std::max(1, G.n_edges)); G.d_ekb = palloc<int>(std::max(1, G.n_edges));
    G.d_eh = palloc<double>(std::max(1, G.n_edges)); G.d_ear = palloc<double>(std::max(1, G.n_edges));
    G.d_eoff = palloc<int>(G.n_cons + 1);
    if (G.n_edges) {
        HC(hipMemcpyAsync(G.d_ek, ek.data(), ek.size() * 4, hipMemcpyHostToDevice, s));
        HC(hipMemcpyAsync(G.d_ekb, ekb.data(), ekb.size() * 4, hipMemcpyHostToDevice, s));
        HC(hipMemcpyAsync(G.d_eh, eh.data(), eh.size() * 8, hipMemcpyHostToDevice, s));
        HC(hipMemcpyAsync(G.d_ear, ear.data(), ear.size() * 8, hipMemcpyHostToDevice, s));
    }
    HC(hipMemcpyAsync(G.d_eoff, eoff.data(), eoff.size() * 4, hipMemcpyHostToDevice, s));
    if (G.graph) {
        if (!gr_prepare_edges()) return 0;
    } else tun_upload();
    eng_sync();
    G.dirty = 0;
    G.ff_seen[0] = G.ff_seen[1] = 0;
    G.ff_off = 0;
    return eng_ok();
}

// device contact table at the applied voltages Va
void eng_upload_cons(const double *Va) {
    ConDev cd[SD_NC_MAX];
    for (int c = 0; c < G.n_cons; c++) {
        const HostCon &cn = G.cons[c];
        cd[c].V = Va[c];
        cd[c].phi_m = cn.phi_m;
        cd[c].Cox = SD_EPS_OX / (cn.tox > 0. ? cn.tox : 10e-9);
        double pm = cn.phi_m > 0. ? cn.phi_m : G.Cref;
        cd[c].VGp = Va[c] + G.Cref - pm;
        cd[c].qox = cn.qox;
        cd[c].bc = cn.bc;
        cd[c].te = cn.te;
    }
    if (G.n_cons) HC(hipMemcpyAsync(G.d_cons, cd, sizeof(ConDev) * G.n_cons, hipMemcpyHostToDevice, G.st));
    // the host array is a stack temporary: finish the copy before returning
    eng_sync();
}

void eng_set_vrange(const double *Va) {
    int any = 0;
    double lo = 0., hi = 0.;
    for (int c = 0; c < G.n_cons; c++) {
        if (G.cons[c].bc == 2 && !G.cons[c].ntp) continue;
        if (!any) { lo = hi = Va[c]; any = 1; }
        else { if (Va[c] < lo) lo = Va[c]; if (Va[c] > hi) hi = Va[c]; }
    }
    G.vlo = lo; G.vhi = hi;
}
void eng_clamp_qf() { ph_clamp_qf(G.st, eng_view(), G.vlo, G.vhi); }
void eng_rebuild_np() { ph_rebuild_np(G.st, eng_view()); }
void eng_upd_mob() {
    if (G.graph) gr_upd_mob(G.d_phin, G.d_phip);
    else ph_upd_mob(G.st, eng_view(), G.d_phin, G.d_phip);
}
void eng_apply_bc() { ph_apply_bc(G.st, eng_view(), G.have_solution); }
static void apply_bc(const double *Va) {
    eng_upload_cons(Va);
    eng_apply_bc();
}

// ------------------------------------------------------------ multigrid
// Multigrid precision the solvers run: the FP32 hierarchy when the FP32
// method is selected, except on a device whose continuity solves found a
// floating majority block of more than 1% of the nodes.  That block's
// near-null mode leaves its balance, and the leakage currents through it,
// sensitive to the preconditioner's rounding (measured on a floating guard
// ring: 2.6 floors from the tight reference sweep with the FP32 hierarchy,
// 0.13 with the FP64 one), so such a device runs the FP64 hierarchy.
static inline int mg_prec() { return G.precision == 1 && !G.ff_off ? 1 : 0; }
template <class T> static Hier<T> &hier();
template <> Hier<double> &hier<double>() { return G.H64; }
template <> Hier<float> &hier<float>() { return G.H32; }

static int g_rec16 = 1;         // 1: the level-0 sweeps read the packed 16-byte coefficient records
extern int g_cont_lds;          // 1: the LDS-staged continuity assembly (k_physics.hip)
static inline int line_axis0(int L);
// BF16 copies of level 0's off-diagonals and line factors for the smoother
// (FP32 hierarchy, when enabled); else the smoother reads the FP32 arrays
static void level0_bf16(Hier<float> &H) {
    MGLevel<float> &v0 = H.lev[0];
    const bool on = G.mg_bf16 && G.n_axes > 0 && G.line_levels > 0 && H.lev.size() > 1;
    if (on && !H.c16) {
        H.c16 = dalloc<__bf16>((size_t)15 * v0.N);
        H.r16 = dalloc<uint4>((size_t)3 * v0.N);
        if (g_err) { H.c16 = nullptr; H.r16 = nullptr; g_err = 0; G.mg_bf16 = 0; return level0_bf16(H); }
    }
    for (int q = 0; q < 6; q++) v0.o16[q] = on ? H.c16 + (size_t)q * v0.N : nullptr;
    for (int a = 0; a < 3; a++)
        for (int q = 0; q < 3; q++) v0.lu16[a][q] = on ? H.c16 + (size_t)(6 + 3 * a + q) * v0.N : nullptr;
    for (int a = 0; a < 3; a++) v0.rec16[a] = nullptr;
    if (!on) return;
    // the separate BF16 arrays are read only by the sweeps that do not take
    // the packed records: the generic line kernel (several lines per slot,
    // lines of at most 32 nodes) or with the records switched off
    bool sep16 = !g_rec16 || !g_line_scan1 || !g_line_scan;
    for (int a = 0; a < G.n_axes; a++) {
        const int ax = G.mg_axes[a], n = ax == 0 ? v0.Nx : (ax == 1 ? v0.Ny : v0.Nz), P = (n + 63) / 64;
        if (P <= 2 && 64 * P / n != 1) sep16 = true;
    }
    if (sep16) mg_o16(G.st, v0);
    for (int a = 0; a < G.n_axes; a++) {
        const int ax = G.mg_axes[a];
        if (sep16) mg_lu16(G.st, v0, ax);
        if (g_rec16) {
            uint4 *r = H.r16 + (size_t)ax * v0.N;
            mg_rec16(G.st, v0, ax, r);
            v0.rec16[ax] = r;
        }
    }
    // level 1, when line-smoothed: its sweeps are latency-bound, and one
    // 16-byte load per node instead of seven 4-byte loads shortens their
    // gather phase
    for (int a = 0; a < 3; a++) if (H.lev.size() > 1) H.lev[1].rec16[a] = nullptr;
    const bool l1 = G.mg_bf16_l1 && g_rec16 && G.line_levels > 1 && (int)H.lev.size() > 2 &&
                    (H.tail < 0 || H.tail > 1);
    if (!l1) return;
    MGLevel<float> &v1 = H.lev[1];
    if (!H.r16b) {
        H.r16b = dalloc<uint4>((size_t)3 * v1.N);
        if (g_err) { H.r16b = nullptr; g_err = 0; G.mg_bf16_l1 = 0; return; }
    }
    for (int a = line_axis0(1); a < G.n_axes; a++) {
        const int ax = G.mg_axes[a];
        uint4 *r = H.r16b + (size_t)ax * v1.N;
        mg_rec16(G.st, v1, ax, r);
        v1.rec16[ax] = r;
    }
}

template <class T> static void mg_setup_T() {
    Hier<T> &H = hier<T>();
    hipStream_t s = G.st;
    DevView d = eng_view();
    int nl = (int)H.lev.size();
    hier_layout(H);
    // level 0: the FP64 operator itself (natural FP64 storage) or a copy in
    // the level's storage order and precision; restriction weights likewise
    MGLevel<T> &v0 = H.lev[0];
    const bool own = v0.px || sizeof(T) != sizeof(double);
    if (own) {
        mg_sten_in(s, G.A, v0, H.A0);
        v0.A = H.A0;
    } else {
        v0.A = *reinterpret_cast<Sten7<T> *>(&G.A);
    }
    if (G.mg_w && v0.px) {
        mg_rs_in(s, G.mg_w, v0, H.rs0);
        v0.rs = H.rs0;
    } else v0.rs = const_cast<double *>(G.mg_w);
    if (G.mg_qsign == 0) {
        if (!H.pw_one) {
            for (int L = 0; L < nl; L++) mg_set_pw_one(s, H.lev[L]);
            H.pw_one = true;
        }
    } else {
        double VT = G.mats[G.ref_mat].VT;
        for (int L = 0; L < nl - 1; L++) mg_weights_level(s, H.lev[L], H.lev[L + 1], L, d, G.mg_qsign, VT);
        H.pw_one = false;
    }
    for (int L = 0; L < nl - 1; L++) mg_galerkin(s, H.lev[L], H.lev[L + 1]);
    // line factors of the levels the line smoother runs on (the others use
    // point Gauss-Seidel)
    for (int L = 0; L < nl - 1 && L < G.line_levels; L++)
        for (int a = 0; a < G.n_axes; a++) mg_line_factor(s, H.lev[L], G.mg_axes[a]);
    if constexpr (std::is_same_v<T, float>) level0_bf16(H);
    mg_coarse_factor(s, H.lev[nl - 1], H.inv);
}
static void mg_setup() {
    double t0 = wtime();
    if (G.graph) gr_mg_setup();
    else if (mg_prec() == 1) mg_setup_T<float>();
    else mg_setup_T<double>();
    G.t_mg += wtime() - t0;
}

// ------------------------------------------------ sampled kernel timer
// HIP events on the solver stream around every 8th level-0 launch of one
// kernel class (0: smoother sweep, 1: Krylov SpMV), for live roofline
// figures; off unless sdx_ktimer_enable is called.
struct KTimer {
    int on = 0, kind = 0, every = 8;
    int armed = 0;                     // enabled and not yet read (pause/resume toggles `on`)
    bool sampling = false;             // kind 0: the running V-cycle is timed
    long long count = 0, sampled = 0, vcycles = 0;
    double total_ms = 0.;
    double bytes = 0.;                 // algorithmic bytes of the sampled launches (kind 0)
    std::vector<hipEvent_t> pool, pend;
    std::vector<int> pend_n;           // launches inside each bracketed pair of pend
};
static KTimer KT;
static hipEvent_t kt_event() {
    hipEvent_t e;
    if (!KT.pool.empty()) { e = KT.pool.back(); KT.pool.pop_back(); return e; }
    HC(hipEventCreate(&e));
    return e;
}
static void kt_flush() {
    for (size_t q = 0; q + 1 < KT.pend.size(); q += 2) {
        float ms = 0.f;
        HC(hipEventSynchronize(KT.pend[q + 1]));
        HC(hipEventElapsedTime(&ms, KT.pend[q], KT.pend[q + 1]));
        KT.total_ms += ms;
        KT.sampled += KT.pend_n[q / 2];
        KT.pool.push_back(KT.pend[q]);
        KT.pool.push_back(KT.pend[q + 1]);
    }
    KT.pend.clear();
    KT.pend_n.clear();
}
static bool kt_begin(int kind) {
    if (!KT.on || KT.kind != kind) return false;
    KT.count++;
    if (kind == 0 ? !KT.sampling : ((KT.count - 1) % KT.every) != 0) return false;
    hipEvent_t e = kt_event();
    HC(hipEventRecord(e, G.st));
    KT.pend.push_back(e);
    return true;
}
// n: kernel launches between the two events; by: their algorithmic bytes
static void kt_end(bool started, int n = 1, double by = 0.) {
    if (!started) return;
    KT.bytes += by;
    hipEvent_t e = kt_event();
    HC(hipEventRecord(e, G.st));
    KT.pend.push_back(e);
    KT.pend_n.push_back(n);
    if (KT.pend.size() >= 512) kt_flush();
}

// Smoother: zebra line Gauss-Seidel along the mesh axes with the strongest
// couplings (G.mg_axes, chosen in prepare), or red-black point Gauss-Seidel
// (G.n_axes == 0).  Pre-smoothing starts from x = 0; returns the restriction
// skip code of the rows the last sweep solved exactly.
// first smoothed axis (index into G.mg_axes) of line level L: coarse line
// levels smooth at most G.coarse_axes axes, the strongest ones (last in order)
static inline int line_axis0(int L) {
    return L > 0 && G.coarse_axes < G.n_axes ? G.n_axes - G.coarse_axes : 0;
}
// Algorithmic bytes of one level-0 line sweep along axis ax (one colour, N/2
// nodes): b, the other colour's x and the x written in T, and 7 coefficients
// (3 line factors, 4 off-line couplings) in BF16 where the sweep reads the
// BF16 copies (FP32 levels, lines of at most 128), else in T
template <class T> static double sweep_bytes(const MGLevel<T> &v, int ax) {
    const double w = sizeof(T);
    const int n = ax == 0 ? v.Nx : (ax == 1 ? v.Ny : v.Nz);
    const bool h16 = sizeof(T) == 4 && (n + 63) / 64 <= 2 && (v.rec16[ax] || v.o16[0]);
    return 0.5 * v.N * (3.0 * w + 7.0 * (h16 ? 2.0 : w));
}
template <class T> static int smooth_pre(const MGLevel<T> &v, const T *b, T *x, int L = 0) {
    hipStream_t s = G.st;
    if (G.n_axes == 0 || L >= G.line_levels) {
        if (g_rb_fuse) mg_rbgs(s, v, b, x, 0, 2);        // both colours from x = 0, one launch
        else {
            mg_rbgs(s, v, b, x, 0, 1);
            mg_rbgs(s, v, b, x, 1, 0);
        }
        return 1;
    }
    // the level-0 sweeps are timed as one bracket (back-to-back launches, as
    // in an untimed V-cycle), the time divided over its launches
    int first = 1;
    bool t = v.N == G.N && kt_begin(0);
    double by = 0.;
    for (int a = line_axis0(L); a < G.n_axes; a++)
        for (int color = 0; color < 2; color++) {
            mg_line(s, v, b, x, G.mg_axes[a], color, first);
            by += sweep_bytes(v, G.mg_axes[a]);
            first = 0;
        }
    kt_end(t, 2 * (G.n_axes - line_axis0(L)), by);
    // rows of the last sweep's colour-1 lines have no residual
    return 2 + G.mg_axes[G.n_axes - 1];
}
// Post-smoothing: the pre-smoother's sweeps in reverse order (symmetric
// V-cycle for CG).  With c != nullptr the first sweep applies the coarse
// correction x += pw * x_parent to the neighbours it reads; the sweeps then
// overwrite every node, so the correction is never stored.
// Returns where the smoothed x lies: x, or the level's second array xs when a
// coarse point-GS level ran its sweep as one launch.
template <class T> static T *smooth_post(const MGLevel<T> &v, const T *b, T *x, const MGLevel<T> *c, int L = 0) {
    hipStream_t s = G.st;
    if (G.n_axes == 0 || L >= G.line_levels) {
        if (g_rb_fuse_post && c && v.xs && x != v.xs) {
            mg_rbgs_post2(s, v, b, x, v.xs, *c);
            return v.xs;
        }
        mg_rbgs(s, v, b, x, 1, 0, c);
        mg_rbgs(s, v, b, x, 0, 0);
        return x;
    }
    bool t = v.N == G.N && kt_begin(0);
    double by = 0.;
    for (int a = G.n_axes - 1; a >= line_axis0(L); a--)
        for (int color = 1; color >= 0; color--) {
            mg_line(s, v, b, x, G.mg_axes[a], color, 0, c);
            by += sweep_bytes(v, G.mg_axes[a]);
            c = nullptr;
        }
    kt_end(t, 2 * (G.n_axes - line_axis0(L)), by);
    return x;
}
// V-cycle from level L; returns where level L's result lies (x, or the
// level's xs after a one-launch post-smoothing sweep)
template <class T> static T *vcycle_T(int L, const T *b, T *x) {
    Hier<T> &H = hier<T>();
    hipStream_t s = G.st;
    int nl = (int)H.lev.size();
    MGLevel<T> &v = H.lev[L];
    if (L == nl - 1) { mg_coarse_solve(s, v, H.inv, b, x); return x; }
    if (L == H.tail) { mg_tail(s, H.d_tail, &H.lev[L], nl - L, H.inv, G.tail_sweeps); return x; }   // lev[L].b -> .x
    MGLevel<T> c = H.lev[L + 1];
    int skip = smooth_pre(v, b, x, L);
    // coarse right-hand side; the coarse-correction scale may differ for the
    // Poisson (CG) and continuity (BiCGSTAB) operators
    mg_resid_restrict(s, v, c, b, x, skip, G.mg_qsign == 0 ? G.mg_alpha_p : G.mg_alpha);
    c.x = vcycle_T<T>(L + 1, c.b, c.x);
    return smooth_post(v, b, x, &c, L);
}
// smoother only (the reference's ILU-only retry): one symmetric smoothing from 0
template <class T> static void smooth_only_T(const T *b, T *x) {
    Hier<T> &H = hier<T>();
    smooth_pre(H.lev[0], b, x);
    smooth_post<T>(H.lev[0], b, x, nullptr);
}
// The level-0 staging of the current hierarchy (off when level 0 works on the
// Krylov vectors directly)
static Stage0 stage0() {
    Stage0 st{nullptr, nullptr, 0, 0, 0, nullptr};
    if (!G.stage_on || G.graph) return st;
    if (mg_prec() == 1 && !G.H32.lev.empty()) {
        const MGLevel<float> &v0 = G.H32.lev[0];
        st = Stage0{G.d_b32, G.d_x32, v0.Nx, v0.px, 1, G.d_x32b};
    } else if (mg_prec() == 0 && !G.H64.lev.empty() && G.H64.lev[0].px) {
        const MGLevel<double> &v0 = G.H64.lev[0];
        st = Stage0{G.H64.b0, G.H64.x0, v0.Nx, v0.px, 0, G.H64.x0b};
    }
    return st;
}
// z = M^-1 r.  sin: r is already staged in the level-0 b (r unused); sout:
// the result stays in the level-0 x (z unused), for a consumer that reads it
// from there.  Both need stage0() on.
// second: the result goes to the second level-0 solution array (Stage0::x2)
static void precond_launch(const double *r, double *z, bool sin = false, bool sout = false, bool second = false) {
    if (mg_prec() == 1) {
        const MGLevel<float> &v0 = G.H32.lev[0];
        float *x0 = second ? G.d_x32b : G.d_x32;
        if (!sin) mg_vec_in(G.st, r, v0, G.d_b32);
        if (G.mg_off || G.H32.lev.size() < 2) smooth_only_T<float>(G.d_b32, x0);
        else vcycle_T<float>(0, G.d_b32, x0);
        if (!sout) mg_vec_out(G.st, x0, v0, z);
    } else if (G.H64.lev[0].px) {
        const MGLevel<double> &v0 = G.H64.lev[0];
        double *x0 = second ? G.H64.x0b : G.H64.x0;
        if (!sin) mg_vec_in(G.st, r, v0, G.H64.b0);
        if (G.mg_off || G.H64.lev.size() < 2) smooth_only_T<double>(G.H64.b0, x0);
        else vcycle_T<double>(0, G.H64.b0, x0);
        if (!sout) mg_vec_out(G.st, x0, v0, z);
    } else {
        if (G.mg_off || G.H64.lev.size() < 2) smooth_only_T<double>(r, z);
        else vcycle_T<double>(0, r, z);
    }
}
// The preconditioner as the Krylov solvers call it.  With the level-0
// smoother timer on, every n-th V-cycle has its level-0 sweeps bracketed by
// HIP events; kind 2 times whole applications.
static void precond(const double *r, double *z, bool sin = false, bool sout = false, bool second = false) {
    if (G.graph) { gr_precond(r, z); return; }
    G.n_vcycles++;
    const bool sample = KT.on && KT.kind == 0 && (KT.vcycles++ % KT.every) == 0;
    KT.sampling = sample;
    const bool whole = kt_begin(2);          // kind 2: the whole preconditioner application
    precond_launch(r, z, sin, sout, second);
    kt_end(whole);
    KT.sampling = false;
}
static inline void matvec(const double *x, double *y) {
    G.n_spmv++;
    if (G.graph) gr_spmv(x, nullptr, y, false);
    else sd_spmv(G.st, eng_view().g, G.A, x, nullptr, y);
}
// y = A x fused with dots (sd_spmv_dots); the graph operator takes no staging
static inline void spmv_dots(const Geo &g, const Sten7<double> &A, const double *x, double *y, const double *u0,
                             const double *u1, int nv, int slot, int post, Stage0 st = Stage0{}) {
    if (G.graph) gr_spmv_dots(x, y, u0, u1, nv, slot, post, A.c == nullptr);
    else sd_spmv_dots(G.st, g, A, x, y, u0, u1, nv, G.rb, slot, post, st);
}

// ------------------------------------------------ FP32 inner Krylov solves
// The FP32 method's linear solves.  The residual b - A x is formed in FP64
// with the FP64 operator and handed to a PCG or BiCGSTAB whose vectors are
// FP32 in the FP32 multigrid's level-0 storage, so that the V-cycle reads its
// right-hand side and writes its result in place (no conversion passes) and
// the products read the FP32 level-0 operator; every dot is accumulated in
// FP64.  The correction is added to x in FP64 and the FP64 residual formed
// again (defect correction), which wins back the accuracy the FP32 operator
// and vectors lose.  Same stopping rules as the FP64 solvers.
// On once both carriers' continuity solves have run on the prepared device,
// and off for good (until the next prepare) once one of them found a
// floating majority block of more than 1% of the nodes (mg_prec() then
// selects the FP64 hierarchy).  Such a block makes its operator
// near-singular along the block's near-null mode, which amplifies the FP32
// operator's rounding into the block balance: on a floating guard ring (a
// block of 30% of the nodes) the leakage currents moved by more than their
// own value.  The pockets float_fix finds now and then in the MOSFET sweeps
// (64-384 nodes, under 0.1%) leave every current within one floor of the
// reference's tight sweep.  G.inner32 = 2 (a test hook) does not wait for
// the carriers' first solves.
static bool inner32_on() {
    return mg_prec() == 1 && G.inner32 && !G.graph && G.d_k32 && G.H32.lev.size() > 1 && G.H32.lev[0].A.xm &&
           (G.inner32 == 2 || (G.ff_seen[0] && G.ff_seen[1]));
}
static L0Geo l0geo() {
    const MGLevel<float> &v0 = G.H32.lev[0];
    return L0Geo{v0.Nx, v0.Ny, v0.Nz, v0.px};
}
struct K32Vec {
    float *r, *rh, *p, *v, *ph, *sh, *t, *d;
};
static K32Vec k32vec() {
    float *b = G.d_k32;
    const size_t N = G.N;
    return K32Vec{b, b + N, b + 2 * N, b + 3 * N, b + 4 * N, b + 5 * N, b + 6 * N, b + 7 * N};
}
// z = M^-1 r on level-0 storage (the V-cycle reads r and writes z in place)
static void precond32(const float *r, float *z) {
    G.n_vcycles++;
    const bool sample = KT.on && KT.kind == 0 && (KT.vcycles++ % KT.every) == 0;
    KT.sampling = sample;
    const bool whole = kt_begin(2);
    if (G.mg_off) smooth_only_T<float>(r, z);
    else vcycle_T<float>(0, r, z);
    kt_end(whole);
    KT.sampling = false;
}
// PCG for the Newton step (the Newton iteration is the outer correction):
// the same M^-1-norm stop as pcg()
static int pcg32(const double *b, double *x, int maxit, double rtol) {
    const int N = G.N;
    hipStream_t s = G.st;
    const K32Vec K = k32vec();
    float *r = K.r, *z = K.ph, *p = K.p, *q = K.v, *d = K.d;
    const L0Geo g0 = l0geo();
    const MGLevel<float> &v0 = G.H32.lev[0];
    G.n_inner++;
    G.n_outer++;
    k32_resid_launch(s, g0, G.A, false, b, x, r, nullptr, nullptr, nullptr, nullptr, G.rb, S_TMP1);
    k32_fill_launch(s, d, 0.f, N);
    precond32(r, z);
    k32_dot_launch(s, r, z, N, G.rb, S_RZ);
    rd(S_RZ, S_RZ + 1);
    const double rz0 = S(S_RZ);
    if (!(rz0 > 0.)) return 0;
    k32_copy_launch(s, p, z, N);
    int it;
    for (it = 1; it <= maxit; it++) {
        G.n_spmv++;
        { bool kt = kt_begin(1);
          k32_spmv_dots_launch(s, g0, v0.A, false, p, q, p, nullptr, 1, G.rb, S_PQ, SOP_PCG_AL);
          kt_end(kt); }
        k32_pcg_xr_launch(s, d, r, p, q, N, G.rb);
        precond32(r, z);
        k32_dot_launch(s, r, z, N, G.rb, S_TMP0, SOP_PCG_BE);
        rd(S_RZ, S_TMP0 + 1);
        const double pq = S(S_PQ), rzn = S(S_TMP0);
        if (!(pq > 0.)) {
            if (G.verbose > 3) fprintf(stderr, "      pcg32 breakdown pq=%.3e it=%d\n", pq, it);
            break;
        }
        if (!(rzn > rtol * rtol * rz0)) break;
        k32_pcg_p_launch(s, p, z, N, G.rb);
    }
    k32_add_out_launch(s, g0, x, d);
    if (G.verbose > 3) fprintf(stderr, "      pcg32 its=%d rz0=%.3e\n", it, rz0);
    G.pcg_its += it;
    return it;
}
// Inner BiCGSTAB from d = 0 on the FP32 residual r (rh = r, p = v = 0 set by
// k32_resid), stopping where its recurrence residual reaches tgt; the rules
// of bicgstab_core (shadow restart, 40-iteration stall).  Returns the
// recurrence residual norm; its iterations in *its.
static double bicgstab32_inner(double tgt, double rstart, int maxit, int *its) {
    const int N = G.N;
    hipStream_t s = G.st;
    const K32Vec K = k32vec();
    const L0Geo g0 = l0geo();
    const MGLevel<float> &v0 = G.H32.lev[0];
    const bool unit = G.unit_diag;
    sd_scalar_op(s, SOP_BI_RESTART, G.rb);       // rho = al = om = 1
    double rr = rstart * rstart, rbest = rstart;
    int stall = 0, it, restarted = 0;
    for (it = 1; it <= maxit; it++) {
        if (it == 1) k32_dot_launch(s, K.rh, K.r, N, G.rb, S_RHO1, SOP_BI_BE);
    again:
        k32_bi_p_launch(s, K.p, K.r, K.v, N, G.rb);
        precond32(K.p, K.ph);
        G.n_spmv++;
        { bool kt = kt_begin(1);
          k32_spmv_dots_launch(s, g0, v0.A, unit, K.ph, K.v, K.rh, nullptr, 1, G.rb, S_RV, SOP_BI_AL);
          kt_end(kt); }
        k32_bi_s_launch(s, K.r, K.v, N, G.rb, S_SS);
        rd(S_AL, S_SS + 1);
        const double rho1 = S(S_RHO1), rv = S(S_RV), ss = S(S_SS);
        if (fabs(rho1) < 1e-300 && !restarted) {
            k32_bi_restart_launch(s, K.rh, K.p, K.v, K.r, N);
            sd_scalar_op(s, SOP_BI_RESTART, G.rb);
            k32_dot_launch(s, K.r, K.r, N, G.rb, S_RHO1, SOP_BI_BE);
            rd(S_RHO1, S_RHO1 + 1);
            if (!(S(S_RHO1) > 0.)) break;
            restarted = 1;
            goto again;
        }
        restarted = 0;
        if (fabs(rv) < 1e-300) break;
        if (sqrt(ss) <= tgt) {
            k32_bi_x1_launch(s, K.d, K.ph, N, G.rb);
            rr = ss;
            break;
        }
        precond32(K.r, K.sh);
        G.n_spmv++;
        { bool kt = kt_begin(1);
          k32_spmv_dots_launch(s, g0, v0.A, unit, K.sh, K.t, nullptr, K.r, 2, G.rb, S_TT, SOP_BI_OM);
          kt_end(kt); }
        k32_bi_xr2_launch(s, K.d, K.r, K.ph, K.sh, K.t, K.rh, N, G.rb);
        rd(S_OM, S_RR + 1);
        const double om = S(S_OM);
        rr = S(S_RR);
        const double rn = sqrt(rr);
        if (G.verbose > 5) fprintf(stderr, "        it32 %d rn=%.3e om=%.3e\n", it, rn, om);
        if (rn <= tgt || om == 0.) break;
        if (rn < 0.99 * rbest) { rbest = rn; stall = 0; }
        else if (++stall > 40) break;
    }
    if (it > maxit) it = maxit;
    *its = it;
    return sqrt(rr);
}
// Continuity solve of the FP32 method: the target of bicgstab_core (rtol |r0|
// or the rounding floor), then FP64 defect correction around inner FP32
// BiCGSTAB solves.  Returns the FP64 residual norm; *ok = false when a
// correction failed to bring it down (the caller then solves with the FP64
// Krylov form from the saved x).
static double bicgstab_mixed(const double *b, double *x, int maxit, double rtol, double atol, int *its,
                             double *r0out, bool *ok) {
    hipStream_t s = G.st;
    const K32Vec K = k32vec();
    const L0Geo g0 = l0geo();
    const bool unit = G.unit_diag;
    Sten7<double> A64 = G.A;
    if (unit) A64.c = nullptr;
    Geo g = eng_view().g;
    *ok = true;
    *its = 0;
    G.n_inner++;
    if (!G.mg_off) mg_setup();
    sd_floor(s, g, G.A, x, b, G.rb, S_FL);
    k32_resid_launch(s, g0, A64, unit, b, x, K.r, K.rh, K.p, K.v, K.d, G.rb, S_RR);
    rd(S_RR, S_FL + 1);
    const double r0 = sqrt(S(S_RR)), fl = S(S_FL);
    const double tgt = fmax(rtol * r0, atol * sqrt(fl));
    *r0out = r0;
    if (G.verbose > 3) fprintf(stderr, "      r0=%.3e tgt=%.3e (fp32 inner)\n", r0, tgt);
    if (!(r0 > tgt)) return r0;
    double rn = r0;
    for (int outer = 0; outer < 8 && *its < maxit; outer++) {
        // an inner solve reduces its residual by at most 1e-5, about what the
        // FP32 recurrence follows; a tighter target takes more corrections
        int it_in = 0;
        const double rin = bicgstab32_inner(fmax(tgt, 1e-5 * rn), rn, maxit - *its, &it_in);
        *its += it_in;
        G.bicg_its += it_in;
        G.n_outer++;
        if (!std::isfinite(rin)) { *ok = false; return INFINITY; }
        k32_add_out_launch(s, g0, x, K.d);
        k32_resid_launch(s, g0, A64, unit, b, x, K.r, K.rh, K.p, K.v, K.d, G.rb, S_RR);
        rd(S_RR, S_RR + 1);
        const double rn2 = sqrt(S(S_RR));
        if (G.verbose > 3) fprintf(stderr, "      outer %d: inner its=%d |r|=%.3e -> %.3e\n", outer, it_in, rn, rn2);
        if (!std::isfinite(rn2)) { *ok = false; return INFINITY; }
        if (rn2 <= tgt) return rn2;
        // the correction must contract the FP64 residual
        if (!(rn2 < 0.5 * rn)) { *ok = false; return rn2; }
        rn = rn2;
    }
    return rn;
}

// ------------------------------------------------------------ Krylov
// Preconditioned CG on the SPD Poisson Jacobian; stops on the M^-1 norm
static int pcg(const double *b, double *x, int maxit, double rtol) {
    const int N = G.N;
    hipStream_t s = G.st;
    double *r = G.d_wk[0], *z = G.d_wk[1], *p = G.d_wk[2], *q = G.d_wk[3];
    mg_setup();
    // the Newton step's PCG (rtol 1e-3 by default) in FP32; a tighter
    // M^-1-norm reduction than the FP32 recurrence can follow stays FP64
    if (inner32_on() && rtol >= 1e-5) return pcg32(b, x, maxit, rtol);
    Geo g = eng_view().g;
    if (G.graph) gr_spmv(x, b, r, false);  // r = b - A x
    else sd_spmv(s, g, G.A, x, b, r);
    precond(r, z);
    sd_dot(s, r, z, N, G.rb, S_RZ);
    rd(S_RZ, S_RZ + 1);
    double rz0 = S(S_RZ);
    if (!(rz0 > 0.)) return 0;
    sd_copy(s, p, z, N);
    const Stage0 st = stage0();
    const bool sg = st.b != nullptr;
    const bool sz = sg && G.stage_read;       // z is read where the V-cycle left it
    int it;
    for (it = 1; it <= maxit; it++) {
        G.n_spmv++;
        { bool kt = kt_begin(1);
          spmv_dots(g, G.A, p, q, p, nullptr, 1, S_PQ, SOP_PCG_AL);
          kt_end(kt); }
        sd_pcg_update_xr(s, x, r, p, q, N, G.rb, st);      // r, staged for the V-cycle
        precond(r, z, sg, sg);
        if (sg) sd_dot_stage(s, r, sz ? nullptr : z, st, N, G.rb, S_TMP0, SOP_PCG_BE);   // z from the staged result
        else sd_dot(s, r, z, N, G.rb, S_TMP0, SOP_PCG_BE);
        rd(S_RZ, S_TMP0 + 1);
        double pq = S(S_PQ), rzn = S(S_TMP0);
        if (!(pq > 0.)) {
            if (G.verbose > 3) fprintf(stderr, "      pcg breakdown pq=%.3e it=%d\n", pq, it);
            break;
        }
        if (!(rzn > rtol * rtol * rz0)) break;
        if (sz) sd_pcg_update_p_st(s, p, st, N, G.rb);
        else sd_pcg_update_p(s, p, z, N, G.rb);
    }
    if (G.verbose > 3) fprintf(stderr, "      pcg its=%d rz0=%.3e\n", it, rz0);
    G.pcg_its += it;
    return it;
}

static double bicgstab_core(const double *b, double *x, int maxit, double rtol, double atol, int *its, double *r0out) {
    const int N = G.N;
    hipStream_t s = G.st;
    double *r = G.d_wk[0], *rh = G.d_wk[1], *p = G.d_wk[2], *v = G.d_wk[3], *ph = G.d_wk[4], *sh = G.d_wk[5],
           *t = G.d_wk[6];
    Geo g = eng_view().g;
    // the products inside the loop skip the diagonal when it is known to be 1
    // (multiplying by 1 is exact, so the iterates are the same)
    Sten7<double> Ak = G.A;
    if (G.unit_diag) Ak.c = nullptr;
    if (!G.mg_off) mg_setup();
    matvec(x, t);
    sd_bi_init(s, r, rh, p, v, b, t, N, G.rb);
    if (G.graph) gr_floor(x, b, S_FL);
    else sd_floor(s, g, G.A, x, b, G.rb, S_FL);
    rd(S_RR, S_FL + 1);
    double rr = S(S_RR), fl = S(S_FL);
    double r0 = sqrt(rr), tgt = fmax(rtol * r0, atol * sqrt(fl));
    if (G.verbose > 3) fprintf(stderr, "      r0=%.3e tgt=%.3e\n", r0, tgt);
    *its = 0;
    *r0out = r0;
    if (!(r0 > tgt)) return r0;
    sd_scalar_op(s, SOP_BI_RESTART, G.rb);       // rho = al = om = 1
    double rbest = r0;
    int stall = 0, it, restarted = 0;
    // the V-cycles read p and s staged by the kernels that update them, and
    // the SpMV after each reads the V-cycle's result where it lies
    const Stage0 st = stage0();
    const bool sg = st.b != nullptr;
    // ph and sh stay where their V-cycles left them (two level-0 arrays): the
    // x update reads them there, and the products write no FP64 copies
    const bool sr = sg && G.stage_read;
    Stage0 st2 = st;
    st2.x = st.x2;
    for (it = 1; it <= maxit; it++) {
        // rho1 = rh.r and beta: here on the first iteration, later by the
        // previous iteration's x/r update
        if (it == 1 || !G.bi_fuse) sd_dot(s, rh, r, N, G.rb, S_RHO1, SOP_BI_BE);
    again:
        sd_bi_p(s, p, r, v, N, G.rb, st);
        precond(p, ph, sg, sg);
        G.n_spmv++;
        { bool kt = kt_begin(1);
          spmv_dots(g, Ak, sr ? nullptr : ph, v, rh, nullptr, 1, S_RV, SOP_BI_AL, st);   // (also writes ph)
          kt_end(kt); }
        sd_bi_r1(s, r, v, N, G.rb, st);
        rd(S_AL, S_SS + 1);
        double rho1 = S(S_RHO1), rv = S(S_RV), ss = S(S_SS);
        if (fabs(rho1) < 1e-300 && !restarted) {
            // breakdown: restart the shadow residual (al was ~0, r unchanged)
            sd_bi_restart(s, rh, p, v, r, N);
            sd_scalar_op(s, SOP_BI_RESTART, G.rb);
            sd_dot(s, r, r, N, G.rb, S_RHO1, SOP_BI_BE);
            rd(S_RHO1, S_RHO1 + 1);
            if (!(S(S_RHO1) > 0.)) break;
            restarted = 1;
            goto again;
        }
        restarted = 0;
        if (fabs(rv) < 1e-300) break;
        if (sqrt(ss) <= tgt) {
            if (sr) sd_bi_x1_st(s, x, st, N, G.rb);
            else sd_bi_x1(s, x, ph, N, G.rb);
            rr = ss;
            break;
        }
        precond(r, sh, sg, sg, sr);
        G.n_spmv++;
        { bool kt = kt_begin(1);
          spmv_dots(g, Ak, sr ? nullptr : sh, t, nullptr, r, 2, S_TT, SOP_BI_OM, sr ? st2 : st);
          kt_end(kt); }
        if (sr) sd_bi_xr2_st(s, x, r, st, t, G.bi_fuse ? rh : nullptr, N, G.rb);
        else sd_bi_xr2(s, x, r, ph, sh, t, G.bi_fuse ? rh : nullptr, N, G.rb);
        rd(S_OM, S_RR + 1);
        double om = S(S_OM);
        rr = S(S_RR);
        double rn = sqrt(rr);
        if (G.verbose > 5) fprintf(stderr, "        it %d rn=%.3e om=%.3e\n", it, rn, om);
        if (rn <= tgt || om == 0.) break;
        if (rn < 0.99 * rbest) { rbest = rn; stall = 0; }
        else if (++stall > 40) break;
    }
    if (it > maxit) it = maxit;
    G.bicg_its += it;
    *its = it;
    return sqrt(rr);
}

static double bicgstab(const double *b, double *x, int maxit, double rtol, double atol, int *its) {
    const int N = G.N;
    sd_copy(G.st, G.d_xsave, x, N);
    double r0, rn;
    if (inner32_on()) {
        bool ok;
        rn = bicgstab_mixed(b, x, maxit, rtol, atol, its, &r0, &ok);
        if (ok) return rn;
        // the FP32 correction did not contract the FP64 residual: solve again
        // with FP64 Krylov (FP32 V-cycle) from the saved start
        G.n_inner_fb++;
        if (G.verbose) fprintf(stderr, "semisim-hip: FP32 inner BiCGSTAB did not converge (%.2e), FP64 Krylov\n", rn);
        sd_copy(G.st, x, G.d_xsave, N);
        int its0 = *its;
        rn = bicgstab_core(b, x, maxit, rtol, atol, its, &r0);
        *its += its0;
    } else rn = bicgstab_core(b, x, maxit, rtol, atol, its, &r0);
    int nl = G.graph ? gr_nlev() : (int)(mg_prec() == 1 ? G.H32.lev.size() : G.H64.lev.size());
    if (!(rn <= fmax(r0, 1e-300)) && nl > 1) {
        if (G.verbose) fprintf(stderr, "semisim-hip: MG-BiCGSTAB diverged (%.2e > %.2e), retry smoother only\n", rn, r0);
        sd_copy(G.st, x, G.d_xsave, N);
        int its2;
        G.mg_off = 1;
        rn = bicgstab_core(b, x, 4 * maxit, rtol, atol, &its2, &r0);
        G.mg_off = 0;
        *its += its2;
    }
    if (!(rn <= fmax(r0, 1e-300)) || !std::isfinite(rn)) {
        sd_copy(G.st, x, G.d_xsave, N);
        return INFINITY;
    }
    return rn;
}

// ------------------------------------------------------------ Poisson
static double poisson_eval(const double *phi, int build, double *F) {
    if (G.graph) gr_poisson(phi, build, F, S_NRM);
    else ph_poisson(G.st, eng_view(), phi, build, F, G.A, G.d_sc, G.rb, S_NRM);
    rd(S_NRM, S_NRM + 1);
    return sqrt(S(S_NRM));
}

static int poisson_newton(double ntol, int maxit) {
    const int N = G.N;
    hipStream_t s = G.st;
    double *dx = G.d_sol, *tr = G.d_tr, *Ft = G.d_Ft;
    for (int it = 0; it < maxit; it++) {
        G.newton_its++;
        double f0 = poisson_eval(G.d_psi, 1, G.d_rhs);
        if (!(f0 > 0.)) return f0 == 0.;
        sd_fill(s, dx, 0.0, N);
        G.mg_w = nullptr;
        G.mg_qsign = 0;
        pcg(G.d_rhs, dx, G.max_iter_p > 0 ? G.max_iter_p : 2000, inner32_on() ? G.pcg_rtol32 : G.pcg_rtol);
        sd_maxabs(s, dx, N, G.rb, S_MAX0);
        rd(S_MAX0, S_MAX0 + 1);
        double dm = S(S_MAX0);
        if (dm != dm) return 0;
        if (dm < ntol) {
            ph_update_psi(s, N, G.d_psi, dx, 1.0, G.d_psi);
            return 1;
        }
        double lam = 1.;
        for (int ls = 0; ls < 40; ls++) {
            ph_update_psi(s, N, G.d_psi, dx, lam, tr);
            double f1 = poisson_eval(tr, 0, Ft);
            if (f1 <= (1. - 1e-4 * lam) * f0) break;
            lam *= 0.5;
        }
        sd_copy(s, G.d_psi, tr, N);
        if (G.verbose > 4) fprintf(stderr, "      newton it=%d f0=%.3e dm=%.3e lam=%.3e\n", it, f0, dm, lam);
        if (lam * dm < ntol) return 1;
    }
    return 0;
}

// ------------------------------------------------------------ continuity
// Floating majority regions (semiconductor_3d.c:3159-3228).  The device marks
// majority nodes, spreads "anchored" from the contact neighbours along
// majority connectivity, and lists the nodes left unanchored; only those are
// gathered for the host, which labels their blocks in the reference's order
// (ascending start node, neighbour order x-,x+,y-,y+,z-,z+) and solves each
// block's balance in ln(lambda).
static int float_fix_host(int hole, int m);
// Spread "anchored" (flag 2) through the majority nodes (flag 1) along the
// mesh axes until nothing changes: afterwards a majority node is anchored
// exactly when its 6-connected block of majority nodes touches a contact.
// Returns the rounds used.
static int maj_closure() {
    DevView d = eng_view();
    hipStream_t s = G.st;
    // two passes over the axes per round; the round ends the propagation when
    // its second pass changed nothing (the flags are then the closure)
    int round;
    for (round = 0; round < 64; round++) {
        HC(hipMemsetAsync(G.d_changed, 0, 8, s));
        for (int rep = 0; rep < 2; rep++)
            for (int ax = 0; ax < 3; ax++) ph_maj_propagate(s, d, G.d_flag, ax, G.d_changed + rep);
        int ch = 0;
        HC(hipMemcpyAsync(&G.h_scal[S_TMP3], G.d_changed + 1, 4, hipMemcpyDeviceToHost, s));
        eng_sync();
        memcpy(&ch, &G.h_scal[S_TMP3], 4);
        if (!ch) break;
    }
    return round + 1;
}
static int float_fix(int hole) {
    if (G.graph) return gr_float_fix(hole);
    DevView d = eng_view();
    hipStream_t s = G.st;
    ph_maj_flags(s, d, hole, G.d_sol, G.d_sc, G.d_flag);
    maj_closure();
    HC(hipMemsetAsync(G.d_changed + 2, 0, 4, s));
    ph_maj_list(s, G.N, G.d_flag, G.d_changed + 2, G.d_mlist, G.N);
    HC(hipMemcpyAsync(&G.h_scal[S_TMP3], G.d_changed + 2, 4, hipMemcpyDeviceToHost, s));
    eng_sync();
    int m = 0;
    memcpy(&m, &G.h_scal[S_TMP3], 4);
    G.ff_seen[hole ? 1 : 0] = 1;
    if ((long)m * 100 > (long)G.N) G.ff_off = 1;          // a large block: until the next prepare
    if (m <= 0 || g_err) return 0;
    return float_fix_host(hole, m > G.N ? G.N : m);
}

// block balance of the m listed (unanchored majority) nodes
static int float_fix_host(int hole, int m) {
    const int N = G.N;
    hipStream_t s = G.st;
    if ((size_t)m * SD_MAJ_REC > G.mrec_cap) {
        if (G.d_mrec) HC(hipFree(G.d_mrec));
        G.mrec_cap = (size_t)m * SD_MAJ_REC;
        if (hipMalloc(&G.d_mrec, G.mrec_cap * 8) != hipSuccess) { G.d_mrec = nullptr; G.mrec_cap = 0; g_err = 1; return 0; }
    }
    std::vector<int> list(m);
    std::vector<double> rec((size_t)m * SD_MAJ_REC);
    ph_maj_gather(s, eng_view(), hole, G.d_sol, G.d_sc, G.d_mlist, m, G.d_mrec);
    HC(hipMemcpyAsync(list.data(), G.d_mlist, (size_t)m * 4, hipMemcpyDeviceToHost, s));
    HC(hipMemcpyAsync(rec.data(), G.d_mrec, rec.size() * 8, hipMemcpyDeviceToHost, s));
    eng_sync();
    if (g_err) return 0;
    std::unordered_map<int, int> pos;          // node -> row of list/rec
    pos.reserve((size_t)m * 2);
    for (int t = 0; t < m; t++) pos[list[t]] = t;
    std::vector<int> order(list);
    std::sort(order.begin(), order.end());
    std::vector<int> lab(m, 0), q(m);
    std::vector<double> lamv(m, 1.0);
    auto R = [&](int k) { return &rec[(size_t)pos[k] * SD_MAJ_REC]; };
    int nfix = 0, blk = 0;
    for (int k0 : order) {
        int t0 = pos[k0];
        if (lab[t0]) continue;
        blk++;
        int head = 0, tail = 0;
        q[tail++] = k0;
        lab[t0] = blk;
        while (head < tail) {
            int k = q[head++], i, j, l;
            host_unidx(k, i, j, l);
            for (int dir = 0; dir < 6; dir++) {
                int kb; double h, ar;
                if (!host_nbr(i, j, l, k, dir, kb, h, ar) || G.kind[kb] != K_SEMI || G.con_mask[kb]) continue;
                auto it = pos.find(kb);
                if (it != pos.end() && !lab[it->second]) { lab[it->second] = blk; q[tail++] = kb; }
            }
        }
        double Aout = 0., Bin = 0.;
        for (int t = 0; t < tail; t++) {
            int k = q[t], i, j, l;
            host_unidx(k, i, j, l);
            const double *r = R(k);
            const HostMat &ma = G.mats[G.mat8[k]];
            double VT = ma.VT;
            for (int dir = 0; dir < 6; dir++) {
                int kb; double h, ar;
                if (!host_nbr(i, j, l, k, dir, kb, h, ar) || G.kind[kb] != K_SEMI) continue;
                auto it = pos.find(kb);
                if (it != pos.end() && lab[it->second] == blk) continue;
                const double *nb3 = r + 4 + 3 * dir;
                double dl = hole ? -(G.lnvK[kb] - G.lnvK[k]) : (G.lncK[kb] - G.lncK[k]);
                double D = (nb3[0] - r[3]) / VT + dl;
                double gg = ar * nb3[1] * VT / h;
                double self = hole ? gg * host_B(D) : gg * host_B(-D), nb = hole ? gg * host_B(-D) : gg * host_B(D);
                Aout += self * r[1] * r[0];
                Bin += nb * nb3[2];
            }
        }
        double sl = 0.;
        for (int it = 0; it < 60; it++) {
            double lam = exp(sl), Gv = lam * Aout - Bin, dG = lam * Aout;
            for (int t = 0; t < tail; t++) {
                int k = q[t], i, j, l;
                host_unidx(k, i, j, l);
                const double *r = R(k);
                const HostMat &mt = G.mats[G.mat8[k]];
                double c = lam * r[1] * r[0], Rr, dRn, dRp;
                double tf = ((G.models & 2) && mt.tau_Nref > 0.) ? 1. / (1. + (G.Nd[k] + G.Na[k]) / mt.tau_Nref) : 1.;
                host_recomb(mt, G.niK[k], tf, hole ? r[2] : c, hole ? c : r[2], &Rr, &dRn, &dRp);
                double vol = G.Lx[i] * G.Ly[j] * G.Lz[l];
                Gv += vol * Rr;
                dG += vol * (hole ? dRp : dRn) * c;
                Gv += r[23] * c - r[24] - r[22];     // tunnelling sink/source, impact ionisation
                dG += r[23] * c;
            }
            if (!(dG > 0.)) break;
            double ds = -Gv / dG;
            if (ds > 2.) ds = 2.; else if (ds < -2.) ds = -2.;
            sl += ds;
            if (fabs(ds) < 1e-13) break;
        }
        if (G.verbose > 3) fprintf(stderr, "      block %d: %d nodes, ln lambda %.3e\n", blk, tail, sl);
        if (sl == sl && fabs(sl) > 1e-15 && fabs(sl) < 200.) {
            double lam = exp(sl);
            for (int t = 0; t < tail; t++) lamv[pos[q[t]]] = lam;
            nfix++;
        }
    }
    if (nfix) {
        double *d_lam = G.d_wk[6];             // free after the Krylov solve
        HC(hipMemcpyAsync(d_lam, lamv.data(), (size_t)m * 8, hipMemcpyHostToDevice, s));
        ph_scale_list(s, G.d_sol, G.d_mlist, d_lam, m);
        eng_sync();                             // lamv is pageable host memory
    }
    (void)N;
    G.ff_stat[0]++; G.ff_stat[1] += blk; G.ff_stat[2] += nfix;
    if (G.verbose > 2) fprintf(stderr, "    float_fix(%c): %d unanchored nodes, %d blocks, %d rescaled\n",
                               hole ? 'p' : 'n', m, blk, nfix);
    return nfix;
}

static double CONT_ATOL = 2e-14;
static int g_unit_diag = 1;            // 0: the continuity products read the stored unit diagonal
static int solve_cont(int hole, double *dq, double lo, double hi, double *dmax) {
    const int N = G.N;
    hipStream_t s = G.st;
    DevView d = eng_view();
    if (G.tun.np) ph_tun_gather(s, G.tun, hole, G.d_tk, G.d_tg);
    if (G.graph) gr_cont(hole);
    else ph_cont(s, d, hole, G.A, G.d_rhs, G.d_sc, G.d_dsc);
    G.mg_w = G.d_dsc;
    G.mg_qsign = hole ? -1 : 1;
    sd_fill(s, G.d_sol, 1.0, N);
    double atol = CONT_ATOL * G.cont_atol_f;
    int its;
    G.unit_diag = g_unit_diag;            // k_cont writes every diagonal entry as 1
    const double rtol = inner32_on() && G.gummel_r > G.loose_r32 ? G.bicg_rtol32 : G.bicg_rtol;
    double rn = bicgstab(G.d_rhs, G.d_sol, G.max_iter_c > 0 ? G.max_iter_c : 1000, rtol, atol, &its);
    G.unit_diag = 0;
    if (G.verbose > 2) fprintf(stderr, "    bicgstab(%c) its=%d |r|=%.3e\n", hole ? 'p' : 'n', its, rn);
    int ok = std::isfinite(rn);
    if (ok) float_fix(hole);
    ph_cont_update(s, d, hole, G.d_sol, G.d_sc, dq, lo, hi, G.rb, S_MAX1);
    rd(S_MAX1, S_MAX1 + 1);
    *dmax = S(S_MAX1);
    return ok;
}

static void terminal_currents(double *I, double *Sv) {
    if (G.graph) {
        gr_currents();
        rd(S_CUR, S_CUR + 3 * std::max(1, G.n_cons));
        for (int c = 0; c < G.n_cons; c++) {
            I[c] = S(S_CUR + 3 * c);
            Sv[c] = S(S_CUR + 3 * c + 1);
            G.cons[c].In = S(S_CUR + 3 * c + 2);
        }
        return;
    }
    CurEdges e;
    e.k = G.d_ek; e.kb = G.d_ekb; e.h = G.d_eh; e.ar = G.d_ear; e.off = G.d_eoff; e.ncon = G.n_cons;
    e.poff = G.tun.np ? G.d_tp_off : nullptr;
    e.pidx = G.tun.np ? G.d_tp_idx : nullptr;
    e.tun = G.tun;
    ph_currents(G.st, eng_view(), e, G.rb);
    rd(S_CUR, S_CUR + 3 * std::max(1, G.n_cons));
    for (int c = 0; c < G.n_cons; c++) {
        I[c] = S(S_CUR + 3 * c);
        Sv[c] = S(S_CUR + 3 * c + 1);
        G.cons[c].In = S(S_CUR + 3 * c + 2);
    }
}

// ------------------------------------------------------------ Anderson
static int AA_DEPTH = 12;
static void aa_reset() { G.aa_n = 0; G.aa_pos = 0; G.aa_have = 0; }
static void aa_step(double *const x[3], double *const g[3]) {
    const int N = G.N;
    const long L = 3L * N;
    hipStream_t s = G.st;
    DevView d = eng_view();
    int M = G.aa_m < AA_DEPTH ? G.aa_m : AA_DEPTH;
    if (M > 0 && G.aa_have) {
        int sl = G.aa_pos;
        ph_aa_diff(s, d, x[0], x[1], x[2], g[0], g[1], g[2], G.d_aaf, G.d_aag, G.d_aadF[sl], G.d_aadG[sl], 1);
        if (G.aa_n < M) G.aa_n++;
        ph_aa_dots(s, (int)L, G.d_aadF, G.aa_n, G.d_aadF[sl], nullptr, G.rb, S_AA0);
        rd(S_AA0, S_AA0 + G.aa_n);
        for (int i = 0; i < G.aa_n; i++) G.aa_gram[i][sl] = G.aa_gram[sl][i] = S(S_AA0 + i);
        G.aa_pos = (sl + 1) % M;
    } else {
        ph_aa_diff(s, d, x[0], x[1], x[2], g[0], g[1], g[2], G.d_aaf, G.d_aag, nullptr, nullptr, 0);
    }
    G.aa_have = 1;
    int m = G.aa_n;
    if (m == 0) return;
    double A[12][12], bb[12], gam[12];
    ph_aa_dots(s, (int)L, G.d_aadF, m, nullptr, G.d_aaf, G.rb, S_AA0);
    rd(S_AA0, S_AA0 + m);
    for (int i = 0; i < m; i++) { bb[i] = S(S_AA0 + i); for (int j = 0; j < m; j++) A[i][j] = G.aa_gram[i][j]; }
    double tr = 0.;
    for (int i = 0; i < m; i++) tr += A[i][i];
    for (int i = 0; i < m; i++) A[i][i] += 1e-9 * tr + 1e-300;
    for (int c = 0; c < m; c++) {
        int pv = c;
        for (int r = c + 1; r < m; r++) if (fabs(A[r][c]) > fabs(A[pv][c])) pv = r;
        if (pv != c) {
            for (int q = 0; q < m; q++) std::swap(A[c][q], A[pv][q]);
            std::swap(bb[c], bb[pv]);
        }
        for (int r = c + 1; r < m; r++) {
            double f = A[r][c] / A[c][c];
            for (int q = c; q < m; q++) A[r][q] -= f * A[c][q];
            bb[r] -= f * bb[c];
        }
    }
    for (int r = m - 1; r >= 0; r--) {
        double t = bb[r];
        for (int q = r + 1; q < m; q++) t -= A[r][q] * gam[q];
        gam[r] = t / A[r][r];
    }
    for (int i = 0; i < m; i++) if (!(gam[i] == gam[i])) return;
    ph_aa_mix(s, d, g[0], g[1], g[2], G.d_aadG, gam, m);
}

// ------------------------------------------------------------ Gummel
// The outer iteration of a bias point (semiconductor_3d.c:3781-3930): refresh
// the carrier-dependent models, solve both continuity equations for the
// quasi-Fermi potentials, re-solve Poisson from them, mix with Anderson.  The
// stopping rules decide the converged state, so they are the reference's:
// both updates below gtol and every terminal current settled (plus, with
// impact ionisation, the generation series settled); or, once the residual
// has stopped halving for 20 iterations, a state within 1e3 gtol whose
// currents are settled; giving up after 60 iterations without progress.
static double newton_tol(const double *Va) {
    double sc = 1.;
    for (int c = 0; c < G.n_cons; c++) if (fabs(Va[c]) > sc) sc = fabs(Va[c]);
    return 2e-14 + 2e-15 * sc;
}

// The impact-ionisation generation enters each iteration from the previous
// one's carriers, so its total builds up as a geometric series whose ratio
// (the ionisation integral) nears 1 close to breakdown.  `observe` takes the
// total of one iteration and estimates what the series still has to add from
// the ratio of its last two changes (capped at 0.99).
struct AvalancheSeries {
    double total = -1.;         // last total, < 0 before the first
    double change = 0.;         // |last change|
    double remainder = 0.;      // estimated rest of the series
    double rel = 1.;            // remainder / total
    double rel_best = 1e300;    // smallest rel that counted as progress
    bool observe(double g, double tol) {
        if (!(g > 0.)) { remainder = 0.; rel = 0.; return true; }
        bool settled = false;
        if (total < 0.) remainder = g;
        else {
            const double step = fabs(g - total);
            double ratio = change > 0. ? step / change : 0.;
            if (ratio > 0.99) ratio = 0.99;
            remainder = step + step * ratio / (1. - ratio);
            settled = step <= tol * g && remainder <= tol * g;
            change = step;
        }
        rel = remainder / g;
        total = g;
        return settled;
    }
};

// Terminal currents of successive iterations: settled when no contact's
// current moved by more than rtol of itself or its round-off floor 3e-14 S.
struct CurrentTrace {
    int nc = 0, have = 0;
    double I[SD_NC_MAX], S[SD_NC_MAX], last[SD_NC_MAX];
    int settled(double rtol) const {
        if (!have) return 0;
        for (int c = 0; c < nc; c++)
            if (fabs(I[c] - last[c]) > fmax(rtol * fabs(I[c]), 3e-14 * S[c])) return 0;
        return 1;
    }
    double flux_max() const {
        double m = 0.;
        for (int c = 0; c < nc; c++) if (S[c] > m) m = S[c];
        return m;
    }
    void roll() { memcpy(last, I, sizeof(double) * nc); have = 1; }
};

// Progress of the outer iteration: the residual halving, or the avalanche
// remainder halving relative to its total.
struct OuterProgress {
    double halved = 1e300;      // residual at the last halving
    double lowest = 1e300;      // lowest residual (Anderson restart reference)
    int at = 0;                 // iteration of the last progress
    void note(int it, double r, AvalancheSeries *ii) {
        if (r < 0.5 * halved) { halved = r; at = it; }
        if (ii && ii->rel < 0.5 * ii->rel_best) { ii->rel_best = ii->rel; at = it; }
    }
    int idle(int it) const { return it - at; }
};

// models that follow the carriers: clamp the quasi-Fermi levels, keep the
// iteration's starting values, carriers, mobilities, tunnelling; returns the
// total impact-ionisation generation (0 without the model)
static double outer_refresh(int avalanche) {
    const int N = G.N;
    hipStream_t s = G.st;
    eng_clamp_qf();
    sd_copy(s, G.d_qn0, G.d_phin, N);
    sd_copy(s, G.d_qp0, G.d_phip, N);
    eng_rebuild_np();
    eng_upd_mob();
    if (G.tun.np) ph_tun_update(s, eng_view(), G.tun);
    if (!avalanche) return 0.;
    ph_ii_update(s, eng_view());
    sd_dot(s, G.d_gii, nullptr, N, G.rb, S_TMP1);              // sum of the generation
    rd(S_TMP1, S_TMP1 + 1);
    return S(S_TMP1);
}

// Poisson from the new quasi-Fermi levels, starting from the neutral-region
// prediction; the previous potential stays in d_xk.  Returns max |dpsi|.
static double outer_potential(double ntol) {
    const int N = G.N;
    hipStream_t s = G.st;
    sd_copy(s, G.d_xk, G.d_psi, N);
    ph_neutral_pred(s, eng_view(), G.d_dqn, G.d_dqp);
    double t = wtime();
    poisson_newton(ntol, 100);
    G.t_pois += wtime() - t;
    sd_maxabsdiff(s, G.d_psi, G.d_xk, N, G.rb, S_MAX2);
    rd(S_MAX2, S_MAX2 + 1);
    return S(S_MAX2);
}

static int gummel(const double *Va, int maxit, double gtol, int *iters, double *res) {
    const int N = G.N;
    const double nt0 = newton_tol(Va), lo = G.vlo, hi = G.vhi;
    const double itol = G.tol_c > 0. ? G.tol_c : 1e-6;
    const int avalanche = (G.models & 256) != 0;
    const double ii_tol = fmax(itol, 1e-4);
    // with impact ionisation the continuity solves run at 1e-8 for this point
    const double brtol_saved = G.bicg_rtol, brtol32_saved = G.bicg_rtol32;
    if (avalanche && G.bicg_rtol > 1e-8) G.bicg_rtol = 1e-8;
    if (avalanche && G.bicg_rtol32 > 1e-8) G.bicg_rtol32 = 1e-8;
    AvalancheSeries ii;
    CurrentTrace cur;
    cur.nc = G.n_cons;
    OuterProgress prog;
    G.ii_err = 0.;
    apply_bc(Va);
    aa_reset();
    poisson_newton(fmax(nt0, 1e-6), 100);      // psi consistent with the predicted quasi-Fermi levels
    int it, conv = 0;
    double dpsi = 1., dphi = 1., r_prev = 1.;
    G.gummel_r = 1.;
    for (it = 1; it <= maxit && !g_stop; it++) {
        double t0 = wtime();
        const double gen = outer_refresh(avalanche);
        int ii_ok = avalanche ? ii.observe(gen, ii_tol) : 1;
        double t1 = wtime();
        G.t_mob += t1 - t0;
        double dn = 0., dp = 0.;
        if (!solve_cont(0, G.d_dqn, lo, hi, &dn) || !solve_cont(1, G.d_dqp, lo, hi, &dp)) break;
        double t2 = wtime();
        G.t_cont += t2 - t1;
        dphi = fmax(dn, dp);
        terminal_currents(cur.I, cur.S);
        G.t_cur += wtime() - t2;
        const int I_ok = cur.settled(itol);
        // a remainder below the currents' round-off floor leaves nothing to converge
        if (!ii_ok && SD_Q * ii.remainder <= 3e-14 * cur.flux_max()) ii_ok = 1;
        cur.roll();
        dpsi = outer_potential(fmax(nt0, G.ntol_fac * fmin(r_prev, 1.)));
        if (G.verbose > 1)
            fprintf(stderr, "semisim-hip:   outer %d: max|dpsi| %.3e max|dphi| %.3e I[0] %.6e (Anderson %d)\n", it, dpsi,
                    dphi, G.n_cons ? cur.I[0] : 0., G.aa_n);
        if (dpsi != dpsi || dphi != dphi) break;
        const double r = fmax(dpsi, dphi);
        g_prog[1] = it; g_prog[2] = r; g_prog[5] = (double)(G.gummel_its + it);
        const int settled = I_ok && ii_ok;
        if (dphi < gtol && dpsi < gtol && settled) {
            sd_copy(G.st, G.d_psi, G.d_xk, N);       // psi of this iteration's carriers
            conv = 1;
            break;
        }
        r_prev = r;
        G.gummel_r = r;
        prog.note(it, r, avalanche ? &ii : nullptr);
        if (prog.idle(it) > 20 && r <= 1e3 * gtol && settled) {
            sd_copy(G.st, G.d_psi, G.d_xk, N);
            conv = 1;
            if (G.verbose)
                fprintf(stderr, "semisim-hip: residual %.2e stalled within 1e3 x tolerance, currents settled: accepted\n", r);
            break;
        }
        if (prog.idle(it) > 60) break;
        if (r < prog.lowest) prog.lowest = r;
        else if (r > 30. * prog.lowest) aa_reset();  // the mixing diverged: drop its history
        double t4 = wtime();
        double *xx[3] = {G.d_xk, G.d_qn0, G.d_qp0}, *gg[3] = {G.d_psi, G.d_phin, G.d_phip};
        aa_step(xx, gg);
        G.t_aa += wtime() - t4;
    }
    if (it > maxit) it = maxit;
    G.gummel_r = 0.;                 // continuity solves outside the loop use bicg_rtol
    if (avalanche && brtol_saved > 1e-8 && G.bicg_rtol == 1e-8) G.bicg_rtol = brtol_saved;
    if (avalanche && brtol32_saved > 1e-8 && G.bicg_rtol32 == 1e-8) G.bicg_rtol32 = brtol32_saved;
    if (avalanche) G.ii_err = SD_Q * ii.remainder;
    G.gummel_its += it;
    *iters = it;
    *res = fmax(dpsi, dphi);
    return conv;
}

// ------------------------------------------------------------ continuation
static void save_state(double **h) {
    sd_copy(G.st, h[0], G.d_psi, G.N); sd_copy(G.st, h[1], G.d_phin, G.N); sd_copy(G.st, h[2], G.d_phip, G.N);
}
static void load_state(double **h) {
    sd_copy(G.st, G.d_psi, h[0], G.N); sd_copy(G.st, G.d_phin, h[1], G.N); sd_copy(G.st, G.d_phip, h[2], G.N);
}

// quasi-neutral regions of ohmic contacts move rigidly with their voltage (host BFS)
static void region_shift(const double *dV) {
    if (G.graph) { gr_region_shift(dV); return; }
    const int N = G.N;
    bool any = false;
    for (int c = 0; c < G.n_cons; c++) if (G.cons[c].bc == 0 && fabs(dV[c]) >= 1e-15) any = true;
    if (!any) return;
    std::vector<double> n(N), p(N), psi(N), qn(N), qp(N);
    eng_download(n.data(), G.d_n, N); eng_download(p.data(), G.d_p, N);
    eng_download(psi.data(), G.d_psi, N); eng_download(qn.data(), G.d_phin, N); eng_download(qp.data(), G.d_phip, N);
    std::vector<unsigned char> mark(N, 0);
    std::vector<int> q(N);
    for (int c = 0; c < G.n_cons; c++) {
        if (G.cons[c].bc != 0 || fabs(dV[c]) < 1e-15) continue;
        int head = 0, tail = 0, type = 0;
        for (int k = 0; k < N; k++)
            if (G.con_mask[k] == c + 1 && G.kind[k] == K_SEMI && !mark[k]) {
                mark[k] = (unsigned char)(c + 1);
                q[tail++] = k;
                if (!type) { double D = G.Nd[k] - G.Na[k]; type = D > 0. ? 1 : (D < 0. ? -1 : 0); }
            }
        if (!type) continue;
        while (head < tail) {
            int k = q[head++], i, j, l;
            host_unidx(k, i, j, l);
            for (int dir = 0; dir < 6; dir++) {
                int kb; double h, ar;
                if (!host_nbr(i, j, l, k, dir, kb, h, ar) || mark[kb] || G.kind[kb] != K_SEMI) continue;
                if (G.con_mask[kb] && G.con_mask[kb] != c + 1) continue;
                double D = G.Nd[kb] - G.Na[kb];
                if (D * type <= 0.) continue;
                double maj = type > 0 ? n[kb] : p[kb];
                if (maj < 0.5 * fabs(D)) continue;
                mark[kb] = (unsigned char)(c + 1);
                q[tail++] = kb;
            }
        }
        for (int t = 0; t < tail; t++) {
            int k = q[t];
            if (G.con_mask[k]) continue;
            psi[k] += dV[c]; qn[k] += dV[c]; qp[k] += dV[c];
        }
    }
    eng_upload(G.d_psi, psi.data(), N); eng_upload(G.d_phin, qn.data(), N); eng_upload(G.d_phip, qp.data(), N);
}

int eng_equilibrium() {
    double Va[SD_NC_MAX];
    G.e_n = 0;
    for (int c = 0; c < G.n_cons; c++) Va[c] = G.cons[c].bc == 2 ? G.cons[c].V : 0.;
    DevView d = eng_view();
    eng_upload_cons(Va);
    ph_init_eq(G.st, d);
    eng_set_vrange(Va);
    eng_apply_bc();
    g_prog[4] = 1.;
    int ok = poisson_newton(newton_tol(Va), 300);
    eng_rebuild_np();
    eng_upd_mob();
    for (int c = 0; c < G.n_cons; c++) G.cons[c].Vprev = Va[c];
    G.have_solution = 1;
    if (G.verbose) fprintf(stderr, "semisim-hip: equilibrium %s (newton its %lld)\n", ok ? "ok" : "NOT converged", G.newton_its);
    return ok;
}

static int PRED_ON = 1;
static double GROW_SLOW = 2.2;
static double gummel_tol() {
    double g = G.tol_p;
    if (!(g > 0.)) g = 1e-9;
    if (g > 1e-4) g = 1e-4;
    if (g < 1e-13) g = 1e-13;
    return g;
}

// ------------------------------------------------------------ learned warm start
// The plain continuation's predicted state of a bias point (secant or region
// shift, boundary values applied) is the network's input; its output is the
// correction toward the converged state.  Mode 1 records (prediction,
// converged state) pairs as training samples; mode 2 applies the correction
// and keeps it only when it lowers the Poisson residual of the starting
// state below half of the plain prediction's (otherwise the plain prediction
// is restored).
static void ws_save_pred() {
    if (!G.d_ws[0])
        for (int a = 0; a < 3; a++) {
            G.d_ws[a] = dalloc<double>(G.N);
            if (!G.d_ws[a]) return;
        }
    sd_copy(G.st, G.d_ws[0], G.d_psi, G.N);
    sd_copy(G.st, G.d_ws[1], G.d_phin, G.N);
    sd_copy(G.st, G.d_ws[2], G.d_phip, G.N);
    G.ws_have = 1;
}
static void ws_before(const double *Va, double dvn) {
    if (!nn_ready()) return;
    hipStream_t s = G.st;
    const int Gg = nn_grid();
    ws_save_pred();
    if (!G.ws_have) return;
    // d_h1: the last converged state (saved at the start of the bias point)
    nn_mesh_to_grid(s, Gg, G.Nx, G.Ny, G.Nz, G.d_psi, G.d_phin, G.d_phip, G.d_Nd, G.d_Na, (float)dvn, G.d_h1[0],
                    G.d_h1[1], G.d_h1[2], NN_CIN, nn_input_buf());
    if (G.ws_mode < 2) return;
    G.ws_stat[0]++;
    double r0 = poisson_eval(G.d_psi, 0, G.d_Ft);
    const float *y = nn_infer(s);
    nn_grid_to_mesh_add(s, Gg, G.Nx, G.Ny, G.Nz, y, NN_COUT, nn_target_scale(), G.d_kind, G.d_con_mask, G.d_psi,
                        G.d_phin, G.d_phip, G.ws_chmask, (float)G.ws_thr);
    eng_clamp_qf();
    apply_bc(Va);
    eng_rebuild_np();
    double r1 = poisson_eval(G.d_psi, 0, G.d_Ft);
    if (G.verbose > 1) fprintf(stderr, "semisim-hip: warm start residual %.3e (plain %.3e)\n", r1, r0);
    if (r1 < G.ws_accept * r0 || G.ws_mode == 3) { G.ws_stat[1]++; return; }
    G.ws_stat[2]++;
    sd_copy(s, G.d_psi, G.d_ws[0], G.N);
    sd_copy(s, G.d_phin, G.d_ws[1], G.N);
    sd_copy(s, G.d_phip, G.d_ws[2], G.N);
    eng_rebuild_np();
}
static void ws_fill_target(hipStream_t s, float *tgt, void *) {
    nn_target_to_grid(s, nn_grid(), G.Nx, G.Ny, G.Nz, G.d_ws[0], G.d_ws[1], G.d_ws[2], G.d_psi, G.d_phin, G.d_phip,
                      1.0f / nn_target_scale(), tgt);
}
static void ws_capture() {
    if (!nn_ready()) return;
    if (nn_capture(G.st, ws_fill_target, nullptr)) G.ws_stat[3]++;
}

// The sweep secant predictor (semiconductor_3d.c:4045-4064).  When this bias
// change continues the previous one along the same line (the next point of a
// sweep, or a secant step of a floating-terminal search, r < 0), the first
// sub-step starts from x1 + r (x1 - x2), x1 and x2 the states after and
// before the previous point, instead of from the neutral-region shift.
// Returns r, or 0 when the change does not continue the last one.
static double sweep_ratio(const double *from, const double *to, double span) {
    const int nc = G.n_cons;
    if (!PRED_ON || G.e_n < 1 || !(span > 0.)) return 0.;
    double ab = 0., bb = 0.;
    for (int c = 0; c < nc; c++) {
        if (fabs(from[c] - G.Ve1[c]) > 1e-12) return 0.;     // the last point did not end here
        const double a = to[c] - from[c], b = from[c] - G.Ve2[c];
        ab += a * b;
        bb += b * b;
    }
    if (!(bb > 0.)) return 0.;
    const double r = ab / bb;
    double off = 0.;                                          // distance from the line
    for (int c = 0; c < nc; c++) off = fmax(off, fabs((to[c] - from[c]) - r * (from[c] - G.Ve2[c])));
    return (r >= -2. && r <= 3. && fabs(r) > 1e-12 && off <= 1e-9 + 1e-6 * span) ? r : 0.;
}

// quadratic form of the predictor (optional): the last two sweep steps were
// equal, so the last three converged states are equally spaced
static int sweep_is_even(const double *from) {
    if (!G.pred2 || G.e_n < 3 || !G.e3_ok) return 0;
    for (int c = 0; c < G.n_cons; c++) {
        const double s1 = from[c] - G.Ve2[c], s2 = G.Ve2[c] - G.Ve3[c];
        if (fabs(s1 - s2) > 1e-9 + 1e-6 * fabs(s1)) return 0;
    }
    return 1;
}

// sub-step length factor after a converged sub-step of `its` outer
// iterations; a slow step does not shorten the next one (the iteration count
// follows the Gummel contraction rate there, not the step)
static double step_growth(int its) { return its <= 8 ? 2.0 : (its <= 25 ? 1.5 : GROW_SLOW); }

// One bias ramp from the voltages of the last solution (Vprev) to the
// requested ones (V), all contacts together, in adaptive sub-steps t in (0, 1]
// (semiconductor_3d.c:4018-4129).
struct BiasRamp {
    int nc = 0;
    double from[SD_NC_MAX], to[SD_NC_MAX];
    double span = 0.;                       // largest |V - Vprev|
    double secant = 0.;                     // sweep predictor ratio, 0 when off
    int quadratic = 0;
    void bias_at(double t, double *v) const {
        for (int c = 0; c < nc; c++) v[c] = from[c] + t * (to[c] - from[c]);
    }
    // starting state of the sub-step t -> t_next: extrapolation of the last
    // two converged sub-steps, else the sweep predictor, else the last state
    // with the quasi-neutral contact regions shifted by the bias change
    void start(double t, double t_prev, double t_next) const {
        const int N = G.N;
        hipStream_t s = G.st;
        if (t_prev >= 0.) {
            ph_extrap(s, N, G.d_psi, G.d_phin, G.d_phip, G.d_h1[0], G.d_h1[1], G.d_h1[2], G.d_h2[0], G.d_h2[1],
                      G.d_h2[2], (t_next - t) / (t - t_prev));
        } else if (secant != 0. && quadratic) {
            double *dst[3] = {G.d_psi, G.d_phin, G.d_phip};
            const double *x1[3] = {G.d_h1[0], G.d_h1[1], G.d_h1[2]}, *x2[3] = {G.d_e2[0], G.d_e2[1], G.d_e2[2]};
            const double *x3[3] = {G.d_e3[0], G.d_e3[1], G.d_e3[2]};
            ph_extrap2(s, N, dst, x1, x2, x3, t_next * secant);
        } else if (secant != 0.) {
            ph_extrap(s, N, G.d_psi, G.d_phin, G.d_phip, G.d_h1[0], G.d_h1[1], G.d_h1[2], G.d_e2[0], G.d_e2[1],
                      G.d_e2[2], t_next * secant);
        } else {
            load_state(G.d_h1);
            if (span > 0.) {
                double dV[SD_NC_MAX];
                for (int c = 0; c < nc; c++) dV[c] = (t_next - t) * (to[c] - from[c]);
                region_shift(dV);
            }
        }
    }
};

// the sweep history (x2 of the predictor) takes this point's start state;
// with the quadratic predictor the previous x2 moves to x3
static void sweep_history_push() {
    if (G.pred2) {
        if (!G.d_e3[0])
            for (int a = 0; a < 3; a++) G.d_e3[a] = dalloc<double>(G.N);
        for (int a = 0; a < 3; a++) std::swap(G.d_e3[a], G.d_e2[a]);
        G.e3_ok = G.e_n >= 1;
    }
    for (int a = 0; a < 3; a++) sd_copy(G.st, G.d_e2[a], G.d_h1[a], G.N);
}

static int cont_path() {
    BiasRamp R;
    R.nc = G.n_cons;
    for (int c = 0; c < R.nc; c++) {
        R.from[c] = G.cons[c].Vprev;
        R.to[c] = G.cons[c].V;
        R.span = fmax(R.span, fabs(R.to[c] - R.from[c]));
    }
    const int max_outer = G.max_gummel > 0 ? G.max_gummel : 100;
    const double gtol = gummel_tol();
    eng_upd_mob();
    save_state(G.d_h1);
    R.secant = sweep_ratio(R.from, R.to, R.span);
    R.quadratic = R.secant != 0. && sweep_is_even(R.from);
    // bias step of the contact that moves most, in units of 50 mV (warm-start input)
    double step50 = 0.;
    for (int c = 0; c < R.nc; c++)
        if (R.span > 0. && fabs(R.to[c] - R.from[c]) >= R.span) step50 = (R.to[c] - R.from[c]) / 0.05;
    // with impact ionisation a failure near breakdown is physical: three in a
    // row end the ramp instead of shrinking the step to 2e-5
    const int fail_limit = (G.models & 256) ? 3 : (1 << 30);
    double t = 0., t_prev = -1., dt = R.span > 0. ? fmin(1., 0.1 / R.span) : 1.;
    int reached = 0, its = 0, failures = 0, in_row = 0, steps = 0;
    double res = 0., v[SD_NC_MAX];
    g_prog[0] = 0.; g_prog[3] = 0.;
    for (int attempt = 0; attempt < 5000 && !g_stop; attempt++) {
        const double t_next = fmin(1., t + dt);
        R.bias_at(t_next, v);
        eng_set_vrange(v);
        R.start(t, t_prev, t_next);
        if (attempt == 0) sweep_history_push();
        eng_clamp_qf();
        apply_bc(v);
        eng_rebuild_np();
        if (attempt == 0 && t_prev < 0. && R.span > 0. && G.ws_mode) ws_before(v, step50);
        const int ok = gummel(v, max_outer, gtol, &its, &res);
        G.conv_iter = its;
        G.conv_res = res;
        if (G.verbose)
            fprintf(stderr, "semisim-hip: sub-step to t = %.4f (V[0] = %.4f) %s, %d outer iterations, residual %.2e\n",
                    t_next, R.nc ? v[0] : 0., ok ? "converged" : "failed", its, res);
        if (!ok) {
            failures++;
            in_row++;
            R.secant = 0.;                              // retry from the neutral-region shift
            if (R.span == 0. || dt * R.span < 2e-5 || in_row >= fail_limit) break;
            dt *= 0.3;
            continue;
        }
        for (int a = 0; a < 3; a++) std::swap(G.d_h1[a], G.d_h2[a]);   // x2 <- x1, x1 <- this state
        save_state(G.d_h1);
        t_prev = t;
        t = t_next;
        steps++;
        g_prog[0] = t; g_prog[3] = steps;
        for (int c = 0; c < R.nc; c++) G.cons[c].Vprev = v[c];
        if (t >= 1.) { reached = 1; break; }
        dt *= step_growth(its);
        in_row = 0;
    }
    // a bias point reached in one sub-step from the saved prediction is a training sample
    if (reached && R.span > 0. && G.ws_mode == 1 && steps == 1 && failures == 0 && G.ws_have) ws_capture();
    G.ws_have = 0;
    if (reached && R.span > 0.) {
        for (int c = 0; c < R.nc; c++) { G.Ve3[c] = G.Ve2[c]; G.Ve2[c] = R.from[c]; G.Ve1[c] = R.to[c]; }
        G.e_n = G.e_n < 3 ? G.e_n + 1 : 3;
    } else if (!reached) {
        G.e_n = 0;
        G.e3_ok = 0;
    }
    if (!reached && (R.span > 0. || g_stop)) {          // keep the last converged sub-step
        load_state(G.d_h1);
        for (int c = 0; c < R.nc; c++) v[c] = G.cons[c].Vprev;
        eng_set_vrange(v);
        apply_bc(v);
        eng_rebuild_np();
        eng_upd_mob();
        if (G.verbose) fprintf(stderr, "semisim-hip: bias ramp ended at t = %.4f\n", t);
    }
    return reached;
}

// Terminal currents and their floors (semiconductor_3d.c:4151-4171).  The
// flux out of a contact is a small difference of large drift and diffusion
// terms, so its round-off floor is 3e-14 of the flux magnitudes S.  A carrier
// contact whose own floor exceeds the other terminals' together reports minus
// their sum (Kirchhoff) with their floor instead; every floor widens when the
// currents fail to sum to zero within the floors; a converged point with
// impact ionisation adds its series remainder.
static void report_currents(int ok) {
    const int nc = G.n_cons;
    double I[SD_NC_MAX], Sv[SD_NC_MAX], own[SD_NC_MAX];
    terminal_currents(I, Sv);
    double net = 0., floors = 0.;
    int carriers = 0;
    for (int c = 0; c < nc; c++) {
        net += I[c];
        own[c] = 3e-14 * Sv[c];
        floors += own[c];
        if (G.cons[c].bc != 2) carriers++;
    }
    const double widen = floors > 0. ? fmax(1., 2. * fabs(net) / floors) : 1.;
    for (int c = 0; c < nc; c++) {
        double Ic = I[c], fc = own[c];
        const double others = floors - own[c];
        if (G.cons[c].bc != 2 && carriers > 1 && others < fc) { Ic = -(net - I[c]); fc = others; }
        G.cons[c].I = Ic;
        G.cons[c].Iflr = fc * widen + (ok ? G.ii_err : 0.);
        G.cons[c].Iraw = I[c];
    }
}

// When gates and carrier terminals both change, the gates ramp first with
// the terminals held, then the terminals (semiconductor_3d.c:4135-4150).
int eng_run_continuation() {
    const int nc = G.n_cons;
    int gates = 0, terminals = 0;
    for (int c = 0; c < nc; c++) {
        if (fabs(G.cons[c].V - G.cons[c].Vprev) <= 1e-12) continue;
        if (G.cons[c].bc == 2) gates = 1; else terminals = 1;
    }
    int ok;
    if (gates && terminals) {
        double target[SD_NC_MAX];
        for (int c = 0; c < nc; c++) {
            target[c] = G.cons[c].V;
            if (G.cons[c].bc != 2) G.cons[c].V = G.cons[c].Vprev;
        }
        g_prog[4] = 2.;
        ok = cont_path();
        for (int c = 0; c < nc; c++) G.cons[c].V = target[c];
        g_prog[4] = 3.;
        ok = ok && cont_path();
    } else {
        g_prog[4] = gates ? 2. : 3.;
        ok = cont_path();
    }
    report_currents(ok);
    return ok;
}

void eng_display() {
    if (G.graph) { gr_display(); eng_sync(); return; }
    DevView d = eng_view();
    ph_upd_J(G.st, d, G.d_Jx, G.d_Jy, G.d_Jz, G.d_Rt);
    ph_mob_display(G.st, d, G.d_mudn, G.d_mudp);
    eng_sync();
}

// environment knobs (read by api3_init, semiconductor_3d.c:4578-4596)
void eng_read_env() {
    const char *e;
    // inner (inexact Newton / Gummel) Krylov tolerances, the reference's
    // defaults and knobs (PCG 1e-3, BiCGSTAB 1e-4).  Looser ones (3e-3 /
    // 3e-4) measured 8-10% faster on the 3.1M-node MOSFET sweep, but left the
    // floating guard ring of a reverse-biased diode 1.5 current floors from
    // the reference in the FP32-multigrid method
    G.pcg_rtol = (e = getenv("SEMISIM_PCGRTOL")) ? atof(e) : 1e-3;
    G.ntol_fac = (e = getenv("SEMISIM_NTOLF")) ? atof(e) : 0.1;
    G.bicg_rtol = (e = getenv("SEMISIM_BRTOL")) ? atof(e) : 1e-4;
    G.pcg_rtol32 = (e = getenv("SEMISIM_PCGRTOL32")) ? atof(e) : 1e-2;
    G.bicg_rtol32 = (e = getenv("SEMISIM_BRTOL32")) ? atof(e) : 1e-3;
    G.loose_r32 = (e = getenv("SEMISIM_LOOSE_R32")) ? atof(e) : 1e-6;
    G.verbose = (e = getenv("SEMISIM_VERBOSE")) ? atoi(e) : 0;
    if ((e = getenv("SEMISIM_AAM"))) { int q = atoi(e); AA_DEPTH = q < 0 ? 0 : (q > 12 ? 12 : q); }
    if ((e = getenv("SEMISIM_PRED"))) PRED_ON = atoi(e) != 0;
    G.pred2 = (e = getenv("SEMISIM_PRED2")) ? atoi(e) != 0 : 0;
    G.e3_ok = 0;
    if ((e = getenv("SEMISIM_GROW"))) { double g = atof(e); if (g >= 1. && g <= 4.) GROW_SLOW = g; }
    if ((e = getenv("SEMISIM_CATOL"))) CONT_ATOL = atof(e);
    if ((e = getenv("SEMISIM_MG_FP32"))) G.precision = atoi(e) ? 1 : 0;
    if ((e = getenv("SEMISIM_MG_LINELEVELS"))) G.line_levels = atoi(e);
    if ((e = getenv("SEMISIM_MG_TAILN"))) G.tail_n = atoi(e);
    if ((e = getenv("SEMISIM_MG_TAILSW"))) { int q = atoi(e); G.tail_sweeps = q > 0 ? q : 1; }
    if ((e = getenv("SEMISIM_MG_COARSEAXES"))) { int q = atoi(e); G.coarse_axes = q > 0 ? q : 1; }
    if ((e = getenv("SEMISIM_MG_ALPHA_P"))) { double q = atof(e); G.mg_alpha_p = q > 0. ? q : 1.; }
    if ((e = getenv("SEMISIM_MG_ALPHA"))) { double q = atof(e); G.mg_alpha = q > 0. ? q : 1.; }
    G.inner32 = (e = getenv("SEMISIM_INNER32")) ? atoi(e) != 0 : 1;
    G.mg_perm = (e = getenv("SEMISIM_MG_PERM")) ? atoi(e) != 0 : 1;
    G.mg_bf16 = (e = getenv("SEMISIM_MG_BF16")) ? atoi(e) != 0 : G.mg_bf16;
    G.mg_bf16_l1 = (e = getenv("SEMISIM_MG_BF16_L1")) ? atoi(e) != 0 : 1;
}
int eng_aa_depth() { return G.aa_m < AA_DEPTH ? G.aa_m : AA_DEPTH; }
void eng_set_predictor(int on) { PRED_ON = on != 0; }
int eng_set_precision(int p) {
    // a graph device keeps its FP64 aggregation multigrid in either setting
    if (p == 1 && G.allocated && !G.graph && !ensure_h32()) return 0;
    G.precision = p ? 1 : 0;
    return 1;
}
int eng_prepare_precision() {
    if (G.precision == 1 && !G.graph) return ensure_h32();
    return 1;
}

// ------------------------------------------------ kernel-level test hooks
// (declared in include/semisim_hip_ext.h; operate on the prepared device)
#include "../../include/semisim_hip_ext.h"
template <class T> static void level_op_T(Hier<T> &H, int L, double *A7) {
    MGLevel<T> &v = H.lev[L];
    size_t n = v.N;
    std::vector<T> tmp(n);
    std::vector<double> rs(n, 1.0);
    if (L > 0) eng_download(rs.data(), v.rs, n);
    T *a[7] = {v.A.c, v.A.xm, v.A.xp, v.A.ym, v.A.yp, v.A.zm, v.A.zp};
    // storage order -> natural order
    std::vector<size_t> pos(n);
    for (size_t k = 0; k < n; k++) {
        int i = (int)(k % v.Nx), j = (int)((k / v.Nx) % v.Ny), l = (int)(k / ((size_t)v.Nx * v.Ny));
        pos[k] = (size_t)mg_idx(v, i, j, l);
    }
    for (int q = 0; q < 7; q++) {
        if (!a[q]) { for (size_t k = 0; k < n; k++) A7[q * n + k] = rs[pos[k]]; continue; }
        HC(hipMemcpyAsync(tmp.data(), a[q], n * sizeof(T), hipMemcpyDeviceToHost, G.st));
        eng_sync();
        for (size_t k = 0; k < n; k++) A7[q * n + k] = (double)tmp[pos[k]] * rs[pos[k]];
    }
}
extern "C" {
int sdx_test_prepare(void) { return eng_prepare() && eng_prepare_precision(); }
// Kernel microbenchmark on the current device and multigrid hierarchy (after
// a solve): `iters` back-to-back launches timed with HIP events on the solver
// stream.  what: 0/1/2 line sweep along x/y/z (level 0, one colour), 3 point
// red-black sweep, 4 Krylov SpMV+dot, 5 residual+restriction level 0->1,
// 6 full V-cycle, 7 copy of N doubles, 8 Poisson assembly, 9 continuity
// assembly.  out[0] = ms per launch.
double sdx_microbench(int what, int iters) {
    if (!G.allocated) return -1.;
    hipStream_t s = G.st;
    hipEvent_t e0, e1;
    HC(hipEventCreate(&e0)); HC(hipEventCreate(&e1));
    DevView d = eng_view();
    // what 10: copy of 8M doubles (64 MB each way), larger than the summed L2
    // of all XCDs, as the calibration reference of the HBM byte counters
    const long BIG_N = 8L << 20;
    double *big = nullptr;
    if (what == 10) HC(hipMalloc(&big, 2 * BIG_N * sizeof(double)));
    if (what == 10) HC(hipMemsetAsync(big, 0, 2 * BIG_N * sizeof(double), s));
    // the restriction as the V-cycle runs it on level 0: after a line sweep
    // only the rows the last sweep did not solve carry a residual
    const int skip0 = G.n_axes > 0 && G.line_levels > 0 ? 2 + G.mg_axes[G.n_axes - 1] : 1;
    auto one = [&]() {
        if (G.precision == 1) {
            MGLevel<float> &v = G.H32.lev[0];
            switch (what) {
            case 0: case 1: case 2: mg_line(s, v, G.d_b32, G.d_x32, what, 0, 0); break;
            case 3: mg_rbgs(s, v, G.d_b32, G.d_x32, 1, 0); break;
            case 5: mg_resid_restrict(s, v, G.H32.lev[1], G.d_b32, G.d_x32, skip0); break;
            case 6: vcycle_T<float>(0, G.d_b32, G.d_x32); break;
            case 13: case 14: case 15:
                if (G.H32.lev.size() > 1) mg_line(s, G.H32.lev[1], G.H32.lev[1].b, G.H32.lev[1].x, what - 13, 0, 0);
                break;
            default: break;
            }
        } else {
            MGLevel<double> &v = G.H64.lev[0];
            switch (what) {
            case 0: case 1: case 2: mg_line(s, v, G.d_wk[0], G.d_wk[1], what, 0, 0); break;
            case 3: mg_rbgs(s, v, G.d_wk[0], G.d_wk[1], 1, 0); break;
            case 5: mg_resid_restrict(s, v, G.H64.lev[1], G.d_wk[0], G.d_wk[1], skip0); break;
            case 6: vcycle_T<double>(0, G.d_wk[0], G.d_wk[1]); break;
            case 13: case 14: case 15:
                if (G.H64.lev.size() > 1) mg_line(s, G.H64.lev[1], G.H64.lev[1].b, G.H64.lev[1].x, what - 13, 0, 0);
                break;
            default: break;
            }
        }
        switch (what) {
        case 4: sd_spmv_dots(s, d.g, G.A, G.d_wk[2], G.d_wk[3], G.d_wk[4], nullptr, 1, G.rb, S_TMP1, -1); break;
        case 16: sd_spmv_dots(s, d.g, G.A, G.d_wk[2], G.d_wk[3], G.d_wk[4], nullptr, 1, G.rb, S_TMP1, -1, stage0()); break;
        case 7: sd_copy_bw(s, G.d_wk[4], G.d_wk[5], G.N); break;
        case 10: sd_copy_bw(s, big, big + BIG_N, BIG_N); break;
        case 11: precond(G.d_wk[0], G.d_wk[1]); break;
        case 12: sd_dot(s, G.d_wk[4], G.d_wk[4], G.N, G.rb, S_TMP1); rd(S_TMP1, S_TMP1 + 1); break;
        case 8: ph_poisson(s, d, G.d_psi, 1, G.d_wk[5], G.A, G.d_sc, G.rb, S_TMP1); break;
        case 9: ph_cont(s, d, 0, G.A, G.d_wk[5], G.d_wk[6], G.d_dsc); break;
        default: break;
        }
    };
    for (int w = 0; w < 3; w++) one();
    HC(hipStreamSynchronize(s));
    double h0 = wtime();
    HC(hipEventRecord(e0, s));
    for (int it = 0; it < iters; it++) one();
    HC(hipEventRecord(e1, s));
    HC(hipEventSynchronize(e1));
    float ms = 0.f;
    HC(hipEventElapsedTime(&ms, e0, e1));
    if (what == 12) ms = (float)((wtime() - h0) * 1e3);   // host round trip: wall clock
    HC(hipEventDestroy(e0)); HC(hipEventDestroy(e1));
    if (big) HC(hipFree(big));
    return (double)ms / iters;
}
void sdx_tune(int what, int value) {
    if (what == 0) g_line_nl_max = value;
    if (what == 4) g_line_scan = value;
    if (what == 5) g_line_slots = value >= 4 ? 4 : (value >= 2 ? 2 : 1);
    if (what == 7) g_line_minwg = value > 0 ? value : 1;
    if (what == 8) g_tail_lds = value != 0;
    if (what == 9) G.mg_perm = value != 0;
    if (what == 10) g_line_xcd = value != 0;
    if (what == 12) g_restrict_act = value != 0;
    if (what == 13) G.bi_fuse = value != 0;
    if (what == 14) G.stage_on = value != 0;
    if (what == 16) G.mg_alpha = G.mg_alpha_p = value > 0 ? value / 100.0 : 1.0;   // (defaults 1.1 / 1.3)
    if (what == 17) G.mg_bf16 = value != 0;
    if (what == 18) g_line_scan1 = value != 0;
    if (what == 19) g_spmv_nt = value != 0;
    if (what == 20) g_unit_diag = value != 0;
    if (what == 21) g_rb_fuse = value != 0;
    if (what == 22) g_rb_fuse_post = value != 0;
    if (what == 23) G.stage_read = value != 0;
    if (what == 24) g_rec16 = value != 0;
    if (what == 25) g_cont_lds = value != 0;
    if (what == 26) { G.pred2 = value != 0; G.e3_ok = 0; }
    if (what == 28) G.mg_bf16_l1 = value != 0;
    if (what == 27) G.inner32 = value >= 0 && value <= 2 ? value : 1;
    if (what == 11) {
        const int m = value >= 0 && value <= 2 ? value : SD_TILE_MODE_DEFAULT;
        hipDeviceSynchronize();
        sd_set_tile_mode_linalg(m);
        ph_set_tile_mode(m);
    }
    if (what == 1) G.line_levels = value;
    if (what == 2) G.tail_sweeps = value > 0 ? value : 1;
    if (what == 3) {
        G.tail_n = value;
        if (G.H64.alloc) hier_tail(G.H64);
        if (G.H32.alloc) hier_tail(G.H32);
    }
}
// learned warm start: 0 off, 1 record training samples, 2 apply the network
// (kept only if it lowers the starting Poisson residual), 3 apply it always
void sdx_ws_mode(int mode) { G.ws_mode = mode >= 1 && mode <= 3 ? mode : 0; }
// out[0] predictions tried, [1] kept, [2] rejected (plain prediction used), [3] samples recorded
void sdx_ws_stats(double *out) { for (int i = 0; i < 4; i++) out[i] = (double)G.ws_stat[i]; }
void sdx_ws_reset_stats(void) { for (int i = 0; i < 4; i++) G.ws_stat[i] = 0; }
// floating majority blocks since the last reset: [0] continuity solves that
// listed unanchored majority nodes, [1] blocks found, [2] blocks rescaled
void sdx_float_stats(double *out) { for (int i = 0; i < 3; i++) out[i] = (double)G.ff_stat[i]; }
void sdx_float_reset_stats(void) { for (int i = 0; i < 3; i++) G.ff_stat[i] = 0; }
// which of psi (1), phi_n (2), phi_p (4) the correction changes, and the
// smallest correction applied (V)
void sdx_ws_config(int chmask, double thr) { G.ws_chmask = chmask & 7; G.ws_thr = thr > 0. ? thr : 0.; }
// mode 2 keeps a correction when it lowers the starting Poisson residual below
// ratio x the plain prediction's (default 0.5)
void sdx_ws_accept(double ratio) { G.ws_accept = ratio > 0. ? ratio : 0.5; }
void sdx_ktimer_enable(int kind, int every) {
    kt_flush();
    KT.on = KT.armed = 1; KT.kind = kind; KT.every = every > 0 ? every : 8;
    KT.count = KT.sampled = 0; KT.total_ms = 0.; KT.bytes = 0.;
}
// pause (1) or resume (0) the sampling without resetting what it gathered
void sdx_ktimer_pause(int paused) {
    if (paused) { KT.on = 0; KT.sampling = false; }
    else KT.on = KT.armed;
}
// out: [sampled launches, total ms of them, algorithmic bytes per launch, launches seen]
int sdx_ktimer_read(double *out) {
    eng_sync();
    kt_flush();
    double bytes;
    if (KT.kind == 0) {
        // mean over the sampled level-0 sweeps of their algorithmic bytes,
        // each counted when it ran (sweep_bytes: its precision, axis and
        // coefficient storage)
        bytes = KT.sampled > 0 ? KT.bytes / (double)KT.sampled : 0.;
    } else if (KT.kind == 2) {
        bytes = 0.;                              // a whole V-cycle: time only
    } else if (G.graph) {
        // CSR product fused with one dot: row pointers, per directed edge its
        // column and value, per node the diagonal, x, y and the dot operand
        bytes = 4.0 * (G.N + 1) + 12.0 * G.nE + 32.0 * G.N;
    } else {
        bytes = (double)G.N * 11.0 * 8.0;   // 7 diagonals + x + dot vector(s) + y (FP64)
    }
    out[0] = (double)KT.sampled; out[1] = KT.total_ms; out[2] = bytes; out[3] = (double)KT.count;
    KT.on = KT.armed = 0;
    return 1;
}
// majority-node flags of carrier `hole` for the quantities u, sc: before
// (0 none, 1 majority, 2 majority next to a contact) and after the closure;
// returns the propagation rounds
int sdx_test_maj_flags(int hole, const double *u, const double *sc, unsigned char *f0, unsigned char *f1) {
    eng_upload(G.d_sol, u, G.N);
    eng_upload(G.d_sc, sc, G.N);
    ph_maj_flags(G.st, eng_view(), hole, G.d_sol, G.d_sc, G.d_flag);
    HC(hipMemcpyAsync(f0, G.d_flag, G.N, hipMemcpyDeviceToHost, G.st));
    int rounds = maj_closure();
    HC(hipMemcpyAsync(f1, G.d_flag, G.N, hipMemcpyDeviceToHost, G.st));
    eng_sync();
    return rounds;
}
void sdx_test_set_state(const double *psi, const double *phin, const double *phip) {
    eng_upload(G.d_psi, psi, G.N); eng_upload(G.d_phin, phin, G.N); eng_upload(G.d_phip, phip, G.N);
    eng_rebuild_np();
    eng_sync();
}
void sdx_test_get_node(int which, double *out) {
    const double *src[] = {G.d_psi, G.d_phin, G.d_phip, G.d_n, G.d_p, G.d_mun, G.d_mup};
    if (which >= 0 && which < 7) eng_download(out, src[which], G.N);
}
static void get_A7(double *A7) {
    double *a[7] = {G.A.c, G.A.xm, G.A.xp, G.A.ym, G.A.yp, G.A.zm, G.A.zp};
    for (int q = 0; q < 7; q++) eng_download(A7 + (size_t)q * G.N, a[q], G.N);
}
double sdx_test_poisson(int build, const double *Va, double *F, double *A7, double *sc) {
    eng_upload_cons(Va);
    double r = poisson_eval(G.d_psi, build, G.d_rhs);
    eng_download(F, G.d_rhs, G.N);
    if (build) { get_A7(A7); eng_download(sc, G.d_sc, G.N); }
    return r;
}
void sdx_test_upd_mob(double *fmu6) {
    eng_upd_mob();
    eng_download(fmu6, G.d_fmu, 6 * (size_t)G.N);
}
void sdx_test_set_fmu(const double *fmu6) { eng_upload(G.d_fmu, fmu6, 6 * (size_t)G.N); }
void sdx_test_cont(int hole, double *A7, double *rhs, double *sc, double *dsc) {
    if (G.tun.np) ph_tun_gather(G.st, G.tun, hole, G.d_tk, G.d_tg);
    if (G.graph) gr_cont(hole);
    else ph_cont(G.st, eng_view(), hole, G.A, G.d_rhs, G.d_sc, G.d_dsc);
    get_A7(A7);
    eng_download(rhs, G.d_rhs, G.N); eng_download(sc, G.d_sc, G.N); eng_download(dsc, G.d_dsc, G.N);
}
// CSR of the device operator last assembled (Poisson Jacobian or scaled
// continuity rows) into host arrays rowptr (N + 1), col and val (cap
// entries); returns the number of nonzeros (< 0: error)
long long sdx_test_csr(int *rowptr, int *col, double *val, long long cap) {
    const size_t N = (size_t)G.N;
    int *d_rp = nullptr, *d_col = nullptr, *d_s = nullptr;
    double *d_val = nullptr;
    long long nnz = SD_CSR_EDEVICE;
    const size_t c = cap > 0 ? (size_t)cap : 1;
    if (hipMalloc(&d_rp, (N + 1) * sizeof(int)) == hipSuccess && hipMalloc(&d_col, c * sizeof(int)) == hipSuccess &&
        hipMalloc(&d_val, c * sizeof(double)) == hipSuccess &&
        hipMalloc(&d_s, (size_t)csr_scratch_words((long)N) * sizeof(int)) == hipSuccess) {
        const double *diag[7] = {G.A.c, G.A.xm, G.A.xp, G.A.ym, G.A.yp, G.A.zm, G.A.zp};
        nnz = csr_from_stencil(G.st, G.Nx, G.Ny, G.Nz, diag, d_rp, d_col, d_val, cap, d_s);
        if (nnz >= 0) {
            HC(hipMemcpyAsync(rowptr, d_rp, (N + 1) * sizeof(int), hipMemcpyDeviceToHost, G.st));
            const size_t m = (size_t)std::min<long long>(nnz, cap);
            if (m) {
                HC(hipMemcpyAsync(col, d_col, m * sizeof(int), hipMemcpyDeviceToHost, G.st));
                HC(hipMemcpyAsync(val, d_val, m * sizeof(double), hipMemcpyDeviceToHost, G.st));
            }
            eng_sync();
        }
    }
    for (void *p : {(void *)d_rp, (void *)d_col, (void *)d_val, (void *)d_s}) if (p) hipFree(p);
    return nnz;
}
// graph device: the operator last assembled (diagonal, one value per
// directed edge); returns the edge count (-1: not a graph device)
int sdx_test_graph_op(double *dg, double *va) {
    if (!G.graph) return -1;
    eng_download(dg, G.A.c, G.N);
    eng_download(va, G.d_gval, G.nE);
    return G.nE;
}
// load a graph operator (diagonal, edge values) as sdx_test_load_op does
int sdx_test_graph_load_op(const double *dg, const double *va, int qsign, const double *w) {
    if (!G.graph) return -1;
    eng_upload(G.A.c, dg, G.N);
    eng_upload(G.d_gval, va, G.nE);
    if (w) { eng_upload(G.d_dsc, w, G.N); G.mg_w = G.d_dsc; } else G.mg_w = nullptr;
    G.mg_qsign = qsign;
    return G.nE;
}
// edge mobilities of a graph device (electrons then holes per undirected
// edge): read (set = 0) or replace (set = 1); returns the edge count
int sdx_test_graph_gmu(double *gmu, int set) {
    if (!G.graph) return -1;
    if (set) eng_upload(G.d_gmu, gmu, (size_t)2 * G.nUE);
    else eng_download(gmu, G.d_gmu, (size_t)2 * G.nUE);
    return G.nUE;
}
// levels of the graph multigrid (0: not a graph device)
int sdx_test_graph_levels(void) { return G.graph ? gr_nlev() : 0; }
void sdx_test_load_op(const double *A7, int qsign, const double *w) {
    double *a[7] = {G.A.c, G.A.xm, G.A.xp, G.A.ym, G.A.yp, G.A.zm, G.A.zp};
    for (int q = 0; q < 7; q++) eng_upload(a[q], A7 + (size_t)q * G.N, G.N);
    if (w) { eng_upload(G.d_dsc, w, G.N); G.mg_w = G.d_dsc; } else G.mg_w = nullptr;
    G.mg_qsign = qsign;
}
void sdx_test_matvec(const double *x, double *y) {
    eng_upload(G.d_wk[0], x, G.N);
    matvec(G.d_wk[0], G.d_wk[1]);
    eng_download(y, G.d_wk[1], G.N);
}
int sdx_test_mg_setup(void) {
    mg_setup();
    eng_sync();
    return (int)(G.precision == 1 ? G.H32.lev.size() : G.H64.lev.size());
}
void sdx_test_level_dims(int L, int *dims) {
    if (G.precision == 1) { auto &v = G.H32.lev[L]; int d[6] = {v.Nx, v.Ny, v.Nz, v.cx, v.cy, v.cz}; memcpy(dims, d, sizeof d); }
    else { auto &v = G.H64.lev[L]; int d[6] = {v.Nx, v.Ny, v.Nz, v.cx, v.cy, v.cz}; memcpy(dims, d, sizeof d); }
}
void sdx_test_level_op(int L, double *A7) {
    if (G.precision == 1) level_op_T(G.H32, L, A7); else level_op_T(G.H64, L, A7);
}
// every multigrid solution array (each level's x and xs, level-0 FP32
// vectors, the FP32 Krylov vectors) filled with value: the V-cycle must not
// read what a previous use left there
extern "C++" template <class T> static void fill_levels(Hier<T> &H, double value) {
    for (auto &v : H.lev) {
        if constexpr (std::is_same_v<T, float>) {
            if (v.x) k32_fill_launch(G.st, v.x, (float)value, v.N);
            if (v.xs) k32_fill_launch(G.st, v.xs, (float)value, v.N);
        } else {
            if (v.x) sd_fill(G.st, v.x, value, v.N);
            if (v.xs) sd_fill(G.st, v.xs, value, v.N);
        }
    }
}
void sdx_test_fill_mg(double value) {
    fill_levels(G.H32, value);
    fill_levels(G.H64, value);
    if (G.d_x32) k32_fill_launch(G.st, G.d_x32, (float)value, G.N);
    if (G.d_x32b) k32_fill_launch(G.st, G.d_x32b, (float)value, G.N);
    if (G.d_k32) k32_fill_launch(G.st, G.d_k32, (float)value, 8 * G.N);
    eng_sync();
}
void sdx_test_precond(const double *r, double *z) {
    eng_upload(G.d_wk[2], r, G.N);
    precond(G.d_wk[2], G.d_wk[3]);
    eng_download(z, G.d_wk[3], G.N);
}
int sdx_test_pcg(const double *b, double *x, int maxit, double rtol) {
    eng_upload(G.d_rhs, b, G.N); eng_upload(G.d_sol, x, G.N);
    int it = pcg(G.d_rhs, G.d_sol, maxit, rtol);
    eng_download(x, G.d_sol, G.N);
    return it;
}
double sdx_test_bicgstab(const double *b, double *x, int maxit, double rtol, double atol, int *its) {
    eng_upload(G.d_rhs, b, G.N); eng_upload(G.d_sol, x, G.N);
    double rn = bicgstab(G.d_rhs, G.d_sol, maxit, rtol, atol, its);
    eng_download(x, G.d_sol, G.N);
    return rn;
}
void sdx_test_currents(const double *Va, double *I, double *Sv, double *In) {
    eng_upload_cons(Va);
    terminal_currents(I, Sv);
    for (int c = 0; c < G.n_cons; c++) In[c] = G.cons[c].In;
}
// impact-ionisation generation at the present state and face mobilities,
// evaluated as with models bit 8 set (the device's own bits are restored)
void sdx_test_ii_update(double *gii) {
    int m = G.models;
    G.models |= 256;
    ph_ii_update(G.st, eng_view());
    G.models = m;
    eng_download(gii, G.d_gii, G.N);
}
// gate tunnelling paths of the prepared device: returns the number of paths,
// *ncol the number of column entries; arrays (any may be null): per path
// node, contact, first column entry, column length, area; per column entry
// node, cell length, spacing
int sdx_test_tun_paths(int *ncol, int *pk, int *pc, int *pc0, int *pcn, double *pA, int *ck, double *cl, double *ch) {
    if (ncol) *ncol = (int)G.col_k.size();
    auto put = [](auto *dst, const auto &v) { if (dst) std::copy(v.begin(), v.end(), dst); };
    put(pk, G.tp_k); put(pc, G.tp_c); put(pc0, G.tp_c0); put(pcn, G.tp_cn); put(pA, G.tp_A);
    put(ck, G.col_k); put(cl, G.col_l); put(ch, G.col_h);
    return (int)G.tp_k.size();
}
// one tunnelling update at the injected state and contact voltages Va: the
// linearised escape cK (2 per column entry) and injection pG (2 per path),
// then the gathered sink tk and source tg of carrier `hole` (N each)
int sdx_test_tun(const double *Va, int hole, double *cK, double *pG, double *tk, double *tg) {
    eng_upload_cons(Va);
    ph_tun_update(G.st, eng_view(), G.tun);
    ph_tun_gather(G.st, G.tun, hole, G.d_tk, G.d_tg);
    if (G.tun.np) {
        eng_download(cK, G.tun.cK, 2 * G.col_k.size());
        eng_download(pG, G.tun.pG, 2 * G.tp_k.size());
    }
    eng_download(tk, G.d_tk, G.N);
    eng_download(tg, G.d_tg, G.N);
    return G.tun.np;
}
void sdx_test_aa_step(const double *x3, double *g3, int reset) {
    const int N = G.N;
    if (reset) aa_reset();
    for (int a = 0; a < 3; a++) { eng_upload(G.d_h1[a], x3 + (size_t)a * N, N); eng_upload(G.d_h2[a], g3 + (size_t)a * N, N); }
    aa_step(G.d_h1, G.d_h2);
    for (int a = 0; a < 3; a++) eng_download(g3 + (size_t)a * N, G.d_h2[a], N);
}
double sdx_test_cont_update(int hole, const double *u, const double *sc, double *dq, double lo, double hi) {
    eng_upload(G.d_sol, u, G.N); eng_upload(G.d_sc, sc, G.N);
    ph_cont_update(G.st, eng_view(), hole, G.d_sol, G.d_sc, G.d_wk[0], lo, hi, G.rb, S_MAX1);
    rd(S_MAX1, S_MAX1 + 1);
    eng_download(dq, G.d_wk[0], G.N);
    return S(S_MAX1);
}
}
