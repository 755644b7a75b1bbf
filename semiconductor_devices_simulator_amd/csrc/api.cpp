// api.cpp - the api3_* C ABI of the MI355X drift-diffusion core.
//
// Same entry points, signatures, units and return conventions as the
// reference core's public API (semiconductor_3d.c:4567-5353), so the
// reference Python front end loads this library through SEMISIM_SO
// unchanged.  Gate tunnelling (through Robin gate stacks and through meshed
// insulator layers), MLDA quantum confinement, Lombardi surface mobility,
// impact ionisation, the C-V charge partition, prism meshes (the graph path)
// and shaped contact node lists are implemented.  Physics outside this core's
// scope (deep traps, metal-contact thermionic emission, incomplete ionisation,
// the Coulomb term of 4H-SiC) is rejected with a diagnostic and a failure
// return, never approximated.
#include "engine.h"
#include "../../include/semisim_api.h"
#include "../../include/semisim_hip_ext.h"
#include <cstdlib>
#include <cstring>
#include <vector>
#include <algorithm>

void eng_read_env();
int eng_aa_depth();
void eng_set_predictor(int on);
int eng_set_precision(int p);
int eng_prepare_precision();

static void mark_changed() { G.dirty = 1; G.have_solution = 0; }
static int disp_valid = 0;
static void unsupported(const char *what) {
    fprintf(stderr, "semisim-hip: %s is outside this core (not supported)\n", what);
}

extern "C" {

void api3_init(void) {
    eng_free();
    for (int c = 0; c < SD_NC_MAX; c++) { G.con_nodes[c].clear(); G.con_areas[c].clear(); }
    G.n_mats = 0; G.n_cons = 0; G.Nx = G.Ny = G.Nz = G.N = 0;
    G.converged = 0; G.conv_iter = 0; G.conv_res = 0.;
    G.max_iter_p = 2000; G.max_iter_c = 1000; G.max_gummel = 500;
    G.tol_p = 1e-9; G.tol_c = 1e-6; G.omega_p = G.omega_c = 1.;
    G.pcg_its = G.bicg_its = G.newton_its = G.gummel_its = 0;
    G.t_pois = G.t_cont = G.t_mob = G.t_cur = G.t_aa = G.t_mg = 0.;
    G.cont_atol_f = 1.;
    G.models = 7; G.caxis = 1; G.e_n = 0;
    G.precision = 0;
    eng_read_env();
    disp_valid = 0;
}
void api3_set_threads(int n) { G.nthreads = n < 1 ? 1 : n; }
void api3_set_verbose(int v) { G.verbose = v; }

int api3_trap_species(int, int, double, double, double) { unsupported("deep-level traps"); return -1; }
int api3_trap_add(int, int, const int *, const double *) { unsupported("deep-level traps"); return 0; }
void api3_traps_clear(void) {}
int api3_charge_add(int n, const int *nodes, const double *rho) {
    if (!G.allocated) return 0;
    for (int q = 0; q < n; q++) { long k = nodes[q]; if (k >= 0 && k < G.N) G.Nf[k] += rho[q]; }
    mark_changed();
    return 1;
}
int api3_n_trap_entries(void) { return 0; }
void api3_set_models(int flags) { G.models = flags; G.have_solution = 0; G.dirty = 1; }
void api3_set_caxis(int ax) { G.caxis = (ax >= 0 && ax <= 2) ? ax : -1; G.have_solution = 0; }
int api3_get_caxis(void) { return G.caxis; }

int api3_add_material(int id, double T) {
    if (G.n_mats >= SD_NM_MAX) return -1;
    if (!host_mat_init(&G.mats[G.n_mats], id, T)) {
        fprintf(stderr, "semisim-hip: WARNING unknown material id %d, using Si\n", id);
        host_mat_init(&G.mats[G.n_mats], 0, T);
    }
    mark_changed();
    return G.n_mats++;
}
int api3_set_mat_tau(int mi, double tn, double tp) {
    if (mi < 0 || mi >= G.n_mats) return 0;
    G.mats[mi].tau_n0 = tn > 0. ? tn : 1e30;
    G.mats[mi].tau_p0 = tp > 0. ? tp : 1e30;
    mark_changed();
    return 1;
}

static double fv_len(const std::vector<double> &c, int i, int N) {
    if (N < 2) return 1.;
    if (i == 0) return 0.5 * (c[1] - c[0]);
    if (i == N - 1) return 0.5 * (c[N - 1] - c[N - 2]);
    return 0.5 * (c[i + 1] - c[i - 1]);
}
int api3_set_grid_coords(int Nx, int Ny, int Nz, const double *xs, const double *ys, const double *zs) {
    for (int i = 1; i < Nx; i++) if (!(xs[i] > xs[i - 1])) return 0;
    for (int j = 1; j < Ny; j++) if (!(ys[j] > ys[j - 1])) return 0;
    for (int l = 1; l < Nz; l++) if (!(zs[l] > zs[l - 1])) return 0;
    if (Nx < 2 || Ny < 2 || Nz < 2) { eng_free(); return 0; }
    if ((double)Nx * Ny * Nz > 2.0e9) return 0;
    if (!eng_alloc(Nx, Ny, Nz)) return 0;
    G.xs.assign(xs, xs + Nx); G.ys.assign(ys, ys + Ny); G.zs.assign(zs, zs + Nz);
    G.Lx.resize(Nx); G.Ly.resize(Ny); G.Lz.resize(Nz);
    for (int i = 0; i < Nx; i++) G.Lx[i] = fv_len(G.xs, i, Nx);
    for (int j = 0; j < Ny; j++) G.Ly[j] = fv_len(G.ys, j, Ny);
    for (int l = 0; l < Nz; l++) G.Lz[l] = fv_len(G.zs, l, Nz);
    size_t N = (size_t)Nx * Ny * Nz;
    G.mat_idx.assign(N, 0); G.Nd.assign(N, 0.); G.Na.assign(N, 0.); G.Nf.assign(N, 0.);
    if (G.precision == 1 && !eng_prepare_precision()) { eng_free(); return 0; }
    G.dirty = 1;
    disp_valid = 0;
    return 1;
}
void api3_set_grid(int Nx, int Ny, int Nz, double Lx, double Ly, double Lz) {
    if (Nx < 2) Nx = 2; if (Ny < 2) Ny = 2; if (Nz < 2) Nz = 2;
    std::vector<double> xs(Nx), ys(Ny), zs(Nz);
    for (int i = 0; i < Nx; i++) xs[i] = i * Lx / (Nx - 1.);
    for (int j = 0; j < Ny; j++) ys[j] = j * Ly / (Ny - 1.);
    for (int l = 0; l < Nz; l++) zs[l] = l * Lz / (Nz - 1.);
    (void)api3_set_grid_coords(Nx, Ny, Nz, xs.data(), ys.data(), zs.data());
}
void api3_set_solver(int mp, int mc, int mg, double tp, double tc, double op, double oc) {
    G.max_iter_p = mp; G.max_iter_c = mc; G.max_gummel = mg;
    G.tol_p = tp; G.tol_c = tc; G.omega_p = op; G.omega_c = oc;
}
static void clip_box(int *x0, int *y0, int *z0, int *x1, int *y1, int *z1) {
    if (*x0 < 0) *x0 = 0; if (*y0 < 0) *y0 = 0; if (*z0 < 0) *z0 = 0;
    if (*x1 >= G.Nx) *x1 = G.Nx - 1; if (*y1 >= G.Ny) *y1 = G.Ny - 1; if (*z1 >= G.Nz) *z1 = G.Nz - 1;
}
void api3_set_region(int x0, int y0, int z0, int x1, int y1, int z1, int midx, double Nd, double Na) {
    if (!G.allocated || midx < 0 || midx >= G.n_mats) return;
    clip_box(&x0, &y0, &z0, &x1, &y1, &z1);
    for (int l = z0; l <= z1; l++)
        for (int j = y0; j <= y1; j++)
            for (int i = x0; i <= x1; i++) {
                size_t k = i + (size_t)G.Nx * (j + (size_t)G.Ny * l);
                G.mat_idx[k] = midx; G.Nd[k] = Nd; G.Na[k] = Na;
                if (G.mats[midx].insulator) G.Nd[k] = G.Na[k] = 0.;
            }
    mark_changed();
}
void api3_add_fixed_charge(int x0, int y0, int z0, int x1, int y1, int z1, double Nf) {
    if (!G.allocated) return;
    clip_box(&x0, &y0, &z0, &x1, &y1, &z1);
    for (int l = z0; l <= z1; l++)
        for (int j = y0; j <= y1; j++)
            for (int i = x0; i <= x1; i++) G.Nf[i + (size_t)G.Nx * (j + (size_t)G.Ny * l)] += Nf;
    mark_changed();
}
void api3_add_sheet_charge(int axis, int idx, int a0, int a1, int b0, int b1, double sigma) {
    if (!G.allocated) return;
    int x0, x1, y0, y1, z0, z1;
    double L;
    if (axis == 0) { if (idx < 0 || idx >= G.Nx) return; x0 = x1 = idx; y0 = a0; y1 = a1; z0 = b0; z1 = b1; L = G.Lx[idx]; }
    else if (axis == 1) { if (idx < 0 || idx >= G.Ny) return; y0 = y1 = idx; x0 = a0; x1 = a1; z0 = b0; z1 = b1; L = G.Ly[idx]; }
    else { if (idx < 0 || idx >= G.Nz) return; z0 = z1 = idx; x0 = a0; x1 = a1; y0 = b0; y1 = b1; L = G.Lz[idx]; }
    api3_add_fixed_charge(x0, y0, z0, x1, y1, z1, sigma / L);
}
int api3_add_contact_ex(int face, int i0, int i1, int j0, int j1, int k0, int k1, int bc, double V, double pm,
                        double tox, double Nox) {
    if (G.n_cons >= SD_NC_MAX) return -1;
    HostCon &c = G.cons[G.n_cons];
    memset(&c, 0, sizeof c);
    G.con_nodes[G.n_cons].clear();
    G.con_areas[G.n_cons].clear();
    c.face = face < 0 ? 0 : (face > 6 ? 6 : face);
    c.i0 = i0; c.i1 = i1; c.j0 = j0; c.j1 = j1; c.k0 = k0; c.k1 = k1;
    c.bc = bc < 0 ? 0 : (bc > 2 ? 2 : bc);
    c.V = V; c.phi_m = pm;
    c.tox = tox > 0. ? tox : 10e-9;
    c.qox = Nox * SD_Q;
    mark_changed();
    return G.n_cons++;
}
void api3_add_contact(int face, int i0, int i1, int j0, int j1, int bc, double V, double pm) {
    (void)api3_add_contact_ex(face, i0, i1, j0, j1, 0, 0, bc, V, pm, 10e-9, 0.);
}
void api3_clear_contacts(void) {
    for (int c = 0; c < SD_NC_MAX; c++) { G.con_nodes[c].clear(); G.con_areas[c].clear(); }
    G.n_cons = 0;
    mark_changed();
}
int api3_set_gate_stack(int cid, int nl, const int *mat, const double *t) {
    if (cid < 0 || cid >= G.n_cons || nl < 0 || nl > 6) return 0;
    for (int i = 0; i < nl; i++) if (mat[i] < 0 || mat[i] >= G.n_mats || !(t[i] > 0.)) return 0;
    HostCon &cn = G.cons[cid];
    cn.snl = nl;
    for (int i = 0; i < nl; i++) { cn.smat[i] = mat[i]; cn.st[i] = t[i]; }
    mark_changed();
    return 1;
}
// A device on a general box-method mesh (api3_set_mesh_graph :4795-4804): N
// nodes, E directed edges as a symmetric CSR graph (rp[N+1], cj[E]); per
// directed edge its length h and dual face area ar, per node its control
// volume vol and position pos (3N).  Replaces the tensor grid; 0 on a bad
// graph (not symmetric, a self edge, h <= 0, ar < 0, vol <= 0) or no memory.
// The graph's aggregation multigrid leaves smooth continuity errors larger
// than the tensor multigrid does for the same residual, so its BiCGSTAB
// tolerance is 1e-6 unless SEMISIM_BRTOL sets one (as the reference's).
int api3_set_mesh_graph(int N, int E, const int *rp, const int *cj, const double *h, const double *ar,
                        const double *vol, const double *pos) {
    if (!rp || !cj || !h || !ar || !vol || !pos) return 0;
    const int ok = eng_set_graph(N, E, rp, cj, h, ar, vol, pos);
    if (ok && !getenv("SEMISIM_BRTOL")) G.bicg_rtol = 1e-6;
    disp_valid = 0;
    return ok;
}
int api3_is_graph(void) { return G.graph; }
void api3_set_node_data(const int *mat, const double *Nd, const double *Na, const double *Nf) {
    if (!G.allocated) return;
    for (int k = 0; k < G.N; k++) {
        int mi = mat[k];
        if (mi < 0 || mi >= G.n_mats) mi = 0;
        G.mat_idx[k] = mi;
        int ins = G.mats[mi].insulator;
        G.Nd[k] = ins ? 0. : Nd[k];
        G.Na[k] = ins ? 0. : Na[k];
        G.Nf[k] = Nf ? Nf[k] : 0.;
    }
    mark_changed();
}
// nodes of contact cid as an explicit list (a shaped contact on the tensor
// mesh; api3_set_con_nodes :4821-4834): they replace the contact's index box
// wherever prepare visits its nodes; areas = the boundary face area of each
// node for a face contact (ignored, 0, for a box contact); n = 0 restores the box
int api3_set_con_nodes(int cid, int n, const int *nodes, const double *areas) {
    if (cid < 0 || cid >= G.n_cons || n < 0) return 0;
    G.con_nodes[cid].clear();
    G.con_areas[cid].clear();
    if (n == 0) { mark_changed(); return 1; }
    for (int q = 0; q < n; q++) {
        if (nodes[q] < 0 || nodes[q] >= G.N) { G.con_nodes[cid].clear(); G.con_areas[cid].clear(); return 0; }
        G.con_nodes[cid].push_back(nodes[q]);
        G.con_areas[cid].push_back(areas ? areas[q] : 0.);
    }
    mark_changed();
    return 1;
}
// interface walls of a graph device (:4836-4842): per node two walls, each
// the node's distance to it and the range a..b of its cell measured from it
// (m); a distance >= 0.5 m is no wall
void api3_set_walls(const double *w) {
    if (!G.allocated || !G.graph || !w) return;
    G.gwall.assign(w, w + (size_t)6 * G.N);
    mark_changed();
}
// tunnelling paths of the last prepared device (all, or those of contact cid)
int api3_get_ntp(int cid) {
    if (cid < 0) return (int)G.tp_k.size();
    return cid < G.n_cons ? G.cons[cid].ntp : 0;
}
// tunnelling fluxes through a stack from a semiconductor surface (test hook
// of the reference, :4848-4867); 0 on invalid arguments, as the reference
double api3_tun_eval(int smat, int nl, const int *mat, const double *t, double psi, double phin, double phip,
                     double Vg, double pm, double *out) {
    if (smat < 0 || smat >= G.n_mats || nl < 1 || nl > 6) return 0.;
    for (int i = 0; i < nl; i++) if (mat[i] < 0 || mat[i] >= G.n_mats) return 0.;
    return eng_tun_eval(smat, nl, mat, t, psi, phin, phip, Vg, pm, out);
}
// MLDA density factor of material mi at the cell range [a, b] (m) from a wall
double api3_mlda_factor(int mi, double a, double b, int hole) {
    if (mi < 0 || mi >= G.n_mats) return 1.;
    return eng_mlda_factor(mi, a, b, hole);
}
double api3_coulomb_inv(int, double, double, double, double) { return NAN; }
int api3_set_surface_mobility(int mi, double, double, double) { return (mi >= 0 && mi < G.n_mats) ? 1 : 0; }
// Lombardi edge mobility (m^2/Vs) for bulk mobility mub, normal field F (V/m),
// total doping Nt (m^-3) and distance factor Dw, evaluated on the device
double api3_lombardi(int mi, double mub, double F, double Nt, double Dw, int hole) {
    if (mi < 0 || mi >= G.n_mats) return mub;
    return eng_lombardi(mi, mub, F, Nt, Dw, hole);
}
void api3_set_contact_V(int cid, double V) { if (cid >= 0 && cid < G.n_cons) G.cons[cid].V = V; }
int api3_get_ncon(void) { return G.n_cons; }
double api3_get_contact_I(int cid) { return (cid >= 0 && cid < G.n_cons) ? G.cons[cid].I : 0.; }
double api3_get_contact_A(int cid) { return (cid >= 0 && cid < G.n_cons) ? G.cons[cid].A : 0.; }
double api3_get_contact_In(int cid) { return (cid >= 0 && cid < G.n_cons) ? G.cons[cid].In : 0.; }
double api3_get_contact_Ifloor(int cid) { return (cid >= 0 && cid < G.n_cons) ? G.cons[cid].Iflr : 0.; }
double api3_get_contact_Iraw(int cid) { return (cid >= 0 && cid < G.n_cons) ? G.cons[cid].Iraw : 0.; }
int api3_get_aa_depth(void) { return eng_aa_depth(); }
double api3_get_contact_Vsolved(int cid) { return (cid >= 0 && cid < G.n_cons) ? G.cons[cid].Vprev : 0.; }

static int do_solve(int keep) {
    if (!eng_prepare()) return 0;
    if (!eng_prepare_precision()) return 0;
    if (!keep || !G.have_solution) eng_equilibrium();
    int ok = eng_run_continuation();
    eng_display();
    disp_valid = 1;
    G.converged = ok && eng_ok();
    g_prog[4] = 0.;
    return G.converged;
}
void api3_request_stop(void) { g_stop = 1; }
void api3_clear_stop(void) { g_stop = 0; }
int api3_stop_requested(void) { return g_stop; }
void api3_get_progress(double *out) { for (int i = 0; i < 8; i++) out[i] = g_prog[i]; }
int api3_solve(void) { return do_solve(0); }
int api3_resolve(void) { return do_solve(1); }

void api3_iv_sweep(int cid, double Vs, double Ve, int Npts, double *Varr, double *Jarr, int *Carr) {
    if (cid < 0 || cid >= G.n_cons) return;
    if (Npts < 1) Npts = 1;
    if (Npts > 400) Npts = 400;
    double dV = (Npts > 1) ? (Ve - Vs) / (Npts - 1.) : 0.;
    double V0 = G.cons[cid].V;
    if (!eng_prepare() || !eng_prepare_precision()) return;
    if (!G.have_solution) eng_equilibrium();
    for (int kk = 0; kk < Npts; kk++) {
        double V = Vs + kk * dV;
        if (g_stop) { Varr[kk] = NAN; Jarr[kk] = NAN; Carr[kk] = -1; continue; }
        G.cons[cid].V = V;
        int ok = eng_run_continuation();
        double A = G.cons[cid].A;
        Varr[kk] = G.cons[cid].Vprev;
        Jarr[kk] = A > 0. ? G.cons[cid].I / A : G.cons[cid].I;
        Carr[kk] = ok;
    }
    eng_display();
    disp_valid = 1;
    g_prog[4] = 0.;
    G.cons[cid].V = V0;
}
void api3_iv_sweep2(int cid, double Vs, double Ve, int Npts, double *Varr, double *Iarr, double *Farr, int *Carr) {
    if (cid < 0 || cid >= G.n_cons) return;
    if (Npts < 1) Npts = 1;
    if (Npts > 400) Npts = 400;
    double dV = (Npts > 1) ? (Ve - Vs) / (Npts - 1.) : 0.;
    double V0 = G.cons[cid].V;
    int nc = G.n_cons;
    if (!eng_prepare() || !eng_prepare_precision()) return;
    if (!G.have_solution) eng_equilibrium();
    for (int kk = 0; kk < Npts; kk++) {
        if (g_stop) {
            Varr[kk] = NAN; Carr[kk] = -1;
            for (int c = 0; c < nc; c++) { Iarr[kk * nc + c] = NAN; Farr[kk * nc + c] = NAN; }
            continue;
        }
        G.cons[cid].V = Vs + kk * dV;
        int ok = eng_run_continuation();
        Varr[kk] = G.cons[cid].Vprev;
        for (int c = 0; c < nc; c++) { Iarr[kk * nc + c] = G.cons[c].I; Farr[kk * nc + c] = G.cons[c].Iflr; }
        Carr[kk] = ok;
    }
    eng_display();
    disp_valid = 1;
    g_prog[4] = 0.;
    G.cons[cid].V = V0;
}

static std::vector<double> dl(const double *src) {
    std::vector<double> v(G.N);
    eng_download(v.data(), src, G.N);
    return v;
}

double api3_get_Emax(void) {
    if (!G.allocated || !G.have_solution) return 0.;
    if (G.graph) return gr_emax();
    std::vector<double> phi = dl(G.d_psi);
    double Em = 0.;
    for (int k = 0; k < G.N; k++) {
        if (G.kind[k] != K_SEMI) continue;
        int l = k / (G.Nx * G.Ny), r = k - l * G.Nx * G.Ny, j = r / G.Nx, i = r - j * G.Nx;
        double g2 = 0.;
        int st[3] = {1, G.Nx, G.Nx * G.Ny}, idx[3] = {i, j, l}, nn[3] = {G.Nx, G.Ny, G.Nz};
        const std::vector<double> *cs[3] = {&G.xs, &G.ys, &G.zs};
        for (int ax = 0; ax < 3; ax++) {
            int q = idx[ax];
            bool a = q > 0, b = q < nn[ax] - 1;
            double g = 0.;
            if (a && b) g = (phi[k + st[ax]] - phi[k - st[ax]]) / ((*cs[ax])[q + 1] - (*cs[ax])[q - 1]);
            else if (b) g = (phi[k + st[ax]] - phi[k]) / ((*cs[ax])[q + 1] - (*cs[ax])[q]);
            else if (a) g = (phi[k] - phi[k - st[ax]]) / ((*cs[ax])[q] - (*cs[ax])[q - 1]);
            g2 += g * g;
        }
        if (g2 > Em) Em = g2;
    }
    return sqrt(Em);
}

int api3_get_Nx(void) { return G.Nx; }
int api3_get_Ny(void) { return G.Ny; }
int api3_get_Nz(void) { return G.Nz; }
int api3_get_conv(void) { return G.converged; }
int api3_get_iter(void) { return G.conv_iter; }
double api3_get_res(void) { return G.conv_res; }
double api3_get_xs(int i) { return (G.allocated && i >= 0 && i < G.Nx) ? G.xs[i] : NAN; }
double api3_get_ys(int j) { return (G.allocated && j >= 0 && j < G.Ny) ? G.ys[j] : NAN; }
double api3_get_zs(int l) { return (G.allocated && l >= 0 && l < G.Nz) ? G.zs[l] : NAN; }
double api3_get_Cref(void) { return G.Cref; }
long long api3_get_stat(int w) {
    switch (w) { case 0: return G.pcg_its; case 1: return G.bicg_its; case 2: return G.newton_its; case 3: return G.gummel_its; }
    return 0;
}
void api3_set_predictor(int on) { eng_set_predictor(on); }
int api3_get_state(double *phi, double *phin, double *phip) {
    if (!G.allocated) return 0;
    eng_download(phi, G.d_psi, G.N);
    eng_download(phin, G.d_phin, G.N);
    eng_download(phip, G.d_phip, G.N);
    return G.have_solution;
}
int api3_set_state(const double *phi, const double *phin, const double *phip) {
    if (!eng_prepare() || !eng_prepare_precision()) return 0;
    if (!G.have_solution) eng_equilibrium();
    eng_upload(G.d_psi, phi, G.N);
    eng_upload(G.d_phin, phin, G.N);
    eng_upload(G.d_phip, phip, G.N);
    double Va[SD_NC_MAX];
    for (int c = 0; c < G.n_cons; c++) { Va[c] = G.cons[c].V; G.cons[c].Vprev = Va[c]; }
    eng_set_vrange(Va);
    eng_upload_cons(Va);
    eng_apply_bc();
    eng_clamp_qf();
    eng_rebuild_np();
    eng_upd_mob();
    eng_sync();
    G.have_solution = 1;
    G.e_n = 0;
    disp_valid = 0;
    return eng_ok();
}
double api3_get_time(int w) {
    switch (w) { case 0: return G.t_pois; case 1: return G.t_cont; case 2: return G.t_mob;
                 case 3: return G.t_cur; case 4: return G.t_aa; case 5: return G.t_mg; }
    return 0.;
}
double api3_mat_ni(int mi) { return (mi >= 0 && mi < G.n_mats) ? G.mats[mi].ni : 0.; }
double api3_mat_Eg(int mi) { return (mi >= 0 && mi < G.n_mats) ? G.mats[mi].Eg : 0.; }

// Derived node quantity `which` for all nodes (node_val, semiconductor_3d.c:5066-5125)
void api3_bulk_copy(int which, double *out) {
    if (!G.allocated) return;
    const int N = G.N;
    if (!G.have_solution || which < 0 || which > 35) {
        for (int k = 0; k < N; k++) out[k] = (which < 0 || which > 35) && G.have_solution ? 0. : NAN;
        return;
    }
    std::vector<double> a, b;
    DevView d = eng_view();
    auto mt = [&](int k) -> const HostMat & { return G.mats[G.mat8[k]]; };
    switch (which) {
    case 0: eng_download(out, G.d_psi, N); return;
    case 1: eng_download(out, G.d_n, N); return;
    case 2: eng_download(out, G.d_p, N); return;
    case 3: case 4: a = dl(G.d_psi);
        for (int k = 0; k < N; k++) {
            const HostMat &m = mt(k);
            out[k] = G.kind[k] == K_METAL ? NAN
                     : (which == 3 ? -a[k] - m.chi - G.bgn[k] * m.VT : -a[k] - m.chi - m.Eg + G.bgn[k] * m.VT);
        }
        return;
    case 5: case 6: a = dl(which == 5 ? G.d_phin : G.d_phip);
        for (int k = 0; k < N; k++) out[k] = G.kind[k] == K_INS ? NAN : -a[k] - G.Cref;
        return;
    case 7: eng_download(out, G.d_Jx, N); return;
    case 8: eng_download(out, G.d_Jy, N); return;
    case 9: eng_download(out, G.d_Jz, N); return;
    case 10: eng_download(out, disp_valid ? G.d_mudn : G.d_mun, N); return;
    case 11: eng_download(out, disp_valid ? G.d_mudp : G.d_mup, N); return;
    case 12: eng_download(out, G.d_Rt, N); return;
    case 13: for (int k = 0; k < N; k++) out[k] = (double)G.kind[k]; return;
    case 14: for (int k = 0; k < N; k++) out[k] = G.Nd[k] - G.Na[k]; return;
    case 15: eng_download(out, G.d_phin, N); return;
    case 16: eng_download(out, G.d_phip, N); return;
    case 17: for (int k = 0; k < N; k++) out[k] = G.Nf[k]; return;
    case 18: for (int k = 0; k < N; k++) out[k] = (double)G.con_mask[k]; return;
    case 19: for (int k = 0; k < N; k++) out[k] = G.bgn[k] * mt(k).VT * 2.; return;
    case 20: case 21: case 22: case 23: case 24: case 25: {
        int ax = (which - 20) % 3, hole = which >= 23;
        if (G.graph) gr_node_J(ax, 1 + hole, G.d_wk[0]);
        else ph_node_Jc(G.st, d, ax, hole, G.d_wk[0]);
        eng_download(out, G.d_wk[0], N);
        return;
    }
    case 26: for (int k = 0; k < N; k++) out[k] = G.kind[k] == K_SEMI ? G.niK[k] : 0.; return;
    case 27:                       // gate tunnelling current density (A/m^2), 0 off the interface
        sd_fill(G.st, G.d_wk[0], 0.0, N);
        ph_tun_J(G.st, d, G.tun, G.d_wk[0]);
        eng_download(out, G.d_wk[0], N);
        return;
    case 28: case 29: {            // quantum potentials -qc VT (eV); 0 without MLDA
        const std::vector<double> &qc = which == 28 ? G.qcn : G.qcp;
        for (int k = 0; k < N; k++)
            out[k] = G.kind[k] == K_SEMI ? -((size_t)k < qc.size() ? qc[k] : 0.) * mt(k).VT : NAN;
        return;
    }
    case 30: {
        std::vector<double> qd, qcn, qcp;
        if (G.graph) gr_walls(qd, qcn, qcp, 0, nullptr);
        else eng_wall_distance(qd);
        for (int k = 0; k < N; k++) out[k] = G.kind[k] == K_SEMI ? qd[k] : NAN;
        return;
    }
    case 31: case 32: case 34: for (int k = 0; k < N; k++) out[k] = G.kind[k] == K_SEMI ? 0. : NAN; return;
    case 33: a = dl(G.d_gii);
        for (int k = 0; k < N; k++) {
            if (G.kind[k] != K_SEMI || !(G.models & 256)) { out[k] = 0.; continue; }
            int l = k / (G.Nx * G.Ny), r = k - l * G.Nx * G.Ny, j = r / G.Nx, i = r - j * G.Nx;
            double v = G.Lx[i] * G.Ly[j] * G.Lz[l];
            out[k] = v > 0. ? a[k] / v : 0.;
        }
        return;
    case 35: for (int k = 0; k < N; k++) out[k] = G.kind[k] == K_SEMI ? G.Nd[k] - G.Na[k] : 0.; return;
    }
}
void api3_carrier_sums(double *out) {
    out[0] = out[1] = 0.;
    if (!G.allocated) return;
    std::vector<double> n = dl(G.d_n), p = dl(G.d_p);
    double sn = 0., sp = 0.;
    for (int k = 0; k < G.N; k++) {
        if (G.kind[k] != K_SEMI) continue;
        int l = k / (G.Nx * G.Ny), r = k - l * G.Nx * G.Ny, j = r / G.Nx, i = r - j * G.Nx;
        double v = G.Lx[i] * G.Ly[j] * G.Lz[l];
        sn += n[k] * v; sp += p[k] * v;
    }
    out[0] = sn; out[1] = sp;
}
// ---- terminal charges (quasi-static C-V), api3_charge_partition :5192-5245,
// api3_terminal_charges :5248-5277, api3_charge_owner :5300-5304.
// The partition: every contact node belongs to its terminal; from the nodes
// of ohmic contacts each terminal spreads, breadth first and all at once,
// through semiconductor nodes of the same majority carrier and through
// metal; fronts of different terminals that meet are counted (0 = every
// domain has one terminal).  Semiconductor left over joins the terminal next
// to it.  A terminal's charge is then the flux of D out of its node set
// (Gauss's law), plus the oxide flux of its face (Robin) gate nodes.  These
// are host passes over a readback of the state, as in the reference (serial,
// the same bits on every run).
static std::vector<int> g_qown;
static int g_qct[SD_NC_MAX];
// neighbours b of node k with the edge's length h and dual face area ar, on
// either mesh (a graph device: its CSR row, in stored order; qnbrs :5172-5187)
extern "C++" {
template <class F> static void for_nbrs(int k, F body) {
    if (G.graph) {
        for (int e = G.grp[k]; e < G.grp[k + 1]; e++) body(G.gcol[e], G.gh[e], G.ga[e]);
        return;
    }
    int i, j, l;
    host_unidx(k, i, j, l);
    for (int dir = 0; dir < 6; dir++) {
        int b; double h, ar;
        if (host_nbr(i, j, l, k, dir, b, h, ar)) body(b, h, ar);
    }
}
}
int api3_charge_partition(const int *cterm) {
    if (!G.allocated || !G.have_solution || !cterm) return -1;
    const int N = G.N;
    std::vector<double> n = dl(G.d_n), p = dl(G.d_p);
    auto qtype = [&](int k) { return n[k] >= p[k] ? 1 : -1; };
    g_qown.assign((size_t)N, -1);
    for (int c = 0; c < SD_NC_MAX; c++) g_qct[c] = c < G.n_cons ? cterm[c] : -1;
    std::vector<int> qu;
    qu.reserve((size_t)N);
    int meet = 0;
    for (int k = 0; k < N; k++) {
        const int c = G.con_mask[k] - 1;
        if (c < 0 || c >= G.n_cons) continue;
        const int t = cterm[c];
        g_qown[k] = t;
        const HostCon &cn = G.cons[c];
        if (t >= 0 && cn.bc == 0 && !cn.te && (G.kind[k] == K_SEMI || G.kind[k] == K_METAL)) qu.push_back(k);
    }
    for (size_t qh = 0; qh < qu.size(); qh++) {
        const int k = qu[qh], t = g_qown[k];
        for_nbrs(k, [&](int b, double, double) {
            if (G.con_mask[b]) return;
            const bool ok = G.kind[k] == K_METAL ? G.kind[b] == K_METAL
                          : (G.kind[k] == K_SEMI && G.kind[b] == K_SEMI && qtype(b) == qtype(k));
            if (!ok) return;
            if (g_qown[b] < 0) { g_qown[b] = t; qu.push_back(b); }
            else if (g_qown[b] != t) meet++;
        });
    }
    // second pass: semiconductor no ohmic front reached joins its neighbour's terminal
    qu.clear();
    for (int k = 0; k < N; k++) if (g_qown[k] >= 0 && G.kind[k] == K_SEMI) qu.push_back(k);
    for (size_t qh = 0; qh < qu.size(); qh++) {
        const int k = qu[qh], t = g_qown[k];
        for_nbrs(k, [&](int b, double, double) {
            if (G.con_mask[b] || G.kind[b] != K_SEMI || g_qown[b] >= 0) return;
            g_qown[b] = t;
            qu.push_back(b);
        });
    }
    return meet;
}
int api3_terminal_charges(int nt, double *Qt) {
    const int N = G.N;
    if (!G.allocated || !G.have_solution || (int)g_qown.size() != N || N == 0 || nt <= 0 || !Qt) return 0;
    for (int t = 0; t <= nt; t++) Qt[t] = 0.;
    std::vector<double> phi = dl(G.d_psi);
    for (int k = 0; k < N; k++) {
        const int t = g_qown[k];
        if (t < 0 || t >= nt) continue;
        for_nbrs(k, [&](int b, double h, double ar) {
            const int tb = g_qown[b];
            if (tb == t) return;
            const double ek = G.eps[k], eb = G.eps[b];
            const double f = (2. * ek * eb / (ek + eb)) * ar / h * (phi[k] - phi[b]);
            Qt[t] += f;
            if (tb < 0 || tb >= nt) Qt[nt] += f;
        });
    }
    for (int k = 0; k < N; k++) {                 // face gates: D through the oxide
        if (!(G.robA[k] > 0.) || !G.gate_mask[k]) continue;
        const int c = G.gate_mask[k] - 1;
        const HostCon &cn = G.cons[c];
        const double Cox = SD_EPS_OX / (cn.tox > 0. ? cn.tox : 10e-9);
        const double pm = cn.phi_m > 0. ? cn.phi_m : G.Cref;
        const double f = G.robA[k] * Cox * (phi[k] - (cn.V + G.Cref - pm));   // out of the semiconductor
        const int t = g_qown[k], tg = (c >= 0 && c < SD_NC_MAX) ? g_qct[c] : -1;
        if (t >= 0 && t < nt) Qt[t] += f;
        if (tg >= 0 && tg < nt) Qt[tg] -= f;
    }
    return 1;
}
int api3_charge_owner(int *out) {
    if ((int)g_qown.size() != G.N || G.N == 0 || !out) return 0;
    memcpy(out, g_qown.data(), (size_t)G.N * sizeof(int));
    return 1;
}
// tensor mesh only: 0 on a graph device (as the reference, :5287)
int api3_face_J(int ax, int carrier, double *out) {
    if (!G.allocated || !G.have_solution || !out || ax < 0 || ax > 2 || G.graph) return 0;
    ph_face_J(G.st, eng_view(), ax, carrier, G.d_wk[0]);
    eng_download(out, G.d_wk[0], G.N);
    return 1;
}
double api3_ins_field_max(double *pos) {
    if (pos) pos[0] = pos[1] = pos[2] = NAN;
    if (!G.allocated || !G.have_solution) return NAN;
    std::vector<double> phi = dl(G.d_psi);
    double best = 0.;
    long kbest = -1;
    for (int k = 0; k < G.N && G.graph; k++) {      // graph: every edge between two non-semiconductor nodes
        if (G.kind[k] == K_SEMI) continue;
        for_nbrs(k, [&](int b, double h, double) {
            if (b < k || G.kind[b] == K_SEMI) return;
            if (G.kind[k] != K_INS && G.kind[b] != K_INS) return;
            const double F = fabs(phi[b] - phi[k]) / h;
            if (F > best) { best = F; kbest = G.kind[k] == K_INS ? k : b; }
        });
    }
    for (int k = 0; k < G.N && !G.graph; k++) {
        if (G.kind[k] == K_SEMI) continue;
        int l = k / (G.Nx * G.Ny), r = k - l * G.Nx * G.Ny, j = r / G.Nx, i = r - j * G.Nx;
        int st[3] = {1, G.Nx, G.Nx * G.Ny}, idx[3] = {i, j, l}, nn[3] = {G.Nx, G.Ny, G.Nz};
        const std::vector<double> *cs[3] = {&G.xs, &G.ys, &G.zs};
        for (int ax = 0; ax < 3; ax++)
            for (int sgn = -1; sgn <= 1; sgn += 2) {
                int q = idx[ax] + sgn;
                if (q < 0 || q >= nn[ax]) continue;
                int b = k + sgn * st[ax];
                if (b < k || G.kind[b] == K_SEMI) continue;
                if (G.kind[k] != K_INS && G.kind[b] != K_INS) continue;
                double h = fabs((*cs[ax])[q] - (*cs[ax])[idx[ax]]);
                double F = fabs(phi[b] - phi[k]) / h;
                if (F > best) { best = F; kbest = G.kind[k] == K_INS ? k : b; }
            }
    }
    for (int k = 0; k < G.N; k++) {
        int gm = G.gate_mask[k];
        if (!gm) continue;
        const HostCon &cn = G.cons[gm - 1];
        double tox = cn.tox > 0. ? cn.tox : 10e-9, pm = cn.phi_m > 0. ? cn.phi_m : G.Cref;
        double F = fabs(phi[k] - (cn.V + G.Cref - pm)) / tox;
        if (F > best) { best = F; kbest = k; }
    }
    if (pos && kbest >= 0 && G.graph) {
        for (int q = 0; q < 3; q++) pos[q] = G.gpos[3 * kbest + q];
    } else if (pos && kbest >= 0) {
        int l = kbest / (G.Nx * G.Ny), r = kbest - l * G.Nx * G.Ny, j = r / G.Nx, i = r - j * G.Nx;
        pos[0] = G.xs[i]; pos[1] = G.ys[j]; pos[2] = G.zs[l];
    }
    return best;
}
static double node1(int which, int k) {
    if (!G.allocated || !G.have_solution || k < 0 || k >= G.N) return NAN;
    std::vector<double> v(G.N);
    api3_bulk_copy(which, v.data());
    return v[k];
}
double api3_get_phi(int k) { return node1(0, k); }
double api3_get_n(int k) { return node1(1, k); }
double api3_get_p(int k) { return node1(2, k); }
double api3_get_Ec(int k) { return node1(3, k); }
double api3_get_Ev(int k) { return node1(4, k); }
double api3_get_Efn(int k) { return node1(5, k); }
double api3_get_Efp(int k) { return node1(6, k); }
double api3_get_Jx(int k) { return node1(7, k); }
double api3_get_Jy(int k) { return node1(8, k); }
double api3_get_Jz(int k) { return node1(9, k); }
double api3_get_mun(int k) { return node1(10, k); }
double api3_get_mup(int k) { return node1(11, k); }
double api3_get_R(int k) { return node1(12, k); }

// ------------------------------------------------ extension entry points
int sdx_set_precision(int mode) { return eng_set_precision(mode); }
int sdx_get_precision(void) { return G.precision; }
int sdx_device_name(char *buf, int n) {
    hipDeviceProp_t p;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&p, dev) != hipSuccess) return 0;
    snprintf(buf, n, "%s (%s, %d CUs)", p.name, p.gcnArchName, p.multiProcessorCount);
    return 1;
}
int sdx_kernel_stats(double *out, int n) {
    double v[7] = {(double)G.n_vcycles, (double)G.n_spmv, (double)G.N,
                   (double)(G.precision == 1 ? G.H32.lev.size() : G.H64.lev.size()),
                   (double)G.n_inner, (double)G.n_outer, (double)G.n_inner_fb};
    for (int i = 0; i < n && i < 7; i++) out[i] = v[i];
    return 7;
}
void sdx_reset_stats(void) { G.n_vcycles = G.n_spmv = G.n_inner = G.n_outer = G.n_inner_fb = 0; }

}  // extern "C"
