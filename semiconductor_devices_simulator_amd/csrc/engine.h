// engine.h - host-side state of the MI355X drift-diffusion core.
//
// One device per process, held in a singleton (the reference ABI keeps a
// global `gdev`, semiconductor_3d.c:366).  Setup data live on the host and are
// uploaded by prepare(); the solver state (psi, phi_n, phi_p, carriers,
// operators, multigrid hierarchy, Krylov and Anderson work vectors) lives in
// HBM for the whole solve and only scalars (norms, dots, terminal currents)
// come back to the host to steer the iteration.
#pragma once
#include "sd_physics.h"
#include "sd_graph.h"
#include "sd_host.h"
#include <vector>

struct HostCon {
    int face, i0, i1, j0, j1, k0, k1, bc;
    double V, phi_m, tox, qox;
    double Vprev, I, A, In, Iflr, Iraw;
    int te, ntp;
    int snl, smat[6];                   // gate stack for tunnelling (api3_set_gate_stack), semiconductor side first
    double st[6];
};

template <class T> struct Hier {
    std::vector<MGLevel<T>> lev;
    double *inv = nullptr;         // coarsest-level inverse (row-major)
    bool alloc = false;
    bool pw_one = false;
    int tail = -1;                 // first level handled by the single-workgroup tail kernel (-1 none)
    MGLevel<T> *d_tail = nullptr;  // device copies of the tail levels' descriptors
    // level-0 storage of the hierarchy's own (operator, restriction weights,
    // vectors), used when level 0 is stored parity-split or in FP32
    Sten7<T> A0{};
    double *rs0 = nullptr;
    T *b0 = nullptr, *x0 = nullptr, *x0b = nullptr;
    // BF16 smoother coefficients of level 0 (FP32 hierarchy; allocated on
    // first use): 6 off-diagonals, then 3 factors for each of the 3 axes
    __bf16 *c16 = nullptr;
    uint4 *r16 = nullptr;          // the same packed per node and axis (3 N records)
    uint4 *r16b = nullptr;         // packed BF16 records of level 1 (3 N1), when it is line-smoothed
};

struct Engine {
    // ------------------------------------------------------ setup (host)
    int Nx = 0, Ny = 0, Nz = 0, N = 0;
    std::vector<double> xs, ys, zs, Lx, Ly, Lz;
    std::vector<int> mat_idx;
    std::vector<double> Nd, Na, Nf;
    HostMat mats[SD_NM_MAX];
    int n_mats = 0;
    HostCon cons[SD_NC_MAX];
    int n_cons = 0;
    // shaped contacts: an explicit node list (and per-node boundary areas)
    // replacing the contact's index box (api3_set_con_nodes); empty = the box
    std::vector<int> con_nodes[SD_NC_MAX];
    std::vector<double> con_areas[SD_NC_MAX];
    int models = 7, caxis = 1;
    int max_iter_p = 2000, max_iter_c = 1000, max_gummel = 500;
    double tol_p = 1e-9, tol_c = 1e-6, omega_p = 1., omega_c = 1.;
    double pcg_rtol = 1e-3, ntol_fac = 0.1, bicg_rtol = 1e-4, cont_atol_f = 1.;
    // relative tolerances of the FP32 method's inner solves: the Newton-step
    // PCG (the Newton iteration corrects an inexact step), and the continuity
    // BiCGSTAB while the Gummel update exceeds loose_r32 (V); closer to
    // convergence the continuity solves return to bicg_rtol, which sets how
    // closely a converged point's currents approach the exact discrete
    // solution (measured on the pn diode at 0.5 V: 1e-3 to the end left the
    // current 4 floors from it, 1e-4 one floor)
    double pcg_rtol32 = 1e-2, bicg_rtol32 = 1e-3, loose_r32 = 1e-6;
    double gummel_r = 0.;              // the running Gummel loop's last update max(|dpsi|, |dphi|) (V)
    int converged = 0, conv_iter = 0;
    double conv_res = 0.;
    int allocated = 0, have_solution = 0, dirty = 1, verbose = 0, nthreads = 1;
    double vlo = 0., vhi = 0.;
    long long pcg_its = 0, bicg_its = 0, newton_its = 0, gummel_its = 0;
    double t_pois = 0, t_cont = 0, t_mob = 0, t_cur = 0, t_aa = 0, t_mg = 0;
    double ii_err = 0.;
    int precision = 0;                 // 0 FP64 multigrid, 1 FP32 multigrid inside FP64 Krylov
    int unsupported = 0;               // a requested feature outside this core's scope

    // derived by prepare()
    std::vector<unsigned char> kind, mat8;
    std::vector<int> con_mask, gate_mask;
    std::vector<double> robA, eps, bgn, lncK, lnvK, niK, mu0n, mu0p;
    double Cref = 0.;
    int ref_mat = 0, n_rob = 0;

    // sweep history (continuation predictor)
    double Ve1[SD_NC_MAX], Ve2[SD_NC_MAX], Ve3[SD_NC_MAX];
    int e_n = 0;
    int pred2 = 0;                     // quadratic sweep predictor through the last three points (off: the reference's secant)
    int e3_ok = 0;                     // d_e3 holds the state of the bias point before the one in d_e2
    double *d_e3[3] = {nullptr, nullptr, nullptr};

    // ------------------------------------------------------ device state
    hipStream_t st = nullptr;
    double *d_geo = nullptr;           // xs|ys|zs|Lx|Ly|Lz
    unsigned char *d_kind = nullptr, *d_mat = nullptr, *d_flag = nullptr;
    int *d_con_mask = nullptr, *d_gate_mask = nullptr, *d_changed = nullptr;
    int *d_mlist = nullptr;            // unanchored majority nodes (floating blocks)
    double *d_mrec = nullptr;          // their gathered records, grown on demand
    size_t mrec_cap = 0;
    double *d_Nd, *d_Na, *d_Nf, *d_eps, *d_bgn, *d_robA, *d_lncK, *d_lnvK, *d_niK;
    double *d_psi, *d_phin, *d_phip, *d_n, *d_p, *d_mun, *d_mup, *d_fmu, *d_gii;
    Sten7<double> A;
    double *d_rhs, *d_sc, *d_dsc, *d_sol, *d_wk[7], *d_xsave, *d_tr, *d_Ft;
    double *d_xk, *d_dqn, *d_dqp, *d_qn0, *d_qp0;
    double *d_h1[3], *d_h2[3], *d_e2[3];
    float *d_aadF[12], *d_aadG[12];
    double *d_aaf, *d_aag;
    double *d_Jx, *d_Jy, *d_Jz, *d_Rt, *d_mudn, *d_mudp;
    double *d_qd = nullptr;            // distance to the nearest interface (Lombardi)
    double *d_tk = nullptr, *d_tg = nullptr;     // gate tunnelling sink/source of the carrier being solved
    double *d_qcn = nullptr, *d_qcp = nullptr;   // MLDA terms (tunnelling supply)
    TunDev tun{};                      // gate tunnelling paths (np = 0: none)
    int *d_tp_off = nullptr;           // paths of each contact (n_cons+1) in d_tp_idx
    int *d_tp_idx = nullptr;           // path indices grouped by contact (each group ascending)
    // host copies of the paths (test hooks): node, contact, first column entry,
    // column length; column nodes, lengths, spacings
    std::vector<int> tp_k, tp_c, tp_c0, tp_cn, col_k;
    std::vector<double> tp_A, col_l, col_h;
    // each path's layers (semiconductor half cell first on a meshed path):
    // count, SD_TUN_MAXL materials and thicknesses (m)
    std::vector<int> tp_nl, tp_m;
    std::vector<double> tp_t;
    std::vector<double> qd;            // its host copy (source of the upload)
    std::vector<double> qcn, qcp;      // MLDA confinement terms of ln Nc, ln Nv (0 without MLDA)
    MatDev *d_mats = nullptr;
    ConDev *d_cons = nullptr;
    int *d_ek = nullptr, *d_ekb = nullptr, *d_eoff = nullptr;
    double *d_eh = nullptr, *d_ear = nullptr;
    int n_edges = 0;
    RedBuf rb{};
    double *h_scal = nullptr;          // pinned mirror of the scalar slots
    std::vector<void *> dev_allocs;
    std::vector<void *> prep_allocs;   // buffers sized by prepare (contact edges, tunnelling paths): freed by the next prepare

    Hier<double> H64;
    Hier<float> H32;
    float *d_b32 = nullptr, *d_x32 = nullptr;   // level-0 FP32 vectors
    float *d_x32b = nullptr;                    // second level-0 FP32 solution (BiCGSTAB sh)
    // FP32 method: each linear solve is FP64 defect correction around an FP32
    // Krylov solve on level-0 storage (8 FP32 vectors: r, rh, p, v, ph, sh, t, d)
    int inner32 = 1;
    float *d_k32 = nullptr;
    const double *mg_w = nullptr;
    int mg_qsign = 0;
    int mg_off = 0;
    int n_axes = 2, mg_axes[3] = {1, 2, 0};   // line-smoother axes (0: point red-black GS)
    int line_levels = 2;                      // levels smoothed with lines (the finest ones)
    int tail_sweeps = 1;                      // red-black sweeps per side in the tail kernel
    int coarse_axes = 3;                      // axes line-smoothed on the coarse line levels (strongest)
    int tail_n = 1024;                        // levels of at most this many nodes run in the tail kernel
    int bi_fuse = 1;                          // BiCGSTAB: rh.r computed with the x/r update
    int unit_diag = 0;                        // the operator's diagonal is 1 (scaled continuity rows):
                                              // the Krylov products skip reading it
    int stage_on = 1;                         // Krylov kernels stage the V-cycle's input/output
    int stage_read = 1;                       // Krylov updates read the V-cycle results where they lie
    // coarse-correction scale above the tail: over-correction of the
    // piecewise-constant aggregation coarse space, continuity (BiCGSTAB) and
    // Poisson (CG) operators (measured on the 3.1M-node MOSFET sweep: 7%
    // fewer BiCGSTAB and 18% fewer PCG iterations than 1.0)
    double mg_alpha = 1.1;
    double mg_alpha_p = 1.3;
    int mg_bf16 = 1;                          // FP32 method: BF16 level-0 smoother coefficients
    int mg_bf16_l1 = 1;                       // and level 1's (packed records; SEMISIM_MG_BF16_L1)
    int mg_perm = 1;                          // line-smoothed levels stored parity-split along x (SEMISIM_MG_PERM)
    // learned warm start: mode, saved prediction of the running bias point
    int ws_mode = 0, ws_have = 0;
    double *d_ws[3] = {nullptr, nullptr, nullptr};
    long long ws_stat[4] = {0, 0, 0, 0};
    int ws_chmask = 7;                 // fields the correction changes (psi 1, phi_n 2, phi_p 4)
    double ws_thr = 0.;                // smallest correction applied (V)
    double ws_accept = 0.5;            // residual ratio a kept correction must reach

    // Anderson
    int aa_m = 12, aa_n = 0, aa_pos = 0, aa_have = 0;
    double aa_gram[12][12];

    // graph device (api3_set_mesh_graph): node arrays as above with Nx = N,
    // Ny = Nz = 1; the CSR graph of directed edges, its reverse and undirected
    // edge ids, edge lengths and dual areas, control volumes, positions and the
    // walls the host supplied (api3_set_walls, 6 per node)
    int graph = 0;
    int nE = 0, nUE = 0;
    std::vector<int> grp, gcol, grev, gue;
    std::vector<double> gh, ga, gvol, gpos, gwall;
    GraphDev gd{};
    double *d_gval = nullptr;          // operator off-diagonals, one per directed edge (diagonal in A.c)
    double *d_gmu = nullptr;           // edge mobilities: electrons [nUE], then holes [nUE]
    int *d_gck = nullptr, *d_gce = nullptr, *d_gcoff = nullptr;   // terminal-current edges per contact

    // statistics of the kernels (see sdx_kernel_stats)
    long long n_vcycles = 0, n_spmv = 0;
    // FP32 inner solves: solves, outer FP64 corrections, fallbacks to the
    // FP64-Krylov form (an inner solve that did not bring the FP64 residual down)
    long long n_inner = 0, n_outer = 0, n_inner_fb = 0;
    // since the last prepare: whether a continuity solve of each carrier (0
    // electrons, 1 holes) has run, and whether one found a floating majority
    // block of more than 1% of the nodes (a near-singular operator: the FP32
    // method then keeps FP64 Krylov)
    int ff_seen[2] = {0, 0}, ff_off = 0;
    // floating majority blocks: continuity solves that listed unanchored
    // nodes, blocks found, blocks rescaled (see sdx_float_stats)
    long long ff_stat[3] = {0, 0, 0};
};

extern Engine G;
extern volatile int g_stop;
extern volatile double g_prog[8];

// engine API used by api.cpp
int eng_alloc(int Nx, int Ny, int Nz, int graph = 0);
// device memory released with the device (eng_free) or by the next prepare
void *eng_dalloc_bytes(size_t bytes);
void *eng_palloc_bytes(size_t bytes);
void eng_set_error();
// graph path (graph.cpp)
int eng_set_graph(int N, int E, const int *rp, const int *cj, const double *h, const double *ar, const double *vol,
                  const double *pos);
void gr_free();
int gr_nlev();
int gr_prepare_edges();
void gr_walls(std::vector<double> &qd, std::vector<double> &qcn, std::vector<double> &qcp, int qc,
              double (*factor)(const HostMat &, double, double, int));
void gr_mg_setup();
void gr_precond(const double *r, double *z);
void gr_spmv(const double *x, const double *b, double *y, bool unit);
void gr_spmv_dots(const double *x, double *y, const double *u0, const double *u1, int nv, int slot, int post,
                  bool unit);
void gr_floor(const double *x, const double *b, int slot);
void gr_poisson(const double *phi, int build, double *F, int slot);
void gr_upd_mob(const double *qn, const double *qp);
void gr_cont(int hole);
void gr_currents();
void gr_display();
void gr_node_J(int ax, int carrier, double *out);
int gr_float_fix(int hole);
void gr_region_shift(const double *dV);
double gr_emax();
void eng_free();
int eng_prepare();
int eng_equilibrium();
int eng_run_continuation();
void eng_display();
DevView eng_view();
// host geometry: neighbour kb of node k = (i, j, l) in direction dir (0..5:
// -x +x -y +y -z +z), spacing h and dual face area ar; false at the boundary
bool host_nbr(int i, int j, int l, int k, int dir, int &kb, double &h, double &ar);
void host_unidx(int k, int &i, int &j, int &l);
// distance of every semiconductor node to the nearest interface (1 m: none)
void eng_wall_distance(std::vector<double> &qd);
// the same, with the MLDA confinement terms per carrier (qcn, qcp)
void eng_walls(std::vector<double> &qd, std::vector<double> *qcn, std::vector<double> *qcp);
double eng_tun_eval(int smat, int nl, const int *mat, const double *t, double psi, double phin, double phip,
                    double Vg, double pm, double *F);
double eng_lombardi(int mi, double mub, double Fv, double Nt, double Dw, int hole);
double eng_mlda_factor(int mi, double a, double b, int hole);
void eng_sync();
void eng_download(double *dst, const double *src, size_t n);
void eng_upload(double *dst, const double *src, size_t n);
void eng_upload_cons(const double *Va);
void eng_set_vrange(const double *Va);
void eng_clamp_qf();
void eng_rebuild_np();
void eng_upd_mob();
void eng_apply_bc();
int eng_ok();
