// sd_physics.h - launchers of the drift-diffusion physics kernels (k_physics.hip).
#pragma once
#include "sd_common.h"
#include "sd_linalg.h"

// Terminal-current edges: (contact node, semiconductor neighbour outside the
// contact), grouped by contact: edges [off[c], off[c+1]) belong to contact c.
struct CurEdges {
    const int *k, *kb;
    const double *h, *ar;
    const int *off;       // device, n_cons+1
    int ncon;
    const int *poff;      // gate tunnelling paths of each contact in pidx (n_cons+1; nullptr: none)
    const int *pidx;      // path indices grouped by contact
    TunDev tun;
};

void ph_tun_update(hipStream_t s, const DevView &d, const TunDev &T);
void ph_tun_gather(hipStream_t s, const TunDev &T, int hole, double *tk, double *tg);
void ph_tun_J(hipStream_t s, const DevView &d, const TunDev &T, double *out);
void ph_lombardi_eval(hipStream_t s, const MatDev &m, double mub, double Fv, double Nt, double Dw, int hole,
                      double *out);
void ph_tun_eval(hipStream_t s, const MatDev *mats, int smat, int nl, const int *mi, const double *t, double psi,
                 double phin, double phip, double Vg, double pm, double Cref, double *F);
void ph_set_tile_mode(int m);          // c_tile_mode of k_physics.hip
void ph_poisson(hipStream_t s, const DevView &d, const double *phi, int build, double *F, Sten7<double> A,
                double *sc, RedBuf rb, int slot);
void ph_rebuild_np(hipStream_t s, const DevView &d);
void ph_upd_mob(hipStream_t s, const DevView &d, const double *qn, const double *qp);
void ph_cont(hipStream_t s, const DevView &d, int hole, Sten7<double> A, double *rhs, double *sc, double *dsc);
void ph_cont_update(hipStream_t s, const DevView &d, int hole, const double *sol, const double *sc, double *dq,
                    double lo, double hi, RedBuf rb, int slot);
void ph_clamp_qf(hipStream_t s, const DevView &d, double lo, double hi);
void ph_apply_bc(hipStream_t s, const DevView &d, int have_solution);
void ph_init_eq(hipStream_t s, const DevView &d);
void ph_neutral_pred(hipStream_t s, const DevView &d, const double *dqn, const double *dqp);
void ph_currents(hipStream_t s, const DevView &d, CurEdges e, RedBuf rb);
void ph_extrap2(hipStream_t s, int N, double *const d[3], const double *const a[3], const double *const b[3],
                const double *const c[3], double r);
void ph_extrap(hipStream_t s, int N, double *dst0, double *dst1, double *dst2, const double *a0, const double *a1,
               const double *a2, const double *b0, const double *b1, const double *b2, double r);
void ph_update_psi(hipStream_t s, int N, double *phi, const double *dx, double lam, double *out);

// float_fix support: flags 0 (not majority), 1 (majority, floating so far),
// 2 (majority, anchored to a contact)
void ph_maj_flags(hipStream_t s, const DevView &d, int hole, const double *u, const double *sc, unsigned char *flag);
void ph_maj_propagate(hipStream_t s, const DevView &d, unsigned char *flag, int axis, int *changed);
void ph_maj_count(hipStream_t s, int N, const unsigned char *flag, RedBuf rb, int slot);
#define SD_MAJ_REC 25   // doubles per node record of ph_maj_gather
void ph_maj_list(hipStream_t s, int N, const unsigned char *flag, int *cnt, int *list, int cap);
void ph_maj_gather(hipStream_t s, const DevView &d, int hole, const double *u, const double *sc, const int *list,
                   int m, double *out);
void ph_scale_list(hipStream_t s, double *u, const int *list, const double *lam, int m);

// Anderson mixing of x = (psi, phi_n, phi_p)
void ph_aa_diff(hipStream_t s, const DevView &d, const double *x0, const double *x1, const double *x2,
                const double *g0, const double *g1, const double *g2, double *f, double *gp, float *dF, float *dG,
                int have_prev);
void ph_aa_dots(hipStream_t s, int L, float *const *vecs, int nv, const float *y, const double *yd, RedBuf rb,
                int slot0);
void ph_aa_mix(hipStream_t s, const DevView &d, double *g0, double *g1, double *g2, float *const *dG, const double *gam,
               int m);

// display quantities after a solve
void ph_upd_J(hipStream_t s, const DevView &d, double *Jx, double *Jy, double *Jz, double *R);
void ph_mob_display(hipStream_t s, const DevView &d, double *mun, double *mup);
void ph_node_Jc(hipStream_t s, const DevView &d, int ax, int hole, double *out);
void ph_face_J(hipStream_t s, const DevView &d, int ax, int carrier, double *out);
void ph_ii_update(hipStream_t s, const DevView &d);
