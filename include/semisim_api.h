/*
 * semisim_api.h - the api3_* C ABI of the MI355X drift-diffusion core
 * (libsemisim_hip.so).  Each entry point replaces the reference core's
 * function of the same name and signature in semiconductor_3d.c (line cited);
 * the reference front end binds them through ctypes (main.py:172-217,
 * 3485-3513; validate.py:31-49) and loads the library selected by SEMISIM_SO
 * (semibuild.py:516-522), checking the sentinel api3_set_caxis.
 *
 * Units are SI (m, m^-3, V, A).  Return conventions follow the reference:
 * 1/0 success/failure, -1 on capacity overflow, NaN from getters before a
 * solve.  One solver instance per process; calls are not reentrant, except
 * api3_request_stop / api3_clear_stop / api3_get_progress, which may be called
 * from another thread while a solve runs.
 */
#ifndef SEMISIM_API_H
#define SEMISIM_API_H
#ifdef __cplusplus
extern "C" {
#endif

/* setup */
void   api3_init(void);                                                   /* semiconductor_3d.c:4570 */
void   api3_set_threads(int n);                                           /* :4599 (no effect: device core) */
void   api3_set_verbose(int v);                                           /* :4607 */
int    api3_trap_species(int type, int ref, double E, double sn, double sp); /* :4618 (rejected: traps) */
int    api3_trap_add(int sid, int n, const int *nodes, const double *dens);  /* :4629 (rejected) */
void   api3_traps_clear(void);                                            /* :4645 */
int    api3_charge_add(int n, const int *nodes, const double *rho);       /* :4646 */
int    api3_n_trap_entries(void);                                         /* :4653 */
void   api3_set_models(int flags);                                        /* :4657 */
void   api3_set_caxis(int ax);                                            /* :4660 */
int    api3_get_caxis(void);                                              /* :4661 */
int    api3_add_material(int id, double T);                               /* :4663 */
int    api3_set_mat_tau(int mi, double tn, double tp);                    /* :4671 */
int    api3_set_grid_coords(int Nx, int Ny, int Nz, const double *xs, const double *ys, const double *zs); /* :4679 */
void   api3_set_grid(int Nx, int Ny, int Nz, double Lx, double Ly, double Lz); /* :4693 */
void   api3_set_solver(int mp, int mc, int mg, double tp, double tc, double op, double oc); /* :4708 */
void   api3_set_region(int x0, int y0, int z0, int x1, int y1, int z1, int midx, double Nd, double Na); /* :4718 */
void   api3_add_fixed_charge(int x0, int y0, int z0, int x1, int y1, int z1, double Nf); /* :4731 */
void   api3_add_sheet_charge(int axis, int idx, int a0, int a1, int b0, int b1, double sigma); /* :4742 */
int    api3_add_contact_ex(int face, int i0, int i1, int j0, int j1, int k0, int k1,
                           int bc, double V, double pm, double tox, double Nox); /* :4757 */
void   api3_add_contact(int face, int i0, int i1, int j0, int j1, int bc, double V, double pm); /* :4772 */
void   api3_clear_contacts(void);                                         /* :4775 */
int    api3_set_gate_stack(int cid, int nl, const int *mat, const double *t); /* :4779 (tunnelling, models bit 5) */
int    api3_set_mesh_graph(int N, int E, const int *rp, const int *cj, const double *h,
                           const double *ar, const double *vol, const double *pos); /* :4795 (prism mesh: graph path) */
int    api3_is_graph(void);                                               /* :4806 */
void   api3_set_node_data(const int *mat, const double *Nd, const double *Na, const double *Nf); /* :4808 */
int    api3_set_con_nodes(int cid, int n, const int *nodes, const double *areas); /* :4821 (shaped contacts) */
void   api3_set_walls(const double *w);                                   /* :4838 */
int    api3_get_ntp(int cid);                                             /* :4844 */
double api3_tun_eval(int smat, int nl, const int *mat, const double *t, double psi,
                     double phin, double phip, double Vg, double pm, double *out); /* :4855 */
double api3_mlda_factor(int mi, double a, double b, int hole);           /* :4869 */
double api3_coulomb_inv(int mi, double F, double n, double NC, double NA); /* :4878 (NaN: no SiC Coulomb term) */
int    api3_set_surface_mobility(int mi, double delta, double eta, double mu1); /* :4885 */
double api3_lombardi(int mi, double mub, double F, double Nt, double Dw, int hole); /* :4894 */
/* contacts and results */
void   api3_set_contact_V(int cid, double V);                             /* :4898 */
int    api3_get_ncon(void);                                               /* :4899 */
double api3_get_contact_I(int cid);                                       /* :4900 */
double api3_get_contact_A(int cid);                                       /* :4901 */
double api3_get_contact_In(int cid);                                      /* :4902 */
double api3_get_contact_Ifloor(int cid);                                  /* :4904 */
double api3_get_contact_Iraw(int cid);                                    /* :4906 */
int    api3_get_aa_depth(void);                                           /* :4907 */
double api3_get_contact_Vsolved(int cid);                                 /* :4908 */
/* solve: the hot path */
void   api3_request_stop(void);                                           /* :4922 */
void   api3_clear_stop(void);                                             /* :4923 */
int    api3_stop_requested(void);                                         /* :4924 */
void   api3_get_progress(double *out);                                    /* :4926 */
int    api3_solve(void);                                                  /* :4928 */
int    api3_resolve(void);                                                /* :4930 */
void   api3_iv_sweep(int cid, double Vs, double Ve, int Npts, double *Varr, double *Jarr, int *Carr); /* :4934 */
void   api3_iv_sweep2(int cid, double Vs, double Ve, int Npts, double *Varr,
                      double *Iarr, double *Farr, int *Carr);             /* :4958 */
double api3_get_Emax(void);                                               /* :4982 */
int    api3_get_Nx(void);                                                 /* :5003 */
int    api3_get_Ny(void);                                                 /* :5004 */
int    api3_get_Nz(void);                                                 /* :5005 */
int    api3_get_conv(void);                                               /* :5006 */
int    api3_get_iter(void);                                               /* :5007 */
double api3_get_res(void);                                                /* :5008 */
double api3_get_xs(int i);                                                /* :5009 */
double api3_get_ys(int j);                                                /* :5010 */
double api3_get_zs(int l);                                                /* :5011 */
double api3_get_Cref(void);                                               /* :5012 */
long long api3_get_stat(int w);                                           /* :5013 */
void   api3_set_predictor(int on);                                        /* :5019 */
int    api3_get_state(double *phi, double *phin, double *phip);           /* :5023 */
int    api3_set_state(const double *phi, const double *phin, const double *phip); /* :5030 */
double api3_get_time(int w);                                              /* :5043 */
double api3_mat_ni(int mi);                                               /* :5048 */
double api3_mat_Eg(int mi);                                               /* :5049 */
void   api3_bulk_copy(int which, double *out);                            /* :5127 */
void   api3_carrier_sums(double *out);                                    /* :5137 */
int    api3_charge_partition(const int *cterm);                           /* :5192 */
int    api3_terminal_charges(int nt, double *Qt);                         /* :5248 */
int    api3_face_J(int ax, int carrier, double *out);                     /* :5284 */
int    api3_charge_owner(int *out);                                       /* :5300 */
double api3_ins_field_max(double *pos);                                   /* :5308 */
double api3_get_phi(int k);                                               /* :5341 */
double api3_get_n(int k);                                                 /* :5342 */
double api3_get_p(int k);                                                 /* :5343 */
double api3_get_Ec(int k);                                                /* :5344 */
double api3_get_Ev(int k);                                                /* :5345 */
double api3_get_Efn(int k);                                               /* :5346 */
double api3_get_Efp(int k);                                               /* :5347 */
double api3_get_Jx(int k);                                                /* :5348 */
double api3_get_Jy(int k);                                                /* :5349 */
double api3_get_Jz(int k);                                                /* :5350 */
double api3_get_mun(int k);                                               /* :5351 */
double api3_get_mup(int k);                                               /* :5352 */
double api3_get_R(int k);                                                 /* :5353 */

#ifdef __cplusplus
}
#endif
#endif
