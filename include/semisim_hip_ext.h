/*
 * semisim_hip_ext.h - entry points of libsemisim_hip.so with no counterpart
 * in the reference core.  They select the multigrid precision, report device
 * and kernel statistics, and expose single kernels on the device state for
 * parity tests.  Plain C types only.
 */
#ifndef SEMISIM_HIP_EXT_H
#define SEMISIM_HIP_EXT_H
#ifdef __cplusplus
extern "C" {
#endif

/* multigrid precision: 0 FP64 hierarchy (the plain method), 1 FP32 hierarchy,
   smoothers and preconditioner SpMV inside FP64 Krylov (mixed precision).
   Also selected by SEMISIM_MG_FP32=1 at api3_init.  1 = done. */
int  sdx_set_precision(int mode);
int  sdx_get_precision(void);
int  sdx_device_name(char *buf, int n);
/* out[0] V-cycles, out[1] level-0 SpMVs, out[2] nodes, out[3] MG levels,
   out[4] FP32 inner solves, out[5] their FP64 corrections, out[6] inner
   continuity solves redone with FP64 Krylov; returns the count filled (7) */
int  sdx_kernel_stats(double *out, int n);
void sdx_reset_stats(void);

/* sampled HIP-event timer on the solver stream: kind 0 = level-0 smoother
   sweep, 1 = Krylov SpMV, 2 = whole preconditioner application (V-cycle);
   every n-th launch is timed.  read: out[0] timed
   launches, out[1] their total ms, out[2] algorithmic bytes per launch,
   out[3] launches seen.  Reading stops the timer. */
void sdx_ktimer_enable(int kind, int every);
void sdx_ktimer_pause(int paused);        /* 1: stop sampling, keep what was gathered; 0: resume */
int  sdx_ktimer_read(double *out);

/* kernel microbenchmark on the solved device: ms per launch of `what`
   (0/1/2 line sweep x/y/z, 3 point GS sweep, 4 SpMV+dot, 5 residual+restriction,
   6 V-cycle, 7 copy of N doubles, 8 Poisson assembly, 9 continuity assembly,
   10 copy of 8M doubles, 11 preconditioner application as the solvers run it,
   12 dot product read back by the host, timed by the host clock, 13/14/15 line
   sweep x/y/z on level 1, 16 SpMV+dot reading the staged level-0 solution) */
double sdx_microbench(int what, int iters);
void   sdx_tune(int what, int value);      /* 0: lines per workgroup of the LDS line kernel, 1: line-smoothed
                                              levels, 2: tail sweeps, 3: tail size, 4: wavefront-scan line
                                              kernel on (1) or off (0), 5: most line slots per wavefront,
                                              7: fewest workgroups a wider slot choice may leave,
                                              8: LDS-resident coarse tail kernel on (1) or off (0),
                                              9: parity-split storage of the line levels, 10: XCD-contiguous
                                              line tiles, 11: tile order of the other stencil kernels (0
                                              contiguous node range per workgroup, 1 XCD-contiguous tiles
                                              with grid-stride node loops, 2 launch order; -1 default),
                                              12: restriction with lanes only for rows not solved exactly
                                              (1) or for every fine node (0), 13: BiCGSTAB computes the
                                              next rh.r with its x/r update (1) or separately (0),
                                              14: Krylov kernels stage the V-cycle's input and output in
                                              level-0 storage (1) or the V-cycle converts them (0),
                                              16: coarse-correction scale above the tail in percent, for both operators
                                              (100 = none), 17: BF16 level-0 smoother coefficients in the
                                              FP32 method (1) or FP32 (0), 18: axis-specialised line
                                              kernel for one line per slot (1) or the generic one (0),
                                              19: non-temporal loads of the operators streamed once per
                                              use (Krylov SpMV, level-0 restriction) (1) or plain (0),
                                              20: continuity BiCGSTAB products skip the unit diagonal
                                              (1) or read it (0), 21: the pre-smoothing red-black sweep
                                              from x = 0 as one launch (1) or two half sweeps (0),
                                              22: the post-smoothing red-black sweep of the coarse
                                              point-GS levels as one launch (1) or two (0),
                                              23: Krylov updates read the V-cycle results from the
                                              level-0 arrays they were left in (1) or from FP64 copies
                                              written by the products (0), 24: the level-0 line sweeps
                                              of the FP32 method read each node's seven BF16
                                              coefficients as one packed 16-byte record (1) or as
                                              seven 2-byte loads (0), 25: continuity assembly over
                                              LDS-staged tiles (1) or per node (0),
                                              26: quadratic sweep predictor (1) or the secant (0),
                                              27: FP32 method solves with FP32 Krylov inside FP64
                                              defect correction (1, on devices without a large
                                              floating block; 2 on any device) or with FP64 Krylov (0),
                                              28: level-1 line sweeps of the FP32 method read packed
                                              BF16 records (1) or the FP32 arrays (0) */

/* ---- learned warm start ----
   A 3-D convolutional U-Net on a G^3 grid resampled from the mesh (index
   space), BF16 matrix-core kernels with FP32 accumulation.  Input (8
   channels): the plain continuation's predicted psi, phi_n, phi_p, signed
   doping log-magnitude, bias step, and the predicted change of psi, phi_n,
   phi_p from the last converged state; output: the correction toward the
   converged state.
   sdx_ws_mode: 0 off, 1 record (prediction, converged) training samples at
   each bias point reached in one step, 2 apply the network's correction and
   keep it only if it lowers the starting Poisson residual (sdx_ws_accept),
   3 apply it always.
   stats: [0] corrections tried, [1] kept, [2] rejected, [3] samples recorded. */
void   sdx_ws_mode(int mode);
void   sdx_ws_stats(double *out);
void   sdx_ws_reset_stats(void);
void   sdx_ws_config(int chmask, double thr);   /* fields corrected: psi 1, phi_n 2, phi_p 4; smallest correction (V) */
void   sdx_ws_accept(double ratio);              /* mode 2 keeps a correction whose starting Poisson residual is
                                                    below ratio x the plain prediction's (default 0.5) */
/* floating majority blocks (float_fix, semiconductor_3d.c:3159-3228) since
   the last reset: out[0] continuity solves that listed unanchored majority
   nodes, out[1] blocks found, out[2] blocks rescaled to their balance */
void   sdx_float_stats(double *out);
void   sdx_float_reset_stats(void);
/* network: G a power of two >= 8, channels multiples of 32 up to 64 with c3 == c2 */
int    sdx_nn_init(int G, int c0, int c1, int c2, int c3, unsigned seed);
void   sdx_nn_free(void);
int    sdx_nn_nparams(void);
int    sdx_nn_get_params(float *out, int n);            /* per layer W[27][N][K] then b[N] */
int    sdx_nn_set_params(const float *in, int n);
void   sdx_nn_set_scale(double volts);                  /* target scale: output x scale = correction */
double sdx_nn_get_scale(void);
int    sdx_nn_forward_host(const float *in8, float *out3);          /* planar [c][x + G (y + G z)] */
double sdx_nn_grad_host(const float *in8, const float *tgt3, float *grad);   /* returns the MSE loss */
int    sdx_nn_dataset_alloc(int nmax);
int    sdx_nn_dataset_size(void);
int    sdx_nn_dataset_get(int i, float *in8, float *tgt3);
int    sdx_nn_dataset_put(const float *in8, const float *tgt3);   /* append a host sample; its index or -1 */
double sdx_nn_train(int steps, int n_train, double lr, unsigned seed, float *loss);  /* Adam; mean loss */
double sdx_nn_eval(int i0, int i1);                     /* mean loss over samples [i0, i1) */

/* device memory for callers without a HIP runtime of their own: allocate
   (null on failure), free, copy (dir 1 host->device, 2 device->host, 3
   device->device; 1 = done) */
void  *sdx_dev_alloc(long long bytes);
void   sdx_dev_free(void *p);
int    sdx_dev_copy(void *dst, const void *src, long long bytes, int dir);

/* ---- CSR export of a 7-point stencil operator ----
   The reference stores tensor-mesh operators as seven diagonals (Sten,
   semiconductor_3d.c:210) with no CSR form; this lists every nonzero entry of
   such an operator in CSR, columns ascending within a row (z-, y-, x-,
   diagonal, x+, y+, z+), so that its sparsity pattern can be compared bit for
   bit with the reference's assembled operator.  diag: a host array of 7
   device pointers c, xm, xp, ym, yp, zm, zp (N = Nx Ny Nz doubles each, node
   k = i + Nx (j + Ny l); c null: unit diagonal); rowptr (N + 1), col and val
   (cap entries) are device arrays.  Returns the number of nonzeros (entries
   past cap are not written), -1 bad arguments, -2 more nonzeros than an int
   row pointer holds, -3 a HIP call failed.  Allocates a scratch of one int
   per 1024 rows for the call; synchronises the null stream. */
long long sdx_csr_from_stencil(int Nx, int Ny, int Nz, const double *const *diag, int *rowptr, int *col,
                               double *val, long long cap);
/* The same export without allocating anything: no counterpart in the
   reference; for callers that cannot allocate device memory (code under HIP
   graph capture, tight serving loops, harnesses that forbid hipMalloc).  Its
   scan partials live in a fixed device array of the library (256 KB per GPU,
   held until the process ends), so it takes at most 67,108,864 rows: more
   return -4 before any device work.  Calls that overlap on one GPU share that
   array and race.  Same arguments and results otherwise. */
long long sdx_csr_from_stencil_noalloc(int Nx, int Ny, int Nz, const double *const *diag, int *rowptr, int *col,
                                       double *val, long long cap);

/* ---- kernel-level hooks on the prepared device (tests) ----
   A7: 7 blocks of N doubles (c, xm, xp, ym, yp, zm, zp). */
int    sdx_test_prepare(void);
void   sdx_test_set_state(const double *psi, const double *phin, const double *phip);
void   sdx_test_get_node(int which, double *out);   /* 0 psi 1 phin 2 phip 3 n 4 p 5 mu_n0 6 mu_p0 */
double sdx_test_poisson(int build, const double *Va, double *F, double *A7, double *sc);
void   sdx_test_upd_mob(double *fmu6);
void   sdx_test_set_fmu(const double *fmu6);
void   sdx_test_cont(int hole, double *A7, double *rhs, double *sc, double *dsc);
void   sdx_test_load_op(const double *A7, int qsign, const double *w);
/* graph devices (api3_set_mesh_graph): the operator last assembled as its
   diagonal and one value per directed edge (returns the edge count, -1 on a
   tensor device); loading such an operator; the edge mobilities, electrons
   then holes per undirected edge (set = 0 read, 1 replace; returns their
   count); the number of levels of the graph multigrid */
int    sdx_test_graph_op(double *dg, double *va);
int    sdx_test_graph_load_op(const double *dg, const double *va, int qsign, const double *w);
int    sdx_test_graph_gmu(double *gmu, int set);
int    sdx_test_graph_levels(void);
void   sdx_test_matvec(const double *x, double *y);
int    sdx_test_mg_setup(void);                     /* number of levels */
void   sdx_test_level_dims(int L, int *dims);      /* Nx Ny Nz cx cy cz */
void   sdx_test_level_op(int L, double *A7);        /* coarse operator, unscaled (rs * Ahat) */
void   sdx_test_precond(const double *r, double *z);
void   sdx_test_fill_mg(double value);              /* every multigrid solution array := value */
int    sdx_test_pcg(const double *b, double *x, int maxit, double rtol);
double sdx_test_bicgstab(const double *b, double *x, int maxit, double rtol, double atol, int *its);
void   sdx_test_currents(const double *Va, double *I, double *S, double *In);
double sdx_test_cont_update(int hole, const double *u, const double *sc, double *dq, double lo, double hi);
void   sdx_test_ii_update(double *gii);             /* impact-ionisation generation, 1/s per node */
int    sdx_test_tun_paths(int *ncol, int *pk, int *pc, int *pc0, int *pcn, double *pA, int *ck, double *cl,
                          double *ch);              /* gate tunnelling paths and supply columns */
int    sdx_test_tun(const double *Va, int hole, double *cK, double *pG, double *tk, double *tg);
                                                   /* one tunnelling update + the gathered sink/source */
void   sdx_test_aa_step(const double *x3, double *g3, int reset);
                                                   /* one Anderson step: x3, g3 the 3 blocks psi|phin|phip,
                                                      g3 overwritten by the mixed iterate; reset clears the history */
long long sdx_test_csr(int *rowptr, int *col, double *val, long long cap);
                                                   /* CSR of the operator last assembled (host arrays: N + 1
                                                      row pointers, cap entries); returns the nonzeros */
int    sdx_test_maj_flags(int hole, const double *u, const double *sc, unsigned char *f0, unsigned char *f1);
                                                   /* majority flags before/after the anchoring closure */

#ifdef __cplusplus
}
#endif
#endif
