// sd_linalg.h - launchers of the linear-algebra kernels (k_linalg.hip).
#pragma once
#include "sd_common.h"

// Deterministic reductions: every launch uses a fixed grid for a given n;
// block partials are summed in a fixed order by the last block to finish, so
// results are bit-identical from run to run.
struct RedBuf {
    double *part;      // [SD_RED_MAXV][SD_RED_MAXB]
    unsigned *cnt;     // completion counter (zero between launches)
    double *scal;      // result slots
};
#define SD_RED_MAXB 1024
#define SD_RED_MAXV 16

// Scalar slots used by the Krylov drivers.
enum {
    S_RZ = 0, S_RZ0, S_PQ, S_AL, S_BE, S_RHO, S_RHO1, S_RV, S_OM, S_SS, S_TT, S_TS, S_RR,
    S_FL, S_TMP0, S_TMP1, S_TMP2, S_TMP3, S_NRM, S_MAX0, S_MAX1, S_MAX2,
    S_AA0 = 32,             // 32..(32+2*12) Anderson dots
    S_CUR = 64,             // 64..(64+3*32) terminal currents I, S, In
    S_NSLOT = 192
};

// Multigrid level in value type T.  Level 0 optionally aliases the FP64
// operator; coarse levels are row-scaled to unit diagonal (c not stored) with
// the FP64 row scale rs (the coarse operator is diag(rs) * Ahat).
template <class T> struct MGLevel {
    int Nx, Ny, Nz, N;
    int cx, cy, cz;
    // storage of the level's node arrays: px = 0 natural order k = i + Nx (j +
    // Ny l); px = 1 each x-row split by parity (even i first: position
    // (i & 1) * ((Nx + 1) / 2) + i / 2), so that the lines of one zebra colour
    // across x are contiguous
    int px;
    Sten7<T> A;          // level 0: full stencil; coarse: c == nullptr (unit diagonal)
    double *rs;          // row scale used by restriction into the next level (nullptr = 1)
    T *pw;               // prolongation weight of each node to its parent
    double *ph;          // Slotboom reference potential / VT (coarse levels)
    T *x, *b;            // level vectors
    T *xs;               // second x: the fused post-smoothing sweep of a point-GS level reads x and
                         // writes here (coarse levels only)
    T *lu[3][3];         // line LU factors per axis: a/w, 1/w, c/w (nullptr: axis not factored)
    // BF16 copies of the off-diagonals (xm, xp, ym, yp, zm, zp) and of the
    // line factors, read by the line smoother instead of the T arrays when
    // set (FP32 level 0 with BF16 smoother coefficients)
    __bf16 *o16[6];
    __bf16 *lu16[3][3];
    // the same BF16 values packed per node into one 16-byte record for the
    // sweeps along axis a (rec16[a][k] = ai, iw, cp, then the four off-line
    // couplings in the order the sweep subtracts them, then 0), so that a
    // lane fetches a node's seven coefficients with one 16-byte load
    const uint4 *rec16[3];
};

// storage position of node (i, j, l) of a level
template <class T> __host__ __device__ inline int mg_xpos(const MGLevel<T> &v, int i) {
    return v.px ? (i & 1) * ((v.Nx + 1) >> 1) + (i >> 1) : i;
}
template <class T> __host__ __device__ inline int mg_idx(const MGLevel<T> &v, int i, int j, int l) {
    return mg_xpos(v, i) + v.Nx * (j + v.Ny * l);
}

// vector / reduction launchers (stream-ordered, no host sync)
void sd_dot(hipStream_t s, const double *a, const double *b, int n, RedBuf rb, int slot, int post = -1);
void sd_dot2(hipStream_t s, const double *a, const double *b, const double *c, int n, RedBuf rb,
             int slot_ab, int slot_ac, int post = -1);   // a.b and a.c
void sd_maxabs(hipStream_t s, const double *a, int n, RedBuf rb, int slot);
void sd_maxabsdiff(hipStream_t s, const double *a, const double *b, int n, RedBuf rb, int slot);
void sd_fill(hipStream_t s, double *a, double v, int n);
void sd_copy(hipStream_t s, double *dst, const double *src, int n);
void sd_axpy(hipStream_t s, double *y, const double *x, double a, int n);     // y += a x

// Level-0 staging of the preconditioner's input and output.  The Krylov
// kernels that produce the V-cycle's right-hand side also write it into the
// hierarchy's level-0 b (its precision and x order), and the kernels that
// consume the V-cycle's result read it from the level-0 x, so the V-cycle
// needs no conversion passes.  b == nullptr: staging off.
struct Stage0 {
    void *b;            // level-0 right-hand side (float when f32, else double)
    const void *x;      // level-0 solution, same storage
    int Nx, px, f32;    // x-row length, parity-split x order, precision
    const void *x2;     // a second level-0 solution (the second V-cycle of a BiCGSTAB iteration)
};

// y = A x, or y = b - A x (b != nullptr), FP64 level-0 operator
void sd_spmv(hipStream_t s, const Geo &g, Sten7<double> A, const double *x, const double *b, double *y);
// y = A x fused with dots: slot <- y.(u0 or y), slot+1 <- y.u1 (nv = 1 or 2); post: SOP_*
// in.x != nullptr: x is read from the staged level-0 solution instead, and
// copied to x (FP64, natural order) on the way unless x == nullptr
void sd_spmv_dots(hipStream_t s, const Geo &g, Sten7<double> A, const double *x, double *y, const double *u0,
                  const double *u1, int nv, RedBuf rb, int slot, int post = -1, Stage0 in = Stage0{});
// slot <- r.z with z the staged level-0 solution, also copied to z (FP64,
// natural order) unless z == nullptr
void sd_dot_stage(hipStream_t s, const double *r, double *z, Stage0 st, int n, RedBuf rb, int slot, int post);
// bicgstab rounding floor sum_k (sum_j |A_kj x_j| + |b_k|)^2 -> slot
void sd_floor(hipStream_t s, const Geo &g, Sten7<double> A, const double *x, const double *b,
              RedBuf rb, int slot);

// Krylov fused updates (scalars read from rb.scal on the device)
void sd_pcg_update_xr(hipStream_t s, double *x, double *r, const double *p, const double *q, int n, RedBuf rb,
                      Stage0 st = Stage0{});                                   // stages r
void sd_pcg_update_p(hipStream_t s, double *p, const double *z, int n, RedBuf rb);
void sd_pcg_update_p_st(hipStream_t s, double *p, Stage0 st, int n, RedBuf rb);   // z = staged st.x
void sd_scalar_op(hipStream_t s, int op, RedBuf rb);
enum { SOP_PCG_AL = 0, SOP_PCG_BE, SOP_BI_BE, SOP_BI_AL, SOP_BI_OM, SOP_BI_NEXT, SOP_BI_RESTART, SOP_BI_NEXT_BE };
void sd_bi_p(hipStream_t s, double *p, const double *r, const double *v, int n, RedBuf rb,
             Stage0 st = Stage0{});                                            // stages p
void sd_bi_r1(hipStream_t s, double *r, const double *v, int n, RedBuf rb,
              Stage0 st = Stage0{});                                           // r -= al v; SS; stages r
void sd_bi_x1(hipStream_t s, double *x, const double *ph, int n, RedBuf rb);   // x += al ph
void sd_bi_x1_st(hipStream_t s, double *x, Stage0 st, int n, RedBuf rb);       // ph = staged st.x
// x += al ph + om sh; r -= om t; RR = r.r.  With rh: also rho = rho1, rho1 =
// rh.r and the next iteration's beta (the dot of its first step)
void sd_bi_xr2(hipStream_t s, double *x, double *r, const double *ph, const double *sh, const double *t,
               const double *rh, int n, RedBuf rb);
// the same with ph and sh read from the staged level-0 solutions st.x and st.x2
void sd_bi_xr2_st(hipStream_t s, double *x, double *r, Stage0 st, const double *t, const double *rh, int n,
                  RedBuf rb);
void sd_bi_init(hipStream_t s, double *r, double *rh, double *p, double *v, const double *b,
                const double *t, int n, RedBuf rb);                             // RR
void sd_bi_restart(hipStream_t s, double *rh, double *p, double *v, const double *r, int n);

// multigrid (T = double or float)
template <class T> void mg_weights_level(hipStream_t s, const MGLevel<T> &f, MGLevel<T> &c, int L,
                                         const DevView &d, int qsign, double VT);
template <class T> void mg_galerkin(hipStream_t s, const MGLevel<T> &f, MGLevel<T> &c);
// red-black Gauss-Seidel: zero 0 one half sweep (colour `color`), 1 that half
// sweep from x = 0, 2 the whole sweep from x = 0 (`color` first) in one launch
template <class T> void mg_rbgs(hipStream_t s, const MGLevel<T> &v, const T *b, T *x, int color, int zero,
                                const MGLevel<T> *c = nullptr);
// post-smoothing red-black sweep in one launch: colour 1 from its neighbours
// in x with the coarse correction pw * c.x added, then colour 0 from the new
// colour-1 values; reads x, writes xo (every node)
template <class T> void mg_rbgs_post2(hipStream_t s, const MGLevel<T> &v, const T *b, const T *x, T *xo,
                                      const MGLevel<T> &c);
// alpha scales the coarse right-hand side, and so the coarse correction
// (over-correction of the aggregation coarse space; 1 = plain)
template <class T> void mg_resid_restrict(hipStream_t s, const MGLevel<T> &f, const MGLevel<T> &c,
                                          const T *b, const T *x, int skip, double alpha = 1.0);
// levels d_levels[0..nt-1] (device copies; the last one the coarsest) in one workgroup
template <class T> void mg_tail(hipStream_t s, const MGLevel<T> *d_levels, const MGLevel<T> *h_levels, int nt,
                                const double *inv, int nsw);
// LU factors of the lines along axis ax (once per multigrid setup)
template <class T> void mg_line_factor(hipStream_t s, const MGLevel<T> &v, int ax);
// zebra line Gauss-Seidel along axis ax on lines of one colour (0 = cannot: line too long)
template <class T> int mg_line(hipStream_t s, const MGLevel<T> &v, const T *b, T *x, int ax, int color, int zero,
                               const MGLevel<T> *c = nullptr);
template <class T> void mg_prolong(hipStream_t s, const MGLevel<T> &f, const MGLevel<T> &c, T *x);
template <class T> void mg_coarse_factor(hipStream_t s, const MGLevel<T> &v, double *inv);
template <class T> void mg_coarse_solve(hipStream_t s, const MGLevel<T> &v, const double *inv, const T *b, T *x);
template <class T> void mg_set_pw_one(hipStream_t s, MGLevel<T> &v);
extern int g_line_nl_max;
extern int g_line_scan;
extern int g_line_slots;
extern int g_line_minwg;
extern int g_line_xcd;
extern int g_tail_lds;
extern int g_line_scan1;        // axis-specialised line kernel for one line per slot
extern int g_spmv_nt;            // non-temporal operator loads in the Krylov SpMV
extern int g_restrict_act;      // active-lane restriction after an exact half sweep
extern int g_rb_fuse;           // pre-smoothing red-black sweep from x = 0 in one launch
extern int g_rb_fuse_post;      // post-smoothing red-black sweep in one launch (coarse levels)
void sd_set_tile_mode_linalg(int m);   // c_tile_mode of k_linalg.hip
void sd_copy_bw(hipStream_t s, const double *a, double *b, long n);   // plain copy (bandwidth reference)
// level-0 data in the level's storage order and precision: the right-hand
// side in, the correction out, the operator and the restriction weights
template <class T> void mg_vec_in(hipStream_t s, const double *r, const MGLevel<T> &v, T *b);
template <class T> void mg_vec_out(hipStream_t s, const T *x, const MGLevel<T> &v, double *z);
template <class T> void mg_sten_in(hipStream_t s, Sten7<double> A, const MGLevel<T> &v, Sten7<T> B);
// BF16 copies of a level's off-diagonals and of the line factors of axis ax
void mg_o16(hipStream_t s, const MGLevel<float> &v);
void mg_lu16(hipStream_t s, const MGLevel<float> &v, int ax);
void mg_rec16(hipStream_t s, const MGLevel<float> &v, int ax, uint4 *rec);
template <class T> void mg_rs_in(hipStream_t s, const double *w, const MGLevel<T> &v, double *rs);
template <class T> void sd_spmv_T(hipStream_t s, const Geo &g, Sten7<T> A, const T *x, const T *b, T *y);

// ------------------------------------------------ FP32 inner Krylov (k_krylov32.hip)
// Level-0 geometry and storage order of the FP32 multigrid: the inner solves'
// vectors are stored as the V-cycle reads and writes them.
struct L0Geo {
    int Nx, Ny, Nz, px;
};
// y = A x with the FP32 level-0 operator (unit: diagonal not read); slot <-
// y.(u0 or y), slot+1 <- y.u1 (nv 2); FP64 accumulation; then scalar op post
void k32_spmv_dots_launch(hipStream_t s, const L0Geo &g, Sten7<float> A, bool unit, const float *x, float *y,
                          const float *u0, const float *u1, int nv, RedBuf rb, int slot, int post);
// r = b - A x in FP64 (FP64 operator, natural order), stored FP32 at level-0
// positions; r.r -> slot.  rh != nullptr: also rh = r, p = v = d = 0
void k32_resid_launch(hipStream_t s, const L0Geo &g, Sten7<double> A, bool unit, const double *b, const double *x,
                      float *r, float *rh, float *p, float *v, float *d, RedBuf rb, int slot);
void k32_add_out_launch(hipStream_t s, const L0Geo &g, double *x, const float *d);      // x += d (FP64)
void k32_dot_launch(hipStream_t s, const float *a, const float *b, int n, RedBuf rb, int slot, int post = -1);
void k32_fill_launch(hipStream_t s, float *a, float v, int n);
void k32_copy_launch(hipStream_t s, float *d, const float *a, int n);
void k32_pcg_xr_launch(hipStream_t s, float *d, float *r, const float *p, const float *q, int n, RedBuf rb);
void k32_pcg_p_launch(hipStream_t s, float *p, const float *z, int n, RedBuf rb);
void k32_bi_p_launch(hipStream_t s, float *p, const float *r, const float *v, int n, RedBuf rb);
void k32_bi_s_launch(hipStream_t s, float *r, const float *v, int n, RedBuf rb, int slot);
void k32_bi_x1_launch(hipStream_t s, float *d, const float *ph, int n, RedBuf rb);
void k32_bi_xr2_launch(hipStream_t s, float *d, float *r, const float *ph, const float *sh, const float *t,
                       const float *rh, int n, RedBuf rb);
void k32_bi_restart_launch(hipStream_t s, float *rh, float *p, float *v, const float *r, int n);
