// sd_common.h - device data layout and per-node physics of the MI355X
// drift-diffusion core.  Everything here is shared by the kernel files and
// the host engine.
//
// Layout: every node quantity is a structure-of-arrays FP64 vector over the
// tensor mesh, node index k = i + Nx*(j + Ny*l) (x fastest, as the reference
// IDX macro, semiconductor_3d.c:114).  Matrices are 7-point stencils stored as
// 7 diagonal vectors (c, xm, xp, ym, yp, zm, zp), semiconductor_3d.c:210.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cmath>

#define SD_Q      1.602176634e-19
#define SD_KB     1.380649e-23
#define SD_EPS0   8.8541878128e-12
#define SD_EPS_OX (3.9 * SD_EPS0)
#define SD_LN_MIN (-230.0)
#define SD_LN_MAX 150.0
#define SD_NFLOOR 1e-100
#define SD_NM_MAX 16
#define SD_NC_MAX 32
#define SD_BLOCK  256

enum { K_SEMI = 0, K_INS = 1, K_METAL = 2 };

// Material parameters the hot path needs (subset of the reference Mat,
// semiconductor_3d.c:127-188), in SI units.
struct MatDev {
    double VT, lnc, lnv, ni, Eg, chi;
    double tau_n0, tau_p0, Cn, Cp, Brad, tau_Nref;
    double vsat_n, vsat_p, beta_n, beta_p;
    double ani_n, ani_p;
    double ii_a[2][2], ii_b[2][2], ii_E0[2], ii_m[2], ii_g;
    int insulator, ii_on;
    double T;                            // K
    int lb_on;                           // Lombardi surface mobility: bit 0 electrons, bit 1 holes
    double lbB[2], lbC[2], lbL[2], lbD[2], lbE[2];
    // gate tunnelling: relative permittivity, raw band-edge densities of
    // states (m^-3), supply masses of a semiconductor and tunnelling masses
    // of an insulator (units of m0)
    double eps_r, Nc, Nv, ms_e, ms_h, mt_e, mt_h;
};

// Gate tunnelling paths (semiconductor node -> gate metal through a layer
// stack) and their supply columns, semiconductor_3d.c:1147-1247, 3230-3390.
// Paths are stored grouped by contact (poff), in the reference's order.
#define SD_TUN_MAXL 6
struct TunDev {
    int np, ntn;
    const int *pk, *pc, *pnl, *pc0, *pcn;     // node, contact, layers, first column entry, column length
    const int *pm;                            // SD_TUN_MAXL material indices per path
    const double *pA, *pt;                    // area (m^2); SD_TUN_MAXL layer thicknesses (m) per path
    const int *ck;                            // column entry: node
    const double *cl, *ch;                    // column entry: cell length along the path, spacing to the next
    double *cK;                               // linearised escape (m^3/s): 2 per column entry (e, h)
    double *pG;                               // injection from the metal (1/s): 2 per path (e, h)
    const int *tn;                            // nodes touched by a column, ascending
    const int *offK, *entK;                   // per touched node: its column entries (ascending)
    const int *offG, *entG;                   // per touched node: the paths starting there (ascending)
    const double *qcn, *qcp;                  // MLDA terms (nullptr: none)
};

// Per-contact data read by kernels (semiconductor_3d.c:191-208).
struct ConDev {
    double V;        // applied voltage of the current (sub)step
    double phi_m;    // work function (eV), 0 = reference
    double Cox;      // EPS_OX / tox
    double VGp;      // V + Cref - phi_m (Robin gate)
    double qox;      // fixed oxide charge (C/m^2)
    int bc, te;
};

// Tensor mesh geometry (device pointers, Nx+Ny+Nz doubles each).
struct Geo {
    int Nx, Ny, Nz, N;
    const double *xs, *ys, *zs, *Lx, *Ly, *Lz;
};

// Everything the physics kernels read or write, as raw device pointers.
struct DevView {
    Geo g;
    const unsigned char *kind;
    const int *con_mask, *gate_mask;
    const unsigned char *mat;
    const double *Nd, *Na, *Nf, *eps, *bgn, *robA;
    const double *lncK, *lnvK, *niK;     // per-node ln Nc_eff, ln Nv_eff, n_i,eff
    double *psi, *phin, *phip, *n, *p;
    double *mu_n, *mu_p;                 // low-field node mobility
    double *fmu;                         // 6 face mobilities, block q at fmu + q*N
    double *gii;                         // impact-ionisation generation (1/s per node)
    const double *qd;                    // distance to the nearest interface (m); nullptr: no Lombardi
    const double *tk, *tg;               // gate tunnelling of the carrier being solved: sink rate
                                         // (m^3/s, times the density) and source (1/s); nullptr: none
    const MatDev *mats;
    const ConDev *cons;
    int models, caxis, ref_mat;
    double Cref;
};

// 7-point stencil (level 0 in FP64, coarse levels in T).
template <class T> struct Sten7 {
    T *c, *xm, *xp, *ym, *yp, *zm, *zp;
};

__device__ __forceinline__ void sd_unidx(const Geo &g, int k, int &i, int &j, int &l) {
    int NxNy = g.Nx * g.Ny;
    l = k / NxNy;
    int r = k - l * NxNy;
    j = r / g.Nx;
    i = r - j * g.Nx;
}

// Bernoulli pair B(x) = x/(e^x - 1) and B(-x) = B(x) + x, evaluated from the
// smaller of the two (one expm1 per edge, no cancellation).  Branches follow
// the reference B(), semiconductor_3d.c:670-675: series below 1e-6, 0 above
// 700.
__device__ __forceinline__ void sd_bern2(double x, double &bp, double &bm) {
    double a = fabs(x), s;
    if (a < 1e-6) {
        bp = 1.0 - x * 0.5 + x * x / 12.0;
        bm = 1.0 + x * 0.5 + x * x / 12.0;
        return;
    }
    s = (a > 700.0) ? 0.0 : a / expm1(a);     // B(|x|)
    if (x > 0) { bp = s; bm = s + a; }
    else       { bm = s; bp = s + a; }
}

__device__ __forceinline__ double sd_clampe(double e) {
    return e < SD_LN_MIN ? SD_LN_MIN : (e > SD_LN_MAX ? SD_LN_MAX : e);
}

// SRH (midgap) + Auger + radiative recombination and its derivatives,
// semiconductor_3d.c:686-697.
__device__ __forceinline__ void sd_recomb(const MatDev &m, double ni, double tf, double n, double p,
                                          double &R, double &dRn, double &dRp) {
    double ni2 = ni * ni, np = n * p - ni2;
    double den = tf * (m.tau_p0 * (n + ni) + m.tau_n0 * (p + ni)) + 1e-300;
    double Rs = np / den;
    double dsn = (p * den - np * tf * m.tau_p0) / (den * den);
    double dsp = (n * den - np * tf * m.tau_n0) / (den * den);
    double Au = m.Cn * n + m.Cp * p;
    R = Rs + Au * np + m.Brad * np;
    dRn = dsn + m.Cn * np + Au * p + m.Brad * p;
    dRp = dsp + m.Cp * np + Au * n + m.Brad * n;
}

// SRH lifetime doping factor, semiconductor_3d.c:787-791.
__device__ __forceinline__ double sd_tauf(const DevView &d, int k, const MatDev &m) {
    if (!(d.models & 2) || !(m.tau_Nref > 0.)) return 1.0;
    return 1.0 / (1.0 + (d.Nd[k] + d.Na[k]) / m.tau_Nref);
}

// Caughey-Thomas velocity saturation, semiconductor_3d.c:664-668.
__device__ __forceinline__ double sd_mob_hf(double mu0, double E, double vsat, double beta) {
    double r = mu0 * fabs(E) / vsat;
    if (r < 1e-12) return mu0;
    return mu0 / pow(1.0 + pow(r, beta), 1.0 / beta);
}

// Driving field of an edge: harmonic mean of the quasi-Fermi and
// electrostatic drops when they have the same sign, semiconductor_3d.c:2882-2886.
__device__ __forceinline__ double sd_heat_field(double dq, double dps, double h) {
    if (dq * dps <= 0.) return 0.;
    double a = fabs(dq), b = fabs(dps);
    return 2. * a * b / ((a + b) * h);
}

// Lombardi surface mobility combined with the bulk value mub (m^2/Vs) by
// Matthiessen's rule, weighted by the interface proximity Dw; F normal field
// (V/m), Nt total doping (m^-3); lombardi, semiconductor_3d.c:2924-2934
// (without the Coulomb term, which only 4H-SiC carries)
__device__ __forceinline__ double sd_lombardi(const MatDev &m, double mub, double Fv, double Nt, double Dw, int hole) {
    double F = Fv * 1e-2; if (F < 1.) F = 1.;                       // V/cm
    double N = Nt * 1e-6; if (N < 1.) N = 1.;                       // cm^-3
    const double Bc = m.lbB[hole], Cc = m.lbC[hole], Lc = m.lbL[hole], Dc = m.lbD[hole], Ec = m.lbE[hole];
    const double muac = Bc / F + Cc * pow(N, Lc) / (cbrt(F) * (m.T / 300.));
    const double musr = Ec > 0. ? 1. / (F * F / Dc + F * F * F / Ec) : Dc / (F * F);
    return 1. / (1. / mub + Dw * 1e4 * (1. / muac + 1. / musr));    // cm^2/Vs -> m^2/Vs
}
#define SD_LB_LCRIT 1e-8      // interface-proximity length of the Lombardi terms (m)

// carrier densities at node k (qcn = qcp = 0: no quantum confinement)
__device__ __forceinline__ double sd_nK(const DevView &d, int k, double psi, double qn, double VT) {
    return exp(sd_clampe(d.lncK[k] + (psi - qn) / VT));
}
__device__ __forceinline__ double sd_pK(const DevView &d, int k, double psi, double qp, double VT) {
    return exp(sd_clampe(d.lnvK[k] + (qp - psi) / VT));
}

// Index of a neighbour load: the neighbour's, or the node's own where there
// is none, so that every neighbour load can be issued before the first use.
// The empty asm hides the value from the compiler, which would otherwise fold
// the clamped load into the node's own load of the same array and move the
// neighbour's back into a branch.
__device__ __forceinline__ int sd_nb_idx(bool has, int nb, int self) {
    int v = has ? nb : self;
    asm("" : "+v"(v));
    return v;
}
// The six neighbours of node k = (i, j, l) at once (directions as sd_nbr):
// whether each exists, its index (k itself where there is none, so loads
// through it stay in bounds), and the spacing h and dual face area ar, with
// sd_nbr's values (the same expressions) where the neighbour exists.
struct Nbr6 {
    bool has[6];
    int kb[6];
    double h[6], ar[6];
};
__device__ __forceinline__ void sd_nbr6(const Geo &g, int i, int j, int l, int k, Nbr6 &n) {
    const int NxNy = g.Nx * g.Ny;
    n.has[0] = i > 0; n.has[1] = i < g.Nx - 1; n.has[2] = j > 0;
    n.has[3] = j < g.Ny - 1; n.has[4] = l > 0; n.has[5] = l < g.Nz - 1;
    n.kb[0] = sd_nb_idx(n.has[0], k - 1, k); n.kb[1] = sd_nb_idx(n.has[1], k + 1, k);
    n.kb[2] = sd_nb_idx(n.has[2], k - g.Nx, k); n.kb[3] = sd_nb_idx(n.has[3], k + g.Nx, k);
    n.kb[4] = sd_nb_idx(n.has[4], k - NxNy, k); n.kb[5] = sd_nb_idx(n.has[5], k + NxNy, k);
    const int im = n.has[0] ? i - 1 : i, ip = n.has[1] ? i + 1 : i, jm = n.has[2] ? j - 1 : j;
    const int jp = n.has[3] ? j + 1 : j, lm = n.has[4] ? l - 1 : l, lp = n.has[5] ? l + 1 : l;
    n.h[0] = g.xs[i] - g.xs[im]; n.h[1] = g.xs[ip] - g.xs[i];
    n.h[2] = g.ys[j] - g.ys[jm]; n.h[3] = g.ys[jp] - g.ys[j];
    n.h[4] = g.zs[l] - g.zs[lm]; n.h[5] = g.zs[lp] - g.zs[l];
    n.ar[0] = n.ar[1] = g.Ly[j] * g.Lz[l];
    n.ar[2] = n.ar[3] = g.Lx[i] * g.Lz[l];
    n.ar[4] = n.ar[5] = g.Lx[i] * g.Ly[j];
}

// neighbour geometry along direction dir (0 x-,1 x+,2 y-,3 y+,4 z-,5 z+),
// semiconductor_3d.c:703-713.  Returns false outside the domain.
__device__ __forceinline__ bool sd_nbr(const Geo &g, int i, int j, int l, int k, int dir,
                                       int &kb, double &h, double &ar) {
    switch (dir) {
    case 0: if (i == 0) return false; kb = k - 1; h = g.xs[i] - g.xs[i - 1]; ar = g.Ly[j] * g.Lz[l]; return true;
    case 1: if (i == g.Nx - 1) return false; kb = k + 1; h = g.xs[i + 1] - g.xs[i]; ar = g.Ly[j] * g.Lz[l]; return true;
    case 2: if (j == 0) return false; kb = k - g.Nx; h = g.ys[j] - g.ys[j - 1]; ar = g.Lx[i] * g.Lz[l]; return true;
    case 3: if (j == g.Ny - 1) return false; kb = k + g.Nx; h = g.ys[j + 1] - g.ys[j]; ar = g.Lx[i] * g.Lz[l]; return true;
    case 4: if (l == 0) return false; kb = k - g.Nx * g.Ny; h = g.zs[l] - g.zs[l - 1]; ar = g.Lx[i] * g.Ly[j]; return true;
    default: if (l == g.Nz - 1) return false; kb = k + g.Nx * g.Ny; h = g.zs[l + 1] - g.zs[l]; ar = g.Lx[i] * g.Ly[j]; return true;
    }
}

// deterministic block reduction helpers (wavefront shuffle, then LDS)
template <class T>
__device__ __forceinline__ T sd_wave_sum(T v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    return v;
}
// NaN-propagating maximum (a NaN anywhere in a max reduction survives)
template <class T>
__device__ __forceinline__ T sd_nanmax(T a, T b) { return (b != b || b > a) ? b : a; }
template <class T>
__device__ __forceinline__ T sd_wave_max(T v) {
    for (int o = 32; o > 0; o >>= 1) { T w = __shfl_down(v, o, 64); v = sd_nanmax(v, w); }
    return v;
}
// sum over a 256-thread block; result valid in thread 0
template <class T>
__device__ __forceinline__ T sd_block_sum(T v, T *lds) {
    v = sd_wave_sum(v);
    int w = threadIdx.x >> 6, ln = threadIdx.x & 63;
    if (ln == 0) lds[w] = v;
    __syncthreads();
    T s = 0;
    if (threadIdx.x == 0) { for (int q = 0; q < (int)(blockDim.x >> 6); q++) s += lds[q]; }
    __syncthreads();
    return s;
}
template <class T>
__device__ __forceinline__ T sd_block_max(T v, T *lds) {
    v = sd_wave_max(v);
    int w = threadIdx.x >> 6, ln = threadIdx.x & 63;
    if (ln == 0) lds[w] = v;
    __syncthreads();
    T s = lds[0];
    if (threadIdx.x == 0) { for (int q = 1; q < (int)(blockDim.x >> 6); q++) s = sd_nanmax(s, lds[q]); }
    __syncthreads();
    return s;
}

#ifndef SD_TILE_MODE_DEFAULT
#define SD_TILE_MODE_DEFAULT 1
#endif
inline int sd_nblocks(long n, int bs = SD_BLOCK) { return (int)((n + bs - 1) / bs); }

// Workgroup b runs on XCD b % 8, and each of the eight XCDs has its own L2.
// sd_tile() renumbers the workgroups so that each XCD takes a contiguous
// range of tiles: stencil neighbours read across tile boundaries (the next
// row, the next plane) then mostly come from the same L2.  Identity when the
// grid is not a multiple of 8.
//
// c_tile_mode (one copy per source file, set by sd_set_tile_mode_*):
//   0  renumbered tiles; node loops cover one contiguous range per workgroup
//   1  renumbered tiles; node loops are grid-stride over the renumbered tiles
//   2  launch order; node loops grid-stride
static __constant__ int c_tile_mode = SD_TILE_MODE_DEFAULT;
__device__ __forceinline__ int sd_tile() {
    return (c_tile_mode == 2 || (gridDim.x & 7)) ? (int)blockIdx.x
                                                 : (int)((blockIdx.x & 7) * (gridDim.x >> 3) + (blockIdx.x >> 3));
}
// coordinates (i, j, l) of the node of a strided node loop, advanced without
// a division per node: the stride is split into (di, dj, dl) once
struct NodeWalk {
    int i, j, l, di, dj, dl, Nx, Ny;
    __device__ __forceinline__ NodeWalk(int k, int step, int nx, int ny) : Nx(nx), Ny(ny) {
        const int NxNy = nx * ny;
        l = k / NxNy;
        int r = k - l * NxNy;
        j = r / nx;
        i = r - j * nx;
        dl = step / NxNy;
        r = step - dl * NxNy;
        dj = r / nx;
        di = r - dj * nx;
    }
    __device__ __forceinline__ void next() {
        i += di; j += dj; l += dl;
        if (i >= Nx) { i -= Nx; j++; }
        if (j >= Ny) { j -= Ny; l++; }
    }
};
// this thread's first node, stride and end for a loop over n nodes by a grid
// of gridDim.x workgroups of SD_BLOCK threads
__device__ __forceinline__ void sd_node_loop(int n, int &k0, int &step, int &k1) {
    if (c_tile_mode == 0) {
        const int per = ((n + (int)gridDim.x - 1) / (int)gridDim.x + SD_BLOCK - 1) / SD_BLOCK * SD_BLOCK;
        const int t0 = sd_tile() * per;
        k0 = t0 + (int)threadIdx.x;
        k1 = t0 + per < n ? t0 + per : n;
        step = SD_BLOCK;
    } else {
        k0 = sd_tile() * SD_BLOCK + (int)threadIdx.x;
        k1 = n;
        step = (int)gridDim.x * SD_BLOCK;
    }
}
