// sd_graph.h - the graph (general box-method mesh) path: device layout and
// kernel launchers (k_graph.hip).
//
// A graph device (api3_set_mesh_graph, semiconductor_3d.c:4788-4804) is N
// nodes and a symmetric CSR graph of directed edges; every edge carries its
// length h and its dual face area ar, every node its control volume.  Node
// arrays keep the tensor path's layout with Nx = N, Ny = Nz = 1 and the
// control volume in Lx (Ly = Lz = 1), so every node-local kernel of the
// tensor path applies unchanged; the kernels here are the ones that walk
// edges (assembly, edge mobilities, terminal currents, operator products
// and the aggregation multigrid).
#pragma once
#include "sd_physics.h"

struct GraphDev {
    int N, E, nue;                 // nodes, directed edges, undirected edges
    const int *rp, *cj;            // row pointers (N+1), neighbour of each directed edge
    const int *ue;                 // undirected edge id of each directed edge
    const double *h, *ar;          // edge length (m), dual face area (m^2)
    const double *vol;             // control volume (m^3)
    const double *pos;             // node positions, 3 per node (m)
    const double *lsq;             // least-squares gradient matrix M^-1 per node (xx xy xz yy yz zz)
    double *grad;                  // gradient of psi per node (3), Lombardi surface mobility
};

// Operator on a graph level: diagonal dg (nullptr: unit diagonal, not read)
// and one off-diagonal value per directed edge, in the CSR order of rp/cj.
struct GOp {
    int N;
    const int *rp, *cj;
    double *va, *dg;
};

// One level of the aggregation multigrid.  Aggregation into the next level:
// agg[i] is the aggregate of node i; the members of aggregate I are
// mem[mrp[I] .. mrp[I+1]); emap[e] is the coarse entry an edge adds to (-1:
// both ends in one aggregate, the edge adds to the coarse diagonal).
// Multicolour order of the smoother: perm lists the nodes grouped by colour.
struct GLev {
    GOp A;
    const int *agg, *mrp, *mem, *emap;
    double *pw;                    // prolongation weight of each node to its aggregate
    double *ph;                    // Slotboom reference potential / VT (coarse levels)
    const int *perm;
    const int *coff;               // colour c: perm[coff[c] .. coff[c+1]) (device copy)
    int ncol;
    double *x, *b;
};

// The small levels of the hierarchy (levels L0 .. nl-1), run by one
// workgroup in one launch (gla_tail): the coarsest by its explicit inverse
// (dense) or by Gauss-Seidel sweeps
#define GTAIL_MAXL 16
struct GTail {
    GLev lev[GTAIL_MAXL];
    int nl, dense;
    const double *inv;
};

// physics (one launch each; the reductions are deterministic)
void gph_poisson(hipStream_t s, const DevView &d, const GraphDev &g, const double *phi, int build, double *F,
                 GOp A, double *sc, RedBuf rb, int slot);
void gph_upd_mob(hipStream_t s, const DevView &d, const GraphDev &g, const double *qn, const double *qp,
                 double *gmu);
void gph_cont(hipStream_t s, const DevView &d, const GraphDev &g, int hole, const double *gmu, GOp A, double *rhs,
              double *sc, double *dsc);
// terminal currents: edges (node ck[q], directed edge ce[q]) of contact c are
// q in [off[c], off[c+1]); results in rb.scal[S_CUR + 3c ..] as k_currents
void gph_currents(hipStream_t s, const DevView &d, const GraphDev &g, const double *gmu, const int *ck,
                  const int *ce, const int *off, int ncon, RedBuf rb);
// 1: majority node of a floating-block candidate, 0 otherwise (float_fix)
void gph_maj_flags(hipStream_t s, const DevView &d, int hole, const double *u, const double *sc,
                   unsigned char *flag);
// display quantities: node current density (least squares over the edge
// fluxes), recombination, node mobility as the mean over its edges
void gph_display(hipStream_t s, const DevView &d, const GraphDev &g, const double *gmu, double *Jx, double *Jy,
                 double *Jz, double *R, double *mun, double *mup);
// the electron (carrier 1) or hole (carrier 2) node current density along ax
void gph_node_J(hipStream_t s, const DevView &d, const GraphDev &g, const double *gmu, int ax, int carrier,
                double *out);

// linear algebra
void gla_spmv(hipStream_t s, GOp A, const double *x, const double *b, double *y);
// lanes: lanes per row (4, 8, 16, 32; 1: one lane per row)
void gla_spmv_dots(hipStream_t s, GOp A, const double *x, double *y, const double *u0, const double *u1, int nv,
                   RedBuf rb, int slot, int post, int lanes = 1);
// lanes per row suited to a level of n rows and nnz off-diagonal entries
int gla_lanes(int n, int nnz);
void gla_floor(hipStream_t s, GOp A, const double *x, const double *b, RedBuf rb, int slot);
void gla_weights(hipStream_t s, const DevView &d, int L, int nc, const GLev &f, const GLev &c, int qsign, double VT);
void gla_pw_one(hipStream_t s, int n, double *pw);
void gla_galerkin(hipStream_t s, int nc, const GLev &f, const GLev &c, const double *w);
void gla_gs(hipStream_t s, GOp A, const double *b, double *x, const int *perm, int n0, int n1, int lanes = 1);
// one l1-Jacobi sweep from xi (nullptr: 0) into xo
void gla_jacobi(hipStream_t s, GOp A, const double *b, const double *xi, double *xo, int lanes);
// r != nullptr: the residual is formed one lane per row into r (member order,
// N values) and summed per aggregate by a second launch
void gla_resid_restrict(hipStream_t s, int nc, const GLev &f, const double *b, const double *x, const double *w,
                        double *bc, double *r = nullptr, int lanes = 8);
void gla_prolong(hipStream_t s, int n, const GLev &f, const double *xc, double *x, double alpha);
void gla_dense_inverse(hipStream_t s, GOp A, double *work, double *inv);
void gla_dense_apply(hipStream_t s, int n, const double *inv, const double *b, double *x);
// V-cycle of levels L0 .. nl-1 from level L0's right-hand side b0 into x0
void gla_tail(hipStream_t s, const GTail &T, int L0, const double *b0, double *x0);
