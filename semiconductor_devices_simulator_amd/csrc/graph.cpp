// graph.cpp - host side of the graph (general box-method mesh) path.
//
// api3_set_mesh_graph (semiconductor_3d.c:4795-4804, alloc_dev_graph
// :981-1040) hands over a symmetric CSR graph; the engine keeps every node
// array of the tensor path with Nx = N, Ny = Nz = 1 and the control volumes in
// Lx, so the node-local kernels and the solver drivers (Newton, Gummel,
// continuation, Anderson) run unchanged, and the edge-walking work goes to
// k_graph.hip.
//
// Multigrid (replaces gmg_alloc/gmg_setup/gvcycle :2467-2705): aggregation by
// pairwise matching on the geometric couplings ar/h, three matching passes
// per level (aggregates of up to 8 nodes that follow the strong direction of
// anisotropic cells; Notay, "An aggregation-based algebraic multigrid method",
// ETNA 37 (2010), whose pairwise matching the reference also uses), built
// once per mesh on the host with the coarse sparsity patterns, the
// fine-entry -> coarse-entry maps and a greedy colouring of every level.  Per
// linear solve the device computes the Slotboom weights and the Galerkin
// products; the V-cycle smooths with multicolour Gauss-Seidel (forward before
// the coarse correction, backward after it, so the cycle is symmetric for CG)
// and solves the coarsest level with an explicit inverse.
#include "engine.h"
#include "sd_graph.h"
#include <algorithm>
#include <cstring>

static const int GMG_COARSE = 160;      // coarsest level: at most this many unknowns (explicit inverse)
static const int GMG_DENSE_MAX = 320;   // an aggregation that stalls above this keeps Gauss-Seidel sweeps
static const int GMG_TAIL_N = 4096;     // levels of at most this many unknowns run in one workgroup (gla_tail)
static int GMG_JACOBI = 1;              // levels 1.. above the tail smoothed by two l1-Jacobi sweeps (0: Gauss-Seidel)
static int GMG_PASSES = 3;              // matching passes per level (aggregates of up to 2^passes nodes)
static double GMG_BETA = 0.5;           // a pair joins at >= beta of the node's strongest coupling

struct GLevHost {
    int N = 0, nnz = 0;
    int lanes = 8;                      // lanes per row of the smoother and residual kernels
    std::vector<int> rp, cj;            // pattern (level 0: the mesh graph)
    std::vector<double> wg;             // geometric coupling per entry (aggregation only)
    std::vector<int> agg, mrp, mem, emap;
    std::vector<int> perm, coff;        // nodes grouped by colour, colour offsets
    GLev dev{};
};
static struct {
    std::vector<GLevHost> lev;
    double *inv = nullptr, *work = nullptr;   // coarsest inverse and its elimination workspace
    double *rmem = nullptr;                   // residual of the finest level in member order (restriction)
    bool dense = false;
    bool pw_one = false;
    int tail = -1;                            // first level run by the single-workgroup tail (-1: none)
    GTail T{};
} GM;

static void gfail(const char *what, hipError_t e) {
    if (e == hipSuccess) return;
    fprintf(stderr, "semisim-hip: %s failed: %s\n", what, hipGetErrorString(e));
    eng_set_error();
}
template <class T> static T *gput(const std::vector<T> &v) {
    T *p = (T *)eng_dalloc_bytes(std::max<size_t>(1, v.size()) * sizeof(T));
    if (p && !v.empty()) gfail("graph upload", hipMemcpyAsync(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, G.st));
    return p;
}
template <class T> static T *galloc(size_t n) { return (T *)eng_dalloc_bytes(std::max<size_t>(1, n) * sizeof(T)); }

// ------------------------------------------------------------ aggregation
// one matching pass: every free node joins its strongest free neighbour whose
// coupling is at least beta of the node's strongest; returns the group count
static int pair_pass(int n, const std::vector<int> &rp, const std::vector<int> &cj, const std::vector<double> &w,
                     std::vector<int> &grp, double beta) {
    grp.assign(n, -1);
    int ng = 0;
    for (int i = 0; i < n; i++) {
        if (grp[i] >= 0) continue;
        double mx = 0.;
        for (int e = rp[i]; e < rp[i + 1]; e++) mx = std::max(mx, w[e]);
        int best = -1;
        double bw = -1.;
        for (int e = rp[i]; e < rp[i + 1]; e++) {
            const int j = cj[e];
            if (j == i || grp[j] >= 0 || !(w[e] > 0.) || w[e] < beta * mx) continue;
            if (w[e] > bw) { bw = w[e]; best = j; }
        }
        grp[i] = ng;
        if (best >= 0) grp[best] = ng;
        ng++;
    }
    return ng;
}
// the graph of the groups: entries summed, self entries dropped
static void contract(int n, const std::vector<int> &rp, const std::vector<int> &cj, const std::vector<double> &w,
                     const std::vector<int> &grp, int ng, std::vector<int> &orp, std::vector<int> &ocj,
                     std::vector<double> &ow) {
    std::vector<int> start(ng + 1, 0), mem(n);
    for (int i = 0; i < n; i++) start[grp[i] + 1]++;
    for (int a = 0; a < ng; a++) start[a + 1] += start[a];
    std::vector<int> fill(start.begin(), start.end() - 1);
    for (int i = 0; i < n; i++) mem[fill[grp[i]]++] = i;
    std::vector<int> mark(ng, -1), slot(ng, 0);
    orp.assign(ng + 1, 0); ocj.clear(); ow.clear();
    for (int I = 0; I < ng; I++) {
        for (int q = start[I]; q < start[I + 1]; q++) {
            const int i = mem[q];
            for (int e = rp[i]; e < rp[i + 1]; e++) {
                const int J = grp[cj[e]];
                if (J == I) continue;
                if (mark[J] != I) { mark[J] = I; slot[J] = (int)ocj.size(); ocj.push_back(J); ow.push_back(0.); }
                ow[slot[J]] += w[e];
            }
        }
        orp[I + 1] = (int)ocj.size();
    }
}
// aggregates of level f: GMG_PASSES matching passes on successively
// contracted graphs; returns their number
static int aggregate(const GLevHost &f, std::vector<int> &agg, double beta) {
    const int N = f.N;
    agg.resize(N);
    for (int i = 0; i < N; i++) agg[i] = i;
    std::vector<int> rp = f.rp, cj = f.cj, grp;
    std::vector<double> w = f.wg;
    int n = N, na = N;
    for (int pass = 0; pass < GMG_PASSES && n > 1; pass++) {
        na = pair_pass(n, rp, cj, w, grp, beta);
        for (int i = 0; i < N; i++) agg[i] = grp[agg[i]];
        if (pass == GMG_PASSES - 1 || na == n) break;
        std::vector<int> nrp, ncj;
        std::vector<double> nw;
        contract(n, rp, cj, w, grp, na, nrp, ncj, nw);
        rp.swap(nrp); cj.swap(ncj); w.swap(nw);
        n = na;
    }
    return na;
}
// coarse pattern of c from the aggregation of f, member lists and entry map
static void coarse_pattern(GLevHost &f, GLevHost &c) {
    const int N = f.N, na = c.N;
    f.mrp.assign(na + 1, 0);
    f.mem.resize(N);
    f.emap.assign(std::max(1, f.nnz), -1);
    for (int i = 0; i < N; i++) f.mrp[f.agg[i] + 1]++;
    for (int a = 0; a < na; a++) f.mrp[a + 1] += f.mrp[a];
    std::vector<int> fill(f.mrp.begin(), f.mrp.end() - 1);
    for (int i = 0; i < N; i++) f.mem[fill[f.agg[i]]++] = i;
    std::vector<int> mark(na, -1), slot(na, 0);
    c.rp.assign(na + 1, 0); c.cj.clear();
    for (int I = 0; I < na; I++) {
        for (int q = f.mrp[I]; q < f.mrp[I + 1]; q++) {
            const int i = f.mem[q];
            for (int e = f.rp[i]; e < f.rp[i + 1]; e++) {
                const int J = f.agg[f.cj[e]];
                if (J == I) { f.emap[e] = -1; continue; }
                if (mark[J] != I) { mark[J] = I; slot[J] = (int)c.cj.size(); c.cj.push_back(J); }
                f.emap[e] = slot[J];
            }
        }
        c.rp[I + 1] = (int)c.cj.size();
    }
    c.nnz = (int)c.cj.size();
    c.wg.assign(std::max(1, c.nnz), 0.);
    for (int i = 0; i < N; i++)
        for (int e = f.rp[i]; e < f.rp[i + 1]; e++)
            if (f.emap[e] >= 0) c.wg[f.emap[e]] += f.wg[e];
}
// greedy colouring in node order; nodes grouped by colour
static void colour(GLevHost &v) {
    std::vector<int> col(v.N, -1), used;
    int nc = 0;
    for (int i = 0; i < v.N; i++) {
        used.assign(nc + 1, 0);
        for (int e = v.rp[i]; e < v.rp[i + 1]; e++) {
            const int c = col[v.cj[e]];
            if (c >= 0) used[c] = 1;
        }
        int c = 0;
        while (used[c]) c++;
        col[i] = c;
        if (c == nc) nc++;
    }
    v.coff.assign(nc + 1, 0);
    for (int i = 0; i < v.N; i++) v.coff[col[i] + 1]++;
    for (int c = 0; c < nc; c++) v.coff[c + 1] += v.coff[c];
    std::vector<int> fill(v.coff.begin(), v.coff.end() - 1);
    v.perm.resize(v.N);
    for (int i = 0; i < v.N; i++) v.perm[fill[col[i]]++] = i;
}

int gr_nlev() { return (int)GM.lev.size(); }
// the device arrays are the engine's (released by eng_free); this drops the host side
void gr_free() {
    GM.lev.clear();
    GM.inv = GM.work = GM.rmem = nullptr;
    GM.dense = GM.pw_one = false;
    GM.tail = -1;
    GM.T = GTail{};
}

static void gmg_build() {
    GM.lev.clear();
    GM.inv = GM.work = nullptr;
    GM.pw_one = false;
    if (const char *e = getenv("SEMISIM_GPASS")) GMG_PASSES = std::max(1, atoi(e));
    if (const char *e = getenv("SEMISIM_GJACOBI")) GMG_JACOBI = atoi(e) != 0;
    if (const char *e = getenv("SEMISIM_GBETA")) GMG_BETA = atof(e);
    GM.lev.emplace_back();
    GLevHost &f0 = GM.lev[0];
    f0.N = G.N; f0.nnz = G.nE;
    f0.rp = G.grp; f0.cj = G.gcol;
    f0.wg.resize(std::max(1, G.nE));
    for (int e = 0; e < G.nE; e++) f0.wg[e] = G.gh[e] > 0. ? G.ga[e] / G.gh[e] : 0.;
    while ((int)GM.lev.size() < GTAIL_MAXL) {
        GLevHost &f = GM.lev.back();
        if (f.N <= GMG_COARSE) break;
        std::vector<int> agg;
        // where strongly anisotropic couplings leave most nodes without a
        // partner above beta of their strongest one, the level is matched
        // again with any free neighbour, so the hierarchy reaches a coarsest
        // level small enough for its explicit inverse
        int na = aggregate(f, agg, GMG_BETA);
        if (na < 1 || (double)na > 0.7 * (double)f.N) na = aggregate(f, agg, 0.0);
        if (na < 1 || (double)na > 0.7 * (double)f.N) break;
        f.agg = agg;
        GLevHost c;
        c.N = na;
        coarse_pattern(f, c);
        GM.lev.push_back(std::move(c));
    }
    for (auto &v : GM.lev) {
        colour(v);
        v.lanes = gla_lanes(v.N, v.nnz);
    }
    const int nl = (int)GM.lev.size();
    // device arrays: level 0 works on the engine's operator (diagonal A.c,
    // values d_gval) and on the Krylov vectors
    for (int L = 0; L < nl; L++) {
        GLevHost &v = GM.lev[L];
        GLev &d = v.dev;
        d.A.N = v.N;
        if (L == 0) {
            d.A.rp = G.gd.rp; d.A.cj = G.gd.cj; d.A.va = G.d_gval; d.A.dg = G.A.c;
        } else {
            d.A.rp = gput(v.rp); d.A.cj = gput(v.cj);
            d.A.va = galloc<double>(v.nnz); d.A.dg = galloc<double>(v.N);
            d.x = galloc<double>(v.N); d.b = galloc<double>(v.N);
            d.ph = galloc<double>(v.N);
        }
        d.perm = gput(v.perm);
        d.coff = gput(v.coff);
        d.ncol = (int)v.coff.size() - 1;
        if (L < nl - 1) {
            d.agg = gput(v.agg); d.mrp = gput(v.mrp); d.mem = gput(v.mem); d.emap = gput(v.emap);
            d.pw = galloc<double>(v.N);
        }
    }
    const int nz = GM.lev.back().N;
    GM.dense = nz <= GMG_DENSE_MAX;
    if (GM.dense) {
        GM.inv = galloc<double>((size_t)nz * nz);
        GM.work = galloc<double>((size_t)2 * nz * nz);
    }
    GM.rmem = galloc<double>(GM.lev[0].N);
    GM.tail = -1;
    for (int L = 1; L < nl && GM.tail < 0; L++) if (GM.lev[L].N <= GMG_TAIL_N) GM.tail = L;
    GM.T = GTail{};
    for (int L = 0; L < nl; L++) GM.T.lev[L] = GM.lev[L].dev;
    GM.T.nl = nl;
    GM.T.dense = GM.dense;
    GM.T.inv = GM.inv;
    if (G.verbose) {
        fprintf(stderr, "semisim-hip: graph multigrid %d levels:", nl);
        for (auto &v : GM.lev) fprintf(stderr, " %d (%d colours)", v.N, (int)v.coff.size() - 1);
        fprintf(stderr, "%s; one-workgroup tail from level %d\n", GM.dense ? "" : "; coarsest by Gauss-Seidel sweeps",
                GM.tail);
    }
}

// ------------------------------------------------------------ set-up of a graph device
int eng_set_graph(int N, int E, const int *rp, const int *cj, const double *h, const double *ar, const double *vol,
                  const double *pos) {
    // validation of alloc_dev_graph :984-1019
    if (N < 2 || E < 1 || rp[0] != 0 || rp[N] != E) return 0;
    for (int k = 0; k < N; k++) {
        if (rp[k + 1] < rp[k] || !(vol[k] > 0.)) return 0;
        for (int e = rp[k]; e < rp[k + 1]; e++)
            if (cj[e] < 0 || cj[e] >= N || cj[e] == k || !(h[e] > 0.) || !(ar[e] >= 0.)) return 0;
    }
    std::vector<int> rev(E), ue(E);
    int nue = 0;
    for (int k = 0; k < N; k++)
        for (int e = rp[k]; e < rp[k + 1]; e++) {
            const int j = cj[e];
            int r = -1;
            for (int t = rp[j]; t < rp[j + 1]; t++) if (cj[t] == k) { r = t; break; }
            if (r < 0) return 0;                   // not symmetric
            rev[e] = r;
            if (k < j) ue[e] = nue++;
        }
    for (int k = 0; k < N; k++)
        for (int e = rp[k]; e < rp[k + 1]; e++) if (cj[e] < k) ue[e] = ue[rev[e]];
    if (!eng_alloc(N, 1, 1, 1)) return 0;
    G.nE = E; G.nUE = nue;
    G.grp.assign(rp, rp + N + 1); G.gcol.assign(cj, cj + E); G.grev = rev; G.gue = ue;
    G.gh.assign(h, h + E); G.ga.assign(ar, ar + E); G.gvol.assign(vol, vol + N); G.gpos.assign(pos, pos + 3 * N);
    G.gwall.assign((size_t)6 * N, 1.);
    // tensor-shaped host geometry: x positions and control volumes on the one "axis"
    G.xs.resize(N); G.Lx.assign(vol, vol + N);
    for (int k = 0; k < N; k++) G.xs[k] = pos[3 * k];
    G.ys.assign(1, 0.); G.zs.assign(1, 0.); G.Ly.assign(1, 1.); G.Lz.assign(1, 1.);
    G.mat_idx.assign(N, 0); G.Nd.assign(N, 0.); G.Na.assign(N, 0.); G.Nf.assign(N, 0.);
    GraphDev &g = G.gd;
    g.N = N; g.E = E; g.nue = nue;
    g.rp = gput(G.grp); g.cj = gput(G.gcol); g.ue = gput(G.gue);
    g.h = gput(G.gh); g.ar = gput(G.ga); g.vol = gput(G.gvol); g.pos = gput(G.gpos);
    g.lsq = galloc<double>((size_t)6 * N);
    g.grad = galloc<double>((size_t)3 * N);
    G.d_gval = galloc<double>(E);
    G.d_gmu = galloc<double>((size_t)2 * std::max(1, nue));
    gfail("graph mobility clear", hipMemsetAsync(G.d_gmu, 0, (size_t)2 * std::max(1, nue) * 8, G.st));
    gmg_build();
    eng_sync();
    if (!eng_ok()) { eng_free(); return 0; }
    G.dirty = 1;
    return 1;
}

// prepare's graph part: box-contact areas from the dual faces to the
// semiconductor outside the contact (prepare :1686-1697), the
// terminal-current edges grouped by contact in node order, and the
// least-squares gradient matrices of the semiconductor nodes over their
// semiconductor neighbours (g_lsq_setup :2651-2676, weights 1/|dr|^2, the
// inverse with 1e-6 of the trace added to the diagonal)
int gr_prepare_edges() {
    const int N = G.N;
    {
        std::vector<double> lsq((size_t)6 * N, 0.);
        for (int k = 0; k < N; k++) {
            double m[6] = {0, 0, 0, 0, 0, 0};
            double *o = &lsq[(size_t)6 * k];
            if (G.kind[k] == K_SEMI)
                for (int e = G.grp[k]; e < G.grp[k + 1]; e++) {
                    const int b = G.gcol[e];
                    if (G.kind[b] != K_SEMI) continue;
                    const double dx = G.gpos[3 * b] - G.gpos[3 * k], dy = G.gpos[3 * b + 1] - G.gpos[3 * k + 1],
                                 dz = G.gpos[3 * b + 2] - G.gpos[3 * k + 2];
                    const double w = 1. / (dx * dx + dy * dy + dz * dz + 1e-40);
                    m[0] += w * dx * dx; m[1] += w * dx * dy; m[2] += w * dx * dz;
                    m[3] += w * dy * dy; m[4] += w * dy * dz; m[5] += w * dz * dz;
                }
            const double tr = m[0] + m[3] + m[5], eps = 1e-6 * tr + 1e-30;
            m[0] += eps; m[3] += eps; m[5] += eps;
            const double a = m[0], b = m[1], c = m[2], dd = m[3], ee = m[4], f = m[5];
            const double A = dd * f - ee * ee, Bq = -(b * f - c * ee), Cc = b * ee - c * dd;
            const double D = a * f - c * c, Ei = -(a * ee - b * c), F = a * dd - b * b;
            const double det = a * A + b * Bq + c * Cc;
            if (!(fabs(det) > 0.)) continue;
            o[0] = A / det; o[1] = Bq / det; o[2] = Cc / det; o[3] = D / det; o[4] = Ei / det; o[5] = F / det;
        }
        gfail("gradient matrices", hipMemcpyAsync((void *)G.gd.lsq, lsq.data(), lsq.size() * 8, hipMemcpyHostToDevice, G.st));
        eng_sync();
    }
    for (int c = 0; c < G.n_cons; c++) {
        HostCon &cn = G.cons[c];
        if (cn.face <= 5) continue;
        double A = 0.;
        for (int k : G.con_nodes[c]) {
            if (G.con_mask[k] != c + 1) continue;
            for (int e = G.grp[k]; e < G.grp[k + 1]; e++) {
                const int kb = G.gcol[e];
                if (G.kind[kb] == K_SEMI && G.con_mask[kb] != c + 1) A += G.ga[e];
            }
        }
        cn.A = A;
    }
    std::vector<std::vector<int>> ck(G.n_cons), ce(G.n_cons);
    for (int k = 0; k < N; k++) {
        const int c = G.con_mask[k];
        if (!c || G.kind[k] != K_SEMI) continue;
        for (int e = G.grp[k]; e < G.grp[k + 1]; e++) {
            const int kb = G.gcol[e];
            if (G.kind[kb] != K_SEMI || G.con_mask[kb] == c) continue;
            ck[c - 1].push_back(k); ce[c - 1].push_back(e);
        }
    }
    std::vector<int> fk, fe, off(G.n_cons + 1, 0);
    for (int c = 0; c < G.n_cons; c++) {
        off[c] = (int)fk.size();
        fk.insert(fk.end(), ck[c].begin(), ck[c].end());
        fe.insert(fe.end(), ce[c].begin(), ce[c].end());
    }
    off[G.n_cons] = (int)fk.size();
    G.d_gck = (int *)eng_palloc_bytes(std::max<size_t>(1, fk.size()) * 4);
    G.d_gce = (int *)eng_palloc_bytes(std::max<size_t>(1, fe.size()) * 4);
    G.d_gcoff = (int *)eng_palloc_bytes(off.size() * 4);
    if (!G.d_gck || !G.d_gce || !G.d_gcoff) return 0;
    if (!fk.empty()) {
        gfail("contact edges", hipMemcpyAsync(G.d_gck, fk.data(), fk.size() * 4, hipMemcpyHostToDevice, G.st));
        gfail("contact edges", hipMemcpyAsync(G.d_gce, fe.data(), fe.size() * 4, hipMemcpyHostToDevice, G.st));
    }
    gfail("contact edges", hipMemcpyAsync(G.d_gcoff, off.data(), off.size() * 4, hipMemcpyHostToDevice, G.st));
    eng_sync();
    return eng_ok();
}

// MLDA confinement terms from the walls the host supplied per node (two
// walls: distance, cell range a..b; >= 0.5 m: none), walls_graph :1252-1270
void gr_walls(std::vector<double> &qd, std::vector<double> &qcn, std::vector<double> &qcp, int qc,
              double (*factor)(const HostMat &, double, double, int)) {
    const int N = G.N;
    qd.assign(N, 1.); qcn.assign(N, 0.); qcp.assign(N, 0.);
    for (int k = 0; k < N; k++) {
        if (G.kind[k] != K_SEMI) continue;
        const HostMat &m = G.mats[G.mat8[k]];
        for (int w = 0; w < 2; w++) {
            const double *g = &G.gwall[(size_t)6 * k + 3 * w];
            if (!(g[0] < 0.5)) continue;
            if (g[0] < qd[k]) qd[k] = g[0];
            if (qc && g[1] < 40e-9) {
                qcn[k] += log(fmax(factor(m, g[1], g[2], 0), 1e-30));
                qcp[k] += log(fmax(factor(m, g[1], g[2], 1), 1e-30));
            }
        }
        if (qcn[k] < -60.) qcn[k] = -60.;
        if (qcp[k] < -60.) qcp[k] = -60.;
    }
}

// ------------------------------------------------------------ multigrid per solve
static GOp level0_op() { return GOp{G.N, G.gd.rp, G.gd.cj, G.d_gval, G.A.c}; }

void gr_mg_setup() {
    const int nl = (int)GM.lev.size();
    hipStream_t s = G.st;
    DevView d = eng_view();
    GM.lev[0].dev.A = level0_op();
    if (G.mg_qsign == 0) {
        if (!GM.pw_one) {
            for (int L = 0; L < nl - 1; L++) gla_pw_one(s, GM.lev[L].N, GM.lev[L].dev.pw);
            GM.pw_one = true;
        }
    } else {
        const double VT = G.mats[G.ref_mat].VT;
        for (int L = 0; L < nl - 1; L++) gla_weights(s, d, L, GM.lev[L + 1].N, GM.lev[L].dev, GM.lev[L + 1].dev, G.mg_qsign, VT);
        GM.pw_one = false;
    }
    for (int L = 0; L < nl - 1; L++)
        gla_galerkin(s, GM.lev[L + 1].N, GM.lev[L].dev, GM.lev[L + 1].dev, L == 0 ? G.mg_w : nullptr);
    if (GM.dense) gla_dense_inverse(s, GM.lev[nl - 1].dev.A, GM.work, GM.inv);
}

static void sweep(const GLevHost &v, const GOp &A, const double *b, double *x, bool forward) {
    const int nc = (int)v.coff.size() - 1;
    for (int q = 0; q < nc; q++) {
        const int c = forward ? q : nc - 1 - q;
        gla_gs(G.st, A, b, x, v.dev.perm, v.coff[c], v.coff[c + 1], v.lanes);
    }
}
static void gvcycle(int L, const double *b, double *x) {
    GLevHost &v = GM.lev[L];
    const int nl = (int)GM.lev.size();
    hipStream_t s = G.st;
    if (L == GM.tail) { gla_tail(s, GM.T, L, b, x); return; }
    const GOp A = L == 0 ? level0_op() : v.dev.A;
    if (L == nl - 1) {
        if (GM.dense) { gla_dense_apply(s, v.N, GM.inv, b, x); return; }
        sd_fill(s, x, 0.0, v.N);
        for (int it = 0; it < 4; it++) { sweep(v, A, b, x, true); sweep(v, A, b, x, false); }
        return;
    }
    GLevHost &c = GM.lev[L + 1];
    // coarse levels: two l1-Jacobi sweeps (one launch each, through the
    // member-order scratch) instead of one launch per colour
    const bool jac = GMG_JACOBI && L > 0;
    if (jac) {
        gla_jacobi(s, A, b, nullptr, GM.rmem, v.lanes);
        gla_jacobi(s, A, b, GM.rmem, x, v.lanes);
    } else {
        sd_fill(s, x, 0.0, v.N);
        sweep(v, A, b, x, true);
    }
    GLev f = v.dev;
    f.A = A;
    // levels above the tail: residual one lane per row, then per-aggregate sums
    gla_resid_restrict(s, c.N, f, b, x, L == 0 ? G.mg_w : nullptr, c.dev.b, GM.rmem, v.lanes);
    gvcycle(L + 1, c.dev.b, c.dev.x);
    gla_prolong(s, v.N, f, c.dev.x, x, 1.0);
    if (jac) {
        gla_jacobi(s, A, b, x, GM.rmem, v.lanes);
        gla_jacobi(s, A, b, GM.rmem, x, v.lanes);
    } else sweep(v, A, b, x, false);
}
void gr_precond(const double *r, double *z) {
    G.n_vcycles++;
    if (G.mg_off || GM.lev.size() < 2) {       // smoother only: one symmetric sweep from 0
        const GOp A = level0_op();
        sd_fill(G.st, z, 0.0, G.N);
        sweep(GM.lev[0], A, r, z, true);
        sweep(GM.lev[0], A, r, z, false);
        return;
    }
    gvcycle(0, r, z);
}

// ------------------------------------------------------------ operator products
void gr_spmv(const double *x, const double *b, double *y, bool unit) {
    GOp A = level0_op();
    if (unit) A.dg = nullptr;
    gla_spmv(G.st, A, x, b, y);
}
void gr_spmv_dots(const double *x, double *y, const double *u0, const double *u1, int nv, int slot, int post,
                  bool unit) {
    GOp A = level0_op();
    if (unit) A.dg = nullptr;
    // one lane per row: the rows of the Krylov product are consecutive, and
    // measured faster so than by lane groups (28 against 31 us at 531k nodes)
    gla_spmv_dots(G.st, A, x, y, u0, u1, nv, G.rb, slot, post, 1);
}
void gr_floor(const double *x, const double *b, int slot) { gla_floor(G.st, level0_op(), x, b, G.rb, slot); }

// ------------------------------------------------------------ physics
void gr_poisson(const double *phi, int build, double *F, int slot) {
    gph_poisson(G.st, eng_view(), G.gd, phi, build, F, level0_op(), G.d_sc, G.rb, slot);
}
void gr_upd_mob(const double *qn, const double *qp) { gph_upd_mob(G.st, eng_view(), G.gd, qn, qp, G.d_gmu); }
void gr_cont(int hole) { gph_cont(G.st, eng_view(), G.gd, hole, G.d_gmu, level0_op(), G.d_rhs, G.d_sc, G.d_dsc); }
void gr_currents() {
    gph_currents(G.st, eng_view(), G.gd, G.d_gmu, G.d_gck, G.d_gce, G.d_gcoff, G.n_cons, G.rb);
}
void gr_display() {
    gph_display(G.st, eng_view(), G.gd, G.d_gmu, G.d_Jx, G.d_Jy, G.d_Jz, G.d_Rt, G.d_mudn, G.d_mudp);
}
void gr_node_J(int ax, int carrier, double *out) { gph_node_J(G.st, eng_view(), G.gd, G.d_gmu, ax, carrier, out); }

// Floating majority blocks on the graph (float_fix_g :4349-4409): blocks of
// majority nodes connected through semiconductor edges, in node order; a
// block touching a contact node is anchored; every other block's carrier is
// rescaled by exp(sl) so that its balance (outflow, recombination, gate
// tunnelling, impact ionisation) closes.  The flags come from the device; the
// blocks and the balance are traced on the host over the graph.
int gr_float_fix(int hole) {
    const int N = G.N;
    hipStream_t s = G.st;
    DevView d = eng_view();
    gph_maj_flags(s, d, hole, G.d_sol, G.d_sc, G.d_flag);
    std::vector<unsigned char> maj(N);
    gfail("majority flags", hipMemcpyAsync(maj.data(), G.d_flag, N, hipMemcpyDeviceToHost, s));
    eng_sync();
    if (!eng_ok()) return 0;
    std::vector<int> lab(N, 0), q;
    std::vector<std::vector<int>> blocks;
    int blk = 0;
    for (int k0 = 0; k0 < N; k0++) {
        if (lab[k0] || !maj[k0]) continue;
        blk++;
        q.assign(1, k0);
        lab[k0] = blk;
        bool anch = false;
        for (size_t head = 0; head < q.size(); head++) {
            const int k = q[head];
            for (int e = G.grp[k]; e < G.grp[k + 1]; e++) {
                const int kb = G.gcol[e];
                if (G.kind[kb] != K_SEMI) continue;
                if (G.con_mask[kb]) { anch = true; continue; }
                if (!lab[kb] && maj[kb]) { lab[kb] = blk; q.push_back(kb); }
            }
        }
        if (!anch) blocks.push_back(q);
    }
    int nfix = 0;
    if (!blocks.empty()) {
        std::vector<double> u(N), sc(N), n(N), p(N), psi(N), mu((size_t)2 * std::max(1, G.nUE)), tk, tg, gii;
        eng_download(u.data(), G.d_sol, N); eng_download(sc.data(), G.d_sc, N);
        eng_download(n.data(), G.d_n, N); eng_download(p.data(), G.d_p, N); eng_download(psi.data(), G.d_psi, N);
        eng_download(mu.data(), G.d_gmu, mu.size());
        if (d.gii) { gii.resize(N); eng_download(gii.data(), G.d_gii, N); }
        const std::vector<double> &oth = hole ? n : p;
        const std::vector<double> &car = hole ? p : n;
        for (auto &B : blocks) {
            const int id = lab[B[0]];
            double Aout = 0., Bin = 0.;
            for (int k : B) {
                const double VT = G.mats[G.mat8[k]].VT;
                for (int e = G.grp[k]; e < G.grp[k + 1]; e++) {
                    const int kb = G.gcol[e];
                    if (G.kind[kb] != K_SEMI || lab[kb] == id) continue;
                    const double dl = hole ? -(G.lnvK[kb] - G.lnvK[k]) : (G.lncK[kb] - G.lncK[k]);
                    const double D = (psi[kb] - psi[k]) / VT + dl;
                    const double gg = G.ga[e] * mu[(size_t)hole * G.nUE + G.gue[e]] * VT / G.gh[e];
                    const double self = hole ? gg * host_B(D) : gg * host_B(-D), nb = hole ? gg * host_B(-D) : gg * host_B(D);
                    const double cb = G.con_mask[kb] ? car[kb] : sc[kb] * u[kb];
                    Aout += self * sc[k] * u[k];
                    Bin += nb * cb;
                }
            }
            double sl = 0.;
            for (int it = 0; it < 60; it++) {
                const double lam = exp(sl);
                double Gv = lam * Aout - Bin, dG = lam * Aout;
                for (int k : B) {
                    const HostMat &mt = G.mats[G.mat8[k]];
                    const double c = lam * sc[k] * u[k];
                    double Rr, dRn, dRp;
                    const double tf = ((G.models & 2) && mt.tau_Nref > 0.) ? 1. / (1. + (G.Nd[k] + G.Na[k]) / mt.tau_Nref) : 1.;
                    host_recomb(mt, G.niK[k], tf, hole ? oth[k] : c, hole ? c : oth[k], &Rr, &dRn, &dRp);
                    Gv += G.gvol[k] * Rr;
                    dG += G.gvol[k] * (hole ? dRp : dRn) * c;
                    if (!gii.empty()) Gv -= gii[k];
                }
                if (!(dG > 0.)) break;
                double ds = -Gv / dG;
                if (ds > 2.) ds = 2.; else if (ds < -2.) ds = -2.;
                sl += ds;
                if (fabs(ds) < 1e-13) break;
            }
            if (sl == sl && fabs(sl) > 1e-15 && fabs(sl) < 200.) {
                const double lam = exp(sl);
                for (int k : B) u[k] *= lam;
                nfix++;
            }
        }
        if (nfix) eng_upload(G.d_sol, u.data(), N);
    }
    G.ff_stat[0] += blocks.empty() ? 0 : 1;
    G.ff_stat[1] += (long long)blocks.size();
    G.ff_stat[2] += nfix;
    if (G.verbose > 2 && !blocks.empty())
        fprintf(stderr, "    float_fix(%c): %d blocks, %d rescaled\n", hole ? 'p' : 'n', (int)blocks.size(), nfix);
    return nfix;
}

// quasi-neutral regions of ohmic contacts shifted rigidly with their voltage
// change, traced over the graph (region_shift_g :4525-4555)
void gr_region_shift(const double *dV) {
    const int N = G.N;
    bool any = false;
    for (int c = 0; c < G.n_cons; c++) if (G.cons[c].bc == 0 && fabs(dV[c]) >= 1e-15) any = true;
    if (!any) return;
    std::vector<double> n(N), p(N), psi(N), qn(N), qp(N);
    eng_download(n.data(), G.d_n, N); eng_download(p.data(), G.d_p, N);
    eng_download(psi.data(), G.d_psi, N); eng_download(qn.data(), G.d_phin, N); eng_download(qp.data(), G.d_phip, N);
    std::vector<unsigned char> mark(N, 0);
    std::vector<int> q;
    for (int c = 0; c < G.n_cons; c++) {
        if (G.cons[c].bc != 0 || fabs(dV[c]) < 1e-15) continue;
        q.clear();
        int type = 0;
        for (int k = 0; k < N; k++)
            if (G.con_mask[k] == c + 1 && G.kind[k] == K_SEMI && !mark[k]) {
                mark[k] = (unsigned char)(c + 1);
                q.push_back(k);
                if (!type) { const double D = G.Nd[k] - G.Na[k]; type = D > 0. ? 1 : (D < 0. ? -1 : 0); }
            }
        if (!type) continue;
        for (size_t head = 0; head < q.size(); head++) {
            const int k = q[head];
            for (int e = G.grp[k]; e < G.grp[k + 1]; e++) {
                const int kb = G.gcol[e];
                if (mark[kb] || G.kind[kb] != K_SEMI) continue;
                if (G.con_mask[kb] && G.con_mask[kb] != c + 1) continue;
                const double D = G.Nd[kb] - G.Na[kb];
                if (D * type <= 0.) continue;
                const double m = type > 0 ? n[kb] : p[kb];
                if (m < 0.5 * fabs(D)) continue;
                mark[kb] = (unsigned char)(c + 1);
                q.push_back(kb);
            }
        }
        for (int k : q) {
            if (G.con_mask[k]) continue;
            psi[k] += dV[c]; qn[k] += dV[c]; qp[k] += dV[c];
        }
    }
    eng_upload(G.d_psi, psi.data(), N); eng_upload(G.d_phin, qn.data(), N); eng_upload(G.d_phip, qp.data(), N);
}

// largest field in the semiconductor: least-squares gradient of psi over the
// semiconductor neighbours of every node (g_lsq_setup/g_grad :2651-2691,
// Emax_g :4557-4566)
double gr_emax() {
    const int N = G.N;
    std::vector<double> psi(N);
    eng_download(psi.data(), G.d_psi, N);
    double Em = 0.;
    for (int k = 0; k < N; k++) {
        if (G.kind[k] != K_SEMI) continue;
        double m[6] = {0, 0, 0, 0, 0, 0}, r[3] = {0, 0, 0};
        for (int e = G.grp[k]; e < G.grp[k + 1]; e++) {
            const int b = G.gcol[e];
            if (G.kind[b] != K_SEMI) continue;
            const double dx = G.gpos[3 * b] - G.gpos[3 * k], dy = G.gpos[3 * b + 1] - G.gpos[3 * k + 1],
                         dz = G.gpos[3 * b + 2] - G.gpos[3 * k + 2];
            const double w = 1. / (dx * dx + dy * dy + dz * dz + 1e-40), f = (psi[b] - psi[k]) * w;
            m[0] += w * dx * dx; m[1] += w * dx * dy; m[2] += w * dx * dz; m[3] += w * dy * dy; m[4] += w * dy * dz; m[5] += w * dz * dz;
            r[0] += f * dx; r[1] += f * dy; r[2] += f * dz;
        }
        const double tr = m[0] + m[3] + m[5], eps = 1e-6 * tr + 1e-30;
        m[0] += eps; m[3] += eps; m[5] += eps;
        const double a = m[0], b = m[1], c = m[2], dd = m[3], e = m[4], f = m[5];
        const double A = dd * f - e * e, Bq = -(b * f - c * e), Cc = b * e - c * dd, D = a * f - c * c, Ei = -(a * e - b * c),
                     F = a * dd - b * b;
        const double det = a * A + b * Bq + c * Cc;
        if (!(fabs(det) > 0.)) continue;
        const double g0 = (A * r[0] + Bq * r[1] + Cc * r[2]) / det, g1 = (Bq * r[0] + D * r[1] + Ei * r[2]) / det,
                     g2 = (Cc * r[0] + Ei * r[1] + F * r[2]) / det;
        Em = std::max(Em, g0 * g0 + g1 * g1 + g2 * g2);
    }
    return sqrt(Em);
}
