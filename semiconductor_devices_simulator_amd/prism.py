"""Devices on general box-method meshes (the CSR graph path of the core,
api3_set_mesh_graph), and the api3_* call sequence that loads them.

A graph device is N nodes with positions, a symmetric CSR graph of directed
edges (row pointers rp, columns cj) carrying each edge's length h and dual
face area ar, and each node's control volume vol (semiconductor_3d.c:4788-
4804).  The front end builds such meshes from triangular prisms
(semimesh.PrismMesh, used by main.py:3363-3410); two builders stand in for
it here:

  tensor_as_graph   a tensor-mesh DeviceSpec re-expressed as a graph: the
                    six neighbours of every node with the tensor box-method
                    geometry.  The graph path then solves the same discrete
                    equations as the tensor path (the identity the reference's
                    validation checks with a prism mesh holding every tensor
                    node, validate.py:652-660).
  prism_mosfet      a planar n-MOSFET on a triangular-prism mesh: an
                    isosceles triangular lattice in the x-y cross-section
                    (alternate rows offset by half a spacing), extruded along
                    z, with the circumcentric (Voronoi) dual of the
                    triangulation as control volumes.

load_graph_device follows _setup_tri (main.py:3363-3410): materials, solver,
the graph, models, node data, then every contact as a node list with its
boundary areas (api3_set_con_nodes) and the walls (api3_set_walls).
"""
from dataclasses import dataclass, field

import numpy as np

from .core import dptr, iptr
from .devices import MAT_SI, Contact, Implant, span_idx


@dataclass
class GraphSpec:
    name: str
    pos: np.ndarray                 # (N, 3) node positions (m)
    rp: np.ndarray                  # (N+1,) row pointers
    cj: np.ndarray                  # (E,) neighbour of every directed edge
    h: np.ndarray                   # (E,) edge length (m)
    ar: np.ndarray                  # (E,) dual face area (m^2)
    vol: np.ndarray                 # (N,) control volume (m^3)
    mat: np.ndarray                 # per-node index into `materials`
    Nd: np.ndarray
    Na: np.ndarray
    contacts: list                  # [(Contact, node indices, per-node boundary areas (m^2))]
    materials: list = field(default_factory=lambda: [MAT_SI])
    models: int = 7
    T: float = 300.0
    solver: tuple = (2000, 1000, 500, 1e-10, 1e-9, 1.0, 1.0)
    tau: tuple = None
    sweep: list = field(default_factory=list)
    Nf: np.ndarray = None
    walls: np.ndarray = None        # (N, 6): two walls per node, each (distance, cell range a, b) (m); None: none

    @property
    def N(self):
        return len(self.vol)

    @property
    def E(self):
        return len(self.cj)

    def bias_points(self):
        return sum(s[3] for s in self.sweep)


def _fv_len(c):
    """Box-method cell length of every tensor line (half the spacing on each
    side; semiconductor_3d.c:700)."""
    c = np.asarray(c, float)
    L = np.empty(len(c))
    L[0] = 0.5 * (c[1] - c[0])
    L[-1] = 0.5 * (c[-1] - c[-2])
    L[1:-1] = 0.5 * (c[2:] - c[:-2])
    return L


def _tensor_face_nodes(con, xs, ys, zs):
    """Nodes and boundary areas of a contact of a tensor mesh, by the index
    ranges the core derives from the contact (v6 convention, clipped)."""
    Nx, Ny, Nz = len(xs), len(ys), len(zs)
    Lx, Ly, Lz = _fv_len(xs), _fv_len(ys), _fv_len(zs)
    i0, i1, j0, j1, _, _ = con.to_idx6(xs, ys, zs)
    ia, ib, ja, jb, la, lb = 0, Nx - 1, 0, Ny - 1, 0, Nz - 1
    f = con.face
    if f in (0, 1):
        ia = ib = 0 if f == 0 else Nx - 1; ja, jb, la, lb = i0, i1, j0, j1
    elif f in (2, 3):
        ja = jb = 0 if f == 2 else Ny - 1; ia, ib, la, lb = i0, i1, j0, j1
    elif f in (4, 5):
        la = lb = 0 if f == 4 else Nz - 1; ia, ib, ja, jb = i0, i1, j0, j1
    else:
        ia, ib, ja, jb, la, lb = i0, i1, j0, j1, 0, 0
    ia, ja, la = max(ia, 0), max(ja, 0), max(la, 0)
    ib, jb, lb = min(ib, Nx - 1), min(jb, Ny - 1), min(lb, Nz - 1)
    nodes, areas = [], []
    for l in range(la, lb + 1):
        for j in range(ja, jb + 1):
            for i in range(ia, ib + 1):
                nodes.append(i + Nx * (j + Ny * l))
                areas.append(Ly[j] * Lz[l] if f <= 1 else (Lx[i] * Lz[l] if f <= 3 else (Lx[i] * Ly[j] if f <= 5 else 0.0)))
    return np.asarray(nodes, np.int32), np.asarray(areas)


def tensor_as_graph(spec):
    """The tensor-mesh device `spec` (devices.DeviceSpec) as a graph device:
    node k = i + Nx (j + Ny l) at (xs[i], ys[j], zs[l]); an edge to each of
    its up to six neighbours with the spacing as length and the tensor dual
    face as area; the box as control volume; every contact as its node list
    with the face areas.  Shaped contacts (spec.con_nodes) keep their lists."""
    xs, ys, zs = (np.asarray(c, float) for c in (spec.xs, spec.ys, spec.zs))
    Nx, Ny, Nz = len(xs), len(ys), len(zs)
    Lx, Ly, Lz = _fv_len(xs), _fv_len(ys), _fv_len(zs)
    N = Nx * Ny * Nz
    Z, Y, X = np.meshgrid(zs, ys, xs, indexing="ij")
    pos = np.c_[X.ravel(), Y.ravel(), Z.ravel()]
    Lz3, Ly3, Lx3 = np.meshgrid(Lz, Ly, Lx, indexing="ij")
    vol = (Lx3 * Ly3 * Lz3).ravel()
    k = np.arange(N).reshape(Nz, Ny, Nx)
    ea, eb, eh, ear = [], [], [], []
    # x edges
    ea.append(k[:, :, :-1].ravel()); eb.append(k[:, :, 1:].ravel())
    eh.append(np.broadcast_to(np.diff(xs)[None, None, :], (Nz, Ny, Nx - 1)).ravel())
    ear.append(np.broadcast_to((Lz[:, None] * Ly[None, :])[:, :, None], (Nz, Ny, Nx - 1)).ravel())
    # y edges
    ea.append(k[:, :-1, :].ravel()); eb.append(k[:, 1:, :].ravel())
    eh.append(np.broadcast_to(np.diff(ys)[None, :, None], (Nz, Ny - 1, Nx)).ravel())
    ear.append(np.broadcast_to((Lz[:, None] * Lx[None, :])[:, None, :], (Nz, Ny - 1, Nx)).ravel())
    # z edges
    ea.append(k[:-1, :, :].ravel()); eb.append(k[1:, :, :].ravel())
    eh.append(np.broadcast_to(np.diff(zs)[:, None, None], (Nz - 1, Ny, Nx)).ravel())
    ear.append(np.broadcast_to((Ly[:, None] * Lx[None, :])[None, :, :], (Nz - 1, Ny, Nx)).ravel())
    a, b = np.concatenate(ea), np.concatenate(eb)
    h, ar = np.concatenate(eh), np.concatenate(ear)
    rp, cj, hh, aa = _csr_fast(N, a, b, h, ar)
    cons = []
    for ci, c in enumerate(spec.contacts):
        if ci in spec.con_nodes:
            n = np.asarray(spec.con_nodes[ci], np.int32)
            cons.append((c, n, np.zeros(len(n))))
        else:
            cons.append((c,) + _tensor_face_nodes(c, xs, ys, zs))
    return GraphSpec("graph_" + spec.name, pos, rp, cj, hh, aa, vol, np.asarray(spec.mat, np.int32),
                     np.asarray(spec.Nd, float), np.asarray(spec.Na, float), cons, list(spec.materials), spec.models,
                     spec.T, spec.solver, spec.tau, list(spec.sweep))


def _csr_fast(N, a, b, h, ar):
    """Symmetric CSR (ascending columns) of undirected edges a-b."""
    src = np.concatenate([a, b]).astype(np.int64)
    dst = np.concatenate([b, a]).astype(np.int64)
    hh = np.concatenate([h, h])
    aa = np.concatenate([ar, ar])
    o = np.lexsort((dst, src))
    src, dst, hh, aa = src[o], dst[o], hh[o], aa[o]
    rp = np.zeros(N + 1, np.int32)
    np.add.at(rp, src + 1, 1)
    rp = np.cumsum(rp).astype(np.int32)
    return rp, dst.astype(np.int32), np.ascontiguousarray(hh), np.ascontiguousarray(aa)


def _lattice(Lx, Ly, nx, ny):
    """Points of an isosceles triangular lattice on [0, Lx] x [0, Ly]: rows
    y_j = j Ly/(ny-1); even rows at x = i a (a = Lx/(nx-1)), odd rows offset by
    a/2 with the two boundary points x = 0 and Lx added.  Returns the points
    (x fastest within a row, rows in order) and the row of every point."""
    a = Lx / (nx - 1)
    pts, row = [], []
    for j in range(ny):
        y = j * Ly / (ny - 1)
        xr = [i * a for i in range(nx)] if j % 2 == 0 else [0.0] + [(i + 0.5) * a for i in range(nx - 1)] + [Lx]
        for x in xr:
            pts.append((x, y)); row.append(j)
    return np.asarray(pts), np.asarray(row)


def _lattice_triangles(nx, ny):
    """Triangles of the lattice of _lattice, counter-clockwise."""
    start, cnt = [], []
    s = 0
    for j in range(ny):
        n = nx if j % 2 == 0 else nx + 1
        start.append(s); cnt.append(n); s += n
    tris = []
    for j in range(ny - 1):
        e, o = (j, j + 1) if j % 2 == 0 else (j + 1, j)
        E = [start[e] + i for i in range(nx)]            # even row: nx points
        O = [start[o] + i for i in range(nx + 1)]        # odd row: 0, a/2, 3a/2, ..., Lx
        up = e > o                                       # even row above the odd one
        for i in range(nx - 1):
            # between E[i], E[i+1] and the odd point O[i+1] (at (i + 1/2) a)
            tris.append((E[i], E[i + 1], O[i + 1]) if not up else (E[i + 1], E[i], O[i + 1]))
        for i in range(nx):
            # between O[i], O[i+1] and E[i]
            tris.append((O[i + 1], O[i], E[i]) if not up else (O[i], O[i + 1], E[i]))
    return np.asarray(tris)


def _voronoi_dual(P, tris):
    """Circumcentric dual of a triangulation without obtuse triangles: for
    every edge the length of its dual segment (sum over its triangles of the
    distance from the edge midpoint to the circumcentre), for every node its
    control area."""
    dual, area = {}, np.zeros(len(P))
    for t in tris:
        A, B, C = P[t[0]], P[t[1]], P[t[2]]
        d = 2.0 * (A[0] * (B[1] - C[1]) + B[0] * (C[1] - A[1]) + C[0] * (A[1] - B[1]))
        ux = ((A @ A) * (B[1] - C[1]) + (B @ B) * (C[1] - A[1]) + (C @ C) * (A[1] - B[1])) / d
        uy = ((A @ A) * (C[0] - B[0]) + (B @ B) * (A[0] - C[0]) + (C @ C) * (B[0] - A[0])) / d
        cc = np.array([ux, uy])
        for p, q, r in ((t[0], t[1], t[2]), (t[1], t[2], t[0]), (t[2], t[0], t[1])):
            m = 0.5 * (P[p] + P[q])
            ev = P[q] - P[p]
            nrm = np.array([-ev[1], ev[0]])
            s = np.dot(cc - m, nrm) / np.linalg.norm(nrm)
            if np.dot(P[r] - m, nrm) < 0:
                s = -s
            key = (min(p, q), max(p, q))
            dual[key] = dual.get(key, 0.0) + s
            half = 0.5 * np.linalg.norm(ev)
            area[p] += 0.5 * half * s
            area[q] += 0.5 * half * s
    return dual, area


def prism_mesh(Lx, Ly, zs, nx, ny):
    """Triangular-prism mesh of [0, Lx] x [0, Ly] x [zs[0], zs[-1]] (m):
    the lattice of _lattice in x-y, extruded through the z lines zs.  Node
    (p, l) = p + P l.  Returns positions, CSR (rp, cj, h, ar), volumes, the
    2-D points, their rows and the 2-D boundary length of every point on the
    rectangle's sides y = 0 and y = Ly (its dual length along the side)."""
    P, row = _lattice(Lx, Ly, nx, ny)
    tris = _lattice_triangles(nx, ny)
    dual, area = _voronoi_dual(P, tris)
    zs = np.asarray(zs, float)
    Lz = _fv_len(zs)
    Np, Nz = len(P), len(zs)
    N = Np * Nz
    ea, eb, eh, ear = [], [], [], []
    for (p, q), s in dual.items():
        L = float(np.linalg.norm(P[q] - P[p]))
        for l in range(Nz):
            ea.append(p + Np * l); eb.append(q + Np * l); eh.append(L); ear.append(max(s, 0.0) * Lz[l])
    for l in range(Nz - 1):
        for p in range(Np):
            ea.append(p + Np * l); eb.append(p + Np * (l + 1)); eh.append(zs[l + 1] - zs[l]); ear.append(area[p])
    rp, cj, h, ar = _csr_fast(N, np.asarray(ea), np.asarray(eb), np.asarray(eh), np.asarray(ear))
    pos = np.c_[np.tile(P, (Nz, 1)), np.repeat(zs, Np)]
    vol = np.tile(area, Nz) * np.repeat(Lz, Np)
    side = np.zeros(Np)
    for j in (0, ny - 1):
        idx = np.nonzero(row == j)[0]
        xs = P[idx, 0]
        o = np.argsort(xs)
        idx, xs = idx[o], xs[o]
        L = np.empty(len(xs))
        L[0] = 0.5 * (xs[1] - xs[0]); L[-1] = 0.5 * (xs[-1] - xs[-2]); L[1:-1] = 0.5 * (xs[2:] - xs[:-2])
        side[idx] = L
    return pos, rp, cj, h, ar, vol, P, row, side, Lz


def prism_mosfet(nx=31, ny=17, nz=5, npts=7, vd=0.05, lg=200.0, tox=2.0, nsub=1e17, nsd=1e20, sigma=40.0,
                 models=7):
    """Planar n-MOSFET of devices.mosfet (x channel, y depth with the top at
    y = 0, z width; 600 x 300 x 200 nm at lg = 200 nm) on the triangular-prism
    mesh of prism_mesh: nx points per even lattice row along x, ny rows, nz
    z lines.  Source, drain and the Robin gate (tox nm SiO2, work function
    4.65 eV) are node lists on the top row, the body on the bottom row, each
    node with its boundary area (side length x cell length along z).  Id-Vg
    0 -> 1.2 V in npts points at Vd = vd."""
    Lx, Ly, Lz = 400.0 + lg, 300.0, 200.0
    zs = np.linspace(0.0, Lz * 1e-9, nz)
    pos, rp, cj, h, ar, vol, P, row, side, Lzc = prism_mesh(Lx * 1e-9, Ly * 1e-9, zs, nx, ny)
    Np, N = len(P), len(vol)
    L3 = (Lx, Ly, Lz)
    Pn = pos * 1e9
    nd = Implant("n", nsd, "y0", 0.0, 200.0, 0.0, 1e9, "gauss", sigma, 0.0, 0.8).conc(Pn, L3)
    nd += Implant("n", nsd, "y0", 200.0 + lg, 1e9, 0.0, 1e9, "gauss", sigma, 0.0, 0.8).conc(Pn, L3)
    Nd = nd * 1e6
    Na = np.full(N, nsub * 1e6)
    mat = np.zeros(N, np.int32)

    def face(rowj, x0, x1):
        """nodes of lattice row rowj with x in the half-open physical span
        [x0, x1] (span_idx rule), every z line"""
        idx = np.nonzero(row == rowj)[0]
        xs = P[idx, 0]
        o = np.argsort(xs)
        idx, xs = idx[o], xs[o]
        i0, i1 = span_idx(xs, x0, x1)
        sel = idx[i0:i1 + 1]
        nodes = np.concatenate([sel + Np * l for l in range(nz)]).astype(np.int32)
        areas = np.concatenate([side[sel] * Lzc[l] for l in range(nz)])
        return nodes, areas

    top, bot = 0, ny - 1
    a, b = 200e-9, (200.0 + lg) * 1e-9
    cons = [(Contact(2, 0, 0, 0, 0, 0, 0.0, label="Source"),) + face(top, 0.0, a),
            (Contact(2, 0, 0, 0, 0, 0, vd, label="Drain"),) + face(top, b, Lx * 1e-9),
            (Contact(2, 0, 0, 0, 0, 2, 0.0, pm=4.65, tox=tox, label="Gate"),) + face(top, a, b),
            (Contact(3, 0, 0, 0, 0, 0, 0.0, label="Body"),) + face(bot, 0.0, Lx * 1e-9)]
    sweep = [(2, 0.0, 1.2, npts, {1: vd})]
    # with MLDA or Lombardi (models bits 3, 4) one wall per node: the top
    # surface (the gate oxide interface), at the node's depth y, its cell
    # spanning half a row spacing either side; otherwise none
    walls = None
    if models & (8 | 16):
        hy = Ly * 1e-9 / (ny - 1)
        y = pos[:, 1]
        walls = np.ones((N, 6))
        walls[:, 0] = y
        walls[:, 1] = np.maximum(y - 0.5 * hy, 0.0)
        walls[:, 2] = y + 0.5 * hy
    name = "prism_mosfet_%dx%dx%d" % (nx, ny, nz) + ("" if models == 7 else "_m%d" % models)
    return GraphSpec(name, pos, rp, cj, h, ar, vol, mat, Nd, Na, cons, models=models, sweep=sweep, walls=walls)


def load_graph_device(lib, spec, solver=None):
    """Send the graph device to an api3 core in the order of _setup_tri
    (main.py:3363-3410).  Returns the contact ids."""
    lib.api3_init()
    mids = [lib.api3_add_material(m, spec.T) for m in spec.materials]
    if min(mids) < 0:
        raise ValueError("too many materials")
    lib.api3_set_solver(*(solver or spec.solver))
    arrs = [np.ascontiguousarray(spec.rp, np.int32), np.ascontiguousarray(spec.cj, np.int32),
            np.ascontiguousarray(spec.h, np.float64), np.ascontiguousarray(spec.ar, np.float64),
            np.ascontiguousarray(spec.vol, np.float64), np.ascontiguousarray(spec.pos, np.float64).ravel()]
    if not lib.api3_set_mesh_graph(spec.N, spec.E, iptr(arrs[0]), iptr(arrs[1]), dptr(arrs[2]), dptr(arrs[3]),
                                   dptr(arrs[4]), dptr(arrs[5])):
        raise ValueError(f"the core refused the graph ({spec.N} nodes, {spec.E} edges)")
    lib.api3_set_models(spec.models)
    lib.api3_set_caxis(1)
    mat = np.ascontiguousarray(np.asarray(mids, np.int32)[spec.mat], np.int32)
    Nf = np.zeros(spec.N) if spec.Nf is None else np.ascontiguousarray(spec.Nf, np.float64)
    lib.api3_set_node_data(iptr(mat), dptr(np.ascontiguousarray(spec.Nd, np.float64)),
                           dptr(np.ascontiguousarray(spec.Na, np.float64)), dptr(Nf))
    lib.api3_clear_contacts()
    cids = []
    for c, nodes, areas in spec.contacts:
        cid = lib.api3_add_contact_ex(c.face, 0, 0, 0, 0, 0, 0, c.bc, c.V, c.pm, c.tox * 1e-9, c.nox * 1e4)
        if cid < 0:
            raise ValueError("too many contacts")
        n = np.ascontiguousarray(nodes, np.int32)
        a = np.ascontiguousarray(areas, np.float64)
        if not lib.api3_set_con_nodes(cid, len(n), iptr(n), dptr(a)):
            raise ValueError(f"bad node list for contact {c.label}")
        cids.append(cid)
    walls = np.ones(6 * spec.N) if spec.walls is None else np.ascontiguousarray(spec.walls, np.float64).ravel()
    lib.api3_set_walls(dptr(walls))
    if spec.tau is not None:
        lib.api3_set_mat_tau(mids[0], spec.tau[0], spec.tau[1])
    return cids
