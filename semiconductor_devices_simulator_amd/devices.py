"""Synthetic devices of the drift-diffusion workloads, and the api3_* call
sequence that loads them into a solver core.

The mesh, doping and contact helpers restate the reference front end's rules
so that a device built here reaches any api3 core (the HIP core or the
reference OpenMP core) exactly as the reference GUI would send it:

  graded_grid   main.py:316-358   Gaussian-density CDF inversion
  span_idx      main.py:363-384   half-open physical index ranges
  Implant.conc  main.py:785-798   Gaussian / erfc implant with erf lateral mask edges
  Contact.to_idx6  main.py:666-692  contact ranges placed by physical position
  load_device   main.py:3286-3361 setup_device (node-data path)

Workloads (sizes in nodes):
  pn_diode      16x16x64 Si abrupt p-n diode, 1x1x4 um, forward 0-0.8 V
  mosfet        planar n-MOSFET on an Nx x Ny x Nz tensor mesh (96x64x64,
                192x128x128), Id-Vg 0-1.2 V at Vd 0.05 V and 1.0 V
  avalanche_diode  the pn diode in reverse bias with impact ionisation
  vdmos         55 V VDMOS-like cell (256x128x256), drain sweep 0-60 V at
                Vgs 0 with impact ionisation, Id-Vds at Vgs 10 V
"""
import math
from dataclasses import dataclass, field

import numpy as np

from .core import dptr, iptr

Q = 1.602176634e-19
MAT_SI, MAT_SIO2 = 0, 8          # reference material ids (semiconductor_3d.c:431, 600)


def graded_grid(L_nm, N, dense_nm, ratio=8.0, sigma_nm=None):
    """N nodes in [0, L_nm] nm (returned in m), about `ratio` times denser
    near the points dense_nm (main.py:316-358)."""
    L = L_nm * 1e-9
    if not dense_nm or N < 3:
        return np.linspace(0, L, N)
    pts = sorted(set(dense_nm))
    if sigma_nm is None:
        sig = []
        for p in pts:
            others = [q for q in pts if q != p]
            dmin = L_nm * 0.10 if not others else min(abs(p - q) for q in others) * 0.5
            sig.append(min(dmin, L_nm * 0.05) * 1e-9)
    else:
        sig = [sigma_nm * 1e-9] * len(pts)
    sig = [max(s, L / N * 1.5) for s in sig]
    xt = np.linspace(0, L, 20000)
    dens = np.ones(20000)
    for d, s in zip(pts, sig):
        dens += (ratio - 1.0) * np.exp(-0.5 * ((xt - d * 1e-9) / s) ** 2)
    cdf = np.cumsum(dens)
    cdf -= cdf[0]
    cdf /= cdf[-1]
    xs = np.interp(np.linspace(0, 1, N), cdf, xt)
    xs[0] = 0.0
    xs[-1] = L
    for i in range(1, N):
        if xs[i] <= xs[i - 1]:
            xs[i] = xs[i - 1] + 1e-13
    return xs


def span_idx(c, a, b):
    """Inclusive index range of the half-open physical interval [a, b)
    (main.py:363-384)."""
    c = np.asarray(c)
    N = len(c)
    tol = 1e-9 * max(c[-1] - c[0], 1e-30)
    i0 = 0 if a <= c[0] + tol else int(np.searchsorted(c, a - tol, side="left"))
    i1 = N - 1 if b >= c[-1] - tol else int(np.searchsorted(c, b - tol, side="left")) - 1
    i0 = max(0, min(i0, N - 1))
    i1 = max(-1, min(i1, N - 1))
    if i1 < i0:
        k = int(np.argmin(np.abs(c - 0.5 * (a + b))))
        i0 = i1 = k
    return i0, i1


@dataclass
class Implant:
    """Doping profile added on semiconductor nodes (main.py:750-798).
    kind 'gauss': peak*exp(-((d-Rp)/Lv)^2); 'erfc': peak*erfc(d/Lv); the mask
    window [a0,a1] x [b0,b1] (nm) on the entry face has erf edges of width
    lat*Lv.  Concentrations in cm^-3, lengths in nm."""
    dtype: str
    peak: float
    face: str = "y0"
    a0: float = 0.0
    a1: float = 1e9
    b0: float = 0.0
    b1: float = 1e9
    kind: str = "gauss"
    Lv: float = 100.0
    Rp: float = 0.0
    lat: float = 0.8

    def axes(self):
        q = "xyz".index(self.face[0])
        o = [a for a in range(3) if a != q]
        return q, o[0], o[1]

    def conc(self, P, L3):
        P = np.asarray(P, float).reshape(-1, 3)
        q, ia, ib = self.axes()
        d = P[:, q] if self.face[1] == "0" else L3[q] - P[:, q]
        Lv = max(self.Lv, 1e-6)
        if self.kind == "erfc":
            v = self.peak * np.vectorize(math.erfc)(np.maximum(d, 0.0) / Lv)
        else:
            v = self.peak * np.exp(-((d - self.Rp) / Lv) ** 2)
        w = max(self.lat * Lv, 1e-9)
        erf = np.vectorize(math.erf)
        for ax, lo, hi in ((ia, self.a0, self.a1), (ib, self.b0, self.b1)):
            if lo <= 0.0 and hi >= L3[ax]:
                continue
            s = P[:, ax]
            fl = erf((s - lo) / w) if lo > 0.0 else 1.0
            fh = erf((hi - s) / w) if hi < L3[ax] else 1.0
            v = v * 0.5 * (fl + fh)
        return v


@dataclass
class Contact:
    """A contact on face 0..5 (x-,x+,y-,y+,z-,z+) over fractional in-plane
    ranges (main.py:640-692).  bc 0 ohmic, 1 Schottky, 2 gate (Robin on
    semiconductor); pm work function (eV), tox (nm)."""
    face: int
    i0p: float
    i1p: float
    j0p: float
    j1p: float
    bc: int = 0
    V: float = 0.0
    pm: float = 4.05
    tox: float = 10.0
    nox: float = 0.0
    label: str = ""

    def to_idx6(self, xs, ys, zs):
        # a box contact (face 6) here spans the x and y fractions on the z = 0
        # plane; a shaped box contact replaces it by its node list
        ax = {0: (ys, zs), 1: (ys, zs), 2: (xs, zs), 3: (xs, zs), 4: (xs, ys), 5: (xs, ys), 6: (xs, ys)}[self.face]
        c1, c2 = np.asarray(ax[0]), np.asarray(ax[1])
        i0, i1 = span_idx(c1, c1[0] + self.i0p * (c1[-1] - c1[0]), c1[0] + self.i1p * (c1[-1] - c1[0]))
        j0, j1 = span_idx(c2, c2[0] + self.j0p * (c2[-1] - c2[0]), c2[0] + self.j1p * (c2[-1] - c2[0]))
        return i0, i1, j0, j1, 0, 0


@dataclass
class DeviceSpec:
    name: str
    L3: tuple                      # domain (nm)
    xs: np.ndarray                 # node coordinates (m)
    ys: np.ndarray
    zs: np.ndarray
    mat: np.ndarray                # per-node index into `materials` (int32, x fastest)
    Nd: np.ndarray                 # m^-3
    Na: np.ndarray
    contacts: list
    materials: list = field(default_factory=lambda: [MAT_SI])
    models: int = 7
    T: float = 300.0
    solver: tuple = (2000, 1000, 500, 1e-10, 1e-9, 1.0, 1.0)
    tau: tuple = None              # (tau_n0, tau_p0) of material 0, or None
    sweep: list = field(default_factory=list)   # [(contact id, V start, V end, npts, {cid: V})]
    # gate stacks for tunnelling (models bit 5): {contact index: [(material id, thickness nm), ...]},
    # semiconductor side first (main.py:3203-3211)
    stacks: dict = field(default_factory=dict)
    # shaped contacts: {contact index: node indices} sent with api3_set_con_nodes
    # (main.py:3352-3358: the staircase of a box contact's outline)
    con_nodes: dict = field(default_factory=dict)

    @property
    def shape(self):
        return len(self.xs), len(self.ys), len(self.zs)

    @property
    def N(self):
        return len(self.xs) * len(self.ys) * len(self.zs)

    def bias_points(self):
        return sum(s[3] for s in self.sweep)


def tight_solver(solver):
    """The solver settings of a sweep converged far below the default
    tolerances (api3_set_solver): tol_p 1e-12 V and tol_c 1e-11 relative
    current, or a tenth of tol_c where that is already 1e-11 or tighter."""
    tol_c = solver[4]
    return tuple(solver[:3]) + (1e-12, 1e-11 if tol_c > 1e-11 else 0.1 * tol_c) + tuple(solver[5:])


def node_positions_nm(xs, ys, zs):
    Z, Y, X = np.meshgrid(np.asarray(zs) * 1e9, np.asarray(ys) * 1e9, np.asarray(xs) * 1e9, indexing="ij")
    return np.c_[X.ravel(), Y.ravel(), Z.ravel()]


def pn_diode(Nx=16, Ny=16, Nz=64):
    """Si abrupt p-n diode in a 1x1x4 um box: N_A 1e17 cm^-3 for z < 2 um,
    N_D 1e17 above; ohmic anode on z = 0, cathode on z = 4 um; SRH tau 1e-7 s;
    forward sweep 0 -> 0.8 V in 0.05 V steps.  Models 0 (no BGN, no tau(N),
    no velocity saturation)."""
    xs = np.linspace(0, 1e-6, Nx)
    ys = np.linspace(0, 1e-6, Ny)
    zs = graded_grid(4000.0, Nz, [2000.0])
    N = Nx * Ny * Nz
    mat = np.zeros(N, np.int32)
    Nd = np.zeros(N)
    Na = np.zeros(N)
    l0, l1 = span_idx(zs, 0.0, 2e-6)
    zl = np.repeat(np.arange(Nz), Nx * Ny)
    Na[(zl >= l0) & (zl <= l1)] = 1e23
    Nd[zl > l1] = 1e23
    cons = [Contact(4, 0.0, 1.0, 0.0, 1.0, 0, 0.0, label="Anode"),
            Contact(5, 0.0, 1.0, 0.0, 1.0, 0, 0.0, label="Cathode")]
    return DeviceSpec("pn_diode_%dx%dx%d" % (Nx, Ny, Nz), (1000.0, 1000.0, 4000.0), xs, ys, zs, mat, Nd, Na,
                      cons, models=0, tau=(1e-7, 1e-7), sweep=[(0, 0.0, 0.8, 17, {})])


def meshed_gate_diode(Nx=8, Ny=8, Nz=32, V1=2.0, npts=6):
    """The pn diode of pn_diode() with its planes x = 0 and x = 1 nm made
    SiO2 and a metal gate (work function 4.65 eV) over x = 0, y and z in
    [0.25, 0.75] of the box: the first silicon plane (x = 2.5 nm) sees the gate
    through one meshed oxide plane, so gate tunnelling (models bit 5) builds
    paths through meshed insulator nodes (tun_build :1214-1237), 176 of them
    at 8x8x32.  Models 7 | 32, anode and cathode at 0 V, gate 0 -> V1 in
    npts points."""
    spec = pn_diode(Nx, Ny, Nz)
    spec.name = "meshed_gate_diode_%dx%dx%d" % (Nx, Ny, Nz)
    spec.xs = np.concatenate([[0.0, 1.0e-9, 2.5e-9], np.linspace(0.2e-6, 1e-6, Nx - 3)])
    spec.materials = [MAT_SI, MAT_SIO2]
    mat = spec.mat.reshape(Nz, Ny, Nx).copy()
    mat[:, :, :2] = 1
    spec.mat = mat.ravel()
    ins = spec.mat == 1
    spec.Nd = np.where(ins, 0.0, spec.Nd)
    spec.Na = np.where(ins, 0.0, spec.Na)
    spec.contacts.append(Contact(0, 0.0, 1.0, 0.25, 0.75, 2, 0.0, pm=4.65, label="Gate"))
    spec.models = 7 | 32
    spec.sweep = [(2, 0.0, V1, npts, {0: 0.0, 1: 0.0})]
    return spec


def avalanche_diode(Nx=8, Ny=8, Nz=48, V1=-19.0, npts=20):
    """The pn diode of pn_diode() in reverse bias with impact ionisation on
    (models bit 8, van Overstraeten-de Man coefficients): anode 0 -> V1 in
    npts points.  The junction breaks down near -20 V; from about -13 V the
    current shows avalanche multiplication."""
    spec = pn_diode(Nx, Ny, Nz)
    spec.name = "avalanche_diode_%dx%dx%d" % (Nx, Ny, Nz)
    spec.models = 256
    spec.sweep = [(0, 0.0, V1, npts, {})]
    return spec


def mosfet(Nx=96, Ny=64, Nz=64, vds=(0.05, 1.0), npts=25, nsub=1e17, nsd=1e20, sigma=40.0, tox=2.0, lg=200.0):
    """Planar n-MOSFET, x = channel, y = depth (0.3 um, top y = 0), z = width
    (0.2 um).  p-substrate nsub cm^-3; n+ source/drain Gaussian implants (peak
    nsd cm^-3 at the surface, vertical sigma `sigma` nm, lateral 0.8 of it)
    through 200 nm windows on either side of a gate of length lg nm over tox
    nm of SiO2 (Robin face gate, work function 4.65 eV: about midgap, V_th ~
    0.5 V at the defaults); ohmic source and drain on the top face, ohmic body
    on the bottom.  Id-Vg 0 -> 1.2 V (npts points) at every drain voltage in
    vds.  The defaults are the 0.6 um device of the benchmark."""
    Lx, Ly, Lz = 400.0 + lg, 300.0, 200.0
    xs = graded_grid(Lx, Nx, [200.0, 200.0 + lg])
    ys = graded_grid(Ly, Ny, [0.0, 40.0])
    zs = np.linspace(0, Lz * 1e-9, Nz)
    N = Nx * Ny * Nz
    mat = np.zeros(N, np.int32)
    Na = np.full(N, nsub * 1e6)
    P = node_positions_nm(xs, ys, zs)
    L3 = (Lx, Ly, Lz)
    nd = Implant("n", nsd, "y0", 0.0, 200.0, 0.0, 1e9, "gauss", sigma, 0.0, 0.8).conc(P, L3)
    nd += Implant("n", nsd, "y0", 200.0 + lg, 1e9, 0.0, 1e9, "gauss", sigma, 0.0, 0.8).conc(P, L3)
    Nd = nd * 1e6
    a, b = 200.0 / Lx, (200.0 + lg) / Lx
    cons = [Contact(2, 0.0, a, 0.0, 1.0, 0, 0.0, label="Source"),
            Contact(2, b, 1.0, 0.0, 1.0, 0, vds[0], label="Drain"),
            Contact(2, a, b, 0.0, 1.0, 2, 0.0, pm=4.65, tox=tox, label="Gate"),
            Contact(3, 0.0, 1.0, 0.0, 1.0, 0, 0.0, label="Body")]
    sweep = [(2, 0.0, 1.2, npts, {1: vd}) for vd in vds]
    return DeviceSpec("mosfet_%dx%dx%d" % (Nx, Ny, Nz), L3, xs, ys, zs, mat, Nd, Na, cons,
                      models=7, sweep=sweep)


def vdmos(Nx=256, Ny=128, Nz=256, vd_max=60.0, npts=61, vgs_on=10.0, npts_on=0, vd_on_max=20.0, tol_c=1e-9):
    """55 V VDMOS-like cell, 20 x 4 x 8 um (x lateral, y depth with the top
    at y = 0, z width), after the reference's VDMOS template (main.py:984-1059).
    n- drift 5e15 cm^-3; n+ substrate 1e19 cm^-3 below y = 3.5 um (drain on the
    bottom face); p-body Gaussian implants (peak 1e17 cm^-3, Lv 0.6 um) under
    both outer thirds of the top face, leaving a JFET window in the middle;
    n+ source (1e20 cm^-3, Lv 0.1 um) and p+ body tap (1e19 cm^-3, Lv 0.2 um)
    under the source contacts at both outer edges; two gates (Robin, 50 nm
    oxide, work function 4.1 eV) over the channels.  Models 7 | 256
    (impact ionisation on).  Sweep: drain 0 -> vd_max in npts points at
    Vgs = 0 (breakdown), then, when npts_on > 0, the drain back at 0 V and
    Id-Vds 0 -> vd_on_max in npts_on points at Vgs = vgs_on.  tol_c: the
    Gummel loop's relative current tolerance (api3_set_solver)."""
    Lx, Ly, Lz = 20000.0, 4000.0, 8000.0
    tap_w, src_w, body_x = 1500.0, 2500.0, 7000.0      # from each outer edge
    sub_top = 3500.0
    xs = graded_grid(Lx, Nx, [tap_w + src_w, body_x, Lx - body_x, Lx - tap_w - src_w])
    ys = graded_grid(Ly, Ny, [0.0, 1000.0, sub_top])
    zs = np.linspace(0, Lz * 1e-9, Nz)
    N = Nx * Ny * Nz
    mat = np.zeros(N, np.int32)
    P = node_positions_nm(xs, ys, zs)
    L3 = (Lx, Ly, Lz)
    Nd = np.full(N, 5e15) + np.where(P[:, 1] >= sub_top, 1e19, 0.0)
    Na = np.zeros(N)
    for a0, a1 in ((0.0, body_x), (Lx - body_x, 1e9)):
        Na += Implant("p", 1e17, "y0", a0, a1, 0.0, 1e9, "gauss", 600.0, 0.0, 0.8).conc(P, L3)
    for a0, a1 in ((0.0, tap_w), (Lx - tap_w, 1e9)):
        Na += Implant("p", 1e19, "y0", a0, a1, 0.0, 1e9, "gauss", 200.0, 0.0, 0.8).conc(P, L3)
    for a0, a1 in ((tap_w, tap_w + src_w), (Lx - tap_w - src_w, Lx - tap_w)):
        Nd += Implant("n", 1e20, "y0", a0, a1, 0.0, 1e9, "gauss", 100.0, 0.0, 0.8).conc(P, L3)
    sf = (tap_w + src_w) / Lx
    g0, g1 = (tap_w + src_w - 200.0) / Lx, (body_x + 500.0) / Lx
    cons = [Contact(2, 0.0, sf, 0.0, 1.0, 0, 0.0, label="Source-L"),
            Contact(2, 1.0 - sf, 1.0, 0.0, 1.0, 0, 0.0, label="Source-R"),
            Contact(2, g0, g1, 0.0, 1.0, 2, 0.0, pm=4.1, tox=50.0, label="Gate-L"),
            Contact(2, 1.0 - g1, 1.0 - g0, 0.0, 1.0, 2, 0.0, pm=4.1, tox=50.0, label="Gate-R"),
            Contact(3, 0.0, 1.0, 0.0, 1.0, 0, 0.0, label="Drain")]
    sweep = [(4, 0.0, vd_max, npts, {2: 0.0, 3: 0.0})]
    if npts_on > 0:
        # the drain returns to 0 V before the gates turn on (the gates move
        # first within a resolve, and at full drain voltage an on-state
        # device would be driven into avalanche)
        sweep.append((4, 0.0, 0.0, 1, {2: 0.0, 3: 0.0}))
        sweep.append((4, 0.0, vd_on_max, npts_on, {2: vgs_on, 3: vgs_on}))
    return DeviceSpec("vdmos_%dx%dx%d" % (Nx, Ny, Nz), L3, xs, ys, zs, mat, Nd * 1e6, Na * 1e6, cons,
                      models=7 | 256, sweep=sweep, solver=(2000, 1000, 500, 1e-10, tol_c, 1.0, 1.0))


def soi_mosfet(Nx=24, Ny=16, Nz=8, tsi=80.0, tbox=100.0, lg=300.0, nbody=1e18, nsd=1e20, tox=1.5,
               vds=(0.05, 1.0), npts=7, tunnelling=True):
    """Floating-body SOI n-MOSFET: a tsi nm silicon film (p-body nbody cm^-3,
    n+ source/drain implants from the top reaching through the film) on a
    tbox nm buried oxide whose bottom face is a metal back gate at 0 V (an
    ohmic contact on insulator nodes), a Robin front gate of tox nm SiO2
    (work function 4.65 eV) and no body contact.  The p-body is a hole-majority
    block touching no contact, so the continuity solves rescale it to its
    global balance (the reference's float_fix); with `tunnelling` the gate
    stack's tunnelling sink/source (models bit 5) enters that balance.
    Id-Vg 0 -> 1.2 V in npts points at every drain voltage in vds."""
    Lx, Ly, Lz = 400.0 + lg, tsi + tbox, 200.0
    xs = graded_grid(Lx, Nx, [200.0, 200.0 + lg])
    ys = graded_grid(Ly, Ny, [0.0, tsi])
    zs = np.linspace(0, Lz * 1e-9, Nz)
    N = Nx * Ny * Nz
    P = node_positions_nm(xs, ys, zs)
    L3 = (Lx, Ly, Lz)
    j0, j1 = span_idx(ys, 0.0, tsi * 1e-9)
    yl = np.tile(np.repeat(np.arange(Ny), Nx), Nz)
    film = yl <= j1
    mat = np.where(film, 0, 1).astype(np.int32)
    Na = np.where(film, nbody * 1e6, 0.0)
    nd = Implant("n", nsd, "y0", 0.0, 200.0, 0.0, 1e9, "gauss", tsi, 0.0, 0.8).conc(P, L3)
    nd += Implant("n", nsd, "y0", 200.0 + lg, 1e9, 0.0, 1e9, "gauss", tsi, 0.0, 0.8).conc(P, L3)
    Nd = np.where(film, nd * 1e6, 0.0)
    a, b = 200.0 / Lx, (200.0 + lg) / Lx
    cons = [Contact(2, 0.0, a, 0.0, 1.0, 0, 0.0, label="Source"),
            Contact(2, b, 1.0, 0.0, 1.0, 0, vds[0], label="Drain"),
            Contact(2, a, b, 0.0, 1.0, 2, 0.0, pm=4.65, tox=tox, label="Gate"),
            Contact(3, 0.0, 1.0, 0.0, 1.0, 0, 0.0, label="BackGate")]
    sweep = [(2, 0.0, 1.2, npts, {1: vd}) for vd in vds]
    spec = DeviceSpec("soi_mosfet_%dx%dx%d" % (Nx, Ny, Nz), L3, xs, ys, zs, mat, Nd, Na, cons,
                      materials=[MAT_SI, MAT_SIO2], models=7, sweep=sweep)
    if tunnelling:
        spec = with_gate_tunnelling(spec, gates=[2])
    return spec


def shaped_anode_diode(Nx=16, Ny=16, Nz=64, radius=0.35):
    """The pn diode of pn_diode() with its anode a shaped box contact: the
    nodes of the two lowest planes (z = 0 and the next) within `radius` um of
    the face centre, a staircase disc sent as a node list
    (api3_set_con_nodes, as the front end sends shaped contacts); the rest of
    the bottom face is uncontacted.  Forward sweep 0 -> 0.8 V."""
    spec = pn_diode(Nx, Ny, Nz)
    spec.name = "shaped_anode_diode_%dx%dx%d" % (Nx, Ny, Nz)
    X, Y = np.meshgrid(spec.xs, spec.ys, indexing="xy")
    disc = ((X - 0.5e-6) ** 2 + (Y - 0.5e-6) ** 2 <= (radius * 1e-6) ** 2).ravel()
    plane = np.nonzero(disc)[0]
    nodes = np.concatenate([plane, plane + Nx * Ny])
    spec.contacts[0] = Contact(6, 0.0, 1.0, 0.0, 1.0, 0, 0.0, label="Anode")
    spec.con_nodes = {0: nodes}
    return spec


def guard_ring_diode(Nx=32, Ny=16, Nz=4, v_end=-8.0, npts=9, gate_v=None, tox=2.0, ring_contact=False):
    """Planar p+/n- diode with a floating guard ring, 8 x 4 x 1 um (x lateral,
    y depth with the top at y = 0, z width).  n- drift 1e15 cm^-3 over an n+
    substrate (1e19 cm^-3 below y = 3.5 um, cathode on the bottom face); p+
    main junction (Gaussian, peak 1e19 cm^-3, Lv 0.5 um) under the anode
    window x < 2 um; a second p+ implant, the guard ring, at 4.5-5.3 um with no
    contact.  The ring is a hole-majority block that touches no contact, so
    every hole continuity solve rescales it to its global balance (the
    reference's float_fix).  With gate_v set, a Robin gate of tox nm SiO2 over
    the ring (work function 4.65 eV) is held at gate_v with gate tunnelling on
    (models bit 5), so that the tunnelling sink/source enters the ring's
    balance.  Reverse sweep of the anode 0 -> v_end in npts points (beyond
    about -8 V the ungated ring's potential is no longer determined to the
    solver's tolerances: its balance currents sink below their floor)."""
    Lx, Ly, Lz = 8000.0, 4000.0, 1000.0
    xs = graded_grid(Lx, Nx, [2000.0, 4500.0, 5300.0])
    ys = graded_grid(Ly, Ny, [0.0, 600.0, 3500.0])
    zs = np.linspace(0, Lz * 1e-9, Nz)
    N = Nx * Ny * Nz
    mat = np.zeros(N, np.int32)
    P = node_positions_nm(xs, ys, zs)
    L3 = (Lx, Ly, Lz)
    Nd = np.full(N, 1e15) + np.where(P[:, 1] >= 3500.0, 1e19, 0.0)
    Na = Implant("p", 1e19, "y0", 0.0, 2000.0, 0.0, 1e9, "gauss", 500.0, 0.0, 0.8).conc(P, L3)
    Na += Implant("p", 1e19, "y0", 4500.0, 5300.0, 0.0, 1e9, "gauss", 500.0, 0.0, 0.8).conc(P, L3)
    cons = [Contact(2, 0.0, 1500.0 / Lx, 0.0, 1.0, 0, 0.0, label="Anode"),
            Contact(3, 0.0, 1.0, 0.0, 1.0, 0, 0.0, label="Cathode")]
    spec = DeviceSpec("guard_ring_diode_%dx%dx%d" % (Nx, Ny, Nz), L3, xs, ys, zs, mat, Nd * 1e6, Na * 1e6, cons,
                      models=7, sweep=[(0, 0.0, v_end, npts, {})])
    if ring_contact:
        # an ohmic contact on the ring, for the front end's floating-terminal
        # search (sweep.floating_sweep finds its zero-current voltage)
        spec.contacts.append(Contact(2, 4700.0 / Lx, 5100.0 / Lx, 0.0, 1.0, 0, 0.0, label="Ring"))
        spec.name = spec.name.replace("guard_ring_diode", "guard_ring_contact_diode")
    if gate_v is not None:
        spec.contacts.append(Contact(2, 4500.0 / Lx, 5300.0 / Lx, 0.0, 1.0, 2, gate_v, pm=4.65, tox=tox,
                                     label="Gate"))
        spec.sweep = [(0, 0.0, v_end, npts, {2: gate_v})]
        spec = with_gate_tunnelling(spec, gates=[2])
        spec.name = spec.name.replace("guard_ring_diode", "guard_ring_gate_diode")
    return spec


def random_mosfet(seed, Nx=96, Ny=64, Nz=64, npts=25):
    """A MOSFET of the warm-start training family, drawn with
    numpy.random.default_rng(seed): substrate log-uniform 1e15-1e18 cm^-3,
    source/drain peak log-uniform 1e19-1e20 cm^-3 with sigma uniform 20-60 nm,
    oxide uniform 1.5-5 nm, gate length uniform 0.1-0.5 um."""
    rng = np.random.default_rng(seed)
    p = dict(nsub=10 ** rng.uniform(15, 18), nsd=10 ** rng.uniform(19, 20), sigma=rng.uniform(20, 60),
             tox=rng.uniform(1.5, 5.0), lg=rng.uniform(100, 500))
    spec = mosfet(Nx, Ny, Nz, npts=npts, **p)
    spec.name = "random_mosfet_%d" % seed
    spec.params = p
    return spec


WORKLOADS = {
    "pn_diode": lambda: pn_diode(),
    "mosfet_96x64x64": lambda: mosfet(96, 64, 64),
    "mosfet_192x128x128": lambda: mosfet(192, 128, 128),
}


def load_device(lib, spec, solver=None):
    """Send the device to an api3 core (main.py:3286-3361, node-data path).
    Returns the contact ids."""
    lib.api3_init()
    mids = [lib.api3_add_material(m, spec.T) for m in spec.materials]
    if min(mids) < 0:
        raise ValueError("too many materials")
    lib.api3_set_solver(*(solver or spec.solver))
    xs = np.ascontiguousarray(spec.xs, np.float64)
    ys = np.ascontiguousarray(spec.ys, np.float64)
    zs = np.ascontiguousarray(spec.zs, np.float64)
    Nx, Ny, Nz = spec.shape
    if not lib.api3_set_grid_coords(Nx, Ny, Nz, dptr(xs), dptr(ys), dptr(zs)):
        raise MemoryError(f"could not allocate {Nx}x{Ny}x{Nz} nodes")
    lib.api3_set_models(spec.models)
    lib.api3_set_caxis(1)
    mat = np.ascontiguousarray(np.asarray(mids, np.int32)[spec.mat], np.int32)
    Nd = np.ascontiguousarray(spec.Nd, np.float64)
    Na = np.ascontiguousarray(spec.Na, np.float64)
    Nf = np.zeros(spec.N)
    lib.api3_set_node_data(iptr(mat), dptr(Nd), dptr(Na), dptr(Nf))
    lib.api3_clear_contacts()
    cids = []
    for c in spec.contacts:
        i0, i1, j0, j1, k0, k1 = c.to_idx6(spec.xs, spec.ys, spec.zs)
        cid = lib.api3_add_contact_ex(c.face, i0, i1, j0, j1, k0, k1, c.bc, c.V, c.pm,
                                      c.tox * 1e-9, c.nox * 1e4)
        if cid < 0:
            raise ValueError("too many contacts")
        cids.append(cid)
    for ci, nodes in spec.con_nodes.items():
        nodes = np.ascontiguousarray(nodes, np.int32)
        if not lib.api3_set_con_nodes(cids[ci], len(nodes), iptr(nodes), dptr(np.zeros(len(nodes)))):
            raise ValueError(f"bad node list for contact {ci}")
    if spec.tau is not None:
        lib.api3_set_mat_tau(mids[0], spec.tau[0], spec.tau[1])
    known = dict(zip(spec.materials, mids))
    for ci, layers in spec.stacks.items():
        ids = []
        for m, _ in layers:
            if m not in known:
                known[m] = lib.api3_add_material(m, spec.T)
                if known[m] < 0:
                    raise ValueError("too many materials")
            ids.append(known[m])
        mi = np.asarray(ids, np.int32)
        th = np.asarray([t * 1e-9 for _, t in layers], np.float64)
        if not lib.api3_set_gate_stack(cids[ci], len(layers), iptr(mi), dptr(th)):
            raise ValueError(f"bad gate stack for contact {ci}")
    return cids


MAT_SIO2 = 8


def with_gate_tunnelling(spec, gates=None):
    """The device with gate tunnelling on (models bit 5): every Robin gate
    (or the contacts listed in `gates`) gets a SiO2 stack of its oxide
    thickness (main.py:3203-3211 with the gate's oxide layer)."""
    gates = [i for i, c in enumerate(spec.contacts) if c.bc == 2] if gates is None else gates
    spec.models |= 32
    spec.stacks = {i: [(MAT_SIO2, spec.contacts[i].tox)] for i in gates}
    spec.name += "_tun"
    return spec
