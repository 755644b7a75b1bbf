"""CPU restatement of the reference drift-diffusion kernels - TEST
INFRASTRUCTURE ONLY.

This module is the checker of the HIP core: only tests/, __graft_entry__.smoke()
and bench.py's cpu_baseline leg may import it, and only to compute expected
values.  It is never the thing measured or shipped; the product path
(semiconductor_devices_simulator_amd) has no CPU fallback.

Vectorised numpy over node arrays in the reference layout (node index
k = i + Nx*(j + Ny*l), semiconductor_3d.c:114).  Each function cites the
reference function it restates.  Pinning (see tests/test_oracle_*.py):
  * bernoulli, recomb, mob_low, bgn_half, neutral_psi: closed forms and
    identities, and the reference harness (oracle/_ref/libsemisim_refh.so)
    point values;
  * prepare, poisson_eval, face_mobility, cont_assemble, matvec,
    mg_hierarchy/galerkin/restrict/prolong, terminal_currents, cont_update,
    ii_update, walls (interface distance, MLDA), face_mobility with Lombardi,
    charge_partition, terminal_charges, readback, Anderson, tun_paths,
    tun_update, tun_gather, tun_gate_currents, and on graph meshes
    GraphDevice, graph_poisson, graph_mobility, graph_cont, graph_currents:
    compared with the reference harness on small devices (golden fixtures
    tests/golden/*.npz written by tests/golden/make_golden.py) and with
    invariants (Poisson Jacobian symmetric on free rows, row sums of the
    scaled SG operator at equilibrium, Galerkin of the constant vector).
"""
import math

import numpy as np

Q = 1.602176634e-19
KB = 1.380649e-23
EPS0 = 8.8541878128e-12
EPS_OX = 3.9 * EPS0
LN_MIN, LN_MAX = -230.0, 150.0
NFLOOR = 1e-100
K_SEMI, K_INS, K_METAL = 0, 1, 2


# ------------------------------------------------------------ scalar physics
def bernoulli(x):
    """B(x) = x/(e^x - 1), semiconductor_3d.c:670-675."""
    x = np.asarray(x, dtype=np.float64)
    out = np.empty_like(x)
    a = np.abs(x)
    small = a < 1e-6
    big = x > 700.0
    neg = x < -700.0
    mid = ~(small | big | neg)
    out[small] = 1.0 - x[small] / 2.0 + x[small] ** 2 / 12.0
    out[big] = 0.0
    out[neg] = -x[neg]
    out[mid] = x[mid] / np.expm1(x[mid])
    return out


def silicon(T=300.0):
    """Si parameters of mat_init (semiconductor_3d.c:431-473), SI units."""
    tr = T / 300.0
    m = dict(id=0, insulator=0, T=T, VT=KB * T / Q)
    m["Eg"] = 1.170 - 4.73e-4 * T * T / (T + 636.0)
    m["chi"], m["eps_r"] = 4.05, 11.7
    m["Nc"], m["Nv"] = 2.86e25 * tr ** 1.58, 3.10e25 * tr ** 1.85
    m["mu_max"] = (0.1417, 0.04705)
    m["mt_max"] = (-2.5, -2.2)
    m["mas"] = dict(min1=(0.00522, 0.00449), min2=(0.00522, 0.0), mu1=(0.00434, 0.00290), Pc=(0.0, 9.23e22),
                    Cr=(9.68e22, 2.23e23), Cs=(3.43e26, 6.10e26), a=(0.680, 0.719), b=(2.0, 2.0))
    m["vsat"] = (1.07e5 * tr ** -0.87, 8.37e4 * tr ** -0.52)
    m["beta"] = (1.109 * tr ** 0.66, 1.213 * tr ** 0.17)
    m["tau"] = (1e-5, 1e-5)
    m["Cn"], m["Cp"], m["Brad"], m["tau_Nref"] = 2.8e-43, 9.9e-44, 4.73e-21, 1e22
    m["bgn"] = (6.92e-3, 1.3e23, 0.5)
    m["ni"] = math.sqrt(m["Nc"] * m["Nv"]) * math.exp(-m["Eg"] / (2 * m["VT"]))
    # van Overstraeten-de Man impact ionisation (:470-472): a, b per carrier
    # and field range (below / above E0), optical-phonon energy hw (:626)
    m["ii_a"] = ((7.03e7, 7.03e7), (1.582e8, 6.71e7))
    m["ii_b"] = ((1.231e8, 1.231e8), (2.036e8, 1.693e8))
    m["ii_E0"], m["ii_m"] = (4e7, 4e7), (1.0, 1.0)
    hw = 0.063
    m["ii_g"] = math.tanh(hw / (2.0 * KB * 300.0 / Q)) / math.tanh(hw / (2.0 * m["VT"]))
    # Lombardi surface mobility (:462-465), cm units, (electrons, holes)
    m["lbB"], m["lbC"] = (4.75e7, 9.925e6), (580.0, 2947.0)
    m["lbL"], m["lbD"], m["lbE"] = (0.125, 0.0317), (5.82e14, 2.0546e14), (0.0, 0.0)
    # MLDA confinement masses (m0) and valley weights (:457): (e1, e2, w_e1), (h1, h2, w_h1)
    m["mq_e"], m["mq_h"] = (0.916, 0.19, 1.0 / 3.0), (0.29, 0.20, 0.84)
    return m


def mob_low(m, Ntot, hole):
    """Masetti low-field mobility, semiconductor_3d.c:646-656."""
    N = np.maximum(np.asarray(Ntot, dtype=np.float64), 1.0)
    c = 1 if hole else 0
    s = m["mas"]
    muL = m["mu_max"][c] * (m["T"] / 300.0) ** m["mt_max"][c]
    mu = (s["min1"][c] * np.exp(-s["Pc"][c] / N) + (muL - s["min2"][c]) / (1 + (N / s["Cr"][c]) ** s["a"][c])
          - s["mu1"][c] / (1 + (s["Cs"][c] / N) ** s["b"][c]))
    return np.maximum(mu, 1e-6)


def mob_hf(mu0, E, vsat, beta):
    """Caughey-Thomas saturation, semiconductor_3d.c:664-668."""
    mu0 = np.asarray(mu0, dtype=np.float64)
    r = mu0 * np.abs(E) / vsat
    out = mu0 / (1.0 + r ** beta) ** (1.0 / beta)
    return np.where(r < 1e-12, mu0, out)


def heat_field(dq, dps, h):
    """Velocity-saturation driving field, semiconductor_3d.c:2882-2886."""
    a, b = np.abs(dq), np.abs(dps)
    with np.errstate(divide="ignore", invalid="ignore"):
        F = 2.0 * a * b / ((a + b) * h)
    return np.where(dq * dps <= 0.0, 0.0, F)


def bgn_half(m, N):
    """Klaassen band-gap narrowing / (2 VT), semiconductor_3d.c:839-843."""
    N = np.asarray(N, dtype=np.float64)
    E, Nr, C = m["bgn"]
    with np.errstate(divide="ignore", invalid="ignore"):
        L = np.log(N / Nr)
        v = 0.5 * E * (L + np.sqrt(L * L + C)) / m["VT"]
    return np.where(N < 1.0, 0.0, v)


def recomb(m, ni, tf, n, p):
    """SRH + Auger + radiative and derivatives, semiconductor_3d.c:686-697."""
    tn, tp = m["tau"]
    ni2 = ni * ni
    np_ = n * p - ni2
    den = tf * (tp * (n + ni) + tn * (p + ni)) + 1e-300
    Rs = np_ / den
    dsn = (p * den - np_ * tf * tp) / (den * den)
    dsp = (n * den - np_ * tf * tn) / (den * den)
    Au = m["Cn"] * n + m["Cp"] * p
    R = Rs + Au * np_ + m["Brad"] * np_
    dRn = dsn + m["Cn"] * np_ + Au * p + m["Brad"] * p
    dRp = dsp + m["Cp"] * np_ + Au * n + m["Brad"] * n
    return R, dRn, dRp


# ------------------------------------------------------------ device
class Device:
    """Prepared device: the node data prepare() derives
    (semiconductor_3d.c:1638-1775) for a tensor mesh of one semiconductor
    (material 0) with face contacts.  contacts: list of dicts with face,
    i0, i1, j0, j1, bc, V, pm, tox (m), qox (C/m^2)."""

    def __init__(self, xs, ys, zs, Nd, Na, contacts, models=7, tau=None, T=300.0, Nf=None):
        self.xs, self.ys, self.zs = (np.asarray(a, dtype=np.float64) for a in (xs, ys, zs))
        self.Nx, self.Ny, self.Nz = len(xs), len(ys), len(zs)
        self.N = self.Nx * self.Ny * self.Nz
        self.Lx, self.Ly, self.Lz = (fv_lengths(c) for c in (self.xs, self.ys, self.zs))
        self.m = silicon(T)
        if tau is not None:
            self.m["tau"] = tuple(t if t > 0 else 1e30 for t in tau)
        self.models = models
        self.Nd = np.asarray(Nd, dtype=np.float64)
        self.Na = np.asarray(Na, dtype=np.float64)
        self.Nf = np.zeros(self.N) if Nf is None else np.asarray(Nf, dtype=np.float64)
        self.contacts = contacts
        m = self.m
        VT = m["VT"]
        self.Cref = m["chi"] + 0.5 * m["Eg"] + 0.5 * VT * math.log(m["Nc"] / m["Nv"])      # mat_prepare :634
        self.lnc = math.log(m["Nc"]) + (m["chi"] - self.Cref) / VT
        self.lnv = math.log(m["Nv"]) + (self.Cref - m["chi"] - m["Eg"]) / VT
        self.kind = np.zeros(self.N, np.uint8)
        self.eps = np.full(self.N, m["eps_r"] * EPS0)
        self.bgn = bgn_half(m, self.Nd + self.Na) if (models & 1) else np.zeros(self.N)
        self.con_mask = np.zeros(self.N, np.int32)
        self.gate_mask = np.zeros(self.N, np.int32)
        self.robA = np.zeros(self.N)
        self.areas = []
        for c, cn in enumerate(contacts):                        # :1656-1668
            A = 0.0
            for k, area in self.contact_nodes(cn):
                A += area
                if cn["bc"] == 2 and self.kind[k] == K_SEMI:
                    self.gate_mask[k] = c + 1
                elif cn["bc"] == 2 or self.kind[k] != K_SEMI:
                    self.kind[k] = K_METAL
                    self.con_mask[k] = c + 1
                    self.gate_mask[k] = 0
                else:
                    self.con_mask[k] = c + 1
                    self.gate_mask[k] = 0
            self.areas.append(A)
        for c, cn in enumerate(contacts):                        # Robin list :1676-1686
            if cn["bc"] != 2:
                continue
            for k, area in self.contact_nodes(cn):
                if self.gate_mask[k] == c + 1 and not self.con_mask[k]:
                    self.robA[k] = area
        # MLDA (models bit 3): confinement terms of ln Nc, ln Nv near interfaces (:736-738)
        if models & 8:
            _, self.qcn, self.qcp = walls(self, mlda=True)
        else:
            self.qcn, self.qcp = np.zeros(self.N), np.zeros(self.N)
        self.lncK = self.lnc + self.bgn + self.qcn
        self.lnvK = self.lnv + self.bgn + self.qcp
        self.niK = m["ni"] * np.exp(self.bgn + 0.5 * (self.qcn + self.qcp))
        semi = self.kind == K_SEMI
        self.mu0n = np.where(semi, mob_low(m, self.Nd + self.Na, 0), 0.0)
        self.mu0p = np.where(semi, mob_low(m, self.Nd + self.Na, 1), 0.0)
        if (models & 2) and m["tau_Nref"] > 0:
            self.tauf = 1.0 / (1.0 + (self.Nd + self.Na) / m["tau_Nref"])
        else:
            self.tauf = np.ones(self.N)

    def ijl(self):
        k = np.arange(self.N)
        NxNy = self.Nx * self.Ny
        l = k // NxNy
        r = k - l * NxNy
        j = r // self.Nx
        return k - j * self.Nx - l * NxNy, j, l

    def contact_nodes(self, cn):
        """FOR_CON_NODES, semiconductor_3d.c:1047-1065 (face contacts)."""
        Nx, Ny, Nz = self.Nx, self.Ny, self.Nz
        f = cn["face"]
        ia, ib, ja, jb, la, lb = 0, Nx - 1, 0, Ny - 1, 0, Nz - 1
        if f in (0, 1):
            ia = ib = 0 if f == 0 else Nx - 1
            ja, jb, la, lb = cn["i0"], cn["i1"], cn["j0"], cn["j1"]
        elif f in (2, 3):
            ja = jb = 0 if f == 2 else Ny - 1
            ia, ib, la, lb = cn["i0"], cn["i1"], cn["j0"], cn["j1"]
        else:
            la = lb = 0 if f == 4 else Nz - 1
            ia, ib, ja, jb = cn["i0"], cn["i1"], cn["j0"], cn["j1"]
        ia, ja, la = max(ia, 0), max(ja, 0), max(la, 0)
        ib, jb, lb = min(ib, Nx - 1), min(jb, Ny - 1), min(lb, Nz - 1)
        for l in range(la, lb + 1):
            for j in range(ja, jb + 1):
                for i in range(ia, ib + 1):
                    k = i + Nx * (j + Ny * l)
                    if f <= 1:
                        area = self.Ly[j] * self.Lz[l]
                    elif f <= 3:
                        area = self.Lx[i] * self.Lz[l]
                    else:
                        area = self.Lx[i] * self.Ly[j]
                    yield k, area

    def edges(self, ax):
        """Edges k -> k + e_ax: (k, kp, h, area) arrays (nbr(), semiconductor_3d.c:703-713)."""
        i, j, l = self.ijl()
        k = np.arange(self.N)
        if ax == 0:
            sel = i < self.Nx - 1
            kp = k + 1
            h = self.xs[np.minimum(i + 1, self.Nx - 1)] - self.xs[i]
            ar = self.Ly[j] * self.Lz[l]
        elif ax == 1:
            sel = j < self.Ny - 1
            kp = k + self.Nx
            h = self.ys[np.minimum(j + 1, self.Ny - 1)] - self.ys[j]
            ar = self.Lx[i] * self.Lz[l]
        else:
            sel = l < self.Nz - 1
            kp = k + self.Nx * self.Ny
            h = self.zs[np.minimum(l + 1, self.Nz - 1)] - self.zs[l]
            ar = self.Lx[i] * self.Ly[j]
        return k[sel], kp[sel], h[sel], ar[sel]

    def volume(self):
        i, j, l = self.ijl()
        return self.Lx[i] * self.Ly[j] * self.Lz[l]

    def carriers(self, psi, phin, phip):
        """rebuild_np, semiconductor_3d.c:2826-2834."""
        VT = self.m["VT"]
        semi = self.kind == K_SEMI
        n = np.exp(np.clip(self.lncK + (psi - phin) / VT, LN_MIN, LN_MAX))
        p = np.exp(np.clip(self.lnvK + (phip - psi) / VT, LN_MIN, LN_MAX))
        return np.where(semi, n, 0.0), np.where(semi, p, 0.0)

    def neutral_psi(self, V):
        """Neutral potential of every node (closed form, no traps), :1785."""
        D = self.Nd - self.Na
        return V + self.m["VT"] * (np.arcsinh(D / (2.0 * self.niK)) - 0.5 * (self.lncK - self.lnvK))


def fv_lengths(c):
    """Finite-volume lengths (half cells on faces), fv_len :700-705."""
    N = len(c)
    L = np.empty(N)
    L[1:-1] = 0.5 * (c[2:] - c[:-2])
    L[0] = 0.5 * (c[1] - c[0])
    L[-1] = 0.5 * (c[-1] - c[-2])
    return L


DIRS = ((0, -1), (0, 1), (1, -1), (1, 1), (2, -1), (2, 1))   # (axis, side): x-, x+, y-, y+, z-, z+


def poisson_eval(dev, psi, phin, phip, Va, build=True, sc_prev=None):
    """Poisson residual F and Jacobian (7 diagonals) and ||F/sc||,
    poisson_eval semiconductor_3d.c:2732-2789."""
    N = dev.N
    F = np.zeros(N)
    dg = np.zeros(N)
    A = np.zeros((7, N))
    free = dev.con_mask == 0
    for ax in range(3):
        k, kp, h, ar = dev.edges(ax)
        ek, eb = dev.eps[k], dev.eps[kp]
        cf = (2.0 * ek * eb / (ek + eb)) * ar / h
        flux = cf * (psi[kp] - psi[k])
        np.add.at(F, k, flux)
        np.add.at(F, kp, -flux)
        np.add.at(dg, k, cf)
        np.add.at(dg, kp, cf)
        A[2 + 2 * ax, k] = np.where(free[kp], -cf, 0.0)       # + neighbour of k
        A[1 + 2 * ax, kp] = np.where(free[k], -cf, 0.0)       # - neighbour of kp
    vol = dev.volume()
    VT = dev.m["VT"]
    semi = dev.kind == K_SEMI
    n, p = dev.carriers(psi, phin, phip)
    F += np.where(semi, vol * Q * (p - n + dev.Nd - dev.Na + dev.Nf), vol * Q * dev.Nf)
    dg += np.where(semi, vol * Q * (n + p) / VT, 0.0)
    for c, cn in enumerate(dev.contacts):
        if cn["bc"] != 2:
            continue
        rob = (dev.gate_mask == c + 1) & free & (dev.robA > 0)
        Cox = EPS_OX / (cn["tox"] if cn["tox"] > 0 else 10e-9)
        pm = cn["pm"] if cn["pm"] > 0 else dev.Cref
        VG = Va[c] + dev.Cref - pm
        F[rob] += dev.robA[rob] * (Cox * (VG - psi[rob]) + cn.get("qox", 0.0))
        dg[rob] += Cox * dev.robA[rob]
    F[~free] = 0.0
    A[0] = dg
    A[:, ~free] = 0.0
    A[0, ~free] = 1.0
    sc = A[0].copy() if build else sc_prev
    return F, A, sc, float(np.sqrt(np.sum((F / sc) ** 2)))


HBAR, M0 = 1.054571817e-34, 9.1093837015e-31


def mlda_pop(a, b, lam):
    """Fraction of the bulk density kept, averaged over the cell [a, b] from
    a wall, for thermal length lam: 1 - exp(-(z/lam)^2) averaged, mlda_pop
    semiconductor_3d.c:1073-1081 (series for small cells)."""
    u, v = a / lam, b / lam
    if v < 0.05:
        d = v - u
        if d < 1e-12:
            return u * u
        return ((v ** 3 - u ** 3) / 3.0 - (v ** 5 - u ** 5) / 10.0 + (v ** 7 - u ** 7) / 42.0) / d
    if v - u < 1e-9 * v:
        return 1.0 - math.exp(-u * u)
    return 1.0 - 0.5 * math.sqrt(math.pi) * (math.erf(v) - math.erf(u)) / (v - u)


def mlda_factor(m, a, b, hole):
    """Cell-averaged MLDA factor of two valley masses, mlda_factor :1082-1089."""
    m1, m2, w = m["mq_h"] if hole else m["mq_e"]
    kT = KB * m["T"]
    a = max(a, 0.0)
    b = max(b, a)
    f = w * mlda_pop(a, b, HBAR / math.sqrt(2.0 * m1 * M0 * kT))
    if w < 1.0 and m2 > 0.0:
        f += (1.0 - w) * mlda_pop(a, b, HBAR / math.sqrt(2.0 * m2 * M0 * kT))
    return f


def wall_distance(dev):
    """Distance (m) of every semiconductor node to the nearest insulator
    interface or Robin gate face along the mesh lines, 1 where there is
    none: the qd part of walls_tensor, semiconductor_3d.c:1100-1135."""
    return walls(dev)[0]


def walls(dev, mlda=False):
    """The wall scan of walls_tensor (semiconductor_3d.c:1100-1146): qd, and
    with mlda the sums over the walls within 40 nm of ln(cell-averaged MLDA
    factor) per carrier (qcn, qcp, clamped at -60)."""
    qd = np.ones(dev.N)
    qcn, qcp = np.zeros(dev.N), np.zeros(dev.N)
    Nx, Ny, Nz = dev.Nx, dev.Ny, dev.Nz
    coords = (dev.xs, dev.ys, dev.zs)
    for ax in range(3):
        n, c = (Nx, Ny, Nz)[ax], coords[ax]
        st = (1, Nx, Nx * Ny)[ax]
        others = [q for q in range(3) if q != ax]
        for b in range((Nx, Ny, Nz)[others[1]]):
            for a in range((Nx, Ny, Nz)[others[0]]):
                idx = [0, 0, 0]
                idx[others[0]], idx[others[1]] = a, b
                k0 = idx[0] + Nx * (idx[1] + Ny * idx[2])
                for side in (0, 1):
                    w, have = 0.0, False
                    for t in range(n):
                        q = n - 1 - t if side else t
                        k = k0 + q * st
                        if dev.kind[k] != K_SEMI:
                            have = False
                            continue
                        qp = q + 1 if side else q - 1
                        if qp < 0 or qp >= n:
                            g, have = dev.gate_mask[k], False
                            if g:
                                f = dev.contacts[g - 1]["face"]
                                if f // 2 == ax and f % 2 == side:
                                    w, have = c[q], True
                        else:
                            kp = k0 + qp * st
                            if dev.kind[kp] == K_INS:
                                w, have = 0.5 * (c[q] + c[qp]), True
                            elif dev.kind[kp] != K_SEMI:
                                have = False
                        if not have:
                            continue
                        lo = 0.5 * (c[q - 1] + c[q]) if q > 0 else c[q]
                        hi = 0.5 * (c[q] + c[q + 1]) if q < n - 1 else c[q]
                        za, zb = (w - hi, w - lo) if side else (lo - w, hi - w)
                        za = max(za, 0.0)
                        qd[k] = min(qd[k], abs(c[q] - w))
                        if mlda and za < 40e-9:
                            qcn[k] += math.log(max(mlda_factor(dev.m, za, zb, 0), 1e-30))
                            qcp[k] += math.log(max(mlda_factor(dev.m, za, zb, 1), 1e-30))
    return qd, np.maximum(qcn, -60.0), np.maximum(qcp, -60.0)


def semi_gradient(dev, psi):
    """Potential gradient per axis at every node, inside the semiconductor:
    central over two semiconductor neighbours, one-sided at an interface, 0
    with none (grad_semi, semiconductor_3d.c:2909-2917); shape (3, N)."""
    N = dev.N
    out = np.zeros((3, N))
    semi = dev.kind == K_SEMI
    for ax in range(3):
        k, kp, h, _ = dev.edges(ax)
        hm, hp, pm, pp = np.ones(N), np.ones(N), np.zeros(N), np.zeros(N)
        has_m, has_p = np.zeros(N, bool), np.zeros(N, bool)
        sp = semi[kp]                      # node k has its + neighbour kp in the semiconductor
        has_p[k[sp]], hp[k[sp]], pp[k[sp]] = True, h[sp], psi[kp[sp]]
        sm = semi[k]                       # node kp has its - neighbour k in the semiconductor
        has_m[kp[sm]], hm[kp[sm]], pm[kp[sm]] = True, h[sm], psi[k[sm]]
        g = np.zeros(N)
        both = has_m & has_p
        g = np.where(both, (pp - pm) / (hm + hp), g)
        g = np.where(has_p & ~has_m, (pp - psi) / hp, g)
        g = np.where(has_m & ~has_p, (psi - pm) / hm, g)
        out[ax] = g
    return out


def lombardi(m, mub, Fv, Nt, Dw, hole):
    """Lombardi surface mobility combined with the bulk mobility mub (m^2/Vs)
    by Matthiessen's rule, weighted by the interface proximity Dw; Fv normal
    field (V/m), Nt total doping (m^-3); lombardi semiconductor_3d.c:2924-2934
    (no Coulomb term: silicon)."""
    F = np.maximum(Fv * 1e-2, 1.0)
    N = np.maximum(Nt * 1e-6, 1.0)
    B, C, L, D, E = (m[key][hole] for key in ("lbB", "lbC", "lbL", "lbD", "lbE"))
    muac = B / F + C * N ** L / (np.cbrt(F) * (m["T"] / 300.0))
    musr = 1.0 / (F * F / D + F ** 3 / E) if E > 0 else D / (F * F)
    return 1.0 / (1.0 / mub + Dw * 1e4 * (1.0 / muac + 1.0 / musr))


def face_mobility(dev, psi, qn, qp):
    """Edge mobilities fmu[6][N] (face k -> k+e_ax stored at k), upd_mob_ref
    semiconductor_3d.c:2962-3006 (Lombardi surface mobility with models bit
    16, no metal contacts)."""
    fmu = np.zeros((6, dev.N))
    m = dev.m
    if dev.models & 16:
        qd = wall_distance(dev)
        gr = semi_gradient(dev, psi)
    for ax in range(3):
        k, kp, h, ar = dev.edges(ax)
        ok = (dev.kind[k] == K_SEMI) & (dev.kind[kp] == K_SEMI)
        mn = 0.5 * (dev.mu0n[k] + dev.mu0n[kp])
        mp = 0.5 * (dev.mu0p[k] + dev.mu0p[kp])
        if dev.models & 16:
            Dw = np.exp(-0.5 * (qd[k] + qd[kp]) / 1e-8)
            fn2 = gr[(ax + 1) % 3] ** 2 + gr[(ax + 2) % 3] ** 2
            F = np.sqrt(0.5 * (fn2[k] + fn2[kp]))
            Nt = 0.5 * (dev.Nd[k] + dev.Na[k] + dev.Nd[kp] + dev.Na[kp])
            near = Dw > 1e-8
            mn = np.where(near, lombardi(m, mn, F, Nt, Dw, 0), mn)
            mp = np.where(near, lombardi(m, mp, F, Nt, Dw, 1), mp)
        if dev.models & 4:
            dps = psi[kp] - psi[k]
            sn = mob_hf(mn, heat_field(qn[kp] - qn[k], dps, h), m["vsat"][0], m["beta"][0])
            sp = mob_hf(mp, heat_field(qp[kp] - qp[k], dps, h), m["vsat"][1], m["beta"][1])
            mn = mn + 1.0 * (sn - mn)
            mp = mp + 1.0 * (sp - mp)
        fmu[ax, k] = np.where(ok, mn, 0.0)
        fmu[3 + ax, k] = np.where(ok, mp, 0.0)
    return fmu


def cont_assemble(dev, hole, psi, n, p, fmu, gii=None):
    """Scaled Scharfetter-Gummel continuity rows, cont_assemble
    semiconductor_3d.c:3067-3106.  Returns A (7,N), rhs, sc, dsc."""
    N = dev.N
    VT = dev.m["VT"]
    car = p if hole else n
    semi = dev.kind == K_SEMI
    fix = dev.con_mask != 0
    eq = semi & ~fix
    lnk = dev.lnvK if hole else dev.lncK
    sk = np.maximum(car, NFLOOR)
    dg = np.zeros(N)
    b = np.zeros(N)
    off = np.zeros((6, N))
    for ax in range(3):
        k, kp, h, ar = dev.edges(ax)
        both = semi[k] & semi[kp]
        g = ar * fmu[(3 if hole else 0) + ax, k] * VT / h
        # row k, neighbour kp (direction + along ax)
        dl = lnk[kp] - lnk[k]
        D = (psi[kp] - psi[k]) / VT + (-dl if hole else dl)
        Bp, Bm = bernoulli(D), bernoulli(-D)
        self_k = g * (Bp if hole else Bm)
        nb_k = g * (Bm if hole else Bp)
        # row kp, neighbour k (direction - along ax): D' = -D
        self_kp = g * (Bm if hole else Bp)
        nb_kp = g * (Bp if hole else Bm)
        rk = both & eq[k]
        rkp = both & eq[kp]
        np.add.at(dg, k[rk], self_k[rk])
        np.add.at(dg, kp[rkp], self_kp[rkp])
        fk = rk & fix[kp]
        np.add.at(b, k[fk], (nb_k * car[kp])[fk])
        ok = rk & ~fix[kp]
        off[1 + 2 * ax, k[ok]] = -(nb_k * np.maximum(car[kp], NFLOOR))[ok]
        fkp = rkp & fix[k]
        np.add.at(b, kp[fkp], (nb_kp * car[k])[fkp])
        okp = rkp & ~fix[k]
        off[2 * ax, kp[okp]] = -(nb_kp * np.maximum(car[k], NFLOOR))[okp]
    R, dRn, dRp = recomb(dev.m, dev.niK, dev.tauf, n, p)
    dR = np.maximum(dRp if hole else dRn, 0.0)
    vol = dev.volume()
    dg = dg + vol * dR
    b = b + vol * (dR * car - R)
    if gii is not None:
        b = b + gii
    w = 1.0 / (dg * sk)
    A = np.zeros((7, N))
    A[0] = 1.0
    with np.errstate(invalid="ignore", over="ignore", divide="ignore"):
        A[1:] = np.where(eq, off * w, 0.0)
        rhs = np.where(eq, b * w, np.where(semi, 1.0, 0.0))
    sc = np.where(eq, sk, np.where(semi, car, 0.0))
    dsc = np.where(eq, dg * sk, 0.0)
    return A, rhs, sc, dsc


def matvec(A, dims, x, b=None):
    """y = A x (or b - A x), 7-point stencil, matvec_L semiconductor_3d.c:1975-1993."""
    Nx, Ny, Nz = dims
    c, xm, xp, ym, yp, zm, zp = (a.reshape(Nz, Ny, Nx) for a in A)
    X = x.reshape(Nz, Ny, Nx)
    y = c * X
    y[:, :, 1:] += xm[:, :, 1:] * X[:, :, :-1]
    y[:, :, :-1] += xp[:, :, :-1] * X[:, :, 1:]
    y[:, 1:, :] += ym[:, 1:, :] * X[:, :-1, :]
    y[:, :-1, :] += yp[:, :-1, :] * X[:, 1:, :]
    y[1:, :, :] += zm[1:, :, :] * X[:-1, :, :]
    y[:-1, :, :] += zp[:-1, :, :] * X[1:, :, :]
    y = y.ravel()
    return y if b is None else b - y


def stencil_csr(A, dims):
    """CSR (rowptr, col, val) of a 7-point stencil operator given as its seven
    diagonals (c, xm, xp, ym, yp, zm, zp in the reference's Sten layout,
    semiconductor_3d.c:210): every stored entry that is nonzero, columns
    ascending within a row.  The reference keeps no CSR of tensor-mesh
    operators; its matrix (the operator matvec_L :1975-1993 applies) holds
    exactly these entries."""
    Nx, Ny, Nz = dims
    N = Nx * Ny * Nz
    k = np.arange(N)
    i, j, l = k % Nx, (k // Nx) % Ny, k // (Nx * Ny)
    c, xm, xp, ym, yp, zm, zp = (np.asarray(a, np.float64).reshape(N) for a in A)
    slots = [(l > 0, k - Nx * Ny, zm), (j > 0, k - Nx, ym), (i > 0, k - 1, xm), (np.ones(N, bool), k, c),
             (i < Nx - 1, k + 1, xp), (j < Ny - 1, k + Nx, yp), (l < Nz - 1, k + Nx * Ny, zp)]
    present = np.stack([ok & (v != 0.0) for ok, _, v in slots], axis=1)          # N x 7, column order
    cols = np.stack([cc for _, cc, _ in slots], axis=1)
    vals = np.stack([v for _, _, v in slots], axis=1)
    rowptr = np.concatenate([[0], np.cumsum(present.sum(axis=1))]).astype(np.int64)
    return rowptr, cols[present].astype(np.int64), vals[present]


def mg_hierarchy(dims):
    """Level shapes: 2x2x2 aggregation of every axis longer than 2 until
    N <= 64, mg_alloc semiconductor_3d.c:1894-1921.  Returns [(Nx,Ny,Nz,cx,cy,cz)]."""
    levs = []
    Nx, Ny, Nz = dims
    while True:
        N = Nx * Ny * Nz
        cx, cy, cz = int(Nx > 2), int(Ny > 2), int(Nz > 2)
        if N <= 64 or not (cx or cy or cz) or len(levs) >= 23:
            levs.append((Nx, Ny, Nz, 0, 0, 0))
            return levs
        levs.append((Nx, Ny, Nz, cx, cy, cz))
        Nx, Ny, Nz = (Nx + 1) // 2 if cx else Nx, (Ny + 1) // 2 if cy else Ny, (Nz + 1) // 2 if cz else Nz


def parent_map(lev):
    """coarse index of every fine node (AGG_RANGE, semiconductor_3d.c:1995-1998)."""
    Nx, Ny, Nz, cx, cy, cz = lev
    cNx, cNy = ((Nx + 1) // 2 if cx else Nx), ((Ny + 1) // 2 if cy else Ny)
    l, j, i = np.meshgrid(np.arange(Nz), np.arange(Ny), np.arange(Nx), indexing="ij")
    I, J, L = (i >> 1 if cx else i), (j >> 1 if cy else j), (l >> 1 if cz else l)
    return (I + cNx * (J + cNy * L)).ravel()


def mg_weights(levs, ph0):
    """Slotboom prolongation weights per level (mg_weights :2099-2117);
    ph0: level-0 reference potential / VT, -1e300 where none.  Returns pw
    per level (all but the coarsest)."""
    pws = []
    ph = np.asarray(ph0, dtype=np.float64)
    for L in range(len(levs) - 1):
        par = parent_map(levs[L])
        nc = int(par.max()) + 1
        mx = np.full(nc, -1e300)
        np.maximum.at(mx, par, ph)
        e = np.where(ph <= -0.5e300, -800.0, ph - mx[par])
        pws.append(np.where(e < -700.0, 0.0, np.exp(np.maximum(e, -745.0))))
        ph = mx
    return pws


def galerkin(lev, A, pw, w=None):
    """Coarse operator P^T W A P of a 7-point stencil (galerkin :2003-2027)."""
    Nx, Ny, Nz, cx, cy, cz = lev
    N = Nx * Ny * Nz
    par = parent_map(lev)
    nc = int(par.max()) + 1
    cNx = (Nx + 1) // 2 if cx else Nx
    cNy = (Ny + 1) // 2 if cy else Ny
    w = np.ones(N) if w is None else w
    C = np.zeros((7, nc))
    k = np.arange(N)
    strides = (1, Nx, Nx * Ny)
    cstr = (1, cNx, cNx * cNy)
    i, j, l = k % Nx, (k // Nx) % Ny, k // (Nx * Ny)
    idx = (i, j, l)
    dims = (Nx, Ny, Nz)
    np.add.at(C[0], par, w * A[0] * pw)
    for ax in range(3):
        for side, q in ((-1, 1 + 2 * ax), (1, 2 + 2 * ax)):
            ok = (idx[ax] > 0) if side < 0 else (idx[ax] < dims[ax] - 1)
            kk = k[ok]
            kb = kk + side * strides[ax]
            t = w[kk] * A[q, kk] * pw[kb]
            inside = par[kb] == par[kk]
            np.add.at(C[0], par[kk][inside], t[inside])
            np.add.at(C[q], par[kk][~inside], t[~inside])
    zero = np.all(C == 0.0, axis=0)
    C[0, zero] = 1.0
    return C, (cNx, cNy, ((Nz + 1) // 2 if cz else Nz))


def restrict(lev, r, w=None):
    """r_c = sum over the aggregate of w r (restrict_sum :2028-2037)."""
    par = parent_map(lev)
    out = np.zeros(int(par.max()) + 1)
    np.add.at(out, par, r if w is None else w * r)
    return out


def prolong_add(lev, pw, xc, x):
    """x += pw * x_parent (prolong_add :2038-2048)."""
    return x + pw * xc[parent_map(lev)]


def face_flux(dev, ka, kb, h, ar, psi, n, p, fmu, ax):
    """SG edge current (A) split in electrons and holes, and the magnitude
    scales, face_flux semiconductor_3d.c:2843-2854."""
    VT = dev.m["VT"]
    dps = (psi[kb] - psi[ka]) / VT
    Dn = dps + (dev.lncK[kb] - dev.lncK[ka])
    Dp = dps - (dev.lnvK[kb] - dev.lnvK[ka])
    kl = np.minimum(ka, kb)
    gn = Q * ar * fmu[ax, kl] * VT / h
    gp = Q * ar * fmu[3 + ax, kl] * VT / h
    a1, a2 = gn * n[kb] * bernoulli(Dn), gn * n[ka] * bernoulli(-Dn)
    b1, b2 = gp * p[ka] * bernoulli(Dp), gp * p[kb] * bernoulli(-Dp)
    return a1 - a2, b1 - b2, a1 + a2, b1 + b2


def ii_alpha(m, F, c):
    """Ionisation coefficient (1/m) of carrier c (0 electrons, 1 holes) at
    field F (V/m), Chynoweth form, ii_alpha semiconductor_3d.c:4478-4485."""
    F = np.asarray(F, dtype=np.float64)
    hi = F >= m["ii_E0"][c]
    a = np.where(hi, m["ii_a"][c][1], m["ii_a"][c][0])
    b = np.where(hi, m["ii_b"][c][1], m["ii_b"][c][0])
    with np.errstate(divide="ignore", over="ignore"):
        x = m["ii_g"] * b / np.where(F > 0, F, 1.0)
        if m["ii_m"][c] != 1.0:
            x = x ** m["ii_m"][c]
        out = m["ii_g"] * a * np.exp(-np.minimum(x, 800.0))
    return np.where((F > 1e5) & (x <= 700.0), out, 0.0)


def node_field_current(dev, psi, n, p, fmu):
    """Node current-density vectors of both carriers (mean of the node's
    semiconductor faces per axis, A/m^2) and the field E = -grad psi
    (central difference over semiconductor neighbours, one-sided at their
    ends), as ii_update semiconductor_3d.c:4500-4512 with grad_semi :2909-2917."""
    N = dev.N
    Jn, Jp, E = np.zeros((3, N)), np.zeros((3, N)), np.zeros((3, N))
    semi = dev.kind == K_SEMI
    for ax in range(3):
        k, kp, h, ar = dev.edges(ax)
        s = semi[k] & semi[kp]
        k, kp, h, ar = k[s], kp[s], h[s], ar[s]
        a, b, _, _ = face_flux(dev, k, kp, h, ar, psi, n, p, fmu, ax)
        sn, sp, cnt = np.zeros(N), np.zeros(N), np.zeros(N)
        for node in (k, kp):                 # the face flux k -> kp seen from both ends
            np.add.at(sn, node, a / ar)
            np.add.at(sp, node, b / ar)
            np.add.at(cnt, node, 1.0)
        Jn[ax] = np.where(cnt > 0, sn / np.maximum(cnt, 1.0), 0.0)
        Jp[ax] = np.where(cnt > 0, sp / np.maximum(cnt, 1.0), 0.0)
        # grad_semi: neighbours on both sides -> central, else one-sided
        hm, hp = np.zeros(N), np.zeros(N)
        pm, pp = np.zeros(N), np.zeros(N)
        has_m, has_p = np.zeros(N, bool), np.zeros(N, bool)
        has_p[k], hp[k], pp[k] = True, h, psi[kp]
        has_m[kp], hm[kp], pm[kp] = True, h, psi[k]
        both = has_m & has_p
        g = np.where(both, (pp - pm) / np.where(both, hm + hp, 1.0), 0.0)
        g = np.where(has_p & ~has_m, (pp - psi) / np.where(has_p, hp, 1.0), g)
        g = np.where(has_m & ~has_p, (psi - pm) / np.where(has_m, hm, 1.0), g)
        E[ax] = -g
    return Jn, Jp, E


def ii_update(dev, psi, n, p, fmu):
    """Impact-ionisation pair generation per node (1/s over the control
    volume): G = (alpha_n(F_n)|J_n| + alpha_p(F_p)|J_p|)/q * vol with
    F = |E.J|/|J|, on semiconductor nodes that are not contacts,
    ii_update semiconductor_3d.c:4486-4523."""
    Jn, Jp, E = node_field_current(dev, psi, n, p, fmu)
    G = np.zeros(dev.N)
    for c, J in enumerate((Jn, Jp)):
        jj = np.sqrt((J * J).sum(axis=0))
        F = np.abs((E * J).sum(axis=0)) / np.where(jj > 0, jj, 1.0)
        G += np.where(jj > 0, ii_alpha(dev.m, F, c) * jj, 0.0)
    on = (dev.kind == K_SEMI) & (dev.con_mask == 0)
    return np.where(on, G / Q * dev.volume(), 0.0)


def terminal_currents(dev, psi, n, p, fmu):
    """Flux out of each contact's nodes into semiconductor nodes of other
    contacts or none (A), its scale and electron part, terminal_currents
    semiconductor_3d.c:3598-3623."""
    nc = len(dev.contacts)
    I, S, In = np.zeros(nc), np.zeros(nc), np.zeros(nc)
    for ax in range(3):
        k, kp, h, ar = dev.edges(ax)
        for a, b in ((k, kp), (kp, k)):
            c = dev.con_mask[a]
            sel = (c > 0) & (dev.kind[a] == K_SEMI) & (dev.kind[b] == K_SEMI) & (dev.con_mask[b] != c)
            Jn, Jp, Sn, Sp = face_flux(dev, a[sel], b[sel], h[sel], ar[sel], psi, n, p, fmu, ax)
            np.add.at(I, c[sel] - 1, Jn + Jp)
            np.add.at(S, c[sel] - 1, Sn + Sp)
            np.add.at(In, c[sel] - 1, Jn)
    return I, S, In


def cont_update(dev, hole, psi, n, p, phin, phip, u, sc, lo, hi):
    """Quasi-Fermi update from u = c/s and the weighted max change,
    cont_update semiconductor_3d.c:3112-3146.  Returns (q, n, p, dq, dmax)."""
    VT = dev.m["VT"]
    eq = (dev.kind == K_SEMI) & (dev.con_mask == 0)
    uu = np.where(~(u > 1e-30), 1e-30, np.minimum(u, 1e30))
    c = sc * uu
    with np.errstate(divide="ignore", invalid="ignore"):
        if hole:
            q = psi + VT * (np.log(c) - dev.lnvK)
        else:
            q = psi - VT * (np.log(c) - dev.lncK)
    q = np.clip(q, lo, hi)
    qf = phip if hole else phin
    dq = np.where(eq, q - qf, 0.0)
    qnew = np.where(eq, q, qf)
    nn, pp = dev.carriers(psi, phin if hole else qnew, qnew if hole else phip)
    n2 = np.where(eq & (not hole), nn, n) if not hole else n
    p2 = np.where(eq, pp, p) if hole else p
    cn = p2 if hole else n2
    scale = n2 + p2 + np.abs(dev.Nd - dev.Na)
    with np.errstate(divide="ignore", invalid="ignore"):
        r = np.where(scale > 0, cn / scale, 0.0)
    w = np.where(r >= 1e-3, 1.0, r * 1e3)
    a = np.where(eq & (r > 1e-6), w * np.abs(dq), 0.0)
    return qnew, n2, p2, dq, float(a.max(initial=0.0))


# ------------------------------------------------ terminal charges (C-V)
def _nbr(dev, i, j, l, d):
    """Neighbour of node (i, j, l) in direction d (0..5: -x +x -y +y -z +z):
    (kb, h, area) or None at the boundary, nbr() semiconductor_3d.c:703-713."""
    Nx, Ny, Nz = dev.Nx, dev.Ny, dev.Nz
    k = i + Nx * (j + Ny * l)
    if d == 0:
        return None if i == 0 else (k - 1, dev.xs[i] - dev.xs[i - 1], dev.Ly[j] * dev.Lz[l])
    if d == 1:
        return None if i == Nx - 1 else (k + 1, dev.xs[i + 1] - dev.xs[i], dev.Ly[j] * dev.Lz[l])
    if d == 2:
        return None if j == 0 else (k - Nx, dev.ys[j] - dev.ys[j - 1], dev.Lx[i] * dev.Lz[l])
    if d == 3:
        return None if j == Ny - 1 else (k + Nx, dev.ys[j + 1] - dev.ys[j], dev.Lx[i] * dev.Lz[l])
    if d == 4:
        return None if l == 0 else (k - Nx * Ny, dev.zs[l] - dev.zs[l - 1], dev.Lx[i] * dev.Ly[j])
    return None if l == Nz - 1 else (k + Nx * Ny, dev.zs[l + 1] - dev.zs[l], dev.Lx[i] * dev.Ly[j])


def charge_partition(dev, n, p, cterm):
    """Owner terminal of every node (-1 none) and the number of meeting
    edges, api3_charge_partition semiconductor_3d.c:5192-5245: fronts from
    the nodes of ohmic contacts spread breadth first, all at once, through
    semiconductor nodes of the same majority carrier (n >= p) and through
    metal; then semiconductor no front reached joins its neighbour's."""
    from collections import deque
    N, Nx, Ny = dev.N, dev.Nx, dev.Ny
    own = np.full(N, -1, np.int64)
    qtype = np.where(n >= p, 1, -1)
    qu = deque()
    for k in range(N):
        c = dev.con_mask[k] - 1
        if c < 0 or c >= len(dev.contacts):
            continue
        t = cterm[c]
        own[k] = t
        if t >= 0 and dev.contacts[c]["bc"] == 0 and dev.kind[k] in (K_SEMI, K_METAL):
            qu.append(k)
    meet = 0

    def ijl(k):
        l, r = divmod(k, Nx * Ny)
        j, i = divmod(r, Nx)
        return i, j, l

    while qu:
        k = qu.popleft()
        t = own[k]
        i, j, l = ijl(k)
        for d in range(6):
            nb = _nbr(dev, i, j, l, d)
            if nb is None:
                continue
            b = nb[0]
            if dev.con_mask[b]:
                continue
            if dev.kind[k] == K_METAL:
                ok = dev.kind[b] == K_METAL
            else:
                ok = dev.kind[k] == K_SEMI and dev.kind[b] == K_SEMI and qtype[b] == qtype[k]
            if not ok:
                continue
            if own[b] < 0:
                own[b] = t
                qu.append(b)
            elif own[b] != t:
                meet += 1
    qu = deque(k for k in range(N) if own[k] >= 0 and dev.kind[k] == K_SEMI)
    while qu:
        k = qu.popleft()
        t = own[k]
        i, j, l = ijl(k)
        for d in range(6):
            nb = _nbr(dev, i, j, l, d)
            if nb is None:
                continue
            b = nb[0]
            if dev.con_mask[b] or dev.kind[b] != K_SEMI or own[b] >= 0:
                continue
            own[b] = t
            qu.append(b)
    return own, meet


def terminal_charges(dev, psi, own, cterm, Va):
    """Charge of every terminal (C): the flux of D out of its node set, plus
    the oxide flux of its face (Robin) gate nodes; the last entry is the
    flux into the nodes of no terminal.  api3_terminal_charges
    semiconductor_3d.c:5248-5277; Va: the contact voltages."""
    nt = int(max(cterm)) + 1
    Q = np.zeros(nt + 1)
    Nx, Ny = dev.Nx, dev.Ny
    for k in range(dev.N):
        t = own[k]
        if t < 0 or t >= nt:
            continue
        l, r = divmod(k, Nx * Ny)
        j, i = divmod(r, Nx)
        for d in range(6):
            nb = _nbr(dev, i, j, l, d)
            if nb is None:
                continue
            b, h, ar = nb
            tb = own[b]
            if tb == t:
                continue
            ek, eb = dev.eps[k], dev.eps[b]
            f = (2.0 * ek * eb / (ek + eb)) * ar / h * (psi[k] - psi[b])
            Q[t] += f
            if tb < 0 or tb >= nt:
                Q[nt] += f
    for k in np.nonzero(dev.robA > 0)[0]:
        c = dev.gate_mask[k] - 1
        if c < 0:
            continue
        cn = dev.contacts[c]
        Cox = EPS_OX / (cn["tox"] if cn["tox"] > 0 else 10e-9)
        pm = cn["pm"] if cn["pm"] > 0 else dev.Cref
        f = dev.robA[k] * Cox * (psi[k] - (Va[c] + dev.Cref - pm))
        t, tg = own[k], cterm[c]
        if 0 <= t < nt:
            Q[t] += f
        if 0 <= tg < nt:
            Q[tg] -= f
    return Q


# ------------------------------------------------------------- readback
def readback(dev, which, psi, phin, phip, fmu=None, gii=None):
    """Node quantity `which` of api3_bulk_copy at a state (node_val,
    semiconductor_3d.c:5066-5123) for the quantities that follow from the
    state, the device and (20-25) the face mobilities: 0 psi, 1 n, 2 p,
    3 Ec, 4 Ev, 5 Efn, 6 Efp (eV, vacuum level -psi), 13 kind, 14 net doping,
    15 phi_n, 16 phi_p, 17 fixed charge, 18 contact id, 19 band-gap
    narrowing (eV), 20-25 node current densities (Jn x y z, Jp x y z),
    26 effective n_i, 27 gate tunnelling (0), 28-29 quantum potentials (eV),
    30 interface distance, 31-32 traps (0), 33 impact ionisation per volume,
    34 interface charge (0), 35 ionised net doping.  No traps, tunnelling
    or incomplete ionisation (outside the scope)."""
    m, VT = dev.m, dev.m["VT"]
    semi, metal, ins = dev.kind == K_SEMI, dev.kind == K_METAL, dev.kind == K_INS
    n, p = dev.carriers(psi, phin, phip)
    nan = np.full(dev.N, np.nan)
    if which == 0:
        return psi.copy()
    if which in (1, 2):
        return n if which == 1 else p
    if which == 3:
        return np.where(metal, nan, -psi - m["chi"] - dev.bgn * VT)
    if which == 4:
        return np.where(metal, nan, -psi - m["chi"] - m["Eg"] + dev.bgn * VT)
    if which in (5, 6):
        return np.where(ins, nan, -(phin if which == 5 else phip) - dev.Cref)
    if which == 13:
        return dev.kind.astype(np.float64)
    if which == 14:
        return dev.Nd - dev.Na
    if which in (15, 16):
        return (phin if which == 15 else phip).copy()
    if which == 17:
        return dev.Nf.copy()
    if which == 18:
        return dev.con_mask.astype(np.float64)
    if which == 19:
        return dev.bgn * VT * 2.0
    if 20 <= which <= 25:
        Jn, Jp, _ = node_field_current(dev, psi, n, p, fmu)
        J = Jn if which < 23 else Jp
        return np.where(semi, J[(which - 20) % 3], 0.0)
    if which == 26:
        return np.where(semi, dev.niK, 0.0)
    if which == 27:
        return np.zeros(dev.N)
    if which in (28, 29):
        return np.where(semi, -(dev.qcn if which == 28 else dev.qcp) * VT, nan)
    if which == 30:
        return np.where(semi, wall_distance(dev), nan)
    if which in (31, 32, 34):
        return np.where(semi, 0.0, nan)
    if which == 33:
        g = np.zeros(dev.N) if gii is None else gii
        return np.where(semi, g / dev.volume(), 0.0)
    if which == 35:
        return np.where(semi, dev.Nd - dev.Na, 0.0)
    raise ValueError(f"readback {which} is not a state quantity")


# ------------------------------------------------------------ Anderson mixing
class Anderson:
    """Anderson mixing of the Gummel map, aa_step semiconductor_3d.c:3709-3757
    (reset :3695).  x, g: the three blocks (psi, phi_n, phi_p) stacked, 3N.
    The history keeps FP32 differences dF, dG of the residual f = g - x and
    of the image g over the last m steps (a ring), their Gram matrix in FP64;
    the mixing solves the regularised normal equations (diagonal + 1e-9 trace)
    by Gaussian elimination with partial pivoting and moves only the free
    nodes: psi off every contact node, phi_n/phi_p on the nodes where the
    carrier is significant (aa_sig :3689-3694)."""

    def __init__(self, dev, n, p, m=12):
        N = dev.N
        self.m, self.N = m, N
        te = np.zeros(self.N, bool)
        for c, cn in enumerate(dev.contacts):
            if cn.get("te", 0):
                te |= dev.con_mask == c + 1
        semi = dev.kind == K_SEMI
        carfix = (dev.con_mask != 0) & ~(te & semi)                      # car_fix :796-799
        base = semi & ~carfix
        tot = n + p + np.abs(dev.Nd - dev.Na)
        sig_n = base & (n > 1e-6 * tot)
        sig_p = base & (p > 1e-6 * tot)
        # the residual's mask (psi everywhere) and the mixing's mask
        self.fmask = np.concatenate([np.ones(N, bool), sig_n, sig_p])
        self.mix = np.concatenate([dev.con_mask == 0, sig_n & ~carfix, sig_p & ~carfix])
        self.reset()

    def reset(self):
        self.n = self.pos = self.have = 0
        self.dF = [None] * self.m
        self.dG = [None] * self.m
        self.gram = np.zeros((self.m, self.m))
        self.f = self.gp = None

    def step(self, x, g):
        """One step: returns the mixed iterate (the new g)."""
        M = self.m
        f = np.where(self.fmask, g - x, 0.0)
        if M > 0 and self.have:
            s = self.pos
            self.dF[s] = (f - self.f).astype(np.float32)
            self.dG[s] = (g - self.gp).astype(np.float32)
            if self.n < M:
                self.n += 1
            dFs = self.dF[s].astype(np.float64)
            for i in range(self.n):
                v = float(np.dot(self.dF[i].astype(np.float64), dFs))
                self.gram[i, s] = self.gram[s, i] = v
            self.pos = (s + 1) % M
        self.f, self.gp = f, g.copy()
        self.have = 1
        m = self.n
        if m == 0:
            return g.copy()
        A = self.gram[:m, :m].copy()
        b = np.array([np.dot(self.dF[i].astype(np.float64), f) for i in range(m)])
        A[np.diag_indices(m)] += 1e-9 * np.trace(A) + 1e-300
        for c in range(m):                                   # elimination, partial pivoting
            pv = c + int(np.argmax(np.abs(A[c:, c])))
            if abs(A[pv, c]) <= abs(A[c, c]):
                pv = c
            if pv != c:
                A[[c, pv]] = A[[pv, c]]
                b[[c, pv]] = b[[pv, c]]
            for r in range(c + 1, m):
                fct = A[r, c] / A[c, c]
                A[r, c:] -= fct * A[c, c:]
                b[r] -= fct * b[c]
        gam = np.zeros(m)
        for r in range(m - 1, -1, -1):
            gam[r] = (b[r] - np.dot(A[r, r + 1:], gam[r + 1:])) / A[r, r]
        if np.any(np.isnan(gam)):
            return g.copy()
        out = g.copy()
        corr = np.zeros_like(g)
        for i in range(m):
            corr += gam[i] * self.dG[i].astype(np.float64)
        out[self.mix] -= corr[self.mix]
        return out


# ------------------------------------------------------------ gate tunnelling
HPLANCK = 6.62607015e-34
TUN_COL, TUN_NE, TUN_EMAX = 5e-9, 64, 3.0
# insulators of a gate stack (mat_init semiconductor_3d.c:599-614): eps_r,
# affinity (eV), gap (eV), tunnelling masses of electrons/holes (m0)
INSULATORS = {8: dict(eps_r=3.9, chi=0.95, Eg=9.0, mt_e=0.42, mt_h=0.33),       # SiO2
              9: dict(eps_r=9.0, chi=1.35, Eg=6.8, mt_e=0.35, mt_h=0.35),       # Al2O3
              10: dict(eps_r=22.0, chi=2.55, Eg=5.8, mt_e=0.18, mt_h=0.20),     # HfO2
              11: dict(eps_r=7.5, chi=2.05, Eg=5.1, mt_e=0.50, mt_h=0.50)}      # Si3N4
SI_SUPPLY_MASS = (0.50, 0.50)                                                    # ms_e, ms_h of Si (:458)


def tun_paths(dev, stacks):
    """Paths of the Robin gates with a layer stack (tun_build :1238-1244,
    tun_push :1172-1210): one per gate node in the order of the gate nodes,
    each with its supply column, the semiconductor nodes inward from the
    interface over TUN_COL (stopping at contacts and at the domain edge).
    stacks: {contact index: [(insulator id, thickness nm), ...]}."""
    P = dict(pk=[], pc=[], pA=[], pc0=[], pcn=[], ck=[], cl=[], ch=[], layers=[])
    for c, cn in enumerate(dev.contacts):
        if cn["bc"] != 2 or cn["face"] > 5 or not stacks.get(c):
            continue
        lay = [(INSULATORS[m], t * 1e-9) for m, t in stacks[c]]
        face = cn["face"]
        inward, ax = face ^ 1, face // 2
        for k, area in dev.contact_nodes(cn):
            if dev.gate_mask[k] != c + 1 or dev.con_mask[k]:
                continue
            P["pk"].append(k); P["pc"].append(c); P["pA"].append(area); P["layers"].append(lay)
            P["pc0"].append(len(P["ck"]))
            cnt, cur, depth = 0, k, 0.0
            while True:
                l, r = divmod(cur, dev.Nx * dev.Ny)
                j, i = divmod(r, dev.Nx)
                L = (dev.Lx[i], dev.Ly[j], dev.Lz[l])[ax]
                nb = _nbr(dev, i, j, l, inward)
                more = nb is not None and dev.kind[nb[0]] == K_SEMI and not dev.con_mask[nb[0]]
                P["ck"].append(cur); P["cl"].append(L); P["ch"].append(nb[1] if more else 0.0)
                cnt += 1
                depth += L
                if not more or depth >= TUN_COL:
                    break
                cur = nb[0]
            P["pcn"].append(cnt)
    for key in ("pk", "pc", "pc0", "pcn", "ck"):
        P[key] = np.asarray(P[key], np.int64)
    for key in ("pA", "cl", "ch"):
        P[key] = np.asarray(P[key], np.float64)
    return P


def _softplus(x):
    return x + math.log1p(math.exp(-x)) if x > 35.0 else math.log1p(math.exp(x))


def tun_wkb(m, L, u0, u1):
    """2 Int kappa dx over a layer of length L whose barrier above the carrier
    energy runs linearly from u0 to u1 eV (wkb_seg :3262-3270)."""
    a, b = max(u0, 0.0), max(u1, 0.0)
    if a == 0.0 and b == 0.0:
        return 0.0
    c = 2.0 * math.sqrt(2.0 * m * M0 * Q) / HBAR
    du = u1 - u0
    if abs(du) < 1e-9 * (a + b):
        I = L * math.sqrt(0.5 * (a + b))
    else:
        I = L * (2.0 / 3.0) * (b * math.sqrt(b) - a * math.sqrt(a)) / du
    return c * I


def tun_stack(layers, psik, psim):
    """Potentials along a stack from the node (psik) to the metal (psim):
    D = (psim - psik) / sum t/eps, psi linear in every layer; a first layer
    that is not an insulator (the semiconductor half cell) puts the interface
    at its far end (tun_stack :3300-3312).  Returns (D, psis, [(mat, t, p0, dp)])."""
    R = sum(t / (m["eps_r"] * EPS0) for m, t in layers)
    D = (psim - psik) / R if R > 0 else 0.0
    out, ps, psis = [], psik, psik
    for i, (m, t) in enumerate(layers):
        dp = D * t / (m["eps_r"] * EPS0)
        out.append((m, t, ps, dp))
        ps += dp
        if i == 0 and not m.get("insulator", True):
            psis = ps
    return D, psis, out


def tun_T(stack, E, hole):
    """WKB transmission at energy E (eV) (tun_T :3275-3288)."""
    s = 0.0
    for m, t, p0, dp in stack:
        if not m.get("insulator", True):
            continue
        mt = m["mt_h"] if hole else m["mt_e"]
        ea, eb = -p0 - m["chi"], -(p0 + dp) - m["chi"]
        if hole:
            u0, u1 = E - (ea - m["Eg"]), E - (eb - m["Eg"])
        else:
            u0, u1 = ea - E, eb - E
        s += tun_wkb(mt, t, u0, u1)
        if s > 745.0:
            return 0.0
    return math.exp(-s)


def tun_int(stack, E0, EF, kT, hole):
    """Simpson's rule over 64 intervals of T(E) ln(1 + exp(+-(EF-E)/kT)) from
    the band edge E0, over min(max(+-(EF-E0), 0) + 25 kT, 3 eV) (tun_int :3291-3303)."""
    sg = -1.0 if hole else 1.0
    span = min(max(sg * (EF - E0), 0.0) + 25.0 * kT, TUN_EMAX)
    h = span / TUN_NE
    s = 0.0
    for q in range(TUN_NE + 1):
        E = E0 + sg * q * h
        w = 1.0 if q in (0, TUN_NE) else (4.0 if q & 1 else 2.0)
        s += tun_T(stack, E, hole) * _softplus(sg * (EF - E) / kT) * w
    return s * h / 3.0


def tun_flux(dev, stack, Ecp, Evp, EFn, EFp, EFm):
    """Electrons s->g, g->s, holes s->g, g->s (1/m^2 s) (tun_flux :3316-3322)."""
    kT = dev.m["VT"]
    h3 = HPLANCK ** 3
    ce = 4.0 * math.pi * SI_SUPPLY_MASS[0] * M0 * Q * Q * kT / h3
    ch = 4.0 * math.pi * SI_SUPPLY_MASS[1] * M0 * Q * Q * kT / h3
    return (ce * tun_int(stack, Ecp, EFn, kT, 0), ce * tun_int(stack, Ecp, EFm, kT, 0),
            ch * tun_int(stack, Evp, EFp, kT, 1), ch * tun_int(stack, Evp, EFm, kT, 1))


def tun_eval(m, Cref, layers, psi, phin, phip, Vg, pm):
    """Fluxes through a stack from a semiconductor (material m) whose surface
    is at psi, with local supply at the band edges of the surface
    (api3_tun_eval :4855-4867): (e s->g, e g->s, h s->g, h g->s) in 1/m^2 s
    and the net current density into the semiconductor (A/m^2)."""
    D, psis, stack = tun_stack(layers, psi, Vg + Cref - (pm if pm > 0 else Cref))
    Ec = -psis - m["chi"]
    dev = type("S", (), {"m": m})()
    F = tun_flux(dev, stack, Ec, Ec - m["Eg"], -phin - Cref, -phip - Cref, -Vg - Cref)
    return F, Q * ((F[0] - F[1]) - (F[2] - F[3]))


def tun_update(dev, P, psi, phin, phip, n, p, Va):
    """Per path (tun_path :3326-3366): the effective supply densities from the
    column (contact theorem for carriers pressed against the interface by
    the field, the local density extrapolated to the interface otherwise),
    the four fluxes, and the linearised exchange: cK (escape per unit density
    of each column entry, e/h interleaved) and pG (injection, e/h)."""
    VT = dev.m["VT"]
    cK = np.zeros(2 * len(P["ck"]))
    pG = np.zeros(2 * len(P["pk"]))
    eps_s = dev.m["eps_r"] * EPS0
    for t in range(len(P["pk"])):
        k, c = int(P["pk"][t]), int(P["pc"][t])
        cn = dev.contacts[c]
        pm = cn["pm"] if cn["pm"] > 0 else dev.Cref
        D, psis, stack = tun_stack(P["layers"][t], psi[k], Va[c] + dev.Cref - pm)
        q0, nq = int(P["pc0"][t]), int(P["pcn"][t])
        Fs = D / eps_s
        neff = [0.0, 0.0]
        for hole in (0, 1):
            sg = -1.0 if hole else 1.0
            car = p if hole else n
            if sg * Fs > 0:
                P0, Fl = 0.0, sg * Fs
                for q in range(q0, q0 + nq):
                    kq = P["ck"][q]
                    Fr = sg * (psi[kq] - psi[P["ck"][q + 1]]) / P["ch"][q] if (q < q0 + nq - 1 and P["ch"][q] > 0) else Fl
                    w = max(0.5 * (Fl + Fr), 0.0) * P["cl"][q] / VT
                    cK[2 * q + hole] = w
                    P0 += w * car[kq]
                    Fl = Fr
                kl = P["ck"][q0 + nq - 1]
                cK[2 * (q0 + nq - 1) + hole] += 1.0
                P0 += car[kl]
                neff[hole] = P0
            else:
                qc = (dev.qcp if hole else dev.qcn)[k]
                neff[hole] = car[k] * math.exp(-qc + sg * (psis - psi[k]) / VT)
                cK[2 * q0 + hole] = neff[hole] / max(car[k], NFLOOR)
            if not neff[hole] > 1e-100:
                neff[hole] = 1e-100
        EFn, EFp, EFm = -phin[k] - dev.Cref, -phip[k] - dev.Cref, -Va[c] - dev.Cref
        Ecp = EFn - VT * math.log(neff[0] / dev.m["Nc"])
        Evp = EFp + VT * math.log(neff[1] / dev.m["Nv"])
        F = tun_flux(dev, stack, Ecp, Evp, EFn, EFp, EFm)
        A = P["pA"][t]
        cK[2 * q0:2 * (q0 + nq):2] *= A * F[0] / neff[0]
        cK[2 * q0 + 1:2 * (q0 + nq):2] *= A * F[2] / neff[1]
        pG[2 * t], pG[2 * t + 1] = A * F[1], A * F[3]
    return cK, pG


def tun_gather(dev, P, cK, pG, hole):
    """Per node sink tk (sum of the column entries' escape) and source tg (sum
    of the injection of the paths starting there) (tun_scatter :3373-3380)."""
    tk, tg = np.zeros(dev.N), np.zeros(dev.N)
    np.add.at(tk, P["ck"], cK[hole::2])
    np.add.at(tg, P["pk"], pG[hole::2])
    return tk, tg


def tun_gate_currents(dev, P, cK, pG, n, p):
    """Gate current of every contact (A, into the device): electrons leaving
    to the gate minus their injection, the reverse for holes, with the flux
    sum S and the electron part (terminal_currents :3614-3621)."""
    I, S, In = (np.zeros(len(dev.contacts)) for _ in range(3))
    for t in range(len(P["pk"])):
        q = slice(int(P["pc0"][t]), int(P["pc0"][t] + P["pcn"][t]))
        ks = P["ck"][q]
        se = float(np.sum(cK[0::2][q] * n[ks]))
        sh = float(np.sum(cK[1::2][q] * p[ks]))
        c = int(P["pc"][t])
        Ie, Ih = Q * (se - pG[2 * t]), -Q * (sh - pG[2 * t + 1])
        I[c] += Ie + Ih
        S[c] += Q * (se + pG[2 * t] + sh + pG[2 * t + 1])
        In[c] += Ie
    return I, S, In


# ------------------------------------------------------- graph (prism) meshes
class GraphDevice:
    """Prepared graph device (alloc_dev_graph :981-1040 and prepare
    :1638-1775 on an unstructured mesh) of one semiconductor (material 0):
    N nodes, directed edges in CSR (rp, cj) with length h and dual face area
    ar, control volumes vol, positions pos (N, 3).  Contacts are node lists:
    dicts with bc, V, pm, tox (m), qox (C/m^2), face, and `nodes`, `areas`
    (the boundary area of each node for a face contact).  walls (N, 6): per
    node two interface walls (distance of the node, range a..b of its cell;
    1 = none), api3_set_walls :4838-4841, read by MLDA and Lombardi."""

    def __init__(self, rp, cj, h, ar, vol, pos, Nd, Na, contacts, models=7, tau=None, T=300.0, Nf=None,
                 walls=None):
        self.rp, self.cj = np.asarray(rp, np.int64), np.asarray(cj, np.int64)
        self.h, self.ar, self.vol = (np.asarray(a, np.float64) for a in (h, ar, vol))
        self.pos = np.asarray(pos, np.float64).reshape(-1, 3)
        self.N, self.E = len(self.vol), len(self.cj)
        self.row = np.repeat(np.arange(self.N), np.diff(self.rp))        # source node of every directed edge
        key = self.row * self.N + self.cj                                 # reverse edge and undirected id
        order = np.argsort(key)
        rkey = self.cj * self.N + self.row
        self.rev = order[np.searchsorted(key[order], rkey)]
        self.ue = np.full(self.E, -1, np.int64)
        lo = self.row < self.cj
        self.ue[lo] = np.arange(int(lo.sum()))
        self.ue[~lo] = self.ue[self.rev[~lo]]
        self.nue = int(lo.sum())
        self.m = silicon(T)
        if tau is not None:
            self.m["tau"] = tuple(t if t > 0 else 1e30 for t in tau)
        self.models = models
        self.Nd, self.Na = np.asarray(Nd, np.float64), np.asarray(Na, np.float64)
        self.Nf = np.zeros(self.N) if Nf is None else np.asarray(Nf, np.float64)
        self.contacts = contacts
        m = self.m
        VT = m["VT"]
        self.Cref = m["chi"] + 0.5 * m["Eg"] + 0.5 * VT * math.log(m["Nc"] / m["Nv"])
        self.lnc = math.log(m["Nc"]) + (m["chi"] - self.Cref) / VT
        self.lnv = math.log(m["Nv"]) + (self.Cref - m["chi"] - m["Eg"]) / VT
        self.kind = np.zeros(self.N, np.uint8)
        self.eps = np.full(self.N, m["eps_r"] * EPS0)
        self.bgn = bgn_half(m, self.Nd + self.Na) if (models & 1) else np.zeros(self.N)
        self.con_mask = np.zeros(self.N, np.int32)
        self.gate_mask = np.zeros(self.N, np.int32)
        self.robA = np.zeros(self.N)
        for c, cn in enumerate(contacts):                        # FOR_CON_ANY over the node lists :1656-1668
            for k in cn["nodes"]:
                if cn["bc"] == 2 and cn["face"] <= 5 and self.kind[k] == K_SEMI:
                    self.gate_mask[k] = c + 1
                elif cn["bc"] == 2 or self.kind[k] != K_SEMI:
                    self.kind[k] = K_METAL
                    self.con_mask[k] = c + 1
                    self.gate_mask[k] = 0
                else:
                    self.con_mask[k] = c + 1
                    self.gate_mask[k] = 0
        for c, cn in enumerate(contacts):                        # Robin list :1676-1686
            if cn["bc"] != 2 or cn["face"] > 5:
                continue
            for k, a in zip(cn["nodes"], cn["areas"]):
                if self.gate_mask[k] == c + 1 and not self.con_mask[k]:
                    self.robA[k] = a
        self.gwall = np.ones((self.N, 6)) if walls is None else np.asarray(walls, np.float64).reshape(-1, 6)
        self.qd, self.qcn, self.qcp = graph_walls(self, (models & 8) != 0)
        self.lncK = self.lnc + self.bgn + self.qcn
        self.lnvK = self.lnv + self.bgn + self.qcp
        self.niK = m["ni"] * np.exp(self.bgn + 0.5 * (self.qcn + self.qcp))
        if models & 16:
            self.lsq = graph_lsq(self)
        semi = self.kind == K_SEMI
        self.mu0n = np.where(semi, mob_low(m, self.Nd + self.Na, 0), 0.0)
        self.mu0p = np.where(semi, mob_low(m, self.Nd + self.Na, 1), 0.0)
        if (models & 2) and m["tau_Nref"] > 0:
            self.tauf = 1.0 / (1.0 + (self.Nd + self.Na) / m["tau_Nref"])
        else:
            self.tauf = np.ones(self.N)

    def carriers(self, psi, phin, phip):
        VT = self.m["VT"]
        semi = self.kind == K_SEMI
        n = np.exp(np.clip(self.lncK + (psi - phin) / VT, LN_MIN, LN_MAX))
        p = np.exp(np.clip(self.lnvK + (phip - psi) / VT, LN_MIN, LN_MAX))
        return np.where(semi, n, 0.0), np.where(semi, p, 0.0)


def graph_walls(dev, mlda):
    """Interface distance qd and, with mlda, the MLDA terms qcn, qcp of every
    semiconductor node from its two host-supplied walls, walls_graph
    semiconductor_3d.c:1252-1270 (walls within 40 nm count)."""
    qd, qcn, qcp = np.ones(dev.N), np.zeros(dev.N), np.zeros(dev.N)
    for k in np.nonzero(dev.kind == K_SEMI)[0]:
        for w in range(2):
            g = dev.gwall[k, 3 * w:3 * w + 3]
            if not g[0] < 0.5:
                continue
            qd[k] = min(qd[k], g[0])
            if mlda and g[1] < 40e-9:
                qcn[k] += math.log(max(mlda_factor(dev.m, g[1], g[2], 0), 1e-30))
                qcp[k] += math.log(max(mlda_factor(dev.m, g[1], g[2], 1), 1e-30))
    return qd, np.maximum(qcn, -60.0), np.maximum(qcp, -60.0)


def graph_lsq(dev):
    """Inverse least-squares gradient matrix (N, 3, 3) of every semiconductor
    node over its semiconductor neighbours, weights 1/|dr|^2, 1e-6 of the
    trace added to the diagonal, g_lsq_setup semiconductor_3d.c:2682-2701."""
    r, j = dev.row, dev.cj
    ok = (dev.kind[r] == K_SEMI) & (dev.kind[j] == K_SEMI)
    dr = dev.pos[j] - dev.pos[r]
    w = np.where(ok, 1.0 / (np.sum(dr * dr, axis=1) + 1e-40), 0.0)
    M = np.zeros((dev.N, 3, 3))
    np.add.at(M, r, w[:, None, None] * dr[:, :, None] * dr[:, None, :])
    eps = 1e-6 * np.trace(M, axis1=1, axis2=2) + 1e-30
    M += eps[:, None, None] * np.eye(3)
    out = np.zeros_like(M)
    det = np.linalg.det(M)
    good = np.abs(det) > 0
    out[good] = np.linalg.inv(M[good])
    return out


def graph_gradient(dev, fld):
    """Least-squares node gradient (N, 3) of a node field over the
    semiconductor neighbours, g_grad semiconductor_3d.c:2702-2714."""
    r, j = dev.row, dev.cj
    ok = (dev.kind[r] == K_SEMI) & (dev.kind[j] == K_SEMI)
    dr = dev.pos[j] - dev.pos[r]
    w = np.where(ok, (fld[j] - fld[r]) / (np.sum(dr * dr, axis=1) + 1e-40), 0.0)
    rhs = np.zeros((dev.N, 3))
    np.add.at(rhs, r, w[:, None] * dr)
    return np.einsum("kab,kb->ka", dev.lsq, rhs)


def graph_poisson(dev, psi, phin, phip, Va):
    """Poisson residual F, Jacobian (diagonal dg, one value per directed edge
    va), row scale and ||F/sc|| on a graph, poisson_core_g :4181-4213 with the
    Robin gate terms of poisson_eval :2776-2782."""
    r, j = dev.row, dev.cj
    ek, eb = dev.eps[r], dev.eps[j]
    cf = (2.0 * ek * eb / (ek + eb)) * dev.ar / dev.h
    F = np.zeros(dev.N)
    dg = np.zeros(dev.N)
    np.add.at(F, r, cf * (psi[j] - psi[r]))
    np.add.at(dg, r, cf)
    free = dev.con_mask == 0
    va = np.where(free[j] & free[r], -cf, 0.0)
    VT = dev.m["VT"]
    semi = dev.kind == K_SEMI
    n, p = dev.carriers(psi, phin, phip)
    F += np.where(semi, dev.vol * Q * (p - n + dev.Nd - dev.Na + dev.Nf), dev.vol * Q * dev.Nf)
    dg += np.where(semi, dev.vol * Q * (n + p) / VT, 0.0)
    for c, cn in enumerate(dev.contacts):
        if cn["bc"] != 2 or cn["face"] > 5:
            continue
        rob = (dev.gate_mask == c + 1) & free & (dev.robA > 0)
        Cox = EPS_OX / (cn["tox"] if cn["tox"] > 0 else 10e-9)
        pm = cn["pm"] if cn["pm"] > 0 else dev.Cref
        VG = Va[c] + dev.Cref - pm
        F[rob] += dev.robA[rob] * (Cox * (VG - psi[rob]) + cn.get("qox", 0.0))
        dg[rob] += Cox * dev.robA[rob]
    F[~free] = 0.0
    dg[~free] = 1.0
    return F, dg, va, dg.copy(), float(np.sqrt(np.sum((F / dg) ** 2)))


def graph_mobility(dev, psi, qn, qp):
    """Edge mobilities (electrons, holes) per undirected edge, upd_mob_ref_g
    :4241-4292 (isotropic material, no metal contacts): Lombardi with models
    bit 16 from the node gradients of psi (nodes within 0.5 m of a wall) and
    the field normal to the edge, then velocity saturation with bit 4."""
    lo = dev.row < dev.cj
    k, kp, h = dev.row[lo], dev.cj[lo], dev.h[lo]
    ok = (dev.kind[k] == K_SEMI) & (dev.kind[kp] == K_SEMI)
    mn = 0.5 * (dev.mu0n[k] + dev.mu0n[kp])
    mp = 0.5 * (dev.mu0p[k] + dev.mu0p[kp])
    m = dev.m
    if dev.models & 16:
        gr = graph_gradient(dev, psi)
        gr[~((dev.kind == K_SEMI) & (dev.qd < 0.5))] = 0.0
        Dw = np.exp(-0.5 * (dev.qd[k] + dev.qd[kp]) / 1e-8)
        t = (dev.pos[kp] - dev.pos[k]) / h[:, None]
        ga, gb = gr[k], gr[kp]
        pa, pb = np.sum(ga * t, axis=1), np.sum(gb * t, axis=1)
        Fa = np.sum(ga * ga, axis=1) - pa * pa
        Fb = np.sum(gb * gb, axis=1) - pb * pb
        F = np.sqrt(np.maximum(0.5 * (Fa + Fb), 0.0))
        Nt = 0.5 * (dev.Nd[k] + dev.Na[k] + dev.Nd[kp] + dev.Na[kp])
        near = Dw > 1e-8
        mn = np.where(near, lombardi(m, mn, F, Nt, Dw, 0), mn)
        mp = np.where(near, lombardi(m, mp, F, Nt, Dw, 1), mp)
    if dev.models & 4:
        dps = psi[kp] - psi[k]
        sn = mob_hf(mn, heat_field(qn[kp] - qn[k], dps, h), m["vsat"][0], m["beta"][0])
        sp = mob_hf(mp, heat_field(qp[kp] - qp[k], dps, h), m["vsat"][1], m["beta"][1])
        mn = mn + 1.0 * (sn - mn)
        mp = mp + 1.0 * (sp - mp)
    gmu = np.zeros((2, dev.nue))
    gmu[0, dev.ue[lo]] = np.where(ok, mn, 0.0)
    gmu[1, dev.ue[lo]] = np.where(ok, mp, 0.0)
    return gmu


def graph_cont(dev, hole, psi, n, p, gmu):
    """Scaled Scharfetter-Gummel rows on a graph, cont_assemble_g :4313-4346:
    unit diagonal, edge values va, rhs, sc, dsc."""
    VT = dev.m["VT"]
    car = p if hole else n
    semi = dev.kind == K_SEMI
    fix = dev.con_mask != 0
    eq = semi & ~fix
    r, j = dev.row, dev.cj
    lnk = dev.lnvK if hole else dev.lncK
    dl = lnk[j] - lnk[r]
    D = (psi[j] - psi[r]) / VT + (-dl if hole else dl)
    g = dev.ar * gmu[1 if hole else 0, dev.ue] * VT / dev.h
    self_ = g * (bernoulli(D) if hole else bernoulli(-D))
    nb = g * (bernoulli(-D) if hole else bernoulli(D))
    act = eq[r] & semi[j]
    dg = np.zeros(dev.N)
    b = np.zeros(dev.N)
    np.add.at(dg, r[act], self_[act])
    fj = act & fix[j]
    np.add.at(b, r[fj], (nb * car[j])[fj])
    va = np.where(act & ~fix[j], -nb * np.maximum(car[j], NFLOOR), 0.0)
    R, dRn, dRp = recomb(dev.m, dev.niK, dev.tauf, n, p)
    dR = np.maximum(dRp if hole else dRn, 0.0)
    dg = dg + dev.vol * dR
    b = b + dev.vol * (dR * car - R)
    sk = np.maximum(car, NFLOOR)
    with np.errstate(invalid="ignore", over="ignore", divide="ignore"):
        w = 1.0 / (dg * sk)
        va = np.where(eq[r], va * w[r], 0.0)
        rhs = np.where(eq, b * w, np.where(semi, 1.0, 0.0))
    sc = np.where(eq, sk, np.where(semi, car, 0.0))
    dsc = np.where(eq, dg * sk, 0.0)
    return np.ones(dev.N), va, rhs, sc, dsc


def graph_currents(dev, psi, n, p, gmu):
    """Terminal currents on a graph (A), their flux scales and electron parts,
    terminal_flux_g :4412-4424 with face_flux_g :4215-4226."""
    nc = len(dev.contacts)
    I, S, In = np.zeros(nc), np.zeros(nc), np.zeros(nc)
    r, j = dev.row, dev.cj
    c = dev.con_mask[r]
    sel = (c > 0) & (dev.kind[r] == K_SEMI) & (dev.kind[j] == K_SEMI) & (dev.con_mask[j] != c)
    ka, kb, e = r[sel], j[sel], np.nonzero(sel)[0]
    VT = dev.m["VT"]
    dps = (psi[kb] - psi[ka]) / VT
    Dn = dps + (dev.lncK[kb] - dev.lncK[ka])
    Dp = dps - (dev.lnvK[kb] - dev.lnvK[ka])
    gn = Q * dev.ar[e] * gmu[0, dev.ue[e]] * VT / dev.h[e]
    gp = Q * dev.ar[e] * gmu[1, dev.ue[e]] * VT / dev.h[e]
    a1, a2 = gn * n[kb] * bernoulli(Dn), gn * n[ka] * bernoulli(-Dn)
    b1, b2 = gp * p[ka] * bernoulli(Dp), gp * p[kb] * bernoulli(-Dp)
    np.add.at(I, c[sel] - 1, (a1 - a2) + (b1 - b2))
    np.add.at(S, c[sel] - 1, (a1 + a2) + (b1 + b2))
    np.add.at(In, c[sel] - 1, a1 - a2)
    return I, S, In
