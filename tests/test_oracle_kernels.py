"""The oracle's assembled operators against the reference core's, dumped at
a perturbed non-equilibrium state of small devices (tests/golden/k_*.npz),
plus invariants of the box-method operators."""
import numpy as np
import pytest

import semisim_oracle as O
from conftest import load_golden, oracle_device

FIXTURES = ["k_mosfet_12x8x6", "k_diode_6x5x16", "k_mosfet_20x14x9", "k_mlda_mosfet_12x8x6", "k_mlda_mosfet_20x14x9"]


@pytest.fixture(scope="module", params=FIXTURES)
def case(request):
    g = load_golden(request.param)
    return g, oracle_device(g)


def close(a, b, rtol=1e-12, scale=None):
    a, b = np.asarray(a), np.asarray(b)
    s = np.max(np.abs(b)) if scale is None else scale
    assert np.all(np.abs(a - b) <= rtol * (np.abs(b) + 1e-300) + 1e-14 * s), np.max(np.abs(a - b))


def test_prepare(case):
    g, dev = case
    np.testing.assert_array_equal(dev.kind, g["kind"])
    np.testing.assert_array_equal(dev.con_mask, g["con_mask"])
    np.testing.assert_array_equal(dev.gate_mask, g["gate_mask"])
    assert dev.Cref == pytest.approx(float(g["Cref"]), rel=1e-15)
    close(dev.bgn, g["bgn"])
    n, p = dev.carriers(g["psi"], g["phin"], g["phip"])
    close(n, g["n"], 1e-12)
    close(p, g["p"], 1e-12)


def test_poisson(case):
    g, dev = case
    F, A, sc, nrm = O.poisson_eval(dev, g["psi"], g["phin"], g["phip"], g["Va"])
    close(F, g["poisson_F"], 1e-9)
    for q in range(7):
        close(A[q], g["poisson_A"][q], 1e-12)
    assert nrm == pytest.approx(float(g["poisson_norm"]), rel=1e-9)


def test_poisson_jacobian_symmetric(case):
    g, dev = case
    A = g["poisson_A"]
    N = dev.N
    free = dev.con_mask == 0
    for ax, (m, p), st in ((0, (1, 2), 1), (1, (3, 4), dev.Nx), (2, (5, 6), dev.Nx * dev.Ny)):
        k, kp, _, _ = dev.edges(ax)
        both = free[k] & free[kp]
        np.testing.assert_allclose(A[p][k[both]], A[m][kp[both]], rtol=1e-15)
    # diagonally dominant on free rows (Poisson is an M-matrix plus a positive charge term)
    assert np.all(A[0][free] >= np.sum(np.abs(A[1:, free]), axis=0) * (1 - 1e-14))


def test_face_mobility(case):
    g, dev = case
    close(dev.mu0n, g["mu0n"], 1e-13)
    close(dev.mu0p, g["mu0p"], 1e-13)
    fmu = O.face_mobility(dev, g["psi"], g["phin"], g["phip"])
    for q in range(6):
        close(fmu[q], g["fmu"][q], 1e-12)


@pytest.mark.parametrize("hole", [0, 1])
def test_continuity(case, hole):
    g, dev = case
    t = "p" if hole else "n"
    n, p = dev.carriers(g["psi"], g["phin"], g["phip"])
    A, rhs, sc, dsc = O.cont_assemble(dev, hole, g["psi"], n, p, g["fmu"])
    for q in range(7):
        close(A[q], g[f"cont{t}_A"][q], 1e-11, scale=1.0)
    close(rhs, g[f"cont{t}_rhs"], 1e-10, scale=1.0)
    close(sc, g[f"cont{t}_sc"], 1e-12)
    close(dsc, g[f"cont{t}_dsc"], 1e-11)


@pytest.mark.parametrize("hole", [0, 1])
def test_continuity_conservation(case, hole):
    """Scaled SG rows have unit diagonal and non-positive couplings, and the
    unscaled operator conserves current: for a node b whose neighbours are all
    equation rows, dg_b - sum_a nb_ab = vol_b dR_b (the flux leaving b through
    an edge is what the neighbour gains), i.e.
    dsc_b + sum_a A_ab dsc_a = vol_b dR_b s_b."""
    g, dev = case
    t = "p" if hole else "n"
    A, dsc, sc = g[f"cont{t}_A"], g[f"cont{t}_dsc"], g[f"cont{t}_sc"]
    assert np.all(A[0] == 1.0) and np.all(A[1:] <= 0.0)
    eq = (dev.kind == 0) & (dev.con_mask == 0)
    n, p = dev.carriers(g["psi"], g["phin"], g["phip"])
    _, dRn, dRp = O.recomb(dev.m, dev.niK, dev.tauf, n, p)
    dR = np.maximum(dRp if hole else dRn, 0.0)
    col = dsc.copy()
    inner = eq.copy()
    for ax in range(3):
        k, kp, _, _ = dev.edges(ax)
        col[kp] += A[2 + 2 * ax][k] * dsc[k]        # row k couples to kp (+ direction)
        col[k] += A[1 + 2 * ax][kp] * dsc[kp]       # row kp couples to k (- direction)
        semi_nb = dev.kind[kp] == 0
        inner[k[semi_nb & ~eq[kp]]] = False
        inner[kp[(dev.kind[k] == 0) & ~eq[k]]] = False
    want = dev.volume() * dR * sc
    scale = np.abs(dsc)
    assert np.all(np.abs(col - want)[inner] <= 1e-10 * scale[inner] + 1e-300)


def test_matvec(case):
    g, dev = case
    y = O.matvec(g["poisson_A"], (dev.Nx, dev.Ny, dev.Nz), g["matvec_x"])
    close(y, g["matvec_y"], 1e-13)


def _struct_pattern(dev, op):
    """Structural pattern of the reference's operators from the node kinds and
    contact masks alone (N x 7 in the order c, xm, xp, ym, yp, zm, zp):
    Poisson (poisson_eval :2732-2789) - Dirichlet rows identity, a coupling
    toward every existing neighbour that is not Dirichlet; continuity
    (cont_assemble :3067-3106) - identity rows off the free semiconductor
    nodes, couplings toward free semiconductor neighbours only."""
    N = dev.N
    k = np.arange(N)
    i, j, l = k % dev.Nx, (k // dev.Nx) % dev.Ny, k // (dev.Nx * dev.Ny)
    nb = [(i > 0, k - 1), (i < dev.Nx - 1, k + 1), (j > 0, k - dev.Nx), (j < dev.Ny - 1, k + dev.Nx),
          (l > 0, k - dev.Nx * dev.Ny), (l < dev.Nz - 1, k + dev.Nx * dev.Ny)]
    if op == "poisson":
        row = dev.con_mask == 0
        ok = lambda kb: dev.con_mask[kb] == 0          # noqa: E731
    else:
        row = (dev.kind == 0) & (dev.con_mask == 0)
        ok = lambda kb: (dev.kind[kb] == 0) & (dev.con_mask[kb] == 0)   # noqa: E731
    P = np.zeros((7, N), bool)
    P[0] = True
    for q, (has, kb) in enumerate(nb):
        kbc = np.where(has, kb, 0)
        P[q + 1] = row & has & ok(kbc)
    return P


@pytest.mark.parametrize("op", ["poisson", "n", "p"])
def test_operator_sparsity_pattern(case, op):
    """The nonzero pattern of the reference's assembled operators is exactly
    the structural pattern of the kinds and contact masks, and the oracle's
    operators have the same pattern (integer data: compared exactly)."""
    g, dev = case
    A_ref = g["poisson_A"] if op == "poisson" else g[f"cont{op}_A"]
    np.testing.assert_array_equal(A_ref != 0.0, _struct_pattern(dev, op))
    if op == "poisson":
        A = O.poisson_eval(dev, g["psi"], g["phin"], g["phip"], g["Va"])[1]
    else:
        n, p = dev.carriers(g["psi"], g["phin"], g["phip"])
        A = O.cont_assemble(dev, op == "p", g["psi"], n, p, g["fmu"])[0]
    np.testing.assert_array_equal(np.asarray(A) != 0.0, A_ref != 0.0)


def test_stencil_csr(case):
    """CSR of the reference's Poisson operator: the rows applied to a vector
    give the reference's own matvec (matvec_L :1975-1993), columns ascend
    within rows, and the entry count equals the nonzero count of the
    diagonals."""
    g, dev = case
    A = g["poisson_A"]
    rp, col, val = O.stencil_csr(A, (dev.Nx, dev.Ny, dev.Nz))
    assert rp[0] == 0 and rp[-1] == np.count_nonzero(A) == len(col)
    x = g["matvec_x"]
    y = np.zeros(dev.N)
    np.add.at(y, np.repeat(np.arange(dev.N), np.diff(rp)), val * x[col])
    close(y, g["matvec_y"], 1e-13)
    for r in range(dev.N):
        c = col[rp[r]:rp[r + 1]]
        assert np.all(np.diff(c) > 0) and r in c


def test_galerkin_poisson(case):
    g, dev = case
    levs = O.mg_hierarchy((dev.Nx, dev.Ny, dev.Nz))
    assert len(levs) == int(g["mgP_nlev"])
    for L, lev in enumerate(levs):
        np.testing.assert_array_equal(np.array(lev, dtype=np.int64), g[f"mgP_L{L}_dims"].astype(np.int64))
    A = g["poisson_A"]
    for L in range(1, min(len(levs), 3)):
        A, _ = O.galerkin(levs[L - 1], A, np.ones(A.shape[1]))
        for q in range(7):
            close(A[q], g[f"mgP_L{L}_A"][q], 1e-12)


def test_galerkin_constant_vector(case):
    """P^T A P applied to the constant equals P^T (A 1): row sums are summed."""
    g, dev = case
    lev = O.mg_hierarchy((dev.Nx, dev.Ny, dev.Nz))[0]
    A = g["poisson_A"]
    C, cdims = O.galerkin(lev, A, np.ones(dev.N))
    lhs = O.matvec(C, cdims, np.ones(C.shape[1]))
    rhs = O.restrict(lev, O.matvec(A, (dev.Nx, dev.Ny, dev.Nz), np.ones(dev.N)))
    zero = np.all(C[1:] == 0, axis=0) & (C[0] == 1)
    np.testing.assert_allclose(lhs[~zero], rhs[~zero], rtol=1e-12, atol=1e-12 * np.abs(A).max())


@pytest.mark.parametrize("hole", [0, 1])
def test_galerkin_continuity(case, hole):
    g, dev = case
    t = "p" if hole else "n"
    VT = dev.m["VT"]
    semi = (dev.kind == 0) & (dev.con_mask == 0)
    ph = (-g["phip"] / VT) if hole else (g["phin"] / VT)
    ph0 = np.where(semi, ph, -1e300)
    levs = O.mg_hierarchy((dev.Nx, dev.Ny, dev.Nz))
    pws = O.mg_weights(levs, ph0)
    close(pws[0], g[f"mg{t}_pw0"], 1e-13)
    A, _ = O.galerkin(levs[0], g[f"cont{t}_A"], pws[0], g[f"cont{t}_dsc"])
    if int(g[f"mg{t}_nlev"]) > 1:
        for q in range(7):
            close(A[q], g[f"mg{t}_L1_A"][q], 1e-11)


def test_terminal_currents(case):
    g, dev = case
    n, p = dev.carriers(g["psi"], g["phin"], g["phip"])
    I, S, In = O.terminal_currents(dev, g["psi"], n, p, g["fmu"])
    close(S, g["cur_S"], 1e-12)
    close(I, g["cur_I"], 1e-9, scale=np.max(g["cur_S"]))
    close(In, g["cur_In"], 1e-9, scale=np.max(g["cur_S"]))


@pytest.mark.parametrize("hole", [0, 1])
def test_cont_update(case, hole):
    g, dev = case
    t = "p" if hole else "n"
    n, p = dev.carriers(g["psi"], g["phin"], g["phip"])
    lo, hi = min(g["Va"]), max(g["Va"])
    lo, hi = _vrange(g)
    q, n2, p2, dq, dmax = O.cont_update(dev, hole, g["psi"], n, p, g["phin"], g["phip"], g[f"cont{t}_u"],
                                        g[f"cont{t}_sc"], lo, hi)
    # dq is a difference of quasi-Fermi levels of about 1 V: an absolute
    # floor of 1e-15 V is a few units in their last place
    close(dq, g[f"cont{t}_dq"], 1e-9, scale=0.1)
    assert dmax == pytest.approx(float(g[f"cont{t}_dmax"]), rel=1e-8)


def _vrange(g):
    """Carrier-terminal voltage span: gates (bc 2) excluded (set_vrange)."""
    from conftest import golden_contacts
    V = [v for v, c in zip(g["Va"], golden_contacts(g)) if c["bc"] != 2]
    return min(V), max(V)
