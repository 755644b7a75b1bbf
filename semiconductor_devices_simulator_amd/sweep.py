"""Bias sweeps through the api3 ABI: one api3_resolve per bias point, as the
reference experiment engine drives the core (run_sweep, main.py:3638-3734:
non-swept contacts held, the swept contact set, api3_resolve, currents and
floors read back), and the front end's floating-terminal search
(solve_floating and its evaluation _fl_eval, main.py:3576-3636) with the
floating-node sweep of run_sweep around it."""
import time

import numpy as np


def run_sweep(lib, spec, cids, segments=None, on_point=None, max_points=None):
    """Solve every bias point of spec.sweep (or `segments`), starting from the
    core's present state.  Returns dict of arrays: V (swept voltage), I and F
    (n_points x n_contacts: current A and resolution floor A), conv, t (s per
    point), gummel (Gummel iterations per point)."""
    segs = segments if segments is not None else spec.sweep
    nc = len(cids)
    R = dict(V=[], I=[], F=[], conv=[], t=[], gummel=[], seg=[])
    done = 0
    for si, (tgt, vs, ve, npts, hold) in enumerate(segs):
        for c, v in hold.items():
            lib.api3_set_contact_V(cids[c], v)
        for V in (np.linspace(vs, ve, npts) if npts > 1 else [vs]):
            if max_points is not None and done >= max_points:
                break
            lib.api3_set_contact_V(cids[tgt], float(V))
            g0 = lib.api3_get_stat(3)
            t0 = time.perf_counter()
            ok = lib.api3_resolve()
            dt = time.perf_counter() - t0
            R["V"].append(float(V))
            R["I"].append([lib.api3_get_contact_I(cids[c]) for c in range(nc)])
            R["F"].append([lib.api3_get_contact_Ifloor(cids[c]) for c in range(nc)])
            R["conv"].append(bool(ok))
            R["t"].append(dt)
            R["gummel"].append(int(lib.api3_get_stat(3) - g0))
            R["seg"].append(si)
            done += 1
            if on_point:
                on_point(len(R["V"]) - 1, R)
    out = {k: np.asarray(v) for k, v in R.items()}
    return out


def _fl_eval(lib, cids, fidx, V):
    """Put the floating node (contacts fidx) at V and solve (main.py:3576-
    3587): returns (net current into the device through the node, the
    current tolerance max(1e-3 max|I|, 3 x the node's floors), converged,
    the voltage reached)."""
    for c in fidx:
        lib.api3_set_contact_V(cids[c], V)
    ok = bool(lib.api3_resolve())
    Vr = lib.api3_get_contact_Vsolved(cids[fidx[0]])
    if Vr != Vr:
        Vr = V
    nc = lib.api3_get_ncon()
    Inet = sum(lib.api3_get_contact_I(cids[c]) for c in fidx)
    Imax = max([abs(lib.api3_get_contact_I(c)) for c in range(nc)] + [0.0])
    Fl = sum(lib.api3_get_contact_Ifloor(cids[c]) for c in fidx)
    return Inet, max(1e-3 * Imax, 3.0 * Fl, 1e-30), ok, Vr


def solve_floating(lib, cids, fidx, V0, lo, hi, slope=None, vtol=1e-9, maxev=40, max_bad=0):
    """Zero-current voltage of a floating node in [lo, hi] (main.py:3589-
    3636): the current is non-decreasing in V, so the root lies between the
    lowest and highest applied voltages; safeguarded secant, Illinois regula
    falsi once both signs are known, bisection as the fallback, every
    evaluation continuing from the previous solution.  Returns (V, I_net,
    converged, evaluations, dI/dV estimate)."""
    a, b = float(lo), float(hi)
    fa = fb = None
    side = 0
    x = min(max(float(V0), a), b)
    hist, ok, fx, xr, found, bad = [], False, 0.0, x, False, 0
    for _ in range(maxev):
        fx, tol, ok, xr = _fl_eval(lib, cids, fidx, x)
        hist.append((xr, fx))
        if ok and abs(fx) <= tol:
            found = True
            break
        bad = 0 if ok else bad + 1
        if max_bad and bad >= max_bad:
            break
        if fx > 0.0:
            b, fb = xr, fx
            side = side - 1 if side < 0 else -1
        else:
            a, fa = xr, fx
            side = side + 1 if side > 0 else 1
        if b - a <= vtol:
            found = ok
            break
        cand = None
        if len(hist) >= 2:
            (x1, f1), (x2, f2) = hist[-2], hist[-1]
            if f2 != f1 and x2 != x1:
                cand = x2 - f2 * (x2 - x1) / (f2 - f1)
        elif slope and slope > 0.0:
            cand = xr - fx / slope
        if fa is not None and fb is not None and not (cand is not None and a < cand < b):
            fa2, fb2 = fa, fb
            if side >= 2:
                fb2 = 0.5 * fb
            if side <= -2:
                fa2 = 0.5 * fa
            cand = b - fb2 * (b - a) / (fb2 - fa2) if fb2 != fa2 else 0.5 * (a + b)
        if cand is None or not (a < cand < b):
            cand = 0.5 * (a + b)
        if abs(cand - xr) < 0.25 * vtol:
            cand = xr + (0.25 * vtol if fx < 0 else -0.25 * vtol)
        x = min(max(cand, a), b)
    sl = None
    if len(hist) >= 2:
        (x1, f1), (x2, f2) = hist[-2], hist[-1]
        if x2 != x1 and (f2 - f1) / (x2 - x1) > 0:
            sl = (f2 - f1) / (x2 - x1)
    return xr, fx, found, len(hist), (sl or slope)


def floating_sweep(lib, spec, cids, tgt, Vs, Ve, npts, flt):
    """Sweep contact tgt from Vs to Ve in npts points with the node of the
    contacts flt floating (run_sweep with flt, main.py:3638-3734, without the
    bisection refinement): at every point the node's zero-current voltage,
    searched from the linear extrapolation of the last two points' node
    voltages within the range of the applied voltages.  Returns dict of
    arrays: V, Vf (node voltage), I, F, conv, nev (evaluations), tol (the
    search's current tolerance at the root), slope (its dI/dV estimate)."""
    nc = len(cids)
    vapp = [c.V for c in spec.contacts]
    fixed = [vapp[c] for c in range(nc) if c not in flt and c != tgt]
    for c in range(nc):
        if c != tgt and c not in flt:
            lib.api3_set_contact_V(cids[c], vapp[c])
    R = dict(V=[], Vf=[], I=[], F=[], conv=[], nev=[], tol=[], slope=[])
    slope = None
    for Vt in (np.linspace(Vs, Ve, npts) if npts > 1 else [Vs]):
        lib.api3_set_contact_V(cids[tgt], float(Vt))
        if not R["V"]:
            guess = vapp[flt[0]]
        elif len(R["V"]) >= 2 and R["V"][-1] != R["V"][-2]:
            (V1, F1), (V2, F2) = (R["V"][-1], R["Vf"][-1]), (R["V"][-2], R["Vf"][-2])
            guess = F1 + (F1 - F2) * (Vt - V1) / (V1 - V2)
        else:
            guess = R["Vf"][-1]
        f = fixed + [float(Vt)]
        lo, hi = min(f), max(f)
        Vx, _, ok, ne, slope = solve_floating(lib, cids, flt, min(max(guess, lo), hi), lo, hi, slope)
        R["V"].append(float(Vt))
        R["Vf"].append(Vx)
        R["I"].append([lib.api3_get_contact_I(cids[c]) for c in range(nc)])
        R["F"].append([lib.api3_get_contact_Ifloor(cids[c]) for c in range(nc)])
        R["conv"].append(bool(ok))
        R["nev"].append(ne)
        Imax = max(abs(v) for v in R["I"][-1])
        R["tol"].append(max(1e-3 * Imax, 3.0 * sum(R["F"][-1][c] for c in flt), 1e-30))
        R["slope"].append(slope if slope else np.nan)
    return {k: np.asarray(v) for k, v in R.items()}
