"""Bias sweeps through the api3 ABI: one api3_resolve per bias point, as the
reference experiment engine drives the core (run_sweep, main.py:3638-3734:
non-swept contacts held, the swept contact set, api3_resolve, currents and
floors read back)."""
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
