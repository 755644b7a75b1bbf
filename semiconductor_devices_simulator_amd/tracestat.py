"""Summary of a rocprofv3 kernel trace (--kernel-trace --output-format csv).

Prints the window the kernels span, the time at least one kernel ran (busy),
the gaps between kernels, and the kernels by total time.  --from NAME starts
the window at the first launch of a kernel whose name starts with NAME (the
sweep's first continuity assembly, k_cont, skips the set-up and the
equilibrium solve).

    python -m semiconductor_devices_simulator_amd.tracestat kernel_trace.csv --from k_cont --top 40

--by-grid keeps launches of one kernel with different grid sizes apart (the
layers of the warm-start network); --match keeps only matching kernels.

--pmc reads a counter collection instead (rocprofv3 --pmc, csv output) and
prints, per kernel and grid size, the dispatches and each counter's mean per
dispatch:

    python -m semiconductor_devices_simulator_amd.tracestat counter_collection.csv --pmc --match k_line_scan
"""
import argparse
import csv
import sys
from collections import defaultdict


def load(path, by_grid=False):
    rows = []
    with open(path) as f:
        for r in csv.DictReader(f):
            name = r["Kernel_Name"]
            if by_grid:
                name = name.split("(")[0] + " grid " + r.get("Grid_Size", r.get("Grid_Size_X", "?"))
            rows.append((int(r["Start_Timestamp"]), int(r["End_Timestamp"]), name))
    rows.sort()
    return rows


def summarise(rows, start=None):
    if start:
        t = min((s for s, _, n in rows if n.startswith(start)), default=None)
        if t is None:
            raise SystemExit(f"no kernel named {start}*")
        rows = [r for r in rows if r[0] >= t]
    busy, gaps = 0, []
    cur_s = cur_e = None
    for s, e, _ in rows:
        if cur_e is None:
            cur_s, cur_e = s, e
        elif s > cur_e:
            busy += cur_e - cur_s
            gaps.append(s - cur_e)
            cur_s, cur_e = s, e
        else:
            cur_e = max(cur_e, e)
    busy += cur_e - cur_s
    per = defaultdict(lambda: [0, 0])
    for s, e, n in rows:
        k = n.split("(")[0]
        per[k][0] += 1
        per[k][1] += e - s
    return dict(window=cur_e - rows[0][0], busy=busy, launches=len(rows), gaps=sorted(gaps), per=per)


def load_pmc(path, match=None):
    """Mean counter value per (kernel, grid size, counter) and the number of
    dispatches it covers."""
    acc = defaultdict(lambda: [0, 0.0])
    with open(path) as f:
        for r in csv.DictReader(f):
            name = r["Kernel_Name"].split("(")[0]
            if match and match not in name:
                continue
            grid = r.get("Grid_Size", r.get("Grid_Size_X", "?"))
            a = acc[(name, grid, r["Counter_Name"])]
            a[0] += 1
            a[1] += float(r["Counter_Value"])
    return {k: (n, v / n) for k, (n, v) in acc.items()}


def print_pmc(P):
    counters = sorted({c for _, _, c in P})
    print("| kernel | grid | dispatches | " + " | ".join(counters) + " |")
    print("|---|---:|---:|" + "---:|" * len(counters))
    for name, grid in sorted({(n, g) for n, g, _ in P}):
        cells = [P.get((name, grid, c)) for c in counters]
        n = max(x[0] for x in cells if x)
        print(f"| `{name[:70]}` | {grid} | {n} | " +
              " | ".join(f"{x[1]:.0f}" if x else "" for x in cells) + " |")


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("trace")
    ap.add_argument("--from", dest="start", default=None)
    ap.add_argument("--top", type=int, default=30)
    ap.add_argument("--by-grid", action="store_true", help="one line per kernel and grid size")
    ap.add_argument("--match", default=None, help="only kernels whose name contains this")
    ap.add_argument("--pmc", action="store_true", help="the file is a counter collection: mean per dispatch")
    a = ap.parse_args(argv)
    if a.pmc:
        print_pmc(load_pmc(a.trace, a.match))
        return 0
    rows = load(a.trace, a.by_grid)
    if a.match:
        rows = [r for r in rows if a.match in r[2]]
    S = summarise(rows, a.start)
    g = S["gaps"]
    med = g[len(g) // 2] / 1e3 if g else 0.0
    print(f"window {S['window'] / 1e9:.3f} s, kernels busy {S['busy'] / 1e9:.3f} s "
          f"({100 * S['busy'] / S['window']:.1f}%), {S['launches']} launches, {len(g)} gaps "
          f"(median {med:.2f} us, total {sum(g) / 1e9:.3f} s)")
    print("| kernel | launches | avg us | % of busy |")
    print("|---|---:|---:|---:|")
    for k, (c, t) in sorted(S["per"].items(), key=lambda x: -x[1][1])[:a.top]:
        print(f"| `{k[:80]}` | {c} | {t / c / 1e3:.2f} | {100 * t / S['busy']:.1f} |")
    return 0


if __name__ == "__main__":
    sys.exit(main())
