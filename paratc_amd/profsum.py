"""Summaries of rocprofv3 output for the kernels of this library.

    python -m paratc_amd.profsum trace <run_results.db>
        per-kernel dispatch count and durations from a --kernel-trace run
    python -m paratc_amd.profsum lds <lds.csv>
        per-kernel mean SQ_LDS_BANK_CONFLICT and SQ_ACTIVE_INST_LDS from one
        --pmc run of both counters, and their ratio
    python -m paratc_amd.profsum sq <sq.csv>
        per-kernel shares of wave time from one --pmc run of SQ_WAVE_CYCLES,
        SQ_WAIT_ANY, SQ_WAIT_INST_ANY, SQ_WAIT_INST_LDS, SQ_ACTIVE_INST_VALU and
        SQ_ACTIVE_INST_LDS (with SQ_WAVES and SQ_BUSY_CYCLES)
    python -m paratc_amd.profsum totals <fetch.csv> <write.csv> <copy_bytes> [--max-grid N] [--skip S ...]
        HBM bytes summed over every dispatch of the run (k_copy excluded),
        scaled as in the traffic table; --max-grid leaves out launches of more
        than N threads and --skip kernels whose name contains S
    python -m paratc_amd.profsum traffic <fetch.csv> <write.csv> <copy_bytes>
        per-kernel mean FETCH_SIZE / WRITE_SIZE (KiB per dispatch) from two
        --pmc runs, scaled to HBM bytes by what the same counters read for
        k_copy, the plain device copy of copy_bytes bytes (read once, written
        once; bench.py --calib-copy copies one field of nx * (ny + 2) * nz
        doubles)

The traffic table is what bench.py reads to fill roofline.traffic.  Scaling:
read_factor = copy bytes / copy FETCH_SIZE, write_factor = copy bytes /
copy WRITE_SIZE.
"""
from __future__ import annotations

import collections
import csv
import sqlite3
import sys
from typing import Dict, List, Optional, Tuple

TRAFFIC_HEADER = ("| kernel @ grid (threads) | dispatches | FETCH_SIZE KiB | WRITE_SIZE KiB | read bytes | write bytes "
                  "| traffic bytes |")


def short_name(name: str) -> str:
    """Kernel name without the argument list (template arguments kept)."""
    name = name.replace("(anonymous namespace)::", "")
    depth = 0
    for q, ch in enumerate(name):
        if ch == "<":
            depth += 1
        elif ch == ">":
            depth -= 1
        elif ch == "(" and depth == 0:
            name = name[:q]
            break
    for pre in ("void ", "ptc::"):
        name = name.replace(pre, "")
    return name.strip()


def trace_table(db: str) -> str:
    """Per kernel and grid size (launches of one kernel over different extents,
    such as slabs, are listed apart)."""
    con = sqlite3.connect(db)
    rows = con.execute("select name, grid_x * grid_y * grid_z, count(*), sum(end-start), avg(end-start), "
                       "min(end-start), max(end-start) from kernels group by name, grid_x * grid_y * grid_z "
                       "order by sum(end-start) desc").fetchall()
    tot = sum(r[3] for r in rows) or 1
    out = ["| kernel | grid (threads) | calls | total us | avg us | min us | max us | % |",
           "|---|---:|---:|---:|---:|---:|---:|---:|"]
    for name, grid, n, s, a, lo, hi in rows:
        out.append(f"| {short_name(name)} | {grid} | {n} | {s / 1e3:.1f} | {a / 1e3:.2f} | {lo / 1e3:.2f} "
                   f"| {hi / 1e3:.2f} | {100 * s / tot:.1f} |")
    return "\n".join(out)


def _counter_means(path: str, counter: Optional[str] = None) -> Dict[str, Tuple[int, float]]:
    """Mean counter value per kernel and grid size ("name @ grid"); with
    `counter`, only that counter's rows."""
    acc: Dict[str, List[float]] = collections.defaultdict(list)
    with open(path) as fh:
        for r in csv.DictReader(fh):
            if counter is not None and r["Counter_Name"] != counter:
                continue
            acc[f"{short_name(r['Kernel_Name'])} @ {r['Grid_Size']}"].append(float(r["Counter_Value"]))
    return {k: (len(v), sum(v) / len(v)) for k, v in acc.items()}


def lds_table(path: str) -> str:
    """LDS bank-conflict cycles against LDS-active cycles per kernel (kernels
    with LDS traffic only)."""
    conf = _counter_means(path, "SQ_LDS_BANK_CONFLICT")
    act = _counter_means(path, "SQ_ACTIVE_INST_LDS")
    out = ["| kernel @ grid (threads) | dispatches | SQ_LDS_BANK_CONFLICT | SQ_ACTIVE_INST_LDS | ratio |",
           "|---|---:|---:|---:|---:|"]
    for k in sorted(act, key=lambda k: -act[k][1]):
        if act[k][1] <= 0 or k not in conf:
            continue
        out.append(f"| {k} | {act[k][0]} | {conf[k][1]:.0f} | {act[k][1]:.0f} | {conf[k][1] / act[k][1]:.3f} |")
    return "\n".join(out)


SQ_SHARES = ("SQ_WAIT_ANY", "SQ_WAIT_INST_ANY", "SQ_WAIT_INST_LDS", "SQ_ACTIVE_INST_VALU", "SQ_ACTIVE_INST_LDS")


def sq_table(path: str, min_wave_cycles: float = 1e9) -> str:
    """Each counter as a share of SQ_WAVE_CYCLES (the summed lifetime of the
    kernel's waves), per kernel with at least min_wave_cycles of wave time."""
    means = {c: _counter_means(path, c) for c in ("SQ_WAVE_CYCLES", "SQ_WAVES") + SQ_SHARES}
    wc = means["SQ_WAVE_CYCLES"]
    out = ["| kernel @ grid (threads) | dispatches | waves | wave cycles | " +
           " | ".join(c[3:].lower() for c in SQ_SHARES) + " |",
           "|---|---:|---:|---:|" + "---:|" * len(SQ_SHARES)]
    for k in sorted(wc, key=lambda k: -wc[k][1]):
        n, cyc = wc[k]
        if cyc < min_wave_cycles:
            continue
        cells = [f"{means[c][k][1] / cyc:.2f}" if k in means[c] else "-" for c in SQ_SHARES]
        waves = means["SQ_WAVES"][k][1] if k in means["SQ_WAVES"] else 0
        out.append(f"| {k} | {n} | {waves:.0f} | {cyc:.3g} | " + " | ".join(cells) + " |")
    return "\n".join(out)


def traffic_table(fetch_csv: str, write_csv: str, nbytes: float) -> str:
    fetch, write = _counter_means(fetch_csv), _counter_means(write_csv)
    copy = [k for k in fetch if k.startswith("k_copy @") or k.startswith("k_copy_nt @")]
    if not copy or copy[0] not in write:
        raise SystemExit("k_copy is not in the counter files: profile bench.py --calib-copy")
    rf = nbytes / (fetch[copy[0]][1] * 1024.0)
    wf = nbytes / (write[copy[0]][1] * 1024.0)
    out = [f"read_factor {rf:.4f} write_factor {wf:.4f} (k_copy of {nbytes:.0f} B each way)", "",
           TRAFFIC_HEADER, "|---|---:|---:|---:|---:|---:|---:|"]
    for k in sorted(fetch, key=lambda k: -fetch[k][1]):
        if k not in write:
            continue
        fr, wr = fetch[k][1], write[k][1]
        rb, wb = fr * 1024.0 * rf, wr * 1024.0 * wf
        out.append(f"| {k} | {fetch[k][0]} | {fr:.1f} | {wr:.1f} | {rb:.0f} | {wb:.0f} | {rb + wb:.0f} |")
    return "\n".join(out)


def _counter_sums(path: str, max_grid: Optional[int] = None,
                  skip: Tuple[str, ...] = ()) -> Tuple[Dict[str, float], Dict[str, int]]:
    """Summed counter value and dispatch count per kernel name, leaving out
    launches of more than max_grid threads and names that contain a skip
    string."""
    tot: Dict[str, float] = collections.defaultdict(float)
    cnt: Dict[str, int] = collections.defaultdict(int)
    with open(path) as fh:
        for r in csv.DictReader(fh):
            k = short_name(r["Kernel_Name"])
            if (max_grid is not None and int(r["Grid_Size"]) > max_grid) or any(s in k for s in skip):
                continue
            tot[k] += float(r["Counter_Value"])
            cnt[k] += 1
    return tot, cnt


def totals(fetch_csv: str, write_csv: str, nbytes: float, max_grid: Optional[int] = None,
           skip: Tuple[str, ...] = ()) -> Tuple[int, float, float]:
    """(dispatches, read bytes, write bytes) over the run, k_copy left out
    (and launches past max_grid threads or with a skip string in the name);
    the copy's own counters give the scale factors."""
    fetch, write = _counter_means(fetch_csv), _counter_means(write_csv)
    copy = [k for k in fetch if k.startswith("k_copy @") or k.startswith("k_copy_nt @")]
    if not copy or copy[0] not in write:
        raise SystemExit("k_copy is not in the counter files: profile bench.py --calib-copy")
    rf = nbytes / (fetch[copy[0]][1] * 1024.0)
    wf = nbytes / (write[copy[0]][1] * 1024.0)
    (fs, fc), (ws, _) = _counter_sums(fetch_csv, max_grid, skip), _counter_sums(write_csv, max_grid, skip)
    keep = [k for k in fs if not k.startswith("k_copy")]
    n = sum(fc[k] for k in keep)
    return n, sum(fs[k] for k in keep) * 1024.0 * rf, sum(ws.get(k, 0.0) for k in keep) * 1024.0 * wf


def read_value(md_path: str, key: str) -> Optional[float]:
    """The number after `key` on a line of its own (``key value``) in a
    summary, or None."""
    try:
        for ln in open(md_path):
            cells = ln.split()
            if len(cells) == 2 and cells[0] == key:
                return float(cells[1])
    except (OSError, ValueError):
        return None
    return None


def read_traffic(md_path: str, kernel: str) -> Optional[float]:
    """Traffic bytes per launch of `kernel` (a short-name prefix) from a traffic
    table, or None; with several grid sizes, the largest launch."""
    try:
        lines = open(md_path).read().splitlines()
    except OSError:
        return None
    best = None
    for ln in lines:
        cells = [c.strip() for c in ln.strip().strip("|").split("|")]
        if len(cells) == 7 and cells[0].startswith(kernel) and cells[6].replace(".", "").isdigit():
            grid = int(cells[0].rsplit("@", 1)[1]) if "@" in cells[0] else 0
            if best is None or grid > best[0]:
                best = (grid, float(cells[6]))
    return None if best is None else best[1]


def main(argv: List[str]) -> None:
    if len(argv) >= 2 and argv[0] == "trace":
        print(trace_table(argv[1]))
    elif len(argv) >= 2 and argv[0] == "lds":
        print(lds_table(argv[1]))
    elif len(argv) >= 2 and argv[0] == "sq":
        print(sq_table(argv[1]))
    elif len(argv) >= 4 and argv[0] == "totals":
        max_grid, skip, rest = None, [], argv[4:]
        while rest:
            if rest[0] == "--max-grid" and len(rest) > 1:
                max_grid, rest = int(rest[1]), rest[2:]
            elif rest[0] == "--skip" and len(rest) > 1:
                skip, rest = skip + [rest[1]], rest[2:]
            else:
                raise SystemExit(__doc__)
        n, rb, wb = totals(argv[1], argv[2], float(argv[3]), max_grid, tuple(skip))
        print(f"dispatches {n}")
        print(f"read_bytes {rb:.0f}")
        print(f"write_bytes {wb:.0f}")
        print(f"traffic_bytes {rb + wb:.0f}")
    elif len(argv) >= 4 and argv[0] == "traffic":
        print(traffic_table(argv[1], argv[2], float(argv[3])))
    else:
        raise SystemExit(__doc__)


if __name__ == "__main__":
    main(sys.argv[1:])
