"""Summaries of rocprofv3 result databases (the SQLite files rocprofv3 writes with -d/-o).

python -m fvm_amd.rocprof_summary trace DB [N]
    per-kernel dispatch count, total / mean / min / max duration (top N by total time)
python -m fvm_amd.rocprof_summary pmc DB
    per-kernel mean of every collected counter per dispatch
python -m fvm_amd.rocprof_summary traffic FETCH_DB WRITE_DB COPY_FETCH_DB COPY_WRITE_DB COPY_MIB
        [CELLS [NAME=ROCPROF_NAME ...]]
    HBM bytes per dispatch of every kernel: FETCH_SIZE and WRITE_SIZE (KiB per dispatch), each
    scaled so that the plain copy kernel of COPY_MIB MiB reads and writes exactly that many
    bytes; prints JSON {"fetch_scale", "write_scale", "all_kernels": {short name: {...}}}.
    With CELLS every entry also gets bytes_per_cell, and each NAME=ROCPROF_NAME pair adds
    kernels[NAME] (the name bench.py times the kernel under) for that rocprof kernel.
"""
from __future__ import annotations

import json
import re
import sqlite3
import sys


def short_name(name: str) -> str:
    """Kernel name without namespace, template arguments and parameter list."""
    n = name.replace("(anonymous namespace)::", "")
    n = re.sub(r"\(.*$", "", n)
    n = re.sub(r"^void ", "", n)
    if n.startswith("Cijk_"):  # rocBLAS/Tensile GEMM kernels
        return "rocblas_gemm:" + n.split("_MT")[1].split("_")[0] if "_MT" in n else n[:24]
    tmpl = re.search(r"<([^<>]*)>$", n)
    base = re.sub(r"<.*$", "", n)
    return f"{base}<{tmpl.group(1)}>" if tmpl else base


def trace(db: str, top: int = 30):
    con = sqlite3.connect(db)
    rows = con.execute(
        "select name, count(*), sum(end-start), avg(end-start), min(end-start), max(end-start) "
        "from kernels group by name order by sum(end-start) desc").fetchall()
    tot = sum(r[2] for r in rows)
    print(f"| kernel | calls | total ms | mean us | min us | max us | % |")
    print(f"|---|---:|---:|---:|---:|---:|---:|")
    for r in rows[:top]:
        print(f"| {short_name(r[0])} | {r[1]} | {r[2] / 1e6:.3f} | {r[3] / 1e3:.1f} | "
              f"{r[4] / 1e3:.1f} | {r[5] / 1e3:.1f} | {100.0 * r[2] / tot:.1f} |")
    print(f"\nall kernels: {tot / 1e6:.3f} ms over {sum(r[1] for r in rows)} dispatches")


def pmc_means(db: str):
    con = sqlite3.connect(db)
    rows = con.execute("select name, counter_name, count(*), avg(counter_value) from pmc_events "
                       "group by name, counter_name").fetchall()
    out = {}
    for name, counter, n, avg in rows:
        out.setdefault(short_name(name), {})[counter] = (n, avg)
    return out


def pmc(db: str):
    for k, v in sorted(pmc_means(db).items()):
        print(k, {c: round(a, 1) for c, (n, a) in v.items()})


def copy_kib(db: str, counter: str) -> float:
    m = pmc_means(db)
    vals = [v[counter][1] for k, v in m.items() if "copyBuffer" in k and counter in v]
    if not vals:
        raise SystemExit(f"no copy kernel with {counter} in {db}")
    return vals[0]


def traffic(fdb: str, wdb: str, cfdb: str, cwdb: str, mib: float, cells: int = 0,
            aliases=()):
    kib = mib * 1024.0
    fs = kib / copy_kib(cfdb, "FETCH_SIZE")
    ws = kib / copy_kib(cwdb, "WRITE_SIZE")
    f, w = pmc_means(fdb), pmc_means(wdb)
    ks = {}
    for k in sorted(set(f) | set(w)):
        fb = f.get(k, {}).get("FETCH_SIZE", (0, None))[1]
        wb = w.get(k, {}).get("WRITE_SIZE", (0, None))[1]
        if fb is None or wb is None:
            continue
        ks[k] = {"fetch_bytes": fb * 1024.0 * fs, "write_bytes": wb * 1024.0 * ws,
                 "bytes_per_launch": (fb * fs + wb * ws) * 1024.0}
        if cells:
            ks[k]["bytes_per_cell"] = ks[k]["bytes_per_launch"] / cells
    out = {"fetch_scale": fs, "write_scale": ws, "copy_mib": mib}
    named = {}
    for a in aliases:
        name, rp = a.split("=", 1)
        if rp in ks:
            named[name] = dict(ks[rp], rocprof_name=rp)
    if named:
        out["kernels"] = named
    out["all_kernels"] = ks
    print(json.dumps(out, indent=1))


if __name__ == "__main__":
    cmd = sys.argv[1]
    if cmd == "trace":
        trace(sys.argv[2], int(sys.argv[3]) if len(sys.argv) > 3 else 30)
    elif cmd == "pmc":
        pmc(sys.argv[2])
    elif cmd == "traffic":
        traffic(sys.argv[2], sys.argv[3], sys.argv[4], sys.argv[5], float(sys.argv[6]),
                int(sys.argv[7]) if len(sys.argv) > 7 else 0, sys.argv[8:])
    else:
        raise SystemExit(__doc__)
