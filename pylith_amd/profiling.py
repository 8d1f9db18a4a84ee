"""Summaries of rocprofv3 result databases (the rocpd SQLite files rocprofv3 writes).

    python -m pylith_amd.profiling stats  <results.db>            per-kernel dispatch statistics
    python -m pylith_amd.profiling counter <results.db> [kernel]  per-kernel counter averages
    python -m pylith_amd.profiling copy                           the calibration workload

The Python-side counterpart of PyLith's profiling helpers (pylith/utils/profiling.py): it reads
what the profiler recorded and prints compact tables.  `copy` is the one command that runs GPU
work: three device-to-device copies of 2^27 doubles (1 GiB read and 1 GiB written each), run
under `rocprofv3 --pmc FETCH_SIZE` (then WRITE_SIZE) to scale those counters to bytes.
"""
import sqlite3
import sys


def kernel_stats(path):
    con = sqlite3.connect(path)
    rows = con.execute(
        "select name, count(*), avg(duration), min(duration), max(duration), max(vgpr_count), "
        "max(accum_vgpr_count), max(sgpr_count), max(lds_size), max(scratch_size), max(workgroup_x), "
        "max(grid_x) from kernels group by name order by sum(duration) desc").fetchall()
    total = sum(r[1] * r[2] for r in rows) or 1.0
    out = []
    for name, n, avg, lo, hi, vgpr, agpr, sgpr, lds, scratch, wg, grid in rows:
        out.append({"kernel": name, "calls": n, "avg_ns": avg, "min_ns": lo, "max_ns": hi,
                    "pct": 100.0 * n * avg / total, "vgpr": vgpr, "agpr": agpr, "sgpr": sgpr,
                    "lds": lds, "scratch": scratch, "workgroup": wg, "grid": grid})
    return out


def counter_stats(path, kernel_like="%"):
    con = sqlite3.connect(path)
    rows = con.execute(
        "select kernel_name, counter_name, count(*), avg(value) from counters_collection "
        "where kernel_name like ? group by kernel_name, counter_name", (kernel_like,)).fetchall()
    return [{"kernel": k, "counter": c, "dispatches": n, "avg": v} for k, c, n, v in rows]


def _short(name, width=70):
    return name if len(name) <= width else name[:width - 3] + "..."


def copy_workload(n=1 << 27, reps=3):
    """Device copies of n doubles (the counters' calibration: bytes moved are known)."""
    import torch
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    print(f"copied {reps} x {8 * n} bytes", flush=True)


def main(argv):
    if len(argv) == 2 and argv[1] == "copy":
        copy_workload()
        return 0
    if len(argv) < 3 or argv[1] not in ("stats", "counter"):
        print(__doc__)
        return 2
    if argv[1] == "stats":
        print("| kernel | calls | avg us | min us | max us | % | VGPR | SGPR | LDS B | scratch B |")
        print("|---|---|---|---|---|---|---|---|---|---|")
        for r in kernel_stats(argv[2]):
            print(f"| {_short(r['kernel'])} | {r['calls']} | {r['avg_ns'] / 1e3:.2f} | "
                  f"{r['min_ns'] / 1e3:.2f} | {r['max_ns'] / 1e3:.2f} | {r['pct']:.1f} | {r['vgpr']} | "
                  f"{r['sgpr']} | {r['lds']} | {r['scratch']} |")
    else:
        like = f"%{argv[3]}%" if len(argv) > 3 else "%"
        print("| kernel | counter | dispatches | average |")
        print("|---|---|---|---|")
        for r in counter_stats(argv[2], like):
            print(f"| {_short(r['kernel'])} | {r['counter']} | {r['dispatches']} | {r['avg']:.6g} |")
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv))
