"""Summaries of rocprofv3 CSV output (`--output-format csv`) for the tables
under profiles/.

    python -m coffee_sim_amd.profile_summary trace DIR SUBSTEPS
        per-kernel calls, average and total device time per substep over the
        last SUBSTEPS substeps of DIR's kernel trace (a substep is counted from
        its k_p2g_bin dispatch, less the two clears launched before it)
    python -m coffee_sim_amd.profile_summary counters DIR SUBSTEPS
        per-kernel counter values per dispatch (FETCH_SIZE / WRITE_SIZE in
        KiB) averaged over the dispatches of the last SUBSTEPS substeps
"""
from __future__ import annotations

import argparse
import collections
import csv
import glob
import sys


def _one(directory: str, suffix: str) -> str:
    paths = sorted(glob.glob(f"{directory}/**/*{suffix}", recursive=True))
    if not paths:
        raise SystemExit(f"no *{suffix} under {directory}")
    return paths[0]


def kernel_name(raw: str) -> str:
    k = raw.replace("void ", "").replace("(anonymous namespace)::", "")
    k = k.split("(")[0]
    return k if k.startswith("k_rbgs_multi") else k.split("<")[0]


def _window(rows: list[dict], substeps: int, marker: str) -> list[dict]:
    rows.sort(key=lambda r: int(r["Start_Timestamp"]))
    idx = [i for i, r in enumerate(rows) if kernel_name(r["Kernel_Name"]) == marker]
    if len(idx) < substeps:
        raise SystemExit(f"only {len(idx)} {marker} dispatches for {substeps} substeps")
    return rows[max(idx[-substeps] - 2, 0):]


def trace(directory: str, substeps: int, marker: str) -> None:
    rows = list(csv.DictReader(open(_one(directory, "kernel_trace.csv"))))
    sel = _window(rows, substeps, marker)
    tot: dict[str, float] = collections.defaultdict(float)
    cnt: collections.Counter = collections.Counter()
    for r in sel:
        k = kernel_name(r["Kernel_Name"])
        tot[k] += (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3
        cnt[k] += 1
    span = (int(sel[-1]["End_Timestamp"]) - int(sel[0]["Start_Timestamp"])) / 1e3
    busy = sum(tot.values())
    print(f"substeps {substeps}; dispatches {len(sel)}; kernel busy {busy / substeps / 1e3:.3f} ms/substep; "
          f"span {span / substeps / 1e3:.3f} ms/substep (includes the profiler's per-dispatch gaps)")
    print(f"{'kernel':40s} {'calls/substep':>13s} {'avg us':>9s} {'us/substep':>11s} {'share':>6s}")
    for k, v in sorted(tot.items(), key=lambda kv: -kv[1]):
        print(f"{k[:40]:40s} {cnt[k] / substeps:13.1f} {v / cnt[k]:9.2f} {v / substeps:11.1f} {100 * v / busy:5.1f}%")
    mg = [k for k in tot if k.startswith("k_mg_") or k.startswith("k_residual")]
    print(f"multigrid start (k_mg_* + residual tests): {sum(cnt[k] for k in mg) / substeps:.1f} dispatches, "
          f"{sum(tot[k] for k in mg) / substeps:.1f} us/substep, {100 * sum(tot[k] for k in mg) / busy:.1f}%")


def counters(directory: str, substeps: int) -> None:
    rows = list(csv.DictReader(open(_one(directory, "counter_collection.csv"))))
    # one value per dispatch and counter (summed over the rows rocprofv3 splits it into)
    per: dict[tuple[str, str, str], float] = collections.defaultdict(float)
    names: dict[str, str] = {}
    for r in rows:
        d = r["Dispatch_Id"]
        names[d] = kernel_name(r["Kernel_Name"])
        per[(d, r["Counter_Name"], names[d])] += float(r["Counter_Value"])
    dispatches = sorted({int(d) for d, _, _ in per})
    counters_seen = sorted({c for _, c, _ in per})
    acc: dict[tuple[str, str], list[float]] = collections.defaultdict(list)
    for (d, c, k), v in per.items():
        acc[(k, c)].append(v)
    kernels = sorted({k for k, _ in acc})
    print(f"dispatches {len(dispatches)}; counters {', '.join(counters_seen)}; values per dispatch, "
          f"averaged (KiB for the *_SIZE counters)")
    print(f"{'kernel':32s} {'calls/substep':>13s} " + " ".join(f"{c:>14s}" for c in counters_seen))
    for k in kernels:
        n = max(len(acc[(k, c)]) for c in counters_seen)
        vals = []
        for c in counters_seen:
            xs = acc.get((k, c), [])
            vals.append(f"{(sum(xs) / len(xs) if xs else float('nan')):14.1f}")
        print(f"{k[:32]:32s} {n / substeps:13.1f} " + " ".join(vals))


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("mode", choices=("trace", "counters"))
    ap.add_argument("directory")
    ap.add_argument("substeps", type=int)
    ap.add_argument("--marker", default="k_p2g_bin", help="the kernel that opens a substep")
    a = ap.parse_args()
    if a.mode == "trace":
        trace(a.directory, a.substeps, a.marker)
    else:
        counters(a.directory, a.substeps)
    return 0


if __name__ == "__main__":
    sys.exit(main())
