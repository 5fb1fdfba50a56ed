"""Summarise rocprofv3 CSV output of a bench.py run as a markdown table.

    python -m opentbl_amd.prof_summary --trace DIR [--fetch DIR] [--write DIR]
        [--lds DIR] [--cmd "..."] > summary.md

--trace holds the *_kernel_stats.csv of a --kernel-trace --stats run; --fetch
and --write hold the *_counter_collection.csv of --pmc FETCH_SIZE and --pmc
WRITE_SIZE runs; --lds one of --pmc SQ_LDS_BANK_CONFLICT SQ_INSTS_LDS ....
The byte counters are scaled by what they read for the largest plain device
copy in the same run (__amd_rocclr_copyBuffer, bench.py's 1 GiB copy): the
scale is copy bytes / counter bytes, printed with the table.
"""
import argparse
import collections
import csv
import glob
import os
import re


def _short(name):
    """Kernel name without return type, namespace and argument list, with
    its template arguments: "rk_final_y<20, float>"."""
    n = re.sub(r"^void\s+", "", name.strip())
    n = n.replace("otbl::", "")
    depth = 0
    for k, ch in enumerate(n):
        if ch == "<":
            depth += 1
        elif ch == ">":
            depth -= 1
        elif ch == "(" and depth == 0:
            return n[:k].strip()
    return n


def _one(d, pattern):
    hits = sorted(glob.glob(os.path.join(d, "**", pattern), recursive=True))
    if not hits:
        raise SystemExit(f"no {pattern} under {d}")
    return hits[0]


def counters(d):
    """kernel -> counter -> list of per-dispatch values, and the raw rows."""
    rows = list(csv.DictReader(open(_one(d, "*counter_collection.csv"))))
    out = collections.defaultdict(lambda: collections.defaultdict(list))
    for r in rows:
        out[_short(r["Kernel_Name"])][r["Counter_Name"]].append(float(r["Counter_Value"]))
    return out


def copy_scale(cnt, counter, copy_bytes):
    vals = cnt.get("__amd_rocclr_copyBuffer", {}).get(counter, [])
    if not vals:
        return None, None
    kib = max(vals)
    return copy_bytes / (kib * 1024.0), kib


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--trace", required=True)
    ap.add_argument("--fetch")
    ap.add_argument("--write")
    ap.add_argument("--lds")
    ap.add_argument("--copy-bytes", type=float, default=float(1 << 30))
    ap.add_argument("--cmd", action="append", default=[])
    a = ap.parse_args()

    stats = list(csv.DictReader(open(_one(a.trace, "*kernel_stats.csv"))))
    print("# rocprofv3 summary\n")
    for c in a.cmd:
        print(f"    {c}")
    print()
    fe = counters(a.fetch) if a.fetch else {}
    wr = counters(a.write) if a.write else {}
    fs, fk = copy_scale(fe, "FETCH_SIZE", a.copy_bytes) if fe else (None, None)
    ws, wk = copy_scale(wr, "WRITE_SIZE", a.copy_bytes) if wr else (None, None)
    if fs:
        print(f"FETCH_SIZE scale {fs:.3f} (a {a.copy_bytes / 2**30:g} GiB copy read {fk:.0f} KiB)  ")
    if ws:
        print(f"WRITE_SIZE scale {ws:.3f} (a {a.copy_bytes / 2**30:g} GiB copy read {wk:.0f} KiB)\n")
    print("| kernel | calls | avg ms | total ms | % | fetch GB/launch | write GB/launch | HBM GB/s |")
    print("|---|---|---|---|---|---|---|---|")
    for r in stats:
        k = _short(r["Name"])
        avg = float(r["AverageNs"]) / 1e6
        f = w = None
        if fs and "FETCH_SIZE" in fe.get(k, {}):
            v = fe[k]["FETCH_SIZE"]
            f = sum(v) / len(v) * 1024 * fs / 1e9
        if ws and "WRITE_SIZE" in wr.get(k, {}):
            v = wr[k]["WRITE_SIZE"]
            w = sum(v) / len(v) * 1024 * ws / 1e9
        bw = (f + w) / (avg * 1e-3) if f is not None and w is not None and avg > 0 else None
        print(f"| {k} | {r['Calls']} | {avg:.4f} | {float(r['TotalDurationNs']) / 1e6:.3f} | "
              f"{float(r['Percentage']):.2f} | {'' if f is None else f'{f:.3f}'} | "
              f"{'' if w is None else f'{w:.3f}'} | {'' if bw is None else f'{bw:.0f}'} |")
    if a.lds:
        lc = counters(a.lds)
        names = sorted({c for v in lc.values() for c in v})
        print("\n| kernel | " + " | ".join(names) + " |")
        print("|---|" + "---|" * len(names))
        for k, cs in lc.items():
            cells = [f"{sum(cs[n]) / len(cs[n]):.4g}" if n in cs else "" for n in names]
            print(f"| {k} | " + " | ".join(cells) + " |")
        print("\n(counter values averaged over dispatches)")


if __name__ == "__main__":
    main()
