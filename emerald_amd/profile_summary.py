"""Summaries of rocprofv3 output for the profiles kept with the project.

  python -m emerald_amd.profile_summary stats <dir>
      kernel statistics (calls, total ms, average us, share of GPU time) from
      the *_kernel_stats.csv of a `rocprofv3 --kernel-trace --stats` run.

  python -m emerald_amd.profile_summary traffic <stats_dir> <fetch_dir> <write_dir>
          --particles N [--fetch-scale F] [--write-scale W] [--json out.json --method M]
      HBM bytes per launch of each kernel from two `rocprofv3 --pmc FETCH_SIZE`
      / `--pmc WRITE_SIZE` runs (KiB per dispatch, scaled by the factors a copy
      kernel of known size gives), with the average durations of the stats run:
      bytes per particle, GB/s and the fraction of the 8 TB/s HBM peak.

  python -m emerald_amd.profile_summary pmc <dir> [kernel-substring ...]
      mean value per dispatch of every counter of a --pmc run, per kernel.
"""

from __future__ import annotations

import argparse
import collections
import csv
import glob
import json
import os
import sys

HBM_PEAK_GBS = 8000.0

# Kernel -> the timing class bench.py reports it under, per method.
CLASSES = {
    "k_build": "build",
    "k_iisph_aii_beta0": "iisph_aii_beta",
    "k_iisph_sdp": "iisph_sdp",
    "k_iisph_beta_p": "iisph_beta_p",
    "k_ap_diag0": "ap_diag",
    "k_ap_accel": "ap_accel",
    "k_ap_iterate": "ap_iterate",
    "k_ap_final": "ap_final",
    "k_dfsph_div": "dfsph_div",
    "k_dfsph_dens": "dfsph_dens",
    "k_finalize": "finalize",
    "k_pci_init": "pci_init",
    "k_pci_density": "pci_density",
    "k_pci_force": "pci_force",
    "k_pci_finalize": "pci_finalize",
    "k_keys": "keys",
}


def _csv(d: str, suffix: str) -> list:
    f = glob.glob(os.path.join(d, "**", f"*{suffix}"), recursive=True)
    if not f:
        raise SystemExit(f"no *{suffix} under {d}")
    with open(f[0]) as fh:
        return list(csv.DictReader(fh))


def _short(name: str) -> str:
    name = name.split("(")[0].replace("void ", "").strip()
    if "::" in name:
        name = name.split("::")[-1]
    return name.split("<")[0]


def kernel_stats(d: str) -> dict:
    out = {}
    for r in _csv(d, "kernel_stats.csv"):
        k = _short(r["Name"])
        e = out.setdefault(k, {"calls": 0, "total_ns": 0.0})
        e["calls"] += int(r["Calls"])
        e["total_ns"] += float(r["TotalDurationNs"])
    for e in out.values():
        e["avg_us"] = e["total_ns"] / e["calls"] / 1e3
    return out


def counters(d: str) -> dict:
    acc = collections.defaultdict(lambda: collections.defaultdict(list))
    for r in _csv(d, "counter_collection.csv"):
        acc[_short(r["Kernel_Name"])][r["Counter_Name"]].append(float(r["Counter_Value"]))
    return {k: {c: sum(v) / len(v) for c, v in cs.items()} for k, cs in acc.items()}


def cmd_stats(a) -> None:
    st = kernel_stats(a.dir)
    tot = sum(e["total_ns"] for e in st.values())
    print("| kernel | calls | total ms | average us | % of GPU time |")
    print("|---|---|---|---|---|")
    for k, e in sorted(st.items(), key=lambda kv: -kv[1]["total_ns"]):
        print(f"| {k} | {e['calls']} | {e['total_ns'] / 1e6:.2f} | {e['avg_us']:.1f} | "
              f"{100 * e['total_ns'] / tot:.1f} |")


def cmd_traffic(a) -> None:
    st = kernel_stats(a.stats_dir)
    fetch = counters(a.fetch_dir)
    write = counters(a.write_dir)
    rows = []
    for k, e in st.items():
        if k not in fetch or k not in write:
            continue
        rd = fetch[k].get("FETCH_SIZE", 0.0) * 1024.0 * a.fetch_scale
        wr = write[k].get("WRITE_SIZE", 0.0) * 1024.0 * a.write_scale
        b = rd + wr
        gbs = b / (e["avg_us"] * 1e-6) / 1e9
        rows.append((k, rd, wr, b, e["avg_us"], gbs))
    rows.sort(key=lambda r: -r[4] * st[r[0]]["calls"])
    print("| kernel | HBM read MB | HBM write MB | B / particle | avg us | HBM GB/s | % of 8 TB/s |")
    print("|---|---|---|---|---|---|---|")
    for k, rd, wr, b, us, gbs in rows:
        print(f"| {k} | {rd / 1e6:.1f} | {wr / 1e6:.1f} | {b / a.particles:.0f} | {us:.1f} | "
              f"{gbs:.0f} | {100 * gbs / HBM_PEAK_GBS:.1f} |")
    if a.json:
        data = {}
        if os.path.exists(a.json):
            with open(a.json) as fh:
                data = json.load(fh)
        m = data.setdefault(a.method, {})
        for k, rd, wr, b, us, gbs in rows:
            cls = CLASSES.get(k)
            if cls:
                m[cls] = b
            if k == "k_apply":
                for c in ("apply_integrate", "apply", "apply_gravity"):
                    m[c] = b
        data["note"] = ("HBM bytes per launch of the named kernel class (rocprofv3 FETCH_SIZE/"
                        "WRITE_SIZE, copy-probe scaled); see *_hbm_traffic.md")
        with open(a.json, "w") as fh:
            json.dump(data, fh, indent=1, sort_keys=True)


def cmd_pmc(a) -> None:
    for k, cs in sorted(counters(a.dir).items()):
        if a.kernels and not any(x in k for x in a.kernels):
            continue
        print(k, {c: round(v) for c, v in sorted(cs.items())})


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="python -m emerald_amd.profile_summary")
    sub = ap.add_subparsers(dest="cmd", required=True)
    s = sub.add_parser("stats")
    s.add_argument("dir")
    t = sub.add_parser("traffic")
    t.add_argument("stats_dir")
    t.add_argument("fetch_dir")
    t.add_argument("write_dir")
    t.add_argument("--particles", type=int, required=True)
    t.add_argument("--fetch-scale", type=float, default=2.0)
    t.add_argument("--write-scale", type=float, default=1.0)
    t.add_argument("--json")
    t.add_argument("--method", default="iisph")
    p = sub.add_parser("pmc")
    p.add_argument("dir")
    p.add_argument("kernels", nargs="*")
    a = ap.parse_args(argv)
    {"stats": cmd_stats, "traffic": cmd_traffic, "pmc": cmd_pmc}[a.cmd](a)
    return 0


if __name__ == "__main__":
    sys.exit(main())
