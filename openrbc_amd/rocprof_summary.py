"""Text summaries of rocprofv3 output (`--output-format csv`), as committed
under profiles/.

    python -m openrbc_amd.rocprof_summary trace DIR --title T --command C [--window KERNEL N]
        per-kernel table of a --kernel-trace run (all dispatches, and the
        window from the N-th last dispatch of KERNEL to the end of the last)
    python -m openrbc_amd.rocprof_summary traffic --kernel K --workload W --fetch DIR --write DIR
            --cal-fetch DIR --cal-write DIR --cal-mib M --last N --command C [--command C ...] --out FILE.json
        HBM bytes per launch of kernel K from FETCH_SIZE and WRITE_SIZE runs,
        each scaled by what the counter reads for a plain device copy of M MiB
        (bench.py --copy-calibration M); the last N launches of K (the timing
        run that follows bench.py's timed region).  Writes FILE.json (read by
        bench.py) and FILE.md.

The JSON records the sha256 of the kernel sources it was measured on, so that
bench.py uses the figure only for the build it describes.
"""
import argparse
import collections
import csv
import glob
import hashlib
import json
import os
import re
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# the sources that decide the fused step kernel and the list layout it reads
TRAFFIC_SOURCES = ["openrbc_amd/csrc/k_force.hip", "openrbc_amd/csrc/k_build.hip", "openrbc_amd/csrc/orbc_internal.h"]


def sources_sha256(paths):
    h = hashlib.sha256()
    for p in paths:
        with open(os.path.join(ROOT, p), "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()


def short_name(name: str) -> str:
    s = re.sub(r"\(anonymous namespace\)::|orbc::", "", name)
    s = re.sub(r"^void ", "", s)
    return s if len(s) <= 70 else s[:67] + "..."


def base_name(name: str) -> str:
    s = re.sub(r"\(anonymous namespace\)::|orbc::|^void ", "", name)
    return re.sub(r"[<(].*", "", s)


def _rows(d: str, suffix: str):
    files = sorted(glob.glob(os.path.join(d, "**", f"*{suffix}"), recursive=True))
    if not files:
        raise SystemExit(f"no *{suffix} under {d}")
    out = []
    for f in files:
        with open(f, newline="") as fh:
            out += list(csv.DictReader(fh))
    return out


def trace(args):
    rows = _rows(args.dir, "kernel_trace.csv")
    rows.sort(key=lambda r: int(r["Start_Timestamp"]))

    def table(rs, top):
        agg = collections.defaultdict(lambda: [0, 0.0, 1e30, 0.0])
        for r in rs:
            d = (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3
            a = agg[short_name(r["Kernel_Name"])]
            a[0] += 1; a[1] += d; a[2] = min(a[2], d); a[3] = max(a[3], d)
        tot = sum(a[1] for a in agg.values())
        span = (int(rs[-1]["End_Timestamp"]) - int(rs[0]["Start_Timestamp"])) / 1e3
        lines = [f"{len(rs)} dispatches, GPU busy {tot / 1e3:.2f} ms over a span of {span / 1e3:.2f} ms", "",
                 "| kernel | calls | total ms | avg us | min us | max us | % of busy |",
                 "|---|---:|---:|---:|---:|---:|---:|"]
        for nm, a in sorted(agg.items(), key=lambda kv: -kv[1][1])[:top]:
            lines.append(f"| `{nm}` | {a[0]} | {a[1] / 1e3:.2f} | {a[1] / a[0]:.1f} | {a[2]:.1f} | {a[3]:.1f} | "
                         f"{100 * a[1] / tot:.1f} |")
        return lines

    out = [f"# Kernel trace: {args.title}", "",
           "Command (repository root on the MI355X host, after `cd /tmp && export TMPDIR=/tmp`):", "", "```"]
    out += args.command + ["```", "",
                           f"Summarised with `python -m openrbc_amd.rocprof_summary trace <output dir> ...`.", ""]
    if args.window:
        k, n = args.window[0], int(args.window[1])
        idx = [i for i, r in enumerate(rows) if base_name(r["Kernel_Name"]) == k]
        if len(idx) >= n:
            w = rows[idx[-n]:]
            last_end = int(rows[idx[-1]]["End_Timestamp"])
            w = [r for r in w if int(r["Start_Timestamp"]) <= last_end]
            out += [f"## Window: the last {n} launches of `{k}` and everything between them", ""]
            d = [(int(rows[i]["End_Timestamp"]) - int(rows[i]["Start_Timestamp"])) / 1e3 for i in idx[-n:]]
            ran = [x for x in d if x > 0.1 * max(d)]
            out += [f"`{k}`: {len(ran)} of these launches ran a step (the others found the list spent and "
                    f"returned at once); their average duration is {sum(ran) / len(ran):.1f} us.", ""]
            out += table(w, 40) + [""]
    out += ["## Whole run (set-up, minimisation, autotuning, timed region, timing run), 30 largest", ""]
    out += table(rows, 30)
    print("\n".join(out))


def _counter(d: str, kernel: str, name: str):
    vals = collections.defaultdict(float)
    names = {}
    for r in _rows(d, "counter_collection.csv"):
        if r["Counter_Name"] != name:
            continue
        key = int(r["Dispatch_Id"])
        vals[key] += float(r["Counter_Value"])
        names[key] = r["Kernel_Name"]
    return [(k, vals[k]) for k in sorted(vals) if kernel is None or base_name(names[k]) == kernel]


def traffic(args):
    def ran(vals):
        # launches that ran a step (a launch that found the list spent returns at once)
        return [v for v in vals if v > 0.1 * max(vals)] if vals else []

    fetch = ran([v for _, v in _counter(args.fetch, args.kernel, "FETCH_SIZE")][-args.last:])
    write = ran([v for _, v in _counter(args.write, args.kernel, "WRITE_SIZE")][-args.last:])
    if not fetch or not write:
        raise SystemExit(f"no FETCH_SIZE / WRITE_SIZE dispatches of {args.kernel}")
    # the copy kernel: the largest dispatches of the calibration run
    cf = max(v for _, v in _counter(args.cal_fetch, None, "FETCH_SIZE"))
    cw = max(v for _, v in _counter(args.cal_write, None, "WRITE_SIZE"))
    kib = args.cal_mib * 1024.0
    fs, ws = kib / cf, kib / cw
    f_kib, w_kib = sum(fetch) / len(fetch), sum(write) / len(write)
    hbm = (f_kib * fs + w_kib * ws) * 1024.0
    prof = {
        "kernel": args.kernel, "workload": args.workload, "launches_measured": min(len(fetch), len(write)),
        "fetch_kib_per_launch": f_kib, "write_kib_per_launch": w_kib, "fetch_scale": fs, "write_scale": ws,
        "hbm_bytes_per_launch": hbm,
        "calibration": f"a {args.cal_mib} MiB device-to-device copy reads back as FETCH_SIZE {cf:.0f} KiB and "
                       f"WRITE_SIZE {cw:.0f} KiB",
        "commands": args.command, "sources": TRAFFIC_SOURCES, "sources_sha256": sources_sha256(TRAFFIC_SOURCES),
        "measured_with": args.measured_with,
    }
    with open(args.out, "w") as fh:
        json.dump(prof, fh, indent=1)
    md = os.path.splitext(args.out)[0] + ".md"
    with open(md, "w") as fh:
        fh.write("\n".join([
            f"# HBM traffic of `{args.kernel}` ({args.workload})", "",
            "Commands (repository root on the MI355X host, after `cd /tmp && export TMPDIR=/tmp`):", "", "```",
            *args.command, "```", "",
            f"* launches measured: {prof['launches_measured']} among the last {args.last} of the run (the timing "
            f"run after the timed region, same knobs) that ran a step",
            f"* FETCH_SIZE per launch: {f_kib:,.0f} KiB; WRITE_SIZE per launch: {w_kib:,.0f} KiB",
            f"* calibration: {prof['calibration']}, so reads are scaled by {fs:.3f} and writes by {ws:.3f}",
            f"* HBM traffic per launch: {hbm / 1e9:.3f} GB (reads {f_kib * fs * 1024 / 1e9:.3f} GB, "
            f"writes {w_kib * ws * 1024 / 1e9:.3f} GB)",
            f"* kernel sources: {', '.join(TRAFFIC_SOURCES)} (sha256 {prof['sources_sha256'][:16]}...)",
            f"* run knobs: {args.measured_with}" if args.measured_with else "", ""]))
    print(json.dumps(prof))


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m openrbc_amd.rocprof_summary")
    sub = ap.add_subparsers(dest="what", required=True)
    t = sub.add_parser("trace")
    t.add_argument("dir")
    t.add_argument("--title", required=True)
    t.add_argument("--command", action="append", required=True)
    t.add_argument("--window", nargs=2, metavar=("KERNEL", "N"))
    f = sub.add_parser("traffic")
    for k in ("kernel", "workload", "fetch", "write", "cal_fetch", "cal_write", "out"):
        f.add_argument("--" + k.replace("_", "-"), dest=k, required=True)
    f.add_argument("--cal-mib", type=int, required=True)
    f.add_argument("--last", type=int, required=True)
    f.add_argument("--command", action="append", required=True)
    f.add_argument("--measured-with", dest="measured_with", default="",
                   help="the run's knobs (skin, block halos, ...), recorded beside the figure")
    a = ap.parse_args(argv)
    (trace if a.what == "trace" else traffic)(a)


if __name__ == "__main__":
    sys.exit(main())
