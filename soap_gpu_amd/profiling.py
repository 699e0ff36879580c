"""Measurement helpers for rocprofv3 runs of the series benchmark.

    python -m soap_gpu_amd.profiling copy
        a plain streaming kernel of known size (1 GiB read + 1 GiB written per
        dispatch, 3 dispatches) to calibrate the FETCH_SIZE / WRITE_SIZE counters;
    python -m soap_gpu_amd.profiling summarize OUT_PREFIX WORKLOAD \\
            --fetch DIR --write DIR --copy-fetch DIR --copy-write DIR
        averages the per-dispatch counters of each kernel from the CSV output of
        `rocprofv3 --pmc FETCH_SIZE` / `--pmc WRITE_SIZE --output-format csv`,
        scales each by (known bytes / counter reading) of the copy kernel and
        writes OUT_PREFIX.json (read by bench.py) and OUT_PREFIX.md;
    python -m soap_gpu_amd.profiling sq OUT.md TITLE --pass COMMAND DIR [--pass COMMAND DIR ...]
            [--kernels k1,k2]
        per-dispatch averages of the SQ counters of each pass (`rocprofv3 --pmc
        SQ_... --output-format csv`) for the named kernels, with the FP64 FMA
        share of the VALU instructions and the SALU / VALU ratio;
    python -m soap_gpu_amd.profiling stats OUT.md TITLE COMMAND DB [--top N] [--match REGEX]
        kernel statistics (calls, average / min / max microseconds, share of the
        total) from the SQLite output of `rocprofv3 --kernel-trace --stats`,
        written as a markdown table headed by the title and the command.
"""
from __future__ import annotations

import csv
import glob
import json
import os
import re
import sys

COPY_BYTES = 1 << 30
COPY_KERNEL_HINT = "mulfunctor"


def copy_kernel(reps: int = 3):
    import torch
    n = COPY_BYTES // 8
    x = torch.ones(n, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    for _ in range(reps):
        torch.mul(x, 1.0, out=y)
    torch.cuda.synchronize()
    print(f"copy kernel: {reps} dispatches of {COPY_BYTES} bytes read + {COPY_BYTES} bytes written", flush=True)


def _counter_rows(d):
    files = glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)
    if not files:
        raise SystemExit(f"no counter_collection.csv under {d}")
    rows = []
    for f in files:
        with open(f) as fh:
            rows += list(csv.DictReader(fh))
    return rows


def per_kernel(d, counter):
    """{kernel name: mean counter value per dispatch (KiB)}"""
    acc = {}
    for r in _counter_rows(d):
        if r.get("Counter_Name") != counter:
            continue
        name = r.get("Kernel_Name", "")
        key = (name, r.get("Dispatch_Id"))
        acc.setdefault(name, {}).setdefault(key, 0.0)
        acc[name][key] += float(r["Counter_Value"])
    return {k: sum(v.values()) / len(v) for k, v in acc.items()}


def _short(name):
    for k in ("k_fir_tile_dpp", "k_fir_prog", "k_fir_tile", "k_fir_active_grp", "k_conv_active", "k_norm_ccf", "k_norm",
              "k_ccf_reduce", "k_ccf_lds",
              "k_fir_quiet_pipe", "k_fir_quiet_lds", "k_conv_quiet", "k_taps_active", "k_taps_quiet", "k_minmax_reduce", "k_fit",
              "k_mask_build", "k_mask_overwrite", "k_mask_wsum", "k_mask_footprint", "k_disk_init",
              "k_taps_finalize", "k_fill_u64",
              "k_gather_active_linear", "k_gather_quiet_cols", "k_gather_quiet_linear", "k_quiet_columns_maps",
              "k_lin_groups", "k_cells_count", "k_cells_write", "k_conv_linear"):
        if (name.endswith(k) or re.search(r"soap::" + k + r"[(<]", name)
                or ("4soap" + str(len(k)) + k + "E") in name or ("4soap" + str(len(k)) + k + "I") in name):
            return k
    return name


def summarize(out_prefix, workload, fetch, write, copy_fetch, copy_write):
    cf = {k: v for k, v in per_kernel(copy_fetch, "FETCH_SIZE").items() if COPY_KERNEL_HINT in k.lower()}
    cw = {k: v for k, v in per_kernel(copy_write, "WRITE_SIZE").items() if COPY_KERNEL_HINT in k.lower()}
    if not cf or not cw:
        raise SystemExit("copy kernel not found in the calibration runs")
    known_kib = COPY_BYTES / 1024.0
    f_fetch = known_kib / max(cf.values())
    f_write = known_kib / max(cw.values())
    fe = per_kernel(fetch, "FETCH_SIZE")
    wr = per_kernel(write, "WRITE_SIZE")
    kern = {}
    for name in sorted(set(fe) | set(wr)):
        s = _short(name)
        fb = fe.get(name, 0.0) * 1024 * f_fetch
        wb = wr.get(name, 0.0) * 1024 * f_write
        kern[s] = {"fetch_kib_raw": round(fe.get(name, 0.0), 1), "write_kib_raw": round(wr.get(name, 0.0), 1),
                   "fetch_bytes": fb, "write_bytes": wb, "traffic_bytes": fb + wb}
    summary = {"workload": workload, "copy_factor_fetch": f_fetch, "copy_factor_write": f_write,
               "copy_kernel": list(cf)[0], "kernels": kern}
    with open(out_prefix + ".json", "w") as f:
        json.dump(summary, f, indent=1)
    lines = [f"# HBM traffic per dispatch (rocprofv3 PMC), {workload}", "",
             "Counters: FETCH_SIZE and WRITE_SIZE (KiB per dispatch), one rocprofv3 pass each,",
             f"scaled by the copy calibration (1 GiB read + 1 GiB written): fetch x{f_fetch:.4f}, "
             f"write x{f_write:.4f}.", "",
             "| kernel | FETCH_SIZE KiB | WRITE_SIZE KiB | HBM bytes (scaled) |", "|---|---:|---:|---:|"]
    for k, v in sorted(kern.items(), key=lambda kv: -kv[1]["traffic_bytes"]):
        lines.append(f"| {k} | {v['fetch_kib_raw']:.1f} | {v['write_kib_raw']:.1f} | {v['traffic_bytes']:.4g} |")
    with open(out_prefix + ".md", "w") as f:
        f.write("\n".join(lines) + "\n")
    print("\n".join(lines))


def sq_summary(out_md, title, passes, kernels):
    """passes: [(command, dir)]; kernels: short names (None: every kernel seen)"""
    data = {}
    for _, d in passes:
        acc = {}
        for r in _counter_rows(d):
            key = (_short(r.get("Kernel_Name", "")), r["Counter_Name"])
            acc.setdefault(key, []).append(float(r["Counter_Value"]))
        for (k, c), v in acc.items():
            data.setdefault(k, {})[c] = sum(v) / len(v)
    names = kernels or sorted(data)
    counters = sorted({c for k in names for c in data.get(k, {})})
    lines = [f"# rocprofv3 SQ counters per dispatch: {title}", "", "Commands (repository root on the MI355X box), one pass each:", ""]
    lines += [f"    {cmd}" for cmd, _ in passes]
    lines += ["", "| counter | " + " | ".join(f"`{k}`" for k in names) + " |",
              "|---|" + "---:|" * len(names)]
    for c in counters:
        lines.append(f"| {c} | " + " | ".join(f"{data.get(k, {}).get(c, float('nan')):.4g}" for k in names) + " |")
    derived = []
    for k in names:
        d = data.get(k, {})
        v, f, sa = d.get("SQ_INSTS_VALU"), d.get("SQ_INSTS_VALU_FMA_F64"), d.get("SQ_INSTS_SALU")
        parts = []
        if v and f is not None:
            parts.append(f"FP64 FMA {100.0 * f / v:.0f} % of VALU")
        if v and sa is not None:
            parts.append(f"SALU / VALU {sa / v:.2f}")
        if parts:
            derived.append(f"- `{k}`: " + ", ".join(parts))
    if derived:
        lines += ["", *derived]
    with open(out_md, "w") as fh:
        fh.write("\n".join(lines) + "\n")
    print("\n".join(lines))


def stats(out_md, title, command, db, top=25, match=None):
    import sqlite3
    con = sqlite3.connect(db)
    rows = con.execute("select name, count(*), avg(duration), min(duration), max(duration), sum(duration) "
                       "from kernels group by name order by sum(duration) desc").fetchall()
    if match:
        rows = [r for r in rows if re.search(match, r[0])]
    total = sum(r[5] for r in rows) or 1.0
    lines = [f"# rocprofv3 kernel statistics: {title}", "", "Command (repository root on the MI355X box):", "",
             f"    {command}", "",
             "| kernel | calls | average us | min us | max us | total % |", "|---|---:|---:|---:|---:|---:|"]
    for name, n, avg, mn, mx, tot in rows[:top]:
        short = name.split("(")[0]
        lines.append(f"| `{short}` | {n} | {avg / 1e3:.1f} | {mn / 1e3:.1f} | {mx / 1e3:.1f} | "
                     f"{100.0 * tot / total:.2f} |")
    with open(out_md, "w") as f:
        f.write("\n".join(lines) + "\n")
    print("\n".join(lines))


def main(argv):
    if not argv:
        raise SystemExit(__doc__)
    if argv[0] == "copy":
        copy_kernel()
    elif argv[0] == "summarize":
        import argparse
        ap = argparse.ArgumentParser()
        ap.add_argument("out_prefix")
        ap.add_argument("workload")
        ap.add_argument("--fetch", required=True)
        ap.add_argument("--write", required=True)
        ap.add_argument("--copy-fetch", required=True)
        ap.add_argument("--copy-write", required=True)
        a = ap.parse_args(argv[1:])
        summarize(a.out_prefix, a.workload, a.fetch, a.write, a.copy_fetch, a.copy_write)
    elif argv[0] == "sq":
        import argparse
        ap = argparse.ArgumentParser()
        ap.add_argument("out_md")
        ap.add_argument("title")
        ap.add_argument("--pass", dest="passes", nargs=2, action="append", metavar=("COMMAND", "DIR"), required=True)
        ap.add_argument("--kernels", default=None)
        a = ap.parse_args(argv[1:])
        sq_summary(a.out_md, a.title, [tuple(x) for x in a.passes], a.kernels.split(",") if a.kernels else None)
    elif argv[0] == "stats":
        import argparse
        ap = argparse.ArgumentParser()
        ap.add_argument("out_md")
        ap.add_argument("title")
        ap.add_argument("command")
        ap.add_argument("db")
        ap.add_argument("--top", type=int, default=25)
        ap.add_argument("--match", default=None, help="regular expression on kernel names")
        a = ap.parse_args(argv[1:])
        stats(a.out_md, a.title, a.command, a.db, a.top, a.match)
    else:
        raise SystemExit(__doc__)


if __name__ == "__main__":
    main(sys.argv[1:])
