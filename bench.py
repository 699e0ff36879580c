"""Benchmark: phase-spectra/s of the full SOAP-GPU forward series on one MI355X.

Workload (synthetic, seeded): grid 1000x1000, lambda 3800-7000 A log-uniform at
0.25 km/s (732,588 pixels), 8000 lines, 20 mu bins, 2 spots + 2 faculae
(overlapping), 100 phases, R=115000 resolution, 4000-line mask CCF on 161
velocities, FP64 throughout.

One step = the whole series on device (inputs resident in HBM): disk grid,
quiet star (normalised + SED), active-region deltas of all 100 phases,
instrumental convolution + normalisation of 3x100 + 2 spectra, mask CCF of
303 spectra and the Gaussian fit + bisector of 303 CCFs.  value = 100 x steps
/ wall time.

Prints one JSON line (rank 0, N=1).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(w, eng_maps, frac: float = 0.04, nthreads: int | None = None):
    """Oracle (FP64 C, OpenMP) on a bounded sample of the same workload: phase 0's
    active-region integration + resolution of its three spectra on a contiguous
    lambda sub-range of `frac` of the pixels (quiet star, CCF and fit excluded),
    scaled to the full grid.  Returns phase-spectra/s."""
    import oracle as O
    n = w.n
    m = int(n * frac)
    j0 = (n - m) // 2
    sl = slice(j0, j0 + m)
    sub = {k: np.ascontiguousarray(v[:, sl]) for k, v in w.cubes.items()}
    wave = np.ascontiguousarray(w.wave[sl])
    nt = nthreads or int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    a, s = w.act, w.star
    maps = O.disk_init(s["grid"], s["prot"], 0.0, s["incl"], s["radius_km"], s["u1"], s["u2"], s["lb"], w.mu_bounds)
    t0 = time.perf_counter()
    temp, bi = O.phase_maps(s["grid"], s["incl"], a["nrho"], a["t_star"], a["t_diff_spot"], a["t_diff_fac"],
                            w.phases[0], w.regions[0])
    raw = O.integrate_phase(O.GRID_LOG, wave, maps, temp, bi, sub, s["grid"], a["t_star"], a["t_diff_fac"],
                            nthreads=nt)
    q = np.full(m, 6.6e5)
    for t in range(3):
        O.lower_resolution(O.GRID_LOG, wave, w.sigma_inst, q, raw[t], 1, nthreads=nt)
    dt = time.perf_counter() - t0
    per_phase = dt / frac
    return dict(value=1.0 / per_phase, unit="phase-spectra/s", cores=nt, kind="port",
                sample=f"oracle on phase 0: integration + resolution of 3 spectra over {m} of {n} pixels "
                       f"({frac:.0%}), {dt:.2f} s measured, scaled x{1 / frac:.1f}; quiet star, CCF, fit excluded",
                seconds=dt)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--phases", type=int, default=100)
    ap.add_argument("--grid", type=int, default=1000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-frac", type=float, default=0.04)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU only (phases shard across GPUs by design, "
              "but no multi-GPU run is part of it); use --gpus 1", file=sys.stderr)
        sys.exit(2)

    from soap_gpu_amd import synth
    from soap_gpu_amd.engine import Engine, setup_workload

    t0 = time.time()
    w = synth.config1(n_phase=args.phases, grid=args.grid)
    log(f"[bench] workload built in {time.time() - t0:.1f} s: n={w.n} grid={args.grid} phases={len(w.phases)}")
    eng = Engine(0)
    setup_workload(eng, w)
    log("[bench] inputs resident on device")

    def step():
        eng.run_series(w.star, w.sigma_inst)

    for i in range(args.warmup):
        step()
        eng.sync()
        log(f"[bench] warmup {i + 1}/{args.warmup} done")
    eng.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    eng.sync()
    dt = time.perf_counter() - t0
    ms_step = dt / args.steps * 1e3
    value = len(w.phases) * args.steps / dt
    log(f"[bench] timed {args.steps} steps: {ms_step:.3f} ms/step")

    # per-stage HIP-event timing over the same steps, on the engine's stream
    eng.set_timing(True)
    for _ in range(args.steps):
        step()
    st = eng.stage_ms()
    eng.set_timing(False)
    per = {k: v / args.steps for k, v in st.items()}
    alg_bytes, phase_bytes, rows = eng.integrate_stats()
    fir_ms = per["fir_active"]
    achieved = alg_bytes / (fir_ms * 1e-3) / 1e9
    log("[bench] stage ms/step: " + ", ".join(f"{k}={v:.3f}" for k, v in per.items()))
    log(f"[bench] fir_active: {alg_bytes / 1e6:.1f} MB minimum traffic (union of rows + outputs), "
        f"{phase_bytes / 1e6:.1f} MB counted per phase, sum R_psi = {rows}")

    cpu = None
    if not args.no_cpu_baseline:
        maps = None
        cpu = cpu_baseline(w, maps, frac=args.cpu_frac)
        log(f"[bench] cpu baseline {cpu['value']:.4f} phase-spectra/s ({cpu['seconds']:.1f} s)")
        cpu.pop("seconds")

    out = {
        "metric": "phase-spectra/s (1000^2 grid, 733k-pt lambda, FP64)",
        "value": round(value, 3),
        "unit": "phase-spectra/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded lines, log-uniform grid; no reference seed spectra exist)",
        "config": {"workload": "config1: grid 1000, 732588-pt log lambda 3800-7000 A, 8000 lines, 20 mu bins, "
                               "2 spots + 2 faculae, 100 phases, R=115000, 4000-line mask, 161 velocities",
                   "grid": args.grid, "n_lambda": w.n, "phases": len(w.phases)},
        "roofline": {"kernel": "k_fir_active", "bound": "hbm", "achieved": round(achieved, 1),
                     "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": round(achieved / PEAK_HBM_GBS, 4),
                     "traffic": None, "alg_bytes_per_launch": alg_bytes, "bytes_per_phase_sum": phase_bytes,
                     "rows_sum": rows, "launch_ms": round(fir_ms, 4)},
        "stage_ms": {k: round(v, 4) for k, v in per.items()},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    eng.close()


if __name__ == "__main__":
    main()
