#!/usr/bin/env python3
"""Benchmark of the D2Q9 BGK time step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--algorithm nptl]
                    [--layout collision] [--storage f64] [--sweep OUT.json]

Workload: the 8192 x 2048 porous channel (seeded non-overlapping discs,
radii U[4, 16), porosity 0.7), tau 0.6, inlet density 1.05, outlet 1.0.
The default line is nptl in FP64 in its fastest layout on this workload,
the collision layout (the sweep times every algorithm x layout x storage).
A step is one fused time step of the whole lattice with the inputs resident
in HBM.  The unit of work is a fluid-node update; the line reports MFLUPS
(10^6 fluid-node updates per second), the kernel's achieved HBM bandwidth
against the MI355X peak (algorithmic bytes: 2 x 9 x sizeof(storage) per
fluid node), and the CPU oracle timed on the host cores as a baseline.

Prints one JSON line.  Optional studies, each written to its own JSON file
before the line is printed:
  --sweep OUT            every algorithm x layout x storage combination
                         (--sweep-cpu-seconds S: the CPU oracle beside each
                         algorithm x layout for about S seconds)
  --precision-study OUT  FP32/FP16/BF16 storage against FP64 after a fixed
                         number of steps (--study-scenario, --study-taus,
                         --study-steps, --study-average)
  --steady-study OUT     the same at the steady state of a 1024x256 channel
  --batched OUT          512 uneven channels: per-domain launches, one
                         persistent launch per step, autotuned graph replay
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

import numpy as np  # noqa: E402

import portlbm_amd as pa  # noqa: E402
from portlbm_amd import workloads  # noqa: E402

HBM_PEAK_GBPS = 8000.0
INNER_H, INNER_V = 8192, 2048
TRAFFIC_FILE = ROOT / "profiles" / "traffic.json"


def workload(algorithm: str, layout: str, storage: str, steps: int, scenario: str = "porous-circles") -> dict:
    wg = 1024 if algorithm in ("npol", "nsol") else 256
    return pa.settings(algorithm, layout, inner_h=INNER_H, inner_v=INNER_V, scenario=scenario,
                       work_group_size=wg, time_steps=steps, tau=0.6, rho_in=1.05, rho_out=1.0,
                       storage=storage, seed=20261202, porosity=0.7, radiusMin=4.0, radiusMax=16.0,
                       obstacleMargin=32)


def time_gpu(doc: dict, steps: int, warmup: int) -> dict:
    lat = pa.Lattice.from_settings(doc)
    try:
        fluid = int((lat.node_codes() & pa.CODE_FLUID).astype(bool).sum())
        if warmup:
            lat.run(warmup, 0)
        lat.sync()
        t0 = time.perf_counter()
        dev_ms = lat.run(steps, 0)  # HIP events on the lattice's stream around the step launches
        lat.sync()
        wall_ms = (time.perf_counter() - t0) * 1e3
        red = lat.reduce()
        return {"fluid": fluid, "dev_ms": dev_ms, "wall_ms": wall_ms, "mass": red["mass"]}
    finally:
        lat.close()


def storage_bytes(storage: str) -> int:
    return pa.STORAGE_BYTES[storage]


def cpu_baseline(algorithm: str, layout: str, budget_s: float = 15.0) -> dict:
    """The CPU oracle (the reference's pass structure: bounce-back emplacement,
    outlet, stream-collide with macroscopic outputs every step) on the same
    workload, OpenMP over rows on the host cores, for a bounded number of steps."""
    from oracle import binding as ob

    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    cores = max(1, min(cores, 16))
    ob.set_threads(cores)
    doc = workload(algorithm, layout, "f64", 1)
    sim = ob.Simulation(doc)
    try:
        fluid = int((sim.phase() == 0).sum())
        sim.run(1)  # warm-up
        t0 = time.perf_counter()
        sim.run(1)
        one = time.perf_counter() - t0
        steps = max(1, min(200, int(budget_s / max(one, 1e-3))))
        t0 = time.perf_counter()
        sim.run(steps)
        dt = time.perf_counter() - t0
    finally:
        sim.close()
    return {"value": round(fluid * steps / dt / 1e6, 2), "unit": "MFLUPS", "cores": cores, "kind": "port",
            "sample": f"oracle {algorithm}/{layout} FP64 on the same 8192x2048 porous channel, "
                      f"{steps} steps after 1 warm-up step, {dt:.1f} s"}


def measured_traffic(algorithm: str, layout: str, storage: str, scenario: str = "porous-circles") -> float | None:
    """HBM bytes per launch from the committed rocprofv3 PMC summary, if it
    covers this kernel and workload (the summaries are of the porous
    channel)."""
    if scenario != "porous-circles":
        return None
    try:
        data = json.loads(TRAFFIC_FILE.read_text())
    except (OSError, ValueError):
        return None
    rec = data.get(f"{algorithm}/{layout}/{storage}/{INNER_H}x{INNER_V}")
    return None if rec is None else float(rec["bytes_per_launch"])


def kernel_of(algorithm: str, layout: str, storage: str) -> str:
    """The step kernel the library launches for a cell of the default
    workload (work-group 1024 for npol / nsol: 32 x 32 and 30 x 31 tiles,
    the staged block kernel in their FP64/FP32 collision layout, the tile
    sweep elsewhere; 256 for lptl's 1-D kernel).
    lptl runs nptl's kernels where those are faster: the staged kernel in the
    collision layout and the FP64/FP32 bundle layout, the row kernel in
    16-bit storage; its own 1-D kernel only in the FP64/FP32 stream layout."""
    sixteen = storage in ("f16-shifted", "bf16-shifted")
    if algorithm in ("nptl", "lptl"):
        if layout == "collision" or (layout == "bundle" and not sixteen):
            return "k_step_nptl_staged"
        if algorithm == "lptl" and not sixteen:
            return "k_step_lptl"
        return "k_step_nptl"
    if layout == "collision" and not sixteen:
        return "k_step_tile_staged"
    return "k_step_tile_sweep"


def sweep(out_path: str, steps: int, warmup: int, cpu_budget_s: float = 0.0) -> None:
    """Every algorithm x layout x storage on the default workload.  With
    cpu_budget_s > 0 the CPU oracle (FP64, the reference's pass structure) is
    timed beside each algorithm x layout for about that many seconds."""
    cpu = {}
    if cpu_budget_s > 0:
        for algorithm in ("nptl", "lptl", "npol", "nsol"):
            for layout in ("collision", "stream", "bundle"):
                cpu[(algorithm, layout)] = cpu_baseline(algorithm, layout, cpu_budget_s)
                print(json.dumps({"cpu": [algorithm, layout], **cpu[(algorithm, layout)]}), file=sys.stderr,
                      flush=True)
    rows = []
    for storage in ("f64", "f32", "f16-shifted", "bf16-shifted"):
        for algorithm in ("nptl", "lptl", "npol", "nsol"):
            for layout in ("collision", "stream", "bundle"):
                r = time_gpu(workload(algorithm, layout, storage, steps), steps, warmup)
                mflups = r["fluid"] * steps / (r["dev_ms"] * 1e-3) / 1e6
                gbps = mflups * 1e6 * 2 * 9 * storage_bytes(storage) / 1e9
                row = {"algorithm": algorithm, "layout": layout, "storage": storage,
                       "kernel": kernel_of(algorithm, layout, storage), "mflups": round(mflups, 1),
                       "mlups_all": round(INNER_H * INNER_V * steps / (r["dev_ms"] * 1e-3) / 1e6, 1),
                       "gbps": round(gbps, 1), "hbm_frac": round(gbps / HBM_PEAK_GBPS, 4),
                       "ms_per_step": round(r["dev_ms"] / steps, 4)}
                if (algorithm, layout) in cpu:
                    c = cpu[(algorithm, layout)]
                    row["cpu_mflups_fp64"] = c["value"]
                    row["cpu_cores"] = c["cores"]
                rows.append(row)
                print(json.dumps(row), file=sys.stderr, flush=True)
    Path(out_path).write_text(json.dumps({"workload": f"porous-circles {INNER_H}x{INNER_V}", "steps": steps,
                                          "rows": rows}, indent=1))


def precision_study(out_path: str, steps: int, taus=(0.52, 0.6, 0.8), scenario: str = "cylinder",
                    average: int = 0, sample_every: int = 100) -> None:
    """Reduced-precision storage against FP64 on the 8192 x 2048 cylinder
    channel (disc r = 64 at x = 1024, centred), inlet 1.02 / outlet 1.0, or
    (scenario "porous") on the porous channel of the default workload.
    Errors are over fluid nodes, relative to the FP64 run of the same tau:
    L2 = ||a - a64||_2 / ||a64||_2 and Linf = max|a - a64| / max|a64| for the
    density and the momentum vector j; mass drift = sum rho(T) / sum rho(0) - 1
    (FP64 sums of the populations on the GPU); drag = momentum-exchange force
    on the solids.  With average > 0 the fields and the drag are averaged over
    the last `average` steps (a sample every `sample_every` steps): the
    cylinder's wake sheds vortices, and instantaneous fields of two runs drift
    out of phase.  A run is finite when the FP64 sums of its fluid
    populations are (the reported fields clamp NaN to 0); a tau whose FP64 run
    is not finite is reported as FP64-unstable, without error figures."""
    rows = []
    for tau in taus:
        ref = None
        for variant in ("f64", "f32", "f16-shifted", "f16-shifted/stochastic", "bf16-shifted",
                        "bf16-shifted/stochastic"):
            storage, _, rounding = variant.partition("/")
            if scenario == "porous":
                doc = workloads.porous_channel(storage=storage, steps=steps, tau=tau)
            else:
                doc = workloads.cylinder_channel(tau, storage=storage, steps=steps)
            doc["mi355x"]["rounding"] = rounding or "nearest"
            with pa.Lattice.from_settings(doc) as lat:
                m0 = lat.reduce()["mass"]
                window = min(average, steps)
                ms = lat.run(steps - window, 0) if steps > window else 0.0
                acc, fx, fy, samples = None, 0.0, 0.0, 0
                if window:
                    done = 0
                    while done < window:
                        k = min(sample_every, window - done)
                        ms += lat.run(k, 0)
                        done += k
                        fields = np.stack(lat.macros()[:3])
                        red = lat.reduce()
                        acc = fields if acc is None else acc + fields
                        fx, fy, samples = fx + red["fx"], fy + red["fy"], samples + 1
                    rho, jx, jy = acc / samples
                    fx, fy = fx / samples, fy / samples
                red = lat.reduce()
                if not window:
                    rho, jx, jy, _ = lat.macros()
                    fx, fy = red["fx"], red["fy"]
            finite = bool(np.isfinite([red["mass"], red["jx"], red["jy"]]).all())
            fluid = rho > -1
            row = {"tau": tau, "storage": storage, "rounding": rounding or "nearest", "steps": steps,
                   "averaged_steps": window, "fluid_nodes": int(fluid.sum()),
                   "mflups": round(int(fluid.sum()) * steps / (ms * 1e-3) / 1e6, 1),
                   "mass_drift": red["mass"] / m0 - 1.0, "drag_x": fx, "drag_y": fy, "finite": finite}
            if storage == "f64":
                ref = (rho, jx, jy, fx) if finite else None
                if not finite:
                    row["fp64_unstable"] = True
            elif ref is not None and finite:
                r64, jx64, jy64, fx64 = ref
                f = fluid & (r64 > -1)
                dj = np.hypot(jx[f] - jx64[f], jy[f] - jy64[f])
                j64 = np.hypot(jx64[f], jy64[f])
                row.update({
                    "rho_l2": float(np.linalg.norm(rho[f] - r64[f]) / np.linalg.norm(r64[f])),
                    "rho_linf": float(np.abs(rho[f] - r64[f]).max() / np.abs(r64[f]).max()),
                    "u_l2": float(np.linalg.norm(dj) / np.linalg.norm(j64)),
                    "u_linf": float(dj.max() / j64.max()),
                    "drag_x_rel": float(abs(fx - fx64) / abs(fx64)),
                })
            rows.append(row)
            print(json.dumps(row), file=sys.stderr, flush=True)
    workload = (f"porous-circles {INNER_H}x{INNER_V} (seed 20261202, radii U[4,16), porosity 0.7), "
                "rho_in 1.05 / rho_out 1.0" if scenario == "porous" else
                f"cylinder r=64 at x=1024 in {INNER_H}x{INNER_V}, rho_in 1.02 / rho_out 1.0")
    Path(out_path).write_text(json.dumps({"workload": workload, "steps": steps, "averaged_steps": average,
                                          "rows": rows}, indent=1))


def steady_precision_study(out_path: str, inner_h: int = 1024, inner_v: int = 256, taus=(0.6, 0.8),
                           tol: float = 1e-10, max_steps: int = 400000, block: int = 1000) -> None:
    """Reduced-precision storage against FP64 at the steady state of a porous
    channel (porous-circles, radii U[4,16), porosity 0.7, seed 20261202,
    inlet 1.05 / outlet 1.0, inner_h x inner_v nodes).  FP64 runs in blocks
    of `block` steps until the residual over the last 100 steps of a block,
    ||j(t) - j(t-100)|| / ||j(t)||, falls below `tol`; every other storage then
    runs the same number of steps.  Errors as in precision_study, plus each
    run's final residual and the first block end at which its residual was
    below 1e-6."""
    rows = []
    for tau in taus:
        ref = None
        total = None
        for variant in ("f64", "f32", "f16-shifted", "f16-shifted/stochastic", "bf16-shifted",
                        "bf16-shifted/stochastic"):
            storage, _, rounding = variant.partition("/")
            doc = workloads.porous_channel(storage=storage, steps=max_steps, tau=tau, inner_h=inner_h,
                                           inner_v=inner_v)
            doc["mi355x"]["rounding"] = rounding or "nearest"
            with pa.Lattice.from_settings(doc) as lat:
                m0 = lat.reduce()["mass"]
                steps, res, first_1e6, ms = 0, 1.0, None, 0.0
                while (steps < total) if total is not None else (res >= tol and steps < max_steps):
                    ms += lat.run(block - 100, 0)
                    lat.residual()
                    ms += lat.run(100, 0)
                    res = lat.residual()
                    steps += block
                    if first_1e6 is None and res < 1e-6:
                        first_1e6 = steps
                red = lat.reduce()
                rho, jx, jy, _ = lat.macros()
            fluid = rho > -1
            row = {"tau": tau, "storage": storage, "rounding": rounding or "nearest", "steps": steps,
                   "fluid_nodes": int(fluid.sum()), "residual": res, "steps_to_1e-6": first_1e6,
                   "mflups": round(int(fluid.sum()) * steps / (ms * 1e-3) / 1e6, 1),
                   "mass_drift": red["mass"] / m0 - 1.0, "drag_x": red["fx"]}
            if storage == "f64":
                ref, total = (rho, jx, jy, red), steps
            else:
                r64, jx64, jy64, red64 = ref
                f = fluid & (r64 > -1)
                dj = np.hypot(jx[f] - jx64[f], jy[f] - jy64[f])
                j64 = np.hypot(jx64[f], jy64[f])
                row.update({
                    "rho_l2": float(np.linalg.norm(rho[f] - r64[f]) / np.linalg.norm(r64[f])),
                    "rho_linf": float(np.abs(rho[f] - r64[f]).max() / np.abs(r64[f]).max()),
                    "u_l2": float(np.linalg.norm(dj) / np.linalg.norm(j64)),
                    "u_linf": float(dj.max() / j64.max()),
                    "drag_x_rel": float(abs(red["fx"] - red64["fx"]) / abs(red64["fx"])),
                })
            rows.append(row)
            print(json.dumps(row), file=sys.stderr, flush=True)
    workload = (f"porous-circles {inner_h}x{inner_v} (seed 20261202, radii U[4,16), porosity 0.7), "
                f"rho_in 1.05 / rho_out 1.0, run to the FP64 steady state (residual < {tol:g})")
    Path(out_path).write_text(json.dumps({"workload": workload, "rows": rows}, indent=1))


def batched(out_path: str, steps: int, count: int = 512,
            modes=("per-domain", "persistent", "graph")) -> dict:
    """512 uneven porous channels in FP32 (portlbm_amd.workloads.batch_channels):
    per-domain launches vs one persistent launch per step vs the whole run in
    one hipGraph with autotuned tiles (`modes` picks which are timed)."""
    docs = workloads.batch_channels(count, steps=steps)
    rows = {}
    with pa.Batch(docs, storage="f32") as b:
        fluid = b.fluid_nodes()
        nodes = sum((d["domain"]["horizontalNodes"]) * (d["domain"]["verticalNodes"]) for d in docs)
        for mode in ("per-domain", "persistent"):
            if mode not in modes:
                continue
            b.init()
            ms = b.run(steps, mode)
            rows[mode] = {"ms": round(ms, 2), "mflups": round(fluid * steps / (ms * 1e-3) / 1e6, 1)}
            print(json.dumps({mode: rows[mode]}), file=sys.stderr, flush=True)
        if "graph" in modes:
            tune_ms = b.autotune(steps_per_trial=20)
            b.run(steps, "graph")  # capture + first replay
            b.init()
            ms = b.run(steps, "graph")
            rows["graph+autotune"] = {"ms": round(ms, 2), "mflups": round(fluid * steps / (ms * 1e-3) / 1e6, 1),
                                      "tile_widths_by_class": b.tiles(),
                                      "autotune_best_ms_per_step": round(tune_ms, 4)}
            print(json.dumps({"graph+autotune": rows["graph+autotune"]}), file=sys.stderr, flush=True)
    result = {"workload": f"{count} porous channels nx U[64,1024] ny U[32,256], FP32, {steps} steps",
              "domains": count, "inner_nodes": nodes, "fluid_nodes": fluid, "modes": rows}
    if "per-domain" in rows and "graph+autotune" in rows:
        result["speedup_graph_vs_per_domain"] = round(rows["per-domain"]["ms"] / rows["graph+autotune"]["ms"], 2)
    if "graph+autotune" in rows:
        result["gbps_graph"] = round(rows["graph+autotune"]["mflups"] * 1e6 * 72 / 1e9, 1)
    Path(out_path).write_text(json.dumps(result, indent=1))
    return result


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--algorithm", default="nptl", choices=sorted(pa.ALGORITHMS))
    ap.add_argument("--layout", default="collision", choices=sorted(pa.LAYOUTS))
    ap.add_argument("--storage", default="f64", choices=sorted(pa.STORAGES))
    ap.add_argument("--scenario", default="porous-circles",
                    help="porous-circles (the benchmark workload) or any other scenario, for comparison")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--sweep", metavar="OUT.json")
    ap.add_argument("--sweep-cpu-seconds", type=float, default=0.0,
                    help="time the CPU oracle beside each algorithm x layout of the sweep for about this long")
    ap.add_argument("--precision-study", metavar="OUT.json")
    ap.add_argument("--study-scenario", default="cylinder", choices=("cylinder", "porous"))
    ap.add_argument("--study-taus", default="0.52,0.6,0.8")
    ap.add_argument("--study-steps", type=int, default=20000)
    ap.add_argument("--study-average", type=int, default=None,
                    help="average fields and drag over the last N steps (default: 5000 for the cylinder, 0 for porous)")
    ap.add_argument("--steady-study", metavar="OUT.json",
                    help="reduced-precision storage against FP64 at the steady state of a 1024x256 porous channel")
    ap.add_argument("--batched", metavar="OUT.json")
    ap.add_argument("--batch-steps", type=int, default=2000)
    ap.add_argument("--batch-modes", default="per-domain,persistent,graph",
                    help="comma list of per-domain, persistent, graph")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(multi-GPU sharding is a design note only).", file=sys.stderr)
        return 2

    if args.sweep:
        sweep(args.sweep, min(args.steps, 500), min(args.warmup, 20), args.sweep_cpu_seconds)
    if args.precision_study:
        taus = tuple(float(t) for t in args.study_taus.split(","))
        average = args.study_average
        if average is None:
            average = 5000 if args.study_scenario == "cylinder" else 0
        precision_study(args.precision_study, args.study_steps, taus, args.study_scenario, average)
    if args.steady_study:
        steady_precision_study(args.steady_study)
    if args.batched:
        batched(args.batched, args.batch_steps, modes=tuple(args.batch_modes.split(",")))

    doc = workload(args.algorithm, args.layout, args.storage, args.steps, args.scenario)
    r = time_gpu(doc, args.steps, args.warmup)
    secs = r["dev_ms"] * 1e-3
    mflups = r["fluid"] * args.steps / secs / 1e6
    bytes_per_node = 2 * 9 * storage_bytes(args.storage)
    launch_ms = r["dev_ms"] / args.steps  # one kernel launch per step for nptl/lptl
    achieved = r["fluid"] * bytes_per_node / (launch_ms * 1e-3) / 1e9
    line = {
        "metric": "MLUPS & % HBM roofline at 8192x2048; steps-to-steady warm-start vs plain",
        "value": round(mflups, 1),
        "unit": "MFLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r["wall_ms"] / args.steps, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"f64": "f64", "f32": "f32", "f16-shifted": "f16-storage/f32", "bf16-shifted":
                  "bf16-storage/f32"}[args.storage],
        "data": "synthetic",
        "config": {"workload": (f"porous-circles channel {INNER_H}x{INNER_V}, discs r~U[4,16) porosity 0.7 "
                                f"seed 20261202, tau 0.6, rho_in 1.05 / rho_out 1.0"
                                if args.scenario == "porous-circles" else
                                f"{args.scenario} channel {INNER_H}x{INNER_V}, tau 0.6, rho_in 1.05 / rho_out 1.0"),
                   "algorithm": args.algorithm, "layout": args.layout, "storage": args.storage,
                   "fluid_nodes": r["fluid"], "inner_nodes": INNER_H * INNER_V},
        "mlups_all_nodes": round(INNER_H * INNER_V * args.steps / secs / 1e6, 1),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBPS, 4),
                     "traffic": measured_traffic(args.algorithm, args.layout, args.storage, args.scenario)},
        "kernel_ms": round(launch_ms, 5),
        "bytes_per_fluid_node": bytes_per_node,
        "mass": r["mass"],
    }
    try:
        line["copy_bandwidth_gbps"] = round(pa.copy_bandwidth_gbps(1 << 30, 5), 1)
    except pa.LbmError:
        line["copy_bandwidth_gbps"] = None
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(args.algorithm, args.layout)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
