"""Benchmark of the coupled SPH + wave step on one MI355X.

Headline workload: 1 048 576 particles + 2048^2 heightfield dam break, the
at-scale configuration; one "step" is one full coupled step (grid build,
density, forces + integration, particle -> wave flux, wave EVOLVE).
Inputs are resident in HBM before timing starts; the timed region is K
steps launched as replays of a 6-step hipGraph captured before it (K mod 6
steps eagerly), bracketed by device synchronisation, repeated from the
initial state (value: the median repeat).

Prints one JSON line: particle-updates/s (value), wave cell-updates/s, the
roofline of the dominant kernel (per-kernel average device time from HIP
events recorded on the scene's stream; HBM bytes per launch from the
rocprofv3 FETCH_SIZE / WRITE_SIZE summary in profiles/traffic.json), the
measured device-copy ceiling, the forward+adjoint / forward ratio, the
batched 2D form (256 scenes, tuned, graph-replayed), the batched 3D
parameter fit (64 scenes, forward + adjoint in one hipGraph per iteration)
and the CPU oracle (OpenMP) timed on bounded samples of the same workloads.

`--write-traffic FETCH_CSV WRITE_CSV` turns two rocprofv3 --pmc counter
collections of a bench run into profiles/traffic.json.
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

PEAK_HBM_GBS = 8000.0

# Algorithmic bytes per unit of work (fp32, SoA float4): each array that
# crosses a kernel boundary counted once written and once read, neighbour
# re-reads cache-served.
BYTES_PER_PARTICLE = {
    "hash": 20,              # pos 16 + key 4
    "reorder": 8,            # key 4 + index 4
    "density": 24,           # pos 16 + rho, p 8
    "force_integrate": 104,  # read pos, vel, F, rho/p 56; write pos, vel, F 48
}
BYTES_PER_CELL = 8          # scan: read count + write offset
BYTES_PER_WAVE_CELL = 12    # read h^n, h^(n-1); write h^(n+1)

# Batched 2D step (SoA: xv float4, rp float2, idf int, q float4, aj float2).
BYTES_PER_PARTICLE_2D = {
    "hash": 24,      # read xv; write key, run record
    "scatter": 8,    # read key, run record
    "reorder": 64,   # read key, record, xv, rp, idf; write xv, rp, idf
    "density": 68,   # read xv, rp, idf; write q, rp, aj, idf, flux
    "force": 44,     # read q, aj, idf; write xv
}
BYTES_PER_WAVE_CELL_2D = 16  # read and write both levels (substeps stay in LDS)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=30)
    ap.add_argument("--workload", choices=["large", "small"], default="large")
    ap.add_argument("--eager", action="store_true", help="launch kernels without hipGraph")
    ap.add_argument("--profile-steps", type=int, default=0,
                    help="steps of the per-kernel event pass (0: the same K as timed)")
    ap.add_argument("--cpu-steps", type=int, default=80,
                    help="timed oracle steps of the CPU baseline, about 10 s on 16 host "
                         "cores (0 disables it)")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="0: every core in the affinity mask, bounded by OMP_NUM_THREADS "
                         "when set")
    ap.add_argument("--adjoint-steps", type=int, default=256,
                    help="steps of the forward+adjoint measurement (0 disables it)")
    ap.add_argument("--batched-steps", type=int, default=500,
                    help="steps of the batched 2D measurement (0 disables it)")
    ap.add_argument("--batched-tuning", default=None, metavar="JSON",
                    help="fixed launch shapes of the batched form instead of the autotuner "
                         "(keys as the tuner reports them), e.g. for counter runs whose "
                         "collection perturbs the tuner's timings")
    ap.add_argument("--batched-cpu-steps", type=int, default=16,
                    help="oracle steps of the batched CPU baseline, about 10 s on 16 host "
                         "cores (0 disables it)")
    ap.add_argument("--fit-iters", type=int, default=50,
                    help="gradient-descent iterations of the batched fit (0 disables it)")
    ap.add_argument("--fit-scenes", type=int, default=64)
    ap.add_argument("--repeats", type=int, default=5,
                    help="timed repeats of the K steps (value: the median)")
    ap.add_argument("--copy-mib", type=int, default=2048,
                    help="size of the device copy that measures the HBM ceiling (0: skip)")
    ap.add_argument("--write-traffic", nargs=2, metavar=("FETCH_CSV", "WRITE_CSV"),
                    help="derive profiles/traffic.json from two rocprofv3 --pmc counter "
                         "collections (FETCH_SIZE, WRITE_SIZE) of a bench run, then exit")
    return ap.parse_args()


# rocprofv3 kernel name prefix -> (workload key, kernel key of kernel_ms)
KERNEL_KEYS = [
    ("cwa::b2d::k2_density<", "batched", "density"), ("cwa::b2d::k2_force<", "batched", "force"),
    ("cwa::b2d::k2_hash(", "batched", "hash"), ("cwa::b2d::k2_scatter(", "batched", "scatter"),
    ("cwa::b2d::k2_reorder(", "batched", "reorder"), ("cwa::b2d::k2_wave(", "batched", "wave"),
    ("cwa::k_density(", "plain", "density"), ("cwa::k_force_integrate(", "plain", "force_integrate"),
    ("cwa::k_hash_count(", "plain", "hash"), ("cwa::k_scatter_runs(", "plain", "scatter"),
    ("cwa::k_stable_reorder(", "plain", "reorder"), ("cwa::k_wave4(", "plain", "wave"),
]


def kernel_sources_digest() -> str:
    """sha256 over the HIP sources and headers of the library, in name
    order: traffic.json records the digest of the code it was measured on,
    and a bench run of other code reports no traffic figure."""
    import hashlib

    h = hashlib.sha256()
    src = ROOT / "coupledwateranimation_amd" / "csrc"
    for p in sorted(list(src.glob("*.hip")) + list(src.glob("*.hpp"))):
        h.update(p.name.encode())
        h.update(p.read_bytes())
    return h.hexdigest()


def stored_traffic(workload: str, kernel: str):
    """(bytes per launch, note) of `kernel` from profiles/traffic.json, or
    (None, reason) when absent or measured on other kernel sources."""
    tf = ROOT / "profiles" / "traffic.json"
    if not tf.exists():
        return None, "no profiles/traffic.json"
    try:
        d = json.loads(tf.read_text())
        if d.get("_sources") != kernel_sources_digest():
            return None, "profiles/traffic.json was measured on other kernel sources"
        return d[workload][kernel]["bytes"], ("rocprofv3 FETCH_SIZE + WRITE_SIZE passes of "
                                              "these kernel sources (profiles/traffic.json)")
    except (ValueError, KeyError, TypeError):
        return None, f"no entry for {kernel} in profiles/traffic.json"


def write_traffic(args) -> int:
    """HBM bytes per launch from rocprofv3 FETCH_SIZE and WRITE_SIZE (KiB per
    dispatch), each scaled by what the same counter reads for the calibration
    copy (measure_copy: `--copy-mib` MiB read and written), averaged over
    every dispatch of each kernel."""
    import csv
    from collections import defaultdict

    def load(path, counter):
        per = defaultdict(list)
        for r in csv.DictReader(open(path)):
            if r["Counter_Name"] == counter:
                per[r["Kernel_Name"]].append(float(r["Counter_Value"]))
        return per

    fetch_csv, write_csv = args.write_traffic
    fetch, write = load(fetch_csv, "FETCH_SIZE"), load(write_csv, "WRITE_SIZE")
    copy_kib = args.copy_mib * 1024.0
    # the calibration copy: the dispatch with the largest FETCH_SIZE
    cal = max(fetch, key=lambda k: max(fetch[k]))
    f_fetch = copy_kib / max(fetch[cal])
    cal_w = max(write, key=lambda k: max(write[k]))
    f_write = copy_kib / max(write[cal_w])
    plain = "dam_break_1048576_2048"
    batched = "batched 2D dam-break, 256 scenes"
    out = {"_calibration": {"kernel": cal[:80], "copy_kib": copy_kib,
                            "fetch_factor": f_fetch, "write_factor": f_write},
           "_sources": kernel_sources_digest()}
    for name in fetch:
        for prefix, wl, key in KERNEL_KEYS:
            if name.startswith(prefix) or name.startswith("void " + prefix):
                fk = sum(fetch[name]) / len(fetch[name])
                wk = sum(write.get(name, [0.0])) / max(1, len(write.get(name, [0.0])))
                w = out.setdefault(plain if wl == "plain" else batched, {})
                w[key] = {"bytes": (fk * f_fetch + wk * f_write) * 1024.0,
                          "fetch_kib": fk, "write_kib": wk, "dispatches": len(fetch[name])}
    dst = ROOT / "profiles" / "traffic.json"
    dst.write_text(json.dumps(out, indent=1) + "\n")
    print(f"wrote {dst}")
    return 0


def measure_copy(mib: int) -> dict:
    """HBM ceiling on this box: a device-to-device copy of `mib` MiB (well
    above the 256 MB last-level cache), read + write bytes over the median
    of 5 copies.  The same copy is the calibration kernel of the rocprofv3
    FETCH_SIZE / WRITE_SIZE runs."""
    import torch

    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    ts = []
    for _ in range(5):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        b.copy_(a)
        e1.record()
        torch.cuda.synchronize()
        ts.append(e0.elapsed_time(e1))
    ts.sort()
    ms = ts[len(ts) // 2]
    del a, b
    torch.cuda.empty_cache()
    return {"bytes": 2 * n * 4, "ms": ms, "gbs": 2 * n * 4 / (ms * 1e-3) / 1e9}


def measure_fit(args) -> dict:
    """Batched forward+adjoint parameter fits: `--fit-scenes` 3D scenes
    (N_s log-uniform in [8k, 128k], heightfield sides uniform in
    [128, 1024], seed 4), 128 steps checkpointed every 16, all scenes'
    forward, loss and adjoint in one hipGraph replayed per iteration; six
    scalars per scene fitted toward seeded target heightfields by L-BFGS
    with per-scene step control (a proposal that does not lower the loss is
    rejected and the step halved), the attenuation as log(1 - atten)."""
    import numpy as np

    from coupledwateranimation_amd import fit

    t0 = time.perf_counter()
    specs, targets, factors = fit.fit_workload(nscenes=args.fit_scenes)
    t_setup = time.perf_counter() - t0
    t0 = time.perf_counter()
    run = fit.FitRun(specs, targets)
    t_capture = time.perf_counter() - t0
    res = run.descend(iters=args.fit_iters)
    truth = fit.true_theta(run, factors)
    err0 = fit.theta_errors(run.theta0, truth)
    err1 = fit.theta_errors(res.best_theta, truth)
    run.close()
    steps = specs[0].steps
    n = np.array([sp.n for sp in specs])
    evals = np.array(res.eval_s[1:] if len(res.eval_s) > 1 else res.eval_s)
    t_eval = float(np.median(evals))
    ratio = res.accepted[-1] / np.maximum(res.accepted[0], 1e-300)
    per_scene = res.tape_bytes + res.scene_bytes
    names = fit.THETA_NAMES
    ident = [names.index("wave_speed"), names.index("atten")]
    print(f"bench: fit {len(specs)} scenes, {int(n.sum())} particles, {args.fit_iters} iterations: "
          f"eval {1e3 * t_eval:.1f} ms, final/initial loss median {np.median(ratio):.3e}, "
          f"max {ratio.max():.3e}", flush=True)
    return {
        "workload": f"{len(specs)} 3D scenes, 8k-128k particles, 128^2-1024^2 heightfields, "
                    f"{steps} steps, checkpoint every 16",
        "scenes": len(specs), "particles": int(n.sum()), "steps": steps,
        "iterations": args.fit_iters, "setup_s": t_setup, "capture_s": t_capture,
        "eval_ms": 1e3 * t_eval,
        "fwd_adjoint_particle_steps_per_s": float(n.sum()) * steps / t_eval,
        "loss_ratio_final": {"median": float(np.median(ratio)), "max": float(ratio.max()),
                             "min": float(ratio.min())},
        "rejected_proposals": {"median": float(np.median(res.rejected)),
                               "max": int(res.rejected.max())},
        "param_error": {
            "per_scalar_initial": {k: float(err0[:, i].mean()) for i, k in enumerate(names)},
            "per_scalar_final": {k: float(err1[:, i].mean()) for i, k in enumerate(names)},
            "identifiable_initial": float(err0[:, ident].mean()),
            "identifiable_final": float(err1[:, ident].mean()),
            "note": "mean over scenes of |theta / theta_true - 1| (atten: of 1 - atten); "
                    "identifiable = wave_speed and atten, the scalars that shape the "
                    "heightfield directly"},
        "hbm_bytes_per_scene": {"median": float(np.median(per_scene)),
                                "max": float(per_scene.max()), "total": float(per_scene.sum())},
    }


TUNING_KEYS = ("block_density", "block_force", "grid_density", "grid_force", "xcd_queues",
               "wave_block")


def fixed_tuning(text: str) -> dict:
    """--batched-tuning: a JSON object of launch shapes with the autotuner's
    keys (its ms_per_step, if copied along, is dropped); unknown keys or
    non-integer values are an error."""
    d = json.loads(text)
    if not isinstance(d, dict):
        raise ValueError("--batched-tuning takes a JSON object")
    d.pop("ms_per_step", None)
    bad = [k for k in d if k not in TUNING_KEYS]
    if bad:
        raise ValueError(f"--batched-tuning: unknown keys {bad}; known: {list(TUNING_KEYS)}")
    if not all(isinstance(v, int) for v in d.values()):
        raise ValueError("--batched-tuning: values must be integers")
    return d


def measure_batched(args) -> dict:
    """The batched 2D form: 256 independent 2D SPH + 1D-wave scenes (N_s
    log-uniform in [1k, 64k], columns log-uniform in [256, 4096] cells),
    tuned by the library's autotuner, K steps replayed from a hipGraph from
    the initial state with inputs resident; per-kernel device time from HIP
    events over the same trajectory; the CPU oracle (OpenMP over scenes) on
    a bounded sample of the same workload."""
    import numpy as np

    from coupledwateranimation_amd import batch2d

    steps = args.batched_steps
    spec = batch2d.batch_c3(steps=steps)
    n, nc = spec.n_total, spec.wave_total
    cell_sub = int((spec.wave_len * spec.wave_substeps).sum())
    b = spec.make()
    t_tune = time.perf_counter()
    if args.batched_tuning:
        b.set_tuning(**fixed_tuning(args.batched_tuning))
        tuning = b.tuning()
    else:
        tuning = b.tune(spec.dt, 3)
    t_tune = time.perf_counter() - t_tune
    print(f"bench: batched {spec.nscenes} scenes, {n} particles, {nc} wave cells; "
          f"{'fixed' if args.batched_tuning else 'tuned'} in {t_tune:.2f} s: {tuning}",
          flush=True)
    b.step(spec.dt, 5, graph=True)  # warm-up (graph instantiated, caches warm)
    times = []
    for _ in range(max(1, min(args.repeats, 3))):
        b.reinit(spec.xv, spec.enabled, spec.h0, spec.h1)
        b.sync()
        t0 = time.perf_counter()
        b.step(spec.dt, steps, graph=True)
        b.sync()
        times.append(time.perf_counter() - t0)
    t = sorted(times)[len(times) // 2]
    g = b.slots()
    finite = bool(np.isfinite(g["xv"]).all() and np.isfinite(g["h0"]).all())
    # per-kernel times over the same trajectory
    b.reinit(spec.xv, spec.enabled, spec.h0, spec.h1)
    acc = {}
    for _ in range(steps):
        for k, v in b.step_profiled(spec.dt).items():
            acc[k] = acc.get(k, 0.0) + v
    kernel_ms = {k: v / steps for k, v in acc.items()}
    b.close()
    alg = {k: BYTES_PER_PARTICLE_2D[k] * n for k in BYTES_PER_PARTICLE_2D}
    alg["wave"] = BYTES_PER_WAVE_CELL_2D * nc + 4 * n
    alg["scan"] = BYTES_PER_CELL * (spec.nscenes * g["nx"] * g["ny"])
    dominant = max(kernel_ms, key=lambda k: kernel_ms[k])
    achieved = alg[dominant] / (kernel_ms[dominant] * 1e-3) / 1e9
    step_bytes = sum(alg.values())
    ms_step = 1e3 * t / steps
    print(f"bench: batched {steps} steps in {t:.3f} s ({ms_step:.4f} ms/step); kernels "
          + ", ".join(f"{k} {v:.4f}" for k, v in kernel_ms.items()), flush=True)
    cpu = None
    if args.batched_cpu_steps > 0:
        from oracle import oracle2d

        threads = cpu_threads(args)
        o = oracle2d.OracleBatch(spec, threads=threads)
        c0 = time.perf_counter()
        o.step(spec.dt, args.batched_cpu_steps)
        ct = time.perf_counter() - c0
        o.close()
        cpu = {"value": n * args.batched_cpu_steps / ct, "unit": "particle-updates/s",
               "cores": threads, "kind": "port", "cpu": cpu_info(),
               "sample": f"{args.batched_cpu_steps} step(s) of the same 256-scene batch from "
                         f"its initial state, fp32 oracle, OpenMP over scenes, {ct:.2f} s"}
        print(f"bench: batched CPU oracle {threads} threads: {ct:.2f} s", flush=True)
    traffic, traffic_src = stored_traffic(spec.name, dominant)
    return {
        "workload": spec.name, "scenes": spec.nscenes, "particles": n, "wave_cells": nc,
        "steps": steps, "dt": spec.dt, "ms_per_step": ms_step,
        "repeats_ms_per_step": [1e3 * x / steps for x in times],
        "particle_updates_per_s": n * steps / t,
        "wave_cell_updates_per_s": cell_sub * steps / t,
        "kernel_ms": kernel_ms, "tuning": tuning, "tune_s": t_tune,
        "step_bytes": step_bytes, "step_gbs": step_bytes / (ms_step * 1e-3) / 1e9,
        "roofline": {"bound": "hbm", "kernel": dominant, "achieved": achieved,
                     "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": achieved / PEAK_HBM_GBS,
                     "traffic": traffic, "traffic_source": traffic_src,
                     "alg_bytes": alg[dominant]},
        "cpu_baseline": cpu, "finite": finite,
    }


def search_stats(p, pos, nb) -> dict:
    """Neighbour-search figures of the state `pos` (n x 4, any order) on
    the grid of `p`: candidates per particle of the density pass's column
    search (numpy float32 restatement of its trimmed column ranges, without
    the float margin), the per-wave cost (for 64 consecutive cell-sorted
    particles, the longest candidate range of each of the 9 columns, summed:
    what one wave's loop runs), and the neighbours per particle the device
    found (`nb`, self included)."""
    import numpy as np

    f32 = np.float32
    dims = [int(d) for d in p.grid_dims]
    gmin = [f32(g) for g in p.grid_min]
    cell = f32(p.cell_size)
    cy = f32(cell / f32(p.grid_ydiv))
    T = f32(p.h) * f32(p.h)

    def axis(v, g, c, d):
        return np.clip(np.floor((v - g) / c), 0, d - 1).astype(np.int64)

    x, y, z = (pos[:, a].astype(f32) for a in range(3))
    ix, iy, iz = axis(x, gmin[0], cell, dims[0]), axis(y, gmin[1], cy, dims[1]), \
        axis(z, gmin[2], cell, dims[2])
    def column(cx, cz):  # the device's z-tiled column order (16-column tiles)
        t = cz // 16
        w = np.minimum(16, dims[2] - t * 16)
        return t * 16 * dims[0] + cx * w + (cz - t * 16)

    key = column(ix, iz) * dims[1] + iy
    order = np.argsort(key, kind="stable")
    offset = np.concatenate([[0], np.cumsum(np.bincount(key, minlength=dims[0] * dims[1] *
                                                        dims[2]))])
    x, y, z, ix, iz = x[order], y[order], z[order], ix[order], iz[order]
    n = len(x)
    total = np.zeros(n, np.int64)
    pad = (-n) % 64
    wave_cost = 0
    for dxc in (-1, 0, 1):
        for dzc in (-1, 0, 1):
            cx, cz = ix + dxc, iz + dzc
            ok = (cx >= 0) & (cx < dims[0]) & (cz >= 0) & (cz < dims[2])
            dx = np.where(cx < ix, x - (gmin[0] + (cx + 1).astype(f32) * cell),
                          np.where(cx > ix, (gmin[0] + cx.astype(f32) * cell) - x, f32(0)))
            dz = np.where(cz < iz, z - (gmin[2] + (cz + 1).astype(f32) * cell),
                          np.where(cz > iz, (gmin[2] + cz.astype(f32) * cell) - z, f32(0)))
            d2 = dx * dx + dz * dz
            ok &= d2 < T
            ry = np.sqrt(np.maximum(T - d2, f32(0)))
            ylo, yhi = axis(y - ry, gmin[1], cy, dims[1]), axis(y + ry, gmin[1], cy, dims[1])
            cxc, czc = np.clip(cx, 0, dims[0] - 1), np.clip(cz, 0, dims[2] - 1)
            base = column(cxc, czc) * dims[1]
            ln = np.where(ok, offset[base + yhi + 1] - offset[base + ylo], 0)
            total += ln
            wave_cost += int(np.concatenate([ln, np.zeros(pad, np.int64)]).reshape(-1, 64)
                             .max(axis=1).sum())
    return {"candidates_per_particle": float(total.mean()),
            "wave_loop_per_64": wave_cost / ((n + pad) / 64),
            "neighbours_per_particle": {"mean": float(nb.mean()), "max": int(nb.max())},
            "note": "after the timed steps; candidates from a host restatement of the "
                    "density pass's column search, neighbours (self included) as the "
                    "device found them"}


def cpu_info() -> dict:
    """Host cores this process may use and the CPU model (/proc/cpuinfo)."""
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"affinity_cores": len(os.sched_getaffinity(0)), "model": model}


def cpu_threads(args) -> int:
    """Threads of the CPU baselines: --cpu-threads, else every core in the
    affinity mask, bounded by OMP_NUM_THREADS when the environment sets it
    (the host's share of cores per GPU)."""
    if args.cpu_threads:
        return args.cpu_threads
    n = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS", "")
    return min(n, int(env)) if env.isdigit() and int(env) > 0 else n


def measure_adjoint(args) -> dict:
    """fwd+adjoint / fwd on the differentiable workload: 65 536 particles +
    512^2 heightfield, checkpoint every 16 steps.  fwd = the same steps
    without a tape, launched as hipGraph replays (the ratio's denominator)
    and eagerly; fwd+adjoint = checkpointed forward + adjoint sweep (segment
    recompute included) for an MSE height loss."""
    import numpy as np

    from coupledwateranimation_amd import configs
    from coupledwateranimation_amd.gradient import Tape

    spec = configs.dam_break_diff(steps=args.adjoint_steps)
    steps, every = args.adjoint_steps, 16
    scene = spec.make_scene()
    fwd = {}
    for mode in ("eager", "graph", "graph"):     # the first graph run captures
        scene.reinit(spec.pos, spec.vel, None, spec.wave_size, spec.height0, spec.height1)
        scene.step(spec.dt, 2, graph=mode == "graph")  # warm-up (kernels loaded, caches warm)
        scene.reinit(spec.pos, spec.vel, None, spec.wave_size, spec.height0, spec.height1)
        scene.sync()
        t0 = time.perf_counter()
        scene.step(spec.dt, steps, graph=mode == "graph")
        scene.sync()
        fwd[mode] = time.perf_counter() - t0
    t_fwd = fwd["graph"]
    target = np.zeros((spec.wave_size[1], spec.wave_size[0]), np.float32)
    scene.reinit(spec.pos, spec.vel, None, spec.wave_size, spec.height0, spec.height1)
    tape = Tape(scene, steps, every)
    scene.sync()
    t0 = time.perf_counter()
    tape.forward(spec.dt)
    h = scene.readback()["height0"]
    grad, _, _ = tape.backward(dheight0=2.0 * (h - target) / h.size)
    t_fa = time.perf_counter() - t0
    tape.close()
    scene.close()
    print(f"bench: adjoint {steps} steps: fwd {1e3 * t_fwd:.2f} ms (graph), "
          f"{1e3 * fwd['eager']:.2f} ms (eager), fwd+adjoint {1e3 * t_fa:.2f} ms", flush=True)
    return {"workload": spec.name, "steps": steps, "ckpt_every": every,
            "fwd_ms": 1e3 * t_fwd, "fwd_eager_ms": 1e3 * fwd["eager"],
            "fwd_adjoint_ms": 1e3 * t_fa, "ratio": t_fa / t_fwd,
            "ratio_vs_eager_fwd": t_fa / fwd["eager"],
            "note": "ratio = fwd+adjoint / graph-replayed forward",
            "finite": bool(np.isfinite(list(grad.values())).all())}


def main() -> int:
    args = parse()
    if args.write_traffic:
        return write_traffic(args)
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: the coupled step runs on a single device "
              "(multi-GPU sharding is a design note, not implemented); use --gpus 1",
              file=sys.stderr)
        return 2

    import numpy as np

    import coupledwateranimation_amd as pkg
    from coupledwateranimation_amd import configs

    if pkg.lib().cwa_device_count() < 1:
        print("no HIP device visible", file=sys.stderr)
        return 1
    spec = configs.dam_break_large() if args.workload == "large" else configs.dam_break_small()
    n = spec.n
    W, H = spec.wave_size
    p = spec.params
    ncell = p.grid_dims[0] * p.grid_dims[1] * p.grid_dims[2]
    scene = spec.make_scene()
    graph = not args.eager

    print(f"bench: {spec.name}: {n} particles, {W}x{H} wave, {ncell} grid cells", flush=True)
    import torch  # noqa: F401  (device synchronisation on both sides)

    # The same trajectory `repeats` times: back to the initial state, W
    # untimed warm-up steps, the step graph captured (no step runs), then K
    # timed steps; the value is the median.
    times = []
    for r in range(max(1, args.repeats)):
        if r:
            scene.reinit(spec.pos, spec.vel, None, spec.wave_size, spec.height0, spec.height1)
        scene.step(spec.dt, args.warmup, graph=graph)
        if graph:  # capture the step graph here, outside the timed steps
            scene.prepare_graph(spec.dt)
        scene.sync()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        scene.step(spec.dt, args.steps, graph=graph)
        scene.sync()
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
    t = sorted(times)[len(times) // 2]
    ms_step = 1e3 * t / args.steps
    print(f"bench: timed {args.steps} steps x {len(times)} repeats, median {t:.3f} s "
          f"({ms_step:.4f} ms/step), min {1e3 * min(times) / args.steps:.4f}, "
          f"max {1e3 * max(times) / args.steps:.4f}", flush=True)

    st = scene.readback()
    finite = bool(np.isfinite(st["pos"]).all() and np.isfinite(st["height0"]).all())
    state_stats = {"rho_min": float(st["rho"].min()), "rho_max": float(st["rho"].max()),
                   "rho_mean": float(st["rho"].mean()),
                   "vmax": float(np.abs(st["vel"][:, :3]).max()),
                   "hmax": float(np.abs(st["height0"]).max())}
    search = search_stats(p, st["pos"], scene.neighbour_counts())
    print(f"bench: search: {search['candidates_per_particle']:.1f} candidates / particle, "
          f"{search['wave_loop_per_64']:.1f} per wave, neighbours "
          f"{search['neighbours_per_particle']}", flush=True)

    # Per-kernel device time from HIP events on the scene's stream, over the
    # same trajectory: back to the initial state, the same warm-up, then the
    # same K steps launched eagerly with an event pair around each kernel.
    scene.reinit(spec.pos, spec.vel, None, spec.wave_size, spec.height0, spec.height1)
    scene.step(spec.dt, args.warmup)
    prof = scene.step_profiled(spec.dt, min(args.steps, args.profile_steps)
                               if args.profile_steps > 0 else args.steps)
    kernel_ms = {k: v[0] for k, v in prof.items()}
    scene.close()

    units = {"force_integrate": n, "density": n, "hash": n, "reorder": n}
    alg = {k: BYTES_PER_PARTICLE[k] * units[k] for k in units}
    alg["wave"] = BYTES_PER_WAVE_CELL * W * H
    alg["scan"] = BYTES_PER_CELL * ncell
    dominant = max(kernel_ms, key=lambda k: kernel_ms[k])
    dom_bytes = alg.get(dominant)
    achieved = dom_bytes / (kernel_ms[dominant] * 1e-3) / 1e9 if dom_bytes else None
    step_bytes = sum(BYTES_PER_PARTICLE.values()) * n + BYTES_PER_CELL * ncell + \
        BYTES_PER_WAVE_CELL * W * H
    traffic, traffic_src = stored_traffic(spec.name, dominant)

    copy = measure_copy(args.copy_mib) if args.copy_mib > 0 else None
    if copy:
        print(f"bench: device copy {copy['bytes'] / 2**30:.1f} GiB moved in {copy['ms']:.3f} ms: "
              f"{copy['gbs']:.0f} GB/s", flush=True)
    adjoint = measure_adjoint(args) if args.adjoint_steps > 0 else None
    batched = measure_batched(args) if args.batched_steps > 0 else None
    fitres = measure_fit(args) if args.fit_iters > 0 else None

    cpu = None
    if args.cpu_steps > 0:
        from oracle import oracle as O

        threads = cpu_threads(args)
        print(f"bench: CPU oracle, {threads} threads, {args.cpu_steps} steps", flush=True)
        o = O.OracleScene(p, spec.pos, spec.vel, None, spec.wave_size, spec.height0,
                          spec.height1, threads=threads)
        o.step(spec.dt, 1)  # warm-up
        c0 = time.perf_counter()
        o.step(spec.dt, args.cpu_steps)
        ct = time.perf_counter() - c0
        o.close()
        cpu = {"value": n * args.cpu_steps / ct, "unit": "particle-updates/s",
               "cores": threads, "kind": "port", "cpu": cpu_info(),
               "sample": f"{args.cpu_steps} coupled steps of the same {n}-particle + "
                         f"{W}x{H} workload after 1 warm-up step, fp32 oracle, OpenMP, "
                         f"{ct:.2f} s",
               "cell_updates_per_s": W * H * args.cpu_steps / ct}

    out = {
        "metric": "particle-updates/s + wave cell-updates/s per coupled step; "
                  "fwd+adjoint ÷ fwd time",
        "value": n * args.steps / t,
        "unit": "particle-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": spec.name, "particles": n, "wave": [W, H],
                   "grid_cells": ncell, "dt": spec.dt, "graph": graph},
        "cell_updates_per_s": W * H * args.steps / t,
        "step_bytes": step_bytes,
        "step_gbs": step_bytes / (ms_step * 1e-3) / 1e9,
        "kernel_ms": kernel_ms,
        "repeats_ms_per_step": [1e3 * x / args.steps for x in times],
        "copy_ceiling": copy,
        "roofline": {"bound": "hbm", "kernel": dominant, "achieved": achieved,
                     "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": achieved / PEAK_HBM_GBS if achieved else None,
                     "traffic": traffic, "traffic_source": traffic_src,
                     "alg_bytes": dom_bytes},
        "cpu_baseline": cpu,
        "finite": finite,
        "state_after": state_stats,
        "search": search,
        "adjoint": adjoint,
        "batched": batched,
        "fit": fitres,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
