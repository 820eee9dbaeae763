"""Benchmark of the coupled SPH + wave step on one MI355X.

Workload (default): 1 048 576 particles + 2048^2 heightfield dam break,
the at-scale configuration; one "step" is one full coupled step (grid
build, density, forces + integration, particle -> wave flux, wave
EVOLVE).  Inputs are resident in HBM before timing starts; the timed region
is K steps launched as hipGraph replays, bracketed by device
synchronisation.

Prints one JSON line: particle-updates/s (value), wave cell-updates/s, the
roofline of the dominant kernel (per-kernel average device time from HIP
events recorded on the scene's stream), and the CPU oracle (OpenMP) timed
on a bounded sample of the same workload.
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


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=30)
    ap.add_argument("--workload", choices=["large", "small"], default="large")
    ap.add_argument("--eager", action="store_true", help="launch kernels without hipGraph")
    ap.add_argument("--profile-steps", type=int, default=0,
                    help="steps of the per-kernel event pass (0: the same K as timed)")
    ap.add_argument("--cpu-steps", type=int, default=3, help="0 disables the CPU baseline")
    ap.add_argument("--cpu-threads", type=int, default=0, help="0: all host cores")
    ap.add_argument("--adjoint-steps", type=int, default=256,
                    help="steps of the forward+adjoint measurement (0 disables it)")
    return ap.parse_args()


def measure_adjoint(args) -> dict:
    """fwd+adjoint / fwd on the differentiable workload: 65 536 particles +
    512^2 heightfield, checkpoint every 16 steps.  fwd = the same steps
    launched eagerly without a tape; fwd+adjoint = checkpointed forward +
    adjoint sweep (segment recompute included) for an MSE height loss."""
    import numpy as np

    from coupledwateranimation_amd import configs
    from coupledwateranimation_amd.gradient import Tape

    spec = configs.dam_break_diff(steps=args.adjoint_steps)
    steps, every = args.adjoint_steps, 16
    scene = spec.make_scene()
    scene.step(spec.dt, 2)  # warm-up (kernels loaded, caches warm)
    scene.reinit(spec.pos, spec.vel, None, spec.wave_size, spec.height0, spec.height1)
    scene.sync()
    t0 = time.perf_counter()
    scene.step(spec.dt, steps)
    scene.sync()
    t_fwd = time.perf_counter() - t0
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
    print(f"bench: adjoint {steps} steps: fwd {1e3 * t_fwd:.2f} ms, fwd+adjoint "
          f"{1e3 * t_fa:.2f} ms", flush=True)
    return {"workload": spec.name, "steps": steps, "ckpt_every": every,
            "fwd_ms": 1e3 * t_fwd, "fwd_adjoint_ms": 1e3 * t_fa, "ratio": t_fa / t_fwd,
            "finite": bool(np.isfinite(list(grad.values())).all())}


def main() -> int:
    args = parse()
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
    scene.step(spec.dt, args.warmup, graph=graph)
    scene.sync()
    import torch  # noqa: F401  (device synchronisation on both sides)

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    scene.step(spec.dt, args.steps, graph=graph)
    scene.sync()
    torch.cuda.synchronize()
    t = time.perf_counter() - t0
    ms_step = 1e3 * t / args.steps
    print(f"bench: timed {args.steps} steps in {t:.3f} s ({ms_step:.4f} ms/step)", flush=True)

    st = scene.readback()
    finite = bool(np.isfinite(st["pos"]).all() and np.isfinite(st["height0"]).all())
    state_stats = {"rho_min": float(st["rho"].min()), "rho_max": float(st["rho"].max()),
                   "rho_mean": float(st["rho"].mean()),
                   "vmax": float(np.abs(st["vel"][:, :3]).max()),
                   "hmax": float(np.abs(st["height0"]).max())}

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
    traffic = None
    tf = ROOT / "profiles" / "traffic.json"
    if tf.exists():
        try:
            traffic = json.loads(tf.read_text()).get(spec.name, {}).get(dominant)
        except (ValueError, AttributeError):
            traffic = None

    adjoint = measure_adjoint(args) if args.adjoint_steps > 0 else None

    cpu = None
    if args.cpu_steps > 0:
        from oracle import oracle as O

        threads = args.cpu_threads or len(os.sched_getaffinity(0))
        threads = min(threads, 16)
        print(f"bench: CPU oracle, {threads} threads, {args.cpu_steps} steps", flush=True)
        o = O.OracleScene(p, spec.pos, spec.vel, None, spec.wave_size, spec.height0,
                          spec.height1, threads=threads)
        o.step(spec.dt, 1)  # warm-up
        c0 = time.perf_counter()
        o.step(spec.dt, args.cpu_steps)
        ct = time.perf_counter() - c0
        o.close()
        cpu = {"value": n * args.cpu_steps / ct, "unit": "particle-updates/s",
               "cores": threads, "kind": "port",
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
        "roofline": {"bound": "hbm", "kernel": dominant, "achieved": achieved,
                     "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": achieved / PEAK_HBM_GBS if achieved else None,
                     "traffic": traffic, "alg_bytes": dom_bytes},
        "cpu_baseline": cpu,
        "finite": finite,
        "state_after": state_stats,
        "adjoint": adjoint,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
