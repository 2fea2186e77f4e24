#!/usr/bin/env python
"""Benchmark of the DUGKS time-step engine on one MI355X.

Workload: the 3-D lid-driven cavity, 48^3 hexes x 24^3 velocities (Kn 0.1,
Ma 0.16 lid, diffuse walls at 273 K), FP64 plain path, fixed dt at Courant 0.5.
A step is one full DUGKS time step (fvDVM::evolution) over the whole mesh with
the state resident in HBM.  value = cell-velocity updates per second (G/s).

Prints one JSON line on stdout; progress goes to stderr.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def bytes_model(case, storage_bytes: int, n_blocks: int, n_blocks_cell: int = None, face_values: bool = False,
                fused_cell: bool = False):
    """Algorithmic bytes per launch of each step kernel: every array the kernel must
    touch, read once and written once.

    k_face reads the resident gbar rows (each cell's row once, however many faces use it)
    and the lagged wall-gradient rows, writes the new wall-gradient rows and the block
    partial sums.  In face-value mode k_face_values does the same reads and writes the face
    values instead of partial sums, and k_face_moments reads the face values and writes the
    partial sums.  k_cell reads gbar (and the wall rows, or in face-value mode each face value
    once) and writes g~ and its partial sums;
    k_cell_s1 reads g~ and writes the next gbar; the two _fin kernels touch O(faces) and
    O(cells) data.  With fused_cell, k_cell also forms the macros and the next gbar in place
    (k_cell_fin and k_cell_s1 do not run; the re-read of its own row adds nothing to the floor).
    h is carried only when K + 3 - D != 0.  The whole-step figure is the
    floor of the scheme: gbar read twice and written once (3 s per update), plus the face
    values written and read once in face-value mode."""
    s = storage_bytes
    nd = 1 if (case.K + 3 - case.dim) == 0 else 2
    nc, nv = case.n_cells, case.nv
    nf = case.n_faces
    nw = sum(p["size"] for p in case.patches if p["type"] == 1)
    nm_f = (case.dim + 2 + case.dim + case.dim * (case.dim + 1) // 2) + 1
    nm_c = nm_f - 1
    fpart = nf * n_blocks * nm_f * 8
    cpart = nc * (n_blocks if n_blocks_cell is None else n_blocks_cell) * nm_c * 8
    fvb = nd * s * nf * nv if face_values else 0
    if face_values:
        kb = {"k_face_values": nd * s * (nc * nv + 2 * nw * nv) + fvb,
              "k_face_moments": fvb + fpart}
    else:
        kb = {"k_face": nd * s * (nc * nv + 2 * nw * nv) + fpart}
    kb.update({
        "k_face_fin": fpart + nf * 8 * (3 + 8 + 16),
        "k_cell": (nd * s * 2 * nc * nv + fvb if face_values else nd * s * (2 * nc * nv + nw * nv))
        + cpart + nf * 8 * 16,
        "k_cell_fin": cpart + nc * 8 * 9,
        "k_cell_s1": nd * s * 2 * nc * nv,
    })
    if fused_cell:
        kb["k_cell"] += nc * 8 * 9 - cpart
        del kb["k_cell_fin"], kb["k_cell_s1"]
    step_min = nd * s * (3 * nc * nv + 3 * nw * nv) + 2 * fvb
    return kb, step_min


def _counters(workload: str = "cavity3d"):
    p = os.path.join(HERE, "profiles", f"counters_{workload}.json")
    if not os.path.exists(p):
        return {}
    try:
        return json.load(open(p))
    except Exception:
        return {}


def traffic_from_profile(kernel_prefix: str, storage: str, workload: str = "cavity3d"):
    """HBM bytes per launch measured with rocprofv3 --pmc (FETCH_SIZE + WRITE_SIZE, each scaled
    by the factor a plain device copy of known size reads back), recorded in
    profiles/counters_<workload>.json, else None."""
    return _counters(workload).get(storage, {}).get(kernel_prefix)


def step_traffic_from_profile(storage: str, kernels, workload: str = "cavity3d"):
    """Measured HBM bytes of one whole step: the sum over the step's kernels (each launched once
    per step) of their calibrated counter bytes, or None if a kernel was not profiled."""
    d = _counters(workload).get(storage, {})
    tot = 0
    for k in kernels:
        if k not in d:
            return None
        tot += d[k]
    return tot


# Algorithmic bytes per cell-velocity update of the whole step (the contract's B = 6 s): g and h
# each streamed in twice and out once between the step's two velocity-space reductions; s = 8
# for FP64 storage, 4 for FP32.  With h identically zero (D = 3, KInner = 0) the g-only floor is
# 3 s.
def contract_bytes_per_update(storage_bytes: int) -> int:
    return 6 * storage_bytes


def host_description(threads: int):
    """The host the CPU baseline ran on: logical CPUs, the CPUs this process may use, the cgroup
    CPU quota, the CPU model and the RAM."""
    info = {"threads_used": threads, "nproc": os.cpu_count()}
    try:
        info["affinity_cpus"] = len(os.sched_getaffinity(0))
    except Exception:
        pass
    try:
        q, per = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        info["cgroup_cpu_quota"] = None if q == "max" else round(int(q) / int(per), 2)
    except Exception:
        pass
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                info["cpu_model"] = line.split(":", 1)[1].strip()
                break
    except Exception:
        pass
    try:
        for line in open("/proc/meminfo"):
            if line.startswith("MemTotal"):
                info["ram_GiB"] = round(int(line.split()[1]) / 2 ** 20, 1)
                break
    except Exception:
        pass
    return info


def run_gpu(case, storage: str, steps: int, warmup: int, tune: bool = True, workload: str = "cavity3d"):
    from dugksfoam_amd.engine import Engine
    e = Engine(case, storage=storage)
    e.init_equilibrium()
    tune_steps = 0
    if tune:
        # the engine's launch-shape tuner: real steps, one candidate shape per step, untimed
        e.autotune(True)
        tune_steps = e.autotune_steps_left()
        e.step(nsteps=tune_steps)
    e.step(nsteps=warmup)
    e.set_profiling(True)
    t0 = time.perf_counter()
    e.step(nsteps=steps)          # returns after the engine stream has drained
    t1 = time.perf_counter()
    kms, n = e.profile()
    n_blocks, n_blocks_cell = e.blocks()[2], e.cell_blocks()[2]
    shapes = {"k_face": e.blocks()[:3], "k_cell": e.cell_blocks()}
    fv = e.face_values()
    fused = e.fused_cell()
    m = e.macro()
    ok = bool(np.all(np.isfinite(m["rho"])) and np.all(m["rho"] > 0))
    e.close()
    wall = t1 - t0
    upd = case.n_cells * case.nv * steps
    sb = 8 if storage == "fp64" else 4
    kb, step_min = bytes_model(case, sb, n_blocks, n_blocks_cell, fv, fused)
    kern = {k: (kms[k] / n, kb[k]) for k in kb if kms.get(k, 0.0) > 0}
    dom = max(kern, key=lambda k: kern[k][0])
    step_ms = sum(v[0] for v in kern.values())     # HIP events around each kernel, per step
    ncv = case.n_cells * case.nv
    # the contract's figure: B = 6 s bytes per update over the whole step's kernel time
    bpu = contract_bytes_per_update(sb)
    achieved = bpu * ncv / (step_ms * 1e-3) / 1e9
    h_elided = (case.K + 3 - case.dim) == 0
    roof = {"bound": "hbm", "kernel": "whole step (" + ", ".join(kern) + ")",
            "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": step_traffic_from_profile(storage, list(kern), workload),
            "bytes_per_update": bpu, "updates_per_step": ncv, "bytes_per_step": bpu * ncv,
            "step_ms": round(step_ms, 4)}
    if h_elided:
        # h is identically zero here and never moved: g read twice and written once, 3 s per update
        gb = 3 * sb
        ga = gb * ncv / (step_ms * 1e-3) / 1e9
        roof["g_only_floor"] = {"bytes_per_update": gb, "achieved": round(ga, 1), "frac": round(ga / HBM_PEAK_GBS, 4)}
    avg_ms, bytes_ = kern[dom]
    roof["dominant_kernel"] = {"name": dom, "avg_launch_ms": round(avg_ms, 4), "design_bytes_per_launch": bytes_,
                               "design_GBs": round(bytes_ / (avg_ms * 1e-3) / 1e9, 1),
                               "traffic": traffic_from_profile(dom, storage, workload)}
    # the engine's own byte model (stored face values counted as traffic it must move)
    roof["design_model"] = {
        "kernels": {k: {"avg_launch_ms": round(v[0], 4), "bytes_per_launch": v[1],
                        "GBs": round(v[1] / (v[0] * 1e-3) / 1e9, 1)} for k, v in kern.items()},
        "bytes_per_step": step_min, "bytes_per_update": round(step_min / ncv, 2),
        "GBs": round(step_min / (step_ms * 1e-3) / 1e9, 1)}
    roof["launch_shapes"] = {k: {"threads": v[0], "iters": v[1], "velocity_blocks": v[2]} for k, v in shapes.items()}
    roof["launch_shapes"]["tuning_steps"] = tune_steps
    roof["face_values"] = fv
    roof["fused_cell"] = fused
    return dict(value=upd / wall / 1e9, ms_per_step=wall / steps * 1e3, kernel_ms=step_ms,
                roofline=roof, finite=ok)


def cpu_baseline(target_s: float = 12.0):
    """The FP64 oracle (a CPU restatement of the reference's time step) on a bounded
    sample of the workload: the same 48^3 mesh with a 12^3 velocity grid."""
    from dugksfoam_amd import cases
    from oracle.oracle import Oracle
    # every CPU this process may use, at most the OMP_NUM_THREADS the host sets for it (16 per
    # GPU on a one-GPU share of the GPU host)
    try:
        avail = len(os.sched_getaffinity(0))
    except Exception:
        avail = os.cpu_count() or 1
    env = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    threads = min(avail, env) if env > 0 else avail
    c = cases.cavity(48, 12, dim=3, kn=0.1)
    o = Oracle(c, threads=threads)
    o.init_equilibrium()
    n, t = 0, 0.0
    while t < target_s and n < 5:
        t0 = time.perf_counter()
        o.step(nsteps=1)
        t += time.perf_counter() - t0
        n += 1
    o.close()
    rate = c.n_cells * c.nv * n / t / 1e9
    out = {"value": rate, "unit": "G cell-velocity updates/s", "cores": threads, "kind": "port",
           "sample": f"FP64 oracle, 48^3 cells x 12^3 velocities, {n} step(s), {t:.1f} s",
           "host": host_description(threads)}
    # one core, on the 2-D cavity (60^2 cells x 41^2 velocities), about 10 s
    c1 = cases.config("cavity2d")
    o1 = Oracle(c1, threads=1)
    o1.init_equilibrium()
    n1, t1 = 0, 0.0
    while t1 < 0.5 * target_s and n1 < 20:
        t0 = time.perf_counter()
        o1.step(nsteps=1)
        t1 += time.perf_counter() - t0
        n1 += 1
    o1.close()
    out["single_core"] = {"value": c1.n_cells * c1.nv * n1 / t1 / 1e9, "cores": 1,
                          "sample": f"FP64 oracle, 60^2 cells x 41^2 velocities, {n1} step(s), {t1:.1f} s"}
    return out


# the named workloads of cases.config, as the JSON line's config.workload describes them
WORKLOADS = {
    "cavity3d": "cavity3d: 48^3 hexes x 24^3 velocities, Kn 0.1, lid Ma 0.16, diffuse walls 273 K",
    "cavity2d": "cavity2d: 60^2 quads x 41^2 velocities, Kn 0.075, lid Ma 0.16, diffuse walls 273 K",
    "delaunay": "delaunay: lid-driven cavity on 203582 Delaunay triangles (orthogonal wall row) x 64^2 "
                "velocities, Kn 1",
    "batch": "batch: 256 independent 2-D cavities (sides 20..80, Kn 0.01..10, per-case dt and muRef) "
             "x 28^2 velocities",
    "channel": "channel: 200x40 quads x 32^2 velocities, Kn 0.5, cyclic in x, diffuse walls",
    "cavity3d_share": "cavity3d_share: 48^3 hexes x 12^3 velocities (one GPU's velocity share of "
                      "cavity3d over 8 GPUs)",
}

SWEEP = (("cavity2d", 2000), ("delaunay", 1000), ("cavity3d", 200), ("batch", 1000), ("channel", 2000),
         ("cavity3d_share", 200))


def sweep():
    """Every workload of the scope table in FP64 and in the FP32-storage form: throughput over
    the stated step count (launch-shape tuning steps included in the timed run) and the
    relative L2 distance of the FP32-form fields from the FP64 fields."""
    from dugksfoam_amd import cases
    from dugksfoam_amd.engine import Engine

    def rel(a, b):
        return float(np.linalg.norm(a - b) / np.linalg.norm(b))

    for name, n in SWEEP:
        c = cases.config(name)
        log(f"[sweep] {name} cells={c.n_cells} nv={c.nv} steps={n}")
        out = {}
        for st in ("fp64", "fp32"):
            e = Engine(c, storage=st)
            e.init_equilibrium()
            e.autotune(True)
            t0 = time.perf_counter()
            done = 0
            while done < n:
                k = min(100, n - done)
                e.step(nsteps=k)
                done += k
            m = e.macro()
            out[st] = (m, time.perf_counter() - t0)
            e.close()
        m64, m32 = out["fp64"][0], out["fp32"][0]
        r = {"workload": name, "cells": c.n_cells, "velocities": c.nv, "steps": n,
             "fp64_G_updates_per_s": round(c.n_cells * c.nv * n / out["fp64"][1] / 1e9, 2),
             "fp32_G_updates_per_s": round(c.n_cells * c.nv * n / out["fp32"][1] / 1e9, 2),
             "fp32_vs_fp64_rel_l2": {k: float(f"{rel(m32[k], m64[k]):.3g}") for k in ("rho", "U", "T", "q")}}
        if c.n_cases > 1:
            # each field's worst case over the batch
            cc = np.asarray(c.cell_case)
            r["fp32_vs_fp64_worst_case"] = {
                k: float(f"{max(rel(m32[k][cc == i], m64[k][cc == i]) for i in range(c.n_cases)):.3g}")
                for k in ("rho", "U", "T", "q")}
        print(json.dumps(r), flush=True)


def adjoint_bench(steps: int, avg_last: int, every: int):
    """The discrete adjoint on the micro-channel (200 x 40 cells, 32^2 velocities, 16 wall-temperature
    modes): one forward + reverse sweep giving dJ/da of the mid-plane mass flow averaged over the
    last avg_last of `steps` steps, timed; beside it the FP64 engine's time for one forward run of
    the same steps, from which central differences would need 2 n_modes + 1 runs."""
    from dugksfoam_amd import cases, sensitivity
    from dugksfoam_amd.adjoint import Adjoint
    from dugksfoam_amd.engine import Engine
    c = cases.config("channel")
    n_modes = int(np.asarray(c.meta["a"]).size)
    log(f"[adjoint] channel cells={c.n_cells} nv={c.nv} steps={steps} avg_last={avg_last} every={every}")
    e = Engine(c)
    e.init_equilibrium()
    e.autotune(True)
    e.step(nsteps=e.autotune_steps_left())
    e.init_equilibrium()
    t0 = time.perf_counter()
    e.step(nsteps=steps)
    engine_s = time.perf_counter() - t0
    e.close()
    log(f"[adjoint] engine forward {engine_s:.3f} s")
    A = Adjoint(c)
    t0 = time.perf_counter()
    J, g = A.run(steps, steps - avg_last + 1, sensitivity.objective_coef(c, avg_last), checkpoint_every=every)
    wall = time.perf_counter() - t0
    fwd, rev = A.last_ms()
    A.close()
    b, dT = sensitivity.dTw_da(c)
    ga = dT.T @ g[b]
    log(f"[adjoint] J {J:.6e} forward {fwd:.0f} ms reverse {rev:.0f} ms wall {wall:.2f} s")
    fd_s = (2 * n_modes + 1) * engine_s
    return {
        "metric": "adjoint gradient wall time (dJ/da, 16 modes)", "value": round(wall, 3), "unit": "s",
        "higher_is_better": False, "n_gpus": 1, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "channel: 200x40 quads x 32^2 velocities, Kn 0.5, cyclic in x, lower wall "
                               "T(x) = 273 (1 + sum a_k sin(2 pi k x / 5)), a_k ~ U[-0.02, 0.02] (seed 2)",
                   "steps": steps, "objective_steps": avg_last, "checkpoint_every": every,
                   "cells": c.n_cells, "velocities": c.nv},
        "forward_sweep_ms": round(fwd, 1), "reverse_sweep_ms": round(rev, 1),
        "engine_forward_run_s": round(engine_s, 3),
        "central_differences_s": round(fd_s, 2), "speedup_vs_central_differences": round(fd_s / wall, 2),
        "J": J, "dJ_da": [float(f"{x:.9e}") for x in ga]}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-fp32", action="store_true", help="skip the FP32-storage form")
    ap.add_argument("--no-tune", action="store_true", help="keep the default launch shapes (no tuning steps)")
    ap.add_argument("--workload", default="cavity3d", choices=list(WORKLOADS),
                    help="the workload measured (default cavity3d); another one runs without the "
                         "CPU-baseline and adjoint legs")
    ap.add_argument("--sweep", action="store_true",
                    help="only run every workload in FP64 and FP32 storage (throughput, FP32 error)")
    ap.add_argument("--copy-calibration", type=float, default=0.0, metavar="GIB",
                    help="only run a plain device copy of GIB GiB (rocprof counter calibration)")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint measurement of the default run")
    ap.add_argument("--adjoint", action="store_true",
                    help="only time the discrete adjoint's gradient on the micro-channel")
    ap.add_argument("--adj-steps", type=int, default=2000)
    ap.add_argument("--adj-avg", type=int, default=500)
    ap.add_argument("--adj-every", type=int, default=100)
    a = ap.parse_args()
    if a.adjoint:
        print(json.dumps(adjoint_bench(a.adj_steps, min(a.adj_avg, a.adj_steps), min(a.adj_every, a.adj_steps))),
              flush=True)
        return
    if a.copy_calibration > 0:
        import torch
        n = int(a.copy_calibration * 2 ** 30) // 8
        x = torch.ones(n, dtype=torch.float64, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": n * 8, "copy_bytes_written": n * 8, "launches": 3}))
        return
    if a.sweep:
        sweep()
        return
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the DUGKS engine runs on a single MI355X; "
              "velocity-space sharding over several GPUs is not built", file=sys.stderr)
        sys.exit(2)
    from dugksfoam_amd import cases
    wl = a.workload
    case = cases.config(wl)
    log(f"[bench] {wl} cells={case.n_cells} nv={case.nv} steps={a.steps} warmup={a.warmup}")
    r64 = run_gpu(case, "fp64", a.steps, a.warmup, tune=not a.no_tune, workload=wl)
    log(f"[bench] fp64 {r64['value']:.2f} G upd/s, {r64['ms_per_step']:.2f} ms/step")
    out = {
        "metric": "G cell-velocity updates/s (FP64 plain path; FP32-storage form in fp32)",
        "value": round(r64["value"], 3),
        "unit": "G cell-velocity updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(r64["ms_per_step"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": WORKLOADS[wl], "cells": case.n_cells, "velocities": case.nv,
                   "dt": case.dt if case.n_cases == 1 else "per case"},
        "roofline": r64["roofline"],
    }
    if case.K + 3 - case.dim == 0:
        out["config"]["resident"] = "g only (h is identically zero for D=3, KInner=0)"
    if not a.no_fp32:
        r32 = run_gpu(case, "fp32", a.steps, a.warmup, tune=not a.no_tune, workload=wl)
        log(f"[bench] fp32 {r32['value']:.2f} G upd/s")
        out["fp32"] = {"value": round(r32["value"], 3), "ms_per_step": round(r32["ms_per_step"], 3),
                       "speedup_vs_fp64": round(r32["value"] / r64["value"], 3), "roofline": r32["roofline"]}
    if not a.no_adjoint and wl == "cavity3d":
        # the discrete adjoint on the micro-channel (its own workload; not part of value)
        try:
            ad = adjoint_bench(a.adj_steps, min(a.adj_avg, a.adj_steps), min(a.adj_every, a.adj_steps))
            out["adjoint"] = {k: ad[k] for k in ("metric", "value", "unit", "config", "forward_sweep_ms",
                                                  "reverse_sweep_ms", "central_differences_s",
                                                  "speedup_vs_central_differences")}
        except Exception as exc:   # reported, never allowed to hide the main line
            out["adjoint"] = {"error": repr(exc)}
    if not a.no_cpu and wl == "cavity3d":
        cb = cpu_baseline()
        log(f"[bench] cpu baseline {cb['value']:.4f} G upd/s on {cb['cores']} threads")
        out["cpu_baseline"] = cb
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
