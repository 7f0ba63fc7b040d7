"""Benchmark: batched elastic FWI gradient on one MI355X.

Workload (the 500x1700 configuration): grid 500 x 1700 at 10 m plus a
40-cell C-PML (608 x 1780 padded), nt = 4000, dt = 1 ms, 64 shots, 1700
receivers, 12 Hz Ricker, seeded layered model with 3 elliptical anomalies.
One "step" = one full gradient evaluation (calc_id 1: forward with boundary
store, residual, time-reversed re-materialisation + adjoint + gradient) of all
64 shots, with the model, source functions and observed data already resident
in HBM.  value = shot-gradients / s.

Prints ONE JSON line.  Extra fields: roofline (dominant kernel, HIP-event
timed inside the timed region), job_roofline (whole gradient against the
algorithmic bytes of a shot-gradient), cpu_baseline (the CPU oracle on a
bounded sample of the same workload).
"""
import argparse
import json
import os
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "fwiflow.jl_amd"))

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0

# Algorithmic bytes per cell and shot for each kernel (FP32): every array the
# kernel must touch, read or written once; model parameters are shared by all
# shots of a launch and amortised to zero; C-PML memory variables add 8 B per
# read-modify-write where the kernel updates them.
KERNEL_BYTES = {
    "stress_fwd": 32.0,   # read vz vx szz sxx sxz, write szz sxx sxz
    "vel_fwd": 28.0,      # read szz sxx sxz vz vx, write vz vx
    "rev_vel": 36.0,      # read 3 stress, 2 velocity, 2 adjoint velocity, write 2 velocity
    "rev_stress": 44.0,   # read 2 velocity, 3 stress, 3 adjoint stress, write 3 stress
    "adj_vel": 28.0,      # read 3 adjoint stress, 2 adjoint velocity, write 2
    "adj_stress": 32.0,   # read 2 adjoint velocity, 3 adjoint stress, write 3
    # exact adjoint (--adjoint exact), each with the reverse step of its field;
    # outside the C-PML frame 3 of the 4 stress operands and 2 of the 4
    # velocity operands are distinct (KERNEL_FRAME adds the others)
    "exact_a": 64.0,      # read 2 adjoint velocity, 3 stress operands, 3 stress, 2 velocity;
                          # write 2 adjoint velocity, 2 velocity operands, 2 velocity
    "exact_b": 76.0,      # read 3 adjoint stress, 2 velocity operands, 2 velocity, 3 stress;
                          # write 3 adjoint stress, 3 stress operands, 3 stress
}
# exact adjoint: the operands stored only in the C-PML frame, per frame cell
# (k_exact_a reads 1 and writes 2, k_exact_b reads 2 and writes 1)
KERNEL_FRAME = {"exact_a": 12.0, "exact_b": 12.0}
# kernels that run two of the above as one grid (disjoint workgroup ranges)
KERNEL_PARTS = {"bwd_pair_a": ("rev_stress", "adj_vel"), "bwd_pair_b": ("adj_stress", "rev_vel")}
# Measured HBM-side bytes per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE, each
# scaled by a 1 GiB copy's counters, averaged over the launches) at this
# workload's shot groups: profiles/stencil_kernels_608x1780_counters.md.
# Reported as roofline.traffic when the dominant kernel and its mean shots
# per launch match the profiled run.
MEASURED_TRAFFIC = {("bwd_pair_a", 2.91): 231.5e6, ("bwd_pair_b", 2.91): 226.9e6,
                    ("stress_fwd", 4.92): 204.4e6, ("vel_fwd", 4.92): 180.9e6,
                    ("exact_a", 2.0): 161.8e6, ("exact_b", 2.0): 199.1e6}
# kernels of the backward sweep (their launches cover the backward shot groups)
BACKWARD = ("rev_vel", "rev_stress", "adj_vel", "adj_stress", "bwd_pair_a", "bwd_pair_b",
            "exact_a", "exact_b")
# C-PML memory-variable traffic per PML cell-direction (2 variables RW = 16 B)
KERNEL_PML = {"stress_fwd": 16.0, "vel_fwd": 16.0, "adj_vel": 16.0, "adj_stress": 16.0,
              "exact_a": 16.0, "exact_b": 16.0}


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=2)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--nt", type=int, default=4000)
    p.add_argument("--nshots", type=int, default=64)
    p.add_argument("--cpu-sample-steps", type=int, default=1500)
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--checkpoint", type=int, default=0,
                   help="checkpoint interval of the backward sweep (0: automatic)")
    p.add_argument("--memory-budget", type=float, default=None,
                   help="device bytes for the boundary ring and checkpoints (default 32e9)")
    p.add_argument("--workload", choices=["c3", "timelapse"], default="c3",
                   help="c3: the 500x1700 benchmark (default); timelapse: the six config-4 "
                        "surveys (200x600, 300x900, 450x1350, baseline and monitor, 16 shots "
                        "each, nt 3000) as one job")
    p.add_argument("--adjoint", choices=["reference", "exact"], default="reference",
                   help="the reference's adjoint (default, the benchmark) or the exact one "
                        "(the transpose of the forward step)")
    p.add_argument("--abi", action="store_true",
                   help="also time one end-to-end fwiflow_cufd gradient call (files, set-up)")
    p.add_argument("--autotune", action="store_true",
                   help="time launch shapes on a 200-step record first (fwiflow.tune) and "
                        "run with the fastest")
    p.add_argument("--copy-calibration", action="store_true",
                   help="only run a plain 1 GiB device-to-device copy, 4 times (the "
                        "reference kernel for scaling rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    return p.parse_args()


def copy_calibration():
    import torch

    n = 1 << 28  # 2^28 float32 = 1 GiB
    x = torch.ones(n, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    y.copy_(x)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(4):
        y.copy_(x)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / 4
    print(json.dumps({"copy_calibration_bytes_each_way": n * 4, "copies": 5,
                      "copy_ms": ms, "copy_rw_GBs": 2 * n * 4 / (ms * 1e-3) / 1e9}), flush=True)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def geometry_counts(c):
    """Active cells, PML cell-directions and per-kernel domain sizes."""
    nzA = c.NZ - c.nPad
    NX = c.NX
    active = nzA * NX
    pml_z = 2 * c.nPml * NX   # z-direction C-PML cells (all columns)
    pml_x = 2 * c.nPml * nzA  # x-direction C-PML cells (all active rows; corners have both)
    upd = (nzA - 4) * (NX - 4)  # stencil update region [2, n-3]^2
    ring_rect = (nzA - 2 * c.nPml + 4) * (NX - 2 * c.nPml + 4)
    frame = upd - (nzA - 2 * c.nPml) * (NX - 2 * c.nPml)  # update-region cells in the C-PML
    return dict(active=active, pml_dirs=pml_z + pml_x, update=upd, ring_rect=ring_rect,
                pml_z=pml_z, pml_x=pml_x, frame=frame)


def algorithmic_bytes_per_shot_gradient(c, nt):
    """168 B per active cell per step (forward 52 + re-materialised forward 52 +
    adjoint with gradient 64) plus 96 B per C-PML cell-direction per step."""
    g = geometry_counts(c)
    return (168.0 * g["active"] + 96.0 * g["pml_dirs"]) * nt


def kernel_launch_bytes(name, c, ns):
    if name in KERNEL_PARTS:
        return sum(kernel_launch_bytes(k, c, ns) for k in KERNEL_PARTS[name])
    g = geometry_counts(c)
    cells = g["ring_rect"] if name.startswith("rev_") else g["update"]
    b = KERNEL_BYTES[name] * cells * ns
    b += KERNEL_PML.get(name, 0.0) * g["pml_dirs"] * ns
    b += KERNEL_FRAME.get(name, 0.0) * g["frame"] * ns
    return b


def host_cores():
    """Threads the CPU baseline uses: the cores this process may run on
    (what `nproc` reports), capped by OMP_NUM_THREADS when the host sets it
    (a shared GPU host gives each GPU's jobs that share of its cores)."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    omp = os.environ.get("OMP_NUM_THREADS", "")
    if omp.isdigit() and int(omp) > 0:
        n = min(n, int(omp))
    return max(1, n or 1)


def cpu_model():
    """`lscpu`'s model name (with its CPU family / model / stepping numbers,
    which identify the part where a virtual machine shortens the name);
    /proc/cpuinfo when lscpu is absent."""
    import subprocess

    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        f = dict(line.split(":", 1) for line in out.splitlines() if ":" in line)
        f = {k.strip(): v.strip() for k, v in f.items()}
        if f.get("Model name"):
            ids = ", ".join(f"{k.lower()} {f[k]}" for k in ("CPU family", "Model", "Stepping")
                            if k in f)
            return f"{f['Model name']} ({ids})" if ids else f["Model name"]
    except (OSError, subprocess.SubprocessError, ValueError):
        pass
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def abi_seconds(c, ws, ids, models, fwiflow):
    """One end-to-end `fwiflow_cufd` gradient call (calc_id 1) as the
    reference's callers make it: JSON parsing, plan set-up and device
    allocation, host model transposes, the observed data read from the 64
    `Shot{id}.bin` files, the gradient, the host copies back, teardown."""
    para, stf = models.write_case(c, os.path.join(ws, "abi"))
    lam, mu, den = models.moduli(c, "true")
    fwiflow.fwi_obs_op(lam, mu, den, stf, 0, ids, para)
    lam0, mu0, den0 = models.moduli(c, "start")
    t0 = time.perf_counter()
    fwiflow.fwi_op_grad(lam0, mu0, den0, stf, 0, ids, para)
    return time.perf_counter() - t0


def cpu_baseline(c, ws, sample_steps):
    from fwiflow import models
    from oracle import oracle

    oracle.build()
    cores = host_cores()
    cs = models.case3(nSteps=sample_steps, nshots=c.src_x.size)
    para, stf = models.write_case(cs, os.path.join(ws, "cpu"))
    lam, mu, den = models.moduli(cs, "true")
    oracle.cufd(lam, mu, den, stf, 2, [0], para, precision=32, nthreads=cores)
    lam0, mu0, den0 = models.moduli(cs, "start")
    t0 = time.perf_counter()
    oracle.cufd(lam0, mu0, den0, stf, 1, [0], para, precision=32, nthreads=cores)
    t = time.perf_counter() - t0
    per_grad = t * (c.nSteps / sample_steps)
    # the 100x200 case's full gradient (1 shot, nt 1000), not extrapolated
    c1 = models.case1()
    p1, s1 = models.write_case(c1, os.path.join(ws, "cpu_c1"))
    oracle.cufd(*models.moduli(c1, "true"), s1, 2, [0], p1, precision=32, nthreads=cores)
    t1 = time.perf_counter()
    oracle.cufd(*models.moduli(c1, "start"), s1, 1, [0], p1, precision=32, nthreads=cores)
    t1 = time.perf_counter() - t1
    return dict(value=1.0 / per_grad, unit="shot-gradients/s", cores=cores, kind="port",
                cpu=cpu_model(), nproc=os.cpu_count(),
                c1_full_gradient_s=round(t1, 3),
                sample=f"1 shot of the same grid, first {sample_steps} of {c.nSteps} steps "
                       f"(calc_id 1, FP32 reference arithmetic, OpenMP on {cores} threads), "
                       f"{t:.1f} s, scaled linearly to {c.nSteps} steps")


def timelapse(a):
    """One JSON line for the time-lapse workload: six surveys (three grids,
    baseline and monitor, 16 shots each, nt 3000) whose gradients run as one
    job (every time step one launch across the grids), against the same
    plans run one after another."""
    import fwiflow
    from fwiflow import models

    nt = 3000
    ws = tempfile.mkdtemp(prefix="fwiflow_bench_tl_")
    plans, cases = [], []
    for nz, nx in ((200, 600), (300, 900), (450, 1350)):
        for monitor in (False, True):
            c = models.case4(nz, nx, monitor=monitor, nSteps=nt, nshots=16)
            para, stf = models.write_case(c, os.path.join(ws, c.name))
            P = fwiflow.Plan(para, list(range(16)))
            P.set_model(*models.moduli(c, "true"))
            P.set_stf(stf)
            P.run(2)
            P.set_obs(P.data())
            P.set_model(*models.moduli(c, "start"))
            plans.append(P)
            cases.append(c)
    log(f"bench: time-lapse surveys {[c.name for c in cases]}, nt {nt}")
    J = fwiflow.Job(plans)
    for _ in range(a.warmup):
        J.run(1)
    t0 = time.perf_counter()
    for i in range(a.steps):
        J.run(1)
        log(f"bench: job step {i} done at {time.perf_counter() - t0:.2f} s")
    elapsed = time.perf_counter() - t0
    fused = J.fused()
    J.close()
    t1 = time.perf_counter()
    for P in plans:
        P.run(1)
    separate = time.perf_counter() - t1
    nshots = 16 * len(plans)
    value = nshots * a.steps / elapsed
    grad_bytes = sum(16 * algorithmic_bytes_per_shot_gradient(c, nt) for c in cases)
    job_gbs = grad_bytes * a.steps / elapsed / 1e9
    out = {
        "metric": "shot-gradients/s, six time-lapse surveys in one job",
        "value": round(value, 4), "unit": "shot-gradients/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": round(elapsed / a.steps * 1e3, 2), "higher_is_better": True,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": "config-4 time-lapse: 200x600, 300x900, 450x1350 grids "
                               "(+32 C-PML), baseline and monitor, 16 shots each, nt 3000, "
                               "calc_id 1, observed data resident in HBM",
                   "grids": [[c.NZ, c.NX] for c in cases[::2]], "nt": nt, "shots": nshots},
        "job_roofline": {"algorithmic_bytes_per_job_gradient": grad_bytes,
                         "achieved": round(job_gbs, 1), "frac": round(job_gbs / HBM_PEAK_GBS, 4),
                         "unit": "GB/s"},
        "fused": fused,
        "plans_one_after_another_ms": round(separate * 1e3, 2),
        "job_speedup": round(separate / (elapsed / a.steps), 3),
    }
    for P in plans:
        P.close()
    print(json.dumps(out), flush=True)


def main():
    a = parse()
    if a.copy_calibration:
        copy_calibration()
        return
    if a.gpus != 1:
        print(f"bench.py: measures one GPU; --gpus {a.gpus} is not supported (shots are "
              "independent: each GPU would run a plan for its own shot range and the "
              "gradients would be summed once per evaluation)", file=sys.stderr)
        sys.exit(2)
    if a.workload == "timelapse":
        timelapse(a)
        return
    import fwiflow
    from fwiflow import models

    c = models.case3(nSteps=a.nt, nshots=a.nshots)
    ws = tempfile.mkdtemp(prefix="fwiflow_bench_")
    para, stf = models.write_case(c, ws)
    ids = list(range(a.nshots))
    log(f"bench: grid {c.NZ}x{c.NX} nt {c.nSteps} shots {a.nshots}")
    P = fwiflow.Plan(para, ids)
    P.set_adjoint(a.adjoint)
    if a.checkpoint:
        P.set_checkpoint(a.checkpoint)
    if a.memory_budget:
        P.set_memory_budget(a.memory_budget)
    lam, mu, den = models.moduli(c, "true")
    P.set_model(lam, mu, den)
    P.set_stf(stf)
    t0 = time.perf_counter()
    P.run(2)
    log(f"bench: observed data (forward, {a.nshots} shots) {time.perf_counter() - t0:.2f} s")
    P.set_obs(P.data())
    lam0, mu0, den0 = models.moduli(c, "start")
    P.set_model(lam0, mu0, den0)
    tuned = None
    if a.autotune:
        from fwiflow import tune

        t0 = time.perf_counter()
        tuned = tune.autotune(para, ids, lam0, mu0, den0, stf, steps=200,
                              path=os.path.join(ws, "launch.json"), log=log)
        P.set_launch(tuned["tile_x"], tuned["shots_fwd"], tuned["shots_bwd"])
        log(f"bench: autotuned launch {P.launch()} in {time.perf_counter() - t0:.1f} s")
    # timing is on from the warmup so that the time loops the engine captures
    # as graphs there (with their event records) are the ones the timed steps
    # replay; set_timing(True) again clears the warmup's totals
    P.set_timing(True)
    for i in range(a.warmup):
        t0 = time.perf_counter()
        P.run(1)
        log(f"bench: warmup {i} {time.perf_counter() - t0:.2f} s")
    P.set_timing(True)
    t0 = time.perf_counter()
    step_s = []
    for i in range(a.steps):
        t1 = time.perf_counter()
        P.run(1)  # synchronises the plan's stream before returning
        step_s.append(time.perf_counter() - t1)
        log(f"bench: step {i} done at {time.perf_counter() - t0:.2f} s")
    elapsed = time.perf_counter() - t0
    ms_per_step = elapsed / a.steps * 1e3
    value = a.nshots * a.steps / elapsed

    # per-kernel HIP-event timings: every 64th launch of each kind is launched
    # with hipExtLaunchKernel and events that the kernel itself stamps
    kt = {}
    for name in list(KERNEL_BYTES) + list(KERNEL_PARTS):
        ms, n = P.timing(name)
        if n:
            kt[name] = (ms / n, n)
    # every kernel runs once per time step per shot group: a launch covers
    # nshots / groups shots on average
    groups_f, groups_b = P.launch_groups()
    dominant = max(kt, key=lambda k: kt[k][0])
    mean_spl = a.nshots / (groups_b if dominant in BACKWARD else groups_f)
    avg_ms = kt[dominant][0]
    lb = kernel_launch_bytes(dominant, c, mean_spl)
    achieved = lb / (avg_ms * 1e-3) / 1e9
    grad_bytes = algorithmic_bytes_per_shot_gradient(c, c.nSteps)
    job_gbs = grad_bytes * value / 1e9
    g = geometry_counts(c)
    out = {
        "metric": "shot-gradients/s at 500x1700 nt=4000, % HBM peak; time-to-misfit vs plain",
        "value": round(value, 4),
        "unit": "shot-gradients/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms_per_step, 2),
        "ms_per_step_median": round(float(np.median(step_s)) * 1e3, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"elastic FWI gradient, 500x1700 grid (+40 C-PML -> {c.NZ}x{c.NX}), "
                        f"nt={c.nSteps}, dt=1 ms, {a.nshots} shots x 1700 receivers, "
                        "calc_id 1 with observed data resident in HBM",
            "grid": [c.NZ, c.NX],
            "nt": c.nSteps,
            "shots": a.nshots,
            "adjoint": a.adjoint,
        },
        "gcell_updates_per_s": round(g["active"] * (c.nSteps - 1) * value / 1e9, 3),
        "roofline": {
            "bound": "hbm",
            "kernel": dominant,
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": (MEASURED_TRAFFIC.get((dominant, round(mean_spl, 2)))
                        if c.NZ == 608 and c.NX == 1780 and c.nSteps == 4000 else None),
            "avg_launch_ms": round(avg_ms, 4),
            "bytes_per_launch": lb,
            "shots_per_launch": round(mean_spl, 3),
        },
        "job_roofline": {
            "algorithmic_bytes_per_shot_gradient": grad_bytes,
            "achieved": round(job_gbs, 1),
            "frac": round(job_gbs / HBM_PEAK_GBS, 4),
            "unit": "GB/s",
        },
        "kernel_avg_ms": {k: round(v[0], 4) for k, v in kt.items()},
        "launch": dict(zip(("tile_x", "shots_fwd", "shots_bwd"), P.launch()),
                       autotuned=tuned is not None),
        "launch_groups": {"forward": groups_f, "backward": groups_b},
        "device_bytes": P.device_bytes(),
        "memory_plan": dict(zip(("checkpoint_every", "shots_per_sweep", "ring_bytes",
                                 "checkpoint_bytes"), P.memory_plan())),
    }
    P.close()
    if a.abi:
        log("bench: one end-to-end fwiflow_cufd gradient call ...")
        t = abi_seconds(c, ws, ids, models, fwiflow)
        out["abi_end_to_end"] = {
            "seconds": round(t, 2), "shot_gradients_per_s": round(a.nshots / t, 4),
            "what": "one fwiflow_cufd calc_id 1 call: JSON, set-up and allocation, host "
                    "transposes, 64 Shot files read, gradient, copies back"}
    if not a.no_cpu_baseline:
        log("bench: CPU baseline (oracle) ...")
        out["cpu_baseline"] = cpu_baseline(c, ws, a.cpu_sample_steps)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
