"""Benchmark of the MI355X split-step path.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME] [--precision P] [--no-cpu]
(default workload: yoshida3d_spin6_ddi_128 in FP64, with the FP32-bulk run of the same
workload and its error against FP64 reported beside it under "fp32_bulk")

Prints ONE JSON line.  A step is one accepted time step of the whole grid
(all 2F+1 components); value = site-steps/s with psi resident in HBM.
Workloads (sizes from the reference benchmark configurations):
  rt3d_spin1_128 : 3-D spin-1 real-time leapfrog (run_simulation!), 128^3, box 20,
                   c0=10, c1=+0.5, omega=(1,1,1), q=0.1, dt=1e-3, saves every 10 steps
                   (energy/norm/Mz on the device); TF polar + 1e-3 seeded noise in m=+-1
  itp1d_spin1_64 : 1-D spin-1 imaginary-time split_step!, 64 points, box 20, c0=10,
                   c1=-0.5, omega=1, dt=0.005 (ITP renormalisation every step)
  yoshida3d_spin6_ddi_128 : 3-D spin-6 dipolar (13 components) 128^3, box 24, c0=10,
                   tensor channels from c2..c12 ~ U[-1,1] (seed 42), c_dd=1 along z (full
                   Q_ab), omega=(1,1,1), run_simulation_yoshida! adaptive (dt_init 1e-3, PI
                   controller, tol 1e-3); psi0 = m=+6 Gaussian + 1e-3 seeded noise.
                   A step = one ACCEPTED Yoshida step (rejected attempts included in time).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "bec-simulation_amd"))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

HBM_PEAK = 8.0e12  # B/s, MI355X


def _cfg2_inputs(G, n=128, precision="fp64"):
    gc = G.GridConfig((n, n, n), (20.0, 20.0, 20.0))
    grid = G.make_grid(gc)
    sys_ = G.SpinSystem(1)
    trap = G.HarmonicTrap(1.0, 1.0, 1.0)
    psi = G.init_psi_thomas_fermi(grid, sys_, trap, 10.0)
    psi = G.add_noise(psi, 1e-3, 42, components=[0, 2])
    return grid, trap, psi


FP32_PEAK = 157.3e12  # FLOP/s, MI355X FP32 (vector with packed FMA, or matrix)
FP64_PEAK = 78.6e12   # FLOP/s, MI355X FP64 vector / matrix
HBM_CLASSES = ("xpass", "fwd", "kin", "inv", "ddi_fft", "ddi_conv")


def _traffic_path(workload, precision):
    return os.path.join(ROOT, "profiles", f"{workload}_{precision}_traffic.json")


def _traffic(workload, kernel, precision):
    """HBM bytes per launch of `kernel` measured with rocprofv3 --pmc (FETCH_SIZE and
    WRITE_SIZE passes, scaled by a copy kernel of known size), if a summary for this
    workload at this precision is committed under profiles/ as JSON; otherwise None."""
    try:
        with open(_traffic_path(workload, precision)) as f:
            d = json.load(f)
        if d.get("precision") != precision:
            return None
        return d["kernels"][kernel]["hbm_bytes_per_launch"]
    except (OSError, KeyError, ValueError):
        return None


def _kernel_class(name):
    """Bench kernel class of a HIP kernel symbol (as the profiled pass groups launches)."""
    if "xsite_kernel" in name:
        return "xsite"
    if "site_kernel" in name:
        return "sites"
    if "xpass_kernel" in name:
        return "xpass"
    if "energy_kernel" in name:
        return "energy"
    if "strided_kernel" in name:
        args = [a.strip() for a in name.split("<", 1)[1].split(">", 1)[0].split(",")]
        mode, fields = int(args[2]), int(args[4])
        if fields == 3:
            return "ddi_conv" if mode == 3 else "ddi_fft"
        return {0: "fwd", 1: "inv", 2: "kin", 4: "ke", 5: "deriv"}.get(mode)
    return None


def pmc_summary(fetch_csv, write_csv, copy_bytes, workload, out_path, precision):
    """HBM bytes per launch by kernel class from two rocprofv3 --pmc runs (FETCH_SIZE, then
    WRITE_SIZE; KiB per dispatch), each scaled so that the largest psi copy of the run (the
    slot copy kernel or the runtime's copyBuffer), a copy of `copy_bytes`, reads and writes
    exactly that many bytes."""
    import csv
    import collections

    def load(path, counter):
        d = collections.defaultdict(list)
        with open(path) as f:
            for r in csv.DictReader(f):
                if r["Counter_Name"] == counter:
                    d[r["Kernel_Name"]].append(float(r["Counter_Value"]))
        return d

    fe, wr = load(fetch_csv, "FETCH_SIZE"), load(write_csv, "WRITE_SIZE")
    is_copy = lambda k: "copyBuffer" in k or "copy_slot_kernel" in k
    cf = max(max(v) for k, v in fe.items() if is_copy(k))
    cw = max(max(v) for k, v in wr.items() if is_copy(k))
    sf, sw = copy_bytes / 1024 / cf, copy_bytes / 1024 / cw
    cls = collections.defaultdict(lambda: [0, 0.0, 0.0])
    for k, v in fe.items():
        c = _kernel_class(k)
        # an FP32-bulk workload runs FP64 kernels only inside the energy evaluation
        if c is not None and precision == "fp32" and "<double" in k:
            c = "energy"
        if "widen_kernel" in k:
            c = "energy"
        if c is None or k not in wr:
            continue
        cls[c][0] += len(v)
        cls[c][1] += sum(v) * 1024 * sf
        cls[c][2] += sum(wr[k]) * 1024 * sw
    out = {"workload": workload, "precision": precision, "copy_bytes": copy_bytes,
           "scale_fetch": sf, "scale_write": sw,
           "kernels": {c: {"launches": n, "fetch_bytes_per_launch": f / n,
                           "write_bytes_per_launch": w / n, "hbm_bytes_per_launch": (f + w) / n}
                       for c, (n, f, w) in sorted(cls.items())}}
    with open(out_path, "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(out, indent=1))


def _rooflines(prof, work, precision):
    """Per-class achieved figures from the profiled pass: HBM classes in GB/s of
    algorithmic bytes; the lane-per-site pass in TFLOP/s of executed flops."""
    out = {}
    op_class = "xsite" if prof.get("xsite", (0, 0))[1] > 0 else "sites"
    for k, (ms, cnt) in prof.items():
        if cnt == 0 or ms <= 0:
            continue
        secs = ms * 1e-3
        if k in HBM_CLASSES and work["bytes"].get(k, 0) > 0:
            gbs = work["bytes"][k] / secs / 1e9
            out[k] = {"bound": "hbm", "achieved": round(gbs, 1), "peak": 8000.0, "unit": "GB/s",
                      "frac": round(gbs / 8000.0, 4), "avg_launch_us": round(ms / cnt * 1e3, 2),
                      "launches": cnt, "alg_bytes_per_launch": work["bytes"][k] / cnt}
        elif k == op_class and work["flops"] > 0:
            # per-site operator flops (tensor matvecs, rotations) land in whichever of the
            # two kernels ran them; the fused one adds its x-line FFTs
            fft = work.get("fft_flops", 0.0) if k == "xsite" else 0.0
            flops = work["flops"] + fft
            peak = FP32_PEAK if precision == "fp32" else FP64_PEAK
            tf = flops / secs / 1e12
            engine = ("VALU FP64 FMA (the FP64 vector and matrix peaks are equal)"
                      if precision == "fp64" else
                      "VALU FP32 FMA without packed FP32 (issue caps it at half the peak)")
            out[k] = {"bound": "valu", "achieved": round(tf, 2), "peak": peak / 1e12,
                      "unit": "TFLOP/s", "frac": round(tf * 1e12 / peak, 4), "engine": engine,
                      "avg_launch_us": round(ms / cnt * 1e3, 2), "launches": cnt,
                      "alg_flops_per_launch": flops / cnt, "fft_flops_per_launch": fft / cnt,
                      "tensor_matvecs": work["tensor_matvecs"], "rotations": work["rotations"],
                      "alg_bytes_per_launch": work["bytes"][k] / cnt,
                      "alg_GB/s": round(work["bytes"][k] / secs / 1e9, 1)}
    return out


def run_rt3d(G, A, steps, warmup, precision, n=128):
    grid, trap, psi = _cfg2_inputs(G, n=n, precision=precision)
    save_every = 10
    sp = G.SimParams(dt=1e-3, n_steps=max(steps, 1), imaginary_time=False, save_every=save_every)
    ws = G.make_workspace(grid=grid, atom=G.Rb87, interactions=G.InteractionParams(10.0, 0.5),
                          zeeman=G.ZeemanParams(0.0, 0.1), potential=trap, sim_params=sp,
                          psi_init=psi, precision=precision)
    lib = G.api.L.lib()

    def segment(k):
        G.api.L.check(lib.bec_leapfrog_segment(ws.handle, k), "leapfrog")
        ws.energy_terms()  # save point: energy, norm, Mz reductions on the device

    def run(nsteps):
        done = 0
        while done < nsteps:
            k = min(save_every, nsteps - done)
            segment(k)
            done += k

    run(warmup)
    ws.synchronize()
    t0 = time.perf_counter()
    run(steps)
    ws.synchronize()
    t1 = time.perf_counter()
    # profiled pass (HIP events around every launch) for per-kernel durations
    ws.profile(True)
    run(steps)
    prof = ws.profile_read()
    work = ws.profile_work()
    ws.profile(False)
    return ws, grid, (t1 - t0), prof, psi, work


def run_itp1d(G, A, steps, warmup, precision):
    gc = G.GridConfig((64,), (20.0,))
    grid = G.make_grid(gc)
    psi = G.init_psi(grid, G.SpinSystem(1), "polar")
    sp = G.SimParams(dt=0.005, n_steps=max(steps, 1), imaginary_time=True)
    ws = G.make_workspace(grid=grid, atom=G.Rb87, interactions=G.InteractionParams(10.0, -0.5),
                          potential=G.HarmonicTrap(1.0), sim_params=sp, psi_init=psi,
                          precision=precision)
    A.split_step(ws, warmup)
    ws.synchronize()
    t0 = time.perf_counter()
    A.split_step(ws, steps)
    ws.synchronize()
    t1 = time.perf_counter()
    ws.profile(True)
    A.split_step(ws, steps)
    prof = ws.profile_read()
    work = ws.profile_work()
    ws.profile(False)
    return ws, grid, (t1 - t0), prof, psi, work


def cfg3_inputs(G, n=128):
    L = 24.0
    grid = G.make_grid(G.GridConfig((n, n, n), (L, L, L)))
    gen = G.physics.SplitMix64(42)
    u = gen.uniform(6)
    c_extra = [0.0] * 11
    for i, k in enumerate(range(2, 13, 2)):
        c_extra[k - 2] = float(2 * u[i] - 1)          # c_k, k = 2..12, U[-1, 1]
    psi = G.init_psi(grid, G.SpinSystem(6), "ferromagnetic")
    psi = G.add_noise(psi, 1e-3, 42)
    inter = G.InteractionParams(10.0, 0.0, 0.0, tuple(c_extra))
    return grid, inter, psi, c_extra


def run_yoshida3d(G, A, steps, warmup, precision, padding=False, profiled=True):
    grid, inter, psi, _ = cfg3_inputs(G)
    sp = G.SimParams(dt=1e-3, n_steps=500)
    ws = G.make_workspace(grid=grid, atom=G.Eu151, interactions=inter,
                          potential=G.HarmonicTrap(1.0, 1.0, 1.0), sim_params=sp, psi_init=psi,
                          enable_ddi=True, c_dd=1.0, precision=precision, ddi_padding=padding)
    ad = G.AdaptiveDtParams(dt_init=1e-3)
    r = A.run_simulation_yoshida(ws, adaptive=ad, t_end=1e9, save_interval=0.05,
                                 keep_snapshots=False, max_accepted=warmup)
    ws.synchronize()
    ad2 = G.AdaptiveDtParams(dt_init=r["final_dt"])
    t0 = time.perf_counter()
    r2 = A.run_simulation_yoshida(ws, adaptive=ad2, t_end=1e9, save_interval=0.05,
                                  keep_snapshots=False, max_accepted=steps)
    ws.synchronize()
    t1 = time.perf_counter()
    if not profiled:
        return ws, grid, (t1 - t0), None, psi, r2, None
    ws.profile(True)
    A.run_simulation_yoshida(ws, adaptive=G.AdaptiveDtParams(dt_init=r2["final_dt"]), t_end=1e9,
                             save_interval=0.05, keep_snapshots=False,
                             max_accepted=max(2, steps // 4))
    prof = ws.profile_read()
    work = ws.profile_work()
    ws.profile(False)
    # one whole-psi copy after the measurements: the adaptive driver copies no psi, and the
    # --pmc traffic summary calibrates its counters on a copy of known size
    ws.copy_slot(1, 0)
    ws.synchronize()
    return ws, grid, (t1 - t0), prof, psi, r2, work


def fp32_companion(G, A, steps, warmup, padding, nsites, n_err=10):
    """The same spin-6 workload in FP32-bulk: its rate over the same number of accepted
    steps, and its error against the FP64 path on a fixed-dt run (n_err Yoshida cores at
    dt = 1e-3 from the same initial state in both precisions; relative L2 of psi over all
    components and relative energy difference)."""
    ws32, _, secs32, _, _, r32, _ = run_yoshida3d(G, A, steps, warmup, "fp32", padding=padding,
                                                   profiled=False)
    ws32.close()
    grid, inter, psi, _ = cfg3_inputs(G)
    out = {}
    for prec in ("fp64", "fp32"):
        ws = G.make_workspace(grid=grid, atom=G.Eu151, interactions=inter,
                              potential=G.HarmonicTrap(1.0, 1.0, 1.0),
                              sim_params=G.SimParams(dt=1e-3, n_steps=n_err), psi_init=psi,
                              enable_ddi=True, c_dd=1.0, precision=prec, ddi_padding=padding)
        for _ in range(n_err):
            A.yoshida_core(ws, 1e-3)
        out[prec] = (ws.get_psi(), A.total_energy(ws))
        ws.close()
    (p64, e64), (p32, e32) = out["fp64"], out["fp32"]
    rel_l2 = float(np.sqrt(np.sum(np.abs(p32 - p64) ** 2) / np.sum(np.abs(p64) ** 2)))
    return {"value": steps * nsites / secs32, "unit": "site-steps/s",
            "ms_per_step": secs32 / steps * 1e3, "accepted": r32["n_accepted"],
            "rejected": r32["n_rejected"],
            "dtype": "f32 (FP64 reductions, phase arguments, DDI kernel, renormalisation)",
            "err_vs_fp64": {"yoshida_cores": n_err, "dt": 1e-3, "rel_l2_psi": rel_l2,
                            "rel_energy": abs(e32 - e64) / abs(e64)}}


def cfg4_jobs(G, n_jobs=256, seed=42, n_steps=10000, sizes=(64, 128, 256)):
    """256 independent 2-D spin-1 ground states: sides uniform in {64, 128, 256} (seeded),
    box 20^2, c0 = 10, c1 ~ U[-1, 1], starts cycling polar / ferromagnetic / uniform /
    antiferromagnetic / seeded random, dt = 0.005, tol = 1e-8, n_steps = 10000 (the
    reference's default), no trap (none is stated): the ground state of the periodic box."""
    from becgpu.batch import ITPJob
    gen = G.physics.SplitMix64(seed)
    u = gen.uniform(2 * n_jobs)
    kinds = ("polar", "ferromagnetic", "uniform", "antiferromagnetic", "random")
    grids = {n: G.make_grid(G.GridConfig((n, n), (20.0, 20.0))) for n in sizes}
    jobs = []
    for k in range(n_jobs):
        n = sizes[min(len(sizes) - 1, int(u[2 * k] * len(sizes)))]
        kind = kinds[k % len(kinds)]
        psi0 = G.init_psi(grids[n], G.SpinSystem(1), kind, seed=seed + k)
        jobs.append(ITPJob(grid=grids[n], F=1, interactions=G.InteractionParams(10.0, 2 * u[2 * k + 1] - 1),
                           psi0=psi0, dt=0.005, n_steps=n_steps, tol=1e-8, tag=(n, kind)))
    return jobs


def converged_error(G, n_jobs=24, n_steps=200000, refine_steps=2000):
    """FP32-bulk against FP64 on relaxed ground states: the first n_jobs solves of the same
    seeded set on its 64^2 grid, run for n_steps in both precisions through the batched path
    with the convergence test switched off (tol = 0), so that both reach their fixed points
    instead of stopping at different checks (tol = 1e-8 on |dE| is a relative 1e-6 on these
    energies).  The error is taken over the solves whose FP64 energy changed by less than
    1e-10 (relative) over the last check interval, for the FP32-bulk solves and for the same
    solves finished by refine_steps FP64 steps (batch.solve_batch_refined)."""
    from becgpu.batch import solve_batch, solve_batch_refined
    jobs = cfg4_jobs(G, n_jobs, n_steps=n_steps, sizes=(64,))
    for j in jobs:
        j.tol = 0.0
    r32, _ = solve_batch(jobs, precision="fp32")
    r32r, _ = solve_batch_refined(jobs, refine_steps=refine_steps)
    r64, _ = solve_batch(jobs, precision="fp64")
    relaxed = [k for k, b in enumerate(r64) if b["dE"] < 1e-10 * abs(b["energy"])]
    rel = lambda a, b: abs(a["energy"] - b["energy"]) / abs(b["energy"])
    errs = [rel(r32[k], r64[k]) for k in relaxed]
    errs_r = [rel(r32r[k], r64[k]) for k in relaxed]
    return dict(max_rel_energy_err=max(errs) if errs else None,
                median_rel_energy_err=float(np.median(errs)) if errs else None,
                max_rel_energy_err_refined=max(errs_r) if errs_r else None,
                refine_steps_fp64=refine_steps,
                n_relaxed=len(relaxed), n_jobs=n_jobs, grid=[64, 64], n_steps=n_steps,
                fp64_last_rel_dE_max=max(b["dE"] / abs(b["energy"]) for b in r64))


def run_batch2d(G, A, n_jobs, n_plain, graphs=False, grouped=False):
    """Batched FP32-bulk and FP64 solves/s of the 256-job set, against the plain per-solve
    FP64 path (find_ground_state one job at a time) on an evenly spaced subsample."""
    from becgpu.batch import solve_batch
    jobs = cfg4_jobs(G, n_jobs)
    from becgpu.batch import solve_batch_refined
    r32, s32 = solve_batch(jobs, precision="fp32", graphs=graphs, grouped=grouped)
    _, s32r = solve_batch_refined(jobs, refine_steps=2000, graphs=graphs, grouped=grouped)
    r64, s64 = solve_batch(jobs, precision="fp64", graphs=graphs, grouped=grouped)
    sub = jobs[::max(1, n_jobs // n_plain)][:n_plain]
    t0 = time.perf_counter()
    for j in sub:
        r = A.find_ground_state(grid=j.grid, atom=G.Rb87, interactions=j.interactions,
                                potential=j.potential, dt=j.dt, n_steps=j.n_steps, tol=j.tol,
                                psi_init=j.psi0)
        r["workspace"].close()
    plain = len(sub) / (time.perf_counter() - t0)
    rel = [abs(a["energy"] - b["energy"]) / abs(b["energy"]) for a, b in zip(r32, r64)]
    conv = [e for e, a, b in zip(rel, r32, r64) if a["converged"] and b["converged"]]
    err = max(conv) if conv else None   # null in the JSON line when no pair converged
    lat = np.array([r["time_s"] for r in r32])
    conv_err = converged_error(G)
    return dict(fp32=s32, fp32_refined=s32r, fp64=s64, plain_fp64_solves_per_s=plain,
                plain_sample=len(sub),
                converged_error=conv_err,
                max_rel_energy_err_fp32=err, max_rel_energy_err_all=max(rel),
                n_both_converged=len(conv), converged_fp32=sum(r["converged"] for r in r32),
                steps_fp32=[r["steps"] for r in r32], p50_s=float(np.percentile(lat, 50)),
                p99_s=float(np.percentile(lat, 99)))


def _cpu_worker(job):
    """One single-threaded oracle worker of the CPU baseline: (units of work, seconds)."""
    kind, budget_s, kw = job
    from threadpoolctl import threadpool_limits
    with threadpool_limits(limits=1):
        return {"yoshida3d": _cpu_yoshida3d, "rt3d": _cpu_rt3d,
                "itp1d": _cpu_itp1d}[kind](budget_s, **kw)


def cpu_pool(kind, budget_s, **kw):
    """Run the oracle sample `kind` on every host core this job may use (at most 16, the
    host share of one GPU), one single-threaded process per core, all at once; the
    throughput is the sum over the processes.  Called before anything touches the GPU, so
    the worker processes are plain forks of a process without GPU state."""
    import multiprocessing as mp
    cores = min(len(os.sched_getaffinity(0)), 16)
    with mp.get_context("fork").Pool(cores) as pool:
        res = pool.map(_cpu_worker, [(kind, budget_s, kw)] * cores)
    units = sum(r[0] for r in res)
    wall = max(r[1] for r in res)
    return cores, units, wall, res[0][2]


def _cpu_yoshida3d(budget_s):
    """FP64 oracle on a bounded sample of the spin-6 dipolar workload: Yoshida attempts
    (Strang core + Yoshida core + L2 error, yoshida.jl:50-56) on the same configuration at
    32^3 (same box and couplings); one attempt is one accepted step when none is rejected."""
    import becgpu as G
    from oracle import spinor_oracle as O
    n = 32
    grid, inter, psi, c_extra = cfg3_inputs(G, n)
    og = O.make_grid((n,) * 3, (24.0,) * 3)
    ows = O.make_workspace(og, 6, psi, 1e-3, False, omega=(1, 1, 1), c0=10.0, c_extra=c_extra,
                           ddi=True, cdd=1.0)
    dt = 1e-3
    att = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < budget_s:
        old = ows.psi.copy()
        O.strang_core(ows, dt)
        ps = ows.psi.copy()
        ows.psi = old
        O.yoshida_core(ows, dt)
        O.l2_change(ows.psi, ps)
        att += 1
    return att * n ** 3, time.perf_counter() - t0, "Yoshida attempts (dt=1e-3) at 32^3"


def cpu_baseline_yoshida3d(budget_s=20.0):
    cores, units, wall, what = cpu_pool("yoshida3d", budget_s)
    return dict(value=units / wall, unit="site-steps/s", cores=cores, kind="port",
                sample=f"{units // 32 ** 3} {what} of the spin-6 dipolar workload (same box, "
                       f"couplings and DDI), FP64 numpy oracle (pocketfft, LAPACK eigh per "
                       f"site), {cores} single-threaded processes for {wall:.1f} s")


def _cpu_rt3d(budget_s, grid_n=128):
    """FP64 oracle (numpy/pocketfft, single thread) on a bounded sample of the spin-1
    real-time workload: leapfrog steps on the same input until ~budget_s."""
    import becgpu as G
    from oracle import spinor_oracle as O
    _, _, psi = _cfg2_inputs(G, n=grid_n)
    og = O.make_grid((grid_n,) * 3, (20.0,) * 3)
    ows = O.make_workspace(og, 1, psi, 1e-3, False, omega=(1, 1, 1), q=0.1, c0=10.0, c1=0.5,
                           n_steps=1000, save_every=10)
    O.half_potential_step(ows, 0.5e-3, False)
    n = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < budget_s:
        ows.psi = O.kinetic_step(ows.psi, ows.kphase)
        O.half_potential_step(ows, 1e-3, False)
        n += 1
    return n * grid_n ** 3, time.perf_counter() - t0, f"leapfrog steps (K + merged V(dt)) at {grid_n}^3"


def cpu_baseline_rt3d(grid_n=128, budget_s=15.0):
    cores, units, wall, what = cpu_pool("rt3d", budget_s, grid_n=grid_n)
    return dict(value=units / wall, unit="site-steps/s", cores=cores, kind="port",
                sample=f"{units // grid_n ** 3} {what} of the spin-1 real-time workload, FP64 "
                       f"numpy oracle, {cores} single-threaded processes for {wall:.1f} s")


def _cpu_itp1d(budget_s):
    from oracle import spinor_oracle as O
    og = O.make_grid((64,), (20.0,))
    psi = O.init_psi(og, 1, "polar")
    ows = O.make_workspace(og, 1, psi, 0.005, True, omega=(1.0,), c0=10.0, c1=-0.5, n_steps=5000)
    n = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < budget_s:
        O.split_step(ows)
        n += 1
    return n * 64, time.perf_counter() - t0, "ITP split_step! at 64 points"


def cpu_baseline_itp1d(budget_s=10.0):
    cores, units, wall, what = cpu_pool("itp1d", budget_s)
    return dict(value=units / wall, unit="site-steps/s", cores=cores, kind="port",
                sample=f"{units // 64} {what} of the 1-D spin-1 workload, FP64 numpy oracle, "
                       f"{cores} single-threaded processes for {wall:.1f} s")


def run_serve(G, n_req, rate, precision, n_check=10, sequential=False):
    """The request stream served by becgpu.serve, then an evenly spaced sample of the cheapest
    served requests re-run alone on the plain FP64 path: their energy error and their
    one-at-a-time service time."""
    from becgpu.serve import Server, poisson_mix, solve_alone
    reqs = poisson_mix(n_req, rate=rate, seed=42, precision=precision)
    srv = Server(keep_psi=False)
    res, stats = srv.run(reqs, realtime=True)
    lat = np.array([r.latency for r in res])
    wait = np.array([r.admitted - r.arrival for r in res])
    by = {r.rid: r for r in res}
    cost = sorted(reqs, key=lambda q: q.sites * (2 * q.F + 1) * q.n_steps)
    sample = cost[:max(1, len(cost) // 2)][::max(1, len(cost) // (2 * n_check))][:n_check]
    errs, alone_s, err_rows = [], [], []
    for q in sample:
        q64 = type(q)(**{**q.__dict__, "precision": "fp64"})
        t0 = time.perf_counter()
        a = solve_alone(q64, keep_psi=False)
        alone_s.append(time.perf_counter() - t0)
        errs.append(abs(by[q.rid].energy - a.energy) / abs(a.energy))
        err_rows.append([q.rid, q.kind, len(q.n), q.n[0], q.F, q.n_steps, errs[-1]])
    seq = None
    if sequential:
        t0 = time.perf_counter()
        for q in reqs:
            solve_alone(q, keep_psi=False)
        seq = time.perf_counter() - t0
    kinds = {}
    for nd in (1, 2, 3):
        sel = [r.latency for r in res if len(reqs[r.rid].n) == nd]
        if sel:
            kinds[f"{nd}d"] = dict(n=len(sel), p50_s=float(np.percentile(sel, 50)),
                                   p99_s=float(np.percentile(sel, 99)))
    return dict(stats=stats, solves_per_s=len(res) / stats["wall_s"],
                p50_s=float(np.percentile(lat, 50)), p99_s=float(np.percentile(lat, 99)),
                wait_p50_s=float(np.percentile(wait, 50)), wait_p99_s=float(np.percentile(wait, 99)),
                by_dim=kinds, arrivals_span_s=reqs[-1].arrival,
                converged_itp=sum(1 for r in res if r.converged),
                n_itp=sum(1 for q in reqs if q.kind == "itp"),
                errors=sum(1 for r in res if r.error),
                max_rel_energy_err_vs_fp64=max(errs) if errs else None,
                err_rows=err_rows, sample_alone_fp64_s=alone_s, sequential_s=seq)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="yoshida3d_spin6_ddi_128",
                    choices=["yoshida3d_spin6_ddi_128", "yoshida3d_spin6_ddipad_128",
                             "rt3d_spin1_128", "itp1d_spin1_64", "batch2d_itp_256", "serve_mix"])
    ap.add_argument("--jobs", type=int, default=256, help="batch2d_itp_256: number of solves")
    ap.add_argument("--requests", type=int, default=64, help="serve_mix: requests in the stream")
    ap.add_argument("--rate", type=float, default=200.0, help="serve_mix: mean arrivals per s")
    ap.add_argument("--graphs", action="store_true",
                    help="batch2d_itp_256: issue each group's steps as captured HIP graphs")
    ap.add_argument("--grouped", action="store_true",
                    help="batch2d_itp_256: one launch per pass over all grid shapes")
    ap.add_argument("--sequential", action="store_true",
                    help="serve_mix: also time every request alone, one after another")
    ap.add_argument("--precision", default=None, choices=["fp64", "fp32"])
    ap.add_argument("--side", type=int, default=128,
                    help="rt3d_spin1_128: cube side (the named workload is 128; other built "
                         "sides, e.g. 96 or 160, measure the mixed-radix FFTs)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp32", action="store_true",
                    help="yoshida3d_*: skip the FP32-bulk companion run of the FP64 line")
    ap.add_argument("--pmc-summary", nargs=2, metavar=("FETCH_CSV", "WRITE_CSV"),
                    help="write profiles/<workload>_traffic.json from two rocprofv3 --pmc runs")
    ap.add_argument("--copy-bytes", type=float, default=None,
                    help="--pmc-summary: bytes of the largest copyBuffer dispatch")
    args = ap.parse_args()
    if args.pmc_summary:
        prec = args.precision or ("fp32" if args.workload.startswith("serve") else "fp64")
        pmc_summary(args.pmc_summary[0], args.pmc_summary[1], args.copy_bytes, args.workload,
                    _traffic_path(args.workload, prec), prec)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; sharding across GPUs is a "
                                   "design note only, not built", "n_gpus": args.gpus}))
        sys.exit(2)
    if args.precision is None:
        args.precision = "fp32" if args.workload.startswith("serve") else "fp64"
    # the CPU baseline runs first, in forked single-threaded processes, before this process
    # touches the GPU
    cpu = None
    if not args.no_cpu:
        if args.workload == "rt3d_spin1_128":
            cpu = cpu_baseline_rt3d(grid_n=args.side)
        elif args.workload == "itp1d_spin1_64":
            cpu = cpu_baseline_itp1d()
        elif args.workload.startswith("yoshida3d_spin6"):
            cpu = cpu_baseline_yoshida3d()
    import becgpu as G
    from becgpu import api as A

    extra = {}
    if args.workload == "serve_mix":
        b = run_serve(G, args.requests, args.rate, args.precision, sequential=args.sequential)
        line = {
            "metric": "served solves/s with p50/p99 latency (mixed ITP and real-time requests)",
            "value": b["solves_per_s"], "unit": "solves/s", "n_gpus": 1, "steps": None,
            "warmup": 0, "ms_per_step": None, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None,
            "dtype": "f64" if args.precision == "fp64" else "f32 (FP64 reductions and phases)",
            "data": "synthetic",
            "config": {"workload": "Poisson request stream (becgpu.serve.poisson_mix, seed 42)",
                       "requests": args.requests, "rate_per_s": args.rate,
                       "dims": "1/2/3 w.p. 0.5/0.35/0.15", "F": "1/2/6 w.p. 0.6/0.25/0.15",
                       "sides": "every even side: 1-D/2-D 64..256, 3-D 48..128", "c0": "U[5,20]",
                       "c1": "U[-1,1]", "n_steps": "U[500,5000]", "itp_tol": 1e-8,
                       "slots_per_pool": 8, "hbm_budget_GiB": 200,
                       "scheduling": "8 pools in flight, shortest remaining work first"},
            "latency_p50_s": b["p50_s"], "latency_p99_s": b["p99_s"],
            "admission_wait_p50_s": b["wait_p50_s"], "admission_wait_p99_s": b["wait_p99_s"],
            "latency_by_dim": b["by_dim"], "arrivals_span_s": b["arrivals_span_s"],
            "wall_s": b["stats"]["wall_s"], "pools_created": b["stats"]["pools_created"],
            "peak_pool_GiB": b["stats"]["peak_pool_bytes"] / 2 ** 30,
            "itp_converged": f'{b["converged_itp"]}/{b["n_itp"]}', "errors": b["errors"],
            "max_rel_energy_err_vs_fp64": b["max_rel_energy_err_vs_fp64"],
            "err_sample": {"columns": ["rid", "kind", "dim", "side", "F", "n_steps",
                                       "rel_energy_err"], "rows": b["err_rows"]},
            "host_issue_s": b["stats"]["host_issue_s"],
            "sequential_alone_s": b["sequential_s"],
            "sample_alone_fp64_s": [round(v, 4) for v in b["sample_alone_fp64_s"]],
        }
        print(json.dumps(line))
        return
    if args.workload == "batch2d_itp_256":
        b = run_batch2d(G, A, args.jobs, n_plain=12, graphs=args.graphs, grouped=args.grouped)
        line = {
            "metric": "batched ITP solves/s (256 x 2-D spin-1, mixed 64^2/128^2/256^2)",
            "value": b["fp32"]["solves_per_s"], "unit": "solves/s", "n_gpus": 1,
            "steps": None, "warmup": 0, "ms_per_step": None, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f32 (FP64 reductions and phases)",
            "data": "synthetic",
            "config": {"workload": "batched find_ground_state, 2-D spin-1", "jobs": args.jobs,
                       "sides": [64, 128, 256], "box": 20.0, "c0": 10.0, "c1": "U[-1,1] seed 42",
                       "trap": "none",
                       "dt": 0.005, "tol": 1e-8, "n_steps": 10000, "graphs": bool(args.graphs),
                       "grouped": bool(args.grouped)},
            "grouped_intervals": b["fp32"].get("grouped_intervals", 0),
            "fp64_batched_solves_per_s": b["fp64"]["solves_per_s"],
            "fp32_refined_solves_per_s": b["fp32_refined"]["solves_per_s"],
            "fp32_refined_note": "FP32-bulk solves finished by 2000 FP64 steps "
                                 "(batch.solve_batch_refined); error in max_rel_energy_err_note",
            "max_rel_energy_err_fp32_refined_vs_fp64":
                b["converged_error"]["max_rel_energy_err_refined"],
            "plain_fp64_solves_per_s": b["plain_fp64_solves_per_s"],
            "speedup_vs_plain_fp64": b["fp32"]["solves_per_s"] / b["plain_fp64_solves_per_s"],
            "max_rel_energy_err_fp32_vs_fp64": b["converged_error"]["max_rel_energy_err"],
            "max_rel_energy_err_note": "converged ground states: " + json.dumps(b["converged_error"]),
            "max_rel_energy_err_same_jobs": b["max_rel_energy_err_fp32"],
            "max_rel_energy_err_incl_unconverged": b["max_rel_energy_err_all"],
            "n_converged_both": b["n_both_converged"],
            "latency_p50_s": b["p50_s"], "latency_p99_s": b["p99_s"],
            "converged": b["converged_fp32"], "groups": b["fp32"]["group_sizes"],
            "plain_sample": b["plain_sample"],
        }
        print(json.dumps(line))
        return
    if args.workload.startswith("yoshida3d_spin6"):
        pad = args.workload == "yoshida3d_spin6_ddipad_128"
        ws, grid, secs, prof, psi, r2, work = run_yoshida3d(G, A, args.steps, args.warmup,
                                                            args.precision, padding=pad)
        # Yoshida attempt = 6V + 4K + 3 psi copies + L2 reduction (B_alg per site-step)
        b_alg = (10852.5 if args.precision == "fp64" else 5426.3) if not pad else None
        cfg = {"workload": "3D spin-6 dipolar 128^3, adaptive Yoshida-4 (run_simulation_yoshida!)"
                           + (", zero-padded 256^3 DDI" if pad else ""),
               "grid": [128, 128, 128], "F": 6, "box": 24.0, "c0": 10.0, "c_dd": 1.0,
               "tensor": "c2..c12 ~ U[-1,1] seed 42", "dt_init": 1e-3, "ddi_padding": pad}
        extra = {"accepted": r2["n_accepted"], "rejected": r2["n_rejected"],
                 "final_dt": r2["final_dt"]}
        if args.precision == "fp64" and not args.no_fp32:
            extra["fp32_bulk"] = fp32_companion(G, A, args.steps, args.warmup, pad, nsites=128 ** 3)
    elif args.workload == "rt3d_spin1_128":
        n = args.side
        ws, grid, secs, prof, psi, work = run_rt3d(G, A, args.steps, args.warmup, args.precision,
                                                   n=n)
        b_alg = 288.0 if args.precision == "fp64" else 144.0
        cfg = {"workload": f"3D spin-1 real-time leapfrog {n}^3 (run_simulation!, save_every=10)",
               "grid": [n, n, n], "F": 1, "dt": 1e-3, "c0": 10.0, "c1": 0.5, "q": 0.1}
    else:
        ws, grid, secs, prof, psi, work = run_itp1d(G, A, args.steps, args.warmup, args.precision)
        b_alg = 384.0 if args.precision == "fp64" else 192.0
        cfg = {"workload": "1D spin-1 ITP split_step! 64 points", "grid": [64], "F": 1,
               "dt": 0.005, "c0": 10.0, "c1": -0.5}
    nsites = int(np.prod(grid.config.n_points))
    value = args.steps * nsites / secs
    roofs = _rooflines(prof, work, args.precision)
    # dominant kernel class by total time in the profiled pass
    dom = max((k for k in prof if prof[k][1] > 0 and k != "energy"), key=lambda k: prof[k][0])
    roof = dict(roofs.get(dom, {"bound": "hbm", "achieved": None, "peak": 8000.0, "unit": "GB/s",
                                "frac": None}))
    roof["kernel"] = dom
    # measured traffic is per workload at its named size
    tr = _traffic(args.workload, dom, args.precision) if args.side == 128 else None
    roof["traffic"] = tr
    if tr is not None:
        roof["traffic_source"] = os.path.relpath(_traffic_path(args.workload, args.precision),
                                                 ROOT)
    total_ms = sum(v[0] for v in prof.values())
    kernels = {k: {"avg_us": round(v[0] / v[1] * 1e3, 2), "launches": v[1],
                   "share": round(v[0] / max(1e-9, total_ms), 3)}
               for k, v in prof.items() if v[1] > 0}
    hbm = {k: {"GB/s": r["achieved"], "frac": r["frac"]} for k, r in roofs.items()
           if r["bound"] == "hbm"}
    line = {
        "metric": "spinor site-steps/s & %HBM peak",
        "value": value, "unit": "site-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": secs / args.steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32 (FP64 reductions and phases)",
        "data": "synthetic", "config": cfg,
        "step_hbm_frac": round(value * b_alg / HBM_PEAK, 4) if b_alg else None,
        "b_alg_per_site_step": b_alg,
        "roofline": roof, "hbm_kernels": hbm, "kernels": kernels, "cpu_baseline": cpu,
    }
    line.update(extra)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
