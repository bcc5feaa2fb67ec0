"""Benchmark of the explicit step on one MI355X.

Workload: 3D 128^3 periodic (mw 8 x ppw 16), X0 = (0.1,0.2,0.3,0.2), T = 873 K,
kappa = 5e-16, elasticity on (c_ij 240/150/120 GPa, V_m 7.1e-6), synthetic
regular-solution table (step 0.025).  One step = interpolation + kappa lap(X) +
rank-1 elastic FFT pair + face-flux divergence + adaptive dt + update/clamp of
all 2,097,152 cells.  Inputs resident in HBM before the timed region.

Reported:
  value        FP32 fused device step (FP64 mass / dt / k=0 mode), cell-updates/s
  fp64_fused   same step, FP64 device-resident fused path
  fp64_plain   the per-function drop-in FP64 path (one device call per reference
               function, host dt from abs_max each step)
  error        max / RMS |X_fp32 - X_fp64| after the timed steps (same init)
  fp32_batched_128  eight 128^3 simulations in one FP32 engine (stepped in
               groups that fit the last-level cache), and its ratio to fp64_plain
  fp32_developed_field  the FP32 step on a decomposed-looking field (smooth
               modes of amplitude 0.05 per component), as later in a run
  fp32_fused_stencil  the same step with the fused marching stencil form
               (k_stencilmarch, the potentials pass on the tiles' boundary cells
               only; same bits), its in-step K1 time and 36 B/cell roofline
  roofline     the fused stencil stage K1 (interpolation + kappa lap + mu + face
               fluxes + divergence + block max) against its compulsory traffic,
               36 B/cell in FP32 (X 16 + psi 4 read, dX/dt 16 written): achieved =
               36 B x cells / the stage's duration inside the step (HIP events on
               the engine stream; with the two-pass form, the potentials pass plus
               the flux pass)
  roofline_step, roofline_fft  the whole step at 96 B/cell (24 words) and the FFT
               pair at 8 B/cell (s read, psi written), the same way
  roofline_passes  each pass of the two-pass stencil against its own bytes,
               which count the mu / Mob intermediates: 56 B/cell potentials (X,
               psi in; mu, Mob out), 68 B/cell flux (X, mu, Mob in; dX/dt out)
  valu         VALU instructions per launch of each kernel (rocprofv3
               SQ_INSTS_VALU, profiles/step_3d128_fp32_counters.json) over the
               launch's duration, as a fraction of the issue peak (one wave64
               instruction per CU per cycle at 2.4 GHz, 256 CUs)
  rocfft_ms    the same transform pair (rfftn, x Bhat, irfftn of the 128^3 misfit)
               through torch.fft (hipFFT / rocFFT) on the device, beside the
               hand-written passes' in-step time
  cpu_baseline C/OpenMP FP64 restatement of the reference loop + scipy.fft, on all
               host cores; cpu_baseline_1thread the same on one core
"""
import argparse
import dataclasses
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "cell-updates/s batched FP32 vs plain FP64 at <=1e-5 err; p50/p99 sim latency"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def cpu_baseline(pb, seconds=15.0, threads=None):
    """Time the C/OpenMP restatement of the reference loop on a bounded sample of the workload."""
    from oracle.cpu_path import CPath
    threads = threads or os.cpu_count()
    cp = CPath(pb.data_grid, pb.cfg.resolution, pb.tree, pb.calphad_values, pb.dx, pb.kappa, pb.X0, pb.l_vec,
               pb.Bhat, pb.ctld, pb.dt_cfl, pb.cfg.lower_limit, pb.cfg.upper_limit, threads=threads)
    X = pb.X_init.copy()
    dt = pb.dt0
    dt, _ = cp.step(X, dt)  # warm-up step
    n, t0 = 0, time.perf_counter()
    while True:
        dt, _ = cp.step(X, dt)
        n += 1
        el = time.perf_counter() - t0
        if el >= seconds or n >= 200:
            break
    return {"value": pb.ncell * n / el, "unit": "cell-updates/s", "cores": cp.threads(), "kind": "port",
            "sample": f"{n} steps of the 3D 128^3 FP64 step (C/OpenMP restatement on {cp.threads()} threads, "
                      f"scipy.fft 1 worker) after 1 warm-up step, {el:.1f} s"}


def rocfft_ms(pb, precision, reps=20):
    """The step's transform pair, psi = irfftn(Bhat rfftn(s)), through torch.fft
    (hipFFT / rocFFT) on the device: average ms per pair over ``reps`` (CUDA events)."""
    import torch
    dt = torch.float32 if precision == "fp32" else torch.float64
    s = ((pb.X_init - pb.X0) * pb.l_vec).sum(axis=-1)
    sd = torch.from_numpy(s).to(device="cuda", dtype=dt)
    bd = torch.from_numpy(pb.Bhat).to(device="cuda", dtype=dt)
    dims = tuple(range(s.ndim))

    def pair():
        return torch.fft.irfftn(torch.fft.rfftn(sd, dim=dims) * bd, s=tuple(s.shape), dim=dims)
    for _ in range(3):
        pair()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(reps):
        pair()
    b.record()
    b.synchronize()
    return a.elapsed_time(b) / reps


def batch_leg(steps, warmup, log_fn):
    """64 uneven simulations (32 2D, 32 3D) in one FP32 engine, one graph per step."""
    from phasefield5d_amd import autotune
    from phasefield5d_amd.engine import Engine
    from phasefield5d_amd.problem import build_problem
    from phasefield5d_amd.workloads import batch_mix
    import torch
    t0 = time.perf_counter()
    pbs = [build_problem(c, table=tab) for c, tab in batch_mix()]
    log_fn(f"batch mix set up in {time.perf_counter() - t0:.1f} s")
    eng = Engine("fp32")
    for pb in pbs:
        eng.add_problem(pb)
    tuned = autotune.tune(eng, chunks=(8, 16), reps=3)
    eng.step(warmup)
    eng.prepare(steps)  # the graphs step(steps) replays, captured outside the timed region
    eng.sync()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    eng.step(steps)
    eng.sync()
    el = time.perf_counter() - t0
    cells = sum(pb.ncell for pb in pbs)
    sts = eng.all_states()
    active = sum(1 for s in sts.values() if s["status"] & 1)
    # FP64 audit of a sample: the first 2D and the first and largest 3D simulation,
    # each run alone on an FP64 engine for the same steps
    slots = sorted(eng.slots)
    pick = [next(i for i, p in enumerate(pbs) if len(p.shape) == 2),
            next(i for i, p in enumerate(pbs) if len(p.shape) == 3),
            max(range(len(pbs)), key=lambda i: pbs[i].ncell)]
    audit = []
    for i in dict.fromkeys(pick):
        X32 = eng.field(slots[i], "X")
        e64 = Engine("fp64")
        s64 = e64.add_problem(pbs[i])
        e64.step(warmup + steps)
        d = np.abs(X32 - e64.field(s64, "X"))
        audit.append({"shape": list(pbs[i].shape), "steps": warmup + steps, "max": float(d.max()),
                      "rms": float(np.sqrt((d ** 2).mean()))})
        e64.close()
    eng.close()
    return {"sims": len(pbs), "cells": cells, "steps": steps, "cell_updates_per_s": cells * steps / el,
            "ms_per_step": 1e3 * el / steps, "active_after": active, "tiling": tuned,
            "audit_vs_fp64": audit}


def serve_leg(duration, rate, max_steps, log_fn):
    """Poisson arrivals of mixed 1D/2D/3D requests into one FP32 continuous-batching server."""
    from phasefield5d_amd.serve import Server, audit, poisson_requests, summarize
    reqs = poisson_requests(rate, duration, seed=0, max_steps=max_steps, audit_frac=0.1)
    srv = Server("fp32", poll_steps=25, setup_workers=6)
    done = srv.run(reqs, max_wall=duration * 6 + 60)
    srv.close()
    for r in done:
        if r.audit:
            audit(r)
    out = summarize(done, srv.wall)
    out.update({"arrival_rate_per_s": rate, "arrival_window_s": duration, "submitted": len(reqs),
                "max_steps_per_request": max_steps, "gpu_busy_s": srv.busy,
                "audited_fraction": sum(1 for r in done if r.audit) / max(1, len(done)),
                "variant": f"reduced serving workload: Poisson arrivals at {rate}/s for {duration:g} s (the "
                           f"workload names 10 min), each request ending at the earlier of 50 CTLD and "
                           f"{max_steps} CFL-limited steps (50 CTLD is 1e5-1e6 steps at the reference's dx); "
                           f"audited requests start from the host's seeded draw and are re-run alone in FP64"})
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--plain-steps", type=int, default=10)
    ap.add_argument("--error-steps", type=int, default=5000,
                    help="steps of the untimed FP32-vs-FP64 precision run (0: only the timed steps)")
    ap.add_argument("--variant", type=int, default=None)
    ap.add_argument("--cx", type=int, default=None)
    ap.add_argument("--calibration-copy", action="store_true",
                    help="end with a 256 MiB device-to-device copy (scales FETCH_SIZE / WRITE_SIZE counters)")
    ap.add_argument("--fp32-only", action="store_true", help="only the FP32 fused leg (profiling runs)")
    ap.add_argument("--no-developed", action="store_true",
                    help="skip the developed-field FP32 leg (profiling runs of the headline workload)")
    ap.add_argument("--no-mix", action="store_true", help="skip the 64-run batched mix leg")
    ap.add_argument("--mix-steps", type=int, default=20)
    ap.add_argument("--no-serve", action="store_true", help="skip the serving leg")
    ap.add_argument("--serve-seconds", type=float, default=20.0)
    ap.add_argument("--serve-rate", type=float, default=4.0)
    ap.add_argument("--serve-max-steps", type=int, default=300)
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the explicit step of one simulation "
              "and the batched engine both live on a single device in this build.", file=sys.stderr)
        sys.exit(2)

    import torch
    from phasefield5d_amd import _lib
    from phasefield5d_amd.engine import Engine
    from phasefield5d_amd.plain import PlainStepper
    from phasefield5d_amd.problem import build_problem
    from phasefield5d_amd.workloads import workload_config

    _lib.require_gpu()
    cfg = workload_config("3d_128", seed=0)
    pb = build_problem(cfg)
    ncell = pb.ncell
    log(f"workload 3D {pb.shape} ncell={ncell} ctld={pb.ctld:.4g} dt0={pb.dt0:.3g} dt_cfl={pb.dt_cfl:.3g}")

    def run_engine(precision, steps, warmup):
        eng = Engine(precision)
        slot = eng.add_problem(pb)
        if args.variant is not None or args.cx is not None:
            eng.set_tiling(args.variant or 0, args.cx or 16)
        else:
            from phasefield5d_amd import autotune
            tuned = autotune.tune(eng, chunks=(4, 8, 16, 32), reps=5)
            log(f"{precision} autotuned: {tuned}")
        eng.step(warmup)
        eng.prepare(steps)  # the graphs step(steps) replays, captured outside the timed region
        eng.sync()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        eng.step(steps)
        eng.sync()
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        return eng, slot, el

    # FP32 fused (headline)
    log("fp32 fused: warmup", args.warmup, "steps", args.steps)
    e32, s32, el32 = run_engine("fp32", args.steps, args.warmup)
    v32 = ncell * args.steps / el32
    k1_ms = e32.time_kernel("k1", reps=20)
    fft_ms = e32.time_kernel("fft", reps=10)
    k2_ms = e32.time_kernel("k2prime", reps=10)
    dt_ms = e32.time_kernel("dt", reps=10)
    info = e32.launch_info()
    # every stencil form has a potentials and a flux pass: by == 0 (the potentials
    # pass on every cell) or -1 (the fused marching form, the potentials pass on
    # the boundary cells)
    kp_ms = e32.time_kernel("k1_potentials", reps=20)
    kf_ms = e32.time_kernel("k1_flux", reps=20)
    log(f"fp32: {v32:.4g} cell-updates/s, {1e3 * el32 / args.steps:.3f} ms/step; K1 {k1_ms:.3f} ms, FFT {fft_ms:.3f} ms",
        info)
    # per-stage durations inside whole steps (HIP events between the stages of steps
    # launched one by one), on a second engine holding the same workload, so the
    # headline simulation keeps its step count
    march = e32.variants()[info["variant"]][3] > 0
    in_step = in_step_sep = None
    if march:
        et = Engine("fp32")
        et.add_problem(pb)
        et.set_tiling(info["variant"], info["cx"])
        et.set_fft(info["fft_budget"] // 1024, info["fft_nt"], info["fft_parts"], info["fft_plane_nt"],
                   wave=info["fft_wave"])
        et.step(args.warmup)
        et.sync()
        in_step = et.time_step_stages(min(max(args.steps, 50), 4096))
        in_step_sep = et.time_step_stages(50, separate_dt=True)
        et.close()
        log("fp32 stages in the step (ms)", in_step, "dt as its own launch", in_step_sep)
    st32 = e32.state(s32)
    if not (st32["status"] & 1) or st32["steps"] != args.warmup + args.steps:
        raise RuntimeError(f"fp32 simulation stopped early: {st32}")

    # the same step on a decomposed-looking field (neighbouring cells in different
    # table hypercubes, some past the linear region), as later in a run
    from phasefield5d_amd.workloads import developed_field
    developed = None
    ed = None if args.no_developed else Engine("fp32")
    if ed is not None:
        pbd = dataclasses.replace(pb, X_init=developed_field(pb, amp=0.05, seed=4))
        sd_ = ed.add_problem(pbd)
        ed.set_tiling(info["variant"], info["cx"])
        ed.set_fft(info["fft_budget"] // 1024, info["fft_nt"], info["fft_parts"], info["fft_plane_nt"],
                   wave=info["fft_wave"])
        ed.step(5)
        ed.prepare(50)
        ed.sync()
        t0 = time.perf_counter()
        ed.step(50)
        ed.sync()
        eld = time.perf_counter() - t0
        developed = {"amplitude": 0.05, "steps": 50, "ms_per_step": 1e3 * eld / 50, "cell_updates_per_s": ncell * 50 / eld,
                     "k1_potentials_ms": ed.time_kernel("k1_potentials", reps=10),
                     "active": bool(ed.state(sd_)["status"] & 1)}
        ed.close()
        log("fp32 developed field", developed)

    # the fused marching stencil form (k_stencilmarch; the potentials pass on the
    # tiles' boundary rows and planes only), same FFT batching, runs of 8 planes:
    # whole steps and the two K1 passes inside them
    fused_leg = None
    if not args.fp32_only:
        ef = Engine("fp32")
        sf = ef.add_problem(pb)
        fv = next((i for i, v in enumerate(ef.variants()) if v[0] == -1), None)
        if fv is not None:
            ef.set_tiling(fv, 8)
            ef.set_fft(info["fft_budget"] // 1024, info["fft_nt"], info["fft_parts"], info["fft_plane_nt"],
                       wave=info["fft_wave"])
            ef.step(args.warmup)
            ef.prepare(args.steps)
            ef.sync()
            t0 = time.perf_counter()
            ef.step(args.steps)
            ef.sync()
            elf = time.perf_counter() - t0
            taken = ef.launch_info()["variant"] == fv
            stf = ef.time_step_stages(50) if taken else None
            fused_leg = {"variant": fv, "cx": 8, "taken": taken, "steps": args.steps,
                         "ms_per_step": 1e3 * elf / args.steps, "cell_updates_per_s": ncell * args.steps / elf,
                         "stage_ms_in_step": stf,
                         "k1_ms_in_step": (stf["k1_potentials"] + stf["k1_flux"]) if stf else None}
            # (the trajectory is the two-pass form's, bit for bit: tests/test_gpu_stencilmarch.py)
        ef.close()
        log("fp32 fused marching stencil", fused_leg)

    # batched FP32 at 128^3: eight simulations of the workload in one engine (the
    # engine steps them in groups that fit the last-level cache)
    batched128 = None
    if not args.fp32_only:
        nb = 8
        eb = Engine("fp32")
        for _ in range(nb):
            eb.add_problem(pb)
        eb.set_tiling(info["variant"], info["cx"])
        eb.set_fft(info["fft_budget"] // 1024, info["fft_nt"], info["fft_parts"], info["fft_plane_nt"],
                   wave=info["fft_wave"])
        eb.step(8)
        kb = 48
        eb.prepare(kb)
        eb.sync()
        t0 = time.perf_counter()
        eb.step(kb)
        eb.sync()
        elb = time.perf_counter() - t0
        batched128 = {"sims": nb, "steps": kb, "groups": eb.num_groups(), "cell_updates_per_s": nb * ncell * kb / elb,
                      "ms_per_step": 1e3 * elb / kb}
        eb.close()
        log("fp32 batched 8 x 128^3", batched128)

    mass0 = pb.X_init.reshape(-1, 4).sum(axis=0)
    drift32 = np.abs(st32["mass"] - mass0) / mass0
    v64 = k1_64 = el64 = vplain = fft_64 = None
    diff = np.zeros(1)
    long_err = None
    if not args.fp32_only:
        # FP64 fused, same number of steps from the same state (error reference)
        log("fp64 fused")
        e64, s64, el64 = run_engine("fp64", args.steps, args.warmup)
        v64 = ncell * args.steps / el64
        k1_64 = e64.time_kernel("k1", reps=10)
        fft_64 = e64.time_kernel("fft", reps=10)
        X32, X64 = e32.field(s32, "X"), e64.field(s64, "X")
        diff = np.abs(X32 - X64)
        log(f"fp64 fused: {v64:.4g} cell-updates/s; max|dX| {diff.max():.3e} rms {np.sqrt((diff**2).mean()):.3e}")
        # the precision run: both engines continue, untimed, to the run length of the
        # 3D 128^3 precision workload (5000 steps); composition error and mass drift there
        done = args.warmup + args.steps
        if args.error_steps > done:
            t0 = time.perf_counter()
            for eng in (e32, e64):
                left = args.error_steps - done
                while left > 0:
                    eng.step(min(left, 500))
                    left -= 500
                eng.sync()
            st32e, st64e = e32.state(s32), e64.state(s64)
            X32e, X64e = e32.field(s32, "X"), e64.field(s64, "X")
            de = np.abs(X32e - X64e)
            long_err = {"steps": args.error_steps, "steps_fp32": st32e["steps"], "steps_fp64": st64e["steps"],
                        "max": float(de.max()), "rms": float(np.sqrt((de ** 2).mean())),
                        "t_fp32": st32e["t"], "t_fp64": st64e["t"],
                        "active": bool(st32e["status"] & 1 and st64e["status"] & 1),
                        "mass_drift_fp32": (np.abs(st32e["mass"] - mass0) / mass0).tolist(),
                        "mass_drift_fp64": (np.abs(st64e["mass"] - mass0) / mass0).tolist(),
                        "field_mass_drift_fp32": (np.abs(X32e.reshape(-1, 4).sum(axis=0) - mass0) / mass0).tolist(),
                        "seconds": time.perf_counter() - t0}
            log("fp32 vs fp64 over the precision run", long_err)
        del e64

        # plain FP64 drop-in path
        log("fp64 plain drop-in")
        ps = PlainStepper(pb)
        ps.step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.plain_steps):
            ps.step()
        torch.cuda.synchronize()
        elp = time.perf_counter() - t0
        vplain = ncell * args.plain_steps / elp
        log(f"fp64 plain: {vplain:.4g} cell-updates/s")
        del ps

    if args.calibration_copy:
        a = torch.ones(64 * 1024 * 1024, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b

    # per-kernel counters of the profiled step (rocprofv3; the file names its commands)
    counters = {}
    cpath = os.path.join(ROOT, "profiles", "step_3d128_fp32_counters.json")
    if os.path.exists(cpath):
        with open(cpath) as f:
            counters = json.load(f).get("kernels", {})

    def traffic_of(*stages, table=None):
        table = counters if table is None else table
        vals = [table[s_]["hbm_bytes"] for s_ in stages if s_ in table and "hbm_bytes" in table[s_]]
        return sum(vals) if len(vals) == len(stages) else None

    def roofline(kname, bpc, kms, traffic=None, alone_ms=None):
        """achieved = algorithmic bytes (bpc per cell) / the stage's duration inside the
        step (kms, HIP events on the engine stream); traffic: the HBM bytes rocprofv3
        counted for it (FETCH_SIZE + WRITE_SIZE, calibrated)"""
        achieved = bpc * ncell / (kms * 1e-3) / 1e9
        r = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
             "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "kernel": kname, "bytes_per_cell": bpc,
             "avg_ms": kms}
        if alone_ms is not None:
            r["alone_ms"] = alone_ms
            r["frac_alone"] = bpc * ncell / (alone_ms * 1e-3) / 1e9 / HBM_PEAK_GBS
        return r

    # the fused form's counters (its own profiled step)
    fcounters = {}
    fpath = os.path.join(ROOT, "profiles", "step_3d128_fp32_fused_stencil_counters.json")
    if os.path.exists(fpath):
        with open(fpath) as f:
            fcounters = json.load(f).get("kernels", {})
    fused_default = e32.variants()[info["variant"]][0] == -1
    K1_TWO_PASS = ("K1 stencil stage: k_potentials (interpolation from the LDS node box + kappa lap + psi -> mu, "
                   "Mob) + k_fluxmarch (face fluxes + divergence + block max; dt controller in its last block)")
    K1_FUSED = ("K1 stencil stage, fused marching form: k_potentials on the tiles' boundary rows and planes + "
                "k_stencilmarch (mu, Mob of the other cells, face fluxes, divergence, block max, dt controller)")
    if fused_leg and fused_leg.get("k1_ms_in_step"):
        fused_leg["roofline_k1"] = roofline(K1_FUSED, 36, fused_leg["k1_ms_in_step"],
                                            traffic_of("potentials", "flux", table=fcounters))
    kp_in = in_step["k1_potentials"] if in_step else kp_ms
    kf_in = in_step["k1_flux"] if in_step else kf_ms
    roof = roofline(K1_FUSED if fused_default else K1_TWO_PASS, 36, kp_in + kf_in,
                    traffic_of("potentials", "flux", table=fcounters if fused_default else None),
                    alone_ms=kp_ms + kf_ms)
    passes = [roofline("k_potentials (its bytes include the mu, Mob it writes)", 56, kp_in,
                       traffic_of("potentials"), kp_ms),
              roofline("k_fluxmarch (its bytes include the mu, Mob it reads)", 68, kf_in, traffic_of("flux"), kf_ms)]
    step_ms = in_step["step"] if in_step else 1e3 * el32 / args.steps
    roof_step = roofline("whole step (FFT pair, K1, dt, update)", 96, 1e3 * el32 / args.steps,
                         traffic_of("fft_fwd", "fft_x", "fft_inv", "potentials", "flux", "update"))
    roof_step["in_step_events_ms"] = step_ms
    fft_in = in_step["fft"] if in_step else fft_ms
    roof_fft = roofline("FFT pair (forward plane, x with Bhat, inverse plane passes)", 8, fft_in,
                        traffic_of("fft_fwd", "fft_x", "fft_inv"))
    # VALU issue fraction: instructions per launch (counters) over the in-step duration
    peak_inst = 256 * 2.4e9
    live = {"potentials": in_step["k1_potentials"] if in_step else kp_ms,
            "flux": in_step["k1_flux"] if in_step else kf_ms,
            "update": in_step["k2"] if in_step else k2_ms}
    valu = {}
    for stage_, c in counters.items():
        if "SQ_INSTS_VALU" not in c:
            continue
        ms = live.get(stage_) or c.get("avg_us", 0) * 1e-3
        valu[stage_] = {"kernel": c.get("kernel"), "insts": c["SQ_INSTS_VALU"], "avg_ms": ms,
                        "valu_frac": c["SQ_INSTS_VALU"] / (ms * 1e-3 * peak_inst) if ms else None,
                        "duration_from": "in-step HIP events" if stage_ in live else "profiled trace"}
    roc = None
    if not args.fp32_only:
        roc = {"fp32_ms": rocfft_ms(pb, "fp32"), "fp64_ms": rocfft_ms(pb, "fp64"), "hand_fp32_in_step_ms": fft_in,
               "hand_fp32_alone_ms": fft_ms, "hand_fp64_alone_ms": fft_64}
        log("rocFFT pair", roc)
    out = {
        "metric": METRIC, "value": v32, "unit": "cell-updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * el32 / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "3D 128^3 periodic explicit step, FP32 fused device path (FP64 mass, dt, k=0 mode), "
                               "elasticity on, synthetic regular-solution table; value is one simulation (a batch "
                               "of one), fp32_batched_128 holds eight in one engine",
                   "cells": ncell, "shape": list(pb.shape), "precision": "fp32", "elastic": True},
        "roofline": roof,
        "roofline_step": roof_step,
        "roofline_fft": roof_fft,
        "roofline_passes": passes,
        "valu": valu,
        "cache_note": "one 128^3 FP32 simulation's per-step working set (about 200 MB) fits the 256 MB Infinity "
                      "Cache, so part of the traffic these fractions count against 8 TB/s HBM is served there",
        "rocfft": roc,
        "fp64_fused": {"value": v64, "ms_per_step": 1e3 * el64 / args.steps if el64 else None, "k1_ms": k1_64},
        "fp64_plain": {"value": vplain, "steps": args.plain_steps},
        "ratio_fp32_vs_plain_fp64": v32 / vplain if vplain else None,
        "fp32_batched_128": batched128,
        "ratio_batched_128_vs_plain_fp64": (batched128["cell_updates_per_s"] / vplain
                                            if batched128 and vplain else None),
        "error_vs_fp64": {"max": float(diff.max()), "rms": float(np.sqrt((diff ** 2).mean())),
                          "steps": args.warmup + args.steps},
        "mass_drift_fp32": drift32.tolist(),
        "error_vs_fp64_long": long_err,
        "fp32_developed_field": developed,
        "fp32_fused_stencil": fused_leg,
        "stage_ms": {"k1": k1_ms, "k1_potentials": kp_ms, "k1_flux": kf_ms, "fft": fft_ms, "k2": k2_ms,
                     "dt": dt_ms},
        "stage_ms_in_step": in_step,
        "stage_ms_in_step_dt_launch": in_step_sep,
        "launch": info,
    }
    if not args.fp32_only and not args.no_mix:
        log("batched mix (64 runs)")
        out["batched_mix"] = batch_leg(args.mix_steps, 3, log)
        log("mix", out["batched_mix"])
    if not args.fp32_only and not args.no_serve:
        log("serving")
        out["serving"] = serve_leg(args.serve_seconds, args.serve_rate, args.serve_max_steps, log)
        log("serving", out["serving"])
    if not args.no_cpu:
        log("cpu baseline")
        out["cpu_baseline"] = cpu_baseline(pb, seconds=args.cpu_seconds)
        log("cpu", out["cpu_baseline"])
        out["cpu_baseline_1thread"] = cpu_baseline(pb, seconds=min(args.cpu_seconds, 10.0), threads=1)
        log("cpu 1 thread", out["cpu_baseline_1thread"])
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
