"""Run the explicit step of one named workload for a profiler.

    python -m phasefield5d_amd.profile_step --workload 3d_128 --precision fp32 --steps 20

Builds the workload's problem, sets the stencil tiling and FFT batching given on
the command line (or the autotuner's cached pick with ``--tuned``; otherwise the
engine's defaults), then runs ``--warmup`` and ``--steps`` graph-replayed steps
and, with ``--stages N``, N steps launched stage by stage.  Under
``rocprofv3 --kernel-trace --stats`` or ``--pmc`` every dispatch is a kernel of
the step, so per-kernel durations and counters need no filtering.  Prints the
launch configuration and the wall time per step as one JSON line.
"""
import argparse
import json
import sys
import time


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--workload", default="3d_128")
    ap.add_argument("--precision", default="fp32", choices=("fp32", "fp64"))
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--stages", type=int, default=0, help="steps launched stage by stage after the graph steps")
    ap.add_argument("--variant", type=int, default=None)
    ap.add_argument("--cx", type=int, default=None)
    ap.add_argument("--fft", type=int, nargs=4, metavar=("KB", "NT", "PARTS", "PLANE_NT"), default=None)
    ap.add_argument("--wave", type=int, default=None, help="FFT wave-resident mode (0 = off, 1..3)")
    ap.add_argument("--tuned", action="store_true", help="use the autotuner (cached result if present)")
    ap.add_argument("--sims", type=int, default=1, help="copies of the workload in the engine")
    ap.add_argument("--developed", type=float, default=0.0, metavar="AMP",
                    help="start from a developed field: smooth modes of amplitude AMP per component")
    ap.add_argument("--calibration-copy", action="store_true",
                    help="end with a 256 MiB device-to-device copy (to scale FETCH_SIZE / WRITE_SIZE)")
    a = ap.parse_args(argv)

    from . import _lib
    from .engine import Engine
    from .problem import build_problem
    from .workloads import workload_config

    _lib.require_gpu()
    pb = build_problem(workload_config(a.workload, seed=0))
    if a.developed > 0:
        import dataclasses
        from .workloads import developed_field
        pb = dataclasses.replace(pb, X_init=developed_field(pb, amp=a.developed, seed=4))
    eng = Engine(a.precision)
    for _ in range(a.sims):
        eng.add_problem(pb)
    if a.tuned:
        from . import autotune
        autotune.tune(eng, chunks=(4, 8, 16, 32), reps=5)
    if a.variant is not None:
        eng.set_tiling(a.variant, a.cx or 8)
    if a.fft is not None:
        kb, nt, parts, pnt = a.fft
        eng.set_fft(kb, nt, parts, pnt)
    if a.wave is not None:
        info = eng.launch_info()
        eng.set_fft(info["fft_budget"] // 1024, info["fft_nt"], wave=a.wave)
    eng.step(a.warmup)
    eng.sync()
    t0 = time.perf_counter()
    eng.step(a.steps)
    eng.sync()
    el = time.perf_counter() - t0
    stages = eng.time_step_stages(a.stages) if a.stages > 0 else None
    out = {"workload": a.workload, "precision": a.precision, "sims": a.sims, "shape": list(pb.shape),
           "steps": a.steps, "ms_per_step": 1e3 * el / a.steps, "launch": eng.launch_info(), "stages_ms": stages}
    eng.close()
    if a.calibration_copy:
        import torch
        src = torch.ones(64 * 1024 * 1024, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        out["calibration_copy_bytes"] = 2 * src.numel() * 4
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    sys.exit(main())
