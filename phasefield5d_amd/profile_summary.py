"""Summarise rocprofv3 output of ``python -m phasefield5d_amd.profile_step`` per step kernel.

    python -m phasefield5d_amd.profile_summary --dir DIR --steps 20 \\
        --json profiles/step_3d128_fp32_counters.json --md profiles/step_3d128_fp32_counters.md

DIR holds one rocprofv3 run per subfolder, each written with ``-o run --output-format csv``:
``trace`` (``--kernel-trace --stats``) and any number of ``--pmc`` passes (``sq``, ``sq2``,
``fetch``, ``write`` ...).  Every kernel of the explicit step is mapped to its stage
(``fft_fwd``, ``fft_x``, ``fft_inv``, ``potentials``, ``flux``, ``update``, ``stencil``, ``dt``);
durations are the trace's average over the last ``--steps`` dispatches of that kernel (the
timed steps), counters the average over its dispatches.  ``FETCH_SIZE`` and ``WRITE_SIZE``
(KiB per dispatch) are scaled by what the same counter reads for the largest device-to-device
copy of the run (``profile_step --calibration-copy`` adds one of 256 MiB, which reads and
writes 256 MiB) into ``hbm_bytes``.  ``valu_frac`` is ``SQ_INSTS_VALU`` (wave instructions per
dispatch) over one VALU issue per CU per cycle: 256 CUs at 2.4 GHz for the dispatch's duration.
"""
import argparse
import collections
import csv
import json
import os
import re
import sys

CUS = 256
CLOCK_HZ = 2.4e9

# stage of each kernel of the step (name prefixes of the HIP kernels)
STAGES = (
    ("k_frp_fwd", "fft_fwd"), ("k_fftp_fwd", "fft_fwd"),
    ("k_frx", "fft_x"), ("k_fft<", "fft_x"),
    ("k_frp_inv", "fft_inv"), ("k_fftp_inv", "fft_inv"),
    ("k_potentials", "potentials"), ("k_fluxmarch", "flux"), ("k_fluxdiv", "flux"), ("k_stencilmarch", "flux"),
    ("k_update<", "update"), ("k_dtctl", "dt"),
)
COPY_KERNEL = "__amd_rocclr_copyBuffer"


def short_name(name):
    """``void pf5d::k_update<float, 256>(args)`` -> ``k_update<float, 256>``"""
    m = re.search(r"(k_\w+(<[^(]*>)?|__amd_rocclr_\w+)", name)
    return m.group(1) if m else name.split("(")[0][:80]


def stage_of(kname):
    for prefix, stage in STAGES:
        if kname.startswith(prefix):
            return stage
    return None


def read_csv(path):
    with open(path, newline="") as f:
        return list(csv.DictReader(f))


def summarise(d, steps):
    trace = collections.defaultdict(list)
    for r in read_csv(os.path.join(d, "trace", "run_kernel_trace.csv")):
        trace[short_name(r["Kernel_Name"])].append(int(r["End_Timestamp"]) - int(r["Start_Timestamp"]))
    counters = collections.defaultdict(lambda: collections.defaultdict(list))
    meta = {}
    copies = {}  # counter -> (grid, value) of the largest copy
    passes = []
    for sub in sorted(os.listdir(d)):
        path = os.path.join(d, sub, "run_counter_collection.csv")
        if sub == "trace" or not os.path.exists(path):
            continue
        passes.append(sub)
        per = collections.defaultdict(dict)
        for r in read_csv(path):
            k = short_name(r["Kernel_Name"])
            per[(k, r["Dispatch_Id"], int(r["Grid_Size"]))][r["Counter_Name"]] = float(r["Counter_Value"])
            # (rocprofv3's own fields, in its units)
            meta[k] = {c: int(r[c]) for c in ("VGPR_Count", "Accum_VGPR_Count", "SGPR_Count", "Scratch_Size",
                                              "LDS_Block_Size", "Workgroup_Size", "Grid_Size")}
        for (k, _, grid), cv in per.items():
            for c, v in cv.items():
                counters[k][c].append(v)
                if k == COPY_KERNEL and (c not in copies or grid > copies[c][0]):
                    copies[c] = (grid, v)
    # the calibration copy moves COPY_BYTES each way: 256 MiB
    copy_bytes = 256 * 1024 * 1024
    scale = {}
    for c in ("FETCH_SIZE", "WRITE_SIZE"):
        # (a run without the calibration copy has only small copies: no scale then)
        if c in copies and copies[c][1] * 1024.0 > copy_bytes / 8:
            scale[c] = copy_bytes / (copies[c][1] * 1024.0)
    kernels = {}
    for k, ds in trace.items():
        stage = stage_of(k)
        if stage is None:
            continue
        timed = ds[-steps:] if steps and len(ds) > steps else ds
        avg_us = sum(timed) / len(timed) / 1e3
        line = {"kernel": k, "dispatches": len(ds), "avg_us": round(avg_us, 3),
                "min_us": round(min(timed) / 1e3, 3), "max_us": round(max(timed) / 1e3, 3),
                # fewer dispatches than timed steps: a set-up launch (k_dtctl of a prime), not the step's
                "in_step": len(ds) >= steps}
        for c, vs in sorted(counters.get(k, {}).items()):
            line[c] = sum(vs) / len(vs)
        if "SQ_INSTS_VALU" in line:
            line["valu_frac"] = line["SQ_INSTS_VALU"] / (avg_us * 1e-6 * CUS * CLOCK_HZ)
        if "FETCH_SIZE" in line and "WRITE_SIZE" in line and len(scale) == 2:
            line["fetch_bytes"] = line["FETCH_SIZE"] * 1024.0 * scale["FETCH_SIZE"]
            line["write_bytes"] = line["WRITE_SIZE"] * 1024.0 * scale["WRITE_SIZE"]
            line["hbm_bytes"] = line["fetch_bytes"] + line["write_bytes"]
        if k in meta:
            line.update(meta[k])
        if stage in kernels:  # one kernel per stage is expected; keep the longer-running one
            if kernels[stage]["avg_us"] * kernels[stage]["dispatches"] >= avg_us * len(ds):
                continue
        kernels[stage] = line
    return {"kernels": kernels, "counter_passes": passes,
            "calibration": {"copy_bytes_each_way": copy_bytes,
                            "raw_kib": {c: v for c, (_, v) in copies.items() if c in scale},
                            "scale": scale},
            "valu_peak": f"{CUS} CUs x {CLOCK_HZ / 1e9:g} GHz, one wave64 VALU issue per CU per cycle"}


ORDER = ("fft_fwd", "fft_x", "fft_inv", "stencil", "potentials", "flux", "dt", "update")


def markdown(s, title, commands):
    out = [f"# {title}", "", "Commands (MI355X, ROCm 7.2):", "", "```"] + commands + ["```", ""]
    ks = s["kernels"]
    out += ["Per launch, averaged over the timed steps (durations) and over every dispatch (counters).",
            f"HBM bytes = FETCH_SIZE x {s['calibration']['scale'].get('FETCH_SIZE', float('nan')):.4f} + "
            f"WRITE_SIZE x {s['calibration']['scale'].get('WRITE_SIZE', float('nan')):.4f} "
            "(the factors that make the run's 256 MiB device-to-device copy read and write 256 MiB). "
            f"valu_frac = SQ_INSTS_VALU / (duration x {s['valu_peak']}).", ""]
    cols = ("avg_us", "hbm_bytes", "SQ_INSTS_VALU", "valu_frac", "SQ_BUSY_CYCLES", "SQ_WAIT_ANY",
            "SQ_LDS_BANK_CONFLICT", "SQ_WAVES", "VGPR_Count", "Scratch_Size", "LDS_Block_Size")
    out.append("| stage | kernel | " + " | ".join(cols) + " |")
    out.append("|---|---|" + "---|" * len(cols))

    def fmt(v):
        if v is None:
            return "-"
        if isinstance(v, float):
            return f"{v:.4g}" if abs(v) < 1e4 else f"{v:.4e}"
        return str(v)
    for st in ORDER:
        if st in ks:
            k = ks[st]
            name = f"`{k['kernel']}`" + ("" if k.get("in_step", True) else " (set-up launch, not in the step)")
            out.append(f"| {st} | {name} | " + " | ".join(fmt(k.get(c)) for c in cols) + " |")
    out.append("")
    return "\n".join(out)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--dir", required=True)
    ap.add_argument("--steps", type=int, default=20, help="timed steps of the profiled run")
    ap.add_argument("--json", required=True)
    ap.add_argument("--md", default=None)
    ap.add_argument("--title", default="Step kernels")
    ap.add_argument("--workload", default="")
    ap.add_argument("--command", action="append", default=[], help="a command that made the data (repeatable)")
    a = ap.parse_args(argv)
    s = summarise(a.dir, a.steps)
    s["workload"] = a.workload
    s["commands"] = a.command
    with open(a.json, "w") as f:
        json.dump(s, f, indent=1)
        f.write("\n")
    if a.md:
        with open(a.md, "w") as f:
            f.write(markdown(s, a.title, a.command))
    for st in ORDER:
        if st in s["kernels"]:
            k = s["kernels"][st]
            print(st, k["kernel"], k["avg_us"], k.get("hbm_bytes"), k.get("valu_frac"))
    return 0


if __name__ == "__main__":
    sys.exit(main())
