"""Autotuner for the stencil stage and the FFT passes.

For an engine holding its simulations, every compiled K1 variant -- the
two-pass forms (potentials tiles, with the flux pass's cells per lane and
block size; the marching flux pass combined with x-chunk lengths for 3D) and
the fused marching form (also with x-chunk lengths; only the marching forms
when the simulations need more than one cache-sized group) -- is timed with HIP events on the engine
stream (``Engine.time_kernel``, between the FFT passes and an update launch as
in a step); the fastest is kept.  Then the FFT: the line
passes' batching (LDS per item, block size), the plane passes of 3D fields
with 128 x 128 planes (parts per plane, block size), and the line batching
again for the x pass the planes leave.  Results are cached per (precision,
multiset of simulation shapes) in a JSON file next to the library so later
engines with the same mix start tuned.
"""
import json
import os
import threading

from . import _lib

CACHE_PATH = os.path.join(_lib.HERE, "autotune_cache.json")
_lock = threading.Lock()


def signature(engine):
    """Cache key: precision, the multiset of shapes, and the library's table of
    stencil variants (an entry names a variant by its index, which a rebuilt
    library with other variants would give another meaning)."""
    import hashlib
    shapes = sorted(tuple(pb.shape) for pb, _ in engine.slots.values())
    table = hashlib.sha1(repr(engine.variants()).encode()).hexdigest()[:12]
    return f"v9:{table}:{engine.precision}:" + ";".join("x".join(map(str, s)) for s in shapes)


def _load():
    try:
        with open(CACHE_PATH) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store(cache):
    tmp = CACHE_PATH + ".tmp"
    try:
        with open(tmp, "w") as f:
            json.dump(cache, f, indent=1, sort_keys=True)
        os.replace(tmp, CACHE_PATH)
    except OSError:
        pass


def tune(engine, chunks=(8, 16, 32), reps=10, use_cache=True, verbose=False, fft_budgets=(8, 16, 32, 64, 160),
         quick=False):
    """Pick the fastest (variant, x-chunk) and FFT batching for the engine's current simulations.

    ``quick`` narrows the FFT search to a few line and plane settings (for callers
    that tune on every change of their mix, such as the server)."""
    sig = signature(engine)
    with _lock:
        cache = _load() if use_cache else {}
    if sig in cache:
        ent = cache[sig]
        engine.set_tiling(ent["variant"], ent["cx"])
        if "fft_kb" in ent:
            engine.set_fft(ent["fft_kb"], ent["fft_nt"], ent.get("fft_parts", 0), ent.get("fft_plane_nt", 256),
                           wave=ent.get("fft_wave", 0))
        return ent
    has3d = any(len(pb.shape) == 3 for pb, _ in engine.slots.values())
    variants = engine.variants()
    # Only the marching forms (fcpl > 0) step simulations in cache-sized groups.
    # When the mix needs more than one group, each group's fields then stay in
    # the last-level cache through the step's passes, which timing the stencil
    # stage by itself does not see (64-run mix: 3.2 ms per step with a marching
    # form, 3.8 with the per-plane form whose stencil stage alone is faster), so
    # the candidates are the marching forms.
    march0 = next((v for v, t in enumerate(variants) if t[3] > 0), None)
    grouped = False
    if march0 is not None:
        engine.set_tiling(march0, chunks[0])
        grouped = engine.num_groups() > 1
    results = []
    for v, (by, bz, nt, fcpl, fnt) in enumerate(variants):
        # the x-chunk length applies to the marching passes (fcpl > 0)
        marching = fcpl > 0
        if grouped and not marching:
            continue
        for cx in (chunks if has3d and marching else chunks[:1]):
            try:
                engine.set_tiling(v, cx)
                # the stencil stage between the FFT passes and the update, as in a
                # step (what the other passes leave in the caches decides part of
                # its time); those passes are the same for every candidate
                ms = engine.time_kernel("step_like", reps=reps)
            except (ValueError, RuntimeError) as e:  # tiling does not fit this device
                if verbose:
                    print(f"  variant {v} ({by}x{bz}/{nt}) cx={cx}: skipped ({e})")
                continue
            results.append((ms, v, cx))
            if verbose:
                print(f"  variant {v} ({by}x{bz}/{nt}) cx={cx}: {ms:.4f} ms")
    if not results:
        raise RuntimeError("no stencil tiling could be launched")
    ms, v, cx = min(results)
    engine.set_tiling(v, cx)
    entry = {"variant": v, "cx": cx, "step_like_ms": ms, "k1_ms": engine.time_kernel("k1", reps=reps),
             "tile": list(engine.variants()[v])}
    # FFT passes: lines per block (LDS budget) and block size of the line passes,
    # then the plane passes (parts per plane and block size) of 3D fields with
    # 128 x 128 planes, then the line passes again for what the planes leave
    if any(pb.Bhat is not None for pb, _ in engine.slots.values()):
        def trial(kb, nt, parts, pnt, wave=0):
            try:
                engine.set_fft(kb, nt, parts, pnt, wave=wave)
                fms = engine.time_kernel("fft", reps=reps)
            except (ValueError, RuntimeError) as e:
                if verbose:
                    print(f"  fft {kb} KiB / {nt}, planes {parts} / {pnt}, wave {wave}: skipped ({e})")
                return None
            if verbose:
                print(f"  fft {kb} KiB / {nt} threads, planes {parts} / {pnt}, wave {wave}: {fms:.4f} ms")
            return (fms, kb, nt, parts, pnt, wave)

        budgets, nts = ((16,), (128, 256)) if quick else (fft_budgets, (128, 256, 512))
        plane_cfgs = ((2, 1024), (4, 1024)) if quick else tuple((p, n) for p in (2, 4, 8) for n in (256, 512, 1024))
        fres = [r for r in (trial(kb, nt, 0, 256) for kb in budgets for nt in nts) if r]
        if fres:
            best = min(fres)
            planes = [r for r in (trial(best[1], best[2], parts, pnt) for parts, pnt in plane_cfgs)
                      if r and engine.launch_info()["fft_parts"]]
            if planes and min(planes) < best:
                best = min(planes)
                if not quick:
                    again = [r for r in (trial(kb, nt, best[3], best[4]) for kb in (16, 32, 64)
                                         for nt in (128, 256)) if r]
                    best = min(again + [best])
            # the wave-resident passes of 128-point planes and x lines, where the mix has them
            waves = []
            for mode in ((1, 7) if quick else range(1, 8)):
                r = trial(best[1], best[2], best[3], best[4], mode)
                if r and engine.launch_info()["fft_wave"]:
                    waves.append(r)
            if waves and min(waves) < best:
                best = min(waves)
            fms, kb, nt, parts, pnt, wave = best
            engine.set_fft(kb, nt, parts, pnt, wave=wave)
            entry.update({"fft_kb": kb, "fft_nt": nt, "fft_parts": parts, "fft_plane_nt": pnt, "fft_wave": wave,
                          "fft_ms": fms})
    with _lock:
        cache = _load()
        cache[sig] = entry
        _store(cache)
    return entry
