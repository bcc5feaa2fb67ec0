"""Autotuner for the stencil stage (and, through it, the FFT line batching).

For an engine holding its simulations, every compiled K1 variant -- the
single-pass fused tilings (BY x BZ tile, threads, register budget) combined
with x-chunk lengths for 3D marching, and the two-pass form (reported as
tile 0 x 64) -- is timed with HIP events on the engine stream
(``Engine.time_kernel``); the fastest is kept.  Results are cached per
(precision, multiset of simulation shapes) in a JSON file next to the library
so later engines with the same mix start tuned.
"""
import json
import os
import threading

from . import _lib

CACHE_PATH = os.path.join(_lib.HERE, "autotune_cache.json")
_lock = threading.Lock()


def signature(engine):
    shapes = sorted(tuple(pb.shape) for pb, _ in engine.slots.values())
    return f"{engine.precision}:" + ";".join("x".join(map(str, s)) for s in shapes)


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


def tune(engine, chunks=(8, 16, 32), reps=10, use_cache=True, verbose=False, fft_budgets=(8, 16, 32, 64, 160)):
    """Pick the fastest (variant, x-chunk) and FFT batching for the engine's current simulations."""
    sig = signature(engine)
    with _lock:
        cache = _load() if use_cache else {}
    if sig in cache:
        ent = cache[sig]
        engine.set_tiling(ent["variant"], ent["cx"])
        if "fft_kb" in ent:
            engine.set_fft(ent["fft_kb"], ent["fft_nt"])
        return ent
    has3d = any(len(pb.shape) == 3 for pb, _ in engine.slots.values())
    results = []
    for v, (by, bz, nt, fcpl, fnt) in enumerate(engine.variants()):
        # the x-chunk length applies to the single-pass form (by > 0) and to the
        # marching flux pass of the two-pass form (fcpl > 0)
        marching = by > 0 or fcpl > 0
        for cx in (chunks if has3d and marching else chunks[:1]):
            try:
                engine.set_tiling(v, cx)
                ms = engine.time_kernel("k1", reps=reps)
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
    entry = {"variant": v, "cx": cx, "k1_ms": ms, "tile": list(engine.variants()[v])}
    # FFT passes: lines per block (LDS budget) and block size
    if any(pb.Bhat is not None for pb, _ in engine.slots.values()):
        fres = []
        for kb in fft_budgets:
            for nt in (128, 256, 512):
                try:
                    engine.set_fft(kb, nt)
                    fms = engine.time_kernel("fft", reps=reps)
                except (ValueError, RuntimeError) as e:
                    if verbose:
                        print(f"  fft {kb} KiB / {nt}: skipped ({e})")
                    continue
                fres.append((fms, kb, nt))
                if verbose:
                    print(f"  fft {kb} KiB / {nt} threads: {fms:.4f} ms")
        if fres:
            fms, kb, nt = min(fres)
            engine.set_fft(kb, nt)
            entry.update({"fft_kb": kb, "fft_nt": nt, "fft_ms": fms})
    with _lock:
        cache = _load()
        cache[sig] = entry
        _store(cache)
    return entry
