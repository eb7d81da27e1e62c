"""Autotuner of the stiffness kernel per (element type, p, points, precision).

For every group of a BatchedAssembly it times each tile-shape variant the library
offers for the group's padded node count (FL_TUNE_STIFF_VARIANT_BASE keys: 1 to 5 tile
shapes per class, fl_stiffness_variant_count) and, for FP64 and FP32 tensor-product
hexes, the sum-factorised kernel (FL_TUNE_STIFF_SF) with its two
data paths (element blocks + gather, or the colour-ordered direct scatter), using the
library's own per-phase HIP events summed over the whole assembly, and keeps the
fastest.  Results are cached as JSON keyed "hex-p2-G27-fp64" -> {"variant": v,
"sum_factorised": b, "direct": b, "assembly_ms": [...]}; the cache file is
$FLORENCE_AMD_TUNING_CACHE or ~/.cache/florence_amd/tuning.json.
"""
import json
import os

import numpy as np

from . import _lib

__all__ = ["cache_path", "load_cache", "save_cache", "tune_batch"]


def cache_path():
    return os.environ.get("FLORENCE_AMD_TUNING_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "florence_amd", "tuning.json"))


def load_cache(path=None):
    path = path or cache_path()
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        return json.load(f)


def save_cache(cache, path=None):
    path = path or cache_path()
    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path, "w") as f:
        json.dump(cache, f, indent=1, sort_keys=True)


def _label(sf, v, direct):
    if not sf:
        return "dense v{}".format(v)
    return "sum-factorised, direct scatter" if direct else "sum-factorised"


def tune_batch(batched, repeats=3, cache=None, force=False):
    """Pick each group's stiffness kernel and data path (mutates batched.groups[*]
    ['variant'], ['sum_factorised'] and ['direct']); returns the cache.  Candidates:
    every tile variant of the dense contraction and, for FP64/FP32 tensor-product hexes, the
    sum-factorised kernel with element blocks + gather and, where the group has a colour
    plan, with the direct scatter.  Timed: the sum of the assembly's phases."""
    import torch
    from .batch import _np_class
    cache = dict(cache or {})
    for g in batched.groups:
        key = batched.tuning_key(g["key"], g["fs"])
        npc = _np_class(g["dm"].npe)
        nvar = _lib.STIFF_VARIANTS[npc] if batched.precision != "bf16" else 1
        cands = [(False, v, False) for v in range(nvar)]
        if g["dm"].tensor_basis is not None and batched.precision in ("fp64", "fp32"):
            cands.append((True, 0, False))
            if batched.direct and g["dm"].scatter is not None:
                cands.append((True, 0, True))
        if key in cache and not force and "direct" in cache[key]:
            g["variant"] = int(cache[key]["variant"])
            g["sum_factorised"] = bool(cache[key]["sum_factorised"])
            g["direct"] = bool(cache[key]["direct"]) and batched.direct
            continue
        times = []
        for sf, v, direct in cands:
            g["variant"], g["sum_factorised"], g["direct"] = v, sf, direct
            batched._launch_group(g)            # warm-up (first-touch, kernel attributes)
            torch.cuda.synchronize()
            _lib.profile_enable(True)
            for _ in range(repeats):
                batched._launch_group(g)
            torch.cuda.synchronize()
            calls = _lib.profile_read()
            _lib.profile_enable(False)
            st = [sum(c[1]) for c in calls if c[0] == 1]
            times.append(float(np.median(st)))
        best = int(np.argmin(times))
        g["sum_factorised"], g["variant"], g["direct"] = cands[best]
        cache[key] = {"variant": int(cands[best][1]), "sum_factorised": bool(cands[best][0]),
                      "direct": bool(cands[best][2]), "candidates": [_label(*c) for c in cands],
                      "assembly_ms": [round(t, 4) for t in times], "nelem": g["dm"].nelem}
    return cache
