"""ORACLE -- test infrastructure, not product code.

Comparison of a simulation result with the oracle's, used by tests/ and by bench.py's CPU
baseline leg (accuracy of the sampled rays).

calc_diffs restates the cross-backend report of
apps/rmagine_simulation_diff_calc/Main.cpp:26-147: per attribute, rays where both results
hit, only one of them hits, neither hits, the rays where both hit with a difference above
the tolerance (1e-4), the largest and the average difference where both hit.  A "hit" is
hits == 1 for the hit mask, range != range.max + 1 for ranges (the invalid value,
sensor_models.h:59-63; the tool hard-codes max_range = 130) and a non-NaN x for points and
normals; the difference is != for hits, |a - b| for ranges and the Euclidean distance for
vectors.

compare() adds what the parity bar needs: exact mismatch counts per attribute (bits, NaN
patterns included), the rays the oracle flags as ambiguous for an FP32 intersector (per
flag bit), and how many mismatching rays are NOT flagged (0 is the bar).
"""
from __future__ import annotations

import numpy as np

from . import FLAG_EDGE, FLAG_NEAR_MISS, FLAG_RANGE, FLAG_TIE

DIFF_TOLERANCE = 1e-4


def _has_hit(kind: str, x: np.ndarray, range_max: float) -> np.ndarray:
    if kind == "hits":
        return x == 1
    if kind == "ranges":
        return x != np.float32(np.float32(range_max) + np.float32(1.0))
    return ~np.isnan(x[:, 0])


def _diff(kind: str, a: np.ndarray, b: np.ndarray) -> np.ndarray:
    if kind == "hits":
        return (a != b).astype(np.float64)
    if kind == "ranges":
        return np.abs(a.astype(np.float64) - b.astype(np.float64))
    return np.linalg.norm(a.astype(np.float64) - b.astype(np.float64), axis=1)


def calc_diffs(kind: str, a: np.ndarray, b: np.ndarray, range_max: float, tol: float = DIFF_TOLERANCE) -> dict:
    """Main.cpp:83-147 for one attribute ("hits", "ranges", "points" or "normals") of two
    results a (e.g. the GPU's) and b (e.g. the oracle's)."""
    if a.shape != b.shape:
        raise ValueError("sizes must be equal")
    ha, hb = _has_hit(kind, a, range_max), _has_hit(kind, b, range_max)
    both = ha & hb
    d = _diff(kind, a[both], b[both])
    return {"both_hit": int(both.sum()), "only_a": int((ha & ~hb).sum()), "only_b": int((~ha & hb).sum()),
            "neither": int((~ha & ~hb).sum()), "both_hit_high_diff": int((d > tol).sum()),
            "max_diff": float(d.max()) if d.size else 0.0, "avg_diff": float(d.mean()) if d.size else 0.0}


def _bits(x: np.ndarray) -> np.ndarray:
    x = np.ascontiguousarray(x)
    if x.dtype == np.float32:
        return x.view(np.uint32)
    return x


def compare(gpu: dict, ora: dict, range_max: float) -> dict:
    """Exact per-attribute comparison over the attributes both hold; ora must carry
    "flags" (oracle.simulate(..., flags=True)).  Returns mismatch counts, flag counts, the
    mismatching rays outside the flagged set and calc_diffs per attribute."""
    flags = ora["flags"]
    n = len(flags)
    bad = np.zeros(n, bool)
    out = {"rays": n, "mismatch": {}, "diff": {}}
    for k in ("hits", "ranges", "points", "normals", "face_ids", "geom_ids", "object_ids"):
        if k not in gpu or k not in ora:
            continue
        a, b = np.asarray(gpu[k]), np.asarray(ora[k])
        ne = _bits(a) != _bits(b)
        if ne.ndim > 1:
            # NaN payloads may differ: compare NaN patterns, then bits of the finite values
            na, nb = np.isnan(a), np.isnan(b)
            ne = (na != nb) | (~na & ~nb & ne)
            ne = ne.any(axis=1)
        out["mismatch"][k] = int(ne.sum())
        bad |= ne
        if k in ("hits", "ranges", "points", "normals"):
            out["diff"][k] = calc_diffs(k, a, b, range_max)
    out["flagged"] = {"any": int((flags != 0).sum()), "edge": int((flags & FLAG_EDGE != 0).sum()),
                      "near_miss": int((flags & FLAG_NEAR_MISS != 0).sum()),
                      "tie": int((flags & FLAG_TIE != 0).sum()), "range_limit": int((flags & FLAG_RANGE != 0).sum())}
    out["mismatching_rays"] = int(bad.sum())
    out["mismatching_unflagged"] = int((bad & (flags == 0)).sum())
    return out


def merge(acc: dict | None, r: dict) -> dict:
    """Sum of two compare() results (max/avg differences combined by weight)."""
    if acc is None:
        return {k: (dict(v) if isinstance(v, dict) else v) for k, v in r.items()} | {
            "diff": {k: dict(v) for k, v in r["diff"].items()}}
    keys = list(dict.fromkeys([*acc["mismatch"], *r["mismatch"]]))
    out = {"rays": acc["rays"] + r["rays"], "mismatching_rays": acc["mismatching_rays"] + r["mismatching_rays"],
           "mismatching_unflagged": acc["mismatching_unflagged"] + r["mismatching_unflagged"],
           "mismatch": {k: acc["mismatch"].get(k, 0) + r["mismatch"].get(k, 0) for k in keys},
           "flagged": {k: acc["flagged"].get(k, 0) + r["flagged"].get(k, 0)
                       for k in dict.fromkeys([*acc["flagged"], *r["flagged"]])},
           "diff": {k: dict(v) for k, v in acc["diff"].items()}}
    for k, d in r["diff"].items():
        a = acc["diff"].get(k)
        if a is None:
            out["diff"][k] = dict(d)
            continue
        nb = a["both_hit"] + d["both_hit"]
        out["diff"][k] = {"both_hit": nb, "only_a": a["only_a"] + d["only_a"], "only_b": a["only_b"] + d["only_b"],
                          "neither": a["neither"] + d["neither"],
                          "both_hit_high_diff": a["both_hit_high_diff"] + d["both_hit_high_diff"],
                          "max_diff": max(a["max_diff"], d["max_diff"]),
                          "avg_diff": (a["avg_diff"] * a["both_hit"] + d["avg_diff"] * d["both_hit"]) / nb if nb else 0.0}
    return out
