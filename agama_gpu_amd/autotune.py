"""Per-lmax launch tuning for the batched Multipole kernel.

For a potential class -- lmax, active harmonics, and whether its table fits
LDS -- the tuner times the batched kernel on that potential alone for every
feasible (path, particles per work item) pair:

  path  AGX_BATCH_LANE    per-lane coefficient reads from L2
        AGX_BATCH_LDS     whole table staged in LDS (only if it fits)
  rows per item  2048, 4096, 8192, 16384

and keeps the fastest.  Results are cached in memory and, when
``AGAMA_GPU_TUNE_CACHE`` names a file, in that JSON file, keyed by the GPU
architecture and the shape.  Without tuning, ``default_choice`` gives the
rule the measurements on gfx950 support: LDS staging whenever the table fits,
8192 rows per item.
"""
from __future__ import annotations

import json
import os
import threading
from typing import Dict, Tuple

import numpy as np

from . import _lib

ITEM_ROWS = (2048, 4096, 8192, 16384)
_cache: Dict[str, Tuple[int, int]] = {}
_lock = threading.Lock()


def shape_key(pot) -> Tuple[int, int, int]:
    """Tuning class: lmax, active harmonics and whether the table fits LDS
    (potentials of one class share a choice)."""
    return int(pot.lmax), int(pot.n_lm), int(lds_fits(pot))


def _arch(device) -> str:
    import torch
    try:
        return torch.cuda.get_device_properties(device).gcnArchName.split(":")[0]
    except Exception:   # noqa: BLE001 -- arch name is only a cache key
        return "gpu"


def lds_fits(pot) -> bool:
    return pot.n_intervals * pot.n_lm * 6 <= int(_lib.load().agx_batch_lds_table_doubles())


def default_choice(pot) -> Tuple[int, int]:
    if lds_fits(pot):
        return _lib.AGX_BATCH_LDS, 8192
    return _lib.AGX_BATCH_LANE, 4096


def candidates(pot):
    out = []
    for rows in ITEM_ROWS:
        out.append((_lib.AGX_BATCH_LANE, rows))
        if lds_fits(pot):
            out.append((_lib.AGX_BATCH_LDS, rows))
    return out


def _cache_file():
    return os.environ.get("AGAMA_GPU_TUNE_CACHE")


def _load_file():
    f = _cache_file()
    if not f or not os.path.exists(f):
        return {}
    try:
        with open(f) as fh:
            return {k: tuple(v) for k, v in json.load(fh).items()}
    except (OSError, ValueError):
        return {}


def _save_file():
    f = _cache_file()
    if not f:
        return
    tmp = f + ".tmp"
    with open(tmp, "w") as fh:
        json.dump({k: list(v) for k, v in _cache.items()}, fh, indent=1, sort_keys=True)
    os.replace(tmp, f)


_unit_pts = {}


def _points_for(pot, n: int):
    """Device positions: Hernquist radii (scale 1, seed 0) rescaled to the
    potential's grid (scale = 0.2 x geometric mean of the grid ends), clipped
    to half / twice its ends so both extrapolations are exercised."""
    import torch
    dk = (str(pot.device), n)
    if dk not in _unit_pts:
        rng = np.random.default_rng(0)
        su = np.sqrt(rng.random(n))
        r = su / (1.0 - su)
        ct = rng.uniform(-1.0, 1.0, n)
        ph = rng.uniform(0.0, 2.0 * np.pi, n)
        st = np.sqrt(1.0 - ct * ct)
        u = np.stack([st * np.cos(ph), st * np.sin(ph), ct], axis=1)
        _unit_pts.clear()
        _unit_pts[dk] = (torch.as_tensor(u, device=pot.device), torch.as_tensor(r, device=pot.device))
    u, r = _unit_pts[dk]
    d = pot._data
    r_lo = float(np.exp(d["logr_min"]))
    r_hi = float(np.exp(d["logr_min"] + d["n_intervals"] * d["dlogr"]))
    rr = torch.clamp(r * (0.2 * np.sqrt(r_lo * r_hi)), r_lo * 0.5, r_hi * 2.0)
    return (u * rr[:, None]).contiguous()


def tune(pot, n: int = 1 << 23, reps: int = 3, order: int = 1) -> Tuple[int, int]:
    """Measure every candidate on ``pot`` and return the fastest (path, rows).
    n is large enough (1024+ items) to fill the chip, so the comparison is of
    per-particle throughput, not of the latency of a few items."""
    import torch
    from .batch import MultipoleEnsembleGPU
    key = f"{_arch(pot.device)}:{'pfh'[order]}:" + ":".join(map(str, shape_key(pot)))
    with _lock:
        if not _cache:
            _cache.update(_load_file())
        if key in _cache:
            return _cache[key]
    x = _points_for(pot, n)
    phi = torch.empty(n, dtype=torch.float64, device=pot.device)
    force = torch.empty((n, 3), dtype=torch.float64, device=pot.device)
    deriv = torch.empty((n, 6), dtype=torch.float64, device=pot.device) if order == 2 else None
    best, best_ms = None, float("inf")
    for cand in candidates(pot):
        ens = MultipoleEnsembleGPU([pot], [n], choices=[cand])
        ens.eval_into(x, phi=phi, force=force, deriv=deriv)
        torch.cuda.synchronize(pot.device)
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record()
        for _ in range(reps):
            ens.eval_into(x, phi=phi, force=force, deriv=deriv)
        b.record()
        torch.cuda.synchronize(pot.device)
        ms = a.elapsed_time(b) / reps
        if ms < best_ms:
            best, best_ms = cand, ms
    with _lock:
        _cache[key] = best
        _save_file()
    return best


def tune_ensemble(pots, order: int = 1):
    """(path, rows) per potential, tuning each distinct shape once."""
    seen = {}
    out = []
    for p in pots:
        k = shape_key(p)
        if k not in seen:
            seen[k] = tune(p, order=order)
        out.append(seen[k])
    return out
