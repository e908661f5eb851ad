"""``PotentialGPU`` factory and Agama INI front end.

Same entry points and semantics as the reference (gpu_potential.py:
``PotentialGPU`` 2377-2502, ``_build_single`` 2307-2374, ``_load_potential_ini``
2154-2304, ``_is_potential_ini`` 2140-2151, ``_coerce`` 2127-2137,
``_apply_modifiers`` 2097-2124, type map 2060-2094).

Types the reference builds by calling Agama on the CPU (Spheroid, King, Disk,
triaxial Dehnen; gpu_potential.py:1935-2052) are built here by this package's
own code instead (expansion.py), since Agama is not available: Spheroid and
triaxial Dehnen as Multipole tables from their densities by the Poisson
solver, King as a monopole table from its Poisson ODE, Disk as DiskAnsatz plus
an axisymmetric lmax-32 Multipole of the residual density (Agama's own
construction, reference gpu_potential.py:1984-2052).
"""
from __future__ import annotations

import re
from pathlib import Path
from typing import Union

from .analytic import ANALYTIC_TYPES, DehnenSphericalPotentialGPU, AnalyticPotentialGPU
from .coefs import CylSplineCoefs, MultipoleCoefs, read_coefs
from .composite import (CompositePotentialGPU, EvolvingPotentialGPU, ScaledPotentialGPU,
                        ShiftedPotentialGPU)
from .multipole import MultipolePotentialGPU


def _key(s) -> str:
    return str(s).lower().replace(" ", "").replace("_", "")


def _kl(kw: dict) -> dict:
    return {_key(k): v for k, v in kw.items()}


def _coerce(v: str):
    for f in (int, float):
        try:
            return f(v)
        except (ValueError, TypeError):
            pass
    return v


def _apply_modifiers(pot, center, scale, ampl):
    """Wrap in Scaled (inner) and Shifted (outer), as gpu_potential.py:2097-2124."""
    if scale is not None or ampl != 1.0:
        if isinstance(scale, str):
            parts = scale.strip().split()
            if len(parts) == 1:
                scale = float(parts[0])
            elif len(parts) == 2:
                ampl, scale = float(parts[0]), float(parts[1])
            else:
                raise ValueError(f"Cannot parse scale= value {scale!r}. Expected a float, "
                                 "'ampl scale' pair, or (T,2)/(T,3) array.")
        pot = ScaledPotentialGPU(pot, scale=scale if scale is not None else 1.0, ampl=ampl)
    if center is not None:
        if isinstance(center, str):
            center = [float(v) for v in center.replace(",", " ").split()]
        pot = ShiftedPotentialGPU(pot, center=center)
    return pot


def _expansion_from_file(path, **kw):
    c = read_coefs(Path(path))
    if isinstance(c, MultipoleCoefs):
        return MultipolePotentialGPU(c, **kw)
    if isinstance(c, CylSplineCoefs):
        from .cylspline import CylSplinePotentialGPU
        return CylSplinePotentialGPU(c, **kw)
    raise ValueError(f"{path}: unsupported coefficient type")


def _dehnen(**kw):
    kl = _kl(kw)
    gamma = float(kl.get("gamma", 1.0))
    yr, zr = float(kl.get("axisratioy", 1.0)), float(kl.get("axisratioz", 1.0))
    if not 0.0 <= gamma <= 2.0:
        raise ValueError(f"Dehnen gamma must be in [0, 2], got {gamma}")
    mass, a = float(kl.get("mass", 1.0)), float(kl.get("scaleradius", 1.0))
    if abs(yr - 1.0) < 1e-10 and abs(zr - 1.0) < 1e-10:
        return DehnenSphericalPotentialGPU(mass=mass, scaleRadius=a, gamma=gamma)
    # triaxial: Multipole expansion from this package's Poisson solver
    from .expansion import dehnen_density, multipole_from_density
    lmax = int(kl.get("lmax", 8))
    nr = int(kl.get("gridsizer", 50))
    rmin = float(kl.get("rmin", 1e-3 * a))
    rmax = float(kl.get("rmax", 1e3 * a))
    coefs = multipole_from_density(dehnen_density(mass, a, gamma, yr, zr), lmax, rmin, rmax, nr,
                                   symmetry="triaxial")
    return MultipolePotentialGPU(coefs)


def _grid_kw(kl: dict, lmax: int, nr: int) -> dict:
    out = dict(lmax=int(kl.get("lmax", lmax)), gridSizeR=int(kl.get("gridsizer", nr)))
    if "rmin" in kl:
        out["rmin"] = float(kl["rmin"])
    if "rmax" in kl:
        out["rmax"] = float(kl["rmax"])
    return out


def _spheroid(**kw):
    """Agama's Spheroid (reference _build_spheroid_gpu, gpu_potential.py:1935-1939):
    densityNorm or mass, scaleRadius, alpha, beta, gamma, outerCutoffRadius,
    cutoffStrength, axisRatioY, axisRatioZ; lmax (default 8), gridSizeR (50),
    rMin / rMax of the table."""
    from .expansion import spheroid_coefs
    kl = _kl(kw)
    names = {"densitynorm": "densityNorm", "scaleradius": "scaleRadius", "alpha": "alpha",
             "beta": "beta", "gamma": "gamma", "outercutoffradius": "outerCutoffRadius",
             "cutoffstrength": "cutoffStrength", "axisratioy": "axisRatioY",
             "axisratioz": "axisRatioZ"}
    args = {v: float(kl[k]) for k, v in names.items() if k in kl}
    if "mass" in kl:
        args["mass"] = float(kl["mass"])
    return MultipolePotentialGPU(spheroid_coefs(**_grid_kw(kl, 8, 50), **args))


def _king(**kw):
    """King (lowered isothermal) sphere (reference _build_king_gpu,
    gpu_potential.py:1942-1945): mass, scaleRadius (the King radius), W0,
    trunc (1: King 1966); gridSizeR of the monopole table (default 60)."""
    from .expansion import king_coefs
    kl = _kl(kw)
    if "w0" not in kl:
        raise ValueError("King potential needs W0 (the dimensionless central potential)")
    return MultipolePotentialGPU(king_coefs(mass=float(kl.get("mass", 1.0)),
                                            scaleRadius=float(kl.get("scaleradius", 1.0)),
                                            W0=float(kl["w0"]), trunc=float(kl.get("trunc", 1.0)),
                                            gridSizeR=int(kl.get("gridsizer", 60))))


def _disk(**kw):
    """Exponential disk as DiskAnsatz + axisymmetric Multipole (lmax 32) of the
    residual density (reference _build_disk_gpu, gpu_potential.py:1984-2052):
    surfaceDensity (or densityNorm, or mass), scaleRadius, scaleHeight
    (> 0 exponential, < 0 isothermal, 0 thin), innerCutoffRadius; lmax,
    gridSizeR, rMin, rMax of the Multipole part."""
    from .analytic import DiskAnsatzPotentialGPU
    from .expansion import disk_potential_parts
    kl = _kl(kw)
    sig = kl.get("surfacedensity", kl.get("densitynorm"))
    ansatz, coefs = disk_potential_parts(
        surfaceDensity=None if sig is None else float(sig),
        mass=float(kl["mass"]) if "mass" in kl else None,
        scaleRadius=float(kl.get("scaleradius", 1.0)), scaleHeight=float(kl.get("scaleheight", 0.1)),
        innerCutoffRadius=float(kl.get("innercutoffradius", 0.0)),
        sersicIndex=float(kl.get("sersicindex", 1.0)), **_grid_kw(kl, 32, 50))
    return CompositePotentialGPU([DiskAnsatzPotentialGPU(**ansatz), MultipolePotentialGPU(coefs)])


def type_map() -> dict:
    m = dict(ANALYTIC_TYPES)
    m.update({"dehnen": _dehnen, "disk": _disk, "king": _king, "spheroid": _spheroid})
    return m


def _is_potential_ini(p: Path) -> bool:
    if not p.exists():
        return False
    if p.suffix.lower() in (".ini", ".pot"):
        return True
    try:
        head = p.read_text(encoding="utf-8", errors="ignore")[:512]
    except OSError:
        return False
    return bool(re.search(r"^\[Potential", head, re.IGNORECASE | re.MULTILINE))


def _load_potential_ini(p: Path):
    """One or more [Potential ...] sections -> a potential or a Composite."""
    base = p.parent
    lines = p.read_text(encoding="utf-8").splitlines()
    starts = [i for i, ln in enumerate(lines) if re.match(r"^\s*\[Potential", ln, re.IGNORECASE)]
    if not starts:
        raise ValueError(f"No [Potential] sections found in {p}.")
    built = []
    for idx, start in enumerate(starts):
        end = starts[idx + 1] if idx + 1 < len(starts) else len(lines)
        sec = lines[start:end]
        params, data_start, data_kind = {}, None, None
        for j, ln in enumerate(sec[1:], start=1):
            s = ln.strip()
            if not s or s.startswith("#") or s.startswith(";"):
                continue
            if s.lower() == "coefficients":
                data_start, data_kind = j, "coef"
                break
            if s.lower() == "timestamps":
                data_start, data_kind = j, "ts"
                break
            if "=" in s:
                k, _, v = s.partition("=")
                params[k.strip()] = _coerce(v.split("#")[0].strip())
        type_ = _key(params.get("type") or params.get("Type", ""))
        if type_ == "diskansatz" and not any(_key(k) in ("surfacedensity", "scaleradius")
                                             for k in params):
            continue     # Agama exports DiskAnsatz without its parameters
        if type_ in ("multipole", "cylspline"):
            if data_kind == "coef":
                c = read_coefs("\n".join(sec))
                if isinstance(c, MultipoleCoefs):
                    built.append(MultipolePotentialGPU(c))
                else:
                    from .cylspline import CylSplinePotentialGPU
                    built.append(CylSplinePotentialGPU(c))
            else:
                f = str(params.get("file") or params.get("File", ""))
                if not f:
                    raise ValueError(f"{type_} section in {p} has neither inline Coefficients "
                                     "nor a 'file' key.")
                fp = Path(f)
                built.append(_expansion_from_file(fp if fp.is_absolute() else base / fp))
            continue
        if type_ == "evolving":
            if data_kind != "ts":
                raise ValueError(f"Evolving section in {p} is missing a 'Timestamps' block.")
            times, pots = [], []
            for ln in sec[data_start + 1:]:
                s = ln.strip()
                if not s or s.startswith("#") or s.startswith(";"):
                    continue
                parts = s.split()
                if len(parts) < 2:
                    continue
                try:
                    tv = float(parts[0])
                except ValueError:
                    continue
                sp = Path(parts[1])
                times.append(tv)
                pots.append(_build_single(sp if sp.is_absolute() else base / sp, {}))
            if not pots:
                raise ValueError(f"Evolving section in {p} has no valid Timestamps entries.")
            lin = bool(params.get("linearInterp", params.get("interpLinear", False)))
            built.append(EvolvingPotentialGPU(pots, times, interpolate=lin))
            continue
        built.append(_build_single(params, {}))
    if not built:
        raise ValueError(f"No buildable components found in {p}.")
    return built[0] if len(built) == 1 else CompositePotentialGPU(built)


def _build_single(source, pot_kw: dict):
    if isinstance(source, MultipoleCoefs):
        return MultipolePotentialGPU(source, **pot_kw)
    if isinstance(source, CylSplineCoefs):
        from .cylspline import CylSplinePotentialGPU
        return CylSplinePotentialGPU(source, **pot_kw)
    if isinstance(source, dict):
        d = dict(source)
        type_s = d.pop("type", None) or d.pop("Type", None)
        if type_s is None:
            raise ValueError("Component dict must include a 'type' key.")
        cen, sc, am = d.pop("center", None), d.pop("scale", None), float(d.pop("ampl", 1.0))
        tk = _key(type_s)
        if tk in ("cylspline", "multipole"):
            ref = str(d.pop("file", None) or d.pop("File", None) or "")
            if not ref:
                raise ValueError(f"{type_s} component dict requires a 'file' key.")
            pot = _expansion_from_file(ref)
        else:
            pot = PotentialGPU(type=type_s, **d)
        return _apply_modifiers(pot, cen, sc, am)
    if isinstance(source, (str, Path)):
        p = Path(source)
        if _is_potential_ini(p):
            return _load_potential_ini(p)
        return _expansion_from_file(p, **pot_kw)
    if callable(getattr(source, "potential", None)) and callable(getattr(source, "force", None)):
        return source
    raise TypeError(
        f"Cannot build a GPU potential from {type(source).__name__!r}. Pass a file path, "
        "coefficient object, GPU potential object, or use type= for analytic potentials.")


def PotentialGPU(*args, type: str = None, file: Union[str, Path] = None, center=None,
                 scale=None, ampl: float = 1.0, **kw):
    """GPU potential factory mirroring ``agama.Potential``:

    * ``PotentialGPU(type='NFW', mass=1e12, scaleRadius=20)`` -- analytic;
    * ``PotentialGPU(file='halo.coef_mul_DR')`` or ``PotentialGPU('halo.ini')``
      -- coefficient file or multi-section INI (Multipole, CylSpline,
      Evolving with Timestamps, analytic sections);
    * ``PotentialGPU(pot_a, coefs_b, ...)`` -- a Composite;
    * ``center=``, ``scale=``, ``ampl=`` -- Shifted / Scaled modifiers.
    """
    pot_kw = kw
    if len(args) > 1:
        pot = CompositePotentialGPU([_build_single(a, pot_kw) for a in args])
        return _apply_modifiers(pot, center, scale, ampl)
    if type is not None:
        key = _key(type)
        if key in ("cylspline", "multipole"):
            if file is None:
                raise ValueError(f"type={type!r} requires file= to specify the coefficient file.")
            return _apply_modifiers(_expansion_from_file(file, **pot_kw), center, scale, ampl)
        m = type_map()
        cls = m.get(key)
        if cls is None:
            raise ValueError(f"Unknown GPU potential type {type!r}. Available: {sorted(m)}")
        pot = AnalyticPotentialGPU(type, **kw) if key in ANALYTIC_TYPES and key != "dehnen" \
            else cls(**kw)
        return _apply_modifiers(pot, center, scale, ampl)
    if file is not None:
        p = Path(file)
        pot = _load_potential_ini(p) if _is_potential_ini(p) else _expansion_from_file(p, **pot_kw)
        return _apply_modifiers(pot, center, scale, ampl)
    if len(args) == 1:
        return _apply_modifiers(_build_single(args[0], pot_kw), center, scale, ampl)
    raise TypeError("PotentialGPU() requires at least one argument, type=, or file=.")
