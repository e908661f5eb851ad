"""Agama expansion-coefficient text format: parser and writer.

The reference reads coefficient files through an un-vendored helper
(``nbody_streams.agama_helper.read_coefs``; call sites in the reference at
gpu_potential.py:683-686, 2219-2223, 2246-2251).  This module re-implements
the format from the files themselves (tests/600.dark.none_8.coef_mul_DR,
tests/100.LMC.none_8.coef_mult, tests/600.bar.none_8.coef_cylsp_DR):

Multipole::

    [Potential]
    type=Multipole
    gridSizeR=<nR>
    lmax=<L>
    symmetry=<...>
    Coefficients
    #Phi
    #radius  l=0,m=0  l=1,m=-1 ...          (one column per (l,m), l-major, m=-l..l)
    r_0      phi_00   ...                     (nR rows)

    #dPhi/dr
    #radius  l=0,m=0 ...
    r_0      dphi_00  ...

CylSpline::

    [Potential]
    type=CylSpline
    gridSizeR=<nR>
    gridSizez=<nz>
    mmax=<M>
    Coefficients
    #Phi
    <m>  #m
    #R(row)\\z(col)  z_0 z_1 ...
    R_0  phi(R_0,z_0) ...                     (nR rows per m block)
"""
from __future__ import annotations

import io
import re
from dataclasses import dataclass, field
from pathlib import Path
from typing import Dict, List, Optional, Tuple, Union

import numpy as np

__all__ = [
    "MultipoleCoefs", "CylSplineCoefs", "read_coefs", "write_coefs",
    "parse_header",
]

_LM_RE = re.compile(r"l\s*=\s*(-?\d+)\s*,\s*m\s*=\s*(-?\d+)")


@dataclass
class MultipoleCoefs:
    """Raw Multipole table: Phi_lm(r_i) and dPhi_lm/dr(r_i).

    ``phi`` and ``dphi_dr`` have shape (nR, n_lm); column j holds harmonic
    ``lm_labels[j] = (l, m)`` (m<0 are sin modes, m>=0 cos modes).
    """
    R_grid: np.ndarray
    phi: np.ndarray
    dphi_dr: Optional[np.ndarray]
    lm_labels: List[Tuple[int, int]]
    lmax: int
    symmetry: str = "None"
    params: Dict[str, object] = field(default_factory=dict)

    def zeroed(self, keep) -> "MultipoleCoefs":
        """Copy with every harmonic not accepted by ``keep(l, m)`` set to 0."""
        phi = self.phi.copy()
        dphi = None if self.dphi_dr is None else self.dphi_dr.copy()
        for j, (l, m) in enumerate(self.lm_labels):
            if not keep(l, m):
                phi[:, j] = 0.0
                if dphi is not None:
                    dphi[:, j] = 0.0
        return MultipoleCoefs(self.R_grid.copy(), phi, dphi, list(self.lm_labels),
                              self.lmax, self.symmetry, dict(self.params))


@dataclass
class CylSplineCoefs:
    """Raw CylSpline table: Phi_m(R_i, z_j) per azimuthal harmonic m."""
    R_grid: np.ndarray
    z_grid: np.ndarray
    phi: Dict[int, np.ndarray]          # m -> (nR, nz)
    mmax: int
    symmetry: str = "None"
    params: Dict[str, object] = field(default_factory=dict)


def parse_header(lines: List[str]) -> Tuple[Dict[str, str], int]:
    """Parse ``key=value`` lines up to the ``Coefficients`` marker.

    Returns the parameter dict (keys as written) and the index of the first
    line after ``Coefficients`` (``len(lines)`` when absent).
    """
    params: Dict[str, str] = {}
    for i, raw in enumerate(lines):
        s = raw.strip()
        if not s or s.startswith("[") or s.startswith(";"):
            continue
        if s.lower() == "coefficients":
            return params, i + 1
        if s.startswith("#"):
            continue
        if "=" in s:
            k, _, v = s.partition("=")
            params[k.strip()] = v.split("#")[0].strip()
    return params, len(lines)


def _get(params: Dict[str, str], key: str, default=None):
    for k, v in params.items():
        if k.lower() == key.lower():
            return v
    return default


def _numeric_rows(lines: List[str], start: int) -> Tuple[List[List[float]], int]:
    """Read consecutive numeric rows starting at ``start``; stop at a blank or
    '#' line.  Returns the rows and the index of the stopping line."""
    rows: List[List[float]] = []
    i = start
    while i < len(lines):
        s = lines[i].strip()
        if not s or s.startswith("#"):
            break
        toks = s.split()
        try:
            rows.append([float(t) for t in toks])
        except ValueError:
            break
        i += 1
    return rows, i


def _parse_multipole(lines: List[str], params: Dict[str, str], body: int) -> MultipoleCoefs:
    blocks: Dict[str, np.ndarray] = {}
    labels: Optional[List[Tuple[int, int]]] = None
    i = body
    current = None
    while i < len(lines):
        s = lines[i].strip()
        if not s:
            i += 1
            continue
        if s.startswith("#"):
            tag = s[1:].strip()
            low = tag.lower()
            if low.startswith("radius"):
                lab = [(int(a), int(b)) for a, b in _LM_RE.findall(tag)]
                if labels is None:
                    labels = lab
                elif lab and lab != labels:
                    raise ValueError("inconsistent (l,m) column labels between blocks")
                rows, i = _numeric_rows(lines, i + 1)
                if current is None:
                    raise ValueError("numeric block without a '#Phi' or '#dPhi/dr' tag")
                blocks[current] = np.asarray(rows, dtype=np.float64)
                current = None
                continue
            if low == "phi":
                current = "phi"
            elif low.replace(" ", "") in ("dphi/dr", "dphidr"):
                current = "dphi"
            i += 1
            continue
        i += 1
    if "phi" not in blocks:
        raise ValueError("Multipole coefficients: missing '#Phi' block")
    phi_blk = blocks["phi"]
    R = phi_blk[:, 0].copy()
    phi = phi_blk[:, 1:].copy()
    if labels is None or len(labels) != phi.shape[1]:
        lmax_h = int(_get(params, "lmax", -1))
        labels = [(l, m) for l in range(lmax_h + 1) for m in range(-l, l + 1)]
        if len(labels) != phi.shape[1]:
            raise ValueError("cannot infer (l,m) labels for %d columns" % phi.shape[1])
    dphi = None
    if "dphi" in blocks:
        d = blocks["dphi"]
        if d.shape != phi_blk.shape or not np.allclose(d[:, 0], R, rtol=1e-12, atol=0):
            raise ValueError("'#dPhi/dr' block does not match the '#Phi' grid")
        dphi = d[:, 1:].copy()
    n_expect = _get(params, "gridSizeR")
    if n_expect is not None and int(float(n_expect)) != len(R):
        raise ValueError(f"gridSizeR={n_expect} but {len(R)} rows found")
    lmax = max(l for l, _ in labels)
    return MultipoleCoefs(R, phi, dphi, labels, lmax,
                          str(_get(params, "symmetry", "None")), dict(params))


def _parse_cylspline(lines: List[str], params: Dict[str, str], body: int) -> CylSplineCoefs:
    phi: Dict[int, np.ndarray] = {}
    R = None
    z = None
    i = body
    m_cur = None
    while i < len(lines):
        s = lines[i].strip()
        if not s:
            i += 1
            continue
        toks = s.split()
        if len(toks) >= 2 and toks[1] == "#m":
            m_cur = int(float(toks[0]))
            i += 1
            continue
        if s.startswith("#R") or s.startswith("#r"):
            zs = np.asarray([float(t) for t in toks[1:]], dtype=np.float64)
            rows, i = _numeric_rows(lines, i + 1)
            arr = np.asarray(rows, dtype=np.float64)
            if m_cur is None:
                raise ValueError("CylSpline block without an '<m> #m' line")
            if z is None:
                z = zs
                R = arr[:, 0].copy()
            elif zs.shape != z.shape or not np.allclose(zs, z) or not np.allclose(arr[:, 0], R):
                raise ValueError("CylSpline blocks use different grids")
            phi[m_cur] = arr[:, 1:].copy()
            m_cur = None
            continue
        i += 1
    if not phi:
        raise ValueError("CylSpline coefficients: no '#m' blocks found")
    if 0 not in phi:
        raise ValueError("CylSpline coefficients: the m=0 harmonic is required")
    mmax = max(abs(m) for m in phi)
    return CylSplineCoefs(R, z, phi, mmax, str(_get(params, "symmetry", "None")), dict(params))


def read_coefs(source: Union[str, Path]) -> Union[MultipoleCoefs, CylSplineCoefs]:
    """Parse an Agama coefficient section from a path or from section text."""
    text = None
    if isinstance(source, Path):
        text = source.read_text()
    elif isinstance(source, str) and ("\n" not in source) and Path(source).exists():
        text = Path(source).read_text()
    elif isinstance(source, str) and "\n" in source:
        text = source
    else:
        raise FileNotFoundError(f"coefficient file not found: {source!r}")
    lines = text.splitlines()
    params, body = parse_header(lines)
    typ = str(_get(params, "type", "")).lower().replace("_", "").replace(" ", "")
    if typ == "multipole":
        return _parse_multipole(lines, params, body)
    if typ == "cylspline":
        return _parse_cylspline(lines, params, body)
    raise ValueError(f"unsupported coefficient type {typ!r}")


def _fmt(v: float) -> str:
    return repr(float(v))


def write_coefs(coefs: Union[MultipoleCoefs, CylSplineCoefs], path=None) -> str:
    """Serialise to the Agama text format (full double precision).  Returns
    the text; also writes it to ``path`` when given."""
    out = io.StringIO()
    if isinstance(coefs, MultipoleCoefs):
        out.write("[Potential]\ntype=Multipole\n")
        out.write(f"gridSizeR={len(coefs.R_grid)}\nlmax={coefs.lmax}\n")
        out.write(f"symmetry={coefs.symmetry}\nCoefficients\n")
        head = "#radius\t" + "\t".join(f"l={l},m={m}" for l, m in coefs.lm_labels)
        blocks = [("#Phi", coefs.phi)]
        if coefs.dphi_dr is not None:
            blocks.append(("#dPhi/dr", coefs.dphi_dr))
        for bi, (tag, arr) in enumerate(blocks):
            if bi:
                out.write("\n")
            out.write(tag + "\n" + head + "\n")
            for i, r in enumerate(coefs.R_grid):
                out.write(_fmt(r) + "\t" + "\t".join(_fmt(v) for v in arr[i]) + "\n")
    elif isinstance(coefs, CylSplineCoefs):
        out.write("[Potential]\ntype=CylSpline\n")
        out.write(f"gridSizeR={len(coefs.R_grid)}\ngridSizez={len(coefs.z_grid)}\n")
        out.write(f"mmax={coefs.mmax}\nsymmetry={coefs.symmetry}\nCoefficients\n#Phi\n")
        for m in sorted(coefs.phi):
            out.write(f"{m}\t#m\n#R(row)\\z(col)\t" + "\t".join(_fmt(z) for z in coefs.z_grid) + "\n")
            tab = coefs.phi[m]
            for i, R in enumerate(coefs.R_grid):
                out.write(_fmt(R) + "\t" + "\t".join(_fmt(v) for v in tab[i]) + "\n")
    else:
        raise TypeError(f"cannot write {type(coefs).__name__}")
    text = out.getvalue()
    if path is not None:
        Path(path).write_text(text)
    return text
