"""Differentiable Multipole (+ time-dependent shift) for fitting potentials.

``FittableMultipole`` holds the raw Agama table -- Phi_lm(r_i) and
dPhi_lm/dr(r_i), the ``2 x n_lm x nR`` numbers of a coefficient file -- as
torch parameters.  Its forward pass rebuilds the kernel tables from them with
torch operations (the same steps as ``multipole_prep.build_multipole_data``:
column order and pruning, inner/outer l=0 laws, log-scaling, the C2 quintic
solve, Horner coefficients; the discrete choices -- pruning, law modes --
frozen at construction) and evaluates potential and force with the gfx950
kernel.  The backward pass is one launch of the adjoint kernel
(``agx_multipole_adjoint``): a Hessian-vector product for dL/dxyz and a
wavefront-reduced scatter of dL/d(table) into the spline intervals; torch
carries the table gradient back through the rebuild to the raw entries.

``FittableCylSpline`` does the same for a CylSpline table Phi_m(R_i, z_j):
the rebuild (z-mirroring, log-scaling, the 2-D spline node derivatives, the
outer multipole's boundary least-squares fit) is linear algebra with
matrices fixed at construction, and the backward pass is one launch of
``agx_cylspline_adjoint``.

``SplineCenter`` is the differentiable counterpart of ShiftedPotentialGPU's
trajectory: the centre at time t is linear in the knot positions, with
weights from the same cubic spline (not-a-knot, linear extrapolation), so
gradients reach the knots through x - c(t).
"""
from __future__ import annotations

import ctypes as C
from typing import Optional

import numpy as np
import torch

from . import _lib
from .base import norm_device
from .coefs import CylSplineCoefs, MultipoleCoefs
from .composite import _linear_extrapolating_spline
from .cylspline_prep import MUL0, MUL1, build_cylspline_data, ptilde
from .multipole import eval_table
from .multipole_prep import build_multipole_data, pack_program, program_flags

F64 = torch.float64


def _quintic_operator(logr: np.ndarray):
    """A (n x n) and the linear maps B_f, B_d (n x n) with d2 = A^-1 (B_f f + B_d df):
    the tridiagonal C2-quintic system of multipole_prep.quintic_second_derivs."""
    n = len(logr)
    A = np.zeros((n, n))
    Bf = np.zeros((n, n))
    Bd = np.zeros((n, n))
    for i in range(n):
        if i > 0:
            h = 1.0 / (logr[i] - logr[i - 1])
            A[i, i] += 3.0 * h
            A[i, i - 1] = -h
            Bf[i, i] -= 20.0 * h ** 3
            Bf[i, i - 1] += 20.0 * h ** 3
            Bd[i, i] += 12.0 * h * h
            Bd[i, i - 1] += 8.0 * h * h
        if i < n - 1:
            h = 1.0 / (logr[i + 1] - logr[i])
            A[i, i] += 3.0 * h
            A[i, i + 1] = -h
            Bf[i, i + 1] += 20.0 * h ** 3
            Bf[i, i] -= 20.0 * h ** 3
            Bd[i, i] -= 12.0 * h * h
            Bd[i, i + 1] -= 8.0 * h * h
    h0 = 1.0 / (logr[1] - logr[0])
    A[0, :] = 0.0
    A[0, 0] = 3.0 * h0
    A[0, 1] = -2.0 * h0
    Bf[0, :] = 0.0
    Bd[0, :] = 0.0
    Bf[0, 1], Bf[0, 0] = 30.0 * h0 ** 3, -30.0 * h0 ** 3
    Bd[0, 1], Bd[0, 0] = -14.0 * h0 * h0, -16.0 * h0 * h0
    hn = 1.0 / (logr[n - 1] - logr[n - 2])
    A[n - 1, :] = 0.0
    A[n - 1, n - 1] = 3.0 * hn
    A[n - 1, n - 2] = -2.0 * hn
    Bf[n - 1, :] = 0.0
    Bd[n - 1, :] = 0.0
    Bf[n - 1, n - 1], Bf[n - 1, n - 2] = -30.0 * hn ** 3, 30.0 * hn ** 3
    Bd[n - 1, n - 2], Bd[n - 1, n - 1] = 14.0 * hn * hn, 16.0 * hn * hn
    Ainv = np.linalg.inv(A)
    return Ainv @ Bf, Ainv @ Bd


class _MultipoleEval(torch.autograd.Function):
    """(xyz, center, poly, scal) -> (phi, force) at xyz - center (center (3,)
    or None) with the adjoint kernel as backward.  The kernels subtract the
    centre themselves (agx_multipole.origin), and the centre's gradient,
    minus the sum over particles of dL/dx, comes from the adjoint kernel too
    (AGX_ADJ_CENTER_GRAD) instead of a separate column reduction."""

    @staticmethod
    def forward(ctx, xyz, center, poly, scal, owner):
        if center is not None:
            center = center.contiguous()
        desc = owner._desc_for(poly, scal, center)
        n = xyz.shape[0]
        phi = torch.empty(n, dtype=F64, device=xyz.device)
        force = torch.empty((n, 3), dtype=F64, device=xyz.device)
        # the sorted path leaves the interval sort of xyz in its workspace:
        # the backward traverses the same order
        ctx.work = eval_table(desc, owner._d, xyz, phi, force, None,
                              _lib.AGX_PHI | _lib.AGX_FORCE, xyz.device)
        ctx.save_for_backward(xyz, poly, scal, center)
        ctx.owner = owner
        ctx.set_materialize_grads(False)      # an unused output reaches the kernel as NULL
        return phi, force

    @staticmethod
    def backward(ctx, gphi, gforce):
        if gphi is None and gforce is None:
            return None, None, None, None, None
        xyz, poly, scal, center = ctx.saved_tensors
        owner = ctx.owner
        desc = owner._desc_for(poly, scal, center)
        n = xyz.shape[0]
        want_c = ctx.needs_input_grad[1]
        gx = torch.empty_like(xyz) if ctx.needs_input_grad[0] else None
        # one buffer (one fill) for the table, law and centre gradients
        gbuf = torch.zeros(poly.numel() + 8, dtype=F64, device=xyz.device)
        gpoly = gbuf[:poly.numel()].view(poly.shape)
        gscal = gbuf[poly.numel():] if want_c else gbuf[poly.numel():poly.numel() + 5]
        gphi = None if gphi is None else gphi.contiguous()
        gforce = None if gforce is None else gforce.contiguous()
        L = _lib.load()
        nwork = int(L.agx_multipole_work_size(n))
        work, flags = ctx.work, _lib.AGX_WORK_SORTED
        ctx.work = None
        if work is None:
            work, flags = torch.empty(nwork, dtype=torch.uint8, device=xyz.device), 0
        if want_c:
            flags |= _lib.AGX_ADJ_CENTER_GRAD
        rc = L.agx_multipole_adjoint(
            C.byref(desc), xyz.data_ptr(), n, None if gphi is None else gphi.data_ptr(),
            None if gforce is None else gforce.data_ptr(), None if gx is None else gx.data_ptr(),
            gpoly.data_ptr(), gscal.data_ptr(), work.data_ptr(), nwork, flags,
            _lib.stream_handle(xyz.device))
        _lib.check(rc, "agx_multipole_adjoint")
        gc = gscal[5:] if want_c else None
        return gx, gc, gpoly, gscal[:5], None


class _MultipoleBuild(torch.autograd.Function):
    """(phi_table, dphi_table) -> (poly_km, law) by agx_multipole_build on the
    device, with agx_multipole_build_adjoint as backward: one launch each way
    instead of the ~100 small torch operations of the same rebuild."""

    @staticmethod
    def forward(ctx, phi_table, dphi_table, owner):
        d = owner._d
        src = owner._source(phi_table, dphi_table)
        dev = phi_table.device
        poly = torch.empty((d["n_intervals"], d["n_lm"], 6), dtype=F64, device=dev)
        law = torch.empty(5, dtype=F64, device=dev)
        rc = _lib.load().agx_multipole_build(C.byref(src), poly.data_ptr(), law.data_ptr(),
                                             _lib.stream_handle(dev))
        _lib.check(rc, "agx_multipole_build")
        ctx.save_for_backward(phi_table, dphi_table)
        ctx.owner = owner
        return poly, law

    @staticmethod
    def backward(ctx, gpoly, glaw):
        phi_table, dphi_table = ctx.saved_tensors
        dev = phi_table.device
        src = ctx.owner._source(phi_table, dphi_table)
        d = ctx.owner._d
        if gpoly is None:
            gpoly = torch.zeros((d["n_intervals"], d["n_lm"], 6), dtype=F64, device=dev)
        if glaw is None:
            glaw = torch.zeros(5, dtype=F64, device=dev)
        gpoly, glaw = gpoly.contiguous(), glaw.contiguous()
        gtab = torch.zeros(2 * phi_table.numel(), dtype=F64, device=dev)   # one fill
        gphi = gtab[:phi_table.numel()].view(phi_table.shape)
        gdphi = gtab[phi_table.numel():].view(dphi_table.shape)
        work = torch.empty(1, dtype=F64, device=dev)
        rc = _lib.load().agx_multipole_build_adjoint(
            C.byref(src), gpoly.data_ptr(), glaw.data_ptr(), gphi.data_ptr(), gdphi.data_ptr(),
            work.data_ptr(), _lib.stream_handle(dev))
        _lib.check(rc, "agx_multipole_build_adjoint")
        return gphi, gdphi, None


class FittableMultipole(torch.nn.Module):
    """Multipole potential with its raw coefficient table as torch parameters.

    ``phi_table`` and ``dphi_table`` have the file's layout (nR, n_columns),
    columns ``lm_labels``.  forward(xyz) -> (phi, force), differentiable with
    respect to xyz and both tables.  No column is pruned by default
    (prune_threshold=0): a harmonic that is zero in the table still has a
    nonzero gradient."""

    def __init__(self, coefs: MultipoleCoefs, prune_threshold: float = 0.0, device=None):
        super().__init__()
        dev = norm_device(device) if device is not None else _lib.require_gpu()
        d = build_multipole_data(coefs, prune_threshold=prune_threshold)
        if d["resampled"]:
            raise ValueError("FittableMultipole needs a log-uniform radial grid")
        self._d = d
        self.device = dev
        r = np.asarray(coefs.R_grid, np.float64)
        self.register_buffer("r", torch.as_tensor(r, device=dev))
        self.phi_table = torch.nn.Parameter(torch.as_tensor(np.asarray(coefs.phi, np.float64), device=dev))
        self.dphi_table = torch.nn.Parameter(torch.as_tensor(np.asarray(coefs.dphi_dr, np.float64), device=dev))
        self.register_buffer("cols", torch.as_tensor(d["columns"], device=dev))
        Mf, Md = _quintic_operator(np.log(r))
        self.register_buffer("Mf", torch.as_tensor(Mf, device=dev))
        self.register_buffer("Md", torch.as_tensor(Md, device=dev))
        prog = d["program"]
        self._program = torch.as_tensor(pack_program(prog).view(np.uint8), device=dev)
        self._n_groups, self._n_steps = len(prog.grp_m), len(prog.step_l)
        self._prog_flags = program_flags(prog)
        self._cols32 = torch.as_tensor(np.asarray(d["columns"], np.int32), device=dev)
        self._modes = ({"log": 1, "s2": 2}.get(d["inner_mode"], 0), int(bool(d["inner_reset"])),
                       1 if d["outer_mode"] == "log" else 0)
        self._device_build = dev.type == "cuda" and len(r) <= 128

    # ---- differentiable table build -------------------------------------------
    def _source(self, phi_table: torch.Tensor, dphi_table: torch.Tensor):
        d = self._d
        s = _lib.agx_multipole_source()
        s.phi, s.dphi, s.r = phi_table.data_ptr(), dphi_table.data_ptr(), self.r.data_ptr()
        s.cols, s.Mf, s.Md = self._cols32.data_ptr(), self.Mf.data_ptr(), self.Md.data_ptr()
        s.nR, s.n_cols, s.n_lm = phi_table.shape[0], phi_table.shape[1], d["n_lm"]
        s.log_scaling = int(d["log_scaling"])
        s.inner_mode, s.inner_reset, s.outer_mode = self._modes
        s.dlogr = d["dlogr"]
        return s

    def tables(self):
        """(poly_km (n_int, n_lm, 6), scal [invPhi0, inner U, W, outer U, W]):
        on the GPU one agx_multipole_build launch (and one adjoint launch pair
        in backward), elsewhere the same steps as torch operations."""
        if self._device_build:
            return _MultipoleBuild.apply(self.phi_table.contiguous(), self.dphi_table.contiguous(),
                                         self)
        return self.tables_torch()

    def tables_torch(self):
        """The rebuild as torch operations (any device)."""
        d = self._d
        r = self.r
        P = self.phi_table[:, self.cols]
        Dr = self.dphi_table[:, self.cols]
        dl = Dr * r[:, None]
        p0, d0 = P[:, 0], Dr[:, 0]
        zero = torch.zeros((), dtype=F64, device=r.device)
        log_scaling = bool(d["log_scaling"])
        invPhi0 = zero
        U_in, W_in = zero, p0[0]
        if log_scaling:
            mode = d["inner_mode"]
            if mode == "log":
                U_in = r[0] * d0[0]
            elif mode == "s2":
                U_in = 0.5 * r[0] * d0[0]
                W_in = p0[0] - U_in
                if not d["inner_reset"]:
                    invPhi0 = 1.0 / W_in
        if d["outer_mode"] == "log":
            U_out, W_out = r[-1] * d0[-1] + p0[-1], p0[-1]
        else:
            U_out, W_out = zero, p0[-1]
        if log_scaling:
            f0 = torch.log(invPhi0 - 1.0 / p0)
            df0 = dl[:, 0] / (p0 * (invPhi0 * p0 - 1.0))
            ratio = P[:, 1:] / p0[:, None]
            f = torch.cat([f0[:, None], ratio], dim=1)
            df = torch.cat([df0[:, None], (dl[:, 1:] - ratio * dl[:, 0:1]) / p0[:, None]], dim=1)
        else:
            f, df = P, dl
        d2 = self.Mf @ f + self.Md @ df
        h = float(d["dlogr"])
        f0_, f1_ = f[:-1], f[1:]
        m0, m1 = df[:-1] * h, df[1:] * h
        q0, q1 = d2[:-1] * h * h, d2[1:] * h * h
        dlt = f1_ - f0_
        poly = torch.stack([f0_, m0, 0.5 * q0,
                            10.0 * dlt - 6.0 * m0 - 4.0 * m1 - 1.5 * q0 + 0.5 * q1,
                            -15.0 * dlt + 8.0 * m0 + 7.0 * m1 + 1.5 * q0 - q1,
                            6.0 * dlt - 3.0 * m0 - 3.0 * m1 + 0.5 * (q1 - q0)], dim=-1)
        scal = torch.stack([invPhi0, U_in, W_in, U_out, W_out])
        return poly.contiguous(), scal

    def _desc_for(self, poly: torch.Tensor, scal: torch.Tensor,
                  center: Optional[torch.Tensor] = None):
        """Descriptor of the rebuilt table; the five law scalars stay in device
        memory (agx_multipole.law), so no launch waits for a copy to the host,
        and so does a centre (agx_multipole.origin: the kernels evaluate at
        x - center without a shifted copy of the positions)."""
        d = self._d
        s = _lib.agx_multipole()
        s.poly_km = poly.data_ptr()
        s.program = self._program.data_ptr()
        s.n_int, s.n_lm = d["n_intervals"], d["n_lm"]
        s.n_groups, s.n_steps, s.lmax = self._n_groups, self._n_steps, d["lmax"]
        s.prog_flags = self._prog_flags
        s.log_scaling, s.inner_log, s.outer_log = d["log_scaling"], d["inner_log"], d["outer_log"]
        s.logr_min, s.dlogr, s.inv_dlogr = d["logr_min"], d["dlogr"], d["inv_dlogr"]
        s.inner_s, s.outer_s = d["inner_s"], d["outer_s"]
        s.law = scal.data_ptr()
        s.origin = center.data_ptr() if center is not None else None
        return s

    def forward(self, xyz: torch.Tensor, center: Optional[torch.Tensor] = None):
        """(phi, force) at xyz, or at xyz - center for a centre (3,)."""
        poly, scal = self.tables()
        return _MultipoleEval.apply(xyz.contiguous(), center, poly, scal.contiguous(), self)


def _spline_slope_matrix(x: np.ndarray, bc) -> np.ndarray:
    """D (n x n) with D @ f = the slopes at the nodes of the cubic spline
    through values f (boundary conditions ``bc`` are homogeneous, so the map
    is linear): the spline of each unit vector, differentiated."""
    from scipy.interpolate import CubicSpline
    return CubicSpline(x, np.eye(len(x)), bc_type=bc)(x, 1)


def _outer_fit_operator(R, z, ms, lmax_fit, r0, nR, nz):
    """The outer multipole fit of cylspline_prep.outer_asymptote as a linear map:
    W = G @ tab[:, iRp, izp].ravel(), with the boundary points (iRp, izp) and
    G ((lmax_fit+1)^2 x n_harm*npts) built from the least-squares
    pseudo-inverse of each harmonic's design matrix."""
    from scipy.linalg import lstsq
    zsym = z[0] == 0      # never after mirroring; kept for the same point set
    pts = []
    for iR in range(nR - 1):
        pts.append((R[iR], z[nz - 1], iR, nz - 1))
        pts.append((R[iR], -z[nz - 1], iR, nz - 1) if zsym else (R[iR], z[0], iR, 0))
    for iz in range(nz):
        pts.append((R[nR - 1], z[iz], nR - 1, iz))
        if zsym and iz > 0:
            pts.append((R[nR - 1], -z[iz], nR - 1, iz))
    Rp = np.array([p[0] for p in pts])
    zp = np.array([p[1] for p in pts])
    iRp = np.array([p[2] for p in pts], np.int64)
    izp = np.array([p[3] for p in pts], np.int64)
    rp = np.hypot(Rp, zp)
    ct, st = zp / rp, Rp / rp
    npts = len(pts)
    G = np.zeros(((lmax_fit + 1) ** 2, len(ms) * npts))
    for k, m in enumerate(ms):
        am = abs(m)
        if am > lmax_fit:
            continue
        ncol = lmax_fit - am + 1
        mul = MUL0 if m == 0 else MUL1
        M = np.zeros((npts, ncol))
        for p in range(npts):
            P = ptilde(lmax_fit, am, ct[p], st[p])
            M[p] = P * (rp[p] / r0) ** (-(am + np.arange(ncol) + 1.0)) * mul
        pinv = lstsq(M, np.eye(npts))[0]          # (ncol, npts)
        for li in range(ncol):
            l = am + li
            G[l * (l + 1) + m, k * npts:(k + 1) * npts] = pinv[li]
    return G, iRp, izp


class _CylSplineEval(torch.autograd.Function):
    """(xyz, nodes, W) -> (phi, force) with the adjoint kernel as backward."""

    @staticmethod
    def forward(ctx, xyz, nodes, W, owner):
        desc = owner._desc_for(nodes, W)
        n = xyz.shape[0]
        phi = torch.empty(n, dtype=F64, device=xyz.device)
        force = torch.empty((n, 3), dtype=F64, device=xyz.device)
        rc = _lib.load().agx_cylspline_eval(C.byref(desc), xyz.data_ptr(), n, phi.data_ptr(),
                                            force.data_ptr(), None, _lib.AGX_PHI | _lib.AGX_FORCE,
                                            _lib.stream_handle(xyz.device))
        _lib.check(rc, "agx_cylspline_eval")
        ctx.save_for_backward(xyz, nodes, W)
        ctx.owner = owner
        ctx.set_materialize_grads(False)      # an unused output reaches the kernel as NULL
        return phi, force

    @staticmethod
    def backward(ctx, gphi, gforce):
        if gphi is None and gforce is None:
            return None, None, None, None
        xyz, nodes, W = ctx.saved_tensors
        desc = ctx.owner._desc_for(nodes, W)
        n = xyz.shape[0]
        gx = torch.empty_like(xyz) if ctx.needs_input_grad[0] else None
        gnodes = torch.zeros_like(nodes)
        gW = torch.zeros_like(W)
        gphi = None if gphi is None else gphi.contiguous()
        gforce = None if gforce is None else gforce.contiguous()
        L = _lib.load()
        nwork = int(L.agx_cylspline_adjoint_work_size(C.byref(desc), n))
        work = torch.empty(max(nwork, 1), dtype=torch.uint8, device=xyz.device)
        rc = L.agx_cylspline_adjoint(
            C.byref(desc), xyz.data_ptr(), n, None if gphi is None else gphi.data_ptr(),
            None if gforce is None else gforce.data_ptr(), None if gx is None else gx.data_ptr(),
            gnodes.data_ptr(), gW.data_ptr(), work.data_ptr(), nwork,
            _lib.stream_handle(xyz.device))
        _lib.check(rc, "agx_cylspline_adjoint")
        return gx, gnodes, gW, None


class FittableCylSpline(torch.nn.Module):
    """CylSpline potential with its raw table Phi_m(R_i, z_j) as a torch parameter.

    ``phi_table`` is (n_harm, nR, nz) in the order ``m_values`` (|m|
    ascending, cos before sin), on the file's (R, z) grid.  forward(xyz) ->
    (phi, force), differentiable with respect to xyz and the table.  Frozen at
    construction, as the reference fixes them when it builds the tables
    (gpu_potential.py:1143-1253): the grid scale Rscale = -Mtot/Phi0, the
    outer radius r0 and whether the m=0 harmonic is log-scaled; so a gradient
    step is exact for this model and a rebuilt CylSplinePotentialGPU of the
    stepped table differs by the change of Rscale alone."""

    def __init__(self, coefs: CylSplineCoefs, device=None):
        super().__init__()
        dev = norm_device(device) if device is not None else _lib.require_gpu()
        d = build_cylspline_data(coefs)
        self._d = d
        self.device = dev
        ms = [int(m) for m in d["m_arr"]]
        self.m_values = ms
        R = np.asarray(coefs.R_grid, np.float64)
        z0 = np.asarray(coefs.z_grid, np.float64)
        self._R, self._z0 = R, z0
        self._mirror = bool(z0[0] == 0)
        z = np.concatenate([-z0[:0:-1], z0]) if self._mirror else z0
        nR, nz = len(R), len(z)
        raw = np.stack([np.asarray(coefs.phi[m], np.float64) for m in ms])
        self.phi_table = torch.nn.Parameter(torch.as_tensor(raw, device=dev))
        G, iRp, izp = _outer_fit_operator(R, z, ms, d["lmax_outer"], d["r0_outer"], nR, nz)
        W_np = G @ self._mirrored_np(raw)[:, iRp, izp].ravel()
        if not np.allclose(W_np, d["W_outer"], rtol=1e-8, atol=1e-12 * np.abs(d["W_outer"]).max()):
            raise ValueError("the outer multipole fit is degenerate for this grid")
        self.register_buffer("G", torch.as_tensor(G, device=dev))
        self.register_buffer("iRp", torch.as_tensor(iRp, device=dev))
        self.register_buffer("izp", torch.as_tensor(izp, device=dev))
        self.register_buffer("Dr", torch.as_tensor(
            _spline_slope_matrix(d["lR_grid"], ((1, np.zeros(nR)), "natural")), device=dev))
        self.register_buffer("Dz", torch.as_tensor(
            _spline_slope_matrix(d["lz_grid"], "natural"), device=dev))
        self._lR = torch.as_tensor(d["lR_grid"], device=dev)
        self._lz = torch.as_tensor(d["lz_grid"], device=dev)
        self._m = torch.as_tensor(d["m_arr"].astype(np.int32), device=dev)
        self._nR, self._nz = nR, nz

    def _mirrored_np(self, raw):
        if not self._mirror:
            return raw
        return np.concatenate([raw[:, :, :0:-1], raw], axis=2)

    # ---- differentiable table build -------------------------------------------
    def tables(self):
        """(nodes (n_harm, nR, nz, 4), W_outer [(lmax_outer+1)^2])."""
        t = self.phi_table
        full = torch.cat([t.flip(-1)[..., :-1], t], dim=-1) if self._mirror else t
        W = self.G @ full[:, self.iRp, self.izp].reshape(-1)
        if self._d["log_scaling"]:
            f = torch.cat([torch.log(-full[:1]), full[1:] / full[:1]], dim=0)
        else:
            f = full
        fy = f @ self.Dz.T
        fx = self.Dr @ f
        fxy = self.Dr @ fy
        return torch.stack([f, fx, fy, fxy], dim=-1).contiguous(), W

    def _desc_for(self, nodes: torch.Tensor, W: torch.Tensor):
        d = self._d
        s = _lib.agx_cylspline()
        s.lR_grid, s.lz_grid = self._lR.data_ptr(), self._lz.data_ptr()
        s.nodes, s.m_arr, s.W_outer = nodes.data_ptr(), self._m.data_ptr(), W.data_ptr()
        s.nR, s.nz = self._nR, self._nz
        s.n_harm, s.log_scaling = d["n_harm"], d["log_scaling"]
        s.lmax_outer, s.mmax_outer = d["lmax_outer"], d["mmax_outer"]
        s.Rscale, s.r0_outer = d["Rscale"], d["r0_outer"]
        return s

    def forward(self, xyz: torch.Tensor):
        nodes, W = self.tables()
        return _CylSplineEval.apply(xyz.contiguous(), nodes, W.contiguous(), self)

    def to_coefs(self) -> CylSplineCoefs:
        """The current table as a coefficient set (for write_coefs)."""
        tab = self.phi_table.detach().cpu().numpy()
        return CylSplineCoefs(R_grid=self._R.copy(), z_grid=self._z0.copy(),
                              phi={m: tab[k].copy() for k, m in enumerate(self.m_values)},
                              mmax=int(self._d["mmax"]))


class SplineCenter(torch.nn.Module):
    """Differentiable trajectory centre c(t) from knots (T, 3) at times (T,):
    the cubic spline of ShiftedPotentialGPU (not-a-knot, linear extrapolation
    outside [t0, tN]); c(t) = w(t) @ knots with spline weights w(t)."""

    def __init__(self, times, knots, device=None):
        super().__init__()
        dev = norm_device(device) if device is not None else _lib.require_gpu()
        self.times = np.asarray(times, np.float64)
        self.knots = torch.nn.Parameter(torch.as_tensor(np.asarray(knots, np.float64), device=dev))
        self._dev = dev
        self._w_fn = None
        self._w_cache = {}

    def weights(self, t: float) -> torch.Tensor:
        """Knot weights at t, kept on the device for the last 64 times asked
        (a fit loop at fixed times copies nothing to the device per step)."""
        w = self._w_cache.get(t)
        if w is not None:
            return w
        if self._w_fn is None:
            # the spline is linear in its values: interpolating the identity
            # gives the weight of every knot
            self._w_fn = _linear_extrapolating_spline(self.times, np.eye(len(self.times)))
        w = torch.as_tensor(np.asarray(self._w_fn(t), np.float64), device=self._dev)
        if len(self._w_cache) >= 64:
            self._w_cache.pop(next(iter(self._w_cache)))
        self._w_cache[t] = w
        return w

    def forward(self, t: float) -> torch.Tensor:
        return self.weights(float(t)) @ self.knots


class _Shift(torch.autograd.Function):
    """x - c for x (N, 3) and a centre c (3,).  The centre's gradient -sum_i g_i
    is summed as a (rows, 3 * 4096) block, column sums folded to 3: torch's
    column reduction of a (N, 3) tensor runs on a handful of workgroups
    (0.34 ms at N = 1M on MI355X against 0.03 ms this way), and a
    (1 x N) @ (N x 3) product goes to a GEMM tile without split-K."""

    @staticmethod
    def forward(ctx, x, c):
        return x - c[None, :]

    @staticmethod
    def backward(ctx, g):
        gc = None
        if ctx.needs_input_grad[1]:
            n = g.shape[0]
            m = n - n % 4096
            gc = g[m:].sum(dim=0)
            if m:
                gc = gc + g[:m].reshape(m // 4096, 3 * 4096).sum(dim=0).view(4096, 3).sum(dim=0)
            gc = -gc
        return g, gc


class FittableShiftedMultipole(torch.nn.Module):
    """Expansion at x - c(t) (a FittableMultipole or FittableCylSpline):
    gradients reach the table and the centre knots."""

    def __init__(self, multipole: FittableMultipole, center: SplineCenter):
        super().__init__()
        self.multipole = multipole
        self.center = center

    def forward(self, xyz: torch.Tensor, t: float = 0.0):
        if isinstance(self.multipole, FittableMultipole):
            return self.multipole(xyz, self.center(t))   # the shift inside the evaluation
        return self.multipole(_Shift.apply(xyz, self.center(t)))
