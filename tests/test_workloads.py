"""Synthetic tables of the parity configurations, checked on the CPU.

* The CylSpline disk + bar table's m=0 harmonic in the z=0 plane equals the
  Miyamoto-Nagai potential plus the bar's azimuthal mean (zero): the 16-point
  FFT of a potential with m = 0, 2, 4, 6 only is exact, so the harmonics
  present are exactly m=0 and the cos/sin pairs of 2, 4, 6.
* The high-order table: lmax 32, all 1089 columns, the l=0 column of a
  Hernquist-like profile (negative, increasing), l>0 columns a perturbation.
"""
import math

import numpy as np


def test_disk_bar_table():
    from agama_gpu_amd.workloads import mw_disk_bar_coefs
    from agama_gpu_amd.multipole_prep import G_AGAMA
    c = mw_disk_bar_coefs()
    assert sorted(c.phi) == [-6, -4, -2, 0, 2, 4, 6]
    R, z = c.R_grid, c.z_grid
    mn = -G_AGAMA * 6e10 / np.sqrt(R[:, None] ** 2 + (3.0 + np.sqrt(z[None, :] ** 2 + 0.09)) ** 2)
    assert np.allclose(c.phi[0], mn, rtol=1e-12, atol=0)
    # bar angle 0.3 rad for m=2: Phi_2 cos 2phi + Phi_-2 sin 2phi = A cos 2(phi - 0.3)
    dev = c.phi[-2] - math.tan(0.6) * c.phi[2]
    assert np.max(np.abs(dev)) < 1e-12 * np.max(np.abs(c.phi[2]))


def test_high_order_table_shape():
    from agama_gpu_amd.workloads import high_order_coefs
    c = high_order_coefs()
    assert c.lmax == 32 and c.phi.shape == (80, 1089)
    p0 = c.phi[:, 0]
    assert np.all(p0 < 0) and np.all(np.diff(p0) > 0)
    assert np.abs(c.phi[:, 1:]).max() < 0.1 * np.abs(p0).max()
