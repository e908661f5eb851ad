"""The Agama coefficient text format (agama_gpu_amd/coefs.py; the reference
reads it with nbody_streams' un-vendored read_coefs, gpu_potential.py:
683-686, 2219-2223): write_coefs followed by read_coefs returns every number
bit for bit (full-precision repr), for the reference's three fixtures and for
generated Multipole and CylSpline tables, and the re-read tables build the
same device tables."""
import numpy as np
import pytest

from conftest import FIXTURES
from agama_gpu_amd.coefs import CylSplineCoefs, MultipoleCoefs, read_coefs, write_coefs

FILES = ["600.dark.none_8.coef_mul_DR", "100.LMC.none_8.coef_mult", "600.bar.none_8.coef_cylsp_DR"]


def _same(a, b):
    assert type(a) is type(b)
    assert np.array_equal(a.R_grid, b.R_grid)
    if isinstance(a, MultipoleCoefs):
        assert a.lm_labels == b.lm_labels and a.lmax == b.lmax
        assert np.array_equal(a.phi, b.phi)
        assert (a.dphi_dr is None) == (b.dphi_dr is None)
        if a.dphi_dr is not None:
            assert np.array_equal(a.dphi_dr, b.dphi_dr)
    else:
        assert np.array_equal(a.z_grid, b.z_grid) and a.mmax == b.mmax
        assert sorted(a.phi) == sorted(b.phi)
        for m in a.phi:
            assert np.array_equal(a.phi[m], b.phi[m]), m


@pytest.mark.parametrize("fname", FILES)
def test_fixture_round_trip(fname, tmp_path):
    c = read_coefs(FIXTURES / fname)
    path = tmp_path / "out.coef"
    text = write_coefs(c, path)
    assert path.read_text() == text
    c2 = read_coefs(path)
    _same(c, c2)
    # and once more: the written text is a fixed point
    assert write_coefs(c2) == text
    # read from section text (the INI path's inline Coefficients)
    _same(c, read_coefs(text))


def test_generated_tables_round_trip(tmp_path):
    from agama_gpu_amd.workloads import halo_lmax8_coefs, mw_disk_bar_coefs
    from agama_gpu_amd.multipole_prep import build_multipole_data
    from agama_gpu_amd.cylspline_prep import build_cylspline_data
    for c in (halo_lmax8_coefs(), mw_disk_bar_coefs()):
        c2 = read_coefs(write_coefs(c))
        _same(c, c2)
        if isinstance(c, MultipoleCoefs):
            d1, d2 = build_multipole_data(c), build_multipole_data(c2)
            assert np.array_equal(d1["poly"], d2["poly"])
        else:
            d1, d2 = build_cylspline_data(c), build_cylspline_data(c2)
            assert np.array_equal(d1["node_arr"], d2["node_arr"])
            assert np.array_equal(d1["W_outer"], d2["W_outer"])


def test_special_values_round_trip():
    """Zeros (pruned harmonics), negative zero, subnormal and extreme values
    survive the text form exactly; a table without dPhi/dr writes no second
    block."""
    r = np.exp(np.linspace(-1.0, 2.0, 6))
    labels = [(0, 0), (1, -1), (1, 0), (1, 1)]
    phi = np.array([[-1.0, 0.0, -0.0, 5e-324], [-2.5e300, 1e-300, 3.0, -7.0]] * 3)
    c = MultipoleCoefs(r, phi, phi * 0.5, labels, 1)
    c2 = read_coefs(write_coefs(c))
    _same(c, c2)
    assert np.array_equal(np.signbit(c.phi), np.signbit(c2.phi))
    c3 = MultipoleCoefs(r, phi, None, labels, 1)
    t = write_coefs(c3)
    assert "#dPhi/dr" not in t and read_coefs(t).dphi_dr is None
    cs = CylSplineCoefs(np.array([0.0, 1.0, 2.0]), np.array([-1.0, 0.0, 1.0]),
                        {0: -np.arange(9.0).reshape(3, 3) - 1, -1: np.full((3, 3), 1e-310)}, 1)
    _same(cs, read_coefs(write_coefs(cs)))
