"""The learned traveltime's t0 lattice (pm_amd.lattice, CPU).

The node law is the gather's own (lattice_next in pstm_hip.hip): the GPU
read-back tests (tests/test_learned.py, tests/test_learned_curved.py) compare
the gather's learned times with pm_amd.surrogate.lattice_times, which uses
these nodes, for uniform and growing lattices.  Here: the law's closed form
for uniform spacing, its invariants, the argument checks, and that the
growing lattice interpolates the exact VTI factor (anisotropy_vti.py:181-220)
to the stated error with far fewer nodes than the uniform one.
"""

from __future__ import annotations

import numpy as np
import pytest

from pm_amd.lattice import GEOMETRIC, Lattice, as_lattice


@pytest.mark.parametrize("nt", [2, 3, 17, 500, 1500, 2001])
@pytest.mark.parametrize("k", [4, 12, 16, 64, 512])
def test_uniform_lattice_closed_form(nt, k):
    want = np.unique(np.r_[np.arange(0, nt, k), nt - 1])
    np.testing.assert_array_equal(Lattice(k).nodes(nt), want)


@pytest.mark.parametrize("lat", [GEOMETRIC, Lattice(4, 256, 1024), Lattice(5, 300, 7), Lattice(16, 77, 128)])
@pytest.mark.parametrize("nt", [2, 9, 100, 1500, 2001])
def test_growing_lattice_invariants(lat, nt):
    n = lat.nodes(nt)
    assert n[0] == 0 and n[-1] == nt - 1
    gap = np.diff(n)
    assert (gap > 0).all() and (gap <= lat.kmax).all()
    # every gap but the last follows the law exactly
    want = np.minimum(np.maximum((n[:-1] * lat.growth) >> 8, lat.spacing), lat.kmax)
    np.testing.assert_array_equal(gap[:-1], want[:-1])
    assert gap[-1] <= want[-1]


def test_lattice_arguments():
    for bad in (dict(spacing=3), dict(spacing=513), dict(growth=-1), dict(growth=1025),
                dict(spacing=16, max_spacing=8), dict(max_spacing=2000)):
        with pytest.raises(ValueError):
            Lattice(**bad)
    with pytest.raises(ValueError):
        Lattice().nodes(1)
    assert as_lattice(12) == Lattice(12) and as_lattice(GEOMETRIC) is GEOMETRIC


def _leg_error(nodes, nt=1500, dt=0.002):
    """Largest one-leg time error (samples) of linear interpolation of the
    exact VTI factor c = -2 eta x / (tau^2 + (1 + 2 eta) x), x = (d / v)^2,
    over the nodes, across leg distances, velocities and eta = 0.2 (the
    factor's largest excursion)."""
    k = np.arange(nt)
    j = np.clip(np.searchsorted(nodes, k, side="right") - 1, 0, len(nodes) - 2)
    f = (k - nodes[j]) / (nodes[j + 1] - nodes[j])
    worst = 0.0
    for d in np.r_[np.linspace(2, 200, 60), np.linspace(200, 4000, 40)]:
        for v in (1500.0, 2500.0, 4500.0):
            x, eta = (d / v) ** 2, 0.2
            c = lambda tau: -2 * eta * x / (tau ** 2 + (1 + 2 * eta) * x)  # noqa: E731
            tau, taun = k * dt / 2, nodes * dt / 2
            ci = c(taun)[j] + f * (c(taun)[j + 1] - c(taun)[j])
            t = np.sqrt(tau ** 2 + x * (1 + c(tau)))
            ti = np.sqrt(tau ** 2 + x * np.maximum(1 + ci, 0))
            worst = max(worst, float(np.abs(t - ti).max() / dt))
    return worst


def test_growing_lattice_beats_uniform_on_the_exact_factor():
    uni, geo = Lattice(16).nodes(1500), GEOMETRIC.nodes(1500)
    e_uni, e_geo = _leg_error(uni), _leg_error(geo)
    # well under half the nodes at half the interpolation error
    assert len(geo) * 2.3 < len(uni)
    assert e_geo < 0.6 * e_uni < 0.06
