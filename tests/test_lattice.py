"""The learned traveltime's t0 lattice (pm_amd.lattice, CPU).

The node law is the gather's own (lattice_next in pstm_hip.hip): the GPU
read-back tests (tests/test_learned.py, tests/test_learned_curved.py) compare
the gather's learned times with pm_amd.surrogate.lattice_times, which uses
these nodes, for uniform and growing lattices.  Here: the law's closed form
for uniform spacing, its invariants, the argument checks, the orders of the
node differences behind the gather's cubic Hermite slopes (exact for a
quadratic with three nodes on a side), and that the cubic on the growing
lattice interpolates the exact VTI factor (anisotropy_vti.py:181-220) better
than lines on a uniform lattice with three times the nodes.
"""

from __future__ import annotations

import numpy as np
import pytest

from pm_amd.lattice import GEOMETRIC, SPARSE, Lattice, as_lattice


@pytest.mark.parametrize("nt", [2, 3, 17, 500, 1500, 2001])
@pytest.mark.parametrize("k", [4, 12, 16, 64, 512])
def test_uniform_lattice_closed_form(nt, k):
    want = np.unique(np.r_[np.arange(0, nt, k), nt - 1])
    np.testing.assert_array_equal(Lattice(k).nodes(nt), want)


@pytest.mark.parametrize("lat", [GEOMETRIC, SPARSE, Lattice(4, 256, 1024), Lattice(5, 300, 7), Lattice(16, 77, 128)])
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


def _hermite(gn, nodes, k):
    """g at samples k from node values gn as the gather interpolates it: the
    cubic Hermite polynomial of each segment with node_derivative slopes."""
    import torch

    from pm_amd.surrogate import node_derivative

    nj = len(nodes)
    nodes_t = torch.as_tensor(nodes)
    kt = torch.as_tensor(k)
    j = (torch.searchsorted(nodes_t, kt, right=True) - 1).clamp(0, nj - 2)
    jj = [j - 1, j, j + 1, j + 2]
    ok = [(q >= 0) & (q < nj) for q in jj]
    x = [nodes_t[q.clamp(0, nj - 1)].double() for q in jj]
    g = [torch.as_tensor(gn)[q.clamp(0, nj - 1)] for q in jj]
    L = x[2] - x[1]
    b0, b1 = L * node_derivative(g, ok, x, 1), L * node_derivative(g, ok, x, 2)
    f = (kt - nodes_t[j]).double() / L
    return (g[1] + f * (b0 + f * ((3 * (g[2] - g[1]) - 2 * b0 - b1) + f * (2 * (g[1] - g[2]) + b0 + b1)))).numpy()


def _leg_error(nodes, cubic, nt=1500, dt=0.002):
    """Largest one-leg time error (samples) of interpolating the exact VTI
    factor g = -2 x / (tau^2 + (1 + 2 eta) x), x = (d / v)^2, over the nodes
    (cubic: the gather's Hermite; else linear), across leg distances,
    velocities and eta = 0.2 (the factor's largest excursion)."""
    k = np.arange(nt)
    j = np.clip(np.searchsorted(nodes, k, side="right") - 1, 0, len(nodes) - 2)
    f = (k - nodes[j]) / (nodes[j + 1] - nodes[j])
    worst = 0.0
    for d in (50.0, 150.0, 400.0, 1000.0, 2500.0):
        for v in (1500.0, 2500.0, 4500.0):
            x, eta = (d / v) ** 2, 0.2
            g = lambda tau: -2 * x / (tau ** 2 + (1 + 2 * eta) * x)  # noqa: E731
            tau, taun = k * dt / 2, nodes * dt / 2
            gi = _hermite(g(taun), nodes, k) if cubic else g(taun)[j] + f * (g(taun)[j + 1] - g(taun)[j])
            t = np.sqrt(tau ** 2 + x * (1 + eta * g(tau)))
            ti = np.sqrt(tau ** 2 + x * np.maximum(1 + eta * gi, 0))
            worst = max(worst, float(np.abs(t - ti).max() / dt))
    return worst


def test_growing_cubic_lattice_beats_uniform_linear_on_the_exact_factor():
    """The gather's cubic on the benchmark's growing lattice (31 nodes over
    1500 samples, under a third of the uniform 16-sample lattice's 95)
    interpolates the factor to under half the worst error of lines through
    the uniform lattice's nodes, and to under three quarters of that of lines
    through its own."""
    uni, geo = Lattice(16).nodes(1500), SPARSE.nodes(1500)
    e_uni, e_geo, e_geo_lin = _leg_error(uni, False), _leg_error(geo, True), _leg_error(geo, False)
    assert len(geo) * 3 < len(uni)
    assert e_geo < 0.5 * e_uni < 0.05
    assert e_geo < 0.75 * e_geo_lin


@pytest.mark.parametrize("ok", [(True, True, True, True), (False, True, True, True), (True, True, True, False),
                                (False, True, True, False)])
def test_node_derivative_orders(ok):
    """Slopes at nodes 1 and 2 of four nonuniform nodes: exact for a quadratic
    with three nodes on a side (central or one-sided), for a line with two."""
    import torch

    from pm_amd.surrogate import node_derivative

    x = [torch.tensor([v], dtype=torch.float64) for v in (3.0, 11.0, 16.0, 31.0)]
    okt = [torch.tensor([o]) for o in ok]
    a, b, c = 0.7, -1.3, 0.045
    quad = [a + b * xi + c * xi * xi for xi in x]
    line = [a + b * xi for xi in x]
    for i in (1, 2):
        three = ok[i - 1] and ok[i + 1] or (i == 1 and ok[3]) or (i == 2 and ok[0])
        f = quad if three else line
        want = b + (2 * c * x[i] if three else torch.zeros(1, dtype=torch.float64))
        assert torch.allclose(node_derivative(f, okt, x, i), want, rtol=0, atol=1e-12)
