"""Pins the oracle's Newell demag tensor.

* against golden vectors written by the reference's own Boris/DemagTFunc*.cpp (compiled
  into oracle/_ref by oracle/Makefile.ref; fixtures made by tests/golden/make_fixtures.py):
  exact region (distance < 40 cells) and far-field series, cubic and non-cubic cells, 2D/3D;
* against closed forms: cube self-demag -1/3, trace -1, far-field dipole limit, symmetries.
"""
from pathlib import Path

import numpy as np
import pytest

import oracle_lib as O

G = np.load(Path(__file__).parent / "golden" / "tensor_golden.npz")


@pytest.mark.parametrize("name", ["c3a", "c3b", "c3c", "c3d"])
def test_tensor3d_matches_reference_build(name):
    n, N, h = tuple(G[name + "_n"]), tuple(G[name + "_N"]), tuple(G[name + "_h"])
    Dd, Do = O.tensor3d(n, N, h)
    idx = G[name + "_idx"]
    got_d = Dd[idx[:, 0], idx[:, 1], idx[:, 2]]
    got_o = Do[idx[:, 0], idx[:, 1], idx[:, 2]]
    np.testing.assert_allclose(got_d, G[name + "_diag"], rtol=1e-14, atol=1e-300)
    np.testing.assert_allclose(got_o, G[name + "_off"], rtol=1e-14, atol=1e-18)
    # the oracle restates the reference's operation order: bit-identical on this toolchain
    assert np.array_equal(got_d, G[name + "_diag"]) and np.array_equal(got_o, G[name + "_off"])


@pytest.mark.parametrize("name", ["c2a", "c2b"])
def test_tensor2d_matches_reference_build(name):
    n, N, h = tuple(G[name + "_n"]), tuple(G[name + "_N"]), tuple(G[name + "_h"])
    Dd, Dxy = O.tensor2d(n, N, h)
    idx = G[name + "_idx"]
    assert np.array_equal(Dd[idx[:, 0], idx[:, 1]], G[name + "_diag"])
    assert np.array_equal(Dxy[idx[:, 0], idx[:, 1]], G[name + "_off"])


def test_self_demag_matches_reference_build_and_closed_forms():
    L = O.lib()
    for h, ref in zip(G["self_h"], G["self_val"]):
        out = np.zeros(3)
        L.or_self_demag(*h, out)
        np.testing.assert_allclose(out, ref, rtol=1e-15)
        assert abs(out.sum() + 1.0) < 1e-13          # trace N = 1, stored as -N
    out = np.zeros(3)
    L.or_self_demag(1.0, 1.0, 1.0, out)
    np.testing.assert_allclose(out, [-1 / 3] * 3, rtol=1e-14)


def _kn_sum(v):
    s = c = 0.0
    for x in v:
        t = s + x
        c += ((s - t) + x) if abs(s) > abs(x) else ((x - t) + s)
        s = t
    return s + c


def _stencil(fn, x, y, z, hx, hy, hz):
    w = {0: 8, 1: -4, 2: 2, 3: -1}
    offs = [(0, 0, 0), (1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1),
            (-1, -1, 0), (-1, 1, 0), (1, -1, 0), (1, 1, 0), (-1, 0, -1), (-1, 0, 1), (1, 0, -1), (1, 0, 1),
            (0, -1, -1), (0, -1, 1), (0, 1, -1), (0, 1, 1),
            (-1, -1, -1), (-1, -1, 1), (-1, 1, -1), (1, -1, -1), (-1, 1, 1), (1, -1, 1), (1, 1, -1), (1, 1, 1)]
    terms = [w[abs(a) + abs(b) + abs(c)] * fn(x + a * hx, y + b * hy, z + c * hz) for a, b, c in offs]
    return _kn_sum(terms) / (4 * np.pi * hx * hy * hz)


def test_single_value_newell_matches_reference_build():
    """f and g off the grid (the reference's Ldia_single / Lodia_single, DemagTFunc_Test.cpp)."""
    L = O.lib()
    for row, ref in zip(G["single_in"], G["single_val"]):
        x, y, z, hx, hy, hz = row
        dia = [-_stencil(L.or_newell_f, x, y, z, hx, hy, hz), -_stencil(L.or_newell_f, y, x, z, hy, hx, hz),
               -_stencil(L.or_newell_f, z, y, x, hz, hy, hx)]
        off = [-_stencil(L.or_newell_g, x, y, z, hx, hy, hz), -_stencil(L.or_newell_g, x, z, y, hx, hz, hy),
               -_stencil(L.or_newell_g, y, z, x, hy, hz, hx)]
        np.testing.assert_allclose(dia + off, ref, rtol=1e-12, atol=1e-15)


def test_far_field_is_point_dipole():
    """At large distance the tensor tends to the dipole kernel V/(4 pi) (3 r r^T - R^2) / R^5."""
    n, N, h = (120, 100, 2), (240, 200, 4), (1.0, 1.0, 1.0)
    Dd, Do = O.tensor3d(n, N, h)
    for (i, j, k) in [(110, 3, 0), (60, 90, 1), (119, 99, 1)]:
        r = np.array([i, j, k], float)
        R = np.linalg.norm(r)
        dip = (3 * np.outer(r, r) - R * R * np.eye(3)) / (4 * np.pi * R ** 5)
        got = np.array([[Dd[k, j, i, 0], Do[k, j, i, 0], Do[k, j, i, 1]],
                        [Do[k, j, i, 0], Dd[k, j, i, 1], Do[k, j, i, 2]],
                        [Do[k, j, i, 1], Do[k, j, i, 2], Dd[k, j, i, 2]]])
        assert np.abs(got - dip).max() <= 2e-3 * np.abs(dip).max()


def test_asymptotic_series_continuous_with_exact_at_switch_distance():
    """The far-field series (used from 40 cells) agrees with the exact Newell stencil there, to
    the exact stencil's own cancellation error (f, g ~ R^3 summed to a ~1/R^3 result)."""
    L = O.lib()
    h = (1.0, 1.0, 0.75)
    for (x, y, z) in [(40.0, 0.0, 0.0), (28.0, 28.0, 2.0), (0.0, 39.0, 11.0)]:
        exact = _stencil(L.or_newell_f, x, y, z, *h)
        asym = L.or_asympt_diag(*h, x, y, z)
        assert abs(exact - asym) <= 2e-5 * abs(exact)
        exact_o = _stencil(L.or_newell_g, x + 1.0, y + 1.0, z, *h)
        asym_o = L.or_asympt_offdiag(*h, x + 1.0, y + 1.0, z)
        assert abs(exact_o - asym_o) <= 2e-5 * abs(exact_o)


def test_tensor_symmetries():
    n, N, h = (7, 5, 4), (14, 10, 8), (1.0, 0.8, 0.6)
    Dd, Do = O.tensor3d(n, N, h)
    for (i, j, k) in [(1, 2, 3), (3, 1, 2), (6, 4, 1)]:
        for si, sj, sk in [(-1, 1, 1), (1, -1, 1), (1, 1, -1), (-1, -1, -1)]:
            a = Dd[k, j, i]; b = Dd[(sk * k) % N[2], (sj * j) % N[1], (si * i) % N[0]]
            np.testing.assert_array_equal(a, b)
            o = Do[(sk * k) % N[2], (sj * j) % N[1], (si * i) % N[0]]
            np.testing.assert_array_equal(o, Do[k, j, i] * np.array([si * sj, si * sk, sj * sk]))
    # padding planes between n and N - n + 1 are zero
    assert np.all(Dd[:, :, n[0]] == 0) and np.all(Do[n[2]:N[2] - n[2] + 1] == 0)
