"""Pins the oracle's layer stack (one ferromagnetic mesh per layer, coupled by the multilayered
demag, as the reference's SDemag stack): per-layer material values and in-layer exchange reduce
to independent single-layer meshes, a one-layer stack to the plain 2D mesh, and two adjacent
equal layers without exchange to the plain two-plane 3D mesh."""
import numpy as np

import oracle_lib as O
from boris2_amd.api import (BM_EVAL_RK4, BM_EVAL_RKDP54, BM_MOD_ANIS_UNI, BM_MOD_DEMAG, BM_MOD_EXCHANGE,
                            BM_MOD_ZEEMAN, FMParams, random_magnetisation)

NX, NY, HX, HY = 12, 8, 4e-9, 4e-9


def tilted(ncell, Ms, seed, tilt=0.3):
    rng = np.random.default_rng(seed)
    th = tilt * rng.random(ncell)
    ph = 2 * np.pi * rng.random(ncell)
    m = np.stack([np.cos(th), np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph)], 1)
    return m * np.asarray(Ms, dtype=np.float64).reshape(-1, 1)


def test_layer_values_without_demag_match_independent_meshes():
    """Demag off: a 3-layer stack with per-layer Ms, A, alpha, K1 evolves each layer exactly as
    a single 2D mesh with that layer's values (no exchange across layers)."""
    mods = BM_MOD_EXCHANGE | BM_MOD_ANIS_UNI | BM_MOD_ZEEMAN
    base = FMParams(Ms=8e5, A=1.3e-11, alpha=0.02, K1=0.0, ea=(0.0, 0.0, 1.0), Ha=(2e4, 5e3, 0.0), modules=mods)
    Ms, A, al, K1 = [6e5, 9e5, 1.1e6], [1.0e-11, 1.3e-11, 2.0e-11], [0.02, 0.1, 0.3], [0.0, 3e5, 1e5]
    hz = [8e-9, 4e-9, 16e-9]
    nl = NX * NY
    M0 = tilted(3 * nl, np.repeat(Ms, nl), seed=3)
    st = O.OracleSim(BM_EVAL_RK4, (NX, NY, 3), (HX, HY, 4e-9), base)
    st.set_layers(HX, HY, hz, [0.0, 12e-9, 20e-9], with_demag=False, Ms=Ms, A=A, alpha=al, K1=K1)
    st.set_M(M0); st.set_dt(1e-13); st.run_to_iteration(60)
    Mst = st.get_M()
    for l in range(3):
        p = FMParams(**{**base.__dict__, "Ms": Ms[l], "A": A[l], "alpha": al[l], "K1": K1[l]})
        s = O.OracleSim(BM_EVAL_RK4, (NX, NY, 1), (HX, HY, hz[l]), p)
        s.set_M(M0[l * nl:(l + 1) * nl]); s.set_dt(1e-13); s.run_to_iteration(60)
        np.testing.assert_allclose(Mst[l * nl:(l + 1) * nl], s.get_M(), rtol=0, atol=1e-9 * Ms[l])


def test_one_layer_stack_is_the_plain_2d_mesh():
    p = FMParams(Ms=8e5, A=1.3e-11, alpha=0.05, Ha=(-2e4, 3e3, 0.0), modules=BM_MOD_DEMAG | BM_MOD_EXCHANGE | BM_MOD_ZEEMAN)
    M0 = tilted(NX * NY, np.full(NX * NY, p.Ms), seed=5)
    plain = O.OracleSim(BM_EVAL_RKDP54, (NX, NY, 1), (HX, HY, 5e-9), p)
    plain.set_M(M0); plain.run_to_iteration(40)
    q = FMParams(**{**p.__dict__, "modules": p.modules & ~BM_MOD_DEMAG})
    st = O.OracleSim(BM_EVAL_RKDP54, (NX, NY, 1), (HX, HY, 5e-9), q)
    st.set_layers(HX, HY, [5e-9], [0.0], with_demag=True)
    st.set_M(M0); st.run_to_iteration(40)
    np.testing.assert_allclose(st.get_M(), plain.get_M(), rtol=0, atol=1e-9 * p.Ms)
    assert abs(st.state()["time"] - plain.state()["time"]) < 1e-9 * plain.state()["time"]


def test_two_adjacent_layers_are_the_two_plane_mesh():
    """Two equal adjacent layers, exchange off: the multilayered demag couples them exactly as
    the 3D demag of a two-plane mesh (sign and size of the z shift, both directions)."""
    p = FMParams(Ms=8e5, alpha=0.1, Ha=(1e4, 0.0, 2e4), modules=BM_MOD_DEMAG | BM_MOD_ZEEMAN)
    hz = 3e-9
    M0 = tilted(2 * NX * NY, np.full(2 * NX * NY, p.Ms), seed=9, tilt=1.0)
    plain = O.OracleSim(BM_EVAL_RK4, (NX, NY, 2), (HX, HY, hz), p)
    plain.set_M(M0); plain.update_fields()
    q = FMParams(**{**p.__dict__, "modules": p.modules & ~BM_MOD_DEMAG})
    st = O.OracleSim(BM_EVAL_RK4, (NX, NY, 2), (HX, HY, hz), q)
    st.set_layers(HX, HY, [hz, hz], [0.0, hz], with_demag=True)
    st.set_M(M0); st.update_fields()
    np.testing.assert_allclose(st.heff(), plain.heff(), rtol=0, atol=1e-9 * np.abs(plain.heff()).max())
    for s in (plain, st):
        s.set_dt(2e-13); s.run_to_iteration(30)
    np.testing.assert_allclose(st.get_M(), plain.get_M(), rtol=0, atol=1e-8 * p.Ms)


def test_one_mesh_of_touching_planes_is_the_3d_mesh():
    """Stacks of 3D meshes: four touching planes grouped as one mesh (exchange across the
    planes, 5 nm spacing) integrate exactly as the plain four-plane 3D mesh with demag,
    exchange and a field; per-layer fields all equal to the field change nothing."""
    p = FMParams(Ms=8e5, A=1.3e-11, alpha=0.05, Ha=(2e4, 3e3, 0.0), modules=BM_MOD_DEMAG | BM_MOD_EXCHANGE | BM_MOD_ZEEMAN)
    hz, L = 5e-9, 4
    M0 = tilted(L * NX * NY, np.full(L * NX * NY, p.Ms), seed=21, tilt=0.8)
    plain = O.OracleSim(BM_EVAL_RK4, (NX, NY, L), (HX, HY, hz), p)
    q = FMParams(**{**p.__dict__, "modules": p.modules & ~BM_MOD_DEMAG})
    st = O.OracleSim(BM_EVAL_RK4, (NX, NY, L), (HX, HY, hz), q)
    st.set_layers(HX, HY, [hz] * L, [k * hz for k in range(L)], with_demag=True)
    st.set_layer_meshes([0] * L, np.tile(p.Ha, (L, 1)))
    for s in (plain, st):
        s.set_M(M0); s.set_dt(1e-13); s.run_to_iteration(40)
    np.testing.assert_allclose(st.get_M(), plain.get_M(), rtol=0, atol=1e-8 * p.Ms)


def test_meshes_of_a_stack_exchange_only_within_themselves():
    """Demag off, three meshes of 2, 1, 2 planes with their own fields: each evolves as the
    separate 3D mesh (or 2D mesh) with its field."""
    p = FMParams(Ms=8e5, A=1.3e-11, alpha=0.05, modules=BM_MOD_EXCHANGE | BM_MOD_ZEEMAN)
    hz = 4e-9
    groups, fields = [0, 0, 1, 2, 2], np.array([[2e4, 1e3, 0], [2e4, 1e3, 0], [-5e3, 4e3, 0], [1e4, -2e4, 3e3], [1e4, -2e4, 3e3]], float)
    z0 = [0.0, 4e-9, 9e-9, 14e-9, 18e-9]
    nl = NX * NY
    M0 = tilted(5 * nl, np.full(5 * nl, p.Ms), seed=4, tilt=0.9)
    st = O.OracleSim(BM_EVAL_RK4, (NX, NY, 5), (HX, HY, hz), p)
    st.set_layers(HX, HY, [hz] * 5, z0, with_demag=False)
    st.set_layer_meshes(groups, fields)
    st.set_M(M0); st.set_dt(1e-13); st.run_to_iteration(50)
    for first, count in ((0, 2), (2, 1), (3, 2)):
        q = FMParams(**{**p.__dict__, "Ha": tuple(fields[first])})
        s = O.OracleSim(BM_EVAL_RK4, (NX, NY, count), (HX, HY, hz), q)
        s.set_M(M0[first * nl:(first + count) * nl]); s.set_dt(1e-13); s.run_to_iteration(50)
        np.testing.assert_allclose(st.get_M()[first * nl:(first + count) * nl], s.get_M(), rtol=0, atol=1e-9 * p.Ms)
