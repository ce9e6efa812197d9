"""Pinning the oracle's Bloch (complex) path and TF/SF box (CPU only).

Bloch invariants follow the reference's tests/test_bloch.py: k = 0 complex == real run
(:176-212), k = 1/2 == antiperiodic doubled domain (:217-250), discrete Bloch dispersion
(:256-308).  TF/SF: the incident field stays inside the box (scattered-field region empty to
fp64 roundoff) for 6-face boxes and for an oblique box on a Bloch cell, which pins the face
signs, extents and the 1D line's consistency with the 3D operator.
"""

from dataclasses import replace

import numpy as np
import pytest

from openem_amd import cpml, scenes, tables, waveform
from openem_amd.grid import Grid, uniform_axis
from openem_amd.model import PointDipole, Scene
from oracle.yee64 import Oracle64


def test_k0_complex_path_equals_real_path():
    sc = scenes.mini_scene(num_steps=120, plane_wave_on=True, magnetic_dipole=True)
    a = Oracle64(tables.build(sc)).run(120)
    b = Oracle64(tables.build(sc, force_complex=True)).run(120)
    for f_r, f_c in ((a.E, b.E), (a.H, b.H)):
        assert np.abs(f_c.imag).max() <= 1e-12 * np.abs(f_r).max()
        assert np.allclose(f_c.real, f_r, rtol=0, atol=1e-12 * np.abs(f_r).max())


def _line_scene(nx, bloch_kx, dipoles, n_steps, dl=20e-9):
    g = Grid(uniform_axis(nx, dl, 0, "BlochBoundary" if bloch_kx else "Periodic"),
             uniform_axis(4, dl, 0, "Periodic"), uniform_axis(30, dl, 0, "PML"))
    dt = cpml.courant_dt((dl, dl, dl))
    one = np.ones(g.shape)
    return Scene(grid=g, dt=dt, num_time_steps=n_steps, eps_ex=one, eps_ey=one, eps_ez=one,
                 pml=scenes.pml_faces(dt, (dl, dl, dl), [(2, "lo"), (2, "hi")], 6),
                 bloch_k=(bloch_kx, 0.0, 0.0), dipoles=dipoles)


def test_half_k_equals_antiperiodic_doubled_domain():
    nx, n = 10, 150
    dt = cpml.courant_dt((20e-9,) * 3)
    wf = waveform.sample(waveform.GaussianPulse(freq0=3e14, fwidth=1e14), dt, n)
    d = lambda i, s: PointDipole(2, np.array([[i, 1, 15]]), np.array([s * 1e5]), wf)  # noqa
    a = Oracle64(tables.build(_line_scene(nx, 0.5, [d(3, 1.0)], n))).run(n)
    b = Oracle64(tables.build(_line_scene(2 * nx, 0.0, [d(3, 1.0), d(3 + nx, -1.0)], n))).run(n)
    ref = np.abs(b.E).max()
    # the Bloch path injects the complex waveform; with real masks (-1 at the wrap) its real
    # part is the response to the real waveform the doubled real run injects
    assert np.allclose(a.E.real, b.E[:, :nx], atol=1e-11 * ref)
    assert np.allclose(a.E.real, -b.E[:, nx:], atol=1e-11 * ref)


@pytest.mark.parametrize("m,k", [(0, 0.25), (1, 0.3), (2, -0.2)])
def test_bloch_mode_follows_discrete_dispersion(m, k):
    nx, dx = 16, 20e-9
    g = Grid(uniform_axis(nx, dx, 0, "BlochBoundary"), uniform_axis(1, dx, 0, "Periodic"),
             uniform_axis(1, dx, 0, "Periodic"))
    dt = cpml.courant_dt((dx, dx, dx))
    one = np.ones(g.shape)
    sc = Scene(grid=g, dt=dt, num_time_steps=8, eps_ex=one, eps_ey=one, eps_ez=one,
               bloch_k=(k, 0.0, 0.0))
    o = Oracle64(tables.build(sc, 8))
    km = 2 * np.pi * (m + k) / (nx * dx)
    o.E[2, :, 0, 0] = np.exp(1j * km * g.x.edges[:-1])
    e = [o.E[2, 3, 0, 0]]
    for _ in range(3):
        o.step()
        e.append(o.E[2, 3, 0, 0])
    cos_wdt = (e[3] + e[1]) / (2 * e[2])
    s = cpml.C_0 * dt / dx * abs(np.sin(km * dx / 2))
    assert abs(cos_wdt - (1 - 2 * s * s)) < 1e-12


def _outside_inside(o, sc):
    src = sc.tfsf_sources[0]
    lo, hi, opn = src.box_lo, src.box_hi, src.open_axes
    E = np.abs(o.E).max(0)
    inner = [slice(None) if a in opn else slice(lo[a] + 1, hi[a] - 1) for a in range(3)]
    far = np.ones(sc.shape, bool)
    far[tuple(slice(None) if a in opn else slice(max(lo[a] - 2, 0), hi[a] + 3)
              for a in range(3))] = False
    return E[far].max(), E[tuple(inner)].max()


@pytest.mark.parametrize("jxy", [(1.0, 0.0), (0.0, 1.0), (0.6, 0.8)])
def test_tfsf_box_confines_the_incident_field(jxy):
    sc = scenes.tfsf_box_scene(num_steps=200, jxy=jxy)
    o = Oracle64(tables.build(sc)).run(200)
    out, inside = _outside_inside(o, sc)
    assert inside > 1.0 and out < 1e-12 * inside


def test_tfsf_open_axes_and_reverse_direction():
    sc = scenes.tfsf_box_scene(num_steps=200, periodic_xy=True, direction=-1)
    o = Oracle64(tables.build(sc)).run(200)
    out, inside = _outside_inside(o, sc)
    assert inside > 1.0 and out < 1e-12 * inside


def test_oblique_tfsf_on_bloch_cell_confines_the_incident_field():
    sc = scenes.bloch_cell_scene(nxy=16, nz=80, n_pml=8, num_steps=400, with_metal=False, nf=3)
    o = Oracle64(tables.build(sc)).run(400)
    out, inside = _outside_inside(o, sc)
    assert inside > 1.0 and out < 1e-11 * inside


def test_tfsf_scatterer_scatters():
    """Sanity: with a dielectric block inside the box the scattered-field region is not empty."""
    sc = scenes.tfsf_box_scene(num_steps=200, scatterer=True)
    o = Oracle64(tables.build(sc)).run(200)
    out, inside = _outside_inside(o, sc)
    assert out > 1e-3 * inside
