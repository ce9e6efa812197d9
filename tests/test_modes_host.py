"""CPU tests of the Yee-grid mode solver and of the mode source in the float64 oracle."""

import numpy as np
import pytest

from openem_amd import cpml, flux, modes, scenes, tables
from openem_amd.grid import Grid, uniform_axis
from openem_amd.model import Scene
from oracle.yee64 import Oracle64

C0 = modes.C0


def test_homogeneous_mode_is_the_discrete_plane_wave():
    """Periodic homogeneous cross-section: beta~ = k0~ sqrt(eps) exactly, so
    n_eff = 2/(dl k0) asin(k0~ sqrt(eps) dl / 2) with k0~ = (2/dt) sin(w dt / 2) / c."""
    dl, eps = 30e-9, 3.7
    g = Grid(uniform_axis(6, dl, 0, "PML"), uniform_axis(5, dl, 0, "Periodic"),
             uniform_axis(7, dl, 0, "Periodic"))
    e = np.full(g.shape, eps)
    sc = Scene(grid=g, dt=cpml.courant_dt((dl, dl, dl)), num_time_steps=10,
               eps_ex=e, eps_ey=e, eps_ez=e)
    f = 1.9e14
    m = modes.solve_mode(sc, 0, 3, f)
    w = 2 * np.pi * f
    k0t = 2.0 / sc.dt * np.sin(0.5 * w * sc.dt) / C0
    want = 2.0 / dl * np.arcsin(0.5 * k0t * np.sqrt(eps) * dl) / (w / C0)
    assert abs(m.n_eff - want) < 1e-10
    assert abs(modes.mode_power(m, g) - 1.0) < 1e-12
    # plane wave: transverse impedance E_u / H_v = Z0 / n along the discrete relation
    ratio = np.abs(m.e_u).max() / np.abs(m.h_v).max() if np.abs(m.e_u).max() > 0 else \
        np.abs(m.e_v).max() / np.abs(m.h_u).max()
    bt = 2.0 / dl * np.sin(0.5 * m.beta * dl)
    assert abs(ratio - 2.0 / sc.dt * np.sin(0.5 * w * sc.dt) * cpml.MU_0 / bt) < 1e-9 * ratio


def _slab_te_neff(d, e1, e2, lam):
    """Fundamental TE mode of a symmetric slab: kappa tan(kappa d/2) = gamma (bisection)."""
    k0 = 2 * np.pi / lam
    lo, hi = np.sqrt(e2) + 1e-12, np.sqrt(e1) - 1e-12

    def g(n):
        ka = k0 * np.sqrt(e1 - n * n)
        ga = k0 * np.sqrt(n * n - e2)
        return ka * np.tan(ka * d / 2) - ga
    # the fundamental branch: kappa d / 2 < pi / 2
    lo = max(lo, np.sqrt(max(e1 - (np.pi / (k0 * d)) ** 2, e2)) + 1e-9)
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if g(mid) > 0:
            lo = mid
        else:
            hi = mid
    return 0.5 * (lo + hi)


def test_slab_te0_matches_the_analytic_dispersion():
    """Fine grid (dl = d/24) with exact 1D averaging at the faces: the TE0 effective index of
    a 240 nm slab (eps 12.1 in 2.1) at 1.55 um agrees with the transcendental equation to the
    discretisation error (< 2e-3)."""
    dl, d = 10e-9, 240e-9
    n_v = 120
    g = Grid(uniform_axis(4, dl, 0, "PML"), uniform_axis(4, dl, 0, "Periodic"),
             uniform_axis(n_v, dl, -n_v * dl / 2, "PML"))
    ex, ey, ez = scenes.layer_eps_z(g, -d / 2, d / 2, 12.1, 2.1)
    sc = Scene(grid=g, dt=cpml.courant_dt((dl, dl, dl)), num_time_steps=10,
               eps_ex=ex, eps_ey=ey, eps_ez=ez)
    lam = 1.55e-6
    m = modes.solve_mode(sc, 0, 2, C0 / lam)
    assert np.abs(m.e_u).max() > 10 * np.abs(m.e_v).max()        # TE: E along the slab (y)
    want = _slab_te_neff(d, 12.1, 2.1, lam)
    assert abs(m.n_eff - want) < 2e-3 * want, (m.n_eff, want)


@pytest.mark.parametrize("direction", [+1, -1])
def test_mode_source_is_one_directional_in_the_oracle(direction):
    """Forward amplitude |a+| = |S|/2 (the real part of the complex drive puts half of it at
    +f0), backward leakage below 1e-3 of |S| at both monitors, and the downstream flux equals
    |a+|^2 of the unit-power mode."""
    n = 900
    sc = scenes.waveguide_scene(shape=(50, 40, 36), n_pml=8, width_cells=10, thick_cells=6,
                                fwidth_rel=0.2, num_steps=n, nf=1, direction=direction)
    t = tables.build(sc)
    o = Oracle64(t).run(n)
    ms = sc.mode_sources[0]
    steps = np.arange(n)
    f0 = t.monitors[0].freqs[0]
    S = abs(np.sum(ms.amp_e[1:n + 1] * np.exp(2j * np.pi * f0 * (steps + 1) * sc.dt)))
    amp = {}
    for i, m in enumerate(t.monitors):
        if m.modes:
            ap, am = modes.amplitudes(o.acc[i], m.modes[0], sc.grid, modes.orientation(m.axis))
            amp[m.name] = (abs(ap[0]) / S, abs(am[0]) / S)
        else:
            amp[m.name] = float(flux.monitor_flux(o.acc[i], sc.grid, sc.flux_monitors[i])[0]) / S ** 2
    fwd, back = (0, 1) if direction > 0 else (1, 0)
    assert abs(amp["fwd"][fwd] - 0.5) < 5e-3, amp                # Re(profile amp): half at +f0
    # tidy3d-convention normalisation: the unit-power mode source launches unit amplitude
    from openem_amd import normalize
    m0 = next(i for i, m in enumerate(t.monitors) if m.name == "fwd")
    a_fwd = modes.amplitudes(o.acc[m0], t.monitors[m0].modes[0], sc.grid,
                             modes.orientation(t.monitors[m0].axis))[fwd][0]
    assert abs(abs(a_fwd / normalize.field_norm(sc, [f0])[0]) - 1.0) < 1e-2
    assert amp["fwd"][back] < 1e-3 and amp["back"][back] < 1e-3, amp
    # the plane's whole flux also counts the little power radiated where the mode's evanescent
    # tail meets the transverse PML, which the plain-Yee mode operator does not model
    assert abs(direction * amp["fwd_flux"] / amp["fwd"][fwd] ** 2 - 1.0) < 2e-2, amp


def test_mode_point_tables_equal_the_oracle_plane_corrections():
    """The device applies each point as F += scale * coef * Re(phase * amp[n]) with scale =
    dt/mu0 on the H side and the cell's cb on the E side (k_points); applying the tables that
    way reproduces the oracle's plane restatement of inject_mode_h / inject_mode_e."""
    sc = scenes.waveguide_scene(shape=(30, 28, 26), n_pml=6, width_cells=8, thick_cells=4,
                                num_steps=300, nf=1)
    t = tables.build(sc)
    o = Oracle64(t)
    mt = t.modes[0]
    for n in (0, 137):
        for magnetic in (True, False):
            F = o.H if magnetic else o.E
            F[:] = 0.0
            o._mode(sc.mode_sources[0], n, magnetic)
            want = F.copy()
            got = np.zeros_like(F)
            comp, cell, coef, ph, amp = ((mt.h_comp, mt.h_cell, mt.h_coef, mt.h_phase, mt.amp_e)
                                         if magnetic else
                                         (mt.e_comp, mt.e_cell, mt.e_coef, mt.e_phase, mt.amp_h))
            for c in range(3):
                sel = comp == c
                scale = t.ch if magnetic else t.cb[c].reshape(-1)[cell[sel]]
                np.add.at(got[c].reshape(-1), cell[sel],
                          scale * coef[sel] * np.real(ph[sel] * amp[n]))
            assert np.abs(want).max() > 0
            assert np.allclose(got, want, rtol=1e-12, atol=1e-12 * np.abs(want).max())


def test_plane_weights_map_every_plane_monitor_to_its_own_spec():
    """Mode monitors are planes too and follow the flux monitors in the table list: each plane
    monitor's face weights come from its own scene monitor (a flux monitor's transverse window;
    a mode monitor's whole plane), never from ``sc.flux_monitors`` by table index."""
    import dataclasses
    from openem_amd.solver import plane_weights
    sc = scenes.waveguide_scene(shape=(30, 20, 18), n_pml=5, width_cells=6, thick_cells=4,
                                num_steps=20, nf=1)
    win = ((2, 15), (3, 12))
    fl = [dataclasses.replace(sc.flux_monitors[0], transverse=win), sc.flux_monitors[1]]
    sc = dataclasses.replace(sc, flux_monitors=fl)
    t = tables.build(sc)
    kinds = [(m.name, m.kind, m.modes is not None) for m in t.monitors]
    assert kinds == [("fwd_flux", "plane", False), ("back_flux", "plane", False),
                     ("fwd", "plane", True), ("back", "plane", True)]
    full = flux.area_weights(sc.grid, 0)
    for i, m in enumerate(t.monitors):
        assert m.spec.name == m.name
        w1, w2 = plane_weights(t, i)
        if i == 0:
            mask = np.zeros(full[0].shape)
            mask[2:16, 3:13] = 1.0
            assert np.array_equal(w1, full[0] * mask) and np.array_equal(w2, full[1] * mask)
            u1, _ = plane_weights(t, i, transverse=False)
            assert np.array_equal(u1, full[0])
        else:
            assert np.array_equal(w1, full[0]) and np.array_equal(w2, full[1])
