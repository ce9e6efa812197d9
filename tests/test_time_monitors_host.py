"""CPU tests of the time-domain field monitor in the float64 oracle."""

from dataclasses import replace

import numpy as np

from openem_amd import scenes, tables
from openem_amd.model import FieldMonitor, FieldTimeMonitor
from oracle.yee64 import Oracle64


def _scene(n, tmons, fmons=()):
    sc = scenes.mini_scene(shape=(10, 12, 40), num_steps=n)
    return replace(sc, field_time_monitors=list(tmons),
                   field_monitors=list(sc.field_monitors) + list(fmons))


def test_samples_follow_the_definition():
    """Against the definition recorded step by step: slot s holds E^m and (H^{m-1/2} +
    H^{m+1/2}) / 2 at m = begin + s * interval (E^0 = H^{-1/2} = 0)."""
    n = 37
    tm = FieldTimeMonitor("t", (0, 2, 3, 4), (2, 3, 5), (4, 5, 20), 3, 31, 4, 6)
    t = tables.build(_scene(n, [tm]))
    o = Oracle64(t)
    Es, Hs = [o.E.copy()], [np.zeros_like(o.H)]       # E^m and H^{m-1/2} for m = 0..n
    for _ in range(n):
        o.step()
        Es.append(o.E.copy())
        Hs.append(o.H.copy())
    (i0, j0, k0), (ni, nj, nk) = tm.origin, tm.box
    sl = (slice(i0, i0 + ni), slice(j0, j0 + nj), slice(k0, k0 + nk))
    filled = tm.filled(n)
    assert filled == 6
    for s in range(filled):
        m = tm.step_begin + s * tm.interval
        for q, c in enumerate(tm.comps):
            want = Es[m][c][sl] if c < 3 else 0.5 * (Hs[m][c - 3][sl] + Hs[m + 1][c - 3][sl])
            assert np.allclose(o.tsamp[0][q, s], want, rtol=0, atol=1e-14 * np.abs(want).max() + 1e-300)
    assert np.abs(o.tsamp[0][:, 0]).max() > 0


def test_dft_of_samples_equals_the_box_monitor_phasors():
    """Every step sampled: the plain DFT of the E samples is the box monitor's E phasor (E at
    (n+1) dt), and the DFT of the averaged H samples is cos(w dt / 2) times the H phasor (H at
    (n+1/2) dt) minus the half-step end term -H^{N-1/2} e^{i w N dt} / 2, exactly."""
    n = 160
    box = ((2, 3, 5), (3, 4, 10))
    tm = FieldTimeMonitor("t", tuple(range(6)), box[0], box[1], 0, n + 1, 1, n + 1)
    fm = FieldMonitor("b", np.array([1.2e14, 1.6e14, 2.0e14]), box[0], box[1])
    t = tables.build(_scene(n, [tm], [fm]))
    o = Oracle64(t).run(n)
    mi = [i for i, m in enumerate(t.monitors) if m.name == "b"][0]
    acc = o.acc[mi]                                   # (6, nf, ni, nj, nk)
    w = 2 * np.pi * t.monitors[mi].freqs
    m = np.arange(n + 1)
    ph = np.exp(1j * w[:, None] * m[None, :] * t.dt)  # (nf, n+1)
    (i0, j0, k0), (ni, nj, nk) = box
    sl = (slice(i0, i0 + ni), slice(j0, j0 + nj), slice(k0, k0 + nk))
    S = o.tsamp[0]                                    # (6, n+1, ni, nj, nk)
    for c in range(6):
        dft = np.tensordot(ph[:, :n], S[c, :n], axes=(1, 0))
        if c < 3:
            dft = dft + ph[:, n][:, None, None, None] * S[c, n]   # E^N, written by the last step
            want = acc[c]
        else:
            want = (np.cos(0.5 * w * t.dt)[:, None, None, None] * acc[c]
                    - 0.5 * ph[:, n][:, None, None, None] * o.H[c - 3][sl][None])
        scale = np.abs(want).max()
        assert scale > 0
        assert np.abs(dft - want).max() < 1e-12 * scale, (c, np.abs(dft - want).max() / scale)


def _ftm(name, axis, km, ndir, g, frac=0.0, beg=2, end=90, iv=3, slots=30):
    from openem_amd.grid import transverse_axes
    from openem_amd.model import FluxTimeMonitor
    t1, t2 = transverse_axes(axis)
    tb = tuple((float(g.axes[t].edges[0]), float(g.axes[t].edges[-1])) for t in (t1, t2))
    return FluxTimeMonitor(name, axis, km, ndir, beg, end, iv, slots, t_bounds=tb, frac=frac)


def test_flux_time_equals_twice_the_colocated_flux_of_the_time_planes():
    """On a uniform normal axis the instantaneous flux of slot m is 2 normal_dir times the
    colocated plane flux (the pinned frequency-domain readout) of the real planes (E^m,
    Hbar^m) with H colocated along the normal as the DFT plane does: the tangential colocation,
    trapezoid weights, the H time average and the signs of the two paths must agree."""
    from openem_amd import flux
    from openem_amd.model import FluxMonitor
    n = 96
    sc0 = scenes.mini_scene(shape=(10, 12, 40), num_steps=n)
    g = sc0.grid
    mons = [_ftm("z", 2, 22, +1, g), _ftm("y", 1, 5, -1, g, iv=4)]
    sc = replace(sc0, flux_time_monitors=mons)
    t = tables.build(sc)
    o = Oracle64(t)
    Es, Hs = [o.E.copy()], [np.zeros_like(o.H)]
    for _ in range(n):
        o.step()
        Es.append(o.E.copy())
        Hs.append(o.H.copy())
    for q, m in enumerate(mons):
        from openem_amd.grid import transverse_axes
        t1, t2 = transverse_axes(m.axis)
        fm = FluxMonitor("p", m.axis, m.plane_index, m.normal_dir, np.ones(1), t_bounds=m.t_bounds)
        assert m.filled(n) > 3
        for s in range(m.filled(n)):
            mm = m.step_begin + s * m.interval
            Hbar = 0.5 * (Hs[mm] + Hs[mm + 1])
            take = lambda F, k: np.take(F, k, axis=m.axis)
            km = m.plane_index
            data = np.stack([take(Es[mm][t1], km), take(Es[mm][t2], km),
                             0.5 * (take(Hbar[t1], km) + take(Hbar[t1], km - 1)),
                             0.5 * (take(Hbar[t2], km) + take(Hbar[t2], km - 1))])[:, None]
            want = 2.0 * m.normal_dir * flux.monitor_flux_colocated(data, g, fm)[0]
            got = o.ftsamp[q][s]
            assert abs(got - want) <= 1e-12 * max(abs(want), 1e-30) + 1e-300, (m.name, s, got, want)
        assert np.abs(o.ftsamp[q][:m.filled(n)]).max() > 0


def test_flux_time_between_planes_is_the_weighted_plane_sum():
    """frac between primal planes combines the two planes' fluxes linearly (reference: the flux
    scalar is interpolated, not the fields)."""
    n = 60
    sc0 = scenes.mini_scene(shape=(10, 12, 40), num_steps=n, nonuniform_z=True)
    g = sc0.grid
    mons = [_ftm("a", 2, 20, +1, g, frac=0.3), _ftm("lo", 2, 20, +1, g), _ftm("hi", 2, 21, +1, g)]
    o = Oracle64(tables.build(replace(sc0, flux_time_monitors=mons))).run(n)
    f = mons[0].filled(n)
    assert np.allclose(o.ftsamp[0][:f], 0.7 * o.ftsamp[1][:f] + 0.3 * o.ftsamp[2][:f],
                       rtol=1e-12, atol=1e-12 * np.abs(o.ftsamp[1][:f]).max())
