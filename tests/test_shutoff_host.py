"""CPU tests of the shutoff policy and of the oracle's shutoff sums (no GPU)."""

import numpy as np

from openem_amd import flux, scenes, shutoff, tables, waveform
from oracle.yee64 import energy_sums, plane_flux_terms


def test_decay_time_fit_recovers_an_exponential():
    r = shutoff.RatioToPeak(threshold=1e-9)
    tau = 1234.5
    for s in range(0, 8000, 400):
        r.update(s, 7.0 * np.exp(-s / tau))
    assert abs(r.tau - tau) < 1e-6 * tau
    g = shutoff.RatioToPeak(threshold=1e-9)
    for s in range(0, 8000, 400):
        g.update(s, 1.0 + s)                       # growing: no decay time
    assert g.tau == float("inf")


def test_ratio_is_to_the_running_peak_and_fires_below_threshold():
    r = shutoff.RatioToPeak(threshold=0.1)
    vals = [1.0, 4.0, 2.0, 0.5, 0.3]
    fired = [r.update(i, v) for i, v in enumerate(vals)]
    assert fired == [False, False, False, False, True]      # 0.3 / 4 < 0.1
    assert r.peak == 4.0 and abs(r.ratio - 0.075) < 1e-15
    assert r.deepest == (4, 0.075)


def test_flux_extrapolation_tracks_the_true_remaining_change():
    """flux(s) = F (1 - exp(-s/tau)): per-window change x tau / window approximates the
    relative change still to come, F exp(-s/tau) / flux(s), to O(window/tau)."""
    tau, w, F = 5000.0, 400, np.array([2.0, -3.0])
    pc = shutoff.FluxConvergence(tol=1e-12, patience=1, window_steps=w)
    for s in range(w, 20 * w + 1, w):
        pc.update(s, F * (1.0 - np.exp(-s / tau)), tau=tau)
        true = np.exp(-s / tau) / (1.0 - np.exp(-s / tau))
        if s > w:
            assert abs(pc.rel / true - 1.0) < 0.1, (s, pc.rel, true)


def test_policy_needs_patience_and_names_the_reason():
    pol = shutoff.Policy(phasor=shutoff.FluxConvergence(tol=1e-2, patience=2, window_steps=1))
    seq = [np.array([1.0]), np.array([1.001]), np.array([1.0015]), np.array([1.0016])]
    out = [pol.check(i, 1.0, 1.0, f) for i, f in enumerate(seq)]
    assert out == [False, False, True, True]
    assert pol.reason.startswith("phasor_converged")
    assert not shutoff.Policy(phasor=shutoff.FluxConvergence()).check(0, 1.0, 1.0, np.zeros(0))


def test_source_end_matches_the_gaussian_envelope():
    p = waveform.GaussianPulse(freq0=2e14, fwidth=1e14)
    dt = 2e-17
    wf = waveform.sample(p, dt, 20000)

    class S:                                               # a scene with one dipole
        sources, tfsf_sources = [], []
        dipoles = [type("D", (), {"waveform": wf})()]
    end = shutoff.source_end_step(S)
    tw = p.twidth
    t_edge = p.offset * tw + tw * np.sqrt(2.0 * np.log(1e3))   # envelope = 1e-3 of its peak
    period = 1.0 / p.freq0
    assert t_edge - 1.5 * period <= end * dt <= t_edge + 2 * dt


def test_energy_sums_closed_form_and_brute_force():
    sc = scenes.mini_scene(shape=(6, 7, 30), nonuniform_z=True, n_pml=4)
    E = np.ones((3,) + sc.shape)
    H = 2.0 * np.ones((3,) + sc.shape)
    out = energy_sums(sc, E, H)
    ax = sc.grid.axes
    # sum over cells of dl_x * dd_y * dd_z (and cyclic) for uniform unit fields
    want = sum(np.prod([(a.dl if c == q else a.dl_dual).sum() for q, a in enumerate(ax)])
               for c in range(3))
    assert abs(out[0] / want - 1.0) < 1e-6                   # float32 volumes
    assert abs(out[2] / (4.0 * want) - 1.0) < 1e-6
    rng = np.random.default_rng(0)
    E = rng.standard_normal((3,) + sc.shape)
    H = rng.standard_normal((3,) + sc.shape)
    out = energy_sums(sc, E, H)
    vols = sc.grid.yee_dual_volumes()
    eps = (sc.eps_ex, sc.eps_ey, sc.eps_ez)
    bf = np.zeros(3)
    nx, ny, nz = sc.shape
    for c in range(3):
        for i in range(nx):
            for j in range(ny):
                for k in range(nz):
                    v = float(np.float32(vols[c][i, j, k]))
                    ve = float(np.float32(v) * np.float32(eps[c][i, j, k]))
                    bf[0] += v * E[c, i, j, k] ** 2
                    bf[1] += ve * E[c, i, j, k] ** 2
                    bf[2] += v * H[c, i, j, k] ** 2
    assert np.allclose(out, bf, rtol=1e-12, atol=0)


def test_flux_terms_match_the_plane_flux():
    sc = scenes.slab_scene(num_steps=10)
    rng = np.random.default_rng(1)
    n1, n2 = sc.shape[0], sc.shape[1]
    data = rng.standard_normal((4, 3, n1, n2)) + 1j * rng.standard_normal((4, 3, n1, n2))
    w1, w2 = flux.area_weights(sc.grid, 2)
    t = plane_flux_terms(data, w1, w2)
    assert np.allclose(0.5 * (t[:, 0] - t[:, 1]), flux.plane_flux(data, sc.grid, axis=2),
                       rtol=1e-13, atol=0)
