"""Pinning the oracle's point-dipole injection (CPU only).

A sheet of point dipoles filling one plane of a 2 x 2 periodic line radiates a discrete plane
wave each way, and the semi-discrete Yee equations fix that wave in closed form.  With the
phasors the monitors accumulate (E^m at t = m dt, H^{m+1/2} at (m+1/2) dt, z = exp(i w dt)) and
w~ = (2/dt) sin(w dt/2), the 1D update of (Ex at z edges k, Hy at z centres k+1/2) with an
electric current density J^{n+1/2} at edge k0, or a magnetic current density M^n on the Hy line
at centre k0 + 1/2, reads

    i w~ mu Hy[k]  = (Ex[k+1] - Ex[k]) / dz + M delta(k, k0)
    i w~ eps Ex[k] = (Hy[k] - Hy[k-1]) / dz + J delta(k, k0)        (Maxwell: eps dE/dt = curl H - J)

whose outgoing solution is the 1D Yee Green's function

    Ex[k] = -J w~ mu dz^2 / (2 sin(kappa dz)) exp(i kappa |k - k0| dz),
    Hy[k] = -M w~ eps dz^2 / (2 sin(kappa dz)) exp(i kappa |k - k0| dz),

kappa from the Yee dispersion relation sin(kappa dz/2) = sqrt(eps_r) dz/(c dt) sin(w dt/2), and
J, M the transforms of the injected samples at their own times: J = coef sum_n amp_half[n]
z^(n+1/2) (the electric current sampled at (n+1/2) dt, between E^n and E^{n+1}), M = coef sum_n
amp_int[n] z^n (the magnetic current at n dt, between H^{n-1/2} and H^{n+1/2}).  The same
equations give the discrete Poynting balance: the power the sheet delivers, -1/2 Re sum F J* dV,
equals the flux through two planes on either side (H colocated on the E plane).

A sign flip, cb in place of ch (or the reverse), an amplitude sample shifted by one step or
by half a step, or a dropped coefficient changes the radiated phasor by at least 3 %; each is
planted in a copy of the oracle below and must fail both checks.
"""

import numpy as np
import pytest

from openem_amd import cpml, flux, scenes, tables, waveform
from openem_amd.grid import Grid, uniform_axis
from openem_amd.model import FieldMonitor, FluxMonitor, PointDipole, Scene
from oracle.yee64 import Oracle64

NZ, DL, K0, EPS_R = 120, 20e-9, 60, 2.0
F0 = 3e14
FREQS = F0 * np.array([0.8, 1.0, 1.2])
K_UP, K_DN = K0 + 25, K0 - 25
N_STEPS = 1600
# Residuals of the correct oracle (measured 2.2e-5 and 2e-9): the CPML's reflection and the
# DFT window's truncation; the smallest planted error moves them by 4.6e-2 and 1.1e-3.
TOL_G = 2e-4      # Green's function, relative to the radiated amplitude
TOL_P = 1e-7      # power balance, relative to the delivered power


def _sheet_scene(magnetic: bool) -> Scene:
    g = Grid(uniform_axis(2, DL, 0, "Periodic"), uniform_axis(2, DL * 1.25, 0, "Periodic"),
             uniform_axis(NZ, DL, 0, "PML"))
    dt = cpml.courant_dt((DL, DL * 1.25, DL))
    e = np.full(g.shape, EPS_R)
    wf = waveform.sample(waveform.GaussianPulse(freq0=F0, fwidth=0.4 * F0), dt, N_STEPS)
    # Ex dipoles on the edge plane k0, or Hy dipoles on the centre line k0 + 1/2
    coef = 3.0e8 if magnetic else 1.0e6
    dips = [PointDipole(0 if not magnetic else 1, np.array([[i, j, K0]]), np.array([coef]), wf,
                        magnetic=magnetic)
            for i in range(2) for j in range(2)]
    fm = FieldMonitor("line", FREQS, (0, 0, 20), (2, 2, 80))
    planes = [FluxMonitor("up", 2, K_UP, +1, FREQS), FluxMonitor("down", 2, K_DN, +1, FREQS)]
    return Scene(grid=g, dt=dt, num_time_steps=N_STEPS, eps_ex=e, eps_ey=e, eps_ez=e,
                 pml=scenes.pml_faces(dt, (DL, DL * 1.25, DL), [(2, "lo"), (2, "hi")]),
                 dipoles=dips, field_monitors=[fm], flux_monitors=planes)


def _kappa_wt(sc):
    dt = sc.dt
    wt = 2.0 / dt * np.sin(np.pi * FREQS * dt)
    s = np.sqrt(EPS_R) * DL / (cpml.C_0 * dt) * np.sin(np.pi * FREQS * dt)
    return 2.0 / DL * np.arcsin(s), wt


def _source_transform(sc, magnetic: bool):
    """Transform of the current density the sheet carries, at its own sample times."""
    d = sc.dipoles[0]
    n = np.arange(N_STEPS)
    if magnetic:
        samples, t = d.waveform.amp_int[n], n * sc.dt
    else:
        samples, t = d.waveform.amp_half[n], (n + 0.5) * sc.dt
    ph = np.exp(2j * np.pi * FREQS[:, None] * t[None, :])
    return float(d.coef[0]) * (samples[None, :] * ph).sum(axis=1)


def _check(sc, o, magnetic: bool):
    """Relative errors of (Green's function over the monitored line, Poynting balance)."""
    kappa, wt = _kappa_wt(sc)
    S = _source_transform(sc, magnetic)
    box = o.acc[2]             # (6, nf, 2, 2, 80), k = 20 + index (flux planes come first)
    ks = np.arange(20, 100)
    if magnetic:
        got = box[4]                             # Hy, centre k + 1/2
        amp = -S * wt * (cpml.EPSILON_0 * EPS_R) * DL ** 2 / (2.0 * np.sin(kappa * DL))
    else:
        got = box[0]                             # Ex, edge k
        amp = -S * wt * cpml.MU_0 * DL ** 2 / (2.0 * np.sin(kappa * DL))
    want = amp[:, None] * np.exp(1j * kappa[:, None] * DL * np.abs(ks - K0)[None, :])
    # uniform over the 2 x 2 cross-section
    err_g = float(np.abs(got - want[:, None, None, :]).max() / np.abs(want).max())
    # Poynting balance: delivered power of the four sheet points against the outgoing flux
    if magnetic:
        dV = sc.grid.yee_h_dual_volumes()[1][:, :, K0]
        F = box[4][:, :, :, K0 - 20]
    else:
        dV = sc.grid.yee_dual_volumes()[0][:, :, K0]
        F = box[0][:, :, :, K0 - 20]
    delivered = -0.5 * np.real((F * np.conj(S)[:, None, None] * dV[None]).sum(axis=(1, 2)))
    up = flux.plane_flux(o.acc[0], sc.grid, axis=2)
    down = flux.plane_flux(o.acc[1], sc.grid, axis=2)
    err_p = float(np.abs(delivered - (up - down)).max() / np.abs(delivered).max())
    return err_g, err_p, delivered, up, down


@pytest.fixture(scope="module", params=[False, True], ids=["electric", "magnetic"])
def sheet(request):
    magnetic = request.param
    sc = _sheet_scene(magnetic)
    t = tables.build(sc)
    return magnetic, sc, t, Oracle64(t).run(N_STEPS)


def test_sheet_radiates_the_yee_greens_function(sheet):
    magnetic, sc, _, o = sheet
    err_g, _, _, up, down = _check(sc, o, magnetic)
    print(f"{'magnetic' if magnetic else 'electric'} sheet: Green's function error {err_g:.2e}")
    assert err_g < TOL_G
    # symmetric: the same power leaves on both sides
    assert np.abs(up + down).max() < TOL_G * np.abs(up).max()


def test_sheet_power_balance(sheet):
    magnetic, sc, _, o = sheet
    _, err_p, delivered, _, _ = _check(sc, o, magnetic)
    print(f"{'magnetic' if magnetic else 'electric'} sheet: power balance error {err_p:.2e}")
    assert np.all(delivered > 0)
    assert err_p < TOL_P


class _Planted(Oracle64):
    """The oracle with one mistake in its dipole injection.  ``other`` holds, per dipole table,
    the amplitudes sampled at the other half step (whole steps for electric currents, half
    steps for magnetic ones)."""

    mistake = ""
    other: dict = {}

    def _dipoles(self, n, magnetic):
        t = self.t
        F = self.H if magnetic else self.E
        for dp in t.dipoles:
            if bool(dp.magnetic) != magnetic:
                continue
            amp = self.other[bool(dp.magnetic)] if self.mistake == "other_half_step" else dp.amp
            for c in range(3):
                sel = dp.comp == c
                cells = dp.cell[sel].astype(np.int64)
                cb = t.cb[c].reshape(-1)[cells]
                scale = t.ch if magnetic else cb
                m = n
                coef = dp.coef[sel]
                if self.mistake == "sign":
                    scale = -scale
                elif self.mistake == "swap_ch_cb":
                    scale = cb if magnetic else t.ch
                elif self.mistake == "next_sample":
                    m = min(n + 1, amp.shape[1] - 1)
                elif self.mistake == "previous_sample":
                    m = max(n - 1, 0)
                elif self.mistake == "no_coef":
                    coef = np.ones_like(coef)
                v = scale * coef * amp[dp.src_of[sel], m]
                np.subtract.at(F[c].reshape(-1), cells, v)


@pytest.mark.parametrize("mistake", ["sign", "swap_ch_cb", "next_sample", "previous_sample",
                                     "no_coef", "other_half_step"])
def test_planted_dipole_mistakes_fail_the_checks(sheet, mistake):
    magnetic, sc, t, _ = sheet
    src = [d for d in sc.dipoles if bool(d.magnetic) == magnetic]
    other = {magnetic: np.stack([(d.waveform.amp_half if magnetic else d.waveform.amp_int)
                                 [:N_STEPS + 1] for d in src])}
    o = type("P", (_Planted,), {"mistake": mistake, "other": other})(t).run(N_STEPS)
    err_g, err_p, _, _, _ = _check(sc, o, magnetic)
    print(f"{mistake}: Green's function error {err_g:.2e}, power balance error {err_p:.2e}")
    assert err_g > 50 * TOL_G and err_p > 100 * TOL_P
