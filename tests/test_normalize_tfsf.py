"""TF/SF normalisation (reference ``openem/normalize.py:123-135``): the divisor is the spectrum
of the incident E on the box's injection face over sqrt(2 eta0), the |E| in V/um of a plane
wave of unit intensity (1 W/um^2).

Pinned on the float64 oracle with the closed forms of a discrete plane wave in an empty,
x/y-periodic TF/SF box: the normalised incident field is sqrt(2 eta0) at the injection face
and carries the Yee grid's own wavenumber k_d (sin(k_d dl / 2) / dl = sin(w dt / 2) / (c dt))
from there, so at a plane m cells into the box it is sqrt(2 eta0) exp(i k_d m dl); the
normalised flux through a plane of area A um^2 is A cos(k dl / 2), the discrete Poynting flux
of a unit-intensity wave.  A dropped sqrt(2), a one-step shift of the DFT time base (a phase
error of w dt, 0.07 rad here) or a flux divisor off by the um^2 conversion fails these.
"""

import numpy as np
import pytest

from openem_amd import cpml, flux, normalize, scenes, tables
from oracle.yee64 import Oracle64

N_STEPS = 700
DL = 20e-9


def _run(direction):
    sc = scenes.tfsf_box_scene(shape=(4, 4, 44), num_steps=N_STEPS, periodic_xy=True,
                               direction=direction)
    tabs = tables.build(sc)
    return sc, Oracle64(tabs).run(N_STEPS)


def _k_yee(sc, f):
    w = 2.0 * np.pi * np.asarray(f)
    return 2.0 / DL * np.arcsin(DL / (cpml.C_0 * sc.dt) * np.sin(w * sc.dt / 2.0))


@pytest.mark.parametrize("direction", [+1, -1])
def test_normalised_incident_field_is_unit_intensity_at_phase_zero(direction):
    sc, o = _run(direction)
    src = sc.tfsf_sources[0]
    m = sc.flux_monitors[0]                     # "zin", inside the total-field region
    face = src.box_lo[2] if direction > 0 else src.box_hi[2]
    depth = abs(m.plane_index - face)
    assert 5 <= depth
    en = o.acc[0][0] / normalize.field_norm(sc, m.freqs)[:, None, None]   # Ex on the plane
    want = normalize.SQRT_2ETA0 * np.exp(1j * _k_yee(sc, m.freqs) * depth * DL)
    err = np.abs(en - want[:, None, None]).max() / normalize.SQRT_2ETA0
    print(f"direction {direction}: normalised Ex {depth} cells into the box, error {err:.2e}")
    assert err < 1e-6


def test_normalised_flux_is_area_times_discrete_factor():
    sc, o = _run(+1)
    m = sc.flux_monitors[0]
    p = flux.monitor_flux(o.acc[0], sc.grid, m) * m.normal_dir
    area_um2 = normalize.monitor_area(sc, m) * 1e12
    got = p / normalize.flux_norm(sc, m.freqs)
    k = 2.0 * np.pi * np.asarray(m.freqs) / cpml.C_0
    want = area_um2 * np.cos(k * DL / 2.0)
    print("normalised flux / area:", got / area_um2)
    assert np.allclose(got, want, rtol=1e-4)
    # the normalised_flux readout takes the same divisor for a TF/SF scene
    from openem_amd.results import Phasors, Result
    res = Result(phasors={m.name: Phasors(data=o.acc[0], freqs=m.freqs)})
    assert np.allclose(normalize.normalized_flux(sc, res, m.name), got, rtol=1e-12)


def test_source_table_is_the_face_series():
    sc = scenes.tfsf_box_scene(shape=(4, 4, 44), num_steps=N_STEPS, periodic_xy=True)
    s = normalize.source_table(sc)
    assert s.shape == (N_STEPS + 1,)
    d = normalize.norm_spectrum(sc, sc.flux_monitors[0].freqs)
    # normal incidence: E along the driving current only, so the spectrum is the divisor's
    assert np.allclose(d / normalize.SQRT_2ETA0,
                       normalize.field_norm(sc, sc.flux_monitors[0].freqs), rtol=1e-12)
