"""Normalisation of raw phasors and fluxes to the tidy3d convention (reference
``openem/normalize.py``), for the readout only: the stepper itself is untouched.

The divisor is the plain DFT of the table the normalisation source injects (the monitor DFT's
convention, t = n dt, no dt factor), converted V/m -> V/um (x 1e6) where the source is a field
source:

* plane wave: field divisor ``D 1e6 sqrt(n A / (2 eta0)) sqrt(cos(k dz_h / 2))``, flux divisor
  ``|D|^2 n A / (2 eta0) cos(k dz_h / 2)`` (unit power through the source cross-section, in
  the discrete Poynting flux a Yee grid conserves);
* mode source: ``D`` of ``amp_e / 2`` (unit-power modes, no unit conversion);
* point dipole: ``D 1e6`` of the whole-step amplitude table;
* TF/SF box (takes priority, as in the reference): ``D / sqrt(2 eta0)`` with ``D`` the spectrum
  of the incident E on the box's injection face.  tidy3d normalises a TF/SF source to unit
  intensity, 1 W/um^2, whose plane wave has ``|E| = sqrt(2 eta0)`` V/um; the reference divides
  by the same ``sqrt(2 eta0)`` (fitted there against tidy3d runs) and by a 3.5-step phase that
  aligns its injection table with its monitor DFT (``openem/normalize.py:123-135``).  Here ``D``
  is taken from the line that drives the box, on the injection face itself, so no time shift
  is needed: the normalised incident field on that face is ``sqrt(2 eta0)`` at phase 0.
"""

from __future__ import annotations

import numpy as np

from . import flux as flux_mod
from . import tfsf as tfsf_mod
from .grid import transverse_axes

UNIT_E = 1.0e6                  # V/m -> V/um
ETA0 = 376.730313668
C_0 = 299792458.0
#: |E| in V/um of a plane wave of unit intensity (1 W/um^2)
SQRT_2ETA0 = float(np.sqrt(2.0 * ETA0))


def background_index(sc) -> float:
    """sqrt of the median eps three cells upstream of the plane-wave source plane (reference
    ``openem/normalize.py:265-277``)."""
    src = sc.sources[0]
    k = int(src.plane_index) - 3 * int(src.direction)
    k = min(max(k, 0), sc.eps_ex.shape[src.axis] - 1)
    sl = [slice(None)] * 3
    sl[src.axis] = k
    return float(np.sqrt(np.median(np.asarray(sc.eps_ex)[tuple(sl)])))


def source_table(sc) -> np.ndarray:
    """Injection table of the normalisation source (reference ``openem/normalize.py:92-108``):
    the first TF/SF box's incident E on its injection face along the driving current's
    direction, else the first plane wave's incident E, else the first mode source's
    ``amp_e / 2``, else the first dipole's whole-step amplitudes."""
    if sc.tfsf_sources:
        src = sc.tfsf_sources[0]
        e = tfsf_mod.incident_at_face(sc, src, int(sc.num_time_steps))
        p = np.asarray(src.jxy, dtype=np.float64)
        p = p / np.linalg.norm(p)
        return e[:, 0] * p[0] + e[:, 1] * p[1]
    if sc.sources:
        return np.asarray(sc.sources[0].ex_inc, dtype=np.float64)
    if sc.mode_sources:
        return np.asarray(sc.mode_sources[0].amp_e, dtype=np.complex128) / 2.0
    if sc.dipoles:
        return np.asarray(sc.dipoles[0].waveform.amp_int, dtype=np.float64)
    raise NotImplementedError("the scene has no normalisation source")


def norm_spectrum(sc, freqs) -> np.ndarray:
    """Plain DFT ``sum_n s[n] exp(2 pi i f n dt)`` of :func:`source_table` (reference
    ``openem/normalize.py:111-120``)."""
    f = np.atleast_1d(np.asarray(freqs, dtype=np.float64))
    s = source_table(sc)
    t = np.arange(s.size, dtype=np.float64) * sc.dt
    return np.exp(2j * np.pi * f[:, None] * t[None, :]) @ s


def half_cell_cos(sc, freqs) -> np.ndarray:
    """``cos(k dz_h / 2)``: discrete over continuous injected flux of the plane wave, dz_h the
    cell holding the H correction (reference ``openem/normalize.py:33-70``)."""
    src = sc.sources[0]
    dl = np.asarray(sc.grid.axes[int(src.axis)].dl, dtype=np.float64)
    kh = int(src.plane_index) - 1 if int(src.direction) > 0 else int(src.plane_index)
    dz = float(dl[min(max(kh, 0), dl.size - 1)])
    k = 2.0 * np.pi * np.atleast_1d(np.asarray(freqs, dtype=np.float64)) / C_0 * background_index(sc)
    return np.cos(0.5 * k * dz)


def snap_phase(sc, freqs) -> np.ndarray:
    """Phase between the user's source centre and the snapped injection plane (reference
    ``openem/normalize.py:73-89``); 1 when the source carries no ``center``."""
    src = sc.sources[0]
    f = np.atleast_1d(np.asarray(freqs, dtype=np.float64))
    zc = float(getattr(src, "center", float("nan")))
    if not np.isfinite(zc):
        return np.ones(f.shape, dtype=np.complex128)
    ze = float(sc.grid.axes[src.axis].edges[src.plane_index])
    k = 2.0 * np.pi * f / C_0 * background_index(sc)
    return np.exp(1j * float(src.direction) * k * (zc - ze))


def _source_area(sc) -> float:
    t1, t2 = transverse_axes(int(sc.sources[0].axis))
    return float(sc.grid.axes[t1].dl.sum() * sc.grid.axes[t2].dl.sum())


def tfsf_spectrum(sc, freqs) -> np.ndarray:
    """Spectrum of the TF/SF box's incident plane wave on its injection face: the plain DFT
    (t = n dt) of the incident E along the driving current, with the magnitude of the whole
    vector (an oblique wave also has E normal to the face)."""
    f = np.atleast_1d(np.asarray(freqs, dtype=np.float64))
    src = sc.tfsf_sources[0]
    e = tfsf_mod.incident_at_face(sc, src, int(sc.num_time_steps))
    t = np.arange(e.shape[0], dtype=np.float64) * sc.dt
    D = np.exp(2j * np.pi * f[:, None] * t[None, :]) @ e            # (nf, 3)
    p = np.asarray(src.jxy, dtype=np.float64)
    p = p / np.linalg.norm(p)
    dt_ = D[:, 0] * p[0] + D[:, 1] * p[1]
    mag = np.sqrt(np.sum(np.abs(D) ** 2, axis=1))
    return mag * np.exp(1j * np.angle(dt_))


def field_norm(sc, freqs) -> np.ndarray:
    """Divisor of field phasors per frequency (reference ``openem/normalize.py:127-194``)."""
    if sc.tfsf_sources:
        return tfsf_spectrum(sc, freqs) / SQRT_2ETA0
    D = norm_spectrum(sc, freqs)
    if sc.sources:
        return (D * UNIT_E * np.sqrt(background_index(sc) * _source_area(sc) / (2.0 * ETA0))
                * np.sqrt(half_cell_cos(sc, freqs)) * snap_phase(sc, freqs))
    if sc.mode_sources:
        return D
    return D * UNIT_E


def flux_norm(sc, freqs) -> np.ndarray:
    """Flux divisor ``|field_norm|^2 / 1e12`` (reference ``openem/normalize.py:232-234``)."""
    return np.abs(field_norm(sc, freqs)) ** 2 / 1.0e12


def monitor_area(sc, m) -> float:
    """Transverse area of a flux monitor including its transverse crop (reference
    ``openem/normalize.py:280-296``)."""
    t1, t2 = transverse_axes(m.axis)
    d1, d2 = sc.grid.axes[t1].dl, sc.grid.axes[t2].dl
    if m.transverse is not None:
        (a1, b1), (a2, b2) = m.transverse
        d1, d2 = d1[a1:b1], d2[a2:b2]
    return float(d1.sum() * d2.sum())


def plane_flux_divisor(sc, m, freqs) -> np.ndarray:
    """``|D|^2 n A / (2 eta0) cos(k dz_h / 2)`` for a plane-wave scene (reference
    ``openem/normalize.py:299-303``)."""
    D = norm_spectrum(sc, freqs)
    return (np.abs(D) ** 2 * background_index(sc) * monitor_area(sc, m) / (2.0 * ETA0)
            * half_cell_cos(sc, freqs))


def normalized_flux(sc, result, name: str) -> np.ndarray:
    """The flux monitor ``name`` of ``result`` in the tidy3d convention: outward along its
    ``normal_dir``, per unit source power (reference ``openem/nb/backend.py:963-982``)."""
    m = next(mm for mm in sc.flux_monitors if mm.name == name)
    vals = flux_mod.monitor_flux(result.phasors[name].data, sc.grid, m) * float(m.normal_dir)
    if sc.sources and not sc.tfsf_sources:
        return vals / plane_flux_divisor(sc, m, m.freqs)
    return vals / flux_norm(sc, m.freqs)


def box_flux(sc, result, name: str) -> np.ndarray:
    """Net outward flux of the :class:`~openem_amd.model.BoxFluxMonitor` ``name``: each face's
    flux times its ``normal_dir``, summed (reference ``openem/nb/backend.py:858-878``); raw
    units."""
    box = next(b for b in sc.box_flux_monitors if b.name == name)
    faces = {m.name: m for m in sc.flux_monitors}
    out = 0.0
    for fn in box.face_names:
        m = faces[fn]
        out = out + flux_mod.monitor_flux(result.phasors[fn].data, sc.grid, m) * float(m.normal_dir)
    return np.asarray(out)


def normalized_box_flux(sc, result, name: str) -> np.ndarray:
    """:func:`box_flux` in the tidy3d convention (divided by :func:`flux_norm`)."""
    box = next(b for b in sc.box_flux_monitors if b.name == name)
    freqs = next(m.freqs for m in sc.flux_monitors if m.name == box.face_names[0])
    return box_flux(sc, result, name) / flux_norm(sc, freqs)
