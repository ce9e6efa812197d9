"""``scene.npz``: a Scene on disk in the reference's layout (``openem/serialize.py``, format
version 10), numpy only.

The reference cuts its pipeline at this file: a machine with tidy3d turns a ``td.Simulation``
into a ``Scene`` and writes ``scene.npz``; the GPU job reads it back and solves.  :func:`load`
reads such a file into this package's :class:`~openem_amd.model.Scene` and :func:`save` writes
one in the same key layout, so scenes move between the two in both directions.

Covered one for one (key names and meanings of ``openem/serialize.py:120-420``): the grid, eps /
sigma / PEC arrays, CPML faces, plane-wave sources, mode sources, point dipoles, the pole-residue
dispersion CSR, flux, field (with its write-back stride), field-time, flux-time, permittivity,
box-flux and projection monitors, ``dt``, ``num_time_steps``, ``shutoff`` and
``bloch_k``.

Differences, each failing closed rather than silently:

* entities this stepper does not model (mixed-interface dispersion, time modulation, tensor
  media, absorber slabs, symmetry planes, broadband mode-source terms, 2D flat axes) raise
  ``NotImplementedError`` on load;
* TF/SF boxes: the reference stores its own 1D line (grid, eps, drive tables); this package
  builds its line from the scene (``openem_amd.tfsf``), so a loaded box keeps the geometry,
  direction, polarisation, open axes, amplitude scale and waveform, and oblique boxes are
  refused.  A box written here carries those fields plus ``oem_tfsf*`` keys of its own line
  and no reference line tables, so the reference cannot read a TF/SF scene written here;
* mode monitors: the reference names the planar field monitor its phasors come from
  (``modemon_planes``); a file written here stores the plane itself (``oem_modemon*``), and a
  reference file is mapped through the field monitor of that name (normal = its thinnest box
  axis, the plane one spare cell above the box origin).
"""

from __future__ import annotations

import os
import pathlib

import numpy as np

from . import cpml, waveform
from .grid import Axis, Grid
from .model import (BoxFluxMonitor, Dispersion, FieldMonitor, FieldTimeMonitor, FluxMonitor,
                    FluxTimeMonitor, ModeMonitor, ModeSource, PermittivityMonitor, PlaneWaveSource,
                    PointDipole, ProjectionMonitor, Scene, TFSFSource)

FORMAT_VERSION = 10

_PML_FIELDS = ("a_E", "b_E", "inv_kappa_E", "a_H", "b_H", "inv_kappa_H")
_MODE_FIELDS = ("ey_inc", "ez_inc", "hy_inc", "hz_inc", "amp_e", "amp_h")
_UNSUPPORTED_PREFIXES = {"dmix_": "mixed-interface dispersion", "mod_": "time modulation",
                         "tensor_": "tensor media"}
_UNSUPPORTED_LISTS: dict = {}


def _opt(z, key):
    """An optional array: missing, or stored as an object / non-numeric 0-d value -> None."""
    if key not in z.files:
        return None
    try:
        a = z[key]
    except ValueError:
        return None
    if a.dtype == object or (a.shape == () and a.dtype.kind not in "fc"):
        return None
    return a


def _count(z, key) -> int:
    return int(z[key]) if key in z.files else 0


def _names(z, key):
    return [str(n) for n in z[key]] if key in z.files else []


def _apod_pack(apod) -> np.ndarray:
    return np.array([np.nan if v is None else float(v) for v in apod])


def _apod_unpack(a):
    return tuple(None if np.isnan(v) else float(v) for v in a)


def _waveform(z, prefix: str, dt: float) -> waveform.Waveform:
    """Real tables plus the optional ``_im`` parts (missing: a purely real waveform)."""
    a_int = np.asarray(z[f"{prefix}_amp_int"], dtype=np.float64)
    a_half = np.asarray(z[f"{prefix}_amp_half"], dtype=np.float64)
    i_int = _opt(z, f"{prefix}_amp_int_im")
    i_half = _opt(z, f"{prefix}_amp_half_im")
    return waveform.Waveform(
        amp_int=a_int, amp_half=a_half,
        amp_int_complex=a_int + 1j * (0.0 if i_int is None else i_int),
        amp_half_complex=a_half + 1j * (0.0 if i_half is None else i_half), dt=dt)


# ------------------------------------------------------------------------------------------
# save
# ------------------------------------------------------------------------------------------

def _put_waveform(d: dict, prefix: str, w: waveform.Waveform) -> None:
    """Real tables under the reference's keys; the imaginary parts (the complex field path
    injects the whole complex waveform) under the ``_im`` keys the reference uses for dipoles
    (``openem/serialize.py:295-310``) and ignores elsewhere."""
    d[f"{prefix}_amp_int"] = w.amp_int
    d[f"{prefix}_amp_half"] = w.amp_half
    d[f"{prefix}_amp_int_im"] = np.imag(w.amp_int_complex)
    d[f"{prefix}_amp_half_im"] = np.imag(w.amp_half_complex)


def save(sc: Scene, path) -> pathlib.Path:
    """Write ``sc`` as one compressed npz (written to a temporary name and renamed, so a reader
    never sees half a file; reference ``openem/serialize.py:625-655``)."""
    d: dict[str, np.ndarray] = {
        "format_version": np.array(FORMAT_VERSION),
        "dt": np.array(float(sc.dt)),
        "num_time_steps": np.array(int(sc.num_time_steps)),
        "shutoff": np.array(float(sc.shutoff)),
        "symmetry": np.array([0, 0, 0]),
        "bloch_k": np.array(sc.bloch_k, dtype=np.float64),
        "eps_ex": np.asarray(sc.eps_ex), "eps_ey": np.asarray(sc.eps_ey),
        "eps_ez": np.asarray(sc.eps_ez),
    }
    if sc.sigma_ex is not None:
        for c in ("ex", "ey", "ez"):
            d[f"sigma_{c}"] = np.asarray(getattr(sc, f"sigma_{c}"))
    if sc.pec_ex is not None:
        for c in ("ex", "ey", "ez"):
            d[f"pec_{c}"] = np.asarray(getattr(sc, f"pec_{c}"))
    for name, ax in zip("xyz", sc.grid.axes):
        d[f"edges_{name}"] = np.asarray(ax.edges)
        d[f"bnd_{name}"] = np.array([ax.boundary_lo, ax.boundary_hi])
    d["pml_keys"] = np.array([f"{a}:{s}" for (a, s) in sorted(sc.pml)])
    for (a, s), c in sc.pml.items():
        for f in _PML_FIELDS:
            d[f"pml_{a}_{s}_{f}"] = np.asarray(getattr(c, f))
    d["n_sources"] = np.array(len(sc.sources))
    for i, src in enumerate(sc.sources):
        d[f"src{i}_meta"] = np.array([src.axis, src.plane_index, src.direction, src.pol_axis,
                                      src.num_freqs])
        d[f"src{i}_angle_theta"] = np.array(float(src.angle_theta))
        d[f"src{i}_ex_inc"] = np.asarray(src.ex_inc)
        d[f"src{i}_hy_inc"] = np.asarray(src.hy_inc)
        _put_waveform(d, f"src{i}", src.waveform)
        d[f"src{i}_center"] = np.array(float(getattr(src, "center", float("nan"))))
    if sc.dispersion is not None and sc.dispersion.n_entry > 0:
        dd = sc.dispersion
        for f in ("comp", "cell", "pole_ofs", "am1", "b", "g"):
            d[f"disp_{f}"] = np.asarray(getattr(dd, f))
    d["n_absorbers"] = np.array(0)
    d["n_mode_sources"] = np.array(len(sc.mode_sources))
    for i, m in enumerate(sc.mode_sources):
        d[f"msrc{i}_meta"] = np.array([m.plane_index, m.direction, m.axis])
        d[f"msrc{i}_n_eff"] = np.array([complex(m.n_eff).real, complex(m.n_eff).imag])
        for f in _MODE_FIELDS:
            d[f"msrc{i}_{f}"] = np.asarray(getattr(m, f))
    d["n_tfsf"] = np.array(len(sc.tfsf_sources))
    for i, t in enumerate(sc.tfsf_sources):
        d[f"tfsf{i}_meta"] = np.array([t.direction, *t.box_lo, *t.box_hi, -1, -1, 1, t.axis,
                                       sum(1 << q for q in t.open_axes)])
        d[f"tfsf{i}_pol"] = np.array(t.jxy, dtype=np.float64)
        _put_waveform(d, f"tfsf{i}", t.waveform)
        d[f"oem_tfsf{i}_kxy"] = np.array(t.kxy, dtype=np.float64)
        d[f"oem_tfsf{i}_amp_scale"] = np.array(float(t.amp_scale))
        if t.eps_line is not None:
            d[f"oem_tfsf{i}_eps_edge"] = np.asarray(t.eps_line[0], dtype=np.float64)
            d[f"oem_tfsf{i}_eps_ctr"] = np.asarray(t.eps_line[1], dtype=np.float64)
    d["mon_names"] = np.array([m.name for m in sc.flux_monitors])
    for i, m in enumerate(sc.flux_monitors):
        d[f"mon{i}_meta"] = np.array([m.axis, m.plane_index, m.normal_dir])
        d[f"mon{i}_freqs"] = np.asarray(m.freqs, dtype=np.float64)
        d[f"mon{i}_spectrum"] = np.ones(np.size(m.freqs), dtype=np.complex128)
        if m.transverse is not None:
            d[f"mon{i}_transverse"] = np.array(m.transverse)
        if m.apodization is not None:
            d[f"mon{i}_apod"] = _apod_pack(m.apodization)
        if m.t_bounds is not None:
            d[f"mon{i}_tbounds"] = np.array(m.t_bounds, dtype=np.float64)
    d["fmon_names"] = np.array([m.name for m in sc.field_monitors])
    for i, m in enumerate(sc.field_monitors):
        d[f"fmon{i}_meta"] = np.array([*m.origin, *m.box, *m.interval_space])
        d[f"fmon{i}_freqs"] = np.asarray(m.freqs, dtype=np.float64)
        if m.apodization is not None:
            d[f"fmon{i}_apod"] = _apod_pack(m.apodization)
    d["modemon_names"] = np.array([m.name for m in sc.mode_monitors])
    d["modemon_planes"] = np.array([m.name for m in sc.mode_monitors])
    for i, m in enumerate(sc.mode_monitors):
        d[f"oem_modemon{i}_meta"] = np.array([m.axis, m.plane_index, m.mode_index])
        d[f"oem_modemon{i}_freqs"] = np.asarray(m.freqs, dtype=np.float64)
        d[f"oem_modemon{i}_n_guess"] = np.array(np.nan if m.n_guess is None else float(m.n_guess))
    d["tmon_names"] = np.array([m.name for m in sc.field_time_monitors])
    for i, m in enumerate(sc.field_time_monitors):
        d[f"tmon{i}_meta"] = np.array([m.step_begin, m.step_end, m.interval, m.num_slots])
        d[f"tmon{i}_comps"] = np.array(m.comps)
        d[f"tmon{i}_origin"] = np.array(m.origin)
        d[f"tmon{i}_box"] = np.array(m.box)
    d["ftmon_names"] = np.array([m.name for m in sc.flux_time_monitors])
    for i, m in enumerate(sc.flux_time_monitors):
        d[f"ftmon{i}_meta"] = np.array([m.axis, m.plane_index, m.normal_dir, m.step_begin,
                                        m.step_end, m.interval, m.num_slots])
        d[f"ftmon{i}_frac"] = np.array(float(m.frac))
        d[f"ftmon{i}_tbounds"] = np.array(m.t_bounds, dtype=np.float64)
    d["pmon_names"] = np.array([m.name for m in sc.permittivity_monitors])
    for i, m in enumerate(sc.permittivity_monitors):
        d[f"pmon{i}_meta"] = np.array([*m.origin, *m.box])
        d[f"pmon{i}_freqs"] = np.asarray(m.freqs, dtype=np.float64)
    d["projmon_names"] = np.array([m.name for m in sc.projection_monitors])
    for i, m in enumerate(sc.projection_monitors):
        d[f"projmon{i}_surfaces"] = np.array(m.surface_names)
        d[f"projmon{i}_dirs"] = np.array(m.normal_dirs)
    d["bfmon_names"] = np.array([m.name for m in sc.box_flux_monitors])
    for i, m in enumerate(sc.box_flux_monitors):
        d[f"bfmon{i}_faces"] = np.array(m.face_names)
    d["n_dipoles"] = np.array(len(sc.dipoles))
    for i, dp in enumerate(sc.dipoles):
        d[f"dip{i}_component"] = np.array(int(dp.component))
        d[f"dip{i}_indices"] = np.asarray(dp.indices)
        d[f"dip{i}_coef"] = np.asarray(dp.coef)
        _put_waveform(d, f"dip{i}", dp.waveform)
        d[f"dip{i}_magnetic"] = np.array(int(dp.magnetic))
    p = pathlib.Path(path)
    p.parent.mkdir(parents=True, exist_ok=True)
    tmp = p.with_name(f"{p.name}.{os.getpid()}.part")
    with open(tmp, "wb") as f:
        np.savez_compressed(f, **d)
    os.replace(tmp, p)
    return p


# ------------------------------------------------------------------------------------------
# load
# ------------------------------------------------------------------------------------------

def _refuse_unsupported(z) -> None:
    for pre, what in _UNSUPPORTED_PREFIXES.items():
        if any(k.startswith(pre) for k in z.files):
            raise NotImplementedError(f"scene.npz holds {what}, which this stepper does not model")
    for key, what in _UNSUPPORTED_LISTS.items():
        if _names(z, key):
            raise NotImplementedError(f"scene.npz holds {what}, which this stepper does not model")
    if _count(z, "n_absorbers"):
        raise NotImplementedError("scene.npz holds absorber slabs, which this stepper does not model")
    if "symmetry" in z.files and any(int(v) for v in z["symmetry"]):
        raise NotImplementedError("symmetry planes are not modelled; export the full domain")
    for name in "xyz":
        if f"flat_{name}" in z.files and bool(z[f"flat_{name}"]):
            raise NotImplementedError(f"axis {name} is a 2D flat axis, which is not modelled")


def _load_tfsf(z, i: int, dt: float) -> TFSFSource:
    m = [int(v) for v in z[f"tfsf{i}_meta"]]
    if f"tfsf{i}_obl" in z.files:
        raise NotImplementedError("oblique TF/SF boxes from scene.npz are not supported")
    axis = m[10] if len(m) > 10 else 2
    open_axes = tuple(q for q in range(3) if len(m) > 11 and (m[11] >> q) & 1)
    pol = z[f"tfsf{i}_pol"] if f"tfsf{i}_pol" in z.files else np.array([1.0, 0.0])
    kxy = (tuple(float(v) for v in z[f"oem_tfsf{i}_kxy"])
           if f"oem_tfsf{i}_kxy" in z.files else (0.0, 0.0))
    scale = float(z[f"oem_tfsf{i}_amp_scale"]) if f"oem_tfsf{i}_amp_scale" in z.files else 1.0
    eps_line = None
    if f"oem_tfsf{i}_eps_edge" in z.files:
        eps_line = (z[f"oem_tfsf{i}_eps_edge"], z[f"oem_tfsf{i}_eps_ctr"])
    return TFSFSource(box_lo=tuple(m[1:4]), box_hi=tuple(m[4:7]), direction=m[0],
                      waveform=_waveform(z, f"tfsf{i}", dt), open_axes=open_axes, kxy=kxy,
                      jxy=(float(pol[0]), float(pol[1])), eps_line=eps_line,
                      amp_scale=scale, axis=axis)


def load(path) -> Scene:
    """Read a ``scene.npz`` (format versions 1..10) into a :class:`Scene`; nothing in the file
    is executed (``allow_pickle=False``)."""
    with np.load(path, allow_pickle=False) as z:
        ver = int(z["format_version"])
        if not 1 <= ver <= FORMAT_VERSION:
            raise ValueError(f"scene.npz format version {ver}; this reader accepts "
                             f"1..{FORMAT_VERSION}")
        _refuse_unsupported(z)
        dt = float(z["dt"])
        axes = [Axis(np.asarray(z[f"edges_{n}"], dtype=np.float64), *(str(v) for v in z[f"bnd_{n}"]))
                for n in "xyz"]
        grid = Grid(*axes)
        pml = {}
        for key in z["pml_keys"]:
            a, side = str(key).split(":")
            pml[(int(a), side)] = cpml.PMLCoeffs(
                **{f: np.asarray(z[f"pml_{a}_{side}_{f}"]) for f in _PML_FIELDS})
        sources = []
        for i in range(_count(z, "n_sources")):
            axis, plane, direction, pol, nf = (int(v) for v in z[f"src{i}_meta"])
            center = float(z[f"src{i}_center"]) if f"src{i}_center" in z.files else float("nan")
            sources.append(PlaneWaveSource(
                axis=axis, plane_index=plane, direction=direction, pol_axis=pol,
                waveform=_waveform(z, f"src{i}", dt), ex_inc=np.asarray(z[f"src{i}_ex_inc"]),
                hy_inc=np.asarray(z[f"src{i}_hy_inc"]),
                angle_theta=float(z[f"src{i}_angle_theta"]), num_freqs=nf, center=center))
        mode_sources = []
        for i in range(_count(z, "n_mode_sources")):
            if any(k.startswith(f"msrc{i}_bb_") for k in z.files):
                raise NotImplementedError("broadband mode-source correction terms are not modelled")
            meta = [int(v) for v in z[f"msrc{i}_meta"]]
            ne = z[f"msrc{i}_n_eff"]
            mode_sources.append(ModeSource(
                plane_index=meta[0], direction=meta[1], n_eff=complex(ne[0], ne[1]),
                axis=meta[2] if len(meta) > 2 else 0,
                **{f: np.asarray(z[f"msrc{i}_{f}"]) for f in _MODE_FIELDS}))
        tfsf = [_load_tfsf(z, i, dt) for i in range(_count(z, "n_tfsf"))]
        dispersion = None
        if "disp_comp" in z.files:
            dispersion = Dispersion(**{f: np.asarray(z[f"disp_{f}"])
                                       for f in ("comp", "cell", "pole_ofs", "am1", "b", "g")})
        flux_mons = []
        for i, name in enumerate(_names(z, "mon_names")):
            axis, plane, ndir = (int(v) for v in z[f"mon{i}_meta"])
            tr = (tuple(tuple(int(x) for x in r) for r in z[f"mon{i}_transverse"])
                  if f"mon{i}_transverse" in z.files else None)
            tb = (tuple(tuple(float(x) for x in r) for r in z[f"mon{i}_tbounds"])
                  if f"mon{i}_tbounds" in z.files else None)
            ap = _apod_unpack(z[f"mon{i}_apod"]) if f"mon{i}_apod" in z.files else None
            flux_mons.append(FluxMonitor(name=name, axis=axis, plane_index=plane,
                                         normal_dir=ndir, freqs=np.asarray(z[f"mon{i}_freqs"]),
                                         transverse=tr, apodization=ap, t_bounds=tb))
        field_mons = []
        for i, name in enumerate(_names(z, "fmon_names")):
            m = [int(v) for v in z[f"fmon{i}_meta"]]
            ap = _apod_unpack(z[f"fmon{i}_apod"]) if f"fmon{i}_apod" in z.files else None
            field_mons.append(FieldMonitor(name=name, freqs=np.asarray(z[f"fmon{i}_freqs"]),
                                           origin=tuple(m[0:3]), box=tuple(m[3:6]),
                                           apodization=ap,
                                           interval_space=tuple(m[6:9]) if len(m) > 6
                                           else (1, 1, 1)))
        mode_mons = []
        planes = _names(z, "modemon_planes")
        for i, name in enumerate(_names(z, "modemon_names")):
            if f"oem_modemon{i}_meta" in z.files:
                axis, plane, idx = (int(v) for v in z[f"oem_modemon{i}_meta"])
                freqs = np.asarray(z[f"oem_modemon{i}_freqs"])
                ng = float(z[f"oem_modemon{i}_n_guess"])
                mode_mons.append(ModeMonitor(name=name, axis=axis, plane_index=plane,
                                             freqs=freqs, mode_index=idx,
                                             n_guess=None if np.isnan(ng) else ng))
                continue
            # the reference expands a mode monitor into a planar field monitor whose box keeps
            # one spare cell below the plane along the normal (openem/scene/monitors.py:188-215)
            fm = next((m for m in field_mons if m.name == planes[i]), None)
            if fm is None:
                raise ValueError(f"mode monitor {name}: no field plane named {planes[i]}")
            axis = int(np.argmin(fm.box))
            mode_mons.append(ModeMonitor(name=name, axis=axis,
                                         plane_index=fm.origin[axis] + fm.box[axis] - 1,
                                         freqs=fm.freqs))
        time_mons = []
        for i, name in enumerate(_names(z, "tmon_names")):
            beg, end, iv, slots = (int(v) for v in z[f"tmon{i}_meta"])
            time_mons.append(FieldTimeMonitor(
                name=name, comps=tuple(int(v) for v in z[f"tmon{i}_comps"]),
                origin=tuple(int(v) for v in z[f"tmon{i}_origin"]),
                box=tuple(int(v) for v in z[f"tmon{i}_box"]), step_begin=beg, step_end=end,
                interval=iv, num_slots=slots))
        perm_mons = [PermittivityMonitor(name=n, freqs=np.asarray(z[f"pmon{i}_freqs"]),
                                         origin=tuple(int(v) for v in z[f"pmon{i}_meta"][:3]),
                                         box=tuple(int(v) for v in z[f"pmon{i}_meta"][3:6]))
                     for i, n in enumerate(_names(z, "pmon_names"))]
        proj_mons = [ProjectionMonitor(name=n, surface_names=tuple(_names(z, f"projmon{i}_surfaces")),
                                       normal_dirs=tuple(_names(z, f"projmon{i}_dirs")))
                     for i, n in enumerate(_names(z, "projmon_names"))]
        box_mons = [BoxFluxMonitor(name=n, face_names=tuple(_names(z, f"bfmon{i}_faces")))
                    for i, n in enumerate(_names(z, "bfmon_names"))]
        flux_time_mons = []
        for i, name in enumerate(_names(z, "ftmon_names")):
            axis, plane, ndir, beg, end, iv, slots = (int(v) for v in z[f"ftmon{i}_meta"])
            flux_time_mons.append(FluxTimeMonitor(
                name=name, axis=axis, plane_index=plane, normal_dir=ndir, step_begin=beg,
                step_end=end, interval=iv, num_slots=slots,
                t_bounds=tuple(tuple(float(v) for v in r) for r in z[f"ftmon{i}_tbounds"]),
                frac=float(z[f"ftmon{i}_frac"])))
        dipoles = []
        for i in range(_count(z, "n_dipoles")):
            dipoles.append(PointDipole(
                component=int(z[f"dip{i}_component"]), indices=np.asarray(z[f"dip{i}_indices"]),
                coef=np.asarray(z[f"dip{i}_coef"]), waveform=_waveform(z, f"dip{i}", dt),
                magnetic=bool(int(z[f"dip{i}_magnetic"])) if f"dip{i}_magnetic" in z.files
                else False))
        pec = {c: (np.asarray(z[f"pec_{c}"]) if f"pec_{c}" in z.files else None)
               for c in ("ex", "ey", "ez")}
        sig = {c: _opt(z, f"sigma_{c}") for c in ("ex", "ey", "ez")}
        return Scene(
            grid=grid, dt=dt, num_time_steps=int(z["num_time_steps"]),
            eps_ex=np.asarray(z["eps_ex"]), eps_ey=np.asarray(z["eps_ey"]),
            eps_ez=np.asarray(z["eps_ez"]), pml=pml, shutoff=float(z["shutoff"]),
            sigma_ex=sig["ex"], sigma_ey=sig["ey"], sigma_ez=sig["ez"],
            pec_ex=pec["ex"], pec_ey=pec["ey"], pec_ez=pec["ez"], dispersion=dispersion,
            bloch_k=(tuple(float(v) for v in z["bloch_k"]) if "bloch_k" in z.files
                     else (0.0, 0.0, 0.0)),
            sources=sources, tfsf_sources=tfsf, dipoles=dipoles, mode_sources=mode_sources,
            flux_monitors=flux_mons, field_monitors=field_mons, mode_monitors=mode_mons,
            field_time_monitors=time_mons, flux_time_monitors=flux_time_mons,
            permittivity_monitors=perm_mons, projection_monitors=proj_mons,
            box_flux_monitors=box_mons)
