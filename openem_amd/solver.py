"""Time stepping on the MI355X through the C ABI.

:func:`run` mirrors the reference's ``openem.solver.run(sc, num_steps=None, use_shutoff=True,
kernels=None, verbose=True, return_fields=False, policy=None, force_complex=False) -> Result``
(``openem/solver.py:403-511``): it builds the tables on the host, uploads them once, advances
the scene on the GPU and reads back float64 phasors. :class:`Stepper` is the lower-level handle
(one resident scene) the bench and the serving loop use.
"""

from __future__ import annotations

import ctypes as C
import time

import numpy as np

from . import _lib, cpml, flux as flux_mod, shutoff as shutoff_mod, tables as tables_mod
from .model import COMPONENTS, Scene
from .results import Phasors, Result


def _f32(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a).real if np.iscomplexobj(a) else a, dtype=np.float32)


def _i32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.int32)


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def plane_weights(tabs: tables_mod.StepTables, i: int, transverse: bool = True):
    """Face weights ``(w1, w2)`` (float64, the plane's ``(n1, n2)`` layout) of plane monitor
    ``i`` for the device flux reduction.  A flux monitor's transverse index window is applied
    when ``transverse``; a mode monitor always spans its whole plane."""
    m = tabs.monitors[i]
    if m.kind != "plane":
        raise ValueError(f"monitor {m.name} is not a plane monitor")
    w1, w2 = flux_mod.area_weights(tabs.scene.grid, m.axis)
    win = getattr(m.spec, "transverse", None) if m.modes is None else None
    if transverse and win is not None:
        mask = np.zeros(w1.shape)
        (j0, j1), (k0, k1) = win
        mask[j0:j1 + 1, k0:k1 + 1] = 1.0
        w1, w2 = w1 * mask, w2 * mask
    return (np.ascontiguousarray(w1, dtype=np.float64),
            np.ascontiguousarray(w2, dtype=np.float64))


class Stepper:
    """One scene resident on the GPU.  With ``pool`` (a :class:`openem_amd.serving.Pool`) its
    device memory comes from the pool's arena and the pool steps it once attached."""

    #: step forms of oem_set_step_form
    FORMS = {None: 0, "sweeps": 1, "fused": 2}

    def __init__(self, tabs: tables_mod.StepTables, device: int = 0, dft_batch: int = 16,
                 dense_materials: bool = False, pool=None, chunk: int = 0,
                 form: str | None = None):
        self.lib = lib = _lib.load()
        self.pool = pool
        self.cplx = bool(tabs.bloch)
        self.tabs = tabs
        nx, ny, nz = tabs.shape
        d = _lib.Dims()
        d.nx, d.ny, d.nz = nx, ny, nz
        d.n_steps = tabs.n_total
        for a in range(3):
            d.pml_lo[a] = tabs.axes[a].pml_lo
            d.pml_hi[a] = tabs.axes[a].pml_hi
        d.complex_fields = 1 if self.cplx else 0
        d.dft_batch = int(dft_batch)
        d.ch = np.float32(tabs.ch)
        d.cw = np.float32(tabs.cw)
        self.ctx = C.c_void_p()
        if pool is None:
            _lib.check(lib, lib.oem_ctx_create(C.byref(d), int(device), C.byref(self.ctx)))
        else:
            _lib.check(lib, lib.oem_pool_ctx_create(pool.handle, C.byref(d), C.byref(self.ctx)))
        try:
            if chunk:
                _lib.check(lib, lib.oem_set_chunk(self.ctx, int(chunk)))
            if form is not None:
                _lib.check(lib, lib.oem_set_step_form(self.ctx, self.FORMS[form]))
            self._upload(dense_materials)
            _lib.check(lib, lib.oem_prepare(self.ctx))
        except Exception:
            self.close()
            raise

    def _upload(self, dense: bool):
        lib, t, ctx = self.lib, self.tabs, self.ctx
        for a, ax in enumerate(t.axes):
            arrs = [_f32(ax.idl_p), _f32(ax.idl_d), _i32(ax.nxt), _i32(ax.prv), _f32(ax.mnx),
                    _f32(ax.mpv), _f32(ax.pec), _f32(ax.a_e), _f32(ax.b_e), _f32(ax.k_e),
                    _f32(ax.a_h), _f32(ax.b_h), _f32(ax.k_h),
                    np.ascontiguousarray(np.imag(ax.mnx), dtype=np.float32),
                    np.ascontiguousarray(np.imag(ax.mpv), dtype=np.float32)]
            _lib.check(lib, lib.oem_set_axis(ctx, a, *[_lib.ptr(x) for x in arrs]))
        mc = None if dense else t.material_classes()
        if mc is not None:
            cls, lut = mc
            cls = np.ascontiguousarray(cls, dtype=np.uint16)
            lut = np.ascontiguousarray(lut, dtype=np.float32)
            _lib.check(lib, lib.oem_set_materials_lut(ctx, _lib.ptr(cls), lut.shape[0],
                                                      _lib.ptr(lut)))
            self.n_classes = int(lut.shape[0])
        else:
            ca, cb = _f32(t.ca), _f32(t.cb)
            _lib.check(lib, lib.oem_set_materials_dense(ctx, _lib.ptr(ca), _lib.ptr(cb)))
            self.n_classes = 0
        if t.disp is not None:
            dp = t.disp
            arrs = [_i32(dp.comp), _i32(dp.cell), _i32(dp.ofs), _i32(dp.cidx),
                    _f32(dp.lut_am1.real), _f32(dp.lut_am1.imag), _f32(dp.lut_b.real),
                    _f32(dp.lut_b.imag), _i32(dp.hist_box)]
            _lib.check(lib, lib.oem_set_dispersion(
                ctx, dp.comp.size, _lib.ptr(arrs[0]), _lib.ptr(arrs[1]), _lib.ptr(arrs[2]),
                dp.am1.size, _lib.ptr(arrs[3]), dp.lut_am1.size, _lib.ptr(arrs[4]),
                _lib.ptr(arrs[5]), _lib.ptr(arrs[6]), _lib.ptr(arrs[7]), _lib.ptr(arrs[8])))
        for ps in t.plane_srcs:
            ex, hy = _f64(ps.ex_inc[:t.n_total + 1]), _f64(ps.hy_inc[:t.n_total + 1])
            _lib.check(lib, lib.oem_add_plane_source(ctx, ps.pol, ps.kh, ps.ks,
                                                     np.float32(ps.coef_h),
                                                     np.float32(ps.coef_e),
                                                     _lib.ptr(ex), _lib.ptr(hy)))
        for dp in t.dipoles:
            cplx_amp = np.iscomplexobj(dp.amp)
            amp = (np.ascontiguousarray(dp.amp, dtype=np.complex128).view(np.float64)
                   if cplx_amp else _f64(dp.amp))
            arrs = [_i32(dp.comp), _i32(dp.cell), _f32(dp.coef), _i32(dp.src_of), amp]
            _lib.check(lib, lib.oem_add_dipoles(ctx, int(dp.magnetic), dp.comp.size,
                                                _lib.ptr(arrs[0]), _lib.ptr(arrs[1]),
                                                _lib.ptr(arrs[2]), _lib.ptr(arrs[3]),
                                                dp.amp.shape[0], _lib.ptr(arrs[4]),
                                                int(cplx_amp)))
        for tf in t.tfsf:
            ncol = len(tf.cols)
            for magnetic, (comp, cell, coef, col, ph, tab) in (
                    (1, (tf.h_comp, tf.h_cell, tf.h_coef, tf.h_col, tf.h_phase, tf.tab_h)),
                    (0, (tf.e_comp, tf.e_cell, tf.e_coef, tf.e_col, tf.e_phase, tf.tab_e))):
                if comp.size == 0:
                    continue
                phase = np.ascontiguousarray(ph, dtype=np.complex64).view(np.float32)
                table = np.ascontiguousarray(tab, dtype=np.complex128).view(np.float64)
                arrs = [_i32(comp), _i32(cell), _f32(coef), _i32(col), phase, table]
                _lib.check(lib, lib.oem_add_tfsf_points(
                    ctx, magnetic, comp.size, *[_lib.ptr(x) for x in arrs[:5]], ncol,
                    _lib.ptr(arrs[5])))
        for mt in t.modes:
            for magnetic, (comp, cell, coef, ph, amp) in (
                    (1, (mt.h_comp, mt.h_cell, mt.h_coef, mt.h_phase, mt.amp_e)),
                    (0, (mt.e_comp, mt.e_cell, mt.e_coef, mt.e_phase, mt.amp_h))):
                if comp.size == 0:
                    continue
                phase = np.ascontiguousarray(ph, dtype=np.complex64).view(np.float32)
                table = np.ascontiguousarray(amp, dtype=np.complex128).view(np.float64)
                arrs = [_i32(comp), _i32(cell), _f32(coef), phase, table]
                _lib.check(lib, lib.oem_add_mode_points(ctx, magnetic, comp.size,
                                                        *[_lib.ptr(x) for x in arrs]))
        self.mon_ids = []
        for m in t.monitors:
            ph = _f64(m.phase)
            mid = C.c_int32(-1)
            if m.kind == "plane":
                _lib.check(lib, lib.oem_add_flux_plane(ctx, m.axis, m.km, m.nf, _lib.ptr(ph),
                                                       C.byref(mid)))
            else:
                o, b = _i32(m.origin), _i32(m.box)
                _lib.check(lib, lib.oem_add_field_box(ctx, _lib.ptr(o), _lib.ptr(b), m.nf,
                                                      _lib.ptr(ph), C.byref(mid)))
            self.mon_ids.append(int(mid.value))
        self.tmon_ids = []
        for m in t.scene.field_time_monitors:
            mid = C.c_int32(-1)
            comps, o, b = _i32(m.comps), _i32(m.origin), _i32(m.box)
            _lib.check(lib, lib.oem_add_time_monitor(ctx, comps.size, _lib.ptr(comps), _lib.ptr(o),
                                                     _lib.ptr(b), m.step_begin, m.step_end,
                                                     m.interval, m.num_slots, C.byref(mid)))
            self.tmon_ids.append(int(mid.value))
        self.ftmon_ids = []
        for m in t.scene.flux_time_monitors:
            g = flux_mod.flux_time_geometry(t.scene.grid, m)
            (e1, l1, h1, w1, tr1), (e2, l2, h2, w2, tr2) = g["s"]
            pl = g["planes"]
            arrs1 = [_i32(e1), _i32(l1), _i32(h1), _f32(w1), _f64(tr1)]
            arrs2 = [_i32(e2), _i32(l2), _i32(h2), _f32(w2), _f64(tr2)]
            pw = _f64([x[0] for x in pl])
            ke, kh = _i32([x[1] for x in pl]), _i32([x[2] for x in pl])
            wlo, whi = _f32([x[3] for x in pl]), _f32([x[4] for x in pl])
            mid = C.c_int32(-1)
            _lib.check(lib, lib.oem_add_flux_time_monitor(
                ctx, m.axis, e1.size, *[_lib.ptr(x) for x in arrs1], e2.size,
                *[_lib.ptr(x) for x in arrs2], len(pl), _lib.ptr(pw), _lib.ptr(ke), _lib.ptr(kh),
                _lib.ptr(wlo), _lib.ptr(whi), float(g["sign"]), m.step_begin, m.step_end,
                m.interval, m.num_slots, C.byref(mid)))
            self.ftmon_ids.append(int(mid.value))

    @property
    def launch(self) -> dict:
        """The launch plan in use: ``{"form": "sweeps" | "fused", "chunk": x planes}``."""
        f, c = C.c_int32(0), C.c_int32(0)
        _lib.check(self.lib, self.lib.oem_launch_plan(self.ctx, C.byref(f), C.byref(c)))
        return {"form": "fused" if f.value == 2 else "sweeps", "chunk": int(c.value)}

    # -- stepping ---------------------------------------------------------------------------
    def run(self, n_steps: int) -> None:
        _lib.check(self.lib, self.lib.oem_run(self.ctx, int(n_steps)))

    def timed_run(self, n_steps: int) -> float:
        ms = C.c_double(0.0)
        _lib.check(self.lib, self.lib.oem_timed_run(self.ctx, int(n_steps), C.byref(ms)))
        return float(ms.value)

    def profile_run(self, n_steps: int) -> dict:
        ms = np.zeros(_lib.K_COUNT, np.float64)
        calls = np.zeros(_lib.K_COUNT, np.int64)
        _lib.check(self.lib, self.lib.oem_profile_run(self.ctx, int(n_steps), _lib.ptr(ms),
                                                      _lib.ptr(calls)))
        return {k: (float(ms[i]), int(calls[i])) for i, k in enumerate(_lib.KERNELS)}

    @property
    def step_index(self) -> int:
        n = C.c_int64(0)
        _lib.check(self.lib, self.lib.oem_step_index(self.ctx, C.byref(n)))
        return int(n.value)

    # -- readout ----------------------------------------------------------------------------
    def read_field(self, comp: int) -> np.ndarray:
        """Field component ``comp`` (0..5 = Ex..Hz): float32, or complex64 on a Bloch scene."""
        dt = np.complex64 if self.cplx else np.float32
        out = np.empty(self.tabs.shape, dt)
        raw = out.view(np.float32)
        _lib.check(self.lib, self.lib.oem_read_field(self.ctx, comp, _lib.ptr(raw), raw.size))
        return out

    def write_field(self, comp: int, arr: np.ndarray) -> None:
        dt = np.complex64 if self.cplx else np.float32
        a = np.ascontiguousarray(arr, dtype=dt).view(np.float32)
        _lib.check(self.lib, self.lib.oem_write_field(self.ctx, comp, _lib.ptr(a), a.size))

    def read_monitor(self, i: int) -> np.ndarray:
        m = self.tabs.monitors[i]
        if m.kind == "plane":
            shp = (4, m.nf, m.n1, m.n2)
        else:
            shp = (6, m.nf) + tuple(m.box)
        re = np.empty(shp, np.float64)
        im = np.empty(shp, np.float64)
        _lib.check(self.lib, self.lib.oem_read_monitor(self.ctx, self.mon_ids[i], _lib.ptr(re),
                                                       _lib.ptr(im), re.size))
        return re + 1j * im

    def read_time_monitor(self, i: int) -> np.ndarray:
        """The whole buffer of field-time monitor ``i``: ``(n_comps, n_slots, ni, nj, nk)``
        float32, or complex64 on a Bloch scene."""
        m = self.tabs.scene.field_time_monitors[i]
        out = np.empty((len(m.comps), m.num_slots) + tuple(m.box),
                       np.complex64 if self.cplx else np.float32)
        raw = out.view(np.float32)
        _lib.check(self.lib, self.lib.oem_read_time_monitor(self.ctx, self.tmon_ids[i],
                                                            _lib.ptr(raw), raw.size))
        return out

    def read_flux_time_monitor(self, i: int) -> np.ndarray:
        """The float64 flux series (all slots) of flux-time monitor ``i``."""
        m = self.tabs.scene.flux_time_monitors[i]
        out = np.empty(m.num_slots, np.float64)
        _lib.check(self.lib, self.lib.oem_read_flux_time_monitor(self.ctx, self.ftmon_ids[i],
                                                                 _lib.ptr(out), out.size))
        return out

    # -- float64 reductions on the device ----------------------------------------------------
    def setup_energy(self) -> None:
        """Upload the energy weights: per-axis primal and dual lengths and eps per class (or
        dense eps when the scene has no class LUT)."""
        sc = self.tabs.scene
        ls = [np.ascontiguousarray(a.dl, dtype=np.float64) for a in sc.grid.axes]
        ds = [np.ascontiguousarray(a.dl_dual, dtype=np.float64) for a in sc.grid.axes]
        if self.n_classes:
            eps, per_class = np.ascontiguousarray(self.tabs.class_eps(), dtype=np.float32), 1
        else:
            eps = np.ascontiguousarray(np.stack([np.broadcast_to(e, sc.shape) for e in
                                                 (sc.eps_ex, sc.eps_ey, sc.eps_ez)]),
                                       dtype=np.float32)
            per_class = 0
        self._energy_keep = ls + ds + [eps]
        _lib.check(self.lib, self.lib.oem_set_energy_tables(
            self.ctx, *[_lib.ptr(a) for a in ls + ds + [eps]], per_class))

    def energy(self) -> np.ndarray:
        """``[sum V|E|^2, sum V eps|E|^2, sum V|H|^2]`` (float64, reduced on the device)."""
        out = np.zeros(3, np.float64)
        _lib.check(self.lib, self.lib.oem_energy(self.ctx, _lib.ptr(out)))
        return out

    def setup_flux(self, transverse: bool = True) -> None:
        """Upload the face weights of every plane monitor (``transverse``: apply the flux
        monitor's transverse window, as :func:`openem_amd.flux.monitor_flux` does)."""
        for i, m in enumerate(self.tabs.monitors):
            if m.kind != "plane":
                continue
            w1, w2 = plane_weights(self.tabs, i, transverse)
            _lib.check(self.lib, self.lib.oem_set_flux_weights(self.ctx, self.mon_ids[i],
                                                               _lib.ptr(w1), _lib.ptr(w2)))

    def flux_terms(self, i: int) -> np.ndarray:
        """``(nf, 2)``: the two face sums of plane monitor ``i`` (flux = (t1 - t2) / 2)."""
        m = self.tabs.monitors[i]
        out = np.zeros((m.nf, 2), np.float64)
        _lib.check(self.lib, self.lib.oem_monitor_flux(self.ctx, self.mon_ids[i], _lib.ptr(out),
                                                       out.size))
        return out

    def plane_flux(self, i: int) -> np.ndarray:
        """``plane_flux`` of monitor ``i`` (no axis sign), reduced on the device."""
        t = self.flux_terms(i)
        return 0.5 * (t[:, 0] - t[:, 1])

    def setup_modes(self) -> None:
        """Upload the overlap weights of every mode monitor."""
        from . import modes as modes_mod
        for i, m in enumerate(self.tabs.monitors):
            if not m.modes:
                continue
            g = modes_mod.overlap_weight_sets(m.modes, self.tabs.scene.grid)
            g = np.ascontiguousarray(g, dtype=np.complex128).view(np.float64)
            _lib.check(self.lib, self.lib.oem_set_overlap_weights(self.ctx, self.mon_ids[i], 2,
                                                                  _lib.ptr(g)))

    def mode_amplitudes(self, i: int):
        """``(a_plus, a_minus)`` ``(nf,)`` complex of mode monitor ``i`` (overlaps reduced on
        the device from the float64 accumulators; the mode carries unit power)."""
        from . import modes as modes_mod
        m = self.tabs.monitors[i]
        out = np.zeros((2, m.nf, 2), np.float64)
        _lib.check(self.lib, self.lib.oem_monitor_overlap(self.ctx, self.mon_ids[i], _lib.ptr(out),
                                                          out.size))
        s = modes_mod.orientation(m.axis) * (out[..., 0] + 1j * out[..., 1])
        return 0.25 * (s[0] + s[1]), 0.25 * (s[0] - s[1])

    def close(self) -> None:
        if getattr(self, "ctx", None) and self.ctx.value:
            self.lib.oem_ctx_destroy(self.ctx)
            self.ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def run(sc: Scene, num_steps: int | None = None, use_shutoff: bool = True, kernels=None,
        verbose: bool = True, return_fields: bool = False, policy=None,
        force_complex: bool = False, device: int = 0) -> Result:
    """Run one simulation; see the module docstring.  With ``use_shutoff`` the run checks the
    ``policy`` (default: :func:`openem_amd.shutoff.default_policy` at the scene's shutoff
    threshold) every ``policy.interval`` steps and may stop early once the sources are done
    (reference ``openem/solver.py:1165-1213``); ``kernels`` is ignored (the library is loaded
    once per process)."""
    if force_complex and not sc.any_bloch:
        raise NotImplementedError("force_complex needs a Bloch scene (set bloch_k)")
    n_total = int(num_steps if num_steps is not None else sc.num_time_steps)
    t0 = time.time()
    tabs = tables_mod.build(sc, n_total)
    st = Stepper(tabs, device=device)
    triggered, reason = False, ""
    decay, energy_tr = [], []
    src_end = shutoff_mod.source_end_step(sc)
    try:
        t1 = time.time()
        if use_shutoff:
            pol = policy or shutoff_mod.default_policy(sc.shutoff)
            st.setup_energy()
            st.setup_flux(transverse=False)
            plane_ids = [i for i, m in enumerate(tabs.monitors) if m.kind == "plane"]
            step = 0
            while step < n_total:
                k = min(pol.interval - step % pol.interval, n_total - step)
                st.run(k)
                step += k
                if step % pol.interval:
                    continue
                e = st.energy()
                u = 0.5 * (cpml.EPSILON_0 * e[1] + cpml.MU_0 * e[2])
                fl = (np.concatenate([st.plane_flux(i) for i in plane_ids]) if plane_ids
                      else np.zeros(0))
                stop = pol.check(step, e[0], u, fl) and step > src_end
                decay.append((step, pol.field_decay.ratio if pol.field_decay else float("nan")))
                energy_tr.append((step, u))
                if stop:
                    triggered, reason = True, pol.reason
                    break
            steps_run = step
        else:
            st.run(n_total)
            steps_run = n_total
        wall = time.time() - t1
        phasors, fphas, mamps = {}, {}, {}
        if any(m.modes for m in tabs.monitors):
            st.setup_modes()
        for i, m in enumerate(tabs.monitors):
            data = st.read_monitor(i)
            if m.kind == "plane":
                phasors[m.name] = Phasors(data=data, freqs=m.freqs)
            else:
                fphas[m.name] = data
            if m.modes:
                ap, am = st.mode_amplitudes(i)
                mamps[m.name] = dict(a_plus=ap, a_minus=am, freqs=m.freqs,
                                     n_eff=np.array([x.n_eff for x in m.modes]))
        tsamp, tslots = {}, {}
        for i, m in enumerate(sc.field_time_monitors):
            filled = m.filled(steps_run)
            tsamp[m.name] = st.read_time_monitor(i)[:, :filled]
            tslots[m.name] = filled
        ftime = {m.name: st.read_flux_time_monitor(i)[:m.filled(steps_run)]
                 for i, m in enumerate(sc.flux_time_monitors)}
        fields = {}
        if return_fields:
            fields = {name: st.read_field(c) for c, name in enumerate(COMPONENTS)}
    finally:
        st.close()
    if verbose:
        print(f"[openem_amd] {sc.shape} x {steps_run} steps: setup {t1 - t0:.2f}s, "
              f"stepping {wall:.2f}s" + (f", stopped: {reason}" if triggered else ""))
    _warn_nonfinite(phasors, fphas)
    summary = ""
    if use_shutoff:
        summary = (f"stopped at step {steps_run}: {reason}" if triggered else
                   f"ran all {steps_run} steps; last decay ratio "
                   f"{decay[-1][1] if decay else float('nan'):.3g}")
    return Result(phasors=phasors, field_phasors=fphas, mode_amps=mamps, steps_run=steps_run,
                  wall_seconds=wall, fields=fields, shutoff_triggered=triggered,
                  stop_reason=reason, shutoff_summary=summary, source_end=src_end,
                  decay_trace=decay, energy_trace=energy_tr, time_samples=tsamp,
                  time_slots=tslots, flux_time=ftime)


def _warn_nonfinite(ph, fph) -> None:
    """NaN/Inf in the outputs is warned, not raised (reference ``openem/solver.py:347-370``)."""
    bad = [k for k, v in ph.items() if not np.all(np.isfinite(v.data))]
    bad += [k for k, v in fph.items() if not np.all(np.isfinite(v))]
    if bad:
        import warnings
        warnings.warn(f"non-finite values in monitors {bad}", RuntimeWarning, stacklevel=3)
