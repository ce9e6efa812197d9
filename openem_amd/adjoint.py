"""Gradients by the discrete adjoint on the device (real and complex Bloch fields, ADE poles
included).

The reference computes inverse-design gradients through tidy3d's adjoint
(``openem/nb/autograd_hook.py:1-28``).  Here the exact transpose of the time step runs on the
MI355X (``csrc/oem_adjoint.hip``) with segmented checkpoint recomputation:

1. forward run, the state checkpointed every ``checkpoint_every`` steps;
2. the objective: ``J = |a+|^2`` of a mode monitor (the device mode overlap), or ``J = |a|^2``
   of a fixed projection ``a = sum K . phasors`` of any plane monitor;
3. for each segment from the last: reload its checkpoint, recompute it with E recorded at the
   design cells, and step the adjoint back over it, accumulating
   ``dJ/d eps = sum_n lambda_E^{n+1} (ca' E^n + cb' S^n)`` (``Re(conj(lambda_E^{n+1}) (...))``
   on complex fields, whose step transposes to its conjugate transpose).

When E at the design cells for every step fits in ``record_bytes`` of device memory, the forward
run records it directly and the adjoint steps back over the whole run at once: no checkpoints
and no recomputation (the adjoint recursion itself needs no forward fields).

The gradient is of the discrete objective the stepper computes, so it agrees with central
finite differences of the same stepper to their truncation error.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import _lib, cpml, modes as modes_mod, tables as tables_mod
from .grid import transverse_axes
from .solver import Stepper


@dataclass
class Gradient:
    J: float                   # |a+|^2
    a_plus: complex
    grad: np.ndarray           # dJ/d eps per design entry
    steps: int
    segments: int


def _design_g(sc, comp: np.ndarray, cell: np.ndarray) -> np.ndarray:
    """The ADE coupling ``G = sum 2 Re B`` of each design entry's (component, cell), 0 where it
    has no poles (reference ``openem/coeffs.py:98-136`` folds it into ca and cb)."""
    out = np.zeros(comp.size)
    if not sc.any_dispersion:
        return out
    d = sc.dispersion
    ncell = int(np.prod(sc.shape))
    key = d.comp.astype(np.int64) * ncell + d.cell
    order = np.argsort(key, kind="stable")
    ks = key[order]
    want = comp.astype(np.int64) * ncell + cell
    pos = np.minimum(np.searchsorted(ks, want), ks.size - 1)
    hit = ks[pos] == want
    out[hit] = d.g[order[pos[hit]]]
    return out


def design_coefficients(tabs, comp: np.ndarray, cell: np.ndarray):
    """``(dca, dcb, ca, cb)`` float32 per design entry: the E-update coefficients
    ``ca = (eps - s)/(eps + s + G)``, ``cb = (dt/eps0)/(eps + s + G)`` and their eps derivatives
    ``(2 s + G)/den^2``, ``-(dt/eps0)/den^2`` (G, the ADE coupling of an entry with poles, does
    not depend on eps)."""
    sc = tabs.scene
    eps = (sc.eps_ex, sc.eps_ey, sc.eps_ez)
    sig = (sc.sigma_ex, sc.sigma_ey, sc.sigma_ez)
    gd = _design_g(sc, comp, cell)
    out = [np.zeros(comp.size) for _ in range(4)]
    for c in range(3):
        sel = comp == c
        if not sel.any():
            continue
        e = np.broadcast_to(np.asarray(eps[c], dtype=np.float64), sc.shape).reshape(-1)[cell[sel]]
        s = (0.0 if sig[c] is None else
             np.broadcast_to(np.asarray(sig[c]), sc.shape).reshape(-1)[cell[sel]]
             * sc.dt / (2.0 * cpml.EPSILON_0))
        den = e + s + gd[sel]
        out[0][sel] = (2.0 * s + gd[sel]) / den ** 2
        out[1][sel] = -(sc.dt / cpml.EPSILON_0) / den ** 2
        out[2][sel] = tabs.ca[c].reshape(-1)[cell[sel]]
        out[3][sel] = tabs.cb[c].reshape(-1)[cell[sel]]
    return [np.ascontiguousarray(x, dtype=np.float32) for x in out]


def _plane_cells(shape, axis: int, k: int) -> np.ndarray:
    """Flat cells ``(n1, n2)`` of plane ``k`` in the monitors' ascending transverse layout."""
    t1, t2 = transverse_axes(axis)
    idx = [None, None, None]
    idx[axis] = np.full((shape[t1], shape[t2]), k)
    idx[t1], idx[t2] = np.meshgrid(np.arange(shape[t1]), np.arange(shape[t2]), indexing="ij")
    return ((idx[0] * shape[1] + idx[1]) * shape[2] + idx[2]).ravel()


def mode_power_gradient(sc, monitor: str, comp, cell, f_index: int = 0,
                        checkpoint_every: int = 500, n_total: int | None = None,
                        device: int = 0, record_bytes: int = 24 << 30) -> Gradient:
    """``J = |a+|^2`` of mode monitor ``monitor`` at frequency ``f_index`` and ``dJ/d eps`` at the
    design entries (``comp`` 0..2, flat ``cell``).  With ``record_bytes`` too small for the whole
    record, the forward run is checkpointed every ``checkpoint_every`` steps and recomputed
    segment by segment."""
    tabs = tables_mod.build(sc, n_total)
    names = [m.name for m in tabs.monitors]
    mi = names.index(monitor)
    m = tabs.monitors[mi]
    if not m.modes:
        raise ValueError(f"{monitor} is not a mode monitor")
    # a+ = sum K . phasors with K the summed overlap weights
    g = modes_mod.overlap_weight_sets(m.modes, sc.grid)
    K = 0.25 * modes_mod.orientation(m.axis) * (g[0, :, f_index] + g[1, :, f_index])
    return _overlap_gradient(sc, tabs, mi, K, True, comp, cell, f_index, checkpoint_every,
                             device, record_bytes)


def overlap_power_gradient(sc, monitor: str, weights, comp, cell, f_index: int = 0,
                           checkpoint_every: int = 500, n_total: int | None = None,
                           device: int = 0, record_bytes: int = 24 << 30) -> Gradient:
    """``J = |a|^2`` with ``a = sum weights . phasors`` of plane monitor ``monitor`` at frequency
    ``f_index`` (``weights`` (4, n1, n2) complex over its E t1, E t2, H t1, H t2 slots: a
    projection on a diffraction order or any fixed field profile), and ``dJ/d eps`` at the design
    entries, on real or complex (Bloch) fields."""
    tabs = tables_mod.build(sc, n_total)
    names = [m.name for m in tabs.monitors]
    mi = names.index(monitor)
    m = tabs.monitors[mi]
    if m.kind != "plane":
        raise ValueError(f"{monitor} is not a plane monitor")
    K = np.asarray(weights, dtype=np.complex128)
    if K.shape != (4, m.n1, m.n2):
        raise ValueError(f"weights of shape {K.shape}, the monitor needs {(4, m.n1, m.n2)}")
    return _overlap_gradient(sc, tabs, mi, K, False, comp, cell, f_index, checkpoint_every,
                             device, record_bytes)


def _overlap_gradient(sc, tabs, mi: int, K: np.ndarray, mode: bool, comp, cell, f_index: int,
                      checkpoint_every: int, device: int, record_bytes: int) -> Gradient:
    """The forward run (recorded, or checkpointed), ``a = sum K . phasors`` of plane monitor
    ``mi`` (the device mode overlap when ``mode``), the objective's point lists and the backward
    sweep.  A sample x of slot c at step n enters ``a`` as ``K_c w_n x``, so dJ/dx is
    ``2 Re(conj(a) K_c w_n)`` on real fields and ``2 a conj(K_c w_n)`` (``dJ/d Re x + i dJ/d Im
    x``) on complex ones: the device adds ``z tab[n]`` (its real part on real fields), so complex
    fields take ``z`` and the phase table conjugated."""
    m = tabs.monitors[mi]
    cplx = bool(tabs.bloch)
    comp = np.ascontiguousarray(comp, dtype=np.int32)
    cell = np.ascontiguousarray(cell, dtype=np.int32)
    n = tabs.n_total
    one_pass = (n + 1) * comp.size * 4 * (2 if cplx else 1) <= record_bytes
    Kseg = n if one_pass else max(1, int(checkpoint_every))
    n_seg = (n + Kseg - 1) // Kseg
    dca, dcb, ca, cb = design_coefficients(tabs, comp, cell)
    st = Stepper(tabs, device=device)
    lib, ctx = st.lib, st.ctx
    try:
        _lib.check(lib, lib.oem_adjoint_setup(ctx, comp.size, _lib.ptr(comp), _lib.ptr(cell),
                                              _lib.ptr(dca), _lib.ptr(dcb), _lib.ptr(ca),
                                              _lib.ptr(cb), Kseg))
        if one_pass:
            _lib.check(lib, lib.oem_adjoint_forward(ctx, n))
        else:
            _lib.check(lib, lib.oem_checkpoint_alloc(ctx, n_seg))
            for s in range(n_seg):
                _lib.check(lib, lib.oem_checkpoint_save(ctx, s))
                st.run(min(Kseg, n - s * Kseg))
        if mode:
            st.setup_modes()
            ap, _ = st.mode_amplitudes(mi)
            a = complex(ap[f_index])
        else:
            a = complex(np.sum(K * st.read_monitor(mi)[:, f_index]))
        t1, t2 = transverse_axes(m.axis)
        ph = m.phase[:, f_index]
        we = ph[:, 0] + 1j * ph[:, 1]
        wh = ph[:, 2] + 1j * ph[:, 3]
        z = 2.0 * np.conj(a) * K                          # (4, n1, n2)
        if cplx:
            z, we, wh = np.conj(z), np.conj(we), np.conj(wh)
        tab_e = np.ascontiguousarray(we, dtype=np.complex128).view(np.float64)
        tab_h = np.ascontiguousarray(wh, dtype=np.complex128).view(np.float64)
        cells_k = _plane_cells(tabs.shape, m.axis, m.km)
        cells_km = _plane_cells(tabs.shape, m.axis, m.km - 1)
        for magnetic, slot, fcomp, planes, tab, w in (
                (0, 0, t1, (cells_k,), tab_e, 1.0), (0, 1, t2, (cells_k,), tab_e, 1.0),
                (1, 2, t1, (cells_k, cells_km), tab_h, 0.5), (1, 3, t2, (cells_k, cells_km), tab_h, 0.5)):
            zc = (w * z[slot]).ravel()
            keep = np.flatnonzero(zc != 0)
            if keep.size == 0:
                continue
            for cells in planes:
                pc = np.full(keep.size, fcomp, np.int32)
                pcell = np.ascontiguousarray(cells[keep], dtype=np.int32)
                zz = np.ascontiguousarray(zc[keep], dtype=np.complex64).view(np.float32)
                _lib.check(lib, lib.oem_adjoint_objective(ctx, magnetic, keep.size, _lib.ptr(pc),
                                                          _lib.ptr(pcell), _lib.ptr(zz),
                                                          _lib.ptr(tab)))
        if one_pass:
            _lib.check(lib, lib.oem_adjoint_back(ctx, n))
        else:
            for s in range(n_seg - 1, -1, -1):
                ln = min(Kseg, n - s * Kseg)
                _lib.check(lib, lib.oem_checkpoint_load(ctx, s))
                _lib.check(lib, lib.oem_adjoint_record(ctx, ln))
                _lib.check(lib, lib.oem_adjoint_back(ctx, ln))
        out = np.zeros(comp.size, np.float64)
        _lib.check(lib, lib.oem_adjoint_gradient(ctx, _lib.ptr(out), out.size))
    finally:
        st.close()
    return Gradient(J=abs(a) ** 2, a_plus=a, grad=out, steps=n, segments=n_seg)


def mode_power_gradients(jobs, concurrency: int = 4, on_done=None, **kw) -> list:
    """:func:`mode_power_gradient` for many designs, ``concurrency`` of them in flight at once:
    each ``(scene, monitor, comp, cell)`` job has its own context (its own HIP stream, fields,
    checkpoints and adjoint state), so one design's host work and launch gaps overlap another's
    kernels (``concurrency`` 1: one after another on the calling thread).  ``on_done(index,
    gradient)`` is called as each design finishes.  Results are in job order."""
    return _in_flight(mode_power_gradient, jobs, concurrency, on_done, kw)


def overlap_power_gradients(jobs, concurrency: int = 4, on_done=None, **kw) -> list:
    """:func:`overlap_power_gradient` for many ``(scene, monitor, weights, comp, cell)`` jobs,
    ``concurrency`` in flight at once, as :func:`mode_power_gradients`."""
    return _in_flight(overlap_power_gradient, jobs, concurrency, on_done, kw)


def _in_flight(fn, jobs, concurrency: int, on_done, kw: dict) -> list:
    from concurrent.futures import ThreadPoolExecutor

    def one(ij):
        g = fn(*ij[1], **kw)
        if on_done is not None:
            on_done(ij[0], g)
        return g

    if int(concurrency) <= 1:               # one at a time, on the calling thread
        return [one(ij) for ij in enumerate(jobs)]
    with ThreadPoolExecutor(int(concurrency)) as ex:
        return list(ex.map(one, enumerate(jobs)))
