"""Guided modes on the Yee grid: a full-vector finite-difference mode solver, mode-source tables
and mode-monitor overlaps.

The reference takes its profiles from tidy3d's ``ModeSolver`` (``openem/scene/modes.py``),
which is not available here, so this module solves for them itself, on the stepper's own
discretisation: the transverse derivatives are the FDTD's forward (H side: ``nxt``/``mnx`` over
the primal spacing) and backward (E side: ``prv``/``mpv`` over the dual spacing) differences,
the derivative along the normal is the discrete ``i beta~`` with
``beta~ = (2/dl) sin(beta dl / 2)``, and time enters through ``omega~ = (2/dt) sin(omega dt / 2)``.
A mode found here is therefore an exact eigenmode of the time stepper at its frequency (on a
grid uniform along the normal near the plane), which keeps the mode source one-directional.

Local frame: normal ``a``, transverse ``(u, v) = ((a+1)%3, (a+2)%3)`` (right-handed), fields
``~ exp(i (beta s - omega t))``.  On the Yee grid the E plane at normal index ``k`` holds
(E_u, E_v, H_a) and the H plane at ``k + 1/2`` holds (H_u, H_v, E_a).  With
U = forward and V = backward transverse differences, eliminating E_a and H_a gives the
eigenproblem for the transverse H on the H plane (phase referred to the E plane by the
half-cell factor, so the derivative along the normal is exactly ``i beta~``)::

    beta~^2 Hu = k0~^2 eps_v Hu + Vu Uu Hu + Vu Uv Hv + eps_v Uv eps_a^-1 (Vv Hu - Vu Hv)
    beta~^2 Hv = k0~^2 eps_u Hv + Vv Uv Hv + Vv Uu Hu + eps_u Uu eps_a^-1 (Vu Hv - Vv Hu)

and then ``E_a = -(Vu Hv - Vv Hu) / (i w~ eps0 eps_a)``,
``E_u = (i w~ mu0 Hv + Uu E_a) / (i beta~)``, ``E_v = (Uv E_a - i w~ mu0 Hu) / (i beta~)``.
(Maxwell with ``exp(-i w t)``: ``curl E = i w mu0 H``, ``curl H = -i w eps0 eps E``.)
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from . import cpml
from .grid import Grid, transverse_axes
from .model import Scene

C0 = 1.0 / np.sqrt(cpml.EPSILON_0 * cpml.MU_0)


def cyclic(axis: int) -> tuple[int, int]:
    return (axis + 1) % 3, (axis + 2) % 3


@dataclass
class Mode:
    """One guided mode on a plane normal to ``axis`` (arrays ``(n_u, n_v)`` complex128 in the
    cyclic transverse order, each at its own Yee position; H_u/H_v/E_a are referred to the E
    plane, i.e. the physical value at the H plane ``k + 1/2`` is ``h * exp(i beta dl / 2)``).
    Normalised so that the plane-monitor flux of the forward mode (H colocated as the monitors
    do) is 1."""

    axis: int
    plane_index: int
    freq: float
    n_eff: float
    beta: float               # physical propagation constant (rad/m)
    dl_normal: float          # normal spacing at the plane
    e_u: np.ndarray
    e_v: np.ndarray
    h_u: np.ndarray
    h_v: np.ndarray
    e_a: np.ndarray
    h_a: np.ndarray

    @property
    def colocation(self) -> float:
        """Plane monitors average H over the planes k - 1/2 and k + 1/2: cos(beta dl / 2)."""
        return float(np.cos(0.5 * self.beta * self.dl_normal))


def _diff(axis_obj, forward: bool) -> sp.csr_matrix:
    """The stepper's 1D difference along one transverse axis (reference tables
    ``openem/grid.py:219-304``): forward ``(f[nxt]*mnx - f)/dl``, backward
    ``(f - f[prv]*mpv)/dl_dual``."""
    t = axis_obj.index_tables(0.0)
    n = axis_obj.n
    i = np.arange(n)
    if forward:
        inv = 1.0 / axis_obj.dl
        rows = np.concatenate([i, i])
        cols = np.concatenate([t["nxt"], i])
        vals = np.concatenate([np.real(t["mnx"]) * inv, -inv])
    else:
        inv = 1.0 / axis_obj.dl_dual
        rows = np.concatenate([i, i])
        cols = np.concatenate([i, t["prv"]])
        vals = np.concatenate([inv, -np.real(t["mpv"]) * inv])
    return sp.csr_matrix((vals, (rows, cols)), shape=(n, n))


def plane_eps(sc: Scene, axis: int, k: int):
    """(eps_u, eps_v, eps_a) on the plane normal to ``axis`` at index ``k`` in the cyclic
    transverse order, ``(n_u, n_v)`` each."""
    u, v = cyclic(axis)
    eps = (sc.eps_ex, sc.eps_ey, sc.eps_ez)
    out = []
    for c in (u, v, axis):
        e = np.broadcast_to(np.asarray(eps[c], dtype=np.float64), sc.shape)
        sl = np.take(e, k, axis=axis)                 # remaining axes in ascending order
        if u > v:                                      # ascending (v, u) -> (u, v)
            sl = sl.T
        out.append(np.ascontiguousarray(sl))
    return out


_CACHE: dict = {}
_CACHE_MAX = 64


def solve_mode(sc: Scene, axis: int, plane_index: int, freq: float, mode_index: int = 0,
               n_guess: float | None = None, n_modes: int = 4) -> Mode:
    """The ``mode_index``-th guided mode (by decreasing effective index) of the cross-section
    of ``sc`` at ``plane_index`` along ``axis``.  Solutions are cached on everything they
    depend on (the plane's eps, the transverse axes, dt, the frequency and the selection)."""
    import copy
    import hashlib
    u, v = cyclic(axis)
    h = hashlib.sha1()
    for arr in plane_eps(sc, axis, plane_index):
        h.update(np.ascontiguousarray(arr).tobytes())
    for ax in (sc.grid.axes[u], sc.grid.axes[v]):
        h.update(np.ascontiguousarray(ax.edges).tobytes())
        h.update(f"{ax.boundary_lo}/{ax.boundary_hi}".encode())
    key = (h.hexdigest(), axis, float(sc.grid.axes[axis].dl[plane_index]), float(sc.dt),
           float(freq), int(mode_index), n_guess, int(n_modes))
    hit = _CACHE.get(key)
    if hit is None:
        hit = _solve(sc, axis, plane_index, freq, mode_index, n_guess, n_modes)
        if len(_CACHE) >= _CACHE_MAX:
            _CACHE.pop(next(iter(_CACHE)))
        _CACHE[key] = hit
    m = copy.copy(hit)
    m.plane_index = plane_index
    return m


def _solve(sc: Scene, axis: int, plane_index: int, freq: float, mode_index: int,
           n_guess: float | None, n_modes: int) -> Mode:
    g = sc.grid
    u, v = cyclic(axis)
    au, av, aa = g.axes[u], g.axes[v], g.axes[axis]
    nu, nv = au.n, av.n
    eu, ev, ea = plane_eps(sc, axis, plane_index)
    dl = float(aa.dl[plane_index])
    w = 2.0 * np.pi * freq
    wt = 2.0 / sc.dt * np.sin(0.5 * w * sc.dt)          # the stepper's discrete frequency
    k0 = wt / C0
    Iu, Iv = sp.identity(nu, format="csr"), sp.identity(nv, format="csr")
    Uu = sp.kron(_diff(au, True), Iv, format="csr")
    Vu = sp.kron(_diff(au, False), Iv, format="csr")
    Uv = sp.kron(Iu, _diff(av, True), format="csr")
    Vv = sp.kron(Iu, _diff(av, False), format="csr")
    Deu = sp.diags(eu.ravel())
    Dev = sp.diags(ev.ravel())
    Dia = sp.diags(1.0 / ea.ravel())
    A_uu = k0 * k0 * Dev + Vu @ Uu + Dev @ Uv @ Dia @ Vv
    A_uv = Vu @ Uv - Dev @ Uv @ Dia @ Vu
    A_vu = Vv @ Uu - Deu @ Uu @ Dia @ Vv
    A_vv = k0 * k0 * Deu + Vv @ Uv + Deu @ Uu @ Dia @ Vu
    A = sp.bmat([[A_uu, A_uv], [A_vu, A_vv]], format="csc")
    nmax = float(np.sqrt(max(eu.max(), ev.max())))
    target = (k0 * (n_guess if n_guess is not None else nmax)) ** 2
    vals, vecs = spla.eigs(A, k=min(n_modes, A.shape[0] - 2), sigma=target, which="LM")
    order = np.argsort(-vals.real)
    bt2 = vals[order[mode_index]]
    hv = vecs[:, order[mode_index]]
    bt = np.sqrt(bt2.real)                                  # guided: real beta~
    beta = 2.0 / dl * np.arcsin(min(1.0, 0.5 * bt * dl))
    n = nu * nv
    h_u, h_v = hv[:n], hv[n:]
    iw = 1j * wt
    e_a = -(Vu @ h_v - Vv @ h_u) / (iw * cpml.EPSILON_0 * ea.ravel())
    e_u = (iw * cpml.MU_0 * h_v + Uu @ e_a) / (1j * bt)
    e_v = (Uv @ e_a - iw * cpml.MU_0 * h_u) / (1j * bt)
    h_a = (Uu @ e_v - Uv @ e_u) / (iw * cpml.MU_0)
    m = Mode(axis=axis, plane_index=plane_index, freq=freq, n_eff=float(beta / (w / C0)),
             beta=float(beta), dl_normal=dl,
             e_u=e_u.reshape(nu, nv), e_v=e_v.reshape(nu, nv), h_u=h_u.reshape(nu, nv),
             h_v=h_v.reshape(nu, nv), e_a=e_a.reshape(nu, nv), h_a=h_a.reshape(nu, nv))
    return _normalise(m, g)


def _face_weights(g: Grid, axis: int):
    """Face elements in the cyclic order: (dl_u * dd_v at E_u, dd_u * dl_v at E_v)."""
    u, v = cyclic(axis)
    au, av = g.axes[u], g.axes[v]
    return au.dl[:, None] * av.dl_dual[None, :], au.dl_dual[:, None] * av.dl[None, :]


def mode_power(m: Mode, g: Grid) -> float:
    """Flux of the forward mode as the plane monitors measure it:
    ``1/2 Re sum (E_u conj(c H_v) w_u - E_v conj(c H_u) w_v)`` with the colocation factor c."""
    wu, wv = _face_weights(g, m.axis)
    c = m.colocation
    return 0.5 * float(np.real(np.sum(m.e_u * np.conj(c * m.h_v) * wu
                                      - m.e_v * np.conj(c * m.h_u) * wv)))


def _normalise(m: Mode, g: Grid) -> Mode:
    p = mode_power(m, g)
    # fix the global phase: the largest transverse E component real and positive
    big = m.e_u if np.abs(m.e_u).max() >= np.abs(m.e_v).max() else m.e_v
    ph = np.exp(-1j * np.angle(big.ravel()[np.argmax(np.abs(big))]))
    s = ph / np.sqrt(abs(p))
    if p < 0:                                   # a mode found travelling the other way
        m.h_u, m.h_v, m.h_a = -m.h_u, -m.h_v, -m.h_a
    for name in ("e_u", "e_v", "h_u", "h_v", "e_a", "h_a"):
        setattr(m, name, getattr(m, name) * s)
    return m


# ------------------------------------------------------------------------------------------
# mode source
# ------------------------------------------------------------------------------------------

def mode_source(sc: Scene, m: Mode, source_time, direction: int = +1,
                num_steps: int | None = None):
    """A :class:`~openem_amd.model.ModeSource` launching ``m`` along ``direction`` from its
    plane, driven by ``source_time`` (complex ``amp_time``).

    The profiles are the incident field itself: E_u, E_v on the E plane ``ks`` and H_u, H_v on
    the H plane the correction updates (``ks - 1/2`` for +, ``ks + 1/2`` for -), i.e. the
    mode's H times ``direction`` (a backward mode's transverse H is reversed) and the half-cell
    phase ``exp(-i beta dl / 2)`` back to the E plane.  amp_e samples ``amp_time(n dt)`` (the H
    side reads E at whole steps), amp_h ``amp_time((n + 1/2) dt)``."""
    from .model import ModeSource
    n = int(num_steps if num_steps is not None else sc.num_time_steps)
    t = np.arange(n + 1, dtype=np.float64) * sc.dt
    half = np.exp(-0.5j * m.beta * m.dl_normal)
    return ModeSource(plane_index=m.plane_index, direction=int(direction),
                      ey_inc=m.e_u.copy(), ez_inc=m.e_v.copy(),
                      hy_inc=direction * half * m.h_u, hz_inc=direction * half * m.h_v,
                      amp_e=np.asarray(source_time.amp_time(t), dtype=np.complex128),
                      amp_h=np.asarray(source_time.amp_time(t + 0.5 * sc.dt), dtype=np.complex128),
                      n_eff=complex(m.n_eff), axis=m.axis)


#: face signs (H_u, H_v, E_u, E_v) of the mode-source corrections: the Levi-Civita signs of a
#: TF/SF face (``+``: the low face of the total-field side, ``-``: the high face).  The
#: reference's table (``openem/sources_setup.py:40-54``) keeps the E signs for ``-`` because
#: its backward profiles carry the forward H; here the profiles carry the incident H itself.
MODE_SIGNS = {+1: (-1.0, +1.0, +1.0, -1.0), -1: (+1.0, -1.0, -1.0, +1.0)}


def plane_cells(shape, axis: int, k: int) -> np.ndarray:
    """Flat cell indices ``(n_u, n_v)`` of the plane ``k`` normal to ``axis`` (cyclic order)."""
    u, v = cyclic(axis)
    idx = [None, None, None]
    idx[axis] = np.full((shape[u], shape[v]), k)
    idx[u], idx[v] = np.meshgrid(np.arange(shape[u]), np.arange(shape[v]), indexing="ij")
    return (idx[0] * shape[1] + idx[1]) * shape[2] + idx[2]


# ------------------------------------------------------------------------------------------
# overlaps with plane-monitor phasors
# ------------------------------------------------------------------------------------------

def monitor_fields(m: Mode):
    """The mode in a plane monitor's layout ``(E_t1, E_t2, H_t1, H_t2)`` with (t1, t2)
    ascending and H colocated as the monitor colocates it."""
    u, v = cyclic(m.axis)
    c = m.colocation
    comp = {u: (m.e_u, c * m.h_u), v: (m.e_v, c * m.h_v)}
    t1, t2 = transverse_axes(m.axis)
    tr = (lambda a: a) if u < v else (lambda a: a.T)       # cyclic (u, v) -> ascending layout
    return (tr(comp[t1][0]), tr(comp[t2][0]), tr(comp[t1][1]), tr(comp[t2][1]))


def overlap_weights(g: Grid, axis: int):
    """Face weights (w1, w2) of the ascending layout (the flux monitors' area elements)."""
    t1, t2 = transverse_axes(axis)
    a1, a2 = g.axes[t1], g.axes[t2]
    return a1.dl[:, None] * a2.dl_dual[None, :], a1.dl_dual[:, None] * a2.dl[None, :]


def amplitudes(data: np.ndarray, m: Mode, g: Grid, sign: float = 1.0):
    """Forward / backward amplitudes ``(a_plus, a_minus)`` (each ``(nf,)``) of the phasors
    ``data`` ``(4, nf, n1, n2)`` of a plane monitor in the mode ``m`` (unit power):
    ``a+- = (S1 +- S2) / 4`` with ``S1 = sum (E x conj(H_m)).n dA`` and
    ``S2 = sum (conj(E_m) x H).n dA``; ``sign`` is the normal orientation of the ascending
    pair (-1 for a y normal)."""
    e1m, e2m, h1m, h2m = monitor_fields(m)
    w1, w2 = overlap_weights(g, m.axis)
    e1, e2, h1, h2 = data[0], data[1], data[2], data[3]
    s1 = np.sum(e1 * np.conj(h2m)[None] * w1[None] - e2 * np.conj(h1m)[None] * w2[None], axis=(1, 2))
    s2 = np.sum(np.conj(e1m)[None] * h2 * w1[None] - np.conj(e2m)[None] * h1 * w2[None], axis=(1, 2))
    s1, s2 = sign * s1, sign * s2
    return 0.25 * (s1 + s2), 0.25 * (s1 - s2)


def overlap_weight_sets(ms: list, g: Grid) -> np.ndarray:
    """The two weight sets of :func:`amplitudes` for the device reduction
    (``oem_set_overlap_weights``): ``(2, 4, nf, n1, n2)`` complex, set 0 giving
    ``S1 = sum (E x conj(H_m)).n dA`` and set 1 ``S2 = sum (conj(E_m) x H).n dA`` (before the
    orientation sign)."""
    axis = ms[0].axis
    w1, w2 = overlap_weights(g, axis)
    n1, n2 = w1.shape
    out = np.zeros((2, 4, len(ms), n1, n2), np.complex128)
    for f, m in enumerate(ms):
        e1m, e2m, h1m, h2m = monitor_fields(m)
        out[0, 0, f] = np.conj(h2m) * w1
        out[0, 1, f] = -np.conj(h1m) * w2
        out[1, 2, f] = -np.conj(e2m) * w2
        out[1, 3, f] = np.conj(e1m) * w1
    return out


def orientation(axis: int) -> float:
    """+1 when the ascending transverse pair and the normal are right-handed (x, z normals),
    -1 for a y normal."""
    return -1.0 if axis == 1 else 1.0
