"""Phasors -> plane flux (reference ``openem/flux.py:60-110``, ``:269-327``, ``:413-513``).

Two conventions, as in the reference:

* in-place Yee (:func:`plane_flux`): ``P(f) = 1/2 Re sum[E_t1 conj(H_t2) dA_1 - E_t2 conj(H_t1)
  dA_2]`` with (t1, t2) ascending and the staggered dual face elements ``dA_1 = dl_t1 *
  dl_dual_t2``, ``dA_2 = dl_dual_t1 * dl_t2``;
* colocated (:func:`monitor_flux_colocated`, the reference's default readout for monitors that
  carry physical bounds): the four tangential components are interpolated from cell centres onto
  the primal edge points, and ``1/2 Re sum dS (E_u conj(H_v) - E_v conj(H_u))`` is taken with
  trapezoid face elements clipped exactly at the monitor bounds.

For a y normal the ascending order gives -S_y, which :func:`monitor_flux` negates.  The normal
colocation of H is done by the plane DFT already (the monitor stores H on the E plane).
"""

from __future__ import annotations

import os

import numpy as np

from .grid import Grid, transverse_axes

#: boundaries next to which edge samples may be dropped (reference ``openem/grid.py:42``)
ABSORBING = ("PML", "StablePML", "Absorber")


def area_weights(grid: Grid, axis: int = 2):
    t1, t2 = transverse_axes(axis)
    a1, a2 = grid.axes[t1], grid.axes[t2]
    return a1.dl[:, None] * a2.dl_dual[None, :], a1.dl_dual[:, None] * a2.dl[None, :]


def plane_flux(data: np.ndarray, grid: Grid, transverse=None, axis: int = 2) -> np.ndarray:
    dA_1, dA_2 = area_weights(grid, axis)
    if transverse is not None:
        mask = np.zeros(dA_1.shape)
        (j0, j1), (k0, k1) = transverse
        mask[j0:j1 + 1, k0:k1 + 1] = 1.0
        dA_1, dA_2 = dA_1 * mask, dA_2 * mask
    e1, e2, h1, h2 = data[0], data[1], data[2], data[3]
    t1 = (np.real(e1 * np.conj(h2)) * dA_1[None]).sum(axis=(1, 2))
    t2 = (np.real(e2 * np.conj(h1)) * dA_2[None]).sum(axis=(1, 2))
    return 0.5 * (t1 - t2)


def colocated_axis_weights(edges: np.ndarray, lo: float, hi: float):
    """``(i0, i1, w)``: the edge samples ``i0..i1`` (inclusive) whose integration cells meet
    ``[lo, hi]`` and their trapezoid weights (reference ``openem/flux.py:305-327``).

    The integration cell of edge ``i`` runs from the midpoint of its lower neighbour interval to
    the midpoint of its upper one (the grid ends close the first and last cells); its weight is
    the length of that cell inside ``[lo, hi]``.
    """
    e = np.asarray(edges, dtype=np.float64)
    mid = 0.5 * (e[:-1] + e[1:])
    cell_lo = np.empty_like(e)
    cell_hi = np.empty_like(e)
    cell_lo[0], cell_lo[1:] = e[0], mid
    cell_hi[-1], cell_hi[:-1] = e[-1], mid
    w = np.clip(np.minimum(cell_hi, hi) - np.maximum(cell_lo, lo), 0.0, None)
    hit = np.nonzero(w > 0.0)[0]
    if hit.size == 0:
        raise ValueError(f"monitor range [{lo}, {hi}] meets no integration cell on this axis")
    return int(hit[0]), int(hit[-1]), w[hit[0]:hit[-1] + 1]


def _edge_samples(ax, bounds, name: str, label: str):
    """Colocation of one transverse axis (reference ``openem/flux.py:395-477``).

    Returns ``(edge, c_lo, c_hi, w_lo, trap)``: the edge sample indices, the cell-centre indices
    below and above each sample, the weight of the lower centre in the centre -> edge
    interpolation, and the trapezoid weights.  A monitor spanning a purely periodic axis whole
    folds sample n onto sample 0 and takes the wrapped cell n - 1 (one period back) as the
    lower neighbour of sample 0.  Samples on the grid ends are dropped next to an absorbing
    boundary and refused otherwise.
    """
    n = ax.n
    i0, i1, w = colocated_axis_weights(ax.edges, float(bounds[0]), float(bounds[1]))
    wrap = (i0 == 0 and i1 == n and ax.boundary_lo == "Periodic"
            and ax.boundary_hi == "Periodic")
    if not wrap:
        for end, bad in (("lo", i0 < 1), ("hi", i1 > n - 1)):
            bnd = ax.boundary_lo if end == "lo" else ax.boundary_hi
            if bad and bnd not in ABSORBING:
                raise ValueError(
                    f"monitor {name}: samples along {label} reach the grid {end} end and that end "
                    f"({bnd}) is not absorbing")
    edges = np.asarray(ax.edges, dtype=np.float64)
    cen = np.asarray(ax.centers, dtype=np.float64)
    if wrap:
        w = np.concatenate([[w[0] + w[-1]], w[1:-1]])
        edge = np.arange(0, n)
        c_lo = (edge - 1) % n
        x_lo = np.where(edge == 0, cen[-1] - (edges[-1] - edges[0]), cen[c_lo])
    else:
        if i0 < 1:
            w, i0 = w[1 - i0:], 1
        if i1 > n - 1:
            w, i1 = w[:n - 1 - i1], n - 1
        edge = np.arange(i0, i1 + 1)
        c_lo = edge - 1
        x_lo = cen[c_lo]
    c_hi = edge
    w_lo = (cen[c_hi] - edges[edge]) / (cen[c_hi] - x_lo)
    return edge, c_lo, c_hi, w_lo, w


def monitor_flux_colocated(data: np.ndarray, grid: Grid, mon) -> np.ndarray:
    """Plane flux ``(nf,)`` in the colocated convention (reference
    ``openem/flux.py:228-264``, the tidy3d ``FieldData.flux`` convention), as +S along the
    monitor axis.  ``data`` is the whole plane ``(4, nf, n1, n2)``; ``mon.t_bounds`` is needed.

    Along t1, E_u and H_v sit at cell centres (interpolated onto the edges) while E_v and H_u
    sit on the edges; along t2 the other way round.
    """
    if getattr(mon, "t_bounds", None) is None:
        raise ValueError(f"monitor {mon.name} has no physical bounds")
    t1, t2 = transverse_axes(mon.axis)
    a1, a2 = grid.axes[t1], grid.axes[t2]
    if data.shape[2:] != (a1.n, a2.n):
        raise ValueError(f"monitor {mon.name}: colocation needs the whole plane")
    # H is colocated along the normal from the cells on both sides of the plane (reference
    # ``openem/flux.py:334-365``): a plane on the grid end is refused
    if not 1 <= mon.plane_index <= grid.axes[mon.axis].n - 1:
        raise ValueError(f"monitor {mon.name}: plane {mon.plane_index} sits on the grid end")
    e1, l1, h1, w1, tr1 = _edge_samples(a1, mon.t_bounds[0], mon.name, "xyz"[t1])
    e2, l2, h2, w2, tr2 = _edge_samples(a2, mon.t_bounds[1], mon.name, "xyz"[t2])

    def centre_to_edge_1(a):          # along t1, then the edge samples along t2
        return (w1[:, None] * a[:, l1][:, :, e2] + (1.0 - w1)[:, None] * a[:, h1][:, :, e2])

    def centre_to_edge_2(a):          # the edge samples along t1, then along t2
        b = a[:, e1]
        return w2[None, :] * b[:, :, l2] + (1.0 - w2)[None, :] * b[:, :, h2]

    eu, hv = centre_to_edge_1(data[0]), centre_to_edge_1(data[3])
    ev, hu = centre_to_edge_2(data[1]), centre_to_edge_2(data[2])
    dS = np.outer(tr1, tr2)
    sign = -1.0 if mon.axis == 1 else 1.0
    return sign * 0.5 * np.real(((eu * np.conj(hv) - ev * np.conj(hu)) * dS[None])
                                .sum(axis=(1, 2)))


def flux_quadrature() -> str:
    """``"colocated"`` (default) or ``"yee"`` (``OPENEM_FLUX_COLOCATED=0``), the reference's
    switch (``openem/flux.py:261-266``)."""
    return "yee" if os.environ.get("OPENEM_FLUX_COLOCATED", "1") == "0" else "colocated"


def monitor_flux(data: np.ndarray, grid: Grid, mon) -> np.ndarray:
    """Flux of a :class:`~openem_amd.model.FluxMonitor` as +S along its axis (reference
    ``openem/flux.py:269-302``): colocated when the monitor has physical bounds and the data is
    the whole plane, falling back to the in-place Yee sum where the colocation is refused."""
    if (flux_quadrature() == "colocated" and getattr(mon, "t_bounds", None) is not None
            and data.ndim == 4):
        t1, t2 = transverse_axes(mon.axis)
        if data.shape[2:] == (grid.axes[t1].n, grid.axes[t2].n):
            try:
                return monitor_flux_colocated(data, grid, mon)
            except ValueError as err:
                print(f"[openem_amd] monitor {mon.name}: colocated flux refused ({err}); "
                      "using the in-place Yee sum")
    sign = -1.0 if mon.axis == 1 else 1.0
    return sign * plane_flux(data, grid, transverse=mon.transverse, axis=mon.axis)


def normal_planes(grid: Grid, mon):
    """The one or two primal planes of a flux-time monitor (reference ``openem/flux.py:334-365``):
    ``(w, ke, kh, wh_lo, wh_hi)`` per plane, E taken on layer ``ke`` and H interpolated by
    coordinate onto ``edges[ke]`` from the cell centres ``kh = ke - 1`` and ``ke``."""
    ax = grid.axes[mon.axis]
    e, c = np.asarray(ax.edges, np.float64), np.asarray(ax.centers, np.float64)
    km, fr = int(mon.plane_index), float(mon.frac)
    planes = []
    for w, k in ((1.0 - fr, km), (fr, km + 1)):
        if w == 0.0:
            continue
        kh = k - 1
        if not 0 <= kh < c.size - 1:
            raise ValueError(f"monitor {mon.name}: plane {k} on the grid end cannot colocate H")
        lo = (c[kh + 1] - e[k]) / (c[kh + 1] - c[kh])
        planes.append((w, k, kh, lo, 1.0 - lo))
    return planes


def flux_time_geometry(grid: Grid, mon) -> dict:
    """Everything the instantaneous colocated flux needs (reference
    ``openem/flux.py:480-513``): per transverse axis the edge samples with their lower / upper
    cell centres, the centre -> edge weight of the lower one and the trapezoid weights; the
    normal planes; the sign (``normal_dir``, and -1 for a y normal's ascending-pair order)."""
    t1, t2 = transverse_axes(mon.axis)
    s1 = _edge_samples(grid.axes[t1], mon.t_bounds[0], mon.name, "xyz"[t1])
    s2 = _edge_samples(grid.axes[t2], mon.t_bounds[1], mon.name, "xyz"[t2])
    sign = float(mon.normal_dir) * (-1.0 if mon.axis == 1 else 1.0)
    return {"axis": mon.axis, "tax": (t1, t2), "s": (s1, s2), "planes": normal_planes(grid, mon),
            "sign": sign}


def colocated_planes(E, H, g: dict, plane: int):
    """``(eu, ev, hu, hv)`` on the edge samples of one normal plane from 3D fields ``E``, ``H``
    (component-first arrays): E_u and H_v interpolated along t1, E_v and H_u along t2, H along
    the normal by the plane's coordinate weights."""
    a = g["axis"]
    t1, t2 = g["tax"]
    (e1, l1, h1, w1, _), (e2, l2, h2, w2, _) = g["s"]
    _, ke, kh, wlo, whi = g["planes"][plane]

    def layer(F, k):
        return np.take(F, k, axis=a)          # (n_t1, n_t2): the remaining axes stay ascending

    def along1(P):
        return w1[:, None] * P[l1][:, e2] + (1.0 - w1)[:, None] * P[h1][:, e2]

    def along2(P):
        Q = P[e1]
        return w2[None, :] * Q[:, l2] + (1.0 - w2)[None, :] * Q[:, h2]

    eu = along1(layer(E[t1], ke))
    ev = along2(layer(E[t2], ke))
    hu = along2(wlo * layer(H[t1], kh) + whi * layer(H[t1], kh + 1))
    hv = along1(wlo * layer(H[t2], kh) + whi * layer(H[t2], kh + 1))
    return eu, ev, hu, hv


def flux_time_sample(E, H, g: dict) -> float:
    """Instantaneous colocated flux ``sum_p w_p sum dS (E_u H_v - E_v H_u)`` of fields ``E``,
    ``H`` (reference ``openem/flux.py:566-580``)."""
    (_, _, _, _, tr1), (_, _, _, _, tr2) = g["s"]
    dS = g["sign"] * np.outer(tr1, tr2)
    out = 0.0
    for p, (w, *_rest) in enumerate(g["planes"]):
        eu, ev, hu, hv = colocated_planes(E, H, g, p)
        out += w * float(np.sum((eu * hv - ev * hu) * dS))
    return out
