"""TF/SF box: the 1D auxiliary line and the face-correction point lists.

The reference drives a TF/SF box from a 1D auxiliary Yee line that uses the same discrete
operator as the 3D grid (``openem/tfsf1d.py:1-30``, ``:297-330``), and applies face corrections
``F += (cb*) coef * incident`` on the box faces (``openem/kernels/tfsf.cu:34-83``).

Here the line carries all six components as complex amplitudes of a field
``f(z, t) * exp(i (kx x + ky y))`` with (kx, ky) the fixed transverse wave vector.  On a uniform
x/y grid every 3D difference of such a field reduces exactly to a multiplication by
``K = 2i sin(k dl / 2) / dl`` at the component's own Yee position, so the line's field solves the
3D discrete equations exactly (normal incidence: K = 0, the reference's line; oblique incidence
on a Bloch cell: the Bloch phase of the 3D grid, ``exp(i kx Lx)`` at the wrap, is the same
phase).  The line is driven by a tangential current sheet below (+z travel) or above (-z) the
box; the sheet's back-going wave runs into the runway and never reaches the box.  Termination is
a runway long enough that nothing returns within the run, as in the reference
(``openem/tfsf1d.py:143-148``).

Face corrections (derived from the curl with the Levi-Civita sign e_cab, for a face normal to
axis ``a`` at edge index ``p`` of the box, component ``c`` tangential, ``b`` the third axis):

    lo face:  H_c[p-1] += ch * e_cab * E_b^inc[p] / dl_a[p-1]
              E_c[p]   -= cb * e_cab * H_b^inc[p-1] / dl_dual_a[p]
    hi face:  H_c[p]   -= ch * e_cab * E_b^inc[p] / dl_a[p]
              E_c[p]   += cb * e_cab * H_b^inc[p] / dl_dual_a[p]

which reproduce the reference's one-face injection signs (``openem/sources_setup.py:93-119``)
on a z face.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import cpml

EPS3 = {(0, 1, 2): 1, (1, 2, 0): 1, (2, 0, 1): 1, (0, 2, 1): -1, (2, 1, 0): -1, (1, 0, 2): -1}
RUNWAY_MARGIN = 1.05
RUNWAY_PAD = 64

# line components: 0 ex, 1 ey, 2 ez, 3 hx, 4 hy, 5 hz; on z: ex ey hz at edges, ez hx hy at centres
LINE_ON_EDGE = (True, True, False, False, False, True)


@dataclass
class TFSFTables:
    """Point lists for the H-side and E-side corrections plus the incident tables."""

    h_comp: np.ndarray
    h_cell: np.ndarray
    h_coef: np.ndarray
    h_col: np.ndarray
    h_phase: np.ndarray          # complex64
    e_comp: np.ndarray
    e_cell: np.ndarray
    e_coef: np.ndarray
    e_col: np.ndarray
    e_phase: np.ndarray
    cols: list                    # (line component, z index) per column
    tab_h: np.ndarray             # (n_total+1, ncol) complex128: E^inc at t = n dt
    tab_e: np.ndarray             # (n_total+1, ncol) complex128: H^inc at t = (n+1/2) dt
    src: object = None            # the TFSFSource (box geometry, kxy)


def k_factor(k: float, dl: float) -> complex:
    return 2j * np.sin(k * dl / 2.0) / dl


def run_line(dz: np.ndarray, eps_edge: np.ndarray, eps_ctr: np.ndarray, dt: float,
             kx: float, ky: float, dx: float, dy: float, src_edge: int, jxy, amp_half: np.ndarray,
             n_steps: int, cols: list, sigma_edge=None, sigma_ctr=None):
    """Run the complex 1D line over ``dz`` (the 3D z spacing) extended by runways at both ends.

    Returns ``(tab_h, tab_e)``: the requested columns of E at t = n dt (rows 0..n_steps) and of
    H at t = (n+1/2) dt (rows 0..n_steps-1, row n_steps repeated).  ``cols`` index the 3D z
    grid (edges for ex/ey/hz, centres for ez/hx/hy).
    """
    nz = dz.size
    c0 = cpml.C_0
    run = int(np.ceil(RUNWAY_MARGIN * n_steps * c0 * dt / float(min(dz[0], dz[-1])))) + RUNWAY_PAD
    dzl = np.concatenate([np.full(run, dz[0]), dz, np.full(run, dz[-1])])
    n1 = dzl.size
    off = run
    el = np.concatenate([np.full(run, eps_edge[0]), eps_edge, np.full(run, eps_edge[-1])])
    ec = np.concatenate([np.full(run, eps_ctr[0]), eps_ctr, np.full(run, eps_ctr[-1])])
    se = np.zeros(n1) if sigma_edge is None else np.concatenate(
        [np.full(run, sigma_edge[0]), sigma_edge, np.full(run, sigma_edge[-1])])
    sc = np.zeros(n1) if sigma_ctr is None else np.concatenate(
        [np.full(run, sigma_ctr[0]), sigma_ctr, np.full(run, sigma_ctr[-1])])
    dzd = 0.5 * (dzl + np.concatenate([[dzl[0]], dzl[:-1]]))
    ch = dt / cpml.MU_0
    s_e = se * dt / (2 * cpml.EPSILON_0)
    s_c = sc * dt / (2 * cpml.EPSILON_0)
    ca_e, cb_e = (el - s_e) / (el + s_e), (dt / cpml.EPSILON_0) / (el + s_e)
    ca_c, cb_c = (ec - s_c) / (ec + s_c), (dt / cpml.EPSILON_0) / (ec + s_c)
    Kx, Ky = k_factor(kx, dx), k_factor(ky, dy)
    f = np.zeros((6, n1), dtype=np.complex128)
    ex, ey, ez, hx, hy, hz = f
    idz = 1.0 / dzl
    idzd = 1.0 / dzd
    q = off + src_edge
    jx, jy = jxy
    ncol = len(cols)
    comp = np.array([c for c, _ in cols], dtype=np.int64)
    pos = np.array([off + z for _, z in cols], dtype=np.int64)
    tab_h = np.zeros((n_steps + 1, ncol), dtype=np.complex128)
    tab_e = np.zeros((n_steps + 1, ncol), dtype=np.complex128)
    for n in range(n_steps):
        tab_h[n] = f[comp, pos]                         # E^n (and stale H, unused)
        # H half step
        hx[:-1] -= ch * (Ky * ez[:-1] - (ey[1:] - ey[:-1]) * idz[:-1])
        hy[:-1] -= ch * ((ex[1:] - ex[:-1]) * idz[:-1] - Kx * ez[:-1])
        hz[:] -= ch * (Kx * ey - Ky * ex)
        tab_e[n] = f[comp, pos]                         # H^{n+1/2}
        # E full step
        ex[1:] = ca_e[1:] * ex[1:] + cb_e[1:] * (Ky * hz[1:] - (hy[1:] - hy[:-1]) * idzd[1:])
        ey[1:] = ca_e[1:] * ey[1:] + cb_e[1:] * ((hx[1:] - hx[:-1]) * idzd[1:] - Kx * hz[1:])
        ez[:] = ca_c * ez + cb_c * (Kx * hy - Ky * hx)
        a = amp_half[n]
        ex[q] -= cb_e[q] * jx * a * idzd[q]
        ey[q] -= cb_e[q] * jy * a * idzd[q]
    tab_h[n_steps] = f[comp, pos]
    tab_e[n_steps] = tab_e[n_steps - 1]
    tail = max(np.abs(f[:, :8]).max(), np.abs(f[:, -8:]).max())
    peak = max(np.abs(tab_h).max(), 1e-300)
    if tail > 1e-30 * peak:
        raise ValueError(f"TF/SF line runway too short: end value {tail:.3e} vs peak {peak:.3e}")
    return tab_h, tab_e


def _ranges(box_lo, box_hi, shape, open_axes, axis, on_edge):
    if axis in open_axes:
        return range(0, shape[axis])
    lo, hi = box_lo[axis], box_hi[axis]
    return range(lo, hi + 1) if on_edge else range(lo, hi)


def build_points(grid, src, shape):
    """Face correction points; returns (H list, E list) of dicts and the column list."""
    lo, hi, opn = src.box_lo, src.box_hi, tuple(src.open_axes)
    ny, nz = shape[1], shape[2]
    axes = grid.axes
    kx, ky = src.kxy
    h_pts, e_pts, cols, colmap = [], [], [], {}

    def col_of(comp, z):
        key = (comp, z)
        if key not in colmap:
            colmap[key] = len(cols)
            cols.append(key)
        return colmap[key]

    def coords(axis, idx, on_edge):
        a = axes[axis]
        return (a.edges[:-1] if on_edge else a.centers)[idx]

    for a in range(3):
        if a in opn:
            continue
        for side in ("lo", "hi"):
            p = lo[a] if side == "lo" else hi[a]
            for c in range(3):
                if c == a:
                    continue
                b = 3 - a - c
                sgn = EPS3[(c, a, b)]
                # ---- H_c on the H plane: edge along c, centre along b
                hplane = p - 1 if side == "lo" else p
                coef_h = (sgn if side == "lo" else -sgn) / axes[a].dl[hplane]
                # ---- E_c on the E plane: centre along c, edge along b
                coef_e = (-sgn if side == "lo" else sgn) / axes[a].dl_dual[p]
                hplane_inc = p - 1 if side == "lo" else p
                for kind in ("H", "E"):
                    c_edge = kind == "H"
                    rc = _ranges(lo, hi, shape, opn, c, c_edge)
                    rb = _ranges(lo, hi, shape, opn, b, not c_edge)
                    idx = [None, None, None]
                    idx[a] = (hplane if kind == "H" else p)
                    for ic in rc:
                        for ib in rb:
                            idx[c], idx[b] = ic, ib
                            i, j, k = idx
                            cell = (i * ny + j) * nz + k
                            # incident component and its z position on the line
                            if kind == "H":
                                comp_inc = b                    # E_b
                                on_edge_z = LINE_ON_EDGE[b]
                                zpos = p if a == 2 else (ib if b == 2 else ic)
                            else:
                                comp_inc = 3 + b                # H_b
                                on_edge_z = LINE_ON_EDGE[3 + b]
                                zpos = hplane_inc if a == 2 else (ib if b == 2 else ic)
                            del on_edge_z
                            # transverse phase at the incident component's Yee position
                            if kind == "H":
                                xe = (True if c == 0 else (False if b == 0 else None))
                                ye = (True if c == 1 else (False if b == 1 else None))
                            else:
                                xe = (False if c == 0 else (True if b == 0 else None))
                                ye = (False if c == 1 else (True if b == 1 else None))
                            x = coords(0, i, xe) if xe is not None else coords(0, i, True)
                            y = coords(1, j, ye) if ye is not None else coords(1, j, True)
                            ph = np.exp(1j * (kx * x + ky * y))
                            rec = (c, cell, coef_h if kind == "H" else coef_e,
                                   col_of(comp_inc, zpos), ph)
                            (h_pts if kind == "H" else e_pts).append(rec)
    return h_pts, e_pts, cols


def _line_setup(sc, src):
    """The line's background column, source edge and transverse wave vector of ``src``."""
    i0 = src.box_lo[0] if 0 not in src.open_axes else 0
    j0 = src.box_lo[1] if 1 not in src.open_axes else 0
    eps_edge = np.asarray(sc.eps_ex[i0, j0, :], dtype=np.float64)
    eps_ctr = np.asarray(sc.eps_ez[i0, j0, :], dtype=np.float64)
    if src.eps_line is not None:
        eps_edge, eps_ctr = src.eps_line
    k_src = src.box_lo[2] - 3 if src.direction > 0 else src.box_hi[2] + 3
    return eps_edge, eps_ctr, k_src


def incident_at_face(sc, src, n_steps: int) -> np.ndarray:
    """``(n_steps + 1, 3)`` complex: the incident E (x, y, z) at t = n dt on the box's injection
    face (the lower z face for +z travel, the upper for -z), from the same line that drives the
    box, at the transverse origin.  Ex, Ey sit on the face's edge plane; Ez, which the line
    keeps at cell centres, is taken at the centre just inside the box (for a plane wave only
    its magnitude enters the normalisation)."""
    grid = sc.grid
    eps_edge, eps_ctr, k_src = _line_setup(sc, src)
    zf = src.box_lo[2] if src.direction > 0 else src.box_hi[2]
    zc = zf if src.direction > 0 else zf - 1
    cols = [(0, zf), (1, zf), (2, zc)]
    amp = src.waveform.amp_half
    if amp.size < n_steps + 1:
        raise ValueError("TF/SF waveform table shorter than the run")
    tab_h, _ = run_line(grid.z.dl, eps_edge, eps_ctr, sc.dt, src.kxy[0], src.kxy[1],
                        float(grid.x.dl[0]), float(grid.y.dl[0]), k_src, src.jxy,
                        amp * src.amp_scale, n_steps, cols)
    return tab_h


def tables(sc, src, n_total: int) -> TFSFTables:
    """Build the point lists and run the line for one :class:`~openem_amd.model.TFSFSource`."""
    grid = sc.grid
    shape = sc.shape
    if src.axis != 2:
        raise NotImplementedError("TF/SF line along z only")
    for a in range(3):
        if a in src.open_axes:
            if not grid.axes[a].is_periodic:
                raise ValueError(f"open axis {a} must be periodic")
            continue
        lo, hi = src.box_lo[a], src.box_hi[a]
        w_lo = sc.pml.get((a, "lo"))
        w_hi = sc.pml.get((a, "hi"))
        nlo = w_lo.num_layers if w_lo else 0
        nhi = w_hi.num_layers if w_hi else 0
        if lo - 1 < nlo or hi + 1 > grid.axes[a].n - nhi or lo >= hi:
            raise ValueError(f"TF/SF box faces on axis {a} ({lo}, {hi}) touch the PML")
    kx, ky = src.kxy
    if (kx or ky):
        if 0 not in src.open_axes or 1 not in src.open_axes:
            raise NotImplementedError("oblique TF/SF needs x and y open (a Bloch cell)")
        for a in (0, 1):
            d = grid.axes[a].dl
            if np.ptp(d) > 1e-12 * d.mean():
                raise NotImplementedError("oblique TF/SF needs a uniform x/y grid")
    h_pts, e_pts, cols = build_points(grid, src, shape)
    # background column at the box corner (the line's medium)
    eps_edge, eps_ctr, k_src = _line_setup(sc, src)
    amp = src.waveform.amp_half
    if amp.size < n_total + 1:
        raise ValueError("TF/SF waveform table shorter than the run")
    tab_h, tab_e = run_line(grid.z.dl, eps_edge, eps_ctr, sc.dt, kx, ky,
                            float(grid.x.dl[0]), float(grid.y.dl[0]), k_src, src.jxy,
                            amp * src.amp_scale, n_total, cols)

    def arr(pts):
        if not pts:
            z = np.zeros(0)
            return z.astype(np.int32), z.astype(np.int32), z, z.astype(np.int32), z.astype(complex)
        c, cell, coef, col, ph = zip(*pts)
        return (np.array(c, np.int32), np.array(cell, np.int32), np.array(coef),
                np.array(col, np.int32), np.array(ph, np.complex128))

    hc, hcell, hcoef, hcol, hph = arr(h_pts)
    ec, ecell, ecoef, ecol, eph = arr(e_pts)
    return TFSFTables(hc, hcell, hcoef, hcol, hph, ec, ecell, ecoef, ecol, eph, cols, tab_h, tab_e,
                      src=src)
