"""Synthetic scene builders for the benchmark workloads and the parity tests.

The reference builds a ``Scene`` from a tidy3d ``Simulation`` (``openem/scene/build.py:224``);
tidy3d is not available, so these builders produce the same ``Scene`` contract directly from
numpy, the pattern the reference's own tests use for tidy3d-free scenes
(``tests/test_bloch.py:42-99``, ``tests/test_fold.py:74-235``):

- :func:`slab_scene` - 0.6x0.6x3.0 um slab transmission (dl = 20 nm, periodic x/y, 12-layer
  PML on z outside the domain, eps=4 slab of 0.3 um, +z plane wave, flux planes);
- :func:`box_soup_scene` - N^3 grid with 12-layer CPML on all faces (inside the N^3), 32 seeded
  axis-aligned boxes (plain eps or 2-pole Lorentz), a centred Ez dipole, a field monitor on the
  central xy plane;
- :func:`mini_scene` - small grids mixing every feature for oracle parity.
"""

from __future__ import annotations

import numpy as np

from . import cpml, waveform
from .grid import Axis, Grid, nearest_edge_index, uniform_axis
from .model import (Dispersion, FieldMonitor, FluxMonitor, ModeMonitor, PlaneWaveSource,
                    PointDipole, Scene,
                    lorentz_poles, trap_coeffs)

UM = 1e-6


def num_steps_for(run_time: float, dt: float) -> int:
    """Step count covering ``run_time``: ``floor(run_time/dt) + 1``."""
    return int(np.floor(run_time / dt + 1e-9)) + 1


def incident_tables(source_time, dt: float, num_steps: int, dl_half: float, eps_r: float,
                    direction: int = +1, sigma: float = 0.0, freq0: float | None = None):
    """The two TF/SF injection tables (reference ``openem/scene/sources.py:19-75``).

    ``ex_inc[n] = Re a(n dt)``; ``hy_inc[n] = dir * Re(a((n+1/2)dt + dl_half/v) * n_c/eta0 * g)``
    where ``n_c = sqrt(eps_r + i sigma/(w0 eps0))``, ``v = c/Re(n_c)`` and ``g`` the half-cell
    gain ``exp(Im(n_c) w0 dl_half / c)`` of a lossy medium.
    """
    eps_c = complex(eps_r)
    if sigma:
        if freq0 is None:
            raise ValueError("a lossy source plane needs freq0")
        eps_c += 1j * sigma / (2 * np.pi * freq0 * cpml.EPSILON_0)
    n_c = np.sqrt(eps_c)
    v = cpml.C_0 / n_c.real
    n = np.arange(num_steps + 1, dtype=np.float64)
    ex_inc = np.real(np.asarray(source_time.amp_time(n * dt)))
    t_h = (n + 0.5) * dt + dl_half / v
    gain = 1.0
    if sigma:
        gain = float(np.exp(n_c.imag * 2 * np.pi * freq0 * dl_half / cpml.C_0))
    hy_inc = direction * np.real(np.asarray(source_time.amp_time(t_h)) * (n_c / cpml.ETA_0 * gain))
    return np.ascontiguousarray(ex_inc), np.ascontiguousarray(hy_inc)


def plane_wave(grid: Grid, eps_pol: np.ndarray, dt: float, num_steps: int, z: float,
               direction: int, pol_axis: int, source_time) -> PlaneWaveSource:
    ks = nearest_edge_index(grid.z.edges, z)
    kh = ks - 1 if direction > 0 else ks
    ex, hy = incident_tables(source_time, dt, num_steps, float(grid.z.dl[kh]) / 2.0,
                             float(eps_pol[0, 0, ks]), direction)
    return PlaneWaveSource(axis=2, plane_index=ks, direction=direction, pol_axis=pol_axis,
                           waveform=waveform.sample(source_time, dt, num_steps),
                           ex_inc=ex, hy_inc=hy)


def _interval_fill(lo_c: np.ndarray, hi_c: np.ndarray, a: float, b: float) -> np.ndarray:
    """Fraction of each interval ``[lo_c, hi_c]`` covered by ``[a, b]``."""
    ov = np.clip(np.minimum(hi_c, b) - np.maximum(lo_c, a), 0.0, None)
    return ov / (hi_c - lo_c)


def layer_eps_z(grid: Grid, z0: float, z1: float, eps_in: float, eps_out: float = 1.0):
    """``(eps_ex, eps_ey, eps_ez)`` of a z-layered slab ``[z0, z1]`` with exact 1D averaging:
    tangential components (on z edges) take the arithmetic mean over their dual interval, the
    normal component (at z centres) the harmonic mean over its primal cell."""
    e = grid.z.edges
    dl = grid.z.dl
    lo_d = np.concatenate([[e[0] - 0.5 * dl[0]], e[1:-1] - 0.5 * dl[:-1]])
    hi_d = e[:-1] + 0.5 * dl
    f_t = _interval_fill(lo_d, hi_d, z0, z1)
    f_n = _interval_fill(e[:-1], e[1:], z0, z1)
    eps_t = eps_out + (eps_in - eps_out) * f_t
    eps_n = 1.0 / ((1.0 - f_n) / eps_out + f_n / eps_in)
    shp = grid.shape
    full = lambda v: np.broadcast_to(v[None, None, :], shp).copy()  # noqa: E731
    return full(eps_t), full(eps_t), full(eps_n)


def pml_faces(dt: float, dls, axes_sides, n_layers: int = 12) -> dict:
    p = cpml.PMLParams(num_layers=n_layers)
    return {(a, s): cpml.face_coeffs(p, dls[a], s, dt) for a, s in axes_sides}


def slab_scene(dl: float = 0.02 * UM, size=(0.6 * UM, 0.6 * UM, 3.0 * UM), eps_slab: float = 4.0,
               thick: float = 0.3 * UM, f0: float = 2.998e14, fwidth_rel: float = 0.2,
               n_pml: int = 12, src_z: float = -1.0 * UM, t_z: float = 1.0 * UM,
               r_z: float | None = -1.2 * UM, nf: int = 11, run_time: float = 4e-13,
               num_steps: int | None = None, pol_axis: int = 0,
               direction: int = +1, homogeneous: bool = False) -> Scene:
    """Slab transmission scene. ``homogeneous=True`` gives the no-slab reference run."""
    nx, ny, nz = (int(round(s / dl)) for s in size)
    ax = uniform_axis(nx, dl, -size[0] / 2, "Periodic")
    ay = uniform_axis(ny, dl, -size[1] / 2, "Periodic")
    az = uniform_axis(nz + 2 * n_pml, dl, -size[2] / 2 - n_pml * dl, "PML")
    grid = Grid(ax, ay, az)
    dt = cpml.courant_dt((dl, dl, dl))
    n_steps = num_steps if num_steps is not None else num_steps_for(run_time, dt)
    if homogeneous:
        ones = np.ones(grid.shape)
        ex, ey, ez = ones, ones.copy(), ones.copy()
    else:
        ex, ey, ez = layer_eps_z(grid, -thick / 2, thick / 2, eps_slab)
    pulse = waveform.GaussianPulse(freq0=f0, fwidth=fwidth_rel * f0)
    src = plane_wave(grid, ex if pol_axis == 0 else ey, dt, n_steps, src_z, direction,
                     pol_axis, pulse)
    freqs = np.linspace(f0 * 0.9, f0 * 1.1, nf)
    # infinite monitors: their physical bounds are the whole periodic cross-section
    tb = ((float(ax.edges[0]), float(ax.edges[-1])), (float(ay.edges[0]), float(ay.edges[-1])))
    mons = [FluxMonitor("T", 2, nearest_edge_index(az.edges, t_z), +1, freqs, t_bounds=tb)]
    if r_z is not None:
        mons.append(FluxMonitor("R", 2, nearest_edge_index(az.edges, r_z), -1, freqs,
                                t_bounds=tb))
    return Scene(grid=grid, dt=dt, num_time_steps=n_steps, eps_ex=ex, eps_ey=ey, eps_ez=ez,
                 pml=pml_faces(dt, (dl, dl, dl), [(2, "lo"), (2, "hi")], n_pml),
                 sources=[src], flux_monitors=mons)


def dipole_stencil(grid: Grid, comp: int, pos, magnetic: bool = False):
    """Multilinear stencil of a point current at ``pos`` on component ``comp``'s Yee points:
    ``(indices (nw,3), weights/dV (nw,))`` (reference ``openem/model.py:198-229``)."""
    coords = grid.yee_coords(comp, magnetic)
    per_axis = []
    for a in range(3):
        c = coords[a]
        x = float(pos[a])
        i = int(np.argmin(np.abs(c - x)))
        if abs(float(c[i]) - x) <= 1e-12 * max(abs(x), 1e-9):
            per_axis.append([(i, 1.0)])
            continue
        j = i + 1 if c[i] < x else i - 1
        w = (x - float(c[i])) / (float(c[j]) - float(c[i]))
        per_axis.append([(i, 1.0 - w), (j, w)])
    idx, coef = [], []
    for i, wi in per_axis[0]:
        for j, wj in per_axis[1]:
            for k, wk in per_axis[2]:
                idx.append((i, j, k))
                coef.append(wi * wj * wk / grid.dual_volume_at(comp, magnetic, i, j, k))
    return np.array(idx, dtype=np.int64), np.array(coef, dtype=np.float64)


def box_soup_scene(n: int = 512, dl: float = 20e-9, n_pml: int = 12, n_boxes: int = 32,
                   seed: int = 0, f0: float = 2e14, fwidth_rel: float = 0.5,
                   num_steps: int = 4000, nf: int = 8, moment: float = 1e-20,
                   shape: tuple | None = None) -> Scene:
    """The roofline workload: ``shape`` (default ``(n, n, n)``) cells including the CPML on all
    six faces, 32 seeded boxes (half plain eps in [1, 12], half 2-pole Lorentz with
    eps_inf in [1, 4], w0 in [1e14, 8e14] rad/s, delta = 1e13 rad/s, de in [0.5, 3]), an Ez
    point dipole at the centre and a field monitor (2 z layers) on the central xy plane."""
    nx, ny, nz = shape if shape is not None else (n, n, n)
    rng = np.random.default_rng(seed)
    axes = [uniform_axis(m, dl, 0.0, "PML") for m in (nx, ny, nz)]
    grid = Grid(*axes)
    dt = cpml.courant_dt((dl, dl, dl))
    eps = np.ones((nx, ny, nz))
    mat = np.full((nx, ny, nz), -1, dtype=np.int32)      # Lorentz material id per cell
    lorentz = []
    lo_in, hi_in = n_pml, (nx - n_pml, ny - n_pml, nz - n_pml)
    for _ in range(n_boxes):
        sides = rng.integers(4, 65, size=3)
        cen = [int(rng.integers(lo_in, hi_in[a])) for a in range(3)]
        lo = [max(lo_in, cen[a] - int(sides[a]) // 2) for a in range(3)]
        hi = [min(hi_in[a], lo[a] + int(sides[a])) for a in range(3)]
        sl = tuple(slice(lo[a], hi[a]) for a in range(3))
        if rng.random() < 0.5:
            eps[sl] = rng.uniform(1.0, 12.0)
            mat[sl] = -1
        else:
            eps_inf = rng.uniform(1.0, 4.0)
            poles = []
            for _p in range(2):
                poles += lorentz_poles(rng.uniform(0.5, 3.0), rng.uniform(1e14, 8e14), 1e13)
            eps[sl] = eps_inf
            mat[sl] = len(lorentz)
            lorentz.append(poles)
    disp = _staircase_dispersion(mat, lorentz, dt) if lorentz else None
    if disp is not None and disp.n_entry == 0:
        disp = None
    pulse = waveform.GaussianPulse(freq0=f0, fwidth=fwidth_rel * f0)
    wf = waveform.sample(pulse, dt, num_steps)
    centre = [0.5 * (a.edges[0] + a.edges[-1]) for a in axes]
    idx, coef = dipole_stencil(grid, 2, centre)
    dip = PointDipole(component=2, indices=idx, coef=coef * moment, waveform=wf)
    kc = nz // 2 - 1
    freqs = np.linspace(f0 * (1 - fwidth_rel), f0 * (1 + fwidth_rel), nf)
    fm = FieldMonitor("xy_center", freqs, (0, 0, kc), (nx, ny, 2))
    sides_all = [(a, s) for a in range(3) for s in ("lo", "hi")]
    return Scene(grid=grid, dt=dt, num_time_steps=num_steps, eps_ex=eps, eps_ey=eps, eps_ez=eps,
                 pml=pml_faces(dt, (dl, dl, dl), sides_all, n_pml), dispersion=disp,
                 dipoles=[dip], field_monitors=[fm])


def _staircase_dispersion(mat: np.ndarray, materials: list, dt: float) -> Dispersion:
    """CSR entries for every cell whose material id >= 0, all three components, the material's
    poles at full weight."""
    cells = np.flatnonzero(mat.reshape(-1) >= 0)
    mids = mat.reshape(-1)[cells]
    coeffs = [trap_coeffs(p, dt) for p in materials]
    npole = np.array([len(p) for p in materials], dtype=np.int64)
    comp = np.repeat(np.arange(3, dtype=np.int32), cells.size)
    cell = np.tile(cells.astype(np.int32), 3)
    mid3 = np.tile(mids, 3)
    counts = npole[mid3]
    ofs = np.concatenate([[0], np.cumsum(counts)]).astype(np.int32)
    # pole t of entry e is pole (t - ofs[e]) of the entry's material
    cat_am1 = np.concatenate([c[0] for c in coeffs])
    cat_b = np.concatenate([c[1] for c in coeffs])
    first = np.concatenate([[0], np.cumsum(npole)[:-1]])
    src = (np.repeat(first[mid3], counts) + np.arange(int(ofs[-1]))
           - np.repeat(ofs[:-1].astype(np.int64), counts))
    am1, b = cat_am1[src], cat_b[src]
    gm = np.array([2.0 * c[1].real.sum() for c in coeffs])
    return Dispersion(comp=comp, cell=cell, pole_ofs=ofs, am1=am1, b=b, g=gm[mid3])


def mini_scene(shape=(10, 12, 40), dl: float = 20e-9, n_pml: int = 6, seed: int = 1,
               num_steps: int = 300, plane_wave_on: bool = True, dipole_on: bool = True,
               magnetic_dipole: bool = False, lorentz_on: bool = True, sigma_on: bool = False,
               pec_on: bool = False, periodic_xy: bool = True, pml_xy: bool = False,
               nf: int = 4, pol_axis: int = 0, direction: int = +1,
               field_box: bool = True, nonuniform_z: bool = False) -> Scene:
    """Small scene exercising every real-path feature at once, for oracle parity."""
    rng = np.random.default_rng(seed)
    nx, ny, nz = shape
    if pml_xy:
        bx = by = "PML"
    else:
        bx = by = "Periodic" if periodic_xy else "PECBoundary"
    ax = uniform_axis(nx, dl, 0.0, bx)
    ay = uniform_axis(ny, dl * 1.1, 0.0, by)
    if nonuniform_z:
        dz = dl * (1.0 + 0.3 * np.sin(np.linspace(0, 3, nz)) ** 2)
        dz[:n_pml] = dz[n_pml]
        dz[-n_pml:] = dz[-n_pml - 1]
        az = Axis(np.concatenate([[0.0], np.cumsum(dz)]), "PML", "PML")
    else:
        az = uniform_axis(nz, dl, 0.0, "PML")
    grid = Grid(ax, ay, az)
    dt = cpml.courant_dt((dl, dl * 1.1, float(az.dl.min())))
    f0 = 3e14
    pulse = waveform.GaussianPulse(freq0=f0, fwidth=0.3 * f0)
    src_k = n_pml + 3
    mon_k = nz - n_pml - 3
    eps = [np.ones(shape) for _ in range(3)]
    k0, k1 = nz // 2 - 3, nz // 2 + 2
    for c in range(3):
        eps[c][:, :, k0:k1] = 2.0 + rng.uniform(0, 2, size=(nx, ny, k1 - k0))
    sig = None
    if sigma_on:
        sig = [np.zeros(shape) for _ in range(3)]
        for c in range(3):
            sig[c][:, :, k0:k1] = rng.uniform(0, 2e5, size=(nx, ny, k1 - k0))
    disp = None
    if lorentz_on:
        mat = np.full(shape, -1, np.int32)
        mat[1:nx - 1, 2:ny - 2, k1 + 1:k1 + 5] = 0
        mat[2:nx - 2, 1:ny - 1, k0 - 5:k0 - 2] = 1
        mats = [lorentz_poles(1.5, 2.5e15, 2e14) + lorentz_poles(0.7, 1.2e15, 5e13),
                lorentz_poles(2.0, 1.8e15, 1e14)]
        disp = _staircase_dispersion(mat, mats, dt)
        for c in range(3):
            eps[c][mat >= 0] = 1.7
    pec = None
    if pec_on:
        pe = np.flatnonzero((np.arange(nx * ny * nz).reshape(shape) % 97) == 5)
        pec = (pe, pe.copy(), pe.copy())
    srcs = []
    if plane_wave_on:
        srcs.append(plane_wave(grid, eps[pol_axis], dt, num_steps,
                               float(az.edges[src_k if direction > 0 else mon_k + 1]),
                               direction, pol_axis, pulse))
    dips = []
    if dipole_on:
        wf = waveform.sample(waveform.GaussianPulse(freq0=f0 * 1.1, fwidth=0.4 * f0,
                                                    phase=0.3), dt, num_steps)
        pos = (0.5 * (ax.edges[0] + ax.edges[-1]) + 0.3 * dl, 0.4 * (ay.edges[0] + ay.edges[-1]),
               float(az.edges[nz // 2 + 6]) + 0.25 * dl)
        for comp in (0, 2):
            idx, coef = dipole_stencil(grid, comp, pos)
            dips.append(PointDipole(comp, idx, coef * 1e-21 * (1 + comp), wf))
        if magnetic_dipole:
            idx, coef = dipole_stencil(grid, 1, pos, magnetic=True)
            dips.append(PointDipole(1, idx, coef * 1e-21 * 377.0, wf, magnetic=True))
    freqs = np.linspace(0.8 * f0, 1.2 * f0, nf)
    mons = [FluxMonitor("T", 2, mon_k, +1, freqs),
            FluxMonitor("R", 2, src_k - 1, -1, freqs),
            FluxMonitor("X", 0, nx // 2, +1, freqs[:2]),
            FluxMonitor("Y", 1, ny // 2, +1, freqs[1:3])]
    fms = []
    if field_box:
        fms.append(FieldMonitor("box", freqs[:3], (1, 2, nz // 2 - 4), (nx - 2, ny - 3, 9),
                                apodization=(None, dt * num_steps * 0.7, dt * 30)))
    sides = [(2, "lo"), (2, "hi")]
    if pml_xy:
        sides += [(0, "lo"), (0, "hi"), (1, "lo"), (1, "hi")]
    return Scene(grid=grid, dt=dt, num_time_steps=num_steps, eps_ex=eps[0], eps_ey=eps[1],
                 eps_ez=eps[2], pml=pml_faces(dt, (dl, dl * 1.1, float(az.dl[0])), sides,
                                              n_pml if not pml_xy else min(n_pml, 3)),
                 sigma_ex=None if sig is None else sig[0],
                 sigma_ey=None if sig is None else sig[1],
                 sigma_ez=None if sig is None else sig[2],
                 pec_ex=None if pec is None else pec[0], pec_ey=None if pec is None else pec[1],
                 pec_ez=None if pec is None else pec[2],
                 dispersion=disp, sources=srcs, dipoles=dips, flux_monitors=mons,
                 field_monitors=fms)


def metal_poles(wp: float = 1.37e16, gamma: float = 1e14,
                lorentz=((1.0, 4e15, 2e14), (0.6, 6e15, 3e14))):
    """Drude + Lorentz pole pairs (angular units) of the Bloch-cell metal."""
    from .model import drude_poles
    poles = drude_poles(wp, gamma)
    for de, w0, dl in lorentz:
        poles += lorentz_poles(de, w0, dl)
    return poles


def bloch_cell_scene(nxy: int = 256, nz: int = 600, dl: float = 10e-9, n_pml: int = 12,
                     seed: int = 0, theta_deg: float = 30.0, phi_deg: float = 45.0,
                     f0: float = 4e14, fwidth_rel: float = 0.4, num_steps: int = 8000,
                     nf: int = 21, sub_frac: float = 0.35, height: float = 200e-9,
                     eps_sub: float = 2.1, with_metal: bool = True, n_cyl: int = 3,
                     with_tfsf: bool = True) -> Scene:
    """Bloch / oblique periodic unit cell: Bloch x/y for (theta, phi) incidence at f0, CPML on
    z, eps = 2.1 substrate, a 3x3 array of metal cylinders (Drude + 2 Lorentz pole-residue
    metal; the Drude DC pole is a conductivity) of seeded radius 20-60 nm on the substrate, an
    oblique TF/SF box (open in x/y) driven from above, and flux planes of 21 frequencies."""
    from .model import split_dc_pole
    from .tfsf import EPS3  # noqa: F401  (module import check)
    from .model import TFSFSource
    rng = np.random.default_rng(seed)
    ax = uniform_axis(nxy, dl, 0.0, "BlochBoundary")
    ay = uniform_axis(nxy, dl, 0.0, "BlochBoundary")
    az = uniform_axis(nz, dl, 0.0, "PML")
    grid = Grid(ax, ay, az)
    dt = cpml.courant_dt((dl, dl, dl))
    L = nxy * dl
    k0 = 2 * np.pi * f0 / cpml.C_0
    th, ph = np.deg2rad(theta_deg), np.deg2rad(phi_deg)
    kx, ky = k0 * np.sin(th) * np.cos(ph), k0 * np.sin(th) * np.sin(ph)
    bk = (kx * L / (2 * np.pi), ky * L / (2 * np.pi), 0.0)
    shape = grid.shape
    k_sub = int(round(sub_frac * nz))                   # substrate top edge index
    eps = np.ones(shape)
    eps[:, :, :k_sub] = eps_sub                          # Ex/Ey at edges k < k_sub, Ez centres
    eps_ez = np.ones(shape)
    eps_ez[:, :, :k_sub] = eps_sub
    sigma = None
    disp = None
    kc_hi = k_sub + int(round(height / dl))
    if with_metal:
        poles = metal_poles()
        keep, sig_dc = split_dc_pole(poles, dt)
        mat = np.full(shape, -1, np.int32)
        xc = ax.centers
        pitch = L / n_cyl
        for ci in range(n_cyl):
            for cj in range(n_cyl):
                r = rng.uniform(20e-9, 60e-9)
                cx, cy = (ci + 0.5) * pitch, (cj + 0.5) * pitch
                inside = ((xc[:, None] - cx) ** 2 + (xc[None, :] - cy) ** 2) <= r * r
                for k in range(k_sub, kc_hi):
                    mat[:, :, k][inside] = 0
        disp = _staircase_dispersion(mat, [keep], dt)
        sigma = np.where(mat >= 0, sig_dc, 0.0)
        for e_ in (eps, eps_ez):
            e_[mat >= 0] = 1.0
    pulse = waveform.GaussianPulse(freq0=f0, fwidth=fwidth_rel * f0)
    wf = waveform.sample(pulse, dt, num_steps)
    tfsf = []
    lo_k, hi_k = max(n_pml + 4, k_sub - 20), min(nz - n_pml - 8, kc_hi + 30)
    if with_tfsf:
        line_e = np.where(np.arange(nz) < k_sub, eps_sub, 1.0)
        tfsf.append(TFSFSource(box_lo=(0, 0, lo_k), box_hi=(nxy, nxy, hi_k), direction=-1,
                               waveform=wf, open_axes=(0, 1), kxy=(kx, ky),
                               jxy=(np.cos(ph), np.sin(ph)), eps_line=(line_e, line_e)))
    freqs = np.linspace(f0 * (1 - fwidth_rel / 2), f0 * (1 + fwidth_rel / 2), nf)
    mons = [FluxMonitor("R", 2, min(hi_k + 4, nz - n_pml - 2), +1, freqs),
            FluxMonitor("T", 2, max(lo_k - 4, n_pml + 2), -1, freqs),
            FluxMonitor("box_zlo", 2, lo_k + 3, -1, freqs),
            FluxMonitor("box_zhi", 2, hi_k - 3, +1, freqs)]
    return Scene(grid=grid, dt=dt, num_time_steps=num_steps, eps_ex=eps, eps_ey=eps.copy(),
                 eps_ez=eps_ez, sigma_ex=sigma, sigma_ey=None if sigma is None else sigma.copy(),
                 sigma_ez=None if sigma is None else sigma.copy(),
                 pml=pml_faces(dt, (dl, dl, dl), [(2, "lo"), (2, "hi")], n_pml),
                 dispersion=disp, bloch_k=bk, tfsf_sources=tfsf, flux_monitors=mons)


def tfsf_box_scene(shape=(26, 24, 44), dl: float = 20e-9, n_pml: int = 6, num_steps: int = 300,
                   jxy=(1.0, 0.0), direction: int = +1, scatterer: bool = False,
                   periodic_xy: bool = False) -> Scene:
    """Small real-field scene with a 6-face (or z-face-only when periodic_xy) TF/SF box at
    normal incidence along z, for leakage and parity tests."""
    from .model import TFSFSource
    nx, ny, nz = shape
    b = "Periodic" if periodic_xy else "PML"
    grid = Grid(uniform_axis(nx, dl, 0, b), uniform_axis(ny, dl, 0, b),
                uniform_axis(nz, dl, 0, "PML"))
    dt = cpml.courant_dt((dl, dl, dl))
    eps = np.ones(shape)
    if scatterer:
        eps[nx // 2 - 2:nx // 2 + 2, ny // 2 - 2:ny // 2 + 2, nz // 2 - 3:nz // 2 + 3] = 4.0
    f0 = 3e14
    wf = waveform.sample(waveform.GaussianPulse(freq0=f0, fwidth=0.4 * f0), dt, num_steps)
    m = 0 if periodic_xy else n_pml + 3
    lo = (m, m, n_pml + 6)
    hi = (nx - m if not periodic_xy else nx, ny - m if not periodic_xy else ny, nz - n_pml - 6)
    src = TFSFSource(box_lo=lo, box_hi=hi, direction=direction, waveform=wf,
                     open_axes=(0, 1) if periodic_xy else (), jxy=jxy)
    sides = [(2, "lo"), (2, "hi")] + ([] if periodic_xy else
                                      [(0, "lo"), (0, "hi"), (1, "lo"), (1, "hi")])
    freqs = np.linspace(0.8 * f0, 1.2 * f0, 3)
    mons = [FluxMonitor("zin", 2, nz // 2, +1, freqs),
            FluxMonitor("zout", 2, nz - n_pml - 3, +1, freqs)]
    return Scene(grid=grid, dt=dt, num_time_steps=num_steps, eps_ex=eps, eps_ey=eps.copy(),
                 eps_ez=eps.copy(), pml=pml_faces(dt, (dl, dl, dl), sides, n_pml),
                 tfsf_sources=[src], flux_monitors=mons)


# ---------------------------------------------------------------------------------------------
# Serving stream: seeded requests of three scene kinds on cubic grids
# ---------------------------------------------------------------------------------------------

SERVING_KINDS = ("slab", "soup", "sphere")


def serving_requests(n: int = 256, rate: float = 20.0, seed: int = 0, side_lo: int = 64,
                     side_hi: int = 320, steps_lo: int = 2000, steps_hi: int = 8000) -> list:
    """``n`` request specs arriving as a Poisson process of ``rate`` per second: cubic side
    log-uniform in [side_lo, side_hi], step count uniform in [steps_lo, steps_hi], 1-3 flux
    monitors of 5-21 frequencies, kind uniform over :data:`SERVING_KINDS`."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, size=n))
    out = []
    for q in range(n):
        side = int(round(np.exp(rng.uniform(np.log(side_lo), np.log(side_hi)))))
        out.append(dict(id=q, arrival=float(t[q]), kind=SERVING_KINDS[int(rng.integers(3))],
                        side=side, n_steps=int(rng.integers(steps_lo, steps_hi + 1)),
                        n_mon=int(rng.integers(1, 4)), nf=int(rng.integers(5, 22)),
                        seed=int(rng.integers(2 ** 31))))
    return out


def _flux_planes(shape, n_mon: int, n_pml: int, freqs) -> list:
    """Up to three flux planes just inside the high-side PML of z, x, y (or of z alone when
    x and y are periodic)."""
    out = []
    for q in range(n_mon):
        axis = (2, 0, 1)[q]
        km = shape[axis] - n_pml - 3 - 2 * q
        out.append(FluxMonitor(f"flux{q}", axis, km, +1, freqs))
    return out


def serving_scene(spec: dict, dl: float = 20e-9, n_pml: int = 12) -> Scene:
    """The scene of one serving request (see :func:`serving_requests`).

    * ``slab``: periodic x/y, PML z, a z-normal slab of eps in [2, 12] (thickness 10-30% of
      the interior) driven by a +z Ex plane wave (f0 = 2.998e14 Hz, fwidth 0.2 f0); flux planes
      normal to z.
    * ``soup``: the seeded box soup (32 boxes, plain eps or 2-pole Lorentz) with PML on all
      faces and a central Ez dipole (f0 = 2e14 Hz, fwidth 0.5 f0); flux planes normal to z, x, y.
    * ``sphere``: an Ez dipole at the centre of a Lorentz sphere (eps_inf in [1, 4], one or two
      poles) of radius 15-35% of the interior, vacuum elsewhere, PML on all faces.
    """
    kind, side, ns = spec["kind"], int(spec["side"]), int(spec["n_steps"])
    rng = np.random.default_rng(spec["seed"])
    if kind == "slab":
        size = (side * dl, side * dl, (side - 2 * n_pml) * dl)
        thick = rng.uniform(0.1, 0.3) * size[2]
        f0 = 2.998e14
        freqs = np.linspace(0.9 * f0, 1.1 * f0, spec["nf"])
        nz_in = side - 2 * n_pml
        src_z = -size[2] / 2 + 6 * dl
        sc = slab_scene(dl=dl, size=size, eps_slab=float(rng.uniform(2.0, 12.0)), thick=thick,
                        f0=f0, n_pml=n_pml, src_z=src_z, t_z=size[2] / 2 - 4 * dl, r_z=None,
                        nf=spec["nf"], num_steps=ns)
        mons = []
        for q in range(spec["n_mon"]):
            km = n_pml + nz_in - 4 - 3 * q
            mons.append(FluxMonitor(f"flux{q}", 2, km, +1, freqs))
        return _replace_monitors(sc, mons, [])
    if kind == "soup":
        f0, fw = 2e14, 0.5
        sc = box_soup_scene(n=side, dl=dl, n_pml=n_pml, seed=int(spec["seed"]) % (2 ** 31),
                            f0=f0, fwidth_rel=fw, num_steps=ns, nf=spec["nf"])
        freqs = np.linspace(f0 * (1 - fw), f0 * (1 + fw), spec["nf"])
        return _replace_monitors(sc, _flux_planes(sc.shape, spec["n_mon"], n_pml, freqs), [])
    if kind == "sphere":
        return _sphere_scene(side, ns, spec, rng, dl, n_pml)
    raise ValueError(f"unknown serving scene kind {kind!r}")


def _replace_monitors(sc: Scene, flux: list, field_mons: list) -> Scene:
    import dataclasses
    return dataclasses.replace(sc, flux_monitors=flux, field_monitors=field_mons)


def _sphere_scene(side: int, ns: int, spec: dict, rng, dl: float, n_pml: int) -> Scene:
    axes = [uniform_axis(side, dl, 0.0, "PML") for _ in range(3)]
    grid = Grid(*axes)
    dt = cpml.courant_dt((dl, dl, dl))
    interior = side - 2 * n_pml
    rad = rng.uniform(0.15, 0.35) * interior
    c = (side - 1) / 2.0
    i = np.arange(side, dtype=np.float64) - c
    r2 = i[:, None, None] ** 2 + i[None, :, None] ** 2 + i[None, None, :] ** 2
    inside = r2 <= rad * rad
    eps_inf = rng.uniform(1.0, 4.0)
    poles = []
    for _p in range(int(rng.integers(1, 3))):
        poles += lorentz_poles(rng.uniform(0.5, 3.0), rng.uniform(1e14, 8e14), 1e13)
    eps = np.where(inside, eps_inf, 1.0)
    mat = np.where(inside, 0, -1).astype(np.int32)
    disp = _staircase_dispersion(mat, [poles], dt)
    if disp.n_entry == 0:
        disp = None
    f0, fw = 2e14, 0.5
    pulse = waveform.GaussianPulse(freq0=f0, fwidth=fw * f0)
    wf = waveform.sample(pulse, dt, ns)
    centre = [0.5 * (a.edges[0] + a.edges[-1]) for a in axes]
    idx, coef = dipole_stencil(grid, 2, centre)
    dip = PointDipole(component=2, indices=idx, coef=coef * 1e-20, waveform=wf)
    freqs = np.linspace(f0 * (1 - fw), f0 * (1 + fw), spec["nf"])
    sides_all = [(a, s) for a in range(3) for s in ("lo", "hi")]
    return Scene(grid=grid, dt=dt, num_time_steps=ns, eps_ex=eps, eps_ey=eps, eps_ez=eps,
                 pml=pml_faces(dt, (dl, dl, dl), sides_all, n_pml), dispersion=disp,
                 dipoles=[dip], flux_monitors=_flux_planes(grid.shape, spec["n_mon"], n_pml, freqs))


# ---------------------------------------------------------------------------------------------
# Waveguides: mode source + mode monitors
# ---------------------------------------------------------------------------------------------

def waveguide_scene(shape=(60, 40, 30), dl: float = 40e-9, n_pml: int = 8,
                    core_eps: float = 12.1, clad_eps: float = 2.1, width_cells: int = 12,
                    thick_cells: int = 6, f0: float = 1.934e14, fwidth_rel: float = 0.1,
                    src_index: int | None = None, num_steps: int = 1200, nf: int = 3,
                    direction: int = +1, mode_index: int = 0) -> Scene:
    """A strip waveguide along x (core ``width_cells`` x ``thick_cells`` of ``core_eps`` in
    ``clad_eps``), PML on all faces, the fundamental mode launched from ``src_index`` along
    ``direction`` (GaussianPulse f0, fwidth ``fwidth_rel`` f0), and mode + flux monitors on a
    plane ahead of the source ("fwd") and one behind it ("back")."""
    import dataclasses
    from . import modes
    nx, ny, nz = shape
    axes = [uniform_axis(m, dl, 0.0, "PML") for m in shape]
    grid = Grid(*axes)
    dt = cpml.courant_dt((dl, dl, dl))
    eps = np.full(shape, clad_eps)
    yc, zc = ny // 2, nz // 2
    eps[:, yc - width_cells // 2: yc - width_cells // 2 + width_cells,
        zc - thick_cells // 2: zc - thick_cells // 2 + thick_cells] = core_eps
    ks = src_index if src_index is not None else (n_pml + 6 if direction > 0 else nx - n_pml - 7)
    k_fwd = ks + direction * (nx - 2 * n_pml - 14) // 1
    k_fwd = min(max(k_fwd, n_pml + 2), nx - n_pml - 3)
    k_back = ks - direction * 3
    freqs = np.linspace(f0 * (1 - 0.05), f0 * (1 + 0.05), nf) if nf > 1 else np.array([f0])
    sides_all = [(a, s) for a in range(3) for s in ("lo", "hi")]
    base = Scene(grid=grid, dt=dt, num_time_steps=num_steps, eps_ex=eps, eps_ey=eps, eps_ez=eps,
                 pml=pml_faces(dt, (dl, dl, dl), sides_all, n_pml))
    m = modes.solve_mode(base, 0, ks, f0, mode_index=mode_index)
    pulse = waveform.GaussianPulse(freq0=f0, fwidth=fwidth_rel * f0)
    src = modes.mode_source(base, m, pulse, direction, num_steps)
    mons = [ModeMonitor("fwd", 0, k_fwd, freqs, mode_index),
            ModeMonitor("back", 0, k_back, freqs, mode_index)]
    flux = [FluxMonitor("fwd_flux", 0, k_fwd, +1, freqs),
            FluxMonitor("back_flux", 0, k_back, +1, freqs)]
    return dataclasses.replace(base, mode_sources=[src], mode_monitors=mons, flux_monitors=flux)


def bend_design(n_pix: int = 160, seed: int = 0, sigma: float = 4.0, eps_lo: float = 2.1,
                eps_hi: float = 12.1) -> np.ndarray:
    """A seeded Gaussian-smoothed random field mapped into [eps_lo, eps_hi] (``(n_pix, n_pix)``),
    the inverse-design starting point."""
    from scipy.ndimage import gaussian_filter
    rng = np.random.default_rng(seed)
    f = gaussian_filter(rng.standard_normal((n_pix, n_pix)), sigma, mode="reflect")
    f = (f - f.min()) / max(float(np.ptp(f)), 1e-30)
    return eps_lo + (eps_hi - eps_lo) * f


def bend_scene(design: np.ndarray | None = None, n: int = 400, nz: int = 48, dl: float = 20e-9,
               n_pml: int = 12, thick: float = 0.22e-6, width: float = 0.5e-6,
               core_eps: float = 12.1, clad_eps: float = 2.1, f0: float = 1.934e14,
               fwidth_rel: float = 0.1, num_steps: int = 6000, seed: int = 0):
    """Inverse-design 90-degree bend: an x-directed input guide and a y-directed output guide
    (``width`` x ``thick`` strips of ``core_eps`` in ``clad_eps``) meet a square design region
    of ``design.shape`` pixels (one cell each, uniform through the slab thickness, all three eps
    components); PML on all faces; the fundamental mode launched along +x, a mode monitor on the
    output guide along +y.  Returns ``(scene, design_cells)`` where ``design_cells`` is
    ``(n_pix, n_pix, n_layers)`` flat indices of the design cells.  (The two mode solves do not
    depend on the design and are cached across designs.)"""
    import dataclasses
    from . import modes as modes_mod
    if design is None:
        design = bend_design(seed=seed)
    npx = design.shape[0]
    axes = [uniform_axis(n, dl, 0.0, "PML"), uniform_axis(n, dl, 0.0, "PML"),
            uniform_axis(nz, dl, 0.0, "PML")]
    grid = Grid(*axes)
    dt = cpml.courant_dt((dl, dl, dl))
    eps = np.full((n, n, nz), clad_eps)
    nt = int(round(thick / dl))
    k0 = nz // 2 - nt // 2
    nw = int(round(width / dl))
    c = n // 2
    d0 = c - npx // 2                                  # design region [d0, d0 + npx) in x and y
    w0 = c - nw // 2
    eps[:d0, w0:w0 + nw, k0:k0 + nt] = core_eps        # input guide along x, from the left
    eps[w0:w0 + nw, d0 + npx:, k0:k0 + nt] = core_eps  # output guide along y, to the top
    eps[d0:d0 + npx, d0:d0 + npx, k0:k0 + nt] = design[:, :, None]
    ii, jj, kk = np.meshgrid(np.arange(d0, d0 + npx), np.arange(d0, d0 + npx),
                             np.arange(k0, k0 + nt), indexing="ij")
    cells = (ii * n + jj) * nz + kk
    sides_all = [(a, s) for a in range(3) for s in ("lo", "hi")]
    base = Scene(grid=grid, dt=dt, num_time_steps=num_steps, eps_ex=eps, eps_ey=eps, eps_ez=eps,
                 pml=pml_faces(dt, (dl, dl, dl), sides_all, n_pml))
    ks = n_pml + 10
    km = n - n_pml - 10
    m_src = modes_mod.solve_mode(base, 0, ks, f0)
    pulse = waveform.GaussianPulse(freq0=f0, fwidth=fwidth_rel * f0)
    src = modes_mod.mode_source(base, m_src, pulse, +1, num_steps)
    mon = ModeMonitor("out", 1, km, np.array([f0]))
    return dataclasses.replace(base, mode_sources=[src], mode_monitors=[mon]), cells


def with_pole_block(sc: Scene, lo, hi, materials) -> Scene:
    """``sc`` with ADE poles on all three components of the cells ``lo <= (i, j, k) < hi``
    (eps_inf stays the scene's eps): ``materials`` is a list of pole lists, material
    ``(i + j + k) % len(materials)`` at cell (i, j, k).  One material gives every entry the same
    pole count; materials of different counts give the general (CSR) pole layout."""
    import dataclasses
    mat = np.full(sc.shape, -1, dtype=np.int64)
    ii, jj, kk = np.meshgrid(*(np.arange(a, b) for a, b in zip(lo, hi)), indexing="ij")
    mat[ii, jj, kk] = (ii + jj + kk) % len(materials)
    return dataclasses.replace(sc, dispersion=_staircase_dispersion(mat, materials, sc.dt))


def pole_guide_scene(num_steps: int = 260, two_materials: bool = False) -> Scene:
    """A 40 x 24 x 22 strip guide (40 nm cells, source plane x = 12, forward mode monitor
    x = 26) with a block of Lorentz poles on and beside the core's upper half for x in [15, 24):
    one 2-pole material (every entry the same pole count) or, ``two_materials``, a 1-pole one on
    alternate cells as well (the general pole layout)."""
    from .model import lorentz_poles
    sc = waveguide_scene(shape=(40, 24, 22), n_pml=6, width_cells=6, thick_cells=4,
                         fwidth_rel=0.3, num_steps=num_steps, nf=1)
    mats = [lorentz_poles(1.5, 2.5e15, 2e14) + lorentz_poles(0.7, 1.2e15, 5e13)]
    if two_materials:
        mats.append(lorentz_poles(2.0, 1.8e15, 1e14))
    return with_pole_block(sc, (15, 8, 11), (24, 16, 16), mats)
