"""Scene -> stepping tables (numpy, float64 masters).

One builder feeds both consumers: the float64 oracle reads the float64 masters, the HIP library
receives their float32 roundings. The table semantics follow the reference's setup chain:

- per-axis CPML coefficient arrays with identity fill outside the PML
  (``openem/coeffs.py:21-45``), neighbour/mask tables (``openem/grid.py:219-304``);
- ``ca = (eps - s)/(eps + s + G)``, ``cb = (dt/eps0)/(eps + s + G)`` with ``s = sigma dt/(2 eps0)``
  and ``G`` the ADE coupling (``openem/coeffs.py:98-136``), zeroed on PEC cells
  (``openem/coeffs.py:139-171``);
- plane-wave TF/SF coefficients (``openem/sources_setup.py:93-119``);
- dipole stencils with the electric amplitude at half steps and the magnetic at whole steps
  (``openem/kernels/source_dft.cu:147-206``);
- DFT phase tables: E at ``t=(n+1)dt``, H at ``(n+1/2)dt``, apodisation multiplied in
  (``openem/monitors_setup.py:74-101``, ``openem/kernels/source_dft.cu:332-362``).

Differences in representation, not in values:

- CPML psi lives only in the PML slabs: ``pml_lo/pml_hi`` give each axis's slab thickness.
  Outside the slabs the coefficients are the identity, where psi stays 0 and the stretched
  derivative equals the plain one, so the restriction changes no value.
- ``(ca, cb)`` of all three components plus a per-component "dispersive" flag are deduplicated
  into one uint16 *material class* per cell and a LUT (the reference dedups each component into
  its own uint16 index, ``openem/coeffs.py:174-208``); a lookup returns the same float32 values.
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import _lib, cpml
from . import tfsf as tfsf_mod
from .grid import transverse_axes
from .model import Scene

LUT_STRIDE = 8          # ca_x cb_x ca_y cb_y ca_z cb_z flags pad
MAX_CLASSES = 65535


@dataclass
class AxisTables:
    n: int
    idl_p: np.ndarray
    idl_d: np.ndarray
    nxt: np.ndarray
    prv: np.ndarray
    mnx: np.ndarray         # float64, or complex128 on a Bloch axis
    mpv: np.ndarray
    pec: np.ndarray
    a_e: np.ndarray
    b_e: np.ndarray
    k_e: np.ndarray
    a_h: np.ndarray
    b_h: np.ndarray
    k_h: np.ndarray
    pml_lo: int
    pml_hi: int


@dataclass
class PlaneSrcTables:
    pol: int                # 0 -> (Ex, Hy), 1 -> (Ey, Hx)
    kh: int
    ks: int
    coef_h: float
    coef_e: float
    ex_inc: np.ndarray      # (n_total+1,) float64
    hy_inc: np.ndarray


@dataclass
class DipoleTables:
    magnetic: bool
    comp: np.ndarray        # (n_pts,) int32
    cell: np.ndarray        # (n_pts,) int32, flat
    coef: np.ndarray        # (n_pts,) float64
    src_of: np.ndarray      # (n_pts,) int32
    amp: np.ndarray         # (n_src, n_total+1) float64, complex128 on a Bloch scene


@dataclass
class MonitorTables:
    name: str
    kind: str               # "plane" or "box"
    nf: int
    freqs: np.ndarray
    phase: np.ndarray       # (n_total, nf, 4) float64: cos_e, sin_e, cos_h, sin_h (windowed)
    axis: int = 2
    km: int = 0
    n1: int = 0
    n2: int = 0
    origin: tuple = (0, 0, 0)
    box: tuple = (0, 0, 0)
    modes: list | None = None   # mode monitors: one solved Mode per frequency
    # the scene monitor these tables were built from (FluxMonitor, FieldMonitor or ModeMonitor)
    spec: object = None


@dataclass
class DispTables:
    comp: np.ndarray        # (n_entry,) int32
    cell: np.ndarray        # (n_entry,) int32 flat
    ofs: np.ndarray         # (n_entry+1,) int32
    am1: np.ndarray         # (n_pole,) complex128
    b: np.ndarray           # (n_pole,) complex128
    cidx: np.ndarray        # (n_pole,) int32 into the deduplicated coefficient LUT
    lut_am1: np.ndarray     # (n_coef,) complex64 values as stored on the device
    lut_b: np.ndarray
    hist_box: tuple         # (x0, x1, y0, y1, z0, z1)


@dataclass
class ModeTables:
    """Point lists of one mode source (H side after the TF/SF H corrections, E side after the
    TF/SF E corrections) and its two amplitude tables (reference ``_mode_source_entry``,
    ``openem/sources_setup.py:376-416``)."""

    h_comp: np.ndarray
    h_cell: np.ndarray
    h_coef: np.ndarray            # float64: face sign / dl (times dt/mu0 on device)
    h_phase: np.ndarray           # complex128 profile value
    e_comp: np.ndarray
    e_cell: np.ndarray
    e_coef: np.ndarray            # float64: face sign / dl_dual (times the cell's cb on device)
    e_phase: np.ndarray
    amp_e: np.ndarray             # (n_total+1,) complex128, amp_time(n dt)
    amp_h: np.ndarray             # (n_total+1,) complex128, amp_time((n+1/2) dt)


@dataclass
class StepTables:
    shape: tuple
    n_total: int
    dt: float
    ch: float
    cw: float
    axes: list
    scene: Scene            # the source of the material coefficients
    disp: DispTables | None = None
    plane_srcs: list = field(default_factory=list)
    dipoles: list = field(default_factory=list)
    monitors: list = field(default_factory=list)
    tfsf: list = field(default_factory=list)
    modes: list = field(default_factory=list)
    bloch: bool = False
    coords: list | None = None   # per axis (edge coordinates, centre coordinates)

    # filled lazily: float64 (ca, cb) and the material classes
    _cacb: tuple | None = None
    _flag: np.ndarray | None = None
    _cls: np.ndarray | None = None
    _lut: np.ndarray | None = None
    _eps_cls: np.ndarray | None = None

    @property
    def ca(self) -> np.ndarray:
        """(3, nx, ny, nz) float64 ca (reference ``openem/coeffs.py:98-136``)."""
        if self._cacb is None:
            self._cacb = e_coeffs(self.scene)
        return self._cacb[0]

    @property
    def cb(self) -> np.ndarray:
        if self._cacb is None:
            self._cacb = e_coeffs(self.scene)
        return self._cacb[1]

    @property
    def disp_flag(self) -> np.ndarray:
        """(3, nx, ny, nz) bool: the cell of this component carries ADE pole entries."""
        if self._flag is None:
            self._flag = _disp_flags(self.scene)
        return self._flag

    def material_classes(self):
        """``(cls (nx,ny,nz) uint16, lut (n_classes, 8) float32)``, or ``None`` when more than
        65535 distinct classes exist (the caller then uploads dense ca/cb)."""
        if self._cls is None:
            self._cls, self._lut, self._eps_cls = _material_classes(self.scene)
        if self._cls.size == 0:
            return None
        return self._cls, self._lut

    def class_eps(self) -> np.ndarray | None:
        """``(n_classes, 4)`` float32: eps_x, eps_y, eps_z of each material class (the energy
        weights of the shutoff), or ``None`` without classes."""
        return None if self.material_classes() is None else self._eps_cls


def _disp_flags(sc: Scene) -> np.ndarray:
    flag = np.zeros((3,) + sc.shape, dtype=bool)
    if sc.any_dispersion:
        d = sc.dispersion
        for c in range(3):
            flag[c].reshape(-1)[d.cell[d.comp == c]] = True
    return flag


def _pml_axis_arrays(sc: Scene, axis: int, at: str):
    n = sc.grid.axes[axis].n
    a = np.zeros(n)
    b = np.ones(n)
    k = np.ones(n)
    for side in ("lo", "hi"):
        c = sc.pml.get((axis, side))
        if c is None:
            continue
        m = c.num_layers
        sl = slice(0, m) if side == "lo" else slice(n - m, n)
        a[sl] = getattr(c, f"a_{at}")
        b[sl] = getattr(c, f"b_{at}")
        k[sl] = getattr(c, f"inv_kappa_{at}")
    return a, b, k


def axis_tables(sc: Scene) -> list[AxisTables]:
    out = []
    for ax_i, ax in enumerate(sc.grid.axes):
        t = ax.index_tables(sc.bloch_k[ax_i])
        a_e, b_e, k_e = _pml_axis_arrays(sc, ax_i, "E")
        a_h, b_h, k_h = _pml_axis_arrays(sc, ax_i, "H")
        lo = sc.pml.get((ax_i, "lo"))
        hi = sc.pml.get((ax_i, "hi"))
        nlo = lo.num_layers if lo is not None else 0
        nhi = hi.num_layers if hi is not None else 0
        if nlo + nhi > ax.n:
            raise ValueError(f"axis {ax_i}: PML layers {nlo}+{nhi} exceed {ax.n} cells")
        out.append(AxisTables(
            n=ax.n, idl_p=1.0 / ax.dl, idl_d=1.0 / ax.dl_dual,
            nxt=t["nxt"], prv=t["prv"], mnx=t["mnx"], mpv=t["mpv"], pec=t["pec"],
            a_e=a_e, b_e=b_e, k_e=k_e, a_h=a_h, b_h=b_h, k_h=k_h, pml_lo=nlo, pml_hi=nhi))
    return out


def e_coeffs(sc: Scene) -> tuple[np.ndarray, np.ndarray]:
    """``(ca, cb)`` each ``(3, nx, ny, nz)`` float64 (reference ``openem/coeffs.py:98-136``)."""
    shape = sc.shape
    gg = sc.dispersion.g_dense(shape) if sc.any_dispersion else None
    eps = (sc.eps_ex, sc.eps_ey, sc.eps_ez)
    sig = (sc.sigma_ex, sc.sigma_ey, sc.sigma_ez)
    ca = np.empty((3,) + shape)
    cb = np.empty((3,) + shape)
    k0 = sc.dt / cpml.EPSILON_0
    for c in range(3):
        e = np.asarray(eps[c], dtype=np.float64).reshape(shape)
        if sig[c] is None:
            num, den = e, (e + gg[c] if gg is not None else e.copy())
        else:
            s = np.asarray(sig[c]) * (sc.dt / (2.0 * cpml.EPSILON_0))
            num, den = e - s, e + s
            if gg is not None:
                den += gg[c]
        np.divide(num, den, out=ca[c])
        np.divide(k0, den, out=cb[c])
    for c, pidx in enumerate((sc.pec_ex, sc.pec_ey, sc.pec_ez)):
        if pidx is None or np.asarray(pidx).size == 0:
            continue
        pidx = np.asarray(pidx, dtype=np.int64)
        if pidx.min() < 0 or pidx.max() >= ca[c].size:
            raise ValueError(f"pec_e{'xyz'[c]} flat index out of range")
        ca[c].reshape(-1)[pidx] = 0.0
        cb[c].reshape(-1)[pidx] = 0.0
    return ca, cb


def _cell_flags(sc: Scene) -> np.ndarray:
    """Per cell uint32: bit c = component c carries ADE entries, bit 3+c = component c is PEC."""
    fl = np.zeros(sc.grid.num_cells, np.uint32)
    if sc.any_dispersion:
        d = sc.dispersion
        for c in range(3):
            fl[d.cell[d.comp == c]] |= np.uint32(1 << c)
    for c, pidx in enumerate((sc.pec_ex, sc.pec_ey, sc.pec_ez)):
        if pidx is not None and np.asarray(pidx).size:
            pidx = np.asarray(pidx, dtype=np.int64)
            if pidx.min() < 0 or pidx.max() >= fl.size:
                raise ValueError(f"pec_e{'xyz'[c]} flat index out of range")
            fl[pidx] |= np.uint32(8 << c)
    return fl


def _material_classes(sc: Scene):
    """Per-cell material class (uint16) and its LUT rows (ca_x, cb_x, ca_y, cb_y, ca_z, cb_z,
    ADE flags) as float32 bits, computed from the distinct per-cell inputs (eps, sigma, G,
    flags): the float64 ca/cb of each distinct input row follow :func:`e_coeffs`' formula term
    by term, so every class row equals the float32 cast of ``e_coeffs`` at its cells.  Classes
    are numbered in order of first appearance (native hash factorization)."""
    empty = (np.zeros(0, np.uint16), np.zeros((0, LUT_STRIDE), np.float32),
             np.zeros((0, 4), np.float32))
    shape = sc.shape
    gg = sc.dispersion.g_dense(shape) if sc.any_dispersion else None
    eps = (sc.eps_ex, sc.eps_ey, sc.eps_ez)
    sig = (sc.sigma_ex, sc.sigma_ey, sc.sigma_ez)
    cols, layout, eps_col = [], [], []
    for c in range(3):
        eps_col.append(len(cols))          # every column so far is float64: one w64 column each
        cols.append(np.asarray(eps[c], dtype=np.float64))
        has_s, has_g = sig[c] is not None, gg is not None
        if has_s:
            cols.append(np.asarray(sig[c], dtype=np.float64))
        if has_g:
            cols.append(gg[c])
        layout.append((has_s, has_g))
    cols.append(_cell_flags(sc))
    r = _lib.factorize(cols, 1 << 20)
    if r is None:
        return empty
    inv, rows = r
    # decode the distinct input rows (float64 = two words) and evaluate e_coeffs on them
    w64 = rows[:, :-1].copy().view(np.float64)
    fl = rows[:, -1]
    k0 = sc.dt / cpml.EPSILON_0
    lut = np.zeros((rows.shape[0], LUT_STRIDE), np.float32)
    q = 0
    for c in range(3):
        has_s, has_g = layout[c]
        e = w64[:, q]; q += 1
        s = g = None
        if has_s:
            s = w64[:, q] * (sc.dt / (2.0 * cpml.EPSILON_0)); q += 1
        if has_g:
            g = w64[:, q]; q += 1
        if s is None:
            num, den = e, (e + g if g is not None else e.copy())
        else:
            num, den = e - s, e + s
            if g is not None:
                den += g
        ca, cb = num / den, k0 / den
        pec = (fl & np.uint32(8 << c)) != 0
        ca[pec] = 0.0
        cb[pec] = 0.0
        lut[:, 2 * c] = ca
        lut[:, 2 * c + 1] = cb
    lut.view(np.uint32)[:, 6] = fl & np.uint32(7)
    eps32 = np.stack([w64[:, q] for q in eps_col], axis=1).astype(np.float32)
    # distinct input rows can share one coefficient row: deduplicate (LUT row, float32 eps)
    r2 = _lib.factorize([lut[:, i].copy() for i in range(7)]
                        + [eps32[:, c].copy() for c in range(3)], MAX_CLASSES)
    if r2 is None:
        return empty
    inv2, rows2 = r2
    out = np.zeros((rows2.shape[0], LUT_STRIDE), np.uint32)
    out[:, :7] = rows2[:, :7]
    eps_cls = np.zeros((rows2.shape[0], 4), np.uint32)
    eps_cls[:, :3] = rows2[:, 7:]
    return (inv2[inv].reshape(shape).astype(np.uint16), out.view(np.float32),
            eps_cls.view(np.float32))


def disp_tables(sc: Scene) -> DispTables | None:
    if not sc.any_dispersion:
        return None
    d = sc.dispersion
    d.validate(sc.shape)
    am1_32 = d.am1.astype(np.complex64)
    b_32 = d.b.astype(np.complex64)
    cols = [np.ascontiguousarray(x, dtype=np.float32).view(np.uint32)
            for x in (am1_32.real, am1_32.imag, b_32.real, b_32.imag)]
    inv, uniq = _lib.factorize(cols, max(am1_32.size, 1))
    u = uniq.view(np.float32)
    nx, ny, nz = sc.shape
    ii, jj, kk = np.unravel_index(d.cell, sc.shape)
    hb = (int(ii.min()), int(ii.max()) + 1, int(jj.min()), int(jj.max()) + 1,
          int(kk.min()), int(kk.max()) + 1)
    return DispTables(
        comp=d.comp.astype(np.int32), cell=d.cell.astype(np.int32),
        ofs=d.pole_ofs.astype(np.int32), am1=d.am1.astype(np.complex128),
        b=d.b.astype(np.complex128), cidx=inv.reshape(-1).astype(np.int32),
        lut_am1=(u[:, 0] + 1j * u[:, 1]).astype(np.complex64),
        lut_b=(u[:, 2] + 1j * u[:, 3]).astype(np.complex64), hist_box=hb)


SIGNS = {+1: (+1.0, +1.0), -1: (-1.0, -1.0)}


def plane_source_tables(sc: Scene) -> list[PlaneSrcTables]:
    """Reference ``openem/sources_setup.py:57-119``: only z propagation with Ex/Ey polarisation;
    sources of one polarisation on the same plane are summed."""
    if not sc.sources:
        return []
    for s in sc.sources:
        if s.axis != 2 or s.pol_axis not in (0, 1):
            raise NotImplementedError("plane waves support z propagation with Ex/Ey polarisation")
    keys = {(int(s.plane_index), int(s.direction)) for s in sc.sources}
    if len(keys) != 1:
        raise NotImplementedError(f"plane-wave sources must share plane and direction: {keys}")
    az = sc.grid.z
    groups: dict[int, list] = {}
    for s in sc.sources:
        groups.setdefault(int(s.pol_axis), []).append(s)
    out = []
    for pol, grp in sorted(groups.items()):
        s0 = grp[0]
        ks = int(s0.plane_index)
        kh = ks - 1 if s0.direction > 0 else ks
        sh, se = SIGNS[s0.direction]
        sp = -1.0 if pol == 1 else 1.0
        out.append(PlaneSrcTables(
            pol=pol, kh=kh, ks=ks,
            coef_h=sp * sh * sc.dt / cpml.MU_0 * float(1.0 / az.dl[kh]),
            coef_e=se * float(1.0 / az.dl_dual[ks]),
            ex_inc=np.ascontiguousarray(sum(x.ex_inc for x in grp), dtype=np.float64),
            hy_inc=np.ascontiguousarray(sum(x.hy_inc for x in grp), dtype=np.float64)))
    return out


def dipole_tables(sc: Scene, n_total: int) -> list[DipoleTables]:
    """Electric and magnetic dipole stencils (reference ``openem/sources_setup.py:419-457``;
    the complex waveform on a Bloch scene, ``:521``)."""
    out = []
    ny, nz = sc.shape[1], sc.shape[2]
    for magnetic in (False, True):
        ds = [d for d in sc.dipoles if bool(d.magnetic) == magnetic]
        if not ds:
            continue
        comp, cell, coef, src_of, amp = [], [], [], [], []
        for s, d in enumerate(ds):
            idx = np.asarray(d.indices, dtype=np.int64).reshape(-1, 3)
            fl = (idx[:, 0] * ny + idx[:, 1]) * nz + idx[:, 2]
            comp.append(np.full(fl.size, d.component, np.int32))
            cell.append(fl.astype(np.int32))
            coef.append(np.asarray(d.coef, dtype=np.float64))
            src_of.append(np.full(fl.size, s, np.int32))
            if sc.any_bloch:
                w = d.waveform.amp_int_complex if magnetic else d.waveform.amp_half_complex
            else:
                w = d.waveform.amp_int if magnetic else d.waveform.amp_half
            if w.size < n_total + 1:
                raise ValueError(f"dipole amplitude table holds {w.size} samples, "
                                 f"{n_total} steps need {n_total + 1}")
            amp.append(np.asarray(w[:n_total + 1]))
        out.append(DipoleTables(magnetic, np.concatenate(comp), np.concatenate(cell),
                                np.concatenate(coef), np.concatenate(src_of), np.stack(amp)))
    return out


def apod_window(t: np.ndarray, spec) -> np.ndarray:
    """Flat-top window with Gaussian ramps, ``exp(-0.5((t-edge)/width)^2)`` outside
    ``[start, end]`` (reference ``openem/apodization.py:27-45``)."""
    w = np.ones_like(t)
    if spec is None:
        return w
    start, end, width = spec
    if start is not None:
        m = t < start
        w[m] = np.exp(-0.5 * ((t[m] - start) / width) ** 2)
    if end is not None:
        m = t > end
        w[m] = np.exp(-0.5 * ((t[m] - end) / width) ** 2)
    return w


def phase_table(freqs: np.ndarray, dt: float, n_total: int, apod=None) -> np.ndarray:
    """``(n_total, nf, 4)``: ``[cos w t_e, sin w t_e, cos w t_h, sin w t_h]`` times the window,
    ``t_e = (n+1) dt``, ``t_h = (n+1/2) dt``."""
    n = np.arange(n_total, dtype=np.float64)
    te = (n + 1.0) * dt
    th = (n + 0.5) * dt
    we = apod_window(te, apod)
    wh = apod_window(th, apod)
    w = 2.0 * np.pi * np.asarray(freqs, dtype=np.float64)
    out = np.empty((n_total, w.size, 4))
    ae = w[None, :] * te[:, None]
    ah = w[None, :] * th[:, None]
    out[:, :, 0] = np.cos(ae) * we[:, None]
    out[:, :, 1] = np.sin(ae) * we[:, None]
    out[:, :, 2] = np.cos(ah) * wh[:, None]
    out[:, :, 3] = np.sin(ah) * wh[:, None]
    return out


def monitor_tables(sc: Scene, n_total: int) -> list[MonitorTables]:
    out = []
    shape = sc.shape
    for m in sc.flux_monitors:
        t1, t2 = transverse_axes(m.axis)
        if not (1 <= m.plane_index < shape[m.axis]):
            raise ValueError(f"flux monitor {m.name}: plane {m.plane_index} needs a cell below it")
        out.append(MonitorTables(
            name=m.name, kind="plane", nf=len(m.freqs), freqs=np.asarray(m.freqs, np.float64),
            phase=phase_table(m.freqs, sc.dt, n_total, m.apodization),
            axis=int(m.axis), km=int(m.plane_index), n1=shape[t1], n2=shape[t2], spec=m))
    for m in sc.field_monitors:
        o, b = tuple(int(v) for v in m.origin), tuple(int(v) for v in m.box)
        for a in range(3):
            if o[a] < 0 or b[a] < 1 or o[a] + b[a] > shape[a]:
                raise ValueError(f"field monitor {m.name}: box outside the grid")
        out.append(MonitorTables(
            name=m.name, kind="box", nf=len(m.freqs), freqs=np.asarray(m.freqs, np.float64),
            phase=phase_table(m.freqs, sc.dt, n_total, m.apodization), origin=o, box=b,
            spec=m))
    if sc.mode_monitors:
        from . import modes as modes_mod
    for m in sc.mode_monitors:
        t1, t2 = transverse_axes(m.axis)
        if not (1 <= m.plane_index < shape[m.axis]):
            raise ValueError(f"mode monitor {m.name}: plane {m.plane_index} needs a cell below it")
        sol = [modes_mod.solve_mode(sc, m.axis, m.plane_index, float(f), m.mode_index, m.n_guess)
               for f in np.asarray(m.freqs, dtype=np.float64)]
        out.append(MonitorTables(
            name=m.name, kind="plane", nf=len(m.freqs), freqs=np.asarray(m.freqs, np.float64),
            phase=phase_table(m.freqs, sc.dt, n_total, None), axis=int(m.axis),
            km=int(m.plane_index), n1=shape[t1], n2=shape[t2], modes=sol, spec=m))
    return out


def mode_tables(sc: Scene, n_total: int) -> list[ModeTables]:
    """Point lists of every mode source (zero profile entries dropped)."""
    from . import modes as modes_mod
    out = []
    for ms in sc.mode_sources:
        a = int(ms.axis)
        u, v = modes_mod.cyclic(a)
        if ms.direction not in (+1, -1):
            raise ValueError(f"mode source direction {ms.direction}")
        ks = int(ms.plane_index)
        kh = ks - 1 if ms.direction > 0 else ks
        if not (1 <= ks < sc.shape[a] - 1):
            raise ValueError(f"mode source plane {ks} needs a cell on both sides")
        shp = (sc.shape[u], sc.shape[v])
        for arr in (ms.ey_inc, ms.ez_inc, ms.hy_inc, ms.hz_inc):
            if np.shape(arr) != shp:
                raise ValueError(f"mode profile {np.shape(arr)} != plane {shp}")
        if ms.amp_e.size < n_total + 1 or ms.amp_h.size < n_total + 1:
            raise ValueError("mode source amplitude tables are shorter than the run")
        s_hu, s_hv, s_eu, s_ev = modes_mod.MODE_SIGNS[ms.direction]
        ax = sc.grid.axes[a]
        ch = 1.0 / float(ax.dl[kh])          # the device multiplies H-side points by dt/mu0
        ce = 1.0 / float(ax.dl_dual[ks])     # and E-side points by the cell's cb
        cells_h = modes_mod.plane_cells(sc.shape, a, kh).ravel()
        cells_e = modes_mod.plane_cells(sc.shape, a, ks).ravel()
        lists = {"h": ([], [], [], []), "e": ([], [], [], [])}
        for side, comp, prof, coef, cells in (("h", u, ms.ez_inc, s_hu * ch, cells_h),
                                              ("h", v, ms.ey_inc, s_hv * ch, cells_h),
                                              ("e", u, ms.hz_inc, s_eu * ce, cells_e),
                                              ("e", v, ms.hy_inc, s_ev * ce, cells_e)):
            pr = np.asarray(prof, dtype=np.complex128).ravel()
            keep = np.flatnonzero(pr != 0)
            L = lists[side]
            L[0].append(np.full(keep.size, comp, np.int32))
            L[1].append(cells[keep].astype(np.int32))
            L[2].append(np.full(keep.size, coef))
            L[3].append(pr[keep])
        h = [np.concatenate(x) for x in lists["h"]]
        e = [np.concatenate(x) for x in lists["e"]]
        out.append(ModeTables(h_comp=h[0], h_cell=h[1], h_coef=h[2], h_phase=h[3],
                              e_comp=e[0], e_cell=e[1], e_coef=e[2], e_phase=e[3],
                              amp_e=np.ascontiguousarray(ms.amp_e[:n_total + 1], np.complex128),
                              amp_h=np.ascontiguousarray(ms.amp_h[:n_total + 1], np.complex128)))
    return out


def build(sc: Scene, n_total: int | None = None, force_complex: bool = False) -> StepTables:
    """All tables of a scene for ``n_total`` steps.  ``force_complex`` takes the complex
    (Bloch) field path at k = 0 (reference ``solver.run(force_complex=True)``,
    ``openem/solver.py:403-420``)."""
    n_total = int(sc.num_time_steps if n_total is None else n_total)
    for k_ax, a in enumerate(sc.grid.axes):
        if sc.bloch_k[k_ax] != 0.0 and not a.is_periodic:
            raise ValueError(f"bloch_k on axis {k_ax} needs a periodic axis")
    for s in sc.sources:
        if s.ex_inc.size < n_total + 1 or s.hy_inc.size < n_total + 1:
            raise ValueError(f"plane-wave tables hold {s.ex_inc.size} samples, "
                             f"{n_total} steps need {n_total + 1}")
    return StepTables(
        shape=sc.shape, n_total=n_total, dt=sc.dt, ch=sc.dt / cpml.MU_0,
        cw=cpml.EPSILON_0 / sc.dt, axes=axis_tables(sc), scene=sc, disp=disp_tables(sc), plane_srcs=plane_source_tables(sc),
        dipoles=dipole_tables(sc, n_total), monitors=monitor_tables(sc, n_total),
        tfsf=[tfsf_mod.tables(sc, s, n_total) for s in sc.tfsf_sources],
        modes=mode_tables(sc, n_total),
        bloch=sc.any_bloch or force_complex,
        coords=[(a.edges[:-1], a.centers) for a in sc.grid.axes])
