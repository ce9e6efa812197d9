"""Float64 NumPy oracle of the FDTD time step.  TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import this module, and only as the checker / the reported CPU baseline. The product path
(``openem_amd``) never calls it and fails loudly when the HIP library is missing.

It restates the reference's step term by term, on the float64 master tables that
``openem_amd.tables.build`` produces (the HIP library receives their float32 roundings):

- H update: ``openem/kernels/yee.cu:76-137`` - forward difference ``(F[nxt]*mnx - F)/dl``,
  CPML stretch ``s = invk*d + psi`` with ``psi <- b*psi + a*d`` (``yee.cu:60-74``), full-domain
  psi with identity coefficients outside the PML;
- E update: ``yee.cu:154-240`` - backward difference over the dual spacing,
  ``E = (ca*E + cb*(s1 - s2 - cw*hist)) * pec_a * pec_b``;
- ADE poles: ``openem/kernels/dispersion.cu:68-147`` - pre on E^n (``dP = (A-1)P + B E``,
  ``hist = sum 2 Re dP``), post after every E-side injection (``P += B E^{n+1}``), the order of
  ``openem/solver.py:1026-1041``;
- plane-wave TF/SF: ``openem/kernels/source_dft.cu:40-73`` (H side after update_h, E side after
  update_e with the cell's cb);
- point dipoles: ``source_dft.cu:147-206`` (electric ``E -= cb*coef*amp_half[n]``, magnetic
  ``H -= ch*coef*amp_int[n]``);
- running DFT: ``source_dft.cu:216-271`` (box, native Yee positions) and ``:371-430`` (plane,
  H colocated as ``(H[k]+H[k-1])/2``), E at ``(n+1)dt``, H at ``(n+1/2)dt``;
- mode sources (:meth:`Oracle64._mode`): ``inject_mode_h`` / ``inject_mode_e`` of
  ``openem/kernels/source_dft.cu:84-128`` on plane slices (H side after the TF/SF H
  corrections, E side after the TF/SF E corrections, ``openem/solver.py:907-966``);
- launch order of ``openem/solver.py:1145-1162``;
- shutoff sums (:func:`energy_sums`): ``openem/solver.py:107-122`` and ``:162-177`` with the
  float32 dual volumes and eps of ``openem/device_tables.py:320-343``; plane flux
  (:func:`plane_flux_terms`): ``openem/flux.py:75-110``.

Pinned by (see tests/test_oracle.py): discrete div(curl)=0, the Yee dispersion relation, the
CPML reflection bound (``tests/test_pml_reflection.py:57-60`` of the reference), TF/SF
one-sidedness, T+R=1 and the slab transmission against the reference's published value, the
dispersive == equivalent-lossy identity (``tests/test_dispersion.py:224-274``), zero-residue
poles as no-ops, the windowed DFT == tensordot of the time series
(``tests/test_apodization.py:171-205``), brute-force loops on tiny grids, and, for the point
dipoles, the 1D Yee Green's function and the discrete Poynting balance of an electric and of a
magnetic current sheet (tests/test_oracle_dipoles.py, which also plants sign, scale and sample
time mistakes in a copy of :meth:`Oracle64._dipoles` and sees them fail).
"""

from __future__ import annotations

import numpy as np

_EPS3 = {(0, 1, 2): 1, (1, 2, 0): 1, (2, 0, 1): 1, (0, 2, 1): -1, (2, 1, 0): -1, (1, 0, 2): -1}


def _rng(lo, hi, shape, opn, axis, on_edge):
    if axis in opn:
        return np.arange(shape[axis])
    return np.arange(lo[axis], hi[axis] + 1) if on_edge else np.arange(lo[axis], hi[axis])


_PSI = {  # (field comp, derivative axis) pairs, H and E sides
    "hx_y": (0, 1), "hx_z": (0, 2), "hy_z": (1, 2), "hy_x": (1, 0), "hz_x": (2, 0), "hz_y": (2, 1),
}


def _bc(v: np.ndarray, axis: int) -> np.ndarray:
    shp = [1, 1, 1]
    shp[axis] = -1
    return np.asarray(v).reshape(shp)


class Oracle64:
    """Stepper in float64 (real fields) or complex128 (Bloch scenes: the wrap masks carry
    ``exp(+-i 2 pi k)``, reference ``openem/kernels/bloch.cu:1-23``, ``openem/grid.py:219-260``)."""

    def __init__(self, tabs):
        self.t = tabs
        self.sc = tabs.scene
        shp = tabs.shape
        dt = np.complex128 if tabs.bloch else np.float64
        self.E = np.zeros((3,) + shp, dtype=dt)
        self.H = np.zeros((3,) + shp, dtype=dt)
        self.psi_h = {k: np.zeros(shp, dtype=dt) for k in _PSI}
        self.psi_e = {k: np.zeros(shp, dtype=dt) for k in _PSI}
        ax = tabs.axes
        mk = (lambda v: v) if tabs.bloch else (lambda v: np.real(v))
        self._m = [dict(nxt=a.nxt, prv=a.prv, mnx=_bc(mk(a.mnx), i), mpv=_bc(mk(a.mpv), i),
                        idp=_bc(a.idl_p, i), idd=_bc(a.idl_d, i), pec=_bc(a.pec, i),
                        aE=_bc(a.a_e, i), bE=_bc(a.b_e, i), kE=_bc(a.k_e, i),
                        aH=_bc(a.a_h, i), bH=_bc(a.b_h, i), kH=_bc(a.k_h, i))
                   for i, a in enumerate(ax)]
        d = tabs.disp
        self.hist = np.zeros((3,) + shp, dtype=dt)
        if d is not None:
            # Re E and Im E each drive a pole state (every update coefficient is real)
            self.P = np.zeros(d.am1.size, dtype=np.complex128)
            self.Q = np.zeros(d.am1.size, dtype=np.complex128)
            self._pole_entry = np.repeat(np.arange(d.comp.size), np.diff(d.ofs))
            self._idx4 = (d.comp.astype(np.int64),) + np.unravel_index(
                d.cell.astype(np.int64), shp)
        self._edges_ctrs = getattr(tabs, "coords", None) or [(None, None)] * 3
        self.acc = []
        for m in tabs.monitors:
            if m.kind == "plane":
                self.acc.append(np.zeros((4, m.nf, m.n1, m.n2), dtype=np.complex128))
            else:
                self.acc.append(np.zeros((6, m.nf) + tuple(m.box), dtype=np.complex128))
        self.tsamp = [np.zeros((len(m.comps), m.num_slots) + tuple(m.box), dtype=dt)
                      for m in self.sc.field_time_monitors]
        from openem_amd import flux as _flux
        self._ftgeom = [_flux.flux_time_geometry(self.sc.grid, m)
                        for m in self.sc.flux_time_monitors]
        self.ftsamp = [np.zeros(m.num_slots) for m in self.sc.flux_time_monitors]
        self._fte = [None] * len(self.sc.flux_time_monitors)
        self.n = 0

    # -- derivatives ------------------------------------------------------------------------
    def _dfwd(self, F, a):
        m = self._m[a]
        return (np.take(F, m["nxt"], axis=a) * m["mnx"] - F) * m["idp"]

    def _dbwd(self, F, a):
        m = self._m[a]
        return (F - np.take(F, m["prv"], axis=a) * m["mpv"]) * m["idd"]

    def _stretch(self, psi, d, a, side):
        m = self._m[a]
        psi *= m["b" + side]
        psi += m["a" + side] * d
        return m["k" + side] * d + psi

    # -- one step ---------------------------------------------------------------------------
    def step(self):
        t, n = self.t, self.n
        if n >= t.n_total:
            raise ValueError(f"step {n} beyond the {t.n_total}-step tables")
        E, H = self.E, self.H
        ch = t.ch
        ph, st = self.psi_h, self._stretch
        sy = st(ph["hx_y"], self._dfwd(E[2], 1), 1, "H")
        sz = st(ph["hx_z"], self._dfwd(E[1], 2), 2, "H")
        hx_new = H[0] - ch * (sy - sz)
        sz = st(ph["hy_z"], self._dfwd(E[0], 2), 2, "H")
        sx = st(ph["hy_x"], self._dfwd(E[2], 0), 0, "H")
        hy_new = H[1] - ch * (sz - sx)
        sx = st(ph["hz_x"], self._dfwd(E[1], 0), 0, "H")
        sy = st(ph["hz_y"], self._dfwd(E[0], 1), 1, "H")
        H[2] -= ch * (sx - sy)
        H[0] = hx_new
        H[1] = hy_new
        self._dipoles(n, magnetic=True)
        for ps in t.plane_srcs:
            hc = 1 if ps.pol == 0 else 0
            H[hc][:, :, ps.kh] += ps.coef_h * ps.ex_inc[n]
        # ---- E phase ----
        for tf in t.tfsf:
            self._tfsf(tf, n, magnetic=True)
        for ms in self.sc.mode_sources:
            self._mode(ms, n, magnetic=True)
        d = t.disp
        if d is not None:
            en = E[self._idx4]
            dP = d.am1 * self.P + d.b * en.real[self._pole_entry]
            self.P += dP
            w = np.bincount(self._pole_entry, weights=2.0 * dP.real, minlength=d.comp.size)
            if t.bloch:
                dQ = d.am1 * self.Q + d.b * en.imag[self._pole_entry]
                self.Q += dQ
                w = w + 1j * np.bincount(self._pole_entry, weights=2.0 * dQ.real,
                                         minlength=d.comp.size)
            self.hist[self._idx4] = w
        pe = self.psi_e
        m = self._m
        ca, cb, cw, hist = t.ca, t.cb, t.cw, self.hist
        sy = st(pe["hx_y"], self._dbwd(H[2], 1), 1, "E")
        sz = st(pe["hx_z"], self._dbwd(H[1], 2), 2, "E")
        ex_new = (ca[0] * E[0] + cb[0] * (sy - sz - cw * hist[0])) * m[1]["pec"] * m[2]["pec"]
        sz = st(pe["hy_z"], self._dbwd(H[0], 2), 2, "E")
        sx = st(pe["hy_x"], self._dbwd(H[2], 0), 0, "E")
        ey_new = (ca[1] * E[1] + cb[1] * (sz - sx - cw * hist[1])) * m[2]["pec"] * m[0]["pec"]
        sx = st(pe["hz_x"], self._dbwd(H[1], 0), 0, "E")
        sy = st(pe["hz_y"], self._dbwd(H[0], 1), 1, "E")
        E[2] = (ca[2] * E[2] + cb[2] * (sx - sy - cw * hist[2])) * m[0]["pec"] * m[1]["pec"]
        E[0] = ex_new
        E[1] = ey_new
        for ps in t.plane_srcs:
            E[ps.pol][:, :, ps.ks] += cb[ps.pol][:, :, ps.ks] * ps.coef_e * ps.hy_inc[n]
        for tf in t.tfsf:
            self._tfsf(tf, n, magnetic=False)
        for ms in self.sc.mode_sources:
            self._mode(ms, n, magnetic=False)
        self._dipoles(n, magnetic=False)
        if d is not None:
            en = E[self._idx4]
            self.P += d.b * en.real[self._pole_entry]
            if t.bloch:
                self.Q += d.b * en.imag[self._pole_entry]
        self._dft(n)
        self._sample_time(n)
        self._sample_flux_time(n)
        self.n += 1

    def _dipoles(self, n, magnetic):
        """Point dipoles of one kind (``source_dft.cu:147-206``): a magnetic current sample
        ``M^n = coef amp_int[n]`` (``t = n dt``) enters the H update as ``H -= ch M``, an electric
        current sample ``J^{n+1/2} = coef amp_half[n]`` (``t = (n+1/2) dt``) the E update as
        ``E -= cb J`` with the cell's cb; points on one (component, cell) add up."""
        t = self.t
        F = self.H if magnetic else self.E
        for dp in t.dipoles:
            if bool(dp.magnetic) != magnetic:
                continue
            for c in range(3):
                sel = dp.comp == c
                cells = dp.cell[sel].astype(np.int64)
                scale = t.ch if magnetic else t.cb[c].reshape(-1)[cells]
                v = scale * dp.coef[sel] * dp.amp[dp.src_of[sel], n]
                np.subtract.at(F[c].reshape(-1), cells, v)

    def _mode(self, ms, n, magnetic):
        """Mode-source plane corrections: H side ``H_u += s_hu ch/dl Re(E_v amp_e[n])``,
        ``H_v += s_hv ch/dl Re(E_u amp_e[n])`` on the H plane; E side
        ``E_u += cb s_eu/dl_dual Re(H_v amp_h[n])``, ``E_v += cb s_ev/dl_dual Re(H_u amp_h[n])``
        on the source plane; (s_hu, s_hv, s_eu, s_ev) the TF/SF face signs, (-, +, +, -) for a
        forward and (+, -, -, +) for a backward source whose profiles carry the incident H."""
        t = self.t
        a = int(ms.axis)
        u, v = (a + 1) % 3, (a + 2) % 3
        ks = int(ms.plane_index)
        fwd = ms.direction > 0
        kh = ks - 1 if fwd else ks
        s_hu, s_hv, s_eu, s_ev = (-1.0, 1.0, 1.0, -1.0) if fwd else (1.0, -1.0, -1.0, 1.0)

        def plane(F, k, val):                      # F[..] on plane k += val (cyclic layout)
            idx = [slice(None)] * 3
            idx[a] = k
            F[tuple(idx)] += val if u < v else val.T

        if magnetic:
            amp = ms.amp_e[n]
            ch = t.ch * t.axes[a].idl_p[kh]
            take = (lambda z: z) if t.bloch else np.real
            plane(self.H[u], kh, s_hu * ch * take(ms.ez_inc * amp))
            plane(self.H[v], kh, s_hv * ch * take(ms.ey_inc * amp))
        else:
            amp = ms.amp_h[n]
            ce = t.axes[a].idl_d[ks]
            idx = [slice(None)] * 3
            idx[a] = ks
            cbu, cbv = t.cb[u][tuple(idx)], t.cb[v][tuple(idx)]
            if u > v:
                cbu, cbv = cbu.T, cbv.T
            take = (lambda z: z) if t.bloch else np.real
            plane(self.E[u], ks, cbu * s_eu * ce * take(ms.hz_inc * amp))
            plane(self.E[v], ks, cbv * s_ev * ce * take(ms.hy_inc * amp))

    def _tfsf(self, tf, n, magnetic):
        """TF/SF face corrections of ``openem/kernels/tfsf.cu:49-83`` restated per face with
        array slices (lo face: H_c[p-1] += ch e_cab E_b/dl, E_c[p] -= cb e_cab H_b/dl_dual; hi
        face: the opposite signs on H_c[p], E_c[p])."""
        t, src = self.t, tf.src
        grid_axes = [self._coords(a) for a in range(3)]
        colmap = {key: q for q, key in enumerate(tf.cols)}
        row = (tf.tab_h if magnetic else tf.tab_e)[n]
        kx, ky = src.kxy
        lo, hi, opn = src.box_lo, src.box_hi, tuple(src.open_axes)
        for a in range(3):
            if a in opn:
                continue
            for side in ("lo", "hi"):
                p = lo[a] if side == "lo" else hi[a]
                for c in range(3):
                    if c == a:
                        continue
                    b = 3 - a - c
                    sgn = _EPS3[(c, a, b)]
                    if magnetic:
                        plane = p - 1 if side == "lo" else p
                        coef = (sgn if side == "lo" else -sgn) * t.axes[a].idl_p[plane] * t.ch
                        c_edge, b_edge, comp_inc = True, False, b
                        zsrc = p
                    else:
                        plane = p
                        coef = (-sgn if side == "lo" else sgn) * t.axes[a].idl_d[p]
                        c_edge, b_edge, comp_inc = False, True, 3 + b
                        zsrc = p - 1 if side == "lo" else p
                    rc = _rng(lo, hi, t.shape, opn, c, c_edge)
                    rb = _rng(lo, hi, t.shape, opn, b, b_edge)
                    sl = [None, None, None]
                    sl[a] = plane
                    sl[c] = rc
                    sl[b] = rb
                    ii, jj, kk = np.meshgrid(*[np.atleast_1d(np.asarray(v)) for v in sl],
                                             indexing="ij")
                    if a == 2:
                        zidx = np.full(ii.shape, zsrc)
                    else:
                        zidx = kk
                    cols = np.vectorize(lambda z: colmap[(comp_inc, int(z))])(zidx)
                    inc = row[cols]
                    if kx or ky:
                        x_edge = c_edge if c == 0 else b_edge
                        y_edge = c_edge if c == 1 else b_edge
                        xs = grid_axes[0][0 if x_edge else 1][ii]
                        ys = grid_axes[1][0 if y_edge else 1][jj]
                        inc = inc * np.exp(1j * (kx * xs + ky * ys))
                    F = self.H if magnetic else self.E
                    if magnetic:
                        F[c][ii, jj, kk] += coef * (inc if t.bloch else inc.real)
                    else:
                        scale = t.cb[c][ii, jj, kk] * coef
                        F[c][ii, jj, kk] += scale * (inc if t.bloch else inc.real)

    def _coords(self, a):
        return self._edges_ctrs[a]

    def _dft(self, n):
        E, H = self.E, self.H
        for m, acc in zip(self.t.monitors, self.acc):
            p = m.phase[n]                                  # (nf, 4)
            we = (p[:, 0] + 1j * p[:, 1])[:, None, None]
            wh = (p[:, 2] + 1j * p[:, 3])[:, None, None]
            if m.kind == "plane":
                a = m.axis
                t1, t2 = [i for i in range(3) if i != a]
                sl = [slice(None)] * 3
                sl[a] = m.km
                slm = list(sl)
                slm[a] = m.km - 1
                sl, slm = tuple(sl), tuple(slm)
                vals = (E[t1][sl], E[t2][sl], 0.5 * (H[t1][sl] + H[t1][slm]),
                        0.5 * (H[t2][sl] + H[t2][slm]))
                for c in range(4):
                    acc[c] += vals[c][None] * (we if c < 2 else wh)
            else:
                (i0, j0, k0), (ni, nj, nk) = m.origin, m.box
                sl = (slice(i0, i0 + ni), slice(j0, j0 + nj), slice(k0, k0 + nk))
                for c in range(6):
                    F = E[c] if c < 3 else H[c - 3]
                    acc[c] += F[sl][None] * (we[..., None] if c < 3 else wh[..., None])

    def _sample_time(self, n):
        """Field-time monitors at the end of step ``n`` (E^{n+1}, H^{n+1/2} in hand; reference
        ``openem/kernels/source_dft.cu:272-316``): the first shot of ``m = n + 1`` writes
        ``E^m + H^{m-1/2} / 2``, the second shot of ``m = n`` adds ``H^{m+1/2} / 2``."""
        E, H = self.E, self.H
        for m, buf in zip(self.sc.field_time_monitors, self.tsamp):
            (i0, j0, k0), (ni, nj, nk) = m.origin, m.box
            sl = (slice(i0, i0 + ni), slice(j0, j0 + nj), slice(k0, k0 + nk))
            s0, s1 = m.slot_of(n + 1), m.slot_of(n)
            for q, c in enumerate(m.comps):
                F = E[c] if c < 3 else H[c - 3]
                if s0 is not None:
                    buf[q, s0] = F[sl] if c < 3 else 0.5 * F[sl]
                if s1 is not None and c >= 3:
                    buf[q, s1] += 0.5 * F[sl]

    def _sample_flux_time(self, n):
        """Flux-time monitors at the end of step ``n`` (reference
        ``openem/readout.py:20-73``): the second shot of ``m = n`` adds ``sum dS (e_u H_v - e_v
        H_u) / 2`` with the E planes kept from step ``n - 1`` and H^{n+1/2}; then the first shot
        of ``m = n + 1`` keeps the colocated planes of E^{n+1} and writes the same sum with
        H^{n+1/2} = H^{m-1/2}."""
        from openem_amd import flux as _flux
        for q, (m, g) in enumerate(zip(self.sc.flux_time_monitors, self._ftgeom)):
            (_, _, _, _, tr1), (_, _, _, _, tr2) = g["s"]
            dS = g["sign"] * np.outer(tr1, tr2)
            for shot, mref in ((1, n), (0, n + 1)):
                slot = m.slot_of(mref)
                if slot is None:
                    continue
                if shot == 0:
                    self._fte[q] = [_flux.colocated_planes(self.E, self.H, g, p)[:2]
                                    for p in range(len(g["planes"]))]
                v = 0.0
                for p, (w, *_r) in enumerate(g["planes"]):
                    _, _, hu, hv = _flux.colocated_planes(self.E, self.H, g, p)
                    eu, ev = self._fte[q][p]
                    v += w * 0.5 * float(np.sum(np.real(eu * hv - ev * hu) * dS))
                if shot == 0:
                    self.ftsamp[q][slot] = v
                else:
                    self.ftsamp[q][slot] += v

    def run(self, n_steps: int):
        for _ in range(int(n_steps)):
            self.step()
        return self


def energy_sums(sc, E, H) -> np.ndarray:
    """``[sum V|E|^2, sum V eps|E|^2, sum V|H|^2]`` accumulated in float64 (reference ``_sum2``
    over the three components, ``openem/solver.py:107-122``, ``:172-176``): V is each E
    component's Yee dual volume rounded to float32, eps the component's permittivity rounded to
    float32, and H is weighted by the same E-component volumes.  Each term is formed in the
    fields' precision, as the reference forms ``w * a * a`` on its float32 device arrays; with
    float64 fields the terms are float64."""
    vols = [v.astype(np.float32) for v in sc.grid.yee_dual_volumes()]
    eps = [np.broadcast_to(np.asarray(e, dtype=np.float64).astype(np.float32), sc.shape)
           for e in (sc.eps_ex, sc.eps_ey, sc.eps_ez)]
    single = E.dtype in (np.float32, np.complex64)
    rt = np.float32 if single else np.float64

    def term(w, a):
        w = w.astype(rt)
        if np.iscomplexobj(a):
            return (w * (a.real * a.real + a.imag * a.imag)).astype(np.float64)
        return (w * a * a).astype(np.float64)

    out = np.zeros(3)
    for c in range(3):
        ve = vols[c] * eps[c]                      # float32 product, as on the device
        out[0] += float(np.sum(term(vols[c], E[c])))
        out[1] += float(np.sum(term(ve, E[c])))
        out[2] += float(np.sum(term(vols[c], H[c])))
    return out


def plane_flux_terms(data, w1, w2) -> np.ndarray:
    """``(nf, 2)``: ``sum w1 Re(E_t1 conj H_t2)`` and ``sum w2 Re(E_t2 conj H_t1)`` of a plane
    monitor's phasors ``(4, nf, n1, n2)`` (reference ``openem/flux.py:75-110``)."""
    t1 = (np.real(data[0] * np.conj(data[3])) * w1[None]).sum(axis=(1, 2))
    t2 = (np.real(data[1] * np.conj(data[2])) * w2[None]).sum(axis=(1, 2))
    return np.stack([t1, t2], axis=1)
