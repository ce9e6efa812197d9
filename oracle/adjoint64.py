"""Float64 discrete adjoint of the oracle's time step.  TEST INFRASTRUCTURE ONLY.

Only ``tests/`` may import this module (it checks the device adjoint).

The forward step of :class:`oracle.yee64.Oracle64` on a real scene is linear in the state
``(E, H, psi_H, psi_E, P)`` plus additive source terms, with the E update
``E' = pec (ca E + cb (S - cw hist))`` depending on eps through ``ca`` and ``cb``
(reference ``openem/kernels/yee.cu:154-240``, coefficients ``openem/coeffs.py:98-136``).  For an
objective ``J`` of the DFT phasors the exact gradient of the discrete run is

    dJ/deps = sum_n lambda_E^{n+1} . (d ca/d eps E^n + d cb/d eps X^n),
    X^n = (E^{n+1} - ca E^n) / cb,

where ``lambda`` runs backwards through the transposed step: the objective's gradient with
respect to E^{n+1} and H^{n+1/2} is added, then the transpose of the ADE post update, then the
transpose of the E update (its backward differences become forward differences of
``cb pec lambda_E``, the CPML recursion transposes onto the psi adjoints), then the transpose of
the ADE pre update, then the transpose of the H update.  This is the "discrete adjoint" that
tidy3d's ``grad`` stands for in the reference (``openem/nb/autograd_hook.py:1-28``); the
reference hands it to tidy3d, here it is written out.

ADE poles (``openem/kernels/dispersion.cu:68-147``), per pole entry of component c at a cell:
pre ``dP = (A-1) P + B E^n``, ``P_mid = P + dP``, ``hist = sum 2 Re dP``; post
``P^{n+1} = P_mid + B E^{n+1}``.  P is complex and E real, so with the adjoint of P written as
one complex number ``lP = dJ/d Re P + i dJ/d Im P``: a product ``a z`` transposes to
``conj(a) lz``, ``B e`` to ``Re(conj(B) lz)`` and ``2 Re dP`` to ``2 l`` on the real part.

Bloch (complex-field) scenes: every coefficient of the step is real except the wrap masks of
the Bloch axes (the phase ``exp(i k L)`` the neighbour across the boundary carries), so the step
is complex-linear and its adjoint, in the same complex convention ``lambda = dJ/d Re + i dJ/d Im``,
is the conjugate transpose: the transposes above with the masks conjugated.  The pole states
driven by Re E (P) and by Im E (Q) each have their adjoint; hist = sum 2 Re dP + i sum 2 Re dQ
hands the real part of lambda_hist to P and the imaginary part to Q.  The eps gradient is then
``Re(conj(lambda_E^{n+1}) (ca' E^n + cb' X^n))``.
"""

from __future__ import annotations

import numpy as np

from .yee64 import Oracle64, _PSI, _bc

# Terms of each update as (source field component, derivative axis, psi key), first with +,
# second with - (oracle.yee64.Oracle64.step):
#   H_c -= ch (s_first - s_second),  s = stretch(psi_H[key], dfwd(E_src, axis))
#   E_c  = ca E_c + cb (s_first - s_second),  s = stretch(psi_E[key], dbwd(H_src, axis))
_TERMS = {0: ((2, 1, "hx_y"), (1, 2, "hx_z")),
          1: ((0, 2, "hy_z"), (2, 0, "hy_x")),
          2: ((1, 0, "hz_x"), (0, 1, "hz_y"))}
_H_TERMS = _E_TERMS = _TERMS


class Adjoint64:
    """Backward sweep over an :class:`Oracle64` run (real fields, with or without ADE poles)."""

    def __init__(self, tabs):
        self.t = tabs
        self.o = Oracle64(tabs)                      # for the difference operators and masks
        shp = tabs.shape
        dt = np.complex128 if tabs.bloch else np.float64
        self.lE = np.zeros((3,) + shp, dtype=dt)
        self.lH = np.zeros((3,) + shp, dtype=dt)
        self.lpe = {k: np.zeros(shp, dtype=dt) for k in _PSI}
        self.lph = {k: np.zeros(shp, dtype=dt) for k in _PSI}
        d = tabs.disp
        if d is not None:
            self.lP = np.zeros(d.am1.size, dtype=np.complex128)    # driven by Re E
            self.lQ = np.zeros(d.am1.size, dtype=np.complex128)    # driven by Im E (Bloch)
            self._entry = self.o._pole_entry          # pole -> entry
            self._idx4 = self.o._idx4                 # entry -> (comp, i, j, k)

    @staticmethod
    def _scatter_add(acc, idx, src):
        """``acc[idx[i]] += src[i]`` along the first axis, repeated targets summed (the end
        cells of a non-periodic axis name one neighbour twice): one fancy-indexed add over the
        first occurrence of every target, the repeats one by one."""
        _, first = np.unique(idx, return_index=True)
        acc[idx[first]] += src[first]
        for i in np.setdiff1d(np.arange(idx.size), first):
            acc[idx[i]] += src[i]

    # transposes of the oracle's one-sided differences along axis a
    def _dfwd_T(self, g, a):
        m = self.o._m[a]
        idl = _bc(self.t.axes[a].idl_p, a)
        w = g * idl
        out = -w
        src = np.moveaxis(w * np.conj(_bc(m["mnx"], a)), a, 0)
        acc = np.moveaxis(out, a, 0)
        self._scatter_add(acc, m["nxt"], src)
        return np.moveaxis(acc, 0, a)

    def _dbwd_T(self, g, a):
        m = self.o._m[a]
        idl = _bc(self.t.axes[a].idl_d, a)
        w = g * idl
        out = w.copy()
        src = np.moveaxis(-w * np.conj(_bc(m["mpv"], a)), a, 0)
        acc = np.moveaxis(out, a, 0)
        self._scatter_add(acc, m["prv"], src)
        return np.moveaxis(acc, 0, a)

    def _coef(self, a, side):
        m = self.o._m[a]
        return m["a" + side], m["b" + side], m["k" + side]

    def _entry_sum(self, v):
        """Per entry sum of the real per-pole values ``v``."""
        return np.bincount(self._entry, weights=v, minlength=self.t.disp.comp.size)

    def _pole_terms(self, lp, lq):
        """``sum_p Re(conj(B_p) l_p)`` per entry for the Re E driven state (+ i times that of
        the Im E driven one on Bloch scenes)."""
        b = np.conj(self.t.disp.b)
        r = self._entry_sum(np.real(b * lp))
        return r + 1j * self._entry_sum(np.real(b * lq)) if self.t.bloch else r

    def absorb_poles(self):
        """Transpose of the ADE post update ``P^{n+1} = P_mid + B E^{n+1}``: E^{n+1} also feeds
        the pole state, so lambda_E^{n+1} gains ``sum_p Re(conj(B_p) lP_p)`` at every pole entry
        (``lP`` passes to P_mid unchanged).  After it, lambda_E is the full dJ/dE^{n+1} the eps
        gradient takes.  No-op without poles."""
        d = self.t.disp
        if d is not None:
            self.lE[self._idx4] += self._pole_terms(self.lP, self.lQ)

    def step_back(self, gE=None, gH=None):
        """Adjoint of one forward step: ``gE``/``gH`` are dJ/dE^{n+1}, dJ/dH^{n+1/2} of this
        step's samples (added first); leaves lambda for the state before the step."""
        if gE is not None:
            self.lE += gE
        if gH is not None:
            self.lH += gH
        self.absorb_poles()
        self.transpose()

    def transpose(self):
        """The E update, ADE pre and H update transposed (lambda_E already absorbed the ADE
        post, :meth:`absorb_poles`)."""
        t, m = self.t, self.o._m
        lE, lH = self.lE, self.lH
        # E update transpose
        pec = [m[1]["pec"] * m[2]["pec"], m[2]["pec"] * m[0]["pec"], m[0]["pec"] * m[1]["pec"]]
        nu = [t.cb[c] * pec[c] * lE[c] for c in range(3)]
        dH = np.zeros_like(lH)
        new_pe = {}
        for c in range(3):
            for sgn, (b, a, key) in zip((1.0, -1.0), _E_TERMS[c]):
                ac, bc, kc = self._coef(a, "E")
                f = sgn * nu[c]
                q = (kc + ac) * f + ac * self.lpe[key]
                dH[b] += self._dbwd_T(q, a)
                new_pe[key] = bc * (f + self.lpe[key])
        for c in range(3):
            lE[c] = t.ca[c] * pec[c] * lE[c]
        self.lpe = new_pe
        d = t.disp
        if d is not None:
            # ADE pre transposed: hist enters E^{n+1} as -cb cw pec hist, so lhist = -cw nu;
            # dP feeds P_mid (identity) and hist (2 Re); P^n feeds P_mid and dP, E^n feeds dP
            lhist = -t.cw * np.stack(nu)[self._idx4]
            ldp = self.lP + 2.0 * np.real(lhist)[self._entry]
            ldq = self.lQ + 2.0 * np.imag(lhist)[self._entry]
            self.lP = self.lP + np.conj(d.am1) * ldp
            if t.bloch:
                self.lQ = self.lQ + np.conj(d.am1) * ldq
            lE[self._idx4] += self._pole_terms(ldp, ldq)
        lH += dH
        # H update transpose: H_c' = H_c - ch (s1 - s2)
        dE = np.zeros_like(lE)
        new_ph = {}
        for c in range(3):
            for sgn, (b, a, key) in zip((1.0, -1.0), _H_TERMS[c]):
                ac, bc, kc = self._coef(a, "H")
                f = -t.ch * sgn * lH[c]
                q = (kc + ac) * f + ac * self.lph[key]
                dE[b] += self._dfwd_T(q, a)
                new_ph[key] = bc * (f + self.lph[key])
        self.lph = new_ph
        lE += dE


def eps_derivatives(tabs, sc):
    """Per component (d ca/d eps, d cb/d eps) of ``ca = (eps - s)/(eps + s + G)``,
    ``cb = (dt/eps0)/(eps + s + G)``: ``((2 s + G)/den^2, -(dt/eps0)/den^2)`` (G, the ADE
    coupling ``sum 2 Re B`` of a pole entry, does not depend on eps)."""
    from openem_amd import cpml
    out = []
    sig = (sc.sigma_ex, sc.sigma_ey, sc.sigma_ez)
    eps = (sc.eps_ex, sc.eps_ey, sc.eps_ez)
    gg = sc.dispersion.g_dense(sc.shape) if sc.any_dispersion else (0.0, 0.0, 0.0)
    for c in range(3):
        e = np.broadcast_to(np.asarray(eps[c], dtype=np.float64), sc.shape)
        s = 0.0 if sig[c] is None else np.asarray(sig[c]) * sc.dt / (2 * cpml.EPSILON_0)
        den = e + s + gg[c]
        out.append(((2 * s + gg[c]) / den ** 2 + 0 * e,
                    -(sc.dt / cpml.EPSILON_0) / den ** 2 + 0 * e))
    return out


def mode_objective_weights(tabs, mon_index: int, f_index: int = 0):
    """``K`` (4, n1, n2) complex: a+ = sum_n sum_c K_c . sample_c^n e^{i w t_c(n)} for the mode
    monitor ``mon_index`` at frequency ``f_index`` (the set sum of the overlap weights)."""
    from openem_amd import modes
    m = tabs.monitors[mon_index]
    g = modes.overlap_weight_sets(m.modes, tabs.scene.grid)
    sign = modes.orientation(m.axis)
    return 0.25 * sign * (g[0, :, f_index] + g[1, :, f_index])


def gradient_overlap_power(sc, tabs, mon_index: int, K: np.ndarray, design: np.ndarray,
                           f_index: int = 0):
    """J = |a|^2, ``a = sum K . phasors`` of plane monitor ``mon_index`` at frequency ``f_index``
    (``K`` (4, n1, n2) complex over its E t1, E t2, H t1, H t2 slots) after ``tabs.n_total``
    steps, and dJ/d eps at ``design`` ((n, 2) int: component, flat cell), by a float64 adjoint
    over the stored trajectory of E at the design entries (E^n for n = 0 .. n_total, the only
    forward values the gradient needs).  A sample x of slot c at step n enters ``a`` as
    ``K_c w_n x`` (w_n the slot's DFT weight), so dJ/dx = ``2 a conj(K_c w_n)`` in the complex
    convention (its real part for real fields); the H slots average the planes km and km - 1."""
    from openem_amd.grid import transverse_axes
    o = Oracle64(tabs)
    n_total = tabs.n_total
    dc = np.asarray(design[:, 0], dtype=np.int64)
    dcell = np.asarray(design[:, 1], dtype=np.int64)
    flat = dc * int(np.prod(tabs.shape)) + dcell          # into E.reshape(-1)
    traj = np.empty((n_total + 1, len(design)), dtype=o.E.dtype)
    traj[0] = o.E.reshape(-1)[flat]
    for n in range(n_total):
        o.step()
        traj[n + 1] = o.E.reshape(-1)[flat]
    m = tabs.monitors[mon_index]
    ph = m.phase[:, f_index]                              # (n_total, 4): cos/sin at t_e, t_h
    data = o.acc[mon_index][:, f_index]
    a = np.sum(K * data)
    J = abs(a) ** 2
    t1, t2 = transverse_axes(m.axis)
    adj = Adjoint64(tabs)
    dca_dcb = eps_derivatives(tabs, sc)
    ca = np.stack([tabs.ca[c] for c in range(3)]).reshape(-1)[flat]
    cb = np.stack([tabs.cb[c] for c in range(3)]).reshape(-1)[flat]
    dca = np.stack([dca_dcb[c][0] for c in range(3)]).reshape(-1)[flat]
    dcb = np.stack([dca_dcb[c][1] for c in range(3)]).reshape(-1)[flat]
    grad = np.zeros(len(design))
    plane = [slice(None)] * 3
    below = [slice(None)] * 3
    plane[m.axis], below[m.axis] = m.km, m.km - 1
    plane, below = tuple(plane), tuple(below)
    part = (lambda v: v) if tabs.bloch else np.real
    for n in range(n_total - 1, -1, -1):
        we = ph[n, 0] + 1j * ph[n, 1]
        wh = ph[n, 2] + 1j * ph[n, 3]
        # dJ/dE^{n+1} on the monitor plane, dJ/dH^{n+1/2} split over the two planes the
        # colocated H sample averages
        for slot, comp in ((0, t1), (1, t2)):
            adj.lE[comp][plane] += part(2.0 * a * np.conj(K[slot] * we))
        for slot, comp in ((2, t1), (3, t2)):
            v = part(2.0 * a * np.conj(K[slot] * wh)) * 0.5
            adj.lH[comp][plane] += v
            adj.lH[comp][below] += v
        adj.absorb_poles()
        e0, e1 = traj[n], traj[n + 1]
        grad += np.real(np.conj(adj.lE.reshape(-1)[flat]) * (dca * e0 + dcb * (e1 - ca * e0) / cb))
        adj.transpose()
    return J, grad


def gradient_mode_power(sc, tabs, mon_index: int, design: np.ndarray, f_index: int = 0):
    """J = |a+|^2 of a mode monitor (the overlap with its mode's weights) and dJ/d eps at
    ``design``: :func:`gradient_overlap_power` with ``K`` from :func:`mode_objective_weights`."""
    K = mode_objective_weights(tabs, mon_index, f_index)
    return gradient_overlap_power(sc, tabs, mon_index, K, design, f_index)
