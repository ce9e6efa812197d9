"""CPU tests of the float64 discrete adjoint (the checker of the device adjoint)."""

import dataclasses

import numpy as np
import pytest

from openem_amd import scenes, tables
from oracle.adjoint64 import gradient_mode_power
from oracle.yee64 import Oracle64


def _small_guide(n_steps=260):
    return scenes.waveguide_scene(shape=(30, 24, 22), n_pml=6, width_cells=6, thick_cells=4,
                                  fwidth_rel=0.3, num_steps=n_steps, nf=1)


def _J(sc, mon):
    t = tables.build(sc)
    o = Oracle64(t).run(t.n_total)
    from oracle.adjoint64 import mode_objective_weights
    K = mode_objective_weights(t, mon)
    return abs(np.sum(K * o.acc[mon][:, 0])) ** 2


def test_adjoint_gradient_matches_central_differences():
    """dJ/d eps of J = |a+|^2 (downstream mode monitor) at seeded design cells strictly between
    the source and monitor planes (eps on the monitor plane also moves the mode the monitor
    projects on), each eps component on its own, against (J(eps+h) - J(eps-h)) / 2h."""
    sc = _small_guide()
    t = tables.build(sc)
    mon = [m.name for m in t.monitors].index("fwd")
    rng = np.random.default_rng(3)
    nx, ny, nz = sc.shape
    ks = sc.mode_sources[0].plane_index
    km = t.monitors[mon].km
    design = []
    for _ in range(8):
        c = int(rng.integers(3))
        i = int(rng.integers(ks + 1, km))
        j = int(rng.integers(ny // 2 - 4, ny // 2 + 4))
        k = int(rng.integers(nz // 2 - 3, nz // 2 + 3))
        design.append((c, (i * ny + j) * nz + k))
    design = np.array(design)
    J0, grad = gradient_mode_power(sc, t, mon, design)
    h = 1e-3
    for q, (c, cell) in enumerate(design):
        vals = []
        for s in (+1, -1):
            eps = [np.array(np.broadcast_to(e, sc.shape), dtype=np.float64)
                   for e in (sc.eps_ex, sc.eps_ey, sc.eps_ez)]
            eps[c].reshape(-1)[cell] += s * h
            vals.append(_J(dataclasses.replace(sc, eps_ex=eps[0], eps_ey=eps[1],
                                               eps_ez=eps[2]), mon))
        fd = (vals[0] - vals[1]) / (2 * h)
        scale = np.abs(grad).max()
        assert abs(grad[q] - fd) <= 1e-4 * abs(fd) + 1e-6 * scale, (q, grad[q], fd, scale)


def test_design_coefficients_match_the_oracle_derivatives():
    from openem_amd import adjoint
    from oracle.adjoint64 import eps_derivatives
    sc = scenes.mini_scene(sigma_on=True, lorentz_on=False)
    t = tables.build(sc)
    rng = np.random.default_rng(0)
    comp = rng.integers(0, 3, 20)
    cell = rng.integers(0, np.prod(sc.shape), 20)
    dca, dcb, ca, cb = adjoint.design_coefficients(t, comp, cell)
    ref = eps_derivatives(t, sc)
    for q in range(20):
        c, i = int(comp[q]), int(cell[q])
        assert np.isclose(dca[q], ref[c][0].reshape(-1)[i], rtol=1e-6, atol=0)
        assert np.isclose(dcb[q], ref[c][1].reshape(-1)[i], rtol=1e-6)
        assert np.isclose(ca[q], t.ca[c].reshape(-1)[i], rtol=1e-6, atol=1e-7)
        assert np.isclose(cb[q], t.cb[c].reshape(-1)[i], rtol=1e-6)


def _fd_grad(sc, mon, design, h=1e-3):
    out = []
    for c, cell in design:
        vals = []
        for s in (+1, -1):
            eps = [np.array(np.broadcast_to(e, sc.shape), dtype=np.float64)
                   for e in (sc.eps_ex, sc.eps_ey, sc.eps_ez)]
            eps[c].reshape(-1)[cell] += s * h
            vals.append(_J(dataclasses.replace(sc, eps_ex=eps[0], eps_ey=eps[1],
                                               eps_ez=eps[2]), mon))
        out.append((vals[0] - vals[1]) / (2 * h))
    return np.array(out)


def _pole_design(sc):
    """Five design entries between the source (x = 12) and the monitor (x = 26): four in the
    pole block (x 15..23), one before it."""
    ny, nz = sc.shape[1], sc.shape[2]
    ijk = [(2, 23, 12, 13), (1, 20, 10, 12), (0, 17, 13, 14), (1, 14, 12, 12), (2, 22, 11, 10)]
    return np.array([(c, (i * ny + j) * nz + k) for c, i, j, k in ijk])


@pytest.mark.parametrize("two", [False, True], ids=["one-material", "two-materials"])
def test_dispersive_adjoint_matches_central_differences(two):
    """With ADE poles (a Lorentz block on the guide, design entries inside and outside it) the
    adjoint gradient equals central differences of the oracle's J to 1e-6 (measured 1e-9 -
    2e-6): the transposed pole updates and the eps derivative through ca and cb with the pole
    coupling G in their denominator."""
    sc = scenes.pole_guide_scene(two_materials=two)
    t = tables.build(sc)
    mon = [m.name for m in t.monitors].index("fwd")
    design = _pole_design(sc)
    _, grad = gradient_mode_power(sc, t, mon, design)
    fd = _fd_grad(sc, mon, design)
    rel = np.abs(grad - fd) / np.abs(fd)
    print("relative error per entry:", rel)
    assert rel.max() < 1e-5


def test_planted_pole_transpose_mistakes_fail_the_central_differences(monkeypatch):
    """Mistakes planted in a copy of the oracle's pole transposes: lambda_E not absorbing the post
    update, A-1 or B not conjugated, the 2 of hist = sum 2 Re dP dropped.  Each misses the
    central differences by more than 1% on some entry (measured 19-190%)."""
    import oracle.adjoint64 as A
    sc = scenes.pole_guide_scene()
    t = tables.build(sc)
    mon = [m.name for m in t.monitors].index("fwd")
    design = _pole_design(sc)[:3]
    fd = _fd_grad(sc, mon, design)
    orig = A.Adjoint64

    def planted(edit):
        class Planted(orig):
            def __init__(self, tabs):
                super().__init__(tabs)
                edit(self)
        return Planted

    def swap(**fns):
        def edit(self):
            d = self.t.disp
            self.t = dataclasses.replace(
                self.t, disp=dataclasses.replace(d, **{k: f(getattr(d, k)) for k, f in fns.items()}))
        return edit

    cases = {
        "no post transpose": lambda self: setattr(self, "absorb_poles", lambda: None),
        "A-1 not conjugated": swap(am1=np.conj),
        "B not conjugated": swap(b=np.conj),
        "hist without its 2": lambda self: setattr(self, "t",
                                                   dataclasses.replace(self.t, cw=self.t.cw / 2)),
    }
    for name, edit in cases.items():
        monkeypatch.setattr(A, "Adjoint64", planted(edit))
        _, g = A.gradient_mode_power(sc, t, mon, design)
        rel = np.abs(g - fd) / np.abs(fd)
        print(name, rel)
        assert rel.max() > 1e-2, name


def test_design_coefficients_with_poles_match_the_oracle_derivatives():
    """At entries with poles the eps derivatives carry G: (2 s + G)/den^2, -(dt/eps0)/den^2."""
    from openem_amd import adjoint
    from oracle.adjoint64 import eps_derivatives
    sc = scenes.mini_scene(sigma_on=True, lorentz_on=True)
    t = tables.build(sc)
    d = sc.dispersion
    rng = np.random.default_rng(1)
    pick = rng.choice(d.n_entry, 12, replace=False)
    comp = np.concatenate([d.comp[pick], rng.integers(0, 3, 8)])
    cell = np.concatenate([d.cell[pick], rng.integers(0, np.prod(sc.shape), 8)])
    dca, dcb, ca, cb = adjoint.design_coefficients(t, comp, cell)
    ref = eps_derivatives(t, sc)
    for q in range(comp.size):
        c, i = int(comp[q]), int(cell[q])
        assert np.isclose(dca[q], ref[c][0].reshape(-1)[i], rtol=1e-6, atol=0)
        assert np.isclose(dcb[q], ref[c][1].reshape(-1)[i], rtol=1e-6)
        assert np.isclose(cb[q], t.cb[c].reshape(-1)[i], rtol=1e-6)
    # entries with poles: dca differs from the pole-free 2 s / (eps + s)^2 by G
    g = adjoint._design_g(sc, comp[:12], cell[:12])
    assert np.all(g != 0) and np.all(adjoint._design_g(sc, comp[12:], cell[12:])[
        ~np.isin(comp[12:] * np.prod(sc.shape) + cell[12:], d.comp * np.prod(sc.shape) + d.cell)] == 0)


def _bloch_cell(n_steps=260):
    """12 x 12 x 48 Bloch cell (30 deg oblique incidence, CPML z, oblique TF/SF box over z 12-31,
    one Drude + Lorentz metal cylinder on the substrate)."""
    return scenes.bloch_cell_scene(nxy=12, nz=48, n_pml=8, num_steps=n_steps, with_metal=True,
                                   nf=3, n_cyl=1)


def _bloch_objective(t):
    """Seeded complex weights over the transmission plane's four phasor slots."""
    mon = [m.name for m in t.monitors].index("T")
    m = t.monitors[mon]
    rng = np.random.default_rng(2)
    K = rng.standard_normal((4, m.n1, m.n2)) + 1j * rng.standard_normal((4, m.n1, m.n2))
    return mon, K


def _bloch_design(sc):
    """Two metal cells (poles and conductivity), cells on the x wrap planes (i = 0, i = nx - 1)
    and the y wrap plane (j = 0), one interior cell: all inside the TF/SF box, off its faces."""
    nx, ny, nz = sc.shape
    d = sc.dispersion
    metal = [(int(c), int(x)) for c, x in zip(d.comp, d.cell)]
    ijk = [(2, 0, 5, 20), (1, nx - 1, 0, 25), (0, 6, 0, 14), (2, 4, 9, 28)]
    return np.array([metal[7], metal[len(metal) // 2]]
                    + [(c, (i * ny + j) * nz + k) for c, i, j, k in ijk])


def _bloch_fd(sc, mon, K, design, h=1e-3):
    def J(s):
        tt = tables.build(s)
        o = Oracle64(tt).run(tt.n_total)
        return abs(np.sum(K * o.acc[mon][:, 0])) ** 2
    out = []
    for c, cell in design:
        vals = []
        for s in (+1, -1):
            eps = [np.array(np.broadcast_to(e, sc.shape), dtype=np.float64)
                   for e in (sc.eps_ex, sc.eps_ey, sc.eps_ez)]
            eps[c].reshape(-1)[cell] += s * h
            vals.append(J(dataclasses.replace(sc, eps_ex=eps[0], eps_ey=eps[1], eps_ez=eps[2])))
        out.append((vals[0] - vals[1]) / (2 * h))
    return np.array(out)


def test_bloch_adjoint_matches_central_differences():
    """Complex fields: the conjugate-transposed step (wrap masks conjugated, pole adjoints of the
    Re E and Im E driven states) gives the gradient of J = |sum K . phasors|^2 on the Bloch cell
    to 1e-5 of central differences (measured 1e-10 - 1.1e-6), at metal and wrap-plane cells."""
    from oracle.adjoint64 import gradient_overlap_power
    sc = _bloch_cell()
    t = tables.build(sc)
    assert t.bloch
    mon, K = _bloch_objective(t)
    design = _bloch_design(sc)
    _, grad = gradient_overlap_power(sc, t, mon, K, design)
    fd = _bloch_fd(sc, mon, K, design)
    rel = np.abs(grad - fd) / np.abs(fd)
    print("relative error per entry:", rel)
    assert rel.max() < 1e-5


def test_planted_bloch_transpose_mistakes_fail_the_central_differences(monkeypatch):
    """Mistakes planted in a copy of the oracle's complex transposes: the wrap masks not
    conjugated (misses by 3-700%), the Im E driven pole state's terms in lambda_E dropped (the
    transposed step then amplifies lambda by ~20x per step and overflows).  Neither stays within
    1% of the central differences on every entry."""
    import oracle.adjoint64 as A
    sc = _bloch_cell()
    t = tables.build(sc)
    mon, K = _bloch_objective(t)
    design = _bloch_design(sc)[[0, 2, 3]]
    fd = _bloch_fd(sc, mon, K, design)
    orig = A.Adjoint64

    def masks_not_conjugated(self):
        for m in self.o._m:
            m["mnx"], m["mpv"] = np.conj(m["mnx"]), np.conj(m["mpv"])

    def im_pole_adjoint_dropped(self):
        b = np.conj(self.t.disp.b)
        self._pole_terms = lambda lp, lq: self._entry_sum(np.real(b * lp))

    for name, edit in (("masks not conjugated", masks_not_conjugated),
                       ("Im E pole adjoint dropped", im_pole_adjoint_dropped)):
        class Planted(orig):
            def __init__(self, tabs, _edit=edit):
                super().__init__(tabs)
                _edit(self)
        monkeypatch.setattr(A, "Adjoint64", Planted)
        with np.errstate(over="ignore", invalid="ignore"):
            _, g = A.gradient_overlap_power(sc, t, mon, K, design)
            rel = np.abs(g - fd) / np.abs(fd)
        print(name, rel)
        assert not np.all(rel <= 1e-2), name


def test_overlap_gradient_checks_its_arguments():
    """The plane-projection gradient refuses weights of the wrong shape and a monitor that is
    not a plane before it touches the device."""
    from openem_amd import adjoint
    sc = _small_guide(40)
    t = tables.build(sc)
    m = t.monitors[[q.name for q in t.monitors].index("fwd_flux")]
    with pytest.raises(ValueError, match="weights of shape"):
        adjoint.overlap_power_gradient(sc, "fwd_flux", np.ones((4, m.n1, m.n2 + 1)), [0], [0])
    box = scenes.mini_scene()
    name = box.field_monitors[0].name
    with pytest.raises(ValueError, match="not a plane monitor"):
        adjoint.overlap_power_gradient(box, name, np.ones((4, 1, 1)), [0], [0])
