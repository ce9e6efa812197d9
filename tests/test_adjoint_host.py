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
