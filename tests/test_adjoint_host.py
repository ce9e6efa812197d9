"""CPU tests of the float64 discrete adjoint (the checker of the device adjoint)."""

import dataclasses

import numpy as np

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
