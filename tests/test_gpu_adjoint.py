"""GPU tests of the device discrete adjoint.

Tolerances: the float32 device gradient against the float64 oracle adjoint of the same tables,
1e-3 of the largest |gradient| (fp32 stepping of fields and adjoint fields); segmented
checkpoint recomputation against one segment, 1e-12 (the recomputation is deterministic).
"""

import numpy as np
import pytest

from openem_amd import adjoint, scenes, tables
from oracle.adjoint64 import gradient_mode_power

pytestmark = pytest.mark.gpu


def _guide(n=260):
    return scenes.waveguide_scene(shape=(30, 24, 22), n_pml=6, width_cells=6, thick_cells=4,
                                  fwidth_rel=0.3, num_steps=n, nf=1)


def _design(sc, t, mon, count=12, seed=5):
    rng = np.random.default_rng(seed)
    ny, nz = sc.shape[1], sc.shape[2]
    ks, km = sc.mode_sources[0].plane_index, t.monitors[mon].km
    comp, cell = [], []
    for _ in range(count):
        comp.append(int(rng.integers(3)))
        i = int(rng.integers(ks + 1, km))
        j = int(rng.integers(ny // 2 - 4, ny // 2 + 4))
        k = int(rng.integers(nz // 2 - 3, nz // 2 + 3))
        cell.append((i * ny + j) * nz + k)
    return np.array(comp), np.array(cell)


def test_device_adjoint_matches_the_oracle_adjoint(gpu_lib):
    sc = _guide()
    t = tables.build(sc)
    mon = [m.name for m in t.monitors].index("fwd")
    comp, cell = _design(sc, t, mon)
    print(f"adjoint: shape {sc.shape}, {t.n_total} steps, {comp.size} design entries", flush=True)
    g = adjoint.mode_power_gradient(sc, "fwd", comp, cell, checkpoint_every=64)
    J0, ref = gradient_mode_power(sc, t, mon, np.stack([comp, cell], axis=1))
    assert abs(g.J - J0) <= 1e-4 * J0, (g.J, J0)
    scale = np.abs(ref).max()
    assert scale > 0
    err = np.abs(g.grad - ref).max()
    print(f"max |grad| {scale:.4e}, max err {err:.3e}", flush=True)
    assert err <= 1e-3 * scale, (g.grad, ref)


def test_checkpoint_segments_do_not_change_the_gradient(gpu_lib):
    sc = _guide(200)
    t = tables.build(sc)
    mon = [m.name for m in t.monitors].index("fwd")
    comp, cell = _design(sc, t, mon, count=6, seed=9)
    one = adjoint.mode_power_gradient(sc, "fwd", comp, cell, checkpoint_every=200, record_bytes=0)
    seg = adjoint.mode_power_gradient(sc, "fwd", comp, cell, checkpoint_every=48, record_bytes=0)
    assert seg.segments == 5 and one.segments == 1
    assert np.allclose(seg.grad, one.grad, rtol=1e-12, atol=0)
    assert seg.J == one.J


def test_recording_forward_run_equals_checkpoint_recomputation(gpu_lib):
    """Recording E at the design during the forward run (no checkpoints) gives the gradient of
    checkpointed segments recomputed one by one, bit for bit: the same steps, recorded values and
    transposed steps."""
    sc = _guide(200)
    t = tables.build(sc)
    mon = [m.name for m in t.monitors].index("fwd")
    comp, cell = _design(sc, t, mon, count=6, seed=9)
    rec = adjoint.mode_power_gradient(sc, "fwd", comp, cell)
    seg = adjoint.mode_power_gradient(sc, "fwd", comp, cell, checkpoint_every=48, record_bytes=0)
    assert rec.segments == 1 and seg.segments == 5
    assert rec.J == seg.J
    assert np.array_equal(rec.grad, seg.grad)


@pytest.mark.parametrize("chunk", ["3", "30"])
def test_transpose_chunks_do_not_change_the_gradient(gpu_lib, monkeypatch, chunk):
    """The transposed sweeps march x in chunks of planes (8 by default); each chunk recomputes
    the terms it carries across its boundary plane, so 3-plane chunks and one chunk over the
    whole x extent (30 planes) give the default's gradient bit for bit."""
    sc = _guide(200)
    t = tables.build(sc)
    mon = [m.name for m in t.monitors].index("fwd")
    comp, cell = _design(sc, t, mon, count=6, seed=9)
    monkeypatch.delenv("OEM_ADJ_CHUNK", raising=False)
    ref = adjoint.mode_power_gradient(sc, "fwd", comp, cell)
    monkeypatch.setenv("OEM_ADJ_CHUNK", chunk)
    alt = adjoint.mode_power_gradient(sc, "fwd", comp, cell)
    assert np.abs(ref.grad).max() > 0
    assert alt.J == ref.J
    assert np.array_equal(alt.grad, ref.grad)


def test_objective_lists_in_one_launch_equal_one_launch_each(gpu_lib, monkeypatch):
    """The objective's point lists touch disjoint cells, so one launch applies them all per
    backward step; the gradient equals that of one launch per list (OEM_ADJ_OBJ_ALL=0) bit for
    bit."""
    sc = _guide(200)
    t = tables.build(sc)
    mon = [m.name for m in t.monitors].index("fwd")
    comp, cell = _design(sc, t, mon, count=6, seed=9)
    monkeypatch.delenv("OEM_ADJ_OBJ_ALL", raising=False)
    ref = adjoint.mode_power_gradient(sc, "fwd", comp, cell)
    monkeypatch.setenv("OEM_ADJ_OBJ_ALL", "0")
    alt = adjoint.mode_power_gradient(sc, "fwd", comp, cell)
    assert np.abs(ref.grad).max() > 0
    assert alt.J == ref.J
    assert np.array_equal(alt.grad, ref.grad)


def _pole_design(sc):
    """Eight design entries between the source (x = 12) and the monitor (x = 26) of the pole
    guide: six in the pole block (x 15..23), two before it."""
    ny, nz = sc.shape[1], sc.shape[2]
    ijk = [(2, 23, 12, 13), (1, 20, 10, 12), (0, 17, 13, 14), (1, 14, 12, 12), (2, 22, 11, 10),
           (0, 19, 12, 11), (2, 16, 9, 12), (1, 13, 13, 11)]
    return (np.array([c for c, *_ in ijk]),
            np.array([(i * ny + j) * nz + k for _, i, j, k in ijk]))


@pytest.mark.parametrize("form", ["compact", "csr", "fused"])
def test_dispersive_adjoint_matches_the_oracle_adjoint(gpu_lib, monkeypatch, form):
    """ADE poles under the adjoint: a Lorentz block on the guide, the gradient through the
    device's transposed pole updates and checkpoints that carry the pole state (segments of 64
    steps) against the float64 oracle adjoint, 1e-3 of the largest |gradient|.  compact: one
    2-pole material (the uniform pole layout); csr: 1- and 2-pole materials mixed (the general
    layout); fused: the compact scene on the fused H+E step."""
    sc = scenes.pole_guide_scene(two_materials=form == "csr")
    if form == "fused":
        monkeypatch.setenv("OEM_FUSE", "1")
    t = tables.build(sc)
    mon = [m.name for m in t.monitors].index("fwd")
    comp, cell = _pole_design(sc)
    print(f"dispersive adjoint ({form}): shape {sc.shape}, {sc.dispersion.n_entry} pole entries, "
          f"{t.n_total} steps", flush=True)
    g = adjoint.mode_power_gradient(sc, "fwd", comp, cell, checkpoint_every=64, record_bytes=0)
    assert g.segments == 5
    J0, ref = gradient_mode_power(sc, t, mon, np.stack([comp, cell], axis=1))
    assert abs(g.J - J0) <= 1e-4 * J0, (g.J, J0)
    scale = np.abs(ref).max()
    err = np.abs(g.grad - ref).max()
    print(f"max |grad| {scale:.4e}, max err {err:.3e} ({err / scale:.2e} of it)", flush=True)
    assert err <= 1e-3 * scale, (g.grad, ref)


def test_dispersive_recording_equals_checkpoint_recomputation(gpu_lib):
    """With poles, the recording forward run and checkpointed segments recomputed one by one
    give the same gradient bit for bit: a checkpoint restores the pole state as the run left
    it, and a recomputed segment ends with the plain post update as a run does."""
    sc = scenes.pole_guide_scene(200)
    comp, cell = _pole_design(sc)
    rec = adjoint.mode_power_gradient(sc, "fwd", comp, cell)
    seg = adjoint.mode_power_gradient(sc, "fwd", comp, cell, checkpoint_every=48, record_bytes=0)
    assert rec.segments == 1 and seg.segments == 5
    assert np.abs(rec.grad).max() > 0
    assert rec.J == seg.J
    assert np.array_equal(rec.grad, seg.grad)


def _bloch_case():
    """The 12 x 12 x 48 Bloch cell of the CPU adjoint tests (oblique TF/SF, Drude + Lorentz
    metal cylinder), seeded weights over its transmission plane, and design entries at metal
    cells, on the x and y wrap planes and inside."""
    sc = scenes.bloch_cell_scene(nxy=12, nz=48, n_pml=8, num_steps=260, with_metal=True, nf=3,
                                 n_cyl=1)
    t = tables.build(sc)
    mon = [m.name for m in t.monitors].index("T")
    m = t.monitors[mon]
    rng = np.random.default_rng(2)
    K = rng.standard_normal((4, m.n1, m.n2)) + 1j * rng.standard_normal((4, m.n1, m.n2))
    nx, ny, nz = sc.shape
    d = sc.dispersion
    metal = [(int(c), int(x)) for c, x in zip(d.comp, d.cell)]
    ijk = [(2, 0, 5, 20), (1, nx - 1, 0, 25), (0, 6, 0, 14), (2, 4, 9, 28), (0, 11, 11, 16),
           (1, 3, 7, 22)]
    design = np.array([metal[7], metal[len(metal) // 2], metal[-3]]
                      + [(c, (i * ny + j) * nz + k) for c, i, j, k in ijk])
    return sc, t, mon, K, design


@pytest.mark.parametrize("form", ["sweeps", "fused"])
def test_bloch_adjoint_matches_the_oracle_adjoint(gpu_lib, monkeypatch, form):
    """Complex (Bloch) fields under the adjoint: the conjugate-transposed step on the device
    (wrap masks conjugated, the Im E driven pole state's adjoint, complex records and objective)
    against the float64 oracle adjoint, through checkpoint segments of 64 steps, 1e-3 of the
    largest |gradient|; on the two sweeps and on the fused step."""
    from oracle.adjoint64 import gradient_overlap_power
    sc, t, mon, K, design = _bloch_case()
    monkeypatch.setenv("OEM_FUSE", "1" if form == "fused" else "0")
    print(f"Bloch adjoint ({form}): shape {sc.shape}, {sc.dispersion.n_entry} pole entries, "
          f"{t.n_total} steps, {len(design)} design entries", flush=True)
    g = adjoint.overlap_power_gradient(sc, "T", K, design[:, 0], design[:, 1],
                                       checkpoint_every=64, record_bytes=0)
    assert g.segments == 5
    J0, ref = gradient_overlap_power(sc, t, mon, K, design)
    assert abs(g.J - J0) <= 1e-4 * J0, (g.J, J0)
    scale = np.abs(ref).max()
    err = np.abs(g.grad - ref).max()
    print(f"max |grad| {scale:.4e}, max err {err:.3e} ({err / scale:.2e} of it)", flush=True)
    assert err <= 1e-3 * scale, (g.grad, ref)


def test_bloch_recording_equals_checkpoint_recomputation(gpu_lib):
    """Complex fields: the recording forward run and checkpointed segments recomputed one by one
    give the same gradient bit for bit (checkpoints restore both pole states)."""
    sc, t, mon, K, design = _bloch_case()
    rec = adjoint.overlap_power_gradient(sc, "T", K, design[:, 0], design[:, 1])
    seg = adjoint.overlap_power_gradient(sc, "T", K, design[:, 0], design[:, 1],
                                         checkpoint_every=48, record_bytes=0)
    assert rec.segments == 1 and seg.segments == 6
    assert np.abs(rec.grad).max() > 0
    assert rec.J == seg.J
    assert np.array_equal(rec.grad, seg.grad)
