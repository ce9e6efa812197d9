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
