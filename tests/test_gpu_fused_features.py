"""GPU tests of the features that read or restore the field sets, on the fused H+E step.

The fused step ping-pongs between two field sets (after s steps the state is in set s % 2), so
everything that reads the state (time monitors, shutoff reductions, returned fields) or restores
it (adjoint checkpoints) must follow the parity.  Each test forces the fused step on a small
scene (OEM_FUSE=1; by default it starts at 10 M cells) and compares with the two sweeps
(OEM_FUSE=0): bit for bit where no point list corrects H, to float32 rounding where one does
(the waveguide's mode source).
"""

from dataclasses import replace

import numpy as np
import pytest

from openem_amd import adjoint, scenes, shutoff, solver, tables
from openem_amd.model import FieldMonitor, FieldTimeMonitor, FluxMonitor, FluxTimeMonitor
from openem_amd.solver import Stepper

pytestmark = pytest.mark.gpu


def _both(monkeypatch, fn):
    out = []
    for env in ("1", "0"):
        monkeypatch.setenv("OEM_FUSE", env)
        out.append(fn())
    monkeypatch.delenv("OEM_FUSE")
    return out


def test_time_and_flux_time_monitors_follow_the_parity(gpu_lib, monkeypatch):
    sc = scenes.mini_scene(num_steps=121)
    g = sc.grid
    span = tuple((float(g.axes[t].edges[0]), float(g.axes[t].edges[-1])) for t in (0, 1))
    sc = replace(sc, field_time_monitors=[
        FieldTimeMonitor("all", tuple(range(6)), (1, 2, 6), (6, 7, 24), 5, 110, 3, 40)],
        flux_time_monitors=[FluxTimeMonitor("f", 2, 30, 1, 0, 121, 2, 70, t_bounds=span)])
    tabs = tables.build(sc)

    def run():
        st = Stepper(tabs, dft_batch=8)
        try:
            st.run(121)                 # odd: the fused state ends in the second set
            return (st.read_time_monitor(0), st.read_flux_time_monitor(0),
                    np.stack([st.read_field(c) for c in range(6)]))
        finally:
            st.close()
    (t1, f1, e1), (t2, f2, e2) = _both(monkeypatch, run)
    assert np.abs(t2).max() > 0 and np.abs(f2).max() > 0
    assert np.array_equal(t1, t2) and np.array_equal(f1, f2) and np.array_equal(e1, e2)


def test_shutoff_stops_at_the_same_step(gpu_lib, monkeypatch):
    sc = scenes.mini_scene(shape=(14, 14, 40), pml_xy=True, plane_wave_on=False,
                           lorentz_on=False, num_steps=3000)

    def run():
        return solver.run(sc, num_steps=3000, policy=shutoff.default_policy(1e-3, 20),
                          verbose=False, return_fields=True)
    a, b = _both(monkeypatch, run)
    assert a.shutoff_triggered and b.shutoff_triggered
    assert a.steps_run == b.steps_run and a.stop_reason == b.stop_reason
    assert a.energy_trace == b.energy_trace
    for k in b.fields:
        assert np.array_equal(a.fields[k], b.fields[k]), k


def test_adjoint_checkpoints_restore_either_set(gpu_lib, monkeypatch):
    """Segments of 47 steps restore checkpoints at odd and even steps; the gradient equals the
    one-segment gradient of the fused step and agrees with the two sweeps' to float32 rounding
    (the mode source corrects H: 1e-4 of the largest |gradient|)."""
    sc = scenes.waveguide_scene(shape=(30, 24, 22), n_pml=6, width_cells=6, thick_cells=4,
                                fwidth_rel=0.3, num_steps=200, nf=1)
    rng = np.random.default_rng(9)
    ny, nz = sc.shape[1], sc.shape[2]
    comp = rng.integers(0, 3, 6)
    cell = np.array([(int(rng.integers(8, 20)) * ny + int(rng.integers(8, 16))) * nz
                     + int(rng.integers(8, 14)) for _ in range(6)])

    def run(every):
        return adjoint.mode_power_gradient(sc, "fwd", comp, cell, checkpoint_every=every,
                                           record_bytes=0)
    monkeypatch.setenv("OEM_FUSE", "1")
    seg = run(47)
    one = run(200)
    monkeypatch.setenv("OEM_FUSE", "0")
    two = run(200)
    monkeypatch.delenv("OEM_FUSE")
    assert seg.segments == 5 and one.segments == 1
    assert np.allclose(seg.grad, one.grad, rtol=1e-12, atol=0) and seg.J == one.J
    scale = np.abs(two.grad).max()
    assert scale > 0
    assert np.abs(one.grad - two.grad).max() <= 1e-4 * scale
    assert abs(one.J - two.J) <= 1e-5 * two.J


# Monitor sets on the 24 x 22 x 150 box soup (dipole at the centre cell (12, 11, 75)).  The
# fused step snaps monitors as long as no work-group (x chunks x 7-row tiles x 64-wide k tiles)
# meets more than three of them (the first through the LDS stage or its own stores, the second
# and third always by direct stores); the others keep the snapshot kernel; each set names what
# it covers.  Flux planes at
# the first plane of an x chunk (3-plane chunks: 6, 15; 16-plane chunks: 16), of a row tile (7,
# 14) and of a k tile (64, 128) take the H below from the prologue, the halo row and the halo
# column; boxes and planes through the dipole are snapped again after its correction; a
# work-group with more than 64 points of a monitor stores them itself, fewer go through LDS.
SNAP_SETS = {
    # x planes at chunk starts: every work-group of the chunk has 448 points (direct stores)
    "x6_x15_chunk3": ("3", lambda f: [FluxMonitor("x6", 0, 6, +1, f[:2]),
                                      FluxMonitor("x15", 0, 15, -1, f[1:3])], lambda f: []),
    # y planes at row-tile starts: one row of 64 points per work-group (staged)
    "y7_y14": (None, lambda f: [FluxMonitor("y7", 1, 7, -1, f[1:3]),
                                FluxMonitor("y14", 1, 14, +1, f[:2])], lambda f: []),
    # z planes at k-tile starts (7 points per work-group, staged) and z75 sharing z64's tiles
    # (the second monitor of those work-groups: direct stores)
    "z64_z128_z75": (None, lambda f: [FluxMonitor("z64", 2, 64, +1, f[:3]),
                                      FluxMonitor("z128", 2, 128, +1, f[:2]),
                                      FluxMonitor("z75", 2, 75, -1, f[:2])], lambda f: []),
    # four z planes in one k tile: three in every work-group of the tile, the fourth keeps the
    # snapshot kernel
    "z64_z70_z75_z80": (None, lambda f: [FluxMonitor("z64", 2, 64, +1, f[:3]),
                                         FluxMonitor("z70", 2, 70, -1, f[1:3]),
                                         FluxMonitor("z75", 2, 75, +1, f[:2]),
                                         FluxMonitor("z80", 2, 80, -1, f[:2])], lambda f: []),
    # x, z and y planes crossing (one work-group holds all three)
    "x6_z75_y7": (None, lambda f: [FluxMonitor("x6", 0, 6, +1, f[:2]),
                                   FluxMonitor("z75", 2, 75, -1, f[1:3]),
                                   FluxMonitor("y7", 1, 7, +1, f[:2])], lambda f: []),
    # a z plane through the dipole and a box around it (both snapped again after the dipole's
    # correction; the box is the second monitor of their shared work-groups)
    "z75_box_dipole": (None, lambda f: [FluxMonitor("z75", 2, 75, -1, f[:2])],
                       lambda f: [FieldMonitor("dip", f[:2], (10, 9, 70), (5, 6, 10))]),
    # a z plane through the dipole (corrected cells snapped again)
    "z75_dipole": (None, lambda f: [FluxMonitor("z75", 2, 75, -1, f[:2])], lambda f: []),
    # a box around the dipole (staged, corrected cells snapped again) and an x plane at the
    # second 16-plane chunk
    "box_dipole_x16": (None, lambda f: [FluxMonitor("x16", 0, 16, +1, f[:2])],
                       lambda f: [FieldMonitor("dip", f[:2], (10, 9, 70), (5, 6, 10))]),
    # a box with up to 196 points per work-group (direct stores) and a z plane in other k tiles
    "wide_box_z40": (None, lambda f: [FluxMonitor("z40", 2, 40, +1, f[:2])],
                     lambda f: [FieldMonitor("wide", f[1:3], (0, 0, 100), (24, 22, 40))]),
}


@pytest.mark.parametrize("name", sorted(SNAP_SETS))
def test_fused_snapshots_equal_the_sweeps(gpu_lib, monkeypatch, name):
    """The fused step writes the DFT snapshots as it stores the new fields; the phasors equal
    those of the two sweeps and their snapshot kernel bit for bit."""
    chunk, flux, boxes = SNAP_SETS[name]
    n = 120
    sc = scenes.box_soup_scene(shape=(24, 22, 150), n_pml=6, n_boxes=8, num_steps=n, seed=3,
                               moment=1e-18)
    f = sc.field_monitors[0].freqs
    sc = replace(sc, flux_monitors=flux(f), field_monitors=boxes(f))
    tabs = tables.build(sc)

    def run():
        st = Stepper(tabs, dft_batch=8)
        try:
            st.run(n)
            return ([st.read_monitor(i) for i in range(len(tabs.monitors))],
                    np.stack([st.read_field(c) for c in range(6)]))
        finally:
            st.close()
    if chunk:
        monkeypatch.setenv("OEM_CHUNK", chunk)
    (a1, e1), (a2, e2) = _both(monkeypatch, run)
    assert np.array_equal(e1, e2)
    for x, y in zip(a1, a2):
        assert np.abs(y).max() > 0
        assert np.array_equal(x, y)


@pytest.mark.parametrize("fuse", ["1", "0"])
def test_step_tail_equals_separate_launches(gpu_lib, monkeypatch, fuse):
    """A step that ends with one launch (E-side dipole, the monitor points it corrects snapped
    again, the step counter) gives the fields and phasors of the same step ending with separate
    launches (OEM_TAIL=0), bit for bit; on the two sweeps with no monitor the tail applies too."""
    n = 120
    sc = scenes.box_soup_scene(shape=(24, 22, 150), n_pml=6, n_boxes=8, num_steps=n, seed=3,
                               moment=1e-18)
    f = sc.field_monitors[0].freqs
    sc = replace(sc, flux_monitors=[],
                 field_monitors=[FieldMonitor("dip", f[:2], (10, 9, 70), (5, 6, 10))]
                 if fuse == "1" else [])
    tabs = tables.build(sc)
    monkeypatch.setenv("OEM_FUSE", fuse)
    out = []
    for tail in ("1", "0"):
        monkeypatch.setenv("OEM_TAIL", tail)
        st = Stepper(tabs, dft_batch=8)
        try:
            st.run(n - 1)
            prof = st.profile_run(1)
            assert prof["step_tail"][1] == (1 if tail == "1" else 0)
            out.append(([st.read_monitor(i) for i in range(len(tabs.monitors))],
                        np.stack([st.read_field(c) for c in range(6)])))
        finally:
            st.close()
    (a1, e1), (a2, e2) = out
    assert np.abs(e2).max() > 0
    assert np.array_equal(e1, e2)
    for x, y in zip(a1, a2):
        assert np.abs(y).max() > 0
        assert np.array_equal(x, y)
