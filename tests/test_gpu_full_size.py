"""GPU tests at the full sizes the benchmark times: the 512^3 box soup (CPML on six faces, 32
boxes of plain eps and 2-pole Lorentz media, the centred Ez dipole, the 8-frequency xy field
monitor) and the 256 x 256 x 600 complex Bloch cell (Bloch x/y, CPML z, oblique TF/SF box,
Drude + Lorentz metal, four 21-frequency flux planes).

The fused H+E march and the two separate sweeps start from the same seeded random fields in
every cell (so every work-group, the CPML slabs, every material class, the ADE entries and the
far end of the 32-bit buffer offsets do work from the first step) and step the same count.

* 512^3: fields and phasors equal bit for bit (the scene has no H-side point list, so the two
  forms round every expression alike), and the field energy stays finite.
* Bloch cell: the TF/SF box adds H-side corrections, which the march adds before the curl and
  the sweeps after it (one rounding apart at the corrected cells), so fields agree within 1e-5
  of each component's peak and phasors within 1e-5 of each monitor's peak.
"""

import numpy as np
import pytest

from openem_amd import scenes, tables
from openem_amd.solver import Stepper

pytestmark = pytest.mark.gpu


def _random_fields(shape, cplx, seed):
    rng = np.random.default_rng(seed)
    out = []
    for c in range(6):
        scale = 1.0 if c < 3 else 1.0 / 376.730313668   # H ~ E / eta0
        a = rng.standard_normal(shape, dtype=np.float32) * np.float32(scale)
        if cplx:
            b = rng.standard_normal(shape, dtype=np.float32) * np.float32(scale)
            a = a + 1j * b
        out.append(a)
    return out


def _run(tabs, init, steps, fuse, monkeypatch):
    monkeypatch.setenv("OEM_FUSE", fuse)
    st = Stepper(tabs, device=0)
    try:
        for c in range(6):
            st.write_field(c, init[c])
        st.run(steps - 2)
        prof = st.profile_run(2)            # the last two steps eagerly, to see which form ran
        assert prof["update_he"][1] == (2 if fuse == "1" else 0)
        fields = [st.read_field(c) for c in range(6)]
        ph = [st.read_monitor(i) for i in range(len(tabs.monitors))]
    finally:
        st.close()
    return fields, ph


def test_512_cubed_fused_equals_two_sweeps_bitwise(gpu_lib, monkeypatch):
    steps = 40
    sc = scenes.box_soup_scene(n=512, num_steps=steps + 8)
    tabs = tables.build(sc, steps + 8)
    init = _random_fields(sc.shape, False, seed=11)
    print(f"512^3: {steps} steps from random fields, fused then two sweeps", flush=True)
    fa, pa = _run(tabs, init, steps, "1", monkeypatch)
    fb, pb = _run(tabs, init, steps, "0", monkeypatch)
    for c in range(6):
        assert np.isfinite(fa[c]).all(), f"component {c} not finite"
        assert np.array_equal(fa[c], fb[c]), f"component {c} differs"
        assert not np.array_equal(fa[c], init[c])
    for i, (x, y) in enumerate(zip(pa, pb)):
        assert np.abs(x).max() > 0
        assert np.array_equal(x, y), f"monitor {i} differs"
    energy = sum(float(np.sum(f.astype(np.float64) ** 2)) for f in fa[:3])
    assert np.isfinite(energy) and energy > 0


def test_bloch_cell_fused_matches_two_sweeps(gpu_lib, monkeypatch):
    steps = 40
    sc = scenes.bloch_cell_scene(num_steps=400)
    tabs = tables.build(sc, steps + 8)
    assert tabs.bloch
    init = _random_fields(sc.shape, True, seed=12)
    print(f"Bloch cell {sc.shape}: {steps} steps from random fields, fused then two sweeps",
          flush=True)
    fa, pa = _run(tabs, init, steps, "1", monkeypatch)
    fb, pb = _run(tabs, init, steps, "0", monkeypatch)
    for c in range(6):
        assert np.isfinite(fa[c]).all()
        scale = np.abs(fb[c]).max()
        err = np.abs(fa[c] - fb[c]).max() / scale
        print(f"component {c}: max difference {err:.2e} of peak", flush=True)
        assert err <= 1e-5
    for i, (x, y) in enumerate(zip(pa, pb)):
        scale = np.abs(y).max()
        assert scale > 0
        err = np.abs(x - y).max() / scale
        print(f"monitor {tabs.monitors[i].name}: max difference {err:.2e} of peak", flush=True)
        assert err <= 1e-5
