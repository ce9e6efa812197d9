"""GPU tests of mode sources and mode monitors.

Tolerances: fields and monitor phasors of the fp32 stepper against the float64 oracle on the
same tables, 1e-4 of their peak; mode amplitudes against those of the oracle's phasors, 1e-5 of
the larger of |a+| and |a-| (measured at most 2e-6); the device overlap reduction against the
host overlap of the same float64 phasors, 1e-12 relative.
"""

import numpy as np
import pytest

from openem_amd import modes, scenes, solver, tables
from openem_amd.solver import Stepper
from oracle.yee64 import Oracle64

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(b).max(), 1e-300))


@pytest.mark.parametrize("direction", [+1, -1])
def test_waveguide_mode_source_matches_oracle(gpu_lib, direction):
    n = 600
    sc = scenes.waveguide_scene(shape=(44, 36, 32), n_pml=8, width_cells=10, thick_cells=6,
                                fwidth_rel=0.2, num_steps=n, nf=2, direction=direction)
    t = tables.build(sc)
    print(f"waveguide direction {direction} shape {sc.shape} steps {n}", flush=True)
    st = Stepper(t)
    try:
        st.setup_modes()
        st.run(n)
        f = np.stack([st.read_field(c) for c in range(6)]).astype(np.float64)
        acc = [st.read_monitor(i) for i in range(len(t.monitors))]
        amps = {i: st.mode_amplitudes(i) for i, m in enumerate(t.monitors) if m.modes}
    finally:
        st.close()
    o = Oracle64(t).run(n)
    ref = np.concatenate([o.E, o.H])
    for c in range(6):
        assert _rel(f[c], ref[c]) < 1e-4, c
    for i in range(len(t.monitors)):
        assert _rel(acc[i], o.acc[i]) < 1e-4, t.monitors[i].name
    for i, (ap, am) in amps.items():
        m = t.monitors[i]
        for fi in range(m.nf):
            hp, hm = modes.amplitudes(acc[i][:, fi:fi + 1], m.modes[fi], sc.grid,
                                      modes.orientation(m.axis))
            assert abs(ap[fi] - hp[0]) <= 1e-12 * max(abs(hp[0]), abs(hm[0]))
            assert abs(am[fi] - hm[0]) <= 1e-12 * max(abs(hp[0]), abs(hm[0]))
            # against the amplitudes of the oracle's float64 phasors
            op, om = modes.amplitudes(o.acc[i][:, fi:fi + 1], m.modes[fi], sc.grid,
                                      modes.orientation(m.axis))
            big = max(abs(op[0]), abs(om[0]))
            ep, em = abs(ap[fi] - op[0]) / big, abs(am[fi] - om[0]) / big
            print(f"{m.name} f{fi}: |a+| {abs(op[0]):.4e} |a-| {abs(om[0]):.4e}, device error "
                  f"{ep:.2e} / {em:.2e} of the larger", flush=True)
            assert ep < 1e-5 and em < 1e-5


def test_run_reports_mode_amplitudes(gpu_lib):
    n = 600
    sc = scenes.waveguide_scene(shape=(44, 36, 32), n_pml=8, width_cells=10, thick_cells=6,
                                fwidth_rel=0.2, num_steps=n, nf=1)
    res = solver.run(sc, use_shutoff=False, verbose=False)
    ms = sc.mode_sources[0]
    f0 = res.mode_amps["fwd"]["freqs"][0]
    k = np.arange(n)
    S = abs(np.sum(ms.amp_e[1:n + 1] * np.exp(2j * np.pi * f0 * (k + 1) * sc.dt)))
    ap = abs(res.mode_amps["fwd"]["a_plus"][0]) / S
    am = abs(res.mode_amps["back"]["a_minus"][0]) / S
    print(f"|a+|/|S| = {ap:.5f}, backward {am:.2e}", flush=True)
    assert abs(ap - 0.5) < 5e-3 and am < 1e-3


def test_run_with_default_arguments_on_a_mode_scene(gpu_lib):
    """``solver.run`` with its defaults (shutoff on) on a scene holding mode monitors: the flux
    weights of every plane monitor upload, and the mode amplitudes equal those of the run
    without shutoff (the scene stops no earlier than its last step here, so both runs step the
    same number of steps)."""
    n = 600
    sc = scenes.waveguide_scene(shape=(44, 36, 32), n_pml=8, width_cells=10, thick_cells=6,
                                fwidth_rel=0.2, num_steps=n, nf=2)
    print(f"run() defaults on waveguide {sc.shape}, {n} steps", flush=True)
    a = solver.run(sc, verbose=False)
    b = solver.run(sc, use_shutoff=False, verbose=False)
    assert a.steps_run == b.steps_run == n, (a.steps_run, a.stop_reason)
    for name in ("fwd", "back"):
        for key in ("a_plus", "a_minus"):
            assert np.array_equal(a.mode_amps[name][key], b.mode_amps[name][key]), (name, key)
    for name in ("fwd_flux", "back_flux"):
        assert np.array_equal(a.phasors[name].data, b.phasors[name].data), name
