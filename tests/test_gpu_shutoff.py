"""GPU tests of the float64 reductions (shutoff energies, plane flux) and of early termination.

Tolerances: the device reductions sum the same float32 terms as the oracle's restatement of the
reference's ``_sum2`` in another order, so on the GPU's own fields they agree to 1e-12
relative; against the float64 oracle's fields the energies agree to 1e-4 (fp32 stepping).  The
early-stop decision (step and reason) must be the one the oracle-driven check loop takes.
"""

import numpy as np
import pytest

from openem_amd import cpml, flux, scenes, shutoff, solver, tables
from openem_amd.solver import Stepper
from oracle.yee64 import Oracle64, energy_sums, plane_flux_terms

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(b).max(), 1e-300))


@pytest.mark.parametrize("dense", [False, True])
def test_energy_and_flux_reductions(gpu_lib, dense):
    sc = scenes.mini_scene(shape=(12, 14, 40), pml_xy=True, sigma_on=True, num_steps=180)
    t = tables.build(sc)
    st = Stepper(t, dense_materials=dense)
    try:
        st.setup_energy()
        st.setup_flux()
        st.run(180)
        e_gpu = st.energy()
        E = np.stack([st.read_field(c) for c in range(3)])          # float32, as stepped
        H = np.stack([st.read_field(c) for c in range(3, 6)])
        terms = {i: st.flux_terms(i) for i, m in enumerate(t.monitors) if m.kind == "plane"}
        data = {i: st.read_monitor(i) for i in terms}
    finally:
        st.close()
    print(f"energy dense={dense}: {e_gpu}", flush=True)
    assert _rel(e_gpu, energy_sums(sc, E, H)) < 1e-12
    o = Oracle64(t).run(180)
    assert _rel(e_gpu, energy_sums(sc, o.E, o.H)) < 1e-4
    assert terms, "mini scene has plane monitors"
    for i, tt in terms.items():
        w1, w2 = flux.area_weights(sc.grid, t.monitors[i].axis)
        assert _rel(tt, plane_flux_terms(data[i], w1, w2)) < 1e-12


def _oracle_shutoff(sc, pol, n_total):
    t = tables.build(sc, n_total)
    o = Oracle64(t)
    src_end = shutoff.source_end_step(sc)
    planes = [i for i, m in enumerate(t.monitors) if m.kind == "plane"]
    trace = []
    for s in range(1, n_total + 1):
        o.step()
        if s % pol.interval:
            continue
        e = energy_sums(sc, o.E, o.H)
        u = 0.5 * (cpml.EPSILON_0 * e[1] + cpml.MU_0 * e[2])
        fl = []
        for i in planes:
            w1, w2 = flux.area_weights(sc.grid, t.monitors[i].axis)
            tt = plane_flux_terms(o.acc[i], w1, w2)
            fl.append(0.5 * (tt[:, 0] - tt[:, 1]))
        stop = pol.check(s, e[0], u, np.concatenate(fl) if fl else np.zeros(0)) and s > src_end
        trace.append((s, pol.field_decay.ratio, u))
        if stop:
            return s, pol.reason, trace
    return n_total, "", trace


def test_run_stops_where_the_oracle_stops(gpu_lib):
    sc = scenes.mini_scene(shape=(14, 14, 40), pml_xy=True, plane_wave_on=False,
                           lorentz_on=False, num_steps=3000)
    n = 3000
    step_o, reason_o, trace_o = _oracle_shutoff(sc, shutoff.default_policy(1e-3, 20), n)
    res = solver.run(sc, num_steps=n, policy=shutoff.default_policy(1e-3, 20), verbose=True)
    print(f"oracle stop {step_o} ({reason_o}); gpu stop {res.steps_run} ({res.stop_reason})",
          flush=True)
    assert step_o < n and res.shutoff_triggered
    assert res.steps_run == step_o and res.stop_reason == reason_o
    assert step_o > shutoff.source_end_step(sc)
    assert len(res.energy_trace) == len(trace_o)
    for (s_g, u_g), (s_o, r_o, u_o) in zip(res.energy_trace, trace_o):
        assert s_g == s_o
        assert abs(u_g - u_o) <= 1e-4 * max(u for _, _, u in trace_o), s_g


def test_run_without_shutoff_runs_every_step(gpu_lib):
    sc = scenes.mini_scene(shape=(10, 12, 40), pml_xy=True, plane_wave_on=False,
                           lorentz_on=False, num_steps=400)
    res = solver.run(sc, use_shutoff=False, verbose=False, return_fields=True)
    assert res.steps_run == 400 and not res.shutoff_triggered and res.stop_reason == ""
    # the reference's return_fields layout: {"Ex": (nx, ny, nz), ...}
    assert sorted(res.fields) == sorted(("Ex", "Ey", "Ez", "Hx", "Hy", "Hz"))
    assert all(v.shape == sc.shape for v in res.fields.values())
