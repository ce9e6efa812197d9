"""GPU parity of the time-domain field monitors against the float64 oracle."""

from dataclasses import replace

import numpy as np
import pytest

from openem_amd import scenes, solver, tables
from openem_amd.model import FieldTimeMonitor
from openem_amd.solver import Stepper
from oracle.yee64 import Oracle64

pytestmark = pytest.mark.gpu

TMONS = [FieldTimeMonitor("all", tuple(range(6)), (1, 2, 6), (6, 7, 24), 5, 170, 3, 60),
         FieldTimeMonitor("ey_hz", (1, 5), (0, 0, 20), (10, 12, 1), 0, 400, 7, 40)]


def _scene(n):
    return replace(scenes.mini_scene(num_steps=n), field_time_monitors=list(TMONS))


@pytest.mark.parametrize("cplx", [False, True])
def test_time_samples_match_the_oracle(gpu_lib, cplx):
    n = 180
    tabs = tables.build(_scene(n), force_complex=cplx)
    print(f"time monitors, complex={cplx}", flush=True)
    st = Stepper(tabs)
    try:
        st.run(n)
        got = [st.read_time_monitor(i) for i in range(len(TMONS))]
    finally:
        st.close()
    o = Oracle64(tabs).run(n)
    for m, g, want in zip(TMONS, got, o.tsamp):
        f = m.filled(n)
        assert f > 0
        for q, c in enumerate(m.comps):
            w = want[q, :f]
            scale = np.abs(w).max()
            assert scale > 0, (m.name, c)
            assert np.abs(g[q, :f] - w).max() < 1e-4 * scale, (m.name, c)


def test_graph_and_eager_time_samples_are_bitwise_identical(gpu_lib, monkeypatch):
    n = 100
    tabs = tables.build(_scene(n))
    outs = []
    for env in (None, "1"):
        if env:
            monkeypatch.setenv("OEM_NO_GRAPH", env)
        st = Stepper(tabs, dft_batch=8)
        try:
            st.run(n)
            outs.append([st.read_time_monitor(i) for i in range(len(TMONS))])
        finally:
            st.close()
    for a, b in zip(*outs):
        assert np.array_equal(a, b)


def test_run_reports_the_filled_slots(gpu_lib):
    n = 150
    res = solver.run(_scene(n), use_shutoff=False, verbose=False)
    for m in TMONS:
        assert res.time_slots[m.name] == m.filled(n)
        assert res.time_samples[m.name].shape == (len(m.comps), m.filled(n)) + tuple(m.box)
