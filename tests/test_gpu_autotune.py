"""GPU test of the launch autotuner: the tuned plan steps a scene to the same bits as the
library's default plan (a scene without H-side point lists, where every form and chunk gives the
same bits), and a second request for the same shape comes from the cache without timing."""

import numpy as np
import pytest

from openem_amd import autotune, scenes, tables
from openem_amd.solver import Stepper

pytestmark = pytest.mark.gpu


def _run(st, n):
    try:
        st.run(n)
        return (np.stack([st.read_field(c) for c in range(6)]),
                [st.read_monitor(i) for i in range(len(st.tabs.monitors))], st.launch)
    finally:
        st.close()


def test_tuned_plan_gives_the_default_plans_bits(gpu_lib, monkeypatch):
    monkeypatch.delenv("OEM_TUNE_CACHE", raising=False)
    n = 64
    sc = scenes.box_soup_scene(shape=(40, 36, 130), n_pml=6, n_boxes=8, num_steps=n, seed=3,
                               moment=1e-18)
    tabs = tables.build(sc)
    plan = autotune.tune(tabs, steps=8, warmup=2, force=True)
    print(f"plan {plan}", flush=True)
    assert set(plan["timings"]) == {f"{f}/{c}" for f, c in autotune.candidates(tabs, "sweeps")}
    assert plan["timings"][f"{plan['form']}/{plan['chunk']}"] == min(plan["timings"].values())
    assert autotune.tune(tabs) is plan                     # cached: no second timing
    f1, p1, l1 = _run(autotune.stepper(tabs), n)
    f0, p0, l0 = _run(Stepper(tabs), n)
    assert l1 == {"form": plan["form"], "chunk": plan["chunk"]}
    assert np.array_equal(f1, f0)
    for x, y in zip(p1, p0):
        assert np.abs(y).max() > 0 and np.array_equal(x, y)
