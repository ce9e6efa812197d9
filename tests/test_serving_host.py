"""CPU tests of the serving request stream and its scenes (no GPU)."""

import numpy as np

from openem_amd import scenes, serving, tables


def test_request_stream_statistics():
    specs = scenes.serving_requests(4000, rate=20.0, seed=3)
    t = np.array([s["arrival"] for s in specs])
    gaps = np.diff(np.concatenate([[0.0], t]))
    assert np.all(gaps > 0)
    assert abs(gaps.mean() - 1 / 20.0) < 0.005                  # Poisson rate 20 / s
    assert abs(gaps.std() / gaps.mean() - 1.0) < 0.05           # exponential gaps: cv = 1
    side = np.array([s["side"] for s in specs])
    assert side.min() >= 64 and side.max() <= 320
    ls = (np.log(side) - np.log(64)) / (np.log(320) - np.log(64))
    assert abs(np.median(ls) - 0.5) < 0.03                      # log-uniform
    steps = np.array([s["n_steps"] for s in specs])
    assert steps.min() >= 2000 and steps.max() <= 8000
    assert {s["n_mon"] for s in specs} == {1, 2, 3}
    assert min(s["nf"] for s in specs) == 5 and max(s["nf"] for s in specs) == 21
    assert {s["kind"] for s in specs} == set(scenes.SERVING_KINDS)


def test_serving_scenes_build_for_every_kind():
    for kind in scenes.SERVING_KINDS:
        sp = dict(id=0, arrival=0.0, kind=kind, side=64, n_steps=40, n_mon=3, nf=5, seed=11)
        sc = scenes.serving_scene(sp)
        assert sc.shape == (64, 64, 64) and sc.num_time_steps == 40
        assert len(sc.flux_monitors) == 3
        tabs = tables.build(sc)
        assert len(tabs.monitors) == 3
        assert tabs.material_classes() is not None
        nb = serving.device_bytes(tabs)
        assert 64 ** 3 * 24 < nb < 64 ** 3 * 100, (kind, nb)
        if kind == "slab":
            assert tabs.plane_srcs and not tabs.dipoles
        else:
            assert tabs.dipoles and tabs.disp is not None
