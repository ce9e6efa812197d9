"""CPU tests of the launch autotuner's candidate rules and cache (the timing itself needs the
GPU and is tested there)."""

import json

from openem_amd import autotune, scenes, tables


def test_both_forms_are_candidates_only_without_h_point_lists():
    plain = tables.build(scenes.box_soup_scene(shape=(24, 22, 40), n_pml=6, n_boxes=4,
                                               num_steps=40, seed=3))
    assert not autotune.has_h_points(plain)
    c = autotune.candidates(plain, "sweeps")
    assert {f for f, _ in c} == {"fused", "sweeps"}
    assert all(ch <= plain.shape[0] for _, ch in c)
    assert len(c) == len(set(c))
    guide = tables.build(scenes.waveguide_scene(shape=(30, 24, 22), n_pml=6, width_cells=6,
                                                thick_cells=4, num_steps=40, nf=1))
    assert autotune.has_h_points(guide)              # a mode source corrects H
    for default in ("sweeps", "fused"):
        assert {f for f, _ in autotune.candidates(guide, default)} == {default}
    mag = tables.build(scenes.mini_scene(magnetic_dipole=True, num_steps=40))
    assert autotune.has_h_points(mag)


def test_shape_key_separates_what_the_plan_depends_on():
    a = tables.build(scenes.mini_scene(shape=(10, 12, 40), num_steps=20))
    b = tables.build(scenes.mini_scene(shape=(10, 12, 40), num_steps=30, seed=4))
    c = tables.build(scenes.mini_scene(shape=(10, 12, 48), num_steps=20))
    d = tables.build(scenes.mini_scene(shape=(10, 12, 40), num_steps=20, n_pml=5))
    assert autotune.shape_key(a) == autotune.shape_key(b)
    assert len({autotune.shape_key(x) for x in (a, c, d)}) == 3


def test_cache_file_round_trip(tmp_path, monkeypatch):
    f = tmp_path / "plans.json"
    monkeypatch.setenv("OEM_TUNE_CACHE", str(f))
    plan = {"form": "fused", "chunk": 12, "ms_per_step": 1.0, "timings": {}, "key": "k"}
    autotune._store_file("k", plan)
    assert json.loads(f.read_text())["k"] == plan
    assert autotune._load_file()["k"] == plan
    f.write_text("not json")
    assert autotune._load_file() == {}
