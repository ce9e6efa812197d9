"""GPU test: contexts created, stepped (with captured step graphs) and destroyed on several host
threads at once, as the serving lanes and the designs in flight of the gradient batch do.  One
thread's graph capture and another thread's release of device memory never overlap (the library
serialises them), so every run succeeds and gives the bits of the same run alone."""

import threading

import numpy as np
import pytest

from openem_amd import scenes, tables
from openem_amd.solver import Stepper

pytestmark = pytest.mark.gpu


def _run(tabs, n):
    st = Stepper(tabs, dft_batch=8)
    try:
        st.run(n)
        return [st.read_monitor(i) for i in range(len(tabs.monitors))]
    finally:
        st.close()


def test_contexts_on_several_threads_match_the_serial_runs(gpu_lib):
    n = 48                                   # several captured 8-step blocks per run
    tabs = [tables.build(scenes.box_soup_scene(shape=(24 + 4 * q, 22, 40 + 8 * q), n_pml=6,
                                               n_boxes=4, num_steps=n, seed=q, moment=1e-18))
            for q in range(6)]
    ref = [_run(t, n) for t in tabs]
    out = {}
    errors = []

    def lane(ids):
        try:
            for rep in range(3):
                for q in ids:
                    out[(q, rep)] = _run(tabs[q], n)
        except Exception as e:  # noqa: BLE001 - reported below
            errors.append(repr(e))

    threads = [threading.Thread(target=lane, args=(ids,)) for ids in ((0, 1, 2), (3, 4, 5),
                                                                        (1, 3, 5))]
    for th in threads:
        th.start()
    for th in threads:
        th.join()
    assert not errors, errors
    assert len(out) == 18
    for (q, _), got in out.items():
        for x, y in zip(got, ref[q]):
            assert np.abs(y).max() > 0 and np.array_equal(x, y)
