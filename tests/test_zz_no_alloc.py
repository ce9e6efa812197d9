"""GPU test of ``oem_monitor_flux_noalloc``, the plane-flux reduction that allocates nothing
(its partial sums live in a file-scope device array per GPU): called through the C ABI, it
returns the bits of ``oem_monitor_flux`` on two calls in a row, the second on a smaller monitor
(fewer points and frequencies) of another scene, so stale contents of the shared array from the
first call cannot leak into the second; a monitor with more frequencies than the array holds is
refused with OEM_EINVAL."""

import ctypes as C

import numpy as np
import pytest

from openem_amd import _lib, scenes, tables
from openem_amd.solver import Stepper

pytestmark = pytest.mark.gpu


def _plane_indices(tabs):
    return [i for i, m in enumerate(tabs.monitors) if m.kind == "plane"]


def _both(st, i):
    m = st.tabs.monitors[i]
    ref = np.zeros(2 * m.nf, np.float64)
    got = np.full(2 * m.nf, np.nan, np.float64)
    _lib.check(st.lib, st.lib.oem_monitor_flux(st.ctx, st.mon_ids[i], _lib.ptr(ref), ref.size))
    _lib.check(st.lib, st.lib.oem_monitor_flux_noalloc(st.ctx, st.mon_ids[i], _lib.ptr(got),
                                                       got.size))
    return ref, got


def test_noalloc_flux_equals_the_allocating_flux_twice(gpu_lib):
    first = scenes.mini_scene(shape=(24, 40, 48), num_steps=200, n_pml=6, nf=9)
    second = scenes.mini_scene(shape=(6, 8, 32), num_steps=120, n_pml=5, nf=3)
    sizes = []
    for sc in (first, second):
        t = tables.build(sc)
        st = Stepper(t, device=0)
        try:
            st.run(t.n_total)
            st.setup_flux()
            planes = _plane_indices(t)
            assert planes
            i = max(planes, key=lambda q: t.monitors[q].n1 * t.monitors[q].n2 * t.monitors[q].nf)
            m = t.monitors[i]
            sizes.append(m.n1 * m.n2 * m.nf)
            print(f"scene {sc.shape}: monitor {m.name} {m.n1}x{m.n2} points, {m.nf} frequencies",
                  flush=True)
            ref, got = _both(st, i)
            assert np.abs(ref).max() > 0
            assert np.array_equal(ref, got)
            ref2, got2 = _both(st, i)              # again on the same context
            assert np.array_equal(ref2, ref) and np.array_equal(got2, got)
        finally:
            st.close()
    assert sizes[1] < sizes[0]


def test_noalloc_flux_refuses_what_the_array_cannot_hold(gpu_lib):
    sc = scenes.mini_scene(shape=(6, 8, 32), num_steps=40, n_pml=5)
    t = tables.build(sc)
    st = Stepper(t, device=0)
    try:
        st.run(4)
        st.setup_flux()
        i = _plane_indices(t)[0]
        out = np.zeros(2 * t.monitors[i].nf + 2, np.float64)       # wrong count
        rc = st.lib.oem_monitor_flux_noalloc(st.ctx, st.mon_ids[i], _lib.ptr(out), out.size)
        assert rc == _lib.OEM_EINVAL
        rc = st.lib.oem_monitor_flux_noalloc(st.ctx, 999, _lib.ptr(out), out.size)
        assert rc == _lib.OEM_EINVAL
    finally:
        st.close()
    buf = C.create_string_buffer(512)
    st.lib.oem_last_error(buf, 512)
    assert b"monitor" in buf.value
