"""GPU tests of the serving pool: batched stepping gives each scene the result it gets alone.

Tolerances: a scene stepped by the pool runs the same per-cell arithmetic, the same ADE
pre/fused/post sequence and the same DFT flush partition as oem_run, so its phasors and fields
equal the standalone run's to 1e-6 relative (observed: bit-identical); a scene's result in a mix
equals its result alone in the pool bit for bit.  The standalone path itself is pinned to the
float64 oracle by test_gpu_parity; one pool case is also checked against the oracle directly.
"""

import numpy as np
import pytest

from openem_amd import scenes, serving, tables
from openem_amd.solver import Stepper
from oracle.yee64 import Oracle64

pytestmark = pytest.mark.gpu

ARENA = 2 << 30

# (scene kwargs, steps): step counts that are not multiples of the 16-step block
MIX = [
    (dict(shape=(10, 12, 40)), 203),
    (dict(shape=(12, 10, 36), pml_xy=True, sigma_on=True, pec_on=True, magnetic_dipole=True), 150),
    (dict(shape=(8, 14, 44), pol_axis=1, direction=-1, nonuniform_z=True, lorentz_on=False), 97),
    (dict(shape=(16, 16, 48), periodic_xy=False, plane_wave_on=False, seed=4), 260),
]


def _tabs(kw, n):
    return tables.build(scenes.mini_scene(num_steps=n, **kw), n)


def _standalone(tabs):
    st = Stepper(tabs)
    try:
        st.run(tabs.n_total)
        mons = [st.read_monitor(i) for i in range(len(tabs.monitors))]
        f = np.stack([st.read_field(c) for c in range(6)])
    finally:
        st.close()
    return mons, f


def _read(st):
    mons = [st.read_monitor(i) for i in range(len(st.tabs.monitors))]
    f = np.stack([st.read_field(c) for c in range(6)])
    return mons, f


def _rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def test_pool_matches_standalone_and_mix_matches_alone(gpu_lib):
    tabs = [_tabs(kw, n) for kw, n in MIX]
    ref = [_standalone(t) for t in tabs]
    pool = serving.Pool(ARENA, 16)
    try:
        # alone in the pool
        alone = []
        for i, t in enumerate(tabs):
            print(f"pool alone: case {i} shape {t.shape} steps {t.n_total}", flush=True)
            st = Stepper(t, pool=pool)
            pool.run_to_completion([st])
            alone.append(_read(st))
            st.close()
        # staggered admission: a new scene joins every two blocks while the others run
        print("pool mix: staggered admission", flush=True)
        sts = [Stepper(t, pool=pool) for t in tabs]
        for st in sts:
            pool.attach(st)
            pool.launch(2)
        while pool.stats()["running"]:
            pool.launch(1)
        mix = []
        for st in sts:
            assert pool.done(st, wait=True)
            mix.append(_read(st))
            st.close()
        pool.wait(0)
        assert pool.stats()["used"] == 0, "retired scenes left arena memory behind"
    finally:
        pool.close()
    for i in range(len(tabs)):
        (m_ref, f_ref), (m_al, f_al), (m_mx, f_mx) = ref[i], alone[i], mix[i]
        for a, b in zip(m_al, m_mx):
            assert np.array_equal(a, b), f"case {i}: mix differs from alone"
        assert np.array_equal(f_al, f_mx), f"case {i}: fields differ in the mix"
        for a, b in zip(m_al, m_ref):
            assert _rel(a, b) <= 1e-6, (i, _rel(a, b))
        assert _rel(f_al, f_ref) <= 1e-6, (i, _rel(f_al, f_ref))


def test_pool_case_matches_oracle(gpu_lib):
    kw, n = MIX[0]
    t = _tabs(kw, n)
    pool = serving.Pool(ARENA, 16)
    try:
        st = Stepper(t, pool=pool)
        pool.run_to_completion([st])
        mons, f = _read(st)
        st.close()
    finally:
        pool.close()
    o = Oracle64(t).run(n)
    ref = np.concatenate([o.E, o.H])
    for c in range(6):
        assert _rel(f[c].astype(np.float64), ref[c]) < 1e-4, c
    for i in range(len(t.monitors)):
        assert _rel(mons[i], o.acc[i]) < 1e-4, i


def test_pool_rejects_what_it_cannot_step(gpu_lib):
    from openem_amd import _lib
    pool = serving.Pool(64 << 20, 16)
    try:
        t = _tabs(MIX[0][0], 64)
        with pytest.raises(_lib.OpenEMError):                 # dft_batch differs from the pool's
            Stepper(t, pool=pool, dft_batch=8)
        big = tables.build(scenes.box_soup_scene(n=160, num_steps=32), 32)
        with pytest.raises(_lib.OpenEMError) as e:             # does not fit a 64 MiB arena
            Stepper(big, pool=pool)
        assert e.value.code == _lib.OEM_ENOMEM
        assert pool.stats()["used"] == 0
        st = Stepper(t, pool=pool)
        with pytest.raises(_lib.OpenEMError):                 # stepping a pool scene directly
            pool.attach(st)
            st.run(4)
        pool.run_to_completion([])
        while pool.stats()["running"]:
            pool.launch(1)
        pool.wait(0)
        st.close()
    finally:
        pool.close()


def test_serve_stream_alone_equals_mix(gpu_lib):
    """A short seeded stream of small requests: every result in the continuous batch equals
    the same request served alone (bit for bit) and the serial loop's (1e-6 relative)."""
    specs = scenes.serving_requests(6, rate=200.0, seed=7, side_lo=64, side_hi=80,
                                    steps_lo=200, steps_hi=400)
    print("serve: mix of", [(s["kind"], s["side"], s["n_steps"]) for s in specs], flush=True)
    mix = serving.serve(specs, arena_bytes=ARENA, workers=3)
    assert all(r.error is None for r in mix["requests"]), [r.error for r in mix["requests"]]
    ser = serving.serve_serial(specs)
    for r_mix, r_ser, sp in zip(mix["requests"], ser["requests"], specs):
        alone = serving.serve([sp], arena_bytes=ARENA, workers=1)["requests"][0]
        assert len(r_mix.flux) == sp["n_mon"]
        for a, b, c in zip(r_mix.flux, alone.flux, r_ser.flux):
            assert np.array_equal(a, b), sp
            assert _rel(a, c) <= 1e-6, (sp, _rel(a, c))
    s = serving.summary(mix)
    assert s["failed"] == 0 and s["sims_per_s"] > 0
