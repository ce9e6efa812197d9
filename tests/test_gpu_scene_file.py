"""GPU test: the scene file the reference's own ``openem/serialize.py`` wrote
(``tests/golden/scene_v10_small.npz``: non-uniform x spacing, periodic x, PEC y walls, CPML z,
conductivity and PEC cells, a Lorentz medium, a plane-wave TF/SF source, an electric and a
magnetic dipole with complex amplitudes, a mode source, flux, box and plane monitors) is loaded
with :func:`openem_amd.scene_io.load` and stepped on the device through the stepper, with both
step forms, against the float64 oracle on the same tables.

Tolerance: fields and every monitor's phasors within 1e-4 of the peak of the compared array
(float32 stepping over the file's 12 steps, as the other parity tests).
"""

from pathlib import Path

import numpy as np
import pytest

from openem_amd import scene_io, tables
from openem_amd.solver import Stepper
from oracle.yee64 import Oracle64

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden" / "scene_v10_small.npz"


@pytest.mark.parametrize("fuse", ["0", "1"])
def test_reference_written_scene_steps_like_the_oracle(gpu_lib, monkeypatch, fuse):
    monkeypatch.setenv("OEM_FUSE", fuse)
    sc = scene_io.load(GOLDEN)
    t = tables.build(sc)
    n = t.n_total
    print(f"scene file {sc.shape}, {n} steps, fused={fuse}, {len(t.monitors)} monitors",
          flush=True)
    o = Oracle64(t).run(n)
    st = Stepper(t, device=0)
    try:
        st.run(n)
        fields = [st.read_field(c).astype(np.float64) for c in range(6)]
        phasors = [st.read_monitor(i) for i in range(len(t.monitors))]
    finally:
        st.close()
    ref = np.concatenate([o.E, o.H])
    for c in range(6):
        scale = np.abs(ref[c]).max()
        assert scale > 0
        err = np.abs(fields[c] - ref[c]).max() / scale
        assert err < 1e-4, f"component {c}: {err:.3e}"
    for i, (got, want) in enumerate(zip(phasors, o.acc)):
        scale = np.abs(want).max()
        assert scale > 0
        err = np.abs(got - want).max() / scale
        print(f"monitor {t.monitors[i].name}: relative error {err:.2e}", flush=True)
        assert err < 1e-4, f"monitor {t.monitors[i].name}: {err:.3e}"
