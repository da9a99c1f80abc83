"""The autotuner on the device: a sweep plan built with tune=True takes the
fastest (tile, waves) of its trial launches, records the table, serves a
second plan of the same signature from the cache, and gives bit-identical
results to the plan with the built-in shape."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_tuned_plan_equals_default_plan(tmp_path, monkeypatch):
    import bench
    from tvb_hip import _lib as L
    from tvb_hip import autotune
    monkeypatch.setattr(autotune, "CACHE_PATH", str(tmp_path / "tune.json"))
    w, Lm, d, H, G, sig = bench.c2_setup(256)
    outs, shapes = [], []
    for tune in (True, False, True):
        plan = bench.build_c2_plan(w, d, G, sig, 0.05, 600, L.MIXED, L.NOISE_PHILOX, 21, bold=20.0, skip=True,
                                   tune=tune)
        try:
            plan.init_uniform(5, [0.0, -2.0], [2.0, 1.5])
            plan.run(600)
            outs.append([plan.read_all(m)[0] for m in range(2)])
            shapes.append((plan.kernel, plan.tile, plan.waves, plan.tuned))
        finally:
            plan.close()
    print("tuned", shapes[0], "default", shapes[1][:3], flush=True)
    assert shapes[0][3] is not None and len(shapes[0][3]["table_ms_per_step"]) >= 2
    assert (shapes[0][1], shapes[0][2]) == (shapes[0][3]["tile"], shapes[0][3]["waves"])
    assert shapes[2][3] == shapes[0][3]               # second tuned plan: from the cache
    for a, b in zip(outs[0], outs[1]):
        np.testing.assert_array_equal(a, b)
