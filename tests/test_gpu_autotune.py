"""The autotuner of the on-chip FP64 reduction (sg_ctx_autotune): every launch
shape of every size bucket timed, the fastest chosen, and every shape giving
the same SCF results bit for bit (the shapes differ only in how the tiles are
spread over the wavefronts; every sum runs in the same order)."""
import numpy as np
import pytest

from semigpu_amd import api, geometry


@pytest.mark.gpu
def test_autotune_times_every_shape_and_picks_the_fastest(engine):
    out = engine.autotune(molecules=256)
    print("autotune", out, flush=True)
    ms = np.array(out["ms"])
    assert ms.shape == (api.TUNE_BUCKETS, api.TUNE_CANDIDATES)
    for b, shapes in enumerate(out["shapes"]):
        assert len(shapes) >= 1
        t = ms[b, :len(shapes)]
        assert (t > 0).all()
        assert (ms[b, len(shapes):] == -1).all()
        assert out["choice"][b] == int(np.argmin(t))
    assert engine.tri_shapes()["choice"] == out["choice"]


@pytest.mark.gpu
def test_every_shape_gives_the_same_results(engine):
    """Water clusters n = 2 .. 44 (N = 12 .. 264: every bucket, and the head
    above N = 256) through the FP64 SCF with each bucket's shapes forced in
    turn: results equal to the default shapes' bit for bit."""
    mols = [geometry.water_cluster(n, geometry.SplitMix64(100 + n)) for n in range(2, 45, 3)]
    for b in range(api.TUNE_BUCKETS):
        engine.set_tri_variant(b, 0)
    ref = engine.batch(mols).scf(threshold=1e-6)
    shapes = engine.tri_shapes()["shapes"]
    for v in range(1, api.TUNE_CANDIDATES):
        for b, sh in enumerate(shapes):
            engine.set_tri_variant(b, v if v < len(sh) else 0)
        res = engine.batch(mols).scf(threshold=1e-6)
        print("shape set", v, [r["steps"] for r in res], flush=True)
        assert res == ref
    for b in range(api.TUNE_BUCKETS):
        engine.set_tri_variant(b, 0)
    with pytest.raises(api.SemiGpuError):
        engine.set_tri_variant(0, 2)
