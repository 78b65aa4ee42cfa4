"""Batched ensembles (include/mpm/mpm_ensemble.h): scenes advanced together
eagerly or by relaunching one captured HIP graph give exactly the results of
the same scenes stepped one engine at a time; the autotuner returns a
configuration per size bucket; a failing scene is reported by index."""
import numpy as np
import pytest

from mpm_mi355x import capi, scenes

pytestmark = pytest.mark.gpu

KS = (1, 2, 4, 6, 7, 8)  # grids 64, 64, 128, 256, 128, 128; linear elastic and Neo-Hookean


def single(sc, prec, relayout, nsteps):
    e = capi.Engine(sc, precision=prec, relayout_every=relayout)
    e.upload(sc.parts)
    assert e.step(1, nsteps) == 0
    out = e.download(sc.n)
    e.close()
    return out


@pytest.mark.parametrize("prec", [capi.PREC_FP32, capi.PREC_FP64_EXACT])
def test_graph_replay_matches_single_engines(prec):
    scs = [scenes.cfg4_scene(k) for k in KS]
    rel = 16 if prec == capi.PREC_FP32 else 0
    print("ensemble prec=%d scenes=%s" % (prec, [s.n for s in scs]), flush=True)
    ens = capi.Ensemble(scs, precision=prec, relayout_every=rel)
    ens.step(1, 16)                  # eager, one stream per scene
    ens.capture(17, 32)              # steps 17..48 of every scene in one graph
    ens.launch(2)                    # steps 17..80
    got = [e.download(s.n) for e, s in zip(ens.engines, scs)]
    ens.close()
    for s, g in zip(scs, got):
        ref = single(s, prec, rel, 80)
        for f in ("coord", "vel", "stress", "def_grad", "det_def_grad", "vol"):
            assert np.array_equal(getattr(g, f), getattr(ref, f)), (s.name, f)


def test_capture_rejects_blocks_that_do_not_return_to_the_same_buffer():
    scs = [scenes.cfg4_scene(k) for k in KS[:2]]
    ens = capi.Ensemble(scs, precision=capi.PREC_FP32, relayout_every=16)
    with pytest.raises(capi.MPMError):
        ens.capture(1, 16)           # one relayout per block: odd
    ens.close()


def test_autotune_then_graph():
    scs = [scenes.cfg4_scene(k) for k in KS]
    bucket = [{64: 0, 128: 1, 256: 2}[s.nelem[0]] for s in scs]
    ens = capi.Ensemble(scs, bucket=bucket, precision=capi.PREC_FP32, relayout_every=16)
    ens.step(1, 2)
    res = ens.autotune(3, trial_steps=4)
    print(res, flush=True)
    assert sorted(r["bucket"] for r in res) == [0, 1, 2]
    for r in res:
        assert r["lanes_per_cell"] in (1, 2, 4) and r["relayout_every"] in (0, 8, 16)
        assert 0 < r["ms_per_step"] <= r["ms_per_step_worst"]
    ens.capture(3 + 9 * 4, 32)
    ms = ens.launch(3)
    assert ms > 0
    ens.close()


def test_failing_scene_is_named():
    scs = [scenes.cfg4_scene(k) for k in KS[:3]]
    bad = scs[1]
    # scene 1 translates at 100 m/s towards x = 1 and leaves the grid within 200 steps (~0.86 m)
    bad.parts.vel[:, 0] = 100.0
    ens = capi.Ensemble(scs, precision=capi.PREC_FP32, relayout_every=0)
    with pytest.raises(capi.MPMError) as ei:
        ens.step(1, 200)
    assert ei.value.code == capi.MPM_ERR_OUT_OF_GRID and "scene 1" in str(ei.value)
    ens.close()


# max |difference| / max |value| of an atlas scene against its own engine
# after 40 steps with the x-merge on (rounding of the node pair sums only)
# (measured: coord 6e-10, vel 4.6e-7, stress 6.9e-5, F, det F and V 1.2e-10)
ATLAS_XMERGE_BOUND = {"coord": 3e-9, "vel": 3e-6, "stress": 5e-4, "def_grad": 1e-9, "det_def_grad": 1e-9,
                      "vol": 1e-9}


@pytest.mark.parametrize("x_merge", [0, 1])
@pytest.mark.parametrize("per_scene_dt", [False, True])
def test_atlas_matches_individual_scenes(per_scene_dt, x_merge):
    # scenes of one grid size packed into one engine give, scene by scene,
    # the bits of separate engines run with the same time step: the atlas's
    # smallest, or (per-material steps) each scene's own. With the x-merge of
    # the cells kernels on, a node's pair sums follow the grouping of the
    # non-empty cells, which the other scenes of the atlas shift: the results
    # then agree to FP32 rounding (bounds below, 40 steps).
    members = [scenes.cfg4_scene(k) for k in (4, 7, 8)]   # 128 x 128 grids
    a, layout = scenes.atlas(members, per_scene_dt=per_scene_dt)
    print("atlas %s n=%d dt=%.3e per-scene %s x-merge %d" % (a.nelem, a.n, a.dt, per_scene_dt, x_merge), flush=True)
    e = capi.Engine(a, precision=capi.PREC_FP32, relayout_every=16)
    e.upload(a.parts)
    e.set_option(capi.OPT_LANES_PER_CELL, 2)
    e.set_option(capi.OPT_X_MERGE, x_merge)
    assert e.step(1, 40) == 0
    got = e.download(a.n)
    e.close()
    for sc, (first, cnt, origin) in zip(members, layout):
        if per_scene_dt:
            sc.mat_dt = [sc.dt]
        else:
            sc.dt = a.dt
        s1 = capi.Engine(sc, precision=capi.PREC_FP32, relayout_every=16)
        s1.upload(sc.parts)
        s1.set_option(capi.OPT_LANES_PER_CELL, 2)
        s1.set_option(capi.OPT_X_MERGE, x_merge)
        assert s1.step(1, 40) == 0
        ref = s1.download(sc.n)
        s1.close()
        sl = slice(first, first + cnt)
        pairs = [("coord", got.coord[sl] - np.asarray(origin), ref.coord)]
        pairs += [(f, getattr(got, f)[sl], getattr(ref, f)) for f in ("vel", "stress", "def_grad", "det_def_grad", "vol")]
        for f, x, y in pairs:
            if not x_merge:
                assert np.array_equal(x, y), (sc.name, f)
        if x_merge:
            diffs = {f: float(np.max(np.abs(x - y))) / max(float(np.max(np.abs(y))), 1e-30) for f, x, y in pairs}
            print(sc.name, " ".join("%s %.2e" % kv for kv in diffs.items()), flush=True)
            for f, d in diffs.items():
                assert d < ATLAS_XMERGE_BOUND[f], (sc.name, f, d)


def test_relayout_turned_on_after_upload():
    """An engine uploaded with relayout off gets its second particle buffer
    when the option turns relayout on; the run then equals one that had
    relayout on from the upload."""
    s = scenes.tiny(n_cells=10)
    e = capi.Engine(s, precision=capi.PREC_FP32, relayout_every=0)
    e.upload(s.parts)
    e.set_option(capi.OPT_RELAYOUT_EVERY, 2)
    assert e.step(1, 9) == 0
    got = e.download(s.n)
    e.close()
    ref = single(s, capi.PREC_FP32, 2, 9)
    for f in ("coord", "vel", "stress", "def_grad"):
        assert np.array_equal(getattr(got, f), getattr(ref, f)), f


def test_autotune_of_an_ensemble_created_without_relayout():
    scs = [scenes.cfg4_scene(k) for k in KS[:3]]
    ens = capi.Ensemble(scs, precision=capi.PREC_FP32, relayout_every=0)
    ens.step(1, 2)
    res = ens.autotune(3, trial_steps=2)
    assert len(res) == 1
    ens.step(3 + 9 * 2, 16)
    ens.close()
