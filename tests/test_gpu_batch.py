"""Batched frames: per-frame equality with single-frame runs, graph replays
against the eager first execution (with other eager work in between), and
two frames against the oracle."""
import tempfile

import numpy as np
import pytest

import oracle
from thistle_amd import driver, synth
from thistle_amd import scattering as sc

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def frames():
    out = []
    for w in synth.make_batch(nframes=10, seed=61):
        with tempfile.TemporaryDirectory() as d:
            mld, pun = synth.write_inputs(w, d)
            inp = driver.prepare_inputs(mld, pun, w.ncore)
        mat, tot = sc.createtwordm(inp.confs, inp.civs)
        out.append((inp, w.q, mat, tot))
    print("frames: " + ", ".join(f"{i.geom.shape[0]}at/{q.size}q" for i, q, _, _ in out),
          flush=True)
    return out


def _args(inp, q, mat, tot):
    return (inp.ga, inp.l, inp.m, inp.n, inp.xx, inp.yy, inp.zz, inp.mmod, q, inp.group, mat,
            tot)


def test_batch_matches_single_frames_and_replays(frames):
    b = sc.Batch([_args(*fr) for fr in frames])
    r1 = b.execute()
    info = b.info()
    print(f"batch: waves={info['waves']} graph={info['graph']} ms={info['ms']:.3f}", flush=True)
    singles = []
    for (inp, q, mat, tot) in frames:
        singles.append(sc.total_from_twordm(inp.ga, inp.l, inp.m, inp.n, inp.xx, inp.yy, inp.zz,
                                            inp.mmod, q, inp.group, mat, tot,
                                            waves=info["waves"]))
    for f in range(len(frames)):
        assert np.array_equal(r1[f], singles[f]), f
    for rep in range(3):
        r = b.execute()
        for f in range(len(frames)):
            assert np.array_equal(r[f], r1[f]), (rep, f)
        # eager work between replays
        inp, q, mat, tot = frames[rep]
        sc.total_from_twordm(inp.ga, inp.l, inp.m, inp.n, inp.xx, inp.yy, inp.zz, inp.mmod, q,
                             inp.group, mat, tot)
    b.close()


def test_batch_frames_vs_oracle(frames):
    order = np.argsort([fr[0].ngto for fr in frames])[:2]
    b = sc.Batch([_args(*frames[i]) for i in order])
    res = b.execute()
    for k, i in enumerate(order):
        inp, q, mat, tot = frames[i]
        ref = oracle.total_j0(inp.l, inp.m, inp.n, inp.group, inp.ga, inp.xx, inp.yy, inp.zz,
                              inp.mmod, mat, tot, q)
        rel = np.max(np.abs(res[k] - ref) / np.abs(ref))
        print(f"frame {i}: max rel err vs oracle {rel:.3e}", flush=True)
        assert rel < 1e-10
    b.close()


def test_batch_per_frame_paths(frames):
    """Batches that keep per-frame quartet launches: a frame with more than
    512 q points (FP64), and mixed precision; each frame bitwise equal to its
    single-frame run with the batch's waves."""
    from thistle_amd import lib as L
    inp, q, mat, tot = frames[0]
    long_q = np.linspace(1e-10, 8.0, 600)
    sets = [([_args(inp, long_q, mat, tot)] + [_args(*fr) for fr in frames[1:3]], L.FP64),
            ([_args(*fr) for fr in frames[:3]], L.MIXED)]
    for fr, prec in sets:
        b = sc.Batch(fr, precision=prec)
        res = b.execute()
        info = b.info()
        print(f"per-frame batch: precision {prec} waves={info['waves']} rows={info['rows']}",
              flush=True)
        assert info["rows"] == 0 and info["groups"] == 0
        for k, a in enumerate(fr):
            single = sc.total_from_twordm(*a, precision=prec, waves=info["waves"])
            assert np.array_equal(single, res[k]), (prec, k)
        b.close()
