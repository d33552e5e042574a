"""The benchmark's workloads run whole on the GPU.

- XY6 (octahedral, 227 GTOs, 64 shells, 20 core + CAS(10,10), 512 q): the
  forward limit S(0) = N(N-1) (pi_sp/pi)^3, an LL-stratified sample of its
  quartets against the oracle's restatement of tot_integral_k_ijkr /
  bessels0rr (the sample the CPU baseline times), and mixed precision
  within 1e-6 of FP64 at every q point.
- X-Y-X (66 GTOs, 3 core + CAS(8,8), 500 q, 2-RDM screened at 1e-12): total
  and elastic j0 against the oracle, and the screened 2-RDM rows bit-exact.
- The 512-frame batch: every frame's forward limit (to 1e-10: the frames'
  Loewdin MOs are orthonormal only to the conditioning of their random
  overlaps), and a seeded sample of frames bitwise equal to single-frame
  runs.
"""
import math
import tempfile

import numpy as np
import pytest

import oracle
from thistle_amd import driver, synth
from thistle_amd import lib as L
from thistle_amd import scattering as sc

pytestmark = pytest.mark.gpu

PI_SP = float(np.float32(np.pi))          # acos(-1.0000): src/Variables.f90:994
PI_RATIO3 = (PI_SP / math.pi) ** 3


def _inputs(w):
    with tempfile.TemporaryDirectory() as d:
        mld, pun = synth.write_inputs(w, d)
        return driver.prepare_inputs(mld, pun, w.ncore)


def _nelec(inp, mat, tot):
    """N from the 2-RDM trace sum_pq Gamma(p,q,q,p) = N(N-1): rows (p,s,q,r)
    hold Gamma(p,q,r,s), so the trace takes the rows with s == p, q == r."""
    sel = (mat[:, 0] == mat[:, 1]) & (mat[:, 2] == mat[:, 3])
    nn1 = float(np.sum(tot[sel]))
    return (1.0 + math.sqrt(1.0 + 4.0 * nn1)) / 2.0


@pytest.fixture(scope="module")
def xy6():
    w = synth.make_xy6()
    inp = _inputs(w)
    mat, tot = sc.createtwordm(inp.confs, inp.civs, cutoff=w.cutoff2rdm)
    print(f"xy6: ngto={inp.ngto} ng={inp.ng} nmat={tot.size} nq={w.q.size}", flush=True)
    plan = sc.Plan(1, inp.ga, inp.l, inp.m, inp.n, inp.xx, inp.yy, inp.zz, inp.mmod, w.q,
                   inp.group, mat=mat, total=tot)
    S = plan.execute()
    yield w, inp, mat, tot, plan, S
    plan.close()


def test_xy6_forward_limit(xy6):
    w, inp, mat, tot, plan, S = xy6
    N = 2 * w.ncore + 10
    assert abs(_nelec(inp, mat, tot) - N) < 1e-9
    ref = N * (N - 1) * PI_RATIO3
    rel = abs(S[0] - ref) / ref
    print(f"xy6: S(q0={w.q[0]:.0e}) = {S[0]:.12f}, N(N-1)(pi_sp/pi)^3 = {ref:.12f}, "
          f"rel {rel:.2e}", flush=True)
    assert rel < 1e-12


def test_xy6_quartet_sample_vs_oracle(xy6):
    """The oracle's LL-stratified sample (every 200th quartet of the list in
    stable LLmax order) on both sides; partial sums compared per q."""
    w, inp, mat, tot, plan, S = xy6
    stride = 200
    dd, P, e12 = oracle.pair_tables(inp.l, inp.m, inp.n, inp.group, inp.ga, inp.xx, inp.yy,
                                    inp.zz, w.q)
    ql = oracle.quartet_list(P[0], P[1], P[2], 0.001)
    Lg = np.array([inp.l[inp.group == g + 1][0] + inp.m[inp.group == g + 1][0]
                   + inp.n[inp.group == g + 1][0] for g in range(inp.ng)])
    LL = np.minimum(Lg[ql[:, :4] - 1].sum(axis=1), 20)
    pos = np.sort(np.argsort(LL, kind="stable")[::stride])
    plan.set_quartets(pos)
    try:
        part = plan.execute()
    finally:
        plan.set_quartets(None)
    ref, _, ns = oracle.total_j0_sample(inp.l, inp.m, inp.n, inp.group, inp.ga, inp.xx,
                                        inp.yy, inp.zz, inp.mmod, mat, tot, w.q, stride)
    assert ns == pos.size
    print(f"xy6 sample: {ns} quartets, LL classes {np.bincount(LL[pos]).tolist()}", flush=True)
    scale = np.max(np.abs(ref))
    norm = np.max(np.abs(part - ref)) / scale
    big = np.abs(ref) > 1e-3 * scale
    rel = np.max(np.abs(part - ref)[big] / np.abs(ref[big]))
    print(f"xy6 sample: normwise {norm:.2e}, per-q (|S| > 1e-3 max) {rel:.2e}", flush=True)
    assert norm < 1e-12 and rel < 1e-10
    # the whole list is back: the full execution is unchanged
    assert np.array_equal(plan.execute(), S)


def test_xy6_mixed_per_q(xy6):
    w, inp, mat, tot, plan, S = xy6
    pm = sc.Plan(1, inp.ga, inp.l, inp.m, inp.n, inp.xx, inp.yy, inp.zz, inp.mmod, w.q,
                 inp.group, mat=mat, total=tot, precision=L.MIXED)
    Sm = pm.execute()
    idx, A2 = pm.mixed_fallback()
    pm.close()
    rel = np.abs(Sm - S) / np.abs(S)
    k = int(np.argmax(rel))
    print(f"xy6 mixed: max per-q rel {rel[k]:.2e} at q = {w.q[k]:.3f}; "
          f"{idx.size} of {w.q.size} points re-evaluated in FP64", flush=True)
    assert rel.max() <= 1e-6
    # the re-evaluated points are FP64 sums of the same records
    if idx.size:
        assert np.max(np.abs(Sm[idx] - S[idx]) / np.abs(S[idx])) < 1e-12


@pytest.fixture(scope="module")
def xyx():
    w = synth.make_xyx()
    inp = _inputs(w)
    mat, tot = sc.createtwordm(inp.confs, inp.civs, cutoff=w.cutoff2rdm)
    print(f"xyx: ngto={inp.ngto} ng={inp.ng} nconf={inp.confs.shape[0]} nmat={tot.size} "
          f"(screen {w.cutoff2rdm:g}) nq={w.q.size}", flush=True)
    return w, inp, mat, tot


def test_xyx_screened_rows_bit_exact(xyx):
    w, inp, mat, tot = xyx
    om, ot = oracle.createtwordm(inp.confs, inp.civs, cutoff=w.cutoff2rdm)
    assert np.array_equal(mat, om) and np.array_equal(tot, ot)
    assert np.all(np.abs(tot) >= w.cutoff2rdm)


def test_xyx_total_and_elastic_vs_oracle(xyx):
    w, inp, mat, tot = xyx
    args = (inp.ga, inp.l, inp.m, inp.n, inp.xx, inp.yy, inp.zz, inp.mmod)
    res = sc.total_from_twordm(*args, w.q, inp.group, mat, tot)
    ref = oracle.total_j0(inp.l, inp.m, inp.n, inp.group, inp.ga, inp.xx, inp.yy, inp.zz,
                          inp.mmod, mat, tot, w.q)
    rel = np.max(np.abs(res - ref) / np.abs(ref))
    print(f"xyx total: max per-q rel {rel:.2e}", flush=True)
    assert rel < 1e-10
    g = sc.onerdm(inp.confs, inp.civs)
    rese = sc.elastic_from_onerdm(*args, w.q, inp.group, g)
    refe = oracle.elastic_j0(inp.l, inp.m, inp.n, inp.group, inp.ga, inp.xx, inp.yy, inp.zz,
                             oracle.elastic_z(g, inp.mmod), w.q)
    rele = np.max(np.abs(rese - refe) / np.abs(refe))
    print(f"xyx elastic: max per-q rel {rele:.2e}", flush=True)
    assert rele < 1e-10


def test_batch_512_frames():
    frames, nel = [], []
    for w in synth.make_batch(nframes=512):
        inp = _inputs(w)
        mat, tot = sc.createtwordm(inp.confs, inp.civs)
        frames.append((inp.ga, inp.l, inp.m, inp.n, inp.xx, inp.yy, inp.zz, inp.mmod, w.q,
                       inp.group, mat, tot))
        nel.append(_nelec(inp, mat, tot))
    print(f"batch: {len(frames)} frames, {sum(f[8].size for f in frames)} q points", flush=True)
    b = sc.Batch(frames)
    res = b.execute()
    info = b.info()
    print(f"batch: waves={info['waves']} graph={info['graph']} ms={info['ms']:.2f}", flush=True)
    worst = 0.0
    for f, (S, N) in enumerate(zip(res, nel)):
        ref = N * (N - 1) * PI_RATIO3
        worst = max(worst, abs(S[0] - ref) / ref)
    print(f"batch: worst forward-limit rel err {worst:.2e}", flush=True)
    # the frames' random exponents give overlaps of condition up to ~1e5, so
    # their Loewdin MOs are orthonormal to ~1e-11, which bounds the limit
    assert worst < 1e-10
    rng = np.random.default_rng(512)
    for f in sorted(rng.choice(len(frames), size=8, replace=False)):
        single = sc.total_from_twordm(*frames[f], waves=info["waves"])
        assert np.array_equal(single, res[f]), f
    again = b.execute()
    assert all(np.array_equal(a, r) for a, r in zip(again, res))
    b.close()
