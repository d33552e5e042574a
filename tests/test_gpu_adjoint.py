"""Adjoint of the total j0 path against finite differences of the forward.

dL/dR by central differences with a 1e-4 bohr step, dL/d(2-RDM values) by
the exact linearity of S in Gamma, and the torch.autograd wrapper against
the direct adjoint call.  Tolerance: normwise 1e-6 for dL/dR, 1e-9 for
dL/dGamma."""
import tempfile

import numpy as np
import pytest

from thistle_amd import driver, synth
from thistle_amd import scattering as sc

pytestmark = pytest.mark.gpu


def _case(name):
    if name == "spd3":
        # 4 core + CAS(4,4): 8 orbitals, Z in the MO-factorised layout
        w = synth.make_small(seed=41, natom=3, spec={"s": 2, "p": 1, "d": 1}, nact=4, nelec=4,
                             ncore=4, nq=48, qmax=8.0)
    else:  # 10 atoms, L <= 2, CAS(6,6), 256 q
        w = synth.make_mol10()
    with tempfile.TemporaryDirectory() as d:
        mld, pun = synth.write_inputs(w, d)
        inp = driver.prepare_inputs(mld, pun, w.ncore)
    mat, tot = sc.createtwordm(inp.confs, inp.civs)
    print(f"case {name}: natom={inp.geom.shape[0]} ngto={inp.ngto} ng={inp.ng} "
          f"nmat={tot.size} nq={w.q.size}", flush=True)
    return inp, w.q, mat, tot


@pytest.fixture(scope="module", params=["spd3", "mol10"])
def case(request):
    return (request.param,) + _case(request.param)


def _forward(inp, q, mat, tot, geom):
    from thistle_amd.autograd import Basis
    b = Basis.from_inputs(inp, q, mat)
    xx, yy, zz = b.centres(geom)
    return sc.total_from_twordm(inp.ga, inp.l, inp.m, inp.n, xx, yy, zz, inp.mmod, q,
                                inp.group, mat, tot)


def _loss_weights(name, inp, q, mat, tot):
    """dL/dS of L = sum_q (S - S_t)^2 / (nq |S_t|), S_t from coordinates
    displaced by seeded N(0, 0.05^2) noise."""
    rng = np.random.default_rng(7)
    geom_t = inp.geom + rng.normal(scale=0.05, size=inp.geom.shape)
    St = _forward(inp, q, mat, tot, geom_t)
    S0 = _forward(inp, q, mat, tot, inp.geom)
    return 2.0 * (S0 - St) / (q.size * np.abs(St)), St


def test_geometry_gradient_vs_central_differences(case):
    name, inp, q, mat, tot = case
    v, _ = _loss_weights(name, inp, q, mat, tot)
    S, dR, _ = sc.total_adjoint(inp.ga, inp.l, inp.m, inp.n, inp.xx, inp.yy, inp.zz, inp.mmod,
                                q, inp.group, mat, tot, v, geom=inp.geom, want_total=False)
    assert np.array_equal(S, _forward(inp, q, mat, tot, inp.geom))
    h = 1e-4
    fd = np.zeros_like(dR)
    for a in range(inp.geom.shape[0]):
        for c in range(3):
            gp, gm = inp.geom.copy(), inp.geom.copy()
            gp[a, c] += h
            gm[a, c] -= h
            fd[a, c] = v @ (_forward(inp, q, mat, tot, gp) - _forward(inp, q, mat, tot, gm)) / (2 * h)
    err = np.max(np.abs(dR - fd)) / np.max(np.abs(fd))
    print(f"{name}: dL/dR normwise err vs FD {err:.3e}", flush=True)
    # translation invariance: the coordinate gradient sums to zero per axis
    assert np.max(np.abs(dR.sum(axis=0))) <= 1e-9 * np.max(np.abs(dR))
    assert err <= 1e-6


def test_rdm_gradient_is_exact_response(case):
    name, inp, q, mat, tot = case
    v = np.random.default_rng(3).normal(size=q.size)
    _, _, dT = sc.total_adjoint(inp.ga, inp.l, inp.m, inp.n, inp.xx, inp.yy, inp.zz, inp.mmod, q,
                                inp.group, mat, tot, v, want_geom=False)
    rows = np.random.default_rng(4).choice(tot.size, size=min(8, tot.size), replace=False)
    worst = 0.0
    for n in rows:
        tp, tm = tot.copy(), tot.copy()
        d = 0.25 * max(abs(tot[n]), 1e-3)
        tp[n] += d
        tm[n] -= d
        fd = v @ (_forward(inp, q, mat, tp, inp.geom) - _forward(inp, q, mat, tm, inp.geom)) / (2 * d)
        worst = max(worst, abs(dT[n] - fd) / max(abs(fd), 1e-30))
    print(f"{name}: dL/dGamma worst rel err {worst:.3e}", flush=True)
    assert worst <= 1e-9


def test_autograd_wrapper(case):
    import torch

    from thistle_amd.autograd import Basis, total_j0
    name, inp, q, mat, tot = case
    _, St = _loss_weights(name, inp, q, mat, tot)
    basis = Basis.from_inputs(inp, q, mat)
    geom = torch.tensor(inp.geom, dtype=torch.float64, requires_grad=True)
    total = torch.tensor(tot, dtype=torch.float64, requires_grad=True)
    St_t = torch.from_numpy(St)
    S = total_j0(geom, total, basis)
    L = ((S - St_t) ** 2 / St_t.abs()).sum() / q.size
    L.backward()
    gS = (2.0 * (S.detach() - St_t) / (q.size * St_t.abs())).numpy()
    _, dR, dT = sc.total_adjoint(inp.ga, inp.l, inp.m, inp.n, inp.xx, inp.yy, inp.zz, inp.mmod,
                                 q, inp.group, mat, tot, gS, geom=inp.geom)
    np.testing.assert_allclose(geom.grad.numpy(), dR, rtol=1e-12, atol=1e-14 * np.abs(dR).max())
    np.testing.assert_allclose(total.grad.numpy(), dT, rtol=1e-12, atol=1e-14 * np.abs(dT).max())
