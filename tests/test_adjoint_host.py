"""Host side of the adjoint: shell-centre to atom mapping and argument
errors, all decided before any device call."""
import tempfile

import numpy as np
import pytest

from thistle_amd import driver, synth
from thistle_amd import lib as L
from thistle_amd import scattering as sc


@pytest.fixture(scope="module")
def small():
    w = synth.make_small(seed=41, natom=3, spec={"s": 1, "p": 1}, nact=2, nelec=2, ncore=0,
                         nq=8, qmax=4.0)
    with tempfile.TemporaryDirectory() as d:
        mld, pun = synth.write_inputs(w, d)
        inp = driver.prepare_inputs(mld, pun, w.ncore)
    mat, tot = sc.createtwordm(inp.confs, inp.civs)
    return inp, w.q, mat, tot


def test_centres_follow_atoms(small):
    from thistle_amd.autograd import Basis
    inp, q, mat, tot = small
    b = Basis.from_inputs(inp, q, mat)
    xx, yy, zz = b.centres(inp.geom)
    assert np.array_equal(xx, inp.xx) and np.array_equal(yy, inp.yy) and np.array_equal(zz, inp.zz)
    moved = inp.geom.copy()
    moved[1] += [0.1, -0.2, 0.3]
    xx2, yy2, zz2 = b.centres(moved)
    on1 = b.atom_of_gto == 1
    assert np.allclose(xx2[on1] - inp.xx[on1], 0.1) and np.array_equal(xx2[~on1], inp.xx[~on1])
    with pytest.raises(ValueError):
        Basis.from_inputs(inp, q, mat, geom=inp.geom + 0.5)


def test_adjoint_argument_errors(small):
    inp, q, mat, tot = small
    args = (inp.ga, inp.l, inp.m, inp.n, inp.xx, inp.yy, inp.zz, inp.mmod, q, inp.group, mat, tot)
    with pytest.raises(ValueError):
        sc.total_adjoint(*args, np.zeros(q.size + 1), geom=inp.geom)
    with pytest.raises(ValueError):
        sc.total_adjoint(*args, np.zeros(q.size))          # geometry gradient needs geom
    # a shell centre that is not an atom position: rejected before the device
    with pytest.raises(L.ThistleError) as e:
        sc.total_adjoint(*args, np.zeros(q.size), geom=inp.geom + 1e-3)
    assert e.value.code == 1
