"""torch.autograd wrapper of the total j0 path.

``total_j0(geom, total, basis)`` returns S(q) as a CPU float64 tensor that is
differentiable with respect to the nuclear coordinates ``geom`` (natom, 3;
bohr) and the 2-RDM values ``total`` (nmat).  The forward runs
``thistle_total_from_twordm``; the backward runs
``thistle_total_scattering_adjoint``, which recomputes the forward and its
Hermite tables rather than keeping per-quartet state.  Torch only carries
the CPU tensors: the GPU work happens in libthistle_hip.so, so torch's own
GPU runtime is never initialised.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch

from . import scattering as sc


@dataclass
class Basis:
    """Primitive GTOs whose centres sit on atoms, plus MOs, q grid, 2-RDM rows."""
    ga: np.ndarray
    l: np.ndarray
    m: np.ndarray
    n: np.ndarray
    group: np.ndarray
    mmod: np.ndarray
    q: np.ndarray
    mat: np.ndarray
    atom_of_gto: np.ndarray
    cutoffz: float = 1e-15
    cutoffmd: float = 1e-20
    cutoffcentre: float = 0.001

    @classmethod
    def from_inputs(cls, inp, q, mat, geom=None, tol=1e-10):
        """Map every GTO centre of ``inp`` (driver.ScatteringInputs) to the
        atom of ``geom`` (default ``inp.geom``) it sits on."""
        geom = np.asarray(inp.geom if geom is None else geom, dtype=np.float64).reshape(-1, 3)
        cen = np.stack([inp.xx, inp.yy, inp.zz], axis=1)
        d = np.abs(cen[:, None, :] - geom[None, :, :]).max(axis=2)
        atom = d.argmin(axis=1)
        if np.any(d[np.arange(len(atom)), atom] > tol * (1.0 + np.abs(cen).max())):
            raise ValueError("a GTO centre does not coincide with any atom")
        return cls(np.asarray(inp.ga, dtype=np.float64), np.asarray(inp.l), np.asarray(inp.m),
                   np.asarray(inp.n), np.asarray(inp.group), np.asarray(inp.mmod),
                   np.asarray(q, dtype=np.float64), np.asarray(mat, dtype=np.int32),
                   atom.astype(np.int64))

    def centres(self, geom):
        c = np.asarray(geom, dtype=np.float64)[self.atom_of_gto]
        return c[:, 0].copy(), c[:, 1].copy(), c[:, 2].copy()


class TotalJ0(torch.autograd.Function):
    @staticmethod
    def forward(ctx, geom, total, basis):
        g = geom.detach().cpu().numpy().astype(np.float64)
        t = total.detach().cpu().numpy().astype(np.float64)
        xx, yy, zz = basis.centres(g)
        S = sc.total_from_twordm(basis.ga, basis.l, basis.m, basis.n, xx, yy, zz, basis.mmod,
                                 basis.q, basis.group, basis.mat, t, basis.cutoffz,
                                 basis.cutoffmd, basis.cutoffcentre)
        ctx.basis = basis
        ctx.save_for_backward(geom, total)
        return torch.from_numpy(S)

    @staticmethod
    def backward(ctx, gS):
        geom, total = ctx.saved_tensors
        b = ctx.basis
        g = geom.detach().cpu().numpy().astype(np.float64)
        xx, yy, zz = b.centres(g)
        _, dR, dT = sc.total_adjoint(b.ga, b.l, b.m, b.n, xx, yy, zz, b.mmod, b.q, b.group,
                                     b.mat, total.detach().cpu().numpy(),
                                     gS.detach().cpu().numpy(), geom=g, cutoffz=b.cutoffz,
                                     cutoffmd=b.cutoffmd, cutoffcentre=b.cutoffcentre,
                                     want_geom=ctx.needs_input_grad[0],
                                     want_total=ctx.needs_input_grad[1])
        gR = torch.from_numpy(dR).to(geom.dtype) if dR is not None else None
        gT = torch.from_numpy(dT).to(total.dtype) if dT is not None else None
        return gR, gT, None


def total_j0(geom, total, basis):
    """S(q) of the total j0 path as a differentiable function of geom and total."""
    return TotalJ0.apply(geom, total, basis)
