"""torch.autograd binding of the V_xc build: E_xc as a differentiable function
of the density matrix, and of the nuclear coordinates.

The forward pass is one nkx_vxc build on the context's grid and basis; the
backward pass returns dE_xc/dP = V_xc, which the same build produced, so a
gradient costs nothing beyond the forward.  P enters only through its
symmetric part (rho = sum_{mu nu} P_{mu nu} phi_mu phi_nu), so the input is
symmetrised and the returned gradient V_xc is symmetric.

xc_energy_of_geometry is E_xc at fixed P as a function of the atom positions:
the forward pass rebuilds the grid (points move with their atoms, Becke
weights recomputed) and moves the basis centres, then runs nkx_xc_grad (or
nkx_xc_grad_mixed), whose gradient the backward pass returns.  These bindings
are a gradcheck harness; the C ABI stays primary.
"""
import dataclasses

import numpy as np
import torch

from .api import PBE


class _XcEnergy(torch.autograd.Function):
    @staticmethod
    def forward(fctx, P, context, funcs, precision):
        Ps = 0.5 * (P + P.transpose(0, 1))
        r = context.vxc(Ps.detach().contiguous(), funcs=funcs, precision=precision)
        fctx.save_for_backward(r.potential)
        return torch.tensor(r.energy, dtype=torch.float64, device=P.device)

    @staticmethod
    def backward(fctx, g):
        (V,) = fctx.saved_tensors
        return g * V, None, None, None


class _XcEnergyPolarized(torch.autograd.Function):
    @staticmethod
    def forward(fctx, Pa, Pb, context, funcs):
        sa = 0.5 * (Pa + Pa.transpose(0, 1))
        sb = 0.5 * (Pb + Pb.transpose(0, 1))
        r = context.vxc_polarized(sa.detach().contiguous(), sb.detach().contiguous(), funcs=funcs)
        fctx.save_for_backward(r.potential_alpha, r.potential_beta)
        return torch.tensor(r.energy, dtype=torch.float64, device=Pa.device)

    @staticmethod
    def backward(fctx, g):
        Va, Vb = fctx.saved_tensors
        return g * Va, g * Vb, None, None


def xc_energy(context, P, funcs=PBE, precision="fp64"):
    """E_xc[P] on the context's grid and basis; differentiable in P (FP64
    tensor on the context's device)."""
    return _XcEnergy.apply(P, context, funcs, precision)


def xc_energy_polarized(context, Pa, Pb, funcs=PBE):
    """Spin-polarised E_xc[P_alpha, P_beta]; differentiable in both."""
    return _XcEnergyPolarized.apply(Pa, Pb, context, funcs)


class _XcEnergyGeometry(torch.autograd.Function):
    @staticmethod
    def forward(fctx, xyz, context, Z, basis, fatom, P, funcs, grid_kw, partition, precision):
        x = np.ascontiguousarray(xyz.detach().cpu().numpy(), dtype=np.float64).reshape(-1, 3)
        context.make_flat_grid(x, Z, **grid_kw)
        context.set_basis(dataclasses.replace(basis, O=np.ascontiguousarray(x[fatom])))
        E, g = context.xc_grad(P, funcs=funcs, partition=partition, precision=precision)
        fctx.save_for_backward(torch.tensor(g, dtype=torch.float64, device=xyz.device))
        return torch.tensor(E, dtype=torch.float64, device=xyz.device)

    @staticmethod
    def backward(fctx, gout):
        (G,) = fctx.saved_tensors
        return gout * G, None, None, None, None, None, None, None, None, None


def function_atoms(basis, xyz):
    """Index of the atom each basis function sits on (its centre must
    coincide with one of xyz)."""
    xyz = np.asarray(xyz, dtype=np.float64).reshape(-1, 3)
    d = np.linalg.norm(np.asarray(basis.O)[:, None, :] - xyz[None, :, :], axis=2)
    a = d.argmin(1)
    if not np.all(d[np.arange(len(a)), a] < 1e-12):
        raise ValueError("every basis function must be centred on an atom")
    return a


def xc_energy_of_geometry(context, xyz, Z, basis, fatom, P, funcs=PBE, grid=None,
                          partition="becke_capped", precision="fp64"):
    """E_xc at fixed P (device tensor) as a differentiable function of the
    atom positions xyz (float64 tensor, natoms x 3).  Function i of basis
    sits on atom fatom[i] (function_atoms at a reference geometry) and moves
    with it.  grid holds the make_flat_grid keyword arguments.  Each call
    rebuilds the context's grid and basis."""
    return _XcEnergyGeometry.apply(xyz, context, np.asarray(Z), basis, np.asarray(fatom), P,
                                   funcs, dict(grid or {}), partition, precision)
