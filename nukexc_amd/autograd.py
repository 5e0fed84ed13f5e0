"""torch.autograd binding of the V_xc build: E_xc as a differentiable function
of the density matrix.

The forward pass is one nkx_vxc build on the context's grid and basis; the
backward pass returns dE_xc/dP = V_xc, which the same build produced, so a
gradient costs nothing beyond the forward.  P enters only through its
symmetric part (rho = sum_{mu nu} P_{mu nu} phi_mu phi_nu), so the input is
symmetrised and the returned gradient V_xc is symmetric.
"""
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
