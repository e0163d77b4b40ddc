"""torch.autograd over the C ABI: the RHF total energy as a differentiable function of the nuclear
positions. forward runs fock_scf_run at the given geometry (basis functions move with their
nucleus); backward returns the analytic nuclear gradient of fock_rhf_gradient. The accurate
integral policy is required (the reference's tabulated Boys function is not differentiable).
"""
from __future__ import annotations

import torch

from . import abi, molecule as m


class RHFEnergy(torch.autograd.Function):
    @staticmethod
    def forward(fctx, positions: torch.Tensor, fb, fn, context: abi.Context, opts):
        pos = positions.detach().to("cpu", torch.float64).numpy()
        fb2, fn2 = m.at_positions(fb, fn, pos)
        st, r = context.scf_run(fb2, fn2, opts)
        if st != abi.FOCK_OK:
            raise abi.FockError(st, "SCF did not converge at this geometry")
        g = context.rhf_gradient(fb2, fn2, opts, r)
        fctx.save_for_backward(torch.as_tensor(g, dtype=positions.dtype, device=positions.device))
        return positions.new_tensor(r["total_energy"])

    @staticmethod
    def backward(fctx, grad_out):
        (g,) = fctx.saved_tensors
        return grad_out * g, None, None, None, None


def rhf_energy(positions: torch.Tensor, fb, fn, context: abi.Context, opts=None) -> torch.Tensor:
    """Total RHF energy (Eh) at `positions` (Bohr, shape (n_atoms, 3)), differentiable."""
    if opts is None:
        opts = abi.default_opts(max_iters=300, epsilon=1e-10, integral_policy=abi.POLICY_ACCURATE,
                                screen=abi.SCREEN_SCHWARZ, schwarz_threshold=1e-12)
    return RHFEnergy.apply(positions, fb, fn, context, opts)
