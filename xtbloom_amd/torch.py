"""PyTorch autograd operator over the MI355X GFN2 batch path.

:func:`xtbloom_torch` mirrors the reference operator of the same name
(reference ``python/xtbloom/torch.py:838-961``): it runs a packed ragged batch
on torch tensors and returns ``(energies, forces)``, with the analytic
gradient ``dE/dR = -F`` with respect to the positions.  CUDA/ROCm tensors stay
on the device (they are passed to the C ABI as device buffers).

The contract follows the reference:

* ``positions`` may require gradient; gradient flowing into the ``forces``
  output (the force Hessian) raises :class:`XTBloomNotSupportedError`, and so
  does higher-order differentiation;
* every input is detached before the native call.

Extension (no counterpart in the reference operator): optional external point
charges.  ``point_values`` and ``point_positions`` may require gradient; their
gradients are ``dE/dQ_p`` (the library's point-charge potential gradient) and
``dE/dR_p = -F_p`` (the point-charge forces).
"""
from __future__ import annotations

import torch

from . import interface, library

__all__ = ["xtbloom_torch", "XTBloomNotSupportedError"]


class XTBloomNotSupportedError(RuntimeError):
    """Raised for derivatives this operator does not provide."""


def _system_index(offsets: torch.Tensor, count: int, device) -> torch.Tensor:
    """System index of each of `count` items of a ragged [B+1] offset array."""
    sizes = (offsets[1:] - offsets[:-1]).to(device=device, dtype=torch.int64)
    return torch.repeat_interleave(torch.arange(sizes.numel(), device=device), sizes,
                                   output_size=count)


class _GFN2Function(torch.autograd.Function):
    @staticmethod
    def forward(fctx, positions, point_positions, point_values, call):
        ctx, args, opts = call
        pos = positions.detach().contiguous()
        pc = {}
        if point_values is not None:
            pc = dict(point_offsets=args["point_offsets"],
                      point_positions=point_positions.detach().contiguous(),
                      point_values=point_values.detach().contiguous(),
                      point_gammas=args["point_gammas"])
        flags = library.COMPUTE_ENERGY | library.COMPUTE_FORCES
        if pc:
            flags |= library.COMPUTE_POINT_CHARGE_FORCES
        out, _ = interface.compute_tensors(
            ctx, args["atom_offsets"], args["numbers"], pos, args["charges"], args["uhf"], **pc,
            flags=flags, point_charge_potential_gradient=bool(pc), **opts)
        energies, forces = out["energies"], out["forces"]
        saved = [forces]
        if pc:
            saved += [out["point_charge_forces"], out["point_charge_potential_gradient"]]
        fctx.save_for_backward(*saved)
        fctx.has_pc = bool(pc)
        fctx.atom_offsets = args["atom_offsets"]
        fctx.point_offsets = args.get("point_offsets")
        return energies, forces

    @staticmethod
    def backward(fctx, grad_energies, grad_forces):
        if torch.is_grad_enabled():
            raise XTBloomNotSupportedError(
                "higher-order derivatives (create_graph=True) are not available")
        if grad_forces is not None and bool(torch.any(grad_forces != 0)):
            raise XTBloomNotSupportedError(
                "gradient through the forces output (the force Hessian) is not available")
        saved = fctx.saved_tensors
        forces = saved[0]
        dev = forces.device
        g_e = grad_energies.to(dev)
        atom_sys = _system_index(fctx.atom_offsets, forces.shape[0], dev)
        grad_pos = -forces * g_e[atom_sys].unsqueeze(1)
        grad_pc_pos = grad_pc_val = None
        if fctx.has_pc:
            pc_forces, dedq = saved[1], saved[2]
            pc_sys = _system_index(fctx.point_offsets, pc_forces.shape[0], dev)
            grad_pc_pos = -pc_forces * g_e[pc_sys].unsqueeze(1)
            grad_pc_val = dedq.to(dev) * g_e[pc_sys]
        return grad_pos, grad_pc_pos, grad_pc_val, None


def xtbloom_torch(positions, atomic_numbers, atom_offsets, molecular_charges, unpaired_electrons,
                  spin_channels=None, *, method: str = "GFN2-xTB", backend: str = "hip",
                  device_id: int = 0, context: interface.Context | None = None,
                  max_scc_iterations: int = 250, charge_tolerance: float = 1.0e-6,
                  energy_tolerance: float = 1.0e-8, electronic_temperature: float = 300.0,
                  scc_mixer_history: int = 8, scc_mixer_damping: float = 0.4,
                  point_offsets=None, point_positions=None, point_values=None,
                  point_gammas=None):
    """GFN2 energies and forces of a ragged batch with ``dE/dR = -F`` autograd.

    ``positions`` (natoms, 3) float64 bohr; ``atomic_numbers`` int32;
    ``atom_offsets`` int64 [B+1]; ``molecular_charges`` float64 [B];
    ``unpaired_electrons`` int32 [B].  Returns ``(energies [B], forces
    [natoms, 3])``.  ``spin_channels`` must be absent or all 1 (restricted).
    ``context`` reuses an open :class:`xtbloom_amd.interface.Context`; by
    default one is opened per call on ``device_id``.
    """
    if method not in ("GFN2-xTB", "GFN2"):
        raise XTBloomNotSupportedError(f"method {method!r} is not available on this backend")
    if backend not in ("hip", "auto"):
        raise XTBloomNotSupportedError(f"backend {backend!r} is not available in this library")
    if not torch.is_tensor(positions) or positions.dtype != torch.float64 or \
            positions.dim() != 2 or positions.shape[1] != 3:
        raise ValueError("positions must be a float64 tensor of shape (natoms, 3)")
    dev = positions.device

    def as_t(x, dtype):
        t = x if torch.is_tensor(x) else torch.as_tensor(x)
        if t.requires_grad:
            raise XTBloomNotSupportedError("only positions and point charges take gradients")
        return t.detach().to(device=dev, dtype=dtype).contiguous()

    if spin_channels is not None:
        sc = torch.as_tensor(spin_channels)
        if bool(torch.any(sc != 1)):
            raise XTBloomNotSupportedError("unrestricted spin channels are not available")
    args = dict(atom_offsets=as_t(atom_offsets, torch.int64),
                numbers=as_t(atomic_numbers, torch.int32),
                charges=as_t(molecular_charges, torch.float64),
                uhf=as_t(unpaired_electrons, torch.int32))
    if point_values is not None:
        if point_positions is None or point_offsets is None or point_gammas is None:
            raise ValueError("point charges need offsets, positions, values and gammas")
        args["point_offsets"] = as_t(point_offsets, torch.int64)
        args["point_gammas"] = as_t(point_gammas, torch.float64)
        point_positions = point_positions.to(device=dev, dtype=torch.float64)
        point_values = point_values.to(device=dev, dtype=torch.float64)
    opts = dict(max_scc_iterations=max_scc_iterations, charge_tolerance=charge_tolerance,
                energy_tolerance=energy_tolerance, electronic_temperature=electronic_temperature,
                scc_mixer_history=scc_mixer_history, scc_mixer_damping=scc_mixer_damping)
    own = context is None
    ctx = interface.Context("hip", device_id) if own else context
    try:
        return _GFN2Function.apply(positions, point_positions, point_values, (ctx, args, opts))
    finally:
        if own:
            ctx.close()
