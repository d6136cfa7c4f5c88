"""The PyTorch autograd operator (xtbloom_amd.torch.xtbloom_torch).

Mirrors the reference operator's contract (reference
python/tests/test_torch_op.py): energies and forces as the C ABI returns them,
dE/dR = -F through autograd, gradient into the forces output rejected.  The
point-charge gradients are this package's extension and are checked against
central finite differences.
"""
import numpy as np
import pytest
import torch

from xtbloom_amd import interface, library, workloads
from xtbloom_amd.torch import XTBloomNotSupportedError, xtbloom_torch


def _small_batch():
    rng = np.random.default_rng(11)
    return workloads.from_systems([workloads.water(rng),
                                   workloads.alkane(3, rng, substituent="OH")])


def test_argument_checks_cpu():
    b = _small_batch()
    pos32 = torch.zeros((int(b.atom_offsets[-1]), 3), dtype=torch.float32)
    with pytest.raises(ValueError):
        xtbloom_torch(pos32, b.atomic_numbers, b.atom_offsets, b.molecular_charges,
                      b.unpaired_electrons)
    pos = torch.from_numpy(b.positions.copy())
    with pytest.raises(XTBloomNotSupportedError):
        xtbloom_torch(pos, b.atomic_numbers, b.atom_offsets, b.molecular_charges,
                      b.unpaired_electrons, method="GFN1-xTB")
    charges = torch.from_numpy(b.molecular_charges.copy()).requires_grad_()
    with pytest.raises(XTBloomNotSupportedError):
        xtbloom_torch(pos, b.atomic_numbers, b.atom_offsets, charges, b.unpaired_electrons)
    with pytest.raises(XTBloomNotSupportedError):
        xtbloom_torch(pos, b.atomic_numbers, b.atom_offsets, b.molecular_charges,
                      b.unpaired_electrons, spin_channels=[2, 1])


@pytest.mark.gpu
def test_energy_forces_and_position_gradient():
    b = _small_batch()
    dev = torch.device("cuda", 0)
    ctx = interface.Context("hip", 0)
    try:
        print("case torch-op positions", flush=True)
        pos = torch.from_numpy(b.positions.copy()).to(dev).requires_grad_()
        e, f = xtbloom_torch(pos, b.atomic_numbers, b.atom_offsets, b.molecular_charges,
                             b.unpaired_electrons, context=ctx)
        ref = interface.compute_arrays(ctx, b.atom_offsets, b.atomic_numbers, b.positions,
                                       b.molecular_charges, b.unpaired_electrons)
        assert np.abs(e.detach().cpu().numpy() - ref.energies).max() < 1e-12
        assert np.abs(f.detach().cpu().numpy() - ref.forces).max() < 1e-12
        w = torch.tensor([0.3, -1.7], dtype=torch.float64, device=dev)
        (w * e).sum().backward()
        sys_of_atom = np.repeat(np.arange(2), np.diff(b.atom_offsets))
        expect = -ref.forces * w.cpu().numpy()[sys_of_atom][:, None]
        assert np.abs(pos.grad.cpu().numpy() - expect).max() < 1e-12
        # gradient into the forces output and higher-order derivatives are rejected
        pos2 = torch.from_numpy(b.positions.copy()).to(dev).requires_grad_()
        e2, f2 = xtbloom_torch(pos2, b.atomic_numbers, b.atom_offsets, b.molecular_charges,
                               b.unpaired_electrons, context=ctx)
        with pytest.raises(XTBloomNotSupportedError):
            f2.sum().backward()
        pos3 = torch.from_numpy(b.positions.copy()).to(dev).requires_grad_()
        e3, _ = xtbloom_torch(pos3, b.atomic_numbers, b.atom_offsets, b.molecular_charges,
                              b.unpaired_electrons, context=ctx)
        with pytest.raises(XTBloomNotSupportedError):
            torch.autograd.grad(e3.sum(), pos3, create_graph=True)
    finally:
        ctx.close()


@pytest.mark.gpu
def test_point_charge_gradients_fd():
    rng = np.random.default_rng(5)
    s = workloads.water(rng)
    s.update(pc_xyz=np.array([[4.0, 1.0, -0.5], [-3.5, 2.0, 1.0]]), pc_q=np.array([0.4, -0.25]),
             pc_gamma=np.array([999.0, 999.0]))
    b = workloads.from_systems([s])
    dev = torch.device("cuda", 0)
    ctx = interface.Context("hip", 0)
    tight = dict(charge_tolerance=1e-10, energy_tolerance=1e-13)
    try:
        print("case torch-op point charges", flush=True)
        kw = lambda q, r: dict(point_offsets=b.point_offsets, point_positions=r,
                               point_values=q, point_gammas=b.point_gammas)
        q = torch.from_numpy(b.point_charges.copy()).to(dev).requires_grad_()
        r = torch.from_numpy(b.point_positions.copy()).to(dev).requires_grad_()
        pos = torch.from_numpy(b.positions.copy()).to(dev)
        e, _ = xtbloom_torch(pos, b.atomic_numbers, b.atom_offsets, b.molecular_charges,
                             b.unpaired_electrons, context=ctx, **tight, **kw(q, r))
        e.sum().backward()
        h = 1e-4

        def energy(qv, rv):
            out, _ = xtbloom_torch(pos, b.atomic_numbers, b.atom_offsets, b.molecular_charges,
                                   b.unpaired_electrons, context=ctx, **tight,
                                   **kw(torch.from_numpy(qv).to(dev), torch.from_numpy(rv).to(dev)))
            return float(out.sum())

        for p in range(2):
            qp, qm = b.point_charges.copy(), b.point_charges.copy()
            qp[p] += h
            qm[p] -= h
            fd = (energy(qp, b.point_positions) - energy(qm, b.point_positions)) / (2 * h)
            assert abs(fd - float(q.grad[p])) < 1e-6, (p, fd, float(q.grad[p]))
            for c in range(3):
                rp, rm = b.point_positions.copy(), b.point_positions.copy()
                rp[p, c] += h
                rm[p, c] -= h
                fd = (energy(b.point_charges, rp) - energy(b.point_charges, rm)) / (2 * h)
                assert abs(fd - float(r.grad[p, c])) < 1e-6, (p, c, fd, float(r.grad[p, c]))
    finally:
        ctx.close()
