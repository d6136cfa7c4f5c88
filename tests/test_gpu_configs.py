"""The benchmark configurations run on the GPU as their workload text states.

* 128 uniform C20H42 conformers (62 atoms, 122 AOs), plain FP64 path, FRESH
  (cold) start: every member against the oracle (E 1e-7 Eh, F 1e-6 Eh/bohr,
  q 1e-6 e, SCC iterations within +-1), plus size-independent properties of
  the whole batch (zero net force and torque, charge conservation).
* Differentiation through the autograd operator on 16 members of the
  256-system embedded batch (500 external point charges each): dE/dR and
  dE/dq_ext against FP64 central differences (h = 1e-4 bohr / 1e-4 e) on 16
  seeded coordinates and 16 seeded point charges, and -dE/dR against the
  oracle's forces (the reference operator's own check, reference
  python/tests/test_torch_op.py).
"""
import os

import numpy as np
import pytest
import torch

from oracle import oracle as O
from xtbloom_amd import interface, library, workloads
from xtbloom_amd.torch import xtbloom_torch

pytestmark = pytest.mark.gpu

E_TOL, F_TOL, Q_TOL = 1e-7, 1e-6, 1e-6


def _threads():
    return max(1, min(len(os.sched_getaffinity(0)), 16))


def _oracle_batch(b, opts=None, want_point_gradient=False):
    kw = {}
    if b.point_offsets is not None:
        kw = dict(point_offsets=b.point_offsets, pc_xyz=b.point_positions, pc_q=b.point_charges,
                  pc_gamma=b.point_gammas, want_point_gradient=want_point_gradient)
    return O.batch(b.atom_offsets, b.atomic_numbers, b.positions, b.molecular_charges,
                   b.unpaired_electrons, opts=opts, threads=_threads(), **kw)


def test_c20h42_128_cold():
    b = workloads.make("c20h42_128", seed=0)
    assert b.batch_size == 128 and np.all(np.diff(b.atom_offsets) == 62)
    ctx = interface.Context("hip", 0)
    try:
        print("case c20h42_128 FRESH", flush=True)
        out = interface.compute_arrays(ctx, b.atom_offsets, b.atomic_numbers, b.positions,
                                       b.molecular_charges, b.unpaired_electrons)
    finally:
        ctx.close()
    ref = _oracle_batch(b)
    assert np.all(ref["status"] == 0)
    assert np.all(out.per_system_status == 0) and np.all(out.scc_converged)
    de = np.abs(out.energies - ref["energies"]).max()
    df = np.abs(out.forces + ref["gradient"]).max()
    dq = np.abs(out.charges - ref["charges"]).max()
    dit = np.abs(out.scc_iterations.astype(int) - ref["iterations"].astype(int)).max()
    print(f"c20h42_128: |dE| {de:.2e} |dF| {df:.2e} |dq| {dq:.2e} d_iter {dit}", flush=True)
    assert de <= E_TOL and df <= F_TOL and dq <= Q_TOL and dit <= 1
    # size-independent properties of every member.  The net force vanishes to
    # rounding (every pair term is antisymmetric); the net torque vanishes only
    # at exact SCC stationarity, so at the default charge tolerance (1e-6) it
    # is bounded by the force tolerance and must equal the oracle's torque.
    for i in range(b.batch_size):
        a0, a1 = b.atom_offsets[i], b.atom_offsets[i + 1]
        f = out.forces[a0:a1]
        r = b.positions[a0:a1]
        tau = np.cross(r, f).sum(axis=0)
        tau_ref = np.cross(r, -ref["gradient"][a0:a1]).sum(axis=0)
        assert np.abs(f.sum(axis=0)).max() < 1e-9, i                     # translation
        assert np.abs(tau).max() < F_TOL, i                                # rotation
        assert np.abs(tau - tau_ref).max() < 1e-9, i
        assert abs(out.charges[a0:a1].sum()) < 1e-9, i                     # neutral
    assert np.all(np.isfinite(out.energies))


def _perturbed_batch(systems):
    return workloads.from_systems(systems)


def test_embedded256_autograd_fd():
    full = workloads.make("embedded256", seed=0)
    assert full.batch_size == 256
    assert np.all(np.diff(full.point_offsets) == 500)
    rng = np.random.default_rng(2024)
    members = sorted(rng.choice(full.batch_size, size=16, replace=False).tolist())
    b = full.subset(members)
    dev = torch.device("cuda", 0)
    ctx = interface.Context("hip", 0)
    tight = dict(charge_tolerance=1e-10, energy_tolerance=1e-12)
    try:
        print("case embedded256 x16 autograd", members, flush=True)
        pos = torch.from_numpy(b.positions.copy()).to(dev).requires_grad_()
        q = torch.from_numpy(b.point_charges.copy()).to(dev).requires_grad_()
        r = torch.from_numpy(b.point_positions.copy()).to(dev)
        e, f = xtbloom_torch(pos, b.atomic_numbers, b.atom_offsets, b.molecular_charges,
                             b.unpaired_electrons, context=ctx, point_offsets=b.point_offsets,
                             point_positions=r, point_values=q, point_gammas=b.point_gammas,
                             **tight)
        e.sum().backward()
        dedr = pos.grad.cpu().numpy()
        dedq = q.grad.cpu().numpy()
        assert np.abs(dedr + f.detach().cpu().numpy()).max() == 0.0

        # -dE/dR against the oracle's forces at the default SCC policy
        ref = _oracle_batch(b)
        assert np.all(ref["status"] == 0)
        e0, f0 = xtbloom_torch(torch.from_numpy(b.positions.copy()).to(dev), b.atomic_numbers,
                               b.atom_offsets, b.molecular_charges, b.unpaired_electrons,
                               context=ctx, point_offsets=b.point_offsets, point_positions=r,
                               point_values=q.detach(), point_gammas=b.point_gammas)
        dfo = np.abs(f0.cpu().numpy() + ref["gradient"]).max()
        deo = np.abs(e0.cpu().numpy() - ref["energies"]).max()
        print(f"embedded x16 vs oracle: |dE| {deo:.2e} |dF| {dfo:.2e}", flush=True)
        assert deo <= E_TOL and dfo <= F_TOL

        # central differences: every perturbed copy of a member is one system of
        # a single batch (each copy's energy depends on its own coordinates only)
        h = 1e-4
        n_atoms = int(b.atom_offsets[-1])
        coords = rng.choice(3 * n_atoms, size=16, replace=False)
        sys_of_atom = np.repeat(np.arange(b.batch_size), np.diff(b.atom_offsets))
        systems, index = [], []
        for c in coords:
            atom, x = divmod(int(c), 3)
            i = int(sys_of_atom[atom])
            for sgn in (1.0, -1.0):
                s = b.system(i)
                xyz = s["xyz"].copy()
                xyz[atom - b.atom_offsets[i], x] += sgn * h
                systems.append(dict(s, xyz=xyz))
            index.append(("R", atom, x, i))
        n_pc = int(b.point_offsets[-1])
        pcs = rng.choice(n_pc, size=16, replace=False)
        sys_of_pc = np.repeat(np.arange(b.batch_size), np.diff(b.point_offsets))
        for p in pcs:
            i = int(sys_of_pc[p])
            for sgn in (1.0, -1.0):
                s = b.system(i)
                qq = s["pc_q"].copy()
                qq[p - b.point_offsets[i]] += sgn * h
                systems.append(dict(s, pc_q=qq))
            index.append(("Q", int(p), 0, i))
        fdb = _perturbed_batch(systems)
        print(f"case embedded x16 central differences ({fdb.batch_size} systems)", flush=True)
        out = interface.compute_arrays(
            ctx, fdb.atom_offsets, fdb.atomic_numbers, fdb.positions, fdb.molecular_charges,
            fdb.unpaired_electrons, point_offsets=fdb.point_offsets,
            point_positions=fdb.point_positions, point_values=fdb.point_charges,
            point_gammas=fdb.point_gammas, **tight)
        assert np.all(out.per_system_status == 0)
        worst_r = worst_q = 0.0
        for k, (kind, item, x, i) in enumerate(index):
            fd = (out.energies[2 * k] - out.energies[2 * k + 1]) / (2.0 * h)
            if kind == "R":
                err = abs(fd - dedr[item, x])
                worst_r = max(worst_r, err)
            else:
                err = abs(fd - dedq[item])
                worst_q = max(worst_q, err)
            assert err < 1e-6, (kind, item, x, fd)
        print(f"embedded x16 FD: dE/dR {worst_r:.2e}  dE/dq {worst_q:.2e}", flush=True)
    finally:
        ctx.close()
