"""The batched path (XTBLOOM_AMD_PATH=batched) against the plain FP64 path and the oracle.

The batched path runs the SCC rounds as captured HIP graphs of several rounds
each, driven by a device-resident round counter, and the captured graph is
replayed by later calls with the same topology.  Every replay is checked:
the batch is computed several times on the batched path with plain-path calls
(and host synchronizations) in between, and each result is compared with the
plain path's and with the oracle's.  Energies 1e-7 Eh, forces 1e-6 Eh/bohr,
charges 1e-6 e against the oracle; the iteration limit and per-system
statuses behave as on the plain path.
"""
import numpy as np
import pytest

from oracle import oracle as O
from xtbloom_amd import interface, library, workloads

pytestmark = pytest.mark.gpu

E_TOL, F_TOL, Q_TOL = 1e-7, 1e-6, 1e-6


def _call(ctx, b, path, monkeypatch, **kw):
    monkeypatch.setenv("XTBLOOM_AMD_PATH", path)
    return interface.compute_arrays(ctx, b.atom_offsets, b.atomic_numbers, b.positions,
                                    b.molecular_charges, b.unpaired_electrons,
                                    point_offsets=b.point_offsets,
                                    point_positions=b.point_positions,
                                    point_values=b.point_charges,
                                    point_gammas=b.point_gammas, **kw)


def _agree(a, b, e=E_TOL, f=F_TOL, q=Q_TOL, it=1):
    assert np.array_equal(a.per_system_status, b.per_system_status)
    ok = a.per_system_status == 0
    assert np.abs(a.energies[ok] - b.energies[ok]).max(initial=0.0) <= e
    assert np.nanmax(np.abs(a.forces - b.forces), initial=0.0) <= f
    assert np.nanmax(np.abs(a.charges - b.charges), initial=0.0) <= q
    assert np.abs(a.scc_iterations.astype(int) - b.scc_iterations.astype(int)).max() <= it


def test_batched_replays_match_plain_and_oracle(monkeypatch):
    b = workloads.ragged_alkanes(batch=24, kmin=2, kmax=38, seed=3, points=20)
    ctx = interface.Context("hip", 0)
    try:
        print("case batched path: 24 ragged systems, 3 replays", flush=True)
        plain = _call(ctx, b, "plain", monkeypatch)
        outs = []
        for rep in range(3):
            outs.append(_call(ctx, b, "batched", monkeypatch))
            again = _call(ctx, b, "plain", monkeypatch)  # eager work between replays
            _agree(again, plain, 1e-12, 1e-12, 1e-12, 0)
        for rep, out in enumerate(outs):
            print(f"replay {rep}: max |dE| vs plain "
                  f"{np.abs(out.energies - plain.energies).max():.2e}", flush=True)
            _agree(out, plain)
    finally:
        ctx.close()
    ref = O.batch(b.atom_offsets, b.atomic_numbers, b.positions, b.molecular_charges,
                  b.unpaired_electrons, point_offsets=b.point_offsets, pc_xyz=b.point_positions,
                  pc_q=b.point_charges, pc_gamma=b.point_gammas, threads=8)
    for out in outs:
        assert np.all(out.per_system_status == 0)
        assert np.abs(out.energies - ref["energies"]).max() <= E_TOL
        assert np.abs(out.forces + ref["gradient"]).max() <= F_TOL
        assert np.abs(out.charges - ref["charges"]).max() <= Q_TOL
        assert np.abs(out.scc_iterations.astype(int) - ref["iterations"].astype(int)).max() <= 1


def test_batched_iteration_limit_and_statuses(monkeypatch):
    b = workloads.ragged_alkanes(batch=6, kmin=3, kmax=12, seed=8)
    ctx = interface.Context("hip", 0)
    try:
        print("case batched path at an iteration limit of 5", flush=True)
        plain = _call(ctx, b, "plain", monkeypatch, max_scc_iterations=5)
        batched = _call(ctx, b, "batched", monkeypatch, max_scc_iterations=5)
    finally:
        ctx.close()
    assert np.all(plain.per_system_status == library.STATUS_SCC_NOT_CONVERGED)
    assert np.array_equal(batched.per_system_status, plain.per_system_status)
    assert np.all(batched.scc_iterations == 5) and np.all(np.isnan(batched.energies))


def test_batched_reproducible_mode_repeats_bit_for_bit(monkeypatch):
    b = workloads.ragged_alkanes(batch=10, kmin=2, kmax=30, seed=12)
    ctx = interface.Context("hip", 0)
    try:
        print("case batched path, reproducible determinism, two calls", flush=True)
        r = library.DETERMINISM_REPRODUCIBLE
        first = _call(ctx, b, "batched", monkeypatch, determinism=r)
        second = _call(ctx, b, "batched", monkeypatch, determinism=r)
    finally:
        ctx.close()
    for name in ("energies", "forces", "charges", "scc_iterations", "per_system_status"):
        assert np.array_equal(getattr(first, name), getattr(second, name)), name
