"""Split-precision FP32 density products (XTBLOOM_AMD_SPLIT32=1) against the oracle.

P = C f C^T runs on the FP32 matrix cores (hi/lo split operands, FP32
accumulation over 16-deep K stages, FP64 across stages) in the SCC rounds
whose mixer residual is above 1e-3, W = C f eps C^T always; the final SCC
rounds keep the FP64 product, so the converged state is an FP64 one.  Bars:
energies 1e-7 Eh, forces 1e-6 Eh/bohr, charges 1e-6 e, SCC iterations +-1
against the oracle, on both execution paths.  The measured errors are printed.
"""
import numpy as np
import pytest

from oracle import oracle as O
from xtbloom_amd import interface, workloads

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("path", ["plain", "batched"])
def test_split32_density_products_against_oracle(path, monkeypatch):
    b = workloads.ragged_alkanes(batch=16, kmin=2, kmax=38, seed=21, points=30)
    monkeypatch.setenv("XTBLOOM_AMD_PATH", path)
    monkeypatch.setenv("XTBLOOM_AMD_SPLIT32", "1")
    ctx = interface.Context("hip", 0)
    try:
        print(f"case split32 density products, {path} path, 16 ragged systems", flush=True)
        out = interface.compute_arrays(ctx, b.atom_offsets, b.atomic_numbers, b.positions,
                                       b.molecular_charges, b.unpaired_electrons,
                                       point_offsets=b.point_offsets,
                                       point_positions=b.point_positions,
                                       point_values=b.point_charges, point_gammas=b.point_gammas)
    finally:
        ctx.close()
    ref = O.batch(b.atom_offsets, b.atomic_numbers, b.positions, b.molecular_charges,
                  b.unpaired_electrons, point_offsets=b.point_offsets, pc_xyz=b.point_positions,
                  pc_q=b.point_charges, pc_gamma=b.point_gammas, threads=8)
    assert np.all(out.per_system_status == 0) and np.all(ref["status"] == 0)
    de = np.abs(out.energies - ref["energies"]).max()
    df = np.abs(out.forces + ref["gradient"]).max()
    dq = np.abs(out.charges - ref["charges"]).max()
    dit = np.abs(out.scc_iterations.astype(int) - ref["iterations"].astype(int)).max()
    print(f"split32 {path}: |dE| {de:.2e} |dF| {df:.2e} |dq| {dq:.2e} d_iter {dit}", flush=True)
    assert de <= 1e-7 and df <= 1e-6 and dq <= 1e-6 and dit <= 1
