"""Final-round Fermi occupations on the GPU against the oracle's occupation function.

For every system of a batch, the GPU's final-round orbital energies are fed
to the oracle's Fermi solve (the reference's compute_occupations restated:
long-double bisection, binary64 policy for degenerate frontiers) and the
published occupations are compared with the GPU's.  The batch holds a
fluoride ion with molecular charge -1 + 2^-50 (four levels, 2s and 2p): each
spin channel holds 4 - 2^-51 electrons, a hole count of 2^-51 that the
rounded Fermi occupations miss, so both sides publish through the binary64
policy, which must conserve the hole count within 64 eps of it; the other
members exercise the ordinary path.
"""
import ctypes

import numpy as np
import pytest

from oracle import oracle as O
from xtbloom_amd import interface, workloads

pytestmark = pytest.mark.gpu

KT = 300.0 * 3.166808578545117e-6


def oracle_occupations(levels, nel):
    e = np.ascontiguousarray(levels, dtype=float)
    occ = np.zeros(len(e))
    mu, ent = ctypes.c_double(), ctypes.c_double()
    assert O.lib().oracle_fermi(len(e), O._p(e), float(nel), KT, O._p(occ), ctypes.byref(mu),
                                ctypes.byref(ent))
    return occ


def test_occupations_match_the_oracle_function():
    rng = np.random.default_rng(5)
    fluoride = dict(Z=np.array([9]), xyz=np.zeros((1, 3)), charge=-1.0 + 2.0 ** -50, uhf=0)
    b = workloads.from_systems([fluoride, workloads.water(rng),
                                workloads.alkane(5, rng, substituent="OH")])
    ctx = interface.Context("hip", 0)
    try:
        print("case occupations: F (charge -1 + 2^-50), water, C5H11OH", flush=True)
        out = interface.compute_arrays(ctx, b.atom_offsets, b.atomic_numbers, b.positions,
                                       b.molecular_charges, b.unpaired_electrons, diagnostics=True)
    finally:
        ctx.close()
    assert np.all(out.per_system_status == 0)
    lev, occ = out.extension["orbital_energies"], out.extension["occupations"]
    at = 0
    for i in range(b.batch_size):
        s = b.system(i)
        _, n, nel0 = O.basis_counts(s["Z"])
        nel = nel0 - b.molecular_charges[i]
        e = lev[at:at + n]
        f = occ[at:at + n]
        at += n
        assert np.all(np.diff(e) >= 0.0)
        ref = oracle_occupations(e, 0.5 * nel) + oracle_occupations(e, 0.5 * nel)
        print(f"system {i}: n {n} max |d occ| {np.abs(f - ref).max():.2e}", flush=True)
        assert np.abs(f - ref).max() <= 4e-16
        if i == 0:
            print("F levels", e.tolist(), "occupations", f.tolist(), flush=True)
            # each spin channel holds 4 - 2^-51 electrons: the published
            # occupations conserve the hole count 2^-51 within 64 eps of it
            holes = sum(1.0 - x / 2.0 for x in f)
            assert abs(holes - 2.0 ** -51) <= 64 * np.finfo(float).eps * 2.0 ** -51
            assert np.array_equal(f, ref)
