"""Unrestricted (two-channel) systems on the GPU against the oracle.

An unrestricted member runs as its alpha system and a beta copy appended to
the batch: the Hamiltonians H0 + V_q +/- W m, one eigensolve per channel
with its own electron count, the state's charge and magnetization blocks
mixed together, the spin energy m.W.m/2, and the magnetization potential in
the overlap-derivative forces (reference spin.hpp:106-118,
scc_driver.cpp:2593-2640, force.cpp:506-533).  Checks: radicals (OH and
alkyl radicals) mixed with restricted members against the oracle; a closed
shell with two channels equals the restricted result; forces against
central differences of the GPU energies; the batched path; strict WARM with
its spin identity; the Calculator's open-shell default.
"""
import numpy as np
import pytest

from oracle import oracle as O
from xtbloom_amd import interface, library, workloads

pytestmark = pytest.mark.gpu

E_TOL, F_TOL, Q_TOL = 1e-7, 1e-6, 1e-6
TIGHT = dict(charge_tolerance=1e-10, energy_tolerance=1e-12, max_scc_iterations=500)


def radical(k, rng):
    """C_kH_(2k+1): an alkane with its last hydrogen removed (one unpaired electron)."""
    s = workloads.alkane(k, rng)
    return dict(Z=s["Z"][:-1], xyz=s["xyz"][:-1], charge=0.0, uhf=1)


def systems():
    rng = np.random.default_rng(17)
    oh = dict(Z=np.array([8, 1]), xyz=np.array([[0.0, 0.0, 0.0], [1.83, 0.1, -0.05]]), charge=0.0, uhf=1)
    members = [oh, workloads.water(rng), radical(2, rng), radical(4, rng),
               workloads.alkane(3, rng, substituent="NH2"), radical(9, rng)]
    spin = np.array([2, 1, 2, 2, 1, 2], np.int32)
    return members, spin


def run(ctx, b, spin, **kw):
    return interface.compute_arrays(ctx, b.atom_offsets, b.atomic_numbers, b.positions,
                                    b.molecular_charges, b.unpaired_electrons, spin_channels=spin, **kw)


def test_mixed_batch_against_oracle():
    members, spin = systems()
    b = workloads.from_systems(members)
    ctx = interface.Context("hip", 0)
    try:
        print("case unrestricted batch", spin.tolist(), flush=True)
        out = run(ctx, b, spin)
    finally:
        ctx.close()
    ref = O.batch(b.atom_offsets, b.atomic_numbers, b.positions, b.molecular_charges,
                  b.unpaired_electrons, spin_channels=spin, threads=4)
    assert np.all(out.per_system_status == 0) and np.all(ref["status"] == 0)
    dE = np.abs(out.energies - ref["energies"]).max()
    dF = np.abs(out.forces + ref["gradient"]).max()
    dq = np.abs(out.charges - ref["charges"]).max()
    print(f"unrestricted vs oracle: dE {dE:.2e} dF {dF:.2e} dq {dq:.2e} iterations "
          f"{out.scc_iterations.tolist()} oracle {ref['iterations'].tolist()}", flush=True)
    assert dE <= E_TOL and dF <= F_TOL and dq <= Q_TOL
    assert np.abs(out.scc_iterations - ref["iterations"]).max() <= 1
    # spin polarization lowers the radicals' energies below the restricted ones
    restricted = O.batch(b.atom_offsets, b.atomic_numbers, b.positions, b.molecular_charges,
                         b.unpaired_electrons, threads=4)
    assert np.all(out.energies[spin == 2] < restricted["energies"][spin == 2] - 1e-5)


def test_closed_shell_with_two_channels_equals_restricted():
    rng = np.random.default_rng(5)
    b = workloads.from_systems([workloads.water(rng), workloads.alkane(4, rng, substituent="OH")])
    ctx = interface.Context("hip", 0)
    try:
        print("case closed shells with two channels", flush=True)
        two = run(ctx, b, np.array([2, 2], np.int32), **TIGHT)
        one = run(ctx, b, None, **TIGHT)
    finally:
        ctx.close()
    assert np.abs(two.energies - one.energies).max() < 1e-9
    assert np.abs(two.forces - one.forces).max() < 1e-8
    assert np.abs(two.charges - one.charges).max() < 1e-8


def test_forces_central_difference():
    members, _ = systems()
    s = members[3]  # C4H9
    spin = np.array([2], np.int32)
    ctx = interface.Context("hip", 0)
    try:
        print("case unrestricted forces vs central differences", flush=True)
        base = run(ctx, workloads.from_systems([s]), spin, **TIGHT)
        h = 1e-4
        for atom, axis in ((0, 0), (3, 1), (len(s["Z"]) - 1, 2)):
            e = []
            for sign in (1, -1):
                xyz = s["xyz"].copy()
                xyz[atom, axis] += sign * h
                e.append(run(ctx, workloads.from_systems([dict(s, xyz=xyz)]), spin, **TIGHT).energies[0])
            assert abs(-(e[0] - e[1]) / (2 * h) - base.forces[atom, axis]) < 2e-7
        assert np.abs(base.forces.sum(axis=0)).max() < 1e-9
    finally:
        ctx.close()


def test_batched_path_matches_plain(monkeypatch):
    members, spin = systems()
    b = workloads.from_systems(members)
    ctx = interface.Context("hip", 0)
    try:
        print("case unrestricted batch on the batched path", flush=True)
        monkeypatch.setenv("XTBLOOM_AMD_PATH", "plain")
        plain = run(ctx, b, spin)
        monkeypatch.setenv("XTBLOOM_AMD_PATH", "batched")
        outs = [run(ctx, b, spin) for _ in range(2)]
    finally:
        ctx.close()
    for got in outs:
        assert np.array_equal(got.per_system_status, plain.per_system_status)
        assert np.abs(got.energies - plain.energies).max() <= E_TOL
        assert np.abs(got.forces - plain.forces).max() <= F_TOL
        assert np.abs(got.scc_iterations.astype(int) - plain.scc_iterations.astype(int)).max() <= 1


def test_warm_start_and_spin_identity():
    members, spin = systems()
    b = workloads.from_systems(members)
    ctx = interface.Context("hip", 0)
    try:
        print("case unrestricted WARM start", flush=True)
        cold = run(ctx, b, spin)
        warm = run(ctx, b, spin, scc_start_mode=library.SCC_START_WARM)
        assert np.all(warm.per_system_status == 0)
        assert warm.scc_iterations.max() <= 3
        assert np.abs(warm.energies - cold.energies).max() < 1e-7
        other = spin.copy()
        other[0] = 1
        with pytest.raises(interface.XTBloomError) as err:
            run(ctx, b, other, scc_start_mode=library.SCC_START_WARM)
        assert err.value.status == library.STATUS_INVALID_ARGUMENT
        # an absent spin buffer and explicit ones are the same identity
        plain = run(ctx, b, None)
        print("restricted call", plain.per_system_status.tolist(), plain.scc_iterations.tolist(), flush=True)
        assert np.all(plain.per_system_status == 0)
        again = run(ctx, b, None, scc_start_mode=library.SCC_START_WARM)
        assert np.all(again.per_system_status == 0)
        again = run(ctx, b, np.ones(b.batch_size, np.int32), scc_start_mode=library.SCC_START_WARM)
        assert np.all(again.per_system_status == 0)
    finally:
        ctx.close()


def test_calculator_defaults_open_shells_to_unrestricted():
    oh = dict(Z=[8, 1], xyz=np.array([[0.0, 0.0, 0.0], [1.83, 0.0, 0.0]]))
    calc = interface.Calculator("GFN2-xTB", oh["Z"], oh["xyz"], uhf=1)
    restricted = interface.Calculator("GFN2-xTB", oh["Z"], oh["xyz"], uhf=1, spin_channels=1)
    try:
        print("case Calculator open-shell default", flush=True)
        assert calc.spin_channels == 2 and restricted.spin_channels == 1
        r2 = calc.singlepoint()
        r1 = restricted.singlepoint()
        ref = O.singlepoint(oh["Z"], oh["xyz"], uhf=1, opts=O.options(spin_channels=2))
        assert abs(r2.energy - ref.energy) < E_TOL
        assert r2.energy < r1.energy - 1e-4
    finally:
        calc.close()
        restricted.close()
    with pytest.raises(interface.XTBloomValueError):
        interface.Structure(oh["Z"], oh["xyz"], uhf=1, spin_channels=3)
