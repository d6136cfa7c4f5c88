"""Unrestricted systems through every call path, and combined with the other
external terms, on the GPU.

An unrestricted member is expanded into an alpha system and a beta copy
inside each lane's run; these checks cover what sits around that expansion:
concurrent lanes (slices of the batch), fixed-topology plans with WARM
starts, asynchronous enqueue, the no-allocation entry point, the packed
diagnostics, and a radical with point charges, a uniform field and the
b + A q operator at once against the oracle.
"""
import numpy as np
import pytest

from oracle import oracle as O
from xtbloom_amd import interface, library, workloads

pytestmark = pytest.mark.gpu


def _radical(k, rng):
    s = workloads.alkane(k, rng)
    return dict(Z=s["Z"][:-1], xyz=s["xyz"][:-1], charge=0.0, uhf=1)


def _batch(count=24, seed=31):
    rng = np.random.default_rng(seed)
    members, spin = [], []
    for i in range(count):
        k = int(rng.integers(2, 9))
        if i % 2:
            members.append(_radical(k, rng))
            spin.append(2)
        else:
            members.append(workloads.alkane(k, rng, substituent="OH" if i % 4 == 0 else None))
            spin.append(1)
    return workloads.from_systems(members), np.array(spin, np.int32)


def _run(ctx, b, spin, **kw):
    return interface.compute_arrays(ctx, b.atom_offsets, b.atomic_numbers, b.positions,
                                    b.molecular_charges, b.unpaired_electrons, spin_channels=spin, **kw)


def _same(a, b):
    for name in ("energies", "charges", "scc_iterations", "per_system_status"):
        assert np.array_equal(getattr(a, name), getattr(b, name)), name
    assert np.abs(a.forces - b.forces).max() < 1e-11


def test_lanes_plans_enqueue_and_no_alloc(monkeypatch):
    b, spin = _batch()
    ctx = interface.Context("hip", 0)
    try:
        print("case unrestricted: one lane", flush=True)
        ref = _run(ctx, b, spin)
        assert np.all(ref.per_system_status == 0)
        for lanes in ("2", "3"):
            monkeypatch.setenv("XTBLOOM_AMD_LANES", lanes)
            print("case unrestricted: lanes", lanes, flush=True)
            _same(_run(ctx, b, spin), ref)
        monkeypatch.setenv("XTBLOOM_AMD_LANES", "1")
        print("case unrestricted: no-allocation entry point", flush=True)
        _same(_run(ctx, b, spin, no_alloc=True), ref)
        print("case unrestricted: enqueue", flush=True)
        req = interface.Request(ctx)
        try:
            _same(_run(ctx, b, spin, request=req), ref)
        finally:
            req.close()
        print("case unrestricted: plan", flush=True)
        plan = interface.Plan(ctx, b.atom_offsets, b.atomic_numbers, b.positions, b.molecular_charges,
                              b.unpaired_electrons, spin_channels=spin)
        try:
            _same(plan.compute(b.positions), ref)
            warm = plan.compute(b.positions, scc_start_mode=library.SCC_START_WARM)
            assert np.all(warm.per_system_status == 0) and warm.scc_iterations.max() <= 3
            assert np.abs(warm.energies - ref.energies).max() < 1e-7
            # the plan's identity includes the spin channels
            with pytest.raises(interface.XTBloomError):
                interface.compute_arrays(ctx, b.atom_offsets, b.atomic_numbers, b.positions,
                                         b.molecular_charges, b.unpaired_electrons, plan=plan)
        finally:
            plan.close()
    finally:
        ctx.close()


def test_diagnostics_cover_the_members():
    b, spin = _batch(count=4, seed=3)
    ctx = interface.Context("hip", 0)
    try:
        print("case unrestricted diagnostics", flush=True)
        out = _run(ctx, b, spin, diagnostics=True)
    finally:
        ctx.close()
    ext = out.extension
    nn = int(library.load().xtbloom_amd_packed_matrix_elements(
        interface._descriptors(library.load(), b.atom_offsets, b.atomic_numbers, b.positions,
                               b.molecular_charges, b.unpaired_electrons, None, None, None, None,
                               library.MODEL_GFN2_XTB, interface._DEFAULT_FLAGS, 250, 1e-6, 1e-8,
                               300.0, 8, 0.4, library.SCC_START_FRESH)[0]))
    assert ext["overlap"].size == nn and np.isfinite(ext["overlap"]).all()
    for i in range(b.batch_size):
        s = b.system(i)
        ref = O.singlepoint(s["Z"], s["xyz"], uhf=s["uhf"], opts=O.options(spin_channels=int(spin[i])),
                            trace=True)
        assert np.abs(ext["energy_components"][i] - ref.trace["components"]).max() < 1e-8, i
        a0, a1 = b.atom_offsets[i], b.atom_offsets[i + 1]
        assert np.abs(ext["coordination_numbers"][a0:a1] - ref.trace["coordination"]).max() < 1e-12


def test_radical_with_point_charges_field_and_operator():
    rng = np.random.default_rng(8)
    s = _radical(5, rng)
    s.update(workloads.point_charges(s["xyz"], rng, 6, 4.0, 7.0, gamma=0.4))
    nat = len(s["Z"])
    bvec = rng.normal(scale=0.005, size=nat)
    A = rng.normal(scale=0.005, size=(nat, nat))
    A = 0.5 * (A + A.T)
    field = (0.002, -0.001, 0.0015)
    b = workloads.from_systems([s, workloads.water(rng)])
    n2 = 3
    shifts = np.concatenate([bvec, np.zeros(n2)])
    roff = np.array([0, nat * nat, nat * nat + n2 * n2])
    resp = np.concatenate([A.ravel(), np.zeros(n2 * n2)])
    ctx = interface.Context("hip", 0)
    try:
        print("case radical with point charges, field and b + A q", flush=True)
        out = interface.compute_arrays(
            ctx, b.atom_offsets, b.atomic_numbers, b.positions, b.molecular_charges, b.unpaired_electrons,
            point_offsets=b.point_offsets, point_positions=b.point_positions, point_values=b.point_charges,
            point_gammas=b.point_gammas, spin_channels=np.array([2, 1], np.int32),
            electric_field=[field, None], atomic_potential_shifts=shifts, charge_response_offsets=roff,
            charge_response_matrix=resp,
            flags=interface._DEFAULT_FLAGS | library.COMPUTE_POINT_CHARGE_FORCES)
    finally:
        ctx.close()
    ref = O.singlepoint(s["Z"], s["xyz"], uhf=1, pc_xyz=s["pc_xyz"], pc_q=s["pc_q"], pc_gamma=s["pc_gamma"],
                        opts=O.options(spin_channels=2, efield=field, shifts=bvec, response=A))
    assert out.per_system_status[0] == 0 and ref.status == 0
    a1 = b.atom_offsets[1]
    print(f"combined vs oracle: dE {abs(out.energies[0] - ref.energy):.2e}", flush=True)
    assert abs(out.energies[0] - ref.energy) < 1e-7
    assert np.abs(out.forces[:a1] + ref.gradient).max() < 1e-6
    assert np.abs(out.charges[:a1] - ref.charges).max() < 1e-6
    p1 = b.point_offsets[1]
    assert np.abs(out.point_charge_forces[:p1] + ref.pc_gradient).max() < 1e-6
    assert out.result_flags & library.RESULT_FORCES_EXCLUDE_EXTERNAL_OPERATOR_DERIVATIVES
