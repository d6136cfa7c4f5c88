"""Systems above 1024 AOs on the GPU.

Systems of at most 1024 AOs keep their per-system work vectors in LDS.  Larger
ones, up to 8192 AOs, run instantiations of the same kernels with these
vectors in a global slice per system:
- the conditioning estimate;
- the blocked Householder reduction, whose symmetric product and panel dots
  run over row chunks of 16 slots of 64 rows;
- the eigensolver;
- the back-transformation;
- the spin fold;
- the three-body force kernel.
The reference CPU path has no size limit (LAPACK dsyevd per system,
eigensolver.cpp:1221-1254).

Checks:
- a ragged batch against the oracle: C190H382 (1142 AOs), an unrestricted
  C190H381 radical (1141 AOs), and members on the blocked (362 AOs) and
  register (122 AOs) paths;
- C350H702 (2102 AOs, 33 slots: three row chunks) against the committed
  oracle fixture;
- central differences of the GPU's own energies at 1142 AOs;
- the plain and graph-captured paths agree, and reproducible mode repeats
  bit for bit;
- the overlap conditioning gate (status 8 exactly where the oracle's dpocon
  gate fails) after the blocked factorization, at 364 and 1144 AOs;
- two concurrent lanes and a fixed-topology plan with a WARM start give the
  single-lane results;
- large eigenvalue clusters (near-dependent bases) in an unsplit and in a
  split tridiagonal matrix, against the oracle.
"""
import os

import numpy as np
import pytest

from oracle import oracle as O
from xtbloom_amd import interface, library, workloads

pytestmark = pytest.mark.gpu

E_TOL, F_TOL, Q_TOL = 1e-7, 1e-6, 1e-6
TIGHT = dict(charge_tolerance=1e-10, energy_tolerance=1e-12, max_scc_iterations=500)
FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "large_c350.npz")


def _radical(k, rng):
    s = workloads.alkane(k, rng)
    return dict(Z=s["Z"][:-1], xyz=s["xyz"][:-1], charge=0.0, uhf=1)


def _members():
    rng = np.random.default_rng(190)
    members = [workloads.water(rng), workloads.alkane(190, rng), _radical(190, rng),
               workloads.alkane(60, rng, substituent="NH2"), workloads.alkane(20, rng)]
    spin = np.array([1, 1, 2, 1, 1], np.int32)
    return members, spin


def _run(ctx, b, spin=None, **kw):
    return interface.compute_arrays(ctx, b.atom_offsets, b.atomic_numbers, b.positions,
                                    b.molecular_charges, b.unpaired_electrons, spin_channels=spin, **kw)


def test_ragged_batch_against_oracle():
    members, spin = _members()
    b = workloads.from_systems(members)
    naos = [O.basis_counts(m["Z"])[1] for m in members]
    assert naos[1] == 1142 and naos[2] == 1141 and 256 < naos[3] <= 1024 and naos[4] <= 256
    ctx = interface.Context("hip", 0)
    try:
        print("case large ragged batch, AOs", naos, flush=True)
        out = _run(ctx, b, spin)
    finally:
        ctx.close()
    ref = O.batch(b.atom_offsets, b.atomic_numbers, b.positions, b.molecular_charges,
                  b.unpaired_electrons, spin_channels=spin, threads=4)
    assert np.all(out.per_system_status == 0) and np.all(ref["status"] == 0)
    dE = np.abs(out.energies - ref["energies"])
    dF = np.abs(out.forces + ref["gradient"]).max()
    dq = np.abs(out.charges - ref["charges"]).max()
    print(f"large batch vs oracle: dE {dE.max():.2e} dF {dF:.2e} dq {dq:.2e} iterations "
          f"{out.scc_iterations.tolist()} oracle {ref['iterations'].tolist()}", flush=True)
    assert dE.max() <= E_TOL and dF <= F_TOL and dq <= Q_TOL
    assert np.abs(out.scc_iterations - ref["iterations"]).max() <= 1


def test_2102_aos_against_fixture():
    fx = np.load(FIXTURE)
    s = dict(Z=fx["Z"], xyz=fx["xyz"], charge=0.0, uhf=0)
    assert O.basis_counts(s["Z"])[1] == 2102
    b = workloads.from_systems([s, workloads.water(np.random.default_rng(1))])
    ctx = interface.Context("hip", 0)
    try:
        print("case C350H702 (2102 AOs)", flush=True)
        out = _run(ctx, b)
    finally:
        ctx.close()
    a1 = b.atom_offsets[1]
    assert out.per_system_status.tolist() == [0, 0]
    dE = abs(out.energies[0] - float(fx["energy"]))
    dF = np.abs(out.forces[:a1] + fx["gradient"]).max()
    dq = np.abs(out.charges[:a1] - fx["charges"]).max()
    print(f"2102 AOs vs oracle fixture: dE {dE:.2e} dF {dF:.2e} dq {dq:.2e} iterations "
          f"{int(out.scc_iterations[0])} oracle {int(fx['iterations'])}", flush=True)
    assert dE <= E_TOL and dF <= F_TOL and dq <= Q_TOL
    assert abs(int(out.scc_iterations[0]) - int(fx["iterations"])) <= 1
    assert np.abs(out.forces[:a1].sum(axis=0)).max() < 1e-8  # translation invariance


def test_forces_central_difference():
    s = workloads.alkane(190, np.random.default_rng(3))
    ctx = interface.Context("hip", 0)
    try:
        print("case 1142 AOs: forces vs central differences", flush=True)
        base = _run(ctx, workloads.from_systems([s]), **TIGHT)
        assert base.per_system_status[0] == 0
        h = 1e-4
        for atom, axis in ((0, 0), (200, 1), (len(s["Z"]) - 1, 2)):
            e = []
            for sign in (1, -1):
                xyz = s["xyz"].copy()
                xyz[atom, axis] += sign * h
                e.append(_run(ctx, workloads.from_systems([dict(s, xyz=xyz)]), **TIGHT).energies[0])
            fd = -(e[0] - e[1]) / (2 * h)
            print(f"atom {atom} axis {axis}: analytic {base.forces[atom, axis]:.9f} fd {fd:.9f}", flush=True)
            assert abs(fd - base.forces[atom, axis]) < 2e-7
    finally:
        ctx.close()


def test_paths_and_reproducible_mode(monkeypatch):
    members, spin = _members()
    b = workloads.from_systems(members[1:3] + members[:1])
    spin = np.array([1, 2, 1], np.int32)
    ctx = interface.Context("hip", 0)
    try:
        print("case large members: plain and batched paths, reproducible mode", flush=True)
        monkeypatch.setenv("XTBLOOM_AMD_PATH", "plain")
        plain = _run(ctx, b, spin)
        monkeypatch.setenv("XTBLOOM_AMD_PATH", "batched")
        batched = _run(ctx, b, spin)
        r = library.DETERMINISM_REPRODUCIBLE
        first = _run(ctx, b, spin, determinism=r)
        second = _run(ctx, b, spin, determinism=r)
    finally:
        ctx.close()
    assert np.all(plain.per_system_status == 0)
    assert np.array_equal(batched.per_system_status, plain.per_system_status)
    assert np.abs(batched.energies - plain.energies).max() <= E_TOL
    assert np.abs(batched.forces - plain.forces).max() <= F_TOL
    for name in ("energies", "forces", "charges", "scc_iterations"):
        assert np.array_equal(getattr(first, name), getattr(second, name)), name
    assert np.abs(first.energies - plain.energies).max() <= 1e-9
    assert np.abs(first.forces - plain.forces).max() <= 1e-9


def _near_dependent(k, d, rng):
    """CkH(2k+2) with an extra H at distance d from one H (and one more H for parity)."""
    base = workloads.alkane(k, rng)
    Z0, X0 = base["Z"], base["xyz"]
    h = np.where(Z0 == 1)[0]
    xyz = np.vstack([X0, X0[h[0]] + [d, 0.0, 0.0], X0[h[3]] + [0.0, 0.0, 0.5]])
    return dict(Z=np.concatenate([Z0, [1, 1]]), xyz=xyz, charge=0.0, uhf=0)


def test_overlap_gate_on_large_systems():
    rng = np.random.default_rng(11)
    members = [workloads.water(rng), _near_dependent(60, 3e-6, rng), _near_dependent(190, 3e-6, rng),
               _near_dependent(60, 1e-3, rng)]
    assert [O.basis_counts(m["Z"])[1] for m in members[1:]] == [364, 1144, 364]
    b = workloads.from_systems(members)
    iters = 40
    ctx = interface.Context("hip", 0)
    try:
        print("case near-dependent overlaps at 364 and 1144 AOs", flush=True)
        out = _run(ctx, b, max_scc_iterations=iters)
    finally:
        ctx.close()
    ref = O.batch(b.atom_offsets, b.atomic_numbers, b.positions, b.molecular_charges,
                  b.unpaired_electrons, opts=O.options(max_iterations=iters), threads=4)
    print("statuses", out.per_system_status.tolist(), "oracle", ref["status"].tolist(), flush=True)
    failed = library.STATUS_EIGENSOLVER_FAILED
    assert ref["status"][1] == failed and ref["status"][2] == failed and ref["status"][3] != failed
    assert [st == failed for st in out.per_system_status] == [st == failed for st in ref["status"]]
    for i in (1, 2):
        a0, a1 = b.atom_offsets[i], b.atom_offsets[i + 1]
        assert np.isnan(out.energies[i]) and np.all(np.isnan(out.forces[a0:a1]))
    assert out.per_system_status[0] == 0 and abs(out.energies[0] - ref["energies"][0]) <= 1e-10
    if ref["status"][3] == 0 and out.per_system_status[3] == 0:
        assert abs(out.energies[3] - ref["energies"][3]) <= E_TOL


def test_lanes_plans_and_warm_start(monkeypatch):
    rng = np.random.default_rng(21)
    # 64 members: two lanes take 32 or more each
    members = [workloads.alkane(190, rng), workloads.alkane(60, rng)]
    members += [workloads.water(rng) if i % 2 else workloads.alkane(2 + i % 9, rng) for i in range(62)]
    b = workloads.from_systems(members)
    ctx = interface.Context("hip", 0)
    try:
        print("case large members: one lane", flush=True)
        ref = _run(ctx, b)
        assert np.all(ref.per_system_status == 0)
        monkeypatch.setenv("XTBLOOM_AMD_LANES", "2")
        print("case large members: two lanes", flush=True)
        two = _run(ctx, b)
        monkeypatch.setenv("XTBLOOM_AMD_LANES", "1")
        for name in ("energies", "charges", "scc_iterations", "per_system_status"):
            assert np.array_equal(getattr(two, name), getattr(ref, name)), name
        assert np.abs(two.forces - ref.forces).max() < 1e-11
        print("case large members: plan and WARM start", flush=True)
        plan = interface.Plan(ctx, b.atom_offsets, b.atomic_numbers, b.positions, b.molecular_charges,
                              b.unpaired_electrons)
        try:
            got = plan.compute(b.positions)
            assert np.array_equal(got.energies, ref.energies)
            warm = plan.compute(b.positions, scc_start_mode=library.SCC_START_WARM)
            assert np.all(warm.per_system_status == 0) and warm.scc_iterations.max() <= 3
            assert np.abs(warm.energies - ref.energies).max() < 1e-7
        finally:
            plan.close()
    finally:
        ctx.close()


def test_large_clusters_in_a_split_tridiagonal():
    # two copies of a near-dependent C60H122 (+2 H) 120 bohr apart: S and H are
    # block diagonal, T splits between the fragments, and each block holds a
    # cluster of about 360 eigenvalues (block-local eigenvalue order)
    rng = np.random.default_rng(11)
    frag = _near_dependent(60, 1e-3, rng)
    both = dict(Z=np.concatenate([frag["Z"], frag["Z"]]),
                xyz=np.vstack([frag["xyz"], frag["xyz"] + [120.0, 0.0, 0.0]]), charge=0.0, uhf=0)
    b = workloads.from_systems([both, workloads.water(rng)])
    iters = 40
    ctx = interface.Context("hip", 0)
    try:
        print("case two near-dependent fragments in one system (728 AOs)", flush=True)
        out = _run(ctx, b, max_scc_iterations=iters)
    finally:
        ctx.close()
    ref = O.batch(b.atom_offsets, b.atomic_numbers, b.positions, b.molecular_charges,
                  b.unpaired_electrons, opts=O.options(max_iterations=iters), threads=4)
    print("statuses", out.per_system_status.tolist(), "oracle", ref["status"].tolist(),
          "iterations", out.scc_iterations.tolist(), "oracle", ref["iterations"].tolist(), flush=True)
    assert ref["status"].tolist() == [0, 0] and out.per_system_status.tolist() == [0, 0]
    assert np.abs(out.energies - ref["energies"]).max() <= E_TOL
    a1 = b.atom_offsets[1]
    assert np.abs(out.charges[:a1] - ref["charges"][:a1]).max() <= 1e-5
