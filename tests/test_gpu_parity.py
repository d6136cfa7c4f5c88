"""Parity of the MI355X path (through the C ABI) against the CPU oracle.

Bars: energies 1e-7 Eh, forces 1e-6 Eh/bohr, charges
1e-6 e, SCC iteration counts within +-1, integer layouts exact.  Every case
prints its name before the GPU step so that a failing step is identifiable.
"""
import json
import os

import numpy as np
import pytest

from oracle import oracle as O
from tests.conftest import GOLDEN
from xtbloom_amd import interface, library, workloads

pytestmark = pytest.mark.gpu

E_TOL, F_TOL, Q_TOL = 1e-7, 1e-6, 1e-6


@pytest.fixture(scope="module")
def ctx():
    c = interface.Context("hip", 0)
    yield c
    c.close()


def run(ctx, batch, **kw):
    pc = {}
    if batch.point_offsets is not None:
        pc = dict(point_offsets=batch.point_offsets, point_positions=batch.point_positions,
                  point_values=batch.point_charges, point_gammas=batch.point_gammas)
    flags = (library.COMPUTE_ENERGY | library.COMPUTE_FORCES | library.COMPUTE_ATOMIC_CHARGES
             | library.COMPUTE_POINT_CHARGE_FORCES)
    return interface.compute_arrays(ctx, batch.atom_offsets, batch.atomic_numbers,
                                    batch.positions, batch.molecular_charges,
                                    batch.unpaired_electrons, **pc, flags=flags, **kw)


def compare(batch, out, label, opts=None, iter_tol=1, fields=None, opt_kw=None):
    """fields: per-system uniform field (or None); such a system's oracle run
    uses O.options(**opt_kw, efield=field)."""
    worst = dict(E=0.0, F=0.0, q=0.0, PF=0.0, it=0)
    for i in range(batch.batch_size):
        s = batch.system(i)
        kw = {k: s[k] for k in ("pc_xyz", "pc_q", "pc_gamma") if k in s and len(s["pc_q"])}
        o = opts
        if fields is not None and fields[i] is not None:
            o = O.options(**(opt_kw or {}), efield=fields[i])
        ref = O.singlepoint(s["Z"], s["xyz"], charge=s["charge"], uhf=s["uhf"], opts=o, **kw)
        a0, a1 = batch.atom_offsets[i], batch.atom_offsets[i + 1]
        assert ref.status == 0, (label, i)
        assert out.per_system_status[i] == 0, (label, i, out.per_system_status[i])
        worst["E"] = max(worst["E"], abs(out.energies[i] - ref.energy))
        worst["F"] = max(worst["F"], np.abs(out.forces[a0:a1] + ref.gradient).max())
        worst["q"] = max(worst["q"], np.abs(out.charges[a0:a1] - ref.charges).max())
        worst["it"] = max(worst["it"], abs(int(out.scc_iterations[i]) - ref.iterations))
        if kw:
            p0, p1 = batch.point_offsets[i], batch.point_offsets[i + 1]
            worst["PF"] = max(worst["PF"],
                              np.abs(out.point_charge_forces[p0:p1] + ref.pc_gradient).max())
    print(label, {k: (f"{v:.2e}" if isinstance(v, float) else v) for k, v in worst.items()},
          flush=True)
    assert worst["E"] <= E_TOL and worst["F"] <= F_TOL and worst["q"] <= Q_TOL
    assert worst["PF"] <= F_TOL and worst["it"] <= iter_tol
    return worst


def test_water_batch(ctx):
    b = workloads.make("water8", seed=3)
    print("case water8", flush=True)
    compare(b, run(ctx, b), "water8")


def test_intermediates_against_oracle(ctx):
    b = workloads.from_systems([workloads.water(np.random.default_rng(4)),
                                workloads.alkane(3, np.random.default_rng(5), substituent="NH2")])
    print("case intermediates", flush=True)
    out = run(ctx, b, diagnostics=True)
    ext = out.extension
    off = 0
    for i in range(b.batch_size):
        s = b.system(i)
        ref = O.singlepoint(s["Z"], s["xyz"], trace=True)
        nao = ref.trace["overlap"].shape[0]
        sl = slice(off, off + nao * nao)
        off += nao * nao
        assert np.abs(ext["overlap"][sl] - ref.trace["overlap"].ravel()).max() < 1e-12
        assert np.abs(ext["core_hamiltonian"][sl] - ref.trace["h0"].ravel()).max() < 1e-12
        a0, a1 = b.atom_offsets[i], b.atom_offsets[i + 1]
        assert np.abs(ext["coordination_numbers"][a0:a1] - ref.trace["coordination"]).max() < 1e-12
        assert np.abs(ext["energy_components"][i] - ref.trace["components"]).max() < 1e-8
        assert np.abs(ext["density"][sl] - ref.trace["density"].ravel()).max() < 1e-6


def test_conformance_goldens(ctx):
    with open(os.path.join(GOLDEN, "conformance.json")) as fh:
        conf = json.load(fh)
    cases = [c for c in conf["cases"] if "skip" not in c]
    systems, names = [], []
    for c in cases:
        s = dict(Z=np.array(c["Z"]), xyz=np.array(c["xyz"], dtype=float).reshape(-1, 3),
                 charge=c["charge"], uhf=c["unpaired"])
        if "pc_q" in c:
            s.update(pc_xyz=np.array(c["pc_xyz"]).reshape(-1, 3), pc_q=np.array(c["pc_q"]),
                     pc_gamma=np.array(c["pc_gamma"]))
        systems.append(s)
        names.append(c["id"])
    b = workloads.from_systems(systems, names)
    fields = [c.get("efield") for c in cases]
    print("case conformance", names, flush=True)
    out = run(ctx, b, charge_tolerance=1e-8, energy_tolerance=1e-10,
              electric_field=fields if any(f is not None for f in fields) else None)
    tol = conf["tolerances"]
    for i, c in enumerate(cases):
        a0, a1 = b.atom_offsets[i], b.atom_offsets[i + 1]
        assert out.per_system_status[i] == 0, c["id"]
        assert abs(out.energies[i] - c["energy"]) <= tol["energy"], c["id"]
        # a case may pin only its energy (the field case: its forces are
        # checked against the oracle below and against central differences)
        if "forces_hartree_per_bohr" in c.get("oracle_properties", ["forces_hartree_per_bohr"]):
            assert np.abs(out.forces[a0:a1].ravel() - np.array(c["forces"])).max() <= tol["forces"], c["id"]
        if "charges" in c:
            assert np.abs(out.charges[a0:a1] - np.array(c["charges"])).max() <= tol["charges"]
        if "pc_forces" in c:
            p0, p1 = b.point_offsets[i], b.point_offsets[i + 1]
            assert np.abs(out.point_charge_forces[p0:p1].ravel()
                          - np.array(c["pc_forces"])).max() <= tol["point_charge_forces"], c["id"]
    compare(b, out, "conformance-vs-oracle", opts=O.options(charge_tolerance=1e-8, energy_tolerance=1e-10),
            fields=fields, opt_kw=dict(charge_tolerance=1e-8, energy_tolerance=1e-10))


def test_ragged_alkanes(ctx):
    full = workloads.make("ragged512", seed=0)
    n = np.diff(full.atom_offsets)
    # 16 members spanning the size range, both substituents included
    idx = np.argsort(n)[np.linspace(0, full.batch_size - 1, 16).astype(int)]
    b = full.subset(list(idx))
    print("case ragged16", [full.names[i] for i in idx], flush=True)
    compare(b, run(ctx, b), "ragged16")


def test_point_charge_embedding(ctx):
    b = workloads.make("embedded256", seed=1, batch=3)
    print("case embedded3", flush=True)
    out = run(ctx, b)
    compare(b, out, "embedded3")


def test_large_system(ctx):
    b = workloads.make("long64", seed=2, batch=1, kmin=50, kmax=50)
    print("case long1", b.names, flush=True)
    compare(b, run(ctx, b), "long1")


def test_system_beyond_lds_staging(ctx):
    # 272 atoms: pair sums from the global staging slice, eigensolver beyond
    # its register-resident sizes; a small molecule rides along in the batch.
    b = workloads.make("long64", seed=3, batch=1, kmin=90, kmax=90)
    w = workloads.make("water8", seed=4, batch=1)
    both = workloads.from_systems([b.system(0), w.system(0)], b.names + w.names)
    print("case long272+water", both.names, flush=True)
    compare(both, run(ctx, both), "long272+water")


def test_failure_isolation_not_converged(ctx):
    b = workloads.make("water8", seed=9, batch=3)
    print("case max_iter=2", flush=True)
    out = run(ctx, b, max_scc_iterations=2)
    assert np.all(out.per_system_status == library.STATUS_SCC_NOT_CONVERGED)
    assert np.all(out.scc_iterations == 2)
    assert np.all(~out.scc_converged)
    assert np.all(np.isnan(out.energies)) and np.all(np.isnan(out.forces))


def test_mixed_convergence_isolation(ctx):
    # H2O converges in fewer rounds than a substituted alkane; cap in between
    w = workloads.water(np.random.default_rng(1))
    a = workloads.alkane(10, np.random.default_rng(2), substituent="OH")
    b = workloads.from_systems([w, a])
    ref_w = O.singlepoint(w["Z"], w["xyz"])
    ref_a = O.singlepoint(a["Z"], a["xyz"])
    cap = min(ref_w.iterations, ref_a.iterations) + 1
    print("case mixed cap", cap, ref_w.iterations, ref_a.iterations, flush=True)
    out = run(ctx, b, max_scc_iterations=cap)
    for i, ref in enumerate((ref_w, ref_a)):
        a0, a1 = b.atom_offsets[i], b.atom_offsets[i + 1]
        if ref.iterations <= cap - 1:
            assert out.per_system_status[i] == 0
            assert abs(out.energies[i] - ref.energy) <= E_TOL
        elif ref.iterations > cap + 1:
            assert out.per_system_status[i] == library.STATUS_SCC_NOT_CONVERGED
            assert np.all(np.isnan(out.forces[a0:a1]))


def test_batch_equals_sequential(ctx):
    b = workloads.make("ragged512", seed=4, batch=5, kmax=12)
    print("case batch-vs-sequential", flush=True)
    out = run(ctx, b)
    for i in range(b.batch_size):
        one = b.subset([i])
        o1 = run(ctx, one)
        a0, a1 = b.atom_offsets[i], b.atom_offsets[i + 1]
        assert o1.scc_iterations[0] == out.scc_iterations[i]
        assert abs(o1.energies[0] - out.energies[i]) < 1e-11
        assert np.abs(o1.forces - out.forces[a0:a1]).max() < 1e-10


def test_lanes_match_single_lane(ctx, monkeypatch):
    # 96 systems: three concurrent lanes of 32 against one lane, with point
    # charges on every system so that the point-charge slices are cut too
    b = workloads.make("embedded256", seed=11, batch=96, kmin=2, kmax=10, points=4)
    runs = {}
    for lanes in ("1", "2", "3"):
        monkeypatch.setenv("XTBLOOM_AMD_LANES", lanes)
        print("case lanes", lanes, flush=True)
        runs[lanes] = run(ctx, b, diagnostics=True, kernel_stats=True)
    one = runs["1"]
    for lanes in ("2", "3"):
        out = runs[lanes]
        for name in ("energies", "charges", "scc_iterations", "per_system_status"):
            assert np.array_equal(getattr(out, name), getattr(one, name)), (lanes, name)
        # gradients are accumulated with floating-point atomics: equal up to
        # the order of the additions
        for name in ("forces", "point_charge_forces", "point_charge_potential_gradient"):
            assert np.abs(getattr(out, name) - getattr(one, name)).max() < 1e-11, (lanes, name)
        for name in ("overlap", "core_hamiltonian", "density", "coordination_numbers",
                     "energy_components"):
            assert np.array_equal(out.extension[name], one.extension[name]), (lanes, name)
        assert out.extension["scc_rounds"] == one.extension["scc_rounds"]
        st1, st = one.extension["kernel_stats"], out.extension["kernel_stats"]
        # every lane launches the per-call kernels once
        assert st["geometry"]["launches"] == int(lanes) * st1["geometry"]["launches"]
        assert abs(st["tridiag"]["flops"] - st1["tridiag"]["flops"]) <= 1e-9 * st1["tridiag"]["flops"]


def test_invalid_input_leaves_outputs(ctx):
    b = workloads.make("water8", seed=2, batch=2)
    bad = b.positions.copy()
    bad[1, 0] = np.nan
    print("case invalid", flush=True)
    with pytest.raises(interface.XTBloomError) as err:
        interface.compute_arrays(ctx, b.atom_offsets, b.atomic_numbers, bad)
    assert err.value.status == library.STATUS_INVALID_ARGUMENT
    # coincident atoms in the second system (reference coordination.cpp:236-246)
    close = b.positions.copy()
    close[5] = close[3] + np.array([0.0, 3e-7, 0.0])
    print("case coincident", flush=True)
    with pytest.raises(interface.XTBloomError) as err:
        interface.compute_arrays(ctx, b.atom_offsets, b.atomic_numbers, close)
    assert err.value.status == library.STATUS_INVALID_ARGUMENT
    assert "coincident" in str(err.value)


def test_point_charge_potential_gradient_fd(ctx):
    rng = np.random.default_rng(7)
    s = workloads.water(rng)
    s.update(workloads.point_charges(s["xyz"], rng, 4, 5.0, 7.0, gamma=0.5))
    b = workloads.from_systems([s])
    print("case dEdq", flush=True)
    out = run(ctx, b, extension=True, charge_tolerance=1e-10, energy_tolerance=1e-12)
    h = 1e-4
    for p in range(4):
        vals = []
        for sgn in (1, -1):
            q = s["pc_q"].copy()
            q[p] += sgn * h
            bb = workloads.from_systems([dict(s, pc_q=q)])
            vals.append(run(ctx, bb, charge_tolerance=1e-10, energy_tolerance=1e-12).energies[0])
        fd = (vals[0] - vals[1]) / (2 * h)
        assert abs(fd - out.point_charge_potential_gradient[p]) < 1e-6


def test_option_and_electron_count_validation(ctx):
    # reference validation.cpp:426-461 and wavefunction.cpp:606-636; every
    # refusal happens before any output is written
    b = workloads.make("water8", seed=5, batch=2)

    def expect_invalid(label, **kw):
        print("case invalid option", label, flush=True)
        with pytest.raises(interface.XTBloomError) as err:
            run(ctx, b, **kw)
        assert err.value.status == library.STATUS_INVALID_ARGUMENT, label

    def setter(field, value):
        return lambda o: setattr(o, field, value)

    expect_invalid("reserved", options_hook=setter("reserved", 1))
    expect_invalid("reserved_v2", options_hook=setter("reserved_v2", 1))
    expect_invalid("reserved_v3", options_hook=setter("reserved_v3", 1))
    expect_invalid("scc_start_mode 0", options_hook=setter("scc_start_mode", 0))
    expect_invalid("scc_start_mode 3", options_hook=setter("scc_start_mode", 3))
    expect_invalid("scc_mixer 0", options_hook=setter("scc_mixer", 0))
    expect_invalid("determinism 2", determinism=2)
    # charge +6 leaves two electrons, four unpaired need beta = -1
    print("case invalid electron counts", flush=True)
    with pytest.raises(interface.XTBloomError) as err:
        interface.compute_arrays(ctx, b.atom_offsets, b.atomic_numbers, b.positions,
                                 np.array([0.0, 6.0]), np.array([0, 4], np.int32))
    assert err.value.status == library.STATUS_INVALID_ARGUMENT
    assert "beta" in str(err.value)


def test_reproducible_mode_is_bit_exact(ctx):
    # XTBLOOM_DETERMINISM_REPRODUCIBLE: repeated calls agree bit for bit, forces
    # included (order-independent sums of the scattered gradient contributions)
    b = workloads.make("embedded256", seed=12, batch=24, kmin=2, kmax=20, points=8)
    outs = []
    for k in range(3):
        print("case reproducible call", k, flush=True)
        outs.append(run(ctx, b, determinism=library.DETERMINISM_REPRODUCIBLE))
    for o in outs[1:]:
        for name in ("energies", "forces", "charges", "point_charge_forces", "scc_iterations"):
            assert np.array_equal(getattr(o, name), getattr(outs[0], name)), name
    compare(b, outs[0], "reproducible-vs-oracle")
    default = run(ctx, b)
    assert np.abs(default.forces - outs[0].forces).max() < 1e-10
