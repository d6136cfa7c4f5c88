"""Strict WARM starts and fixed-topology plans on the MI355X path.

WARM semantics follow the reference ABI (xtbloom.h:502-523, 939-1003): a WARM
call starts from the converged state of the latest fully converged call with
the same topology and compute policy on the same context (or plan), never
falls back to FRESH, and is rejected with INVALID_ARGUMENT, outputs untouched,
when there is none.  Any accepted call consumes the checkpoint; a fully
converged call leaves a new one.  Parity: the HIP WARM call against the
oracle's WARM call from the oracle's own checkpoint, to the same bars as
FRESH (E 1e-7, F 1e-6, q 1e-6, iterations +-1).
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests.test_gpu_parity import E_TOL, F_TOL, Q_TOL
from xtbloom_amd import interface, library, workloads

pytestmark = pytest.mark.gpu

WARM = library.SCC_START_WARM


@pytest.fixture()
def ctx():
    c = interface.Context("hip", 0)
    yield c
    c.close()


def small_batch(seed=0):
    rng = np.random.default_rng(seed)
    systems = [workloads.water(rng), workloads.alkane(4, rng, substituent="OH"),
               workloads.alkane(7, rng), workloads.alkane(3, rng, substituent="NH2")]
    return workloads.from_systems(systems)


def moved(batch, scale=0.02, seed=1):
    return batch.positions + scale * np.random.default_rng(seed).standard_normal(batch.positions.shape)


def run(ctx, b, positions=None, **kw):
    return interface.compute_arrays(ctx, b.atom_offsets, b.atomic_numbers,
                                    b.positions if positions is None else positions,
                                    b.molecular_charges, b.unpaired_electrons, **kw)


def test_warm_without_checkpoint_is_rejected(ctx):
    import torch

    b = small_batch()
    inp = dict(atom_offsets=torch.from_numpy(b.atom_offsets),
               numbers=torch.from_numpy(b.atomic_numbers),
               positions=torch.from_numpy(b.positions),
               charges=torch.from_numpy(b.molecular_charges),
               uhf=torch.from_numpy(b.unpaired_electrons))
    print("case warm-first-call", flush=True)
    # caller-owned host outputs, filled with a sentinel
    out = dict(energies=torch.full((b.batch_size,), 123.0, dtype=torch.float64),
               forces=torch.full((len(b.atomic_numbers), 3), 123.0, dtype=torch.float64),
               charges=torch.full((len(b.atomic_numbers),), 123.0, dtype=torch.float64),
               scc_iterations=torch.full((b.batch_size,), 77, dtype=torch.int32),
               scc_converged=torch.full((b.batch_size,), 5, dtype=torch.uint8),
               per_system_status=torch.full((b.batch_size,), 99, dtype=torch.int32))
    with pytest.raises(interface.XTBloomError) as err:
        interface.compute_tensors(ctx, **inp, outputs=out, scc_start_mode=WARM)
    assert err.value.status == library.STATUS_INVALID_ARGUMENT
    assert "WARM" in str(err.value)
    for name, t in out.items():
        assert bool((t == {"scc_iterations": 77, "scc_converged": 5,
                           "per_system_status": 99}.get(name, 123.0)).all()), name
    # after a converged FRESH call the same request is accepted
    interface.compute_tensors(ctx, **inp, outputs=out)
    res, _ = interface.compute_tensors(ctx, **inp, outputs=out, scc_start_mode=WARM)
    assert bool((res["per_system_status"] == 0).all())


def test_warm_matches_oracle_warm(ctx):
    b = small_batch()
    pos2 = moved(b)
    print("case warm-parity fresh", flush=True)
    fresh = run(ctx, b)
    assert np.all(fresh.per_system_status == 0)
    print("case warm-parity warm", flush=True)
    warm = run(ctx, b, positions=pos2, scc_start_mode=WARM)
    fresh2 = run(ctx, b, positions=pos2)
    worst = dict(E=0.0, F=0.0, q=0.0, it=0)
    for i in range(b.batch_size):
        s = b.system(i)
        a0, a1 = b.atom_offsets[i], b.atom_offsets[i + 1]
        ref1 = O.singlepoint(s["Z"], s["xyz"])
        ref2 = O.singlepoint(s["Z"], pos2[a0:a1], warm_state=ref1.state)
        assert ref2.status == 0 and warm.per_system_status[i] == 0
        worst["E"] = max(worst["E"], abs(warm.energies[i] - ref2.energy))
        worst["F"] = max(worst["F"], np.abs(warm.forces[a0:a1] + ref2.gradient).max())
        worst["q"] = max(worst["q"], np.abs(warm.charges[a0:a1] - ref2.charges).max())
        worst["it"] = max(worst["it"], abs(int(warm.scc_iterations[i]) - ref2.iterations))
        # a WARM start does less SCC work than FRESH at the same geometry
        assert warm.scc_iterations[i] < fresh2.scc_iterations[i]
    print("warm-parity", worst, flush=True)
    assert worst["E"] <= E_TOL and worst["F"] <= F_TOL and worst["q"] <= Q_TOL
    assert worst["it"] <= 1


def test_same_geometry_warm_takes_two_iterations(ctx):
    b = small_batch(2)
    print("case warm-same-geometry", flush=True)
    fresh = run(ctx, b)
    warm = run(ctx, b, scc_start_mode=WARM)
    assert np.all(warm.per_system_status == 0)
    assert np.all(warm.scc_iterations <= 2)
    assert np.abs(warm.energies - fresh.energies).max() <= 1e-8
    # chained WARM calls keep working: every converged call leaves a checkpoint
    again = run(ctx, b, positions=moved(b, 0.01, 3), scc_start_mode=WARM)
    assert np.all(again.per_system_status == 0)


def test_warm_identity_is_strict(ctx):
    b = small_batch(4)
    print("case warm-identity", flush=True)
    run(ctx, b)
    # a different compute policy is a different identity
    with pytest.raises(interface.XTBloomError) as err:
        run(ctx, b, scc_start_mode=WARM, charge_tolerance=1e-7)
    assert err.value.status == library.STATUS_INVALID_ARGUMENT
    # a rejected request leaves the checkpoint in place
    ok = run(ctx, b, scc_start_mode=WARM)
    assert np.all(ok.per_system_status == 0)
    # a different molecular charge is a different identity
    charged = workloads.from_systems([dict(b.system(i), charge=(1.0 if i == 0 else 0.0),
                                           uhf=(1 if i == 0 else 0))
                                      for i in range(b.batch_size)])
    with pytest.raises(interface.XTBloomError):
        run(ctx, charged, scc_start_mode=WARM)
    # a call that does not fully converge consumes the checkpoint and leaves none
    capped = run(ctx, b, positions=moved(b, 0.05, 6), max_scc_iterations=2)
    assert np.any(capped.per_system_status != 0)
    with pytest.raises(interface.XTBloomError):
        run(ctx, b, scc_start_mode=WARM, max_scc_iterations=2)
    with pytest.raises(interface.XTBloomError):
        run(ctx, b, scc_start_mode=WARM)


def test_fresh_then_warm_replays_exactly(ctx):
    b = small_batch(5)
    pos2 = moved(b, 0.004, 7)
    print("case warm-replay", flush=True)
    seq = []
    for _ in range(2):
        run(ctx, b)
        seq.append(run(ctx, b, positions=pos2, scc_start_mode=WARM))
    assert np.array_equal(seq[0].energies, seq[1].energies)
    assert np.array_equal(seq[0].scc_iterations, seq[1].scc_iterations)
    assert np.array_equal(seq[0].charges, seq[1].charges)


def test_plan_matches_compute_and_rejects_other_batches(ctx):
    b = small_batch(8)
    pos2 = moved(b, 0.01, 9)
    print("case plan", flush=True)
    plan = interface.Plan(ctx, b.atom_offsets, b.atomic_numbers, b.positions,
                          b.molecular_charges, b.unpaired_electrons)
    try:
        ws = plan.query_workspace()
        assert ws["device_bytes"] > 0 and ws["host_bytes"] > 0
        p1 = plan.compute(b.positions)
        c1 = run(ctx, b)
        assert np.array_equal(p1.energies, c1.energies)
        assert np.array_equal(p1.charges, c1.charges)
        assert np.array_equal(p1.scc_iterations, c1.scc_iterations)
        assert np.abs(p1.forces - c1.forces).max() < 1e-11
        # the plan keeps its own checkpoint: WARM through the plan equals WARM
        # through the context from the same converged state
        p2 = plan.compute(pos2, scc_start_mode=WARM)
        c2 = run(ctx, b, positions=pos2, scc_start_mode=WARM)
        assert np.array_equal(p2.energies, c2.energies)
        assert np.array_equal(p2.scc_iterations, c2.scc_iterations)
        # topology or policy other than the plan's fails before any output
        other = b.subset([0, 1, 2])
        with pytest.raises(interface.XTBloomError) as err:
            interface.compute_arrays(ctx, other.atom_offsets, other.atomic_numbers,
                                     other.positions, plan=plan)
        assert err.value.status == library.STATUS_INVALID_ARGUMENT
        with pytest.raises(interface.XTBloomError):
            interface.compute_arrays(ctx, b.atom_offsets, b.atomic_numbers, b.positions,
                                     b.molecular_charges, b.unpaired_electrons, plan=plan,
                                     energy_tolerance=1e-9)
        plan._flags ^= library.COMPUTE_ATOMIC_CHARGES
        with pytest.raises(interface.XTBloomError):
            plan.query_workspace()
    finally:
        plan.close()


def test_plan_warm_across_lanes(ctx, monkeypatch):
    # 64 systems in two lanes: the topology caches and the WARM state are cut
    # per slice
    monkeypatch.setenv("XTBLOOM_AMD_LANES", "2")
    b = workloads.make("ragged512", seed=12, batch=64, kmax=8)
    pos2 = moved(b, 0.01, 13)
    print("case plan-lanes", flush=True)
    plan = interface.Plan(ctx, b.atom_offsets, b.atomic_numbers, b.positions,
                          b.molecular_charges, b.unpaired_electrons)
    try:
        plan.compute(b.positions)
        p2 = plan.compute(pos2, scc_start_mode=WARM)
    finally:
        plan.close()
    monkeypatch.setenv("XTBLOOM_AMD_LANES", "1")
    run(ctx, b)
    c2 = run(ctx, b, positions=pos2, scc_start_mode=WARM)
    assert np.all(p2.per_system_status == 0)
    assert np.array_equal(p2.energies, c2.energies)
    assert np.array_equal(p2.scc_iterations, c2.scc_iterations)


def test_calculator_warm_start_for_dynamics():
    # the reference calculators' opt-in automatic warm start: a trajectory of
    # small steps reuses each converged state; a changed setting falls back to
    # FRESH without an error
    s = workloads.alkane(6, np.random.default_rng(14), substituent="OH")
    print("case calculator-warm", flush=True)
    warm = interface.Calculator("GFN2-xTB", s["Z"], s["xyz"], warm_start=True)
    cold = interface.Calculator("GFN2-xTB", s["Z"], s["xyz"])
    try:
        rng = np.random.default_rng(15)
        xyz = s["xyz"].copy()
        iters = []
        for step in range(4):
            rw, rc = warm.singlepoint(), cold.singlepoint()
            assert abs(rw.get("energy") - rc.get("energy")) < 1e-7
            iters.append((int(rw.get("scc_iterations")), int(rc.get("scc_iterations"))))
            xyz = xyz + 0.005 * rng.standard_normal(xyz.shape)
            warm.update(positions=xyz)
            cold.update(positions=xyz)
        assert all(w < c for w, c in iters[1:]), iters
        warm.set("charge_tolerance", 1e-7)
        assert abs(warm.singlepoint().get("energy") - cold.singlepoint().get("energy")) < 1e-7
    finally:
        warm.close()
        cold.close()


def test_context_topology_cache_follows_the_batch(ctx):
    # one context keeps the host topology of its previous call: alternating
    # topologies and moving geometry must give what a new context gives
    a, b = small_batch(16), workloads.make("ragged512", seed=17, batch=6, kmax=9)
    pos_a2 = moved(a, 0.02, 18)
    print("case topology-cache", flush=True)
    seq = [run(ctx, a), run(ctx, b), run(ctx, a), run(ctx, a, positions=pos_a2), run(ctx, b)]
    other = interface.Context("hip", 0)
    try:
        ref_a, ref_b, ref_a2 = run(other, a), run(other, b), run(other, a, positions=pos_a2)
    finally:
        other.close()
    for got, ref in zip(seq, (ref_a, ref_b, ref_a, ref_a2, ref_b)):
        assert np.array_equal(got.energies, ref.energies)
        assert np.array_equal(got.scc_iterations, ref.scc_iterations)
        assert np.abs(got.forces - ref.forces).max() < 1e-11


def test_enqueue_matches_compute(ctx):
    # xtbloom_compute_enqueue / xtbloom_plan_compute_enqueue (reference
    # xtbloom.h:836-1003): submitted, then waited for
    b = small_batch(20)
    print("case enqueue", flush=True)
    req = interface.Request(ctx)
    try:
        assert req.query() == dict(state=library.REQUEST_IDLE, completion_status=0, result_flags=0)
        assert req.error() == ""
        ref = run(ctx, b)
        got = run(ctx, b, request=req)
        st = req.wait()
        assert st["state"] == library.REQUEST_COMPLETE and st["completion_status"] == 0
        assert req.error() == ""
        assert np.array_equal(got.energies, ref.energies)
        assert np.array_equal(got.scc_iterations, ref.scc_iterations)
        assert np.abs(got.forces - ref.forces).max() < 1e-11
        # an admission error raises directly and leaves the request as it was
        bad = b.positions.copy()
        bad[0, 0] = np.nan
        with pytest.raises(interface.XTBloomError) as err:
            run(ctx, b, positions=bad, request=req)
        assert err.value.status == library.STATUS_INVALID_ARGUMENT
        assert req.query()["state"] == library.REQUEST_COMPLETE
        # strict WARM through the request, then through a plan's enqueue
        warm = run(ctx, b, positions=moved(b, 0.01, 21), scc_start_mode=WARM, request=req)
        assert np.all(warm.per_system_status == 0)
        plan = interface.Plan(ctx, b.atom_offsets, b.atomic_numbers, b.positions,
                              b.molecular_charges, b.unpaired_electrons)
        try:
            pg = interface.compute_arrays(ctx, b.atom_offsets, b.atomic_numbers, b.positions,
                                          b.molecular_charges, b.unpaired_electrons, plan=plan,
                                          request=req)
            assert np.array_equal(pg.energies, ref.energies)
        finally:
            plan.close()
        # a request of another context is rejected at admission
        other = interface.Context("hip", 0)
        try:
            foreign = interface.Request(other)
            with pytest.raises(interface.XTBloomError):
                run(ctx, b, request=foreign)
            foreign.close()
        finally:
            other.close()
    finally:
        req.close()
