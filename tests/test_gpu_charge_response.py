"""External charge operator (periodic QM/MM coupling) on the GPU.

v_A = b_A + sum_B A_AB q_B joins the charge-channel atom potential and the
energy gains q.b + q.A.q/2 (reference xtbloom.h:424-431,
periodic_embedding.hpp:55-60, scc_driver.cpp:2236-2262, 2515-2531); forces
leave out coordinate derivatives of b and A and the result says so
(XTBLOOM_RESULT_FORCES_EXCLUDE_EXTERNAL_OPERATOR_DERIVATIVES).  Checks: a
ragged batch with operators against the oracle, Hellmann-Feynman derivatives
in b and A and central differences of the forces from the GPU's own
energies, the validation rules, device-resident inputs, WARM identity and
plans, and the Calculator's ChargeResponse.
"""
import ctypes

import numpy as np
import pytest
import torch

from oracle import oracle as O
from xtbloom_amd import interface, library, workloads

pytestmark = pytest.mark.gpu

E_TOL, F_TOL, Q_TOL = 1e-7, 1e-6, 1e-6
TIGHT = dict(charge_tolerance=1e-10, energy_tolerance=1e-12, max_scc_iterations=500)
EXCLUDE = library.RESULT_FORCES_EXCLUDE_EXTERNAL_OPERATOR_DERIVATIVES


@pytest.fixture(scope="module")
def ctx():
    c = interface.Context("hip", 0)
    yield c
    c.close()


def operators(batch, seed, skip=()):
    """b and A per system (zero for the members in skip), packed."""
    rng = np.random.default_rng(seed)
    shifts, mats = [], []
    for i, n in enumerate(np.diff(batch.atom_offsets)):
        n = int(n)
        A = rng.normal(scale=0.01, size=(n, n))
        A = 0.5 * (A + A.T) + 0.05 * np.eye(n)
        b = rng.normal(scale=0.01, size=n)
        if i in skip:
            A, b = np.zeros((n, n)), np.zeros(n)
        shifts.append(b)
        mats.append(A.ravel())
    roff = np.zeros(batch.batch_size + 1, np.int64)
    roff[1:] = np.cumsum([m.size for m in mats])
    return np.concatenate(shifts), roff, np.concatenate(mats)


def run(ctx, b, op=None, positions=None, **kw):
    extra = {}
    if op is not None:
        extra = dict(atomic_potential_shifts=op[0], charge_response_offsets=op[1],
                     charge_response_matrix=op[2])
    return interface.compute_arrays(ctx, b.atom_offsets, b.atomic_numbers,
                                    b.positions if positions is None else positions,
                                    b.molecular_charges, b.unpaired_electrons, **extra, **kw)


def test_batch_against_oracle(ctx):
    b = workloads.make("ragged512", seed=6, batch=8, kmax=14)
    op = operators(b, 3, skip=(5,))
    print("case charge operator batch of 8", flush=True)
    out = run(ctx, b, op)
    assert out.result_flags & EXCLUDE
    assert not run(ctx, b).result_flags & EXCLUDE
    ref = O.batch(b.atom_offsets, b.atomic_numbers, b.positions, b.molecular_charges,
                  b.unpaired_electrons, shifts=op[0], response=op[2], threads=4)
    assert np.all(out.per_system_status == 0) and np.all(ref["status"] == 0)
    dE = np.abs(out.energies - ref["energies"]).max()
    dF = np.abs(out.forces + ref["gradient"]).max()
    dq = np.abs(out.charges - ref["charges"]).max()
    print(f"charge operator vs oracle: dE {dE:.2e} dF {dF:.2e} dq {dq:.2e}", flush=True)
    assert dE <= E_TOL and dF <= F_TOL and dq <= Q_TOL
    assert np.abs(out.scc_iterations - ref["iterations"]).max() <= 1
    # the member with zero b and A equals the operator-free call
    plain = run(ctx, b)
    assert out.energies[5] == plain.energies[5]


def test_hellmann_feynman_and_forces(ctx):
    s = workloads.water(np.random.default_rng(4))
    b = workloads.from_systems([s])
    shifts, roff, mat = operators(b, 1)
    print("case charge operator derivatives", flush=True)
    base = run(ctx, b, (shifts, roff, mat), **TIGHT)
    q = base.charges
    h = 1e-5

    def energy(sh=shifts, m=mat, positions=None):
        return run(ctx, b, (sh, roff, m), positions=positions, **TIGHT).energies[0]

    for a in range(3):
        p, m = shifts.copy(), shifts.copy()
        p[a] += h
        m[a] -= h
        assert abs((energy(sh=p) - energy(sh=m)) / (2 * h) - q[a]) < 1e-7
    for a, c in ((0, 1), (2, 2)):
        p, m = mat.copy().reshape(3, 3), mat.copy().reshape(3, 3)
        p[a, c] += h
        m[a, c] -= h
        if a != c:
            p[c, a] += h
            m[c, a] -= h
        want = q[a] * q[c] if a != c else 0.5 * q[a] ** 2
        assert abs((energy(m=p.ravel()) - energy(m=m.ravel())) / (2 * h) - want) < 1e-7
    hx = 1e-4
    for atom, axis in ((0, 0), (1, 1), (2, 2)):
        xp, xm = s["xyz"].copy(), s["xyz"].copy()
        xp[atom, axis] += hx
        xm[atom, axis] -= hx
        fd = -(energy(positions=xp) - energy(positions=xm)) / (2 * hx)
        assert abs(fd - base.forces[atom, axis]) < 2e-7


def _raw(ctx, b, mutate):
    lib = library.load()
    op = operators(b, 2)
    batch, opts, keep, (B, N, _) = interface._descriptors(
        lib, b.atom_offsets, b.atomic_numbers, b.positions, b.molecular_charges,
        b.unpaired_electrons, None, None, None, None, library.MODEL_GFN2_XTB,
        interface._DEFAULT_FLAGS, 250, 1e-6, 1e-8, 300.0, 8, 0.4, library.SCC_START_FRESH,
        atomic_potential_shifts=op[0], charge_response_offsets=op[1], charge_response_matrix=op[2])
    held = mutate(batch, keep[-1])
    e = np.full(B, np.nan)
    f = np.full((N, 3), np.nan)
    q = np.full(N, np.nan)
    it, conv, st = np.zeros(B, np.int32), np.zeros(B, np.uint8), np.zeros(B, np.int32)
    res = library.BatchResult()
    assert lib.xtbloom_batch_result_init(ctypes.byref(res), ctypes.sizeof(res)) == 0
    res.energies, res.forces, res.atomic_charges = (interface._out(x) for x in (e, f, q))
    res.scc_iterations, res.scc_converged, res.per_system_status = (interface._out(x) for x in (it, conv, st))
    status = lib.xtbloom_compute(ctx.handle, ctypes.byref(batch), ctypes.byref(opts), ctypes.byref(res))
    del held
    return status, e, f


def test_validation_and_device_inputs(ctx):
    b = workloads.water_batch(batch=3, seed=6)
    INV = library.STATUS_INVALID_ARGUMENT
    print("case charge operator validation", flush=True)

    def setter(name, value):
        def mutate(batch, op):
            setattr(batch, name, value)
        return mutate

    def poke(index, k, value):
        def mutate(batch, op):
            op[index][k] = value
        return mutate

    def empty(name):
        return setter(name, library.ConstBuffer())

    cases = {
        "negative count": setter("total_charge_response_elements", -1),
        "count mismatch": setter("total_charge_response_elements", 18),
        "no offsets": empty("charge_response_offsets"),
        "no matrix": empty("charge_response_matrix"),
        "nonzero first offset": poke(1, 0, 1),
        "not square blocks": poke(1, 1, 8),
        "asymmetric": poke(2, 1, 0.5),
        "NaN shift": poke(0, 2, np.nan),
        "inf matrix": poke(2, 0, np.inf),
    }
    got = {k: _raw(ctx, b, m)[0] for k, m in cases.items()}
    assert all(v == INV for v in got.values()), got
    st, e_host, f_host = _raw(ctx, b, lambda batch, op: None)
    assert st == 0
    # count zero with no offsets and no matrix: only b
    st, e_b, _ = _raw(ctx, b, lambda batch, op: (setter("total_charge_response_elements", 0)(batch, op),
                                                 empty("charge_response_offsets")(batch, op),
                                                 empty("charge_response_matrix")(batch, op)))
    assert st == 0 and not np.array_equal(e_b, e_host)

    # the same operator from device memory
    def to_device(batch, op):
        dev = [torch.from_numpy(np.ascontiguousarray(x)).cuda() for x in op]
        for name, t in zip(("atomic_potential_shifts", "charge_response_offsets", "charge_response_matrix"), dev):
            setattr(batch, name, library.ConstBuffer(t.data_ptr(), t.numel() * t.element_size(),
                                                     library.MEMORY_ROCM_DEVICE, 0))
        return dev
    st, e_dev, f_dev = _raw(ctx, b, to_device)
    assert st == 0 and np.array_equal(e_dev, e_host) and np.abs(f_dev - f_host).max() < 1e-11


def test_warm_identity_and_plans(ctx):
    b = workloads.water_batch(batch=4, seed=9)
    op = operators(b, 5)
    op2 = (op[0] * 1.01, op[1], op[2])
    print("case charge operator WARM identity and plans", flush=True)
    assert np.all(run(ctx, b, op).per_system_status == 0)
    # new b and A values keep the identity; the operator's presence is part of it
    warm = run(ctx, b, op2, scc_start_mode=library.SCC_START_WARM)
    assert np.all(warm.per_system_status == 0)
    assert np.abs(warm.energies - run(ctx, b, op2).energies).max() < 1e-7
    with pytest.raises(interface.XTBloomError) as err:
        run(ctx, b, scc_start_mode=library.SCC_START_WARM)
    assert err.value.status == library.STATUS_INVALID_ARGUMENT
    plan = interface.Plan(ctx, b.atom_offsets, b.atomic_numbers, b.positions, b.molecular_charges,
                          b.unpaired_electrons, atomic_potential_shifts=op[0], charge_response_offsets=op[1],
                          charge_response_matrix=op[2])
    bare = interface.Plan(ctx, b.atom_offsets, b.atomic_numbers, b.positions, b.molecular_charges,
                          b.unpaired_electrons)
    try:
        got = plan.compute(b.positions, atomic_potential_shifts=op2[0])
        assert np.array_equal(got.energies, run(ctx, b, op2).energies)
        assert got.result_flags & EXCLUDE
        with pytest.raises(interface.XTBloomError):
            interface.compute_arrays(ctx, b.atom_offsets, b.atomic_numbers, b.positions,
                                     b.molecular_charges, b.unpaired_electrons, plan=bare,
                                     atomic_potential_shifts=op[0], charge_response_offsets=op[1],
                                     charge_response_matrix=op[2])
    finally:
        plan.close()
        bare.close()


def test_calculator_charge_response():
    s = workloads.water(np.random.default_rng(12))
    rng = np.random.default_rng(0)
    A = rng.normal(scale=0.01, size=(3, 3))
    resp = interface.ChargeResponse(rng.normal(scale=0.01, size=3), 0.5 * (A + A.T))
    calc = interface.Calculator("GFN2-xTB", s["Z"], s["xyz"], charge_response=resp)
    ctx = interface.Context("hip", 0)
    try:
        print("case Calculator with ChargeResponse", flush=True)
        r = calc.singlepoint()
        b = workloads.from_systems([s])
        ref = run(ctx, b, (resp.shifts, np.array([0, 9]), resp.matrix.ravel()))
        assert r.energy == ref.energies[0]
    finally:
        calc.close()
        ctx.close()
    with pytest.raises(interface.XTBloomValueError):
        interface.ChargeResponse(np.zeros(3), np.zeros((3, 2)))
    with pytest.raises(interface.XTBloomValueError):
        interface.Structure(s["Z"], s["xyz"], charge_response=interface.ChargeResponse(np.zeros(2), np.zeros((2, 2))))
