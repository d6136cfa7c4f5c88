"""Uniform electric field attachments on the GPU (ABI-v3 interaction suffix).

The field E (Hartree per elementary charge per bohr) adds -E.R_A to the
charge potential of atom A and -E to its dipole potential; the energy gains
-sum_A (q_A E.R_A + E.mu_A) and the force on A gains q_A E (reference
xtbloom.h:160-222, gfn2_cpu_execution.cpp:1294-1302, 1422-1430, 1588-1600).
Checks: the two field goldens of the reference (the conformance water case
and tblite's water at (0.001, 0.002, -0.0015)), a ragged batch whose members
carry different fields, none, or an explicit zero field, against the oracle;
forces against central differences of the GPU energies; descriptors and
payload in device memory; the descriptor validation and the reserved tags;
strict WARM identity (the field is part of it) and a plan taking new fields.
"""
import ctypes
import json
import os

import numpy as np
import pytest
import torch

from oracle import oracle as O
from tests.conftest import GOLDEN
from xtbloom_amd import interface, library, workloads

pytestmark = pytest.mark.gpu

TIGHT = dict(charge_tolerance=1e-10, energy_tolerance=1e-12, max_scc_iterations=500)
E_TOL, F_TOL, Q_TOL = 1e-7, 1e-6, 1e-6

# reference python/tests/test_interface.py:459-488
TBLITE_WATER_XYZ = np.array([[0.0, 0.0, -0.2358784530], [0.0, 1.4270063049, 1.0081495306],
                             [0.0, -1.4270063049, 1.0081495306]]) * 1.8897261246257702


@pytest.fixture(scope="module")
def ctx():
    c = interface.Context("hip", 0)
    yield c
    c.close()


def run(ctx, b, field=None, **kw):
    return interface.compute_arrays(ctx, b.atom_offsets, b.atomic_numbers, b.positions,
                                    b.molecular_charges, b.unpaired_electrons,
                                    electric_field=field, **kw)


def oracle_fields(B, field):
    E = np.zeros((B, 3))
    for i, f in enumerate(field):
        if f is not None:
            E[i] = f
    return E


def test_field_goldens(ctx):
    with open(os.path.join(GOLDEN, "conformance.json")) as fh:
        conf = json.load(fh)
    case = next(c for c in conf["cases"] if "efield" in c)
    systems = [dict(Z=np.array(case["Z"]), xyz=np.array(case["xyz"], dtype=float).reshape(-1, 3)),
               dict(Z=np.array([8, 1, 1]), xyz=TBLITE_WATER_XYZ)]
    fields = [case["efield"], (0.001, 0.002, -0.0015)]
    b = workloads.from_systems(systems)
    print("case field goldens", case["id"], "tblite water", flush=True)
    out = run(ctx, b, fields, flags=interface._DEFAULT_FLAGS | library.COMPUTE_DIPOLE_MOMENTS, **TIGHT)
    assert out.per_system_status.tolist() == [0, 0]
    assert abs(out.energies[0] - case["energy"]) <= conf["tolerances"]["energy"]
    assert abs(out.energies[1] - (-4.7652477392228)) <= 1e-7
    assert np.isfinite(out.dipole_moments).all()
    ref = O.batch(b.atom_offsets, b.atomic_numbers, b.positions, b.molecular_charges,
                  b.unpaired_electrons, opts=O.options(charge_tolerance=1e-10, energy_tolerance=1e-12,
                                                       max_iterations=500),
                  efield=np.array(fields))
    assert np.abs(out.energies - ref["energies"]).max() <= E_TOL
    assert np.abs(out.forces + ref["gradient"]).max() <= F_TOL
    assert np.abs(out.charges - ref["charges"]).max() <= Q_TOL


def test_mixed_batch_against_oracle(ctx):
    rng = np.random.default_rng(21)
    b = workloads.make("ragged512", seed=4, batch=10, kmax=12)
    fields = [tuple(rng.normal(scale=0.004, size=3)) if i % 3 != 2 else None for i in range(b.batch_size)]
    fields[4] = (0.0, 0.0, 0.0)  # an explicit zero field is an attachment that changes nothing
    print("case mixed-field batch", [f is not None for f in fields], flush=True)
    out = run(ctx, b, fields)
    plain = run(ctx, b)
    ref = O.batch(b.atom_offsets, b.atomic_numbers, b.positions, b.molecular_charges,
                  b.unpaired_electrons, efield=oracle_fields(b.batch_size, fields), threads=4)
    assert np.all(out.per_system_status == 0) and np.all(ref["status"] == 0)
    dE = np.abs(out.energies - ref["energies"]).max()
    dF = np.abs(out.forces + ref["gradient"]).max()
    dq = np.abs(out.charges - ref["charges"]).max()
    print(f"field batch vs oracle: dE {dE:.2e} dF {dF:.2e} dq {dq:.2e}", flush=True)
    assert dE <= E_TOL and dF <= F_TOL and dq <= Q_TOL
    assert np.abs(out.scc_iterations - ref["iterations"]).max() <= 1
    # members without a field, or with a zero one, equal the field-free call
    for i, f in enumerate(fields):
        a0, a1 = b.atom_offsets[i], b.atom_offsets[i + 1]
        if f is None or not np.any(f):
            assert out.energies[i] == plain.energies[i]
            assert np.abs(out.forces[a0:a1] - plain.forces[a0:a1]).max() < 1e-11
        else:
            assert abs(out.energies[i] - plain.energies[i]) > 1e-7


def test_forces_central_difference(ctx):
    s = workloads.water(np.random.default_rng(4))
    field = [(0.003, -0.004, 0.005)]
    b = workloads.from_systems([s])
    print("case field forces vs central differences", flush=True)
    out = run(ctx, b, field, **TIGHT)
    h = 1e-4
    for atom, axis in ((0, 0), (1, 1), (2, 2), (0, 2)):
        e = []
        for sign in (1, -1):
            xyz = s["xyz"].copy()
            xyz[atom, axis] += sign * h
            e.append(run(ctx, workloads.from_systems([dict(s, xyz=xyz)]), field, **TIGHT).energies[0])
        assert abs(-(e[0] - e[1]) / (2 * h) - out.forces[atom, axis]) < 2e-7
    # a neutral molecule feels no net force
    assert np.abs(out.forces.sum(axis=0)).max() < 1e-9


def _raw_call(ctx, b, desc, payload, count, space=library.MEMORY_HOST):
    lib = library.load()
    batch, opts, keep, (B, N, _) = interface._descriptors(
        lib, b.atom_offsets, b.atomic_numbers, b.positions, b.molecular_charges,
        b.unpaired_electrons, None, None, None, None, library.MODEL_GFN2_XTB,
        interface._DEFAULT_FLAGS, 250, 1e-6, 1e-8, 300.0, 8, 0.4, library.SCC_START_FRESH)
    batch.total_interactions = count
    if desc is not None:
        batch.interaction_descriptors = library.ConstBuffer(desc[0], desc[1], space, 0)
    if payload is not None:
        batch.interaction_payload = library.ConstBuffer(payload[0], payload[1], space, 0)
    e = np.full(B, np.nan)
    f = np.full((N, 3), np.nan)
    q = np.full(N, np.nan)
    it = np.zeros(B, np.int32)
    conv = np.zeros(B, np.uint8)
    st = np.zeros(B, np.int32)
    res = library.BatchResult()
    assert lib.xtbloom_batch_result_init(ctypes.byref(res), ctypes.sizeof(res)) == 0
    res.energies, res.forces, res.atomic_charges = (interface._out(x) for x in (e, f, q))
    res.scc_iterations, res.scc_converged, res.per_system_status = (interface._out(x) for x in (it, conv, st))
    status = lib.xtbloom_compute(ctx.handle, ctypes.byref(batch), ctypes.byref(opts), ctypes.byref(res))
    return status, e, f


def _blocks(items, fields):
    """Descriptor array and payload: items are (type, system, offset, size)."""
    desc = (library.Interaction * max(1, len(items)))()
    for k, (typ, sysi, off, size) in enumerate(items):
        desc[k].type, desc[k].flags, desc[k].system_index = typ, 0, sysi
        desc[k].payload_offset, desc[k].payload_size = off, size
    payload = np.zeros(4 * max(1, len(fields)) + 1, np.float64)
    head = payload.view(np.int32)
    for k, f in enumerate(fields):
        head[8 * k] = 1
        payload[4 * k + 1:4 * k + 4] = f
    return desc, payload


def test_device_resident_attachments_and_validation(ctx):
    b = workloads.water_batch(batch=3, seed=6)
    F = [(0.002, 0.0, -0.001), (0.0, 0.003, 0.001)]
    EF = library.INTERACTION_ELECTRIC_FIELD
    desc, payload = _blocks([(EF, 0, 0, 32), (EF, 2, 32, 32)], F)
    print("case field attachments: host, device, validation", flush=True)
    st, e_host, f_host = _raw_call(ctx, b, (ctypes.addressof(desc), ctypes.sizeof(desc)),
                                   (payload.ctypes.data, payload.nbytes), 2)
    assert st == 0, library.last_error()
    ref = run(ctx, b, [F[0], None, F[1]])
    assert np.array_equal(e_host, ref.energies)
    # the same descriptors and payload in device memory
    d_desc = torch.frombuffer(bytearray(bytes(desc)), dtype=torch.uint8).cuda()
    d_pay = torch.from_numpy(payload.view(np.uint8).copy()).cuda()
    st, e_dev, f_dev = _raw_call(ctx, b, (d_desc.data_ptr(), d_desc.numel()), (d_pay.data_ptr(), d_pay.numel()), 2,
                                 space=library.MEMORY_ROCM_DEVICE)
    assert st == 0, library.last_error()
    assert np.array_equal(e_dev, ref.energies)
    assert np.abs(f_dev - ref.forces).max() < 1e-11

    def status(items, fields=F, count=None, mutate=None, pay_bytes=None):
        d, p = _blocks(items, fields)
        if mutate:
            mutate(d, p)
        nbytes = p.nbytes if pay_bytes is None else pay_bytes
        return _raw_call(ctx, b, (ctypes.addressof(d), ctypes.sizeof(d)), (p.ctypes.data, nbytes),
                         len(items) if count is None else count)[0]

    INV, NI = library.STATUS_INVALID_ARGUMENT, library.STATUS_NOT_IMPLEMENTED
    cases = {
        "negative count": (status([(EF, 0, 0, 32)], count=-1), INV),
        "flags set": (status([(EF, 0, 0, 32)], mutate=lambda d, p: setattr(d[0], "flags", 1)), INV),
        "NONE tag": (status([(library.INTERACTION_NONE, 0, 0, 32)]), INV),
        "unknown tag": (status([(0x0999, 0, 0, 32)]), INV),
        "system outside batch": (status([(EF, 3, 0, 32)]), INV),
        "negative system": (status([(EF, -1, 0, 32)]), INV),
        "block past payload": (status([(EF, 0, 16, 32)], pay_bytes=40), INV),
        "undersized block": (status([(EF, 0, 0, 24)]), INV),
        "oversized block": (status([(EF, 0, 0, 40)]), INV),
        "misaligned block": (status([(EF, 0, 4, 32)]), INV),
        "block_version 2": (status([(EF, 0, 0, 32)], mutate=lambda d, p: p.view(np.int32).__setitem__(0, 2)), INV),
        "reserved word": (status([(EF, 0, 0, 32)], mutate=lambda d, p: p.view(np.int32).__setitem__(1, 7)), INV),
        "NaN field": (status([(EF, 0, 0, 32)], fields=[(np.nan, 0.0, 0.0)]), INV),
        "duplicate": (status([(EF, 1, 0, 32), (EF, 1, 32, 32)]), INV),
        "reserved tag": (status([(library.INTERACTION_ALPB_SOLVATION, 0, 0, 8)]), NI),
        "reserved tag beside a field": (status([(EF, 0, 0, 32), (library.INTERACTION_ALPB_SOLVATION, 1, 32, 4)]), NI),
        "reserved tag without header": (status([(library.INTERACTION_ALPB_SOLVATION, 0, 0, 2)]), INV),
        "no payload": (_raw_call(ctx, b, (ctypes.addressof(desc), ctypes.sizeof(desc)), None, 2)[0], INV),
        "no descriptors": (_raw_call(ctx, b, None, (payload.ctypes.data, payload.nbytes), 2)[0], INV),
    }
    bad = {k: v for k, v in cases.items() if v[0] != v[1]}
    assert not bad, bad
    # zero attachments with a descriptor array present: the plain call
    st, e0, _ = _raw_call(ctx, b, (ctypes.addressof(desc), ctypes.sizeof(desc)), None, 0)
    assert st == 0 and np.array_equal(e0, run(ctx, b).energies)


def test_warm_identity_and_plans(ctx):
    b = workloads.water_batch(batch=4, seed=9)
    F = [(0.002, -0.001, 0.0)] * 4
    G = [(0.002, -0.001, 1e-4)] * 4
    print("case field WARM identity and plan", flush=True)
    cold = run(ctx, b, F)
    assert np.all(cold.per_system_status == 0)
    warm = run(ctx, b, F, scc_start_mode=library.SCC_START_WARM)
    assert np.all(warm.per_system_status == 0)
    assert np.abs(warm.energies - cold.energies).max() < 1e-7
    for other in (G, None):
        with pytest.raises(interface.XTBloomError) as err:
            run(ctx, b, other, scc_start_mode=library.SCC_START_WARM)
        assert err.value.status == library.STATUS_INVALID_ARGUMENT
        run(ctx, b, F)  # a fresh converged call restores the checkpoint
    # a plan made without a field takes fields on its FRESH calls
    plan = interface.Plan(ctx, b.atom_offsets, b.atomic_numbers, b.positions, b.molecular_charges,
                          b.unpaired_electrons)
    try:
        got = plan.compute(b.positions, electric_field=G)
        want = run(ctx, b, G)
        assert np.array_equal(got.energies, want.energies)
        again = plan.compute(b.positions, electric_field=G, scc_start_mode=library.SCC_START_WARM)
        assert np.abs(again.energies - want.energies).max() < 1e-7
        with pytest.raises(interface.XTBloomError):
            plan.compute(b.positions, electric_field=F, scc_start_mode=library.SCC_START_WARM)
    finally:
        plan.close()


def test_calculator_field():
    calc = interface.Calculator("GFN2-xTB", [8, 1, 1], TBLITE_WATER_XYZ, efield=[0.001, 0.002, -0.0015],
                                **TIGHT)
    plain = interface.Calculator("GFN2-xTB", [8, 1, 1], TBLITE_WATER_XYZ, efield=[0.0, 0.0, 0.0], **TIGHT)
    try:
        print("case Calculator with efield", flush=True)
        r = calc.singlepoint()
        assert abs(r.energy - (-4.7652477392228)) <= 1e-7
        assert r.dipole_moments is not None and np.isfinite(r.dipole_moments).all()
        assert plain.efield is None  # a zero field normalizes to no attachment
        p = plain.singlepoint()
        assert p.dipole_moments is None
        assert abs(p.charges.sum()) < 1e-10 and abs(r.charges.sum()) < 1e-10
        assert abs(r.charges[0] - p.charges[0]) > 1e-6
    finally:
        calc.close()
        plain.close()
    with pytest.raises(interface.XTBloomValueError):
        interface.Structure([8, 1, 1], TBLITE_WATER_XYZ, efield=[0.001, 0.002])
    with pytest.raises(interface.XTBloomValueError):
        interface.Structure([8, 1, 1], TBLITE_WATER_XYZ, efield=[0.001, 0.002, float("nan")])
