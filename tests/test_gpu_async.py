"""Asynchronous submission and device result arenas on the GPU.

xtbloom_compute_enqueue (reference include/xtbloom/xtbloom.h:892-925) admits
a call, leaves its request PENDING and returns; the context's worker runs
submissions in order.  Checks: several submissions in flight at once give the
same results as the synchronous calls (energies bit for bit); a synchronous call after
pending submissions waits for them; a pending request cannot be resubmitted;
host inputs may be overwritten once enqueue has returned; destroying a
pending request waits for it.  Then a ROCm device result arena (reference
xtbloom.h:1011-1145) receives a call's outputs and hands them to torch
through DLPack without copying.
"""
import ctypes

import numpy as np
import pytest
import torch

from xtbloom_amd import dlpack, interface, library, workloads

pytestmark = pytest.mark.gpu


def run(ctx, b, positions=None, **kw):
    return interface.compute_arrays(ctx, b.atom_offsets, b.atomic_numbers,
                                    b.positions if positions is None else positions,
                                    b.molecular_charges, b.unpaired_electrons, **kw)


def same(a, b):
    # forces and charges carry the last-bit spread of atomic accumulation
    for name in ("energies", "scc_iterations", "per_system_status"):
        assert np.array_equal(getattr(a, name), getattr(b, name)), name
    assert np.abs(a.forces - b.forces).max() < 1e-11
    assert np.abs(a.charges - b.charges).max() < 1e-11


def test_submissions_in_flight_match_synchronous_calls():
    batches = [workloads.ragged_alkanes(batch=b, kmin=2, kmax=30, seed=s)
               for b, s in ((48, 1), (16, 2), (32, 3))]
    ctx = interface.Context("hip", 0)
    reqs = [interface.Request(ctx) for _ in batches]
    try:
        print("case async: three submissions in flight", flush=True)
        refs = [run(ctx, b) for b in batches]
        pend = []
        for b, r in zip(batches, reqs):
            pos = b.positions.copy()
            pend.append(run(ctx, b, positions=pos, request=r, wait=False))
            pos[:] = np.nan  # host inputs are staged at admission
        states = [r.query()["state"] for r in reqs]
        assert all(s in (library.REQUEST_PENDING, library.REQUEST_COMPLETE) for s in states)
        # a pending request takes no second submission (it may have completed
        # in the meantime, and then the submission is accepted)
        try:
            extra = run(ctx, batches[-1], request=reqs[-1], wait=False)
        except interface.XTBloomError as err:
            assert err.status == library.STATUS_INVALID_ARGUMENT
            extra = None
        # a synchronous call waits for the submissions queued before it
        again = run(ctx, batches[0])
        assert all(r.query()["state"] == library.REQUEST_COMPLETE for r in reqs)
        same(again, refs[0])
        for p, ref, r in zip(pend, refs, reqs):
            same(p.result(), ref)
            info = r.wait()
            assert info["state"] == library.REQUEST_COMPLETE and info["completion_status"] == 0
            assert r.error() == ""
        if extra is not None:
            same(extra.result(), refs[-1])
        # reuse of a completed request; destroying it while pending waits for it
        p = run(ctx, batches[1], request=reqs[1], wait=False)
        reqs[1].close()
        same(p._computed, refs[1])
    finally:
        for r in reqs:
            r.close()
        ctx.close()


def test_failed_submission_reports_through_the_request():
    b = workloads.ragged_alkanes(batch=6, kmin=3, kmax=12, seed=8)
    ctx = interface.Context("hip", 0)
    req = interface.Request(ctx)
    try:
        print("case async: iteration limit through a request", flush=True)
        p = run(ctx, b, request=req, wait=False, max_scc_iterations=5)
        got = p.result()  # per-system failures are data, the call succeeds
        assert np.all(got.per_system_status == library.STATUS_SCC_NOT_CONVERGED)
        bad = b.positions.copy()
        bad[0, 0] = np.nan
        with pytest.raises(interface.XTBloomError) as err:
            run(ctx, b, positions=bad, request=req, wait=False)
        assert err.value.status == library.STATUS_INVALID_ARGUMENT
        assert req.query()["state"] == library.REQUEST_COMPLETE  # unchanged by the rejection
    finally:
        req.close()
        ctx.close()


def test_device_arena_exports_to_torch():
    b = workloads.ragged_alkanes(batch=24, kmin=2, kmax=30, seed=5)
    B, N = b.batch_size, len(b.atomic_numbers)
    ctx = interface.Context("hip", 0)
    try:
        print("case device result arena and DLPack export", flush=True)
        ref = run(ctx, b)
        owner = dlpack.ResultOwner(8 * (B + 4 * N) + 4 * 3 * B + B, device_id=0)
        assert owner.memory_space == library.MEMORY_ROCM_DEVICE
        lib = library.load()
        batch, opts, keep, _ = interface._descriptors(
            lib, b.atom_offsets, b.atomic_numbers, b.positions, b.molecular_charges,
            b.unpaired_electrons, None, None, None, None, library.MODEL_GFN2_XTB,
            interface._DEFAULT_FLAGS, 250, 1e-6, 1e-8, 300.0, 8, 0.4, library.SCC_START_FRESH)
        off = dict(energies=0, forces=8 * B, charges=8 * (B + 3 * N), iters=8 * (B + 4 * N))
        off["status"] = off["iters"] + 4 * B
        off["conv"] = off["status"] + 4 * B
        res = library.BatchResult()
        assert lib.xtbloom_batch_result_init(ctypes.byref(res), ctypes.sizeof(res)) == 0
        res.energies = owner.buffer(off["energies"], 8 * B)
        res.forces = owner.buffer(off["forces"], 24 * N)
        res.atomic_charges = owner.buffer(off["charges"], 8 * N)
        res.scc_iterations = owner.buffer(off["iters"], 4 * B)
        res.per_system_status = owner.buffer(off["status"], 4 * B)
        res.scc_converged = owner.buffer(off["conv"], B)
        st = lib.xtbloom_compute(ctx.handle, ctypes.byref(batch), ctypes.byref(opts), ctypes.byref(res))
        assert st == 0, library.last_error()
        e = torch.from_dlpack(owner.tensor(off["energies"], np.float64, (B,)))
        f = torch.from_dlpack(owner.tensor(off["forces"], np.float64, (N, 3)))
        q = torch.from_dlpack(owner.tensor(off["charges"], np.float64, (N,)))
        it = torch.from_dlpack(owner.tensor(off["iters"], np.int32, (B,)))
        assert e.is_cuda and e.data_ptr() == owner.data  # no copy
        owner.close()  # the imported tensors keep the arena alive
        assert np.array_equal(e.cpu().numpy(), ref.energies)
        assert np.abs(f.cpu().numpy() - ref.forces).max() < 1e-11
        assert np.abs(q.cpu().numpy() - ref.charges).max() < 1e-11
        assert np.array_equal(it.cpu().numpy(), ref.scc_iterations)
        del e, f, q, it
    finally:
        ctx.close()
