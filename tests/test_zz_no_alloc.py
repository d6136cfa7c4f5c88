"""The no-allocation entry point (xtbloom_amd_compute_no_alloc) against xtbloom_compute.

Both calls run the same kernels on the same inputs; only the device workspace
differs (the library's fixed per-GPU array instead of the context's arena).
With XTBLOOM_DETERMINISM_REPRODUCIBLE every output must agree bit for bit;
with the default determinism the gradient sums use floating-point atomics, so
forces agree to rounding.  Two no-allocation calls in a row, the second on a
smaller batch, check that a call does not depend on what the previous call
left in the shared array.  A batch that needs more workspace than the array
holds is rejected before any output is written.
"""
import numpy as np
import pytest

from xtbloom_amd import interface, library, workloads

pytestmark = pytest.mark.gpu

FIELDS = ("energies", "forces", "charges", "scc_iterations", "scc_converged", "per_system_status")


def _call(ctx, b, **kw):
    return interface.compute_arrays(ctx, b.atom_offsets, b.atomic_numbers, b.positions,
                                    b.molecular_charges, b.unpaired_electrons,
                                    point_offsets=b.point_offsets,
                                    point_positions=b.point_positions,
                                    point_values=b.point_charges,
                                    point_gammas=b.point_gammas, **kw)


def _batches():
    rng = np.random.default_rng(11)
    first = workloads.ragged_alkanes(batch=8, kmin=2, kmax=20, seed=5, points=50)
    second = workloads.from_systems([
        workloads.water(rng), workloads.alkane(3, rng, substituent="NH2"), workloads.water(rng)])
    return first, second


def _same(a, b, exact):
    for name in FIELDS:
        x, y = getattr(a, name), getattr(b, name)
        if exact or name not in ("energies", "forces", "charges"):
            assert np.array_equal(x, y), name
        else:
            assert np.abs(x - y).max() <= 1e-12, name


@pytest.mark.parametrize("determinism", [library.DETERMINISM_REPRODUCIBLE, library.DETERMINISM_DEFAULT])
def test_no_alloc_matches_compute_twice(determinism):
    first, second = _batches()
    exact = determinism == library.DETERMINISM_REPRODUCIBLE
    ctx = interface.Context("hip", 0)
    try:
        for name, b in (("first", first), ("second (smaller)", second)):
            print(f"case no_alloc {name}: {b.batch_size} systems, determinism {determinism}",
                  flush=True)
            ref = _call(ctx, b, determinism=determinism)
            got = _call(ctx, b, determinism=determinism, no_alloc=True)
            assert np.all(got.per_system_status == 0)
            _same(got, ref, exact)
            # the context's arena keeps its resident topology across the
            # no-allocation call: a regular call afterwards still agrees
            again = _call(ctx, b, determinism=determinism)
            _same(again, ref, exact)
    finally:
        ctx.close()


def test_no_alloc_first_call_on_a_fresh_context():
    _, second = _batches()
    ctx = interface.Context("hip", 0)
    try:
        print("case no_alloc on a fresh context", flush=True)
        got = _call(ctx, second, determinism=library.DETERMINISM_REPRODUCIBLE, no_alloc=True)
        ref = _call(ctx, second, determinism=library.DETERMINISM_REPRODUCIBLE)
        _same(got, ref, True)
    finally:
        ctx.close()


def test_no_alloc_rejects_a_batch_beyond_the_workspace():
    big = workloads.make("ragged512", seed=0)
    ctx = interface.Context("hip", 0)
    try:
        print("case no_alloc beyond the workspace (512 systems)", flush=True)
        with pytest.raises(interface.XTBloomError) as err:
            _call(ctx, big, no_alloc=True)
        assert err.value.status == library.STATUS_INVALID_ARGUMENT
        assert "no-allocation" in str(err.value)
    finally:
        ctx.close()
