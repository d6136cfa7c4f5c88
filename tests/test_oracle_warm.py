"""The oracle's WARM start (state in/out of the SCC), pinned without a GPU.

A WARM call restores the converged multipole state of an earlier call and
runs the same SCC from it with a reset mixer (reference
gfn2_cpu_execution.cpp:1261-1268, 1313-1325).  The checks below fix that
behaviour by properties a wrong state path would break:

* the FRESH SAD state passed in as a WARM state reproduces FRESH bit for bit
  (the WARM path differs from FRESH only by its initial state);
* a same-geometry WARM start from a converged state reconverges in at most two
  iterations to the same energy and charges (reference
  tests/cpu_public_inference_test.cpp:1481-1492 asserts no more SCC work than
  FRESH and agreement within the tolerances);
* the converged state is a fixed point of one SCC step;
* a WARM start at a moved geometry converges to the FRESH solution there.
"""
import numpy as np
import pytest

from oracle import oracle as O
from xtbloom_amd import workloads


@pytest.fixture(scope="module")
def system():
    return workloads.alkane(5, np.random.default_rng(21), substituent="NH2")


def test_sad_state_as_warm_state_is_fresh(system):
    Z, xyz = system["Z"], system["xyz"]
    sad = O.singlepoint(Z, xyz, opts=O.options(max_iterations=0))
    assert sad.iterations == 0
    fresh = O.singlepoint(Z, xyz)
    warm = O.singlepoint(Z, xyz, warm_state=sad.state)
    assert fresh.status == warm.status == 0
    assert warm.energy == fresh.energy and warm.iterations == fresh.iterations
    assert np.array_equal(warm.gradient, fresh.gradient)
    assert np.array_equal(warm.state, fresh.state)


def test_same_geometry_warm_reconverges(system):
    Z, xyz = system["Z"], system["xyz"]
    opts = O.options()
    fresh = O.singlepoint(Z, xyz, opts=opts)
    warm = O.singlepoint(Z, xyz, opts=opts, warm_state=fresh.state)
    assert warm.status == 0
    assert warm.iterations <= 2 < fresh.iterations
    assert abs(warm.energy - fresh.energy) <= opts.energy_tolerance
    assert np.abs(warm.charges - fresh.charges).max() <= 1e-5
    assert np.abs(warm.gradient - fresh.gradient).max() <= 1e-4


def test_converged_state_is_a_fixed_point(system):
    Z, xyz = system["Z"], system["xyz"]
    tight = O.options(charge_tolerance=1e-10, energy_tolerance=1e-12)
    fresh = O.singlepoint(Z, xyz, opts=tight)
    warm = O.singlepoint(Z, xyz, opts=tight, warm_state=fresh.state)
    assert np.abs(warm.state - fresh.state).max() < 1e-8


def test_moved_geometry_warm_matches_fresh(system):
    Z, xyz = system["Z"], system["xyz"]
    first = O.singlepoint(Z, xyz)
    moved = xyz + 0.02 * np.random.default_rng(5).standard_normal(xyz.shape)
    fresh = O.singlepoint(Z, moved)
    warm = O.singlepoint(Z, moved, warm_state=first.state)
    assert warm.status == fresh.status == 0
    assert warm.iterations < fresh.iterations
    assert abs(warm.energy - fresh.energy) < 1e-7
    assert np.abs(warm.charges - fresh.charges).max() < 1e-5


def test_warm_state_must_fit_the_system(system):
    Z, xyz = system["Z"], system["xyz"]
    with pytest.raises(ValueError):
        O.singlepoint(Z, xyz, warm_state=np.zeros(3))
