"""The oracle's Fermi occupations against the reference's own occupation tests.

Cases and bars are those of reference tests/eigensolver_test.cpp
(test_occupations_degeneracy_and_fractional_filling :439-600,
test_binary64_frontier_retry_controls :577, test_occupation_representability_policy
:696-900): the long-double main path (fractional filling, degenerate levels,
translation invariance, tiny and near-capacity targets, extreme ranges) and
the binary64 rare path of occupation_binary64_policy.hpp (degenerate blocks
whose count cannot be split into equal binary64 occupations, subnormal ties,
frontier retries).  The reference's input-validation cases of
fill_occupations_cpu (descending levels, aliased buffers) belong to its C ABI
wrapper, not to the occupation arithmetic, and are not restated.
"""
import ctypes
import math

import numpy as np

from oracle import oracle as O

KT300 = 300.0 * 3.166808578545117e-6
EPS = np.finfo(float).eps
DENORM = 5e-324
MAX = np.finfo(float).max


def fermi(levels, nel, kt):
    e = np.ascontiguousarray(levels, dtype=float)
    occ = np.full(len(e), -1.0)
    mu, ent = ctypes.c_double(7.0), ctypes.c_double(7.0)
    ok = O.lib().oracle_fermi(len(e), O._p(e), float(nel), float(kt), O._p(occ), ctypes.byref(mu),
                              ctypes.byref(ent))
    return ok, occ, mu.value, ent.value


def published_entropy(occ):
    s = 0.0
    for o in occ:
        if 0.0 < o < 1.0:
            s -= o * math.log(o) + (1.0 - o) * math.log(1.0 - o)
    return s


def test_fractional_filling_and_degeneracy():
    ok, occ, mu, ent = fermi([-1.0, 0.0, 1.0], 1.5, 0.0)
    assert ok and list(occ) == [1.0, 0.5, 0.0] and mu == 0.0 and ent == 0.0
    ok, occ, mu, ent = fermi([0.0, 0.0], 1.0, KT300)
    assert ok and abs(occ[0] - 0.5) <= 2e-14 and abs(occ[1] - 0.5) <= 2e-14
    assert abs(mu) <= 2e-14 and abs(ent - 2.0 * math.log(2.0)) <= 2e-13
    ok, occ, mu, ent = fermi([0.0, 0.0, 0.0], 1.3, KT300)
    assert ok and occ[0] == occ[1] == occ[2] and abs(occ.sum() - 1.3) <= 2e-14
    assert abs(ent - published_entropy(occ)) <= 2e-15
    ok, occ, mu, ent = fermi([-1.0, 0.0, 1.0], 1.3, 0.02)
    assert ok and abs(occ.sum() - 1.3) <= 2e-13 and ent > 0.0 and math.isfinite(mu)


def test_extreme_ranges_and_empty_or_full():
    ok, occ, mu, ent = fermi([MAX / 2.0, MAX], 1.0, MAX / 256.0)
    assert ok and math.isfinite(mu) and math.isfinite(ent) and np.all(np.isfinite(occ))
    assert abs(occ.sum() - 1.0) <= 8e-15
    ok, occ, mu, ent = fermi([-3.0, 7.0], 0.0, 0.01)
    assert ok and list(occ) == [0.0, 0.0] and mu == 0.0 and ent == 0.0
    ok, occ, mu, ent = fermi([-3.0, 7.0], 2.0, 0.01)
    assert ok and list(occ) == [1.0, 1.0] and math.isfinite(mu) and ent == 0.0


def test_small_targets_and_translation_invariance():
    for target in (1e-16, 1e-20, 1e-300):
        ok, occ, mu, ent = fermi([-0.2, 0.3], target, 0.01)
        pop = occ.sum()
        assert ok and pop > 0.0 and abs(pop - target) <= 2e-13 * target
        assert math.isfinite(mu) and math.isfinite(ent)
        ok0, o0, mu0, e0 = fermi([0.0, 1.0], target, 1e-7)
        ok1, o1, mu1, e1 = fermi([100.0, 101.0], target, 1e-7)
        assert ok0 and ok1 and np.array_equal(o0, o1) and e0 == e1
        assert abs(mu1 - (mu0 + 100.0)) <= 2e-13
        assert abs(o1.sum() - target) <= 2e-13 * target
    near_full = np.nextafter(2.0, 0.0)
    ok0, o0, mu0, e0 = fermi([0.0, 1.0], near_full, 1e-7)
    ok1, o1, mu1, e1 = fermi([100.0, 101.0], near_full, 1e-7)
    assert ok0 and ok1 and np.array_equal(o0, o1) and e0 == e1
    assert abs(mu1 - (mu0 + 100.0)) <= 2e-13
    holes = (1.0 - o1[0]) + (1.0 - o1[1])
    assert abs(holes - (2.0 - near_full)) <= 2e-13 * (2.0 - near_full)
    ok, occ, mu, ent = fermi([-0.2, 0.3], near_full, 0.01)
    holes = (1.0 - occ[0]) + (1.0 - occ[1])
    assert ok and holes > 0.0 and abs(holes - (2.0 - near_full)) <= 2e-13 * (2.0 - near_full)


def test_binary64_policy_representability():
    one_hole = np.nextafter(1.0, 0.0)
    nel = np.nextafter(3.0, 0.0)
    ok, occ, mu, ent = fermi([1.0, 1.0, 1.0], nel, KT300)
    assert ok and occ[0] == occ[1] == occ[2] == one_hole
    assert abs(math.fsum(occ) - nel) <= 2.0 * EPS * 3.0
    assert math.isfinite(ent) and abs(ent - published_entropy(occ)) <= 1e-15
    # exact half-subnormal tie: the lower symmetric occupation
    ok, occ, mu, ent = fermi([1.0, 1.0], 9.0 * DENORM, KT300)
    assert ok and occ[0] == occ[1] == 4.0 * DENORM
    # electron-poor: the all-zero symmetric state
    ok, occ, mu, ent = fermi([1.0, 1.0, 1.0], DENORM, KT300)
    assert ok and occ[0] == occ[1] == occ[2]
    assert abs(occ.sum() - DENORM) <= 2.0 * EPS * 3.0 and ent == 0.0 and math.isfinite(mu)
    # the subnormal electron lands on the singleton above the frontier
    ok, occ, mu, ent = fermi([0.0, 0.0, 1.0], DENORM, 1e-7)
    assert ok and list(occ) == [0.0, 0.0, DENORM] and ent > 0.0 and math.isfinite(mu)
    # the canonical subnormal logit keeps the saturated chemical potential
    ok, occ, mu, ent = fermi([MAX, MAX, MAX], DENORM, MAX)
    assert ok and list(occ) == [0.0, 0.0, 0.0] and mu == -MAX and ent == 0.0
    # an exactly representable symmetric target stays exact
    nel2 = np.nextafter(2.0, 0.0)
    ok, occ, mu, ent = fermi([1.0, 1.0], nel2, KT300)
    assert ok and occ[0] == occ[1] and (occ[0] + occ[1]) == nel2
    # a singleton absorbs the one-hole ulp: strict candidates first
    nel4 = np.nextafter(4.0, 0.0)
    ok, occ, mu, ent = fermi([0.0, 1.0, 1.0, 1.0], nel4, KT300)
    assert ok and np.all(occ == one_hole)


def test_binary64_policy_frontier_retry():
    nel = np.nextafter(1.5, 0.0)
    ok, occ, mu, ent = fermi([0.0, 1.0, 1.0], nel, 1e-7)
    assert ok and occ[1] == occ[2]
    assert abs(math.fsum(occ) - nel) <= 64.0 * EPS * nel
    nel3 = np.nextafter(3.0, 0.0)
    ok, occ, mu, ent = fermi([0.0, 1.0, 1.0, 1.0, 1.0, 1.0], nel3, 1e-7)
    assert ok and np.all(occ[1:] == occ[1])
    # no multi-orbital block: the strict singleton correction
    ok, occ, mu, ent = fermi([0.0, 1.0, 2.0], nel3, KT300)
    holes = (1.0 - occ[0]) + (1.0 - occ[1]) + (1.0 - occ[2])
    assert ok and abs(holes - (3.0 - nel3)) <= 64.0 * EPS * (3.0 - nel3)
