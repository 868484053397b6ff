"""svmi_expval_static -- the expectation entry point that allocates no device
memory -- against svmi_expval (the same kernels, so the same result) and the
oracle: two calls in a row on one handle, the second with fewer terms, then
a smaller handle; sizes beyond the static scratch are rejected."""
import ctypes

import numpy as np
import pytest

import svmi
from oracle import oracle as O
from svmi import workloads as W

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _expval_static(e, terms):
    t = svmi.as_terms(terms)
    out = ctypes.c_double(0.0)
    st = svmi.lib().svmi_expval_static(e._h, svmi._ptr(t), len(t), ctypes.byref(out))
    return st, out.value


def _terms(n, count, rng):
    return [(rng.standard_normal(), int(rng.integers(0, 1 << n)), int(rng.integers(0, 1 << n))) for _ in range(count)]


def test_expval_static_matches_expval():
    rng = np.random.default_rng(4242)
    n = 18
    st = W.random_state(n, seed=18)
    with svmi.Engine(n) as e:
        e.set_state(*st)
        for count in (300, 7):      # second call smaller
            terms = _terms(n, count, rng)
            want = e.expval(terms)
            rc, got = _expval_static(e, terms)
            assert rc == 0
            assert abs(got - want) <= 1e-12 * max(1.0, abs(want))
            assert abs(got - O.expval(n, *st, terms)) < TOL
    n = 11
    st = W.random_state(n, seed=11)
    with svmi.Engine(n) as e:
        e.set_state(*st)
        terms = _terms(n, 5, rng)
        rc, got = _expval_static(e, terms)
        assert rc == 0
        assert abs(got - e.expval(terms)) <= 1e-12 * max(1.0, abs(got))
        assert abs(got - O.expval(n, *st, terms)) < TOL


def test_expval_static_rejects_sizes_beyond_its_scratch():
    n = 6
    with svmi.Engine(n) as e:
        terms = [(1.0, 0, 1)] * 16385
        rc, _ = _expval_static(e, terms)
        assert rc == 1 and b"16384" in svmi.lib().svmi_last_error()
        rc, v = _expval_static(e, [(1.0, 0, 0)] * 16384)
        assert rc == 0 and abs(v - 16384.0) < 1e-9
