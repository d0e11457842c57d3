"""bsl_lanczos_noalloc (the Lanczos loop 1 entry point that allocates no
device memory, its per-run state in a static per-GPU array) against
bsl_lanczos on the same contexts: the coefficients must be bitwise equal.
Two calls in a row share the array, the second on a smaller operator with a
different potential kernel and fewer steps, so anything stale from the first
call would show."""
import numpy as np
import pytest

import _cases as cs


def _noalloc(c, niter, start="random", seed=0, v0=None):
    import bsl_amd as B
    from bsl_amd._lib import START, check, lib
    a, b, b2 = np.zeros(niter), np.zeros(niter + 1), np.zeros(niter + 1)
    ptr = None if v0 is None else np.ascontiguousarray(v0, dtype=np.float64).ctypes.data
    check(lib().bsl_lanczos_noalloc(c.h, niter, ptr, START[start], seed, a, b, b2))
    assert isinstance(c, B.Calculator)
    return a, b, b2


@pytest.mark.gpu
def test_noalloc_equals_lanczos_two_calls():
    import bsl_amd as B
    # call 1: lmax 5 full basis (the four-wave potential kernel), 300 steps
    _, k1 = cs.cage_case(7, 5, 12, 14, "All")
    # call 2: a smaller operator: lmax 3 reduced basis, FP32, 120 steps
    _, k2 = cs.cage_case(5, 3, 8, 9, "Odd")
    k2 = dict(k2, basis="reduced", precision="fp32")
    with B.Calculator(**k1) as c1, B.Calculator(**k2) as c2:
        got1 = _noalloc(c1, 300, seed=3)
        got2 = _noalloc(c2, 120, seed=5)
        ref1 = c1.lanczos(300, start="random", seed=3)
        ref2 = c2.lanczos(120, start="random", seed=5)
        for got, ref in ((got1, ref1), (got2, ref2)):
            for g, r in zip(got, ref):
                assert np.array_equal(g, r)
        assert np.all(got1[1][1:] > 0) and np.all(got2[1][1:] > 0)
        # uniform start and a caller's start vector on the smaller context
        ua, ub, _ = _noalloc(c2, 50, start="uniform")
        ra, rb, _ = c2.lanczos(50, start="uniform")
        assert np.array_equal(ua, ra) and np.array_equal(ub, rb)
        v0 = np.random.default_rng(2).uniform(-1, 1, c2.N)
        v0 /= np.linalg.norm(v0)
        va, vb, _ = _noalloc(c2, 40, v0=v0)
        wa, wb, _ = c2.lanczos(40, v0=v0)
        assert np.array_equal(va, wa) and np.array_equal(vb, wb)


@pytest.mark.gpu
def test_noalloc_rejects_runs_beyond_its_scratch():
    import bsl_amd as B
    from bsl_amd._lib import BslError
    _, kw = cs.cage_case(4, 2, 6, 6, "All")
    with B.Calculator(**kw) as c:
        with pytest.raises(BslError) as e:
            _noalloc(c, 32769)
        assert e.value.code == 1 and "32768" in str(e.value)
        # the context still runs afterwards
        a, _, _ = _noalloc(c, 10)
        assert np.array_equal(a, c.lanczos(10, start="random")[0])
