"""The Cullum-Willoughby eigensolve with its Sturm bisections and inverse
iterations on the GPU (bsl_lancbis_device) against the host restatement of
BISEC/ISOEV/INVERR (bsl_lancbis, bitwise equal to the oracle's) on the same
alpha/beta: same good eigenvalues to the bisection tolerance, same
multiplicity/spurious decisions, same converged/unconverged error classes."""
import json
import os

import numpy as np
import pytest

import _cases as cs
import _oracle as O

pytestmark = pytest.mark.gpu


def _compare(a, b, b2, emin, emax):
    import bsl_amd as B
    eh, rh = B.lancbis(a, b, b2, emin, emax)
    ed, rd = B.lancbis(a, b, b2, emin, emax, device=True)
    assert len(eh) == len(ed)
    scale = max(1.0, np.max(np.abs(a)))
    assert np.max(np.abs(eh - ed)) <= 1e-11 * scale
    # error estimates: same sign/zero pattern, close where they are not tiny
    assert np.array_equal(rh == 0, rd == 0)
    big = np.abs(rh) > 1e-6
    assert np.all(np.sign(rh[big]) == np.sign(rd[big]))
    assert np.allclose(rd[big], rh[big], rtol=1e-3, atol=0)
    conv_h = (np.abs(rh) < 1e-10) & (rh != 0)
    assert np.all(np.abs(rd[conv_h]) < 1e-8)
    return eh, ed, rd


def test_saved_run_coefficients():
    """The oracle's alpha/beta of the reference's saved run (cage 11, 10^3 x
    lmax 3, 300 steps): the device eigensolve reproduces the saved levels."""
    import bsl_amd as B  # noqa: F401
    runs = json.load(open(os.path.join(cs.GOLDEN, "runs.json")))
    run = runs["SmallBasisTest2_2013-09-03_19_06_47"]
    hv = run["hv_deck"]
    nx, xmax, lmax, na, nb = int(hv[0]), float(hv[1]), int(hv[6]), int(hv[7]), int(hv[8])
    spec, _ = cs.cage_case(nx, lmax, na, nb, hv[21], x_max=xmax, mass=float(hv[15]), rot=float(hv[16]),
                           ceiling=float(hv[19]))
    a, b, b2 = O.Oracle(spec).lanczos(int(run["lanczos_deck"][0]))
    _, ed, _ = _compare(a, b, b2, float(run["lanczos_deck"][1]), float(run["lanczos_deck"][2]))
    ref = np.array(run["boundstates"])[:, 0]
    for e in ref[:10]:
        assert np.min(np.abs(ed - e)) / abs(e) < 1e-9, e


@pytest.mark.parametrize("niter", [200, 900])
def test_gpu_runs(niter):
    """Longer runs (ghost copies appear): cage and degenerate oscillator."""
    import bsl_amd as B
    _, kw = cs.cage_case(6, 3, 8, 8, "All")
    with B.Calculator(**kw) as c:
        _compare(*c.lanczos(niter, start="random", seed=2), -10.0, 10.0)
    _, kw = cs.iho_case(9, 2, 10, 10, x_max=0.4)
    with B.Calculator(**kw) as c:
        _compare(*c.lanczos(niter, start="random", seed=3), 0.0, 30.0)


def test_empty_window():
    import bsl_amd as B
    a = np.array([1.0, 2.0, 3.0])
    b = np.array([0.0, 0.5, 0.5, 0.1])
    ev, er = B.lancbis(a, b, b * b, 10.0, 20.0, device=True)
    assert len(ev) == 0
