"""GPU parity beyond the sizes of the register and MFMA kernels, through the
C ABI, against the CPU oracle: axes of more than 64 points (the FP64 line
kernels of the kinetic term, up to 256 points) and lmax 7, 8 and 12 (parity
blocks of 28 to 91 planes: the row-parallel block kernel of the potential;
the quadrature calculator up to its 16 m rows).

Tolerances: Hv 1e-12 of the largest magnitude (FP64), 1e-6 (FP32
potential), 2e-5 (BF16x3 quadrature); alpha/beta over a 20-step prefix
1e-9 relative.
"""
import numpy as np
import pytest

import _cases as cs
import _oracle as O

pytestmark = pytest.mark.gpu

import bsl_amd as B  # noqa: E402  (a missing library fails the tests, it never skips them)


def _rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def _axes(spec, kw, nx, ny, nz):
    spec = dict(spec, nx=nx, ny=ny, nz=nz)
    kw = dict(kw, nx=nx, ny=ny, nz=nz)
    return spec, kw


@pytest.mark.parametrize("nx,ny,nz", [(66, 66, 66), (70, 9, 5), (5, 65, 8), (6, 7, 67), (256, 4, 3)])
def test_long_axes_hv_matches_oracle(nx, ny, nz):
    """Every axis above 64 points (even and odd lengths), one at a time, up
    to the 256-point maximum."""
    spec, kw = cs.cage_case(4, 1, 4, 5, "All")
    spec, kw = _axes(spec, kw, nx, ny, nz)
    o = O.Oracle(spec)
    with B.Calculator(**kw) as c:
        v = np.random.default_rng(nx + ny + nz).uniform(-1, 1, o.N)
        assert _rel(c.hv(v), o.hv(v)) < 1e-12


def test_long_axes_recurrence_and_precisions():
    """The recurrence through the line kernels (their v = ca X + cb R pass)
    against the oracle's, the reduced basis, and an FP32 potential."""
    spec, kw = cs.cage_case(4, 2, 6, 7, "Odd")
    spec, kw = _axes(spec, kw, 72, 10, 9)
    o = O.Oracle(spec)
    with B.Calculator(**kw) as c:
        ag, bg, _ = c.lanczos(30, start="random", seed=4)
    ar, br, _ = o.lanczos(30, O.start_vector(o.N, 4))
    assert np.max(np.abs(ag[:20] - ar[:20]) / np.abs(ar[:20])) < 1e-9
    assert np.max(np.abs(bg[1:21] - br[1:21]) / br[1:21]) < 1e-9
    with B.Calculator(basis="reduced", **kw) as c:
        ag, bg, _ = c.lanczos(30, start="random", seed=5)
    ar, br, _ = o.lanczos(30, cs.reduced_start(o, 5, 1))
    assert np.max(np.abs(ag[:20] - ar[:20]) / np.abs(ar[:20])) < 1e-9
    v = np.random.default_rng(6).uniform(-1, 1, o.N)
    with B.Calculator(precision="fp32", **kw) as c:
        assert _rel(c.hv(v), o.hv(v)) < 1e-6


@pytest.mark.parametrize("lmax,sym,basis,kernel", [(7, "All", "full", "k_vblock_rows"),
                                                    (7, "Even", "reduced", "k_vblock<28, 0"),
                                                    (8, "Odd", "reduced", "k_vblock_rows"),
                                                    (8, "All", "full", "k_vblock_rows")])
def test_high_lmax_block_rows_match_oracle(lmax, sym, basis, kernel):
    """Parity blocks of 28 + 36, 36 and 45 + 36 planes (no register kernel)
    and of 28 (the largest register kernel): Hv and the recurrence prefix."""
    spec, kw = cs.cage_case(4, lmax, 10, 12, sym)
    o = O.Oracle(spec)
    with B.Calculator(basis=basis, **kw) as c:
        assert c.potential_kernel().startswith(kernel)
        v = np.random.default_rng(lmax).uniform(-1, 1, o.N)
        if basis == "reduced":
            v = cs.reduced_start(o, lmax, 0 if sym == "Even" else 1)
        assert _rel(c.hv(v), o.hv(v)) < 1e-12
        ag, bg, _ = c.lanczos(30, start="random", seed=7)
    v0 = O.start_vector(o.N, 7) if basis == "full" else cs.reduced_start(o, 7, 0 if sym == "Even" else 1)
    ar, br, _ = o.lanczos(30, v0)
    assert np.max(np.abs(ag[:20] - ar[:20]) / np.abs(ar[:20])) < 1e-9
    assert np.max(np.abs(bg[1:21] - br[1:21]) / br[1:21]) < 1e-9


def test_lmax12_matches_oracle():
    """The largest lmax: 169 planes in blocks of 91 and 78, both on the
    row-parallel kernel (46.6 KB of LDS rows), reduced and full basis."""
    spec, kw = cs.cage_case(3, 12, 14, 14, "All")
    o = O.Oracle(spec)
    with B.Calculator(**kw) as c:
        assert c.potential_kernel().startswith("k_vblock_rows")
        v = np.random.default_rng(12).uniform(-1, 1, o.N)
        assert _rel(c.hv(v), o.hv(v)) < 1e-12
    spec, kw = cs.cage_case(3, 12, 14, 14, "Odd")
    o = O.Oracle(spec)
    with B.Calculator(basis="reduced", **kw) as c:
        v = cs.reduced_start(o, 13, 1)
        assert _rel(c.hv(v), o.hv(v)) < 1e-12


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("fp32", 1e-6), ("bf16x3", 2e-5)])
def test_quadrature_lmax7_matches_oracle(precision, tol):
    """The quadrature calculator at its largest angular size: 15 m rows, 64
    carried planes (both parities); 8 theta nodes (4 per hemisphere, so that
    16 points fill whole column tiles and the group's LDS fits)."""
    spec, kw = cs.cage_case(3, 7, 8, 12, "All")
    o = O.Oracle(spec)
    with B.Calculator(calculator="quad", precision=precision, **kw) as c:
        v = np.random.default_rng(8).uniform(-1, 1, o.N)
        assert _rel(c.hv(v), o.hv(v)) < tol


def test_batch_with_long_axis_member():
    """A member with a 66-point axis runs its own kinetic and potential
    launches inside the lockstep iteration; every member equals its single
    run bitwise."""
    _, k0 = cs.baseline_cage_case(5, 40, 70, "Even", seed=0)
    _, k1 = cs.baseline_cage_case(5, 40, 70, "Odd", seed=1)
    members = [dict(k0, basis="reduced", precision="fp32"),
               dict(k1, basis="reduced", precision="fp32", nx=66, ny=5, nz=5),
               dict(k0, basis="reduced", precision="fp32", nx=5, ny=5, nz=6)]
    niter, seed = 16, 30
    single = []
    for i, kw in enumerate(members):
        with B.Calculator(**kw) as c:
            single.append(c.lanczos(niter, start="random", seed=seed + i))
    with B.Batch(members) as bt:
        for graph in (False, True):
            got = bt.lanczos(niter, start="random", seed=seed, graph=graph)
            for (a, b, b2), (ra, rb, rb2) in zip(got, single):
                assert np.array_equal(a, ra) and np.array_equal(b, rb) and np.array_equal(b2, rb2), graph
