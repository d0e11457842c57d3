"""GPU parity of the fused z-kinetic + potential kernel (k_zpot2, k_zpot1).

A NoQuad step whose potential is the lmax-5 harmonic-form pair of blocks
(full basis) or one of its blocks (reduced basis) can run (bsl_set_fusion)
the x-y kinetic kernel and then one kernel that applies the z contraction (Tv_5D's z term,
linRotCartSphHvRoutines.cpp:1872-1969) and the potential (Vv_5D,
linRotCartSphHvFullVPrepRoutines.cpp:675-734) to tiles of 4 columns x every
z point x the block's planes.  The bar is the FP64 bar of the other
kernels: Hv within 1e-12 of the oracle relative to its largest magnitude,
within 1e-13 of the three-kernel step of the same context (the two differ
in summation order only), the recurrence's alpha/beta prefix within 1e-9;
FP32 within 2e-6 of the oracle.  The grids cover ragged last tiles (x-y
sizes that are not multiples of 4), odd and even z axes, one and two
M-tiles of half indices (z up to 32 and beyond), the largest z axis (64)
and the masked full basis.
"""
import numpy as np
import pytest

import _cases as cs
import _oracle as O

pytestmark = pytest.mark.gpu

import bsl_amd as B  # noqa: E402  (a missing library fails the tests, it never skips them)


def _rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def _aniso(case, nx, ny, nz):
    spec, kw = case
    spec = dict(spec, nx=nx, ny=ny, nz=nz)
    kw = dict(kw, nx=nx, ny=ny, nz=nz)
    return spec, kw


# (nx, ny, nz): x-y sizes 25, 42, 36, 9, 6, 20 (the last tile of 4 columns
# holds 1, 2, 4, 1, 2, 4 of them); z odd and even, with 1 or 2 M-tiles of 16
# half indices (z = 33: the middle point alone in the second)
GRIDS = [(5, 5, 5), (7, 6, 6), (6, 6, 33), (3, 3, 64), (2, 3, 40), (4, 5, 2)]


@pytest.mark.parametrize("dims", GRIDS)
@pytest.mark.parametrize("sym,basis", [("All", "full"), ("Even", "reduced"), ("Odd", "reduced")])
def test_fused_hv_matches_oracle_and_three_kernel_step(dims, sym, basis):
    spec, kw = _aniso(cs.cage_case(dims[0], 5, 8, 9, sym), *dims)
    o = O.Oracle(spec)
    with B.Calculator(basis=basis, **kw) as c:
        assert not c.fused
        c.fused = True
        kname = c.potential_kernel()
        assert kname.startswith("k_zpot2<double" if basis == "full" else
                                ("k_zpot1<4, 0, double" if sym == "Even" else "k_zpot1<5, 1, double"))
        v = np.random.default_rng(31).uniform(-1, 1, o.N)
        u0 = np.random.default_rng(32).uniform(-1, 1, o.N)
        ref = u0.copy()
        O.lib.bso_hv(o.h, v, ref)
        got = c.hv(v, u0)
        assert _rel(got, ref) < 1e-12, kname
        c.fused = False
        got3 = c.hv(v, u0)
        assert _rel(got3, ref) < 1e-12
        assert _rel(got, got3) < 1e-13


@pytest.mark.parametrize("sym", ["Even", "Odd"])
def test_fused_masked_full_basis(sym):
    """Full basis with a symmetry mask: the masked block's rows get no
    potential and no z term (the mask is uniform within a parity block)."""
    spec, kw = cs.cage_case(6, 5, 8, 9, sym)
    o = O.Oracle(spec)
    with B.Calculator(**kw) as c:
        c.fused = True
        v = np.random.default_rng(41).uniform(-1, 1, o.N)
        assert _rel(c.hv(v), o.hv(v)) < 1e-12


def test_fused_recurrence_matches_oracle():
    """alpha/beta of the device recurrence (the fused kernel's partial sums
    feed alpha and beta) against the oracle's, full and reduced bases."""
    for (spec, kw), basis in ((cs.cage_case(7, 5, 8, 9, "All"), "full"), (cs.cage_case(6, 5, 8, 9, "Odd"), "reduced")):
        o = O.Oracle(spec)
        with B.Calculator(basis=basis, **kw) as c:
            c.fused = True
            n = 30
            if basis == "full":
                a_o, b_o, _ = o.lanczos(n)
                a_g, b_g, _ = c.lanczos(n)
            else:
                v0 = cs.reduced_start(o, 5, 1)
                a_o, b_o, _ = o.lanczos(n, v0=v0)
                a_g, b_g, _ = c.lanczos(n, start="random", seed=5)
            assert np.max(np.abs(a_g - a_o) / np.abs(a_o)) < 1e-9
            assert np.max(np.abs(b_g[1:] - b_o[1:]) / np.abs(b_o[1:])) < 1e-9


def test_fused_fp32_close_to_oracle():
    spec, kw = cs.cage_case(6, 5, 12, 14, "All")
    o = O.Oracle(spec)
    v = np.random.default_rng(11).uniform(-1, 1, o.N)
    ref = o.hv(v)
    with B.Calculator(precision="fp32", **kw) as c:
        c.fused = True
        assert c.potential_kernel().startswith("k_zpot2<float")
        h = c.hv(v)
        c.fused = False
        h3 = c.hv(v)
    assert _rel(h, ref) < 2e-6 and _rel(h3, ref) < 2e-6


def test_fusion_applies_only_where_built():
    """BF16x3, the block form, other block shapes and axes of more than 64
    points keep the three-kernel step; asking for the fused one fails."""
    cases = [dict(cs.cage_case(5, 5, 8, 9, "All")[1], precision="bf16x3"),
             cs.cage_case(5, 4, 8, 9, "All")[1],
             dict(cs.cage_case(4, 5, 8, 9, "All")[1], nz=65)]
    for kw in cases:
        with B.Calculator(**kw) as c:
            assert not c.fused
            with pytest.raises(B.BslError):
                c.fused = True
    with B.Calculator(**cs.cage_case(5, 5, 8, 9, "All")[1]) as c:
        c.fused = True
        c.potential_form = "blocks"
        assert not c.fused
        with pytest.raises(B.BslError):
            c.fused = True


def test_fused_levels_match_three_kernel_step():
    """Converged levels of 2000-step runs agree between the two step
    structures to the FP64 bar (1e-9 relative)."""
    spec, kw = cs.cage_case(8, 5, 10, 11, "All")
    conv, good = {}, {}
    for fused in (True, False):
        with B.Calculator(**kw) as c:
            c.fused = fused
            ev, err = B.lancbis(*c.lanczos(2000, start="random", seed=3), -10.0, 10.0)
            good[fused] = np.sort(ev)
            conv[fused] = ev[(np.abs(err) < 1e-10) & (err != 0)]
    for f in (True, False):
        assert len(conv[f]) >= 8
        d = np.min(np.abs(conv[f][:, None] - good[not f][None, :]), axis=1)
        assert np.max(d / np.maximum(np.abs(conv[f]), 1.0)) < 1e-9, f


def test_batch_members_with_and_without_fusion():
    """A batch whose reduced-basis members run the fused table kernel and
    the three-kernel one side by side (FP64 and FP32, odd and even axes)
    gives exactly the coefficients of the members run alone."""
    specs = [(cs.cage_case(5, 5, 8, 9, "Even"), "fp64", True), (cs.cage_case(6, 5, 8, 9, "Odd"), "fp32", True),
             (cs.cage_case(7, 5, 8, 9, "Odd"), "fp64", False), (cs.cage_case(6, 5, 8, 9, "All"), "fp64", True)]
    members = [dict(kw, basis="reduced" if sp["sym"] else "full", precision=prec) for (sp, kw), prec, _ in specs]
    ref = []
    for i, (kw, (_, _, fused)) in enumerate(zip(members, specs)):
        with B.Calculator(**kw) as c:
            c.fused = fused
            ref.append(c.lanczos(25, start="random", seed=7 + i))
    with B.Batch(members) as bt:
        for i, (_, _, fused) in enumerate(specs):
            bt.member(i).fused = fused
        for graph in (False, True):
            got = bt.lanczos(25, start="random", seed=7, graph=graph)
            for (a, b, b2), (ra, rb, rb2) in zip(got, ref):
                assert np.array_equal(a, ra) and np.array_equal(b, rb) and np.array_equal(b2, rb2), graph
