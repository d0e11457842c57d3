"""GPU parity of the harmonic form of the NoQuad potential (BSL_VFORM_HARMONICS).

The kernel rebuilds each per-point block <lm|V|l'm'> from the even-L
harmonic coefficients of the potential at that point through the Gaunt
coefficients of the real harmonics.  That is the same matrix as the
reference's quadrature sum (the product of two harmonics is a finite sum of
harmonics at every node of any grid), so the bar is the FP64 bar of the
block form: Hv within 1e-12 of the oracle relative to its largest
magnitude, and within 1e-13 of the block form of the same context.  Every
block shape the kernel is instantiated for is covered, on quadratures too
coarse to integrate the products exactly (the identity does not need it).
"""
import numpy as np
import pytest

import _cases as cs
import _oracle as O

pytestmark = pytest.mark.gpu

import bsl_amd as B  # noqa: E402  (a missing library fails the tests, it never skips them)


def _rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


# (lmax, symmetry, basis): blocks (lb, parity) = two-block shapes for lmax 1..5
# and every single-block shape of the reduced basis
SHAPES = [(0, "All", "full"), (1, "All", "full"), (2, "Even", "full"), (3, "Odd", "full"), (4, "All", "full"),
          (5, "All", "full"), (5, "Even", "full"), (0, "Even", "reduced"), (1, "Odd", "reduced"),
          (2, "Even", "reduced"), (3, "Odd", "reduced"), (4, "Even", "reduced"), (5, "Odd", "reduced")]


@pytest.mark.parametrize("lmax,sym,basis", SHAPES)
def test_harmonic_hv_matches_oracle_and_blocks(lmax, sym, basis):
    na, nb = max(2, lmax + 1), max(2, lmax + 2)  # coarser than 2 lmax + 1: products not integrated exactly
    spec, kw = cs.cage_case(5, lmax, na, nb, sym)
    o = O.Oracle(spec)
    with B.Calculator(basis=basis, **kw) as c:
        assert c.potential_form == "harmonics" and not c.fused
        assert c.potential_kernel().startswith("k_vgaunt")
        v = np.random.default_rng(100 + lmax).uniform(-1, 1, o.N)
        u0 = np.random.default_rng(200 + lmax).uniform(-1, 1, o.N)
        ref = u0.copy()
        O.lib.bso_hv(o.h, v, ref)
        h = c.hv(v, u0)
        assert _rel(h, ref) < 1e-12
        c.potential_form = "blocks"
        assert c.potential_form == "blocks" and c.potential_kernel().startswith("k_vblock<") and not c.fused
        b = c.hv(v, u0)
        assert _rel(b, ref) < 1e-12
        assert _rel(h, b) < 1e-13
        c.potential_form = "harmonics"
        assert _rel(c.hv(v, u0), h) == 0.0  # rebuilt from a regenerated grid: the same numbers


def test_harmonic_vblocks_match_oracle():
    """The blocks the harmonic form stands for (bsl_get_vblocks expands them
    with the kernel's term list) against the oracle's quadrature sums."""
    spec, kw = cs.cage_case(6, 5, 8, 9, "All")
    o = O.Oracle(spec)
    with B.Calculator(**kw) as c:
        assert c.potential_form == "harmonics"
        vg = c.vblocks()
        c.potential_form = "blocks"
        vb = c.vblocks()
    vr = o.vblocks()
    lsum = np.add.outer(o.qnum()[:, 0], o.qnum()[:, 0])
    same = lsum % 2 == 0
    assert _rel(vg[..., same], vr[..., same]) < 1e-12
    assert _rel(vg, vb) < 1e-13


def test_harmonic_form_at_production_quadrature():
    """lmax 5 on the 40 x 70 quadrature of the 54^3 workload, both blocks
    and ragged last workgroups (5^3 = 125 points: 61 of the second 64-point
    group; 7^3 = 343), other potentials."""
    cases = ((cs.cage_case(5, 5, 40, 70, "All"), "k_vgaunt_split<"), (cs.cage_case(7, 5, 12, 13, "Odd"), "k_vgaunt_split<"),
             (cs.cage_grid_case(5, 5, 12, 14, "Odd"), "k_vgaunt_split<"),
             (cs.iho_case(5, 5, 10, 12, x_max=0.4), "k_vgaunt_split<"), (cs.coulomb_case(5, 4, 9, 10), "k_vgaunt<"),
             ((lambda sp, k: (sp, dict(k, basis="reduced")))(*cs.cage_case(7, 5, 12, 13, "Odd")),
              "k_vgaunt_split1<5, 1"),
             ((lambda sp, k: (sp, dict(k, basis="reduced")))(*cs.cage_case(5, 5, 40, 70, "Even")),
              "k_vgaunt_split1<4, 0"))
    for (spec, kw), kernel in cases:
        o = O.Oracle(spec)
        with B.Calculator(**kw) as c:
            assert c.potential_form == "harmonics" and c.potential_kernel().startswith(kernel)
            v = np.random.default_rng(7).uniform(-1, 1, o.N)
            assert _rel(c.hv(v), o.hv(v)) < 1e-12


def test_lmax6_uses_blocks():
    """The even block of lmax 6 (28 planes, 91 coefficients) stays in block
    form; the harmonic form is refused for it."""
    spec, kw = cs.cage_case(3, 6, 8, 8, "All")
    o = O.Oracle(spec)
    with B.Calculator(**kw) as c:
        assert c.potential_form == "blocks"
        with pytest.raises(B.BslError):
            c.potential_form = "harmonics"
        v = np.random.default_rng(9).uniform(-1, 1, o.N)
        assert _rel(c.hv(v), o.hv(v)) < 1e-12


def test_harmonic_fp32_hv():
    """FP32 coefficients (FP64 arithmetic): within the FP32 bar of the
    oracle, and of the FP32 block form."""
    spec, kw = cs.cage_case(6, 5, 12, 14, "All")
    o = O.Oracle(spec)
    v = np.random.default_rng(11).uniform(-1, 1, o.N)
    ref = o.hv(v)
    with B.Calculator(precision="fp32", **kw) as c:
        assert c.potential_form == "harmonics" and "float" in c.potential_kernel()
        h = c.hv(v)
        c.potential_form = "blocks"
        b = c.hv(v)
    assert _rel(h, ref) < 1e-6 and _rel(b, ref) < 1e-6


def test_harmonic_levels_match_blocks():
    """Levels of a 1500-step run agree between the forms to the FP64 bar:
    every level converged to 1e-10 in one run has a good level of the other
    within 1e-9 relative (which levels pass a fixed error cut is itself
    rounding-dependent near the cut)."""
    spec, kw = cs.cage_case(8, 4, 10, 11, "All")
    good, conv = {}, {}
    for form in ("harmonics", "blocks"):
        with B.Calculator(**kw) as c:
            c.potential_form = form
            ev, err = B.lancbis(*c.lanczos(1500, start="random", seed=3), -10.0, 10.0)
            good[form] = np.sort(ev)
            conv[form] = ev[(np.abs(err) < 1e-10) & (err != 0)]
    for f, g in (("harmonics", "blocks"), ("blocks", "harmonics")):
        assert len(conv[f]) >= 20
        d = np.min(np.abs(conv[f][:, None] - good[g][None, :]), axis=1)
        assert np.max(d / np.maximum(np.abs(conv[f]), 1.0)) < 1e-9, f
