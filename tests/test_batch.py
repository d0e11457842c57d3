"""Batched lockstep recurrences (bsl_batch_*): mixed cages, symmetry blocks,
grid sizes and precisions in one batch give exactly the coefficients of
separate single-context runs, eagerly and through replayed hipGraphs.  The
batch launches each kernel variant once per step over a table of its
members; the cases cover every variant (even and odd axes, FP64 and FP32,
the full and both reduced lmax-5 potential shapes) and members whose
potential is launched on its own (other shapes, the quadrature calculator)."""
import ctypes as C

import numpy as np
import pytest

import _cases as cs


def _members():
    _, k1 = cs.cage_case(5, 2, 6, 6, "All")
    _, k2 = cs.cage_case(6, 3, 8, 8, "Even")
    k2 = dict(k2, basis="reduced")
    _, k3 = cs.cage_case(4, 3, 8, 9, "Odd")
    k3 = dict(k3, precision="fp32")
    _, k4 = cs.iho_case(5, 2, 6, 6, 0.4)
    _, k5 = cs.cage_case(4, 2, 8, 9, "All")
    k5 = dict(k5, calculator="quad", precision="bf16x3")
    return [k1, k2, k3, k4, k5]


def test_batch_rejects_empty():
    from bsl_amd._lib import lib
    h = C.c_void_p()
    assert lib().bsl_batch_create(0, None, C.byref(h)) != 0
    assert b"bad arguments" in lib().bsl_last_error()


@pytest.mark.gpu
@pytest.mark.parametrize("niter", [40, 41])
def test_batch_matches_single_runs(niter):
    import bsl_amd as B
    members = _members()
    ref = []
    for i, kw in enumerate(members):
        with B.Calculator(**kw) as c:
            ref.append(c.lanczos(niter, start="random", seed=7 + i))
    with B.Batch(members) as bt:
        for graph in (False, True):
            got = bt.lanczos(niter, start="random", seed=7, graph=graph)
            for (a, b, b2), (ra, rb, rb2) in zip(got, ref):
                assert np.array_equal(a, ra) and np.array_equal(b, rb) and np.array_equal(b2, rb2), graph


@pytest.mark.gpu
def test_graph_replays_with_eager_work_between():
    """Replayed graphs stay exact when eager work (an Hv through a member
    context, a host synchronisation) runs between the runs."""
    import bsl_amd as B
    members = _members()
    with B.Batch(members) as bt:
        eager = bt.lanczos(30, seed=3, graph=False)
        for rep in range(3):
            m = bt.member(rep % len(members))
            m.hv(np.random.default_rng(rep).uniform(-1, 1, m.N))
            got = bt.lanczos(30, seed=3, graph=True)
            for (a, b, _), (ra, rb, _) in zip(got, eager):
                assert np.array_equal(a, ra) and np.array_equal(b, rb), rep


@pytest.mark.gpu
def test_batch_member_levels():
    """Members are full contexts: converged levels (Loop 2 Ritz vectors and
    Rayleigh refinement) from a member equal those of a separate context."""
    import bsl_amd as B
    members = _members()[:2]
    with B.Batch(members) as bt:
        bt.lanczos(20)
        e, _, r = bt.member(1).levels(600, 2)
    with B.Calculator(**members[1]) as c:
        e1, _, _ = c.levels(600, 2)
    assert np.allclose(e, e1, rtol=1e-12, atol=0)


@pytest.mark.gpu
def test_batch_table_variants_match_single_runs():
    """lmax-5 members of every table potential shape (full basis pair of
    blocks, para and ortho blocks) in FP64 and FP32, on even, odd and
    anisotropic grids, 60 steps (an even and an odd count) eagerly and
    through graphs: bitwise the coefficients of the members run alone."""
    import bsl_amd as B
    members = []
    for i, (n, sym, basis, prec) in enumerate([(6, "All", "full", "fp64"), (5, "All", "full", "fp32"),
                                               (6, "Even", "reduced", "fp64"), (7, "Odd", "reduced", "fp64"),
                                               (5, "Even", "reduced", "fp32"), (6, "Odd", "reduced", "fp32")]):
        _, kw = cs.cage_case(n, 5, 8, 9, sym)
        members.append(dict(kw, basis=basis, precision=prec))
    _, kw = cs.cage_case(6, 5, 8, 9, "Odd")
    members.append(dict(kw, nz=5, basis="reduced"))
    for niter in (60, 61):
        ref = []
        for i, kw in enumerate(members):
            with B.Calculator(**kw) as c:
                ref.append(c.lanczos(niter, start="random", seed=40 + i))
        with B.Batch(members) as bt:
            for graph in (False, True):
                got = bt.lanczos(niter, start="random", seed=40, graph=graph)
                for i, ((a, b, b2), (ra, rb, rb2)) in enumerate(zip(got, ref)):
                    assert np.array_equal(a, ra) and np.array_equal(b, rb) and np.array_equal(b2, rb2), (niter, graph, i)
            ms = bt.time_steps(3, 6, graph=True)
            assert ms > 0
