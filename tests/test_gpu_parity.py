"""GPU parity of the HIP path (libbsl_hip.so, through its C ABI) against the
CPU oracle and the reference's own outputs.

Tolerances: the potential grid and the V blocks are compared elementwise
relative to their largest magnitude (1e-13 / 1e-12: the device sums the
quadrature in a different order); Hv the same way (1e-12); alpha/beta on
the prefix before loss of orthogonality at 1e-9 relative and converged
eigenvalues at 1e-9 relative (the FP64 bar); FP32 levels within 0.01 cm^-1.
"""
import json
import os

import numpy as np
import pytest

import _cases as cs
import _oracle as O

pytestmark = pytest.mark.gpu

import bsl_amd as B  # noqa: E402  (a missing library fails the tests, it never skips them)


def _rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


@pytest.fixture(scope="module")
def cage_small():
    spec, kw = cs.cage_case(6, 3, 8, 7, "All")
    return O.Oracle(spec), B.Calculator(**kw)


def test_vgrid_cage_matches_oracle(cage_small):
    o, c = cage_small
    g1, g2 = c.vgrid()
    r1, r2 = o.vgrid()
    assert _rel(g1, r1) < 1e-13 and _rel(g2, r2) < 1e-13


def test_vblocks_cage_matches_oracle(cage_small):
    o, c = cage_small
    vg = c.vblocks()
    vr = o.vblocks()
    lsum = np.add.outer(o.qnum()[:, 0], o.qnum()[:, 0])
    same = (lsum % 2 == 0)
    # couplings of opposite l parity vanish for the homonuclear rotor; the
    # device does not store them
    assert np.max(np.abs(vr[..., ~same])) < 1e-12 * np.max(np.abs(vr))
    assert _rel(vg[..., same], vr[..., same]) < 1e-12


@pytest.mark.parametrize("sym", ["All", "Even", "Odd"])
def test_hv_full_matches_oracle(sym):
    spec, kw = cs.cage_case(5, 2, 6, 6, sym)
    o, c = O.Oracle(spec), B.Calculator(**kw)
    v = np.random.default_rng(7).uniform(-1, 1, o.N)
    u0 = np.random.default_rng(8).uniform(-1, 1, o.N)
    ref = u0.copy()
    O.lib.bso_hv(o.h, v, ref)
    got = c.hv(v, u0)
    assert _rel(got, ref) < 1e-12


@pytest.mark.parametrize("sym", ["Even", "Odd"])
def test_hv_reduced_matches_oracle(sym):
    spec, kw = cs.cage_case(5, 3, 8, 8, sym)
    o, c = O.Oracle(spec), B.Calculator(basis="reduced", **kw)
    v = np.random.default_rng(3).uniform(-1, 1, o.N)
    ref = o.hv(v)
    got = c.hv(v)
    assert c.N_krylov < c.N
    assert _rel(got, ref) < 1e-12


@pytest.mark.parametrize("dims", [(6, 5, 4), (8, 6, 5), (6, 8, 7), (7, 7, 6)])
def test_hv_anisotropic_grids_match_oracle(dims):
    """Axis lengths and box lengths that differ per axis (the deck's nx, ny,
    nz and x/y/z_max, linRotCartSphHvFullVPrepRoutines.cpp:248-321,
    362-382), odd and even: every pairing of the full-size and the half-size
    (centrosymmetric) kinetic contractions for x-y and for z."""
    nx, ny, nz = dims
    spec, kw = cs.cage_case(nx, 2, 6, 6, "All")
    spec = dict(spec, ny=ny, nz=nz, y_max=0.35, z_max=0.45)
    kw = dict(kw, ny=ny, nz=nz, y_max=0.35, z_max=0.45)
    o, c = O.Oracle(spec), B.Calculator(**kw)
    v = np.random.default_rng(21).uniform(-1, 1, o.N)
    u0 = np.random.default_rng(22).uniform(-1, 1, o.N)
    ref = u0.copy()
    O.lib.bso_hv(o.h, v, ref)
    assert _rel(c.hv(v, u0), ref) < 1e-12


@pytest.mark.parametrize("n", [64, 63])
def test_hv_largest_axis_matches_oracle(n):
    """64 points per axis, the largest the kinetic kernels stage (and 63, its
    odd neighbour); 65 is rejected at creation (test_host)."""
    spec, kw = cs.cage_case(n, 1, 4, 4, "All", x_max=0.8)
    o, c = O.Oracle(spec), B.Calculator(**kw)
    v = np.random.default_rng(23).uniform(-1, 1, o.N)
    assert _rel(c.hv(v), o.hv(v)) < 1e-12


def test_hv_iho_matches_oracle():
    spec, kw = cs.iho_case(9, 2, 6, 8, x_max=0.4)
    o, c = O.Oracle(spec), B.Calculator(**kw)
    v = np.random.default_rng(5).uniform(-1, 1, o.N)
    assert _rel(c.hv(v), o.hv(v)) < 1e-12


def test_hv_matches_reference_fixture():
    """Hv of the reference's own compiled calculator (committed vectors)."""
    fx = np.load(os.path.join(cs.GOLDEN, "ref_hv_cage11.npz"))
    for tag in ("a", "b"):
        nx, lmax, na, nb, sym = [int(x) for x in fx["%s_params" % tag]]
        _, kw = cs.cage_case(nx, lmax, na, nb, ["All", "Even", "Odd"][sym])
        c = B.Calculator(**kw)
        got = c.hv(fx["%s_v" % tag])
        assert _rel(got, fx["%s_hv" % tag]) < 1e-12, tag


def test_hv_device_entry_point():
    """bsl_hv_device on device buffers agrees with bsl_hv on host buffers."""
    import ctypes as C
    spec, kw = cs.cage_case(5, 2, 6, 6, "Even")
    c = B.Calculator(**kw)
    v = np.random.default_rng(9).uniform(-1, 1, c.N)
    u = np.random.default_rng(10).uniform(-1, 1, c.N)
    ref = c.hv(v, u)
    # the HIP runtime the library is linked against, by its soname (a torch
    # import may have loaded a second runtime under the bare file name)
    hip = C.CDLL("libamdhip64.so.7")
    dv, du = C.c_void_p(), C.c_void_p()
    nbytes = 8 * c.N
    assert hip.hipMalloc(C.byref(dv), C.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(C.byref(du), C.c_size_t(nbytes)) == 0
    try:
        assert hip.hipMemcpy(dv, v.ctypes.data_as(C.c_void_p), C.c_size_t(nbytes), 1) == 0
        assert hip.hipMemcpy(du, u.ctypes.data_as(C.c_void_p), C.c_size_t(nbytes), 1) == 0
        B.check(B.lib().bsl_hv_device(c.h, dv, du))
        out = np.zeros(c.N)
        assert hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), du, C.c_size_t(nbytes), 2) == 0
    finally:
        hip.hipFree(dv)
        hip.hipFree(du)
    assert np.array_equal(out, ref)


@pytest.mark.parametrize("start", ["uniform", "random"])
def test_lanczos_prefix_matches_oracle(start):
    spec, kw = cs.cage_case(6, 2, 6, 6, "All")
    o, c = O.Oracle(spec), B.Calculator(**kw)
    niter = 40
    v0 = O.start_vector(o.N, 11) if start == "random" else None
    ar, br, b2r = o.lanczos(niter, v0)
    ag, bg, b2g = c.lanczos(niter, start=start, seed=11)
    assert np.max(np.abs(ag - ar) / np.abs(ar)) < 1e-9
    assert np.max(np.abs(bg[1:] - br[1:]) / br[1:]) < 1e-9
    assert np.max(np.abs(b2g[1:] - b2r[1:]) / b2r[1:]) < 1e-9


def _converged(ev, err, tol=1e-8):
    return ev[(np.abs(err) < tol) & (err != 0)]


def test_eigenvalues_match_oracle():
    spec, kw = cs.cage_case(6, 3, 8, 8, "All")
    o, c = O.Oracle(spec), B.Calculator(**kw)
    niter = 200
    er_o = O.lancbis(*o.lanczos(niter), -10.0, 10.0)
    er_g = B.lancbis(*c.lanczos(niter), -10.0, 10.0)
    conv = _converged(*er_o)
    assert len(conv) >= 3
    for e in conv:
        assert np.min(np.abs(er_g[0] - e)) / abs(e) < 1e-9


def test_saved_run_eigenvalues():
    """cage 11, 10^3 x (lmax 3), 10x10 quadrature, 300 steps, uniform start:
    the levels the reference's saved run converged (boundstates.txt of
    Compilation/SmallBasisTest2_2013-09-03_19_06_47)."""
    runs = json.load(open(os.path.join(cs.GOLDEN, "runs.json")))
    run = runs["SmallBasisTest2_2013-09-03_19_06_47"]
    hv = run["hv_deck"]
    nx, xmax, lmax, na, nb = int(hv[0]), float(hv[1]), int(hv[6]), int(hv[7]), int(hv[8])
    _, kw = cs.cage_case(nx, lmax, na, nb, hv[21], x_max=xmax, mass=float(hv[15]), rot=float(hv[16]),
                         ceiling=float(hv[19]))
    c = B.Calculator(**kw)
    ev, err = B.lancbis(*c.lanczos(int(run["lanczos_deck"][0])), float(run["lanczos_deck"][1]),
                        float(run["lanczos_deck"][2]))
    ref = np.array(run["boundstates"])[:, 0]
    # the lowest levels are converged in the saved run (later ones depend on
    # rounding through the loss of orthogonality)
    for e in ref[:10]:
        assert np.min(np.abs(ev - e)) / abs(e) < 1e-9, e


def test_iho_analytic_levels():
    """(n + 3/2) hbar w + B l(l+1) on the wide box (x_max 0.4 nm, 21 points)."""
    from bsl_amd import physics
    spec, kw = cs.iho_case(21, 3, 20, 30, x_max=0.4)
    c = B.Calculator(**kw)
    ev, err = B.lancbis(*c.lanczos(300, start="random", seed=0), 0.0, 8.0)
    hw = physics.H_BAR * physics.omega_from_cm1(100.0)
    Bc = kw["rot_const"]
    levels = sorted({round((n + 1.5) * hw + Bc * l * (l + 1), 12) for n in range(8) for l in range(4)})
    # the levels are degenerate, so BISEC reports most of them as multiple
    # (error column 0): compare with every good eigenvalue
    good = np.sort(ev)
    low = [x for x in levels if x < 5.0]
    for e in low:
        assert np.min(np.abs(good - e)) / e < 1e-6, (e, good[:10])


def test_ritz_vectors_match_oracle():
    spec, kw = cs.cage_case(5, 2, 6, 6, "All")
    o, c = O.Oracle(spec), B.Calculator(**kw)
    niter = 60
    a, b, b2 = o.lanczos(niter)
    ev, err = O.lancbis(a, b, b2, -10.0, 10.0)
    ev = ev[:3]
    evtr, _ = O.trivec(a, b, b2, ev)
    ref = o.ritz(niter, evtr, len(ev))
    got = c.ritz(niter, evtr, len(ev))
    for g in range(len(ev)):
        assert _rel(got[g], ref[g]) < 1e-7


def test_ritz_windows_equal_single_vector_runs():
    """Several Ritz vectors are summed a window of 16 copied Lanczos vectors
    at a time, one alone step by step from the recurrence's own buffer: the
    same multiply-adds in the same order, so each vector of a 5-vector Loop 2
    (80 steps: five windows) is bitwise the one a Loop 2 for that vector
    alone gives."""
    spec, kw = cs.cage_case(5, 2, 6, 6, "All")
    o, c = O.Oracle(spec), B.Calculator(**kw)
    niter = 80
    a, b, b2 = o.lanczos(niter)
    ev, err = O.lancbis(a, b, b2, -10.0, 10.0)
    ev = ev[:5]
    evtr, _ = O.trivec(a, b, b2, ev)
    got = c.ritz(niter, evtr, len(ev))
    cols = np.asarray(evtr).reshape(len(ev), niter)
    for g in range(len(ev)):
        one = c.ritz(niter, np.ascontiguousarray(cols[g]), 1)
        assert np.array_equal(got[g], one[0]), g


def _oracle_rayleigh(o, y):
    hy = o.hv(y)
    e = float(y @ hy) / float(y @ y)
    return e, float(np.linalg.norm(hy - e * y) / np.linalg.norm(y))


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_rayleigh_refinement_matches_oracle(precision):
    """FP64 Rayleigh quotients and residual norms of Ritz vectors, on the
    device with the context's FP64 operator whatever its precision, against
    the oracle's own <y|H|y>/<y|y> and |Hy - e y|/|y| of the same vectors;
    bsl_ritz_refine (Loop 2 and the refinement on the device) equals
    bsl_ritz followed by bsl_rayleigh, also when its kept workspace is
    reused for fewer vectors."""
    spec, kw = cs.cage_case(5, 2, 6, 6, "All")
    o = O.Oracle(spec)
    c = B.Calculator(precision=precision, **kw)
    niter = 120
    a, b, b2 = c.lanczos(niter, start="random", seed=4)
    ev, err = B.lancbis(a, b, b2, -10.0, 10.0, mult_tol=c.mult_tol)
    evtr, _ = B.trivec(a, b, b2, ev[:3], mult_tol=c.mult_tol)
    y = c.ritz(niter, evtr, 3, start="random", seed=4)
    e, r = c.rayleigh(y)
    for k in range(3):
        eo, ro = _oracle_rayleigh(o, y[k])
        assert abs(e[k] - eo) <= 1e-12 * abs(eo), k
        assert abs(r[k] - ro) <= 1e-11 * abs(eo), k
    # a random vector (residual far from zero)
    z = np.random.default_rng(9).uniform(-1, 1, (1, o.N))
    ez, rz = c.rayleigh(z)
    eo, ro = _oracle_rayleigh(o, z[0])
    assert abs(ez[0] - eo) <= 1e-12 * abs(eo) and abs(rz[0] - ro) <= 1e-12 * ro
    assert abs(e[0] - ev[0]) / abs(ev[0]) < (1e-9 if precision == "fp64" else 1e-6)
    yv, e2, r2 = c.ritz_refine(niter, evtr, 3, start="random", seed=4)
    assert np.array_equal(yv, y) and np.allclose(e2, e, rtol=1e-14, atol=0) and np.allclose(r2, r, rtol=1e-9)
    _, e1, _ = c.ritz_refine(niter, evtr[:niter], 1, start="random", seed=4, vectors=False)
    assert abs(e1[0] - e[0]) <= 1e-14 * abs(e[0])


@pytest.mark.parametrize("precision", ["fp32", "bf16x3"])
def test_fp32_levels_within_target(precision):
    """FP32 potential coefficients with FP32 (or BF16x3) kinetic MFMA
    contractions, FP64 recurrence: lowest levels within 0.01 cm^-1."""
    from bsl_amd import physics
    spec, kw = cs.cage_case(8, 3, 10, 10, "All")
    c64 = B.Calculator(**kw)
    c32 = B.Calculator(precision=precision, **kw)
    e64, r64 = B.lancbis(*c64.lanczos(300), -10.0, 10.0)
    e32, r32 = B.lancbis(*c32.lanczos(300), -10.0, 10.0, mult_tol=c32.mult_tol)
    conv = _converged(e64, r64, 1e-9)[:5]
    assert len(conv) >= 3
    for e in conv:
        d = np.min(np.abs(e32 - e)) / physics.KJMOL_PER_CM1
        assert d < 0.01, (e, d)


def test_mixed_levels_match_fp64_one_to_one():
    """A long FP32 run (3000 steps on a 3456-dimensional operator, many ghost
    copies): with the mixed calculators' multiplicity tolerance every
    converged FP64 level has an FP32 counterpart within 0.01 cm^-1 (no copy
    dropped as spurious), and a second counterpart only where a copy is still
    converging, which its own error estimate (>= 1e-3) marks.  Without the
    tolerance the same coefficients report the ghost copies as hundreds of
    extra levels."""
    from bsl_amd import physics
    spec, kw = cs.cage_case(6, 3, 8, 8, "All")
    c64 = B.Calculator(**kw)
    c32 = B.Calculator(precision="fp32", **kw)
    assert c64.mult_tol == 0.0 and c32.mult_tol == B.MIXED_MULT_TOL
    e64, r64 = B.lancbis(*c64.lanczos(3000, start="random", seed=5), -10.0, 10.0)
    e32, r32 = B.lancbis(*c32.lanczos(3000, start="random", seed=5), -10.0, 10.0, mult_tol=c32.mult_tol)
    conv = e64[np.abs(r64) < 1e-8]
    assert len(conv) >= 20
    extra = 0
    for e in conv:
        near = np.abs(e32 - e) / physics.KJMOL_PER_CM1 < 0.01
        assert near.sum() >= 1, e
        if near.sum() > 1:
            err = np.sort(np.abs(r32[near]))
            assert np.all(err[1:] >= 1e-3), (e, e32[near], r32[near])
            extra += near.sum() - 1
    assert extra <= 0.02 * len(conv)
    e0, _ = B.lancbis(*c32.lanczos(3000, start="random", seed=5), -10.0, 10.0)
    assert len(e0) > 1.5 * len(e64)  # BISEC's FP64 width keeps the spread copies apart
    ed, _ = B.lancbis(*c32.lanczos(3000, start="random", seed=5), -10.0, 10.0, mult_tol=c32.mult_tol, device=True)
    assert len(ed) == len(e32) and np.max(np.abs(ed - e32)) < 1e-10


def test_reduced_basis_drops_spurious_zero():
    """para (Even) block: the reduced basis reproduces the masked full-basis
    levels without the spurious 0 of the masked subspace.

    The masked full-basis operator is not symmetric (masked rows are zero,
    their columns are not), so where its 0 lands is set by rounding alone:
    the same 200-step recurrence on this case puts it at 1e-7 (the oracle),
    2e-7, 1e-6 or 1.5e-5 depending only on how alpha and beta are summed.
    The check is the dichotomy: a level within 1e-3 of 0 in the full basis,
    none in the reduced basis."""
    spec, kw = cs.cage_case(6, 2, 6, 6, "Even")
    full = B.Calculator(**kw)
    red = B.Calculator(basis="reduced", **kw)
    ef, rf = B.lancbis(*full.lanczos(200), -10.0, 10.0)
    er, rr = B.lancbis(*red.lanczos(200), -10.0, 10.0)
    assert np.min(np.abs(ef)) < 1e-3          # the masked subspace's zero
    assert np.min(np.abs(er)) > 1e-3
    conv = _converged(ef, rf)
    conv = conv[np.abs(conv) > 1e-6]
    assert len(conv) >= 2
    for e in conv:
        assert np.min(np.abs(er - e)) / abs(e) < 1e-9


# ------------------------------------------------------ quadrature calculator

@pytest.mark.parametrize("sym,basis", [("All", "full"), ("Odd", "full"), ("Even", "reduced")])
def test_quad_fp64_hv_matches_oracle(sym, basis):
    """The quadrature calculator evaluates the same operator as the NoQuad
    blocks (V_nn'(P) = sum_q Y_n w V Y_n'), so it is checked against the
    oracle's NoQuad Hv."""
    spec, kw = cs.cage_case(5, 3, 8, 9, sym)
    o = O.Oracle(spec)
    c = B.Calculator(calculator="quad", basis=basis, **kw)
    v = np.random.default_rng(21).uniform(-1, 1, o.N)
    assert _rel(c.hv(v), o.hv(v)) < 1e-12


def _tip4p_case(nx, lmax, na, nb, sym="All"):
    pos, q, typ = O.build_sites(cs.cage11_cm_ea(), 1)
    lj = np.array([1 if t == ord("O") else 0 for t in typ], dtype=np.int32)
    return cs.cage_case(nx, lmax, na, nb, sym, sites=(pos, q, typ, lj))


@pytest.mark.parametrize("case", ["tip4p", "cage_grid", "coulomb", "iho", "anisotropic"])
def test_quad_fp64_other_potentials_match_oracle(case):
    """The quadrature calculator over every potential grid the library builds
    (TIP4P waters, the grid-mapped universe, the Coulomb charge, the
    oscillator) and on an anisotropic grid, against the oracle's NoQuad Hv;
    the NoQuad calculator on the same operator too."""
    if case == "tip4p":
        spec, kw = _tip4p_case(5, 3, 8, 9)
    elif case == "cage_grid":
        spec, kw = cs.cage_grid_case(5, 3, 8, 9, "All", universe=(17, 19, 23, 0.44, 0.5, 0.41))
    elif case == "coulomb":
        spec, kw = cs.coulomb_case(6, 2, 6, 7)
    elif case == "iho":
        spec, kw = cs.iho_case(7, 3, 8, 9, x_max=0.4)
    else:
        spec, kw = cs.cage_case(6, 3, 8, 9, "All")
        spec = dict(spec, ny=5, nz=7, y_max=0.35, z_max=0.45)
        kw = dict(kw, ny=5, nz=7, y_max=0.35, z_max=0.45)
    o = O.Oracle(spec)
    v = np.random.default_rng(43).uniform(-1, 1, o.N)
    ref = o.hv(v)
    for calc in ("quad", "noquad"):
        with B.Calculator(calculator=calc, **kw) as c:
            assert _rel(c.hv(v), ref) < 1e-12, calc


def test_quad_fp32_hv_close_to_oracle():
    spec, kw = cs.cage_case(5, 3, 8, 9, "All")
    o = O.Oracle(spec)
    c = B.Calculator(calculator="quad", precision="fp32", **kw)
    v = np.random.default_rng(22).uniform(-1, 1, o.N)
    assert _rel(c.hv(v), o.hv(v)) < 1e-6


QUAD_TOL = {"fp64": 1e-12, "fp32": 1e-6, "bf16x3": 2e-5}


@pytest.mark.parametrize("precision", ["fp64", "fp32", "bf16x3"])
def test_quad_hv_production_shape(precision):
    """lmax 5 (11 m, K padded to 12) and the 40 x 70 quadrature of the 54^3
    workload (140 phi nodes in 9 MFMA tiles of 16, the last one partial) on
    a small grid whose last workgroup holds fewer points."""
    spec, kw = cs.cage_case(3, 5, 40, 70, "All")
    o = O.Oracle(spec)
    c = B.Calculator(calculator="quad", precision=precision, **kw)
    v = np.random.default_rng(23).uniform(-1, 1, o.N)
    assert _rel(c.hv(v), o.hv(v)) < QUAD_TOL[precision]


@pytest.mark.parametrize("na,nb,sym,basis", [(5, 7, "Even", "reduced"), (3, 4, "Odd", "full"),
                                             (12, 16, "All", "full")])
def test_quad_ragged_tiles(na, nb, sym, basis):
    """theta counts that are not multiples of 16 (16 points per workgroup for
    odd na), phi counts below and at one MFMA tile."""
    spec, kw = cs.cage_case(4, 2, na, nb, sym)
    o = O.Oracle(spec)
    c = B.Calculator(calculator="quad", basis=basis, **kw)
    v = np.random.default_rng(24).uniform(-1, 1, o.N)
    assert _rel(c.hv(v), o.hv(v)) < 1e-12


@pytest.mark.parametrize("na,nb,sym,basis", [(5, 7, "Even", "reduced"), (3, 20, "Odd", "full"),
                                             (6, 100, "All", "full"), (4, 128, "All", "reduced")])
def test_quad_bf16x3_phi_tile_pairs(na, nb, sym, basis):
    """The BF16 instruction takes phi tiles in pairs: 14, 40, 200 and 256 phi
    nodes fill 2, 4, 16 and 16 tiles (zero padded), with one or both l
    parities carried."""
    spec, kw = cs.cage_case(4, 3, na, nb, sym)
    o = O.Oracle(spec)
    c = B.Calculator(calculator="quad", precision="bf16x3", basis=basis, **kw)
    v = np.random.default_rng(25).uniform(-1, 1, o.N)
    assert _rel(c.hv(v), o.hv(v)) < QUAD_TOL["bf16x3"]


def test_quad_bf16x3_levels_within_target():
    """BF16x3 split products on MFMA: converged levels within 0.01 cm^-1 of
    the FP64 NoQuad operator's."""
    from bsl_amd import physics
    spec, kw = cs.cage_case(6, 3, 8, 8, "All")
    cq = B.Calculator(calculator="quad", precision="bf16x3", **kw)
    cn = B.Calculator(**kw)
    eq, rq = B.lancbis(*cq.lanczos(200), -10.0, 10.0)
    en, rn = B.lancbis(*cn.lanczos(200), -10.0, 10.0)
    conv = _converged(en, rn)
    assert len(conv) >= 3
    for e in conv:
        assert np.min(np.abs(eq - e)) / physics.KJMOL_PER_CM1 < 0.01


def test_quad_eigenvalues_match_noquad():
    spec, kw = cs.cage_case(6, 3, 8, 8, "All")
    cq = B.Calculator(calculator="quad", **kw)
    cn = B.Calculator(**kw)
    eq, rq = B.lancbis(*cq.lanczos(200), -10.0, 10.0)
    en, rn = B.lancbis(*cn.lanczos(200), -10.0, 10.0)
    conv = _converged(en, rn)
    assert len(conv) >= 3
    for e in conv:
        assert np.min(np.abs(eq - e)) / abs(e) < 1e-9


# ---------------------------------------------- grid-mapped and Coulomb ----

@pytest.mark.parametrize("sym,basis", [("All", "full"), ("Odd", "reduced")])
def test_grid_mapped_potential_and_hv(sym, basis):
    """Potential universe lookups (linRotCartSph_NoQuad_Alavi_SPCE) and Hv."""
    spec, kw = cs.cage_grid_case(5, 3, 8, 9, sym, universe=(17, 19, 23, 0.44, 0.5, 0.41))
    o = O.Oracle(spec)
    c = B.Calculator(basis=basis, **kw)
    g1, g2 = c.vgrid()
    r1, r2 = o.vgrid()
    assert _rel(g1, r1) < 1e-13 and _rel(g2, r2) < 1e-13
    v = np.random.default_rng(41).uniform(-1, 1, o.N)
    assert _rel(c.hv(v), o.hv(v)) < 1e-12


@pytest.mark.parametrize("name,nlev", [("SmallBasisTest2_2013-08-29_19_35_47", 3),
                                       ("SmallBasisTest2_2013-08-29_19_42_01", 8)])
def test_saved_grid_mapped_runs(name, nlev):
    run = json.load(open(os.path.join(cs.GOLDEN, "runs.json")))[name]
    hv = run["hv_deck"]
    _, kw = cs.cage_grid_case(int(hv[0]), int(hv[6]), int(hv[7]), int(hv[8]), hv[21], x_max=float(hv[1]),
                              mass=float(hv[15]), rot=float(hv[16]), ceiling=float(hv[19]),
                              universe=(int(hv[9]), int(hv[11]), int(hv[13]), float(hv[10]), float(hv[12]),
                                        float(hv[14])))
    c = B.Calculator(**kw)
    ev, err = B.lancbis(*c.lanczos(int(run["lanczos_deck"][0])), -10.0, 10.0)
    ref = np.array(run["boundstates"])
    assert np.max(np.abs(ev[:nlev] - ref[:nlev, 0]) / np.abs(ref[:nlev, 0])) < 1e-10


def test_coulomb_potential_and_hv():
    spec, kw = cs.coulomb_case(6, 2, 6, 6)
    o = O.Oracle(spec)
    c = B.Calculator(**kw)
    g1, _ = c.vgrid()
    r1, _ = o.vgrid()
    assert _rel(g1, r1) < 1e-14
    v = np.random.default_rng(42).uniform(-1, 1, o.N)
    assert _rel(c.hv(v), o.hv(v)) < 1e-12


def test_breakdown_is_reported():
    """A start vector that is an eigenvector of the operator (from a dense
    diagonalisation of the oracle's, 4^3 x l <= 1: 256 states): the first
    residual vanishes, beta^2 = |r|^2 - alpha^2 keeps no correct digit, and
    the run reports that step instead of returning coefficients built on it
    (both Loop-1 entry points); a random start on the same context runs."""
    from bsl_amd._lib import START, BslError, lib
    spec, kw = cs.cage_case(4, 1, 4, 4, "All")
    o = O.Oracle(spec)
    n = o.N
    h = np.zeros((n, n))
    e = np.zeros(n)
    for j in range(n):
        e[j] = 1.0
        h[:, j] = o.hv(e)
        e[j] = 0.0
    w, vecs = np.linalg.eigh(0.5 * (h + h.T))
    y = np.ascontiguousarray(vecs[:, 0])
    with B.Calculator(**kw) as c:
        assert c.N == 256
        a, b, _ = c.lanczos(20, start="random", seed=1)
        assert np.all(np.isfinite(a)) and np.all(b[1:] > 0)
        with pytest.raises(BslError) as err:
            c.lanczos(5, v0=y)
        assert err.value.code == 7 and "breakdown at step 1" in str(err.value)
        aa, bb, bb2 = np.zeros(5), np.zeros(6), np.zeros(6)
        rc = lib().bsl_lanczos_noalloc(c.h, 5, y.ctypes.data, START["random"], 0, aa, bb, bb2)
        assert rc == 7 and b"breakdown at step 1" in lib().bsl_last_error()
        assert abs(aa[0] - w[0]) < 1e-12 * abs(w[0])
