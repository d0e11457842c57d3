"""Pins of the CPU oracle (CPU only).

Each part of the oracle is checked against something that does not restate
its own formula: closed forms and library routines (numpy, scipy), brute
force on tiny problems, the reference's saved outputs and known answers, and
-- where the reference sources were mounted when the fixtures were made --
vectors produced by the reference's own compiled calculator."""
import json
import math
import os

import numpy as np
import pytest
import scipy.special as sp

import _cases as cs
import _oracle as O


# ---------------------------------------------------------------- basis ----

def test_lm_index_order():
    q = np.zeros(2 * 9, dtype=np.int32)
    assert O.lib.bso_gen_indices_lm(2, q) == 9
    assert [tuple(x) for x in q.reshape(9, 2)] == [(2, -2), (1, -1), (2, -1), (0, 0), (1, 0), (2, 0),
                                                   (1, 1), (2, 1), (2, 2)]


@pytest.mark.parametrize("n", [3, 10, 20, 40])
def test_gauss_legendre_matches_numpy(n):
    x, w = np.zeros(n), np.zeros(n)
    O.lib.bso_gauleg(-1.0, 1.0, x, w, n)
    xr, wr = np.polynomial.legendre.leggauss(n)
    assert np.max(np.abs(x - xr)) < 1e-14 and np.max(np.abs(w - wr)) < 1e-14


def test_normalised_legendre_matches_scipy():
    lmax = 6
    nn = (lmax + 1) ** 2
    q = np.zeros(2 * nn, dtype=np.int32)
    O.lib.bso_gen_indices_lm(lmax, q)
    q = q.reshape(nn, 2)
    out = np.zeros(nn)
    for x in (-0.93, -0.2, 0.0, 0.41, 0.999):
        O.lib.bso_norm_assoc_legendre(lmax, q.ravel().copy(), nn, x, out)
        for n, (l, m) in enumerate(q):
            am = abs(int(m))
            norm = math.sqrt((2 * l + 1) * math.factorial(l - am) / (2.0 * math.factorial(l + am)))
            ref = norm * sp.lpmv(am, l, x)  # scipy includes the Condon-Shortley phase (-1)^m
            assert abs(out[n] - ref) < 1e-13 * max(1.0, abs(ref)), (l, m, x)


def test_tesseral_harmonics_orthonormal_on_the_quadrature():
    """sum_{a,b} w_a w_b Y_n Y_n' over both phi halves = delta (the reference's
    tesseralTest invariant), lmax 5 on the 40x70 production quadrature."""
    spec, _ = cs.iho_case(3, 5, 40, 70, x_max=0.4)
    o = O.Oracle(spec)
    ct, wt, cp, wp = o.quad()
    L, S1, S2 = o.tess()
    qn = o.qnum()
    lmax = 5
    Y = []
    for n, (l, m) in enumerate(qn):
        y1 = np.outer(L[n], S1[m + lmax])
        y2 = np.outer(L[n], S2[m + lmax])
        Y.append(np.concatenate([y1.ravel(), y2.ravel()]))
    Y = np.array(Y)
    w = np.concatenate([np.outer(wt, wp).ravel()] * 2)
    G = (Y * w) @ Y.T
    assert np.max(np.abs(G - np.eye(len(qn)))) < 1e-13


def test_phi_nodes_are_the_uniform_midpoint_rule():
    spec, _ = cs.iho_case(3, 2, 4, 7, x_max=0.4)
    o = O.Oracle(spec)
    _, _, cp, wp = o.quad()
    phi = np.concatenate([np.arccos(cp), 2 * np.pi - np.arccos(cp)])
    assert np.allclose(np.sort(phi), np.pi * (2 * np.arange(14) + 1) / 14, atol=1e-14)
    assert np.allclose(wp, np.pi / 7)


def test_sinc_dvr_harmonic_oscillator_spectrum():
    """T + m w^2 x^2/2 on a wide 1-D box gives (n + 1/2) hbar w."""
    from bsl_amd import physics
    n, xmax, m = 40, 0.6, 2.016
    kin, grid = np.zeros(n * n), np.zeros(n)
    O.lib.bso_cart_kin_grid(xmax, n, m, kin, grid)
    kin = kin.reshape(n, n)
    assert np.array_equal(kin, kin.T)
    assert np.allclose(grid, np.linspace(-xmax / 2, xmax / 2, n), atol=1e-15)
    w = physics.omega_from_cm1(100.0)
    e = np.linalg.eigvalsh(kin + np.diag(0.5 * m * w * w * grid ** 2))
    hw = physics.H_BAR * w
    assert np.max(np.abs(e[:6] / hw - (np.arange(6) + 0.5))) < 1e-9


# ------------------------------------------------------------- potential ---

def test_point_energies_match_reference_build():
    """Site builders and pair sums against the reference's compiled
    getTIP4Patoms/Q_TIP4P_Eng and getSPCEatoms/Q_SPCE_Eng/LJ_SPCE_Eng_Fast
    (committed values).  The reference's test comment quotes -30.2331 kJ/mol
    for +1 e at (1.0824, 0.2164, 1.0819) nm in this unit cell
    (TIP4P_AH2_EngRoutines.cpp:238); its own code gives -58.3723 with the
    shipped geometry, and so does the oracle."""
    import bsl_amd.cage as cg
    fx = json.load(open(os.path.join(cs.GOLDEN, "point_energies.json")))
    cm_ea = cg.read_xyz(os.path.join(cs.GOLDEN, "sII_unitCelltest_CM_EA_ZYZ.xyz"))
    pos, q, typ = O.build_sites(cm_ea, 1)
    for p, e_ref in fx["tip4p_coulomb_unit_cell"]:
        e = O.lib.bso_coulomb_sum(np.array(p), 1.0, len(q), pos.ravel().copy(), q)
        assert e == e_ref
    pos, q, typ, lj = cs.cage11_sites()
    for p, e_cm, e_h in fx["spce_cage11"]:
        p = np.array(p)
        e = O.lib.bso_coulomb_sum(p, -0.9864, len(q), pos.ravel().copy(), q)
        e += O.lib.bso_lj_sum(p, len(q), pos.ravel().copy(), typ, 1.364e-06, 1.533e-3)
        assert e == e_cm
        assert O.lib.bso_coulomb_sum(p, 0.4932, len(q), pos.ravel().copy(), q) == e_h


def test_cage_grid_matches_direct_pair_sums():
    """V(P, theta_a, phi_b) recomputed with numpy from the site list."""
    sites = cs.cage11_sites()
    spec, _ = cs.cage_case(4, 1, 3, 4, "All", sites=sites)
    o = O.Oracle(spec)
    v1, v2 = o.vgrid()
    pos, q, typ, lj = sites
    g = o.grid(0)
    ct, wt, cp, wp = o.quad()
    rng = np.random.default_rng(0)
    K = 138.9354846111096
    for _ in range(20):
        i, j, k = rng.integers(0, 4, 3)
        a, b = rng.integers(0, 3), rng.integers(0, 4)
        c = np.array([g[i], g[j], g[k]])
        th, ph = np.arccos(ct[a]), np.arccos(cp[b])
        for half, phi in ((0, ph), (1, 2 * np.pi - ph)):
            e = 0.03707 * np.array([np.sin(th) * np.cos(phi), np.sin(th) * np.sin(phi), np.cos(th)])
            v = 0.0
            for h in (c + e, c - e):
                v += np.sum(K * 0.4932 * q / np.linalg.norm(pos - h, axis=1))
            r = np.linalg.norm(pos[lj == 1] - c, axis=1)
            v += np.sum(K * -0.9864 * q / np.linalg.norm(pos - c, axis=1)) + np.sum(1.364e-6 / r ** 12 - 1.533e-3 / r ** 6)
            v = min(v, 18.0)
            got = (v1 if half == 0 else v2)[i, j, k, a, b]
            assert abs(got - v) < 1e-12 * max(1, abs(v))


def test_vblocks_constant_potential_is_identity():
    """V = c everywhere (IHO clamped by a ceiling below it) gives c * delta."""
    spec, _ = cs.iho_case(4, 3, 6, 8, x_max=4.0, ceiling=3.25)
    o = O.Oracle(spec)
    vb = o.vblocks()
    eye = np.eye(16)
    assert np.max(np.abs(vb - 3.25 * eye)) < 1e-13


def test_vblocks_radial_potential_is_diagonal():
    spec, _ = cs.iho_case(5, 2, 6, 6, x_max=0.4)
    o = O.Oracle(spec)
    vb = o.vblocks()
    g = o.grid(0)
    r2 = g[:, None, None] ** 2 + g[None, :, None] ** 2 + g[None, None, :] ** 2
    expect = spec["iho_factor"] * r2
    assert np.max(np.abs(vb - expect[..., None, None] * np.eye(9))) < 1e-12 * expect.max()


def test_vblocks_cage_symmetric_and_parity_block_diagonal():
    spec, _ = cs.cage_case(4, 3, 8, 8, "All")
    o = O.Oracle(spec)
    vb = o.vblocks()
    scale = np.max(np.abs(vb))
    assert np.max(np.abs(vb - np.swapaxes(vb, -1, -2))) < 1e-13 * scale
    l = o.qnum()[:, 0]
    odd = (np.add.outer(l, l) % 2) == 1
    assert np.max(np.abs(vb[..., odd])) < 1e-12 * scale
    assert np.max(np.abs(vb[..., ~odd])) > 1e-3 * scale


# ------------------------------------------------------------------- Hv ----

def _dense_h(o, spec):
    """H assembled with numpy from the oracle's tables (kron products)."""
    nx, nn = spec["nx"], o.nn
    I = np.eye(nx)
    T1 = [o.kin(a) for a in range(3)]
    # vector index ((n*nz + k)*ny + j)*nx + i
    kin = np.kron(np.eye(nn), np.kron(T1[2], np.kron(I, I)) + np.kron(I, np.kron(T1[1], I)) + np.kron(I, np.kron(I, T1[0])))
    rot = np.kron(np.diag(o.rot()), np.eye(nx ** 3))
    vb = o.vblocks()
    V = np.zeros((o.N, o.N))
    P = nx ** 3
    for i in range(nx):
        for j in range(nx):
            for k in range(nx):
                p = (k * nx + j) * nx + i
                idx = np.arange(nn) * P + p
                V[np.ix_(idx, idx)] = vb[i, j, k]
    S = np.kron(np.diag(o.sym()), np.eye(P))
    return S @ (kin + rot + V)


@pytest.mark.parametrize("sym", ["All", "Even"])
def test_hv_matches_dense_operator(sym):
    spec, _ = cs.cage_case(3, 1, 3, 3, sym)
    o = O.Oracle(spec)
    H = _dense_h(o, spec)
    v = np.random.default_rng(2).uniform(-1, 1, o.N)
    assert np.max(np.abs(o.hv(v) - H @ v)) < 1e-12 * np.max(np.abs(H @ v))


def test_lanczos_matches_numpy_lanczos():
    spec, _ = cs.cage_case(3, 1, 3, 3, "All")
    o = O.Oracle(spec)
    H = _dense_h(o, spec)
    m = 20
    a, b, b2 = o.lanczos(m)
    v = np.full(o.N, 1 / np.sqrt(o.N))
    r = np.zeros(o.N)
    scale = np.abs(np.linalg.eigvalsh(0.5 * (H + H.T))).max()
    for j in range(m):
        r = r + H @ v
        al = v @ r
        r = r - al * v
        be = np.linalg.norm(r)
        assert abs(al - a[j]) < 1e-12 * scale and abs(be - b[j + 1]) < 1e-12 * scale, j
        if be < 1e-3 * scale:  # an invariant subspace is nearly exhausted: stop before rounding is amplified
            break
        r, v = -be * v, r / be
    assert j >= 5


def test_eigenvalues_match_dense_diagonalisation():
    spec, _ = cs.cage_case(3, 1, 3, 3, "All")
    o = O.Oracle(spec)
    H = _dense_h(o, spec)
    w = np.linalg.eigvalsh(0.5 * (H + H.T))
    ev, err = O.lancbis(*o.lanczos(100, O.start_vector(o.N, 1)), -10.0, 10.0)
    inwin = w[(w > -10) & (w <= 10)]
    for e in inwin:
        assert np.min(np.abs(ev - e)) < 1e-9 * abs(e)


# --------------------------------------------------- tridiagonal routines --

def test_bisec_spurious_test_by_brute_force():
    """good = eigenvalues of T_m that are not eigenvalues of T-hat (T without
    its first row and column), or that are multiple."""
    spec, _ = cs.cage_case(4, 1, 4, 4, "All")
    o = O.Oracle(spec)
    a, b, b2 = o.lanczos(150)
    vs, mp = O.bisec(a, b, -10.0, 10.0)
    T = np.diag(a) + np.diag(b[1:-1], 1) + np.diag(b[1:-1], -1)
    eT = np.linalg.eigvalsh(T)
    eH = np.linalg.eigvalsh(T[1:, 1:])
    # BISEC's multiplicity tolerance ep0 = (m + 1000) * 2 eps max|T_ij|
    ep0 = (len(a) + 1000) * 2 * 2.220446049250313e-16 * max(np.abs(a).max(), b.max())
    tol = ep0 / 5
    nspur = 0
    for v, m in zip(vs, mp):
        d_hat = np.min(np.abs(eH - v))
        if m == 0:
            assert d_hat < tol and np.sum(np.abs(eT - v) < tol) == 1
            nspur += 1
        elif m == 1:
            assert d_hat > tol
        else:
            assert np.sum(np.abs(eT - v) < ep0) == m
    assert nspur > 0


def test_error_estimates_match_exact_eigenvectors():
    spec, _ = cs.cage_case(4, 1, 4, 4, "All")
    o = O.Oracle(spec)
    m = 60
    a, b, b2 = o.lanczos(m)
    ev, err = O.lancbis(a, b, b2, -10.0, 10.0)
    T = np.diag(a) + np.diag(b[1:-1], 1) + np.diag(b[1:-1], -1)
    w, U = np.linalg.eigh(T)
    checked = 0
    for e, er in zip(ev, err):
        if er <= 0:
            continue
        k = np.argmin(np.abs(w - e))
        exact = b[m] * abs(U[-1, k])
        if exact > 1e-8:  # below that |y_m| is rounding noise in both
            assert abs(er - exact) < 1e-6 * exact
            checked += 1
        else:
            assert er < 1e-7
    assert checked > 0


def test_trivec_vectors_are_eigenvectors():
    spec, _ = cs.cage_case(4, 1, 4, 4, "All")
    o = O.Oracle(spec)
    m = 80
    a, b, b2 = o.lanczos(m)
    ev, err = O.lancbis(a, b, b2, -10.0, 10.0)
    ev = ev[:4]
    evtr, mamin = O.trivec(a, b, b2, ev)
    for g, e in enumerate(ev):
        ma = mamin[g]
        y = evtr[g * m: g * m + ma]
        T = np.diag(a[:ma]) + np.diag(b[1:ma], 1) + np.diag(b[1:ma], -1)
        assert abs(np.linalg.norm(y) - 1) < 1e-12
        assert np.linalg.norm(T @ y - e * y) < 1e-8


def test_ritz_vectors_are_eigenvectors_of_h():
    spec, _ = cs.cage_case(3, 1, 3, 3, "All")
    o = O.Oracle(spec)
    m = 60
    v0 = O.start_vector(o.N, 3)
    a, b, b2 = o.lanczos(m, v0)
    ev, err = O.lancbis(a, b, b2, -10.0, 10.0)
    ev = ev[:2]
    evtr, _ = O.trivec(a, b, b2, ev)
    Y = o.ritz(m, evtr, len(ev), v0)
    H = _dense_h(o, spec)
    for g, e in enumerate(ev):
        y = Y[g] / np.linalg.norm(Y[g])
        assert np.linalg.norm(H @ y - e * y) < 1e-6


def test_park_miller_stream():
    """GNU RAND: the minimal standard generator (seed 1 -> 16807, 282475249,
    1622650073), returned as (s - 1) with its low 9 bits cleared, over 2^31."""
    r = O.genran(1, 3)
    s = np.array([16807, 282475249, 1622650073])
    assert np.array_equal(r, ((s - 1) & ~511) / 2.0 ** 31)
    assert r[0] == 2.0 ** -17


def _tridiagonals():
    """Random tridiagonals (some with repeated diagonal halves) and the
    coefficients of two oracle Lanczos runs (the longer one has ghosts)."""
    out = []
    for seed in range(60):
        rng = np.random.default_rng(seed)
        m = int(5 + seed % 70)
        a = rng.standard_normal(m) * 3
        if seed % 3 == 0:
            a = np.concatenate([a[: m // 2], a[: m - m // 2]])
        b = np.concatenate([[0.0], np.abs(rng.standard_normal(m))])
        out.append((a, b, b * b, -5.0, 5.0))
    spec, _ = cs.cage_case(5, 2, 5, 5, "All")
    o = O.Oracle(spec)
    for m in (120, 700):
        a, b, b2 = o.lanczos(m)
        out.append((a, b, b2, -10.0, 10.0))
    return out


@pytest.mark.skipif(not O.ref_available(), reason="reference build (oracle/_ref) not present")
def test_genran_matches_reference_build():
    """GENRAN against the reference's genran.f built with the GNU runtime's
    SRAND / RAND: identical streams."""
    R = O.ref_lib()
    for seed in (0, 1, 7, 12345, 2 ** 30 + 3):
        r = np.zeros(5000)
        R.ref_genran(seed, r, 5000)
        assert np.array_equal(O.genran(seed, 5000), r), seed


@pytest.mark.skipif(not O.ref_available(), reason="reference build (oracle/_ref) not present")
def test_lancbis_matches_reference_build():
    """Good eigenvalues and INVERR error estimates against the reference's
    own lancbis.cpp with BISEC, ISOEV and INVERR built from its Fortran (GNU
    RAND, LAPACK DLAMCH): identical to the last bit."""
    R = O.ref_lib()
    n = 0
    for a, b, b2, lo, hi in _tridiagonals():
        try:
            e1, r1 = O.lancbis(a, b, b2, lo, hi)
        except RuntimeError:
            continue  # the reference STOPs the process there
        m = len(a)
        ev, er = np.zeros(m), np.zeros(m)
        ng = R.ref_lancbis(m, a, b, b2, lo, hi, ev, er, m)
        assert ng == len(e1)
        assert np.array_equal(ev[:ng], e1) and np.array_equal(er[:ng], r1)
        n += ng
    assert n > 300


@pytest.mark.skipif(not O.ref_available(), reason="reference build (oracle/_ref) not present")
def test_trivec_matches_reference_build():
    """T_m eigenvectors and the sizes MAMIN of the T_m used for them against
    the reference's trivec.f (with STURMI, INVERM, GENRAN) built from
    source: identical to the last bit."""
    R = O.ref_lib()
    n = 0
    for a, b, b2, lo, hi in _tridiagonals():
        try:
            ev, _ = O.lancbis(a, b, b2, lo, hi)
            if len(ev) == 0:
                continue
            v1, m1 = O.trivec(a, b, b2, ev)
        except RuntimeError:
            continue
        m, ng = len(a), len(ev)
        evtr, mamin = np.zeros(m * ng), np.zeros(ng)
        R.ref_trivec(m, a, b, b2, ng, ev, evtr, mamin)
        # MAMIN is an INTEGER array in trivec.f; the reference passes a
        # double buffer, so its first ng 4-byte words hold the values
        assert np.array_equal(evtr, v1)
        assert np.array_equal(mamin.view(np.int32)[:ng], m1)
        n += 1
    assert n > 30


# --------------------------------------------------- reference's outputs ---

def _run_case(name):
    run = json.load(open(os.path.join(cs.GOLDEN, "runs.json")))[name]
    hv = run["hv_deck"]
    spec, _ = cs.cage_case(int(hv[0]), int(hv[6]), int(hv[7]), int(hv[8]), hv[21], x_max=float(hv[1]),
                           mass=float(hv[15]), rot=float(hv[16]), ceiling=float(hv[19]))
    return run, spec


def test_saved_run_300_steps():
    """Converged levels of Compilation/SmallBasisTest2_2013-09-03_19_06_47."""
    run, spec = _run_case("SmallBasisTest2_2013-09-03_19_06_47")
    o = O.Oracle(spec)
    ev, err = O.lancbis(*o.lanczos(int(run["lanczos_deck"][0])), -10.0, 10.0)
    ref = np.array(run["boundstates"])
    assert np.array_equal(ev[:5], ref[:5, 0])  # bit for bit
    assert np.max(np.abs(ev[:11] - ref[:11, 0]) / np.abs(ref[:11, 0])) < 1e-12


def test_saved_run_100_steps_ground_state():
    run, spec = _run_case("SmallBasisTest2_2013-09-03_18_36_00")
    o = O.Oracle(spec)
    ev, err = O.lancbis(*o.lanczos(int(run["lanczos_deck"][0])), -10.0, 10.0)
    ref = np.array(run["boundstates"])
    assert ev[0] == ref[0, 0]
    assert abs(ev[1] - ref[1, 0]) < 1e-10 * abs(ref[1, 0])


def _grid_run_case(name):
    run, spec = _run_case(name)
    hv = run["hv_deck"]
    spec = dict(spec, pot_kind=2, universe=(int(hv[9]), int(hv[11]), int(hv[13]), float(hv[10]), float(hv[12]),
                                            float(hv[14])))
    return run, spec


@pytest.mark.parametrize("name,nlev", [("SmallBasisTest2_2013-08-29_19_35_47", 3),
                                       ("SmallBasisTest2_2013-08-29_19_42_01", 8)])
def test_saved_grid_mapped_runs(name, nlev):
    """The grid-mapped calculator (linRotCartSph_NoQuad_Alavi_SPCE: centre and
    H potentials on a 20^3 universe, nearest-point lookups) reproduces the
    reference's saved runs (Compilation/<name>/boundstates.txt)."""
    run, spec = _grid_run_case(name)
    assert run["lanczos_deck"][5] == "linRotCartSph_NoQuad_Alavi_SPCE"
    o = O.Oracle(spec)
    ev, err = O.lancbis(*o.lanczos(int(run["lanczos_deck"][0])), -10.0, 10.0)
    ref = np.array(run["boundstates"])
    assert np.max(np.abs(ev[:nlev] - ref[:nlev, 0]) / np.abs(ref[:nlev, 0])) < 1e-11


def test_grid_mapped_lookups_recomputed():
    """Nearest-universe-point lookups recomputed with numpy (universe sizes
    and lengths differ per axis, so an axis mix-up or a wrong rounding rule
    fails), with pair sums at the universe points."""
    sites = cs.cage11_sites()
    pos, q, typ, lj = sites
    spec, _ = cs.cage_case(4, 1, 3, 4, "All", sites=sites)
    un, um = (7, 6, 5), (0.52, 0.61, 0.47)
    spec = dict(spec, pot_kind=2, universe=un + um)
    o = O.Oracle(spec)
    v1, v2 = o.vgrid()
    g = o.grid(0)
    ct, wt, cp, wp = o.quad()
    K = 138.9354846111096
    du = [um[d] / (un[d] - 1) for d in range(3)]

    def idx(c, d):
        x = (c - (-0.5 * um[d])) / du[d]
        i = int(math.floor(x + 0.5)) if x >= 0 else -int(math.floor(-x + 0.5))  # half away from zero
        return min(max(i, 0), un[d] - 1)

    def upoint(ix):
        return np.array([ix[d] * du[d] - um[d] / 2.0 for d in range(3)])

    rng = np.random.default_rng(1)
    for _ in range(20):
        i, j, k = rng.integers(0, 4, 3)
        a, b = rng.integers(0, 3), rng.integers(0, 4)
        c = np.array([g[i], g[j], g[k]])
        th, ph = np.arccos(ct[a]), np.arccos(cp[b])
        for half, phi in ((0, ph), (1, 2 * np.pi - ph)):
            e = np.array([np.sin(th) * np.cos(phi), np.sin(th) * np.sin(phi), np.cos(th)])
            u0 = upoint([idx(c[d], d) for d in range(3)])
            r = np.linalg.norm(pos[lj == 1] - u0, axis=1)
            v = np.sum(K * -0.9864 * q / np.linalg.norm(pos - u0, axis=1)) + np.sum(1.364e-6 / r ** 12 - 1.533e-3 / r ** 6)
            for sgn in (1.0, -1.0):
                uh = upoint([idx(c[d] + e[d] * sgn * 0.03707, d) for d in range(3)])
                v += np.sum(K * 0.4932 * q / np.linalg.norm(pos - uh, axis=1))
            v = min(v, 18.0)
            got = (v1 if half == 0 else v2)[i, j, k, a, b]
            assert abs(got - v) < 1e-12 * max(1, abs(v))


def test_coulomb_potential_closed_form():
    """linRotCartSph_NoQuad_Coulomb: V = K q1 q2 / |r| capped at |V| <= |cap|,
    orientation independent, so the V blocks are V(r) times the identity."""
    from bsl_amd import physics
    spec, _ = cs.iho_case(4, 2, 6, 6, x_max=0.4)
    spec = dict(spec, pot_kind=3, coulomb=(1.0, -0.5, -200.0), ceiling=18.0)
    o = O.Oracle(spec)
    v1, v2 = o.vgrid()
    g = o.grid(0)
    for i, x in enumerate(g):
        for j, y in enumerate(o.grid(1)):
            for k, z in enumerate(o.grid(2)):
                v = physics.K_COULOMB * 1.0 * -0.5 / np.sqrt(x * x + y * y + z * z)
                v = -200.0 if abs(v) > 200.0 else v
                assert np.all(v1[i, j, k] == v1[i, j, k].flat[0])
                assert abs(v1[i, j, k].flat[0] - min(v, 18.0)) <= 1e-13 * abs(v)
    vb = o.vblocks()
    nn = vb.shape[-1]
    off = vb - np.einsum("ijkn,nm->ijknm", np.einsum("ijknn->ijkn", vb), np.eye(nn))
    assert np.max(np.abs(off)) < 1e-12 * np.max(np.abs(vb))


def test_reference_calculator_vectors():
    """V blocks and Hv of the reference's own compiled NoQuad ContGrid
    calculator (committed fixture): identical to the last bit."""
    fx = np.load(os.path.join(cs.GOLDEN, "ref_hv_cage11.npz"))
    for tag in ("a", "b"):
        nx, lmax, na, nb, sym = [int(x) for x in fx["%s_params" % tag]]
        spec, _ = cs.cage_case(nx, lmax, na, nb, ["All", "Even", "Odd"][sym])
        o = O.Oracle(spec)
        assert np.array_equal(o.vblocks().ravel(), fx["%s_vblocks" % tag])
        assert np.array_equal(o.hv(fx["%s_v" % tag]), fx["%s_hv" % tag])


@pytest.mark.skipif(not O.ref_available(), reason="reference build (oracle/_ref) not present")
@pytest.mark.parametrize("case", [(4, 2, 5, 5, "Odd"), (5, 3, 7, 9, "All"), (3, 4, 6, 4, "Even")])
def test_reference_build_bitwise(tmp_path, case):
    """The oracle against the reference's own compiled calculator, fresh decks."""
    import sys
    sys.path.insert(0, cs.GOLDEN)
    from make_fixtures import write_hv_deck
    nx, lmax, na, nb, sym = case
    geo = os.path.join(cs.GOLDEN, "sII_smallCage11_CM_EA_ZYZ.xyz")
    deck = str(tmp_path / "hv.txt")
    write_hv_deck(deck, nx, 0.4, lmax, na, nb, geo, sym)
    R = O.ref_lib()
    N = R.ref_setup(b"linRotCartSph_NoQuad_ContGrid_Alavi_SPCE", deck.encode(), None)
    spec, _ = cs.cage_case(nx, lmax, na, nb, sym)
    o = O.Oracle(spec)
    assert N == o.N
    vb = np.zeros(o.vblocks().size)
    R.ref_vblocks(vb, vb.size)
    assert np.array_equal(vb, o.vblocks().ravel())
    v = np.random.default_rng(nx).uniform(-1, 1, N)
    u = np.zeros(N)
    R.ref_hv(v, u)
    assert np.array_equal(u, o.hv(v))


@pytest.mark.skipif(not O.ref_available(), reason="reference build (oracle/_ref) not present")
def test_reference_hv_with_injected_blocks(tmp_path):
    """The reference's Hv on blocks set from outside (ref_set_vblocks, which
    the CPU baseline of bench.py uses to time the reference on the GPU
    context's blocks): a reference prepared on a 1 x 1 quadrature, with the
    oracle's blocks of a 6 x 7 quadrature set in, gives the oracle's Hv of
    the 6 x 7 deck bitwise, and differs from it without them."""
    nx, lmax, na, nb, sym = 5, 3, 6, 7, "All"
    geo = os.path.join(cs.GOLDEN, "sII_smallCage11_CM_EA_ZYZ.xyz")
    deck = str(tmp_path / "hv.txt")
    O.write_ref_hv_deck(deck, nx, 0.4, lmax, 1, 1, cs.MASS, cs.B_ROT, geo, 18.0, sym)
    R = O.ref_lib()
    N = R.ref_setup(b"linRotCartSph_NoQuad_ContGrid_Alavi_SPCE", deck.encode(), None)
    spec, _ = cs.cage_case(nx, lmax, na, nb, sym)
    o = O.Oracle(spec)
    assert N == o.N
    v = np.random.default_rng(17).uniform(-1, 1, N)
    u1 = np.zeros(N)
    R.ref_hv(v, u1)
    ref = o.hv(v)
    assert not np.array_equal(u1, ref)
    vb = np.ascontiguousarray(o.vblocks().ravel())
    assert R.ref_set_vblocks(vb, vb.size) == 0
    u = np.zeros(N)
    R.ref_hv(v, u)
    assert np.array_equal(u, ref)


@pytest.mark.skipif(not os.path.exists(O.REF_LANCZOS), reason="reference program (oracle/_ref) not built")
def test_reference_program_loop1_and_levels(tmp_path):
    """The reference's own Lanczos program (Loop 1 from the uniform start,
    lancbis, output files) on a 10^3 x 16 cage deck against the oracle:
    every alpha and beta to its 16 printed digits, every good eigenvalue and
    error estimate of boundstates.txt to its 16 printed digits."""
    import test_driver as td
    t = str(tmp_path)
    ld = td.lanczos_deck(t, niter=300, calc="linRotCartSph_NoQuad_ContGrid_Alavi_SPCE")
    hd = td.hv_deck(t)
    a_ref, b_ref, d = O.run_ref_lanczos(ld, hd, t)
    spec, _ = cs.cage_case(10, 3, 10, 10, "All")
    o = O.Oracle(spec)
    a, b, b2 = o.lanczos(300)
    assert len(a_ref) == 300
    assert np.max(np.abs(a_ref - a) / np.abs(a)) < 1e-15
    assert np.max(np.abs(b_ref - b[1:]) / np.abs(b[1:])) < 1e-15
    ev, er = O.lancbis(a, b, b2, -10.0, 10.0)
    rows = np.loadtxt(os.path.join(t, d, "boundstates.txt"), skiprows=4)
    assert rows.shape == (len(ev), 3)
    assert np.max(np.abs(rows[:, 1] - ev) / np.abs(ev)) < 1e-15
    assert np.all(np.abs(rows[:, 2] - er) <= 1e-15 * np.abs(er) + 1e-300)


@pytest.mark.skipif(not os.path.exists(O.REF_LANCZOS), reason="reference program (oracle/_ref) not built")
def test_reference_program_iho_workload(tmp_path):
    """The plain-FP64 isotropic oscillator workload (21^3, l <= 3, 20 x 30,
    300 steps) through the reference's own program against the oracle:
    alpha, beta and the good levels to their 16 printed digits."""
    import test_driver as td
    t = str(tmp_path)
    ld = td.lanczos_deck(t, niter=300, calc="linRotCartSph_NoQuad_IsoHarmOsc", word="none", nvec=0)
    hv, pot = td.iho_decks(t)
    a_ref, b_ref, d = O.run_ref_lanczos(ld, hv, t, pot)
    from bsl_amd import physics
    spec, _ = cs.iho_case(21, 3, 20, 30, x_max=0.4, rot=physics.cm1_to_kjmol(59.322))
    o = O.Oracle(spec)
    a, b, b2 = o.lanczos(300)
    assert np.max(np.abs(a_ref - a) / np.abs(a)) < 1e-15
    assert np.max(np.abs(b_ref - b[1:]) / np.abs(b[1:])) < 1e-15
    ev, _ = O.lancbis(a, b, b2, -10.0, 10.0)
    rows = np.loadtxt(os.path.join(t, d, "boundstates.txt"), skiprows=4)
    assert rows.shape[0] == len(ev) > 10
    assert np.max(np.abs(rows[:, 1] - ev) / np.abs(ev)) < 1e-15
