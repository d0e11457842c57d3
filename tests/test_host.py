"""CPU tests of the product library's host side: the C ABI loads and exports
every symbol of include/bsl_hip.h, argument validation, and the
Cullum-Willoughby post-processing (host code) against the oracle."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import _cases as cs
import _oracle as O

import bsl_amd as B  # noqa: E402  (a missing library fails the tests, it never skips them)
from bsl_amd import _lib  # noqa: E402

HEADER = os.path.join(os.path.dirname(cs.HERE), "include", "bsl_hip.h")


def test_exports_every_declared_symbol():
    decl = set(re.findall(r"\b(bsl_[a-z0-9_]+)\s*\(", open(HEADER).read()))
    lib = C.CDLL(_lib.LIB_PATH)
    missing = [s for s in sorted(decl) if not hasattr(lib, s)]
    assert not missing, missing
    assert set(_lib.EXPORTS) == decl


def test_create_rejects_bad_parameters():
    _, kw = cs.iho_case(9, 2, 6, 8, x_max=0.4)
    for bad in (dict(nx=1), dict(nx=257), dict(lmax=13), dict(symmetry="Even", basis="full", x_max=-1.0)):
        k = dict(kw)
        k.update(bad)
        with pytest.raises(B.BslError) as ei:
            B.Calculator(**k)
        assert ei.value.code == 1
    with pytest.raises(KeyError):
        B.Calculator(**dict(kw, symmetry="Both"))


def _lanczos_like(seed, m):
    rng = np.random.default_rng(seed)
    a = rng.standard_normal(m) * 3
    if seed % 3 == 0:
        a = np.concatenate([a[: m // 2], a[: m - m // 2]])  # repeated values
    b = np.concatenate([[0.0], np.abs(rng.standard_normal(m))])
    return a, b, b * b


def test_lancbis_identical_to_oracle():
    for seed in range(150):
        a, b, b2 = _lanczos_like(seed, int(5 + seed % 70))
        e1, r1 = O.lancbis(a, b, b2, -5.0, 5.0)
        e2, r2 = B.lancbis(a, b, b2, -5.0, 5.0)
        assert np.array_equal(e1, e2) and np.array_equal(r1, r2), seed


def test_trivec_identical_to_oracle():
    done = 0
    for seed in range(120):
        a, b, b2 = _lanczos_like(seed, int(6 + seed % 60))
        ev, _ = O.lancbis(a, b, b2, -5.0, 5.0)
        if len(ev) == 0:
            continue
        try:
            v1, m1 = O.trivec(a, b, b2, ev)
        except RuntimeError:
            with pytest.raises(B.BslError):
                B.trivec(a, b, b2, ev)
            continue
        v2, m2 = B.trivec(a, b, b2, ev)
        assert np.array_equal(v1, v2) and np.array_equal(m1, m2), seed
        done += 1
    assert done > 50


def test_lancbis_on_a_real_lanczos_run_identical():
    spec, _ = cs.cage_case(5, 2, 5, 5, "All")
    o = O.Oracle(spec)
    a, b, b2 = o.lanczos(120)
    e1, r1 = O.lancbis(a, b, b2, -10.0, 10.0)
    e2, r2 = B.lancbis(a, b, b2, -10.0, 10.0)
    assert len(e1) > 5
    assert np.array_equal(e1, e2) and np.array_equal(r1, r2)


def test_lancbis_tol_zero_is_lancbis():
    """bsl_lancbis_tol with mult_tol = 0 is the reference's BISEC (bitwise
    the oracle's restatement)."""
    spec, _ = cs.cage_case(5, 2, 5, 5, "All")
    a, b, b2 = O.Oracle(spec).lanczos(300)
    e1, r1 = O.lancbis(a, b, b2, -10.0, 10.0)
    e2, r2 = B.lancbis(a, b, b2, -10.0, 10.0, mult_tol=0.0)
    assert np.array_equal(e1, e2) and np.array_equal(r1, r2)
    from bsl_amd._lib import lib
    import ctypes as C
    ng = C.c_int32()
    ev, er = np.zeros(len(a)), np.zeros(len(a))
    assert lib().bsl_lancbis_tol(len(a), a, b, b2, -10.0, 10.0, -1.0, ev, er, C.byref(ng)) != 0


def _dense_operator(spec):
    o = O.Oracle(spec)
    h = np.zeros((o.N, o.N))
    e = np.zeros(o.N)
    for j in range(o.N):
        e[j] = 1.0
        h[:, j] = o.hv(e)
        e[j] = 0.0
    return 0.5 * (h + h.T)


def _noisy_lanczos(h, niter, noise, seed):
    """Plain Lanczos whose Hv carries a relative rounding of `noise` per
    application (what a reduced-precision operator does), FP64 otherwise."""
    rng = np.random.default_rng(seed)
    n = h.shape[0]
    v, vp, beta = np.ones(n) / np.sqrt(n), np.zeros(n), 0.0
    a, b = np.zeros(niter), np.zeros(niter + 1)
    for j in range(niter):
        u = h @ v
        u = u + noise * np.abs(u) * rng.standard_normal(n)
        r = u - beta * vp
        a[j] = v @ r
        r -= a[j] * v
        beta = np.linalg.norm(r)
        b[j + 1] = beta
        vp, v = v, r / beta
    return a, b


def test_lancbis_tol_merges_noisy_ghosts():
    """With a noisy operator (relative 1e-7 per Hv) the ghost copies of the
    converged levels are spread beyond BISEC's FP64 width: mult_tol = 0
    reports them as extra good levels; mult_tol = 1e-6 (the mixed
    calculators' tolerance) returns exactly the operator's eigenvalues in the
    window, each within 1e-5."""
    h = _dense_operator(cs.cage_case(4, 2, 6, 6, "All")[0])
    w = np.linalg.eigvalsh(h)
    inside = w[(w > -10.0) & (w <= 10.0)]
    a, b = _noisy_lanczos(h, 3 * h.shape[0], 1e-7, 0)
    e0, _ = B.lancbis(a, b, b * b, -10.0, 10.0)
    assert len(e0) > 2 * len(inside)  # the failure mode, for the record
    e1, _ = B.lancbis(a, b, b * b, -10.0, 10.0, mult_tol=B.MIXED_MULT_TOL)
    assert len(e1) == len(inside)
    assert np.max(np.abs(np.sort(e1) - inside)) < 1e-5


def test_trivec_tol_finds_merged_levels():
    """TRIVEC for the merged values of a noisy operator (relative 3e-8 per
    Hv: the copies of a level spread well inside the mixed tolerance).  With
    STURMI's FP64 window some values are found only at orders where several
    copies coexist and the inverse iteration never converges; with the same
    mult_tol as the eigensolve every value gets a T_MA eigenvector (shifted
    by T_MA's own copy), with a residual at rounding level and a Rayleigh
    quotient within the tolerance of the value."""
    h = _dense_operator(cs.cage_case(4, 2, 6, 6, "All")[0])
    failed = 0
    for seed in range(4):
        a, b = _noisy_lanczos(h, 3 * h.shape[0], 3e-8, seed)
        x, _ = B.lancbis(a, b, b * b, -10.0, 10.0, mult_tol=B.MIXED_MULT_TOL)
        x = x[:30]
        for v in x:
            try:
                B.trivec(a, b, b * b, np.array([v]))
            except B.BslError:
                failed += 1
        evtr, mam = B.trivec(a, b, b * b, x, mult_tol=B.MIXED_MULT_TOL)
        m = len(a)
        for g in range(len(x)):
            k = mam[g]
            assert 0 < k <= m
            y = evtr[g * m: g * m + k]
            ty = a[:k] * y
            ty[:-1] += b[1:k] * y[1:]
            ty[1:] += b[1:k] * y[:-1]
            lam = y @ ty / (y @ y)
            assert np.linalg.norm(ty - lam * y) / np.linalg.norm(y) < 1e-10 * np.max(np.abs(a))
            assert abs(lam - x[g]) < 2 * B.MIXED_MULT_TOL
    assert failed > 0  # the failure mode, for the record


@pytest.mark.parametrize("model", [0, 1])
def test_water_sites_identical_to_oracle(model):
    """SPC/E (getSPCEatoms, SPCE_AH2_EngRoutines.cpp:5-108) and TIP4P
    (getTIP4Patoms, TIP4P_AH2_EngRoutines.cpp) sites from the same CM/Euler
    rows, bitwise the oracle's."""
    cm_ea = cs.cage11_cm_ea()
    pos_o, q_o, typ = O.build_sites(cm_ea, model)
    pos, q, lj = B.cage.sites_from_cm_ea(cm_ea, model)
    assert len(q) == len(cm_ea) * (3 if model == 0 else 4)
    assert np.array_equal(pos, pos_o) and np.array_equal(q, q_o)
    assert list(lj) == [1 if t == ord("O") else 0 for t in typ]


def test_synthetic_cage_geometry():
    pos, q, lj = B.cage.synthetic_cage(0)
    assert pos.shape == (60, 3) and abs(q.sum()) < 1e-12 and lj.sum() == 20
    o = pos[0::3]
    r = np.linalg.norm(o, axis=1)
    assert np.allclose(r, 0.395)
    for i in range(20):
        h1, h2 = pos[3 * i + 1] - o[i], pos[3 * i + 2] - o[i]
        assert abs(np.linalg.norm(h1) - 0.1) < 1e-12 and abs(np.linalg.norm(h2) - 0.1) < 1e-12
        ang = np.degrees(np.arccos(np.dot(h1, h2) / 0.01))
        assert abs(ang - 109.47) < 1e-8
        # the first H points along a cage edge (towards another O)
        d = o - o[i]
        cos = d @ h1 / (np.linalg.norm(d, axis=1) + 1e-30) / 0.1
        assert np.max(cos) > 1 - 1e-12
    # seeded: different seeds tilt differently, same seed reproduces
    assert np.array_equal(B.cage.synthetic_cage(3)[0], B.cage.synthetic_cage(3)[0])
    assert not np.array_equal(B.cage.synthetic_cage(3)[0], B.cage.synthetic_cage(4)[0])
