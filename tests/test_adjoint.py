"""Hellmann-Feynman gradients of eigenvalues with respect to the cage
potential's parameters (site charges and positions, Alavi charges, LJ A/B).

CPU: the oracle's gradient (bso_hf_gradient) against central finite
differences of dense-diagonalisation eigenvalues of the oracle operator.
GPU: the library's adjoint kernels against the oracle on the same vectors,
against central finite differences of the GPU's own converged levels, and
through the torch.autograd wrapper (chain rule to water-model charges and
LJ sigma/epsilon)."""
import copy

import numpy as np
import pytest

import _cases as cs
import _oracle as O


def _dense(o):
    n = o.N
    h = np.zeros((n, n))
    e = np.zeros(n)
    for j in range(n):
        e[j] = 1.0
        h[:, j] = o.hv(e)
        e[j] = 0.0
    return 0.5 * (h + h.T)


def _perturb(spec, name, idx, delta):
    s = copy.deepcopy(spec)
    if name in ("site_q", "site_pos"):
        arr = np.array(s[name], dtype=np.float64, copy=True)
        arr.reshape(-1)[idx] += delta
        s[name] = arr
    else:
        s[name] = s[name] + delta
    return s


def _value(spec, name, idx):
    return np.asarray(spec[name], dtype=np.float64).reshape(-1)[idx] if name in ("site_q", "site_pos") \
        else spec[name]


def _grad_entry(g, ns, name, idx):
    if name == "site_q":
        return g[4 * idx]
    if name == "site_pos":
        s, d = divmod(idx, 3)
        return g[4 * s + 1 + d]
    return g[4 * ns + ["q_h2_h", "q_h2_cm", "lj_a", "lj_b"].index(name)]


PARAMS = [("site_q", 0), ("site_q", 1), ("site_pos", 0), ("site_pos", 5), ("site_pos", 14), ("q_h2_h", 0),
          ("q_h2_cm", 0), ("lj_a", 0), ("lj_b", 0)]


def test_oracle_gradient_matches_finite_differences():
    """The oracle's Hellmann-Feynman gradient of the two lowest levels
    against central differences (h = 1e-4 relative) of dense eigenvalues."""
    spec, _ = cs.cage_case(4, 2, 6, 6, "All")
    o = O.Oracle(spec)
    w, v = np.linalg.eigh(_dense(o))
    assert w[1] - w[0] > 1e-3 and w[2] - w[1] > 1e-3
    ns = len(spec["site_q"])
    g0 = o.hf_gradient(v[:, 0], [1.0])
    g1 = o.hf_gradient(v[:, 1], [1.0])
    g01 = o.hf_gradient(np.stack([v[:, 0], v[:, 1]]), [1.0, -2.0])
    assert np.allclose(g01, g0 - 2.0 * g1, rtol=1e-12, atol=1e-12 * np.max(np.abs(g0)))
    for name, idx in PARAMS:
        x = _value(spec, name, idx)
        h = 1e-4 * abs(x)
        wp = np.linalg.eigvalsh(_dense(O.Oracle(_perturb(spec, name, idx, h))))
        wm = np.linalg.eigvalsh(_dense(O.Oracle(_perturb(spec, name, idx, -h))))
        for lev, g in ((0, g0), (1, g1)):
            fd = (wp[lev] - wm[lev]) / (2 * h)
            an = _grad_entry(g, ns, name, idx)
            assert abs(an - fd) <= 1e-5 * max(abs(fd), 1e-3 * abs(x) ** -1), (name, idx, lev, an, fd)


@pytest.mark.gpu
def test_gpu_gradient_matches_oracle():
    """Same quadratic form on the same (random, unit) vectors and weights."""
    import bsl_amd as B
    spec, kw = cs.cage_case(5, 3, 8, 9, "All")
    o = O.Oracle(spec)
    c = B.Calculator(**kw)
    rng = np.random.default_rng(31)
    psi = rng.uniform(-1, 1, (2, o.N))
    psi /= np.linalg.norm(psi, axis=1, keepdims=True)
    g = np.array([0.7, -0.3])
    ref = o.hf_gradient(psi, g)
    ns = len(spec["site_q"])
    got = c.hf_gradient(psi, g)
    flat = np.zeros(4 * ns + 4)
    flat[0:4 * ns:4] = got["site_q"]
    flat[:4 * ns].reshape(ns, 4)[:, 1:4] = got["site_pos"]
    flat[4 * ns:] = [got["q_h2_h"], got["q_h2_cm"], got["lj_a"], got["lj_b"]]
    scale = np.max(np.abs(ref[:4 * ns]))
    assert np.max(np.abs(flat[:4 * ns] - ref[:4 * ns])) / scale < 1e-10
    for e in range(4 * ns, 4 * ns + 4):
        assert abs(flat[e] - ref[e]) <= 1e-10 * abs(ref[e])
    # the clamp mask is found by the first call and kept: a second call with
    # other vectors and weights, on a grid with clamped nodes, still matches
    g1, g2 = c.vgrid()
    assert np.count_nonzero(g1 >= kw["ceiling"]) + np.count_nonzero(g2 >= kw["ceiling"]) > 0
    psi2 = rng.uniform(-1, 1, (3, o.N))
    psi2 /= np.linalg.norm(psi2, axis=1, keepdims=True)
    g_2 = np.array([0.2, 1.1, -0.5])
    ref2 = o.hf_gradient(psi2, g_2)
    got2 = c.hf_gradient(psi2, g_2)
    assert np.max(np.abs(got2["site_q"] - ref2[0:4 * ns:4])) / np.max(np.abs(ref2[:4 * ns])) < 1e-10
    assert np.max(np.abs(got2["site_pos"] - ref2[:4 * ns].reshape(ns, 4)[:, 1:4])) / np.max(np.abs(ref2[:4 * ns])) < 1e-10
    for e, name in zip(range(4 * ns, 4 * ns + 4), ("q_h2_h", "q_h2_cm", "lj_a", "lj_b")):
        assert abs(got2[name] - ref2[e]) <= 1e-10 * abs(ref2[e])


@pytest.mark.gpu
def test_gpu_gradient_sets_equal_single_sets():
    """bsl_hf_gradient_sets over six sets of one to three vectors (two passes
    of up to four sets, the second ragged) gives bitwise the gradients of
    bsl_hf_gradient on each set alone, and the oracle's within 1e-10."""
    import bsl_amd as B
    spec, kw = cs.cage_case(5, 3, 8, 9, "All")
    o = O.Oracle(spec)
    ns = len(spec["site_q"])
    rng = np.random.default_rng(77)
    sets = []
    for k in (1, 2, 3, 1, 2, 1):
        psi = rng.uniform(-1, 1, (k, o.N))
        psi /= np.linalg.norm(psi, axis=1, keepdims=True)
        sets.append((psi, rng.uniform(-1, 1, k)))
    with B.Calculator(**kw) as c:
        single = [c.hf_gradient(p, g) for p, g in sets]
    names = ("site_q", "site_pos", "q_h2_h", "q_h2_cm", "lj_a", "lj_b")
    # a fresh context too: the clamp mask is found by the first call of either entry point
    for fresh in (False, True):
        with B.Calculator(**kw) as c:
            if not fresh:
                c.hf_gradient(*sets[0])
            got = c.hf_gradient_sets(sets)
        assert len(got) == len(sets)
        for gs, g1 in zip(got, single):
            for name in names:
                assert np.array_equal(np.asarray(gs[name]), np.asarray(g1[name])), (fresh, name)
    for (psi, g), gs in zip(sets, got):
        ref = o.hf_gradient(psi, g)
        scale = np.max(np.abs(ref[:4 * ns]))
        assert np.max(np.abs(gs["site_q"] - ref[0:4 * ns:4])) / scale < 1e-10
        assert np.max(np.abs(gs["site_pos"] - ref[:4 * ns].reshape(ns, 4)[:, 1:4])) / scale < 1e-10
        for e, name in zip(range(4 * ns, 4 * ns + 4), names[2:]):
            assert abs(gs[name] - ref[e]) <= 1e-10 * abs(ref[e])


@pytest.mark.gpu
def test_gpu_gradient_sets_split_by_vector_count():
    """At lmax 5 (36 carried planes) a density pass holds at most 56
    eigenvectors: three sets of 20 run as two passes (40, then 20), and a
    set of 60 alone is refused with an error, not launched; every set's
    gradient is bitwise the single-set one."""
    import bsl_amd as B
    spec, kw = cs.cage_case(5, 5, 8, 9, "All")
    rng = np.random.default_rng(9)
    with B.Calculator(**kw) as c:
        n = c.N
        sets = []
        for _ in range(3):
            psi = rng.uniform(-1, 1, (20, n))
            psi /= np.linalg.norm(psi, axis=1, keepdims=True)
            sets.append((psi, rng.uniform(-1, 1, 20)))
        got = c.hf_gradient_sets(sets)
        single = [c.hf_gradient(p, g) for p, g in sets]
        for a, b in zip(got, single):
            for name in a:
                assert np.array_equal(np.asarray(a[name]), np.asarray(b[name])), name
        big = rng.uniform(-1, 1, (60, n))
        with pytest.raises(B.BslError):
            c.hf_gradient(big, np.ones(60))


@pytest.mark.gpu
def test_gpu_gradient_reduced_basis_matches_oracle():
    """Reduced-basis contexts (para: even l, ortho: odd l carried): vectors
    that vanish on the other planes, gradients against the oracle's of the
    same full-length vectors within 1e-10."""
    import bsl_amd as B
    for sym, par in (("Even", 0), ("Odd", 1)):
        spec, kw = cs.cage_case(5, 3, 8, 9, sym)
        o = O.Oracle(spec)
        ns = len(spec["site_q"])
        ls = o.qnum()[:, 0]
        rng = np.random.default_rng(13 + par)
        psi = rng.uniform(-1, 1, (2, o.N))
        for n, l in enumerate(ls):
            if l % 2 != par:
                psi[:, n * o.P:(n + 1) * o.P] = 0.0
        psi /= np.linalg.norm(psi, axis=1, keepdims=True)
        g = np.array([0.6, -1.3])
        ref = o.hf_gradient(psi, g)
        with B.Calculator(basis="reduced", **kw) as c:
            got = c.hf_gradient(psi, g)
        scale = np.max(np.abs(ref[:4 * ns]))
        assert np.max(np.abs(got["site_q"] - ref[0:4 * ns:4])) / scale < 1e-10, sym
        assert np.max(np.abs(got["site_pos"] - ref[:4 * ns].reshape(ns, 4)[:, 1:4])) / scale < 1e-10, sym
        for e, name in zip(range(4 * ns, 4 * ns + 4), ("q_h2_h", "q_h2_cm", "lj_a", "lj_b")):
            assert abs(got[name] - ref[e]) <= 1e-10 * abs(ref[e]), (sym, name)


@pytest.mark.gpu
def test_gpu_clamp_mask_matches_grid_and_oracle():
    """The gradients' clamp mask (a pass of the grid kernel keeping one bit
    per (point, node)) counts exactly the nodes of the context's own clamped
    grid below the ceiling, and those of the oracle's grid, on two grids
    whose P Q nodes end in a part-filled word."""
    import bsl_amd as B
    for n, na, nb in ((5, 8, 9), (6, 7, 5)):
        spec, kw = cs.cage_case(n, 3, na, nb, "All")
        with B.Calculator(**kw) as c:
            frac = c.unclamped_fraction()
            g1, g2 = c.vgrid()
        total = g1.size + g2.size  # 18000 and 15120 nodes: a part-filled last word of 64 bits
        below = np.count_nonzero(g1 < kw["ceiling"]) + np.count_nonzero(g2 < kw["ceiling"])
        assert 0 < below < total
        assert round(frac * total) == below
        o1, o2 = O.Oracle(spec).vgrid()
        assert np.count_nonzero(np.asarray(o1) < kw["ceiling"]) + np.count_nonzero(np.asarray(o2) < kw["ceiling"]) == below


@pytest.mark.gpu
def test_gpu_gradient_more_sites_than_one_block():
    """120 sites (the cage plus a second one 0.9 nm away): the gradient kernel
    takes its sites 64 at a time, so the pair list is walked twice; every
    entry against the oracle's gradient within 1e-10, single and in sets."""
    import bsl_amd as B
    p1, q1, t1, l1 = cs.cage11_sites()
    p2, q2, t2, l2 = cs.synthetic_sites(3)
    sites = (np.concatenate([p1, p2 + np.array([0.9, 0.0, 0.0])]), np.concatenate([q1, q2]),
             bytes(t1) + bytes(t2), np.concatenate([l1, np.asarray(l2, dtype=np.int32)]))
    spec, kw = cs.cage_case(5, 3, 8, 9, "All", sites=sites)
    ns = len(spec["site_q"])
    assert ns == 120
    o = O.Oracle(spec)
    rng = np.random.default_rng(5)
    sets = []
    for k in (2, 1, 1):
        psi = rng.uniform(-1, 1, (k, o.N))
        psi /= np.linalg.norm(psi, axis=1, keepdims=True)
        sets.append((psi, rng.uniform(-1, 1, k)))
    with B.Calculator(**kw) as c:
        got = [c.hf_gradient(p, g) for p, g in sets]
        got_sets = c.hf_gradient_sets(sets)
    for (psi, g), g1, gs in zip(sets, got, got_sets):
        ref = o.hf_gradient(psi, g)
        scale = np.max(np.abs(ref[:4 * ns]))
        for gg in (g1, gs):
            assert np.max(np.abs(gg["site_q"] - ref[0:4 * ns:4])) / scale < 1e-10
            assert np.max(np.abs(gg["site_pos"] - ref[:4 * ns].reshape(ns, 4)[:, 1:4])) / scale < 1e-10
            for e, name in zip(range(4 * ns, 4 * ns + 4), ("q_h2_h", "q_h2_cm", "lj_a", "lj_b")):
                assert abs(gg[name] - ref[e]) <= 1e-10 * abs(ref[e])
        for name in g1:
            assert np.array_equal(np.asarray(g1[name]), np.asarray(gs[name])), name


NITER = 800


@pytest.mark.gpu
def test_gpu_gradient_matches_finite_differences():
    """Gradients of the lowest levels from converged Ritz vectors against
    central differences (h = 1e-4 relative) of the GPU's refined levels."""
    import bsl_amd as B
    spec, kw = cs.cage_case(6, 3, 8, 8, "All")
    ns = len(spec["site_q"])
    c = B.Calculator(**kw)
    e, psi, res = c.levels(NITER, 2)
    assert np.all(res < 1e-6)
    grads = [c.hf_gradient(psi[k], [1.0]) for k in range(2)]
    c.close()
    kwargs_of = {"site_q": lambda s: dict(kw, sites=(s["site_pos"], s["site_q"], kw["sites"][2])),
                 "site_pos": lambda s: dict(kw, sites=(s["site_pos"], s["site_q"], kw["sites"][2]))}
    for name, idx in PARAMS:
        x = _value(spec, name, idx)
        h = 1e-4 * abs(x)
        fds = []
        for sgn in (1, -1):
            sp = _perturb(spec, name, idx, sgn * h)
            kk = kwargs_of[name](sp) if name in kwargs_of else dict(kw, **{name: sp[name]})
            with B.Calculator(**kk) as cp:
                fds.append(cp.levels(NITER, 2)[0])
        for lev in range(2):
            fd = (fds[0][lev] - fds[1][lev]) / (2 * h)
            gr = grads[lev]
            an = {"site_q": lambda: gr["site_q"][idx], "site_pos": lambda: gr["site_pos"].reshape(-1)[idx]}.get(
                name, lambda: gr[name])()
            assert abs(an - fd) <= 1e-4 * max(abs(fd), 1e-3 / abs(x)), (name, idx, lev, an, fd)


@pytest.mark.gpu
def test_gpu_gradient_of_device_ritz_vectors():
    """bsl_hf_gradient_ritz on the Ritz vectors levels() leaves on the device
    equals bsl_hf_gradient_sets of the downloaded unit vectors (within 1e-12
    of the largest entry: only the normalisation rounds differently), for
    sets of one and two levels with arbitrary weights; it refuses a context
    without refined Ritz vectors, one whose vectors a later bsl_ritz
    replaced, and indices beyond them."""
    import bsl_amd as B
    spec, kw = cs.cage_case(6, 3, 8, 8, "All")
    _ritz_gradient_case(B, kw)


@pytest.mark.gpu
def test_gpu_gradient_of_device_ritz_vectors_mixed_precision():
    """The same on an FP32-operator context (FP32 Loop 1 and 2, FP64 Ritz
    vectors and refinement), the adjoint leg's configuration."""
    import bsl_amd as B
    spec, kw = cs.cage_case(6, 3, 8, 8, "All")
    _ritz_gradient_case(B, dict(kw, precision="fp32"))


def _ritz_gradient_case(B, kw):
    with B.Calculator(**kw) as c:
        with pytest.raises(B.BslError):
            c.hf_gradient_ritz([([0], [1.0])])
        e, psi, res = c.levels(NITER, 2)
        sets_idx = [([0], [1.0]), ([1, 0], [0.7, -0.4]), ([1], [2.5]), ([0, 1], [1.0, 1.0]), ([0], [0.3])]
        got = c.hf_gradient_ritz(sets_idx)
        ref = c.hf_gradient_sets([(psi[list(i)], g) for i, g in sets_idx])
        for a, b in zip(got, ref):
            scale = max(np.max(np.abs(b["site_q"])), np.max(np.abs(b["site_pos"])))
            assert np.max(np.abs(a["site_q"] - b["site_q"])) <= 1e-12 * scale
            assert np.max(np.abs(a["site_pos"] - b["site_pos"])) <= 1e-12 * scale
            for name in ("q_h2_h", "q_h2_cm", "lj_a", "lj_b"):
                assert abs(a[name] - b[name]) <= 1e-12 * abs(b[name]), name
        with pytest.raises(B.BslError):
            c.hf_gradient_ritz([([2], [1.0])])
        # levels without downloading the vectors: the same gradients
        e2, none, _ = c.levels(NITER, 2, vectors=False)
        assert none is None and np.array_equal(e2, e)
        got2 = c.hf_gradient_ritz(sets_idx)
        for a, b in zip(got2, got):
            for name in a:
                assert np.array_equal(np.asarray(a[name]), np.asarray(b[name])), name
        a, b, b2 = c.lanczos(NITER, start="random", seed=0)
        evtr, _ = B.trivec(a, b, b2, e[:1], mult_tol=c.mult_tol)
        c.ritz(NITER, evtr, 1, start="random", seed=0)
        with pytest.raises(B.BslError):
            c.hf_gradient_ritz([([0], [1.0])])


@pytest.mark.gpu
def test_autograd_wrapper_chain_rule():
    """torch.autograd through the water-model parameters: dL/d(q_O, q_H,
    sigma_O, eps_O) of L = E0 + 2 E1 against central differences."""
    import torch
    import bsl_amd as B
    from bsl_amd import adjoint
    spec, kw = cs.cage_case(6, 3, 8, 8, "All")
    pos, q, typ, ljf = cs.cage11_sites()
    kinds = ["O" if t == ord("O") else "H" for t in typ]
    calc_kw = {k: v for k, v in kw.items() if k not in ("potential", "sites", "q_h2_h", "q_h2_cm", "lj_a",
                                                        "lj_b")}
    config = {"calc": calc_kw, "lj_sites": ljf, "niter": NITER, "count": 2}
    q_h2 = torch.tensor([0.4932, -0.9864], dtype=torch.float64)

    def loss(charges, lj):
        site_pos, site_q, lj_ab = adjoint.water_model_sites(charges, lj, torch.as_tensor(pos), kinds)
        e = adjoint.cage_levels(site_pos, site_q, q_h2, lj_ab, config)
        return e[0] + 2.0 * e[1]

    charges = torch.tensor([-0.8476, 0.4238], dtype=torch.float64, requires_grad=True)
    lj = torch.tensor([0.3166, 0.650, 0.296, 0.2852], dtype=torch.float64, requires_grad=True)
    L = loss(charges, lj)
    L.backward()
    for which, t, i in (("charges", charges, 0), ("charges", charges, 1), ("lj", lj, 0), ("lj", lj, 1)):
        x0 = t.detach().clone()
        h = 1e-4 * abs(float(x0[i]))
        vals = []
        for sgn in (1, -1):
            x = x0.clone()
            x[i] += sgn * h
            with torch.no_grad():
                vals.append(float(loss(x, lj.detach()) if which == "charges" else loss(charges.detach(), x)))
        fd = (vals[0] - vals[1]) / (2 * h)
        an = float(t.grad[i])
        assert abs(an - fd) <= 1e-4 * max(abs(fd), 1e-3), (which, i, an, fd)
