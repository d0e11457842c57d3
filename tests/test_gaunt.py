"""The harmonic form of the NoQuad potential on the CPU: the Gaunt term list
and values the library builds (bsl_gaunt_table; host code, no device)
against integrals computed here with scipy's spherical harmonics in the
reference's real-harmonic convention (normAssocLegendrePoly and
tesseralTrigTerm, linRotCartSphHvRoutines.cpp:293-434).

* every coupling that is non-zero is in the list, the values agree to
  1e-13, and the listed terms that are not are accidental zeros;
* the identity the form rests on, Y_i Y_j = sum_c G[i][j][c] Yh_c, holds at
  random points of the sphere to 1e-12 with the library's G (a wrong sign,
  index, normalisation or a missing term fails it);
* the term counts of the lmax-5 blocks (1120 odd, 477 even).
"""
import math

import numpy as np
import pytest
import scipy.special as sp

import bsl_amd as B


def real_y(l, m, x, phi):
    am = abs(m)
    norm = math.sqrt(math.factorial(l - am) * (2 * l + 1) / (2.0 * math.factorial(l + am)))
    p = norm * sp.lpmv(am, l, x)  # scipy includes the Condon-Shortley phase (-1)^m, as the reference
    if m == 0:
        t = 1.0 / math.sqrt(2 * math.pi)
    elif m > 0:
        t = np.cos(m * phi) / math.sqrt(math.pi)
    else:
        t = np.sin(am * phi) / math.sqrt(math.pi)
    return p * t


def block(lb, par):
    return [(l, m) for m in range(-lb, lb + 1) for l in range(abs(m), lb + 1) if l % 2 == par]


def columns(lb):
    return [(L, M) for L in range(0, 2 * lb + 1, 2) for M in range(-L, L + 1)]


SHAPES = [(0, 0), (1, 1), (2, 0), (3, 1), (4, 0), (5, 1), (6, 0)]


@pytest.mark.parametrize("lb,par", SHAPES)
def test_gaunt_table_matches_scipy_integrals(lb, par):
    ijc, g = B.gaunt_table(lb, par)
    blk, cols = block(lb, par), columns(lb)
    x, w = np.polynomial.legendre.leggauss(2 * lb + 4)
    nphi = 4 * lb + 6
    phi = np.pi * (2 * np.arange(nphi) + 1) / nphi
    X, PH = np.meshgrid(x, phi, indexing="ij")
    W = np.outer(w, np.full(nphi, 2 * np.pi / nphi))
    yb = np.array([real_y(l, m, X, PH) for l, m in blk])
    yc = np.array([real_y(L, M, X, PH) for L, M in cols])
    gnum = np.einsum("iab,jab,cab,ab->ijc", yb, yb, yc, W)
    listed = np.zeros(gnum.shape, dtype=bool)
    listed[ijc[:, 0], ijc[:, 1], ijc[:, 2]] = True
    assert np.all(ijc[:, 0] <= ijc[:, 1])
    assert listed.sum() == len(g)  # no duplicates
    upper = np.triu(np.ones((len(blk), len(blk)), dtype=bool))[:, :, None] & np.ones(len(cols), dtype=bool)
    missing = upper & ~listed & (np.abs(gnum) > 1e-12)
    assert not missing.any()
    assert np.max(np.abs(g - gnum[ijc[:, 0], ijc[:, 1], ijc[:, 2]])) < 1e-13
    accidental = np.abs(gnum[ijc[:, 0], ijc[:, 1], ijc[:, 2]]) < 1e-12
    assert np.all(np.abs(g[accidental]) < 1e-13)


@pytest.mark.parametrize("lb,par", SHAPES)
def test_product_identity_at_random_points(lb, par):
    ijc, g = B.gaunt_table(lb, par)
    blk, cols = block(lb, par), columns(lb)
    rng = np.random.default_rng(lb)
    x = rng.uniform(-1, 1, 200)
    phi = rng.uniform(0, 2 * np.pi, 200)
    yb = np.array([real_y(l, m, x, phi) for l, m in blk])
    yc = np.array([real_y(L, M, x, phi) for L, M in cols])
    expand = np.zeros((len(blk), len(blk), x.size))
    for (i, j, c), gv in zip(ijc, g):
        expand[i, j] += gv * yc[c]
    iu, ju = np.triu_indices(len(blk))
    assert np.max(np.abs(expand[iu, ju] - yb[iu] * yb[ju])) < 1e-12


def test_term_counts_and_arguments():
    assert len(B.gaunt_table(5, 1)[1]) == 1120 and len(B.gaunt_table(4, 0)[1]) == 477
    for lb, par in ((7, 1), (4, 1), (-1, 1), (2, 2)):
        with pytest.raises(B.BslError) as ei:
            B.gaunt_table(lb, par)
        assert ei.value.code == 1
