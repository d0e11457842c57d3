"""Shared problem definitions for the oracle and the GPU calculator."""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")

MASS = 2.01588
B_ROT = 0.7096487130193394


def cage11_cm_ea():
    import bsl_amd.cage as cg
    return cg.read_xyz(os.path.join(GOLDEN, "sII_smallCage11_CM_EA_ZYZ.xyz"))


def cage11_sites():
    """SPC/E sites of the reference's cage 11 geometry, built by the oracle."""
    import _oracle as O
    pos, q, typ = O.build_sites(cage11_cm_ea(), 0)
    lj = np.array([1 if t == ord("O") else 0 for t in typ], dtype=np.int32)
    return pos, q, typ, lj


def synthetic_sites(seed=0):
    import bsl_amd.cage as cg
    pos, q, lj = cg.synthetic_cage(seed)
    typ = bytes([ord("O") if f else ord("H") for f in lj])
    return pos, q, typ, lj


SYMS = {"All": 0, "Even": 1, "Odd": 2}


def cage_case(nx, lmax, na, nb, sym="All", x_max=0.4, sites=None, lj=(1.364e-06, 1.533e-3), mass=MASS,
              rot=B_ROT, ceiling=18.0):
    """(oracle spec, calculator kwargs) for an Alavi H2 / SPC/E cage."""
    pos, q, typ, ljf = sites if sites is not None else cage11_sites()
    spec = dict(nx=nx, ny=nx, nz=nx, x_max=x_max, y_max=x_max, z_max=x_max, lmax=lmax, na=na, nb=nb, mass=mass,
                rot_const=rot, ceiling=ceiling, sym=SYMS[sym], pot_kind=1, site_pos=pos, site_q=q, site_type=typ,
                q_h2_h=0.4932, q_h2_cm=-0.9864, h2_half_bond=0.03707, lj_a=lj[0], lj_b=lj[1])
    kw = dict(nx=nx, x_max=x_max, lmax=lmax, n_theta=na, n_phi=nb, mass=mass, rot_const=rot, ceiling=ceiling,
              symmetry=sym, potential="cage", sites=(pos, q, ljf), q_h2_h=0.4932, q_h2_cm=-0.9864,
              h2_half_bond=0.03707, lj_a=lj[0], lj_b=lj[1])
    return spec, kw


def cage_grid_case(nx, lmax, na, nb, sym="All", universe=(20, 20, 20, 0.48, 0.48, 0.48), **kw):
    """(oracle spec, calculator kwargs) for the grid-mapped cage potential."""
    spec, ckw = cage_case(nx, lmax, na, nb, sym, **kw)
    spec = dict(spec, pot_kind=2, universe=tuple(universe))
    ckw = dict(ckw, potential="cage_grid", universe=tuple(universe))
    return spec, ckw


def coulomb_case(nx, lmax, na, nb, x_max=0.4, charges=(1.0, -0.5, -200.0), sym="All", ceiling=18.0):
    """(oracle spec, calculator kwargs) for a point charge at the origin."""
    spec, ckw = iho_case(nx, lmax, na, nb, x_max, sym=sym, ceiling=ceiling)
    spec = dict(spec, pot_kind=3, coulomb=tuple(charges))
    ckw = {k: v for k, v in ckw.items() if k not in ("iho_factor", "iho_ceil")}
    ckw = dict(ckw, potential="coulomb", coulomb=tuple(charges))
    return spec, ckw


def iho_case(nx, lmax, na, nb, x_max, omega_cm1=100.0, mass=2.016, rot=0.709648713019339, sym="All",
             ceiling=1.0e6):
    from bsl_amd import physics
    w = physics.omega_from_cm1(omega_cm1)
    factor = 0.5 * mass * w * w
    spec = dict(nx=nx, ny=nx, nz=nx, x_max=x_max, y_max=x_max, z_max=x_max, lmax=lmax, na=na, nb=nb, mass=mass,
                rot_const=rot, ceiling=ceiling, sym=SYMS[sym], pot_kind=0, iho_factor=factor, iho_ceil=ceiling)
    kw = dict(nx=nx, x_max=x_max, lmax=lmax, n_theta=na, n_phi=nb, mass=mass, rot_const=rot, ceiling=ceiling,
              symmetry=sym, potential="iho", iho_factor=factor, iho_ceil=ceiling)
    return spec, kw


def baseline_cage_case(nx, na=40, nb=70, sym="All", seed=0, x_max=0.4, lmax=5):
    """(oracle spec, calculator kwargs) of the benchmark cage workload: Alavi
    H2 (mass 2.016 amu, B = 59.322 cm^-1) in the synthetic 5^12 SPC/E cage of
    `seed`, LJ from the stated sigma/epsilon mixing, ceiling 18 kJ/mol."""
    import bsl_amd.cage as cg
    from bsl_amd import physics
    return cage_case(nx, lmax, na, nb, sym, x_max=x_max, sites=synthetic_sites(seed), lj=cg.baseline_lj(),
                     mass=2.016, rot=physics.cm1_to_kjmol(59.322))


def reduced_start(o, seed, carried_l_parity):
    """The oracle's start vector of a reduced-basis run: the library's random
    start (same generator and seed) on the carried planes only (the l parity
    of the symmetry block), zero on the others, normalised."""
    import _oracle as O
    v = O.start_vector(o.N, seed)
    ls = o.qnum()[:, 0]
    P = o.P
    for n, l in enumerate(ls):
        if l % 2 != carried_l_parity:
            v[n * P:(n + 1) * P] = 0.0
    return v / np.linalg.norm(v)
