"""The standalone driver (lib/bsl_lanczos): the reference's command line,
decks and output files over the HIP calculators.  Error paths run on the CPU
(they stop before any device work); full runs are GPU tests."""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

import _cases as cs

DRIVER = os.path.join(os.path.dirname(cs.HERE), "boundstate-lanczos_amd", "lib", "bsl_lanczos")


def write_deck(path, values):
    with open(path, "w") as f:
        f.write("#\n#\n")
        for i, v in enumerate(values):
            f.write("name%d= %s\n" % (i, v))


def lanczos_deck(tmp, niter=300, calc="linRotCartSph_HIP_NoQuad_ContGrid_Alavi_SPCE", word="partial", nvec=2,
                 window=(-10, 10)):
    p = os.path.join(tmp, "lanczos.txt")
    write_deck(p, [niter, window[0], window[1], word, nvec, calc, tmp + "/"])
    return p


def hv_deck(tmp, nx=10, lmax=3, na=10, nb=10, sym="All", xmax=0.4, pxmax=0.48):
    shutil.copyfile(os.path.join(cs.GOLDEN, "sII_smallCage11_CM_EA_ZYZ.xyz"), os.path.join(tmp, "cage.xyz"))
    p = os.path.join(tmp, "hv.txt")
    write_deck(p, [nx, xmax, nx, xmax, nx, xmax, lmax, na, nb, 793, pxmax, 793, pxmax, 793, pxmax, 2.01588,
                   0.7096487130193394, "cage.xyz", "DriverTest", 18, "FALSE", sym])
    return p


def iho_decks(tmp, nx=21, lmax=3, na=20, nb=30, xmax=0.4, omega_cm1=100.0, mass=2.016, ceiling=1.0e6):
    """Hv deck and potential deck (argv[3]: mass, angular frequency,
    ceiling; IsoHarmOsc_Routines.cpp:38-62) of the isotropic oscillator
    workload: H2 (B = 59.322 cm^-1) in a 100 cm^-1 well, 21 points per axis,
    l <= 3, 20 x 30 quadrature."""
    from bsl_amd import physics
    shutil.copyfile(os.path.join(cs.GOLDEN, "sII_smallCage11_CM_EA_ZYZ.xyz"), os.path.join(tmp, "cage.xyz"))
    hv = os.path.join(tmp, "hv.txt")
    write_deck(hv, [nx, xmax, nx, xmax, nx, xmax, lmax, na, nb, 793, 0.48, 793, 0.48, 793, 0.48, mass,
                    repr(physics.cm1_to_kjmol(59.322)), "cage.xyz", "DriverTest", repr(ceiling), "FALSE", "All"])
    pot = os.path.join(tmp, "pot.txt")
    write_deck(pot, [mass, repr(physics.omega_from_cm1(omega_cm1)), repr(ceiling)])
    return hv, pot


def string_decks(tmp, calc, nx=8, lmax=3, na=8, nb=9):
    """Hv deck (and potential deck where the calculator reads argv[3]) of a
    small cage11 workload for one of the reference's calculator strings; the
    grid-mapped ones get a 17 x 19 x 23 potential universe that holds the
    box."""
    shutil.copyfile(os.path.join(cs.GOLDEN, "sII_smallCage11_CM_EA_ZYZ.xyz"), os.path.join(tmp, "cage.xyz"))
    univ = [17, 0.44, 19, 0.5, 23, 0.41] if "_ContGrid_" not in calc else [793, 0.48, 793, 0.48, 793, 0.48]
    hv = os.path.join(tmp, "hv.txt")
    write_deck(hv, [nx, 0.4, nx, 0.4, nx, 0.4, lmax, na, nb] + univ +
               [2.01588, 0.7096487130193394, "cage.xyz", "DriverTest", 18, "FALSE", "All"])
    pot = None
    if calc.endswith("_Coulomb"):
        pot = os.path.join(tmp, "pot.txt")
        write_deck(pot, [1.0, -0.5, -200.0])
    return hv, pot


def run(args, cwd):
    return subprocess.run([DRIVER] + args, cwd=cwd, capture_output=True, text=True, timeout=600)


pytestmark = pytest.mark.skipif(not os.path.exists(DRIVER), reason="driver not built")


def test_unknown_calculator_rejected(tmp_path):
    t = str(tmp_path)
    r = run([lanczos_deck(t, calc="linRotCartSph_Alavi_SPCE"), hv_deck(t)], t)
    assert r.returncode == 1 and "not recognized" in r.stderr


def test_bad_eigenvector_word_rejected(tmp_path):
    t = str(tmp_path)
    r = run([lanczos_deck(t, word="some"), hv_deck(t)], t)
    assert r.returncode == 1 and "NumberOfEigenvectors" in r.stderr


def test_bad_symmetry_flag_rejected(tmp_path):
    t = str(tmp_path)
    r = run([lanczos_deck(t), hv_deck(t, sym="Both")], t)
    assert r.returncode == 1 and "Symmeterizer" in r.stderr


def test_box_larger_than_potential_universe_rejected(tmp_path):
    t = str(tmp_path)
    r = run([lanczos_deck(t), hv_deck(t, xmax=0.6, pxmax=0.48)], t)
    assert r.returncode == 1 and "potential universe" in r.stderr


def test_missing_deck_rejected(tmp_path):
    t = str(tmp_path)
    r = run([os.path.join(t, "nope.txt"), hv_deck(t)], t)
    assert r.returncode == 1 and "could not be opened" in r.stderr


def _read_boundstates(d):
    rows = []
    with open(os.path.join(d, "boundstates.txt")) as f:
        lines = f.read().split("\n")
    assert lines[0] == ""
    assert lines[1].startswith("The following are the eigenvalues and errors for the simulation DriverTest")
    for line in lines[4:]:
        t = line.split()
        if len(t) == 3:
            rows.append([float(t[1]), float(t[2])])
    return np.array(rows)


@pytest.mark.gpu
@pytest.mark.parametrize("path", ["device", "host"])
def test_driver_reproduces_saved_run(tmp_path, path):
    """The reference's saved run Compilation/SmallBasisTest2_2013-09-03_19_06_47
    through the drop-in: same decks, same output files, same converged levels."""
    t = str(tmp_path)
    env = dict(os.environ)
    if path == "host":
        env["BSL_HV_PATH"] = "host"
    r = subprocess.run([DRIVER, lanczos_deck(t), hv_deck(t)], cwd=t, capture_output=True, text=True, timeout=600,
                       env=env)
    assert r.returncode == 0, r.stderr[-2000:]
    out = [d for d in os.listdir(t) if d.startswith("DriverTest_")]
    assert len(out) == 1
    d = os.path.join(t, out[0])
    for f in ("boundstates.txt", "states_zpe.txt", "lv", "eigVecFile.txt", "lanczos.txt", "hv.txt"):
        assert os.path.exists(os.path.join(d, f)), f
    got = _read_boundstates(d)
    run_fx = json.load(open(os.path.join(cs.GOLDEN, "runs.json")))["SmallBasisTest2_2013-09-03_19_06_47"]
    ref = np.array(run_fx["boundstates"])
    for e in ref[:10, 0]:
        assert np.min(np.abs(got[:, 0] - e)) / abs(e) < 1e-9
    # Ritz vectors: two columns, unit norm for the converged ground state
    vec = np.loadtxt(os.path.join(d, "eigVecFile.txt"), skiprows=4)
    assert vec.shape == (16000, 3)
    assert abs(np.linalg.norm(vec[:, 1]) - 1.0) < 1e-6


@pytest.mark.gpu
@pytest.mark.parametrize("calc", ["linRotCartSph_HIP_NoQuad_ContGrid_Alavi_SPCE_FP32",
                                  "linRotCartSph_HIP_Quad_ContGrid_Alavi_SPCE_BF16X3"])
def test_driver_mixed_precision_long_run(tmp_path, calc):
    """The saved run's decks through the mixed-precision calculators for 1000
    steps (ghost copies of the converged levels exist by then): the driver's
    eigensolve uses the mixed multiplicity tolerance, so each of the saved
    run's 10 lowest levels appears once in boundstates.txt, within
    0.01 cm^-1."""
    from bsl_amd import physics
    t = str(tmp_path)
    r = run([lanczos_deck(t, niter=1000, calc=calc), hv_deck(t)], t)
    assert r.returncode == 0, r.stderr[-2000:]
    d = os.path.join(t, [x for x in os.listdir(t) if x.startswith("DriverTest_")][0])
    got = _read_boundstates(d)[:, 0]
    ref = np.array(json.load(open(os.path.join(cs.GOLDEN, "runs.json")))
                   ["SmallBasisTest2_2013-09-03_19_06_47"]["boundstates"])[:10, 0]
    for e in ref:
        near = np.abs(got - e) / physics.KJMOL_PER_CM1 < 0.01
        assert near.sum() == 1, (e, got[near])
    # the two levels whose Ritz vectors the deck asks for are refined on the
    # device by their FP64 Rayleigh quotient: the saved FP64 run's values
    rows = np.loadtxt(os.path.join(d, "refined_levels.txt"), skiprows=4)
    assert rows.shape == (2, 4)
    for k in range(2):
        assert abs(rows[k, 2] - ref[k]) <= 1e-9 * abs(ref[k]), (k, rows[k], ref[k])
        assert rows[k, 3] < 1e-4


def hv_deck_grid(tmp, run):
    shutil.copyfile(os.path.join(cs.GOLDEN, "sII_smallCage11_CM_EA_ZYZ.xyz"), os.path.join(tmp, "cage.xyz"))
    vals = list(run["hv_deck"])
    vals[17] = "cage.xyz"
    vals[18] = "DriverTest"
    p = os.path.join(tmp, "hv.txt")
    write_deck(p, vals)
    return p


@pytest.mark.gpu
def test_driver_reference_string_grid_mapped(tmp_path):
    """The reference's own calculator string linRotCartSph_NoQuad_Alavi_SPCE
    and decks of Compilation/SmallBasisTest2_2013-08-29_19_35_47 through the
    driver: the saved run's converged levels."""
    t = str(tmp_path)
    run_fx = json.load(open(os.path.join(cs.GOLDEN, "runs.json")))["SmallBasisTest2_2013-08-29_19_35_47"]
    lz = run_fx["lanczos_deck"]
    r = run([lanczos_deck(t, niter=int(lz[0]), calc=lz[5]), hv_deck_grid(t, run_fx)], t)
    assert r.returncode == 0, r.stderr[-2000:]
    d = os.path.join(t, [x for x in os.listdir(t) if x.startswith("DriverTest_")][0])
    got = _read_boundstates(d)
    ref = np.array(run_fx["boundstates"])
    for e in ref[:3, 0]:
        assert np.min(np.abs(got[:, 0] - e)) / abs(e) < 1e-10


@pytest.mark.gpu
def test_driver_matches_reference_program(tmp_path):
    """The same two decks (the reference's own calculator string, 300 steps,
    two Ritz vectors) through the drop-in driver and through the reference's
    Lanczos program built from its sources (oracle/_ref/ref_lanczos): the same
    four output files with the same header lines; every converged level
    (error below 1e-10) of either run within 1e-9 relative of a level of the
    other; the two Ritz vectors equal up to sign within 1e-6."""
    import _oracle as O
    if not os.path.exists(O.REF_LANCZOS):
        pytest.skip("reference program (oracle/_ref) not built")
    dirs = {}
    for side in ("ref", "hip"):
        t = str(tmp_path / side)
        os.makedirs(t)
        ld = lanczos_deck(t, niter=300, calc="linRotCartSph_NoQuad_ContGrid_Alavi_SPCE")
        hd = hv_deck(t)
        if side == "ref":
            dirs[side] = os.path.join(t, O.run_ref_lanczos(ld, hd, t)[2])
        else:
            r = run([ld, hd], t)
            assert r.returncode == 0, r.stderr[-2000:]
            dirs[side] = os.path.join(t, [x for x in os.listdir(t) if x.startswith("DriverTest_")][0])
    for f in ("boundstates.txt", "states_zpe.txt", "lv", "eigVecFile.txt"):
        heads = []
        for side in ("ref", "hip"):
            with open(os.path.join(dirs[side], f)) as fh:
                heads.append([fh.readline() for _ in range(4)])
        # line 2 carries the count, which may differ by a ghost copy
        assert [heads[0][i] for i in (0, 1, 3)] == [heads[1][i] for i in (0, 1, 3)], f
    ref, hip = _read_boundstates(dirs["ref"]), _read_boundstates(dirs["hip"])
    for a, b in ((ref, hip), (hip, ref)):
        conv = a[(a[:, 1] < 1e-10), 0]
        assert len(conv) >= 5
        d = np.min(np.abs(conv[:, None] - b[None, :, 0]), axis=1)
        assert np.max(d / np.abs(conv)) < 1e-9
    vr = np.loadtxt(os.path.join(dirs["ref"], "eigVecFile.txt"), skiprows=4)
    vh = np.loadtxt(os.path.join(dirs["hip"], "eigVecFile.txt"), skiprows=4)
    assert vr.shape == vh.shape == (16000, 3)
    for c in (1, 2):
        s = np.sign(np.dot(vr[:, c], vh[:, c]))
        assert np.max(np.abs(vr[:, c] - s * vh[:, c])) < 1e-6


@pytest.mark.gpu
def test_driver_matches_reference_program_iho(tmp_path):
    """The plain-FP64 isotropic oscillator workload (21^3, l <= 3, 20 x 30,
    300 steps from the uniform start) through the drop-in driver and the
    reference's own program on the same three decks: the good eigenvalues of
    the two boundstates.txt files pair up one to one within 1e-9 relative,
    and the device recurrence's alpha and beta (bsl_lanczos, uniform start)
    are within 1e-9 relative of the ones the reference program prints over
    the first 40 steps (about 3e-13 measured).  From about step 45 the
    ground level has converged and the lost orthogonality amplifies any
    rounding difference between two FP64 runs (the reference's own alpha
    and beta against a device run differ by 1e-3 at step 60 and 0.4 by step 150),
    so the later coefficients are compared through the levels."""
    import _oracle as O
    if not os.path.exists(O.REF_LANCZOS):
        pytest.skip("reference program (oracle/_ref) not built")
    rows = {}
    for side in ("ref", "hip"):
        t = str(tmp_path / side)
        os.makedirs(t)
        ld = lanczos_deck(t, niter=300, calc="linRotCartSph_NoQuad_IsoHarmOsc", word="none", nvec=0)
        hv, pot = iho_decks(t)
        if side == "ref":
            a_ref, b_ref, d = O.run_ref_lanczos(ld, hv, t, pot)
            d = os.path.join(t, d)
        else:
            r = run([ld, hv, pot], t)
            assert r.returncode == 0, r.stderr[-2000:]
            d = os.path.join(t, [x for x in os.listdir(t) if x.startswith("DriverTest_")][0])
        rows[side] = _read_boundstates(d)
    ref, hip = rows["ref"][:, 0], rows["hip"][:, 0]
    assert len(ref) == len(hip) and len(ref) > 10
    assert np.max(np.abs(ref - hip) / np.abs(ref)) < 1e-9
    import bsl_amd as B
    from bsl_amd import physics
    _, kw = cs.iho_case(21, 3, 20, 30, x_max=0.4, rot=physics.cm1_to_kjmol(59.322))
    with B.Calculator(**kw) as c:
        a, b, _ = c.lanczos(300, start="uniform")
    assert len(a_ref) == 300
    assert np.max(np.abs(a[:40] - a_ref[:40]) / np.abs(a_ref[:40])) < 1e-9
    assert np.max(np.abs(b[1:41] - b_ref[:40]) / np.abs(b_ref[:40])) < 1e-9


@pytest.mark.gpu
@pytest.mark.parametrize("calc", ["linRotCartSph_NoQuad_ContGrid_Alavi_TIP4P", "linRotCartSph_NoQuad_Alavi_SPCE",
                                  "linRotCartSph_NoQuad_Alavi_TIP4P", "linRotCartSph_NoQuad_Coulomb"])
def test_driver_matches_reference_program_strings(tmp_path, calc):
    """The reference's other calculator strings (TIP4P sites, the
    grid-mapped potential universe, the Coulomb probe) on an 8^3 x 16
    workload, 300 steps: the drop-in driver's converged levels and the
    reference program's (built from source) agree within 1e-9 relative.

    The Coulomb probe clamps the whole box at its -200 kJ/mol cap, so V is a
    constant and H separates into box x rotor states.  The uniform start has
    even translational parity on every axis; the even-axis kinetic kernels
    contract v(i) + v(n-1-i) and v(i) - v(n-1-i) separately and keep that
    parity exactly, while the reference's plain products let rounding seed
    the odd states, which then converge (levels spaced 3 E1 beside the even
    ones' 8 E1).  The kinetic contractions keep the parity on every axis
    length, so for this string the comparison is made two-way through the
    separable spectrum (1-D sinc-DVR eigenvalues from the oracle's kinetic
    matrix, the constant potential, B l(l+1)), split by translational
    parity: every level of the driver is an all-even state and in the
    reference's list; every level of the reference is a state of that
    spectrum; at least three of those the driver does not list are odd on
    some axis and nowhere even; and a random-start run of the library
    (which reaches every parity) lists each of the reference's levels."""
    import _oracle as O
    if not os.path.exists(O.REF_LANCZOS):
        pytest.skip("reference program (oracle/_ref) not built")
    rows = {}
    for side in ("ref", "hip"):
        t = str(tmp_path / side)
        os.makedirs(t)
        ld = lanczos_deck(t, niter=300, calc=calc, word="none", nvec=0,
                          window=(-400, 400) if calc.endswith("_Coulomb") else (-10, 10))
        hv, pot = string_decks(t, calc)
        if side == "ref":
            d = os.path.join(t, O.run_ref_lanczos(ld, hv, t, pot)[2])
        else:
            r = run([ld, hv] + ([pot] if pot else []), t)
            assert r.returncode == 0, r.stderr[-2000:]
            d = os.path.join(t, [x for x in os.listdir(t) if x.startswith("DriverTest_")][0])
        rows[side] = _read_boundstates(d)
    pairs = [(rows["hip"], rows["ref"])]
    if not calc.endswith("_Coulomb"):
        pairs.append((rows["ref"], rows["hip"]))
    for a, b in pairs:
        conv = a[(a[:, 1] < 1e-10), 0]
        assert len(conv) >= 5
        d = np.min(np.abs(conv[:, None] - b[None, :, 0]), axis=1)
        assert np.max(d / np.abs(conv)) < 1e-9
    if calc.endswith("_Coulomb"):
        _coulomb_two_way(rows["hip"], rows["ref"])


def _coulomb_two_way(hip, ref):
    """The Coulomb string's levels split by translational parity (see
    test_driver_matches_reference_program_strings)."""
    import bsl_amd as B
    import _oracle as O
    spec, kw = cs.coulomb_case(8, 3, 8, 9, 0.4, charges=(1.0, -0.5, -200.0))
    spec = dict(spec, mass=2.01588, rot_const=0.7096487130193394)
    kw = dict(kw, mass=2.01588, rot_const=0.7096487130193394)
    o = O.Oracle(spec)
    v1, v2 = o.vgrid()
    assert np.ptp(v1) == 0 and np.ptp(v2) == 0 and v1.flat[0] == v2.flat[0]
    w, u = np.linalg.eigh(o.kin(0))
    even = np.array([np.allclose(u[:, k], u[::-1, k], atol=1e-10) for k in range(len(w))])
    assert even.sum() == len(w) // 2
    rot = sorted({spec["rot_const"] * l * (l + 1) for l in range(4)})
    lev_even, lev_odd = [], []
    for a in range(len(w)):
        for b in range(len(w)):
            for c in range(len(w)):
                for r in rot:
                    e = w[a] + w[b] + w[c] + v1.flat[0] + r
                    (lev_even if even[a] and even[b] and even[c] else lev_odd).append(e)
    lev_even, lev_odd = np.array(lev_even), np.array(lev_odd)

    def near(x, pool):
        return np.min(np.abs(pool - x)) <= 1e-9 * abs(x)

    hconv = hip[hip[:, 1] < 1e-10, 0]
    rconv = ref[ref[:, 1] < 1e-10, 0]
    assert len(hconv) >= 5 and len(rconv) > len(hconv)
    for e in hconv:
        assert near(e, lev_even), e
    for e in rconv:
        assert near(e, np.concatenate([lev_even, lev_odd])), e
    only_ref = [e for e in rconv if not near(e, hip[:, 0])]
    odd_only = [e for e in only_ref if near(e, lev_odd) and not near(e, lev_even)]
    assert len(odd_only) >= 3
    # 1000 steps: a random start spreads over every symmetry sector, so it
    # needs more steps than the uniform one to converge the same levels
    with B.Calculator(**kw) as c:
        ev, _ = B.lancbis(*c.lanczos(1000, start="random", seed=3), -400.0, 400.0)
    for e in rconv:
        assert near(e, ev), e


@pytest.mark.gpu
def test_reference_program_selects_hip_calculator(tmp_path):
    """The reference's own Lanczos program (lanczosEigenValuesVectors.cpp
    main and interface.cpp, built from source) with the plugin bound into its
    calculator selection (oracle/ref_bind.cpp, the branch INTEGRATION's
    interface.cpp edit adds): the HIP string runs through the reference's
    driver (its Loop 1 and Loop 2, the Hv function pointer on host vectors)
    and reproduces that program's run of its own string on the saved run's
    decks, coefficients over the first 40 steps and every converged level
    both ways within 1e-9; a string neither knows still ends in the
    reference's error branch."""
    import _oracle as O
    if not (os.path.exists(O.REF_LANCZOS) and os.path.exists(O.REF_LANCZOS_HIP)):
        pytest.skip("reference program (oracle/_ref) not built")
    res = {}
    for side, calc, prog in (("ref", "linRotCartSph_NoQuad_ContGrid_Alavi_SPCE", O.REF_LANCZOS),
                             ("hip", "linRotCartSph_HIP_NoQuad_ContGrid_Alavi_SPCE", O.REF_LANCZOS_HIP)):
        t = str(tmp_path / side)
        os.makedirs(t)
        a, b, d = O.run_ref_lanczos(lanczos_deck(t, niter=300, calc=calc), hv_deck(t), t, program=prog)
        res[side] = (a, b, _read_boundstates(os.path.join(t, d)))
    (ar, br, lr), (ah, bh, lh) = res["ref"], res["hip"]
    assert len(ah) == len(ar) == 300
    assert np.max(np.abs(ah[:40] - ar[:40]) / np.abs(ar[:40])) < 1e-9
    assert np.max(np.abs(bh[:40] - br[:40]) / np.abs(br[:40])) < 1e-9
    for x, y in ((lh, lr), (lr, lh)):
        conv = x[x[:, 1] < 1e-10, 0]
        assert len(conv) >= 5
        d = np.min(np.abs(conv[:, None] - y[None, :, 0]), axis=1)
        assert np.max(d / np.abs(conv)) < 1e-9
    t = str(tmp_path / "unknown")
    os.makedirs(t)
    r = subprocess.run([O.REF_LANCZOS_HIP, lanczos_deck(t, calc="linRotCartSph_HIP_NoSuchCalculator"), hv_deck(t)],
                       cwd=t, capture_output=True, text=True, timeout=120)
    assert r.returncode == 1 and "not recognized" in r.stderr
