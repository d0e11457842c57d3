"""Writes the golden fixtures under tests/golden/ from the reference's own
data files (run in the build container, where the reference is mounted).

Fixtures (data only: inputs and expected outputs):
  sII_smallCage11_CM_EA_ZYZ.xyz   cage geometry used by the saved runs
  sII_unitCelltest_CM_EA_ZYZ.xyz  unit cell of the -30.2331 kJ/mol known answer
  runs.json                       decks and boundstates.txt of the saved runs
                                  Compilation/SmallBasisTest2_2013-09-03_{18_36_00,19_06_47}
                                  (ContGrid) and _2013-08-29_{19_35_47,19_42_01} (grid-mapped)
  point_energies.json             point energies of the reference's compiled
                                  site builders and pair sums (TIP4P Coulomb in
                                  the sII unit cell, SPC/E Coulomb + LJ in cage 11)
  ref_hv_cage11.npz               V blocks and Hv of the reference's own compiled
                                  calculator (oracle/_ref) on two small cage11
                                  decks, for a seeded input vector
Usage: python tests/golden/make_fixtures.py [/root/reference]
"""
import json
import os
import shutil
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(REPO, "tests"))

RUNS = ["SmallBasisTest2_2013-09-03_18_36_00", "SmallBasisTest2_2013-09-03_19_06_47",
        # grid-mapped calculator (linRotCartSph_NoQuad_Alavi_SPCE, 20^3 potential universe)
        "SmallBasisTest2_2013-08-29_19_35_47", "SmallBasisTest2_2013-08-29_19_42_01"]


def parse_deck(path):
    vals = []
    with open(path) as f:
        lines = f.read().split("\n")[2:]
    for t in " ".join(lines).split():
        vals.append(t)
    # name/value pairs; names ignored, order is what counts
    return [vals[i + 1] for i in range(0, len(vals) - 1, 2)]


def read_boundstates(path):
    out = []
    with open(path) as f:
        for line in f.read().split("\n")[4:]:
            t = line.split()
            if len(t) == 3:
                out.append([float(t[1]), float(t[2])])
    return out


def write_hv_deck(path, nx, xmax, lmax, na, nb, geo, sym):
    vals = [nx, xmax, nx, xmax, nx, xmax, lmax, na, nb, 2, 0.48, 2, 0.48, 2, 0.48, 2.01588, 0.7096487130193394,
            geo, "fixture", 18.0, "FALSE", sym]
    with open(path, "w") as f:
        f.write("#\n#\n")
        for i, v in enumerate(vals):
            f.write("v%d= %s\n" % (i, repr(v) if isinstance(v, float) else v))


def main(ref):
    comp = os.path.join(ref, "Compilation")
    for name, src in (("sII_smallCage11_CM_EA_ZYZ.xyz", os.path.join(comp, "sII_smallCage11_CM_EA_ZYZ.xyz")),
                      ("sII_unitCelltest_CM_EA_ZYZ.xyz", os.path.join(ref, "Execution", "sII_unitCelltest_CM_EA_ZYZ.xyz"))):
        shutil.copyfile(src, os.path.join(HERE, name))
    runs = {}
    for r in RUNS:
        d = os.path.join(comp, r)
        lz = [f for f in os.listdir(d) if f.startswith("Lanczos")][0]
        hv = [f for f in os.listdir(d) if f.startswith("HvInput")][0]
        runs[r] = {"lanczos_deck": parse_deck(os.path.join(d, lz)), "hv_deck": parse_deck(os.path.join(d, hv)),
                   "boundstates": read_boundstates(os.path.join(d, "boundstates.txt"))}
    with open(os.path.join(HERE, "runs.json"), "w") as f:
        json.dump(runs, f, indent=1)

    import _oracle as O
    if not O.ref_available():
        print("oracle/_ref not built; skipping ref_hv_cage11.npz")
        return
    R = O.ref_lib()
    import ctypes as C
    R.ref_tip4p_coulomb.restype = C.c_double
    R.ref_tip4p_coulomb.argtypes = [C.c_char_p] + [C.c_double] * 4
    geo = os.path.join(HERE, "sII_smallCage11_CM_EA_ZYZ.xyz")
    cell = os.path.join(HERE, "sII_unitCelltest_CM_EA_ZYZ.xyz")
    pts = [(1.0824, 0.2164, 1.0819), (0.6491, 0.2165, 0.6492), (0.8651, 0.8655, 0.8654)]
    energies = {"tip4p_coulomb_unit_cell": [[list(p), R.ref_tip4p_coulomb(cell.encode(), *p, 1.0)] for p in pts],
                "spce_cage11": []}
    rng = np.random.default_rng(5)
    for _ in range(8):
        p = [float(x) for x in rng.uniform(-0.2, 0.2, 3)]
        energies["spce_cage11"].append([p, R.ref_spce_energy(geo.encode(), *p, -0.9864, 1),
                                        R.ref_spce_energy(geo.encode(), *p, 0.4932, 0)])
    with open(os.path.join(HERE, "point_energies.json"), "w") as f:
        json.dump(energies, f, indent=1)
    out = {}
    for tag, (nx, lmax, na, nb, sym) in {"a": (5, 2, 5, 6, "All"), "b": (6, 3, 8, 7, "Even")}.items():
        deck = os.path.join("/tmp", "bsl_fixture_hv_%s.txt" % tag)
        write_hv_deck(deck, nx, 0.4, lmax, na, nb, geo, sym)
        N = R.ref_setup(b"linRotCartSph_NoQuad_ContGrid_Alavi_SPCE", deck.encode(), None)
        nn = (lmax + 1) ** 2
        vb = np.zeros(nx ** 3 * nn * nn)
        R.ref_vblocks(vb, vb.size)
        v = np.random.default_rng(1234).uniform(-1, 1, N)
        u = np.zeros(N)
        R.ref_hv(v, u)
        out["%s_params" % tag] = np.array([nx, lmax, na, nb, {"All": 0, "Even": 1, "Odd": 2}[sym]])
        out["%s_vblocks" % tag] = vb
        out["%s_v" % tag] = v
        out["%s_hv" % tag] = u
    np.savez_compressed(os.path.join(HERE, "ref_hv_cage11.npz"), **out)
    print("fixtures written")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
