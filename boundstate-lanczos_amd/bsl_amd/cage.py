"""Water-cage geometry: the reference's .xyz format (centre of mass + ZY'Z''
Euler angles per molecule) and the synthetic 5^12 cage of rigid SPC/E
waters used for the cage workloads."""
import math

import numpy as np

from . import physics
from ._lib import check, lib


def read_xyz(path):
    """CM/Euler rows of a reference geometry file (SPCE_AH2_EngRoutines.cpp:15-39):
    first token the molecule count, a comment line, then per molecule
    `name CMx phi CMy theta CMz chi` (Angstrom, radian).  Returns [nmol, 6]."""
    with open(path) as f:
        tokens = f.read().split("\n")
    nmol = int(tokens[0].split()[0])
    rows = []
    for line in tokens[2:]:
        t = line.split()
        if len(t) >= 7:
            rows.append([float(x) for x in t[1:7]])
        if len(rows) == nmol:
            break
    if len(rows) != nmol:
        raise ValueError("%s: expected %d molecules, found %d" % (path, nmol, len(rows)))
    return np.array(rows, dtype=np.float64)


def sites_from_cm_ea(cm_ea, model=0):
    """Water sites (positions nm, charges e, LJ flags) built by the library
    (bsl_water_sites) from CM/Euler rows: model 0 SPC/E (3 sites per
    molecule), 1 TIP4P (4 sites)."""
    cm_ea = np.ascontiguousarray(cm_ea, dtype=np.float64)
    nmol = cm_ea.shape[0]
    ns = 3 if model == 0 else 4
    pos = np.zeros(nmol * ns * 3)
    q = np.zeros(nmol * ns)
    lj = np.zeros(nmol * ns, dtype=np.int32)
    check(lib().bsl_water_sites(model, nmol, cm_ea.ravel(), pos, q, lj))
    return pos.reshape(-1, 3), q, lj


def dodecahedron(radius):
    """Vertices (20x3) and edges (30 pairs) of a regular dodecahedron with the
    given circumradius."""
    phi = (1.0 + math.sqrt(5.0)) / 2.0
    v = []
    for sx in (-1, 1):
        for sy in (-1, 1):
            for sz in (-1, 1):
                v.append((sx, sy, sz))
    for s1 in (-1, 1):
        for s2 in (-1, 1):
            v.append((0.0, s1 / phi, s2 * phi))
            v.append((s1 / phi, s2 * phi, 0.0))
            v.append((s1 * phi, 0.0, s2 / phi))
    v = np.array(v, dtype=np.float64)
    v *= radius / math.sqrt(3.0)
    d = np.linalg.norm(v[:, None, :] - v[None, :, :], axis=-1)
    edge = d[d > 1e-9].min()
    edges = [(i, j) for i in range(20) for j in range(i + 1, 20) if abs(d[i, j] - edge) < 1e-6 * radius]
    assert len(edges) == 30
    return v, edges


def _rotate(vec, axis, angle):
    axis = axis / np.linalg.norm(axis)
    return (vec * math.cos(angle) + np.cross(axis, vec) * math.sin(angle)
            + axis * np.dot(axis, vec) * (1.0 - math.cos(angle)))


def synthetic_cage(seed=0, radius_ang=3.95, oh_ang=1.0, hoh_deg=109.47, max_tilt_deg=15.0):
    """Synthetic 5^12 cage: 20 rigid SPC/E waters with O on the vertices of a
    regular dodecahedron (circumradius radius_ang).  Each water donates its
    first H along a cage edge (each edge used once); the second H lies in the
    plane of that edge and the outward radial direction, at the HOH angle;
    the molecule is then turned about its O-H(edge) axis by a seeded angle
    uniform in [-max_tilt, max_tilt] degrees, which keeps the first H on the
    edge.  Returns (positions [60,3] nm in O,H,H order, charges, LJ flags)."""
    rng = np.random.default_rng(seed)
    verts, edges = dodecahedron(radius_ang)
    used = set()
    nbrs = {i: sorted([j for (a, b) in edges for j in (a, b) if i in (a, b) and j != i]) for i in range(20)}
    pos, q, lj = [], [], []
    hoh = math.radians(hoh_deg)
    for i in range(20):
        o = verts[i]
        target = None
        for j in nbrs[i]:
            e = (min(i, j), max(i, j))
            if e not in used:
                used.add(e)
                target = j
                break
        if target is None:
            target = nbrs[i][0]
        u1 = verts[target] - o
        u1 /= np.linalg.norm(u1)
        radial = o / np.linalg.norm(o)
        perp = radial - np.dot(radial, u1) * u1
        perp /= np.linalg.norm(perp)
        u2 = math.cos(hoh) * u1 + math.sin(hoh) * perp
        tilt = math.radians(rng.uniform(-max_tilt_deg, max_tilt_deg))
        u2 = _rotate(u2, u1, tilt)
        h1 = o + oh_ang * u1
        h2 = o + oh_ang * u2
        for site, charge, flag in ((o, physics.Q_O_SPCE, 1), (h1, physics.Q_H_SPCE, 0), (h2, physics.Q_H_SPCE, 0)):
            pos.append(site * 0.1)  # Angstrom -> nm
            q.append(charge)
            lj.append(flag)
    return np.array(pos), np.array(q), np.array(lj, dtype=np.int32)


def baseline_lj():
    """O-H2 LJ of the synthetic cage: O sigma 3.166 A eps 0.650 kJ/mol, H2
    sigma 2.96 A eps 0.2852 kJ/mol, combined with the reference's geometric
    rule."""
    return physics.lj_ab(0.650, 0.3166, 0.2852, 0.296, mixing="geometric")
