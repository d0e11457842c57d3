"""ctypes bindings of the CPU oracle (oracle/_build/libbsl_oracle.so) and of the
reference harness (oracle/_ref/libref_hv.so, only where /root/reference was
present at build time).  Test infrastructure only: used as the checker."""
import ctypes as C
import os

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(REPO, "oracle", "_build", "libbsl_oracle.so")
REF_SO = os.path.join(REPO, "oracle", "_ref", "libref_hv.so")

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


class Problem(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("x_max", C.c_double), ("y_max", C.c_double), ("z_max", C.c_double),
        ("lmax", C.c_int), ("na", C.c_int), ("nb", C.c_int),
        ("mass", C.c_double), ("rot_const", C.c_double), ("ceiling", C.c_double),
        ("sym", C.c_int), ("pot_kind", C.c_int),
        ("iho_factor", C.c_double), ("iho_ceil", C.c_double),
        ("nsites", C.c_int),
        ("site_pos", C.c_void_p), ("site_q", C.c_void_p), ("site_type", C.c_char_p),
        ("q_h2_h", C.c_double), ("q_h2_cm", C.c_double), ("h2_half_bond", C.c_double),
        ("lj_a", C.c_double), ("lj_b", C.c_double),
        ("pnx", C.c_int), ("pny", C.c_int), ("pnz", C.c_int),
        ("px_max", C.c_double), ("py_max", C.c_double), ("pz_max", C.c_double),
        ("coul_q_centre", C.c_double), ("coul_q_mol", C.c_double), ("coul_ceil", C.c_double),
    ]


def _load():
    if not os.path.exists(ORACLE_SO):
        raise RuntimeError("oracle not built: run `make -C oracle` or __graft_entry__.build()")
    lib = C.CDLL(ORACLE_SO)
    lib.bso_prepare.restype = C.c_void_p
    lib.bso_prepare.argtypes = [C.POINTER(Problem)]
    lib.bso_prepare_external.restype = C.c_void_p
    lib.bso_prepare_external.argtypes = [C.POINTER(Problem), _dp]
    lib.bso_free.argtypes = [C.c_void_p]
    lib.bso_basis_size.restype = C.c_longlong
    lib.bso_basis_size.argtypes = [C.c_void_p]
    lib.bso_nn.restype = C.c_int
    lib.bso_nn.argtypes = [C.c_void_p]
    lib.bso_hv.argtypes = [C.c_void_p, _dp, _dp]
    lib.bso_get_vblocks.argtypes = [C.c_void_p, _dp]
    lib.bso_get_vgrid.argtypes = [C.c_void_p, _dp, _dp]
    lib.bso_get_kin.argtypes = [C.c_void_p, C.c_int, _dp]
    lib.bso_get_grid.argtypes = [C.c_void_p, C.c_int, _dp]
    lib.bso_get_rot.argtypes = [C.c_void_p, _dp]
    lib.bso_get_sym.argtypes = [C.c_void_p, _dp]
    lib.bso_get_qnum.argtypes = [C.c_void_p, _ip]
    lib.bso_get_quad.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp]
    lib.bso_get_tess.argtypes = [C.c_void_p, _dp, _dp, _dp]
    lib.bso_lanczos.argtypes = [C.c_void_p, C.c_int, C.c_void_p, _dp, _dp, _dp]
    lib.bso_start_vector.argtypes = [C.c_longlong, C.c_uint64, _dp]
    lib.bso_lancbis.restype = C.c_int
    lib.bso_lancbis.argtypes = [C.c_int, _dp, _dp, _dp, C.c_double, C.c_double, _dp, _dp]
    lib.bso_bisec.restype = C.c_int
    lib.bso_bisec.argtypes = [C.c_int, _dp, _dp, C.c_double, C.c_double, _dp, _ip]
    lib.bso_trivec.restype = C.c_int
    lib.bso_trivec.argtypes = [C.c_int, _dp, _dp, _dp, _dp, C.c_int, _dp, _ip]
    lib.bso_ritz.argtypes = [C.c_void_p, C.c_int, C.c_void_p, _dp, C.c_int, _dp]
    lib.bso_genran.argtypes = [C.c_int, _dp, C.c_int]
    lib.bso_hf_gradient.argtypes = [C.c_void_p, C.c_int, _dp, _dp, _dp]
    lib.bso_gen_indices_lm.restype = C.c_int
    lib.bso_gen_indices_lm.argtypes = [C.c_int, _ip]
    lib.bso_gauleg.argtypes = [C.c_double, C.c_double, _dp, _dp, C.c_int]
    lib.bso_gauss_chebyshev.argtypes = [C.c_int, _dp, _dp]
    lib.bso_norm_assoc_legendre.argtypes = [C.c_int, _ip, C.c_int, C.c_double, _dp]
    lib.bso_tesseral_trig.argtypes = [_ip, C.c_int, C.c_double, _dp]
    lib.bso_cart_kin_grid.argtypes = [C.c_double, C.c_int, C.c_double, _dp, _dp]
    lib.bso_build_sites.restype = C.c_int
    lib.bso_build_sites.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp, C.c_char_p]
    lib.bso_coulomb_sum.restype = C.c_double
    lib.bso_coulomb_sum.argtypes = [_dp, C.c_double, C.c_int, _dp, _dp]
    lib.bso_lj_sum.restype = C.c_double
    lib.bso_lj_sum.argtypes = [_dp, C.c_int, _dp, C.c_char_p, C.c_double, C.c_double]
    return lib


lib = _load()


def build_sites(cm_ea, model=0):
    cm_ea = np.ascontiguousarray(cm_ea, dtype=np.float64)
    nmol = cm_ea.shape[0]
    ns = 3 if model == 0 else 4
    pos = np.zeros(nmol * ns * 3)
    q = np.zeros(nmol * ns)
    typ = C.create_string_buffer(nmol * ns)
    n = lib.bso_build_sites(model, nmol, cm_ea.ravel(), pos, q, typ)
    return pos.reshape(n, 3), q, typ.raw[:n]


class Oracle:
    """A prepared oracle operator (bso_prepare)."""

    def __init__(self, spec, vblocks=None):
        self.spec = dict(spec)
        p = Problem()
        for f in ("nx", "ny", "nz", "lmax", "na", "nb", "sym", "pot_kind"):
            setattr(p, f, int(spec[f]))
        for f in ("x_max", "y_max", "z_max", "mass", "rot_const", "ceiling"):
            setattr(p, f, float(spec[f]))
        p.iho_factor = float(spec.get("iho_factor", 0.0))
        p.iho_ceil = float(spec.get("iho_ceil", 0.0))
        self._keep = []
        if spec["pot_kind"] == 2:
            p.pnx, p.pny, p.pnz = [int(x) for x in spec["universe"][:3]]
            p.px_max, p.py_max, p.pz_max = [float(x) for x in spec["universe"][3:]]
        if spec["pot_kind"] == 3:
            p.coul_q_centre, p.coul_q_mol, p.coul_ceil = [float(x) for x in spec["coulomb"]]
        if spec["pot_kind"] in (1, 2):
            pos = np.ascontiguousarray(spec["site_pos"], dtype=np.float64).ravel()
            q = np.ascontiguousarray(spec["site_q"], dtype=np.float64)
            typ = C.create_string_buffer(bytes(spec["site_type"]), len(spec["site_type"]))
            self._keep += [pos, q, typ]
            p.nsites = len(q)
            p.site_pos = pos.ctypes.data
            p.site_q = q.ctypes.data
            p.site_type = C.cast(typ, C.c_char_p)
            for f in ("q_h2_h", "q_h2_cm", "h2_half_bond", "lj_a", "lj_b"):
                setattr(p, f, float(spec[f]))
        self._p = p
        if vblocks is None:
            self.h = lib.bso_prepare(C.byref(p))
        else:
            self.h = lib.bso_prepare_external(C.byref(p), np.ascontiguousarray(vblocks, dtype=np.float64).ravel())
        if not self.h:
            raise ValueError("bso_prepare rejected the problem")
        self.N = lib.bso_basis_size(self.h)
        self.nn = lib.bso_nn(self.h)
        self.P = spec["nx"] * spec["ny"] * spec["nz"]

    def __del__(self):
        if getattr(self, "h", None):
            lib.bso_free(self.h)
            self.h = None

    def hv(self, v):
        v = np.ascontiguousarray(v, dtype=np.float64)
        u = np.zeros(self.N)
        lib.bso_hv(self.h, v, u)
        return u

    def vblocks(self):
        out = np.zeros(self.P * self.nn * self.nn)
        lib.bso_get_vblocks(self.h, out)
        return out.reshape(self.spec["nx"], self.spec["ny"], self.spec["nz"], self.nn, self.nn)

    def vgrid(self):
        s = self.spec
        n = self.P * s["na"] * s["nb"]
        v1, v2 = np.zeros(n), np.zeros(n)
        lib.bso_get_vgrid(self.h, v1, v2)
        shp = (s["nx"], s["ny"], s["nz"], s["na"], s["nb"])
        return v1.reshape(shp), v2.reshape(shp)

    def kin(self, axis):
        n = self.spec[("nx", "ny", "nz")[axis]]
        out = np.zeros(n * n)
        lib.bso_get_kin(self.h, axis, out)
        return out.reshape(n, n)

    def grid(self, axis):
        n = self.spec[("nx", "ny", "nz")[axis]]
        out = np.zeros(n)
        lib.bso_get_grid(self.h, axis, out)
        return out

    def rot(self):
        out = np.zeros(self.nn)
        lib.bso_get_rot(self.h, out)
        return out

    def sym(self):
        out = np.zeros(self.nn)
        lib.bso_get_sym(self.h, out)
        return out

    def qnum(self):
        out = np.zeros(2 * self.nn, dtype=np.int32)
        lib.bso_get_qnum(self.h, out)
        return out.reshape(self.nn, 2)

    def quad(self):
        na, nb = self.spec["na"], self.spec["nb"]
        ct, wt, cp, wp = np.zeros(na), np.zeros(na), np.zeros(nb), np.zeros(nb)
        lib.bso_get_quad(self.h, ct, wt, cp, wp)
        return ct, wt, cp, wp

    def tess(self):
        na, nb, lmax = self.spec["na"], self.spec["nb"], self.spec["lmax"]
        nm = 2 * lmax + 1
        L, S1, S2 = np.zeros(self.nn * na), np.zeros(nm * nb), np.zeros(nm * nb)
        lib.bso_get_tess(self.h, L, S1, S2)
        return L.reshape(self.nn, na), S1.reshape(nm, nb), S2.reshape(nm, nb)

    def lanczos(self, niter, v0=None):
        a = np.zeros(niter)
        b = np.zeros(niter + 1)
        b2 = np.zeros(niter + 1)
        if v0 is not None:
            v0 = np.ascontiguousarray(v0, dtype=np.float64)
            lib.bso_lanczos(self.h, niter, v0.ctypes.data, a, b, b2)
        else:
            lib.bso_lanczos(self.h, niter, None, a, b, b2)
        return a, b, b2

    def ritz(self, niter, evtr, ngood, v0=None):
        out = np.zeros(self.N * ngood)
        evtr = np.ascontiguousarray(evtr, dtype=np.float64)
        ptr = None if v0 is None else np.ascontiguousarray(v0, dtype=np.float64).ctypes.data
        lib.bso_ritz(self.h, niter, ptr, evtr, ngood, out)
        return out.reshape(ngood, self.N)


    def hf_gradient(self, psi, g):
        """sum_k g[k] dE_k/dtheta (bso_hf_gradient); psi: (nvec, N)."""
        psi = np.ascontiguousarray(np.atleast_2d(psi), dtype=np.float64)
        g = np.ascontiguousarray(np.atleast_1d(g), dtype=np.float64)
        out = np.zeros(4 * len(self.spec["site_q"]) + 4)
        lib.bso_hf_gradient(self.h, psi.shape[0], psi.ravel(), g, out)
        return out


def start_vector(n, seed):
    v = np.zeros(n)
    lib.bso_start_vector(n, seed, v)
    return v


def lancbis(alpha, beta, beta2, emin, emax):
    niter = len(alpha)
    ev = np.zeros(max(niter, 1))
    er = np.zeros(max(niter, 1))
    ng = lib.bso_lancbis(niter, np.ascontiguousarray(alpha), np.ascontiguousarray(beta),
                         np.ascontiguousarray(beta2), emin, emax, ev, er)
    if ng < 0:
        raise RuntimeError("BISEC: estimated number of Sturms exceeds limit")
    return ev[:ng].copy(), er[:ng].copy()


def bisec(alpha, beta, emin, emax):
    niter = len(alpha)
    vs = np.zeros(niter)
    mp = np.zeros(niter, dtype=np.int32)
    nd = lib.bso_bisec(niter, np.ascontiguousarray(alpha), np.ascontiguousarray(beta), emin, emax, vs, mp)
    return vs[:nd].copy(), mp[:nd].copy()


def trivec(alpha, beta, beta2, evals):
    niter = len(alpha)
    ng = len(evals)
    evtr = np.zeros(niter * ng)
    mamin = np.zeros(max(ng, 1), dtype=np.int32)
    st = lib.bso_trivec(niter, np.ascontiguousarray(alpha), np.ascontiguousarray(beta),
                        np.ascontiguousarray(beta2), np.ascontiguousarray(evals, dtype=np.float64), ng,
                        evtr, mamin)
    if st != 0:
        raise RuntimeError("TRIVEC failed (%d)" % st)
    return evtr, mamin[:ng].copy()


def genran(seed, n):
    r = np.zeros(n)
    lib.bso_genran(seed, r, n)
    return r


# --------------------------------------------------------------------------
# reference harness (oracle/_ref: built from the reference sources in the
# build container; the built library travels with the tree)
# --------------------------------------------------------------------------

def write_ref_hv_deck(path, nx, x_max, lmax, na, nb, mass, rot_const, geo, ceiling, sym):
    """The reference's 22-token Hv deck (HvPrep_Internal_NoQuad,
    linRotCartSphHvFullVPrepRoutines.cpp:248-321): a cubic grid; the
    grid-mapped universe fields are unused by the ContGrid calculators but
    must contain the box (:330-348)."""
    u = max(0.48, float(x_max))
    vals = [nx, x_max, nx, x_max, nx, x_max, lmax, na, nb, 2, u, 2, u, 2, u, mass, rot_const, geo,
            "deck", ceiling, "FALSE", sym]
    with open(path, "w") as f:
        f.write("#\n#\n")
        for i, v in enumerate(vals):
            f.write("v%d= %s\n" % (i, repr(float(v)) if isinstance(v, float) else v))


def ref_available():
    return os.path.exists(REF_SO)


# the reference's own Lanczos program (lanczosEigenValuesVectors.cpp), built
# from source into oracle/_ref by oracle/Makefile.ref
REF_LANCZOS = os.path.join(REPO, "oracle", "_ref", "ref_lanczos")
# the same program linked with the HIP plugin binding (oracle/ref_bind.cpp)
REF_LANCZOS_HIP = os.path.join(REPO, "oracle", "_ref", "ref_lanczos_hip")


def run_ref_lanczos(lanczos_deck, hv_deck, cwd, pot_deck=None, program=None):
    """Runs the reference program (or `program`, built from the same
    sources) on its decks in cwd; returns its stdout alpha, beta (Loop 1
    coefficients as printed, 16 digits) and the output directory it made."""
    import subprocess
    args = [program or REF_LANCZOS, lanczos_deck, hv_deck] + ([pot_deck] if pot_deck else [])
    r = subprocess.run(args, cwd=cwd, capture_output=True, text=True, timeout=600,
                       env=dict(os.environ, OMP_NUM_THREADS="4"))
    if r.returncode != 0:
        raise RuntimeError(r.stderr[-2000:])
    lines = r.stdout.split("\n")
    k = lines.index("Iteration j alpha_j beta_j") + 1
    a, b = [], []
    while k < len(lines):
        t = lines[k].split()
        if len(t) != 3 or not t[0].isdigit():
            break
        a.append(float(t[1]))
        b.append(float(t[2]))
        k += 1
    out = [l.split(": ", 1)[1] for l in lines if l.startswith("The output directory is: ")]
    return np.array(a), np.array(b), out[0]


def ref_lib():
    lib = C.CDLL(REF_SO)
    lib.ref_setup.restype = C.c_longlong
    lib.ref_setup.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p]
    lib.ref_hv.argtypes = [_dp, _dp]
    lib.ref_vblocks.restype = C.c_int
    lib.ref_vblocks.argtypes = [_dp, C.c_longlong]
    lib.ref_set_vblocks.restype = C.c_int
    lib.ref_set_vblocks.argtypes = [_dp, C.c_longlong]
    lib.ref_kin.argtypes = [C.c_int, _dp, _dp]
    lib.ref_quad.argtypes = [_dp, _dp, _dp, _dp, _dp, _dp, _dp]
    lib.ref_spce_sites.restype = C.c_int
    lib.ref_spce_sites.argtypes = [C.c_char_p, _dp, C.c_char_p]
    lib.ref_lancbis.restype = C.c_int
    lib.ref_lancbis.argtypes = [C.c_int, _dp, _dp, _dp, C.c_double, C.c_double, _dp, _dp, C.c_int]
    lib.ref_trivec.argtypes = [C.c_int, _dp, _dp, _dp, C.c_int, _dp, _dp, _dp]
    lib.ref_genran.argtypes = [C.c_int, _dp, C.c_int]
    lib.ref_spce_energy.restype = C.c_double
    lib.ref_spce_energy.argtypes = [C.c_char_p, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int]
    return lib
