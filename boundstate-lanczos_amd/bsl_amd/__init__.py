"""bsl_amd -- Python host interface of libbsl_hip.so, the MI355X
implementation of the boundstate-lanczos Hv matvec and Lanczos recurrence.

The Calculator mirrors the reference's HvPrep/Hv plugin (an opaque prepared
operator with an accumulate-style Hv on reference-layout vectors) and adds
the device-resident Lanczos loops.  lancbis/trivec are the Cullum-Willoughby
post-processing of the reference driver."""
import ctypes as C

import numpy as np

from . import cage, physics  # noqa: F401
from ._lib import BASIS, CALC, POT, PREC, START, SYM, VFORM, BslError, Params, check, lib  # noqa: F401


class Calculator:
    """A prepared Hv operator on the GPU (bsl_create).

    Keyword arguments follow the Hv input deck of the reference
    (HvPrep_Internal_NoQuad, linRotCartSphHvFullVPrepRoutines.cpp:248-321):
    nx, ny, nz, x_max, y_max, z_max (nm, full box length), lmax, n_theta,
    n_phi, mass, rot_const, ceiling, symmetry ("All"/"Even"/"Odd"), plus
    basis ("full"/"reduced"), precision ("fp64"/"fp32"), potential
    ("iho"/"cage") and its parameters."""

    def __init__(self, **kw):
        p, keep = make_params(**kw)
        self._keep = keep
        self.params = p
        h = C.c_void_p()
        check(lib().bsl_create(C.byref(p), C.byref(h)))
        self._owned = True
        self._attach(h)

    def _attach(self, h):
        p = self.params
        self.shape = (p.nx, p.ny, p.nz)
        self.nn = (p.lmax + 1) ** 2
        self.h = h
        self.N = lib().bsl_basis_size(h)
        self.N_krylov = lib().bsl_krylov_size(h)

    @classmethod
    def _member(cls, h, params, owner):
        """Non-owning view of a batch member's context."""
        self = cls.__new__(cls)
        self.params, self._keep, self._owned, self._owner = params, [], False, owner
        self._attach(h)
        return self

    def close(self):
        if getattr(self, "h", None):
            if getattr(self, "_owned", False):
                lib().bsl_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # the reference's Hv contract: uec += H vec
    def hv(self, vec, uec=None):
        vec = np.ascontiguousarray(vec, dtype=np.float64)
        if vec.size != self.N:
            raise ValueError("vector length %d, basis size %d" % (vec.size, self.N))
        out = np.zeros(self.N) if uec is None else np.ascontiguousarray(uec, dtype=np.float64).copy()
        check(lib().bsl_hv(self.h, vec, out))
        return out

    def lanczos(self, niter, start="uniform", seed=0, v0=None):
        a, b, b2 = np.zeros(niter), np.zeros(niter + 1), np.zeros(niter + 1)
        ptr = None
        if v0 is not None:
            v0 = np.ascontiguousarray(v0, dtype=np.float64)
            ptr = v0.ctypes.data
        check(lib().bsl_lanczos(self.h, niter, ptr, START[start], seed, a, b, b2))
        return a, b, b2

    def ritz(self, niter, evtr, ngood, start="uniform", seed=0, v0=None):
        out = np.zeros(self.N * max(ngood, 1))
        ptr = None
        if v0 is not None:
            v0 = np.ascontiguousarray(v0, dtype=np.float64)
            ptr = v0.ctypes.data
        evtr = np.ascontiguousarray(evtr, dtype=np.float64)
        check(lib().bsl_ritz(self.h, niter, ptr, START[start], seed, evtr, ngood, out))
        return out[: self.N * ngood].reshape(ngood, self.N)

    def rayleigh(self, y):
        y = np.ascontiguousarray(np.atleast_2d(y), dtype=np.float64)
        e, r = np.zeros(len(y)), np.zeros(len(y))
        check(lib().bsl_rayleigh(self.h, len(y), y.ravel(), e, r))
        return e, r

    @property
    def mult_tol(self):
        """Multiplicity tolerance of the eigensolve for this operator: 0 (the
        reference's BISEC) for FP64, MIXED_MULT_TOL for FP32, BF16X3_MULT_TOL
        for the BF16x3 quadrature transforms and BF16X3_KIN_MULT_TOL for the
        BF16x3 kinetic contractions of the NoQuad calculator."""
        p = self.params.precision
        if p == PREC["fp64"]:
            return 0.0
        if p == PREC["fp32"]:
            return MIXED_MULT_TOL
        return BF16X3_MULT_TOL if self.params.calculator == CALC["quad"] else BF16X3_KIN_MULT_TOL

    def levels(self, niter, count, emin=-1.0e4, emax=1.0e4, start="random", seed=0, tol=1e-8, vectors=True):
        """The `count` lowest converged good levels: Loop 1, Cullum-Willoughby,
        Tm eigenvectors, then Loop 2 rebuilds their Ritz vectors by
        recomputing the Krylov vectors (lanczosEigenValuesVectors.cpp:
        170-242, 251-295, 329-469), and an FP64 Rayleigh quotient refines
        each level.  The Sturm bisections and error estimates run on the GPU
        (the good eigenvalues of the whole window: 0.04 s against 2.1 s on
        the host for 4000 steps at 41^3).  Returns (energies, unit Ritz
        vectors (count, N), residual norms); vectors=False leaves the Ritz
        vectors on the device only (None in their place), where
        hf_gradient_ritz finds level i as index i."""
        a, b, b2 = self.lanczos(niter, start=start, seed=seed)
        ev, er = lancbis(a, b, b2, emin, emax, device=True, mult_tol=self.mult_tol)
        order = np.argsort(ev)[:count]
        if len(order) < count or not np.all(np.abs(er[order]) < tol):
            raise BslError(0, "the %d lowest levels are not all converged after %d steps" % (count, niter))
        sel = ev[order]
        evtr, _ = trivec(a, b, b2, sel, mult_tol=self.mult_tol)
        psi, e, r = self.ritz_refine(niter, evtr, count, start=start, seed=seed, vectors=vectors)
        if psi is not None:
            psi /= np.linalg.norm(psi, axis=1, keepdims=True)
        return e, psi, r

    def ritz_refine(self, niter, evtr, ngood, start="uniform", seed=0, v0=None, vectors=True):
        """Loop 2 and the FP64 Rayleigh refinement of its Ritz vectors on the
        device (bsl_ritz_refine): (Ritz vectors (ngood, N) or None, energies,
        residual norms)."""
        out = np.zeros(self.N * max(ngood, 1)) if vectors else None
        e, r = np.zeros(max(ngood, 1)), np.zeros(max(ngood, 1))
        ptr = None
        if v0 is not None:
            v0 = np.ascontiguousarray(v0, dtype=np.float64)
            ptr = v0.ctypes.data
        evtr = np.ascontiguousarray(evtr, dtype=np.float64)
        check(lib().bsl_ritz_refine(self.h, niter, ptr, START[start], seed, evtr, ngood,
                                    None if out is None else out.ctypes.data, e, r))
        vec = None if out is None else out[: self.N * ngood].reshape(ngood, self.N)
        return vec, e[:ngood], r[:ngood]

    def hf_gradient(self, psi, g):
        """Hellmann-Feynman gradient sum_k g[k] <psi_k|dV/dtheta|psi_k> of the
        cage potential (bsl_hf_gradient).  Returns a dict: site_q (n_sites),
        site_pos (n_sites, 3), q_h2_h, q_h2_cm, lj_a, lj_b."""
        psi = np.ascontiguousarray(np.atleast_2d(psi), dtype=np.float64)
        g = np.ascontiguousarray(np.atleast_1d(g), dtype=np.float64)
        if psi.shape[1] != self.N or len(g) != psi.shape[0]:
            raise ValueError("psi must be (len(g), basis size)")
        ns = self.params.n_sites
        out = np.zeros(4 * ns + 4)
        check(lib().bsl_hf_gradient(self.h, psi.shape[0], psi.ravel(), g, out))
        return split_gradient(out, ns)

    def hf_gradient_sets(self, sets):
        """The gradients of several sets at once (bsl_hf_gradient_sets): sets
        is a list of (psi (k, N), g (k)) pairs, e.g. one per degenerate
        multiplet; returns one dict per set, each equal to hf_gradient of
        that set alone (up to eight sets share a pass over the grid)."""
        psis, gs, counts = [], [], []
        for psi, g in sets:
            psi = np.ascontiguousarray(np.atleast_2d(psi), dtype=np.float64)
            g = np.ascontiguousarray(np.atleast_1d(g), dtype=np.float64)
            if psi.shape[1] != self.N or len(g) != psi.shape[0]:
                raise ValueError("each psi must be (len(g), basis size)")
            psis.append(psi)
            gs.append(g)
            counts.append(psi.shape[0])
        ns = self.params.n_sites
        out = np.zeros(len(sets) * (4 * ns + 4))
        cnt = np.asarray(counts, dtype=np.int32)
        check(lib().bsl_hf_gradient_sets(self.h, len(sets), cnt.ctypes.data, np.concatenate(psis).ravel(),
                                         np.concatenate(gs), out))
        return [split_gradient(out[i * (4 * ns + 4):(i + 1) * (4 * ns + 4)], ns) for i in range(len(sets))]

    def hf_gradient_ritz(self, sets):
        """The gradients of sets of the unit Ritz vectors that the last
        levels() / ritz_refine() left on the device (bsl_hf_gradient_ritz):
        sets is a list of (indices, g) pairs (index i: the i-th level or
        evtr entry of that call), e.g. one per multiplet; one dict per set."""
        idx, gs, counts = [], [], []
        for ind, g in sets:
            ind = np.atleast_1d(np.asarray(ind, dtype=np.int32))
            g = np.atleast_1d(np.asarray(g, dtype=np.float64))
            if len(ind) != len(g) or len(ind) == 0:
                raise ValueError("each set needs as many weights as indices, at least one")
            idx.append(ind)
            gs.append(g)
            counts.append(len(ind))
        ns = self.params.n_sites
        out = np.zeros(len(sets) * (4 * ns + 4))
        cnt = np.asarray(counts, dtype=np.int32)
        ind = np.ascontiguousarray(np.concatenate(idx), dtype=np.int32)
        check(lib().bsl_hf_gradient_ritz(self.h, len(sets), cnt.ctypes.data, ind.ctypes.data,
                                         np.ascontiguousarray(np.concatenate(gs)), out))
        return [split_gradient(out[i * (4 * ns + 4):(i + 1) * (4 * ns + 4)], ns) for i in range(len(sets))]

    def vblocks(self):
        nx, ny, nz = self.shape
        out = np.zeros(nx * ny * nz * self.nn * self.nn)
        check(lib().bsl_get_vblocks(self.h, out))
        return out.reshape(nx, ny, nz, self.nn, self.nn)

    def vgrid(self):
        nx, ny, nz = self.shape
        na, nb = self.params.n_theta, self.params.n_phi
        v1, v2 = np.zeros(nx * ny * nz * na * nb), np.zeros(nx * ny * nz * na * nb)
        check(lib().bsl_get_vgrid(self.h, v1, v2))
        return v1.reshape(nx, ny, nz, na, nb), v2.reshape(nx, ny, nz, na, nb)

    def time_steps(self, warmup, steps):
        ms = np.zeros(4)
        check(lib().bsl_time_steps(self.h, warmup, steps, ms))
        return ms

    def step_bytes(self):
        out = np.zeros(5)
        check(lib().bsl_step_bytes(self.h, out))
        return out

    @property
    def potential_form(self):
        """"blocks" (packed <lm|V|l'm'> blocks per point, as the reference
        stores them) or "harmonics" (their even-L harmonic coefficients,
        expanded by the kernel through the Gaunt coefficients)."""
        r = lib().bsl_get_potential_form(self.h)
        if r < 0:
            check(r)
        return {v: k for k, v in VFORM.items()}[r]

    @potential_form.setter
    def potential_form(self, form):
        """Rebuilds the potential in the other form (clears the Krylov vectors)."""
        check(lib().bsl_set_potential_form(self.h, VFORM[form]))

    @property
    def fused(self):
        """True when a NoQuad step runs the z kinetic term inside the
        potential kernel (x-y kernel, then one fused kernel), False for the
        three-kernel step, the default (bsl_get_fusion)."""
        r = lib().bsl_get_fusion(self.h)
        if r < 0:
            check(r)
        return bool(r)

    @fused.setter
    def fused(self, on):
        """Selects the step structure; raises where the fused kernel does not
        apply (bsl_set_fusion)."""
        check(lib().bsl_set_fusion(self.h, 1 if on else 0))

    def unclamped_fraction(self):
        """Fraction of the (point, node) pairs of the potential grid below the
        ceiling clamp (cage potential; bsl_clamp_stats)."""
        u, t = C.c_int64(), C.c_int64()
        check(lib().bsl_clamp_stats(self.h, C.byref(u), C.byref(t)))
        return u.value / t.value

    def potential_kernel(self):
        """Name of the potential kernel (and autotuned variant) in use."""
        buf = C.create_string_buffer(128)
        check(lib().bsl_potential_kernel(self.h, buf, 128))
        return buf.value.decode()


def gaunt_table(lb, par):
    """Harmonic-form term list of parity block (lb, par) (bsl_gaunt_table):
    (ijc [nt, 3] int32: row i, column j of the block, coefficient column c;
    g [nt]: the Gaunt value)."""
    nt = lib().bsl_gaunt_table(lb, par, 0, None, None)
    if nt < 0:
        check(-nt)
    ijc = np.zeros((nt, 3), dtype=np.int32)
    g = np.zeros(nt)
    r = lib().bsl_gaunt_table(lb, par, nt, ijc.ctypes.data, g.ctypes.data)
    if r != nt:
        raise BslError(7, "bsl_gaunt_table returned %d for %d terms" % (r, nt))
    return ijc, g


def make_params(*, nx, ny=None, nz=None, x_max, y_max=None, z_max=None, lmax, n_theta, n_phi, mass, rot_const,
                ceiling, symmetry="All", basis="full", calculator="noquad", precision="fp64", potential="iho",
                iho_factor=0.0, iho_ceil=0.0, sites=None, q_h2_h=physics.Q_H2_H, q_h2_cm=physics.Q_H2_CM,
                h2_half_bond=physics.H2_HALF_BOND, lj_a=physics.LJ_A_REF, lj_b=physics.LJ_B_REF, universe=None,
                coulomb=None):
    """bsl_params from keyword arguments; returns (params, arrays to keep alive).
    potential: "iho", "cage" (pair sums at the exact H positions), "cage_grid"
    (universe=(pnx, pny, pnz, px_max, py_max, pz_max): tabulated potentials
    looked up at the nearest universe point) or "coulomb"
    (coulomb=(q_centre, q_mol, v_ceil))."""
    p = Params()
    p.nx, p.ny, p.nz = nx, ny or nx, nz or nx
    p.x_max, p.y_max, p.z_max = x_max, y_max or x_max, z_max or x_max
    p.lmax, p.n_theta, p.n_phi = lmax, n_theta, n_phi
    p.mass, p.rot_const, p.ceiling = mass, rot_const, ceiling
    p.symmetry = SYM[symmetry]
    p.basis = BASIS[basis]
    p.calculator = CALC[calculator]
    p.precision = PREC[precision]
    p.potential = POT[potential]
    p.iho_factor, p.iho_ceil = iho_factor, iho_ceil
    keep = []
    if potential == "cage_grid":
        p.pnx, p.pny, p.pnz = universe[0], universe[1], universe[2]
        p.px_max, p.py_max, p.pz_max = universe[3], universe[4], universe[5]
    if potential == "coulomb":
        p.coul_q_centre, p.coul_q_mol, p.coul_ceil = coulomb
    if potential in ("cage", "cage_grid"):
        pos, q, lj = sites
        pos = np.ascontiguousarray(pos, dtype=np.float64).ravel()
        q = np.ascontiguousarray(q, dtype=np.float64)
        lj = np.ascontiguousarray(lj, dtype=np.int32)
        keep = [pos, q, lj]
        p.n_sites = len(q)
        p.site_pos, p.site_q, p.site_lj = pos.ctypes.data, q.ctypes.data, lj.ctypes.data
    p.q_h2_h, p.q_h2_cm, p.h2_half_bond, p.lj_a, p.lj_b = q_h2_h, q_h2_cm, h2_half_bond, lj_a, lj_b
    return p, keep


class Batch:
    """Independent recurrences advanced in lockstep on one GPU (bsl_batch_*):
    `members` is a list of Calculator keyword dicts (shapes, symmetry blocks
    and precisions may differ).  Member i starts from seed + i."""

    def __init__(self, members):
        built = [make_params(**m) for m in members]
        self._keep = [k for _, k in built]
        arr = (Params * len(built))(*[p for p, _ in built])
        self._params = [p for p, _ in built]
        h = C.c_void_p()
        check(lib().bsl_batch_create(len(built), arr, C.byref(h)))
        self.h = h
        self.n = len(built)

    def close(self):
        if getattr(self, "h", None):
            lib().bsl_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def member(self, i):
        h = C.c_void_p()
        check(lib().bsl_batch_member(self.h, i, C.byref(h)))
        return Calculator._member(h, self._params[i], self)

    def lanczos(self, niter, start="random", seed=0, graph=True):
        """Per member (alpha, beta, beta2), as Calculator.lanczos."""
        a = np.zeros((self.n, niter))
        b = np.zeros((self.n, niter + 1))
        b2 = np.zeros((self.n, niter + 1))
        check(lib().bsl_batch_lanczos(self.h, niter, START[start], seed, 1 if graph else 0, a, b, b2))
        return [(a[i], b[i], b2[i]) for i in range(self.n)]

    def time_steps(self, warmup, steps, graph=True):
        """Wall time (ms) of `steps` lockstep iterations of all members."""
        ms = np.zeros(1)
        check(lib().bsl_batch_time_steps(self.h, warmup, steps, 1 if graph else 0, ms))
        return float(ms[0])


def split_gradient(grad, ns):
    """Named parts of a bsl_hf_gradient vector."""
    sg = grad[: 4 * ns].reshape(ns, 4)
    return {"site_q": sg[:, 0].copy(), "site_pos": sg[:, 1:4].copy(), "q_h2_h": float(grad[4 * ns]),
            "q_h2_cm": float(grad[4 * ns + 1]), "lj_a": float(grad[4 * ns + 2]), "lj_b": float(grad[4 * ns + 3])}


def multiplets(energies, tol=1e-6):
    """Groups of indices of (near-)degenerate levels: consecutive sorted
    levels closer than tol (kJ/mol) share a group."""
    order = np.argsort(energies)
    groups, cur = [], [int(order[0])]
    for i, j in zip(order[:-1], order[1:]):
        if energies[j] - energies[i] < tol:
            cur.append(int(j))
        else:
            groups.append(cur)
            cur = [int(j)]
    groups.append(cur)
    return groups


# Multiplicity tolerances (kJ/mol) of the eigensolve for the mixed-precision
# operators (BSL_MIXED_MULT_TOL, BSL_BF16X3_MULT_TOL of bsl_hip.h): their
# rounding spreads the ghost copies of a converged level far beyond FP64's
# BISEC width.  FP32: ~1e-8 relative (3e-8 kJ/mol at 54^3 x 36); 1e-6 kJ/mol
# (8e-5 cm^-1) merges them, two orders below the 0.01 cm^-1 accuracy of the
# mixed levels.  BF16x3 quadrature transforms (two-term bf16 splits, ~16
# significant bits): ~1e-6 relative (6.6e-6 kJ/mol measured); 3e-5 kJ/mol
# (2.5e-3 cm^-1), the level accuracy of that precision.  BF16x3 kinetic
# contractions (the NoQuad BF16x3 operator): levels shift by up to 1.5e-2
# cm^-1 at 54^3; 1e-4 kJ/mol, at which TRIVEC converges for the lowest
# levels and their FP64 refinement is within 1e-4 cm^-1
# (BSL_BF16X3_KIN_MULT_TOL).
MIXED_MULT_TOL = 1.0e-6
BF16X3_MULT_TOL = 3.0e-5
BF16X3_KIN_MULT_TOL = 1.0e-4


def lancbis(alpha, beta, beta2, emin, emax, device=False, mult_tol=0.0):
    """Good eigenvalues and error estimates (lancbis.cpp:5-105); device=True
    runs the Sturm bisections and inverse iterations on the GPU.  mult_tol
    > 0 merges eigenvalues of T_m closer than that into one multiple
    eigenvalue (bsl_lancbis_tol: for the coefficients of a mixed-precision
    operator, see MIXED_MULT_TOL); 0 is the reference's BISEC."""
    niter = len(alpha)
    ev, er = np.zeros(niter), np.zeros(niter)
    ng = C.c_int32()
    fn = lib().bsl_lancbis_device_tol if device else lib().bsl_lancbis_tol
    check(fn(niter, np.ascontiguousarray(alpha, dtype=np.float64),
             np.ascontiguousarray(beta, dtype=np.float64),
             np.ascontiguousarray(beta2, dtype=np.float64), emin, emax, float(mult_tol), ev, er, C.byref(ng)))
    return ev[: ng.value].copy(), er[: ng.value].copy()


def trivec(alpha, beta, beta2, evals, mult_tol=0.0):
    """Tm eigenvectors, column-major [niter][ngood] flattened (trivec.f);
    mult_tol as given to lancbis for these eigenvalues (bsl_trivec_tol)."""
    niter, ng = len(alpha), len(evals)
    evtr = np.zeros(niter * max(ng, 1))
    mamin = np.zeros(max(ng, 1), dtype=np.int32)
    check(lib().bsl_trivec_tol(niter, np.ascontiguousarray(alpha, dtype=np.float64),
                               np.ascontiguousarray(beta, dtype=np.float64),
                               np.ascontiguousarray(beta2, dtype=np.float64),
                               np.ascontiguousarray(evals, dtype=np.float64), ng, float(mult_tol), evtr, mamin))
    return evtr[: niter * ng], mamin[:ng].copy()
