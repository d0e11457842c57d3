"""Lanczos steps/s of the MI355X Hv + recurrence at a 54^3 x 36 basis.

Workload (one GPU): Alavi H2 in a synthetic 5^12 cage of 20 rigid SPC/E
waters (seed 0), sinc-DVR grid of 54 points per axis on a box of full
length x_max (default 0.4 nm, the reference's production decks' box:
[-0.2, 0.2] nm, spacing 0.0075 nm), tesseral harmonics l <= 5 (36 lm, full
basis), 40 x 70 (theta, phi) quadrature, NoQuad potential (its per-point
blocks held as even-L harmonic coefficients and expanded through the Gaunt
coefficients in the kernel), FP64 Hv and FP64 recurrence.  A step is one Hv
plus one update of the three-term recurrence (folded into the Hv kernels
plus a one-workgroup finalize); inputs (potential coefficients, Krylov
vectors) are resident in HBM when the timed region starts.

Also reported: the FP32-potential ("mixed") step rate and its ratio to FP64,
the error of the mixed runs' lowest levels against the FP64 run (cm^-1) at
10000 and 25000 steps, on this box and on two others (the same grid on the
0.8 nm box, and the production decks' own 28^3 grid at 0.4 nm), the
quadrature calculator at FP64, FP32 and BF16x3 with its level errors, the
32-member batch, the adjoint leg on both boxes, the roofline of the dominant
kernel and the reference's CPU path timed on this host on a bounded sample.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import contextlib
import ctypes
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "boundstate-lanczos_amd"))

HBM_PEAK_GBS = 8000.0


@contextlib.contextmanager
def _stdout_to_stderr():
    """Route file descriptor 1 to stderr (the reference's preparation prints
    its progress on stdout; stdout carries only the result line)."""
    sys.stdout.flush()
    saved = os.dup(1)
    os.dup2(2, 1)
    try:
        yield
    finally:
        sys.stdout.flush()
        ctypes.CDLL(None).fflush(None)
        os.dup2(saved, 1)
        os.close(saved)


def make_calc(B, args, precision, basis="full", symmetry="All", calculator="noquad", n=None, x_max=None):
    from bsl_amd import cage, physics
    pos, q, lj = cage.synthetic_cage(0)
    lj_a, lj_b = cage.baseline_lj()
    return B.Calculator(nx=n or args.n, x_max=x_max or args.x_max, lmax=5, n_theta=40, n_phi=70, mass=2.016,
                        rot_const=physics.cm1_to_kjmol(59.322), ceiling=18.0, symmetry=symmetry, basis=basis,
                        precision=precision, potential="cage", sites=(pos, q, lj), lj_a=lj_a, lj_b=lj_b,
                        calculator=calculator)


def quad_leg(B, args, e_ref):
    """On-the-fly quadrature calculator at the same sizes: FP64, FP32 and
    BF16x3 MFMA transforms over the resident weighted V grid; the lowest 20
    levels of each after --quad-eig-steps steps against the FP64 NoQuad
    levels `e_ref` of the same step count (cm^-1)."""
    out = {}
    for prec in ("fp64", "fp32", "bf16x3"):
        t0 = time.perf_counter()
        c = make_calc(B, args, prec, calculator="quad")
        prep = time.perf_counter() - t0
        ms = c.time_steps(args.warmup, args.quad_steps)
        sb = c.step_bytes()
        sps = args.quad_steps / (ms[0] / 1e3)
        ach = sb[0] / (ms[1] * 1e-3) / 1e9
        out[prec] = {"steps_per_s": sps, "prep_s": prep, "potential_ms": ms[1], "kinetic_ms": ms[2],
                     "finalize_ms": ms[3], "kernel": c.potential_kernel(), "algorithmic_bytes_per_launch": sb[0],
                     "achieved_GBs": ach, "hbm_frac": ach / HBM_PEAK_GBS,
                     "flops_per_launch": sb[4] - 2.0 * 3 * args.n * c.N - 4.0 * c.N,
                     "tflops": (sb[4] - 2.0 * 3 * args.n * c.N - 4.0 * c.N) / (ms[1] * 1e-3) / 1e12}
        # algorithmic FLOP/s against the dense MFMA peak of the instruction the
        # mode runs on (BF16x3: three BF16 products per FP32 one)
        peak = {"fp64": 78.6, "fp32": 157.3, "bf16x3": 2500.0 / 3.0}[prec]
        out[prec]["mfma_peak_tflops"] = peak
        out[prec]["mfma_frac"] = out[prec]["tflops"] / peak
        if e_ref is not None and args.quad_eig_steps > 0:
            co = c.lanczos(args.quad_eig_steps, start="random", seed=0)
            e, _ = B.lancbis(*co, -10.0, 10.0, mult_tol=c.mult_tol)
            out[prec]["max_err_cm1_vs_noquad_fp64"] = level_errors(e_ref, e)
            out[prec]["eig_steps"] = args.quad_eig_steps
        print("quad %s: prep %.2f s, %.1f steps/s, potential %.3f ms (%.0f GB/s, %.1f TFLOP/s), level err %s cm^-1"
              % (prec, prep, sps, ms[1], ach, out[prec]["tflops"], out[prec].get("max_err_cm1_vs_noquad_fp64")),
              file=sys.stderr, flush=True)
        c.close()
    out["fp32_vs_fp64"] = out["fp32"]["steps_per_s"] / out["fp64"]["steps_per_s"]
    out["bf16x3_vs_fp64"] = out["bf16x3"]["steps_per_s"] / out["fp64"]["steps_per_s"]
    return out


def adjoint_leg(B, args, x_max):
    """Hellmann-Feynman gradients at a 41^3 grid (lmax 5, 40 x 70, same cage):
    lowest 10 levels from an FP32-Hv run (Loop 1, Cullum-Willoughby, Loop 2
    Ritz vectors recomputed, FP64 Rayleigh refinement), gradients summed over
    degenerate multiplets (from the Ritz vectors left on the device, one
    bsl_hf_gradient_ritz call), checked against central differences (h = 1e-4
    relative) of the refined levels and against the gradient of an FP64-Hv
    run."""
    from bsl_amd import cage, physics
    pos, q, lj = cage.synthetic_cage(0)
    lj_a, lj_b = cage.baseline_lj()
    base = dict(nx=args.adjoint_n, x_max=x_max, lmax=5, n_theta=40, n_phi=70, mass=2.016,
                rot_const=physics.cm1_to_kjmol(59.322), ceiling=18.0, potential="cage", sites=(pos, q, lj),
                lj_a=lj_a, lj_b=lj_b, q_h2_h=physics.Q_H2_H, q_h2_cm=physics.Q_H2_CM)
    niter, count = args.adjoint_steps, 10

    def run(kw, grad):
        with B.Calculator(**kw) as c:
            t0 = time.perf_counter()
            e, _, res = c.levels(niter, count, vectors=False)
            t_lev = time.perf_counter() - t0
            if not grad:
                return e
            groups = B.multiplets(e, tol=1e-6)
            # the context's one-time gradient set-up (potential grid pass,
            # clamp mask, kept node-pair list), then the gradients themselves
            t0 = time.perf_counter()
            unclamped = c.unclamped_fraction()
            t_setup = time.perf_counter() - t0
            t0 = time.perf_counter()
            gr = c.hf_gradient_ritz([(gi, np.ones(len(gi))) for gi in groups])
            t_grad = time.perf_counter() - t0
            return e, res, groups, gr, t_lev, (t_setup, t_grad), unclamped

    e32, res, groups, g32, t_lev, t_grad, unclamped = run(dict(base, precision="fp32"), True)
    _, _, groups64, g64, _, _, _ = run(dict(base, precision="fp64"), True)

    def entry(g, name, idx):
        if name == "site_q":
            return float(g["site_q"][idx])
        if name == "site_pos":
            return float(g["site_pos"].reshape(-1)[idx])
        return float(g[name])

    names = [("q_h2_h", None), ("q_h2_cm", None), ("lj_a", None), ("lj_b", None), ("site_q", 0), ("site_pos", 0)]
    fd_err, fp64_diff = 0.0, 0.0
    for name, idx in names:
        x = base[name] if idx is None else (q[idx] if name == "site_q" else pos.reshape(-1)[idx])
        h = 1e-4 * abs(x)
        vals = []
        for sgn in (1.0, -1.0):
            kw = dict(base, precision="fp32")
            if name == "site_q":
                qq = q.copy()
                qq[idx] += sgn * h
                kw["sites"] = (pos, qq, lj)
            elif name == "site_pos":
                pp = pos.copy()
                pp.reshape(-1)[idx] += sgn * h
                kw["sites"] = (pp, q, lj)
            else:
                kw[name] = x + sgn * h
            vals.append(run(kw, False))
        for gi, g, g_64 in zip(groups, g32, g64):
            fd = (np.sum(vals[0][gi]) - np.sum(vals[1][gi])) / (2 * h)
            an = entry(g, name, idx)
            fd_err = max(fd_err, abs(an - fd) / max(abs(fd), 1e-3 / abs(x)))
            fp64_diff = max(fp64_diff, abs(an - entry(g_64, name, idx)) / max(abs(entry(g_64, name, idx)), 1e-12))
    t_setup, t_grad = t_grad
    out = {"grid": args.adjoint_n, "x_max_nm": x_max, "levels": count, "multiplets": len(groups), "steps": niter,
           "levels_s": t_lev, "gradient_setup_ms": 1e3 * t_setup, "gradient_ms": 1e3 * t_grad,
           "unclamped_fraction": unclamped,
           "max_residual": float(np.max(res)),
           "fd_parameters": [n if i is None else "%s[%d]" % (n, i) for n, i in names],
           "max_rel_err_vs_fd": fd_err, "max_rel_diff_vs_fp64_gradient": fp64_diff,
           "same_multiplets_as_fp64": [list(g) for g in groups] == [list(g) for g in groups64]}
    print("adjoint %d^3, x_max %.2f nm: levels %.2f s, gradient set-up %.1f ms, gradients %.1f ms (%.1f%% of the "
          "pairs unclamped), max rel err vs FD %.2e, vs FP64 gradient %.2e"
          % (args.adjoint_n, x_max, t_lev, 1e3 * t_setup, 1e3 * t_grad, 100 * unclamped, fd_err, fp64_diff),
          file=sys.stderr, flush=True)
    return out


def batch_members(args, calculator, precision, n=32):
    """16 synthetic cages (seeds 0-15) x {para (even l), ortho (odd l)} on
    grids cycled through 41^3, 47^3, 54^3 (same box, lmax 5, 40 x 70)."""
    from bsl_amd import cage, physics
    out = []
    grids = (41, 47, 54)
    for i in range(n):
        seed, block = i // 2, ("Even", "Odd")[i % 2]
        pos, q, lj = cage.synthetic_cage(seed)
        lj_a, lj_b = cage.baseline_lj()
        out.append(dict(nx=grids[i % 3], x_max=args.x_max, lmax=5, n_theta=40, n_phi=70, mass=2.016,
                        rot_const=physics.cm1_to_kjmol(59.322), ceiling=18.0, symmetry=block, basis="reduced",
                        calculator=calculator, precision=precision, potential="cage", sites=(pos, q, lj),
                        lj_a=lj_a, lj_b=lj_b))
    return out


def batch_leg(B, args):
    """32 independent recurrences in lockstep (hipGraph-replayed pairs of
    iterations), FP32 potential, FP64 recurrence: aggregate member steps/s."""
    out = {}
    for calculator in ("quad", "noquad"):
        t0 = time.perf_counter()
        bt = B.Batch(batch_members(args, calculator, "fp32"))
        prep = time.perf_counter() - t0
        ms = bt.time_steps(4, args.batch_steps, graph=True)
        # algorithmic bytes of one lockstep step: every member's potential
        # data and vector passes (bsl_step_bytes)
        nbytes = sum(float(np.sum(bt.member(i).step_bytes()[:2])) for i in range(32))
        bt.close()
        sec = ms / args.batch_steps / 1e3
        out[calculator] = {"members": 32, "member_steps_per_s": 32 / sec, "ms_per_lockstep_step": 1e3 * sec,
                           "prep_s": prep, "steps_timed": args.batch_steps,
                           "algorithmic_bytes_per_lockstep_step": nbytes, "achieved_GBs": nbytes / sec / 1e9}
        print("batch %s fp32: prep %.1f s, %.3f ms per lockstep step, %.0f member-steps/s, %.0f GB/s"
              % (calculator, prep, 1e3 * sec, out[calculator]["member_steps_per_s"], nbytes / sec / 1e9),
              file=sys.stderr, flush=True)
    return out


COPY_PROBE = r"""
import json, torch
a = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
b = torch.empty_like(a)
for _ in range(3):
    torch.mul(a, 1.0, out=b)
e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
torch.cuda.synchronize()
e0.record()
for _ in range(20):
    torch.mul(a, 1.0, out=b)
e1.record()
torch.cuda.synchronize()
print(json.dumps({"GBs": 20 * 2 * a.numel() * 8 / (e0.elapsed_time(e1) * 1e-3) / 1e9}))
"""


def copy_bandwidth():
    """Achievable HBM bandwidth of a plain copy kernel (1 GiB read + 1 GiB
    written, torch elementwise), in a child process so that torch's HIP
    runtime never shares a process with the library's."""
    import subprocess
    try:
        r = subprocess.run([sys.executable, "-c", COPY_PROBE], capture_output=True, text=True, timeout=300)
        return json.loads(r.stdout.strip().splitlines()[-1])["GBs"]
    except Exception:
        return None


def traffic_from_profiles(tag, kernel):
    """HBM bytes per launch of `kernel` measured with two rocprofv3 --pmc
    runs (FETCH_SIZE, then WRITE_SIZE, each scaled by what the counter reads
    for a plain copy kernel of known size), committed under profiles/."""
    path = os.path.join(REPO, "profiles", "traffic_%s.json" % tag)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        k = json.load(f)["kernels"].get("void bsl::" + kernel)
    return None if k is None else k["bytes_per_launch"]


def cpu_baseline(B, calc, args):
    """CPU time of the same operator on this host.  Where the reference's
    own calculator was built (oracle/_ref, compiled from the reference
    sources), that is timed: HvInterfaceSetup / HvPrep / Hv of
    linRotCartSph_NoQuad_ContGrid_Alavi_SPCE on the same grid, basis, mass
    and rotor constant, prepared on a 1 x 1 quadrature and then given the GPU
    context's V blocks (so the CPU does not redo the 40 x 70 preparation),
    one Hv per step.  The oracle (the same algorithm restated: serial kinetic
    term, OpenMP potential term) is timed on full Lanczos steps beside it,
    and the two Hv results are compared."""
    sys.path.insert(0, os.path.join(REPO, "tests"))
    import tempfile

    import _oracle as O
    from bsl_amd import cage, physics
    pos, q, lj = cage.synthetic_cage(0)
    typ = bytes([ord("O") if f else ord("H") for f in lj])
    lj_a, lj_b = cage.baseline_lj()
    mass, rot = 2.016, physics.cm1_to_kjmol(59.322)
    spec = dict(nx=args.n, ny=args.n, nz=args.n, x_max=args.x_max, y_max=args.x_max, z_max=args.x_max, lmax=5,
                na=40, nb=70, mass=mass, rot_const=rot, ceiling=18.0, sym=0,
                pot_kind=1, site_pos=pos, site_q=q, site_type=typ, q_h2_h=physics.Q_H2_H,
                q_h2_cm=physics.Q_H2_CM, h2_half_bond=physics.H2_HALF_BOND, lj_a=lj_a, lj_b=lj_b)
    vb = np.ascontiguousarray(calc.vblocks().ravel())
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    o = O.Oracle(spec, vblocks=vb.reshape(args.n, args.n, args.n, 36, 36))
    steps = args.cpu_steps
    o.lanczos(1)  # warm-up (page faults, thread pool)
    t0 = time.perf_counter()
    o.lanczos(steps)
    dt = time.perf_counter() - t0
    port = {"value": steps / dt, "unit": "steps/s", "cores": cores, "kind": "port",
            "sample": "%d Lanczos steps of the same %d^3 x 36 operator (oracle: serial kinetic term, OpenMP "
                      "potential term, as the reference), %.1f s" % (steps, args.n, dt)}
    if not O.ref_available() or args.ref_hv_calls <= 0:
        return port
    R = O.ref_lib()
    with tempfile.TemporaryDirectory() as tmp, _stdout_to_stderr():
        deck = os.path.join(tmp, "hv.txt")
        geo = os.path.join(REPO, "tests", "golden", "sII_smallCage11_CM_EA_ZYZ.xyz")  # prep only: V is set below
        O.write_ref_hv_deck(deck, args.n, args.x_max, 5, 1, 1, mass, rot, geo, 18.0, "All")
        n = R.ref_setup(b"linRotCartSph_NoQuad_ContGrid_Alavi_SPCE", deck.encode(), None)
    if n != calc.N or R.ref_set_vblocks(vb, vb.size) != 0:
        return port
    del vb
    v = np.random.default_rng(1).uniform(-1, 1, calc.N)
    u = np.zeros(calc.N)
    R.ref_hv(v, u)  # warm-up, and the parity check against the oracle
    ref_diff = float(np.max(np.abs(u - o.hv(v))) / np.max(np.abs(u)))
    k = args.ref_hv_calls
    t0 = time.perf_counter()
    for _ in range(k):
        u[:] = 0.0
        R.ref_hv(v, u)
    dt = time.perf_counter() - t0
    return {"value": k / dt, "unit": "steps/s", "cores": cores, "kind": "reference",
            "sample": "%d Hv calls of the reference's own compiled calculator (oracle/_ref: "
                      "linRotCartSph_NoQuad_ContGrid_Alavi_SPCE, serial kinetic term with its index reshuffles, "
                      "OpenMP potential term) on the same %d^3 x 36 operator, its V blocks set from the GPU "
                      "context, %.1f s; one Hv per step (the recurrence's seven serial N-length passes not "
                      "included)" % (k, args.n, dt),
            "ref_vs_oracle_hv_max_rel": ref_diff, "port": port}


def level_errors(e_ref, e, count=20):
    """Largest distance (cm^-1) from each of the `count` lowest reference
    levels to the nearest level of `e`."""
    from bsl_amd import physics
    low = e_ref[:count]
    return float(max(np.min(np.abs(e - x)) for x in low) / physics.KJMOL_PER_CM1)


def eig_error(B, c64, c32, steps):
    """Lowest 20 good levels of FP64 and mixed runs (same random start) after
    each step count in `steps` (the shorter runs are prefixes of the longest
    one's coefficients), and the Cullum-Willoughby eigensolve timed on the
    host and on the GPU at the first count.  Returns (per-count errors, FP64
    levels of the first count, eigensolve timings)."""
    nmax = max(steps)
    co64 = c64.lanczos(nmax, start="random", seed=0)
    co32 = c32.lanczos(nmax, start="random", seed=0)
    out, low, tri = {}, None, None
    for n in steps:
        p64 = tuple(x[: n + (1 if k else 0)] for k, x in enumerate(co64))
        p32 = tuple(x[: n + (1 if k else 0)] for k, x in enumerate(co32))
        t0 = time.perf_counter()
        e64, _ = B.lancbis(*p64, -10.0, 10.0)
        t_host = time.perf_counter() - t0
        e32, _ = B.lancbis(*p32, -10.0, 10.0, mult_tol=c32.mult_tol)
        out[str(n)] = level_errors(e64, e32)
        if low is None:
            # the mixed run's lowest 20 levels refined on the device: Loop 2
            # rebuilds their Ritz vectors, then the FP64 Rayleigh quotient
            sel = np.sort(e32)[:20]
            try:
                evtr, _ = B.trivec(*p32, sel, mult_tol=c32.mult_tol)
                t0 = time.perf_counter()
                _, e_ref, r_ref = c32.ritz_refine(n, evtr, len(sel), start="random", seed=0, vectors=False)
                out["refined_%d" % n] = level_errors(e64, e_ref)
                out["refined_%d_max_residual_kjmol" % n] = float(np.max(r_ref))
                out["refined_%d_s" % n] = time.perf_counter() - t0
            except B.BslError as e:
                out["refined_%d" % n] = None
                out["refined_%d_error" % n] = str(e)
            t0 = time.perf_counter()
            e64d, _ = B.lancbis(*p64, -10.0, 10.0, device=True)
            t_dev = time.perf_counter() - t0
            low = e64
            tri = {"steps": n, "host_s": t_host, "device_s": t_dev, "good_levels": int(len(e64)),
                   "mixed_mult_tol_kjmol": c32.mult_tol, "good_levels_mixed": int(len(e32)),
                   "max_abs_diff_kjmol": float(np.max(np.abs(e64 - e64d))) if len(e64) == len(e64d) else None}
    return out, low, tri


def eig_box(B, args, n, x_max, steps):
    """The mixed-vs-FP64 level error of another grid or box (same cage)."""
    c64 = make_calc(B, args, "fp64", n=n, x_max=x_max)
    c32 = make_calc(B, args, "fp32", n=n, x_max=x_max)
    err, low, _ = eig_error(B, c64, c32, steps)
    c64.close()
    c32.close()
    return {"grid": n, "x_max_nm": x_max, "max_err_cm1": err, "lowest_kjmol": [float(x) for x in low[:5]]}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=54)
    ap.add_argument("--x-max", type=float, default=0.4,
                    help="full box length (nm) of every leg; 0.4 is the reference's production decks' box")
    ap.add_argument("--eig-steps", type=str, default="10000,25000",
                    help="step counts of the mixed-vs-FP64 level comparison (comma-separated)")
    ap.add_argument("--eig-boxes", type=str, default="54:0.8,28:0.4",
                    help="further grid:x_max pairs for the 10000-step level comparison ('' skips them)")
    ap.add_argument("--quad-eig-steps", type=int, default=10000,
                    help="steps of the quadrature calculator's level comparison (0 skips it)")
    ap.add_argument("--adjoint-boxes", type=str, default="0.4,0.8",
                    help="box lengths (nm) of the adjoint leg (comma-separated)")
    ap.add_argument("--cpu-steps", type=int, default=40)
    ap.add_argument("--ref-hv-calls", type=int, default=32,
                    help="Hv calls of the reference's compiled calculator timed for cpu_baseline (0: the oracle only)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-eig", action="store_true")
    ap.add_argument("--no-mixed", action="store_true", help="skip the FP32-potential context")
    ap.add_argument("--batch-steps", type=int, default=100,
                    help="lockstep iterations timed for the 32-member batch (0 skips the batch leg)")
    ap.add_argument("--quad-steps", type=int, default=200,
                    help="steps timed per quadrature-calculator precision (0 skips the quadrature leg)")
    ap.add_argument("--adjoint-n", type=int, default=41, help="grid of the adjoint leg (0 skips it)")
    ap.add_argument("--adjoint-steps", type=int, default=4000)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain copy kernel over 1 GiB of doubles three times (the reference point "
                         "for the HBM byte counters of rocprofv3 --pmc) and exit")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: every box in this setup has one GPU, and one Lanczos "
              "recurrence per GPU is the unit of work (independent recurrences shard with no "
              "communication). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        import torch
        a = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            torch.mul(a, 1.0, out=b)
        torch.cuda.synchronize()
        print(json.dumps({"copy_kernel_bytes_read": a.numel() * 8, "copy_kernel_bytes_written": b.numel() * 8}))
        return
    copy_gbs = copy_bandwidth() if not args.no_cpu else None
    import bsl_amd as B

    t0 = time.perf_counter()
    c64 = make_calc(B, args, "fp64")
    prep64 = time.perf_counter() - t0
    ms = c64.time_steps(args.warmup, args.steps)
    sb = c64.step_bytes()
    steps_s = args.steps / (ms[0] / 1e3)
    pot_ms = ms[1]
    achieved = sb[0] / (pot_ms * 1e-3) / 1e9
    kname = c64.potential_kernel()
    print("fp64: prep %.2f s, %.1f steps/s, per step %.3f ms (potential %.3f, kinetic %.3f, finalize %.3f)"
          % (prep64, steps_s, ms[0] / args.steps, ms[1], ms[2], ms[3]), file=sys.stderr, flush=True)

    # the same operator with the z kinetic term fused into the potential
    # kernel (bsl_set_fusion; opt-in: fewer bytes per step, slower here)
    fused = None
    c64.fused = True
    msf = c64.time_steps(args.warmup, args.steps)
    sbf = c64.step_bytes()
    fused = {"steps_per_s": args.steps / (msf[0] / 1e3), "kernel": c64.potential_kernel(),
             "per_step_ms": {"fused_z_potential": msf[1], "xy_kinetic": msf[2], "finalize": msf[3]},
             "algorithmic_bytes_per_step": sbf[0] + sbf[1]}
    c64.fused = False
    print("fp64 fused z+potential: %.1f steps/s (fused kernel %.3f ms, x-y %.3f ms)"
          % (fused["steps_per_s"], msf[1], msf[2]), file=sys.stderr, flush=True)

    c32, steps_s32, prep32 = None, None, None
    if not args.no_mixed:
        t0 = time.perf_counter()
        c32 = make_calc(B, args, "fp32")
        prep32 = time.perf_counter() - t0
        ms32 = c32.time_steps(args.warmup, args.steps)
        steps_s32 = args.steps / (ms32[0] / 1e3)
        print("fp32: prep %.2f s, %.1f steps/s, per step %.3f ms (potential %.3f, kinetic %.3f, finalize %.3f)"
              % (prep32, steps_s32, ms32[0] / args.steps, ms32[1], ms32[2], ms32[3]), file=sys.stderr,
              flush=True)

    eig, e_ref = None, None
    steps_list = [int(x) for x in args.eig_steps.split(",") if x]
    if not args.no_eig and c32 is not None and steps_list:
        err, low, tri = eig_error(B, c64, c32, steps_list)
        e_ref = low
        eig = {"levels": 20, "max_err_cm1": err, "x_max_nm": args.x_max, "grid": args.n,
               "lowest_kjmol": [float(x) for x in low[:5]], "cullum_willoughby": tri, "other_boxes": []}
        print("mixed vs fp64 lowest 20 levels: max err %s cm^-1" % err, file=sys.stderr, flush=True)
        for spec in [x for x in args.eig_boxes.split(",") if x]:
            n, xm = spec.split(":")
            r = eig_box(B, args, int(n), float(xm), [steps_list[0]])
            eig["other_boxes"].append(r)
            print("mixed vs fp64 at %s^3, x_max %s nm: %s cm^-1" % (n, xm, r["max_err_cm1"]), file=sys.stderr,
                  flush=True)
    if c32 is not None:
        c32.close()

    # BF16x3 kinetic contractions (K = 32 BF16 MFMA split products), FP32
    # potential coefficients, FP64 recurrence
    bf = None
    if not args.no_mixed:
        t0 = time.perf_counter()
        c16 = make_calc(B, args, "bf16x3")
        prep16 = time.perf_counter() - t0
        ms16 = c16.time_steps(args.warmup, args.steps)
        bf = {"steps_per_s": args.steps / (ms16[0] / 1e3), "prep_s": prep16,
              "per_step_ms": {"potential": ms16[1], "kinetic": ms16[2], "finalize": ms16[3]}}
        bf["vs_fp64"] = bf["steps_per_s"] / steps_s
        if not args.no_eig and steps_list:
            err16, _, _ = eig_error(B, c64, c16, steps_list)
            bf["max_err_cm1"] = err16
        print("bf16x3 kinetic: %.1f steps/s (%.2fx fp64), level err %s cm^-1"
              % (bf["steps_per_s"], bf["vs_fp64"], bf.get("max_err_cm1")), file=sys.stderr, flush=True)
        c16.close()

    quad = quad_leg(B, args, e_ref) if args.quad_steps > 0 else None
    batch = batch_leg(B, args) if args.batch_steps > 0 else None
    adj = None
    if args.adjoint_n > 0:
        adj = []
        for xm in [float(x) for x in args.adjoint_boxes.split(",") if x]:
            try:
                adj.append(adjoint_leg(B, args, xm))
            except B.BslError as e:
                adj.append({"x_max_nm": xm, "error": str(e)})
                print("adjoint leg: %s" % e, file=sys.stderr, flush=True)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(B, c64, args)
        print("cpu baseline (%s): %.4f steps/s on %d threads%s" % (
            cpu["kind"], cpu["value"], cpu["cores"],
            "; oracle %.4f steps/s, reference Hv vs oracle Hv %.1e" % (cpu["port"]["value"],
                                                                      cpu["ref_vs_oracle_hv_max_rel"])
            if "port" in cpu else ""), file=sys.stderr, flush=True)

    line = {
        "metric": "Lanczos steps/s at 54³×36 basis, mixed vs FP64 GPU; eig err cm⁻¹ vs FP64",
        "value": steps_s,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms[0] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "Alavi H2 in a synthetic 5^12 SPC/E cage (seed 0); %d^3 sinc-DVR grid, x_max %.3g nm; "
                        "lmax 5 (36 lm, full basis); 40x70 quadrature; NoQuad potential (%s form); FP64 Hv and "
                        "recurrence" % (args.n, args.x_max, c64.potential_form),
            "basis_size": int(c64.N),
            "mixed_fp32_steps_per_s": steps_s32,
            "mixed_vs_fp64": steps_s32 / steps_s if steps_s32 else None,
            "eig_err_vs_fp64": eig,
            "mixed_bf16x3": bf,
            "prep_s": {"fp64": prep64, "fp32": prep32},
            "per_step_ms": {"potential": ms[1], "kinetic": ms[2], "finalize": ms[3]},
            "fused_z_potential_step": fused,
            "hv_per_s": 1e3 / (ms[1] + ms[2]),
            "quadrature": quad,
            "batch_32": batch,
            "adjoint": adj,
            "algorithmic_bytes_per_step": sb[0] + sb[1],
            "floor_bytes_per_step": sb[3],
            "step_bytes_roofline_frac": sb[3] / (ms[0] / args.steps * 1e-3) / 1e9 / HBM_PEAK_GBS,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": kname,
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic_from_profiles("cage54_lmax5_fp64", kname),
            "algorithmic_bytes_per_launch": sb[0],
            "achievable_copy_GBs": copy_gbs,
            "frac_of_copy": achieved / copy_gbs if copy_gbs else None,
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
