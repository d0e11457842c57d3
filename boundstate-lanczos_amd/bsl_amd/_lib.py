"""ctypes binding of libbsl_hip.so (the C ABI declared in include/bsl_hip.h).

The library is loaded from the package's own lib/ directory; when it is
missing the import fails loudly (there is no CPU fallback for the product
path)."""
import ctypes as C
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "lib", "libbsl_hip.so")

SYM = {"All": 0, "Even": 1, "Odd": 2}
BASIS = {"full": 0, "reduced": 1}
CALC = {"noquad": 0, "quad": 1}
PREC = {"fp64": 0, "fp32": 1, "bf16x3": 2}
POT = {"iho": 0, "cage": 1, "cage_grid": 2, "coulomb": 3}
START = {"uniform": 0, "random": 1}
VFORM = {"blocks": 0, "harmonics": 1}

EXPORTS = (
    "bsl_create", "bsl_destroy", "bsl_last_error", "bsl_basis_size", "bsl_krylov_size", "bsl_hv",
    "bsl_hv_device", "bsl_lanczos", "bsl_lancbis", "bsl_trivec", "bsl_ritz", "bsl_rayleigh",
    "bsl_water_sites", "bsl_get_vblocks", "bsl_get_vgrid", "bsl_time_steps", "bsl_step_bytes",
    "bsl_potential_kernel", "bsl_hf_gradient", "bsl_batch_create", "bsl_batch_destroy", "bsl_batch_size",
    "bsl_batch_member", "bsl_batch_lanczos", "bsl_batch_time_steps", "bsl_lancbis_device", "bsl_lancbis_tol",
    "bsl_lancbis_device_tol", "bsl_trivec_tol", "bsl_get_potential_form", "bsl_set_potential_form",
    "bsl_gaunt_table", "bsl_lanczos_noalloc", "bsl_clamp_stats",
    "bsl_ritz_refine", "bsl_get_fusion", "bsl_set_fusion", "bsl_hf_gradient_sets", "bsl_hf_gradient_ritz",
)


class Params(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
        ("x_max", C.c_double), ("y_max", C.c_double), ("z_max", C.c_double),
        ("lmax", C.c_int32), ("n_theta", C.c_int32), ("n_phi", C.c_int32),
        ("mass", C.c_double), ("rot_const", C.c_double), ("ceiling", C.c_double),
        ("symmetry", C.c_int32), ("basis", C.c_int32), ("calculator", C.c_int32),
        ("precision", C.c_int32), ("potential", C.c_int32),
        ("iho_factor", C.c_double), ("iho_ceil", C.c_double),
        ("n_sites", C.c_int32),
        ("site_pos", C.c_void_p), ("site_q", C.c_void_p), ("site_lj", C.c_void_p),
        ("q_h2_h", C.c_double), ("q_h2_cm", C.c_double), ("h2_half_bond", C.c_double),
        ("lj_a", C.c_double), ("lj_b", C.c_double),
        ("pnx", C.c_int32), ("pny", C.c_int32), ("pnz", C.c_int32),
        ("px_max", C.c_double), ("py_max", C.c_double), ("pz_max", C.c_double),
        ("coul_q_centre", C.c_double), ("coul_q_mol", C.c_double), ("coul_ceil", C.c_double),
    ]


_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


def load():
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            "libbsl_hip.so is not built (%s); run __graft_entry__.build() or `make -C boundstate-lanczos_amd`"
            % LIB_PATH)
    lib = C.CDLL(LIB_PATH)
    lib.bsl_create.argtypes = [C.POINTER(Params), C.POINTER(C.c_void_p)]
    lib.bsl_destroy.argtypes = [C.c_void_p]
    lib.bsl_last_error.restype = C.c_char_p
    lib.bsl_basis_size.restype = C.c_int64
    lib.bsl_basis_size.argtypes = [C.c_void_p]
    lib.bsl_krylov_size.restype = C.c_int64
    lib.bsl_krylov_size.argtypes = [C.c_void_p]
    lib.bsl_hv.argtypes = [C.c_void_p, _dp, _dp]
    lib.bsl_hv_device.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
    lib.bsl_lanczos.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.c_int32, C.c_uint64, _dp, _dp, _dp]
    lib.bsl_lanczos_noalloc.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.c_int32, C.c_uint64, _dp, _dp, _dp]
    lib.bsl_lancbis.argtypes = [C.c_int32, _dp, _dp, _dp, C.c_double, C.c_double, _dp, _dp,
                                C.POINTER(C.c_int32)]
    lib.bsl_lancbis_device.argtypes = [C.c_int32, _dp, _dp, _dp, C.c_double, C.c_double, _dp, _dp,
                                       C.POINTER(C.c_int32)]
    lib.bsl_lancbis_tol.argtypes = [C.c_int32, _dp, _dp, _dp, C.c_double, C.c_double, C.c_double, _dp, _dp,
                                    C.POINTER(C.c_int32)]
    lib.bsl_lancbis_device_tol.argtypes = [C.c_int32, _dp, _dp, _dp, C.c_double, C.c_double, C.c_double, _dp, _dp,
                                           C.POINTER(C.c_int32)]
    lib.bsl_trivec.argtypes = [C.c_int32, _dp, _dp, _dp, _dp, C.c_int32, _dp, _ip]
    lib.bsl_trivec_tol.argtypes = [C.c_int32, _dp, _dp, _dp, _dp, C.c_int32, C.c_double, _dp, _ip]
    lib.bsl_ritz.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.c_int32, C.c_uint64, _dp, C.c_int32, _dp]
    lib.bsl_rayleigh.argtypes = [C.c_void_p, C.c_int32, _dp, _dp, _dp]
    lib.bsl_ritz_refine.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.c_int32, C.c_uint64, _dp, C.c_int32,
                                    C.c_void_p, _dp, _dp]
    lib.bsl_water_sites.argtypes = [C.c_int32, C.c_int32, _dp, _dp, _dp, _ip]
    lib.bsl_get_vblocks.argtypes = [C.c_void_p, _dp]
    lib.bsl_get_vgrid.argtypes = [C.c_void_p, _dp, _dp]
    lib.bsl_time_steps.argtypes = [C.c_void_p, C.c_int32, C.c_int32, _dp]
    lib.bsl_step_bytes.argtypes = [C.c_void_p, _dp]
    lib.bsl_potential_kernel.argtypes = [C.c_void_p, C.c_char_p, C.c_int32]
    lib.bsl_get_potential_form.argtypes = [C.c_void_p]
    lib.bsl_set_potential_form.argtypes = [C.c_void_p, C.c_int32]
    lib.bsl_get_fusion.argtypes = [C.c_void_p]
    lib.bsl_set_fusion.argtypes = [C.c_void_p, C.c_int32]
    lib.bsl_gaunt_table.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.c_void_p, C.c_void_p]
    lib.bsl_hf_gradient.argtypes = [C.c_void_p, C.c_int32, _dp, _dp, _dp]
    lib.bsl_hf_gradient_sets.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, _dp, _dp, _dp]
    lib.bsl_hf_gradient_ritz.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, _dp, _dp]
    lib.bsl_clamp_stats.argtypes = [C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
    lib.bsl_batch_create.argtypes = [C.c_int32, C.c_void_p, C.c_void_p]
    lib.bsl_batch_destroy.argtypes = [C.c_void_p]
    lib.bsl_batch_size.argtypes = [C.c_void_p]
    lib.bsl_batch_member.argtypes = [C.c_void_p, C.c_int32, C.c_void_p]
    lib.bsl_batch_lanczos.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_uint64, C.c_int32, _dp, _dp, _dp]
    lib.bsl_batch_time_steps.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, _dp]
    return lib


class BslError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("bsl error %d: %s" % (code, msg))
        self.code = code


_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        _LIB = load()
    return _LIB


def check(rc):
    if rc != 0:
        raise BslError(rc, lib().bsl_last_error().decode(errors="replace"))
