"""ctypes binding of ``libserenity_xc.so`` (the C ABI in include/serenity_xc.h).

The library is built in-tree by ``serenity_amd.build``; importing this module
without it raises instead of falling back to anything else.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libserenity_xc.so")

XC_OK = 0
XC_BUSY = 6
FUNCTIONALS = {"LDA": 0, "PBE": 1, "TF": 2, "SLATER": 3, "VWN5": 4, "PBEX": 5, "PBEC": 6, "PBE_LIBXC": 7, "BP86": 8}
GGA = {"PBE", "PBEX", "PBEC", "PBE_LIBXC", "BP86"}
PRECISIONS = {"FP64": 0, "FP32": 1, "MIXED": 2}
STAGES = ["convert", "density", "functional", "fock", "reduce", "finalize"]
GRID_TYPES = {"BECKE": 0, "VORONOI": 1, "SSF": 2}
RADIAL_TYPES = {"BECKE": 0, "HANDY": 1, "AHLRICHS": 2, "KNOWLES": 3, "EQUI": 4}

_lib = None

dp = C.POINTER(C.c_double)
ip = C.POINTER(C.c_int)
vp = C.c_void_p
i64 = C.c_int64


class XCError(RuntimeError):
    """Error raised by the XC library (Serenity's SerenityError counterpart)."""

    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code


def _sig(name, res, *args):
    f = getattr(_lib, name)
    f.restype = res
    f.argtypes = list(args)


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            "libserenity_xc.so is not built: run `python -m serenity_amd.build` "
            "(or __graft_entry__.build()) first")
    _lib = C.CDLL(LIB_PATH)
    _sig("xc_last_error", C.c_char_p)
    _sig("xc_version", C.c_int)
    _sig("xc_ctx_create", C.c_int, C.c_int, C.POINTER(vp))
    _sig("xc_ctx_destroy", C.c_int, vp)
    _sig("xc_ctx_synchronize", C.c_int, vp)
    _sig("xc_ctx_stream", vp, vp)
    _sig("xc_ctx_memory", C.c_int, vp, C.POINTER(i64))
    _sig("xc_ctx_set_deterministic", C.c_int, vp, C.c_int)
    _sig("xc_normalize_contraction", C.c_int, C.c_int, C.c_int, dp, dp)
    _sig("xc_load_basis", C.c_int, C.c_char_p, C.c_int, ip, dp, C.c_int, ip,
         C.POINTER(ip), C.POINTER(ip), C.POINTER(ip), C.POINTER(dp), C.POINTER(dp), C.POINTER(dp))
    _sig("xc_free_host", None, vp)
    _sig("xc_molecular_grid", C.c_int, C.c_int, ip, dp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double,
         C.c_int, C.POINTER(i64), C.POINTER(dp), C.POINTER(dp), C.POINTER(i64))
    _sig("xc_molecular_grid_device", C.c_int, vp, C.c_int, ip, dp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double,
         C.c_int, C.POINTER(i64), C.POINTER(dp), C.POINTER(dp), C.POINTER(i64))
    _sig("xc_lebedev", C.c_int, C.c_int, dp, dp, dp, dp)
    _sig("xc_overlap", C.c_int, C.c_int, ip, ip, ip, dp, dp, dp, dp)
    _sig("xc_boys", C.c_int, C.c_int, C.c_double, dp)
    _sig("xc_one_electron", C.c_int, C.c_int, ip, ip, ip, dp, dp, dp, C.c_int, ip, dp, dp, dp)
    _sig("xc_eri_packed", C.c_int, C.c_int, ip, ip, ip, dp, dp, dp, C.c_double, dp)
    _sig("xc_ri_j_integrals", C.c_int, C.c_int, ip, ip, ip, dp, dp, dp, C.c_int, ip, ip, ip, dp, dp, dp,
         C.c_double, dp, dp)
    _sig("xc_ri_metric", C.c_int, C.c_int, ip, ip, ip, dp, dp, dp, dp)
    _sig("xc_coulomb_create", C.c_int, vp, vp, C.c_double, C.POINTER(vp))
    _sig("xc_coulomb_info", C.c_int, vp, C.POINTER(i64), C.POINTER(i64))
    _sig("xc_coulomb_rhs", C.c_int, vp, dp, dp)
    _sig("xc_coulomb_matrix", C.c_int, vp, dp, dp)
    _sig("xc_coulomb_destroy", C.c_int, vp)
    _sig("xc_basis_create", C.c_int, vp, C.c_int, ip, ip, ip, dp, dp, dp, C.POINTER(vp))
    _sig("xc_basis_destroy", C.c_int, vp)
    _sig("xc_basis_nao", C.c_int, vp, ip)
    _sig("xc_grid_create", C.c_int, vp, i64, dp, dp, C.POINTER(vp))
    _sig("xc_grid_destroy", C.c_int, vp)
    _sig("xc_bfogc_create", C.c_int, vp, vp, C.c_int, C.c_double, C.c_int, C.c_int, C.POINTER(vp))
    _sig("xc_bfogc_destroy", C.c_int, vp)
    _sig("xc_bfogc_info", C.c_int, vp, ip, C.POINTER(i64), ip, ip)
    _sig("xc_bfogc_stats", C.c_int, vp, C.POINTER(i64), dp, C.POINTER(i64))
    _sig("xc_bfogc_block_data", C.c_int, vp, C.c_int, dp, dp, dp, dp, ip, ip)
    _sig("xc_bfogc_block_hessian", C.c_int, vp, C.c_int, dp)
    _sig("xc_geom_gradient", C.c_int, vp, C.c_int, dp, C.c_int, ip, dp, dp)
    _sig("xc_geom_gradient_unrestricted", C.c_int, vp, C.c_int, dp, dp, C.c_int, ip, dp, dp)
    _sig("xc_fock_build_unrestricted", C.c_int, vp, C.c_int, C.c_int, C.c_double, dp, dp, dp, dp, dp)
    _sig("xc_fock_build_unrestricted_device", C.c_int, vp, C.c_int, C.c_int, C.c_double, vp, vp, vp, vp, dp)
    _sig("xc_density", C.c_int, vp, dp, dp, dp, ip)
    _sig("xc_functional", C.c_int, vp, C.c_int, i64, C.c_int, dp, dp, dp, dp, dp, dp, dp)
    _sig("xc_functional_device_noalloc", C.c_int, vp, C.c_int, i64, C.c_int, vp, vp, vp, vp, vp, vp, dp)
    _sig("xc_fock_add", C.c_int, vp, C.c_double, dp, dp, dp)
    _sig("xc_fock_build", C.c_int, vp, C.c_int, C.c_int, C.c_double, dp, dp, dp)
    _sig("xc_fock_build_device", C.c_int, vp, C.c_int, C.c_int, C.c_double, vp, vp, dp)
    _sig("xc_fock_build_packed", C.c_int, C.c_int, C.POINTER(vp), C.c_int, C.c_int, C.c_double, C.POINTER(vp),
         C.POINTER(vp), dp)
    _sig("xc_last_energy", C.c_int, vp, dp)
    _sig("xc_nadd_fock_build", C.c_int, vp, C.c_int, C.POINTER(vp), C.c_int, C.c_int, C.c_double, dp,
         C.POINTER(dp), dp, dp)
    _sig("xc_nadd_fock_build_unrestricted", C.c_int, vp, C.c_int, C.POINTER(vp), C.c_int, C.c_int, C.c_double, dp,
         dp, C.POINTER(dp), C.POINTER(dp), dp, dp, dp)
    _sig("xc_server_create", C.c_int, vp, C.c_int, C.c_int, C.POINTER(vp))
    _sig("xc_server_destroy", C.c_int, vp)
    _sig("xc_server_submit", C.c_int, vp, vp, C.c_int, C.c_int, C.c_double, dp, dp, dp, C.POINTER(C.c_int64))
    _sig("xc_server_poll", C.c_int, vp, C.c_int64, C.POINTER(C.c_int))
    _sig("xc_server_wait", C.c_int, vp, C.c_int64)
    _sig("xc_server_stats", C.c_int, vp, C.POINTER(C.c_int), C.POINTER(C.c_int64), C.POINTER(C.c_int64))
    _sig("xc_bfogc_update_density", C.c_int, vp, C.c_int, vp, C.c_int)
    _sig("xc_nadd_fock_build_cached", C.c_int, vp, C.c_int, C.POINTER(vp), C.c_int, C.c_int, C.c_double, vp, dp)
    _sig("xc_nadd_fock_build_device", C.c_int, vp, C.c_int, C.POINTER(vp), C.c_int, C.c_int, C.c_double, vp,
         C.POINTER(vp), vp, dp)
    _sig("xc_profile", C.c_int, vp, C.c_int)
    _sig("xc_profile_collect", C.c_int, vp, dp, ip)
    _sig("xc_hilbert_sort", C.c_int, i64, dp, dp)
    _sig("xc_device_alloc", C.c_int, vp, C.c_size_t, C.POINTER(vp))
    _sig("xc_device_free", C.c_int, vp, vp)
    _sig("xc_memcpy_h2d", C.c_int, vp, vp, vp, C.c_size_t)
    _sig("xc_memcpy_d2h", C.c_int, vp, vp, vp, C.c_size_t)
    _sig("xc_device_copy", C.c_int, vp, vp, vp, C.c_size_t)
    return _lib


def check(rc):
    if rc != XC_OK:
        msg = lib().xc_last_error()
        raise XCError(rc, msg.decode() if msg else "error %d" % rc)


def exported_symbols_declared_in_header():
    """Names of every function declared in include/serenity_xc.h."""
    import re
    hdr = os.path.join(HERE, "..", "include", "serenity_xc.h")
    text = open(hdr).read()
    return sorted(set(re.findall(r"\b(xc_[a-z0-9_]+)\s*\(", text)))
