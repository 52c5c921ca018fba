"""ctypes binding of libplanc_hip.so (the C ABI declared in include/planc_hip.h).

The shared library is built in-tree (planc_amd/libplanc_hip.so) by
`planc_amd.build.build()`.  There is no fallback: importing the binding
without the library raises, so nothing can silently run on the CPU.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libplanc_hip.so")

c_int, c_long, c_double, c_float, c_void_p, c_size_t, c_char_p = (
    ctypes.c_int, ctypes.c_long, ctypes.c_double, ctypes.c_float, ctypes.c_void_p,
    ctypes.c_size_t, ctypes.c_char_p)

PLANC_OK = 0
PLANC_EINVAL = -1
PLANC_EHIP = -2
PLANC_ENOCONV = -3
PLANC_ENOMEM = -4
PLANC_EUNSUPPORTED = -5

F32, BF16, F16 = 0, 1, 2
MU, HALS, ANLSBPP, AOADMM, NESTEROV, GNSYM, R2 = 0, 1, 2, 4, 5, 7, 8


class NMFDesc(ctypes.Structure):
    _fields_ = [("algo", c_int), ("dtype", c_int), ("m", c_long), ("n", c_long), ("k", c_int),
                ("iters", c_int), ("regW", c_double * 2), ("regH", c_double * 2),
                ("compute_error", c_int), ("A16T", c_void_p), ("normA2", c_double),
                ("symm_reg", c_double)]


class HierDesc(ctypes.Structure):
    _fields_ = [("m", c_long), ("n", c_long), ("depth", c_int), ("iters", c_int),
                ("initseed", c_int), ("nodes", c_int)]


class HierNode(ctypes.Structure):
    _fields_ = [("index", c_long), ("level", c_int), ("ncols", c_long), ("col_offset", c_long),
                ("sigma", c_double), ("score", c_double), ("lvalid", c_int), ("rvalid", c_int),
                ("split", c_int), ("parent_side", c_int)]


class NTFDesc(ctypes.Structure):
    _fields_ = [("algo", c_int), ("dtype", c_int), ("nmodes", c_int), ("dims", c_long * 8),
                ("k", c_int), ("iters", c_int), ("dimtree", c_int), ("compute_error", c_int),
                ("X16T", c_void_p), ("normX2", c_double)]


# name -> (restype, argtypes); the exported C ABI.
SIGNATURES = {
    "planc_version": (c_char_p, []),
    "planc_read_bandwidth": (c_int, [c_void_p, c_void_p, c_size_t, c_int, ctypes.POINTER(c_double)]),
    "planc_create": (c_int, [c_int, ctypes.POINTER(c_void_p)]),
    "planc_destroy": (c_int, [c_void_p]),
    "planc_sync": (c_int, [c_void_p]),
    "planc_last_error": (c_char_p, [c_void_p]),
    "planc_set_stream": (c_int, [c_void_p, c_void_p]),
    "planc_get_stream": (c_void_p, [c_void_p]),
    "planc_malloc": (c_int, [c_void_p, ctypes.POINTER(c_void_p), c_size_t]),
    "planc_free": (c_int, [c_void_p, c_void_p]),
    "planc_memcpy_h2d": (c_int, [c_void_p, c_void_p, c_void_p, c_size_t]),
    "planc_memcpy_d2h": (c_int, [c_void_p, c_void_p, c_void_p, c_size_t]),
    "planc_convert": (c_int, [c_void_p, c_int, c_void_p, c_long, c_long, c_long, c_void_p, c_long,
                              c_void_p]),
    "planc_transpose16": (c_int, [c_void_p, c_void_p, c_long, c_long, c_long, c_void_p, c_long]),
    "planc_gen_uniform": (c_int, [c_void_p, c_void_p, c_long, ctypes.c_ulonglong,
                                  ctypes.c_ulonglong]),
    "planc_gen_lowrank": (c_int, [c_void_p, c_long, c_long, c_int, c_void_p, c_void_p, c_double,
                                  ctypes.c_ulonglong, ctypes.c_ulonglong, c_void_p, c_long]),
    "planc_gen_cp_tensor": (c_int, [c_void_p, c_int, ctypes.POINTER(c_long),
                                    ctypes.POINTER(c_void_p), c_int, c_double, ctypes.c_ulonglong,
                                    ctypes.c_ulonglong, c_void_p]),
    "planc_gemm_nn": (c_int, [c_void_p, c_int, c_void_p, c_long, c_long, c_long, c_void_p, c_void_p,
                              c_long, c_int, c_void_p, c_long]),
    "planc_gemm_tn": (c_int, [c_void_p, c_int, c_void_p, c_long, c_long, c_long, c_void_p, c_void_p,
                              c_long, c_int, c_void_p, c_long]),
    "planc_gram": (c_int, [c_void_p, c_void_p, c_long, c_int, c_long, c_void_p]),
    "planc_mu_update": (c_int, [c_void_p, c_void_p, c_long, c_int, c_long, c_void_p, c_long, c_long,
                                c_void_p, c_double]),
    "planc_hals_update_rows": (c_int, [c_void_p, c_void_p, c_long, c_int, c_long, c_void_p, c_long,
                                       c_long, c_void_p]),
    "planc_hals_update_cols": (c_int, [c_void_p, c_void_p, c_long, c_int, c_long, c_void_p, c_long,
                                       c_void_p]),
    "planc_normalize_by_w": (c_int, [c_void_p, c_void_p, c_long, c_long, c_void_p, c_long, c_long,
                                     c_int]),
    "planc_normalize_cols": (c_int, [c_void_p, c_void_p, c_long, c_int, c_long, c_void_p]),
    "planc_hadamard_grams": (c_int, [c_void_p, ctypes.POINTER(c_void_p), c_int, c_int, c_int,
                                     c_void_p]),
    "planc_mttkrp": (c_int, [c_void_p, c_int, c_void_p, c_int, ctypes.POINTER(c_long),
                             ctypes.POINTER(c_void_p), c_int, c_int, c_void_p, c_long]),
    "planc_mttkrp_partial": (c_int, [c_void_p, c_int, c_void_p, c_int, ctypes.POINTER(c_long),
                                     ctypes.POINTER(c_void_p), c_int, c_int, c_int, c_void_p,
                                     c_long]),
    "planc_multi_ttv": (c_int, [c_void_p, c_void_p, c_long, c_int, ctypes.POINTER(c_long),
                                ctypes.POINTER(c_void_p), c_int, c_int, c_void_p, c_long]),
    "planc_cp_loss_fwd": (c_int, [c_void_p, c_void_p, c_int, ctypes.POINTER(c_long),
                                  ctypes.POINTER(c_void_p), c_int, ctypes.POINTER(c_double)]),
    "planc_cp_loss_fwd_noalloc": (c_int, [c_void_p, c_void_p, c_int, ctypes.POINTER(c_long),
                                          ctypes.POINTER(c_void_p), c_int, c_void_p]),
    "planc_cp_loss_bwd": (c_int, [c_void_p, c_void_p, c_int, ctypes.POINTER(c_long),
                                  ctypes.POINTER(c_void_p), c_int, ctypes.POINTER(c_void_p)]),
    "planc_admm_update": (c_int, [c_void_p, c_void_p, c_int, c_void_p, c_long, c_long, c_void_p, c_long,
                                  c_void_p, c_long, c_long]),
    "planc_bpp": (c_int, [c_void_p, c_void_p, c_int, c_void_p, c_long, c_long, c_long, c_void_p,
                          c_long, c_long]),
    "planc_r2_update": (c_int, [c_void_p, c_void_p, c_void_p, c_long, c_long, c_long, c_void_p, c_long,
                                c_long]),
    "planc_rel_error_gram": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_void_p]),
    "planc_csr_col_subset": (c_int, [c_void_p, c_long, c_long, c_void_p, c_void_p, c_void_p,
                                     ctypes.POINTER(c_long), c_long, c_void_p, c_void_p, c_void_p,
                                     ctypes.POINTER(c_long)]),
    "planc_hiernmf_run": (c_int, [c_void_p, ctypes.POINTER(HierDesc), ctypes.POINTER(c_long),
                                  ctypes.POINTER(c_int), ctypes.POINTER(c_float),
                                  ctypes.POINTER(HierNode), ctypes.POINTER(c_int),
                                  ctypes.POINTER(c_long), ctypes.POINTER(c_double),
                                  ctypes.POINTER(c_double)]),
    "planc_csr_transpose": (c_int, [c_void_p, c_long, c_long, c_void_p, c_void_p, c_void_p, c_void_p,
                                    c_void_p, c_void_p]),
    "planc_spmm": (c_int, [c_void_p, c_long, c_long, c_void_p, c_void_p, c_void_p, c_void_p, c_long,
                           c_int, c_void_p, c_long]),
    "planc_sparse_fit_error": (c_int, [c_void_p, c_long, c_long, c_void_p, c_void_p, c_void_p, c_void_p,
                                       c_long, c_void_p, c_long, c_int, ctypes.POINTER(c_double)]),
    "planc_nmf_sparse_run": (c_int, [c_void_p, ctypes.POINTER(NMFDesc), c_void_p, c_void_p, c_void_p,
                                     c_void_p, c_void_p, ctypes.POINTER(c_double)]),
    "planc_nmf_run": (c_int, [c_void_p, ctypes.POINTER(NMFDesc), c_void_p, c_long, c_void_p,
                              c_void_p, c_void_p, c_void_p, ctypes.POINTER(c_double)]),
    "planc_ntf_tree_split": (c_int, [c_int, ctypes.POINTER(c_long)]),
    "planc_ntf_run": (c_int, [c_void_p, ctypes.POINTER(NTFDesc), c_void_p, c_void_p,
                              ctypes.POINTER(c_void_p), c_void_p, ctypes.POINTER(c_double)]),
}

_lib = None


def load():
    """Loads libplanc_hip.so and declares every C ABI signature."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with planc_amd.build.build() "
            "(hipcc --offload-arch=gfx950); there is no CPU fallback")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib
