"""ctypes binding of the C ABI declared in include/dmftgpu.h.

The shared library is built in-tree (python -m dmftpack_amd.build) and loaded
from this directory; loading fails loudly when it is missing -- there is no
CPU fallback.
"""
import ctypes as C
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libdmftgpu.so")
HEADER = os.path.join(HERE, "..", "include", "dmftgpu.h")

OK, ERR_ARG, ERR_STATE, ERR_HIP, ERR_NUMERIC, ERR_UNSUPPORTED = 0, 1, 2, 3, 4, 5
PLAIN_F64, SCHUR_F64, SCHUR_F32R = 0, 1, 2
OPT_REFINE_STEPS, OPT_TARGET_WORKGROUPS, OPT_SCHUR_WG, OPT_SCHUR_LAYOUT, OPT_UPFOLD_CHUNK = 1, 2, 3, 4, 5
OPT_DENSITY_KERNEL = 6
MODE_FWD, MODE_DENS, MODE_VJP = 0, 1, 2


class Lattice(C.Structure):
    _fields_ = [("n", C.c_int32), ("n_corr", C.c_int32), ("n_cluster", C.c_int32), ("n_k", C.c_int32),
                ("n_k_global", C.c_int32), ("n_freq", C.c_int32), ("high_freq", C.c_int32), ("beta", C.c_double)]


class System(C.Structure):
    _fields_ = [("lat", Lattice), ("Hk", C.c_void_p), ("sigma", C.c_void_p), ("moments", C.c_void_p),
                ("mu", C.c_double), ("gloc", C.c_void_p), ("num_matrix", C.c_void_p), ("n_total", C.c_void_p),
                ("gloc_target", C.c_void_p), ("n_target", C.c_double), ("sigma_bar", C.c_void_p),
                ("mu_bar", C.c_void_p)]


BATCH_TUNE, BATCH_GRAPH, BATCH_PREPARE = 1, 2, 4


class DmftGpuError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"dmftgpu error {code}: {msg}")
        self.code = code


_lib = None
P = C.c_void_p
D = C.c_double
I = C.c_int
PD = C.POINTER(C.c_double)
PI = C.POINTER(C.c_int)

_SIGS = {
    "dmftgpu_create": ([C.POINTER(P), P], I),
    "dmftgpu_destroy": ([P], I),
    "dmftgpu_set_stream": ([P, P], I),
    "dmftgpu_last_error": ([P], C.c_char_p),
    "dmftgpu_version": ([], C.c_char_p),
    "dmftgpu_set_option": ([P, I, I], I),
    "dmftgpu_enable_timing": ([P, I], I),
    "dmftgpu_kernel_time": ([P, I, PD], I),
    "dmftgpu_prepare": ([P, C.POINTER(Lattice), P, P], I),
    "dmftgpu_static_eig": ([P, P, P], I),
    "dmftgpu_greenftn_w": ([P, I, P, D, P, P, P, PD], I),
    "dmftgpu_greenftn_w_noalloc": ([P, I, P, D, P, P, P, PD], I),
    "dmftgpu_density_matrix": ([P, P, D, P], I),
    "dmftgpu_density": ([P, I, P, D, PD, PD], I),
    "dmftgpu_set_core_bands": ([P, I, P, PI], I),
    "dmftgpu_adjust_mu": ([P, I, P, D, D, D, PD, PI], I),
    "dmftgpu_gloc_vjp": ([P, I, P, D, P, D, P, PD], I),
    "dmftgpu_gloc_jvp": ([P, I, P, D, P, D, P, PD], I),
    "dmftgpu_gloc_loss_grad": ([P, I, P, D, P, D, P, P, PD, PD, P, PD], I),
    "dmftgpu_moments": ([P, I, I, D, P, I, P, PI], I),
    "dmftgpu_batch_create": ([P, I, P, I, I, C.POINTER(P)], I),
    "dmftgpu_batch_run": ([P], I),
    "dmftgpu_batch_tuning": ([P, I, PI, PI, PI, PD, PD], I),
    "dmftgpu_batch_destroy": ([P], I),
    "dmftgpu_ft_w2tau": ([P, I, I, I, D, P, P, P, PI], I),
    "dmftgpu_ft_tau2w": ([P, I, I, I, D, P, P, PI], I),
    "dmftgpu_ft_w2tau_points": ([P, I, I, D, P, P, I, PD, P], I),
    "dmftgpu_update_moments": ([P, P], I),
    "dmftgpu_greenftn_real": ([P, I, I, D, D, D, P, D, P, P], I),
    "dmftgpu_greenftn_real_k": ([P, I, I, D, D, D, P, D, P, P], I),
    "dmftgpu_hyb_delta": ([P, I, I, D, D, P, P, P, I, PI, P], I),
    "dmftgpu_hk_from_hr": ([P, I, I, PI, PD, I, P, P], I),
    "dmftgpu_upfold_density": ([P, I, I, I, I, D, D, P, P, P, PI, P, P, P, I, PI, P, P, P], I),
}


def header_symbols():
    """Function names declared in include/dmftgpu.h."""
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(dmftgpu_\w+)\s*\(", txt, re.M)))


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OSError(f"{LIB_PATH} is not built; run `python -m dmftpack_amd.build`")
        L = C.CDLL(LIB_PATH)
        for name, (args, res) in _SIGS.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib


def check(handle, rc):
    if rc != OK:
        msg = lib().dmftgpu_last_error(handle)
        raise DmftGpuError(rc, msg.decode() if msg else "")
    return rc
