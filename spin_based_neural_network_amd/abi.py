"""ctypes mirror of the C ABI exported by ``libnqs_hip.so``.

The structures below repeat, field for field, ``include/nqs/nqs_config.h``
(reference: include/nqs/nqs_config.h:82-120), ``nqs_sr_step_info_t``
(include/nqs/nqs_optimizer.h:35-42), ``nqs_marshall_wrapper_t``
(include/nqs/nqs_marshall.h:35-40) and ``nqs_hip_options_t``
(include/nqs_hip/nqs_hip.h).  ``API`` is the signature table of the
reference entry points; ``bind(lib, prefix)`` applies it to any library that
exports them under ``prefix`` (``nqs_`` for this library and for the
reference build, ``orc_`` for the CPU oracle used by the tests).

Loading this module loads nothing; ``load_library()`` opens the in-tree
``libnqs_hip.so`` and raises if it is missing -- there is no fallback.
"""
import ctypes
import os
from ctypes import (POINTER, Structure, c_double, c_int, c_long, c_longlong,
                    c_char_p, c_uint, c_ulonglong, c_void_p)

import numpy as np

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, "libnqs_hip.so")

# nqs_ansatz_kind_t / nqs_hamiltonian_kind_t (include/nqs/nqs_config.h:25-79)
ANSATZ_MF, ANSATZ_VIT, ANSATZ_FVIT, ANSATZ_AR, ANSATZ_KAN = 0, 1, 2, 3, 4
ANSATZ_RBM, ANSATZ_CRBM, ANSATZ_FVIT_C = 5, 6, 7
HAM_TFIM, HAM_HEIS, HAM_J1J2, HAM_KITAEV, HAM_XXZ, HAM_KH, HAM_KAGOME = range(7)
# nqs_hip_move_t / nqs_hip_precision_t / nqs_hip_gram_t
MOVE_FLIP, MOVE_EXCHANGE, MOVE_LEARNED = 0, 1, 2
FP64, FP32, MIXED = 0, 1, 2
GRAM_DIRECT, GRAM_FACTORED, GRAM_AUTO, GRAM_PARAMETER = 0, 1, 2, 3


class NqsConfig(Structure):
    _fields_ = [
        ("ansatz", c_int), ("patch_size", c_int), ("depth", c_int),
        ("width", c_int), ("heads", c_int), ("rbm_hidden_units", c_int),
        ("rbm_init_scale", c_double), ("symmetries", c_int),
        ("hamiltonian", c_int), ("j_coupling", c_double),
        ("j2_coupling", c_double), ("j_z_coupling", c_double),
        ("transverse_field", c_double), ("kh_K", c_double), ("kh_J", c_double),
        ("kagome_pbc", c_int), ("num_samples", c_int),
        ("num_thermalize", c_int), ("num_decorrelate", c_int),
        ("cluster_moves", c_int), ("rng_seed", c_uint),
        ("num_iterations", c_int), ("learning_rate", c_double),
        ("sr_diag_shift", c_double), ("sr_cg_max_iters", c_int),
        ("sr_cg_tol", c_double),
    ]


def config_defaults(**overrides):
    """nqs_config_defaults() (include/nqs/nqs_config.h:123-152) + overrides."""
    c = NqsConfig(
        ansatz=ANSATZ_MF, patch_size=1, depth=3, width=128, heads=4,
        rbm_hidden_units=0, rbm_init_scale=0.01, symmetries=(1 << 1) | 1,
        hamiltonian=HAM_TFIM, j_coupling=1.0, j2_coupling=0.0,
        j_z_coupling=1.0, transverse_field=1.0, kh_K=1.0, kh_J=0.0,
        kagome_pbc=1, num_samples=1024, num_thermalize=256,
        num_decorrelate=4, cluster_moves=0, rng_seed=0xC0FFEE,
        num_iterations=200, learning_rate=1e-2, sr_diag_shift=1e-3,
        sr_cg_max_iters=100, sr_cg_tol=1e-6)
    for k, v in overrides.items():
        if not hasattr(c, k):
            raise AttributeError(f"nqs_config_t has no field {k!r}")
        setattr(c, k, v)
    return c


class SrInfo(Structure):
    _fields_ = [("mean_energy", c_double), ("variance_energy", c_double),
                ("update_norm", c_double), ("acceptance_ratio", c_double),
                ("cg_iterations", c_int), ("converged", c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class LanczosResult(Structure):
    """lanczos_result_t (include/mps/lanczos.h)."""
    _fields_ = [("eigenvalue", c_double), ("iterations", c_int), ("converged", c_int),
                ("residual_norm", c_double)]


class SymprojWrapper(Structure):
    """nqs_symproj_wrapper_t (include/nqs/nqs_symproj.h)."""
    _fields_ = [("base_log_amp", c_void_p), ("base_user", c_void_p), ("num_sites", c_int),
                ("num_group_elements", c_int), ("perm", c_void_p), ("characters", c_void_p)]


class TranslationWrapper(Structure):
    """nqs_translation_wrapper_t (include/nqs/nqs_translation.h)."""
    _fields_ = [("base_log_amp", c_void_p), ("base_user", c_void_p), ("size_x", c_int),
                ("size_y", c_int)]


class MarshallWrapper(Structure):
    _fields_ = [("base_log_amp", c_void_p), ("base_user", c_void_p),
                ("size_x", c_int), ("size_y", c_int)]


class HipOptions(Structure):
    _fields_ = [("num_chains", c_int), ("move", c_int),
                ("sz_zero_start", c_int), ("periodic", c_int),
                ("precision", c_int), ("gram", c_int),
                ("ar_hidden", c_int), ("ar_train_every", c_int),
                ("ar_train_window", c_int), ("ar_epochs", c_int),
                ("ar_batch", c_int), ("ar_lr", c_double), ("ar_local", c_int),
                ("ar_seed", c_uint)]


def options_defaults(**overrides):
    """nqs_hip_options_defaults() + overrides (same values as the C side)."""
    o = HipOptions(num_chains=1, move=MOVE_FLIP, sz_zero_start=0, periodic=0,
                   precision=FP64, gram=GRAM_DIRECT, ar_hidden=64,
                   ar_train_every=10, ar_train_window=16384, ar_epochs=2,
                   ar_batch=2048, ar_lr=3e-3, ar_local=0, ar_seed=0)
    for k, v in overrides.items():
        if not hasattr(o, k):
            raise AttributeError(f"nqs_hip_options_t has no field {k!r}")
        setattr(o, k, v)
    return o


_P = c_void_p
_CFG = POINTER(NqsConfig)
_INFO = POINTER(SrInfo)

# name suffix -> (restype, argtypes); pointers to arrays are c_void_p.
API = {
    "ansatz_create": (c_void_p, [_CFG, c_int]),
    "ansatz_free": (None, [_P]),
    "ansatz_num_params": (c_long, [_P]),
    "ansatz_log_amp": (None, [_P, c_int, _P, _P, _P]),
    "ansatz_logpsi_gradient": (c_int, [_P, _P, c_int, _P]),
    "ansatz_logpsi_gradient_complex": (c_int, [_P, _P, c_int, _P, _P]),
    "ansatz_is_complex": (c_int, [_P]),
    "ansatz_apply_update": (c_int, [_P, _P, c_double]),
    "ansatz_params_raw": (c_void_p, [_P]),
    "marshall_parity": (c_int, [_P, c_int, c_int]),
    "marshall_log_amp": (None, [_P, c_int, _P, _P, _P]),
    "sampler_create": (c_void_p, [c_int, _CFG, _P, _P]),
    "sampler_free": (None, [_P]),
    "sampler_thermalize": (None, [_P]),
    "sampler_next": (c_void_p, [_P]),
    "sampler_batch": (c_int, [_P, c_int, _P]),
    "sampler_acceptance_ratio": (c_double, [_P]),
    "sampler_current": (c_void_p, [_P]),
    "sampler_num_sites": (c_int, [_P]),
    "local_energy": (c_double, [_CFG, c_int, c_int, _P, _P, _P]),
    "local_energy_batch": (None, [_CFG, c_int, c_int, _P, c_int, _P, _P, _P]),
    "local_energy_complex": (None, [_CFG, c_int, c_int, _P, _P, _P, _P, _P]),
    "local_energy_batch_complex": (None, [_CFG, c_int, c_int, _P, c_int, _P, _P, _P, _P]),
    "sr_step": (c_int, [_CFG, c_int, c_int, _P, _P, _INFO]),
    "sr_run": (c_int, [_CFG, c_int, c_int, _P, _P, _P]),
    "sr_step_custom": (c_int, [_CFG, c_int, c_int, _P, _P, _P, _P, _INFO]),
    "sr_run_custom": (c_int, [_CFG, c_int, c_int, _P, _P, _P, _P, _P]),
    "sr_step_custom_full": (c_int, [_CFG, c_int, c_int, _P, _P, _P, _P, _P, _P, _INFO]),
    "sr_run_custom_full": (c_int, [_CFG, c_int, c_int, _P, _P, _P, _P, _P, _P, _P]),
    "sr_step_holomorphic": (c_int, [_CFG, c_int, c_int, _P, _P, _P, _P, _INFO]),
    "sr_run_holomorphic": (c_int, [_CFG, c_int, c_int, _P, _P, _P, _P, _P]),
    "sr_step_holomorphic_full": (c_int, [_CFG, c_int, c_int, _P, _P, _P, _P, _P, _P, _INFO]),
    "sr_run_holomorphic_full": (c_int, [_CFG, c_int, c_int, _P, _P, _P, _P, _P, _P, _P]),
    "sr_step_minsr_full": (c_int, [_CFG, c_int, c_int, _P, _P, _P, _P, _P, _P, _INFO]),
    "sr_run_minsr": (c_int, [_CFG, c_int, c_int, _P, _P, _P, _P, _P]),
    "sr_step_excited": (c_int, [_CFG, c_int, c_int, _P, _P, _P, _P, _P, _P, c_double, _INFO]),
    "compute_chi_F": (c_int, [_CFG, c_int, c_int, _P, _P, _P, _P]),
    "symproj_log_amp": (None, [_P, c_int, _P, _P, _P]),
    "symproj_gradient": (c_int, [_P, _P, _P, c_int, _P]),
    "symproj_gradient_complex": (c_int, [_P, _P, _P, c_int, _P, _P]),
    "translation_log_amp": (None, [_P, c_int, _P, _P, _P]),
    "translation_gradient": (c_int, [_P, _P, _P, c_int, _P]),
    "compute_kagome_bond_phase": (c_int, [_CFG, c_int, c_int, _P, _P, _P, _P, _P]),
    "sr_run_excited": (c_int, [_CFG, c_int, c_int, _P, _P, _P, _P, _P, _P, c_double, _P]),
    "tvmc_step_real_time": (c_int, [_CFG, c_double, c_int, c_int, _P, _P, _P, _P, _INFO]),
    "tvmc_step_heun": (c_int, [_CFG, c_double, c_int, c_int, _P, _P, _P, _P, _INFO]),
}

# Extensions of this library (include/nqs_hip/nqs_hip.h)
_OPT = POINTER(HipOptions)
HIP_API = {
    "nqs_hip_sampler_create_ex": (c_void_p, [c_int, _CFG, _OPT, _P, _P]),
    "nqs_hip_sampler_get_options": (c_int, [_P, _OPT]),
    "nqs_hip_sampler_last_energies": (c_int, [_P, _P, c_int]),
    "nqs_hip_sampler_rounds": (c_int, [_P]),
    "nqs_hip_local_energy_batch_ex": (c_int, [_CFG, c_int, c_int, _OPT, _P, c_int, _P, _P, _P, _P]),
    "nqs_hip_log_amp_batch": (c_int, [_P, _P, c_int, _P, _P]),
    "nqs_hip_logpsi_gradient_batch": (c_int, [_P, _P, c_int, c_int, _P, _P]),
    "nqs_hip_tau_int": (c_double, [_P, c_int]),
    "nqs_hip_last_stage_ms": (c_int, [_P, _P]),
    "nqs_hip_device_launches": (c_longlong, []),
    "nqs_hip_host_callback_calls": (c_longlong, []),
    "nqs_hip_version": (c_char_p, []),
    "nqs_hip_config_layout": (c_int, [_P, c_int]),
    "nqs_hip_gram": (c_int, [_P, c_int, c_long, c_double, _P]),
    "nqs_hip_minsr_solve": (c_int, [_P, _P, c_int, _P, c_double, c_int, _P, _P, _P]),
    "nqs_hip_minsr_solve_ex": (c_int, [_P, _P, c_int, _P, c_double, _OPT, _P, _P, _P]),
    "nqs_hip_kernel_timing": (c_int, [c_char_p]),
    "nqs_hip_copy_probe": (c_double, [c_long, c_int]),
    "nqs_hip_kernel_stats": (c_int, [c_char_p, POINTER(c_longlong), POINTER(c_double)]),
    "nqs_hip_cholesky_solve": (c_int, [_P, c_int, _P, _P]),
    "nqs_hip_cholesky_solve_noalloc": (c_int, [_P, c_int, _P, _P]),
    "nqs_hip_ar_create": (c_void_p, [c_int, c_ulonglong, c_double]),
    "nqs_hip_ar_free": (None, [_P]),
    "nqs_hip_ar_num_params": (c_long, [_P]),
    "nqs_hip_ar_get_params": (c_int, [_P, _P]),
    "nqs_hip_ar_set_params": (c_int, [_P, _P]),
    "nqs_hip_ar_log_prob": (c_int, [_P, _P, c_int, c_int, _P]),
    "nqs_hip_ar_sample": (c_int, [_P, _P, c_int, c_int, _P, _P, _P]),
    "nqs_hip_ar_gradient": (c_int, [_P, _P, c_int, c_int, _P, _P]),
    "nqs_hip_ar_train": (c_int, [_P, _P, c_int, c_int, c_int, c_int, c_double, _P]),
    "nqs_hip_sampler_ar": (c_void_p, [_P]),
    "nqs_hip_sampler_learned_stats": (c_int, [_P, _P, _P, _P, _P]),
    "nqs_hip_multi_create": (c_void_p, [c_int, _P, _P, _P, _P, _P]),
    "nqs_hip_multi_step": (c_int, [_P, _P]),
    "nqs_hip_multi_run": (c_int, [_P, c_int, _P]),
    "nqs_hip_multi_sync": (c_int, [_P]),
    "nqs_hip_multi_tuning": (c_int, [_P, _P, _P, c_int]),
    "nqs_hip_multi_free": (None, [_P]),
    "nqs_hip_multi_use_graph": (c_int, [_P, c_int]),
    "nqs_hip_multi_use_persistent": (c_int, [_P, c_int]),
    "nqs_hip_multi_last_sample_ms": (c_int, [_P, _P]),
    "nqs_hip_multi_last_energies": (c_int, [_P, c_int, _P, c_int]),
}


# exact diagonalisation (include/nqs/nqs_lanczos.h), the reference's names
_LR = POINTER(LanczosResult)
ED_API = {
    "nqs_exact_energy_kagome_heisenberg": (c_int, [_P, c_int, c_int, c_double, c_int, _P]),
    "nqs_exact_energy_kagome_heisenberg_with_cb": (c_int, [_P, _P, c_int, c_int, c_double, c_int,
                                                           _P]),
    "nqs_lanczos_e0_kagome_heisenberg_projected_lean": (c_int, [c_int, c_int, c_double, c_int, _P,
                                                                _P, c_int, c_int, c_double, _P,
                                                                _LR]),
    "nqs_lanczos_e0_kagome_heisenberg_projected_lean_eigvec": (
        c_int, [c_int, c_int, c_double, c_int, _P, _P, c_int, c_int, c_double, _P, _P, _LR]),
    "nqs_materialise_state": (c_int, [_P, c_int, c_int, _P, _P]),
    "nqs_materialise_state_with_cb": (c_int, [_P, _P, c_int, c_int, _P, _P]),
    "nqs_materialise_state_with_cb_N": (c_int, [_P, _P, c_int, _P, _P]),
    "nqs_exact_energy_tfim": (c_int, [_P, c_int, c_int, c_double, c_double, _P]),
    "nqs_exact_energy_heisenberg": (c_int, [_P, c_int, c_int, c_double, c_double, _P]),
    "nqs_lanczos_refine_tfim": (c_int, [_P, c_int, c_int, c_double, c_double, c_int, c_double,
                                        _P, _P, _LR]),
    "nqs_lanczos_refine_heisenberg": (c_int, [_P, c_int, c_int, c_double, c_double, c_int,
                                              c_double, _P, _P, _LR]),
    "nqs_lanczos_refine_kagome_heisenberg": (c_int, [_P, c_int, c_int, c_double, c_int, c_int,
                                                     c_double, _P, _P, _LR]),
    "nqs_lanczos_refine_kagome_heisenberg_with_cb": (c_int, [_P, _P, c_int, c_int, c_double,
                                                             c_int, c_int, c_double, _P, _P,
                                                             _LR]),
    "nqs_lanczos_k_lowest_kagome_heisenberg": (c_int, [_P, c_int, c_int, c_double, c_int, c_int,
                                                       c_int, _P, _LR]),
    "nqs_lanczos_k_lowest_kagome_heisenberg_with_cb": (c_int, [_P, _P, c_int, c_int, c_double,
                                                               c_int, c_int, c_int, _P, _LR]),
    "nqs_lanczos_k_lowest_kagome_heisenberg_projected": (
        c_int, [_P, _P, c_int, c_int, c_double, c_int, _P, _P, c_int, c_int, c_int, _P, _LR]),
    "nqs_lanczos_refine_kagome_heisenberg_projected": (
        c_int, [_P, _P, c_int, c_int, c_double, c_int, _P, _P, c_int, c_int, c_double, _P, _P,
                _LR]),
}


# kagome group tables (include/nqs/nqs_symproj.h), the reference's names
SYM_API = {
    "nqs_kagome_translation_perm": (c_int, [c_int, c_int, _P, _P]),
    "nqs_kagome_p2_perm": (c_int, [c_int, c_int, _P, _P, _P]),
    "nqs_kagome_p3_perm": (c_int, [c_int, _P, _P, _P]),
    "nqs_kagome_p6_perm": (c_int, [c_int, _P, _P, _P]),
    "nqs_kagome_p6m_perm": (c_int, [c_int, _P, _P, _P]),
    "nqs_kagome_p6m_perm_irrep": (c_int, [c_int, c_int, _P, _P, _P]),
    "nqs_kagome_p6m_project_inplace": (None, [_P, c_int, c_int, _P, _P]),
}


def kagome_group(lib, kind, *size, irrep=0):
    """(perm [G][N] int32, characters [G]) from a library's builder:
    kind in translation, p2, p3, p6, p6m (irrep: nqs_symproj_kagome_irrep_t).
    The C arrays are copied and released with libc free."""
    perm_p = ctypes.POINTER(c_int)()
    chi_p = ctypes.POINTER(c_double)()
    G = c_int(0)
    name = f"nqs_kagome_{kind}_perm" + ("_irrep" if kind == "p6m" else "")
    f = getattr(lib, name)
    if kind == "translation":
        rc = f(*size, ctypes.byref(perm_p), ctypes.byref(G))
    elif kind == "p6m":
        rc = f(*size, irrep, ctypes.byref(perm_p), ctypes.byref(chi_p), ctypes.byref(G))
    else:
        rc = f(*size, ctypes.byref(perm_p), ctypes.byref(chi_p), ctypes.byref(G))
    if rc != 0:
        raise RuntimeError(f"{name} returned {rc}")
    N = 3 * int(np.prod(size)) if kind in ("translation", "p2") else 3 * size[0] * size[0]
    perm = np.ctypeslib.as_array(perm_p, (G.value * N,)).reshape(G.value, N).copy()
    chi = (np.ctypeslib.as_array(chi_p, (G.value,)).copy() if kind != "translation"
           else np.ones(G.value))
    libc = ctypes.CDLL(None)
    libc.free(ctypes.cast(perm_p, c_void_p))
    if kind != "translation":
        libc.free(ctypes.cast(chi_p, c_void_p))
    return perm.astype(np.int32), chi


def bind(lib, prefix, table=API, strict=True):
    """Attach restype/argtypes for ``prefix + name`` of every table entry."""
    for name, (res, args) in table.items():
        full = prefix + name
        try:
            fn = getattr(lib, full)
        except AttributeError:
            if strict:
                raise
            continue
        fn.restype = res
        fn.argtypes = args
    return lib


def fnptr(lib, name):
    """Address of an exported function, to pass as a C callback."""
    return ctypes.cast(getattr(lib, name), c_void_p).value


def ptr(a):
    """Data pointer of a C-contiguous numpy array (or None)."""
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "array must be C-contiguous"
    return a.ctypes.data


_LIB = None


def load_library(path=None):
    """Open the in-tree ``libnqs_hip.so``; raise when it was not built."""
    global _LIB
    if _LIB is not None and path is None:
        return _LIB
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(
            f"{p} not found: build it with `python -c 'import __graft_entry__ as g; g.build()'`")
    lib = ctypes.CDLL(p, mode=ctypes.RTLD_LOCAL)
    bind(lib, "nqs_")
    bind(lib, "", HIP_API)
    bind(lib, "", ED_API)
    bind(lib, "", SYM_API)
    lib.nqs_hip_options_defaults.restype = HipOptions
    lib.nqs_hip_options_defaults.argtypes = []
    if path is None:
        _LIB = lib
    return lib


def spins_array(x):
    """int32 C-contiguous copy of a spin array."""
    return np.ascontiguousarray(np.asarray(x, dtype=np.int32))
