"""CPU oracle bindings -- TEST INFRASTRUCTURE ONLY.

Loads ``oracle/liboracle.so`` (the plain-C FP64 restatement of the reference
algorithm, oracle/nqs_oracle.c, and of the learned-proposal network,
oracle/ar_oracle.c) and, where it was built,
``oracle/_ref/libnqs_ref.so`` (the reference's own sources compiled by
oracle/Makefile).  Only tests/, ``__graft_entry__.smoke()`` and the
cpu_baseline leg of bench.py import this package; the product never does.
"""
import ctypes
import os
from ctypes import POINTER, c_double, c_int, c_uint, c_ulonglong, c_void_p

from spin_based_neural_network_amd.abi import (API, ED_API, SYM_API, HipOptions, NqsConfig,
                                               bind)

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_LIB = os.path.join(HERE, "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libnqs_ref.so")

_EXTRA = {
    "orc_ansatz_num_hidden": (c_int, [c_void_p]),
    "orc_sampler_create_ex": (c_void_p, [c_int, POINTER(NqsConfig), POINTER(HipOptions), c_void_p, c_void_p]),
    "orc_sampler_chain_state": (c_void_p, [c_void_p, c_int]),
    "orc_sampler_rng_state": (c_ulonglong, [c_void_p, c_int]),
    "orc_chain_seed": (c_ulonglong, [c_uint, c_int, c_int]),
    "orc_local_energy_batch_ex": (None, [POINTER(NqsConfig), c_int, c_int, c_int, c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_void_p]),
    "orc_cg_solve": (c_int, [c_void_p, ctypes.c_long, ctypes.c_long, c_void_p, c_double, c_void_p, c_int, c_double, c_void_p, c_void_p]),
    "orc_cg_solve_complex": (c_int, [c_void_p, c_void_p, ctypes.c_long, ctypes.c_long, c_void_p, c_void_p, c_double, c_void_p, c_int, c_double, c_void_p, c_void_p]),
    "orc_cholesky_lower": (c_int, [c_void_p, c_int]),
    "orc_cholesky_solve": (None, [c_void_p, c_int, c_void_p, c_void_p]),
    "orc_minsr_solve": (c_int, [c_void_p, ctypes.c_long, ctypes.c_long, c_void_p, c_double, c_void_p, c_void_p, c_void_p]),
    "orc_ed_ground_energy": (c_double, [POINTER(NqsConfig), c_int, c_int, c_int, c_int, c_int, c_double]),
    "orc_exact_energy": (c_double, [POINTER(NqsConfig), c_int, c_int, c_int, c_void_p, c_void_p]),
    "orc_tau_int": (c_double, [c_void_p, c_int]),
    "orc_ar_num_params": (ctypes.c_long, [c_int]),
    "orc_ar_init": (None, [c_int, c_ulonglong, c_double, c_void_p]),
    "orc_ar_log_prob": (c_int, [c_int, c_void_p, c_void_p, c_int, c_int, c_int, c_void_p]),
    "orc_ar_gradient": (c_int, [c_int, c_void_p, c_void_p, c_int, c_int, c_int, c_void_p, c_void_p]),
    "orc_ar_adam_step": (None, [c_int, c_void_p, c_void_p, c_void_p, c_void_p, ctypes.c_long, c_double]),
}

_orc = None


def load_oracle():
    """The oracle library (raises if it was not built)."""
    global _orc
    if _orc is None:
        if not os.path.exists(ORACLE_LIB):
            raise RuntimeError(f"{ORACLE_LIB} missing: run `make -C oracle`")
        lib = ctypes.CDLL(ORACLE_LIB, mode=ctypes.RTLD_LOCAL)
        bind(lib, "orc_")
        bind(lib, "", _EXTRA)
        _orc = lib
    return _orc


def load_reference():
    """The compiled reference, or None where it was not built."""
    if not os.path.exists(REF_LIB):
        return None
    lib = ctypes.CDLL(REF_LIB, mode=ctypes.RTLD_LOCAL)
    bind(lib, "nqs_")
    bind(lib, "", SYM_API, strict=False)
    bind(lib, "", ED_API, strict=False)
    return lib
