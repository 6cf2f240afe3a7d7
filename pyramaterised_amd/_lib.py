"""ctypes binding of ``lib/libpqs.so`` (declarations in include/pqs.h).

Loading the library needs no GPU; every compute entry point does.  There is
no CPU fallback: when the library is missing, ``lib()`` raises.
"""
import ctypes as C
import os
import threading

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libpqs.so")

PQS_OK = 0
PQS_ERR_ARG, PQS_ERR_HIP, PQS_ERR_UNSUPPORTED, PQS_ERR_NOMEM, PQS_ERR_NODEVICE = -1, -2, -3, -4, -5
PQS_C128, PQS_C64 = 0, 1
# pqs_plan_opts.flags
PQS_PLAN_LANES_ASCENDING, PQS_PLAN_NO_SOFT_REGS, PQS_PLAN_JIT_PERSISTENT, PQS_PLAN_ROLLING = 1, 2, 4, 8
PQS_PLAN_NO_WAVE_LOCAL = 16
PQS_PLAN_NO_DIRECT = 32
PQS_PLAN_GREEDY_SWEEPS = 64
PQS_PLAN_GREEDY_PHASES = 128
PQS_PLAN_NO_EXCHANGE = 256

(OP_MAT1, OP_RX, OP_RY, OP_RZ, OP_MAT2, OP_CX, OP_CZ, OP_RXX, OP_RYY, OP_RZZ,
 OP_FSIM, OP_GEN, OP_GEN_FSIM) = range(13)


class Op(C.Structure):
    _fields_ = [("kind", C.c_int32), ("q0", C.c_int32), ("q1", C.c_int32),
                ("slot", C.c_int32), ("slot2", C.c_int32), ("tangent", C.c_int32),
                ("pauli", C.c_int32), ("reserved", C.c_int32),
                ("scale", C.c_double), ("offset", C.c_double), ("m", C.c_double * 32)]


class PlanOpts(C.Structure):
    _fields_ = [("tile_bits", C.c_int32), ("low_bits", C.c_int32),
                ("block", C.c_int32), ("dtype", C.c_int32), ("flags", C.c_int32)]


class Instance(C.Structure):
    _fields_ = [("state_dev", C.c_void_p), ("src_dev", C.c_void_p),
                ("params_dev", C.c_void_p), ("tangent", C.c_int32), ("init", C.c_int32)]


vp, i32, i64, dp = C.c_void_p, C.c_int32, C.c_int64, C.c_void_p

SIGNATURES = {
    "pqs_abi_version": (C.c_int, []),
    "pqs_last_error": (C.c_char_p, []),
    "pqs_range_push": (C.c_int, [C.c_char_p]),
    "pqs_range_pop": (C.c_int, []),
    "pqs_ctx_create": (C.c_int, [C.c_int, C.POINTER(vp)]),
    "pqs_ctx_destroy": (C.c_int, [vp]),
    "pqs_device_info": (C.c_int, [vp, C.POINTER(i32), C.POINTER(i32), C.c_char_p, i32]),
    "pqs_upload": (C.c_int, [vp, vp, vp, i64, vp]),
    "pqs_ctx_set_max_blocks": (C.c_int, [vp, i64]),
    "pqs_plan_create": (C.c_int, [i32, C.POINTER(Op), i32, C.POINTER(PlanOpts), C.POINTER(vp)]),
    "pqs_plan_destroy": (C.c_int, [vp]),
    "pqs_plan_info": (C.c_int, [vp, C.POINTER(i32), C.POINTER(i32), C.POINTER(i32), C.POINTER(i32)]),
    "pqs_plan_sweep": (C.c_int, [vp, i32, C.POINTER(i32), C.POINTER(i32), C.POINTER(i32), C.POINTER(i32)]),
    "pqs_plan_sweep_layout": (C.c_int, [vp, i32, C.POINTER(i32), C.POINTER(i32)]),
    "pqs_batch_table_bytes": (i64, [i64]),
    "pqs_batch_table": (C.c_int, [vp, vp, vp, i64, i64, vp, i64, i32, vp, vp, vp, vp]),
    "pqs_run_table": (C.c_int, [vp, vp, vp, i64, vp, vp]),
    "pqs_plan_tangent_sweep": (C.c_int, [vp, i32, C.POINTER(i32)]),
    "pqs_plan_tangent_parts": (C.c_int, [vp, i32, C.POINTER(i32)]),
    "pqs_run": (C.c_int, [vp, vp, C.POINTER(Instance), i64, vp, vp]),
    "pqs_run_batch": (C.c_int, [vp, vp, vp, i64, i64, vp, i64, vp, i32, vp, vp, vp]),
    "pqs_run_batch_pairs": (C.c_int, [vp, vp, vp, i64, i64, vp, i64, i32, vp, vp, vp, vp]),
    "pqs_apply_dense": (C.c_int, [vp, i32, vp, i64, vp, i64, i64, vp, vp]),
    "pqs_apply_pauli": (C.c_int, [vp, i32, vp, i32, vp, vp, vp, i32, vp, vp]),
    "pqs_pauli_transition": (C.c_int, [vp, i32, vp, vp, i32, vp, vp, vp, i32, vp, vp]),
    "pqs_run_multi": (C.c_int, [vp, vp, i32, vp, vp, vp, i64, vp]),
    "pqs_run_batch_noalloc": (C.c_int, [vp, vp, vp, i64, i64, vp, i64, i32, vp, vp, vp, vp]),
    "pqs_jacobian": (C.c_int, [vp, vp, vp, i32, vp, vp, vp, vp, vp]),
    "pqs_inner": (C.c_int, [vp, i32, vp, i64, vp, vp, i64, vp, vp]),
    "pqs_mw": (C.c_int, [vp, i32, vp, i32, i64, vp, vp, vp]),
    "pqs_gram": (C.c_int, [vp, i32, vp, i64, i64, vp, i64, i64, i64, vp, vp]),
    "pqs_qfim": (C.c_int, [vp, i32, vp, i64, vp, i64, vp, vp]),
    "pqs_expect_pauli": (C.c_int, [vp, i32, vp, i32, i64, vp, vp, vp, i32, vp, vp]),
    "pqs_histogram": (C.c_int, [vp, vp, i64, vp, i32, vp, vp]),
    "pqs_abs2": (C.c_int, [vp, vp, i64, vp, vp]),
    "pqs_magic": (C.c_int, [vp, i32, vp, i32, i64, C.c_double, vp, vp]),
    "pqs_plan_jit": (C.c_int, [vp, vp, i32]),
    "pqs_plan_jit_source": (C.c_int, [vp, C.c_char_p, i64, C.POINTER(i64), C.POINTER(i32)]),
    "pqs_plan_jit_sweep_source": (C.c_int, [vp, i32, C.c_char_p, i64, C.POINTER(i64)]),
    "pqs_jit_prepare": (C.c_int, [C.c_char_p, i64]),
}

_lock = threading.Lock()
_lib = None


class EngineError(RuntimeError):
    pass


_fingerprint = None


def library_fingerprint():
    """Hash of the built library: part of the tile-shape tuner's cache key, so
    a rebuilt planner or kernel set is tuned afresh."""
    global _fingerprint
    if _fingerprint is None:
        import hashlib
        with open(LIB_PATH, "rb") as f:
            _fingerprint = hashlib.sha1(f.read()).hexdigest()[:16]
    return _fingerprint


def lib():
    """The loaded library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise EngineError(
                f"{LIB_PATH} is missing: build it with `python -m pyramaterised_amd.build` "
                "(the engine has no CPU fallback)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        if L.pqs_abi_version() != 1:
            raise EngineError("libpqs.so ABI version mismatch")
        _lib = L
    return _lib


def check(rc):
    if rc != PQS_OK:
        msg = lib().pqs_last_error().decode(errors="replace")
        raise EngineError(f"libpqs error {rc}: {msg}")
    return rc
