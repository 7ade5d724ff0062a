"""ctypes mirror of include/yaohip.h and the loader of libyaohip.so.

The loader raises if the built library is missing: there is no CPU fallback for
any call that touches amplitudes.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libyaohip.so")

# ---- enums (include/yaohip.h) ----
YH_OK = 0
STATUS_NAMES = {
    0: "OK", 1: "INVALID_ARG", 2: "CONTROL_CONFIG_LEN", 3: "LOC_OUT_OF_RANGE",
    4: "OVERLAPPING_LOCS", 5: "MATRIX_SIZE", 6: "NBITS_MISMATCH", 7: "MEASURE_LOC",
    8: "CHANNEL_PARAM", 9: "HIP", 10: "OUT_OF_MEMORY", 11: "UNSUPPORTED", 12: "DTYPE",
    13: "RESET_VALUE", 14: "NO_DEVICE", 15: "IO", 16: "FORMAT",
}
C128, C64 = 0, 1
ELEM_GATE, ELEM_CHANNEL, ELEM_ANNOTATION = 0, 1, 2
(GATE_X, GATE_Y, GATE_Z, GATE_H, GATE_S, GATE_T, GATE_SWAP, GATE_PHASE, GATE_RX, GATE_RY,
 GATE_RZ, GATE_SQRTX, GATE_SQRTY, GATE_SQRTW, GATE_ISWAP, GATE_FSIM, GATE_CUSTOM) = range(17)
GATE_NAMES = ["X", "Y", "Z", "H", "S", "T", "SWAP", "Phase", "Rx", "Ry", "Rz", "SqrtX",
              "SqrtY", "SqrtW", "ISWAP", "FSim", "Custom"]
(CH_BITFLIP, CH_PHASEFLIP, CH_DEPOLARIZING, CH_PAULI, CH_RESET, CH_AMPLITUDE_DAMPING,
 CH_PHASE_DAMPING, CH_PHASE_AMPLITUDE_DAMPING, CH_THERMAL_RELAXATION, CH_COHERENT,
 CH_CUSTOM) = range(11)
OP_I, OP_X, OP_Y, OP_Z, OP_P0, OP_P1, OP_PU, OP_PD = range(8)
MLP_LINEAR, MLP_TRIG = 0, 1
(K_1Q, K_DENSE, K_DIAG, K_X, K_SWAP, K_OOP, K_FUSED, K_TILE, K_DM, K_VQE, K_TBLOCK) = range(11)
KCLASS_NAMES = ["1q", "dense", "diag", "x", "swap", "oop", "fused", "tile", "dm", "vqe", "tblock"]
MAX_TARGETS = 12
MAX_CTRLS = 16


class YhOp(ctypes.Structure):
    """struct yh_op (include/yaohip.h)."""
    _fields_ = [
        ("elem", ctypes.c_int32),
        ("kind", ctypes.c_int32),
        ("nt", ctypes.c_int32),
        ("nc", ctypes.c_int32),
        ("targets", ctypes.c_int32 * MAX_TARGETS),
        ("ctrls", ctypes.c_int32 * MAX_CTRLS),
        ("ctrl_vals", ctypes.c_int32 * MAX_CTRLS),
        ("params", ctypes.c_double * 4),
        ("matrix", ctypes.POINTER(ctypes.c_double)),
        ("nmat", ctypes.c_int32),
        ("dim", ctypes.c_int32),
        ("is_diag", ctypes.c_int32),
        ("nconf", ctypes.c_int32),
    ]


class YhError(RuntimeError):
    """A non-zero status returned through the C ABI."""

    def __init__(self, status, message):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {message}")
        self.status = status
        self.message = message


def dptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def i32ptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))


def i64ptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))


def u64ptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))


def interleave(z):
    """complex128 array -> contiguous float64 (re, im) pairs."""
    z = np.ascontiguousarray(np.asarray(z, dtype=np.complex128))
    return z.view(np.float64).reshape(-1).copy()


def deinterleave(d):
    return np.ascontiguousarray(d).view(np.complex128).copy()


class OpList:
    """A packed yh_op array plus the matrix buffers it points to (kept alive)."""

    def __init__(self, ops, keep):
        self.n = len(ops)
        self.arr = (YhOp * max(1, self.n))(*ops)
        self._keep = keep

    @property
    def ptr(self):
        return self.arr


_lib = None

# (name, restype, argtypes) of every exported entry point declared in include/yaohip.h
P = ctypes.POINTER
c_int, c_size_t, c_double = ctypes.c_int, ctypes.c_size_t, ctypes.c_double
c_void_p = ctypes.c_void_p
SIGNATURES = [
    ("yh_last_error", ctypes.c_char_p, []),
    ("yh_init", c_int, []),
    ("yh_device_name", c_int, [ctypes.c_char_p, c_size_t]),
    ("yh_copy_bandwidth", c_int, [c_size_t, c_int, P(ctypes.c_double)]),
    ("yh_reg_create", c_int, [c_int, c_int, P(c_void_p)]),
    ("yh_reg_destroy", c_int, [c_void_p]),
    ("yh_reg_nbits", c_int, [c_void_p, P(c_int)]),
    ("yh_reg_dtype", c_int, [c_void_p, P(c_int)]),
    ("yh_reg_upload", c_int, [c_void_p, P(c_double), c_size_t]),
    ("yh_reg_download", c_int, [c_void_p, P(c_double), c_size_t]),
    ("yh_reg_clone", c_int, [c_void_p, P(c_void_p)]),
    ("yh_reg_copy", c_int, [c_void_p, c_void_p]),
    ("yh_reg_norm", c_int, [c_void_p, P(c_double)]),
    ("yh_reg_normalize", c_int, [c_void_p]),
    ("yh_reg_fidelity", c_int, [c_void_p, c_void_p, P(c_double)]),
    ("yh_reg_device_ptr", c_int, [c_void_p, P(c_void_p)]),
    ("yh_reg_product_state", c_int, [c_int, c_int, ctypes.c_uint64, P(c_void_p)]),
    ("yh_reg_uniform_state", c_int, [c_int, c_int, P(c_void_p)]),
    ("yh_reg_ghz_state", c_int, [c_int, c_int, P(c_void_p)]),
    ("yh_reg_deterministic_state", c_int, [c_int, c_int, P(c_void_p)]),
    ("yh_reg_rand_state", c_int, [c_int, c_int, P(c_double), c_size_t, P(c_void_p)]),
    ("yh_state_write", c_int, [c_void_p, ctypes.c_char_p]),
    ("yh_state_read", c_int, [ctypes.c_char_p, c_int, P(c_void_p)]),
    ("yh_state_read_header", c_int, [ctypes.c_char_p, P(c_int)]),
    ("yh_validate", c_int, [c_int, P(YhOp), c_size_t]),
    ("yh_apply", c_int, [c_void_p, c_int, P(YhOp), c_size_t]),
    ("yh_apply_fused", c_int, [c_void_p, c_int, P(YhOp), c_size_t, c_int]),
    ("yh_instruct", c_int, [c_void_p, P(YhOp)]),
    ("yh_circuit_num_params", c_int, [P(YhOp), c_size_t, P(c_size_t)]),
    ("yh_expect_grad", c_int, [c_void_p, c_int, P(YhOp), c_size_t, c_size_t, P(c_double), P(ctypes.c_int32),
                               P(ctypes.c_int32), P(ctypes.c_int32), P(c_double), P(c_double), c_size_t]),
    ("yh_plan_stats", c_int, [c_int, P(YhOp), c_size_t, P(ctypes.c_int64)]),
    ("yh_block_plan_stats", c_int, [c_int, P(YhOp), c_size_t, P(ctypes.c_int64)]),
    ("yh_block_pass_stats", c_int, [c_int, P(YhOp), c_size_t, c_int, P(ctypes.c_int64)]),
    ("yh_block_plan", c_int, [c_int, P(YhOp), c_size_t, P(ctypes.c_int32), c_size_t, P(c_size_t)]),
    ("yh_gate_matrix", c_int, [P(YhOp), P(c_double), P(c_int)]),
    ("yh_expect", c_int, [c_void_p, c_size_t, P(c_double), P(ctypes.c_int32),
                          P(ctypes.c_int32), P(ctypes.c_int32), P(c_double)]),
    ("yh_expect_noalloc", c_int, [c_void_p, c_int, c_int, c_size_t, P(c_double), P(ctypes.c_int32),
                                  P(ctypes.c_int32), P(ctypes.c_int32), c_void_p, P(c_double)]),
    ("yh_probs", c_int, [c_void_p, P(ctypes.c_int32), c_int, P(c_double)]),
    ("yh_sample", c_int, [c_void_p, P(ctypes.c_int32), c_int, P(c_double), c_size_t,
                          P(ctypes.c_uint64)]),
    ("yh_measure", c_int, [c_void_p, P(ctypes.c_int32), c_int, c_int, ctypes.c_uint64,
                           c_double, P(ctypes.c_int32)]),
    ("yh_dm_create", c_int, [c_int, P(c_void_p)]),
    ("yh_dm_from_reg", c_int, [c_void_p, P(c_void_p)]),
    ("yh_dm_mixed", c_int, [P(c_double), P(c_void_p), c_size_t, P(c_void_p)]),
    ("yh_dm_from_reg_locs", c_int, [c_void_p, P(ctypes.c_int32), c_int, P(c_void_p)]),
    ("yh_dm_destroy", c_int, [c_void_p]),
    ("yh_dm_nbits", c_int, [c_void_p, P(c_int)]),
    ("yh_dm_upload", c_int, [c_void_p, P(c_double), c_size_t]),
    ("yh_dm_download", c_int, [c_void_p, P(c_double), c_size_t]),
    ("yh_dm_apply", c_int, [c_void_p, c_int, P(YhOp), c_size_t]),
    ("yh_dm_apply_fused", c_int, [c_void_p, c_int, P(YhOp), c_size_t]),
    ("yh_dm_apply_blocks", c_int, [c_void_p, c_int, P(YhOp), c_size_t]),
    ("yh_dm_trace", c_int, [c_void_p, P(c_double)]),
    ("yh_dm_purity", c_int, [c_void_p, P(c_double)]),
    ("yh_dm_partial_tr", c_int, [c_void_p, P(ctypes.c_int32), c_int, P(c_void_p)]),
    ("yh_dm_entropy", c_int, [c_void_p, P(c_double)]),
    ("yh_dm_expect", c_int, [c_void_p, c_size_t, P(c_double), P(ctypes.c_int32),
                             P(ctypes.c_int32), P(ctypes.c_int32), P(c_double)]),
    ("yh_dm_probs", c_int, [c_void_p, P(ctypes.c_int32), c_int, P(c_double)]),
    ("yh_channel_kraus", c_int, [P(YhOp), P(c_double), c_int, P(c_int), P(c_int)]),
    ("yh_reg_timing_enable", c_int, [c_void_p, c_int]),
    ("yh_reg_timing_read", c_int, [c_void_p, c_int, P(ctypes.c_uint64), P(c_double)]),
    ("yh_dm_timing_enable", c_int, [c_void_p, c_int]),
    ("yh_dm_timing_read", c_int, [c_void_p, c_int, P(ctypes.c_uint64), P(c_double)]),
    ("yh_vqe_batch_create", c_int, [c_int, c_int, c_int, P(c_void_p)]),
    ("yh_vqe_batch_destroy", c_int, [c_void_p]),
    ("yh_vqe_batch_configure", c_int, [c_void_p, c_int, c_int, c_int, c_int]),
    ("yh_vqe_batch_config", c_int, [c_void_p, P(ctypes.c_int32)]),
    ("yh_vqe_batch_energies", c_int, [c_void_p, c_int, P(ctypes.c_int32), P(ctypes.c_int32),
                                      P(ctypes.c_int64), P(c_double), c_int, c_double, c_double,
                                      P(c_double)]),
    ("yh_vqe_plain_energies", c_int, [c_int, P(ctypes.c_int32), P(ctypes.c_int32),
                                      P(ctypes.c_int64), P(c_double), c_int, c_double, c_double, c_int,
                                      P(c_double)]),
    ("yh_mlp_create", c_int, [c_int, c_int, c_int, c_int, c_int, ctypes.c_uint64, P(c_void_p)]),
    ("yh_mlp_destroy", c_int, [c_void_p]),
    ("yh_mlp_num_params", c_int, [c_void_p, P(c_size_t)]),
    ("yh_mlp_get_params", c_int, [c_void_p, P(ctypes.c_float), c_size_t]),
    ("yh_mlp_set_params", c_int, [c_void_p, P(ctypes.c_float), c_size_t]),
    ("yh_mlp_predict", c_int, [c_void_p, P(ctypes.c_float), c_size_t, P(ctypes.c_float)]),
    ("yh_mlp_predict_f64", c_int, [c_void_p, P(c_double), c_size_t, P(ctypes.c_float)]),
    ("yh_mlp_loss_grad", c_int, [c_void_p, P(ctypes.c_float), P(ctypes.c_float), c_int, P(ctypes.c_float),
                                 P(ctypes.c_float)]),
    ("yh_mlp_train", c_int, [c_void_p, P(ctypes.c_float), P(ctypes.c_float), c_size_t, c_int, c_int,
                             ctypes.c_float, ctypes.c_uint64, P(ctypes.c_float)]),
]


def lib():
    """Load libyaohip.so once; raise loudly when it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"libyaohip.so not found at {LIB_PATH}: build it first "
                "(python -c 'import __graft_entry__ as g; g.build()')")
        L = ctypes.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(status):
    if status != YH_OK:
        msg = lib().yh_last_error()
        raise YhError(status, msg.decode() if msg else "")
    return status


_inited = False


def init():
    """Bind device 0 (a gfx950 agent)."""
    global _inited
    if not _inited:
        check(lib().yh_init())
        _inited = True
