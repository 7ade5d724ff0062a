"""ctypes binding of the CPU oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker of the HIP backend. It shares the
`struct yh_op` layout of include/yaohip.h (a data format) but no code with the
product library.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "liboracle.so")

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        build()  # make is a no-op when the library is current
        L = ctypes.CDLL(LIB)
        P = ctypes.POINTER
        d, i32, u64, vp = ctypes.c_double, ctypes.c_int32, ctypes.c_uint64, ctypes.c_void_p
        sig = {
            "orc_itercontrol": (ctypes.c_int, [ctypes.c_int, P(i32), P(i32), ctypes.c_int, P(u64)]),
            "orc_gate_matrix": (ctypes.c_int, [vp, P(d), P(ctypes.c_int)]),
            "orc_validate": (ctypes.c_int, [ctypes.c_int, vp, ctypes.c_size_t]),
            "orc_instruct": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, P(d), P(i32), ctypes.c_int, P(d),
                                            P(i32), P(i32), ctypes.c_int]),
            "orc_apply": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, P(d), vp, ctypes.c_size_t]),
            "orc_dispatch": (ctypes.c_int, [ctypes.c_int, P(d), vp]),
            "orc_set_threads": (ctypes.c_int, [ctypes.c_int]),
            "orc_expect": (ctypes.c_int, [ctypes.c_int, P(d), ctypes.c_size_t, P(d), P(i32), P(i32), P(i32), P(d)]),
            "orc_probs": (ctypes.c_int, [ctypes.c_int, P(d), P(i32), ctypes.c_int, P(d)]),
            "orc_sample_from_probs": (u64, [P(d), ctypes.c_size_t, d]),
            "orc_sample": (ctypes.c_int, [ctypes.c_int, P(d), P(i32), ctypes.c_int, P(d), ctypes.c_size_t, P(u64)]),
            "orc_measure": (ctypes.c_int, [ctypes.c_int, P(d), P(i32), ctypes.c_int, ctypes.c_int, u64, d, P(i32),
                                           P(d)]),
            "orc_norm": (d, [ctypes.c_int, P(d)]),
            "orc_fidelity": (d, [ctypes.c_int, P(d), P(d)]),
            "orc_kraus": (ctypes.c_int, [vp, P(d), ctypes.c_int, P(ctypes.c_int), P(ctypes.c_int)]),
            "orc_superop": (ctypes.c_int, [vp, P(d), P(ctypes.c_int)]),
            "orc_dm_from_reg": (ctypes.c_int, [ctypes.c_int, P(d), P(d)]),
            "orc_deterministic_state": (ctypes.c_int, [ctypes.c_int, P(d)]),
            "orc_rand_state": (ctypes.c_int, [ctypes.c_int, P(d), P(d)]),
            "orc_dm_mixed": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, P(d), P(d), P(d)]),
            "orc_dm_apply": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, P(d), vp, ctypes.c_size_t]),
            "orc_dm_trace": (ctypes.c_int, [ctypes.c_int, P(d), P(d)]),
            "orc_dm_purity": (d, [ctypes.c_int, P(d)]),
            "orc_dm_expect": (ctypes.c_int, [ctypes.c_int, P(d), ctypes.c_size_t, P(d), P(i32), P(i32), P(i32),
                                             P(d)]),
            "orc_dm_probs": (ctypes.c_int, [ctypes.c_int, P(d), P(i32), ctypes.c_int, P(d)]),
            "orc_dm_partial_tr": (ctypes.c_int, [ctypes.c_int, P(d), P(i32), ctypes.c_int, P(d),
                                                 P(ctypes.c_int)]),
            "orc_dm_entropy": (ctypes.c_int, [ctypes.c_int, P(d), P(d)]),
            "orc_expect_grad": (ctypes.c_int, [ctypes.c_int, P(d), vp, ctypes.c_size_t, ctypes.c_size_t, P(d), P(i32),
                                               P(i32), P(i32), P(d), P(d), ctypes.c_size_t]),
            "orc_mlp_loss_grad": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, P(d),
                                                 P(ctypes.c_float), P(ctypes.c_float), ctypes.c_int, ctypes.c_int,
                                                 P(d), P(d), P(d)]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _i(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))


def _u(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))


def _il(z):
    return np.ascontiguousarray(np.asarray(z, dtype=np.complex128)).view(np.float64).reshape(-1).copy()


def _cx(d):
    return np.ascontiguousarray(d).view(np.complex128).copy()


class OracleError(RuntimeError):
    def __init__(self, status):
        super().__init__(f"oracle status {status}")
        self.status = status


def _ck(st):
    if st != 0:
        raise OracleError(st)


def itercontrol(nbits, positions, configs):
    m = len(positions)
    out = np.zeros(1 << (nbits - m), dtype=np.uint64)
    _ck(lib().orc_itercontrol(nbits, _i(np.array(positions or [0], dtype=np.int32)),
                              _i(np.array(configs or [0], dtype=np.int32)), m, _u(out)))
    return [int(x) for x in out]


def gate_matrix(op):
    out = np.zeros(2 * 16, dtype=np.float64)
    dim = ctypes.c_int()
    _ck(lib().orc_gate_matrix(ctypes.byref(op), _d(out), ctypes.byref(dim)))
    d = dim.value
    return _cx(out[:2 * d * d]).reshape(d, d)


INSTRUCT = {"1q": 0, "1q_diag": 1, "2q": 2, "2q_diag": 3, "nq": 4, "1q_controlled": 5,
            "2q_controlled": 6, "1q_diag_controlled": 7, "2q_diag_controlled": 8, "x": 9,
            "x_controlled": 10, "swap": 11}


def instruct(fn, nbits, state, locs, gate=None, ctrl_locs=(), ctrl_bits=()):
    """Call one instruct_* function of the restatement directly (in place)."""
    s = _il(state)
    g = _il(np.asarray(gate).reshape(-1)) if gate is not None else np.zeros(2)
    cl = np.array(list(ctrl_locs) or [0], dtype=np.int32)
    cb = np.array(list(ctrl_bits) or [0], dtype=np.int32)
    _ck(lib().orc_instruct(INSTRUCT[fn], nbits, _d(s), _i(np.array(locs, dtype=np.int32)), len(locs), _d(g),
                           _i(cl), _i(cb), len(ctrl_locs)))
    return _cx(s)


def set_threads(k):
    """Threads for the oracle's loops over independent amplitude groups (results do
    not depend on it); returns the previous count. The default is 1."""
    return lib().orc_set_threads(int(k))


def apply(nbits, state, oplist, circuit_nbits=None):
    s = _il(state)
    _ck(lib().orc_apply(nbits, nbits if circuit_nbits is None else circuit_nbits, _d(s), oplist.arr, oplist.n))
    return _cx(s)


def validate(nbits, oplist):
    return lib().orc_validate(nbits, oplist.arr, oplist.n)


def expect(nbits, state, poly):
    T, coeffs, tl, sites, ops = poly.flat()
    out = np.zeros(2)
    _ck(lib().orc_expect(nbits, _d(_il(state)), T, _d(coeffs), _i(tl), _i(sites), _i(ops), _d(out)))
    return complex(out[0], out[1])


def probs(nbits, state, locs=None):
    n = (1 << nbits) if locs is None else (1 << len(locs))
    out = np.zeros(n)
    la = None if locs is None else np.array(list(locs) or [0], dtype=np.int32)
    _ck(lib().orc_probs(nbits, _d(_il(state)), None if la is None else _i(la), 0 if locs is None else len(locs),
                        _d(out)))
    return out


def sample_from_probs(p, r):
    p = np.ascontiguousarray(p, dtype=np.float64)
    return int(lib().orc_sample_from_probs(_d(p), len(p), float(r)))


def sample(nbits, state, locs, uniforms):
    u = np.ascontiguousarray(uniforms, dtype=np.float64)
    out = np.zeros(len(u), dtype=np.uint64)
    la = None if locs is None else np.array(list(locs) or [0], dtype=np.int32)
    _ck(lib().orc_sample(nbits, _d(_il(state)), None if la is None else _i(la), 0 if locs is None else len(locs),
                         _d(u), len(u), _u(out)))
    return out


def measure(nbits, state, locs, post, reset_val, r):
    """measure_with_postprocess; returns (outcome bits, new state)."""
    s = _il(state)
    bits = np.zeros(len(locs), dtype=np.int32)
    newn = nbits - len(locs) if post == 2 else nbits
    ns = np.zeros(2 << newn)
    _ck(lib().orc_measure(nbits, _d(s), _i(np.array(locs, dtype=np.int32)), len(locs), post, reset_val, r, _i(bits),
                          _d(ns)))
    return [int(b) for b in bits], (_cx(ns) if post == 2 else _cx(s))


def norm(nbits, state):
    return lib().orc_norm(nbits, _d(_il(state)))


def fidelity(nbits, a, b):
    return lib().orc_fidelity(nbits, _d(_il(a)), _d(_il(b)))


def kraus(oplist):
    out = np.zeros(2 * 256 * 16)
    nops, dim = ctypes.c_int(), ctypes.c_int()
    _ck(lib().orc_kraus(ctypes.byref(oplist.arr[0]), _d(out), 256, ctypes.byref(nops), ctypes.byref(dim)))
    d, k = dim.value, nops.value
    return [_cx(out[2 * d * d * i: 2 * d * d * (i + 1)]).reshape(d, d) for i in range(k)]


def superop(oplist):
    out = np.zeros(2 * 256 * 256)
    d2 = ctypes.c_int()
    _ck(lib().orc_superop(ctypes.byref(oplist.arr[0]), _d(out), ctypes.byref(d2)))
    d = d2.value
    return _cx(out[:2 * d * d]).reshape(d, d)


def dm_from_reg(nbits, state):
    rho = np.zeros(2 << (2 * nbits))
    _ck(lib().orc_dm_from_reg(nbits, _d(_il(state)), _d(rho)))
    return _cx(rho)


def deterministic_state(nbits):
    """deterministic_state (register.rs:71-84)."""
    out = np.zeros(2 << nbits)
    _ck(lib().orc_deterministic_state(nbits, _d(out)))
    return _cx(out)


def rand_state(nbits, uniforms):
    """rand_state (register.rs:35-45) from 2 * 2^nbits f64 uniforms (re, im per amplitude)."""
    u = np.ascontiguousarray(uniforms, dtype=np.float64)
    assert u.size == 2 << nbits
    out = np.zeros(2 << nbits)
    _ck(lib().orc_rand_state(nbits, _d(u), _d(out)))
    return _cx(out)


def dm_mixed(nbits, weights, states):
    """DensityMatrix::mixed (density_matrix.rs:37-61)."""
    w = np.ascontiguousarray(weights, dtype=np.float64)
    st = np.concatenate([_il(s) for s in states])
    rho = np.zeros(2 << (2 * nbits))
    _ck(lib().orc_dm_mixed(nbits, len(states), _d(w), _d(st), _d(rho)))
    return _cx(rho)


def density_matrix_from_reg(nbits, state, locs):
    """density_matrix_from_reg (density_matrix.rs:266-272): from_reg, then partial_tr
    over the bit positions not in locs."""
    traced = [l for l in range(nbits) if l not in locs]
    return dm_partial_tr(nbits, dm_from_reg(nbits, state), traced)


def dm_apply(nbits, rho, oplist):
    r = _il(rho)
    _ck(lib().orc_dm_apply(nbits, nbits, _d(r), oplist.arr, oplist.n))
    return _cx(r)


def dm_trace(nbits, rho):
    out = np.zeros(2)
    _ck(lib().orc_dm_trace(nbits, _d(_il(rho)), _d(out)))
    return complex(out[0], out[1])


def dm_purity(nbits, rho):
    return lib().orc_dm_purity(nbits, _d(_il(rho)))


def dm_expect(nbits, rho, poly):
    T, coeffs, tl, sites, ops = poly.flat()
    out = np.zeros(2)
    _ck(lib().orc_dm_expect(nbits, _d(_il(rho)), T, _d(coeffs), _i(tl), _i(sites), _i(ops), _d(out)))
    return complex(out[0], out[1])


def dm_probs(nbits, rho, locs=None):
    n = (1 << nbits) if locs is None else (1 << len(locs))
    out = np.zeros(n)
    la = None if locs is None else np.array(list(locs) or [0], dtype=np.int32)
    _ck(lib().orc_dm_probs(nbits, _d(_il(rho)), None if la is None else _i(la), 0 if locs is None else len(locs),
                           _d(out)))
    return out


def dm_partial_tr(nbits, rho, traced):
    out = np.zeros(2 << (2 * nbits))
    kept = ctypes.c_int()
    _ck(lib().orc_dm_partial_tr(nbits, _d(_il(rho)), _i(np.array(list(traced) or [0], dtype=np.int32)), len(traced),
                                _d(out), ctypes.byref(kept)))
    k = kept.value
    return _cx(out[:2 << (2 * k)]).reshape(1 << k, 1 << k)


def mlp_num_params(in_dim, width, hidden, trig=False):
    f = 2 * in_dim if trig else in_dim
    return width * (f + 1) + (hidden - 1) * width * (width + 1) + width + 1


def mlp_loss_grad(in_dim, width, hidden, params, x, y, quant=True, trig=False):
    """FP64 forward/backward of the surrogate MLP; returns (predictions, loss, grads).
    quant rounds to bfloat16 where the GPU stores bfloat16 operands; trig feeds
    (cos, sin) pairs of the inputs."""
    params = np.ascontiguousarray(params, dtype=np.float64)
    assert params.size == mlp_num_params(in_dim, width, hidden, trig)
    x = np.ascontiguousarray(x, dtype=np.float32)
    y = np.ascontiguousarray(y, dtype=np.float32)
    n = y.size
    assert x.size == n * in_dim
    out = np.zeros(n)
    loss = np.zeros(1)
    grads = np.zeros(params.size)
    fp = ctypes.POINTER(ctypes.c_float)
    _ck(lib().orc_mlp_loss_grad(in_dim, int(bool(trig)), width, hidden, _d(params), x.ctypes.data_as(fp),
                                y.ctypes.data_as(fp), n, int(bool(quant)), _d(out), _d(loss), _d(grads)))
    return out, float(loss[0]), grads


def expect_grad(nbits, psi0, oplist, poly, nparams):
    """expect_grad (ad.rs:75-156) on the CPU: returns (value, gradient)."""
    T, coeffs, tl, sites, ops = poly.flat()
    grad = np.zeros(max(1, nparams))
    val = np.zeros(1)
    _ck(lib().orc_expect_grad(nbits, _d(_il(psi0)), oplist.arr, oplist.n, T, _d(coeffs), _i(tl), _i(sites),
                              _i(ops), _d(val), _d(grad), nparams))
    return float(val[0]), grad[:nparams]


def dm_entropy(nbits, rho):
    """von_neumann_entropy (density_matrix.rs:113-122) of a 2^n x 2^n matrix."""
    out = np.zeros(1)
    _ck(lib().orc_dm_entropy(nbits, _d(_il(rho)), _d(out)))
    return float(out[0])
