"""Device-resident registers and the simulation entry points.

Mirrors yao-rs's public functions over the C ABI: `ArrayReg` (src/register.rs),
`apply` / `apply_inplace` (src/apply.rs:123-154), `expect_arrayreg` / `expect_dm`
(src/expect.rs:7,38), `probs` / `measure_with_postprocess` (src/measure.rs:82,243),
`DensityMatrix` (src/density_matrix.rs). Every amplitude-touching call runs on the
GPU through libyaohip.so; there is no host fallback.
"""
import ctypes
import os

import numpy as np

from . import _abi as A
from .circuit import Circuit, OperatorPolynomial


def _lib():
    A.init()
    return A.lib()


class ArrayReg:
    """ArrayReg on the device (register.rs:15-18). dtype: "c128" or "c64"."""

    def __init__(self, handle, dtype):
        self._h = handle
        self.dtype = dtype

    @staticmethod
    def _new(nbits, dtype="c128"):
        h = ctypes.c_void_p()
        A.check(_lib().yh_reg_create(nbits, A.C128 if dtype == "c128" else A.C64, ctypes.byref(h)))
        return ArrayReg(h, dtype)

    @staticmethod
    def zero_state(nbits, dtype="c128"):
        """zero_state (register.rs:21-26)."""
        return ArrayReg._new(nbits, dtype)

    @staticmethod
    def from_vec(nbits, state, dtype="c128"):
        """from_vec (register.rs:65-69)."""
        r = ArrayReg._new(nbits, dtype)
        buf = A.interleave(state)
        if buf.size != 2 << nbits:
            raise ValueError("state length must be 2^nbits")
        A.check(A.lib().yh_reg_upload(r._h, A.dptr(buf), 1 << nbits))
        return r

    @staticmethod
    def _made(fn, dtype, *args):
        h = ctypes.c_void_p()
        A.check(fn(*args[:1], A.C128 if dtype == "c128" else A.C64, *args[1:], ctypes.byref(h)))
        return ArrayReg(h, dtype)

    @staticmethod
    def uniform_state(nbits, dtype="c128"):
        """uniform_state (register.rs:47-54)."""
        return ArrayReg._made(_lib().yh_reg_uniform_state, dtype, nbits)

    @staticmethod
    def ghz_state(nbits, dtype="c128"):
        """ghz_state (register.rs:56-63)."""
        return ArrayReg._made(_lib().yh_reg_ghz_state, dtype, nbits)

    @staticmethod
    def product_state(nbits, bits, dtype="c128"):
        """product_state (register.rs:28-33): BitStr::onehot of `bits`, an integer
        basis index or a bit string written most significant bit first."""
        index = int(bits, 2) if isinstance(bits, str) else int(bits)
        return ArrayReg._made(_lib().yh_reg_product_state, dtype, nbits, ctypes.c_uint64(index))

    @staticmethod
    def deterministic_state(nbits, dtype="c128"):
        """deterministic_state (register.rs:71-84)."""
        return ArrayReg._made(_lib().yh_reg_deterministic_state, dtype, nbits)

    @staticmethod
    def rand_state(nbits, rng, dtype="c128"):
        """rand_state (register.rs:35-45). `rng` is a callable returning an array of k
        f64 uniforms in [0, 1) for rng(k) (the caller owns the stream, as `&mut impl
        Rng`), or a numpy Generator; 2 uniforms per amplitude, (re, im) in order."""
        k = 2 << nbits
        u = rng.random(k) if hasattr(rng, "random") else rng(k)
        u = np.ascontiguousarray(u, dtype=np.float64)
        return ArrayReg._made(_lib().yh_reg_rand_state, dtype, nbits, A.dptr(u), ctypes.c_size_t(k))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and A._lib is not None:
            A._lib.yh_reg_destroy(h)
            self._h = None

    def nqubits(self):
        n = ctypes.c_int()
        A.check(A.lib().yh_reg_nbits(self._h, ctypes.byref(n)))
        return n.value

    nbits = property(nqubits)

    def state_vec(self):
        """state_vec (register.rs:91-93), copied to the host as complex128."""
        n = self.nqubits()
        buf = np.zeros(2 << n)
        A.check(A.lib().yh_reg_download(self._h, A.dptr(buf), 1 << n))
        return A.deinterleave(buf)

    def clone(self):
        h = ctypes.c_void_p()
        A.check(A.lib().yh_reg_clone(self._h, ctypes.byref(h)))
        return ArrayReg(h, self.dtype)

    def norm(self):
        v = ctypes.c_double()
        A.check(A.lib().yh_reg_norm(self._h, ctypes.byref(v)))
        return v.value

    def normalize(self):
        A.check(A.lib().yh_reg_normalize(self._h))

    def fidelity(self, other):
        v = ctypes.c_double()
        A.check(A.lib().yh_reg_fidelity(self._h, other._h, ctypes.byref(v)))
        return v.value

    def apply(self, circuit):
        """Register::apply (register.rs:127-139)."""
        apply_inplace(circuit, self)


def write_state(reg, path):
    """write_state (yao-cli/src/state_io.rs:18-47): yao-state-v1 file."""
    A.check(A.lib().yh_state_write(reg._h, os.fsencode(path)))


def read_state(path, dtype="c128"):
    """read_state_from_file (yao-cli/src/state_io.rs:49-102) into a device register."""
    A.init()
    h = ctypes.c_void_p()
    A.check(A.lib().yh_state_read(os.fsencode(path), A.C128 if dtype == "c128" else A.C64, ctypes.byref(h)))
    return ArrayReg(h, dtype)


def read_state_header(path):
    """Host-only header check of a yao-state-v1 file; returns num_qubits."""
    n = ctypes.c_int()
    A.check(A.lib().yh_state_read_header(os.fsencode(path), ctypes.byref(n)))
    return n.value


def apply_inplace(circuit, reg, fused=None):
    """apply_inplace (apply.rs:123-147). fused=None (default) or 5: gates fused into
    MFMA blocks of up to 5 qubits (yh_apply_fused, the same default as the C ABI's
    Rust binding); 1..4: LDS tile passes; 0: one launch per gate (yh_apply)."""
    ops = circuit.ops
    if fused is None:
        fused = 5
    if fused:
        A.check(A.lib().yh_apply_fused(reg._h, circuit.nbits, ops.arr, ops.n, int(fused)))
    else:
        A.check(A.lib().yh_apply(reg._h, circuit.nbits, ops.arr, ops.n))


def apply(circuit, reg, fused=None):
    """apply (apply.rs:150-154): returns a new register."""
    out = reg.clone()
    apply_inplace(circuit, out, fused)
    return out


def expect_arrayreg(reg, op: OperatorPolynomial):
    """expect_arrayreg (expect.rs:7-11)."""
    T, coeffs, tl, sites, ops = op.flat()
    out = np.zeros(2)
    A.check(A.lib().yh_expect(reg._h, T, A.dptr(coeffs), A.i32ptr(tl), A.i32ptr(sites), A.i32ptr(ops),
                              A.dptr(out)))
    return complex(out[0], out[1])


def expect_grad(observable: OperatorPolynomial, circuit, psi0):
    """expect_grad (ad.rs:75-156): (Re <psi|H|psi>, gradient over circuit.parameters())
    with psi = circuit |psi0>; psi0 is left unchanged."""
    T, coeffs, tl, sites, ops = observable.flat()
    ol = circuit.ops
    n = ctypes.c_size_t()
    A.check(A.lib().yh_circuit_num_params(ol.arr, ol.n, ctypes.byref(n)))
    grad = np.zeros(max(1, n.value))
    val = ctypes.c_double()
    A.check(A.lib().yh_expect_grad(psi0._h, circuit.nbits, ol.arr, ol.n, T, A.dptr(coeffs), A.i32ptr(tl),
                                   A.i32ptr(sites), A.i32ptr(ops), ctypes.byref(val), A.dptr(grad), n.value))
    return val.value, grad[:n.value]


def probs(state, locs=None):
    """probs (measure.rs:82-87) for an ArrayReg or DensityMatrix."""
    n = state.nqubits()
    m = (1 << n) if locs is None else (1 << len(locs))
    out = np.zeros(m)
    la = None if locs is None else np.array(list(locs) or [0], dtype=np.int32)
    lp = None if la is None else A.i32ptr(la)
    nl = 0 if locs is None else len(locs)
    if isinstance(state, DensityMatrix):
        A.check(A.lib().yh_dm_probs(state._h, lp, nl, A.dptr(out)))
    else:
        A.check(A.lib().yh_probs(state._h, lp, nl, A.dptr(out)))
    return out


def sample(reg, locs, uniforms):
    """One outcome index per uniform (sample_from_probs, measure.rs:140-150)."""
    u = np.ascontiguousarray(uniforms, dtype=np.float64)
    out = np.zeros(len(u), dtype=np.uint64)
    la = None if locs is None else np.array(list(locs) or [0], dtype=np.int32)
    A.check(A.lib().yh_sample(reg._h, None if la is None else A.i32ptr(la), 0 if locs is None else len(locs),
                              A.dptr(u), len(u), A.u64ptr(out)))
    return out


class PostProcess:
    """PostProcess (measure.rs:18-23)."""
    NoPostProcess = ("none", 0)

    @staticmethod
    def ResetTo(v):
        return ("reset", int(v))

    RemoveMeasured = ("remove", 0)


def measure_with_postprocess(reg, locs, post, rng):
    """measure_with_postprocess (measure.rs:243-274). `rng` is a callable returning
    the next f64 uniform in [0, 1) (the caller owns the stream, as `&mut impl Rng`).
    Returns the outcome bits; with RemoveMeasured returns (bits, new register)."""
    kind, val = post
    code = {"none": 0, "reset": 1, "remove": 2}[kind]
    bits = np.zeros(len(locs), dtype=np.int32)
    target = reg.clone() if code == 2 else reg
    A.check(A.lib().yh_measure(target._h, A.i32ptr(np.array(locs, dtype=np.int32)), len(locs), code, val,
                               float(rng()), A.i32ptr(bits)))
    out = [int(b) for b in bits]
    return (out, target) if code == 2 else out


class DensityMatrix:
    """DensityMatrix on the device (density_matrix.rs:9-12), complex128."""

    def __init__(self, handle):
        self._h = handle

    @staticmethod
    def zero_state(nbits):
        h = ctypes.c_void_p()
        A.check(_lib().yh_dm_create(nbits, ctypes.byref(h)))
        return DensityMatrix(h)

    @staticmethod
    def from_reg(reg):
        h = ctypes.c_void_p()
        A.check(A.lib().yh_dm_from_reg(reg._h, ctypes.byref(h)))
        return DensityMatrix(h)

    @staticmethod
    def mixed(weights, regs):
        """mixed (density_matrix.rs:37-61): sum_j w_j |psi_j><psi_j|."""
        w = np.ascontiguousarray(weights, dtype=np.float64)
        if len(w) != len(regs):
            raise ValueError("weights and registers differ in length")
        hs = (ctypes.c_void_p * max(1, len(regs)))(*[r._h for r in regs])
        h = ctypes.c_void_p()
        A.check(A.lib().yh_dm_mixed(A.dptr(w), hs, len(regs), ctypes.byref(h)))
        return DensityMatrix(h)

    @staticmethod
    def from_matrix(nbits, rho):
        d = DensityMatrix.zero_state(nbits)
        buf = A.interleave(np.asarray(rho).reshape(-1))
        A.check(A.lib().yh_dm_upload(d._h, A.dptr(buf), 1 << (2 * nbits)))
        return d

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and A._lib is not None:
            A._lib.yh_dm_destroy(h)
            self._h = None

    def nbits(self):
        n = ctypes.c_int()
        A.check(A.lib().yh_dm_nbits(self._h, ctypes.byref(n)))
        return n.value

    nqubits = nbits

    def state_data(self):
        n = self.nbits()
        buf = np.zeros(2 << (2 * n))
        A.check(A.lib().yh_dm_download(self._h, A.dptr(buf), 1 << (2 * n)))
        return A.deinterleave(buf)

    def apply(self, circuit: Circuit, fused=False):
        """Register::apply for DensityMatrix (density_matrix.rs:244-259). fused=True:
        each gate with its following 1-qubit channels in one pass; fused="blocks":
        gates and channels as superoperators fused into 5-bit MFMA blocks."""
        ops = circuit.ops
        if fused == "blocks":
            fn = A.lib().yh_dm_apply_blocks
        else:
            fn = A.lib().yh_dm_apply_fused if fused else A.lib().yh_dm_apply
        A.check(fn(self._h, circuit.nbits, ops.arr, ops.n))

    def trace(self):
        out = np.zeros(2)
        A.check(A.lib().yh_dm_trace(self._h, A.dptr(out)))
        return complex(out[0], out[1])

    def purity(self):
        v = ctypes.c_double()
        A.check(A.lib().yh_dm_purity(self._h, ctypes.byref(v)))
        return v.value

    def partial_tr(self, traced_locs):
        """partial_tr (density_matrix.rs:80-111); traced locs are bit positions
        (LSB-first), as in the reference."""
        t = np.array(list(traced_locs) or [0], dtype=np.int32)
        h = ctypes.c_void_p()
        A.check(A.lib().yh_dm_partial_tr(self._h, A.i32ptr(t), len(traced_locs), ctypes.byref(h)))
        return DensityMatrix(h)

    def von_neumann_entropy(self):
        """von_neumann_entropy (density_matrix.rs:113-122)."""
        v = ctypes.c_double()
        A.check(A.lib().yh_dm_entropy(self._h, ctypes.byref(v)))
        return v.value


def density_matrix_from_reg(reg, locs):
    """density_matrix_from_reg (density_matrix.rs:266-272): the reduced density matrix
    on `locs` (bit positions, as partial_tr reads them), without the full matrix."""
    la = np.array(list(locs) or [0], dtype=np.int32)
    h = ctypes.c_void_p()
    A.check(A.lib().yh_dm_from_reg_locs(reg._h, A.i32ptr(la), len(locs), ctypes.byref(h)))
    return DensityMatrix(h)


def expect_dm(dm, op: OperatorPolynomial):
    """expect_dm (expect.rs:38-42); sites are LSB-first bit positions."""
    T, coeffs, tl, sites, ops = op.flat()
    out = np.zeros(2)
    A.check(A.lib().yh_dm_expect(dm._h, T, A.dptr(coeffs), A.i32ptr(tl), A.i32ptr(sites), A.i32ptr(ops),
                                 A.dptr(out)))
    return complex(out[0], out[1])


def channel_kraus(noise, locs=None):
    """NoiseChannel::kraus_operators (noise.rs:183) computed by the backend's host side."""
    from .circuit import channel, lower
    n = noise.num_qubits()
    ol = lower([channel(list(range(n)) if locs is None else locs, noise)])
    out = np.zeros(2 * 256 * 16)
    nops, dim = ctypes.c_int(), ctypes.c_int()
    A.check(A.lib().yh_channel_kraus(ctypes.byref(ol.arr[0]), A.dptr(out), 256, ctypes.byref(nops),
                                     ctypes.byref(dim)))
    d = dim.value
    return [A.deinterleave(out[2 * d * d * i: 2 * d * d * (i + 1)]).reshape(d, d) for i in range(nops.value)]


def gate_matrix(gate):
    from .circuit import put, lower
    ol = lower([put(list(range(gate.num_sites())), gate)])
    out = np.zeros(2 * 16 * 16 * 4)
    dim = ctypes.c_int()
    A.check(A.lib().yh_gate_matrix(ctypes.byref(ol.arr[0]), A.dptr(out), ctypes.byref(dim)))
    d = dim.value
    return A.deinterleave(out[:2 * d * d]).reshape(d, d)
