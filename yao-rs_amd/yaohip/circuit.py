"""Circuit data model mirroring yao-rs's public types.

Names and argument meaning follow the reference: `Gate` (src/gate.rs:7-35),
`NoiseChannel` (src/noise.rs:16-60), `PositionedGate` / `put` / `control` /
`channel` / `Circuit` (src/circuit.rs:103-465), `Op` / `OperatorString` /
`OperatorPolynomial` (src/operator.rs:8-157). A circuit lowers to a packed array of
`struct yh_op` for the C ABI. Validation errors are raised as `CircuitError` with
the reference's variant name.
"""
import ctypes

import numpy as np

from . import _abi as A


class Gate:
    """Gate enum (gate.rs:7-35). Named gates carry kind + parameters; Custom carries
    its matrix (rows x rows complex) and is_diagonal flag."""

    __slots__ = ("kind", "params", "matrix", "is_diagonal", "label")

    def __init__(self, kind, params=(), matrix=None, is_diagonal=False, label=""):
        self.kind = kind
        self.params = tuple(float(p) for p in params)
        self.matrix = None if matrix is None else np.ascontiguousarray(matrix, dtype=np.complex128)
        self.is_diagonal = bool(is_diagonal)
        self.label = label

    def __repr__(self):
        name = A.GATE_NAMES[self.kind]
        return f"{name}{self.params if self.params else ''}"

    # named constructors, as the Rust variants
    @staticmethod
    def Phase(theta): return Gate(A.GATE_PHASE, (theta,))

    @staticmethod
    def Rx(theta): return Gate(A.GATE_RX, (theta,))

    @staticmethod
    def Ry(theta): return Gate(A.GATE_RY, (theta,))

    @staticmethod
    def Rz(theta): return Gate(A.GATE_RZ, (theta,))

    @staticmethod
    def FSim(theta, phi): return Gate(A.GATE_FSIM, (theta, phi))

    @staticmethod
    def Custom(matrix, is_diagonal=False, label="custom"):
        return Gate(A.GATE_CUSTOM, (), matrix, is_diagonal, label)

    def num_sites(self):
        """Gate::num_sites (gate.rs:71-95)."""
        if self.kind in (A.GATE_SWAP, A.GATE_ISWAP, A.GATE_FSIM):
            return 2
        if self.kind == A.GATE_CUSTOM:
            return int(self.matrix.shape[0]).bit_length() - 1
        return 1

    def is_diag(self):
        """Gate::is_diagonal (gate.rs:98-104)."""
        if self.kind in (A.GATE_Z, A.GATE_S, A.GATE_T, A.GATE_PHASE, A.GATE_RZ):
            return True
        return self.kind == A.GATE_CUSTOM and self.is_diagonal

    def num_params(self):
        """Gate::num_params (gate.rs:107-113)."""
        if self.kind in (A.GATE_RX, A.GATE_RY, A.GATE_RZ, A.GATE_PHASE):
            return 1
        return 2 if self.kind == A.GATE_FSIM else 0


for _k, _n in enumerate(A.GATE_NAMES):
    if _n in ("X", "Y", "Z", "H", "S", "T", "SWAP", "SqrtX", "SqrtY", "SqrtW", "ISWAP"):
        setattr(Gate, _n, Gate(_k))


class NoiseChannel:
    """NoiseChannel enum (noise.rs:16-60)."""

    __slots__ = ("kind", "params", "matrices", "n")

    def __init__(self, kind, params=(), matrices=None, n=1):
        self.kind = kind
        self.params = tuple(float(p) for p in params)
        self.matrices = matrices
        self.n = n

    @staticmethod
    def BitFlip(p): return NoiseChannel(A.CH_BITFLIP, (p,))

    @staticmethod
    def PhaseFlip(p): return NoiseChannel(A.CH_PHASEFLIP, (p,))

    @staticmethod
    def Depolarizing(n, p): return NoiseChannel(A.CH_DEPOLARIZING, (p, n), n=n)

    @staticmethod
    def PauliChannel(px, py, pz): return NoiseChannel(A.CH_PAULI, (px, py, pz))

    @staticmethod
    def Reset(p0, p1): return NoiseChannel(A.CH_RESET, (p0, p1))

    @staticmethod
    def AmplitudeDamping(gamma, excited_population=0.0):
        return NoiseChannel(A.CH_AMPLITUDE_DAMPING, (gamma, excited_population))

    @staticmethod
    def PhaseDamping(gamma): return NoiseChannel(A.CH_PHASE_DAMPING, (gamma,))

    @staticmethod
    def PhaseAmplitudeDamping(amplitude, phase, excited_population=0.0):
        return NoiseChannel(A.CH_PHASE_AMPLITUDE_DAMPING, (amplitude, phase, excited_population))

    @staticmethod
    def ThermalRelaxation(t1, t2, time, excited_population=0.0):
        return NoiseChannel(A.CH_THERMAL_RELAXATION, (t1, t2, time, excited_population))

    @staticmethod
    def Coherent(matrix):
        m = np.ascontiguousarray(matrix, dtype=np.complex128)
        return NoiseChannel(A.CH_COHERENT, (), [m], n=int(m.shape[0]).bit_length() - 1)

    @staticmethod
    def Custom(kraus_ops):
        ms = [np.ascontiguousarray(k, dtype=np.complex128) for k in kraus_ops]
        return NoiseChannel(A.CH_CUSTOM, (), ms, n=int(ms[0].shape[0]).bit_length() - 1)

    def num_qubits(self):
        """NoiseChannel::num_qubits (noise.rs:64-86)."""
        return self.n


class PositionedGate:
    """PositionedGate (circuit.rs:103-122)."""

    __slots__ = ("gate", "target_locs", "control_locs", "control_configs")

    def __init__(self, gate, target_locs, control_locs=(), control_configs=()):
        self.gate = gate
        self.target_locs = list(target_locs)
        self.control_locs = list(control_locs)
        self.control_configs = [bool(c) for c in control_configs]


class PositionedChannel:
    """PositionedChannel (circuit.rs:25-29)."""

    __slots__ = ("channel", "locs")

    def __init__(self, channel_, locs):
        self.channel = channel_
        self.locs = list(locs)


class Annotation:
    """PositionedAnnotation with a Label (circuit.rs:11-22)."""

    __slots__ = ("label", "loc")

    def __init__(self, label, loc):
        self.label = label
        self.loc = loc


def put(target_locs, gate):
    """put (circuit.rs:435)."""
    return PositionedGate(gate, target_locs)


def control(ctrl_locs, target_locs, gate):
    """control (circuit.rs:454): active-high controls."""
    return PositionedGate(gate, target_locs, ctrl_locs, [True] * len(ctrl_locs))


def channel(locs, noise):
    """channel (circuit.rs:460)."""
    return PositionedChannel(noise, locs)


def label(loc, text):
    return Annotation(text, loc)


def _mat_buffer(m):
    buf = A.interleave(np.asarray(m).reshape(-1))
    return buf


def element_to_op(e, keep):
    """Lower one circuit element to a struct yh_op; matrix buffers go to `keep`."""
    op = A.YhOp()
    if isinstance(e, PositionedGate):
        op.elem = A.ELEM_GATE
        op.kind = e.gate.kind
        if len(e.target_locs) > A.MAX_TARGETS or len(e.control_locs) > A.MAX_CTRLS:
            raise CircuitError("Unsupported", "too many locations for the C ABI")
        op.nt = len(e.target_locs)
        op.nc = len(e.control_locs)
        for k, l in enumerate(e.target_locs):
            op.targets[k] = int(l)
        for k, l in enumerate(e.control_locs):
            op.ctrls[k] = int(l)
        op.nconf = len(e.control_configs)
        for k in range(min(op.nc, op.nconf)):
            op.ctrl_vals[k] = int(e.control_configs[k])
        for k, p in enumerate(e.gate.params):
            op.params[k] = p
        if e.gate.kind == A.GATE_CUSTOM:
            buf = _mat_buffer(e.gate.matrix)
            keep.append(buf)
            op.matrix = A.dptr(buf)
            op.dim = e.gate.matrix.shape[0]
            op.is_diag = int(e.gate.is_diagonal)
    elif isinstance(e, PositionedChannel):
        op.elem = A.ELEM_CHANNEL
        op.kind = e.channel.kind
        op.nt = len(e.locs)
        for k, l in enumerate(e.locs):
            op.targets[k] = int(l)
        for k, p in enumerate(e.channel.params):
            op.params[k] = p
        if e.channel.matrices is not None:
            buf = np.concatenate([_mat_buffer(m) for m in e.channel.matrices])
            keep.append(buf)
            op.matrix = A.dptr(buf)
            op.nmat = len(e.channel.matrices)
            op.dim = e.channel.matrices[0].shape[0]
    elif isinstance(e, Annotation):
        op.elem = A.ELEM_ANNOTATION
        op.nt = 1
        op.targets[0] = int(e.loc)
    else:
        raise TypeError(f"not a circuit element: {e!r}")
    return op


def lower(elements):
    keep = []
    ops = [element_to_op(e, keep) for e in elements]
    return A.OpList(ops, keep)


class CircuitError(ValueError):
    """CircuitError (circuit.rs:41-57); `variant` is the Rust variant name."""

    STATUS_TO_VARIANT = {
        2: "ControlConfigLengthMismatch", 3: "LocOutOfRange",
        4: "OverlappingLocs", 5: "MatrixSizeMismatch",
    }

    def __init__(self, variant, message=""):
        super().__init__(f"{variant}: {message}")
        self.variant = variant


class Circuit:
    """Circuit (circuit.rs:139-272); qubit-only, dims must all be 2."""

    def __init__(self, dims, elements):
        self.dims = list(dims)
        self.nbits = len(self.dims)
        self.elements = list(elements)
        self._lowered = lower(self.elements)
        if any(d != 2 for d in self.dims):
            raise CircuitError("Unsupported", "only qubit (d=2) sites are simulated")
        st = A.lib().yh_validate(self.nbits, self._lowered.arr, self._lowered.n)
        if st != A.YH_OK:
            raise CircuitError(CircuitError.STATUS_TO_VARIANT.get(st, A.STATUS_NAMES.get(st)),
                               A.lib().yh_last_error().decode())

    @staticmethod
    def qubits(nbits, elements):
        return Circuit([2] * nbits, elements)

    @property
    def ops(self):
        return self._lowered

    def num_sites(self):
        return self.nbits

    def num_params(self):
        """Circuit::num_params (circuit.rs:311-319)."""
        return sum(e.gate.num_params() for e in self.elements if isinstance(e, PositionedGate))

    def parameters(self):
        """Circuit::parameters (circuit.rs:322-330)."""
        out = []
        for e in self.elements:
            if isinstance(e, PositionedGate):
                out.extend(e.gate.params)
        return out

    def dispatch(self, params):
        """Circuit::dispatch (circuit.rs:332-350): overwrite parameters in order."""
        params = list(params)
        if len(params) != self.num_params():
            raise ValueError("dispatch length does not match num_params")
        off = 0
        for e in self.elements:
            if isinstance(e, PositionedGate):
                k = e.gate.num_params()
                if k:
                    e.gate = Gate(e.gate.kind, params[off:off + k])
                    off += k
        self._lowered = lower(self.elements)


class Op:
    """Op enum (operator.rs:8-17)."""
    I, X, Y, Z, P0, P1, Pu, Pd = range(8)


class OperatorString:
    """OperatorString::new (operator.rs:27-33): sorted by site, identities dropped."""

    def __init__(self, ops):
        ops = sorted(((int(s), int(o)) for s, o in ops), key=lambda x: x[0])
        self.ops = [(s, o) for s, o in ops if o != Op.I]

    @staticmethod
    def identity():
        return OperatorString([])


class OperatorPolynomial:
    """OperatorPolynomial (operator.rs:69-157)."""

    def __init__(self, coeffs, opstrings):
        if len(coeffs) != len(opstrings):
            raise ValueError("coeffs and opstrings differ in length")
        self.coeffs = [complex(c) for c in coeffs]
        self.opstrings = list(opstrings)

    @staticmethod
    def single(site, op, coeff):
        return OperatorPolynomial([coeff], [OperatorString([(site, op)])])

    def __add__(self, other):
        return OperatorPolynomial(self.coeffs + other.coeffs, self.opstrings + other.opstrings)

    def __neg__(self):
        return OperatorPolynomial([-c for c in self.coeffs], self.opstrings)

    def __mul__(self, scalar):
        return OperatorPolynomial([c * scalar for c in self.coeffs], self.opstrings)

    def flat(self):
        """Flattened arrays for yh_expect: coeffs (2*T), term_len (T), sites, ops."""
        T = len(self.coeffs)
        coeffs = np.zeros(2 * max(T, 1), dtype=np.float64)
        for k, c in enumerate(self.coeffs):
            coeffs[2 * k], coeffs[2 * k + 1] = c.real, c.imag
        term_len = np.array([len(s.ops) for s in self.opstrings] or [0], dtype=np.int32)
        sites = np.array([s for st in self.opstrings for s, _ in st.ops] or [0], dtype=np.int32)
        ops = np.array([o for st in self.opstrings for _, o in st.ops] or [0], dtype=np.int32)
        return T, coeffs, term_len, sites, ops


def tfim_hamiltonian(n, J=1.0, h=1.0):
    """H = -J sum_i Z_i Z_{i+1} - h sum_i X_i (open boundary)."""
    coeffs, strings = [], []
    for i in range(n - 1):
        coeffs.append(-J)
        strings.append(OperatorString([(i, Op.Z), (i + 1, Op.Z)]))
    for i in range(n):
        coeffs.append(-h)
        strings.append(OperatorString([(i, Op.X)]))
    return OperatorPolynomial(coeffs, strings)
