"""Batched VQE energies over the C ABI (yh_vqe_batch_*), plus the workload
generator: hardware-efficient ansatz instances (per layer a single-qubit block on
every qubit, then a CZ chain CZ(i, i+1)) on the open transverse-field Ising chain
H = -J sum Z_i Z_{i+1} - h sum X_i. The block is Ry then Rz (2 angles per qubit)
or Rz, Ry, Rz (general_u2, 3 angles per qubit)."""
import ctypes
import math

import numpy as np

from . import _abi as A
from .circuit import Gate, control, put, tfim_hamiltonian

BLOCKS = {2: ("Ry", "Rz"), 3: ("Rz", "Ry", "Rz")}


class VqeInstances:
    """Instance table: qubit counts, depths, angle offsets and the flat angle array,
    ordered per instance as (layer, loc, block gate)."""

    def __init__(self, nq, depth, angles_per_instance, apq=2):
        self.apq = int(apq)
        self.nq = np.ascontiguousarray(nq, dtype=np.int32)
        self.depth = np.ascontiguousarray(depth, dtype=np.int32)
        offs, flat = [], []
        off = 0
        for a in angles_per_instance:
            offs.append(off)
            flat.append(np.asarray(a, dtype=np.float64).reshape(-1))
            off += flat[-1].size
        self.offset = np.ascontiguousarray(offs, dtype=np.int64)
        self.angles = np.ascontiguousarray(np.concatenate(flat) if flat else np.zeros(1))
        self.count = len(self.nq)

    def instance_angles(self, i):
        n, d = int(self.nq[i]), int(self.depth[i])
        return self.angles[self.offset[i]: self.offset[i] + self.apq * n * d].reshape(d, n, self.apq)

    def subset(self, idx):
        return VqeInstances(self.nq[idx], self.depth[idx], [self.instance_angles(i) for i in idx], self.apq)

    def circuit_elements(self, i):
        """The instance as a yao-rs element list (CZ = control([i],[i+1], Z))."""
        n, d = int(self.nq[i]), int(self.depth[i])
        a = self.instance_angles(i)
        mk = {"Rx": Gate.Rx, "Ry": Gate.Ry, "Rz": Gate.Rz}
        els = []
        for l in range(d):
            for q in range(n):
                for g, name in enumerate(BLOCKS[self.apq]):
                    els.append(put([q], mk[name](a[l, q, g])))
            for q in range(n - 1):
                els.append(control([q], [q + 1], Gate.Z))
        return els


def random_instances(count, seed, nmin=10, nmax=20, dmin=4, dmax=8, amin=-math.pi, amax=math.pi, apq=2):
    """Seeded instances: n ~ U{nmin..nmax}, depth ~ U{dmin..dmax}, angles ~ U[amin, amax]."""
    rng = np.random.default_rng(seed)
    nq = rng.integers(nmin, nmax + 1, size=count)
    depth = rng.integers(dmin, dmax + 1, size=count)
    angles = [rng.uniform(amin, amax, size=apq * int(n) * int(d)) for n, d in zip(nq, depth)]
    return VqeInstances(nq, depth, angles, apq)


def fixed_instances(angles, n, depth, apq):
    """Instances that share n and depth; angles is (count, depth*n*apq)."""
    angles = np.asarray(angles, dtype=np.float64)
    c = angles.shape[0]
    inst = VqeInstances.__new__(VqeInstances)
    inst.apq = apq
    inst.nq = np.full(c, n, dtype=np.int32)
    inst.depth = np.full(c, depth, dtype=np.int32)
    inst.offset = np.ascontiguousarray(np.arange(c, dtype=np.int64) * angles.shape[1])
    inst.angles = np.ascontiguousarray(angles.reshape(-1))
    inst.count = c
    return inst


class VqeBatch:
    """yh_vqe_batch handle; dtype "c64" (complex64 amplitudes, FP64 reductions) or "c128"."""

    def __init__(self, max_instances, max_qubits, dtype="c64"):
        A.init()
        self._h = ctypes.c_void_p()
        A.check(A.lib().yh_vqe_batch_create(max_instances, max_qubits, A.C64 if dtype == "c64" else A.C128,
                                            ctypes.byref(self._h)))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and A._lib is not None:
            A._lib.yh_vqe_batch_destroy(h)
            self._h = None

    def configure(self, tile_bits=0, blocks_per_cu=2, resident_grid=True, step_threads=256):
        """Fix the launch configuration (tile_bits 11, 12 or 13); tile_bits=0 autotunes."""
        A.check(A.lib().yh_vqe_batch_configure(self._h, tile_bits, blocks_per_cu, int(bool(resident_grid)),
                                               step_threads))
        return self

    def config(self):
        """Launch configuration of the last energies() call."""
        out = np.zeros(4, dtype=np.int32)
        A.check(A.lib().yh_vqe_batch_config(self._h, A.i32ptr(out)))
        return {"tile_bits": int(out[0]), "blocks_per_cu": int(out[1]), "resident_grid": bool(out[2]),
                "step_threads": int(out[3])}

    def energies(self, inst: VqeInstances, J=1.0, h=1.0):
        out = np.zeros(inst.count)
        A.check(A.lib().yh_vqe_batch_energies(self._h, inst.count, A.i32ptr(inst.nq), A.i32ptr(inst.depth),
                                              A.i64ptr(inst.offset), A.dptr(inst.angles), inst.apq, J, h,
                                              A.dptr(out)))
        return out


def plain_energies(inst: VqeInstances, J=1.0, h=1.0, dtype="c128"):
    """Per-circuit GPU path: yh_apply (one launch per gate) then yh_expect per instance."""
    A.init()
    out = np.zeros(inst.count)
    A.check(A.lib().yh_vqe_plain_energies(inst.count, A.i32ptr(inst.nq), A.i32ptr(inst.depth), A.i64ptr(inst.offset),
                                          A.dptr(inst.angles), inst.apq, J, h, A.C64 if dtype == "c64" else A.C128,
                                          A.dptr(out)))
    return out


def tfim(n, J=1.0, h=1.0):
    return tfim_hamiltonian(n, J, h)
