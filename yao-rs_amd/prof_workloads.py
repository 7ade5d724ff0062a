"""Named workloads for rocprofv3 runs (kernel traces and PMC counters).

usage: python yao-rs_amd/prof_workloads.py WORKLOAD [nbits] [depth]
  tiled      random brickwork through yh_apply_fused, LDS tile passes (one step)
  block      the same circuit as 5-qubit MFMA blocks (one step, complex128)
  block64    the same circuit as 5-qubit MFMA blocks (one step, complex64)
  gbg        the same circuit gate by gate (one step)
  copy       one device-to-device copy of a 2^nbits complex128 state (the
             reference kernel for scaling FETCH_SIZE / WRITE_SIZE)
  vqe        one batch of 4096 ansatz instances (n in 10..20, depth 4..8), complex64
  vqe128     the same batch in complex128
  k1q_bits   Rx on every qubit of an nbits complex128 state (4 gates per qubit),
             printing k_1q's average launch time per target bit (HIP events)
  dm         noisy brickwork on a density matrix (fused gate+channel passes)
  dmblk      the same on the superoperator block path
Each workload prints one line naming itself before it touches the GPU.
"""
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path[:0] = [HERE, os.path.join(ROOT, "tests")]

import numpy as np  # noqa: E402


def main():
    w = sys.argv[1]
    n = int(sys.argv[2]) if len(sys.argv) > 2 else 28
    depth = int(sys.argv[3]) if len(sys.argv) > 3 else 20
    print(f"[prof] workload={w} nbits={n} depth={depth}", flush=True)
    import yaohip
    from yaohip.circuit import Circuit
    from fixture_circuits import brickwork, noisy_brickwork
    if w in ("tiled", "gbg", "block", "block64"):
        reg = yaohip.ArrayReg.zero_state(n, "c64" if w == "block64" else "c128")
        circ = Circuit.qubits(n, brickwork(n, depth, 20))
        t0 = time.perf_counter()
        yaohip.apply_inplace(circ, reg, {"tiled": 4, "gbg": 0, "block": 5, "block64": 5}[w])
        print(f"[prof] {w} {time.perf_counter() - t0:.3f} s norm {reg.norm():.15f}", flush=True)
    elif w == "copy":
        a = yaohip.ArrayReg.zero_state(n)
        b = yaohip.ArrayReg.zero_state(n)
        from yaohip import _abi as A
        for _ in range(3):
            A.check(A.lib().yh_reg_copy(b._h, a._h))
        print(f"[prof] copy of {16 << n} bytes x3 done", flush=True)
    elif w in ("vqe", "vqe128"):
        from yaohip.vqe import VqeBatch, random_instances
        inst = random_instances(4096, seed=4)
        b = VqeBatch(4096, 20, "c64" if w == "vqe" else "c128")
        e = b.energies(inst)
        t0 = time.perf_counter()
        e = b.energies(inst)
        print(f"[prof] vqe batch {time.perf_counter() - t0:.4f} s mean E {e.mean():.6f}", flush=True)
    elif w == "k1q_bits":
        import ctypes
        from yaohip import _abi as A
        from yaohip.circuit import Gate, put
        reg = yaohip.ArrayReg.zero_state(n)
        L = A.lib()
        for loc in range(n):
            circ = Circuit.qubits(n, [put([loc], Gate.Rx(0.3))] * 4)
            yaohip.apply_inplace(circ, reg)
            A.check(L.yh_reg_timing_enable(reg._h, 1))
            yaohip.apply_inplace(circ, reg)
            c, ms = ctypes.c_uint64(), ctypes.c_double()
            A.check(L.yh_reg_timing_read(reg._h, A.K_1Q, ctypes.byref(c), ctypes.byref(ms)))
            A.check(L.yh_reg_timing_enable(reg._h, 0))
            us = 1e3 * ms.value / c.value
            print(f"[prof] k_1q target bit {n - 1 - loc:2d}: {us:8.1f} us, "
                  f"{2 * 16 * (1 << n) / (us * 1e-6) / 1e12:.2f} TB/s", flush=True)
    elif w in ("dm", "dmblk"):
        dm = yaohip.DensityMatrix.zero_state(n)
        dm.apply(Circuit.qubits(n, noisy_brickwork(n, depth, 3)), fused=True if w == "dm" else "blocks")
        print(f"[prof] {w} trace {dm.trace().real:.15f}", flush=True)
    else:
        raise SystemExit(f"unknown workload {w}")


if __name__ == "__main__":
    main()
