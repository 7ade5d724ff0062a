"""Benchmark: 28-qubit random brickwork circuit (depth 20) on a complex128 state.

A step applies the whole circuit (1,950 gates: Rx, Ry, Rz on every qubit per layer,
then CNOTs on even / odd pairs) to the 4 GiB state resident in HBM, as 5-qubit
fused blocks applied with FP64 MFMA (yh_apply_fused, max_fuse_qubits = 5). Prints
one JSON line with whole-job gate applications per second, the dominant kernel's
roofline (live HIP-event timing on the library's stream) and the CPU oracle timed
on a bounded sample of the same workload on this host. The same circuit is also
timed gate by gate and through LDS tile passes, and the other workloads (complex64,
batched VQE, noisy density matrix, learned surrogate) are reported beside it.

usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--skip-extra]
       python bench.py --profile WORKLOAD [--nbits N] [--depth D]
The --profile form runs one named workload once (for rocprofv3 kernel traces and
counter runs) and prints progress lines instead of the JSON line; see PROFILE_WORKLOADS.
"""
import argparse
import ctypes
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "yao-rs_amd"))
sys.path.insert(0, ROOT)

HBM_PEAK_GBPS = 8000.0
FP64_MFMA_PEAK_TFLOPS = 78.6
FP32_MFMA_PEAK_TFLOPS = 157.3
# HBM bytes per launch on the 28-qubit state, from rocprofv3 --pmc FETCH_SIZE and
# WRITE_SIZE (KiB per dispatch, averaged over one step's launches: 18 of k_tblocks,
# 1 of k_block5), each scaled by the factor a plain copy of known size gives for the
# same counter (2 for FETCH_SIZE, 1 for WRITE_SIZE); summary in
# profiles/block_kernels_brickwork28_pmc.md. Keyed by the amplitude size in bytes
# (16 complex128, 8 complex64).
TRAFFIC28_BYTES = {16: {"tblock": int((2 * 2112806 + 4194304) * 1024),
                        "fused": int((2 * 2101370 + 5145837) * 1024)},
                   8: {"tblock": int((2 * 1049172 + 2097152) * 1024),
                       "fused": int((2 * 1084233 + 2350587) * 1024)}}
NBITS = 28
DEPTH = 20
SEED = 20


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def gate_bytes(e, nbits, amp_bytes=16):
    """Algorithmic HBM bytes of one gate applied on its own: every amplitude it can
    change is read and written once (2^n for a 1-qubit gate, 2^(n-c) with c controls)."""
    return 2 * amp_bytes * (1 << (nbits - len(e.control_locs)))


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(els, nbits, locs=(0, 9, 18, 27), pairs=((8, 9), (18, 19))):
    """Oracle (single-threaded C++ restatement of the reference, -O3) on a bounded
    sample of the circuit: layer 0's rotations on `locs` and its CNOTs on `pairs`
    (12 + 2 gates, the 6 : 1 rotation-to-CNOT mix of every layer, targets spread
    over high and low bits), applied to the full-size state and timed whole."""
    from oracle import oracle as O
    from yaohip.circuit import lower
    layer0 = els[:3 * nbits + len(range(0, nbits - 1, 2))]
    sample = [e for e in layer0 if (not e.control_locs and e.target_locs[0] in locs) or
              (e.control_locs and (e.control_locs[0], e.target_locs[0]) in pairs)]
    buf = np.zeros(2 << nbits)
    buf[0] = 1.0
    buf[1:] = 0.0  # touch every page before timing
    prev = O.set_threads(1)
    ol = lower(sample)
    t0 = time.perf_counter()
    st = O.lib().orc_apply(nbits, nbits, buf.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), ol.arr, ol.n)
    dt = time.perf_counter() - t0
    O.set_threads(prev)
    assert st == 0
    return {"value": len(sample) / dt, "unit": "gate-apps/s", "cores": 1, "kind": "port",
            "cpu_model": cpu_model(),
            "sample": (f"oracle (C++ restatement of yao-rs instruct/dispatch, g++ -O3 -ffp-contract=off, one "
                       f"thread) on the full {nbits}-qubit state: layer 0's Rx, Ry, Rz on locs {list(locs)} and "
                       f"its CNOTs on {[list(p) for p in pairs]} ({len(sample)} gates, the circuit's 6:1 "
                       f"rotation-to-CNOT mix), timed whole: {dt:.1f} s")}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--skip-extra", action="store_true", help="only the 28-qubit complex128 brickwork")
    ap.add_argument("--nbits", type=int, default=NBITS)
    ap.add_argument("--depth", type=int, default=DEPTH)
    ap.add_argument("--profile", choices=sorted(PROFILE_WORKLOADS), help="run one workload for a profiler")
    args = ap.parse_args()
    if args.profile:
        profile_workload(args.profile, args.nbits, args.depth)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this backend runs on one GPU; multi-GPU sharding is design only "
                                   "(see DESIGN.md), run with --gpus 1"}))
        sys.exit(2)

    import yaohip
    from yaohip import _abi as A
    from yaohip.circuit import Circuit
    from yaohip.workloads import brickwork

    n, depth = args.nbits, args.depth
    els = brickwork(n, depth, SEED)
    circ = Circuit.qubits(n, els)
    ngates = len(els)
    step_bytes = sum(gate_bytes(e, n) for e in els)
    log(f"workload brickwork n={n} depth={depth} gates={ngates} algorithmic bytes/step={step_bytes/1e9:.1f} GB")

    A.init()
    name = ctypes.create_string_buffer(256)
    A.lib().yh_device_name(name, 256)
    log(f"device {name.value.decode()}")
    reg = yaohip.ArrayReg.zero_state(n)
    L = A.lib()
    ol = circ.ops
    plan = np.zeros(4, dtype=np.int64)
    A.check(L.yh_plan_stats(n, ol.arr, ol.n, A.i64ptr(plan)))

    def run(fused, steps, warmup, label):
        for w in range(warmup):
            t0 = time.perf_counter()
            yaohip.apply_inplace(circ, reg, fused)
            log(f"{label} warmup {w} {time.perf_counter() - t0:.3f} s")
        A.check(L.yh_reg_timing_enable(reg._h, 1))
        t0 = time.perf_counter()
        for s in range(steps):
            yaohip.apply_inplace(circ, reg, fused)
            log(f"{label} step {s} done")
        dt = time.perf_counter() - t0
        kstats = {}
        for k, kn in enumerate(A.KCLASS_NAMES):
            c, ms = ctypes.c_uint64(), ctypes.c_double()
            A.check(L.yh_reg_timing_read(reg._h, k, ctypes.byref(c), ctypes.byref(ms)))
            if c.value:
                kstats[kn] = {"launches": c.value, "avg_us": 1e3 * ms.value / c.value, "total_ms": ms.value}
        A.check(L.yh_reg_timing_enable(reg._h, 0))
        log(f"{label}: {dt / steps * 1e3:.2f} ms/step; norm {reg.norm():.15f}; kernels {kstats}")
        return dt, kstats

    def roofline_of(kstats, amp_bytes=16, dom=None):
        # dominant kernel; every pass type reads and writes every amplitude once
        dom = dom or max(kstats, key=lambda k: kstats[k]["total_ms"])
        bytes_per_launch = 2 * amp_bytes * (1 << n)
        achieved = bytes_per_launch / (kstats[dom]["avg_us"] * 1e-6) / 1e9
        return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBPS, 4), "traffic": None, "kernel": dom,
                "bytes_per_launch": bytes_per_launch, "avg_launch_us": round(kstats[dom]["avg_us"], 2)}

    # algorithmic flops of one block: a 32 x 32 complex matrix on every 32 amplitudes
    # (32 complex multiply-adds of 8 real flops per amplitude), or 16 x 16 on every 16
    F5 = (1 << n) * 32 * 8
    F4 = (1 << n) * 16 * 8

    def mfma_of(kstats, peak, steps):
        # blocks run in k_block5 (one per launch) and k_tblocks (several per launch),
        # so the rate is over both kernels' time
        flops = (blocks5 * F5 + blocks4 * F4) * steps
        t = sum(kstats[k]["total_ms"] for k in ("fused", "tblock") if k in kstats) * 1e-3
        rate = flops / t
        return {"flops_per_block": {"32x32": F5, "16x16": F4}, "blocks_per_step": {"32x32": blocks5, "16x16": blocks4},
                "achieved_tflops": round(rate / 1e12, 2), "peak_tflops": peak, "frac": round(rate / 1e12 / peak, 4)}

    gbg_dt, gbg_k = run(0, 1, 0 if args.warmup == 0 else 1, "gate-by-gate")
    tile_dt, tile_k = run(4, 1, 0 if args.warmup == 0 else 1, "tiled")
    dt, kstats = run(5, args.steps, args.warmup, "blocks")
    bplan = np.zeros(3, dtype=np.int64)
    A.check(L.yh_block_plan_stats(n, ol.arr, ol.n, A.i64ptr(bplan)))
    # the launch plan: passes, multi-block passes, blocks in them, 16 x 16 ones among them
    pst = np.zeros(7, dtype=np.int64)
    A.check(L.yh_block_pass_stats(n, ol.arr, ol.n, 0, A.i64ptr(pst)))
    passes, multi, in_multi, small = (int(x) for x in pst[:4])
    blocks4 = small
    blocks5 = (passes - multi) + (in_multi - small)
    value = ngates * args.steps / dt
    ms_per_step = 1e3 * dt / args.steps
    # The two block kernels: k_block5 (a block alone in its HBM pass, HBM-bound)
    # and k_tblocks (several blocks per pass applied from LDS, MFMA-bound); the
    # roofline reported is the one of the kernel with the larger share of the step.
    gb = ctypes.c_double()
    A.check(L.yh_copy_bandwidth(16 << n, 5, ctypes.byref(gb)))
    copy_gbps = gb.value  # device-to-device copy of a state-sized buffer, HIP events
    log(f"device copy of the state: {copy_gbps:.1f} GB/s")
    def block_rooflines(ks, amp_bytes, peak_tflops, steps):
        """k_block5 (one block per launch) as HBM-bound, k_tblocks (several per
        launch) as MFMA-bound, from the launches' HIP-event averages."""
        traffic = TRAFFIC28_BYTES.get(amp_bytes, {}) if n == 28 else {}
        per = {}
        if "fused" in ks:
            r = roofline_of(ks, amp_bytes=amp_bytes, dom="fused")
            r["traffic"] = traffic.get("fused")
            r["copy_gbps"] = round(copy_gbps, 1)
            r["frac_of_copy"] = round(r["achieved"] / copy_gbps, 4)
            per["k_block5"] = r
        if "tblock" in ks:
            t = ks["tblock"]
            assert t["launches"] == multi * steps, (t["launches"], multi, steps)
            fpl = ((in_multi - small) * F5 + small * F4) / multi
            tf = fpl / (t["avg_us"] * 1e-6) / 1e12
            per["k_tblocks"] = {
                "bound": "mfma", "achieved": round(tf, 2), "peak": peak_tflops, "unit": "TFLOP/s",
                "frac": round(tf / peak_tflops, 4), "traffic": traffic.get("tblock"), "kernel": "tblock",
                "blocks_per_launch": round(in_multi / multi, 3),
                "blocks_16x16_per_launch": round(small / multi, 3), "flops_per_launch": int(fpl),
                # the kernel forms each complex product from three real ones (Gauss),
                # so the MFMA units execute 3/4 of the algorithmic flops
                "mfma_flops_executed_per_launch": int(fpl * 3 // 4),
                "avg_launch_us": round(t["avg_us"], 2),
                "hbm_gbps": round(2 * amp_bytes * (1 << n) / (t["avg_us"] * 1e-6) / 1e9, 1)}
        return per

    per_kernel = block_rooflines(kstats, 16, FP64_MFMA_PEAK_TFLOPS, args.steps)
    dom = max((k for k in ("fused", "tblock") if k in kstats), key=lambda k: kstats[k]["total_ms"])
    roof = dict(per_kernel["k_block5" if dom == "fused" else "k_tblocks"])
    roof["copy_gbps"] = round(copy_gbps, 1)
    out = {
        "metric": "gate-apps/s at 28q c128 (random brickwork depth 20)",
        "value": round(value, 2), "unit": "gate-apps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"random brickwork {n}q depth {depth} complex128 from |0...0>, gates fused "
                               f"into 4- and 5-qubit blocks applied with FP64 MFMA (yh_apply_fused, "
                               f"max_fuse_qubits=5)",
                   "nbits": n, "depth": depth, "gates": ngates, "blocks": int(bplan[0]),
                   "blocks_16x16": blocks4, "hbm_passes": passes},
        "hbm_gbps_effective": round(passes * 2 * 16 * (1 << n) * args.steps / dt / 1e9, 1),
        "kernels": {k: {kk: round(vv, 3) for kk, vv in v.items()} for k, v in kstats.items()},
        "roofline": roof,
        "roofline_kernels": per_kernel,
        "mfma": mfma_of(kstats, FP64_MFMA_PEAK_TFLOPS, args.steps),
        "tiled": {"value": round(ngates / tile_dt, 2), "unit": "gate-apps/s", "ms_per_step": round(1e3 * tile_dt, 3),
                  "tile_passes": int(plan[0]), "gate_groups": int(plan[1]),
                  "hbm_gbps_effective": round(int(plan[0]) * 2 * 16 * (1 << n) / tile_dt / 1e9, 1),
                  "roofline": roofline_of(tile_k),
                  "kernels": {k: {kk: round(vv, 3) for kk, vv in v.items()} for k, v in tile_k.items()}},
        "gate_by_gate": {"value": round(ngates / gbg_dt, 2), "unit": "gate-apps/s",
                         "ms_per_step": round(1e3 * gbg_dt, 3),
                         "hbm_gbps_effective": round(step_bytes / gbg_dt / 1e9, 1),
                         "roofline": roofline_of(gbg_k),
                         "kernels": {k: {kk: round(vv, 3) for kk, vv in v.items()} for k, v in gbg_k.items()}},
    }
    # complex64 variant of the same circuit; error against the FP64 result (the
    # complex128 path is pinned to the oracle at 1e-12 by the parity tests)
    if not args.skip_extra:
        del reg
        reg = yaohip.ArrayReg.zero_state(n)
        yaohip.apply_inplace(circ, reg, 5)
        ref = reg.state_vec()
        reg64 = yaohip.ArrayReg.zero_state(n, "c64")
        yaohip.apply_inplace(circ, reg64, 5)
        A.check(L.yh_reg_timing_enable(reg64._h, 1))
        t0 = time.perf_counter()
        steps64 = max(1, args.steps)
        for _ in range(steps64):
            yaohip.apply_inplace(circ, reg64, 5)
        dt64 = time.perf_counter() - t0
        k64 = {}
        for kc, kn in ((A.K_FUSED, "fused"), (A.K_TBLOCK, "tblock")):
            c, ms = ctypes.c_uint64(), ctypes.c_double()
            A.check(L.yh_reg_timing_read(reg64._h, kc, ctypes.byref(c), ctypes.byref(ms)))
            if c.value:
                k64[kn] = {"launches": c.value, "avg_us": 1e3 * ms.value / c.value, "total_ms": ms.value}
        reg64b = yaohip.ArrayReg.zero_state(n, "c64")
        yaohip.apply_inplace(circ, reg64b, 5)
        got64 = reg64b.state_vec()
        err64 = float(np.max(np.abs(got64 - ref)) / np.max(np.abs(ref)))
        fid64 = float(abs(np.vdot(ref, got64)) ** 2)
        out["complex64"] = {"value": round(ngates * steps64 / dt64, 2), "unit": "gate-apps/s",
                            "ms_per_step": round(1e3 * dt64 / steps64, 3),
                            "max_rel_err_vs_fp64": err64, "fidelity_vs_fp64": fid64,
                            "roofline_kernels": block_rooflines(k64, 8, FP32_MFMA_PEAK_TFLOPS, steps64),
                            "mfma": mfma_of(k64, FP32_MFMA_PEAK_TFLOPS, steps64),
                            "dtype": "f32 amplitudes and FP32 MFMA, block matrices fused in f64 on the host"}
        log(f"complex64: {1e3 * dt64 / steps64:.2f} ms/step, max rel err {err64:.3e}, fidelity {fid64:.12f}")
        del reg64, reg64b, got64, ref
        out["qft12"] = bench_qft12(log)
        out["vqe_batched"] = bench_vqe(log)
        out["noisy_dm"] = bench_dm(log)
        out["surrogate"] = bench_surrogate(log)
    if not args.no_cpu:
        log("cpu baseline (oracle) ...")
        del reg
        out["cpu_baseline"] = cpu_baseline(els, n)
    print(json.dumps(out), flush=True)


def bench_qft12(log, reps=200):
    """12-qubit QFT with SWAPs (84 gates) on a seeded N(0,1) state, then Z(0)Z(1)
    and 1,024 shots from fixed uniforms: wall time per circuit application (the
    state, 64 KiB, is latency-bound work), gate by gate and as fused blocks, plus
    the expectation and sampling calls. Parity is tested against the oracle. At 12
    qubits the whole state is one LDS tile, so the tile path (max_fuse_qubits 1..4)
    runs the circuit in one launch of one workgroup."""
    import yaohip
    from yaohip.circuit import Circuit, OperatorPolynomial, OperatorString, Op
    from yaohip.workloads import qft_with_swaps
    n = 12
    rng = np.random.default_rng(2024)
    psi = rng.standard_normal(1 << n) + 1j * rng.standard_normal(1 << n)
    psi /= np.linalg.norm(psi)
    circ = Circuit.qubits(n, qft_with_swaps(n))
    reg = yaohip.ArrayReg.from_vec(n, psi)
    out = {"nbits": n, "gates": len(circ.elements)}
    for fused, key in ((0, "gate_by_gate"), (5, "fused_blocks"), (4, "lds_tile")):
        yaohip.apply_inplace(circ, reg, fused)
        reg.norm()  # drains the stream
        t0 = time.perf_counter()
        for _ in range(reps):
            yaohip.apply_inplace(circ, reg, fused)
        reg.norm()
        out[key + "_us_per_circuit"] = round((time.perf_counter() - t0) / reps * 1e6, 1)
    # device time of the one-launch tile pass (HIP events around the launch)
    from yaohip import _abi as A
    A.check(A.lib().yh_reg_timing_enable(reg._h, 1))
    for _ in range(reps):
        yaohip.apply_inplace(circ, reg, 4)
    c, ms = ctypes.c_uint64(), ctypes.c_double()
    A.check(A.lib().yh_reg_timing_read(reg._h, A.K_TILE, ctypes.byref(c), ctypes.byref(ms)))
    A.check(A.lib().yh_reg_timing_enable(reg._h, 0))
    out["lds_tile_launches_per_circuit"] = round(c.value / reps, 2)
    out["lds_tile_device_us_per_circuit"] = round(1e3 * ms.value / reps, 1)
    zz = OperatorPolynomial([1.0], [OperatorString([(0, Op.Z), (1, Op.Z)])])
    t0 = time.perf_counter()
    for _ in range(reps):
        yaohip.expect_arrayreg(reg, zz)
    out["expect_zz_us"] = round((time.perf_counter() - t0) / reps * 1e6, 1)
    u = np.random.default_rng(1024).random(1024)
    t0 = time.perf_counter()
    for _ in range(reps):
        yaohip.sample(reg, None, u)
    out["shots_1024_us"] = round((time.perf_counter() - t0) / reps * 1e6, 1)
    log(f"qft12: {out}")
    return out


def bench_vqe(log, count=4096, plain_sample=256):
    """Batched VQE: 4096 ansatz instances, n ~ U{10..20}, depth ~ U{4..8}, angles
    U[-pi, pi], TFIM J = h = 1. evals/s of the one-launch batch (complex64 and
    complex128) against the plain per-circuit GPU path (yh_apply + yh_expect per
    instance), timed on a sample of instances and scaled to the batch."""
    from yaohip.vqe import VqeBatch, plain_energies, random_instances
    inst = random_instances(count, seed=4)
    res = {"instances": count}
    energies = {}
    for dt in ("c128", "c64"):
        b = VqeBatch(count, 20, dt)
        b.energies(inst)  # autotune + graph capture
        reps = 3
        t0 = time.perf_counter()
        for _ in range(reps):
            e = b.energies(inst)
        el = (time.perf_counter() - t0) / reps
        energies[dt] = e
        res[f"batched_{dt}_evals_per_s"] = round(count / el, 1)
        res[f"batched_{dt}_ms"] = round(1e3 * el, 3)
        log(f"vqe batched {dt}: {1e3 * el:.2f} ms for {count} instances")
        del b
    d = np.abs(energies["c64"] - energies["c128"])
    res["c64_abs_dE_vs_fp64"] = {"max": float(d.max()), "median": float(np.median(d)),
                                 "p99": float(np.percentile(d, 99)), "frac_within_1e-6": float(np.mean(d <= 1e-6))}
    # the plain per-circuit path on a sample (same instance mix: every k-th instance)
    idx = np.arange(0, count, count // plain_sample)[:plain_sample]
    from yaohip.vqe import VqeInstances
    sub = VqeInstances(inst.nq[idx], inst.depth[idx], [inst.instance_angles(i) for i in idx])
    plain_energies(VqeInstances(sub.nq[:4], sub.depth[:4], [sub.instance_angles(i) for i in range(4)]))
    t0 = time.perf_counter()
    pe = plain_energies(sub, dtype="c128")
    tp = time.perf_counter() - t0
    res["plain_c128_evals_per_s"] = round(len(idx) / tp, 1)
    res["plain_sample"] = f"{len(idx)} instances (every {count // plain_sample}th), {tp:.2f} s"
    res["plain_vs_batched_c128_max_abs_dE"] = float(np.max(np.abs(pe - energies['c128'][idx])))
    ok64 = res["c64_abs_dE_vs_fp64"]["max"] <= 1e-6
    best = "c64" if ok64 else "c128"
    res["reported_path"] = f"batched {best} (max |dE| vs FP64 {'<=' if ok64 else '>'} 1e-6 for c64)"
    res["evals_per_s_at_dE_1e-6"] = res[f"batched_{best}_evals_per_s"]
    res["speedup_vs_plain"] = round(res[f"batched_{best}_evals_per_s"] / res["plain_c128_evals_per_s"], 1)
    log(f"vqe: {res}")
    return res


def bench_surrogate(log, n=16, depth=6, apq=3, train_pairs=1 << 20, candidates=65536, keep_frac=0.01,
                    schedule=((40, 1e-3), (10, 3e-4)), batch=4096, plain_sample=64):
    """Learned surrogate: 288 angles (16 qubits, depth 6, Rz-Ry-Rz per qubit per
    layer, CZ chain; angles uniform in [-pi/4, pi/4]) -> TFIM energy (J = h = 1).
    MLP 576 trig features -> 4 x 256 GELU -> 1, BF16 MFMA with FP32 accumulation,
    trained by the library's kernels on 2^20 pairs from the batched simulator
    (complex64); screens 65536 candidates and re-evaluates the best 1% exactly
    (batched complex128). Every candidate is also evaluated exactly (the judge):
    reports the surrogate's MAE, recall of the true best 1 % and the regret of the
    best energy it finds, with the screened and exhaustive times."""
    from yaohip.surrogate import SurrogateScreen
    from yaohip.vqe import fixed_instances, plain_energies
    s = SurrogateScreen(n=n, depth=depth, apq=apq, width=256, hidden=4, max_batch=4096, seed=0)
    res = {"ansatz": f"{n}q depth {depth}, Rz-Ry-Rz + CZ chain, {s.n_angles} angles U[-pi/4, pi/4]",
           "network": "576 (cos, sin) features -> 4 x 256 GELU -> 1; bf16 weights/activations, fp32 accum",
           "train_pairs": train_pairs, "candidates": candidates, "keep_frac": keep_frac}
    train = s.sample_angles(train_pairs, seed=101)
    s.exact_energies(s.sample_angles(s.chunk, seed=100), dtype="c64")  # same shape list: autotune, capture
    t0 = time.perf_counter()
    e = s.exact_energies(train, dtype="c64")
    tg = time.perf_counter() - t0
    res["datagen_s"] = round(tg, 3)
    res["datagen_evals_per_s"] = round(train_pairs / tg, 1)
    log(f"surrogate: generated {train_pairs} energies in {tg:.2f} s")
    t0 = time.perf_counter()
    losses = s.fit(train, e, batch=batch, seed=7, schedule=list(schedule))
    tt = time.perf_counter() - t0
    del train, e
    epochs = sum(ep for ep, _ in schedule)
    fwd_flops = 2 * (2 * s.n_angles * 256 + 3 * 256 * 256 + 256)
    res["train"] = {"epochs": epochs, "batch": batch, "schedule": [list(p) for p in schedule],
                    "seconds": round(tt, 3),
                    "epoch_loss_standardised": [round(float(v), 5) for v in losses],
                    "samples_per_s": round(train_pairs * epochs / tt, 1),
                    "tflops_algorithmic": round(3 * fwd_flops * train_pairs * epochs / tt / 1e12, 2)}
    log(f"surrogate: trained {epochs} epochs in {tt:.2f} s, losses {losses}")
    cand = s.sample_angles(candidates, seed=202)
    # warm-up on another pool of the same size: the simulator autotunes and captures
    # its graph once per instance-shape list (here: as many instances as are kept),
    # a one-off per ansatz like training; that first call is reported on its own
    _, _, _, tm0 = s.screen(s.sample_angles(candidates, seed=201), keep_frac=keep_frac, dtype="c128")
    idx, exact, pred, tm = s.screen(cand, keep_frac=keep_frac, dtype="c128")
    t_screen = tm["predict_s"] + tm["reeval_s"]
    # the judge: every candidate evaluated exactly by the batched complex128 simulator
    s.exact_energies(s.sample_angles(candidates, seed=204), dtype="c128")  # same shape list: tune, capture
    t0 = time.perf_counter()
    ex_all = s.exact_energies(cand, dtype="c128")
    t_exh = time.perf_counter() - t0
    k = int(idx.size)
    true_top = np.argsort(ex_all, kind="stable")[:k]
    recall = len(np.intersect1d(idx, true_top)) / k
    regret = float(exact.min() - ex_all.min())
    rng = np.random.default_rng(303)
    rand = rng.choice(candidates, size=k, replace=False)
    res["screen"] = {"predict_s": round(tm["predict_s"], 4), "reeval_s": round(tm["reeval_s"], 4),
                     "reevaluated": k, "first_call_reeval_s": round(tm0["reeval_s"], 4),
                     "predictions_per_s": round(candidates / tm["predict_s"], 1),
                     "exact_evals_per_s": round(k / t_screen, 1),
                     "seconds": round(t_screen, 4)}
    res["exhaustive"] = {"exact_evals": candidates, "seconds": round(t_exh, 4),
                         "exact_evals_per_s": round(candidates / t_exh, 1),
                         "best_exact": round(float(ex_all.min()), 6)}
    res["quality"] = {"mae_all_candidates": round(float(np.mean(np.abs(pred - ex_all))), 5),
                      "energy_std": round(float(ex_all.std()), 5),
                      "recall_at_1pct": round(recall, 4),
                      "best_found_regret": round(regret, 6),
                      "best_found_exact": round(float(exact.min()), 6),
                      "top_mean_exact": round(float(exact.mean()), 5),
                      "random_same_count_mean_exact": round(float(ex_all[rand].mean()), 5),
                      "random_same_count_best_exact": round(float(ex_all[rand].min()), 5),
                      "pool_mean_exact": round(float(ex_all.mean()), 5)}
    res["time_to_best_found_vs_exhaustive"] = round(t_exh / t_screen, 2)
    # the plain per-circuit path on the same ansatz
    sub = fixed_instances(cand[:plain_sample], n, depth, apq)
    plain_energies(fixed_instances(cand[:2], n, depth, apq))
    t0 = time.perf_counter()
    plain_energies(sub, dtype="c128")
    tp = time.perf_counter() - t0
    res["plain_c128_evals_per_s"] = round(plain_sample / tp, 1)
    res["batched_vs_plain"] = round(res["exhaustive"]["exact_evals_per_s"] / res["plain_c128_evals_per_s"], 1)
    res["note"] = ("only exact simulator energies count as evaluations: the screened pipeline "
                   "evaluates the surrogate's best 1% exactly (exact_evals_per_s counts those, over "
                   "prediction plus re-evaluation time); recall_at_1pct and best_found_regret compare its "
                   "choice with exhaustive exact evaluation of the same pool; data generation, training "
                   "and the simulator's first call (autotune + graph capture) are reported separately")
    log(f"surrogate: {res}")
    return res


def bench_dm(log, n=14, depth=10):
    """14-qubit noisy density matrix (2^28 complex128 entries), brickwork depth 10 with
    Depolarizing(1e-3) then AmplitudeDamping(1e-3) after every gate on each touched
    qubit. Reported path: gates and channels as superoperators on the 28-bit vector
    rho, fused into 5-bit MFMA blocks (yh_dm_apply_blocks); beside it the per-gate
    path that fuses each gate with its channels into one pass (yh_dm_apply_fused).
    Outputs <Z_i Z_i+1>, trace, purity, and the largest difference between the paths."""
    import yaohip
    from yaohip import _abi as A
    from yaohip.circuit import Circuit
    from yaohip.workloads import noisy_brickwork
    els = noisy_brickwork(n, depth, 3)
    circ = Circuit.qubits(n, els)
    ngates = sum(1 for e in els if hasattr(e, "gate"))
    nch = len(els) - ngates
    bytes_pass = 2 * 16 * (1 << (2 * n))

    def run(mode):
        dm = yaohip.DensityMatrix.zero_state(n)
        dm.apply(circ, fused=mode)  # warm-up (plan, first launches)
        dm = yaohip.DensityMatrix.zero_state(n)
        A.check(A.lib().yh_dm_timing_enable(dm._h, 1))
        t0 = time.perf_counter()
        dm.apply(circ, fused=mode)
        dt = time.perf_counter() - t0
        kern = {}
        for kc, kn in ((A.K_DM, "dm"), (A.K_FUSED, "fused"), (A.K_TBLOCK, "tblock")):
            c, ms = ctypes.c_uint64(), ctypes.c_double()
            A.check(A.lib().yh_dm_timing_read(dm._h, kc, ctypes.byref(c), ctypes.byref(ms)))
            if c.value:
                kern[kn] = {"launches": int(c.value), "avg_us": round(1e3 * ms.value / c.value, 2),
                            "total_ms": round(ms.value, 3)}
        return dm, dt, kern

    dm, dt, kern = run("blocks")
    tr, pur = dm.trace(), dm.purity()
    zz = [gpu_zz(dm, i) for i in range(n - 1)]
    del dm
    dmf, dtf, kernf = run(True)
    trf, purf = dmf.trace(), dmf.purity()
    zzf = [gpu_zz(dmf, i) for i in range(n - 1)]
    del dmf
    passes = sum(k["launches"] for k in kern.values())
    res = {"nbits": n, "depth": depth, "gates": ngates, "channels": nch,
           "path": "superoperator blocks (yh_dm_apply_blocks)", "passes": passes,
           "elements_per_s": round(len(els) / dt, 1), "seconds": round(dt, 4), "kernels": kern,
           "pass_bytes": bytes_pass, "hbm_gbps_effective": round(passes * bytes_pass / dt / 1e9, 1),
           "trace": [tr.real, tr.imag], "purity": pur, "zz_lsb_sites": zz,
           "per_gate_fused": {"passes": sum(k["launches"] for k in kernf.values()), "seconds": round(dtf, 4),
                              "elements_per_s": round(len(els) / dtf, 1), "kernels": kernf,
                              "pass_gbps": round(bytes_pass / (kernf["dm"]["avg_us"] * 1e-6) / 1e9, 1)},
           "speedup_vs_per_gate_fused": round(dtf / dt, 2),
           "max_abs_diff_vs_per_gate": max([abs(tr - trf), abs(pur - purf)] +
                                           [abs(a - b) for a, b in zip(zz, zzf)])}
    log(f"noisy dm: {res}")
    return res


def gpu_zz(dm, i):
    import yaohip
    from yaohip.circuit import OperatorPolynomial, OperatorString, Op
    return yaohip.expect_dm(dm, OperatorPolynomial([1.0], [OperatorString([(i, Op.Z), (i + 1, Op.Z)])])).real


PROFILE_WORKLOADS = {
    "block": "random brickwork (--nbits, --depth) as fused MFMA blocks, complex128, one application",
    "block64": "the same circuit as fused MFMA blocks, complex64",
    "tiled": "the same circuit through LDS tile passes (max_fuse_qubits 4)",
    "gbg": "the same circuit gate by gate",
    "copy": "three device-to-device copies of a 2^nbits complex128 state (the reference for scaling "
            "FETCH_SIZE / WRITE_SIZE counters)",
    "k1q_bits": "Rx on every qubit of a 2^nbits complex128 state, k_1q launch time per target bit",
    "vqe": "one timed batch of 4096 ansatz instances (n 10..20, depth 4..8), complex64, after the autotuned first",
    "vqe128": "the same batch in complex128",
    "dm": "noisy brickwork on a 2^nbits x 2^nbits density matrix, per-gate fused passes",
    "dmblk": "the same on the superoperator block path",
    "mlp_train": "surrogate training: 2 epochs over 65536 random pairs, 288 angles, batch 4096",
    "mlp_grad": "surrogate loss and gradient (no update, no graph) of one 4096-row batch, 288 angles, 40 times",
    "sample": "1,024 shots from every basis state's probability (sample_from_probs) of a deterministic 12-qubit "
              "and 22-qubit state, 20 times each",
}


def profile_workload(w, n, depth):
    """One named workload, printing a line that names it before it touches the GPU."""
    print(f"[prof] workload={w} nbits={n} depth={depth}", flush=True)
    import yaohip
    from yaohip import _abi as A
    from yaohip.circuit import Circuit
    from yaohip.workloads import brickwork, noisy_brickwork
    if w in ("tiled", "gbg", "block", "block64"):
        reg = yaohip.ArrayReg.zero_state(n, "c64" if w == "block64" else "c128")
        circ = Circuit.qubits(n, brickwork(n, depth, SEED))
        t0 = time.perf_counter()
        yaohip.apply_inplace(circ, reg, {"tiled": 4, "gbg": 0, "block": 5, "block64": 5}[w])
        print(f"[prof] {w} {time.perf_counter() - t0:.3f} s norm {reg.norm():.15f}", flush=True)
    elif w == "copy":
        a = yaohip.ArrayReg.zero_state(n)
        b = yaohip.ArrayReg.zero_state(n)
        for _ in range(3):
            A.check(A.lib().yh_reg_copy(b._h, a._h))
        print(f"[prof] copy of {16 << n} bytes x3 done", flush=True)
    elif w in ("vqe", "vqe128"):
        from yaohip.vqe import VqeBatch, random_instances
        inst = random_instances(4096, seed=4)
        b = VqeBatch(4096, 20, "c64" if w == "vqe" else "c128")
        b.energies(inst)
        t0 = time.perf_counter()
        e = b.energies(inst)
        print(f"[prof] vqe batch {time.perf_counter() - t0:.4f} s mean E {e.mean():.6f}", flush=True)
    elif w == "k1q_bits":
        from yaohip.circuit import Gate, put
        reg = yaohip.ArrayReg.zero_state(n)
        L = A.lib()
        for loc in range(n):
            circ = Circuit.qubits(n, [put([loc], Gate.Rx(0.3))] * 4)
            yaohip.apply_inplace(circ, reg, 0)  # gate by gate: one k_1q launch per gate
            A.check(L.yh_reg_timing_enable(reg._h, 1))
            yaohip.apply_inplace(circ, reg, 0)
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
    elif w == "mlp_train":
        from yaohip.surrogate import Mlp
        rng = np.random.default_rng(5)
        x = rng.uniform(-0.8, 0.8, size=(65536, 288)).astype(np.float32)
        y = np.cos(x[:, :8]).sum(axis=1).astype(np.float32)
        m = Mlp(288, width=256, hidden=4, max_batch=4096, seed=1, encoding="trig")
        t0 = time.perf_counter()
        losses = m.train(x, y, batch=4096, epochs=2, lr=1e-3, seed=3)
        print(f"[prof] mlp_train 2 epochs {time.perf_counter() - t0:.3f} s losses {losses}", flush=True)
    elif w == "sample":
        for nb in (12, 22):
            reg = yaohip.ArrayReg.deterministic_state(nb)
            u = np.random.default_rng(nb).random(1024)
            t0 = time.perf_counter()
            for _ in range(20):
                out = yaohip.sample(reg, None, u)
            print(f"[prof] sample n={nb} 20 x 1024 shots {time.perf_counter() - t0:.4f} s first {int(out[0])}", flush=True)
    elif w == "mlp_grad":
        from yaohip.surrogate import Mlp
        rng = np.random.default_rng(5)
        x = rng.uniform(-0.8, 0.8, size=(4096, 288)).astype(np.float32)
        y = np.cos(x[:, :8]).sum(axis=1).astype(np.float32)
        m = Mlp(288, width=256, hidden=4, max_batch=4096, seed=1, encoding="trig")
        t0 = time.perf_counter()
        for _ in range(40):
            loss, _g = m.loss_grad(x, y)
        print(f"[prof] mlp_grad 40 x {time.perf_counter() - t0:.3f} s loss {loss}", flush=True)


if __name__ == "__main__":
    main()
