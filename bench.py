"""Benchmark of the MI355X solve backend on one GPU.

Headline ("step" = one sweep; the value of the JSON line): the AC sweep of
256 seeded RLC ladders of uneven size (1k-64k sections, 24.9 M unknowns in
all) over 101 log-spaced frequencies from 1 Hz to 1 GHz — 25,856 complex
systems, every one solved by banded LU with partial pivoting after a reverse
Cuthill-McKee ordering, each system's rows cut into parts eliminated in
parallel and joined through the parts' boundary unknowns. The symbolic half
(pattern union, RCM, band rows) is built once and resident in HBM when the
timed region starts; a step is the numeric sweep over all 25,856 systems.
Every system of the last timed sweep is then compared with the oracle's
FP64 sparse LU (ac::run with a direct solve) on the host's cores, and that
run is the CPU baseline. The dominant kernel's HBM roofline position comes
from HIP events on the backend's stream.

The other workloads are measured once each (they are not the headline):
  diode_mesh_256       nonlinear DC, 65,540 unknowns: GPU Newton with
                       ISAI(1)-BiCGSTAB in FP64 and FP32+FP64 refinement,
                       against the committed FP64 LU fixture; the SpMV of
                       its matrix alone and of 64 block-diagonal copies
                       (335 MB, past the Infinity Cache) beside the
                       stream-read probe
  resistor_ladder_10k  linear DC, 10,002 unknowns, Jacobi-BiCGSTAB to 1e-10
  rc_grid_512          transient, 262,148 unknowns, 1,000 backward-Euler
                       steps of 1 ps, warm-started, against the committed
                       SuperLU fixture; plus K Monte Carlo instances batched
                       versus the same instances one by one
  amplifier_adjoint    4,096 stages (8,194 unknowns): operating point and
                       the adjoint gradient of sum (V_out - 0.9)^2 with
                       respect to 12,288 conductances and 2,048 kp
  inverter_chain_netlists  the reference's own benchmark netlists (CMOS
                       inverter chains of 10-200 stages, .DC and .TRAN 1n 100n)
                       from netlist text to CSV, wall time per netlist

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--only headline|ac|all]
"""
from __future__ import annotations

import argparse
import ctypes
import json
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
FP64_PEAK_TFLOPS = 78.6  # MI355X FP64

# HBM traffic of one k_spmv<double> launch on the mesh matrix, from the two
# rocprofv3 counter passes over `bench.py --roofline-only` (FETCH_SIZE, then
# WRITE_SIZE, KiB per dispatch), each scaled by what the counter reads for a
# copy kernel of known size (65,536 KiB each way): FETCH x 1.9994, WRITE x 1.0.
# Summary: profiles/spmv_diode_mesh_256_counters.md.
SPMV_TRAFFIC_BYTES = round((2897.01 * 1.9994 + 512.03 * 1.0) * 1024)
# HBM traffic of one k_part_solve launch of the AC sweep (2,048 rows per
# part), from the same two counter passes over `bench.py --roofline-only`:
# FETCH 11,758,936.7 KiB x 1.99948 + WRITE 39,383,843.2 KiB x 1.0.
# Summary: profiles/ac_sweep_256x101_counters.md.
AC_TRAFFIC_BYTES = round((11758936.7 * 65536.0 / 32776.5 + 39383843.2) * 1024)
# VALU issue of one k_part_solve launch, from a third counter pass over the
# same program (SQ_INSTS_VALU; GRBM_GUI_ACTIVE, summed over the 8 XCDs):
# 3.619e9 wave-level instructions against 1,024 SIMDs x 1.795e8 / 8 cycles,
# one instruction per SIMD per 4 cycles. Summary: the same file.
AC_VALU_ISSUE_FRAC = 3.619e9 / (1024 * 1.795e8 / 8 / 4)


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def spmv_bytes(n: int, nnz: int, v: int) -> int:
    """Algorithmic bytes of one CSR SpMV: values + column indices, row
    pointers, x read once, y written once."""
    return nnz * (v + 4) + 4 * (n + 1) + 2 * n * v


def replicated(g, copies: int):
    """Block-diagonal replication of a CSR pattern (a working set beyond the
    256 MB Infinity Cache)."""
    n, nnz = g.nrows, g.nnz
    rp = np.empty(n * copies + 1, dtype=np.uint64)
    rp[0] = 0
    base = g.row_pointers[1:].astype(np.uint64)
    for c in range(copies):
        rp[1 + c * n: 1 + (c + 1) * n] = base + np.uint64(c * nnz)
    col = (g.col_indices.astype(np.uint64)[None, :] +
           (np.arange(copies, dtype=np.uint64) * np.uint64(n))[:, None]).ravel()
    val = np.tile(g.values, copies)
    return val, col.astype(np.uint32), rp.astype(np.uint32), n * copies


def within(x, ref, rel=1e-6, abs_=1e-9):
    err = np.abs(x - ref)
    return float(err.max(initial=0.0)), bool(np.all(err <= np.maximum(rel * np.abs(ref), abs_)))


# ------------------------------------------------------------- headline --
def ac_wave_counts(sizes, n_freqs: int, packed: bool):
    """Waves of the sweep holding each circuit's frequencies. Circuits
    longest first; one lane per system: ceil(n_freqs / 64) waves each;
    partitioned (packed): a wave of 64 lanes takes the next circuit's
    frequencies once one circuit's run out, up to two circuits per wave."""
    if not packed:
        return [-(-n_freqs // 64) if n > 0 else 0 for n in sizes]
    counts = [0] * len(sizes)
    fill, segs = 64, 2  # the open wave: lanes used, circuits in it
    for c in sorted((c for c in range(len(sizes)) if sizes[c] > 0), key=lambda c: -sizes[c]):
        left = n_freqs
        while left > 0:
            if fill < 64 and segs < 2:
                take = min(64 - fill, left)
                fill, segs = fill + take, segs + 1
            else:
                take = min(64, left)
                fill, segs = take, 1
            counts[c] += 1
            left -= take
    return counts


def ac_bytes(sizes, n_freqs: int, rows_per_part: int):
    """Algorithmic bytes of one sweep's kernels. A partitioned wave holds 64
    frequencies of up to two circuits (ac_wave_counts); the band rows of G
    and C and the permuted b (3 + 3 + 2 doubles = 64 bytes per row at half
    bandwidth 1) of each of its circuits are read by the wave once per pass:
    once per direction by the relations and once by the part solve (the
    identity rows that pad the shorter circuit are not read, and their x and
    checkpoints are not counted). The relations' top-down pass stores the elimination state
    every 8 steps (4 complex = 64 bytes per system: ceil((m - 1) / 8) of them
    in a part of m rows), which the part solve reads back to recompute each
    chunk's pivot rows in registers before writing x (16 bytes per (system,
    row)). The residual pass reads, per part and system, x at the part's
    first and last rows and their neighbours (6 complex) and the part's two
    partial sums."""
    rows = float(sum(sizes))
    band = 64.0
    wc = ac_wave_counts(sizes, n_freqs, packed=bool(rows_per_part))
    band_rows = float(sum(n * w for n, w in zip(sizes, wc)))  # (row, wave) pairs
    if not rows_per_part:  # one lane per system (k_band_lu<1>): U rows of 3 complex
        return {"part_solve": rows * 112.0 * n_freqs + band_rows * band, "relations": 0.0,
                "residual": 0.0}
    parts, ck, part_waves = 0, 0, 0
    for n, w in zip(sizes, wc):
        p = -(-n // rows_per_part)
        parts += p
        part_waves += p * w
        ck += p * -(-(-(-n // p) - 1) // 8)
    return {"part_solve": rows * 16.0 * n_freqs + ck * 64.0 * n_freqs + band_rows * band,
            "relations": band_rows * band * 2 + ck * 64.0 * n_freqs,
            "residual": parts * n_freqs * (6 * 16.0 + 16.0) + part_waves * 2 * band}


def ac_flops(sizes, n_freqs: int, rows_per_part: int):
    """FP64 flops of one sweep's two main kernels (an FMA counted as 2, a
    complex multiply as 6, a complex a - b c as 8). The part solve, per
    (system, row): the forward step recomputed from the checkpoints - on a
    step whose pivot choice all lanes share (86 % on these ladders): omega c
    products 3, the two squared pivot moduli 6, the pivot's reciprocal 11,
    the multiplier 6, the stored pivot row 12 (or the remainder's third
    entry 6) and the carried row's update 16: 48-54; on the others the
    scaled two-sided combination: 71; 55 on average - and the back
    substitution with the row's residual (omega c 3, x 8, or 22 on the rows
    whose pivot was the next row, the residual 28): 44. 99 in all. The
    relations, per (system, row) and direction: the scale 5, omega c 3 and
    the carried row with its two right-hand-side columns (40): 48."""
    rows = float(sum(sizes)) * n_freqs
    if not rows_per_part:
        return {"part_solve": 0.0, "relations": 0.0, "residual": 0.0}
    return {"part_solve": 99.0 * rows, "relations": 2 * 48.0 * rows, "residual": 0.0}


def check_all_systems(plan, mnas, freqs, threads: int):
    """Every (circuit, frequency) solution of the plan's last sweep against
    the oracle's FP64 sparse LU (oracle.ac_sweep: the union pattern's RCM
    once per circuit, one banded LU per frequency), circuits spread over
    `threads` host threads. Returns the parity figures and the oracle's
    wall time."""
    import concurrent.futures as cf

    import oracle

    def job(c):
        t = time.perf_counter()
        ref = oracle.ac_sweep(mnas[c], freqs)
        return ref, time.perf_counter() - t

    n_sys = len(mnas) * len(freqs)
    bad, worst_abs, worst_rel, busy = 0, 0.0, 0.0, 0.0
    order = iter(sorted(range(len(mnas)), key=lambda c: -mnas[c].size))
    pending = {}
    t0 = time.perf_counter()
    with cf.ThreadPoolExecutor(threads) as ex:
        def submit():
            c = next(order, None)
            if c is not None:
                pending[ex.submit(job, c)] = c

        for _ in range(2 * threads):
            submit()
        while pending:
            done, _ = cf.wait(list(pending), return_when=cf.FIRST_COMPLETED)
            for fut in done:
                c = pending.pop(fut)
                ref, sec = fut.result()
                busy += sec
                x = plan.x(c)
                err = np.abs(x - ref)
                bound = np.maximum(1e-6 * np.abs(ref), 1e-9)
                bad += int(np.count_nonzero((err > bound).any(axis=1)))
                worst_abs = max(worst_abs, float(err.max(initial=0.0)))
                worst_rel = max(worst_rel, float((err / np.maximum(np.abs(ref), 1e-300)).max()))
                del x, ref, err, bound
                submit()
    wall = time.perf_counter() - t0
    return {"systems_checked": n_sys, "systems_outside_1e-6_rel_1e-9_abs": bad,
            "max_abs_err": worst_abs, "max_rel_err": worst_rel}, wall, busy


def ac_headline(args, be, N, ir, solver):
    t0 = time.perf_counter()
    ckts, freqs = ir.rlc_sweep(n_circuits=args.ac_circuits, seed=4)
    mnas = [solver.compile(c) for c in ckts]
    del ckts
    t_build = time.perf_counter() - t0
    nf = int(freqs.shape[0])
    n_sys = len(mnas) * nf
    sizes = [m.size for m in mnas]
    log(f"AC workload: {len(mnas)} circuits, {sum(sizes):,} unknowns, {n_sys:,} systems "
        f"(built in {t_build:.1f} s)")
    t0 = time.perf_counter()
    plan = solver.AcPlan(be, mnas)
    t_plan = time.perf_counter() - t0
    first = None
    for w in range(max(args.warmup, 1)):  # the first sweep autotunes the rows per part
        t0 = time.perf_counter()
        plan.solve(freqs)
        if w == 0:
            first = time.perf_counter() - t0
    be.sync()
    t0 = time.perf_counter()
    res, acc = None, {}
    for _ in range(args.steps):
        res = plan.solve(freqs)
        # per-phase device ms of this sweep (HIP events on the backend's stream)
        for k, v in plan.stats()["phase_ms"].items():
            acc[k] = acc.get(k, 0.0) + v
    be.sync()
    dt = time.perf_counter() - t0
    st = plan.stats()
    worst = float(res.residual_norms.max())
    failed = int((~res.converged).sum())
    ms = {k: v / args.steps for k, v in acc.items()}  # average over the timed sweeps
    by = ac_bytes(sizes, nf, st["rows_per_part"])
    fl = ac_flops(sizes, nf, st["rows_per_part"])
    kernel = {"part_solve": "k_part_solve" if st["rows_per_part"] else "k_band_lu<1>",
              "relations": "k_part_rel", "residual": "k_part_resid_sum"}
    dom = max(("part_solve", "relations", "residual"), key=lambda k: ms[k])
    achieved = by[dom] / (ms[dom] * 1e-3) / 1e9
    log(f"AC sweep: {1e3 * dt / args.steps:.1f} ms/sweep ({n_sys * args.steps / dt:,.0f} "
        f"systems/s), device {sum(ms.values()):.1f} ms {ms}, rows/part {st['rows_per_part']}, "
        f"redone {st['redone_systems']}, worst residual {worst:.1e}")
    parity, cpu = None, None
    if not args.no_cpu_baseline:
        import os
        threads = max(1, min(16, os.cpu_count() or 1))
        parity, wall, busy = check_all_systems(plan, mnas, freqs, threads)
        cpu = {"value": n_sys / wall, "unit": "solves/s", "cores": threads, "kind": "port",
               "host": host_info(threads),
               "sample": f"all {n_sys:,} systems of the sweep: oracle.ac_sweep (ac::run with the "
                         f"oracle's FP64 sparse LU: RCM once per circuit, banded LU with partial "
                         f"pivoting per frequency) on {threads} host threads, {wall:.1f} s wall, "
                         f"{busy:.1f} thread-s ({1e3 * busy / n_sys:.2f} ms per system on one "
                         f"thread)"}
        log(f"AC parity: {parity}; CPU oracle {n_sys / wall:,.0f} systems/s on {threads} threads")
    out = {
        "value": n_sys * args.steps / dt, "ms_per_step": 1e3 * dt / args.steps,
        "systems": n_sys, "circuits": len(mnas), "frequencies": nf,
        "unknowns_per_sweep": int(sum(sizes)) * nf,
        "rows_per_part": st["rows_per_part"], "redone_systems": st["redone_systems"],
        "chunks": st["chunks"], "device_ms_per_sweep": ms,
        "worst_residual": worst, "failed_systems": failed,
        "build_s": t_build, "symbolic_plan_s": t_plan, "first_sweep_s_with_autotune": first,
        "end_to_end_s": t_plan + dt / args.steps,
        "roofline": {"bound": "hbm", "kernel": kernel[dom], "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": AC_TRAFFIC_BYTES,
                     "traffic_unit": "HBM bytes per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE, "
                                     "copy-calibrated)" if AC_TRAFFIC_BYTES else None,
                     "bytes_per_launch": by[dom], "avg_launch_ms": ms[dom],
                     "bytes_per_launch_all": by,
                     # what limits the kernel: the issue rate of its vector
                     # instructions (counter pass), beside its FP64 flop rate
                     # against the 78.6 TFLOP/s peak
                     "valu_issue_frac": AC_VALU_ISSUE_FRAC if dom == "part_solve" and
                     st["rows_per_part"] else None,
                     "fp64": {"flops_per_launch": fl[dom],
                              "achieved_tflops": fl[dom] / (ms[dom] * 1e-3) / 1e12,
                              "peak_tflops": FP64_PEAK_TFLOPS,
                              "frac": fl[dom] / (ms[dom] * 1e-3) / 1e12 / FP64_PEAK_TFLOPS}},
        "parity": parity, "cpu_baseline": cpu,
    }
    return out, mnas, freqs


def mesh_newton(args, be, N, ir, solver):
    """Nonlinear DC on the 256x256 diode mesh, FP64 and FP32-IR, against the
    committed FP64 LU-Newton fixture (every 4th unknown)."""
    import os
    ckt = ir.diode_mesh(256, 256, seed=2)
    m = solver.compile(ckt)
    ctx = solver.NewtonContext(be, m.g, m.b_dc, m.diode_descriptors)
    gold = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "tests", "golden",
                                "diode_mesh_256.npz"))
    ref = gold["x_every4"]
    # bicgstab_floor: the sub-solves may stop at floor * ||b||. The assembled
    # ||b|| (3.4e4) is set by the diodes at the 1 V corners, so the default
    # 1e-15 stops Newton at ||F|| ~ 3e-11 and leaves ~5e-7 V on the mesh
    # nodes; 1e-18 reaches ||F|| ~ 3e-14 and errors ~3e-10 V.
    # forcing: Eisenstat-Walker choice 1 with its safeguard at every
    # iteration, so eta cannot collapse on the large residual drops of the
    # diode turn-on phase. eta_max 0.9 is Eisenstat and Walker's published
    # ceiling.
    newton_kw = dict(abs_tol=1e-10, max_iterations=100, bicgstab_tolerance=1e-6,
                     bicgstab_max_iterations=20_000, forcing="ew1", forcing_max=0.9,
                     forcing_safeguard="always", bicgstab_floor=1e-18)
    forms = {}
    steps = max(1, min(args.steps, 5))
    for prec in ("fp64", "fp32_ir", "fp64_fixed_tol"):
        kw = dict(newton_kw, precision=prec)
        if prec == "fp64_fixed_tol":  # every sub-solve to 1e-6 (no forcing)
            kw.update(precision="fp64", forcing="fixed")
        ctx.run(**kw)
        be.sync()
        t0 = time.perf_counter()
        stats = None
        for _ in range(steps):
            stats = ctx.run(**kw)
        be.sync()
        dt = time.perf_counter() - t0
        e, ok = within(ctx.x()[::4], ref)
        forms[prec] = {
            "solves_per_s": steps / dt, "ms_per_solve": 1e3 * dt / steps,
            "newton_iterations": stats.newton_iterations,
            "bicgstab_iterations": stats.bicgstab_iterations,
            "fp64_fallbacks": stats.fp64_fallbacks,
            "max_abs_err_vs_lu_fixture": e, "within_1e-6_rel_1e-9_abs": ok,
            "us_per_bicgstab_iteration": 1e6 * dt / steps / max(stats.bicgstab_iterations, 1)}
        log(f"diode mesh {prec}: {forms[prec]['ms_per_solve']:.2f} ms/solve, "
            f"{stats.newton_iterations} Newton, {stats.bicgstab_iterations} BiCGSTAB its, "
            f"err {e:.1e}")
    it_b = isai_iteration_bytes(m.g)
    us = forms["fp64"]["us_per_bicgstab_iteration"]
    out = {"unknowns": m.size, "nnz": m.g.nnz, "forms": forms,
           "iteration_roofline": {
               "bytes_per_iteration": it_b, "us_per_iteration_incl_newton": us,
               "achieved_gbs": it_b / (us * 1e-6) / 1e9,
               "frac": it_b / (us * 1e-6) / 1e9 / HBM_PEAK_GBS}}
    # stand-alone SpMV of the system matrix, HIP events on the backend stream
    a = be.upload_matrix(m.g.values, m.g.col_indices, m.g.row_pointers, m.size)
    bx, by = be.new_buffer(m.size), be.new_buffer(m.size)
    be.upload_vec(np.ones(m.size), bx)
    msv = ctypes.c_double()
    N.check(N.lib().ohm_spmv_timed(be.h, a.h, bx.h, by.h, 200, 0, ctypes.byref(msv)))
    b = spmv_bytes(m.size, m.g.nnz, 8)
    out["spmv"] = {"bytes_per_launch": b, "avg_launch_us": msv.value * 1e3,
                   "achieved_gbs": b / (msv.value * 1e-3) / 1e9,
                   "frac": b / (msv.value * 1e-3) / 1e9 / HBM_PEAK_GBS,
                   "traffic_bytes_rocprof": SPMV_TRAFFIC_BYTES}
    out["spmv_large_working_set"] = spmv_large(be, N, m.g)
    return out


def spmv_large(be, N, g, copies=64):
    """The mesh matrix replicated block-diagonally 64 times (a 335 MB working
    set, past the 256 MB Infinity Cache): FP64 and FP32 SpMV with the
    library's default row-block shape, HIP events over 100 back-to-back
    launches, against the 8 TB/s peak and against the stream-read probe over
    a buffer of the same size (32-byte loads)."""
    val, col, rp, n = replicated(g, copies)
    a = be.upload_matrix(val, col, rp, n)
    bx, by = be.new_buffer(n), be.new_buffer(n)
    be.upload_vec(np.ones(n), bx)
    out = {"copies": copies, "unknowns": n, "nnz": int(val.shape[0])}
    for name, prec, v in (("fp64", 0, 8), ("fp32", N.FP32_IR, 4)):
        ms = ctypes.c_double()
        N.check(N.lib().ohm_spmv_timed(be.h, a.h, bx.h, by.h, 100, prec, ctypes.byref(ms)))
        b = spmv_bytes(n, int(val.shape[0]), v)
        out[name] = {"bytes_per_launch": b, "avg_launch_us": ms.value * 1e3,
                     "achieved_gbs": b / (ms.value * 1e-3) / 1e9,
                     "frac": b / (ms.value * 1e-3) / 1e9 / HBM_PEAK_GBS}
    ms = ctypes.c_double()
    nbytes = out["fp64"]["bytes_per_launch"]
    N.check(N.lib().ohm_stream_timed(be.h, ctypes.c_int64(nbytes), 0, 50, ctypes.byref(ms)))
    rd = nbytes / (ms.value * 1e-3) / 1e9
    out["stream_read_gbs"] = rd
    out["stream_read_frac"] = rd / HBM_PEAK_GBS
    log(f"SpMV 335 MB: fp64 {out['fp64']['achieved_gbs']:.0f} GB/s "
        f"({out['fp64']['frac']:.3f}), fp32 {out['fp32']['achieved_gbs']:.0f} GB/s "
        f"({out['fp32']['frac']:.3f}); stream read {rd:.0f} GB/s ({rd / HBM_PEAK_GBS:.3f})")
    return out


def isai_iteration_bytes(g) -> float:
    """Algorithmic bytes of one ISAI(1)-BiCGSTAB iteration in FP64: two SpMVs
    with A, two applies of M_U M_L (M_L on pattern(L^2), M_U on
    pattern(U^2) of A's lower / upper parts with the diagonal), and the
    vector traffic of the fused schedule (22 n doubles)."""
    import scipy.sparse as sp
    n = g.nrows
    a = sp.csr_matrix((np.ones(g.nnz), g.col_indices, g.row_pointers), shape=(n, n))
    a = a + sp.identity(n, format="csr")
    lo = sp.tril(a, format="csr")
    up = sp.triu(a, format="csr")
    nl = (lo @ lo).nnz
    nu = (up @ up).nnz
    return float(2 * spmv_bytes(n, g.nnz, 8) + 2 * (spmv_bytes(n, nl, 8) + spmv_bytes(n, nu, 8))
                 + 22 * n * 8)


def host_info(threads: int = 1) -> dict:
    """The CPU the baseline ran on: logical CPUs visible and the model name."""
    import os
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "cpu_model": model, "threads_used": threads}


# ------------------------------------------------------------ workloads --
def ladder(be, N, ir, solver, check):
    """Jacobi-BiCGSTAB at the configuration's 1e-10 relative residual and at
    1e-13, the tolerance at which every node voltage (they decay by orders
    of magnitude along the ladder) is within 1e-6 of LU."""
    mna = solver.compile(ir.resistor_ladder(10_000, seed=1))
    ref = None
    if check:
        import oracle
        ref = oracle.sparse_lu_solve(mna.g, mna.b_dc)
    out = {"unknowns": mna.size}
    for tol in (1e-10, 1e-13):
        s = solver.GpuSolver(be, precond=N.PREC_JACOBI,
                             params=solver.bicgstab_params(tolerance=tol, max_iterations=200_000))
        s.solve_real(mna.g, mna.b_dc)  # warm-up
        t0 = time.perf_counter()
        reps = 5
        for _ in range(reps):
            x = s.solve_real(mna.g, mna.b_dc)
        dt = (time.perf_counter() - t0) / reps
        r = {"ms_per_solve": 1e3 * dt, "solves_per_s": 1 / dt,
             "bicgstab_iterations": s.last_stats.iterations,
             "residual_norm": s.last_stats.residual_norm}
        if ref is not None:
            e, ok = within(x, ref)
            r.update({"max_abs_err_vs_lu": e, "within_1e-6_rel_1e-9_abs": ok})
        out[f"tol_{tol:g}"] = r
        log(f"ladder tol {tol:g}: {r['ms_per_solve']:.2f} ms/solve, {r['bicgstab_iterations']} its")
    return out


def rc_grid(be, N, ir, solver, steps, mc_k, mc_steps):
    from ohmnivore_amd import transient as T
    ckt = ir.rc_grid(512, 512, 1024, seed=3)
    mna = solver.compile(ckt)
    dc = solver.GpuSolver(be, precond=N.PREC_AUTO,
                          params=solver.bicgstab_params(tolerance=1e-12, max_iterations=100_000))
    x0 = dc.solve_real(mna.g, mna.b_dc)
    src, _ = T.source_table(ckt, mna)
    import os
    gold = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "tests", "golden",
                                "rc_grid_512.npz"))
    probes = gold["probes"]
    every = int(gold["every"])
    out = {"unknowns": mna.size, "nnz": mna.g.nnz, "steps": steps, "h": 1e-12}
    it_b = isai_iteration_bytes(mna.g)
    for name, pc, prec in (("fp64_isai", N.PREC_ISAI, "fp64"), ("fp64_jacobi", N.PREC_JACOBI, "fp64"),
                           ("fp32_ir_isai", N.PREC_ISAI, "fp32_ir")):
        tr = T.Transient(be, mna, src)
        prm = solver.bicgstab_params(tolerance=1e-9, max_iterations=20_000, precision=prec)
        t0 = time.perf_counter()
        try:
            r = tr.run(1e-12, steps * 1e-12, method="be", x0=x0, probes=probes, params=prm,
                       precond=pc)
        except N.OhmnivoreError as e:
            out[name] = {"error": str(e)}
            log(f"rc grid {name}: FAILED: {e}")
            continue
        dt = time.perf_counter() - t0
        out[name] = {"steps": r.steps_accepted, "ms_per_step": 1e3 * dt / r.steps_accepted,
                     "steps_per_s": r.steps_accepted / dt,
                     "bicgstab_iters_per_step": r.bicgstab_iterations / r.steps_accepted}
        w = r.waves[:, 0, :]
        if w.shape[0] > every * (gold["waves_every"].shape[0] - 1):
            e, ok = within(w[::every][: gold["waves_every"].shape[0]], gold["waves_every"])
            e2, ok2 = within(w[: gold["waves"].shape[0]], gold["waves"])
            out[name].update({"max_abs_err_vs_superlu_fixture": max(e, e2),
                              "within_1e-6_rel_1e-9_abs": ok and ok2})
        if pc == N.PREC_ISAI and prec == "fp64":
            us = 1e6 * dt / max(r.bicgstab_iterations, 1)
            out[name]["iteration_roofline"] = {
                "bytes_per_iteration": it_b, "us_per_iteration_incl_step": us,
                "achieved_gbs": it_b / (us * 1e-6) / 1e9,
                "frac": it_b / (us * 1e-6) / 1e9 / HBM_PEAK_GBS}
        log(f"rc grid {name}: {out[name]['ms_per_step']:.2f} ms/step, "
            f"{out[name]['bicgstab_iters_per_step']:.1f} its/step, "
            f"parity {out[name].get('within_1e-6_rel_1e-9_abs')}")
    # Monte Carlo: K perturbed instances batched vs one at a time (both FP64, Jacobi)
    cg = T.c_on_g(mna)
    rng = np.random.default_rng(7)
    inst = [(mna.g.values * (1 + 0.05 * (rng.random() - 0.5)),
             cg * (1 + 0.2 * rng.random(cg.shape[0]))) for _ in range(mc_k)]
    prm = solver.bicgstab_params(tolerance=1e-9, max_iterations=20_000)
    trb = T.Transient(be, mna, src, instances=inst)
    t0 = time.perf_counter()
    rb = trb.run(1e-12, mc_steps * 1e-12, method="be", x0=x0, probes=probes, params=prm,
                 precond=N.PREC_JACOBI)
    t_batched = time.perf_counter() - t0
    t0 = time.perf_counter()
    plain_its = 0
    for k in range(mc_k):
        trk = T.Transient(be, mna, src, instances=[inst[k]])
        rk = trk.run(1e-12, mc_steps * 1e-12, method="be", x0=x0, probes=probes, params=prm,
                     precond=N.PREC_JACOBI)
        plain_its += rk.bicgstab_iterations
    t_plain = time.perf_counter() - t0
    solves = mc_k * mc_steps
    out["monte_carlo"] = {
        "instances": mc_k, "steps": mc_steps,
        "batched_solves_per_s": solves / t_batched, "plain_solves_per_s": solves / t_plain,
        "speedup": t_plain / t_batched, "batched_iterations": rb.bicgstab_iterations,
        "plain_iterations": plain_its}
    log(f"rc grid Monte Carlo K={mc_k}: batched {solves / t_batched:.1f} solves/s, plain "
        f"{solves / t_plain:.1f} solves/s, speedup {t_plain / t_batched:.2f}")
    return out


def ac_plain(be, N, ir, solver, mnas, freqs):
    """One system per call on the same device path (a plan per circuit,
    its symbolic half cached, one frequency per solve; device time only),
    on 8 sampled (circuit, frequency) pairs; and the reference's own AC
    method, plain Jacobi-BiCGSTAB on one system (src/solver/gpu.rs:210-735),
    on 2 of them."""
    from ohmnivore_amd import transient as T
    nc, nf = len(mnas), int(freqs.shape[0])
    picks = [(i * (nc // 8), (i * 13) % nf) for i in range(8)]
    dev_ms, wall = [], []
    for c, f in picks:
        p = solver.AcPlan(be, [mnas[c]])
        p.solve(freqs[f:f + 1])  # first solve: autotunes its rows per part
        t0 = time.perf_counter()
        r = p.solve(freqs[f:f + 1])
        wall.append(time.perf_counter() - t0)
        dev_ms.append(sum(p.stats()["phase_ms"].values()))
        del p, r
    gs = solver.GpuSolver(be, precond=N.PREC_JACOBI,
                          params=solver.bicgstab_params(tolerance=1e-10, max_iterations=10_000))
    sample = []
    for c, f in picks[:2]:
        m = mnas[c]
        cg = T.c_on_g(m)
        a = ir.CsrMatrix(m.size, m.size, m.g.values + 2j * np.pi * freqs[f] * cg,
                         m.g.col_indices, m.g.row_pointers)
        t1 = time.perf_counter()
        try:
            gs.solve_complex(a, m.b_ac)
            conv = True
        except N.OhmnivoreError:
            conv = False
        sample.append({"unknowns": m.size, "freq_hz": float(freqs[f]), "converged": conv,
                       "iterations": gs.last_stats.iterations,
                       "seconds": time.perf_counter() - t1})
    per_sys = [d / m.size for d, (c, _) in zip(dev_ms, picks) for m in [mnas[c]]]
    return {"sample": [{"circuit": c, "frequency": f, "unknowns": mnas[c].size,
                        "device_ms": d, "wall_ms": 1e3 * w}
                       for (c, f), d, w in zip(picks, dev_ms, wall)],
            "device_ms_per_unknown": float(np.mean(per_sys)),
            "plain_bicgstab_sample": sample}


def adjoint(be, N, ir, solver):
    from ohmnivore_amd import adjoint as A
    ckt = ir.amplifier_array(4096, seed=5)
    mna = solver.compile(ckt)
    prob = A.DcProblem(be, ckt, mna)
    outs = A.output_nodes(ckt, mna)
    prob.solve()
    t0 = time.perf_counter()
    x = prob.solve()
    t1 = time.perf_counter()
    dj = np.zeros(mna.size)
    dj[outs] = 2 * (x[outs] - 0.9)
    gg, gk = prob.adjoint(dj)
    t2 = time.perf_counter()
    out = {"unknowns": mna.size, "parameters": int(gg.shape[0] + gk.shape[0]),
           "forward_ms": 1e3 * (t1 - t0), "adjoint_ms": 1e3 * (t2 - t1),
           "newton_iterations": prob.last_stats.newton_iterations,
           "adjoint_bicgstab_iterations": prob.last_adjoint.iterations,
           "gradients_per_s": 1 / (t2 - t0)}
    log(f"adjoint: forward {out['forward_ms']:.1f} ms, adjoint {out['adjoint_ms']:.1f} ms")
    return out


def roofline_only(args, be, N, ir, solver) -> int:
    """The headline's kernels alone, for rocprofv3 counter runs: two sweeps
    of the AC workload at the given rows per part (OHM_AC_PART_ROWS; the
    autotuner is off), the stand-alone SpMV on the diode-mesh matrix, and a
    plain copy kernel of known size (64 MiB read + 64 MiB written per
    launch) as the counter calibration."""
    be.set_autotune(False)
    ckts, freqs = ir.rlc_sweep(n_circuits=args.ac_circuits, seed=4)
    mnas = [solver.compile(c) for c in ckts]
    del ckts
    plan = solver.AcPlan(be, mnas)
    for _ in range(2):
        plan.solve(freqs)
    st = plan.stats()
    sizes = [m.size for m in mnas]
    by = ac_bytes(sizes, int(freqs.shape[0]), st["rows_per_part"])
    del plan
    m = solver.compile(ir.diode_mesh(256, 256, seed=2))
    a = be.upload_matrix(m.g.values, m.g.col_indices, m.g.row_pointers, m.size)
    bx, by_ = be.new_buffer(m.size), be.new_buffer(m.size)
    be.upload_vec(np.ones(m.size), bx)
    ms = ctypes.c_double()
    N.check(N.lib().ohm_spmv_timed(be.h, a.h, bx.h, by_.h, 200, 0, ctypes.byref(ms)))
    nc = 8 * 1024 * 1024  # 64 MiB of doubles
    cx, cy = be.new_buffer(nc), be.new_buffer(nc)
    for _ in range(20):
        be.copy(cx, cy)
    be.sync()
    b = spmv_bytes(m.size, m.g.nnz, 8)
    print(json.dumps({"ac": {"stats": st, "bytes_per_launch": by},
                      "spmv": {"kernel": "k_spmv<double>", "bytes_per_launch": b,
                               "avg_launch_us": ms.value * 1e3,
                               "achieved_gbs": b / (ms.value * 1e-3) / 1e9},
                      "copy_kernel": "k_copy", "copy_bytes_per_launch": 2 * 8 * nc}))
    return 0


def spmv_large_only(args, be, N, ir, solver) -> int:
    """The 335 MB SpMV alone (FP64 and FP32 on the 64-copy mesh matrix, the
    stream-read probe over a buffer of the same size) and the copy kernel
    calibration, for rocprofv3 kernel-trace and counter runs."""
    m = solver.compile(ir.diode_mesh(256, 256, seed=2))
    r = spmv_large(be, N, m.g)
    nc = 8 * 1024 * 1024  # 64 MiB of doubles
    cx, cy = be.new_buffer(nc), be.new_buffer(nc)
    for _ in range(20):
        be.copy(cx, cy)
    be.sync()
    r.update({"copy_kernel": "k_copy", "copy_bytes_per_launch": 2 * 8 * nc})
    print(json.dumps(r))
    return 0


def chain_netlist(stages, analysis):
    """The reference's own benchmark circuit family (bench/generate_chain.py,
    run by bench/run_benchmark.sh for 10-200 stages): a CMOS inverter chain
    with VDD = 5, VIN = 5 V DC (.DC) or PULSE(0 5 10n 1n 1n 50n 100n)
    (.TRAN 1n 100n)."""
    lines = [f"* CMOS Inverter Chain - {stages} stages - {analysis.upper()}",
             ".MODEL NMOD NMOS(VTO=0.7 KP=1.1e-4 LAMBDA=0.04)",
             ".MODEL PMOD PMOS(VTO=-0.7 KP=5.5e-5 LAMBDA=0.04)", "VDD vdd 0 DC 5",
             "VIN in 0 DC 5" if analysis == "dc" else "VIN in 0 PULSE(0 5 10n 1n 1n 50n 100n)"]
    for i in range(1, stages + 1):
        gate = "in" if i == 1 else f"out_{i - 1}"
        lines += [f"MP{i} out_{i} {gate} vdd vdd PMOD", f"MN{i} out_{i} {gate} 0 0 NMOD"]
    lines += [".DC" if analysis == "dc" else ".TRAN 1n 100n", ".END", ""]
    return "\n".join(lines)


def inverter_chains(be):
    """Netlist text -> parse -> compile -> GPU analysis -> CSV, wall time per
    netlist, on the reference's benchmark chains (DC through the continuation
    cascade, TRAN with the MOSFETs in the time loop)."""
    import io
    from ohmnivore_amd import cli
    cli.simulate(be, chain_netlist(5, "dc"), io.StringIO())  # first-use setup
    out = {}
    for stages in (10, 50, 100, 200):
        r = {}
        for analysis in ("dc", "tran"):
            t0 = time.perf_counter()
            res = cli.simulate(be, chain_netlist(stages, analysis), io.StringIO())
            ms = 1e3 * (time.perf_counter() - t0)
            a = res[0][1]
            if analysis == "dc":
                r["dc"] = {"ms": ms, "stage": a.stage, "newton_iterations": a.newton_iterations}
            else:
                r["tran"] = {"ms": ms, "steps": a.result.steps_accepted,
                             "newton_iterations": a.result.linear_solves}
        out[f"stages_{stages}"] = r
        log(f"inverter chain {stages}: dc {r['dc']['ms']:.0f} ms ({r['dc']['stage']}), "
            f"tran {r['tran']['ms']:.0f} ms ({r['tran']['steps']} steps)")
    return out


def rc_profile(args, be, N, ir, solver) -> dict:
    """The RC-grid transient alone for rocprofv3 kernel traces: ISAI FP64,
    --tran-steps fixed 1 ps BE steps from the DC operating point (hipGraphs
    are off in this mode, so every launch is traced)."""
    from ohmnivore_amd import transient as T
    ckt = ir.rc_grid(512, 512, 1024, seed=3)
    mna = solver.compile(ckt)
    dc = solver.GpuSolver(be, precond=N.PREC_AUTO,
                          params=solver.bicgstab_params(tolerance=1e-12, max_iterations=100_000))
    x0 = dc.solve_real(mna.g, mna.b_dc)
    src, _ = T.source_table(ckt, mna)
    tr = T.Transient(be, mna, src)
    prm = solver.bicgstab_params(tolerance=1e-9, max_iterations=20_000)
    t0 = time.perf_counter()
    r = tr.run(1e-12, args.tran_steps * 1e-12, method="be", x0=x0, params=prm,
               precond=N.PREC_ISAI)
    dt = time.perf_counter() - t0
    return {"workload": "rc_grid_512 ISAI FP64 (graphs off)", "steps": r.steps_accepted,
            "ms_per_step": 1e3 * dt / r.steps_accepted,
            "bicgstab_iters_per_step": r.bicgstab_iterations / r.steps_accepted}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true",
                    help="skip the oracle run (parity of the headline then unchecked)")
    ap.add_argument("--only", choices=["headline", "all", "ac"], default="all",
                    help="headline/ac: the AC sweep only")
    ap.add_argument("--tran-steps", type=int, default=1000)
    ap.add_argument("--mc-instances", type=int, default=8)
    ap.add_argument("--mc-steps", type=int, default=50)
    ap.add_argument("--ac-circuits", type=int, default=256)
    ap.add_argument("--profile", action="store_true",
                    help="rocprofv3 runs: the diode-mesh workload only, no CPU work, no hipGraphs")
    ap.add_argument("--profile-rc", action="store_true",
                    help="rocprofv3 runs: the RC-grid transient only (ISAI FP64, --tran-steps "
                         "steps), no hipGraphs")
    ap.add_argument("--roofline-only", action="store_true",
                    help="only the roofline kernels: two AC sweeps, 200 SpMVs on the mesh matrix "
                         "and, for counter calibration, 20 copies of a 64 MiB vector")
    ap.add_argument("--spmv-large-only", action="store_true",
                    help="only the 335 MB SpMV (FP64, FP32), its stream-read probe and the copy "
                         "calibration, for rocprofv3 runs")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: the backend has no multi-GPU path (asked for "
              f"--gpus {args.gpus}).", file=sys.stderr)
        return 2

    from ohmnivore_amd import _native as N
    from ohmnivore_amd import circuit as ir
    from ohmnivore_amd import solver

    be = solver.HipBackend()
    if args.roofline_only:
        return roofline_only(args, be, N, ir, solver)
    if args.spmv_large_only:
        return spmv_large_only(args, be, N, ir, solver)
    if args.profile_rc:
        be.set_graphs(False)
        print(json.dumps(rc_profile(args, be, N, ir, solver)))
        return 0
    if args.profile:
        be.set_graphs(False)
        args.steps = min(args.steps, 2)
        print(json.dumps(mesh_newton(args, be, N, ir, solver)))
        return 0
    h, mnas, freqs = ac_headline(args, be, N, ir, solver)
    workloads = {}
    if args.only == "all":
        # a workload that fails is recorded with its error; the headline line
        # is printed regardless
        def run(name, fn):
            try:
                workloads[name] = fn()
            except Exception as e:  # noqa: BLE001 - recorded in the JSON line
                workloads[name] = {"error": f"{type(e).__name__}: {e}"}
                log(f"{name}: FAILED: {e}")

        run("rlc_ac_256x101_one_system_per_call", lambda: ac_plain(be, N, ir, solver, mnas, freqs))
        plain = workloads["rlc_ac_256x101_one_system_per_call"]
        if "error" not in plain:
            plain_ms = plain["device_ms_per_unknown"] * h["unknowns_per_sweep"]
            h["batched_vs_one_system_per_call_device"] = (
                plain_ms / sum(h["device_ms_per_sweep"].values()))
        del mnas
        run("diode_mesh_256", lambda: mesh_newton(args, be, N, ir, solver))
        run("resistor_ladder_10k", lambda: ladder(be, N, ir, solver, check=True))
        run("rc_grid_512", lambda: rc_grid(be, N, ir, solver, args.tran_steps, args.mc_instances,
                                           args.mc_steps))
        run("amplifier_adjoint_4096", lambda: adjoint(be, N, ir, solver))
        run("inverter_chain_netlists", lambda: inverter_chains(be))
    checked = h["parity"] is not None
    line = {
        "metric": "solves/sec at 1e-6 rel err vs FP64 LU; SpMV HBM GB/s; batched-vs-plain speedup",
        "value": h["value"],
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": h["ms_per_step"],
        "higher_is_better": True,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "complex128",
        "data": "synthetic (seeded SplitMix64 circuits)",
        "config": {
            "workload": f"AC sweep: {h['circuits']} uneven RLC ladders (1k-64k sections) x "
                        f"{h['frequencies']} log-spaced frequencies 1 Hz-1 GHz = "
                        f"{h['systems']:,} complex systems, {h['unknowns_per_sweep']:,} unknowns "
                        f"per sweep; banded LU with partial pivoting after RCM, symbolic plan "
                        f"resident in HBM",
            "rows_per_part": h["rows_per_part"],
            "parity_checked": checked,
        },
        "roofline": h["roofline"],
        "cpu_baseline": h["cpu_baseline"],
        "parity": h["parity"],
        "headline_detail": {k: v for k, v in h.items()
                            if k not in ("roofline", "cpu_baseline", "parity", "value")},
        "workloads": workloads or None,
    }
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
