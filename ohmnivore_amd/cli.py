"""Netlist front end: parse -> compile -> every requested analysis on the GPU
-> CSV on stdout (the reference's command line, src/main.rs:73-205).

    python -m ohmnivore_amd circuit.spice [--stats]

Analyses run in netlist order: ``.DC``/``.OP`` through dc.run (the Newton
continuation cascade for nonlinear circuits), ``.AC`` through the
multi-circuit AC sweep (banded direct or batched BiCGSTAB per circuit
shape), ``.TRAN`` through the adaptive TRAP/BE stepper. Every solve runs on
the GPU; there is no CPU solver behind ``--cpu``, which is refused.
"""
from __future__ import annotations

import argparse
import sys
import time
from dataclasses import dataclass, field
from typing import TextIO

import numpy as np

from . import _native as N
from . import dc, netlist, output, transient
from . import solver as S

# Relative residual of every AC system and transient step. The transient's
# is tighter: TRAP carries each step's error forward undamped, and in an
# inductor's node voltage (v_n = -v_n-1 + (2L/h)(i_n - i_n-1)) amplified by
# 2L/h, so per-step residuals near 1e-13 are what keeps an RL circuit's
# samples within 1e-6 of an LU-based run.
AC_TOLERANCE = 1e-10
TRAN_TOLERANCE = 1e-13


@dataclass
class AcAnalysis:
    """AcResult (src/analysis/mod.rs): per-node / per-branch phasors."""
    frequencies: np.ndarray
    node_voltages: list
    branch_currents: list


@dataclass
class Stats:
    """Phase timings and counters (Stats, src/stats.rs:9-127)."""
    phases: list = field(default_factory=list)
    newton_iterations: int = 0
    timesteps_accepted: int = 0
    timesteps_rejected: int = 0
    linear_solves: int = 0
    gpu_dispatches: int = 0
    gpu_readbacks: int = 0
    start: float = field(default_factory=time.perf_counter)

    def add_phase(self, name: str, seconds: float) -> None:
        self.phases.append((name, seconds))

    def display(self, err: TextIO) -> None:
        err.write("\n=== Ohmnivore Performance Stats ===\n")
        for name, sec in self.phases:
            err.write(f"  {name:<24} {sec:>8.3f}s\n")
        if self.newton_iterations:
            err.write(f"  Newton iterations:      {self.newton_iterations}\n")
        if self.timesteps_accepted or self.timesteps_rejected:
            err.write(f"  Timesteps:              accepted={self.timesteps_accepted}  "
                      f"rejected={self.timesteps_rejected}\n")
        if self.linear_solves:
            err.write(f"  Linear solves:          {self.linear_solves}\n")
        err.write("  " + "─" * 33 + "\n")
        err.write(f"  Total:                  {time.perf_counter() - self.start:>8.3f}s\n")
        err.write(f"  GPU dispatches:         {self.gpu_dispatches}\n")
        err.write(f"  GPU readbacks:          {self.gpu_readbacks}\n")


def run_ac(backend: S.HipBackend, mna, a: netlist.Ac) -> AcAnalysis:
    """ac::run (src/analysis/ac.rs:25-77) for one circuit."""
    freqs = S.generate_frequencies(a.sweep, a.n_points, a.f_start, a.f_stop)
    res = S.ac_sweep_circuits(backend, [mna], freqs, method="auto",
                              params=S.bicgstab_params(tolerance=AC_TOLERANCE,
                                                       max_iterations=100_000))
    x = res.x[0]
    nn = mna.n_nodes
    nodes = [(name, x[:, i].copy()) for i, name in enumerate(mna.node_names)]
    branches = [(name, x[:, nn + i].copy()) for i, name in enumerate(mna.branch_names)]
    return AcAnalysis(freqs, nodes, branches)


def simulate(backend: S.HipBackend, text: str, out: TextIO | None = None,
             stats: Stats | None = None, newton: dict | None = None) -> list:
    """Parse and compile ``text``, run its analyses in order and write each
    result's CSV to ``out`` (if given). Returns [(analysis, result), ...].
    ``newton``: keyword arguments of solver.newton_params for every DC
    operating point (the .DC/.OP analyses and a transient's start); the
    default is the reference's NewtonParams (abs_tol 1e-4, BiCGSTAB
    tolerance 1e-5). Errors are OhmnivoreError with the reference's message
    texts."""
    newton = dict(newton or {})
    t = time.perf_counter()
    nl = netlist.parse(text)
    if stats:
        stats.add_phase("Parse", time.perf_counter() - t)
    t = time.perf_counter()
    ckt = nl.to_circuit()
    mna = S.compile(ckt)
    if stats:
        stats.add_phase("Compile", time.perf_counter() - t)
    results = []
    for a in nl.analyses:
        t = time.perf_counter()
        if isinstance(a, netlist.Dc):
            r = dc.run(backend, mna, **newton)
            if stats:
                stats.add_phase("DC analysis", time.perf_counter() - t)
                stats.newton_iterations += r.newton_iterations
            if out is not None:
                output.write_dc_csv(r, out)
        elif isinstance(a, netlist.Ac):
            r = run_ac(backend, mna, a)
            if stats:
                stats.add_phase("AC analysis", time.perf_counter() - t)
                stats.linear_solves += len(r.frequencies)
            if out is not None:
                output.write_ac_csv(r, out)
        else:
            r = transient.run(backend, ckt, mna, a.tstep, a.tstop, a.tstart, a.uic,
                              params=S.bicgstab_params(tolerance=TRAN_TOLERANCE,
                                                       max_iterations=100_000),
                              dc_kw=newton)
            if stats:
                stats.add_phase("Transient analysis", time.perf_counter() - t)
                stats.timesteps_accepted += r.result.steps_accepted
                stats.timesteps_rejected += r.result.steps_rejected
                stats.linear_solves += r.result.linear_solves
            if out is not None:
                output.write_tran_csv(r, out)
        results.append((a, r))
    if stats:
        stats.gpu_dispatches, stats.gpu_readbacks = backend.counters()
    return results


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="ohmnivore_amd",
                                 description="GPU circuit simulation (MI355X)")
    ap.add_argument("netlist", help="SPICE netlist file to simulate")
    ap.add_argument("--cpu", action="store_true",
                    help="not available: every solve runs on the GPU")
    ap.add_argument("--stats", action="store_true", help="print performance stats to stderr")
    args = ap.parse_args(argv)
    if args.cpu:
        print("--cpu: there is no CPU solver; every analysis runs on the GPU",
              file=sys.stderr)
        return 2
    try:
        with open(args.netlist, encoding="utf-8") as f:
            text = f.read()
    except OSError as e:
        print(f"Error reading {args.netlist}: {e}", file=sys.stderr)
        return 1
    stats = Stats() if args.stats else None
    try:
        backend = S.HipBackend()
        simulate(backend, text, sys.stdout, stats)
    except N.OhmnivoreError as e:
        print(str(e), file=sys.stderr)
        return 1
    if stats:
        stats.display(sys.stderr)
    return 0
