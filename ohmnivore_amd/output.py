"""CSV result writers (src/output.rs:7-93).

Numbers are written the way Rust's ``{}`` formats an f64: the shortest
digit string that reads back to the same double, in positional notation
(never an exponent), "NaN", "inf" and "-inf" for the specials. AC phasors
are written as magnitude (hypot of the parts) and phase in degrees
(atan2(im, re) * 180/pi).
"""
from __future__ import annotations

import math
from typing import TextIO

import numpy as np

_RAD_TO_DEG = 180.0 / math.pi


def fmt(x: float) -> str:
    """f64 Display text (shortest round-trip digits, positional)."""
    x = float(x)
    if math.isnan(x):
        return "NaN"
    if math.isinf(x):
        return "inf" if x > 0 else "-inf"
    return np.format_float_positional(x, unique=True, trim="-")


def mag_phase(z: complex) -> tuple[float, float]:
    """(|z|, arg z in degrees) as num_complex's norm() and arg().to_degrees()."""
    z = complex(z)
    return float(np.hypot(z.real, z.imag)), math.atan2(z.imag, z.real) * _RAD_TO_DEG


def write_dc_csv(result, out: TextIO) -> None:
    """write_dc_csv (src/output.rs:16-25): ``Variable,Value`` then V(node)
    and I(branch) rows."""
    out.write("Variable,Value\n")
    for name, v in result.node_voltages:
        out.write(f"V({name}),{fmt(v)}\n")
    for name, i in result.branch_currents:
        out.write(f"I({name}),{fmt(i)}\n")


def write_ac_csv(result, out: TextIO) -> None:
    """write_ac_csv (src/output.rs:35-61): one row per frequency, a
    magnitude and a phase column per node voltage and branch current."""
    head = ["Frequency"]
    for name, _ in result.node_voltages:
        head.append(f"V({name})_mag,V({name})_phase_deg")
    for name, _ in result.branch_currents:
        head.append(f"I({name})_mag,I({name})_phase_deg")
    out.write(",".join(head) + "\n")
    series = [w for _, w in result.node_voltages] + [w for _, w in result.branch_currents]
    for fi, f in enumerate(result.frequencies):
        row = [fmt(f)]
        for w in series:
            m, p = mag_phase(w[fi])
            row.append(f"{fmt(m)},{fmt(p)}")
        out.write(",".join(row) + "\n")


def write_tran_csv(result, out: TextIO) -> None:
    """write_tran_csv (src/output.rs:71-93): ``time`` then one column per
    node voltage and branch current."""
    head = ["time"] + [f"V({n})" for n, _ in result.node_voltages] + \
        [f"I({n})" for n, _ in result.branch_currents]
    out.write(",".join(head) + "\n")
    series = [w for _, w in result.node_voltages] + [w for _, w in result.branch_currents]
    for ti, t in enumerate(result.times):
        out.write(",".join([fmt(t)] + [fmt(w[ti]) for w in series]) + "\n")
