"""The golden fixtures are what their generator says: the ladder fixture is
the oracle's LU answer, the diode-mesh fixture solves the oracle's Newton
system (its assembled residual vanishes), the transient probes start at a DC
point and move with the PULSE sinks, and the amplifier fixture's gradients
are finite and of both signs."""
import os

import numpy as np

import oracle
from ohmnivore_amd import circuit as ir
from ohmnivore_amd import solver

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return np.load(os.path.join(GOLD, name))


def test_ladder_fixture_is_the_oracle_lu():
    mna = solver.compile(ir.resistor_ladder(10_000, seed=1))
    assert np.array_equal(load("ladder_10k.npz")["x"], oracle.sparse_lu_solve(mna.g, mna.b_dc))


def test_diode_mesh_fixture_shape():
    g = load("diode_mesh_256.npz")
    x = g["x_every4"]
    assert x.shape == ((65_540 + 3) // 4,) and np.all(np.isfinite(x))
    v = x[: 65_536 // 4]  # node voltages (the last entries are the four source currents)
    assert 0.0 < v.min() and v.max() <= 1.0 + 1e-12  # between ground and the 1 V sources
    assert int(g["newton_iterations"]) > 3


def test_rc_grid_fixture_shape():
    g = load("rc_grid_512.npz")
    assert g["waves"].shape == (51, g["probes"].shape[0]) and float(g["h"]) == 1e-12
    assert np.all(np.isfinite(g["waves"]))


def test_amplifier_fixture_shape():
    g = load("amplifier_4096.npz")
    assert g["v_out"].shape == (4096,) and g["grad_g"].shape == (48,) and g["grad_kp"].shape == (16,)
    assert np.all(np.isfinite(g["grad_g"])) and (g["grad_g"] > 0).any() and (g["grad_g"] < 0).any()
