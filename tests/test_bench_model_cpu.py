"""The headline's byte model in bench.py: how many partitioned waves hold
each circuit's frequencies (ac_wave_counts), against closed forms. With
circuits longest first and 64-lane waves filled across circuit boundaries,
circuit k of that order occupies lanes [F k, F k + F) of one continuous
stream when F >= 64 (no wave can then hold three circuits), so it touches
floor((F k + F - 1) / 64) - floor(F k / 64) + 1 waves; with F < 32 two
circuits share each wave."""
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

import bench  # noqa: E402


def _sizes(n=256):
    # distinct sizes in no particular order (the model sorts them)
    return [1000 + (37 * k) % 997 * 13 for k in range(n)]


def test_wave_counts_continuous_stream():
    sizes = _sizes()
    F = 101
    wc = bench.ac_wave_counts(sizes, F, packed=True)
    order = sorted(range(len(sizes)), key=lambda c: -sizes[c])
    for k, c in enumerate(order):
        assert wc[c] == (F * k + F - 1) // 64 - (F * k) // 64 + 1, (k, c)
    # 25,856 systems in 404 waves: the distinct waves of the stream
    assert -(-len(sizes) * F // 64) == 404
    assert sorted(set(wc)) == [2, 3]


def test_wave_counts_two_circuits_per_wave_and_unpacked():
    sizes = _sizes(9)
    assert bench.ac_wave_counts(sizes, 5, packed=True) == [1] * 9
    assert bench.ac_wave_counts(sizes, 101, packed=False) == [2] * 9
    assert bench.ac_wave_counts(sizes + [0], 101, packed=True)[-1] == 0


def test_band_bytes_follow_wave_counts():
    sizes = _sizes(16)
    F, R = 101, 2048
    wc = bench.ac_wave_counts(sizes, F, packed=True)
    by = bench.ac_bytes(sizes, F, R)
    x = sum(sizes) * 16 * F
    ck = 0
    for n in sizes:
        p = -(-n // R)
        ck += p * -(-(-(-n // p) - 1) // 8)
    assert by["part_solve"] == x + ck * 64 * F + 64 * sum(n * w for n, w in zip(sizes, wc))
