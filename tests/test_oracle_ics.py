"""Pins the oracle's initial conditions and random streams.

  - std::mt19937_64: the C++ standard's known answer (10000th output of the
    default seed 5489 is 9981545732273789042, [rand.predef]);
  - SingleSeed (include/openpfc/kernel/simulation/initial_conditions/single_seed.hpp:51-83):
    value rho + 12*amp at the origin (all six cosines are 1), 0.0 outside the
    sphere, the six <110>/sqrt(2) wave vectors (|q| = 1, so the pattern has
    period 2*pi*sqrt(2) along x);
  - the counter-based noise stream: its mixing function against the published
    splitmix64 known answers (Vigna's generator seeded with 0 starts
    0xe220a8397b1dcdaf, 0x6e789e6aa1b965f4; seeded with 1234567 it starts
    6457827717110365317, 3203168211198807973, 9817491932198370423,
    4593380528125082431, 16408922859458223821), each value u_i of the field
    against its definition (splitmix64(splitmix64(seed) + i) >> 11) * 2^-53
    built from that pinned function, then range, mean and independence of
    the seed.
"""
import math

import numpy as np

from oracle import oracle as O


def test_mt19937_64_standard_known_answer():
    assert O.mt64_nth(5489, 10000) == 9981545732273789042


def test_uniform_noise_draw_rule():
    a = O.uniform_noise((4, 1, 1), 5489, 0.0, 1.0).reshape(-1)
    # first output of mt19937_64(5489) is 14514284786278117030
    assert a[0] == (14514284786278117030 >> 11) * 2.0 ** -53


def test_single_seed_reference_semantics():
    n = (40, 40, 40)
    dx = 1.0
    origin = (-20.0, -20.0, -20.0)
    a = O.ic_single_seed(n, origin, (dx, dx, dx), rho=-0.1, amp=0.05, radius=10.0)
    # grid point i = 20 is x = 0: all cosines equal 1
    assert abs(a[20, 20, 20] - (-0.1 + 12 * 0.05)) <= 1e-15
    # outside the sphere: exactly 0.0
    assert a[0, 0, 0] == 0.0 and a[20, 20, 31] == 0.0
    # inside, the value is rho + 2 amp sum cos(q.x); at (x, 0, 0):
    x = 3.0
    s = 1 / math.sqrt(2)
    expect = -0.1 + 2 * 0.05 * (4 * math.cos(s * x) + 2 * 1.0)
    assert abs(a[20, 20, 23] - expect) <= 1e-15


def test_single_seed_masked_keeps_outside():
    n = (16, 16, 16)
    a = np.full((16, 16, 16), -0.2)
    O.ic_single_seed_masked(a, (-8.0,) * 3, (1.0,) * 3, -0.2, 0.1, 4.0, 1)
    assert a[0, 0, 0] == -0.2
    assert abs(a[8, 8, 8] - (-0.2 + 12 * 0.1)) <= 1e-15
    b = np.full((16, 16, 16), -0.2)
    O.ic_single_seed_masked(b, (-8.0,) * 3, (1.0,) * 3, -0.2, 0.1, 4.0, 0)
    assert b[0, 0, 0] == 0.0
    ref = O.ic_single_seed(n, (-8.0,) * 3, (1.0,) * 3, -0.2, 0.1, 4.0)
    assert np.array_equal(b, ref)


def test_counter_noise_stream():
    a = np.zeros(1 << 16)
    O.add_counter_noise(a, 1234, -0.01, 0.01)
    assert a.min() >= -0.01 and a.max() < 0.01
    assert abs(a.mean()) < 1e-4
    b = np.zeros(1 << 16)
    O.add_counter_noise(b, 1235, -0.01, 0.01)
    assert abs(np.corrcoef(a, b)[0, 1]) < 0.02
    c = np.zeros(1 << 16)
    O.add_counter_noise(c, 1234, -0.01, 0.01)
    assert np.array_equal(a, c)


GOLDEN_GAMMA = 0x9E3779B97F4A7C15
MASK64 = (1 << 64) - 1


def test_splitmix64_published_known_answers():
    # the n-th output of a generator seeded with S is mix(S + (n-1)*gamma)
    assert O.splitmix64(0) == 0xE220A8397B1DCDAF
    assert O.splitmix64(GOLDEN_GAMMA) == 0x6E789E6AA1B965F4
    want = [6457827717110365317, 3203168211198807973, 9817491932198370423,
            4593380528125082431, 16408922859458223821]
    got = [O.splitmix64((1234567 + i * GOLDEN_GAMMA) & MASK64) for i in range(5)]
    assert got == want


def test_counter_noise_values_follow_their_definition():
    n = 4099
    for seed, lo, hi in ((1234, -0.01, 0.01), (0, 0.0, 1.0), (2 ** 63 + 5, -3.0, 2.0)):
        a = np.zeros(n)
        O.add_counter_noise(a, seed, lo, hi)
        s0 = O.splitmix64(seed)
        # the last indices wrap the 64-bit counter when s0 is near 2^64
        for i in list(range(40)) + [n - 1]:
            u = (O.splitmix64((s0 + i) & MASK64) >> 11) * 2.0 ** -53
            assert a[i] == lo + (hi - lo) * u, (seed, i)
    # base values are added to, not overwritten
    b = np.full(16, -0.2)
    O.add_counter_noise(b, 1234, -0.01, 0.01)
    c = np.zeros(16)
    O.add_counter_noise(c, 1234, -0.01, 0.01)
    assert np.array_equal(b, -0.2 + c)
