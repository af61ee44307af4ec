"""The tree/COSS statistic's runs of empty bins (reference tree.cu:272-275,
387-394 add (0 - avg)^2 once per bin, in FP64, in bin order).

The library adds a run of g equal terms in a few steps per binade
(msj_probe_repeat_add, host code); it must equal g sequential IEEE
additions bit for bit, which Python's float addition performs.  The cases
sit where a shortcut would go wrong: sums past 2^53 where every step rounds,
exact half-ulp terms (ties to even), terms below half an ulp, sums starting
just below a power of two, and long runs that cross several binades.
"""
import math
import random

import pytest


def naive(s, c, g):
    for _ in range(g):
        s = s + c
    return s


def _cases():
    rnd = random.Random(2024)
    out = []
    for it in range(1500):
        e = rnd.randint(40, 64)
        u = math.ldexp(1.0, e - 53)
        s = math.ldexp(float(rnd.getrandbits(52) | (1 << 52)), e - 53)
        kind = it % 7
        if kind == 0:
            c = float(rnd.randint(1, 100000))
        elif kind == 1:
            c = u * rnd.randint(0, 40) + u / 2          # exact tie
        elif kind == 2:
            c = u * rnd.randint(0, 40) + u / 2 + u / 8
        elif kind == 3:
            c = u / 2                                     # stalls at a tie
        elif kind == 4:
            c = u / 4                                     # below half an ulp
        elif kind == 5:
            s = math.ldexp(1.0, e) - u * rnd.randint(0, 40)  # just below 2^e
            c = u * rnd.randint(1, 9) + (u / 2 if rnd.random() < 0.5 else 0.0)
        else:
            s = 0.0
            c = float(rnd.randint(1, 1 << 20)) ** 2
        g = rnd.randint(0, 3000)
        out.append((s, c, g))
    return out


def test_repeat_add_equals_sequential_sum(msj_mod):
    bad = [(s, c, g) for s, c, g in _cases() if msj_mod.probe_repeat_add(s, c, g) != naive(s, c, g)]
    assert not bad, bad[:5]


@pytest.mark.parametrize("s,c,g", [
    (2.0 ** 53 - 64, 3.0, 200000),              # crosses 2^53: steps start rounding
    (2.0 ** 60 - 2.0 ** 10, 2.0 ** 7 + 64, 100000),  # exact ties past 2^60
    (1e15, 1e6, 300000),                        # ends past 2^53
])
def test_repeat_add_long_runs(s, c, g, msj_mod):
    assert msj_mod.probe_repeat_add(s, c, g) == naive(s, c, g)


def test_repeat_add_planted_shortcut_is_caught(msj_mod):
    """s + g*c (one rounding) differs from the sequential sum on the tie
    cases above: the comparison can tell the two apart."""
    s, c, g = 2.0 ** 60 - 2.0 ** 10, 2.0 ** 7 + 64, 100000
    assert s + g * c != naive(s, c, g)
    assert msj_mod.probe_repeat_add(s, c, g) == naive(s, c, g)
