"""The rounding step of the FP64 distance path (ExactDistanceNear in
csrc/ee_exact.h), checked in exact integer arithmetic on the CPU.

On the device the floor of the root estimate r may be m - 1, m or m + 1
(m = floor(sqrt(s))) when sqrt(s) lies within 1/4 of an integer; the kernel
then rounds with rem = s - r^2 and
  EUC_2D:  r + (rem > r)
  CEIL_2D: r + (rem > 0) + (rem > 2r + 1)
with no step moving r onto m.  This test takes every r the estimate can give
and compares the result with the reference's rules (src/common/distance.cpp:
44-52: EUC_2D adds one when s - m^2 > m, CEIL_2D when s != m^2)."""
from __future__ import annotations

import math
import random

EUC_2D, CEIL_2D = 0, 1
S_MAX = 2 * ((1 << 25) - 1) ** 2  # dx, dy below 2^25


def reachable_floors(s: int) -> list[int]:
    """floor(estimate) for every estimate within 1/4 of sqrt(s)."""
    m = math.isqrt(s)
    out = [m]
    if m >= 1 and 16 * s < (4 * m + 1) ** 2:  # sqrt(s) < m + 1/4
        out.append(m - 1)
    if 16 * s > (4 * m + 3) ** 2:  # sqrt(s) > m + 3/4
        out.append(m + 1)
    return out


def kernel_rounding(s: int, r: int, dtype: int) -> int:
    rem = s - r * r
    if dtype == EUC_2D:
        return r + (rem > r)
    return r + (rem > 0) + (rem > 2 * r + 1)


def reference_rounding(s: int, dtype: int) -> int:
    m = math.isqrt(s)
    if dtype == EUC_2D:
        return m + (s - m * m > m)
    return m + (s != m * m)


def adversarial_sums() -> list[int]:
    """Squares, their neighbours and the EUC_2D half-way points, at the
    bottom and the top of the range, plus random sums."""
    out = set(range(0, 4096))
    rng = random.Random(7)
    tops = [math.isqrt(S_MAX) - i for i in range(0, 3000)]
    mids = [rng.randrange(1, math.isqrt(S_MAX)) for _ in range(3000)]
    for m in tops + mids + list(range(1, 300)):
        for s in (m * m - 1, m * m, m * m + 1, m * m + m - 1, m * m + m, m * m + m + 1,
                  m * m + 2 * m - 1, m * m + 2 * m):
            if 0 <= s <= S_MAX:
                out.add(s)
    out.update(rng.randrange(0, S_MAX + 1) for _ in range(20000))
    return sorted(out)


def test_every_reachable_floor_rounds_like_the_reference():
    checked = {0: 0, -1: 0, 1: 0}
    for s in adversarial_sums():
        m = math.isqrt(s)
        for r in reachable_floors(s):
            for dtype in (EUC_2D, CEIL_2D):
                assert kernel_rounding(s, r, dtype) == reference_rounding(s, dtype), (s, r, dtype)
            checked[r - m] += 1
    # all three cases occur, so each formula branch is exercised
    assert all(v > 100 for v in checked.values()), checked


def test_a_floor_two_away_would_be_wrong():
    """The 1/4 bound matters: r = m - 2 or m + 2 gives wrong answers, so the
    check above would catch an estimate that is off by more."""
    s = 1000 ** 2 + 1
    m = math.isqrt(s)
    assert any(kernel_rounding(s, r, t) != reference_rounding(s, t)
               for r in (m - 2, m + 2) for t in (EUC_2D, CEIL_2D))
