"""CPU checks of device numerical methods, emulated exactly on the host.

log_ratio (pcbqs_amd/csrc/pair.h) replaces log(a / b) in the near integrals: the mantissa ratio
reduced into [1/sqrt2, sqrt2], 2 atanh(s) with s = (ma - mb) / (ma + mb) as a degree-19 odd series
(Horner in s^2 with fused multiply-adds), plus the exponent difference times ln 2 in two parts.
Here every rounding step is reproduced in exact rational arithmetic (fused multiply-add = one
rounding of the exact a b + c; IEEE division correctly rounded) and compared with a 50-digit
logarithm. A dropped series term, a wrong ln 2 split or a mis-set reduction threshold shows up as
errors of many ulp. The device kernels themselves are held to the oracle by the GPU fill tests."""
import math
import random
from decimal import Decimal, getcontext
from fractions import Fraction

SQRT2 = 1.41421356237309504880
LN2_HI, LN2_LO = 6.93147180369123816490e-01, 1.90821492927058770002e-10


def fma(a, b, c):
    return float(Fraction(a) * Fraction(b) + Fraction(c))


def log_ratio(a, b, terms=(19, 17, 15, 13, 11, 9, 7, 5, 3), ln2=(LN2_HI, LN2_LO), sqrt2=SQRT2):
    ma, ea = math.frexp(a)
    mb, eb = math.frexp(b)
    e = ea - eb
    if ma > sqrt2 * mb:
        mb *= 2.0
        e += 1
    elif sqrt2 * ma < mb:
        ma *= 2.0
        e -= 1
    s = float(Fraction(ma - mb) / Fraction(ma + mb))
    z = s * s
    p = 1.0 / terms[0]
    for k in terms[1:]:
        p = fma(p, z, 1.0 / k)
    t = 2.0 * s
    r = fma(t * z, p, t)
    de = float(e)
    return fma(de, ln2[0], fma(de, ln2[1], r))


def worst_ulp(fn, cases):
    getcontext().prec = 50
    worst = 0.0
    for a, b in cases:
        exact = (Decimal(a) / Decimal(b)).ln()
        if exact == 0:
            continue
        err = abs(Decimal(fn(a, b)) - exact) / Decimal(math.ulp(float(exact)))
        worst = max(worst, float(err))
    return worst


def cases(seed=7, n=1500):
    rng = random.Random(seed)
    out = []
    for _ in range(n):
        a = math.exp(rng.uniform(-40, 40))
        b = a * math.exp(rng.uniform(-3, 3)) if rng.random() < 0.7 else math.exp(rng.uniform(-40, 40))
        out.append((a, b))
    for _ in range(n // 3):                       # ratios within 1e-6 of 1 (tiny logs)
        a = rng.uniform(0.1, 10)
        out.append((a, a * (1 + rng.uniform(-1e-6, 1e-6))))
    for x in (SQRT2, 1 / SQRT2, 2.0, 0.5, 1.0 + 2 ** -52):   # the reduction's edges
        out.append((x, 1.0))
        out.append((1.0, x))
    return out


def test_log_ratio_is_within_two_ulp():
    assert worst_ulp(log_ratio, cases()) <= 2.0


def test_planted_mistakes_in_the_method_are_caught():
    c = cases(seed=11, n=300)
    # a series cut one term short, the tail of ln 2 dropped, a reduction threshold off
    assert worst_ulp(lambda a, b: log_ratio(a, b, terms=(17, 15, 13, 11, 9, 7, 5, 3)), c) > 2.0
    assert worst_ulp(lambda a, b: log_ratio(a, b, ln2=(LN2_HI, 0.0)), c) > 2.0
    assert worst_ulp(lambda a, b: log_ratio(a, b, sqrt2=1.9), c) > 2.0
