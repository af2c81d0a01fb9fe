"""The reference's own float64 primitives on the host (stereonet.py:29-205):
fsqrt (Newton iteration from a power-of-4 seed), fexp (range reduction +
15-term Taylor series), flog (atanh series), fcos (range reduction +
11-term series), fmin / fmax / fclamp.  The small host-side computations of
the 3D output (camera poses, learning-rate constants, the vertex-grid step)
use them so that their values equal the reference's bit for bit; the device
code has the same steps in csrc/refmath.h."""

LN2 = 0.6931471805599453


def fsqrt(x):
    if x <= 0.0:
        return 0.0
    y = x
    if y > 4.0:
        s = 1.0
        while y > 4.0:
            y *= 0.25
            s *= 2.0
        y = s
    elif y < 0.25:
        s = 1.0
        while y < 0.25:
            y *= 4.0
            s *= 0.5
        y = s
    else:
        y = 1.0
    for _ in range(24):
        y = 0.5 * (y + x / y)
    return y


def fexp(x):
    if x > 700.0:
        return 1.0e308
    if x < -700.0:
        return 0.0
    k = int(x / LN2 + (0.5 if x >= 0.0 else -0.5))
    r = x - k * LN2
    term = 1.0
    total = 1.0
    for n in range(1, 16):
        term = term * r / n
        total += term
    if k >= 0:
        for _ in range(k):
            total *= 2.0
    else:
        for _ in range(-k):
            total *= 0.5
    return total


def flog(x):
    if x <= 0.0:
        return -1.0e308
    k = 0
    while x > 1.5:
        x *= 0.5
        k += 1
    while x < 0.5:
        x *= 2.0
        k -= 1
    z = (x - 1.0) / (x + 1.0)
    z2 = z * z
    total = 0.0
    term = z
    n = 1
    while n < 40:
        total += term / n
        term *= z2
        n += 2
    return 2.0 * total + k * LN2


def fcos(x):
    two_pi = 6.283185307179586
    while x > 3.141592653589793:
        x -= two_pi
    while x < -3.141592653589793:
        x += two_pi
    x2 = x * x
    term = 1.0
    total = 1.0
    for n in range(1, 12):
        term = term * (-x2) / ((2 * n - 1) * (2 * n))
        total += term
    return total


def fmax(a, b):
    return a if a > b else b


def fmin(a, b):
    return a if a < b else b


def fclamp(x, lo, hi):
    if x < lo:
        return lo
    if x > hi:
        return hi
    return x
