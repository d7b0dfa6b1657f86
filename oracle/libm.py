"""Python restatement of the one libm routine the device replicates
bit-for-bit -- test infrastructure only.

x86-64 glibc (>= 2.35) computes hypot(x, y) with the corrected square root
of C. F. Borges, "An improved algorithm for hypot(a, b)" (2019), without
fused multiply-add, rescaling by powers of two far from 1.  The reference's
obscurations call it (src/obscuration.cpp:32-34, 80-83); the GPU's
ref_hypot (batoid_amd/csrc/bt_device.h) follows the same steps.  Python
floats are IEEE binary64 with round-to-nearest and math.sqrt is correctly
rounded, so this function reproduces libm exactly; tests pin it against
libm itself (tests/test_raygen_oracle.py).
"""
import math


def _kernel(ax, ay):
    h = math.sqrt(ax * ax + ay * ay)
    if h <= 2.0 * ay:
        delta = h - ay
        t1 = ax * (2.0 * delta - ax)
        t2 = (delta - 2.0 * (ax - ay)) * delta
    else:
        delta = h - ax
        t1 = 2.0 * delta * (ax - 2.0 * ay)
        t2 = (4.0 * delta - ay) * ay + delta * delta
    return h - (t1 + t2) / (2.0 * h)


def hypot(x, y):
    if not (math.isfinite(x) and math.isfinite(y)):
        if math.isinf(x) or math.isinf(y):
            return math.inf
        return x + y
    x, y = abs(x), abs(y)
    ax, ay = (y, x) if x < y else (x, y)
    scale, eps = 2.0 ** -600, 2.0 ** -54
    if ax > 2.0 ** 511:
        if ay <= ax * eps:
            return ax + ay
        return _kernel(ax * scale, ay * scale) / scale
    if ay < 2.0 ** -511:
        if ax >= ay / eps:
            return ax + ay
        return _kernel(ax / scale, ay / scale) * scale
    if ay <= ax * eps:
        return ax + ay
    return _kernel(ax, ay)
