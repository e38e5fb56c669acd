"""The FP64 return map's elastic pre-check (device.h, surely_elastic) skips
the square root for trial stresses whose squared deviator norm q2 is below
the squared yield radius times (1 - 64 eps).  The claim that this changes
no decision: every stress it passes is also left elastic by the square-root
test sqrt(q2) - k23 t > 0, whether the compiler fuses that subtraction
(exact product, one rounding) or not.  Checked here on random stresses at
and around the tie, in FP64 and (for the margin's sake) FP32."""
from fractions import Fraction as F

import numpy as np
import pytest

K23 = 0.81649658092772603273  # sqrt(2/3)


def _violations(dt, n, seed):
    rng = np.random.default_rng(seed)
    eps = np.finfo(dt).eps
    c = dt(1) - dt(64) * dt(eps)
    k23 = dt(K23)
    bad = 0
    for _ in range(n):
        t = dt(rng.uniform(1e-3, 1e3)) * dt(10.0) ** dt(rng.integers(-3, 9))
        rad = dt(k23 * t)
        r2 = float(rad) ** 2
        if rng.random() < 0.5:
            q2 = dt(r2 * (1 + rng.uniform(-200, 200) * float(eps)))
        else:
            q2 = np.nextafter(dt(r2), dt(np.inf) if rng.random() < 0.5 else dt(0))
            for _ in range(int(rng.integers(0, 40))):
                q2 = np.nextafter(q2, dt(np.inf) if rng.random() < 0.5 else dt(0))
        pre = q2 <= dt(dt(rad * rad) * c)
        s = dt(np.sqrt(q2))
        fused = F(float(s)) - F(float(k23)) * F(float(t)) > 0
        unfused = dt(s - dt(k23 * t)) > 0
        bad += bool(pre and (fused or unfused))
    return bad


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_precheck_never_passes_a_yielding_stress(dt):
    assert _violations(dt, 20000, seed=1) == 0


def test_precheck_catches_a_margin_that_is_too_small():
    """The same check with no margin finds violations, so the test above
    can fail."""
    rng = np.random.default_rng(2)
    k23 = np.float64(K23)
    bad = 0
    for _ in range(20000):
        t = np.float64(rng.uniform(1.0, 10.0))
        rad = np.float64(k23 * t)
        q2 = np.float64(float(rad) ** 2 * (1 + rng.uniform(-4, 4) * 2.2e-16))
        pre = q2 <= np.float64(rad * rad)
        s = np.float64(np.sqrt(q2))
        bad += bool(pre and (F(float(s)) - F(float(k23)) * F(float(t)) > 0))
    assert bad > 0
