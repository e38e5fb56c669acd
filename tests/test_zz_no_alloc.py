"""The allocation-free stable-dt entry point against cpi_stable_dt."""
import numpy as np
import pytest

from cppimpact_amd import _lib as L
from cppimpact_amd import workloads as W

pytestmark = pytest.mark.gpu


def test_stable_dt_noalloc_equals_stable_dt_twice_second_smaller():
    """Two calls in a row on two contexts, the second on a smaller mesh:
    the scratch keeps the first call's minima, so a stale partial would show;
    the results equal cpi_stable_dt's bit for bit (minima are exact)."""
    w1 = W.c0()
    x2, c2 = L.box_mesh(4, 4, 4, 0.1, 0.1, 0.1, origin=(0.0, 0.0, 1e-4), jitter=0.05, seed=3)
    w2 = W.Workload("small", x2, c2, 10 * W.E * 0.025, (0.0, 0.0, -1.0), 10.0, 10)
    for w in (w1, w2):
        d = w.dynamics()
        d.init(0.0)
        d.step(7)
        d.synchronize()
        a = d.stable_dt()
        b = d.stable_dt_noalloc()
        assert a == b, (a, b)
        assert np.isfinite(b[0]) and b[0] > 0 and np.isfinite(b[1])
        d.close()
    # a call before cpi_init is refused
    d = w2.dynamics()
    with pytest.raises(L.CpiError) as ex:
        d.stable_dt_noalloc()
    assert ex.value.code == -7
    d.close()
