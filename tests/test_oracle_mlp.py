"""Pins the MLP-surrogate oracle (oracle/mlp_oracle.py) on the CPU:
finite differences for the gradient, exact bfloat16 rounding cases, the
radial return's consistency conditions and its agreement with the C oracle,
and the closed form of the overshoot-scaled training targets."""
import numpy as np

from oracle import mlp_oracle as M
from oracle import pyoracle as O

from common import E, H, NU, SY, plastic_states


def random_theta(rng, scale=1.0):
    W1 = rng.normal(0, scale / np.sqrt(M.N_IN), (M.N_HID, M.N_IN))
    W2 = rng.normal(0, scale / np.sqrt(M.N_HID), (M.N_HID, M.N_HID))
    W3 = rng.normal(0, scale / np.sqrt(M.N_HID), (M.N_OUT, M.N_HID))
    return M.pack(W1, rng.normal(0, 0.3, M.N_HID), W2, rng.normal(0, 0.3, M.N_HID), W3,
                  rng.normal(0, 0.3, M.N_OUT))


def test_parameter_layout_roundtrip():
    rng = np.random.default_rng(0)
    th = rng.normal(size=M.N_PARAMS)
    assert M.N_PARAMS == 5511
    assert np.array_equal(M.pack(*M.unpack(th)), th)


def test_bf16_rounding_cases():
    one = 1.0
    assert M.bf16(one) == 1.0
    assert M.bf16(1.0 + 2.0 ** -8) == 1.0                       # tie -> even
    assert M.bf16(1.0 + 3 * 2.0 ** -8) == 1.0 + 2.0 ** -6       # tie -> even (up)
    assert M.bf16(1.0 + 2.0 ** -8 + 2.0 ** -12) == 1.0 + 2.0 ** -7
    assert M.bf16(-3.0) == -3.0
    x = np.random.default_rng(1).normal(size=10000)
    r = M.bf16(x)
    assert np.all(np.abs(r - x) <= 2.0 ** -8 * np.abs(x) + 1e-300)


def test_silu_derivative_is_the_derivative():
    z = np.linspace(-8, 8, 101)
    h = 1e-6
    fd = (M.silu(z + h) - M.silu(z - h)) / (2 * h)
    assert np.allclose(M.silu_grad(z), fd, rtol=1e-7, atol=1e-9)


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(2)
    th = random_theta(rng)
    x = rng.normal(size=(37, M.N_IN))
    t = rng.normal(size=(37, M.N_OUT))
    loss, g = M.loss_grad(th, x, t)
    assert abs(loss - 0.5 * np.mean(np.sum((M.forward(th, x) - t) ** 2, axis=1))) < 1e-12
    idx = np.concatenate([rng.integers(lo, hi, 12) for lo, hi in (
        (M.OFF_W1, M.OFF_B1), (M.OFF_B1, M.OFF_W2), (M.OFF_W2, M.OFF_B2),
        (M.OFF_B2, M.OFF_W3), (M.OFF_W3, M.OFF_B3), (M.OFF_B3, M.N_PARAMS))])
    fd = M.finite_difference(th, x, t, idx)
    assert np.allclose(g[idx], fd, rtol=1e-5, atol=1e-8 * np.abs(g).max())


def test_rounded_gradient_is_close_to_exact():
    rng = np.random.default_rng(3)
    th = random_theta(rng)
    x = rng.normal(size=(200, M.N_IN))
    t = rng.normal(size=(200, M.N_OUT))
    _, g = M.loss_grad(th, x, t)
    _, gr = M.loss_grad(th, x, t, rounded=True)
    err = np.linalg.norm(gr - g) / np.linalg.norm(g)
    assert 1e-5 < err < 3e-2  # BF16 moves it, but only by its rounding


def test_radial_return_consistency_and_c_oracle():
    rng = np.random.default_rng(4)
    de, so, ep = plastic_states(rng, 400)
    tr = M.trial(de, so, E, NU)
    q = M.overshoot(tr, ep, SY, H)
    s, ep1, dg = M.radial_return(tr, ep, E, NU, SY, H)
    pl = q > 0
    assert pl.sum() > 100 and (~pl).sum() > 10
    # on the yield surface after return, unchanged when elastic
    assert np.abs(M.yield_residual(s[pl], ep1[pl], SY, H)).max() < 1e-8 * SY
    assert np.array_equal(s[~pl], tr[~pl]) and np.array_equal(ep1[~pl], ep[~pl])
    # pressure kept, deviator scaled along itself
    assert np.allclose(s[:, :3].sum(1), tr[:, :3].sum(1), rtol=0, atol=1e-6 * SY)
    mu = E / (2 * (1 + NU))
    assert np.allclose(dg[pl], q[pl] / (2 * mu + 2 * H / 3), rtol=1e-12)
    # the C oracle's restatement gives the same state
    for i in np.flatnonzero(pl)[:50]:
        sc, epc, dgc = O.radial_return(E, NU, SY, H, de[i], so[i], ep[i])
        assert np.allclose(sc, s[i], rtol=1e-12, atol=1e-6)
        assert abs(epc - ep1[i]) < 1e-15 + 1e-12 * ep1[i]
        assert abs(dgc - dg[i]) <= 1e-12 * dg[i]


def test_targets_have_the_closed_form():
    """For linear hardening (s_trial - s_{n+1}) / q = 2mu/(2mu + 2H/3) n and
    Y0 dep / q = sqrt(2/3) Y0 / (2mu + 2H/3)."""
    rng = np.random.default_rng(5)
    de, so, ep = plastic_states(rng, 500)
    tr = M.trial(de, so, E, NU)
    s1, ep1, _ = M.radial_return(tr, ep, E, NU, SY, H)
    smp = np.concatenate([de, so, ep[:, None], s1, ep1[:, None]], axis=1).astype(np.float32)
    x, t = M.sample_targets(smp, E, NU, SY, H)
    assert len(x) > 100
    # recompute the kept set in float32-rounded inputs
    s64 = smp.astype(np.float64)
    tr32 = M.trial(s64[:, :6], s64[:, 6:12], E, NU)
    q32 = M.overshoot(tr32, s64[:, 12], SY, H)
    keep = q32 > M.MIN_OVERSHOOT * SY
    assert keep.sum() == len(x)
    mu = E / (2 * (1 + NU))
    p = tr32[keep, :3].mean(axis=1, keepdims=True)
    n = np.concatenate([tr32[keep, :3] - p, tr32[keep, 3:]], axis=1) / M.dev_norm(tr32[keep])[:, None]
    c = 2 * mu / (2 * mu + 2 * H / 3)
    # float32 samples: s1 carries ~6e-8 Y0 rounding, q >= 1e-3 Y0
    assert np.abs(t[:, :6] - c * n).max() < 2e-3
    assert np.abs(t[:, 6] - M.K23 * SY / (2 * mu + 2 * H / 3)).max() < 2e-3 * abs(t[:, 6]).max()
    assert np.array_equal(x[:, 12], s64[keep, 12])


def test_surrogate_state_of_exact_targets_is_on_the_yield_surface():
    rng = np.random.default_rng(6)
    de, so, ep = plastic_states(rng, 300)
    tr = M.trial(de, so, E, NU)
    s1, ep1, _ = M.radial_return(tr, ep, E, NU, SY, H)
    q = M.overshoot(tr, ep, SY, H)
    pl = q > 0
    t = np.concatenate([(tr - s1) / q[:, None], (SY * (ep1 - ep) / q)[:, None]], axis=1)[pl]
    mean, std = np.zeros(7), np.ones(7)
    s, e = M.surrogate_state(t, tr[pl], ep[pl], mean, std, SY, H)
    assert np.allclose(s, s1[pl], rtol=0, atol=1e-6 * SY)
    assert np.abs(M.yield_residual(s, e, SY, H)).max() < 1e-6 * SY


def test_fit_scaling_floor():
    x = np.ones((10, M.N_IN))
    x[:, 0] = 0.0
    t = np.zeros((10, M.N_OUT))
    xm, xs, tm, ts = M.fit_scaling(x, t)
    assert np.all(xs[1:] == 1e-3) and xs[0] == 1.0 and np.all(ts == 1.0)


def _unit_deviatoric(rng, n):
    d = rng.normal(size=(n, 6))
    d[:, :3] -= d[:, :3].mean(axis=1, keepdims=True)
    return d / np.sqrt(M.voigt_dot(d, d))[:, None]


def test_acceptance_conditions_pass_the_exact_return_and_fail_planted_mistakes():
    """The exact J2 return meets the four acceptance conditions to rounding,
    and each planted mistake of the size of the tolerance breaks its own
    condition: a pressure shift (pressure), a turn of the returned deviator
    at fixed norm (turn), a plastic-strain increment off by tol / (sqrt(6)
    G) (consistency), a returned deviator scaled along itself off the yield
    surface with the matching plastic strain (yield).  The turn is measured
    against a brute-force minimum over the trial direction's line."""
    rng = np.random.default_rng(7)
    de, so, ep = plastic_states(rng, 600)
    tr = M.trial(de, so, E, NU)
    q = M.overshoot(tr, ep, SY, H)
    pl = q > 1e-3 * SY
    tr, ep, q = tr[pl], ep[pl], q[pl]
    s1, ep1, dg = M.radial_return(tr, ep, E, NU, SY, H)
    tol = 1e-3
    lim = tol * SY
    res = M.acceptance_residuals(tr, s1, ep, ep1, E, NU, SY, H)
    for r in res:
        assert r.max() < 1e-5 * lim
    assert np.all(M.surrogate_accepts(tr, s1, ep, ep1, E, NU, SY, H, tol))
    mu = E / (2 * (1 + NU))
    k = 3.0 * lim
    # pressure
    sp = s1.copy()
    sp[:, :3] += k
    r = M.acceptance_residuals(tr, sp, ep, ep1, E, NU, SY, H)
    assert np.all(r[0] > 2 * lim) and np.all(r[1] < 1e-5 * lim)
    # turn: rotate the returned deviator towards a direction orthogonal to the
    # trial deviator, keeping its norm
    dt = M.deviator(tr)
    nhat = dt / np.sqrt(M.voigt_dot(dt, dt))[:, None]
    u = _unit_deviatoric(rng, len(tr))
    u -= M.voigt_dot(u, nhat)[:, None] * nhat
    u /= np.sqrt(M.voigt_dot(u, u))[:, None]
    d1 = M.deviator(s1)
    n1 = np.sqrt(M.voigt_dot(d1, d1))
    ang = k / n1
    st = s1.copy()
    st += (np.cos(ang)[:, None] - 1.0) * d1 + (np.sin(ang) * n1)[:, None] * u
    r = M.acceptance_residuals(tr, st, ep, ep1, E, NU, SY, H)
    assert np.all(r[1] > 2 * lim)
    # brute force: the distance from the turned deviator to the trial line
    lam = np.linspace(0.9, 1.1, 20001)
    dtt = M.deviator(st)
    for i in range(5):
        dist = np.sqrt(M.voigt_dot(dtt[i] - lam[:, None] * nhat[i] * n1[i], dtt[i] - lam[:, None] * nhat[i] * n1[i]))
        assert abs(dist.min() - r[1][i]) < 1e-3 * r[1][i]
    assert np.all(r[3] < 1e-5 * lim)  # norm kept: still on the surface
    # consistency: dep off at an unchanged stress
    epc = ep1 + k / (np.sqrt(6.0) * mu)
    r = M.acceptance_residuals(tr, s1, ep, epc, E, NU, SY, H)
    assert np.all(r[2] > 2 * lim)
    # yield: deviator scaled along itself by k, plastic strain consistent with
    # the new norm, so only the yield residual fails
    sy_ = s1 - (k / n1)[:, None] * d1
    epy = ep + (np.sqrt(M.voigt_dot(dt, dt)) - (n1 - k)) / (np.sqrt(6.0) * mu)
    r = M.acceptance_residuals(tr, sy_, ep, epy, E, NU, SY, H)
    assert np.all(r[3] > 2 * lim) and np.all(r[2] < 1e-5 * lim) and np.all(r[1] < 1e-5 * lim)
    assert not np.any(M.surrogate_accepts(tr, sy_, ep, epy, E, NU, SY, H, tol))
