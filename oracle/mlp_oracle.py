"""CPU oracle of the MLP surrogate (test infrastructure only).

Only tests/, the smoke check and bench.py's CPU-baseline leg use this module,
as the checker of the HIP library; the library never calls it.

The surrogate has no counterpart in the reference (its stress update is the
linear K_e du of dynamics_kernels.h:101-114), so parity is anchored on
mathematics instead of reference vectors:
  * forward / loss / gradient in float64, pinned by central finite
    differences of the loss (tests/test_oracle_mlp.py);
  * the same with operands rounded to bfloat16 (round-to-nearest-even) at
    the points where the GPU kernels feed the matrix cores -- inputs,
    weights, hidden activations, dL/dy and pre-activation gradients -- so
    the GPU gradient can be compared to a tolerance set by FP32 summation
    order rather than by BF16 rounding;
  * the J2 radial return of Simo & Hughes (1998) Box 3.2 with linear
    isotropic hardening, pinned by its closed-form consistency conditions;
  * the surrogate's acceptance conditions (the four properties that fix a
    J2 return: pressure kept, no turn of the deviator, consistency, yield),
    pinned by the exact return meeting them to rounding and by planted
    mistakes each failing its own condition;
  * the sample -> (x, t) map of the trainer (overshoot-scaled targets).

Parameter order: W1[64][13], b1[64], W2[64][64], b2[64], W3[7][64], b3[7].
"""
import numpy as np

N_IN, N_HID, N_OUT = 13, 64, 7
OFF_W1 = 0
OFF_B1 = OFF_W1 + N_HID * N_IN
OFF_W2 = OFF_B1 + N_HID
OFF_B2 = OFF_W2 + N_HID * N_HID
OFF_W3 = OFF_B2 + N_HID
OFF_B3 = OFF_W3 + N_OUT * N_HID
N_PARAMS = OFF_B3 + N_OUT
K23 = np.sqrt(2.0 / 3.0)
MIN_OVERSHOOT = 1e-3


def bf16(a):
    """Round to bfloat16 (RNE) through float32, returned as float64."""
    f = np.asarray(a, dtype=np.float64).astype(np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(theta):
    t = np.asarray(theta, dtype=np.float64)
    W1 = t[OFF_W1:OFF_B1].reshape(N_HID, N_IN)
    b1 = t[OFF_B1:OFF_W2]
    W2 = t[OFF_W2:OFF_B2].reshape(N_HID, N_HID)
    b2 = t[OFF_B2:OFF_W3]
    W3 = t[OFF_W3:OFF_B3].reshape(N_OUT, N_HID)
    b3 = t[OFF_B3:N_PARAMS]
    return W1, b1, W2, b2, W3, b3


def pack(W1, b1, W2, b2, W3, b3):
    return np.concatenate([W1.ravel(), b1, W2.ravel(), b2, W3.ravel(), b3])


def silu(z):
    return z / (1.0 + np.exp(-z))


def silu_grad(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(theta, x, rounded=False):
    """y = W3 silu(W2 silu(W1 x + b1) + b2) + b3 for rows of x."""
    r = bf16 if rounded else (lambda a: np.asarray(a, dtype=np.float64))
    W1, b1, W2, b2, W3, b3 = unpack(theta)
    x = r(x)
    a1 = r(silu(x @ r(W1).T + b1))
    a2 = r(silu(a1 @ r(W2).T + b2))
    return a2 @ r(W3).T + b3


def loss_grad(theta, x, t, rounded=False):
    """L = 1/(2n) sum (y - t)^2 and dL/dtheta (canonical order)."""
    r = bf16 if rounded else (lambda a: np.asarray(a, dtype=np.float64))
    W1, b1, W2, b2, W3, b3 = unpack(theta)
    x = r(x)
    t = np.asarray(t, dtype=np.float64)
    n = x.shape[0]
    W1r, W2r, W3r = r(W1), r(W2), r(W3)
    z1 = x @ W1r.T + b1
    a1 = r(silu(z1))
    z2 = a1 @ W2r.T + b2
    a2 = r(silu(z2))
    y = a2 @ W3r.T + b3
    res = y - t
    loss = 0.5 * np.sum(res * res) / n
    dy = res / n
    dyr = r(dy)
    gW3 = dyr.T @ a2
    gb3 = dy.sum(axis=0)
    dz2 = (dyr @ W3r) * silu_grad(z2)
    dz2r = r(dz2)
    gW2 = dz2r.T @ a1
    gb2 = dz2.sum(axis=0)
    dz1 = (dz2r @ W2r) * silu_grad(z1)
    dz1r = r(dz1)
    gW1 = dz1r.T @ x
    gb1 = dz1.sum(axis=0)
    return loss, pack(gW1, gb1, gW2, gb2, gW3, gb3)


def finite_difference(theta, x, t, idx, h=1e-6):
    """Central differences of the float64 loss at parameter indices idx."""
    out = np.zeros(len(idx))
    th = np.array(theta, dtype=np.float64)
    for k, i in enumerate(idx):
        p, m = th.copy(), th.copy()
        p[i] += h
        m[i] -= h
        out[k] = (loss_grad(p, x, t)[0] - loss_grad(m, x, t)[0]) / (2 * h)
    return out


def elastic_moduli(E, nu):
    k = E / ((1 + nu) * (1 - 2 * nu))
    return (1 - nu) * k, nu * k, 0.5 * (1 - 2 * nu) * k, E / (2 * (1 + nu))


def trial(de, so, E, nu):
    """sigma_n + D de, Voigt order (xx, yy, zz, xy, yz, xz), engineering shear."""
    c11, c12, c44, _ = elastic_moduli(E, nu)
    de = np.asarray(de, dtype=np.float64)
    s = np.array(so, dtype=np.float64, copy=True)
    s[..., 0] += c11 * de[..., 0] + c12 * (de[..., 1] + de[..., 2])
    s[..., 1] += c11 * de[..., 1] + c12 * (de[..., 0] + de[..., 2])
    s[..., 2] += c11 * de[..., 2] + c12 * (de[..., 0] + de[..., 1])
    s[..., 3:] += c44 * de[..., 3:]
    return s


def dev_norm(s):
    p = (s[..., 0] + s[..., 1] + s[..., 2]) / 3.0
    d = s[..., :3] - p[..., None]
    return np.sqrt(np.sum(d * d, axis=-1) + 2.0 * np.sum(s[..., 3:] ** 2, axis=-1))


def overshoot(s, ep, Y0, H):
    return dev_norm(s) - K23 * (Y0 + H * np.asarray(ep, dtype=np.float64))


def radial_return(s_trial, ep, E, nu, Y0, H):
    """Simo & Hughes Box 3.2 (isotropic hardening): returns (sigma, ep, dgamma)."""
    _, _, _, mu = elastic_moduli(E, nu)
    s = np.array(s_trial, dtype=np.float64, copy=True)
    ep = np.array(ep, dtype=np.float64, copy=True)
    q = overshoot(s, ep, Y0, H)
    nrm = dev_norm(s)
    dg = np.where(q > 0, q / (2 * mu + 2.0 / 3.0 * H), 0.0)
    sc = np.where(q > 0, 2 * mu * dg / np.where(nrm > 0, nrm, 1.0), 0.0)
    p = (s[..., 0] + s[..., 1] + s[..., 2]) / 3.0
    for c in range(3):
        s[..., c] -= sc * (s[..., c] - p)
    for c in range(3, 6):
        s[..., c] -= sc * s[..., c]
    return s, ep + K23 * dg, dg


def sample_targets(samples, E, nu, Y0, H):
    """Raw (x, t) of samples (de, so, ep0, s1, ep1) whose trial overshoot
    exceeds MIN_OVERSHOOT * Y0: x = (de, so, ep0),
    t = ((s_trial - s1) / q, Y0 (ep1 - ep0) / q)."""
    smp = np.asarray(samples, dtype=np.float32).astype(np.float64)
    de, so, ep0, s1, ep1 = smp[:, 0:6], smp[:, 6:12], smp[:, 12], smp[:, 13:19], smp[:, 19]
    tr = trial(de, so, E, nu)
    q = overshoot(tr, ep0, Y0, H)
    keep = q > MIN_OVERSHOOT * Y0
    x = np.concatenate([de, so, ep0[:, None]], axis=1)[keep]
    t = np.concatenate([(tr - s1) / q[:, None], (Y0 * (ep1 - ep0) / q)[:, None]], axis=1)[keep]
    return x, t


def fit_scaling(x, t):
    """Per-column mean and std with the trainer's floor (1e-3 |mean|, else 1)."""
    def st(a):
        mu = a.mean(axis=0)
        sd = a.std(axis=0)
        sd = np.maximum(sd, 1e-3 * np.abs(mu))
        return mu, np.where(sd > 0, sd, 1.0)
    return st(x) + st(t)


def surrogate_state(y_std, s_trial, ep, out_mean, out_std, Y0, H):
    """Stress and plastic strain the surrogate proposes from standardized
    outputs (before the residual check)."""
    q = overshoot(s_trial, ep, Y0, H)
    y = np.asarray(y_std, dtype=np.float64) * out_std + out_mean
    s = s_trial - y[:, :6] * q[:, None]
    dep = np.maximum(y[:, 6] * q / Y0, 0.0)
    return s, ep + dep


def yield_residual(s, ep, Y0, H):
    return dev_norm(s) - K23 * (Y0 + H * np.asarray(ep, dtype=np.float64))


def voigt_dot(a, b):
    """a : b for Voigt vectors with engineering shear stored as stress
    (shear terms counted twice)."""
    return (np.sum(a[..., :3] * b[..., :3], axis=-1)
            + 2.0 * np.sum(a[..., 3:] * b[..., 3:], axis=-1))


def deviator(s):
    d = np.array(s, dtype=np.float64, copy=True)
    p = (d[..., 0] + d[..., 1] + d[..., 2]) / 3.0
    d[..., :3] -= p[..., None]
    return d


def acceptance_residuals(s_trial, s_new, ep, ep_new, E, nu, Y0, H):
    """The four J2 conditions a return map (s_trial, ep) -> (s_new, ep_new)
    meets exactly (Simo & Hughes 1998, Box 3.2: the return is radial,
    deviatoric, consistent and ends on the yield surface), as residuals in
    stress units: (pressure change, turn of the deviator, consistency,
    yield).  The surrogate's state is accepted when all four are within
    tol * Y0."""
    _, _, _, G = elastic_moduli(E, nu)
    s_trial = np.asarray(s_trial, dtype=np.float64)
    s_new = np.asarray(s_new, dtype=np.float64)
    dep = np.asarray(ep_new, dtype=np.float64) - np.asarray(ep, dtype=np.float64)
    c = s_trial - s_new
    pc = (c[..., 0] + c[..., 1] + c[..., 2]) / 3.0
    dt, dc = deviator(s_trial), deviator(c)
    nt = np.sqrt(voigt_dot(dt, dt))
    proj = voigt_dot(dc, dt) / nt
    turn = np.sqrt(np.maximum(voigt_dot(dc, dc) - proj * proj, 0.0))
    nn = np.sqrt(voigt_dot(dt - dc, dt - dc))
    consistency = (nt - nn) - np.sqrt(6.0) * G * dep
    return (np.abs(pc), turn, np.abs(consistency),
            np.abs(yield_residual(s_new, ep_new, Y0, H)))


def surrogate_accepts(s_trial, s_new, ep, ep_new, E, nu, Y0, H, tol):
    """Whether the surrogate's proposal passes its acceptance checks."""
    r = acceptance_residuals(s_trial, s_new, ep, ep_new, E, nu, Y0, H)
    lim = tol * Y0
    return (r[0] <= lim) & (r[1] <= lim) & (r[2] <= lim) & (r[3] <= lim)
