"""MLP surrogate of the J2 return map (cpi_mlp_* in include/cppimpact_hip.h).

The reference has no learned stress update; its element force is the linear
K_e du of dynamics_kernels.h:101-114.  This module wraps the network object
of the HIP library: sample collection from exact-return runs, BF16-MFMA
training on the GPU, and installation into a Dynamics context created with
stress_update=RETURN_MLP.
"""
import ctypes as C

import numpy as np

from . import _lib as L
from .dynamics import Dynamics


class Surrogate:
    """13 -> 64 -> 64 -> 7 SiLU network for one material."""

    def __init__(self, material, seed=0, device=0):
        mat = L.Material(material.E, material.rho, material.nu, getattr(material, "beta", 0.0),
                         getattr(material, "H", 0.0), getattr(material, "Y0", 0.0))
        h = C.c_void_p()
        L.check_mlp(L.lib().cpi_mlp_create(int(device), int(seed), C.byref(mat), C.byref(h)))
        self.handle = h
        self.material = material

    def _chk(self, rc):
        return L.check_mlp(rc, self.handle)

    @property
    def params(self):
        p = np.zeros(L.MLP_NPARAMS, dtype=np.float32)
        self._chk(L.lib().cpi_mlp_get_params(self.handle, L.fptr(p)))
        return p

    @params.setter
    def params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.shape == (L.MLP_NPARAMS,)
        self._chk(L.lib().cpi_mlp_set_params(self.handle, L.fptr(p)))

    def scaling(self):
        a, b = np.zeros(L.MLP_INPUTS, np.float32), np.zeros(L.MLP_INPUTS, np.float32)
        c, d = np.zeros(L.MLP_OUTPUTS, np.float32), np.zeros(L.MLP_OUTPUTS, np.float32)
        self._chk(L.lib().cpi_mlp_get_scaling(self.handle, L.fptr(a), L.fptr(b), L.fptr(c),
                                              L.fptr(d)))
        return a, b, c, d

    def set_scaling(self, in_mean, in_std, out_mean, out_std):
        arrs = [np.ascontiguousarray(v, dtype=np.float32) for v in
                (in_mean, in_std, out_mean, out_std)]
        self._chk(L.lib().cpi_mlp_set_scaling(self.handle, *(L.fptr(v) for v in arrs)))

    def save(self, path):
        """Writes the parameters and the scaling to an .npz file."""
        a, b, c, d = self.scaling()
        np.savez(path, params=self.params, in_mean=a, in_std=b, out_mean=c, out_std=d)

    @classmethod
    def load(cls, material, path, device=0):
        """A network for material with the parameters and scaling of a file
        save() wrote."""
        f = np.load(path)
        net = cls(material, device=device)
        net.set_scaling(f["in_mean"], f["in_std"], f["out_mean"], f["out_std"])
        net.params = f["params"]
        return net

    def fit_scaling(self, samples):
        s = np.ascontiguousarray(samples, dtype=np.float32)
        self._chk(L.lib().cpi_mlp_fit_scaling(self.handle, L.fptr(s), len(s)))

    def targets(self, samples):
        """Standardized (x, t) of the samples the trainer keeps."""
        s = np.ascontiguousarray(samples, dtype=np.float32)
        x = np.zeros((max(len(s), 1), L.MLP_INPUTS), np.float32)
        t = np.zeros((max(len(s), 1), L.MLP_OUTPUTS), np.float32)
        k = C.c_int32()
        self._chk(L.lib().cpi_mlp_targets(self.handle, L.fptr(s), len(s), L.fptr(x), L.fptr(t),
                                          C.byref(k)))
        return x[:k.value].copy(), t[:k.value].copy()

    def forward(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.zeros((len(x), L.MLP_OUTPUTS), np.float32)
        self._chk(L.lib().cpi_mlp_forward(self.handle, L.fptr(x), len(x), L.fptr(y)))
        return y

    def gradient(self, x, t):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        g = np.zeros(L.MLP_NPARAMS, np.float32)
        loss = C.c_float()
        self._chk(L.lib().cpi_mlp_gradient(self.handle, L.fptr(x), L.fptr(t), len(x), L.fptr(g),
                                           C.byref(loss)))
        return float(loss.value), g

    def train(self, samples, epochs=60, batch=1024, lr=2e-3, lr_final=2e-5, seed=1):
        o = L.MlpTrainOpts()
        L.lib().cpi_mlp_default_train_opts(C.byref(o))
        o.epochs, o.batch, o.lr, o.lr_final, o.seed = epochs, batch, lr, lr_final, seed
        s = np.ascontiguousarray(samples, dtype=np.float32)
        hist = np.zeros(epochs, np.float32)
        self._chk(L.lib().cpi_mlp_train(self.handle, L.fptr(s), len(s), C.byref(o),
                                        L.fptr(hist)))
        return hist

    def return_map(self, de, sigma_n, ep_n, tolerance=1e-3):
        """Surrogate return map of the element kernels on given points:
        (sigma_{n+1}, ep_{n+1}, outcome CPI_MLP_*)."""
        de = np.ascontiguousarray(de, dtype=np.float32)
        so = np.ascontiguousarray(sigma_n, dtype=np.float32)
        ep = np.ascontiguousarray(ep_n, dtype=np.float32)
        n = len(ep)
        s = np.zeros((n, 6), np.float32)
        e = np.zeros(n, np.float32)
        how = np.zeros(n, np.int32)
        self._chk(L.lib().cpi_mlp_return_map(self.handle, float(tolerance), n, L.fptr(de),
                                             L.fptr(so), L.fptr(ep), L.fptr(s), L.fptr(e),
                                             L.iptr(how)))
        return s, e, how

    def close(self):
        if getattr(self, "handle", None) is not None and self.handle.value:
            L.lib().cpi_mlp_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def collect_samples(mesh, material, wall, speeds, dt, nsteps, capacity=1 << 21,
                    direction=(0.0, 0.0, -1.0), **options):
    """Runs exact-return impacts (FP32 element kernels unless overridden) at
    each speed and returns the recorded yielding points [k][20], sorted.

    capacity is the first recorder size per speed.  The recorder fills its
    slots in atomic-counter order, so a run that sees more yielding points
    than it holds would keep a subset that depends on that order: such a
    speed is run again with room for every point it saw (the simulation is
    deterministic, so the count is), and the set, and the training run on
    it, are the same in every process."""
    out = []
    opts = dict(physics=L.PHYSICS_EP, precision=L.PREC_F32)
    opts.update(options)

    def run(v, cap):
        dyn = Dynamics(mesh, material, wall, **opts)
        dyn.vel[:] = np.asarray(direction, dtype=np.float64) * float(v)
        dyn.enable_samples(cap)
        dyn.init(dt)
        dyn.step(nsteps)
        dyn.synchronize()
        got = dyn.samples()
        dyn.close()
        return got

    for v in speeds:
        s, seen = run(v, capacity)
        if seen > len(s):
            s, seen2 = run(v, seen)
            if seen2 != seen or len(s) != seen:
                raise RuntimeError(f"collect_samples: {seen} then {seen2} yielding points at "
                                   f"{v} m/s, {len(s)} kept")
        out.append(s)
    if not out:
        return np.zeros((0, L.SAMPLE_FLOATS), np.float32)
    smp = np.concatenate(out)
    # the recorder fills slots in atomic-counter order; sorting the rows makes
    # the set (and so the training run) independent of that order
    return smp[np.lexsort(smp.T[::-1])]


def workload_training_set(workloads, per_workload=3, target=10_000_000, seed=0, low=5.0,
                          high=30.0, precision=L.PREC_F64, log=None):
    """Return-mapping samples of the plain solver on whole workloads.

    Each workload (cppimpact_amd.workloads) runs its full length with the
    exact return at per_workload impact speeds drawn uniformly from
    [low, high] m/s (numpy default_rng(seed)).  A first run of every speed
    only counts the yielding points; each workload that yields then gets an
    equal share of target, which its runs keep as a fixed hash-selected
    fraction of their points (cpi_samples_subsample), so the set is about
    target samples, spread over the workloads and over the whole of each
    run, and the same in every process.  Returns (samples [k][20] sorted,
    info) with info listing the speeds, yielding points and kept samples
    per run."""
    rng = np.random.default_rng(seed)
    plan = [(w, [float(v) for v in rng.uniform(low, high, per_workload)]) for w in workloads]
    runs = []

    def one(dyn, w, v, cap, keep):
        dyn.vel[:] = np.asarray(w.direction, dtype=np.float64) * v
        dyn.reset_samples()
        dyn.subsample_samples(keep, seed=seed * 1000003 + len(runs))
        dyn.init(0.0)
        dyn.step(w.steps)
        dyn.synchronize()
        return dyn.samples()

    # pass 1: yielding points per run (nothing recorded, only counted)
    counts = {}
    for w, speeds in plan:
        dyn = w.dynamics(precision=precision)
        dyn.enable_samples(1)
        for v in speeds:
            _, seen = one(dyn, w, v, 1, 1.0)
            counts[(w.name, v)] = seen
        dyn.close()
    yielding = [w.name for w, speeds in plan if sum(counts[(w.name, v)] for v in speeds) > 0]
    if not yielding:
        return np.zeros((0, L.SAMPLE_FLOATS), np.float32), {"runs": []}
    share = target / len(yielding)
    out = []
    for w, speeds in plan:
        total = sum(counts[(w.name, v)] for v in speeds)
        if total == 0:
            for v in speeds:
                runs.append(dict(workload=w.name, speed=round(v, 3), yielding=0, kept=0))
            continue
        keep = min(1.0, share / total)
        cap = int(max(counts[(w.name, v)] for v in speeds) * keep * 1.05) + 4096
        dyn = w.dynamics(precision=precision)
        dyn.enable_samples(cap)
        for v in speeds:
            s, seen = one(dyn, w, v, cap, keep)
            if seen > len(s):
                # the recorder overflowed: again with room for every kept point
                dyn.enable_samples(seen)
                s, seen2 = one(dyn, w, v, seen, keep)
                if seen2 != seen or len(s) != seen:
                    raise RuntimeError(f"workload_training_set: {seen} then {seen2} kept points "
                                       f"for {w.name} at {v} m/s")
                dyn.enable_samples(cap)
            runs.append(dict(workload=w.name, speed=round(v, 3), yielding=counts[(w.name, v)],
                             kept=len(s), keep_fraction=keep))
            if log:
                log(f"training set: {w.name} at {v:.2f} m/s: {counts[(w.name, v)]} yielding "
                    f"points, {len(s)} kept")
            out.append(s)
        dyn.close()
    smp = np.concatenate(out) if out else np.zeros((0, L.SAMPLE_FLOATS), np.float32)
    smp = smp[np.lexsort(smp.T[::-1])]
    return smp, {"runs": runs, "samples": int(len(smp)), "target": int(target),
                 "seed": int(seed), "speed_range": [low, high]}
