"""The benchmark configurations at their real sizes on the GPU.

c0: 10^3-hex cube, 6,000 tet4, 1,000 steps at -10 m/s, FP64, against the
    CPU oracle over the whole run.
c1: 20^3-hex cube, 48,000 tet10 (4-point rule), -20 m/s: the first steps
    and the whole 2,000-step run against the oracle on the full mesh, the
    energy balance, FP32 against FP64.
c2: 64^3-hex cube, 1,572,864 tet4 jittered by 0.1 of the spacing (seed 0),
    -10 m/s, the throughput workload: its 500 steps against the oracle on
    the full mesh in FP64 and FP32, then 241 steps of FP64, FP32 and FP32
    with the surrogate (the impact stays elastic, so the network is never
    evaluated).
c4: the surrogate on the c2 mesh where the cube yields (30 m/s), trained on
    return-mapping samples of c1-c3 runs, against the FP64 and FP32 exact
    runs over c2's 500 steps.
c3: thin-walled cylinder R 1 m, L 4 m, t 0.02 m, 360 x 185 x 3 cells x 6
    tet10 (1,198,800 elements), oblique 7 m/s impact 30 degrees from the
    wall normal: the first 10 steps against the oracle on the full mesh,
    then 1,000 steps, FP32 against FP64.
Tolerances are written per test."""
import math

import numpy as np
import pytest

import cppimpact_amd as ci
from cppimpact_amd import _lib as L
from oracle import pyoracle as O

from common import E, H, NU, RHO, SY, rel

pytestmark = pytest.mark.gpu


def material():
    return ci.Elastoplastic(E, RHO, NU, 1.0, H, SY)


def dynamics(x, conn, v0, k, **opts):
    d = ci.Dynamics(ci.Mesh.from_arrays(x, conn), material(), ci.Wall("Wall", 0.0, k, None, 0, 1),
                    physics=L.PHYSICS_EP, **opts)
    d.vel[:] = np.asarray(v0, dtype=np.float64)
    return d


def run(d, dt, steps):
    d.init(dt)
    d.step(steps)
    out = dict(zip(("x", "v", "a", "m", "v_i"), d.fetch()))
    out["energy"], out["step"] = d.energy()
    out["dt"] = d.info().dt
    return out


def test_c0_full_run_against_oracle():
    x, conn = ci.box_mesh(10, 10, 10, 0.25, 0.25, 0.25, origin=(0.0, 0.0, 5e-5))
    k = 10 * E * 0.025
    v0 = (0.0, 0.0, -10.0)
    d = dynamics(x, conn, v0, k)
    g = run(d, 0.0, 1000)
    dt = g["dt"]
    o = O.Solver(x, conn, v0, E=E, nu=NU, rho=RHO, sigma_y=SY, H=H, physics=O.EP,
                 wall=dict(location=0.0, stiffness=k), dt=dt).solve(1000)
    print(f"c0: dt {dt:.4e}, disp err {rel(g['x'] - x, o['x'] - x):.2e}, "
          f"v err {rel(g['v'], o['v']):.2e}")
    assert rel(g["x"] - x, o["x"] - x) < 1e-8
    assert rel(g["v"], o["v"]) < 1e-8
    eo = o["energy"][-1]
    for k_ in (O.E_KINETIC, O.E_INTERNAL, O.E_PENALTY, O.E_GRAVITY):
        assert abs(g["energy"][k_] - eo[k_]) <= 1e-7 * eo[O.E_KINETIC]
    d.close()


def c1_mesh():
    x, conn = ci.box_mesh(20, 20, 20, 0.25, 0.25, 0.25, origin=(0.0, 0.0, 1e-6), quadratic=True)
    assert conn.shape == (48000, 10) and x.shape[0] == 68921
    return x, conn, 10 * E * 0.25 / 20


def test_c1_first_steps_against_oracle_on_full_mesh():
    x, conn, k = c1_mesh()
    v0 = (0.0, 0.0, -20.0)
    d = dynamics(x, conn, v0, k)
    g = run(d, 0.0, 3)
    o = O.Solver(x, conn, v0, E=E, nu=NU, rho=RHO, sigma_y=SY, H=H, physics=O.EP,
                 wall=dict(location=0.0, stiffness=k), dt=g["dt"]).solve(3)
    assert rel(g["m"], o["m"]) < 1e-12
    assert rel(g["x"] - x, o["x"] - x) < 1e-9
    assert rel(g["v"], o["v"]) < 1e-9
    d.close()


def test_c1_two_thousand_steps_energy_and_fp32():
    x, conn, k = c1_mesh()
    v0 = (0.0, 0.0, -20.0)
    d64 = dynamics(x, conn, v0, k)
    d32 = dynamics(x, conn, v0, k, precision=L.PREC_F32)
    # control: FP64 with the colour scatter differs from d64 only in the
    # rounding of the nodal force sums
    dcol = dynamics(x, conn, v0, k, scatter=L.SCATTER_COLOR)
    d64.init(0.0)
    d32.init(d64.info().dt)
    dcol.init(d64.info().dt)
    errs, ctrl = {}, {}
    done = 0
    for n in (20, 100, 500, 1000, 2000):
        for d in (d64, d32, dcol):
            d.step(n - done)
        done = n
        x64, x32, xc = d64.fetch()[0], d32.fetch()[0], dcol.fetch()[0]
        _, ep64 = d64.element_state()
        _, ep32 = d32.element_state()
        errs[n] = (rel(x32 - x, x64 - x), rel(ep32, ep64) if ep64.max() > 0 else 0.0)
        ctrl[n] = rel(xc - x, x64 - x)
    m = d64.fetch()[3]
    ke0 = 0.5 * m.sum() * 20.0 ** 2
    e, _ = d64.energy()
    print(f"c1: dt {d64.info().dt:.4e}, balance {e[L.E_BALANCE] / ke0:.3e} KE0, "
          f"plastic work {e[L.E_PLASTIC] / ke0:.3e} KE0, ep max {ep64.max():.3e}, "
          "fp32 vs fp64 (disp, ep) by step: "
          + ", ".join(f"{n}: {a:.1e}/{b:.1e}" for n, (a, b) in errs.items())
          + "; fp64 colour vs block (disp): "
          + ", ".join(f"{n}: {a:.1e}" for n, a in ctrl.items()))
    assert np.isfinite(e).all()
    assert ep64.max() > 0  # 20 m/s yields
    # measured 8.8e-6 KE0 while 33% of KE0 is dissipated plastically
    assert abs(e[L.E_BALANCE]) < 1e-4 * ke0
    # FP32 element arithmetic: the gap starts at rounding level and grows
    # with the plastic flow (the yield overshoot is a small difference of
    # large stresses); measured 2.2e-2 after 2000 steps
    assert errs[20][0] < 1e-5
    assert errs[2000][0] < 5e-2
    # the whole 2,000-step run against the oracle (threaded, bit-identical to
    # the serial one): the node sums differ in order only, a rounding-level
    # gap that the plastic flow amplifies as it does between the colour and
    # block scatters above
    import os

    threads = max(1, min(16, os.cpu_count() or 1))
    o = O.Solver(x, conn, v0, E=E, nu=NU, rho=RHO, sigma_y=SY, H=H, physics=O.EP,
                 wall=dict(location=0.0, stiffness=k), dt=d64.info().dt).solve(2000,
                                                                             threads=threads)
    x64 = d64.fetch()[0]
    derr = rel(x64 - x, o["x"] - x)
    dl2 = float(np.linalg.norm(x64 - o["x"]) / np.linalg.norm(o["x"] - x))
    eperr = rel(ep64, o["epbar"])
    eo = o["energy"][-1]
    de = {k_: abs(e[k_] - eo[k_]) / ke0 for k_ in (O.E_KINETIC, O.E_INTERNAL, O.E_PLASTIC,
                                                    O.E_PENALTY)}
    print(f"c1 2000 steps against the oracle: disp {derr:.2e} (L2 {dl2:.2e}), ep {eperr:.2e}, "
          f"energies / KE0 {de}")
    # measured: disp 1.2e-5 (L2 6.9e-7), ep 3.9e-6, energies <= 9.5e-9 KE0
    # (colour vs block scatter in the same run: 1.5e-5)
    assert derr < 1e-4 and dl2 < 1e-5 and eperr < 1e-4
    assert all(v < 1e-7 for v in de.values())
    d64.close()
    d32.close()
    dcol.close()


def c2_mesh():
    x, conn = ci.box_mesh(64, 64, 64, 0.25, 0.25, 0.25, origin=(0.0, 0.0, 1e-6), jitter=0.1,
                          seed=0)
    assert conn.shape == (1572864, 4) and x.shape[0] == 274625
    return x, conn, 10 * E * 0.25 / 64


def test_c2_whole_run_against_oracle_on_full_mesh():
    """c2 for its stated 500 steps: FP64 against the oracle (its threaded
    solve sums every node in element order, bit-identical to the serial
    one) and the FP32 element kernels against the same oracle.  The two
    sum a node's forces in different orders; over 500 steps on the penalty
    wall a rounding-level difference can flip a node in or out of contact
    for a step, a local jump the max-norm picks up (measured FP64 2.7e-9
    in displacement, 1.8e-7 in velocity; FP32 6.5e-3 in displacement,
    against 2.5e-5 at step 241), while the displacement field as a whole
    (relative L2) stays close.  Energies to 1e-8 of KE0."""
    import os

    x, conn, k = c2_mesh()
    v0 = (0.0, 0.0, -10.0)
    d = dynamics(x, conn, v0, k)
    g = run(d, 0.0, 500)
    d32 = dynamics(x, conn, v0, k, precision=L.PREC_F32)
    g32 = run(d32, g["dt"], 500)
    threads = max(1, min(16, os.cpu_count() or 1))
    o = O.Solver(x, conn, v0, E=E, nu=NU, rho=RHO, sigma_y=SY, H=H, physics=O.EP,
                 wall=dict(location=0.0, stiffness=k), dt=g["dt"]).solve(500, threads=threads)
    ke0 = 0.5 * o["m"].sum() * 10.0 ** 2
    derr, verr = rel(g["x"] - x, o["x"] - x), rel(g["v"], o["v"])
    d32err = rel(g32["x"] - x, o["x"] - x)

    def l2(a, b):
        return float(np.linalg.norm(a - b) / np.linalg.norm(b))

    dl2, d32l2 = l2(g["x"] - x, o["x"] - x), l2(g32["x"] - x, o["x"] - x)
    print(f"c2 500 steps: fp64 disp {derr:.2e} (L2 {dl2:.2e}) v {verr:.2e}; fp32 disp "
          f"{d32err:.2e} (L2 {d32l2:.2e}); balance {g['energy'][L.E_BALANCE] / ke0:.2e} KE0")
    assert rel(g["m"], o["m"]) < 1e-12
    assert derr < 1e-8 and dl2 < 1e-9 and verr < 1e-6
    eo = o["energy"][-1]
    de = {k_: abs(g["energy"][k_] - eo[k_]) / ke0 for k_ in range(5)}
    print(f"c2 500 steps: energy differences / KE0 {de}, contact force "
          f"{abs(g['energy'][L.E_CONTACTF] / eo[O.E_CONTACTF] - 1):.2e}")
    for k_ in (O.E_KINETIC, O.E_INTERNAL, O.E_PENALTY, O.E_GRAVITY):
        assert de[k_] <= 1e-8, k_
    assert abs(g["energy"][L.E_CONTACTF] - eo[O.E_CONTACTF]) <= 1e-8 * abs(eo[O.E_CONTACTF])
    assert d32l2 < 1e-3 and d32err < 2e-2
    d.close()
    d32.close()


def test_c2_fp32_and_surrogate_modes_against_fp64():
    from cppimpact_amd.surrogate import Surrogate

    x, conn, k = c2_mesh()
    v0 = (0.0, 0.0, -10.0)
    d64 = dynamics(x, conn, v0, k)
    d32 = dynamics(x, conn, v0, k, precision=L.PREC_F32)
    dml = dynamics(x, conn, v0, k, precision=L.PREC_F32, stress_update=L.RETURN_MLP)
    # an untrained network with unit scaling: a point that reached it would
    # show in the stats
    net = Surrogate(material(), seed=0)
    net.set_scaling(np.zeros(L.MLP_INPUTS), np.ones(L.MLP_INPUTS), np.zeros(L.MLP_OUTPUTS),
                    np.ones(L.MLP_OUTPUTS))
    dml.set_surrogate(net)
    d64.init(0.0)
    for d in (d32, dml):
        d.init(d64.info().dt)
    for d in (d64, d32, dml):
        d.step(241)
    x64, x32, xml = d64.fetch()[0], d32.fetch()[0], dml.fetch()[0]
    _, ep64 = d64.element_state()
    _, epml = dml.element_state()
    m = d64.fetch()[3]
    ke0 = 0.5 * m.sum() * 10.0 ** 2
    e, _ = d64.energy()
    err32, errml = rel(x32 - x, x64 - x), rel(xml - x, x64 - x)
    print(f"c2: 241 steps, fp32 vs fp64 disp {err32:.2e}, surrogate mode vs fp64 {errml:.2e}, "
          f"surrogate vs fp32 {rel(xml - x, x32 - x):.2e}, balance {e[L.E_BALANCE] / ke0:.2e} KE0, "
          f"stats {dml.surrogate_stats()}")
    assert ep64.max() == 0.0 and epml.max() == 0.0  # elastic at 10 m/s
    assert dml.surrogate_stats() == (0, 0)
    # FP32 element arithmetic with FP64 nodal state; measured 2.5e-5
    assert err32 < 1e-3 and errml < 1e-3
    # the surrogate mode's elastic pass rounds as the FP32 kernel's elastic
    # path does: the same positions bit for bit
    assert np.array_equal(xml, x32)
    # measured 5.8e-5 KE0
    assert np.isfinite(e).all() and abs(e[L.E_BALANCE]) < 5e-4 * ke0
    for d in (d64, d32, dml):
        d.close()


@pytest.mark.parametrize("precision", [L.PREC_F64, L.PREC_F32], ids=["f64", "f32"])
def test_c2_graph_replay_equals_eager_launches(precision):
    """The throughput workload as bench.py times it (steps replayed from
    captured multi-step graphs) against the same steps launched one kernel at
    a time: the same positions and velocities bit for bit."""
    x, conn, k = c2_mesh()
    v0 = (0.0, 0.0, -10.0)
    dg = dynamics(x, conn, v0, k, precision=precision)
    de = dynamics(x, conn, v0, k, precision=precision, use_graph=0)
    dg.init(0.0)
    de.init(dg.info().dt)
    for d in (dg, de):
        d.step(70)
    xg, vg = dg.fetch()[:2]
    xe, ve = de.fetch()[:2]
    assert np.array_equal(xg, xe) and np.array_equal(vg, ve)
    assert np.abs(xg - x).max() > 0  # the cube has moved
    dg.close()
    de.close()


def c3_case():
    nt, nz, nr = 360, 185, 3
    R, Lz, t = 1.0, 4.0, 0.02
    x, conn = ci.cylinder_mesh(nt, nz, nr, R, Lz, t, origin=(0.0, 0.0, 1e-6), quadratic=True)
    assert conn.shape == (nt * nz * nr * 6, 10)
    a = math.radians(30.0)
    v0 = (7.0 * math.sin(a), 0.0, -7.0 * math.cos(a))
    return x, conn, v0, 10 * E * t / nr


def test_c3_first_steps_against_oracle_on_full_mesh():
    """The full cylinder (1,198,800 tet10, grown blocks, the FP64 state
    streamed past the caches) for its first 10 steps, the end ring reaching
    the wall in the second: FP64 and FP32 against the oracle."""
    import os

    x, conn, v0, k = c3_case()
    d = dynamics(x, conn, v0, k)
    g = run(d, 0.0, 10)
    d32 = dynamics(x, conn, v0, k, precision=L.PREC_F32)
    g32 = run(d32, g["dt"], 10)
    threads = max(1, min(16, os.cpu_count() or 1))
    o = O.Solver(x, conn, v0, E=E, nu=NU, rho=RHO, sigma_y=SY, H=H, physics=O.EP,
                 wall=dict(location=0.0, stiffness=k), dt=g["dt"]).solve(10, threads=threads)
    derr, verr = rel(g["x"] - x, o["x"] - x), rel(g["v"], o["v"])
    d32err, v32err = rel(g32["x"] - x, o["x"] - x), rel(g32["v"], o["v"])
    ke0 = 0.5 * o["m"].sum() * 7.0 ** 2
    eo = o["energy"][-1]
    print(f"c3 10 steps: fp64 disp {derr:.2e} v {verr:.2e}; fp32 disp {d32err:.2e} v "
          f"{v32err:.2e}; contact force {eo[O.E_CONTACTF]:.3e}")
    assert eo[O.E_CONTACTF] > 0
    # measured: FP64 3.0e-13 / 5.4e-14, FP32 4.7e-8 / 9.1e-8
    assert rel(g["m"], o["m"]) < 1e-12
    assert derr < 1e-11 and verr < 1e-11
    for k_ in (O.E_KINETIC, O.E_INTERNAL, O.E_PENALTY, O.E_GRAVITY):
        assert abs(g["energy"][k_] - eo[k_]) <= 1e-8 * ke0, k_
    assert d32err < 1e-6 and v32err < 1e-6
    d.close()
    d32.close()


def test_c3_cylinder_oblique_impact_fp32_against_fp64():
    x, conn, v0, k = c3_case()
    d64 = dynamics(x, conn, v0, k)
    g64 = run(d64, 0.0, 1000)
    d32 = dynamics(x, conn, v0, k, precision=L.PREC_F32)
    g32 = run(d32, g64["dt"], 1000)
    ke0 = 0.5 * g64["m"].sum() * 7.0 ** 2
    e = g64["energy"]
    derr = rel(g32["x"] - x, g64["x"] - x)
    verr = rel(g32["v"], g64["v"])
    print(f"c3: {conn.shape[0]} tet10, dt {g64['dt']:.4e}, balance {e[L.E_BALANCE] / ke0:.3e} KE0, "
          f"contact force {e[L.E_CONTACTF]:.3e}, fp32 disp err {derr:.2e}, v err {verr:.2e}")
    assert np.isfinite(e).all()
    assert e[L.E_CONTACTF] > 0  # the end ring is on the wall
    # measured -1.9e-5 KE0 over the 1,000 steps
    assert abs(e[L.E_BALANCE]) < 2e-4 * ke0
    assert derr < 1e-3
    d64.close()
    d32.close()


def test_c4_surrogate_at_work_on_c2_mesh():
    """c4: the surrogate on the c2 mesh at 30 m/s, where the cube yields from
    the first steps, over c2's 500 steps.  The network is trained as c4
    states, on the plain solver's return-mapping samples of c1, c2 and c3
    runs at speeds drawn uniformly from [5, 30] m/s (three per workload, as
    bench.py draws them, and 2M samples instead of its 10M, to keep the
    test short).  The network replaces the return map at nearly every yielding
    point; displacement and plastic strain stay within the stated bounds
    of the plain FP64 run (measured with the 10M-sample network: L2 6.5e-5
    and max 7.8e-3 in displacement, L2 1.4e-3 in plastic strain; the FP32
    exact return alone is at L2 2.1e-6, max 6.6e-4)."""
    from cppimpact_amd import workloads as W
    from cppimpact_amd.surrogate import Surrogate, workload_training_set

    smp, info = workload_training_set([W.c1(), W.c2(), W.c3()], per_workload=3,
                                      target=2_000_000, seed=0)
    kept = {}
    for r in info["runs"]:
        kept[r["workload"]] = kept.get(r["workload"], 0) + r["kept"]
    assert len(smp) > 1_000_000 and min(kept.values()) > 0, kept
    net = Surrogate(material(), seed=0)
    hist = net.train(smp, epochs=10, batch=1024)
    assert hist[-1] < 0.05 * hist[0]
    x, conn, k = c2_mesh()
    v0 = (0.0, 0.0, -30.0)
    d64 = dynamics(x, conn, v0, k)
    d32 = dynamics(x, conn, v0, k, precision=L.PREC_F32)
    dml = dynamics(x, conn, v0, k, precision=L.PREC_F32, stress_update=L.RETURN_MLP)
    dml.set_surrogate(net)
    d64.init(0.0)
    for d in (d32, dml):
        d.init(d64.info().dt)
    for d in (d64, d32, dml):
        d.step(500)
    x64, x32, xml = d64.fetch()[0], d32.fetch()[0], dml.fetch()[0]
    _, ep64 = d64.element_state()
    _, epml = dml.element_state()
    acc, fb = dml.surrogate_stats()

    def l2(a, b):
        return float(np.linalg.norm(a - b) / np.linalg.norm(b))

    dl2, dmx = l2(xml - x, x64 - x), rel(xml - x, x64 - x)
    el2 = l2(epml, ep64)
    print(f"c4 30 m/s 500 steps: {acc} accepted, {fb} fell back; surrogate vs fp64 disp L2 "
          f"{dl2:.2e} max {dmx:.2e}, ep L2 {el2:.2e}; fp32 vs fp64 disp L2 "
          f"{l2(x32 - x, x64 - x):.2e} max {rel(x32 - x, x64 - x):.2e}; ep max {ep64.max():.3e}")
    assert ep64.max() > 1e-3
    assert acc > 10_000_000 and acc >= 0.95 * (acc + fb)
    assert dl2 < 1e-3 and dmx < 3e-2 and el2 < 2e-2
    assert l2(x32 - x, x64 - x) < 1e-4
    assert np.isfinite(dml.energy()[0]).all()
    for d in (d64, d32, dml):
        d.close()
    net.close()
