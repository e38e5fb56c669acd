"""Throughput of the explicit-dynamics step on one MI355X.

Workload: the 1.57M-tet cube -- 64^3 hexes x 6 tet4 (1,572,864 elements,
274,625 nodes), interior nodes jittered by U(-0.1, 0.1) of the spacing from
SplitMix64 seed 0, aluminium J2 elastoplastic (E 68.9 GPa, nu 0.33,
rho 2700, Y0 276 MPa, H 700 MPa), impact at -10 m/s on a rigid wall at z = 0
with penalty k = 10 E h, dt = 0.5 x the stable step.  A step is one loop
iteration of Dynamics::solve over the whole mesh, inputs resident in HBM.

Two runs of the same workload: the plain FP64 step (exact radial return) and
the mixed-precision step (FP32 element kernels, FP64 nodal state, mass, dt
and energies, MLP surrogate of the return map with exact fallback; the
network is trained first on samples of small-cube impacts at speeds uniform
in [5, 30] m/s, --no-surrogate keeps the exact return).  The headline value
is the mixed-precision run; the plain run, the speedup and the displacement
and plastic-strain errors between them ride along.  A supplementary pair of
runs on the thin-walled cylinder (1,198,800 tet10, oblique 7 m/s impact;
--no-c3 skips it) reports the tet10 step in FP64 and FP32 under
"cylinder_c3".  Two CPU lines: "cpu_baseline" is the reference algorithm
on one core, "cpu_baseline_threaded" the FP64 elastoplastic step on up to
16 threads.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
N_HEX = 64
SIDE = 0.25


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--profile-steps", type=int, default=20)
    ap.add_argument("--scatter", default="block", choices=["block", "color", "atomic"])
    ap.add_argument("--no-surrogate", action="store_true",
                    help="mixed-precision run with the exact return instead of the MLP")
    ap.add_argument("--speed", type=float, default=10.0, help="impact speed, m/s")
    ap.add_argument("--train-n", type=int, default=16,
                    help="hexes per side of the cube the training samples come from")
    ap.add_argument("--train-epochs", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-steps", type=int, default=1)
    ap.add_argument("--cpu-ep-steps", type=int, default=40,
                    help="steps of the threaded FP64 elastoplastic CPU line")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="threads of that line (default: min(16, cores))")
    ap.add_argument("--n", type=int, default=N_HEX, help="hexes per side (default 64)")
    ap.add_argument("--no-c3", action="store_true",
                    help="skip the supplementary thin-walled cylinder (tet10) runs")
    ap.add_argument("--c3-steps", type=int, default=100)
    return ap.parse_args()


def make_mesh(n):
    import cppimpact_amd as ci
    return ci.box_mesh(n, n, n, SIDE, SIDE, SIDE, origin=(0.0, 0.0, 1e-6), jitter=0.1, seed=0)


def train_surrogate(args):
    """Samples from exact-return FP32 impacts of a small cube at speeds
    uniform in [5, 30] m/s, then BF16-MFMA training on the GPU."""
    import numpy as np

    import cppimpact_amd as ci
    from cppimpact_amd.surrogate import Surrogate, collect_samples

    E = 68.9e9
    n = args.train_n
    x, conn = ci.box_mesh(n, n, n, SIDE, SIDE, SIDE, origin=(0.0, 0.0, 1e-6), jitter=0.1, seed=1)
    mat = ci.Elastoplastic(E, 2700.0, 0.33, 1.0, 700e6, 276e6)
    wall = ci.Wall("Wall", 0.0, 10 * E * SIDE / n, None, 0, 1)
    speeds = np.random.default_rng(0).uniform(5.0, 30.0, 6)
    t0 = time.time()
    smp = collect_samples(ci.Mesh.from_arrays(x, conn), mat, wall, speeds, 0.0, 400,
                          capacity=1 << 21)
    t1 = time.time()
    net = Surrogate(mat, seed=0)
    hist = net.train(smp, epochs=args.train_epochs, batch=1024)
    t2 = time.time()
    info = {"samples": int(len(smp)), "speeds": [round(float(v), 2) for v in speeds],
            "mesh": f"cube {n}^3 x 6 tet4, 400 steps per speed", "epochs": args.train_epochs,
            "loss_first": float(hist[0]), "loss_last": float(hist[-1]),
            "collect_s": round(t1 - t0, 2), "train_s": round(t2 - t1, 2)}
    return net, info


def formula_elem_bytes(info, precision):
    """Algorithmic bytes per launch of the element kernel from the step's
    compulsory-traffic formula B_step = E (4 npe + 14 nq s) + 224 N:
    connectivity, sigma (6) and ep (1) read and written per point in the
    element precision s, and per node the element pass's x, v gather and
    force scatter (72 B) and the node pass (152 B).  The element kernel's
    launch covers all of it except the node pass of the block-boundary
    nodes, which the boundary-node kernel finishes (152 B each)."""
    from cppimpact_amd import _lib as L
    s = 4.0 if precision == L.PREC_F32 else 8.0
    e, n = info.num_elements, info.num_nodes
    step = e * (4.0 * info.nodes_per_element + 14.0 * info.num_quadrature_pts * s) + 224.0 * n
    return step - 152.0 * info.num_boundary_nodes


def run_mode(x, conn, args, precision, stress_update, dt, net=None):
    import numpy as np

    import cppimpact_amd as ci
    from cppimpact_amd import _lib as L

    E = 68.9e9
    mesh = ci.Mesh.from_arrays(x, conn)
    mat = ci.Elastoplastic(E, 2700.0, 0.33, 1.0, 700e6, 276e6)
    wall = ci.Wall("Wall", 0.0, 10 * E * SIDE / args.n, None, 0, 1)
    scatter = {"block": L.SCATTER_BLOCK, "color": L.SCATTER_COLOR, "atomic": L.SCATTER_ATOMIC}
    dyn = ci.Dynamics(mesh, mat, wall, physics=L.PHYSICS_EP, precision=precision,
                      stress_update=stress_update, scatter=scatter[args.scatter])
    dyn.vel[:] = (0.0, 0.0, -args.speed)
    if net is not None:
        dyn.set_surrogate(net)
    t0 = time.time()
    dyn.init(dt)
    setup = time.time() - t0
    dyn.step(args.warmup)
    dyn.synchronize()
    t0 = time.perf_counter()
    gpu_ms = dyn.time_steps(args.steps)
    dyn.synchronize()
    wall_s = time.perf_counter() - t0
    prof = dyn.profile_kernels(args.profile_steps)
    dyn.synchronize()
    info = dyn.info()
    xg, *_ = dyn.fetch()
    energy, step = dyn.energy()
    _, ep = dyn.element_state()
    stats = dyn.surrogate_stats() if net is not None else None
    dyn.close()
    ne = info.num_elements
    elem_name = [k for k in prof if k.startswith("elem")][0]
    node_name = [k for k in prof if k.startswith("node")][0]
    elem_ms = prof[elem_name][0] / prof[elem_name][1]
    node_ms = prof[node_name][0] / prof[node_name][1]
    # the surrogate mode's element pass is two launches: the elastic pass
    # (every block's traffic) and the deferred-block walk, timed one by one
    pass_ms = prof["elem_pass"][0] / prof["elem_pass"][1] if "elem_pass" in prof else None
    walk_ms = prof["mlp_walk"][0] / prof["mlp_walk"][1] if "mlp_walk" in prof else None
    return dict(
        dt=info.dt, setup_s=setup, wall_s=wall_s, gpu_ms=gpu_ms,
        value=ne * args.steps / wall_s, ms_per_step=1e3 * wall_s / args.steps,
        gpu_ms_per_step=gpu_ms / args.steps,
        elem_kernel=elem_name, elem_ms=elem_ms, node_ms=node_ms, pass_ms=pass_ms, walk_ms=walk_ms,
        elem_bytes=formula_elem_bytes(info, precision), model_bytes=info.elem_kernel_bytes,
        node_bytes=info.node_kernel_bytes, step_bytes=info.bytes_per_step, x=xg, energy=energy.tolist(), step=step,
        epbar_max=float(np.max(ep)), blocks=info.num_blocks,
        boundary_nodes=info.num_boundary_nodes, colors=info.num_colors, ep=ep,
        surrogate_points=stats,
    )


def run_c3(args, precision, dt):
    """The thin-walled cylinder: R 1 m, L 4 m, t 0.02 m, 360 x 185 x 3 cells
    x 6 tet10 (1,198,800 elements), oblique 7 m/s impact 30 degrees from the
    wall normal, J2 aluminium, 4-point rule, exact return."""
    import math

    import cppimpact_amd as ci
    from cppimpact_amd import _lib as L

    E = 68.9e9
    x, conn = ci.cylinder_mesh(360, 185, 3, 1.0, 4.0, 0.02, origin=(0.0, 0.0, 1e-6),
                               quadratic=True)
    a = math.radians(30.0)
    mat = ci.Elastoplastic(E, 2700.0, 0.33, 1.0, 700e6, 276e6)
    wall = ci.Wall("Wall", 0.0, 10 * E * 0.02 / 3, None, 0, 1)
    dyn = ci.Dynamics(ci.Mesh.from_arrays(x, conn), mat, wall, physics=L.PHYSICS_EP,
                      precision=precision)
    dyn.vel[:] = (7.0 * math.sin(a), 0.0, -7.0 * math.cos(a))
    dyn.init(dt)
    dyn.step(10)
    dyn.synchronize()
    t0 = time.perf_counter()
    dyn.time_steps(args.c3_steps)
    dyn.synchronize()
    wall_s = time.perf_counter() - t0
    prof = dyn.profile_kernels(5)
    dyn.synchronize()
    info = dyn.info()
    xg, *_ = dyn.fetch()
    dyn.close()
    elem = [v[0] / v[1] for k, v in prof.items() if k.startswith("elem")][0]
    n = conn.shape[0]
    return dict(x0=x, x=xg, dt=info.dt, elements=n,
                value=n * args.c3_steps / wall_s, ms_per_step=1e3 * wall_s / args.c3_steps,
                elem_ms=elem, elem_gbs=formula_elem_bytes(info, precision) / (elem * 1e-3) / 1e9)


def c3_summary(args):
    import numpy as np

    from cppimpact_amd import _lib as L

    f64 = run_c3(args, L.PREC_F64, 0.0)
    f32 = run_c3(args, L.PREC_F32, f64["dt"])
    disp = np.abs(f64["x"] - f64["x0"]).max()
    derr = float(np.abs(f32["x"] - f64["x"]).max() / disp) if disp > 0 else 0.0
    row = lambda r: {"value": round(r["value"], 1), "ms_per_step": round(r["ms_per_step"], 5),
                     "elem_kernel_ms": round(r["elem_ms"], 5),
                     "elem_algorithmic_gbs": round(r["elem_gbs"], 1)}
    return {"workload": f"cylinder R 1 m, L 4 m, t 0.02 m, 360x185x3 cells x 6 tet10 "
                        f"({f64['elements']} elements), oblique 7 m/s at 30 deg, 4-point rule, "
                        f"J2 ep, dt=0.5*stable, {args.c3_steps} steps after 10",
            "unit": "element-steps/s", "plain_fp64": row(f64), "fp32": row(f32),
            "speedup_fp32_vs_fp64": round(f32["value"] / f64["value"], 4),
            "disp_rel_err_fp32_vs_fp64": derr}


def roofline(r):
    """The dominant kernel against the HBM roofline: the element kernel, or
    in the surrogate mode its elastic pass (the deferred-block walk after it
    reads one flag per block when nothing yields; its time is reported
    beside, and the two together are the element pass)."""
    kernel, ms = r["elem_kernel"], r["elem_ms"]
    if r.get("pass_ms"):
        kernel, ms = "elem_pass", r["pass_ms"]
    achieved = r["elem_bytes"] / (ms * 1e-3) / 1e9
    traffic = None
    path = os.path.join(ROOT, "profiles", "elem_block_cube64_traffic.json")
    if os.path.exists(path):
        try:
            traffic = json.load(open(path)).get(r["elem_kernel_key"])
        except Exception:
            traffic = None
    out = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
           "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
           "kernel": kernel, "kernel_ms": round(ms, 5),
           "algorithmic_bytes": r["elem_bytes"],
           "kernel_model_bytes": r["model_bytes"],
           "kernel_model_gbs": round(r["model_bytes"] / (ms * 1e-3) / 1e9, 1)}
    if r.get("pass_ms"):
        out["element_pass_ms"] = round(r["elem_ms"], 5)
        out["deferred_walk_ms"] = round(r["walk_ms"], 5)
    return out


def cpu_baseline(x, conn, dt, nsteps):
    import numpy as np

    from oracle import pyoracle as O
    E = 68.9e9
    wall = dict(location=0.0, stiffness=10 * E * SIDE / N_HEX)
    s = O.Solver(x, conn, (0.0, 0.0, -10.0), E=E, nu=0.33, rho=2700.0, physics=O.COMPAT,
                 wall=wall, dt=dt)
    out = s.solve(nsteps, with_energy=False)
    secs = out["loop_seconds"]
    return {"value": round(conn.shape[0] * nsteps / secs, 1), "unit": "element-steps/s",
            "cores": 1, "kind": "port",
            "sample": f"{nsteps} loop iteration(s) of the reference algorithm (compat: per-element "
                      f"K_e = sum_q w detJ B^T D B on the 5-point rule, f = K_e du) over the full "
                      f"{conn.shape[0]}-element mesh, 1 thread, initialisation excluded "
                      f"({secs:.1f} s)",
            "host_cpu": _cpu_model(), "host_nproc": os.cpu_count()}


def cpu_baseline_threaded(x, conn, dt, nsteps, threads, speed):
    """The oracle's FP64 elastoplastic step (the same physics as the plain
    FP64 GPU run) with its element and node loops on several threads."""
    from oracle import pyoracle as O
    E = 68.9e9
    wall = dict(location=0.0, stiffness=10 * E * SIDE / N_HEX)
    s = O.Solver(x, conn, (0.0, 0.0, -speed), E=E, nu=0.33, rho=2700.0, sigma_y=276e6,
                 H=700e6, physics=O.EP, wall=wall, dt=dt)
    out = s.solve(nsteps, with_energy=False, threads=threads)
    secs = out["loop_seconds"]
    return {"value": round(conn.shape[0] * nsteps / secs, 1), "unit": "element-steps/s",
            "cores": threads, "kind": "port",
            "sample": f"{nsteps} loop iteration(s) of the FP64 elastoplastic step (J2 radial "
                      f"return, centroid rule, HRZ mass) over the full {conn.shape[0]}-element "
                      f"mesh on {threads} threads, initialisation excluded ({secs:.1f} s)"}


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def main():
    args = parse()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} / WORLD_SIZE {world} "
              "is not supported (the step would shard by element partition with an "
              "interface-force exchange, not built here)", file=sys.stderr)
        sys.exit(2)
    import numpy as np

    from cppimpact_amd import _lib as L

    import ctypes
    gbs = ctypes.c_double(0.0)
    L.check(L.lib().cpi_bandwidth_probe(0, 1 << 30, 20, ctypes.byref(gbs)))
    print(f"bench: stream copy {gbs.value:.0f} GB/s", file=sys.stderr, flush=True)
    x, conn = make_mesh(args.n)
    print(f"bench: mesh {conn.shape[0]} tet4, {x.shape[0]} nodes", file=sys.stderr, flush=True)
    # one dt for both runs: half the stable step of the FP64 context
    plain = run_mode(x, conn, args, L.PREC_F64, L.RETURN_EXACT, 0.0)
    plain["elem_kernel_key"] = "f64"
    print(f"bench: fp64 {plain['value']:.4g} el-steps/s, elem kernel {plain['elem_ms']:.4f} ms",
          file=sys.stderr, flush=True)
    net, train = None, None
    if not args.no_surrogate:
        net, train = train_surrogate(args)
        print(f"bench: surrogate trained on {train['samples']} samples in "
              f"{train['train_s']} s (loss {train['loss_first']:.3g} -> {train['loss_last']:.3g})",
              file=sys.stderr, flush=True)
    ret = L.RETURN_EXACT if args.no_surrogate else L.RETURN_MLP
    mixed = run_mode(x, conn, args, L.PREC_F32, ret, plain["dt"], net)
    mixed["elem_kernel_key"] = "f32" if args.no_surrogate else "f32_mlp"
    print(f"bench: fp32 {mixed['value']:.4g} el-steps/s, elem kernel {mixed['elem_ms']:.4f} ms",
          file=sys.stderr, flush=True)
    disp = np.abs(plain["x"] - x).max()
    derr = float(np.abs(mixed["x"] - plain["x"]).max() / disp) if disp > 0 else 0.0
    epmax = float(np.abs(plain["ep"]).max())
    eperr = float(np.abs(mixed["ep"] - plain["ep"]).max() / epmax) if epmax > 0 else 0.0
    c3 = None
    if not args.no_c3:
        c3 = c3_summary(args)
        print(f"bench: c3 cylinder fp64 {c3['plain_fp64']['value']:.4g}, fp32 "
              f"{c3['fp32']['value']:.4g} el-steps/s", file=sys.stderr, flush=True)
    cpu, cpu_mt = None, None
    if not args.no_cpu:
        cpu = cpu_baseline(x, conn, plain["dt"], args.cpu_steps)
        print(f"bench: cpu baseline {cpu['value']:.4g} el-steps/s", file=sys.stderr, flush=True)
        nth = args.cpu_threads or min(16, os.cpu_count() or 1)
        cpu_mt = cpu_baseline_threaded(x, conn, plain["dt"], args.cpu_ep_steps, nth, args.speed)
        print(f"bench: cpu ep baseline, {nth} threads: {cpu_mt['value']:.4g} el-steps/s",
              file=sys.stderr, flush=True)
    line = {
        "metric": "element-steps/s, 1 MI355X: FP32+surrogate vs FP64 plain at <=1e-3 disp err",
        "value": round(mixed["value"], 1),
        "unit": "element-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(mixed["ms_per_step"], 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"cube {args.n}^3 hexes x 6 tet4 ({conn.shape[0]} elements, "
                        f"{x.shape[0]} nodes), jitter 0.1 seed 0, J2 ep, v0 -{args.speed:g} "
                        f"m/s, wall k=10*E*h, dt=0.5*stable",
            "mode": "fp32 element kernels" + ("" if args.no_surrogate else " + MLP surrogate")
                    + ", fp64 nodal state/mass/dt/energy",
            "scatter": args.scatter,
            "dt": plain["dt"],
        },
        "roofline": roofline(mixed),
        "cpu_baseline": cpu,
        "cpu_baseline_threaded": cpu_mt,
        "plain_fp64": {"value": round(plain["value"], 1),
                       "ms_per_step": round(plain["ms_per_step"], 5),
                       "roofline": roofline(plain)},
        "speedup_vs_fp64": round(mixed["value"] / plain["value"], 4),
        "disp_rel_err_vs_fp64": derr,
        "epbar_rel_err_vs_fp64": eperr,
        "surrogate": None if net is None else dict(
            train, points_accepted=mixed["surrogate_points"][0],
            points_fell_back=mixed["surrogate_points"][1]),
        "gpu_event_ms_per_step": round(mixed["gpu_ms_per_step"], 5),
        "hbm_stream_copy_gbs": round(gbs.value, 1),
        "node_kernel_ms": round(mixed["node_ms"], 5),
        "blocks": mixed["blocks"], "boundary_nodes": mixed["boundary_nodes"],
        "epbar_max": mixed["epbar_max"],
        "cylinder_c3": c3,
    }
    if net is not None:
        net.close()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
