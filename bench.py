"""Benchmark: interactions/sec of one gravity force evaluation on one MI355X.

Workload (default): 1M equal-mass Plummer sphere (G=M=a=1, r <= 10a, seeded),
bucket 16, theta 0.7, softening 0.005, monopole+quadrupole.  One step = one full
force evaluation from positions already resident in HBM: universe box, Morton
keys, radix sort, octree build, upward pass, traversal lists, P-C + P-P
interactions (FP32 bucket-shifted Kahan form = the headline; the plain FP64
form is timed the same way for the ratio).

value = (n_part_ints + n_node_ints) / seconds per step (counter definitions of
src/Traverser.h:15-37), the steps timed with the stage timers off.  The
roofline object covers the interaction kernel: algorithmic FLOPs per launch
(20 per P-P pair, 67 per P-C particle-node interaction) over its average
duration, from HIP events recorded on the stream the kernel runs on during a
second pass of the same number of steps with the stage timers on; its traffic
is the counter pass recorded in profiles/plummer1m_kernels.md, next to the
bytes the kernel must move at least once.  Errors are measured against an
FP64 direct sum on 4096 seeded random targets.  cpu_baseline times the oracle
(CPU restatement) on the host cores through the same phases on the same
input: all threads over the whole step, one thread over a sample.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

CONFIGS = {
    # name: (n, bucket, theta, eps, generator)
    "plummer1m": (1_000_000, 16, 0.7, 0.005, "plummer"),
    "plummer100k": (100_000, 12, 0.7, 0.01, "plummer"),
    "halo16m": (16_000_000, 32, 0.7, 0.01, "halo"),
}
# HBM bytes per dispatch of grouped_gravity_kernel<FP32C> on the 1M workload: FETCH_SIZE
# + WRITE_SIZE from the two rocprofv3 counter passes recorded in this profile
# (KiB per dispatch; the counters are read with factor 1.0, calibrated against
# a device copy of known size).
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "plummer1m_kernels.md")
TRAFFIC_KERNEL = "grouped_gravity_kernel<true, 512, false, false>"
PEAK_FP32_TFLOPS = 157.3
PEAK_FP64_TFLOPS = 78.6
FLOPS_PP = 20.0
FLOPS_PC = 67.0


def profile_traffic(path=TRAFFIC_PROFILE, kernel=TRAFFIC_KERNEL):
    """FETCH_SIZE + WRITE_SIZE (bytes per dispatch) of `kernel` from the
    counter table of a committed profile, or None when it has no such row."""
    try:
        with open(path) as f:
            lines = f.read().splitlines()
    except OSError:
        return None
    head = None
    for line in lines:
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if "FETCH_SIZE (KiB)" in cells and "WRITE_SIZE (KiB)" in cells:
            head = cells
            continue
        if head and len(cells) == len(head) and cells[0] == kernel:
            fetch = float(cells[head.index("FETCH_SIZE (KiB)")])
            write = float(cells[head.index("WRITE_SIZE (KiB)")])
            return {"fetch_bytes": fetch * 1024, "write_bytes": write * 1024,
                    "bytes": (fetch + write) * 1024, "source": os.path.relpath(path, ROOT)}
    return None


def algorithmic_bytes_fp32c(n, nn, nb):
    """Bytes the fused FP32C kernel must move at least once per dispatch:
    the sorted particles' x, y, z, m, eps (40 B), each node's walk record
    (TreeNode, 64 B) and source record (GravNode32, 64 B), each bucket's
    record (128 B), and the results written (a, pot: 32 B per particle)."""
    return 40 * n + 128 * nn + 128 * nb + 32 * n


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def make_input(name):
    from paratreet_amd import ics
    n, bucket, theta, eps, gen = CONFIGS[name]
    if gen == "plummer":
        pos, vel, m = ics.plummer(n, seed=2024)
    else:
        pos, vel, m = ics.halo_mixture(n, n_halos=64, seed=4242)
    return pos, vel, m, np.full(n, eps), bucket, theta


def time_path(d, theta, precision, steps, warmup):
    """Warm-up, then `steps` force evaluations with the stage timers off (the
    step time: nothing but the pipeline in the timed region), then `steps`
    more with them on (HIP events around each stage on the context's stream:
    the per-stage and fused-kernel times).  Returns (ms per step, counters,
    mean force-stage ms, last stage times)."""
    import torch
    for _ in range(warmup):
        d.compute_forces(theta=theta, precision=precision)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    cnt = None
    for _ in range(steps):
        cnt = d.compute_forces(theta=theta, precision=precision)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms_step = (t1 - t0) * 1000.0 / steps
    d.set_timing(True)
    force_ms = []
    stage = None
    for _ in range(steps):
        d.compute_forces(theta=theta, precision=precision)
        tm = d.timings()
        force_ms.append(tm["force"])
        stage = tm
    d.set_timing(False)
    return ms_step, cnt, float(np.mean(force_ms)), stage


def time_list_path(d, theta, precision):
    """One evaluation through the list path (traverse -> lists -> interact),
    timed per stage, to separate the walk from the interactions."""
    from paratreet_amd.gravity import GravityVisitor
    d.remake_universe()
    d.assign_keys()
    d.build_tree()
    d.set_timing(True)
    d.traverse(GravityVisitor((0.0, 0.0, 0.0), theta))
    d.interact(precision)
    tm = d.timings()
    d.set_timing(False)
    return {"traverse_2pass": tm["traverse"], "interact": tm["force"]}


def cpu_baseline(pos, m, eps, bucket, theta, sample_frac=0.2):
    """The oracle (the CPU restatement, kind "port") on the host cores, the
    same phases as a GPU step: universe + keys, sort (with the gather of the
    sorted copies), tree build, moments (single-threaded C), and the force
    evaluation (OpenMP over buckets) with all threads over every bucket and
    with one thread over a seeded sample of the buckets.  value = whole-step
    interactions/s with all threads."""
    from oracle import oracle as O
    nthreads = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count()
    t = {}
    t0 = time.perf_counter()
    box = O.universe(pos[:, 0], pos[:, 1], pos[:, 2])
    keys = O.keys(pos[:, 0], pos[:, 1], pos[:, 2], box)
    t["keys"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    sk, perm = O.sort(keys)
    x, y, z = (np.ascontiguousarray(pos[perm, k]) for k in range(3))
    ms, es = m[perm].copy(), eps[perm].copy()
    t["sort"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    nodes = O.build_tree(sk, bucket)
    t["build"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    data = O.moments(nodes, x, y, z, ms, es)
    t["moments"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    _, _, cnt = O.gravity(nodes, data, x, y, z, ms, es, theta, nthreads=nthreads)
    t["force"] = time.perf_counter() - t0
    ints = float(cnt[0] + cnt[1])
    nb = len(O.buckets(nodes))
    rng = np.random.default_rng(5)
    sel = np.sort(rng.choice(nb, max(1, int(sample_frac * nb)), replace=False)).astype(np.int32)
    t0 = time.perf_counter()
    _, _, scnt = O.gravity(nodes, data, x, y, z, ms, es, theta, bucket_sel=sel, nthreads=1)
    t1 = time.perf_counter() - t0
    ips1 = float(scnt[0] + scnt[1]) / t1
    serial = t["keys"] + t["sort"] + t["build"] + t["moments"]
    step = serial + t["force"]
    return {"value": ints / step, "unit": "interactions/s", "cores": nthreads, "kind": "port",
            "sample": "whole step on the same %d-particle input (keys, sort, build and moments "
                      "single-threaded, %.3g s; force over all %d buckets with %d OpenMP "
                      "threads, %.3g s); one-thread force on %d seeded buckets (%.1f s)" % (
                          len(x), serial, nb, nthreads, t["force"], len(sel), t1),
            "cpu_model": cpu_model(), "logical_cpus": os.cpu_count(),
            "phase_ms": {k: v * 1e3 for k, v in t.items()},
            "force_ips_all_threads": ints / t["force"],
            "force_ips_1_thread": ips1,
            "step_ips_1_thread_est": ints / (serial + ints / ips1)}


ADJOINT_CONFIGS = {
    # name: (n, bucket, theta, eps, generator, steps, dt, checkpoint_every, precision)
    "plummer256k_adjoint": (262_144, 16, 0.7, 0.005, "plummer", 10, 1e-3, 0, "fp64"),
    "halo4m_adjoint": (4_194_304, 32, 0.7, 0.01, "halo16", 50, 5e-4, 10, "fp32c"),
}


def _loss_grad(d, pos, vel, m, eps, steps, dt, theta, precision, mode, every):
    """Forward + reverse sweep of L = sum m |x_T - x_0|^2 with the adjoint in
    the forward's precision; returns the gradients (x0, v0, m) and the
    (forward, adjoint) wall times in ms."""
    import torch
    d.set_adjoint_precision(precision)
    d.load(pos, m, eps, vel)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    d.leapfrog_forward(steps, dt, theta=theta, precision=precision, mode=mode,
                       checkpoint_every=every)
    t1 = time.perf_counter()
    xT = d.state()[0]
    lam = 2.0 * m[:, None] * (xT - pos)
    t2 = time.perf_counter()
    gx, gv, gm = d.leapfrog_adjoint(lam_x=lam)
    t3 = time.perf_counter()
    return (gx - lam, gv, gm + np.sum((xT - pos) ** 2, axis=1)), (t1 - t0) * 1e3, (t3 - t2) * 1e3


def adjoint_bench(name, fd_coords, direct):
    """Differentiable-trajectory workload: timing of the forward (with
    checkpoints) and the reverse sweep, gradient checks against central
    finite differences (h = 1e-6) and against the FP64 direct-sum gradient,
    and the device memory the context holds."""
    import torch
    from paratreet_amd import ics
    from paratreet_amd.gravity import (FORCE_DIRECT, FORCE_TREE, PT_FP32C, PT_FP64,
                                       Configuration, Driver)
    n, bucket, theta, eps0, gen, steps, dt, every, prec = ADJOINT_CONFIGS[name]
    if gen == "plummer":
        pos, vel, m = ics.plummer(n, seed=2025)
    else:
        pos, vel, m = ics.halo_mixture(n, n_halos=16, seed=4243)
    eps = np.full(n, eps0)
    precision = PT_FP32C if prec == "fp32c" else PT_FP64
    free0 = torch.cuda.mem_get_info()[0]
    d = Driver(Configuration(max_particles_per_leaf=bucket))
    _loss_grad(d, pos, vel, m, eps, steps, dt, theta, precision, FORCE_TREE, every)  # warm-up
    g, fwd_ms, adj_ms = _loss_grad(d, pos, vel, m, eps, steps, dt, theta, precision, FORCE_TREE,
                                   every)
    hbm = free0 - torch.cuda.mem_get_info()[0]
    out = {"workload": name, "n": n, "bucket": bucket, "theta": theta, "eps": eps0,
           "steps": steps, "dt": dt, "checkpoint_every": every or steps,
           "forward_precision": prec,
           "adjoint_arithmetic": "f32 pairs, f64 sums" if prec == "fp32c" else "f64",
           "forward_ms": fwd_ms, "adjoint_ms": adj_ms, "device_bytes_in_use": int(hbm)}
    if precision == PT_FP32C:
        # the same trajectory with FP64 forward and FP64 adjoint
        g64, f64_ms, a64_ms = _loss_grad(d, pos, vel, m, eps, steps, dt, theta, PT_FP64,
                                         FORCE_TREE, every)
        out["fp32c_vs_fp64_gradient_rel_l2"] = {
            k: float(np.linalg.norm(a - b) / np.linalg.norm(b))
            for k, a, b in zip(("x0", "v0", "m"), g, g64)}
        out["fp64_ms"] = {"forward": f64_ms, "adjoint": a64_ms}
    if precision == PT_FP32C:
        # FP32 force rounding (~1e-7 of |a| per step) swamps an h = 1e-6
        # difference of the loss; the FD check belongs to the FP64 workload
        fd_coords = 0
    if fd_coords:
        rng = np.random.default_rng(7)
        h = 1e-6
        errs = []
        scale = np.abs(g[0]).max()

        def final(p):
            d.load(p, m, eps, vel)
            d.leapfrog_forward(steps, dt, theta=theta, precision=precision, mode=FORCE_TREE,
                               checkpoint_every=every)
            xT = d.state()[0]
            return m * np.sum((xT - p) ** 2, axis=1)

        for c in rng.choice(pos.size, fd_coords, replace=False):
            i, k = divmod(int(c), 3)
            p1, p2 = pos.copy(), pos.copy()
            p1[i, k] += h
            p2[i, k] -= h
            fd = np.sum(final(p1) - final(p2)) / (2 * h)
            errs.append(abs(fd - g[0][i, k]) / scale)
        out["fd_x0_rel_err"] = {"coords": fd_coords, "h": h, "median": float(np.median(errs)),
                                "max": float(np.max(errs))}
    if direct:
        gd, dfwd, dadj = _loss_grad(d, pos, vel, m, eps, steps, dt, theta, PT_FP64, FORCE_DIRECT,
                                    every)
        out["direct_sum_gradient_rel_l2"] = {
            k: float(np.linalg.norm(a - b) / np.linalg.norm(b))
            for k, a, b in zip(("x0", "v0", "m"), g, gd)}
        out["direct_sum_ms"] = {"forward": dfwd, "adjoint": dadj}
    d.close()
    print(json.dumps(out))


def hexadecapole_bench(args):
    """The HEXADECAPOLE build (moments to fourth order, the fused walk with the
    hexadecapole open() and node()) on the same workload, compensated-FP32
    form and FP64: one JSON line of its own."""
    import torch
    from paratreet_amd.gravity import Configuration, Driver, PT_FP32C, PT_FP64
    pos, vel, m, eps, bucket, theta = make_input(args.config)
    if args.theta is not None:
        theta = args.theta
    n = len(pos)
    d = Driver(Configuration(max_particles_per_leaf=bucket, hexadecapole=True))
    d.load(pos, m, eps, vel)
    ms64, cnt, f64_ms, st64 = time_path(d, theta, PT_FP64, args.steps, args.warmup)
    acc64, _ = d.accelerations()
    ms, cnt32, f_ms, st = time_path(d, theta, PT_FP32C, args.steps, args.warmup)
    acc, _ = d.accelerations()
    ints = float(cnt[0] + cnt[1])
    e32 = np.linalg.norm(acc - acc64, axis=1) / np.linalg.norm(acc64, axis=1)
    err = {"fp32c_vs_fp64_rms": float(np.sqrt(np.mean(e32 * e32))),
           "fp32c_vs_fp64_p99": float(np.percentile(e32, 99))}
    if not args.no_error:
        rng = np.random.default_rng(99)
        tgt = np.sort(rng.choice(n, 4096, replace=False))
        ref = d.direct_sum(tgt)
        for name, a in (("fp32c", acc), ("fp64", acc64)):
            e = np.linalg.norm(a[tgt] - ref, axis=1) / np.linalg.norm(ref, axis=1)
            err[name + "_vs_direct_rms"] = float(np.sqrt(np.mean(e * e)))
            err[name + "_vs_direct_p99"] = float(np.percentile(e, 99))
            err[name + "_vs_direct_max"] = float(e.max())
    d.close()
    out = {
        "metric": "interactions/sec, HEXADECAPOLE build (FP32-comp vs FP64), %s" % args.config,
        "value": ints / (ms / 1000.0), "unit": "interactions/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 interactions (Kahan-compensated), f64 keys/tree/moments",
        "data": "synthetic (seeded generator)",
        "config": {"workload": "%s: n=%d, bucket=%d, theta=%.2f, eps=%g, moments to fourth "
                               "order, full force evaluation per step (fused walk + "
                               "hexadecapole node())" % (args.config, n, bucket, theta, eps[0]),
                   "n": n, "bucket": bucket, "theta": theta, "parallelism": "single GPU"},
        "interactions_per_step": ints, "n_part_ints": int(cnt[1]), "n_node_ints": int(cnt[0]),
        "counters_equal": [int(v) for v in cnt] == [int(v) for v in cnt32],
        "fp64_value": ints / (ms64 / 1000.0), "fp64_ms_per_step": ms64,
        "fp32c_over_fp64": ms64 / ms, "force_ms": {"fp32c": f_ms, "fp64": f64_ms},
        "stage_ms_fp32c": st, "stage_ms_fp64": st64, "errors": err,
    }
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="plummer1m", choices=sorted(CONFIGS))
    ap.add_argument("--theta", type=float, default=None,
                    help="opening angle override (the halo mixture is quoted at 0.5, 0.7, 0.9)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the cpu_baseline leg")
    ap.add_argument("--no-error", action="store_true", help="skip the direct-sum error leg")
    ap.add_argument("--adjoint", choices=sorted(ADJOINT_CONFIGS),
                    help="run the differentiable-trajectory workload instead")
    ap.add_argument("--fd-coords", type=int, default=64)
    ap.add_argument("--no-direct", action="store_true",
                    help="skip the direct-sum gradient leg of --adjoint")
    ap.add_argument("--hexadecapole", action="store_true",
                    help="time the HEXADECAPOLE build (FP64) on the workload instead")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only; --gpus %d is not supported "
              "(the multi-GPU sharding is a design, see the documentation)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)

    import torch
    from paratreet_amd.gravity import Configuration, Driver, PT_FP32C, PT_FP64
    torch.cuda.is_available()

    if args.adjoint:
        adjoint_bench(args.adjoint, args.fd_coords, not args.no_direct)
        return
    if args.hexadecapole:
        hexadecapole_bench(args)
        return

    pos, vel, m, eps, bucket, theta = make_input(args.config)
    if args.theta is not None:
        theta = args.theta
    n = len(pos)
    d = Driver(Configuration(max_particles_per_leaf=bucket))
    d.load(pos, m, eps, vel)

    ms32, cnt32, f32_ms, st32 = time_path(d, theta, PT_FP32C, args.steps, args.warmup)
    acc32, _ = d.accelerations()
    ms64, cnt64, f64_ms, st64 = time_path(d, theta, PT_FP64, args.steps, args.warmup)
    acc64, _ = d.accelerations()
    ints = float(cnt32[0] + cnt32[1])
    value = ints / (ms32 / 1000.0)
    value64 = ints / (ms64 / 1000.0)
    flops = FLOPS_PP * cnt32[1] + FLOPS_PC * cnt32[0]
    achieved32 = flops / (f32_ms / 1000.0) / 1e12
    achieved64 = flops / (f64_ms / 1000.0) / 1e12

    nn, nb, _ = d.tree_size()
    algo_bytes = algorithmic_bytes_fp32c(n, nn, nb)
    traffic = profile_traffic() if args.config == "plummer1m" else None
    list_path = {"fp32c": time_list_path(d, theta, PT_FP32C),
                 "fp64": time_list_path(d, theta, PT_FP64)}

    err = {}
    if not args.no_error:
        rng = np.random.default_rng(99)
        tgt = np.sort(rng.choice(n, 4096, replace=False))
        ref = d.direct_sum(tgt)
        nr = np.linalg.norm(ref, axis=1)
        for name, a in (("fp32c", acc32), ("fp64", acc64)):
            e = np.linalg.norm(a[tgt] - ref, axis=1) / nr
            err[name + "_vs_direct_rms"] = float(np.sqrt(np.mean(e * e)))
            err[name + "_vs_direct_p99"] = float(np.percentile(e, 99))
        e = np.linalg.norm(acc32 - acc64, axis=1) / np.linalg.norm(acc64, axis=1)
        err["fp32c_vs_fp64_rms"] = float(np.sqrt(np.mean(e * e)))
        err["fp32c_vs_fp64_p99"] = float(np.percentile(e, 99))
    d.close()

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(pos, m, eps, bucket, theta)

    out = {
        "metric": "interactions/sec, 1M Plummer θ=0.7, FP32-comp vs FP64 GPU; RMS force err",
        "value": value,
        "unit": "interactions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms32,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 interactions (Kahan-compensated), f64 keys/tree/moments",
        "data": "synthetic (seeded Plummer generator)",
        "config": {"workload": "%s: n=%d, bucket=%d, theta=%.2f, eps=%g, monopole+quadrupole, "
                               "full force evaluation per step" % (args.config, n, bucket, theta,
                                                                   eps[0]),
                   "n": n, "bucket": bucket, "theta": theta, "parallelism": "single GPU"},
        "interactions_per_step": ints,
        "n_part_ints": int(cnt32[1]), "n_node_ints": int(cnt32[0]),
        "fp64_value": value64, "fp64_ms_per_step": ms64,
        "fp32c_over_fp64": value / value64,
        # the definition of t_force in the scope table: upward pass + traversal +
        # interactions device time (keys, sort and build reported separately)
        "t_force_ms": {"fp32c": st32["moments"] + f32_ms, "fp64": st64["moments"] + f64_ms},
        "fp32c_over_fp64_t_force": (st64["moments"] + f64_ms) / (st32["moments"] + f32_ms),
        "force_kernel_ms": {"fp32c": f32_ms, "fp64": f64_ms},
        "force_only_ips": {"fp32c": ints / (f32_ms / 1000.0), "fp64": ints / (f64_ms / 1000.0)},
        "stage_ms_fp32c": st32, "stage_ms_fp64": st64,
        "list_path_ms": list_path,
        "roofline": {"bound": "valu", "achieved": achieved32, "peak": PEAK_FP32_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved32 / PEAK_FP32_TFLOPS,
                     "traffic": traffic["bytes"] if traffic else None,
                     "traffic_source": traffic,
                     "algorithmic_bytes": algo_bytes,
                     "kernel": "grouped_gravity_kernel<FP32C>",
                     "fp64_kernel": {"achieved": achieved64, "peak": PEAK_FP64_TFLOPS,
                                     "frac": achieved64 / PEAK_FP64_TFLOPS}},
        "errors": err,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
