"""Headline benchmark: Poisson solves/s of the FLUPS solve path on one MI355X.

Workload: cell-centred 256^3 on [0,1]^3 (h = 1/256), all six faces unbounded
(zero-padded 512^3 transforms), HEJ_2 kernel, lda = 1, FP64, right-hand side
= 4 seeded Gaussian blobs (centres U[0.3,0.7]^3, sigma U[0.03,0.06],
amplitudes N(0,1)), device-resident buffers. One step = one flups_solve.

Prints one JSON line with value = solves/s, a roofline block for the dominant
kernel (per-kernel HIP events on the solve stream; algorithmic bytes from the
pass shapes) and a CPU baseline (the oracle's C++ FP64 solver on the host
cores, bounded sample).
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM = 8.0e12   # MI355X HBM3E, bytes/s
METRIC = "Poisson solves/s + % HBM roofline at 256³ UNB; p99 latency of served mix"


def blobs(n, seed):
    rng = np.random.default_rng(seed)
    x = (np.arange(n) + 0.5) / n
    f = np.zeros((n, n, n))
    for _ in range(4):
        c = rng.uniform(0.3, 0.7, 3)
        s = rng.uniform(0.03, 0.06)
        a = rng.normal()
        gx = np.exp(-(x - c[0]) ** 2 / (2 * s * s))
        gy = np.exp(-(x - c[1]) ** 2 / (2 * s * s))
        gz = np.exp(-(x - c[2]) ** 2 / (2 * s * s))
        f += a * gz[:, None, None] * gy[None, :, None] * gx[None, None, :]
    return f[None]


def host_cpus():
    """the host's CPU topology (lscpu) and the threads the CPU baseline uses: the CPUs this process may run
    on, capped by OMP_NUM_THREADS when the environment sets it"""
    import subprocess
    info = {}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            k, v = k.strip(), v.strip()
            if k in ("Socket(s)", "Core(s) per socket", "Thread(s) per core", "Model name", "CPU(s)"):
                info[k] = v
    except Exception:
        pass
    avail = len(os.sched_getaffinity(0))
    omp = os.environ.get("OMP_NUM_THREADS")
    used = min(avail, int(omp)) if omp and omp.isdigit() and int(omp) > 0 else avail
    return {"lscpu_sockets": info.get("Socket(s)"), "lscpu_cores_per_socket": info.get("Core(s) per socket"),
            "lscpu_threads_per_core": info.get("Thread(s) per core"), "lscpu_cpus": info.get("CPU(s)"),
            "model": info.get("Model name"), "affinity_cpus": avail, "threads_used": used}


def cpu_baseline(n, kernel, budget_s, threads):
    """oracle solver timed on the host: setup excluded, 1 warm-up, then timed solves for about `budget_s`
    seconds (at least 3)"""
    from tests import oracle_lib as orc
    os.environ["OMP_NUM_THREADS"] = str(threads)   # read by the OpenMP runtime when the oracle loads
    rhs = blobs(n, 2024)
    field = np.zeros_like(rhs)
    fb = orc.flat_bc([[4, 4]] * 3, 1)
    ng = np.array([n] * 3, dtype=np.int32)
    sec = np.zeros(1)
    reps = np.zeros(1, dtype=np.int32)
    orc.check(orc.lib().oracle_solve_timed_for(orc._p(ng, ctypes.c_int), orc._p(fb, ctypes.c_int), 1,
                                               orc._p(np.array([1.0 / n] * 3)), orc._p(np.ones(3)), 1, 0, kernel, 2.0,
                                               orc.KDIR, 0, orc._p(rhs), orc._p(field), 1, budget_s, 3, 1000,
                                               orc._p(reps, ctypes.c_int), orc._p(sec)))
    return int(reps[0]) / sec[0], int(reps[0]), sec[0]


SERVE_SHAPES = [(64, 64, 64), (96, 128, 64), (128, 128, 128), (256, 256, 256), (1024, 1024, 1)]
SERVE_TEMPLATES = {"unbounded": [[4, 4], [4, 4], [4, 4]], "periodic-x": [[3, 3], [4, 4], [4, 4]],
                   "symmetric-x/periodic-y": [[0, 1], [3, 3], [4, 4]], "periodic": [[3, 3], [3, 3], [3, 3]]}
SERVE_KERNELS = {"HEJ_2": 2, "HEJ_6": 4, "LGF_4": 8}


def serve_mix(fl, nreq, rates, workers, seed=7):
    """Poisson arrivals of mixed requests through the library's serving loop (`workers` threads), against the
    plain path (one request at a time), at each arrival rate in `rates`: the same request sequence and
    arrival pattern scaled to the rate. Each server's solver cache is warmed (every request kind set up once)
    before the first rate; solves/s, p50 and p99 latency (submit to completion) per rate."""
    import threading
    rng = np.random.default_rng(seed)
    kinds = {}
    reqs = []
    for _ in range(nreq):
        ng = SERVE_SHAPES[rng.integers(len(SERVE_SHAPES))]
        tpl = list(SERVE_TEMPLATES)[rng.integers(len(SERVE_TEMPLATES))]
        kern = list(SERVE_KERNELS)[rng.integers(len(SERVE_KERNELS))]
        fp32 = bool(rng.integers(2))
        key = (ng, tpl, kern, fp32)
        if key not in kinds:
            bc = [list(b) for b in SERVE_TEMPLATES[tpl]]
            n = max(ng)
            h = [1.0 / n] * 3
            L = [h[0] * m for m in ng]
            if ng[2] == 1:
                bc[2] = [9, 9]
                L[2] = h[2]
            kinds[key] = fl.ServeRequest.make(ng, bc, h, L, green=SERVE_KERNELS[kern], fp32=fp32)
        reqs.append(key)
    gaps = rng.exponential(1.0, nreq)   # unit-rate Poisson gaps, scaled per rate
    # device buffers: one input per shape, a pool of outputs per shape
    shapes = sorted({k[0] for k in reqs})
    inputs, pools = {}, {}
    for ng in shapes:
        n0, n1, n2 = ng
        x = [(np.arange(m) + 0.5) / max(ng) for m in ng]
        f = np.exp(-((x[2][:, None, None] - 0.45) ** 2 + (x[1][None, :, None] - 0.5) ** 2 +
                     (x[0][None, None, :] - 0.55) ** 2) / (2 * 0.04 ** 2))[None]
        inputs[ng] = fl.DeviceArray.from_numpy(f)
        pools[ng] = [fl.DeviceArray(f.shape) for _ in range(workers + 2)]

    def run(srv, nw, rate):
        arrivals = np.cumsum(gaps / rate)
        srv.reset_stats()
        free = {ng: list(range(len(pools[ng]))) for ng in shapes}
        cv = threading.Condition()
        pending = []

        def waiter():
            for i in range(nreq):
                with cv:
                    while len(pending) <= i:
                        cv.wait()
                    t, ng, b = pending[i]
                srv.wait(t)
                with cv:
                    free[ng].append(b)
                    cv.notify_all()

        th = threading.Thread(target=waiter)
        th.start()
        t0 = time.perf_counter()
        for i, key in enumerate(reqs):
            dt = t0 + arrivals[i] - time.perf_counter()
            if dt > 0:
                time.sleep(dt)
            ng = key[0]
            with cv:
                while not free[ng]:
                    cv.wait()
                b = free[ng].pop()
            t = srv.submit(kinds[key], inputs[ng], pools[ng][b])
            with cv:
                pending.append((t, ng, b))
                cv.notify_all()
        th.join()
        st = srv.stats()
        return {"p50_ms": round(st["p50_ms"], 3), "p99_ms": round(st["p99_ms"], 3), "mean_ms": round(st["mean_ms"], 3),
                "solves_per_s": round(st["solves_per_s"], 1), "completed": int(st["completed"]),
                "failed": int(st["failed"]), "workers": nw}

    results = {r: {"rate_per_s": r} for r in rates}
    for label, nw in (("served", workers), ("plain_one_at_a_time", 1)):
        srv = fl.Server(160 << 30, workers=nw, solvers_per_key=1)
        for key, r in kinds.items():   # warm the solver cache
            srv.wait(srv.submit(r, inputs[key[0]], pools[key[0]][0]))
        for r in rates:
            results[r][label] = run(srv, nw, r)
        srv.close()
    return {"requests": nreq, "kinds": len(kinds),
            "mix": "shapes 64^3, 96x128x64, 128^3, 256^3, 2-D 1024^2; BC templates unbounded, periodic-x, "
                   "symmetric-x/periodic-y, periodic; HEJ_2/HEJ_6/LGF_4; FP32 or FP64; Poisson arrivals",
            "sweep": [results[r] for r in rates]}


def rhs_mixed(n, seed):
    """seeded N(0,1) field smoothed by a Gaussian of width 3h, mean removed along the periodic direction (y)"""
    from scipy.ndimage import gaussian_filter
    rng = np.random.default_rng(seed)
    f = gaussian_filter(rng.normal(size=(n, n, n)), sigma=3.0, mode="wrap")
    f -= f.mean(axis=1, keepdims=True)
    return f[None]


def rhs_ring(n, seed, h):
    """vorticity of a vortex ring (radius 0.2, Gaussian core sigma 0.03, seeded centre and axis) on the node grid
    x_i = i h, plus 1e-3 N(0,1) noise; shape (3, n, n, n)"""
    rng = np.random.default_rng(seed)
    c = rng.uniform(0.4, 0.6, 3)
    a = rng.normal(size=3)
    a /= np.linalg.norm(a)
    R, sig = 0.2, 0.03
    x = np.arange(n) * h
    out = 1e-3 * rng.normal(size=(3, n, n, n))
    # the core is negligible (exp(-36)) beyond R + 6 sigma from the centre: work in that box only
    rb = R + 6 * sig
    lo = np.clip(np.floor((c - rb) / h).astype(int), 0, n - 1)
    hi = np.clip(np.ceil((c + rb) / h).astype(int) + 1, 1, n)
    Y, X = np.meshgrid(x[lo[1]:hi[1]], x[lo[0]:hi[0]], indexing="ij")
    for k in range(lo[2], hi[2]):   # one z plane at a time keeps the temporaries small
        d = np.stack([X - c[0], Y - c[1], np.full_like(X, x[k] - c[2])])
        za = np.tensordot(a, d, axes=1)
        rho = d - a[:, None, None] * za
        rn = np.sqrt(np.sum(rho * rho, axis=0)) + 1e-300
        s2 = (rn - R) ** 2 + za ** 2
        w = np.exp(-s2 / sig ** 2) / (np.pi * sig ** 2)
        t = np.cross(a[:, None, None], rho, axis=0) / rn
        out[:, k, lo[1]:hi[1], lo[0]:hi[0]] += w * t
    return out


SETTLE_MS = 150.0   # untimed back-to-back work before every timed block (--settle-ms; see settle())


def run_config(fl, name, nglob, bc, center, kernel, lda, typ, odiff, rhs, h, L, steps, warmup, fp32=False, ref=None):
    """time one workload; with `ref` (the FP64 result on the device) also its relative L2 / Linf error against it"""
    topo = fl.Topology(0, lda, list(nglob), [1, 1, 1], False)
    s = fl.Solver(topo, bc, h, L, order_diff=odiff, center_type=[center] * 3)
    s.set_green_type(kernel)
    if fp32:
        s.set_precision(True)
    t0 = time.time()
    s.setup()
    setup_s = time.time() - t0
    din = fl.DeviceArray.from_numpy(rhs)
    dout = fl.DeviceArray(rhs.shape)
    stream = fl.Stream()   # solves queued back to back on a caller stream, as for the headline
    s.set_stream(stream)
    for _ in range(warmup):
        s.solve(dout, din, typ)
    stream.synchronize()
    nsettle = settle(lambda: s.solve(dout, din, typ), stream, SETTLE_MS)
    t0 = time.perf_counter()
    for _ in range(steps):
        s.solve(dout, din, typ)
    stream.synchronize()
    el = (time.perf_counter() - t0) / steps
    s.set_stream(None)
    by = s.solve_bytes(typ)
    out = {"workload": name, "solves_per_s": round(1.0 / el, 3), "ms_per_step": round(el * 1e3, 3),
           "algorithmic_GB": round(by / 1e9, 3), "achieved_GBs": round(by / el / 1e9, 1),
           "roofline_frac": round(by / el / PEAK_HBM, 4), "path": ["general", "fused-fp64", "fused-fp32", "general+fused-z", "general-fp32",
                    "general-fp32+fused-z"][s.path],
           "setup_s": round(setup_s, 2), "steps": steps, "settle_solves": nsettle}
    if ref is not None:
        a, b = dout.numpy(), ref.numpy()
        out["rel_l2_vs_fp64"] = float(np.sqrt(np.sum((a - b) ** 2) / np.sum(b * b)))
        out["rel_linf_vs_fp64"] = float(np.max(np.abs(a - b)) / np.max(np.abs(b)))
        del a, b
    del s
    return out, dout


def settle(solve, stream, ms):
    """untimed solves queued back to back, in rounds of 10, until at least `ms` milliseconds of them
    have run, so that the timed solves start from the steady state: after a few warm-up solves from
    idle, the kernels' durations drift for tens of milliseconds of sustained load (256^3 kernel
    trace: a solve spans 0.78 ms right after 5 warm-up solves, 0.87 ms a few solves later, 0.73-0.74
    ms later on; 1312-1315 solves/s timed after 5 warm-up solves, 1353-1362 after 100); returns the
    number of solves run"""
    n = 0
    t0 = time.perf_counter()
    while ms > 0:
        for _ in range(10):
            solve()
        n += 10
        stream.synchronize()
        if (time.perf_counter() - t0) * 1e3 >= ms:
            break
    return n


def blob_rhs(shape, seed):
    """lda components, each a sum of 4 Gaussian blobs (centres U[0.3,0.7] of the box, sigma U[0.03,0.06],
    amplitudes N(0,1)) on the cell grid of a box whose longest side is 1"""
    lda, n2, n1, n0 = shape
    rng = np.random.default_rng(seed)
    n = max(n0, n1, n2)
    xs = [(np.arange(m) + 0.5) / n for m in (n0, n1, n2)]
    out = np.zeros(shape)
    for c in range(lda):
        for _ in range(4):
            cen = rng.uniform(0.3, 0.7, 3) * np.array([n0, n1, n2]) / n
            sg = rng.uniform(0.03, 0.06)
            g = [np.exp(-(xs[d] - cen[d]) ** 2 / (2 * sg * sg)) for d in range(3)]
            out[c] += rng.normal() * g[2][:, None, None] * g[1][None, :, None] * g[0][None, None, :]
    return out


def batch_config(fl, nprob, seed, reps, same=False):
    """independent problems solved by one hipGraph launch of grouped kernels (flups_batch_run: one launch per
    kernel, step and grid over the problems), by one hipGraph launch with a branch per problem
    (flups_batch_run_branches), and by the plain path, the same solves one after the other with a
    synchronisation each (flups_batch_run_sequential)"""
    rng = np.random.default_rng(seed)
    sizes = [32, 48, 64, 80, 96, 112, 128]
    tpls = {"unbounded": [[4, 4]] * 3, "periodic-x/unbounded-yz": [[3, 3], [4, 4], [4, 4]],
            "symmetric-x/periodic-y/unbounded-z": [[0, 1], [3, 3], [4, 4]], "periodic": [[3, 3]] * 3}
    names = list(tpls)
    reqs, dins, douts, kinds = [], [], [], {}
    for p in range(nprob):
        ng = [int(rng.choice(sizes)) for _ in range(3)]
        while len(set(ng)) == 1:   # unequal axes
            ng = [int(rng.choice(sizes)) for _ in range(3)]
        tpl = names[int(rng.integers(len(names)))]
        lda = 3 if rng.random() < 0.25 else 1
        fp32 = bool(rng.random() < 0.5)
        if same:   # one problem description, nprob right-hand sides
            ng, tpl, lda, fp32 = [64, 48, 80], "periodic-x/unbounded-yz", 1, False
        typ = 1 if lda == 3 else 0   # three-component problems are Biot-Savart (ROT, spectral derivative)
        n = max(ng)
        h = [1.0 / n] * 3
        L = [h[0] * m for m in ng]
        rhs = blob_rhs((lda, ng[2], ng[1], ng[0]), seed * 1000 + p)
        if tpl == "periodic":
            rhs -= rhs.mean(axis=(1, 2, 3), keepdims=True)
        elif tpl.startswith("symmetric"):
            rhs -= rhs.mean(axis=2, keepdims=True)
        reqs.append(fl.ServeRequest.make(ng, tpls[tpl], h, L, green=2, lda=lda, fp32=fp32, solver_type=typ,
                                         order_diff=1 if typ else 0))
        dins.append(fl.DeviceArray.from_numpy(rhs))
        douts.append(fl.DeviceArray(rhs.shape))
        key = "%s/lda%d/%s" % (tpl, lda, "fp32" if fp32 else "fp64")
        kinds[key] = kinds.get(key, 0) + 1
    t0 = time.time()
    b = fl.Batch(reqs, dins, douts)
    setup_s = time.time() - t0
    for _ in range(2):
        b.run()
        b.run_branches()
        b.run_sequential()

    def settle_batch(fn):   # as settle(): untimed back-to-back batches for at least SETTLE_MS
        t = time.perf_counter()
        while (time.perf_counter() - t) * 1e3 < SETTLE_MS:
            fn()

    settle_batch(b.run)
    t0 = time.perf_counter()
    ev = []
    for _ in range(reps):
        ev.append(b.run())
    graph = (time.perf_counter() - t0) / reps
    settle_batch(b.run_branches)
    t0 = time.perf_counter()
    evb = []
    for _ in range(reps):
        evb.append(b.run_branches())
    branches = (time.perf_counter() - t0) / reps
    t0 = time.perf_counter()
    for _ in range(reps):
        b.run_sequential()
    seq = (time.perf_counter() - t0) / reps
    wl = ("%d problems of one description (64 x 48 x 80, periodic x, unbounded y and z, lda 1, FP64, HEJ_2), "
          "Gaussian-blob right-hand sides" % nprob) if same else (
        "%d independent problems, shapes from {32..128}^3 with unequal axes, 4 BC templates, lda 1 or 3 "
        "(lda 3: ROT), FP32 or FP64, HEJ_2, Gaussian-blob right-hand sides" % nprob)
    out = {"workload": wl,
           "problems": nprob, "mix": kinds, "graph_ms_per_batch": round(graph * 1e3, 3),
           "graph_device_ms": round(float(np.median(ev)), 3), "solves_per_s_graph": round(nprob / graph, 1),
           "launches_grouped": b.launches(True), "launches_per_problem": b.launches(False),
           "branches_ms_per_batch": round(branches * 1e3, 3), "branches_device_ms": round(float(np.median(evb)), 3),
           "sequential_ms_per_batch": round(seq * 1e3, 3), "solves_per_s_sequential": round(nprob / seq, 1),
           "speedup": round(seq / graph, 3), "setup_s": round(setup_s, 2), "reps": reps,
           "device_MB": round(b.device_bytes() / 1e6, 1)}
    b.close()
    return out


def other_configs(fl, which):
    out = {}
    if "mixed" in which:
        n = 256
        out["mixed"], _ = run_config(fl, "cell 256^3, x EVEN-ODD, y PER, z UNB (padded 512), LGF_4, lda 1, FP64",
                                     [n] * 3, [[0, 1], [3, 3], [4, 4]], 1, 8, 1, 0, 0, rhs_mixed(n, 11),
                                     [1.0 / n] * 3, [1.0] * 3, 20, 2)
    if "ring" in which:
        n = 513
        h = 1.0 / (n - 1)
        rhs = rhs_ring(n, 5, h)
        out["ring"], ref = run_config(fl, "node 513^3 all faces UNB (padded 1024^3), HEJ_6, lda 3 ROT (SPE), FP64",
                                      [n] * 3, [[4, 4]] * 3, 0, 4, 3, 1, 1, rhs, [h] * 3, [1.0] * 3, 3, 1)
        # the FP32 transforms and multiply on the same problem, error against the FP64 result
        # (itself within 1e-12 of the oracle: tests/test_gpu_parity.py)
        out["ring_fp32"], d32 = run_config(fl, "node 513^3 all faces UNB (padded 1024^3), HEJ_6, lda 3 ROT (SPE), FP32",
                                           [n] * 3, [[4, 4]] * 3, 0, 4, 3, 1, 1, rhs, [h] * 3, [1.0] * 3, 3, 1,
                                           fp32=True, ref=ref)
        out["ring_fp32"]["speedup_vs_fp64"] = round(out["ring_fp32"]["solves_per_s"] / out["ring"]["solves_per_s"], 3)
        del ref, d32, rhs
    if "batch" in which:
        out["batch"] = batch_config(fl, 64, 7, 10)
        out["batch_same"] = batch_config(fl, 64, 7, 10, same=True)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--settle-ms", type=float, default=150.0,
                    help="after the warm-up, untimed back-to-back solves for at least this long (clock settling)")
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--kernel", type=int, default=2, help="Green's kernel (2 = HEJ_2)")
    ap.add_argument("--cpu-seconds", type=float, default=20.0,
                    help="time budget of the CPU baseline's timed solves (0 = skip)")
    ap.add_argument("--profile-steps", type=int, default=10)
    ap.add_argument("--fp32-steps", type=int, default=50, help="timed solves of the FP32 variant (0 = skip)")
    ap.add_argument("--serve-requests", type=int, default=300, help="requests of the served mix (0 = skip)")
    ap.add_argument("--serve-rates", default="400,2000,5000,10000",
                    help="Poisson arrival rates of the served mix (1/s, comma list)")
    ap.add_argument("--serve-workers", type=int, default=4)
    ap.add_argument("--configs", default="mixed,ring,batch",
                    help="other workloads to time (comma list of mixed, ring, batch; '' = none)")
    ap.add_argument("--copy-calib-mb", type=int, default=0,
                    help="also run one plain device-to-device copy of this many MiB (counter calibration)")
    args = ap.parse_args()
    global SETTLE_MS
    SETTLE_MS = args.settle_ms
    if args.gpus != 1:
        print("bench.py runs on one GPU only (this build has no multi-GPU path); got --gpus %d" % args.gpus,
              file=sys.stderr)
        sys.exit(2)

    import flups_amd as fl

    n = args.n
    topo = fl.Topology(0, 1, [n, n, n], [1, 1, 1], False)
    s = fl.Solver(topo, [[fl.BoundaryType.UNB, fl.BoundaryType.UNB]] * 3, [1.0 / n] * 3, [1.0] * 3,
                  center_type=[fl.CenterType.CELL_CENTER] * 3)
    s.set_green_type(args.kernel)
    t0 = time.time()
    s.setup()
    setup_s = time.time() - t0
    rhs = fl.DeviceArray.from_numpy(blobs(n, 2024))
    field = fl.DeviceArray(rhs.shape)
    assert fl.lib().flups_pointer_is_device(rhs.ptr) == 1

    if args.copy_calib_mb > 0:
        nb = args.copy_calib_mb << 20
        a, b = fl.DeviceArray((nb // 8,)), fl.DeviceArray((nb // 8,))
        fl.lib().flups_memcpy(b.ptr, a.ptr, nb)
        del a, b
    # throughput: solves queued back to back on a caller stream (flups_set_stream: flups_solve on device
    # arrays then returns once queued), the timed block bracketed by stream synchronisations
    stream = fl.Stream()
    s.set_stream(stream)
    for _ in range(args.warmup):
        s.solve(field, rhs, fl.SolverType.STD)
    stream.synchronize()
    settle_n = settle(lambda: s.solve(field, rhs, fl.SolverType.STD), stream, args.settle_ms)
    t_start = time.perf_counter()
    for _ in range(args.steps):
        s.solve(field, rhs, fl.SolverType.STD)
    stream.synchronize()
    elapsed = time.perf_counter() - t_start
    solves_per_s = args.steps / elapsed
    # latency: one solve at a time on the default stream (flups_solve returns with the result complete)
    s.set_stream(None)
    lat = []
    for _ in range(max(args.steps, 20)):
        t1 = time.perf_counter()
        s.solve(field, rhs, fl.SolverType.STD)
        lat.append(time.perf_counter() - t1)
    sync_solves_per_s = len(lat) / sum(lat)
    # host arrays: the same solve with the right-hand side and the field in host memory (PCIe copies inside)
    h_rhs = blobs(n, 2024)
    h_field = np.zeros_like(h_rhs)
    s.solve(h_field, h_rhs, fl.SolverType.STD)
    t0 = time.perf_counter()
    for _ in range(5):
        s.solve(h_field, h_rhs, fl.SolverType.STD)
    host_solves_per_s = 5 / (time.perf_counter() - t0)
    del h_rhs, h_field

    # dominant kernel: per-kernel HIP events on the solve stream over a profiled block
    L = fl.lib()
    L.flups_profile_enable.argtypes = [ctypes.c_void_p, ctypes.c_int]
    L.flups_profile_get.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_char_p, ctypes.c_int,
                                    ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_long),
                                    ctypes.POINTER(ctypes.c_double)]
    L.flups_profile_get.restype = ctypes.c_int
    L.flups_profile_enable(s._solver, 1)
    for _ in range(args.profile_steps):
        s.solve(field, rhs, fl.SolverType.STD)
    kernels = []
    i = 0
    while True:
        name = ctypes.create_string_buffer(64)
        ms, calls, by = ctypes.c_double(), ctypes.c_long(), ctypes.c_double()
        cnt = L.flups_profile_get(s._solver, i, name, 64, ctypes.byref(ms), ctypes.byref(calls), ctypes.byref(by))
        if cnt == 0 or i >= cnt:
            break
        avg = ms.value / max(calls.value, 1) * 1e-3
        kernels.append({"kernel": name.value.decode(), "avg_us": avg * 1e6, "bytes": by.value,
                        "gbs": by.value / avg / 1e9 if avg > 0 else 0.0})
        i += 1
    L.flups_profile_enable(s._solver, 0)
    roof = None
    if kernels:
        dom = max(kernels, key=lambda k: k["avg_us"])
        ach = dom["gbs"]
        traffic = None
        tf = os.path.join(ROOT, "profiles", "traffic_%s_%d_unb_fp64.json" % (dom["kernel"], n))
        if os.path.exists(tf):
            traffic = json.load(open(tf)).get("bytes_per_launch")
        roof = {"bound": "hbm", "achieved": round(ach, 1), "peak": PEAK_HBM / 1e9, "unit": "GB/s",
                "frac": round(ach * 1e9 / PEAK_HBM, 4), "traffic": traffic, "kernel": dom["kernel"],
                "algorithmic_bytes": dom["bytes"], "avg_us": round(dom["avg_us"], 2),
                # traffic is not measured in this run: it is the committed rocprofv3 figure of the same
                # kernel (FETCH_SIZE x 2.000 + WRITE_SIZE x 1.000 per launch, factors from a 1 GiB copy)
                "traffic_source": os.path.relpath(tf, ROOT) if traffic is not None else None}
    total_bytes = sum(k["bytes"] for k in kernels) if kernels else None

    # FP32 transforms + multiply on the same workload (FP64 caller buffers), error against the FP64 path
    fp32 = None
    if args.fp32_steps > 0:
        s32 = fl.Solver(topo, [[fl.BoundaryType.UNB, fl.BoundaryType.UNB]] * 3, [1.0 / n] * 3, [1.0] * 3,
                        center_type=[fl.CenterType.CELL_CENTER] * 3)
        s32.set_green_type(args.kernel)
        s32.set_precision(True)
        s32.setup()
        f32 = fl.DeviceArray(rhs.shape)
        s32.set_stream(stream)
        for _ in range(args.warmup):
            s32.solve(f32, rhs, fl.SolverType.STD)
        stream.synchronize()
        settle32 = settle(lambda: s32.solve(f32, rhs, fl.SolverType.STD), stream, args.settle_ms)
        t0 = time.perf_counter()
        for _ in range(args.fp32_steps):
            s32.solve(f32, rhs, fl.SolverType.STD)
        stream.synchronize()
        el32 = time.perf_counter() - t0
        s32.set_stream(None)
        L.flups_profile_enable(s32._solver, 1)
        for _ in range(args.profile_steps):
            s32.solve(f32, rhs, fl.SolverType.STD)
        k32 = []
        i = 0
        while True:
            name = ctypes.create_string_buffer(64)
            ms, calls, by = ctypes.c_double(), ctypes.c_long(), ctypes.c_double()
            cnt = L.flups_profile_get(s32._solver, i, name, 64, ctypes.byref(ms), ctypes.byref(calls), ctypes.byref(by))
            if cnt == 0 or i >= cnt:
                break
            avg = ms.value / max(calls.value, 1) * 1e-3
            k32.append({"kernel": name.value.decode(), "avg_us": round(avg * 1e6, 3), "bytes": by.value,
                        "gbs": round(by.value / avg / 1e9, 1) if avg > 0 else 0.0})
            i += 1
        L.flups_profile_enable(s32._solver, 0)
        a64, a32 = field.numpy(), f32.numpy()
        e2 = float(np.sqrt(np.sum((a32 - a64) ** 2) / np.sum(a64 ** 2)))
        ei = float(np.max(np.abs(a32 - a64)) / np.max(np.abs(a64)))
        v32 = args.fp32_steps / el32
        fp32 = {"value": round(v32, 3), "unit": "solves/s", "ms_per_step": round(el32 / args.fp32_steps * 1e3, 4),
                "speedup_vs_fp64": round(v32 / solves_per_s, 3), "rel_l2_vs_fp64": e2, "rel_linf_vs_fp64": ei,
                "path": "fused-fp32" if s32.path == 2 else "general", "steps": args.fp32_steps,
                "settle": {"ms": args.settle_ms, "solves": settle32},
                "schedule": list(s32.schedule), "kernels": k32}
        del s32, f32

    rates = [float(r) for r in args.serve_rates.split(",") if r]
    served = serve_mix(fl, args.serve_requests, rates, args.serve_workers) if args.serve_requests > 0 and rates else None
    others = other_configs(fl, [c for c in args.configs.split(",") if c]) if args.configs else None

    cpu = None
    if args.cpu_seconds > 0:
        topo_cpu = host_cpus()
        threads = topo_cpu["threads_used"]
        v, reps, sec = cpu_baseline(n, args.kernel, args.cpu_seconds, threads)
        cpu = {"value": round(v, 4), "unit": "solves/s", "cores": threads, "kind": "port",
               "sample": "%d timed solves of the same %d^3 unbounded HEJ_2 workload (setup excluded) after 1 warm-up, "
                         "%.1f s, OpenMP over pencils with %d threads" % (reps, n, sec, threads),
               "host": topo_cpu}

    lat_ms = np.array(lat) * 1e3
    out = {
        "metric": METRIC, "value": round(solves_per_s, 3), "unit": "solves/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "settle": {"ms": args.settle_ms, "solves": settle_n},
        "ms_per_step": round(elapsed / args.steps * 1e3, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded Gaussian blobs)",
        "config": {"workload": "cell %d^3, all faces UNB (padded %d^3), HEJ_2, lda=1, FP64, device-resident" % (n, 2 * n),
                   "nglob": [n, n, n], "kernel": "HEJ_2", "path": "fused" if s.path == 1 else "general"},
        "roofline": roof,
        "cpu_baseline": cpu,
        "solve_roofline_frac": round(total_bytes * solves_per_s / PEAK_HBM, 4) if total_bytes else None,
        "latency_ms": {"p50": round(float(np.percentile(lat_ms, 50)), 4), "p99": round(float(np.percentile(lat_ms, 99)), 4),
                       "solves": len(lat)},
        "sync_solves_per_s": round(sync_solves_per_s, 3),
        "host_pointer_solves_per_s": round(host_solves_per_s, 3),
        "kernels": [{k: (round(v, 3) if isinstance(v, float) else v) for k, v in kk.items()} for kk in kernels],
        "setup_s": round(setup_s, 3),
        "schedule": list(s.schedule),
        "fp32": fp32,
        "served_mix": served,
        "other_configs": others,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
