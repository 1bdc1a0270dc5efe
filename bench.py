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


def cpu_baseline(n, kernel, reps, threads):
    """oracle solver timed on the host: setup excluded, 1 warm-up + `reps` timed solves"""
    from tests import oracle_lib as orc
    os.environ.setdefault("OMP_NUM_THREADS", str(threads))
    rhs = blobs(n, 2024)
    field = np.zeros_like(rhs)
    fb = orc.flat_bc([[4, 4]] * 3, 1)
    ng = np.array([n] * 3, dtype=np.int32)
    sec = np.zeros(1)
    orc.check(orc.lib().oracle_solve_timed(orc._p(ng, ctypes.c_int), orc._p(fb, ctypes.c_int), 1,
                                           orc._p(np.array([1.0 / n] * 3)), orc._p(np.ones(3)), 1, 0, kernel, 2.0,
                                           orc.KDIR, 0, orc._p(rhs), orc._p(field), 1, reps, orc._p(sec)))
    return reps / sec[0], sec[0]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--kernel", type=int, default=2, help="Green's kernel (2 = HEJ_2)")
    ap.add_argument("--cpu-reps", type=int, default=2, help="timed oracle solves for the CPU baseline (0 = skip)")
    ap.add_argument("--profile-steps", type=int, default=10)
    ap.add_argument("--copy-calib-mb", type=int, default=0,
                    help="also run one plain device-to-device copy of this many MiB (counter calibration)")
    args = ap.parse_args()
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
    for _ in range(args.warmup):
        s.solve(field, rhs, fl.SolverType.STD)
    # each flups_solve returns with its stream synchronised: the timed block starts and ends idle
    lat = []
    t_start = time.perf_counter()
    for _ in range(args.steps):
        t1 = time.perf_counter()
        s.solve(field, rhs, fl.SolverType.STD)
        lat.append(time.perf_counter() - t1)
    elapsed = time.perf_counter() - t_start
    solves_per_s = args.steps / elapsed

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
                "algorithmic_bytes": dom["bytes"], "avg_us": round(dom["avg_us"], 2)}
    total_bytes = sum(k["bytes"] for k in kernels) if kernels else None

    cpu = None
    if args.cpu_reps > 0:
        threads = int(os.environ.get("OMP_NUM_THREADS", str(min(16, os.cpu_count() or 1))))
        v, sec = cpu_baseline(n, args.kernel, args.cpu_reps, threads)
        cpu = {"value": round(v, 4), "unit": "solves/s", "cores": threads, "kind": "port",
               "sample": "%d solves of the same %d^3 unbounded HEJ_2 workload after 1 warm-up, %.1f s" %
                         (args.cpu_reps, n, sec)}

    lat_ms = np.array(lat) * 1e3
    out = {
        "metric": METRIC, "value": round(solves_per_s, 3), "unit": "solves/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(elapsed / args.steps * 1e3, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded Gaussian blobs)",
        "config": {"workload": "cell %d^3, all faces UNB (padded %d^3), HEJ_2, lda=1, FP64, device-resident" % (n, 2 * n),
                   "nglob": [n, n, n], "kernel": "HEJ_2", "path": "fused" if s.path == 1 else "general"},
        "roofline": roof,
        "cpu_baseline": cpu,
        "solve_roofline_frac": round(total_bytes * solves_per_s / PEAK_HBM, 4) if total_bytes else None,
        "latency_ms": {"p50": round(float(np.percentile(lat_ms, 50)), 4), "p99": round(float(np.percentile(lat_ms, 99)), 4)},
        "kernels": [{k: (round(v, 3) if isinstance(v, float) else v) for k, v in kk.items()} for kk in kernels],
        "setup_s": round(setup_s, 3),
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
