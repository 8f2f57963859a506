"""Benchmark: forward + back projection pairs per second at 128x128-crystal panels on one MI355X.

Workload (the full-size plain configuration): two planograms (panels at psi = 0 and 90 degrees),
128 x 128 crystals at 2 mm, 200 mm panel separation -> planograms of shape (101, 101, 128, 228) each;
a 128^3 volume at 2 mm holding a seeded 20-ellipsoid phantom.  One step = one forward projection of the
volume into both planograms followed by one back projection (reference PETL semantics) of both
planograms, on buffers already resident in HBM.

Prints one JSON line.  Besides the headline rate it carries:
  roofline      HBM roofline of the dominant kernel (the fast forward kernel), from HIP events around its
                launches on the stream they run on; algorithmic bytes per launch = 4 * (voxels + bins)
  cpu_baseline  the CPU oracle (C++/OpenMP restatement of the reference projectors) timed on a bounded
                sample of the same workload on this host's cores
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-rows", type=int, default=4, help="v1 rows of each planogram in the CPU sample")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--mode", default="petl", choices=["petl", "matched"])
    ap.add_argument("--calib-copy", action="store_true",
                    help="also run one plain 256 MiB device copy (reference for HBM byte counters)")
    return ap.parse_args()


def cpu_baseline(rows, f):
    """Times the oracle's fp32 forward + PETL back on `rows` v1 rows of both planograms and scales to
    the full pair.  Returns (pairs/s, threads, sample description)."""
    from oracle import oracle as O
    from petl_amd.workloads import GEOMETRIES
    geo = GEOMETRIES["panel128"]
    plans = [O.planogram(psi, geo["D"], geo["L"], geo["H"], geo["T"]) for psi in geo["psis"]]
    vol = O.volume(128, 128, 128, 2.0)
    nt = O.nthreads()
    f = np.ascontiguousarray(f, dtype=np.float32)
    nv1 = plans[0].N_v1
    b0 = nv1 // 2 - rows // 2
    t0 = time.perf_counter()
    for p in plans:
        g = np.zeros(p.shape, np.float32)
        O.project_rows(p, vol, f, g, b0, b0 + rows, nt)
        fb = np.zeros(vol.shape, np.float32)
        O.backproject_petl_rows(p, vol, g, fb, b0, b0 + rows, nt)
    dt = time.perf_counter() - t0
    per_pair = dt * nv1 / rows
    desc = (f"fp32 oracle forward + PETL-mode back over v1 rows {b0}..{b0 + rows - 1} of both planograms "
            f"({rows}/{nv1} of the views, central rows), {dt:.1f} s measured, scaled by {nv1}/{rows}")
    return 1.0 / per_pair, nt, desc


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU per box; multi-GPU scaling is not measured",
                          "n_gpus": a.gpus}))
        sys.exit(2)

    import torch
    from petl_amd.petl_server import PETLserver
    from petl_amd.workloads import GEOMETRIES, ellipsoid_phantom, setup_server

    dev = torch.device("cuda:0")
    s = PETLserver()
    s.print_warnings = False
    geo = setup_server(s, "panel128")
    s.set_backprojector_mode(a.mode)
    fast = all(s.uses_fast_path(i) for i in range(s.get_numPlanograms()))
    shapes = [tuple(int(x) for x in s.get_planogramSize(i)) for i in range(s.get_numPlanograms())]
    n = geo["n"]
    f_host = ellipsoid_phantom(n, geo["voxel"], 20, seed=2)
    f = torch.from_numpy(f_host).to(dev)
    fb = torch.empty_like(f)
    g = [torch.empty(sh, dtype=torch.float32, device=dev) for sh in shapes]
    stream = torch.cuda.current_stream(dev).cuda_stream

    def step(ev=None):
        if ev is not None:
            ev[0].record()
        s.project_gpu(g, f, stream=stream, sync=False)
        if ev is not None:
            ev[1].record()
        s.backproject_gpu(g, fb, stream=stream, sync=False)
        if ev is not None:
            ev[2].record()

    for _ in range(a.warmup):
        step()
    torch.cuda.synchronize()
    if a.calib_copy:
        src = torch.ones(64 << 20, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    events = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(a.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(a.steps):
        step(events[k])
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    fwd_ms = float(np.mean([e[0].elapsed_time(e[1]) for e in events]))
    back_ms = float(np.mean([e[1].elapsed_time(e[2]) for e in events]))

    V = n ** 3
    G = [int(np.prod(sh)) for sh in shapes]
    nplan = len(shapes)
    fwd_launch_ms = fwd_ms / nplan                # the forward call is one fwd kernel launch per planogram
    alg_bytes = 4.0 * (V + G[0])                  # per launch: read the volume, write one planogram
    achieved = alg_bytes / (fwd_launch_ms * 1e-3) / 1e9
    pairs_per_s = a.steps / elapsed

    out = {
        "metric": "fwd+back projections/s at 128² panels; served recons/s + p99 latency",
        "value": pairs_per_s,
        "unit": "fwd+back projection pairs/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": 1e3 * elapsed / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded 20-ellipsoid phantom)",
        "config": {
            "workload": "2 planograms (psi 0 and 90 deg), 128x128 crystals @ 2 mm, D = 200 mm -> "
                        "(101,101,128,228) each; 128^3 voxels @ 2 mm; one step = forward of both planograms "
                        f"+ {a.mode}-mode back projection of both",
            "planogram_shape": list(shapes[0]),
            "planograms": nplan,
            "volume": [n, n, n],
            "back_mode": a.mode,
            "fast_kernels": fast,
        },
        "kernels_ms": {"forward_call": fwd_ms, "back_call": back_ms, "forward_launch": fwd_launch_ms},
        "roofline": {
            "kernel": "fwd_fast_kernel",
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None,
            "alg_bytes_per_launch": alg_bytes,
        },
    }
    if not a.no_cpu:
        v, cores, desc = cpu_baseline(a.cpu_rows, f_host)
        out["cpu_baseline"] = {"value": v, "unit": "fwd+back projection pairs/s", "cores": cores, "kind": "port",
                               "sample": desc}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
