"""Benchmark: forward + back projection pairs per second at 128x128-crystal panels on one MI355X.

Workload (the full-size plain configuration): two planograms (panels at psi = 0 and 90 degrees),
128 x 128 crystals at 2 mm, 200 mm panel separation -> planograms of shape (101, 101, 128, 228) each;
a 128^3 volume at 2 mm holding a seeded 20-ellipsoid phantom.  One step = one forward projection of the
volume into both planograms followed by one back projection (reference PETL semantics) of both
planograms, on buffers already resident in HBM.

Prints one JSON line.  Besides the headline rate it carries:
  roofline      HBM roofline of the dominant kernel (the fast forward kernel), from HIP events around its
                launches on the stream they run on; algorithmic bytes per launch = 4 * (voxels + bins);
                traffic from the committed rocprofv3 counter table
  valu_roofline the same kernel against FP32 vector arithmetic, the bound that actually limits it: a
                dense-equivalent flop rate, and the executed VALU instruction issue rate from the
                committed SQ counters
  lds_roofline  the same kernel's LDS read bandwidth (instruction count from the committed SQ counters)
  storage_forms the same pair with bfloat16 / binary16 volume and planograms (fp32 arithmetic)
  geometries    the same pair at the PET/X preset (fast kernels, and the general kernels forced) and at the
                six-planogram rotating-panel demo geometry (four oblique planograms on the general kernels),
                with the matched back projector's call time beside each
  binning       list-mode binning of 16.8M device-resident coincidences into both planograms
  batched_autograd  8 volumes of 64^3 through the autograd projector, forward + matched backward
  response      scintillator stopping power of every bin through 256 box modules (ray tracing)
  reconstruction  MLEM and RWLS per-iteration times and FBP, host data in and image out
  small_config  10 MLEM iterations at 32x32-crystal panels and a 32^3 volume
  serving       reconstructions served per second and p50/p99 latency for a Poisson arrival stream of
                MLEM requests of mixed panel/volume sizes, through the continuous-batching service (each
                request its own captured iteration, or all of them in batched iterations) and,
                for comparison, one request at a time
  cpu_baseline  the CPU oracle (C++/OpenMP restatement of the reference projectors) timed on a bounded
                sample of the same workload on this host's cores, and on the small configuration
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
    ap.add_argument("--cpu-rows", type=int, default=20, help="v1 rows of each planogram in the CPU sample")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--mode", default="petl", choices=["petl", "matched"])
    ap.add_argument("--calib-copy", action="store_true",
                    help="also run one plain 256 MiB device copy (reference for HBM byte counters)")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving measurement")
    ap.add_argument("--pair-only", action="store_true",
                    help="measure only the projection pair and its 16-bit storage forms (for profiling runs)")
    ap.add_argument("--sections", default="all",
                    help="comma list of the measurements after the projection pair: geometries, binning, autograd, "
                         "response, recon, small, serving, cpu (default all)")
    ap.add_argument("--serve-rates", default="5,20,50",
                    help="request arrival rates (req/s), comma separated: below, near and above the service's "
                         "capacity")
    ap.add_argument("--serve-requests", type=int, default=48)
    ap.add_argument("--serve-slots", type=int, default=8)
    return ap.parse_args()


def cpu_baseline(rows, f):
    """Times the oracle's fp32 forward + PETL back on `rows` v1 rows of both planograms, spread evenly over
    the v1 range (oblique rows mask more bins, so a central block would overstate the work), and scales to
    the full pair.  Returns (pairs/s, threads, sample description)."""
    from oracle import oracle as O
    from petl_amd.workloads import GEOMETRIES
    geo = GEOMETRIES["panel128"]
    plans = [O.planogram(psi, geo["D"], geo["L"], geo["H"], geo["T"]) for psi in geo["psis"]]
    vol = O.volume(128, 128, 128, 2.0)
    nt = O.nthreads()
    f = np.ascontiguousarray(f, dtype=np.float32)
    nv1 = plans[0].N_v1
    sample = sorted(set(int(round(x)) for x in np.linspace(0, nv1 - 1, rows)))
    t0 = time.perf_counter()
    for p in plans:
        g = np.zeros(p.shape, np.float32)
        fb = np.zeros(vol.shape, np.float32)
        for r in sample:
            O.project_rows(p, vol, f, g, r, r + 1, nt)
            O.backproject_petl_rows(p, vol, g, fb, r, r + 1, nt)
    dt = time.perf_counter() - t0
    per_pair = dt * nv1 / len(sample)
    desc = (f"fp32 oracle forward + PETL-mode back over {len(sample)} of the {nv1} v1 rows of both planograms "
            f"(rows {sample[0]}..{sample[-1]}, evenly spread), {dt:.1f} s measured, scaled by {nv1}/{len(sample)}")
    return 1.0 / per_pair, nt, desc


def cpu_small_mlem(y):
    """The fp32 oracle's 10 MLEM iterations (beta 0) at the small configuration on the host cores, median
    of 3 runs, for the GPU's small_config figure."""
    from oracle import oracle as O
    from petl_amd.workloads import GEOMETRIES
    geo = GEOMETRIES["panel32"]
    plans = [O.planogram(psi, geo["D"], geo["L"], geo["H"], geo["T"]) for psi in geo["psis"]]
    vol = O.volume(geo["n"], geo["n"], geo["n"], geo["voxel"])
    ts = []
    for _ in range(3):
        t0 = time.perf_counter()
        O.mlem(plans, vol, y, 10)
        ts.append(time.perf_counter() - t0)
    return {"mlem10_s_median_of_3": float(np.median(ts)), "cores": O.nthreads(), "kind": "port"}


SERVE_SIZES = ("panel48", "panel64", "panel96", "panel128")


def serving_datasets(torch, dev, seed):
    """Two Poisson datasets per panel size (counts drawn log-uniform in [1e5, 1e8]), in page-locked host
    memory so uploads run at full PCIe rate.  Returns {size: [(counts, [planogram arrays])]}."""
    from petl_amd.petl_server import PETLserver
    from petl_amd.workloads import GEOMETRIES, ellipsoid_phantom, setup_server
    rng = np.random.default_rng(seed)
    gen = torch.Generator(device=dev).manual_seed(seed)
    out = {}
    for k, name in enumerate(SERVE_SIZES):
        geo = GEOMETRIES[name]
        s = PETLserver()
        s.print_warnings = False
        setup_server(s, name)
        shapes = [tuple(int(x) for x in s.get_planogramSize(i)) for i in range(s.get_numPlanograms())]
        f = torch.from_numpy(ellipsoid_phantom(geo["n"], geo["voxel"], 20, seed=seed + k)).to(dev)
        q = [torch.empty(sh, dtype=torch.float32, device=dev) for sh in shapes]
        s.project_gpu(q, f)
        tot = float(sum(float(x.double().sum()) for x in q))
        sets = []
        for _ in range(2):
            counts = float(10.0 ** rng.uniform(5.0, 8.0))
            ys = []
            for x in q:
                host = torch.empty(x.shape, dtype=torch.float32, pin_memory=True)
                host.copy_(torch.poisson(x * (counts / tot), generator=gen))
                ys.append(host)
            sets.append((counts, ys))
        out[name] = sets
        del q, f
    torch.cuda.synchronize()
    return out


def serving_run(datasets, lam, nreq, slots, seed, storage="f32", batched=False):
    """Poisson arrivals at rate lam; each request = (panel size, dataset, iterations U{5..20}).  Returns
    recons/s over the run and latency percentiles (submission to completion)."""
    import threading
    from petl_amd.serving import ReconService
    from petl_amd.workloads import GEOMETRIES
    rng = np.random.default_rng(seed)
    reqs = []
    t = 0.0
    for _ in range(nreq):
        t += float(rng.exponential(1.0 / lam))
        name = SERVE_SIZES[int(rng.integers(len(SERVE_SIZES)))]
        which = int(rng.integers(2))
        reqs.append((t, name, which, int(rng.integers(5, 21))))
    lat = [None] * nreq
    with ReconService(max_slots=slots, batched=batched) as svc:
        ids = [None] * nreq
        ready = [threading.Event() for _ in range(nreq)]

        def waiter():
            for i in range(nreq):
                ready[i].wait()
                _, times = svc.wait(ids[i])
                lat[i] = times[2]

        th = threading.Thread(target=waiter)
        t0 = time.perf_counter()
        th.start()
        for i, (ta, name, which, iters) in enumerate(reqs):
            dt = t0 + ta - time.perf_counter()
            if dt > 0:
                time.sleep(dt)
            geo = GEOMETRIES[name]
            ys = [y.numpy() for y in datasets[name][which][1]]
            ids[i] = svc.submit(psis=geo["psis"], D=geo["D"], L=geo["L"], H=geo["H"], T=geo["T"], n=geo["n"],
                                voxel=geo["voxel"], y=ys, num_iter=iters, storage=storage,
                                image_storage="f32")
            ready[i].set()
        th.join()
        elapsed = time.perf_counter() - t0
        st = svc.stats()
    lat = np.array(lat)
    return {"recons_per_s": nreq / elapsed, "p50_s": float(np.percentile(lat, 50)),
            "p99_s": float(np.percentile(lat, 99)), "mean_s": float(lat.mean()), "peak_on_device": st["peak_active"],
            "failed": st["failed"],
            "host_s": {k: round(st[k], 4) for k in ("admit_alloc_s", "admit_setup_s", "admit_capture_s", "release_s")},
            "pool_hits": st["pool_hits"],
            **({"batched_rounds": st["batched_rounds"], "batched_peak_requests": st["batched_peak"]} if batched else {})}


FWD_KERNEL = "void fwd_fast_kernel<7, 4, 8, 1, 0, 0, 0, 128>"  # the fp32 forward instance of the benchmark


def kernel_source_sha():
    """SHA-256 of the fast kernels' source, which the committed counter tables record: a table made from
    other kernel code is reported as not current."""
    import hashlib
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "petl_amd", "csrc", "proj_fast.hip")
    with open(path, "rb") as fh:
        return hashlib.sha256(fh.read()).hexdigest()


def profile_table(name):
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", name)
    try:
        with open(path) as fh:
            return json.load(fh)
    except (OSError, ValueError):
        return None


def table_current(name):
    t = profile_table(name)
    return t is not None and t.get("source_sha256") == kernel_source_sha()


def measured_traffic(kernel):
    """HBM bytes per dispatch of `kernel` from the committed counter summary (copy-calibrated FETCH_SIZE +
    WRITE_SIZE), or None when it has no entry for the kernel."""
    t = profile_table("traffic_panel128.json")
    entry = None if t is None else t.get("kernels", {}).get(kernel)
    return None if entry is None else float(entry["traffic_bytes"])


def lds_roofline(launch_ms):
    insts = measured_counter(FWD_KERNEL, "SQ_INSTS_LDS")
    if insts is None:
        return None
    achieved = insts * 256.0 / (launch_ms * 1e-3) / 1e9
    return {"kernel": "fwd_fast_kernel", "achieved": achieved, "peak": LDS_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / LDS_PEAK_GBS, "lds_instructions_per_launch": insts,
            "source": "profiles/counters_panel128.json (SQ_INSTS_LDS, rocprofv3 --pmc, --pair-only)",
            "table_current": table_current("counters_panel128.json")}


def measured_counter(kernel, counter):
    """Mean per dispatch of an SQ counter of `kernel` from the committed counter summary, or None."""
    t = profile_table("counters_panel128.json")
    entry = None if t is None else t.get("kernels", {}).get(kernel)
    return None if entry is None or counter not in entry else float(entry[counter])


# LDS bandwidth of MI355X: 128 bytes per clock per CU, 256 CUs, 2.4 GHz
LDS_PEAK_GBS = 128 * 256 * 2.4
# VALU issue of MI355X: 256 CUs x 4 SIMDs of 16 lanes at 2.4 GHz, a wave64 instruction every 4 clocks per SIMD
VALU_ISSUE_PEAK = 256 * 4 * 2.4e9 / 4


def valu_issue(launch_ms):
    """Executed VALU wave instructions per second of the fast forward kernel (count per launch from the
    committed SQ_INSTS_VALU counter) against the issue peak."""
    insts = measured_counter(FWD_KERNEL, "SQ_INSTS_VALU")
    if insts is None:
        return None
    rate = insts / (launch_ms * 1e-3)
    return {"achieved": rate / 1e9, "peak": VALU_ISSUE_PEAK / 1e9, "unit": "G wave-instructions/s",
            "frac": rate / VALU_ISSUE_PEAK, "valu_instructions_per_launch": insts,
            "source": "profiles/counters_panel128.json (SQ_INSTS_VALU, rocprofv3 --pmc, --pair-only)",
            "table_current": table_current("counters_panel128.json")}


def main():
    a = parse()
    chosen = set(x.strip() for x in a.sections.split(",") if x.strip())

    def want(name):
        if a.pair_only:
            return False
        if name == "serving" and a.no_serve:
            return False
        if name == "cpu" and a.no_cpu:
            return False
        return "all" in chosen or name in chosen
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
    # the forward's arithmetic: two FMAs (4 flop) per bin and slice (the u0 interpolation sums Q and W);
    # slices = n (the volume axis the view direction crosses)
    flop_per_launch = 4.0 * n * G[0]
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
            "traffic": measured_traffic(FWD_KERNEL),
            "traffic_table_current": table_current("traffic_panel128.json"),
            "traffic_source": "profiles/traffic_panel128.json (table in profiles/kernels_panel128.md): rocprofv3 "
                              "--pmc FETCH_SIZE and WRITE_SIZE of this kernel (one run each, --pair-only), scaled "
                              "by the same counters on a 256 MiB copy",
            "alg_bytes_per_launch": alg_bytes,
        },
        # LDS: the forward's main loop reads one dword per lane per instruction (ds_read_b32, 256 bytes per
        # wave instruction); instruction count per launch from the committed SQ_INSTS_LDS counter
        "lds_roofline": lds_roofline(fwd_launch_ms),
        # the bound that actually limits this kernel: FP32 vector arithmetic (published peak 157.3 TFLOP/s
        # counts packed FP32, which this build does not emit; unpacked FP32 peaks at 78.6)
        "valu_roofline": {
            "kernel": "fwd_fast_kernel",
            "achieved": flop_per_launch / (fwd_launch_ms * 1e-3) / 1e12,
            "peak": 157.3,
            "unit": "TFLOP/s",
            "frac": flop_per_launch / (fwd_launch_ms * 1e-3) / 1e12 / 157.3,
            "flop_per_launch": flop_per_launch,
            # flop_per_launch counts every (bin, slice) pair; the kernel skips slice chunks outside each
            # view's support, so "achieved" is a dense-equivalent rate.  "issue" is what it executes.
            "issue": valu_issue(fwd_launch_ms),
        },
    }
    # the same pair with the volume and planograms stored in bfloat16 / binary16 (fp32 arithmetic, fp32
    # back-projected volume)
    storage_forms = {}
    for name, dt in (("bf16", torch.bfloat16), ("f16", torch.float16)):
        f16 = f.to(dt)
        g16 = [torch.empty(sh, dtype=dt, device=dev) for sh in shapes]
        for _ in range(a.warmup):
            s.project_gpu(g16, f16, stream=stream, sync=False)
            s.backproject_gpu(g16, fb, stream=stream, sync=False)
        ev16 = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(a.steps)]
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for k in range(a.steps):
            ev16[k][0].record()
            s.project_gpu(g16, f16, stream=stream, sync=False)
            ev16[k][1].record()
            s.backproject_gpu(g16, fb, stream=stream, sync=False)
            ev16[k][2].record()
        torch.cuda.synchronize()
        el16 = time.perf_counter() - t1
        storage_forms[name] = {
            "pairs_per_s": a.steps / el16,
            "forward_call_ms": float(np.mean([e[0].elapsed_time(e[1]) for e in ev16])),
            "back_call_ms": float(np.mean([e[1].elapsed_time(e[2]) for e in ev16])),
        }
        del g16, f16
    out["storage_forms"] = storage_forms

    if want("geometries"):
        # the reference's other geometries: the PET/X preset (psi 0 / 90 panels at 1.6 mm, its default
        # 156 x 62 x 96 volume) and the six-planogram rotating-panel demo (psi = 0, 30, ..., 150; four oblique
        # planograms on the general kernels); fwd + PETL back pairs/s on device buffers, and the PET/X preset
        # with the general kernels forced
        import math
        t15 = math.tan(15.0 * math.pi / 180.0)
        geoms = {
            "petx": [(0.0, 101.0, 251.52, 156.72, 1.6, None, None), (90.0, 266.6, 125.12, 156.72, 1.6, None, None)],
            "rotating6": [(k * 30.0, 180.0, 192.0, 144.0, 2.0, t15, t15) for k in range(6)],
        }
        res = {}
        for gname, planos in geoms.items():
            for path in (("auto", "general") if gname == "petx" else ("auto",)):
                sg = PETLserver()
                sg.print_warnings = False
                for psi, D_, L_, H_, T_, m0, m1 in planos:
                    sg.add_planogram(psi, D_, L_, H_, T_, m0, m1)
                sg.set_default_volume()
                sg.set_projector_path("generic" if path == "general" else "auto")
                vs = (sg.numZ(), sg.numY(), sg.numX())
                gsh = [tuple(int(x) for x in sg.get_planogramSize(i)) for i in range(sg.get_numPlanograms())]
                fg = torch.from_numpy(ellipsoid_phantom(vs[2], sg.voxelWidth(), 8, seed=5, shape=vs)).to(dev)
                gg = [torch.empty(sh, dtype=torch.float32, device=dev) for sh in gsh]
                fo = torch.empty_like(fg)
                for _ in range(2):
                    sg.project_gpu(gg, fg, stream=stream, sync=False)
                    sg.backproject_gpu(gg, fo, stream=stream, sync=False)
                eg = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
                reps = 5
                torch.cuda.synchronize()
                eg[0].record()
                for _ in range(reps):
                    sg.project_gpu(gg, fg, stream=stream, sync=False)
                eg[1].record()
                for _ in range(reps):
                    sg.backproject_gpu(gg, fo, stream=stream, sync=False)
                eg[2].record()
                torch.cuda.synchronize()
                fms, bms = eg[0].elapsed_time(eg[1]) / reps, eg[1].elapsed_time(eg[2]) / reps
                # the matched back projector (the forward's exact transpose: gradients, RWLS) on the same data
                sg.set_backprojector_mode("matched")
                sg.backproject_gpu(gg, fo, stream=stream, sync=False)
                torch.cuda.synchronize()
                eg[0].record()
                for _ in range(reps):
                    sg.backproject_gpu(gg, fo, stream=stream, sync=False)
                eg[1].record()
                torch.cuda.synchronize()
                mms = eg[0].elapsed_time(eg[1]) / reps
                res[f"{gname}_{path}"] = {
                    "planograms": [list(x) for x in gsh], "volume": list(vs),
                    "fast_kernels": [sg.uses_fast_path(i) for i in range(len(gsh))],
                    "forward_call_ms": fms, "back_call_ms": bms, "pairs_per_s": 1e3 / (fms + bms),
                    "matched_back_call_ms": mms}
                del fg, gg, fo
        out["geometries"] = res

    if want("binning"):
        # list-mode binning: device-resident coincidences between the two panel pairs into both planograms
        # (bilinear u bins, chest-wall origin: the reference defaults); events/s of one binning call
        nev = 1 << 24
        gen = torch.Generator(device=dev).manual_seed(7)
        half = nev // 2
        R, L, H = geo["D"] / 2, geo["L"] / 2, geo["H"]
        def panel(n, fixed_axis, sign):
            p = torch.empty((n, 3), device=dev)
            p[:, 0].uniform_(-L, L, generator=gen)
            p[:, 1].uniform_(-L, L, generator=gen)
            p[:, 2].uniform_(0.0, H, generator=gen)
            p[:, fixed_axis] = sign * R
            return p
        blue = torch.cat([panel(half, 1, 1.0), panel(nev - half, 0, 1.0)]).contiguous()
        pink = torch.cat([panel(half, 1, -1.0), panel(nev - half, 0, -1.0)]).contiguous()
        gb = [torch.zeros(sh, dtype=torch.float32, device=dev) for sh in shapes]
        s.bin_events_gpu(blue, pink, gb, stream=stream, sync=True)  # warm-up (allocates the sort workspace)
        eb = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        reps = 3
        eb[0].record()
        for _ in range(reps):
            s.bin_events_gpu(blue, pink, gb, stream=stream, sync=False)
        eb[1].record()
        torch.cuda.synchronize()
        bin_ms = eb[0].elapsed_time(eb[1]) / reps
        out["binning"] = {"events": nev, "planograms": nplan, "ms_per_call": bin_ms,
                          "events_per_s": nev / (bin_ms * 1e-3),
                          "workload": "uniform coincidences between the two panel pairs of the benchmark geometry, "
                                      "device-resident; bilinear u bins, chest-wall origin; bit-exact event-order sums"}
        del blue, pink, gb

    if want("autograd"):
        # batched form: a batch of 8 seeded 64^3 volumes through the autograd projector (forward into both
        # 64x64-crystal planograms, then the matched back projection as the backward), volumes per second
        from petl_amd import autograd as pa
        sb = PETLserver()
        sb.print_warnings = False
        setup_server(sb, "panel64")
        fbatch = torch.from_numpy(np.stack([ellipsoid_phantom(64, 2.0, 6, seed=3 + b) for b in range(8)])).to(dev)
        def ag_step():
            x = fbatch.clone().requires_grad_(True)
            q = pa.project(sb, x)
            (gx,) = torch.autograd.grad(q, x, grad_outputs=[torch.ones_like(t) for t in q])
            return gx
        for _ in range(2):
            ag_step()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        nag = 5
        for _ in range(nag):
            ag_step()
        torch.cuda.synchronize()
        ag_s = (time.perf_counter() - t1) / nag
        out["batched_autograd"] = {"workload": "batch of 8 seeded 64^3 volumes, two 64x64-crystal planograms "
                                               "(63,101,64,164); forward + matched backward via torch.autograd",
                                   "ms_per_batch": 1e3 * ag_s, "volumes_per_s": 8 / ag_s}
        del fbatch

    if want("response"):
        # response model at the benchmark geometry: scintillator stopping power through 4 x 64 box modules
        # (30.32 mm blocks, 1.28 mm gaps, 20 mm deep, behind each of the four panels), device planograms
        modules = []
        blk, gap, depth = 30.32, 1.28, 20.0
        npan = int(geo["L"] // (blk + gap))
        for m_ in range(npan):
            cz = (m_ + 0.5) * (blk + gap)
            for n_ in range(npan):
                cx = (n_ - (npan - 1) / 2) * (blk + gap)
                for c_, r_ in (([cx, geo["D"] / 2 + depth / 2, cz], [blk / 2, depth / 2, blk / 2]),
                               ([cx, -geo["D"] / 2 - depth / 2, cz], [blk / 2, depth / 2, blk / 2]),
                               ([geo["D"] / 2 + depth / 2, cx, cz], [depth / 2, blk / 2, blk / 2]),
                               ([-geo["D"] / 2 - depth / 2, cx, cz], [depth / 2, blk / 2, blk / 2])):
                    modules.append((c_, r_))
        for c_, r_ in modules:
            s.add_module(c_, r_)
        resp = {}
        for os_ in (1, 3):
            s.ray_trace_gpu(g, oversampling=os_, stopping=True, stream=stream, sync=True)  # warm-up
            er = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
            er[0].record()
            s.ray_trace_gpu(g, oversampling=os_, stopping=True, stream=stream, sync=False)
            er[1].record()
            torch.cuda.synchronize()
            ms = er[0].elapsed_time(er[1])
            resp[f"oversampling_{os_}"] = {"ms_both_planograms": ms,
                                           "rays_per_s": os_ * os_ * sum(int(np.prod(sh)) for sh in shapes) / (ms * 1e-3)}
        s.clear_modules()
        out["response"] = dict(resp, workload=f"stopping power of every bin of both planograms through {len(modules)} "
                                              "box modules (two half rays each, oversampling^2 sub-rays per bin)")

    if want("recon"):
        # reconstruction loops at the benchmark geometry, host data in and image out as PETLserver.MLEM /
        # PETLserver.RWLS take them: per-iteration cost from the difference of a short and a long run (the
        # upload, weights, sensitivity and preconditioner set-up cancel)
        s.project_gpu(g, f, stream=stream, sync=True)
        tot = float(sum(float(x.double().sum()) for x in g))
        gen_y = torch.Generator(device=dev).manual_seed(11)
        y_host = [torch.poisson(x * (1e8 / tot), generator=gen_y).cpu().numpy() for x in g]
        def timed(fn, reps=2):
            # fastest of `reps` runs: the per-iteration figure is a difference of two run times, and one
            # slow upload in either run would land on it
            best = float("inf")
            for _ in range(reps):
                t = time.perf_counter()
                fn()
                best = min(best, time.perf_counter() - t)
            return best
        s.MLEM(y_host, None, 1)  # warm-up
        k1, k2, k3 = 2, 32, 17
        m1 = timed(lambda: s.MLEM(y_host, None, k1, delta=2.0, beta=0.01))
        m2 = timed(lambda: s.MLEM(y_host, None, k2, delta=2.0, beta=0.01))
        r1 = timed(lambda: s.RWLS(y_host, None, k1, delta=0.5, beta=0.02, preconditioner="SQS"))
        r2 = timed(lambda: s.RWLS(y_host, None, k3, delta=0.5, beta=0.02, preconditioner="SQS"))
        f0 = timed(lambda: s.FBP(y_host), 1)  # the first call also builds the FFT plans (rocFFT compiles at run time)
        f1 = timed(lambda: s.FBP(y_host), 1)
        out["reconstruction"] = {
            "workload": "benchmark geometry, Poisson data at 1e8 counts; MLEM with the relative-difference prior "
                        "(beta 0.01, delta 2); RWLS with default weights, SQS and TV (beta 0.02, delta 0.5)",
            "mlem_ms_per_iteration": 1e3 * (m2 - m1) / (k2 - k1),
            "mlem_kernel_floor_ms": fwd_ms + back_ms,
            "mlem_20_iterations_s": m1 + (m2 - m1) * (20 - k1) / (k2 - k1),
            "rwls_ms_per_iteration": 1e3 * (r2 - r1) / (k3 - k1),
            "rwls_kernel_floor_ms": 2 * fwd_ms + back_ms,
            "fbp_s": f1,
            "fbp_first_call_s": f0,
            "fbp_note": "PFDR of both planograms (hipFFT) + ramp filter + back projection, host arrays in and out; "
                        "the first call includes building the FFT plans",
        }
        del y_host

    small_y = None
    if want("small"):
        # the small plain configuration: one 32x32-crystal planogram (psi 0, D = 100 mm -> (31,51,32,82)),
        # 32^3 volume at 2 mm, seeded 6-ellipsoid phantom, Poisson data at 1e6 counts; 10 MLEM iterations
        # (beta 0) from host data to a host image, as PETLserver.MLEM takes them; median of 5 runs
        geo1 = GEOMETRIES["panel32"]
        s1 = PETLserver()
        s1.print_warnings = False
        setup_server(s1, "panel32")
        f1 = torch.from_numpy(ellipsoid_phantom(geo1["n"], geo1["voxel"], 6, seed=1)).to(dev)
        q1 = [torch.empty(tuple(int(x) for x in s1.get_planogramSize(0)), dtype=torch.float32, device=dev)]
        s1.project_gpu(q1, f1)
        gen1 = torch.Generator(device=dev).manual_seed(101)
        small_y = [torch.poisson(q1[0] * (1e6 / float(q1[0].double().sum())), generator=gen1).cpu().numpy()]
        s1.MLEM(small_y, None, 10)  # warm-up
        ts = []
        for _ in range(5):
            t1 = time.perf_counter()
            s1.MLEM(small_y, None, 10)
            ts.append(time.perf_counter() - t1)
        out["small_config"] = {
            "workload": "one planogram (31,51,32,82) (32x32 crystals @ 2 mm, D = 100 mm), 32^3 volume @ 2 mm, "
                        "Poisson data at 1e6 counts; 10 MLEM iterations (beta 0), host data in, host image out",
            "mlem10_s_median_of_5": float(np.median(ts)),
        }
        del s1, f1, q1

    if want("serving"):
        del g, f, fb
        torch.cuda.empty_cache()
        data = serving_datasets(torch, dev, seed=500)
        runs = []
        for rate in [float(x) for x in a.serve_rates.split(",") if x]:
            mixed = serving_run(data, rate, a.serve_requests, a.serve_slots, seed=501)
            mixed16 = serving_run(data, rate, a.serve_requests, a.serve_slots, seed=501, storage="bf16")
            batched = serving_run(data, rate, a.serve_requests, a.serve_slots, seed=501, batched=True)
            single = serving_run(data, rate, a.serve_requests, 1, seed=501)
            runs.append({"arrival_rate_req_s": rate, "continuous_batching": dict(mixed, slots=a.serve_slots),
                         "batched_iterations": dict(batched, slots=a.serve_slots),
                         "continuous_batching_bf16_planograms": dict(mixed16, slots=a.serve_slots),
                         "one_at_a_time": dict(single, slots=1)})
        out["serving"] = {
            "workload": "Poisson arrivals; panels uniform over 48/64/96/128 crystals (2 orientations, D = 200 mm, "
                        "matched n^3 volumes at 2 mm); data = Poisson at counts log-uniform in [1e5, 1e8]; "
                        "MLEM iterations uniform in 5..20; fp32 (and bf16 planograms with an fp32 image); "
                        "continuous_batching: each request replays its own captured iteration on its own stream; "
                        "batched_iterations: one launch per phase advances all requests on the device",
            "requests": a.serve_requests,
            "runs": runs,
        }
    if want("cpu"):
        v, cores, desc = cpu_baseline(a.cpu_rows, f_host)
        out["cpu_baseline"] = {"value": v, "unit": "fwd+back projection pairs/s", "cores": cores, "kind": "port",
                               "sample": desc}
        if small_y is not None:
            out["cpu_baseline"]["small_config"] = cpu_small_mlem(small_y)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
