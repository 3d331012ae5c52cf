#!/usr/bin/env python3
"""Throughput of the Skwarka ray/photon hot path on one MI355X.

A step is one pass of the hot path over one batch of synthetic input, resident in HBM before the
timed region starts:
  c0 (default line): closest hit of 512x512 FP64 primary rays (pinhole at (0,0,-2.9), 60 deg x-FOV,
      pixel centres) against an 81,920-triangle icosphere (r=1, +-2% radial noise) inside the
      12-triangle box [-3,3]^3 -> Mrays/s.
  c1 (reported beside it): photon gather, 1,000,000 photons on the c0 scene, 262,143 gather points
      at the c0 primary hits, k=100, r=0.05, PhotonFilter, disc estimate + Lambertian rho=0.5
      radiance -> M gathers/s.
Timing: W untimed warm-up steps, then exactly K steps between HIP events recorded on the stream
the kernels run on, with device synchronisation on both sides.
The CPU baseline is the oracle (CPU restatement of the reference path, OpenMP) timed on this host
on the same workload (1 warm-up, median of 5), on rank 0 at N=1 only.
"""
import argparse
import json
import os
import platform
import statistics
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
RAY_BYTES, HIT_BYTES, NODE_BYTES, TRI_BYTES = 64, 32, 32, 36
KD_NODE_BYTES, CODE_BYTES, PHOTON_BYTES = 16, 4, 16


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_info():
    model = platform.processor() or ""
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
    except Exception:
        pass
    return model


def pmc_traffic(name):
    """HBM bytes per launch from a committed rocprofv3 --pmc summary (profiles/<name>.json)."""
    path = os.path.join(ROOT, "profiles", name + ".json")
    if not os.path.exists(path):
        return None
    try:
        with open(path) as f:
            return json.load(f).get("hbm_bytes_per_launch")
    except Exception:
        return None


def time_steps(torch, fn, steps, warmup, stream):
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    start = torch.cuda.Event(enable_timing=True)
    end = torch.cuda.Event(enable_timing=True)
    start.record(stream)
    for _ in range(steps):
        fn()
    end.record(stream)
    torch.cuda.synchronize()
    return start.elapsed_time(end) / steps


def bench_c0(torch, sk, wl, args, stream):
    meshes = wl.scene_c0()
    rays = wl.rays_c0()
    n = len(rays)
    t0 = time.time()
    scene = sk.Scene(meshes)
    build_s = time.time() - t0
    d_rays = torch.from_numpy(rays.view(np.uint8).copy()).to("cuda")
    d_hits = torch.empty(n * 32, dtype=torch.uint8, device="cuda")
    d_pops = torch.empty(n * 4, dtype=torch.uint8, device="cuda")
    d_tests = torch.empty(n * 4, dtype=torch.uint8, device="cuda")
    sptr = stream.cuda_stream
    # Untimed statistics pass: node boxes and triangles tested per ray (same walk, same order).
    scene.closest_hit_stats(d_rays.data_ptr(), n, d_hits.data_ptr(), d_pops.data_ptr(), d_tests.data_ptr(), sptr)
    torch.cuda.synchronize()
    pops = d_pops.cpu().numpy().view(np.uint32).astype(np.int64)
    tests = d_tests.cpu().numpy().view(np.uint32).astype(np.int64)
    ms = time_steps(torch, lambda: scene.closest_hit(d_rays.data_ptr(), n, d_hits.data_ptr(), sptr),
                    args.steps, args.warmup, stream)
    alg_bytes = n * (RAY_BYTES + HIT_BYTES) + NODE_BYTES * int(pops.sum()) + TRI_BYTES * int(tests.sum())
    achieved = alg_bytes / (ms * 1e-3) / 1e9
    hits = d_hits.cpu().numpy().view(sk.HIT_DTYPE)
    return {
        "n": n, "ms": ms, "value": n / (ms * 1e-3) / 1e6, "build_s": build_s,
        "mean_pops": float(pops.mean()), "mean_tests": float(tests.mean()), "alg_bytes": alg_bytes,
        "achieved_gbs": achieved, "hits": int((hits["prim"] != sk.MISS).sum()),
        "info": scene.info(), "scene": meshes, "rays": rays, "hit_records": hits,
    }


def bench_c1(torch, sk, wl, args, stream, c0):
    mesh = wl.merge(c0["scene"])
    ph = wl.photons_on_mesh(mesh, 1_000_000, seed=1)
    t0 = time.time()
    pm = sk.PhotonMap(ph["pos"], ph["weight"], ph["dir"], ph["normal"])
    build_s = time.time() - t0
    # gather points: c0 primary hits and their geometric normals, from the product's hit_geometry
    rays, hits = c0["rays"], c0["hit_records"]
    keep = hits["prim"] != sk.MISS
    scene = sk.Scene(c0["scene"])
    n = len(rays)
    d_rays = torch.from_numpy(rays.view(np.uint8).copy()).to("cuda")
    d_hits = torch.from_numpy(hits.view(np.uint8).copy()).to("cuda")
    d_geom = torch.empty(n * 104, dtype=torch.uint8, device="cuda")
    scene.hit_geometry(d_rays.data_ptr(), d_hits.data_ptr(), n, d_geom.data_ptr(), stream.cuda_stream)
    torch.cuda.synchronize()
    geom = d_geom.cpu().numpy().view(sk.GEOM_DTYPE)[keep]
    p = np.ascontiguousarray(geom["p"])
    nrm = np.ascontiguousarray(geom["ng"])
    nq, k, r = len(p), 100, 0.05
    dp = torch.from_numpy(p.view(np.uint8).reshape(-1).copy()).to("cuda")
    dn = torch.from_numpy(nrm.view(np.uint8).reshape(-1).copy()).to("cuda")
    ee = torch.empty(nq * 12, dtype=torch.uint8, device="cuda")
    ei = torch.empty(nq * 12, dtype=torch.uint8, device="cuda")
    rad = torch.empty(nq * 12, dtype=torch.uint8, device="cuda")
    cnt = torch.empty(nq * 4, dtype=torch.uint8, device="cuda")
    sptr = stream.cuda_stream

    def step():
        pm.irradiance(dp.data_ptr(), dn.data_ptr(), nq, k, r, 1, 0.5, ee.data_ptr(), ei.data_ptr(), rad.data_ptr(),
                      None, None, cnt.data_ptr(), sptr)

    ms = time_steps(torch, step, max(3, args.steps // 2), 2, stream)
    counts = cnt.cpu().numpy().view(np.uint32)
    return {"nq": nq, "ms": ms, "value": nq / (ms * 1e-3) / 1e6, "build_s": build_s,
            "mean_found": float(counts.mean()), "photons": ph, "p": p, "n": nrm}


def cpu_baseline_c0(c0, threads):
    import oracle  # CPU restatement of the reference path: the baseline, never the product
    acc = oracle.Accel(c0["scene"])
    rays = c0["rays"]
    acc.closest_hit(rays, nthreads=threads)  # warm-up
    times = []
    for _ in range(5):
        t = time.perf_counter()
        acc.closest_hit(rays, nthreads=threads)
        times.append(time.perf_counter() - t)
    med = statistics.median(times)
    return {"value": len(rays) / med / 1e6, "unit": "Mrays/s", "cores": threads, "kind": "port",
            "sample": "all %d c0 primary rays per run; 1 warm-up, median of 5 runs (%s)" % (len(rays), cpu_info())}


def cpu_baseline_c1(c1, threads, sample):
    import oracle
    ph = c1["photons"]
    om = oracle.PhotonMap(ph["pos"], ph["weight"], ph["dir"], ph["normal"])
    sel = slice(0, min(sample, len(c1["p"])))
    p, n = c1["p"][sel], c1["n"][sel]
    om.irradiance(p, n, 100, 0.05, with_sets=False, nthreads=threads)
    times = []
    for _ in range(3):
        t = time.perf_counter()
        om.irradiance(p, n, 100, 0.05, with_sets=False, nthreads=threads)
        times.append(time.perf_counter() - t)
    med = statistics.median(times)
    return {"value": len(p) / med / 1e6, "unit": "M gathers/s", "cores": threads, "kind": "port",
            "sample": "first %d of the c1 gather points; 1 warm-up, median of 3" % len(p)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-c1", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (the scene is replicated per "
              "GPU and rays shard by image tile, see DESIGN.md); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    import torch
    if not torch.cuda.is_available():
        print("no HIP device visible", file=sys.stderr)
        sys.exit(1)
    import skwarka_amd as sk
    from skwarka_amd import workloads as wl
    sk.lib()
    stream = torch.cuda.current_stream()
    threads = min(int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)), os.cpu_count() or 1)

    log("c0: closest hit, 512x512 rays, 81,932 triangles")
    c0 = bench_c0(torch, sk, wl, args, stream)
    log("c0: %.3f ms/step, %.1f Mrays/s" % (c0["ms"], c0["value"]))
    c1 = None
    if not args.no_c1:
        log("c1: photon gather, 1M photons, k=100, r=0.05")
        c1 = bench_c1(torch, sk, wl, args, stream, c0)
        log("c1: %.3f ms/step, %.2f M gathers/s" % (c1["ms"], c1["value"]))
    cpu = None
    cpu1 = None
    if not args.no_cpu:
        log("cpu baseline (oracle, %d threads)" % threads)
        cpu = cpu_baseline_c0(c0, threads)
        if c1 is not None:
            cpu1 = cpu_baseline_c1(c1, threads, 16384)
    traffic = pmc_traffic("closest_hit_c0_pmc")
    line = {
        "metric": "Mrays/s closest-hit",
        "value": round(c0["value"], 3),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(c0["ms"], 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "c0: 81,920-tri icosphere (r=1, 2% radial noise) + 12-tri box [-3,3]^3, "
                               "512x512 FP64 primary rays, 1 spp, closest hit",
                   "rays_per_step": c0["n"], "triangles": c0["info"]["triangles"], "nodes": c0["info"]["nodes"]},
        "roofline": {"bound": "hbm", "achieved": round(c0["achieved_gbs"], 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(c0["achieved_gbs"] / HBM_PEAK_GBS, 5), "traffic": traffic,
                     "bytes_per_ray": round(c0["alg_bytes"] / c0["n"], 2), "mean_node_tests": round(c0["mean_pops"], 3),
                     "mean_tri_tests": round(c0["mean_tests"], 3)},
        "cpu_baseline": cpu,
        "components": {},
    }
    if c1 is not None:
        line["components"]["c1_photon_gather"] = {
            "metric": "M photon-gathers/s", "value": round(c1["value"], 4), "ms_per_step": round(c1["ms"], 4),
            "gathers_per_step": c1["nq"], "mean_photons_found": round(c1["mean_found"], 3),
            "cpu_baseline": cpu1,
        }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
