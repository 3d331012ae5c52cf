#!/usr/bin/env python3
"""Throughput of the Skwarka ray/photon hot path on one MI355X.

A step is one pass of the hot path over one batch of synthetic input, resident in HBM before the
timed region starts:
  c2 (the line's value): one direct-lighting frame of the seeded 2,026,260-triangle room (200
      displaced icospheres, 4 quad area lights, matte and plastic materials) at 1024x1024, 16 spp,
      4 light + 4 BSDF samples per hit: camera rays, closest hit, shading, occlusion rays, BSDF
      rays, MIS resolve -> Mrays/s over every ray cast (primary + occlusion + BSDF).
  c0 (component): closest hit of 512x512 FP64 primary rays (pinhole at (0,0,-2.9), 60 deg x-FOV,
      pixel centres) against an 81,920-triangle icosphere (r=1, +-2% radial noise) inside the
      12-triangle box [-3,3]^3 -> Mrays/s.
  c1 (component): photon gather, 1,000,000 photons on the c0 scene, 262,143 gather points at the
      c0 primary hits, k=100, r=0.05, PhotonFilter, disc estimate + Lambertian rho=0.5 radiance
      -> M gathers/s.
  c5 (component): the photon integrator's final gather on the c2 scene at 1024x1024, 4 spp, 8
      gather samples (rounded to 9), over an irradiance photon map built from 100k direct, 400k
      indirect and 20k caustic photons -> Mrays/s over camera + gather rays.
  c3 (component): 64 batched scenes (1.3k-874k triangles) closest hit in one persistent launch
      per frame, alone and with 8 photon-gather jobs -> Mrays/s, per-item latency.
  c4 (component): the 132-parameter direct-lighting gradient at 512x512, 8 spp -> gradients/s.
  c6 (component): photon shooting from the c2 scene's area lights, 4M paths -> M paths/s, and the
      device kd-tree build of the indirect map against the host build.
  c7 (component): the photon-mapping render of the c2 scene (direct + caustic + final gather) at
      1024x1024, 16 spp, over shot maps -> Mrays/s.
  c8 (component): the triangle-tree build of the c2 room on the device against the host build
      -> M triangles/s.
Timing: W untimed warm-up steps, then exactly K steps between HIP events recorded on the stream
the kernels run on, with device synchronisation on both sides.
The CPU baseline is the oracle (CPU restatement of the reference path, OpenMP) timed on this host
on a bounded sample of the same workload, on rank 0 at N=1 only.
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
# Per-ray algorithmic bytes: ray in, hit out, 40 B per node box tested (the reference's node: float
# bbox + child/range words) and 36 B per triangle tested (three float vertices).
RAY_BYTES, HIT_BYTES, NODE_BYTES, TRI_BYTES = 64, 32, 40, 36
# Per-gather algorithmic bytes: query point and normal in (48), E_ext / E_int / radiance out (36),
# 20 B per node visited (8 B node + 12 B point), 2 B per photon reaching the PhotonFilter (its
# normal code), 18 B per photon in the final set (12 B weight + 2 B direction code + 4 B id).
GATHER_IN_BYTES, GATHER_OUT_BYTES, GATHER_VISIT_BYTES, GATHER_FILTER_BYTES, GATHER_FOUND_BYTES = 48, 36, 20, 2, 18


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


def median_runs(fn, runs=5):
    """CPU baseline protocol: one warm-up call, then `runs` timed calls; returns the median time
    and the last call's result."""
    fn()
    times, out = [], None
    for _ in range(runs):
        t = time.perf_counter()
        out = fn()
        times.append(time.perf_counter() - t)
    return statistics.median(times), out


def pmc_summary(name):
    """The committed rocprofv3 --pmc summary profiles/<name>.json (HBM bytes per launch, L2 hit
    rate, VALU figures), or {} when it is absent."""
    path = os.path.join(ROOT, "profiles", name + ".json")
    if not os.path.exists(path):
        return {}
    try:
        with open(path) as f:
            return json.load(f)
    except Exception:
        return {}


def pmc_traffic(name):
    """HBM bytes per launch from a committed rocprofv3 --pmc summary (profiles/<name>.json)."""
    return pmc_summary(name).get("hbm_bytes_per_launch")


def source_sha(files):
    """sha256 (first 16 hex digits) of the given kernel sources, as the PMC summaries record it."""
    import hashlib
    h = hashlib.sha256()
    for f in files:
        with open(os.path.join(ROOT, f), "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()[:16]


def pmc_fields(name):
    """The measured figures of a committed PMC summary, and whether the kernel sources it was taken
    from are the ones this run executes (a summary of older kernels is reported as stale)."""
    p = pmc_summary(name)
    if not p:
        return {"traffic": None, "pmc": None}
    files = p.get("source_files") or []
    current = source_sha(files) if files else None
    return {"traffic": p.get("hbm_bytes_per_launch"),
            "pmc": {"summary": "profiles/%s.json" % name, "l2_hit_rate": p.get("l2_hit_rate"),
                    "valu_issue_utilisation": p.get("valu_issue_utilisation"),
                    "valu_active_lanes_of_64": p.get("valu_active_lanes_of_64"),
                    "hbm_write_bytes_per_launch": p.get("hbm_write_bytes_per_launch"),
                    "profiled_source_sha": p.get("source_sha"), "current_source_sha": current,
                    "matches_current_source": bool(files) and p.get("source_sha") == current}}


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


def dev_bytes(torch, arr):
    return torch.from_numpy(np.ascontiguousarray(arr).view(np.uint8).reshape(-1).copy()).to("cuda")


def bench_c2(torch, sk, wl, steps, warmup, stream):
    s = wl.scene_room()
    t0 = time.time()
    scene = sk.Scene(s["meshes"])
    build_s = time.time() - t0
    dls = sk.DLScene(scene, s["mat"], s["kd"], s["ks"], s["rough"], s["light_prim"], s["light_radiance"])
    w = h = 1024
    spp = 16
    cam, cam15 = sk.camera(*s["camera"], w, h)
    d_rgb = torch.empty(w * h * 12, dtype=torch.uint8, device="cuda")
    sptr = stream.cuda_stream

    def step():
        # the plain frame as a caller issues it: enqueued on the stream, no host wait inside
        dls.render(cam, w, h, spp, d_rgb.data_ptr(), 4, 4, 1234, sptr, count_rays=False)

    ms = time_steps(torch, step, steps, warmup, stream)
    # the same frames again with HIP events at their stage boundaries (on the frame's stream)
    timed = [dls.render_profiled(cam, w, h, spp, d_rgb.data_ptr(), 4, 4, 1234, False, sptr)
             for _ in range(max(1, min(steps, 3)))]
    rays = np.array(timed[-1]["rays"], np.int64)
    total = int(rays.sum())
    stage_ms = {k: float(np.mean([p["stage_ms"][k] for p in timed])) for k in timed[-1]["stage_ms"]}
    # one untimed frame of the counting walk: node boxes and triangles tested per ray class (the
    # same rays as the timed frames: the sample stream is deterministic)
    cnt = dls.render_profiled(cam, w, h, spp, d_rgb.data_ptr(), 4, 4, 1234, True, sptr)
    assert cnt["rays"] == timed[-1]["rays"]
    nodes, tris = np.array(cnt["node_tests"], np.int64), np.array(cnt["tri_tests"], np.int64)
    out_bytes = np.array([HIT_BYTES, 1, HIT_BYTES], np.int64)  # occlusion rays write one byte
    alg_bytes = int((rays * (RAY_BYTES + out_bytes) + NODE_BYTES * nodes + TRI_BYTES * tris).sum())
    trace_ms = stage_ms["primary_trace"] + stage_ms["occlusion_trace"] + stage_ms["bsdf_trace"]
    rgb = d_rgb.cpu().numpy().view(np.float32)
    return {"ms": ms, "value": total / (ms * 1e-3) / 1e6, "rays": [int(x) for x in rays], "build_s": build_s,
            "info": scene.info(), "spec": s, "cam15": cam15, "mean_rgb": float(rgb.mean()), "scene": scene,
            "stage_ms": stage_ms, "trace_ms": trace_ms, "alg_bytes": alg_bytes, "launches": timed[-1]["launches"],
            "achieved_gbs": alg_bytes / (trace_ms * 1e-3) / 1e9,
            "mean_node_tests": (nodes / np.maximum(rays, 1)).tolist(),
            "mean_tri_tests": (tris / np.maximum(rays, 1)).tolist()}


def bench_c0(torch, sk, wl, steps, warmup, stream):
    meshes = wl.scene_c0()
    rays = wl.rays_c0()
    n = len(rays)
    t0 = time.time()
    scene = sk.Scene(meshes)
    build_s = time.time() - t0
    d_rays = dev_bytes(torch, rays)
    d_hits = torch.empty(n * 32, dtype=torch.uint8, device="cuda")
    d_pops = torch.empty(n * 4, dtype=torch.uint8, device="cuda")
    d_tests = torch.empty(n * 4, dtype=torch.uint8, device="cuda")
    sptr = stream.cuda_stream
    # Untimed statistics pass: node boxes and triangles tested per ray (same walk, same order).
    scene.closest_hit_stats(d_rays.data_ptr(), n, d_hits.data_ptr(), d_pops.data_ptr(), d_tests.data_ptr(), sptr)
    torch.cuda.synchronize()
    pops = d_pops.cpu().numpy().view(np.uint32).astype(np.int64)
    tests = d_tests.cpu().numpy().view(np.uint32).astype(np.int64)
    ms = time_steps(torch, lambda: scene.closest_hit(d_rays.data_ptr(), n, d_hits.data_ptr(), sptr), steps, warmup,
                    stream)
    alg_bytes = n * (RAY_BYTES + HIT_BYTES) + NODE_BYTES * int(pops.sum()) + TRI_BYTES * int(tests.sum())
    hits = d_hits.cpu().numpy().view(sk.HIT_DTYPE)
    return {"n": n, "ms": ms, "value": n / (ms * 1e-3) / 1e6, "build_s": build_s, "mean_pops": float(pops.mean()),
            "mean_tests": float(tests.mean()), "alg_bytes": alg_bytes,
            "achieved_gbs": alg_bytes / (ms * 1e-3) / 1e9, "info": scene.info(), "scene": meshes, "rays": rays,
            "hit_records": hits, "scene_handle": scene, "d_rays": d_rays, "d_hits": d_hits}


def bench_c1(torch, sk, wl, steps, warmup, stream, c0):
    mesh = wl.merge(c0["scene"])
    ph = wl.photons_on_mesh(mesh, 1_000_000, seed=1)
    t0 = time.time()
    pm = sk.PhotonMap(ph["pos"], ph["weight"], ph["dir"], ph["normal"])
    build_s = time.time() - t0
    # gather points: c0 primary hits and their geometric normals, from the product's hit_geometry
    n = len(c0["rays"])
    keep = c0["hit_records"]["prim"] != sk.MISS
    d_geom = torch.empty(n * 104, dtype=torch.uint8, device="cuda")
    c0["scene_handle"].hit_geometry(c0["d_rays"].data_ptr(), c0["d_hits"].data_ptr(), n, d_geom.data_ptr(),
                                    stream.cuda_stream)
    torch.cuda.synchronize()
    geom = d_geom.cpu().numpy().view(sk.GEOM_DTYPE)[keep]
    p = np.ascontiguousarray(geom["p"])
    nrm = np.ascontiguousarray(geom["ng"])
    nq, k, r = len(p), 100, 0.05
    dp, dn = dev_bytes(torch, p), dev_bytes(torch, nrm)
    ee = torch.empty(nq * 12, dtype=torch.uint8, device="cuda")
    ei = torch.empty(nq * 12, dtype=torch.uint8, device="cuda")
    rad = torch.empty(nq * 12, dtype=torch.uint8, device="cuda")
    cnt = torch.empty(nq * 4, dtype=torch.uint8, device="cuda")
    sptr = stream.cuda_stream

    def step():
        pm.irradiance(dp.data_ptr(), dn.data_ptr(), nq, k, r, 1, 0.5, ee.data_ptr(), ei.data_ptr(), rad.data_ptr(),
                      None, None, cnt.data_ptr(), sptr)

    ms = time_steps(torch, step, steps, warmup, stream)
    counts = cnt.cpu().numpy().view(np.uint32)
    # untimed counting pass of the same lookups: nodes visited, photons inside the radius (they
    # reach the PhotonFilter), set size per gather
    st = torch.empty(nq * 12, dtype=torch.uint8, device="cuda")
    pm.irradiance_stats(dp.data_ptr(), dn.data_ptr(), nq, k, r, 1, 0.5, ee.data_ptr(), ei.data_ptr(), rad.data_ptr(),
                        st.data_ptr(), sptr)
    torch.cuda.synchronize()
    c = st.cpu().numpy().view(np.uint32).reshape(nq, 3).astype(np.int64)
    n_visit, n_filter, n_found = int(c[:, 0].sum()), int(c[:, 1].sum()), int(c[:, 2].sum())
    assert n_found == int(counts.astype(np.int64).sum())
    alg_bytes = (GATHER_IN_BYTES + GATHER_OUT_BYTES) * nq + GATHER_VISIT_BYTES * n_visit + \
        GATHER_FILTER_BYTES * n_filter + GATHER_FOUND_BYTES * n_found
    return {"nq": nq, "ms": ms, "value": nq / (ms * 1e-3) / 1e6, "build_s": build_s,
            "mean_found": float(counts.mean()), "photons": ph, "p": p, "n": nrm, "alg_bytes": alg_bytes,
            "achieved_gbs": alg_bytes / (ms * 1e-3) / 1e9, "mean_visits": n_visit / nq,
            "mean_inside": n_filter / nq}


def bench_c3(torch, sk, wl, steps, warmup, stream):
    specs = wl.batch_scenes()
    # the 64 scenes' trees built on the device, one timed scene creation each (descriptions made
    # beforehand; one untimed build first)
    descs = [sk.scene_desc(sp[0]) for sp in specs]
    sk.Scene.from_desc(descs[0][0], device_build=True).close()
    t0 = time.time()
    scenes, views, build_ms = [], [], []
    for (desc, _), (meshes, camspec, size, spp, seed) in zip(descs, specs):
        torch.cuda.synchronize()
        t = time.perf_counter()
        scenes.append(sk.Scene.from_desc(desc, device_build=True))
        torch.cuda.synchronize()
        build_ms.append(1e3 * (time.perf_counter() - t))
        cam, _ = sk.camera(*camspec, size, size)
        views.append((cam, size, size, spp, seed))
    build_s = time.time() - t0
    del descs
    batch = sk.Batch(scenes, views)
    t0 = time.time()
    cfg, tune_ms = batch.autotune(frames=1)
    tune_s = time.time() - t0
    n = batch.ray_count()
    sptr = stream.cuda_stream
    ms = time_steps(torch, lambda: batch.frame(sptr), steps, warmup, stream)
    lat = batch.latency_ns() / 1e6  # completion of each scene in the last frame, ms after its start
    tris = [int(s.info()["triangles"]) for s in scenes]
    # The same 64 scenes with 8 photon-gather jobs in the same persistent, graph-replayed launch.
    jobs = wl.batch_gather_jobs(specs)
    gathers, keep = [], []
    for jb in jobs:
        pm = sk.PhotonMap(jb["ph"]["pos"], jb["ph"]["weight"], jb["ph"]["dir"], jb["ph"]["normal"])
        nq = len(jb["p"])
        dp, dn = dev_bytes(torch, jb["p"]), dev_bytes(torch, jb["n"])
        outs = [torch.empty(nq * 12, dtype=torch.uint8, device="cuda") for _ in range(3)]
        gathers.append((pm, dp.data_ptr(), dn.data_ptr(), nq, jb["k"], jb["r"], jb["paths"], jb["rho"],
                        outs[0].data_ptr(), outs[1].data_ptr(), outs[2].data_ptr()))
        keep.append((pm, dp, dn, outs))
    bg = sk.Batch(scenes, views, gathers)
    cfg_g, _ = bg.autotune(frames=1)
    ms_g = time_steps(torch, lambda: bg.frame(sptr), steps, warmup, stream)
    lat_g = bg.latency_ns() / 1e6
    bg.check()
    nq_total = sum(len(jb["p"]) for jb in jobs)
    with_g = {"ms": ms_g, "rays_per_s": n / (ms_g * 1e-3) / 1e6, "gathers_per_s": nq_total / (ms_g * 1e-3) / 1e6,
              "config": cfg_g, "jobs": len(jobs), "queries": nq_total,
              "p50_ms": float(np.percentile(lat_g, 50)), "p99_ms": float(np.percentile(lat_g, 99)),
              "job_p50_ms": float(np.percentile(lat_g[len(scenes):], 50)),
              "job_max_ms": float(lat_g[len(scenes):].max())}
    bg.close()
    for pm, *_ in keep:
        pm.close()
    return {"ms": ms, "value": n / (ms * 1e-3) / 1e6, "rays": n, "config": cfg, "tune_ms": tune_ms,
            "tune_s": tune_s, "build_s": build_s, "p50_ms": float(np.percentile(lat, 50)),
            "p99_ms": float(np.percentile(lat, 99)), "max_ms": float(lat.max()), "n_scenes": len(scenes),
            "tri_range": [min(tris), max(tris)], "tris_total": sum(tris), "specs": specs, "batch": batch,
            "scenes": scenes, "with_gathers": with_g, "build_ms": build_ms, "tris": tris}


def bench_c4(torch, sk, wl, steps, warmup, stream, spec=None, scene=None):
    s = spec if spec is not None else wl.scene_room()
    if scene is None:
        scene = sk.Scene(s["meshes"])
    dls = sk.DLScene(scene, s["mat"], s["kd"], s["ks"], s["rough"], s["light_prim"], s["light_radiance"],
                     prim_group=s["prim_group"], n_groups=s["n_groups"])
    w = h = 512
    spp = 8
    cam, cam15 = sk.camera(*s["camera"], w, h)
    th0 = s["theta0"]
    tt = wl.perturbed_theta(th0)
    sptr = stream.cuda_stream
    target = torch.empty((w * h, 3), dtype=torch.float64, device="cuda")
    dls.render_replay(tt, tt, cam, w, h, spp, target.data_ptr(), 4, 4, 4321, sptr)
    img = torch.empty_like(target)
    adj = torch.empty_like(target)
    grad = torch.empty(dls.n_params, dtype=torch.float64, device="cuda")
    grad_rc = torch.empty_like(grad)
    loss = torch.empty((), dtype=torch.float64, device="cuda")
    tape = [None]

    def step():
        # one gradient evaluation: forward image with its derivative tape, L2 loss and its image
        # gradient, contraction of the tape (the tape's buffers are reused between evaluations)
        tape[0] = dls.render_tape(th0, cam, w, h, spp, img.data_ptr(), 1 << 36, 4, 4, 1234, sptr, tape[0],
                                  count_rays=False)
        torch.sub(img, target, out=adj)
        loss.copy_((adj * adj).sum())
        adj.mul_(2.0)
        tape[0].adjoint(adj.data_ptr(), grad.data_ptr(), sptr)

    def step_recompute():
        # the same evaluation when the tape does not fit: forward image, loss, adjoint replay
        dls.render_replay(th0, th0, cam, w, h, spp, img.data_ptr(), 4, 4, 1234, sptr, count_rays=False)
        torch.sub(img, target, out=adj)
        loss.copy_((adj * adj).sum())
        adj.mul_(2.0)
        dls.adjoint(th0, cam, w, h, spp, adj.data_ptr(), grad_rc.data_ptr(), 4, 4, 1234, sptr, count_rays=False)

    ms_rc = time_steps(torch, step_recompute, steps, warmup, stream)
    ms = time_steps(torch, step, steps, warmup, stream)
    g, g_rc = grad.cpu().numpy(), grad_rc.cpu().numpy()
    tape[0].close()
    return {"ms": ms, "value": 1e3 / ms, "ms_recompute": ms_rc, "n_params": dls.n_params, "loss": float(loss.item()),
            "grad_norm": float(np.linalg.norm(g)), "grad_nonzero": int(np.count_nonzero(g)),
            "tape_vs_recompute": float(np.abs(g - g_rc).max() / max(np.abs(g_rc).max(), 1e-300)), "spec": s,
            "cam15": cam15, "theta0": th0, "target_theta": tt, "target": target, "img": img}


def fg_inputs(sk, wl, s, scene_stream):
    """Photon maps and irradiance map of the final-gather workload: 100k direct, 400k indirect and
    20k caustic photons spread uniformly over the c2 room's surfaces (seed 31)."""
    mesh = wl.merge(s["meshes"])
    sets = wl.photon_maps(mesh, 100_000, 400_000, 20_000, seed=31)
    maps = [sk.PhotonMap(p["pos"], p["weight"], p["dir"], p["normal"]) for p in sets]
    area = wl.scene_area(s["meshes"])
    paths = (200_000, 150_000, 250_000)
    t0 = time.time()
    irr = sk.IrradianceMap(direct=maps[0], indirect=maps[1], caustic=maps[2], caustic_paths=paths[0],
                           direct_paths=paths[1], indirect_paths=paths[2], scene_area=area, stream=scene_stream)
    irr_s = time.time() - t0
    return sets, maps, area, paths, irr, irr_s


def bench_c5(torch, sk, wl, steps, warmup, stream, spec=None, scene=None):
    s = spec if spec is not None else wl.scene_room()
    if scene is None:
        scene = sk.Scene(s["meshes"])
    dls = sk.DLScene(scene, s["mat"], s["kd"], s["ks"], s["rough"], s["light_prim"], s["light_radiance"])
    sptr = stream.cuda_stream
    sets, maps, area, paths, irr, irr_s = fg_inputs(sk, wl, s, sptr)
    w = h = 1024
    spp, gs = 4, 8
    cam, cam15 = sk.camera(*s["camera"], w, h)
    d_rgb = torch.empty(w * h * 12, dtype=torch.uint8, device="cuda")

    def step():
        dls.final_gather(irr, maps[1], area, cam, w, h, spp, gs, d_rgb.data_ptr(), None, 1234, sptr,
                         count_rays=False)

    ms = time_steps(torch, step, steps, warmup, stream)
    rays = dls.final_gather(irr, maps[1], area, cam, w, h, spp, gs, d_rgb.data_ptr(), None, 1234, sptr)
    total = int(rays.sum())
    rgb = d_rgb.cpu().numpy().view(np.float32)
    return {"ms": ms, "value": total / (ms * 1e-3) / 1e6, "rays": [int(x) for x in rays],
            "samples_per_s": w * h * spp / (ms * 1e-3), "irr_n": irr.n, "irr_max_dist": irr.max_lookup_dist,
            "irr_build_s": irr_s, "mean_rgb": float(rgb.mean()), "spec": s, "cam15": cam15, "sets": sets,
            "area": area, "paths": paths}


def cpu_baseline_c5(c5, threads, rows):
    import oracle
    s = c5["spec"]
    acc = oracle.Accel(s["meshes"])
    dls = oracle.DLScene(acc, s["mat"], s["kd"], s["ks"], s["rough"], s["light_prim"], s["light_radiance"])
    om = [oracle.PhotonMap(p["pos"], p["weight"], p["dir"], p["normal"]) for p in c5["sets"]]
    oirr = oracle.IrradianceMap(direct=om[0], indirect=om[1], caustic=om[2], paths=c5["paths"], scene_area=c5["area"],
                                nthreads=threads)
    w = 1024
    rows = max(1, min(rows, 1024))
    r0 = 512 - rows // 2
    dt, (_, grays) = median_runs(lambda: oracle.fg_render(acc, dls, oirr, om[1], c5["area"], c5["cam15"], w, 1024, 4,
                                                          8, 1234, r0 * w, (r0 + rows) * w, nthreads=threads))
    n = rows * w * 4 + grays
    return {"value": n / dt / 1e6, "unit": "Mrays/s", "cores": threads, "kind": "port",
            "sample": "image rows %d-%d of the final-gather frame (%d pixels x 4 spp, %d camera + gather rays), "
                      "1 warm-up, median of 5 runs (%.2f s)" % (r0, r0 + rows - 1, rows * w, n, dt)}


def bench_c6(torch, sk, wl, steps, warmup, stream, spec=None, scene=None):
    s = spec if spec is not None else wl.scene_room()
    if scene is None:
        scene = sk.Scene(s["meshes"])
    dls = sk.DLScene(scene, s["mat"], s["kd"], s["ks"], s["rough"], s["light_prim"], s["light_radiance"])
    sptr = stream.cuda_stream
    n_paths, per_round, big = 4_000_000, 1 << 20, 10 ** 12

    def shoot():
        return dls.shoot_photons(n_paths, (big, big, big), per_round, 77, sptr)

    for _ in range(warmup):
        shoot().close()
    times, ps = [], None
    for _ in range(steps):
        if ps is not None:
            ps.close()
        torch.cuda.synchronize()
        t = time.perf_counter()
        ps = shoot()
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t)
    ms = 1e3 * statistics.median(times)
    photons = sum(ps.counts)
    # the same shooting with the limits a drop-in caller passes by default: the integrator clamps
    # every map to MAX_PHOTONS_IN_MAP = 40M (never reached by 4M paths, but finite)
    lim, lim_times = (40_000_000,) * 3, []
    for _ in range(max(3, steps // 2)):
        torch.cuda.synchronize()
        t = time.perf_counter()
        q = dls.shoot_photons(n_paths, lim, per_round, 77, sptr)
        torch.cuda.synchronize()
        lim_times.append(time.perf_counter() - t)
        assert q.counts == ps.counts
        q.close()
    ms_lim = 1e3 * statistics.median(lim_times)
    # kd build of the indirect map: on the device (from the set) and on the host (from host arrays)
    ps.photon_map(2).close()
    dev_ms, host_ms = [], []
    m = ps.export(2)
    for _ in range(3):
        torch.cuda.synchronize()
        t = time.perf_counter()
        pm = ps.photon_map(2)
        torch.cuda.synchronize()
        dev_ms.append(1e3 * (time.perf_counter() - t))
        pm.close()
        t = time.perf_counter()
        pm = sk.PhotonMap(m["pos"], m["weight"], m["dir"], m["normal"])
        host_ms.append(1e3 * (time.perf_counter() - t))
        pm.close()
    out = {"ms": ms, "value": n_paths / (ms * 1e-3) / 1e6, "photons": photons, "counts": ps.counts,
           "paths": ps.paths, "photons_per_s": photons / (ms * 1e-3), "spec": s,
           "kd_n": len(m["pos"]), "kd_device_ms": statistics.median(dev_ms), "kd_host_ms": statistics.median(host_ms),
           "ms_limits_40M": ms_lim}
    ps.close()
    return out


def bench_c7(torch, sk, wl, steps, warmup, stream, spec=None, scene=None):
    """C2's photon-mapping render: direct lighting + caustic + final gather per sample at 1024x1024,
    16 spp, 4 light + 4 BSDF samples, 8 (9) gather samples, over maps shot from the room's lights
    (2^20 paths), built on the device, and their irradiance map."""
    s = spec if spec is not None else wl.scene_room()
    if scene is None:
        scene = sk.Scene(s["meshes"])
    dls = sk.DLScene(scene, s["mat"], s["kd"], s["ks"], s["rough"], s["light_prim"], s["light_radiance"])
    sptr = stream.cuda_stream
    area = wl.scene_area(s["meshes"])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ps = dls.shoot_photons(1 << 20, (40_000_000,) * 3, 1 << 20, 77, sptr)
    maps = [ps.photon_map(j) if ps.counts[j] > 0 else None for j in range(3)]
    paths = [int(x) for x in ps.paths]
    irr = sk.IrradianceMap(direct=maps[1], indirect=maps[2], caustic=maps[0], caustic_paths=paths[0],
                           direct_paths=paths[1], indirect_paths=paths[2], scene_area=area, stream=sptr)
    torch.cuda.synchronize()
    maps_ms = 1e3 * (time.perf_counter() - t0)
    w = h = 1024
    spp, gs = 16, 8
    cam, cam15 = sk.camera(*s["camera"], w, h)
    d_rgb = torch.empty(w * h * 12, dtype=torch.uint8, device="cuda")

    def step():
        dls.pm_render(irr, maps[0], maps[2], area, cam, w, h, spp, d_rgb.data_ptr(), 4, 4, gs, 1234, sptr,
                      count_rays=False)

    ms = time_steps(torch, step, steps, warmup, stream)
    rays = dls.pm_render(irr, maps[0], maps[2], area, cam, w, h, spp, d_rgb.data_ptr(), 4, 4, gs, 1234, sptr)
    rgb = d_rgb.cpu().numpy().view(np.float32)
    exported = [ps.export(j) if ps.counts[j] > 0 else None for j in range(3)]
    out = {"ms": ms, "value": int(rays.sum()) / (ms * 1e-3) / 1e6, "rays": [int(x) for x in rays],
           "samples_per_s": w * h * spp / (ms * 1e-3), "mean_rgb": float(rgb.mean()), "maps_ms": maps_ms,
           "photons": [int(x) for x in ps.counts], "paths": paths, "irr_n": irr.n, "spec": s, "cam15": cam15,
           "area": area, "exported": exported}
    for m in maps:
        if m is not None:
            m.close()
    irr.close()
    ps.close()
    return out


def cpu_baseline_c7(c7, threads, rows):
    import oracle
    s = c7["spec"]
    acc = oracle.Accel(s["meshes"])
    dls = oracle.DLScene(acc, s["mat"], s["kd"], s["ks"], s["rough"], s["light_prim"], s["light_radiance"])
    om = [oracle.PhotonMap(m["pos"], m["weight"], m["dir"], m["normal"]) if m is not None else None
          for m in c7["exported"]]
    oirr = oracle.IrradianceMap(direct=om[1], indirect=om[2], caustic=om[0], paths=tuple(c7["paths"]),
                                scene_area=c7["area"], nthreads=threads)
    w = 1024
    r0 = 512 - rows // 2
    dt, (_, cnt) = median_runs(lambda: oracle.pm_render(acc, dls, oirr, om[2], c7["area"], c7["cam15"], w, 1024, 16,
                                                        4, 4, 8, 1234, r0 * w, (r0 + rows) * w, nthreads=threads))
    n = int(cnt.sum())
    return {"value": n / dt / 1e6, "unit": "Mrays/s", "cores": threads, "kind": "port",
            "sample": "image rows %d-%d of the photon-mapping frame (%d pixels x 16 spp, %d camera + shadow + BSDF + "
                      "gather rays), 1 warm-up, median of 5 runs (%.2f s)" % (r0, r0 + rows - 1, rows * w, n, dt)}


def bench_c8(torch, sk, wl, steps, warmup, stream, spec=None):
    """Triangle-tree build (scene creation: tree + device tables) of the 2,026,260-triangle room on
    the device (SKW_SCENE_DEVICE_BUILD) and on the host (the same tree), and of the 64 batched
    scenes both ways; scene descriptions are prepared before the timed region."""
    s = spec if spec is not None else wl.scene_room()
    desc, keep = sk.scene_desc(s["meshes"])

    def build(device):
        torch.cuda.synchronize()
        t = time.perf_counter()
        sc = sk.Scene.from_desc(desc, device_build=device)
        torch.cuda.synchronize()
        return 1e3 * (time.perf_counter() - t), sc

    for _ in range(warmup):
        build(True)[1].close()
    dev_ms = []
    for _ in range(steps):
        ms, sc = build(True)
        dev_ms.append(ms)
        info = sc.info()
        sc.close()
    host_ms = []
    for _ in range(2):
        ms, sc = build(False)
        host_ms.append(ms)
        assert sc.info() == info
        sc.close()
    specs = wl.batch_scenes()
    descs = [sk.scene_desc(sp[0]) for sp in specs]
    batch = {}
    for device in (True, False):
        torch.cuda.synchronize()
        t = time.perf_counter()
        for d, _ in descs:
            sk.Scene.from_desc(d, device_build=device).close()
        torch.cuda.synchronize()
        batch[device] = 1e3 * (time.perf_counter() - t)
    ms = statistics.median(dev_ms)
    tris = info["triangles"]
    return {"ms": ms, "value": tris / (ms * 1e-3) / 1e6, "host_ms": statistics.median(host_ms), "info": info,
            "batch_device_ms": batch[True], "batch_host_ms": batch[False],
            "batch_tris": int(sum(sum(len(m["tris"]) for m in sp[0]) for sp in specs)), "spec": s}


def cpu_baseline_c8(c8):
    import oracle
    dt, acc = median_runs(lambda: oracle.Accel(c8["spec"]["meshes"]), runs=3)
    assert acc.info()["nodes"] == c8["info"]["nodes"]
    return {"value": c8["info"]["triangles"] / dt / 1e6, "unit": "M triangles/s", "cores": 1, "kind": "port",
            "sample": "the whole room's tree built by the oracle's serial restatement of the reference builder; "
                      "1 warm-up, median of 3 runs (%.1f s each)" % dt}


def cpu_baseline_c6(c6, threads, n_paths):
    import oracle
    s = c6["spec"]
    acc = oracle.Accel(s["meshes"])
    dls = oracle.DLScene(acc, s["mat"], s["kd"], s["ks"], s["rough"], s["light_prim"], s["light_radiance"])
    dt, out = median_runs(lambda: oracle.photon_shoot(acc, dls, n_paths, (10 ** 12,) * 3, 1 << 20, 77,
                                                      nthreads=threads))
    photons = sum(len(m["pos"]) for m in out["maps"])
    return {"value": n_paths / dt / 1e6, "unit": "M paths/s", "cores": threads, "kind": "port",
            "sample": "the first %d of the 4M paths (%d photons), 1 warm-up, median of 5 runs (%.2f s)"
                      % (n_paths, photons, dt)}


def cpu_baseline_c3(c3, threads, n_sample):
    import oracle
    import torch
    batch, scenes, specs = c3["batch"], c3["scenes"], c3["specs"]
    order = sorted(range(len(scenes)), key=lambda i: c3["specs"][i][2] ** 2 * c3["specs"][i][3])
    pick = [order[int(round(j * (len(order) - 1) / max(n_sample - 1, 1)))] for j in range(n_sample)]
    work = []
    for i in pick:
        _, n = batch.hits(i)
        d = torch.empty(n * 64, dtype=torch.uint8, device="cuda")
        batch.rays(i, d.data_ptr())
        torch.cuda.synchronize()
        work.append((oracle.Accel(specs[i][0]), d.cpu().numpy().view(np.float64).reshape(n, 8)))
    rays_total = sum(len(r) for _, r in work)
    secs, _ = median_runs(lambda: [acc.closest_hit(r, nthreads=threads) for acc, r in work])
    return {"value": rays_total / secs / 1e6, "unit": "Mrays/s", "cores": threads, "kind": "port",
            "sample": "every camera ray of %d of the %d scenes (spread over frame sizes), %d rays; 1 warm-up, "
                      "median of 5 runs" % (len(pick), len(scenes), rays_total)}


def cpu_baseline_c4(c4, threads, rows):
    import oracle
    s = c4["spec"]
    acc = oracle.Accel(s["meshes"])
    dls = oracle.DLScene(acc, s["mat"], s["kd"], s["ks"], s["rough"], s["light_prim"], s["light_radiance"],
                         prim_group=s["prim_group"], n_groups=s["n_groups"])
    w, spp = 512, 8
    r0 = 256 - rows // 2
    p0, p1 = r0 * w, (r0 + rows) * w
    th0 = c4["theta0"]
    target = c4["target"].cpu().numpy()[p0:p1]
    def run():
        img = dls.render_replay(th0, th0, c4["cam15"], w, 512, spp, 4, 4, 1234, p0, p1, nthreads=threads)
        dls.gradient(th0, 2.0 * (img - target), c4["cam15"], w, 512, spp, 4, 4, 1234, p0, p1, nthreads=threads)

    dt, _ = median_runs(run)
    return {"value": (rows / 512.0) / dt, "unit": "gradients/s", "cores": threads, "kind": "port",
            "sample": "image rows %d-%d of the 512x512 x 8 spp frame (forward image + FP64 forward-mode "
                      "gradient of those rows), 1 warm-up, median of 5 runs, time scaled by 512/%d rows"
                      % (r0, r0 + rows - 1, rows)}


def cpu_baseline_c2(c2, threads, rows):
    import oracle  # CPU restatement of the reference path: the baseline, never the product
    s = c2["spec"]
    acc = oracle.Accel(s["meshes"])
    dls = oracle.DLScene(acc, s["mat"], s["kd"], s["ks"], s["rough"], s["light_prim"], s["light_radiance"])
    w, spp = 1024, 16
    # rows around the image centre (the middle half of the frame by default)
    rows = max(1, min(rows, 1024))
    r0 = 512 - rows // 2
    dt, (_, cnt) = median_runs(lambda: dls.render(c2["cam15"], w, 1024, spp, 4, 4, 1234, r0 * w, (r0 + rows) * w,
                                                  nthreads=threads))
    return {"value": float(cnt.sum()) / dt / 1e6, "unit": "Mrays/s", "cores": threads, "kind": "port",
            "sample": "image rows %d-%d of the c2 frame (%d pixels x 16 spp, %d rays), 1 warm-up, median of 5 runs "
                      "(%.2f s; %s)" % (r0, r0 + rows - 1, rows * w, int(cnt.sum()), dt, cpu_info())}


def cpu_baseline_c0(c0, threads):
    import oracle
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
            "sample": "all %d c0 primary rays per run; 1 warm-up, median of 5 runs" % len(rays)}


def cpu_baseline_c1(c1, threads, sample):
    import oracle
    ph = c1["photons"]
    om = oracle.PhotonMap(ph["pos"], ph["weight"], ph["dir"], ph["normal"])
    sel = slice(0, min(sample, len(c1["p"])))
    p, n = c1["p"][sel], c1["n"][sel]
    med, _ = median_runs(lambda: om.irradiance(p, n, 100, 0.05, with_sets=False, nthreads=threads))
    return {"value": len(p) / med / 1e6, "unit": "M gathers/s", "cores": threads, "kind": "port",
            "sample": "first %d of the c1 gather points; 1 warm-up, median of 5 runs" % len(p)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5, help="timed c2 frames")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--only", choices=["c0", "c1", "c2", "c3", "c4", "c5", "c6", "c7", "c8"], default=None,
                    help="run one workload (profiling runs); the printed line still names c2's metric only "
                         "when c2 runs")
    ap.add_argument("--cpu-rows", type=int, default=512,
                    help="c2 image rows in the CPU baseline sample (1024: the whole frame; 1 warm-up + median of "
                         "5 runs)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run a plain device-to-device copy of BYTES (3 times): the reference "
                         "point for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE readings")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        src = torch.ones(args.copy_calibration // 4, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": args.copy_calibration}))
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (the scene is replicated per "
              "GPU and rays shard by image tile); run with --gpus 1", file=sys.stderr)
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
    run = (lambda w: args.only is None or args.only == w)

    c0 = c1 = c2 = None
    if run("c0") or run("c1"):
        log("c0: closest hit, 512x512 rays, 81,932 triangles")
        c0 = bench_c0(torch, sk, wl, 20, 3, stream)
        log("c0: %.3f ms/step, %.1f Mrays/s" % (c0["ms"], c0["value"]))
    if run("c1"):
        log("c1: photon gather, 1M photons, k=100, r=0.05")
        c1 = bench_c1(torch, sk, wl, 10, 2, stream, c0)
        log("c1: %.3f ms/step, %.2f M gathers/s" % (c1["ms"], c1["value"]))
    if run("c2"):
        log("c2: direct lighting, 2,026,260 triangles, 1024x1024 x 16 spp")
        c2 = bench_c2(torch, sk, wl, args.steps, args.warmup, stream)
        log("c2: %.1f ms/frame, %.1f Mrays/s, rays %s" % (c2["ms"], c2["value"], c2["rays"]))
    c3 = c4 = None
    if run("c3"):
        log("c3: batched multi-scene closest hit, 64 scenes")
        c3 = bench_c3(torch, sk, wl, 5, 1, stream)
        log("c3: %.2f ms/frame, %.1f Mrays/s, p50 %.2f ms p99 %.2f ms, config %s" %
            (c3["ms"], c3["value"], c3["p50_ms"], c3["p99_ms"], c3["config"]))
        g = c3["with_gathers"]
        log("c3 + gathers: %.2f ms/frame, %.1f Mrays/s + %.1f M gathers/s, item p50 %.2f ms p99 %.2f ms" %
            (g["ms"], g["rays_per_s"], g["gathers_per_s"], g["p50_ms"], g["p99_ms"]))
    if run("c4"):
        log("c4: direct-lighting gradient, 512x512 x 8 spp, 132 parameters")
        c4 = bench_c4(torch, sk, wl, 3, 1, stream, c2["spec"] if c2 else None, c2["scene"] if c2 else None)
        log("c4: %.1f ms/gradient, %.2f gradients/s" % (c4["ms"], c4["value"]))
    c5 = None
    if run("c5"):
        log("c5: photon final gather, 1024x1024 x 4 spp, 9 gather samples")
        c5 = bench_c5(torch, sk, wl, 3, 1, stream, c2["spec"] if c2 else None, c2["scene"] if c2 else None)
        log("c5: %.1f ms/frame, %.1f Mrays/s, rays %s" % (c5["ms"], c5["value"], c5["rays"]))
    c6 = None
    if run("c6"):
        log("c6: photon shooting, 4M paths from the c2 scene's lights, device kd build")
        c6 = bench_c6(torch, sk, wl, 10, 2, stream, c2["spec"] if c2 else None, c2["scene"] if c2 else None)
        log("c6: %.1f ms per 4M paths, %.1f M paths/s, photons %s, kd build %d photons: device %.1f ms, host %.1f ms"
            % (c6["ms"], c6["value"], c6["counts"], c6["kd_n"], c6["kd_device_ms"], c6["kd_host_ms"]))
    c7 = None
    if run("c7"):
        log("c7: photon-mapping render (direct + caustic + final gather), 1024x1024 x 16 spp, shot maps")
        c7 = bench_c7(torch, sk, wl, 3, 1, stream, c2["spec"] if c2 else None, c2["scene"] if c2 else None)
        log("c7: %.1f ms/frame, %.1f Mrays/s, rays %s, maps %.1f ms (photons %s)"
            % (c7["ms"], c7["value"], c7["rays"], c7["maps_ms"], c7["photons"]))
    c8 = None
    if run("c8"):
        log("c8: triangle-tree build, 2,026,260 triangles (and the 64 batched scenes), device and host")
        c8 = bench_c8(torch, sk, wl, 5, 1, stream, c2["spec"] if c2 else None)
        log("c8: device %.1f ms (%.1f M triangles/s), host %.1f ms; 64 batched scenes: device %.1f ms, host %.1f ms"
            % (c8["ms"], c8["value"], c8["host_ms"], c8["batch_device_ms"], c8["batch_host_ms"]))
    cpu0 = cpu1 = cpu2 = cpu3 = cpu4 = cpu5 = cpu6 = cpu7 = cpu8 = None
    if not args.no_cpu:
        log("cpu baseline (oracle, %d threads)" % threads)
        if c2 is not None:
            cpu2 = cpu_baseline_c2(c2, threads, args.cpu_rows)
        if c0 is not None:
            cpu0 = cpu_baseline_c0(c0, threads)
        if c1 is not None:
            cpu1 = cpu_baseline_c1(c1, threads, 1 << 18)
        if c3 is not None:
            cpu3 = cpu_baseline_c3(c3, threads, 16)
        if c4 is not None:
            cpu4 = cpu_baseline_c4(c4, threads, 128)
        if c5 is not None:
            cpu5 = cpu_baseline_c5(c5, threads, 128)
        if c6 is not None:
            cpu6 = cpu_baseline_c6(c6, threads, 4_000_000)
        if c7 is not None:
            cpu7 = cpu_baseline_c7(c7, threads, 32)
        if c8 is not None:
            cpu8 = cpu_baseline_c8(c8)
    comps = {}
    if c3 is not None:
        comps["c3_batch"] = {
            "metric": "Mrays/s (64 scenes, one persistent launch per frame)", "value": round(c3["value"], 3),
            "ms_per_step": round(c3["ms"], 4), "rays_per_step": c3["rays"],
            "scene_latency_ms": {"p50": round(c3["p50_ms"], 4), "p99": round(c3["p99_ms"], 4),
                                 "max": round(c3["max_ms"], 4)},
            "autotuned": c3["config"], "autotune_s": round(c3["tune_s"], 2),
            "scene_build": {"mode": "tree built on the device (SKW_SCENE_DEVICE_BUILD), scene creation timed",
                            "total_ms": round(sum(c3["build_ms"]), 2),
                            "ms_per_scene": [round(x, 3) for x in c3["build_ms"]],
                            "triangles_per_scene": c3["tris"]},
            "with_photon_gathers": {
                "ms_per_step": round(c3["with_gathers"]["ms"], 4),
                "Mrays_per_s": round(c3["with_gathers"]["rays_per_s"], 3),
                "M_gathers_per_s": round(c3["with_gathers"]["gathers_per_s"], 3),
                "item_latency_ms": {"p50": round(c3["with_gathers"]["p50_ms"], 4),
                                    "p99": round(c3["with_gathers"]["p99_ms"], 4),
                                    "gather_jobs_p50": round(c3["with_gathers"]["job_p50_ms"], 4),
                                    "gather_jobs_max": round(c3["with_gathers"]["job_max_ms"], 4)},
                "autotuned": c3["with_gathers"]["config"],
                "workload": "the 64 scenes' camera rays plus %d photon-gather jobs (%d queries; 250k photons "
                            "each, k=100, radius holding 40 photons on average, disc estimate) in one persistent "
                            "launch per frame" % (c3["with_gathers"]["jobs"], c3["with_gathers"]["queries"])},
            "workload": "%d scenes, %d-%d triangles (%d total), 128/256/512 square frames, 4 spp" %
                        (c3["n_scenes"], c3["tri_range"][0], c3["tri_range"][1], c3["tris_total"]),
            "cpu_baseline": cpu3,
        }
    if c4 is not None:
        comps["c4_adjoint"] = {
            "metric": "gradients/s (132-parameter direct-lighting gradient: forward + L2 loss + adjoint)",
            "value": round(c4["value"], 4), "ms_per_step": round(c4["ms"], 3), "n_params": c4["n_params"],
            "mode": "tape (per-sample derivatives stored by the forward frame)",
            "ms_per_step_recompute": round(c4["ms_recompute"], 3),
            "tape_vs_recompute_max_rel": c4["tape_vs_recompute"],
            "loss": c4["loss"], "grad_norm": c4["grad_norm"], "grad_nonzero": c4["grad_nonzero"],
            "workload": "c2 scene at 512x512, 8 spp, 4 light + 4 BSDF samples; target = perturbed parameters "
                        "(seed 4)",
            "cpu_baseline": cpu4,
        }
    if c6 is not None:
        comps["c6_photon_shooting"] = {
            "metric": "M photon paths/s (photon shooting from the area lights)", "value": round(c6["value"], 3),
            "ms_per_step": round(c6["ms"], 3), "photons_per_step": c6["counts"], "paths_per_map": c6["paths"],
            "photons_per_s": round(c6["photons_per_s"], 1),
            "ms_per_step_limits_40M": round(c6["ms_limits_40M"], 3),
            "kd_build": {"photons": c6["kd_n"], "device_ms": round(c6["kd_device_ms"], 3),
                         "host_ms": round(c6["kd_host_ms"], 3)},
            "workload": "c2 scene, 4,000,000 paths in rounds of 2^20, all maps unbounded (seed 77); kd build "
                        "of the indirect map on the device and on the host (same tree)",
            "cpu_baseline": cpu6,
        }
    if c7 is not None:
        comps["c7_photon_mapping"] = {
            "metric": "Mrays/s (photon-mapping render: camera + shadow + BSDF + gather rays)",
            "value": round(c7["value"], 3), "ms_per_step": round(c7["ms"], 3), "rays_per_step": c7["rays"],
            "samples_per_s": round(c7["samples_per_s"], 1),
            "maps": {"paths": 1 << 20, "photons": c7["photons"], "irradiance_photons": c7["irr_n"],
                     "shoot_build_ms": round(c7["maps_ms"], 2)},
            "workload": "c2 scene at 1024x1024, 16 spp, 4 light + 4 BSDF samples, 8 (9) gather samples; "
                        "Le + direct + caustic + final gather per sample over maps of 2^20 paths shot from the "
                        "room's lights (seed 77), device-built kd-trees and irradiance map",
            "cpu_baseline": cpu7,
        }
    if c8 is not None:
        comps["c8_tree_build"] = {
            "metric": "M triangles/s (triangle-tree build: scene creation with the tree and device tables)",
            "value": round(c8["value"], 3), "ms_per_step": round(c8["ms"], 3), "host_build_ms": round(c8["host_ms"], 3),
            "nodes": c8["info"]["nodes"], "max_depth": c8["info"]["max_depth"],
            "batch_64_scenes": {"triangles": c8["batch_tris"], "device_ms": round(c8["batch_device_ms"], 2),
                                "host_ms": round(c8["batch_host_ms"], 2)},
            "workload": "the c2 room (2,026,260 triangles) built on the device (the same tree as the host "
                        "builder, node for node), host builder timed beside it; the 64 batched scenes both ways",
            "cpu_baseline": cpu8,
        }
    if c5 is not None:
        comps["c5_final_gather"] = {
            "metric": "Mrays/s (photon final gather: camera + gather rays)", "value": round(c5["value"], 3),
            "ms_per_step": round(c5["ms"], 3), "rays_per_step": c5["rays"],
            "samples_per_s": round(c5["samples_per_s"], 1),
            "irradiance_map": {"photons": c5["irr_n"], "max_lookup_dist": c5["irr_max_dist"],
                               "build_s": round(c5["irr_build_s"], 3)},
            "workload": "c2 scene at 1024x1024, 4 spp, 8 gather samples (9, a square); 100k direct / 400k "
                        "indirect / 20k caustic photons uniform over the surfaces (seed 31)",
            "cpu_baseline": cpu5,
        }
    if c0 is not None:
        comps["c0_closest_hit"] = {
            "metric": "Mrays/s closest-hit", "value": round(c0["value"], 3), "ms_per_step": round(c0["ms"], 5),
            "rays_per_step": c0["n"], "cpu_baseline": cpu0,
            "roofline": dict({"bound": "hbm", "achieved": round(c0["achieved_gbs"], 2), "peak": HBM_PEAK_GBS,
                              "unit": "GB/s", "frac": round(c0["achieved_gbs"] / HBM_PEAK_GBS, 5)},
                             **pmc_fields("closest_hit_c0_pmc"),
                             **{"limiter": "launch latency and the tail of the slowest walks (4,096 waves, one "
                                           "round of the chip): VALU issue and active lanes in 'pmc'",
                                "kernel": "k_trace_lds<false,false,false,32> (one launch per step)",
                                "alg_bytes_per_launch": c0["alg_bytes"], "avg_launch_ms": round(c0["ms"], 5),
                                "bytes_per_ray": round(c0["alg_bytes"] / c0["n"], 2),
                                "mean_node_tests": round(c0["mean_pops"], 3),
                                "mean_tri_tests": round(c0["mean_tests"], 3)}),
        }
    if c1 is not None:
        comps["c1_photon_gather"] = {
            "metric": "M photon-gathers/s", "value": round(c1["value"], 4), "ms_per_step": round(c1["ms"], 4),
            "gathers_per_step": c1["nq"], "mean_photons_found": round(c1["mean_found"], 3), "cpu_baseline": cpu1,
            "roofline": dict({"bound": "hbm", "achieved": round(c1["achieved_gbs"], 2), "peak": HBM_PEAK_GBS,
                              "unit": "GB/s", "frac": round(c1["achieved_gbs"] / HBM_PEAK_GBS, 5)},
                             **pmc_fields("photon_gather_c1_pmc"),
                             **{"limiter": "latency of the lookups' dependent node loads with divergent lanes: "
                                           "VALU issue and active lanes in 'pmc'; HBM traffic is a few percent "
                                           "of the algorithmic bytes (L2 hits)",
                                "kernel": "k_gather_lds<0,32,false> (one launch per step)",
                                "alg_bytes_per_launch": c1["alg_bytes"], "avg_launch_ms": round(c1["ms"], 5),
                                "mean_visits": round(c1["mean_visits"], 3),
                                "mean_inside_radius": round(c1["mean_inside"], 3)}),
        }
    if c2 is not None:
        line = {
            "metric": "Mrays/s (direct lighting, all rays: primary + occlusion + BSDF)",
            "value": round(c2["value"], 3), "unit": "Mrays/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": round(c2["ms"], 3), "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f64", "data": "synthetic",
            "config": {"workload": "c2: seeded room, 2,026,260 triangles (200 displaced icospheres, 4 quad area "
                                   "lights, matte + plastic), 1024x1024, 16 spp, 4 light + 4 BSDF samples per hit, "
                                   "direct lighting, f32 radiance out",
                       "rays_per_frame": c2["rays"], "triangles": c2["info"]["triangles"],
                       "nodes": c2["info"]["nodes"]},
            # The frame's dominant kernel is the traversal (primary, occlusion and BSDF launches):
            # algorithmic bytes from the counting frame over the traversal stages' event time.
            "roofline": dict({"bound": "hbm", "achieved": round(c2["achieved_gbs"], 2), "peak": HBM_PEAK_GBS,
                              "unit": "GB/s", "frac": round(c2["achieved_gbs"] / HBM_PEAK_GBS, 5)},
                             **pmc_fields("trace_c2_pmc"),
                             **{"limiter": "FP64 VALU issue of divergent walks: VALU issue utilisation and active "
                                           "lanes per instruction in 'pmc'; most node and triangle reads hit L2, so "
                                           "HBM traffic ('traffic') is a fraction of the algorithmic bytes and "
                                           "'frac' is not a DRAM figure",
                                "kernel": "k_trace_lds_wide<.., 24> (closest + any hit; the walk compiled for five waves "
                                          "per SIMD): per launch of %d sample passes, "
                                          "one primary, one occlusion and one BSDF trace" % (16 // c2["launches"]),
                                "launches_per_frame": 3 * c2["launches"],
                                "alg_bytes_per_launch": c2["alg_bytes"] // (3 * c2["launches"]),
                                "avg_launch_ms": round(c2["trace_ms"] / (3 * c2["launches"]), 4),
                                "trace_ms_per_frame": round(c2["trace_ms"], 3),
                                "alg_bytes_per_frame": c2["alg_bytes"],
                                "mean_node_tests": [round(x, 3) for x in c2["mean_node_tests"]],
                                "mean_tri_tests": [round(x, 3) for x in c2["mean_tri_tests"]]}),
            "stage_ms_per_frame": {k: round(v, 3) for k, v in c2["stage_ms"].items()},
            "cpu_baseline": cpu2,
            "components": comps,
        }
    else:
        c = (comps.get("c0_closest_hit") or comps.get("c1_photon_gather") or comps.get("c3_batch")
             or comps.get("c4_adjoint") or comps.get("c5_final_gather") or comps.get("c6_photon_shooting")
             or comps.get("c7_photon_mapping") or comps.get("c8_tree_build"))
        line = {"metric": c["metric"], "value": c["value"], "unit": c["metric"].split(" ")[0], "n_gpus": 1,
                "steps": 20, "warmup": 3, "ms_per_step": c["ms_per_step"], "higher_is_better": True,
                "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
                "config": {"workload": args.only}, "roofline": c.get("roofline"), "cpu_baseline": c["cpu_baseline"],
                "components": comps}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
