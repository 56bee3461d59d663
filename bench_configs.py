"""Timing of the other workloads with bench.py's harness (`bench.py --config N`).

  2  scene step: 512 rigid entities (boxes, icospheres, noisy tori) in a 50 m
     cube, 100 frames at 1/60 s, broad + mid + narrow each frame (cd_frame);
     also timed with the uncollide rays and the response (the whole frame)
  4  precision stress: 2^24 triangle pairs (80% random, 10% near-coplanar,
     10% touching); plain float32 Möller vs the filtered exact test, each
     checked against the exact oracle on a sample
  5  deformable batch: 64 skinned tori (20,000 triangles, 64 bones, 4 weights
     per vertex) skinned and refit every frame, then tested against a
     316 x 316 fBm terrain in one batched launch
  6  glTF mesh ingest: the three shipped scenes (Apple, Table_room, Sponza
     environmentTest: 50 meshes, 64,842 collision triangles) parsed and
     given their trees, one document after another, as a model load does

Each prints one JSON line in bench.py's format.  The oracle (CPU
restatement) only appears as the cpu_baseline / mismatch checker.
"""
from __future__ import annotations

import sys
import time

import numpy as np


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def _line(metric, value, unit, ms, args, config, dtype="f32", extra=None, data="synthetic"):
    d = {"metric": metric, "value": round(value, 2), "unit": unit, "n_gpus": 1, "steps": args.steps,
         "warmup": args.warmup, "ms_per_step": round(ms, 4), "higher_is_better": True, "scaling": "weak",
         "vs_baseline": None, "dtype": dtype, "data": data, "config": config}
    d.update(extra or {})
    return d


def config2(args):
    from inmyroom_vulkan_amd import ENTITY_DTYPE, MODE_BATCHED, MODE_PLAIN, Collider, scenes

    sc = scenes.Scene(n_entities=512, seed=2)
    gpu = Collider(leaf_max=4)
    tuning = {"use_graph": 1}  # batched frames replay one captured graph whatever their pair count
    for kv in filter(None, getattr(args, "tuning", "").split(",")):
        k, v = kv.split("=")
        tuning[k.strip()] = int(v)
    gpu.set_tuning(**tuning)
    ids = gpu.register_soups(sc.soups)
    frames = []
    for _ in range(100):
        frames.append(sc.entities(ids, ENTITY_DTYPE))
        sc.step()
    out = {}
    # --plain-steps 0 leaves the plain mode out (profiles of the batched frame)
    modes = (("plain", MODE_PLAIN),) * (args.plain_steps > 0) + (("batched", MODE_BATCHED),)
    for name, mode in modes:
        for f in frames[: args.warmup]:
            gpu.frame(f, mode)
        t = time.perf_counter()
        npairs = ncons = 0
        dev = 0.0
        for f in frames:
            p, c = gpu.frame(f, mode)
            npairs += len(p)
            ncons += len(c)
            dev += gpu.timings()["total_ms"]
        wall = (time.perf_counter() - t) * 1e3 / len(frames)
        out[name] = {"ms_per_frame": round(wall, 4), "device_ms_per_frame": round(dev / len(frames), 4),
                     "pairs_per_frame": npairs / len(frames), "contacts_per_frame": ncons / len(frames)}
        log(f"config 2 {name}: {wall:.3f} ms/frame wall, {dev / len(frames):.3f} ms device")
        # the whole ExecuteCollisionDetection: frame, uncollide rays, ray shooting and the split
        t = time.perf_counter()
        nrays = nshot = 0
        for f in frames:
            gpu.frame(f, mode)
            _, rays = gpu.uncollide_rays()
            resp = gpu.shoot_rays()
            nrays += len(rays)
            nshot += int(resp["shot"].sum())
        full = (time.perf_counter() - t) * 1e3 / len(frames)
        out[name].update({"full_ms_per_frame": round(full, 4), "rays_per_frame": nrays / len(frames),
                          "pairs_shot_per_frame": nshot / len(frames)})
        log(f"config 2 {name}: {full:.3f} ms/frame with rays and response")
    cpu = None
    if not args.no_cpu:
        from oracle import Oracle

        orc = Oracle()
        trees = [orc.tree(s, 4) for s in sc.soups]
        t = time.perf_counter()
        n = 0
        for f in frames:
            orc.frame(trees, sc.mesh_of, f["current"], f["should_callback"], mode=0)
            n += 1
            if time.perf_counter() - t > args.cpu_seconds:
                break
        el = time.perf_counter() - t
        cpu = {"value": round(n / el, 2), "unit": "frames/s", "cores": 1, "kind": "port",
               "sample": f"first {n} of 100 frames, reference semantics, single thread"}
    ms = out["batched"]["ms_per_frame"]
    return _line("scene step frames/s, 512 entities, broad + narrow", 1e3 / ms, "frames/s", ms, args,
                 {"workload": "512 rigid entities (box 12 / icosphere 1,280 / torus 8,192 triangles), 50 m cube, "
                              "100 frames at 1/60 s", "parallelism": "single GPU"},
                 extra={"modes": out, "cpu_baseline": cpu})


def config4(args):
    from inmyroom_vulkan_amd import MODE_BATCHED, MODE_PLAIN, Collider, scenes

    n = args.pairs
    t0 = time.time()
    a9, b9, kind = scenes.config4(n, seed=4)
    log(f"config 4: {n:,} pairs generated in {time.time() - t0:.1f}s")
    gpu = Collider(leaf_max=4)
    res = {}
    for name, mode in (("plain", MODE_PLAIN), ("batched", MODE_BATCHED)):
        f, _ = gpu.tritri(a9[:1024], b9[:1024], mode)
        ms = []
        for _ in range(max(1, args.steps)):
            f, _ = gpu.tritri(a9, b9, mode)
            ms.append(gpu.timings()["narrow_ms"])
        c = gpu.counters()
        res[name] = {"flags": f, "kernel_ms": float(np.median(ms)), "fp64": c["fp64_fallbacks"],
                     "exact": c["exact_fallbacks"]}
        log(f"config 4 {name}: {np.median(ms):.3f} ms kernel")
    sample = min(n, args.oracle_pairs)
    from oracle import Oracle

    orc = Oracle()
    t = time.perf_counter()
    ex = orc.tritri_exact(a9[:sample], b9[:sample])
    t_ex = time.perf_counter() - t
    t = time.perf_counter()
    fo, _ = orc.tritri(a9[:sample], b9[:sample])
    t_cpu = time.perf_counter() - t
    truth = ex == 1
    plain_hit = ((res["plain"]["flags"][:sample] & 1) != 0) & ((res["plain"]["flags"][:sample] & 2) == 0)
    batched_hit = res["batched"]["flags"][:sample] == 1
    mism = {"plain_fp32_moller": int((plain_hit != truth).sum()),
            "filtered_fp32_fp64_exact": int((batched_hit != truth).sum()),
            "by_kind_plain": [int((plain_hit != truth)[kind[:sample] == k].sum()) for k in range(3)],
            "by_kind_filtered": [int((batched_hit != truth)[kind[:sample] == k].sum()) for k in range(3)],
            "pairs_checked": sample}
    ms = res["batched"]["kernel_ms"]
    bytes_per_launch = 76.0 * n
    achieved = bytes_per_launch / (ms * 1e-3) / 1e9
    return _line("tri-tri tests/s, precision stress (filtered exact predicates)", n / (ms * 1e-3), "tests/s", ms, args,
                 {"workload": f"{n:,} triangle pairs in [-1e3, 1e3]^3: 80% random, 10% near-coplanar (1e-7 rel), "
                              "10% touching (1 ulp)", "parallelism": "single GPU"},
                 dtype="f32 (f64 predicate fallback)",
                 extra={"plain_tests_per_s": round(n / (res["plain"]["kernel_ms"] * 1e-3), 1),
                        "fp64_fallbacks": res["batched"]["fp64"], "exact_fallbacks": res["batched"]["exact"],
                        "mismatches_vs_exact_oracle": mism,
                        "roofline": {"bound": "hbm", "kernel": "k_tritri_f32 + k_tritri_rerun<1> + k_tritri_rerun<2>",
                                     "achieved": round(achieved, 2), "peak": 8000.0, "unit": "GB/s",
                                     "frac": round(achieved / 8000.0, 5),
                                     "traffic": None, "algorithmic_bytes_per_launch": int(bytes_per_launch)},
                        "cpu_baseline": {"value": round(sample / t_cpu, 1), "unit": "tests/s", "cores": 1,
                                         "kind": "port", "sample": f"{sample:,} pairs, reference Möller, single thread"},
                        "exact_oracle_tests_per_s": round(sample / t_ex, 1)})


def config5(args):
    from inmyroom_vulkan_amd import MODE_BATCHED, QUERY_DTYPE, Collider, scenes

    db = scenes.DeformableBatch(n_meshes=64, seed=5)
    gpu = Collider(leaf_max=4)
    t0 = time.time()
    ids = np.array([gpu.register_skinned(db.P, db.F, db.joints, db.weights, db.n_bones) for _ in range(64)], np.uint32)
    terrain = gpu.register_mesh(db.tP.astype(np.float32), db.tF)
    log(f"config 5: 64 skinned meshes + terrain registered in {time.time() - t0:.1f}s")
    q = np.zeros(64, QUERY_DTYPE)
    q["mesh_a"], q["mesh_b"], q["rel"] = ids, terrain, scenes.identity16()
    frames = [db.bone_matrices() for _ in range(args.warmup + args.steps)]
    tuning = {"use_graph": 0}
    for kv in filter(None, getattr(args, "tuning", "").split(",")):
        k, v = kv.split("=")
        tuning[k.strip()] = int(v)
    gpu.set_tuning(**tuning)
    refit_ms, query_ms, wall, ncon = [], [], [], 0
    for k, bones in enumerate(frames):
        t = time.perf_counter()
        gpu.skin_refit(ids, bones)
        r = gpu.timings()["total_ms"]
        gpu.stage(q)
        n = gpu.run_staged(MODE_BATCHED)
        w = (time.perf_counter() - t) * 1e3
        if k >= args.warmup:
            refit_ms.append(r)
            query_ms.append(gpu.timings()["total_ms"])
            wall.append(w)
            ncon += n
    ms = float(np.mean(wall))
    log(f"config 5: {ms:.3f} ms/frame wall; refit {np.mean(refit_ms):.3f} ms, queries {np.mean(query_ms):.3f} ms")
    cpu = None
    if not args.no_cpu:
        # the reference has no refit: a deformable mesh's tree would be rebuilt from
        # its skinned triangles; per mesh: CPU skinning, the oracle's tree build and
        # its exact query against the (prebuilt) terrain tree
        from oracle import Oracle

        orc = Oracle()
        terrain_tree = orc.tree(db.tP.astype(np.float32)[db.tF].reshape(-1, 9), 4)
        bones = frames[-1].reshape(64, -1)
        rel = scenes.identity16()[None]
        t = time.perf_counter()
        done = 0
        for k in range(64):
            Pk = scenes.DeformableBatch.skin_reference(db.P, db.joints, db.weights, bones[k]).astype(np.float32)
            tree = orc.tree(Pk[db.F].reshape(-1, 9), 4)
            orc.batch([tree, terrain_tree], np.zeros(1, np.uint32), np.ones(1, np.uint32), rel, mode=1)
            done += 1
            if time.perf_counter() - t > args.cpu_seconds:
                break
        per_frame = (time.perf_counter() - t) / done * 64
        cpu = {"value": round(1.0 / per_frame, 4), "unit": "frames/s", "cores": 1, "kind": "port",
               "sample": f"{done} of the 64 meshes of one frame (CPU skinning, tree rebuild, exact query against "
                         "the terrain), extrapolated to 64, single thread"}
    return _line("deformable batch frames/s (skin + refit + collide vs terrain)", 1e3 / ms, "frames/s", ms, args,
                 {"workload": "64 skinned tori x 20,000 triangles (64 bones, 4 weights/vertex) vs 198,450-triangle "
                              "fBm terrain", "parallelism": "single GPU"},
                 dtype="f32 (f64 predicate fallback)",
                 extra={"refit_ms": round(float(np.mean(refit_ms)), 4), "query_ms": round(float(np.mean(query_ms)), 4),
                        "contacts_per_frame": ncon / max(1, len(wall)), "cpu_baseline": cpu})


def config6(args):
    import os

    from inmyroom_vulkan_amd import Collider, GltfDocument

    here = os.path.dirname(os.path.abspath(__file__))
    paths = [os.path.join(here, "tests", "golden", "gltf", p)
             for p in ("apple/scene.gltf", "table_room.gltf", "environment.gltf")]
    gpu = Collider(leaf_max=4)
    parse_ms, reg_ms, build_ms, ntri, tri_levels = [], [], [], 0, 0
    for k in range(args.warmup + args.steps):
        pm = rm = bm = 0.0
        n = lv = 0
        for p in paths:
            t = time.perf_counter()
            doc = GltfDocument(p)
            t1 = time.perf_counter()
            ids = gpu.register_gltf(doc)
            t2 = time.perf_counter()
            bm += gpu.timings()["traverse_ms"]
            n += sum(doc.mesh_info(m)["ntri"] for m in range(len(doc)))
            for i in ids:
                if i != 0xFFFFFFFF:
                    info = gpu.mesh_info(int(i))
                    lv += info["ntri"] * (info["depth"] + 1)  # levels a triangle passes, at most
            doc.close()
            pm += (t1 - t) * 1e3
            rm += (t2 - t1) * 1e3
        if k >= args.warmup:
            parse_ms.append(pm)
            reg_ms.append(rm)
            build_ms.append(bm)
            ntri, tri_levels = n, lv
    ms = float(np.median(np.array(parse_ms) + np.array(reg_ms)))
    log(f"config 6: {ms:.3f} ms per load of {ntri:,} triangles (parse {np.median(parse_ms):.3f}, "
        f"assemble + tree build {np.median(reg_ms):.3f})")
    cpu = None
    if not args.no_cpu:
        from oracle import Oracle
        from oracle import gltf as og

        orc = Oracle()
        t = time.perf_counter()
        n = loads = 0
        while True:
            for p in paths:
                for prims in og.load(p):
                    soup = og.mesh_triangles(prims)[0]
                    if len(soup):
                        orc.tree(soup, 4)
                        n += len(soup)
            loads += 1
            if time.perf_counter() - t > args.cpu_seconds:
                break
        el = time.perf_counter() - t
        cpu = {"value": round(n / el, 1), "unit": "triangles/s", "cores": 1, "kind": "port",
               "sample": f"{loads} loads of the three scenes: numpy decode and the oracle's tree build, single thread"}
    # the build moves every triangle once per level it passes: 40 B read (positions, id), 40 B written
    bms = float(np.median(build_ms))
    build_bytes = 80.0 * tri_levels
    achieved = build_bytes / (bms * 1e-3) / 1e9
    return _line("glTF ingest triangles/s (parse, assemble, OBB-tree build)", ntri / (ms * 1e-3), "triangles/s", ms,
                 args, {"workload": "Apple + Table_room + Sponza environmentTest glTF scenes: 50 meshes, "
                                    f"{ntri:,} collision triangles", "parallelism": "single GPU"},
                 extra={"parse_ms": round(float(np.median(parse_ms)), 4),
                        "assemble_build_ms": round(float(np.median(reg_ms)), 4),
                        "build_device_ms": round(bms, 4),
                        "roofline": {"bound": "hbm", "kernel": "tree build (k_level and its level passes)",
                                     "achieved": round(achieved, 3), "peak": 8000.0, "unit": "GB/s",
                                     "frac": round(achieved / 8000.0, 6), "traffic": None,
                                     "algorithmic_bytes_per_load": int(build_bytes),
                                     "note": "latency bound: the PCA sums are sequential double chains"},
                        "cpu_baseline": cpu},
                 data="the reference's shipped glTF scenes (tests/golden/gltf)")


def run(args):
    if not hasattr(args, "pairs"):
        args.pairs = 1 << 24
    if not hasattr(args, "oracle_pairs"):
        args.oracle_pairs = 1 << 18
    return {2: config2, 4: config4, 5: config5, 6: config6}[args.config](args)
