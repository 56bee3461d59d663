"""Timing of the other workloads with bench.py's harness (`bench.py --config N`).

  1  plain method: two 1,280-triangle unit icospheres, 1,000 seeded poses, at
     leaf sizes 4 and 1; plain and batched paths, every pose checked
  2  scene step: 512 rigid entities (boxes, icospheres, noisy tori) in a 50 m
     cube, 100 frames at 1/60 s, broad + mid + narrow each frame (cd_frame);
     also timed with the uncollide rays and the response (the whole frame)
  4  precision stress: 2^24 triangle pairs (80% random, 10% near-coplanar,
     10% touching); plain float32 Möller vs the filtered exact test, every
     pair checked against the exact oracle
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


def host_env() -> dict:
    """The host the CPU baseline ran on: CPU model, cores, compiler, ROCm."""
    import os
    import subprocess

    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        gxx = subprocess.run(["g++", "--version"], capture_output=True, text=True, check=True).stdout.splitlines()[0]
    except (OSError, subprocess.CalledProcessError, IndexError):
        gxx = "unknown"
    try:
        with open("/opt/rocm/.info/version") as f:
            rocm = f.read().strip()
    except OSError:
        rocm = "unknown"
    return {"cpu_model": model, "nproc": os.cpu_count(), "cpus_usable": len(os.sched_getaffinity(0)),
            "compiler": gxx, "rocm": rocm}


class CpuOracles:
    """The CPU path in two builds: the reference's release flags for this host
    (-O3 -ffast-math -march=native, inMyRoom_vulkan/CMakeLists.txt:30), compiled
    here, which is the headline baseline; and the strict parity build."""

    def __init__(self):
        import os
        import tempfile

        from oracle import NATIVE_FLAGS, STRICT_FLAGS, Oracle, build_native

        t = time.time()
        path = build_native(os.path.join(tempfile.mkdtemp(prefix="oracle_native_"), "liboracle_native.so"))
        log(f"CPU baseline build ({NATIVE_FLAGS}) in {time.time() - t:.1f}s")
        self.native = Oracle(path=path)
        self.strict = Oracle()
        self.flags = {"native": NATIVE_FLAGS, "strict": STRICT_FLAGS}

    def trees(self, which: str, soups, leaf: int = 4, threads: int = 16):
        from concurrent.futures import ThreadPoolExecutor

        orc = getattr(self, which)
        with ThreadPoolExecutor(threads) as ex:  # the oracle's build releases the GIL
            return list(ex.map(lambda s: orc.tree(s, leaf), soups))


def timed_median(fn, warmup: int = 1, reps: int = 5) -> float:
    """Seconds: one untimed warm-up call, then the median of `reps` calls."""
    for _ in range(warmup):
        fn()
    ts = []
    for _ in range(reps):
        t = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t)
    return float(np.median(ts))


def cpu_line(value, unit, sample, native_s, strict_s, orcs, extra=None):
    """The cpu_baseline object: the release build's rate, the strict build's beside it."""
    d = {"value": round(value, 2), "unit": unit, "cores": 1, "kind": "port", "sample": sample,
         "build": orcs.flags["native"] + " (the reference's release flags) at compile time, on this host; linked "
                  "without the fast-math startup file, so the process keeps IEEE denormals",
         "timing": "1 warm-up, median of 5, single thread",
         "seconds_per_sample": round(native_s, 4),
         "strict": {"seconds_per_sample": round(strict_s, 4), "build": orcs.flags["strict"]},
         "host": host_env()}
    d.update(extra or {})
    return d


def _line(metric, value, unit, ms, args, config, dtype="f32", extra=None, data="synthetic"):
    d = {"metric": metric, "value": round(value, 2), "unit": unit, "n_gpus": 1, "steps": args.steps,
         "warmup": args.warmup, "ms_per_step": round(ms, 4), "higher_is_better": True, "scaling": "weak",
         "vs_baseline": None, "dtype": dtype, "data": data, "config": config}
    d.update(extra or {})
    return d


def _keyset(c):
    return set(zip(c["query"].tolist(), c["tri_a"].tolist(), c["tri_b"].tolist()))


def config1(args):
    """Two 1,280-triangle unit icospheres, 1,000 seeded poses, leaf sizes 4
    (the reference's) and 1: plain (reference semantics) and batched paths,
    the reference path on the CPU, mismatches over every pose."""
    from inmyroom_vulkan_amd import MODE_BATCHED, MODE_PLAIN, Collider, scenes

    mesh, rel = scenes.config1(1000, seed=1)
    nq = len(rel)
    z = np.zeros(nq, np.uint32)
    orcs = None if args.no_cpu else CpuOracles()
    per_leaf = {}
    for L in (4, 1):
        gpu = Collider(leaf_max=L)
        mid = gpu.register_soups([mesh])[0]
        q = gpu.make_queries(mid, mid, rel)
        gpu.stage(q)
        r = {}
        for name, mode, graph in (("plain", MODE_PLAIN, 0), ("batched", MODE_BATCHED, 1)):
            gpu.set_tuning(use_graph=graph)
            for _ in range(max(1, args.warmup)):
                gpu.run_staged(mode)
            wall = []
            for _ in range(args.steps):
                t = time.perf_counter()
                gpu.run_staged(mode)
                wall.append((time.perf_counter() - t) * 1e3)
            ms = float(np.median(wall))
            r[name] = {"ms_per_batch": round(ms, 4), "queries_per_s": round(nq / (ms * 1e-3), 1),
                       "counters": gpu.counters(), "contacts": gpu.fetch()}
            log(f"config 1 L={L} {name}: {ms:.4f} ms per batch of {nq}")
        # reference counts (the plain path reproduces them) and eager kernel times of the batched path
        lp = gpu.query_leafpairs(q)
        gpu.set_tuning(use_graph=0)
        gpu.stage(q)
        kt = []
        for _ in range(5):
            gpu.run_staged(MODE_BATCHED)
            kt.append(gpu.timings())
        ref = r["plain"]["counters"]
        trav_b = 64.0 * nq + 96.0 * ref["sat_tests"]
        narrow_b = 36.0 * float(((lp[:, 3] & 0xFFFF) + (lp[:, 3] >> 16)).astype(np.float64).sum()) + 32.0 * ref["contacts"]
        t_trav = float(np.median([k["traverse_ms"] for k in kt]))
        t_nar = float(np.median([k["narrow_ms"] for k in kt]))
        dom, dms, dbytes = ("traversal (k_split_level + k_dfs)", t_trav, trav_b) if t_trav >= t_nar else \
            ("narrow (k_narrow_exact + reruns + k_contacts)", t_nar, narrow_b)
        ach = dbytes / (dms * 1e-3) / 1e9
        out = {"plain": {k: r["plain"][k] for k in ("ms_per_batch", "queries_per_s")},
               "batched": {k: r["batched"][k] for k in ("ms_per_batch", "queries_per_s")},
               "speedup_batched_vs_plain": round(r["batched"]["queries_per_s"] / r["plain"]["queries_per_s"], 2),
               "reference_counts": {k: int(ref[k]) for k in ("sat_tests", "leaf_pairs", "tri_tests", "contacts")},
               "tritests_per_s": round(ref["tri_tests"] / (r["batched"]["ms_per_batch"] * 1e-3), 1),
               "kernel_ms": {"traversal": round(t_trav, 4), "narrow": round(t_nar, 4)},
               "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(ach, 2), "peak": 8000.0, "unit": "GB/s",
                            "frac": round(ach / 8000.0, 5), "traffic": None,
                            "algorithmic_bytes_per_launch": int(dbytes), "launch_ms": round(dms, 4)}}
        if orcs is not None:
            strict_tree = orcs.strict.tree(mesh, L)
            native_tree = orcs.native.tree(mesh, L)
            runs = {}

            def run(which, tree, mode=0):
                runs[which] = getattr(orcs, which).batch([tree], z, z, rel, mode=mode, nthreads=1)

            native_s = timed_median(lambda: run("native", native_tree))
            strict_s = timed_median(lambda: run("strict", strict_tree))
            tt = runs["strict"].total_tritests
            out["cpu_baseline"] = cpu_line(
                nq / native_s, "queries/s",
                f"all {nq} poses, reference semantics, leaf size {L}; tri-tri tests/s {tt / native_s:.0f}",
                native_s, strict_s, orcs, extra={"strict_value": round(nq / strict_s, 2)})
            ex = orcs.strict.batch([strict_tree], z, z, rel, mode=1, nthreads=16)
            out["mismatches"] = {
                "poses_checked": nq,
                "plain_vs_reference_semantics": len(_keyset(r["plain"]["contacts"]) ^ _keyset(runs["strict"].contacts)),
                "batched_vs_exact_oracle": len(_keyset(r["batched"]["contacts"]) ^ _keyset(ex.contacts)),
                "plain_vs_exact_oracle": len(_keyset(r["plain"]["contacts"]) ^ _keyset(ex.contacts))}
        per_leaf[f"L{L}"] = out
        gpu.close()
    head = per_leaf["L4"]
    ms = head["batched"]["ms_per_batch"]
    extra = {"leaf_sizes": per_leaf, "plain_queries_per_s": head["plain"]["queries_per_s"],
             "speedup_batched_vs_plain": head["speedup_batched_vs_plain"], "tritests_per_s": head["tritests_per_s"],
             "roofline": head["roofline"]}
    if "cpu_baseline" in head:
        extra["cpu_baseline"] = head["cpu_baseline"]
    return _line("mesh-pair queries/sec + tri-tri tests/sec, batched vs plain, 1 MI355X", nq / (ms * 1e-3),
                 "queries/s", ms, args,
                 {"workload": "two 1,280-triangle unit icospheres, 1,000 seeded poses (translation U[-2.5, 2.5]^3, "
                              "uniform rotation); leaf sizes 4 (headline) and 1", "queries": nq,
                  "parallelism": "single GPU"}, dtype="f32 (f64 predicate fallback)", extra=extra)


def config2(args):
    from inmyroom_vulkan_amd import ENTITY_DTYPE, MODE_BATCHED, MODE_PLAIN, MODE_REFERENCE, Collider, scenes

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
    modes = (("plain", MODE_PLAIN),) * (args.plain_steps > 0) + (("reference", MODE_REFERENCE),
                                                                 ("batched", MODE_BATCHED))
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
        shoot_ms = boxes = tritests = 0.0
        for f in frames:
            gpu.frame(f, mode)
            _, rays = gpu.uncollide_rays()
            resp = gpu.shoot_rays()
            nrays += len(rays)
            nshot += int(resp["shot"].sum())
            shoot_ms += gpu.timings()["traverse_ms"]
            c = gpu.counters()
            boxes += c["sat_tests"]
            tritests += c["tri_tests"]
        full = (time.perf_counter() - t) * 1e3 / len(frames)
        # k_shoot's roofline: 128 B per box tested (its child record and its
        # slab record), 36 B per ray-triangle test, 24 B per ray read and 56 B
        # per ray result written
        shoot_bytes = 128.0 * boxes + 36.0 * tritests + (24.0 + 56.0) * nrays
        shoot_gbs = shoot_bytes / (shoot_ms * 1e-3) / 1e9
        out[name].update({"full_ms_per_frame": round(full, 4), "rays_per_frame": nrays / len(frames),
                          "pairs_shot_per_frame": nshot / len(frames),
                          "shoot": {"kernel_ms_per_frame": round(shoot_ms / len(frames), 4),
                                    "boxes_per_frame": boxes / len(frames),
                                    "ray_triangle_tests_per_frame": tritests / len(frames),
                                    "roofline": {"bound": "hbm", "kernel": "k_shoot", "achieved": round(shoot_gbs, 2),
                                                 "peak": 8000.0, "unit": "GB/s",
                                                 "frac": round(shoot_gbs / 8000.0, 5), "traffic": None,
                                                 "algorithmic_bytes_per_frame": int(shoot_bytes / len(frames))}}})
        log(f"config 2 {name}: {full:.3f} ms/frame with rays and response")
        # the same step as one call (cd_step): contacts and rays stay on the device
        for f in frames[: args.warmup]:
            gpu.step(f, mode)
        t = time.perf_counter()
        for f in frames:
            gpu.step(f, mode)
        one = (time.perf_counter() - t) * 1e3 / len(frames)
        out[name]["step_ms_per_frame"] = round(one, 4)
        log(f"config 2 {name}: {one:.3f} ms/frame as one cd_step")
    # every frame's pairs and contacts in both modes (for the mismatch counts),
    # and the batched traversal's device time from eager runs (events around
    # the kernels) against the reference's node-pair tests from the plain frames
    keys = {"plain": [], "reference": [], "batched": []}
    trav_bytes = trav_ms = 0.0
    for name, mode in (("plain", MODE_PLAIN), ("reference", MODE_REFERENCE), ("batched", MODE_BATCHED)):
        gpu.set_tuning(**dict(tuning, use_graph=0, step_graph=2))  # launched one by one: kernel events
        for f in frames:
            p, c = gpu.frame(f, mode)
            k = set()
            for pr in p:
                cc = c[pr["contact_begin"]: pr["contact_begin"] + pr["contact_count"]]
                k.update(zip([int(pr["first"])] * len(cc), [int(pr["second"])] * len(cc), cc["tri_a"].tolist(),
                             cc["tri_b"].tolist()))
            keys[name].append(k)
            if name == "plain":
                cnt = gpu.counters()
                trav_bytes += 64.0 * cnt["broad_pairs"] + 96.0 * cnt["sat_tests"]
            elif name == "batched":
                trav_ms += gpu.timings()["traverse_ms"]
    gpu.set_tuning(**tuning)
    ach = trav_bytes / (trav_ms * 1e-3) / 1e9
    roofline = {"bound": "hbm", "kernel": "traversal (k_split_level + k_dfs), summed over the 100 frames",
                "achieved": round(ach, 2), "peak": 8000.0, "unit": "GB/s", "frac": round(ach / 8000.0, 5),
                "traffic": None, "algorithmic_bytes_per_frame": int(trav_bytes / len(frames)),
                "launch_ms_per_frame": round(trav_ms / len(frames), 4)}
    cpu = mism = None
    if not args.no_cpu:
        from concurrent.futures import ThreadPoolExecutor

        orcs = CpuOracles()
        strict_trees = [orcs.strict.tree(s, 4) for s in sc.soups]
        native_trees = [orcs.native.tree(s, 4) for s in sc.soups]
        sample = frames[:5]

        def run(orc, trees):
            for f in sample:
                orc.frame(trees, sc.mesh_of, f["current"], f["should_callback"], mode=0)

        native_s = timed_median(lambda: run(orcs.native, native_trees))
        strict_s = timed_median(lambda: run(orcs.strict, strict_trees))
        cpu = cpu_line(len(sample) / native_s, "frames/s",
                       f"the first {len(sample)} of the 100 frames (broad + mid + narrow, reference semantics)",
                       native_s, strict_s, orcs, extra={"strict_value": round(len(sample) / strict_s, 3)})

        def oracle_keys(f, mode):
            res = orcs.strict.frame(strict_trees, sc.mesh_of, f["current"], f["should_callback"], mode=mode)
            pr = res.pairs[res.contacts["query"]] if len(res.contacts) else np.zeros((0, 2), np.uint32)
            return set(zip(pr[:, 0].tolist(), pr[:, 1].tolist(), res.contacts["tri_a"].tolist(),
                           res.contacts["tri_b"].tolist()))

        t = time.time()
        with ThreadPoolExecutor(16) as ex:  # the oracle releases the GIL
            ref = list(ex.map(lambda f: oracle_keys(f, 0), frames))
            exact = list(ex.map(lambda f: oracle_keys(f, 1), frames))
        log(f"config 2: oracle over the 100 frames, both semantics: {time.time() - t:.1f}s")
        mism = {"frames_checked": len(frames),
                "plain_vs_reference_semantics": sum(len(a ^ b) for a, b in zip(keys["plain"], ref)),
                "reference_mode_vs_reference_semantics": sum(len(a ^ b) for a, b in zip(keys["reference"], ref)),
                "batched_vs_exact_oracle": sum(len(a ^ b) for a, b in zip(keys["batched"], exact)),
                "plain_vs_exact_oracle": sum(len(a ^ b) for a, b in zip(keys["plain"], exact))}
    ms = out["batched"]["ms_per_frame"]
    return _line("scene step frames/s, 512 entities, broad + narrow", 1e3 / ms, "frames/s", ms, args,
                 {"workload": "512 rigid entities (box 12 / icosphere 1,280 / torus 8,192 triangles), 50 m cube, "
                              "100 frames at 1/60 s", "parallelism": "single GPU"},
                 extra={"modes": out, "roofline": roofline, "mismatches": mism, "cpu_baseline": cpu})


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
    mism = cpu = None
    if not args.no_cpu:
        orcs = CpuOracles()
        # every pair against the exact oracle, and the plain path against the
        # reference's Möller (strict build) on every pair
        t = time.perf_counter()
        ex = orcs.strict.tritri_exact(a9, b9)
        t_ex = time.perf_counter() - t
        fo, _ = orcs.strict.tritri(a9, b9)
        truth = ex == 1
        plain_hit = ((res["plain"]["flags"] & 1) != 0) & ((res["plain"]["flags"] & 2) == 0)
        batched_hit = res["batched"]["flags"] == 1
        mism = {"pairs_checked": n,
                "plain_flags_vs_reference_moller": int((res["plain"]["flags"] != fo).sum()),
                "plain_fp32_moller_vs_exact": int((plain_hit != truth).sum()),
                "filtered_fp32_fp64_exact_vs_exact": int((batched_hit != truth).sum()),
                "by_kind_plain": [int((plain_hit != truth)[kind == k].sum()) for k in range(3)],
                "by_kind_filtered": [int((batched_hit != truth)[kind == k].sum()) for k in range(3)],
                "exact_oracle_tests_per_s": round(n / t_ex, 1)}
        sample = min(n, 1 << 22)
        native_s = timed_median(lambda: orcs.native.tritri(a9[:sample], b9[:sample]))
        strict_s = timed_median(lambda: orcs.strict.tritri(a9[:sample], b9[:sample]))
        cpu = cpu_line(sample / native_s, "tests/s", f"the first {sample:,} pairs, reference Möller",
                       native_s, strict_s, orcs, extra={"strict_value": round(sample / strict_s, 1)})
    ms = res["batched"]["kernel_ms"]
    bytes_per_launch = 76.0 * n
    achieved = bytes_per_launch / (ms * 1e-3) / 1e9
    return _line("tri-tri tests/s, precision stress (filtered exact predicates)", n / (ms * 1e-3), "tests/s", ms, args,
                 {"workload": f"{n:,} triangle pairs in [-1e3, 1e3]^3: 80% random, 10% near-coplanar (1e-7 rel), "
                              "10% touching (1 ulp)", "parallelism": "single GPU"},
                 dtype="f32 (f64 predicate fallback)",
                 extra={"plain_tests_per_s": round(n / (res["plain"]["kernel_ms"] * 1e-3), 1),
                        "fp64_fallbacks": res["batched"]["fp64"], "exact_fallbacks": res["batched"]["exact"],
                        "mismatches": mism,
                        "roofline": {"bound": "hbm", "kernel": "k_tritri_f32 + k_tritri_rerun<1> + k_tritri_rerun<2>",
                                     "achieved": round(achieved, 2), "peak": 8000.0, "unit": "GB/s",
                                     "frac": round(achieved / 8000.0, 5),
                                     "traffic": None, "algorithmic_bytes_per_launch": int(bytes_per_launch)},
                        "cpu_baseline": cpu})


def config5(args):
    from inmyroom_vulkan_amd import MODE_BATCHED, QUERY_DTYPE, Collider, scenes

    db = scenes.DeformableBatch(n_meshes=64, seed=5)
    gpu = Collider(leaf_max=4)
    t0 = time.time()
    ids = np.array([gpu.register_skinned(db.P, db.F, db.joints, db.weights, db.n_bones) for _ in range(64)], np.uint32)
    terrain_id = gpu.register_mesh(db.tP.astype(np.float32), db.tF)
    log(f"config 5: 64 skinned meshes + terrain registered in {time.time() - t0:.1f}s")
    q = np.zeros(64, QUERY_DTYPE)
    q["mesh_a"], q["mesh_b"], q["rel"] = ids, terrain_id, scenes.identity16()
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
    # roofline of the dominant step, skinning + refit: per frame 48 B per vertex
    # skinned (bind position, joints, weights in; position out), 48 B per
    # triangle gathered (indices in, corners out), and three passes (mean,
    # covariance, extents) over the 36 B of every triangle of every box
    links = gpu.export_tree(int(ids[0]))["links"]
    sub = {}

    def tris_of(node):  # triangles under internal node `node`, and the sum over its child boxes
        if node in sub:
            return sub[node]
        n = box_sum = 0
        for side in (0, 1):
            cnt = int(links[node, 2 + side])
            c, b = (cnt, 0) if cnt else tris_of(int(links[node, side]))
            n += c
            box_sum += c + b
        sub[node] = (n, box_sum)
        return sub[node]

    ntri = len(db.F)
    box_tris = ntri + (tris_of(0)[1] if len(links) else 0)  # the root box, then every child box
    refit_bytes = 64 * (48.0 * len(db.P) + 48.0 * ntri + 3 * 36.0 * box_tris)
    rms = float(np.mean(refit_ms))
    ach = refit_bytes / (rms * 1e-3) / 1e9
    roofline = {"bound": "hbm", "kernel": "skinning + refit (k_skin, k_gather_tris, k_rf_small, k_rf_chunk, k_rf_big)",
                "achieved": round(ach, 2), "peak": 8000.0, "unit": "GB/s", "frac": round(ach / 8000.0, 5),
                "traffic": None, "algorithmic_bytes_per_launch": int(refit_bytes), "launch_ms": round(rms, 4),
                "box_triangles_per_mesh": int(box_tris)}
    cpu = mism = None
    if not args.no_cpu:
        # the reference has no refit: a deformable mesh's tree would be rebuilt from
        # its skinned triangles; per mesh: CPU skinning, the oracle's tree build and
        # its exact query against the (prebuilt) terrain tree
        from concurrent.futures import ThreadPoolExecutor

        orcs = CpuOracles()
        tsoup = db.tP.astype(np.float32)[db.tF].reshape(-1, 9)
        bones = frames[-1].reshape(64, -1)
        rel = scenes.identity16()[None]
        sample = 4
        terrain = {w: getattr(orcs, w).tree(tsoup, 4) for w in ("native", "strict")}

        def rebuild(which):
            orc = getattr(orcs, which)
            for k in range(sample):
                Pk = scenes.DeformableBatch.skin_reference(db.P, db.joints, db.weights, bones[k]).astype(np.float32)
                tree = orc.tree(Pk[db.F].reshape(-1, 9), 4)
                orc.batch([tree, terrain[which]], np.zeros(1, np.uint32), np.ones(1, np.uint32), rel, mode=1)

        native_s = timed_median(lambda: rebuild("native"))
        strict_s = timed_median(lambda: rebuild("strict"))
        cpu = cpu_line(sample / (native_s * 64), "frames/s",
                       f"{sample} of the 64 meshes of one frame (CPU skinning, tree rebuild, exact query against the "
                       "terrain), extrapolated to 64", native_s, strict_s, orcs,
                       extra={"strict_value": round(sample / (strict_s * 64), 4)})
        # the last frame's pair set, all 64 meshes, against the exact oracle on the GPU's skinned triangles
        got = gpu.fetch()
        soups, gperm = [], []
        for m in ids:
            tri = gpu.export_triangles(int(m))
            perm = gpu.export_tree(int(m))["perm"]
            inp = np.empty_like(tri)
            inp[perm] = tri
            soups.append(inp)
            gperm.append(perm)
        tperm = gpu.export_tree(int(terrain_id))["perm"]
        with ThreadPoolExecutor(16) as ex:
            trees = list(ex.map(lambda s_: orcs.strict.tree(s_, 4), soups + [tsoup]))
        res = orcs.strict.batch(trees, np.arange(64, dtype=np.uint32), np.full(64, 64, np.uint32),
                                np.tile(scenes.identity16(), (64, 1)), mode=1, nthreads=16)
        operm = [t_.export()["perm"] for t_ in trees]
        mine = set(zip(got["query"].tolist(), [int(gperm[q_][a]) for q_, a in zip(got["query"], got["tri_a"])],
                       tperm[got["tri_b"]].tolist()))
        want = set(zip(res.contacts["query"].tolist(),
                       [int(operm[q_][a]) for q_, a in zip(res.contacts["query"], res.contacts["tri_a"])],
                       operm[64][res.contacts["tri_b"]].tolist()))
        mism = {"meshes_checked": 64, "contacts": len(want), "batched_vs_exact_oracle": len(mine ^ want)}
    return _line("deformable batch frames/s (skin + refit + collide vs terrain)", 1e3 / ms, "frames/s", ms, args,
                 {"workload": "64 skinned tori x 20,000 triangles (64 bones, 4 weights/vertex) vs 198,450-triangle "
                              "fBm terrain", "parallelism": "single GPU"},
                 dtype="f32 (f64 predicate fallback)",
                 extra={"refit_ms": round(rms, 4), "query_ms": round(float(np.mean(query_ms)), 4),
                        "contacts_per_frame": ncon / max(1, len(wall)), "roofline": roofline, "mismatches": mism,
                        "cpu_baseline": cpu})


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
        from oracle import gltf as og

        orcs = CpuOracles()
        n = sum(len(og.mesh_triangles(prims)[0]) for p in paths for prims in og.load(p))

        def load(which):
            orc = getattr(orcs, which)
            for p in paths:
                for prims in og.load(p):
                    soup = og.mesh_triangles(prims)[0]
                    if len(soup):
                        orc.tree(soup, 4)

        native_s = timed_median(lambda: load("native"))
        strict_s = timed_median(lambda: load("strict"))
        cpu = cpu_line(n / native_s, "triangles/s", "one load of the three scenes: numpy decode and the tree build",
                       native_s, strict_s, orcs, extra={"strict_value": round(n / strict_s, 1)})
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
    return {1: config1, 2: config2, 4: config4, 5: config5, 6: config6}[args.config](args)
