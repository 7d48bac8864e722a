"""Benchmark of the gfx950 step path: body-steps/s on one MI355X.

Headline workload (N=1): 4,096 independent worlds stepped together (the
batched many-world form): each world a random pile of N ~ logU[8,256] bodies
(boxes, spheres and 16-point convex hulls in equal shares) dropped onto its
own ground, dt 1/60, 8 solver iterations, FP32, warm starting on.  One
bepu_step advances every world; each world's result is bit-identical to the
same world stepped alone.  Inputs are resident on the device before the timed
region; the timed region is K full steps (force, AABB, broadphase,
narrowphase, general convex pair tests, islands and colouring, prestep, warm
start + iterations, integration) bracketed by device synchronisation.

The JSON line also carries
  roofline      the dominant kernel's algorithmic bytes per launch over its
                average duration (HIP events on the world's stream), against
                the MI355X HBM peak of 8 TB/s, with the measured HBM traffic
                of that kernel (rocprofv3 counter passes, profiles/);
  latency_ms    p50 / p99 of single synchronised batched steps;
  plain_path    the same worlds stepped one at a time, each in its own world
                handle (a sample), and the batched/plain ratio;
  cpu_baseline  the oracle (this repository's C++ restatement of the
                reference step) on a sample of the same worlds, one world per
                host thread on up to 16 threads;
  parity        batched worlds against the oracle, each world alone: motion
                states bit for bit and the largest relative impulse error;
  towers        the plain single-world path on 64 x 64 towers of 20 boxes
                (81,920 bodies, 10 iterations) with its own roofline and
                oracle baseline;
  served, stress
                continuous batching of rollouts, and one step of the
                broadphase/narrowphase stress scene;
  mesh          16,384 bodies of all four shapes on a terrain and on a
                static triangle mesh: body-steps/s, the mesh pair kernel's
                time and bytes, the oracle's rate and bit parity.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--legs batched,towers,served,stress,mesh]
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
BATCHED_ITERATIONS = 8

# Algorithmic bytes per unit of work, per stage (FP32, each element counted
# once per stage and once per solver iteration, no cache-reuse credit):
#   body_pre      force/inertia 160 + AABB 80 per body
#   broadphase    Morton 32 + radix sort 64 + tree build 88 + traversal leaf read 24 per body
#   narrowphase   312 per broadphase pair
#   convex        1,120 per general convex pair (a pair with a hull or a
#                 capsule): pair state 336 and tester/supplement state 160,
#                 each read and written, + pose, velocity and shape of both
#                 bodies 128 (the hull vertex pool, a few KB that stays in
#                 cache, is not counted)
#   islands       union-find and island bucketing: 40 per body (parent init 4,
#                 compress 12, count 4, body list 8, local index 4, tier list 8)
#                 + 32 per manifold (union reads 16, count 8, constraint list 8)
#   island_solve  prestep 332 + 72 c, warm start 190 + 70 c and per iteration
#                 316 + 44 c, per manifold of an island solved in one workgroup
#   global_solve  the same per manifold of a large island, plus 64 for its
#                 colouring (pair read 24, incidence 16, colour/priority 16, order 8)
#   integrate     80 per body
# c = contacts of the manifold (the scene's mean is used per tier).
B_CONVEX_PAIR = 1120
#   mesh          per static mesh or terrain pair: pair state 336 and tester /
#                 supplement state 160, each read and written, + both bodies'
#                 pose, velocity and shape 128 (1,120), + 64 per live
#                 triangle tester (32 B read and written); the triangles'
#                 vertices, indices and tree nodes (a few MB that stay in
#                 cache) are not counted
B_MESH_PAIR = 1120
B_MESH_TESTER = 64


def manifold_bytes(m, contacts, iters):
    return (332 + 190) * m + (72 + 70) * contacts + iters * (316 * m + 44 * contacts)


def algorithmic_bytes(stage, n_body, n_pair, n_cons, n_contacts, iters, n_global=0, n_convex=0, n_mesh=0,
                      n_testers=0):
    c_mean = n_contacts / n_cons if n_cons else 0.0
    m_isl = n_cons - n_global
    if stage == "body_pre":
        return 240 * n_body
    if stage == "broadphase":
        return 208 * n_body
    if stage == "narrowphase":
        return 312 * n_pair
    if stage == "convex":
        return B_CONVEX_PAIR * n_convex
    if stage == "mesh":
        return B_MESH_PAIR * n_mesh + B_MESH_TESTER * n_testers
    if stage == "islands":
        return 40 * n_body + 32 * n_cons
    if stage == "island_solve":
        return manifold_bytes(m_isl, c_mean * m_isl, iters)
    if stage == "global_solve":
        return 64 * n_global + manifold_bytes(n_global, c_mean * n_global, iters)
    if stage == "integrate":
        return 80 * n_body
    return 0


def step_bytes(n_body, n_pair, n_cons, n_contacts, iters, n_global=0, n_convex=0, n_mesh=0, n_testers=0):
    """Whole-step algorithmic bytes: B_body*N + B_pair*P + B_convex*P_convex + B_mesh*P_mesh + sum over manifolds."""
    return 528 * n_body + 312 * n_pair + B_CONVEX_PAIR * n_convex + B_MESH_PAIR * n_mesh + B_MESH_TESTER * n_testers + \
        (332 + 190) * n_cons + (72 + 70) * n_contacts + iters * (316 * n_cons + 44 * n_contacts)


# Kernels of each timed stage (name substrings), for the measured HBM traffic
# of the dominant one: profiles/traffic_<workload>.json holds FETCH_SIZE /
# WRITE_SIZE per launch from rocprofv3 counter passes, scaled by a plain copy
# of known size.
STAGE_KERNELS = {
    "island_solve": ["k_island_solve"],
    "narrowphase": ["k_pair_prep", "k_narrowphase"],
    "convex": ["k_convex<"],
    "mesh": ["k_mesh"],
    "broadphase": ["k_grid_insert", "k_grid_query"],
    "global_solve": ["k_global"],
    "body_pre": ["k_body_pre"],
    "integrate": ["k_integrate"],
}


def measured_traffic(stage, workload, window=None):
    """HBM bytes per launch of the stage's kernels from the counter passes in
    profiles/traffic_<workload>.json, when that file was measured over the
    same steps as this run (`window` = (steps, warmup)); else None."""
    path = os.path.join(ROOT, "profiles", "traffic_%s.json" % workload)
    if stage not in STAGE_KERNELS or not os.path.exists(path):
        return None
    doc = json.load(open(path))
    if window is not None and tuple(doc.get("window", ())) != tuple(window):
        return None
    t = doc["bytes_per_launch"]
    ks = [k for k in t if any(s in k for s in STAGE_KERNELS[stage])]
    return sum(t[k] for k in ks) if ks else None


def host_cpu():
    """nproc, the threads this process may run on, and the CPU model."""
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = os.cpu_count()
    return {"nproc": os.cpu_count(), "affinity": affinity, "model": model}


def baseline_threads():
    """Host threads for the oracle baseline: every core this process may use,
    at most 16 (one GPU's share of the host on the benchmark machines)."""
    return max(1, min(16, host_cpu()["affinity"] or 1))


def stage_table(stages, counts):
    out = {}
    for name, ms in stages.items():
        b = algorithmic_bytes(name, **counts)
        out[name] = {"ms": round(ms, 5), "alg_bytes": int(b),
                     "GBps": round(b / (ms * 1e-3) / 1e9, 2) if ms > 0 else None}
    return out


def roofline(stages, counts, workload, window=None):
    dominant = max(stages, key=lambda k: stages[k])
    dom_bytes = algorithmic_bytes(dominant, **counts)
    achieved = dom_bytes / (stages[dominant] * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": measured_traffic(dominant, workload, window),
            "traffic_window": "steps %d-%d after the upload" % (window[1] + 1, window[1] + window[0]) if window else None,
            "kernel": dominant, "kernel_ms": round(stages[dominant], 5), "alg_bytes_per_launch": int(dom_bytes)}


def counts_of(st, iters):
    return dict(n_body=st.bodies, n_pair=st.pairs, n_cons=st.manifolds, n_contacts=st.contacts, iters=iters,
                n_global=st.global_manifolds, n_convex=int(sum(st.convex_pairs)), n_mesh=st.mesh_pairs,
                n_testers=st.mesh_testers)


def latency(sp, steps):
    if steps <= 0:
        return None
    lat = []
    for _ in range(steps):
        t = time.perf_counter()
        sp.Update(1)
        sp.stats()
        lat.append(1e3 * (time.perf_counter() - t))
    return {"p50": round(float(np.percentile(lat, 50)), 4), "p99": round(float(np.percentile(lat, 99)), 4)}


MIX = None  # boxes, spheres and 16-point hulls in equal shares (set in main)


def batched_settings():
    from bepuphysics1_amd._lib import default_settings
    st = default_settings()
    st.iteration_limit = BATCHED_ITERATIONS
    return st


def batched_run(a):
    """The headline: every world of the batch in one world handle, one
    bepu_step per step for all of them."""
    from bepuphysics1_amd import Space, scenes
    worlds = scenes.batched_worlds(a.worlds, seed=0, shapes=MIX)
    n_dyn = int(sum((w["mass"] > 0).sum() for w in worlds))
    sp = Space(capacity=sum(len(w) for w in worlds), settings=batched_settings())
    sp.register_hulls(scenes.hull_library())
    for w in worlds:
        sp.add_world(w)
    tune = dict(sp.tuning(), steps=0)
    forced = {}
    for kv in filter(None, (a.tuning or "").split(",")):
        k, v = kv.split("=")
        forced[k.strip()] = int(v)
    if forced:
        sp.set_tuning(**forced)
        tune = dict(sp.tuning(), steps=0, forced=True)
    if a.autotune_steps > 0 or (forced and a.settled_steps > 0):
        # The autotuner picks the launch configuration on a copy of the batch
        # stepped into its settled state (most of the timed window is settled
        # piles, whose costs differ from the falling ones), untimed; every
        # configuration gives bit-identical results.
        tsp = Space(capacity=sum(len(w) for w in worlds), settings=batched_settings())
        tsp.register_hulls(scenes.hull_library())
        for w in worlds:
            tsp.add_world(w)
        tsp.Update(a.autotune_after)
        if forced:
            tsp.set_tuning(**forced)
            # as many steps as the autotuner runs with every choice open
            tsp.Update(24 * (a.autotune_steps + 1))
            tune["steps"] = 24 * (a.autotune_steps + 1)
        else:
            tune = tsp.autotune(a.autotune_steps)
        tune["tuned_after_steps"] = a.autotune_after
        # The settled phase, which most of a long run is: the tuned copy is
        # now more than 300 steps in; time a fixed number of its steps.
        settled = timed_steps(tsp, a.settled_steps, n_dyn)
        settled["after_steps"] = a.autotune_after + tune["steps"]
        tsp.close()
        sp.set_tuning(**{k: tune[k] for k in Space.TUNING_FIELDS})
    else:
        settled = None
    if a.warmup:
        sp.Update(a.warmup)
    sp.stats()  # synchronise
    t0 = time.perf_counter()
    sp.Update(a.steps)
    st = sp.stats()  # synchronises the world's stream
    elapsed = time.perf_counter() - t0
    counts = counts_of(st, BATCHED_ITERATIONS)
    lat = latency(sp, a.latency_steps)
    sp.close()
    ref_defaults = None if a.no_reference_defaults else reference_defaults_run(a, worlds, n_dyn, tune)
    # Stage times: HIP events between the stages of the same steps as the
    # timed window (W + 1 .. W + K), on a replay of the batch (every step is
    # deterministic, so the replay's steps are the timed steps bit for bit),
    # so that the events' own cost stays out of `value`.
    rsp = Space(capacity=sum(len(w) for w in worlds), settings=batched_settings())
    rsp.register_hulls(scenes.hull_library())
    for w in worlds:
        rsp.add_world(w)
    rsp.set_tuning(**{k: tune[k] for k in Space.TUNING_FIELDS})
    if a.warmup:
        rsp.Update(a.warmup)
    rsp.set_timing(True)
    rsp.Update(a.steps)
    stages = rsp.stage_times()
    rsp.close()
    return worlds, {
        "value": n_dyn * a.steps / elapsed, "ms_per_step": 1e3 * elapsed / a.steps, "bodies": n_dyn,
        "stats": st, "counts": counts, "stages": stages, "latency": lat, "tuning": tune, "settled": settled,
        "reference_defaults": ref_defaults}


def reference_defaults_run(a, worlds, n_dyn, tune):
    """The same batch under the reference's defaults: no body is always
    active, so a settled island sleeps (DeactivationManager) and is not
    stepped until something wakes it.  Timed over the same window as the
    headline (steps W + 1 .. W + K); a body-step counts every body of the
    batch each step, asleep or not, as a Space.Update over the world does."""
    from bepuphysics1_amd import Space, scenes
    sp = Space(capacity=sum(len(w) for w in worlds), settings=batched_settings())
    sp.register_hulls(scenes.hull_library())
    for w in worlds:
        w2 = w.copy()
        w2["flags"] = 0
        sp.add_world(w2)
    sp.set_tuning(**{k: tune[k] for k in Space.TUNING_FIELDS})
    if a.warmup:
        sp.Update(a.warmup)
    sp.stats()
    t0 = time.perf_counter()
    sp.Update(a.steps)
    st = sp.stats()
    dt = time.perf_counter() - t0
    act = sp.activity()
    dyn = np.concatenate([w["mass"] > 0 for w in worlds])
    sp.close()
    return {"value": round(n_dyn * a.steps / dt, 1), "ms_per_step": round(1e3 * dt / a.steps, 4),
            "asleep_at_end": round(float((~act[dyn]).mean()), 4), "manifolds_at_end": st.manifolds}


def timed_steps(sp, steps, n_dyn):
    """`steps` synchronised steps of a world handle as a body-steps/s line."""
    sp.stats()
    t0 = time.perf_counter()
    sp.Update(steps)
    st = sp.stats()
    dt = time.perf_counter() - t0
    return {"value": round(n_dyn * steps / dt, 1), "ms_per_step": round(1e3 * dt / steps, 4), "steps": steps,
            "pairs": st.pairs, "manifolds": st.manifolds, "contacts": st.contacts,
            "convex_pairs_by_state": list(st.convex_pairs)}


def plain_path(worlds, sample_worlds, sample_steps):
    """The same worlds stepped one at a time, each alone in its own world
    handle (a sample of them, scaled to body-steps/s)."""
    from bepuphysics1_amd import Space, scenes
    rng = np.random.default_rng(1)
    pick = rng.choice(len(worlds), size=min(sample_worlds, len(worlds)), replace=False)
    t_sum, b_sum = 0.0, 0
    for k in pick:
        one = Space(capacity=len(worlds[k]), settings=batched_settings())
        one.register_hulls(scenes.hull_library())
        one.add_scene(worlds[k])
        one.Update(2)
        one.stats()
        t0 = time.perf_counter()
        for _ in range(sample_steps):
            one.Update(1)
        one.stats()
        t_sum += (time.perf_counter() - t0) / sample_steps
        b_sum += int((worlds[k]["mass"] > 0).sum())
        one.close()
    return b_sum / t_sum, "%d worlds x %d single steps, one world per handle" % (len(pick), sample_steps)


def oracle_hull_table():
    from bepuphysics1_amd import compute_hull, compute_hull_minimum_radius, scenes
    return [compute_hull(p)[:2] + (compute_hull_minimum_radius(p),) for p in scenes.hull_library()]


def cpu_baseline_batched(worlds, sample_worlds, warmup, steps, threads):
    """The oracle on a sample of the batch's worlds: every world its own oracle
    world, the worlds spread over `threads` host threads (the oracle releases
    the interpreter lock while it steps).  `warmup` untimed steps per world (as
    on the GPU), then `steps` steps of every sampled world timed together."""
    from concurrent.futures import ThreadPoolExecutor
    from oracle import refcpu
    hulls = oracle_hull_table()
    rng = np.random.default_rng(2)
    pick = rng.choice(len(worlds), size=min(sample_worlds, len(worlds)), replace=False)
    st = refcpu.default_settings()
    st.iteration_limit = BATCHED_ITERATIONS
    ows = []
    for k in pick:
        w = refcpu.World(st)
        w.add_scene(worlds[k], hulls)
        ows.append(w)
    with ThreadPoolExecutor(max_workers=threads) as ex:
        list(ex.map(lambda w: w.step(warmup), ows))
        t0 = time.perf_counter()
        list(ex.map(lambda w: w.step(steps), ows))
        dt = time.perf_counter() - t0
    n = int(sum((worlds[k]["mass"] > 0).sum() for k in pick))
    for w in ows:
        w.close()
    return {"value": round(n * steps / dt, 1), "unit": "body-steps/s", "cores": threads, "kind": "port",
            "sample": "oracle (C++ restatement of the reference step), %d of the %d worlds (%d bodies), one "
                      "oracle world per task on %d host threads, %d steps after %d untimed steps, %.1f s"
                      % (len(pick), len(worlds), n, threads, steps, warmup, dt),
            "host": host_cpu()}


def f32bits(a):
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)


def impulse_rel_err(rg, ro):
    x = np.array(list(rg.normal) + list(rg.friction) + list(rg.sliding) + [rg.twist], np.float64)
    y = np.array(list(ro.normal) + list(ro.friction) + list(ro.sliding) + [ro.twist], np.float64)
    return float((np.abs(x - y) / np.maximum(np.maximum(np.abs(x), np.abs(y)), 1e-3)).max())


def parity_batched(n_worlds, steps):
    """Worlds of the batched workload stepped together on the GPU and each
    alone on the oracle: motion states bit for bit, impulses by relative error."""
    from bepuphysics1_amd import Space, scenes
    from oracle import refcpu
    worlds = scenes.batched_worlds(n_worlds, seed=7, shapes=MIX)
    sp = Space(capacity=sum(len(w) for w in worlds), settings=batched_settings())
    sp.register_hulls(scenes.hull_library())
    for w in worlds:
        sp.add_world(w)
    sp.Update(steps)
    ms = sp.motion_states()
    imp = {int(r.key): r for r in sp.impulses()}
    hulls = oracle_hull_table()
    st = refcpu.default_settings()
    st.iteration_limit = BATCHED_ITERATIONS
    exact, err, keys_equal = True, 0.0, True
    for k, sc in enumerate(worlds):
        first, count = sp.world_range(k)
        o = refcpu.World(st)
        o.add_scene(sc, hulls)
        o.step(steps)
        exact &= bool(np.array_equal(f32bits(ms[first:first + count]), f32bits(o.motion_states())))
        shift = (first << 32) | first
        io = {int(r.key) + shift: r for r in o.impulses()}
        mine = {key: r for key, r in imp.items() if first <= (key & 0xffffffff) < first + count}
        keys_equal &= set(io) == set(mine)
        for key, ro in io.items():
            if key in mine:
                err = max(err, impulse_rel_err(mine[key], ro))
        o.close()
    sp.close()
    return {"scene": "%d worlds of the batched workload (seed 7), %d steps, batched on the GPU vs each alone on "
                     "the oracle" % (n_worlds, steps),
            "motion_states_bit_exact": exact, "constraint_sets_equal": keys_equal, "max_impulse_rel_err": err}


def cpu_baseline_towers(scene, steps, threads=1):
    """Oracle on the towers scene; one warm-up step excluded.  With threads > 1
    the body stages, broadphase cells and pair updates run in parallel and the
    solver colour by colour (same results as one thread)."""
    from oracle import refcpu
    w = refcpu.World()
    w.add_scene(scene)
    w.set_threads(threads)
    w.step(1)
    t0 = time.perf_counter()
    w.step(steps)
    dt = time.perf_counter() - t0
    dyn = int((scene["mass"] > 0).sum())
    w.close()
    return dyn * steps / dt, dt


def parity_towers():
    """Small towers scene stepped on the GPU and on the oracle."""
    from bepuphysics1_amd import Space, scenes
    from oracle import refcpu
    sc = scenes.towers(nx=8, nz=8, height=20, seed=1)
    sp = Space(capacity=len(sc))
    sp.add_scene(sc)
    sp.Update(60)
    w = refcpu.World()
    w.add_scene(sc)
    w.step(60)
    pairs_exact = bool(np.array_equal(sp.pairs(), w.pairs()))
    ig = {int(r.key): r for r in sp.impulses()}
    io = {int(r.key): r for r in w.impulses()}
    same_keys = set(ig) == set(io)
    err = max((impulse_rel_err(ig[k], ro) for k, ro in io.items()), default=0.0) if same_keys else None
    motion_exact = bool(np.array_equal(f32bits(sp.motion_states()), f32bits(w.motion_states())))
    sp.close()
    return {"scene": "towers 8x8x20, 60 steps", "pairs_bit_exact": pairs_exact, "constraint_sets_equal": same_keys,
            "max_impulse_rel_err": err, "motion_states_bit_exact": motion_exact}


def towers_leg(a, cpu, parity):
    """The plain single-world path on the large stacking scene."""
    from bepuphysics1_amd import Space, scenes
    scene = scenes.towers(nx=a.nx, nz=a.nz, height=a.height, seed=0)
    n_dyn = int((scene["mass"] > 0).sum())
    sp = Space(capacity=len(scene))
    sp.add_scene(scene)
    iters = sp.Solver.IterationLimit
    sp.Update(10)
    sp.stats()
    t0 = time.perf_counter()
    sp.Update(a.tower_steps)
    st = sp.stats()
    elapsed = time.perf_counter() - t0
    sp.set_timing(True)
    sp.Update(min(a.tower_steps, 50))
    stages = sp.stage_times()
    sp.set_timing(False)
    lat = latency(sp, a.latency_steps)
    ms_state = sp.motion_states()
    top = ms_state[1:].reshape(a.nx * a.nz, a.height, 13)[:, -1, :]
    start_top = scene["state"][1:].reshape(a.nx * a.nz, a.height, 13)[:, -1, :]
    # A tower stands while its top box has drifted less than half a box sideways
    # and sits within one box height of where it started (contacts are soft:
    # 20 stacked contacts compress by about 20 x allowed penetration).
    standing = int(((np.abs(top[:, 0:3:2] - start_top[:, 0:3:2]).max(axis=1) < 0.5) &
                    (np.abs(top[:, 1] - start_top[:, 1]) < 1.0)).sum())
    sp.close()
    counts = counts_of(st, iters)
    value = n_dyn * a.tower_steps / elapsed
    out = {"workload": "towers %dx%dx%d boxes (plain single-world GPU path), dt 1/60, %d iterations"
                       % (a.nx, a.nz, a.height, iters),
           "bodies": n_dyn, "steps": a.tower_steps, "value": round(value, 1),
           "ms_per_step": round(1e3 * elapsed / a.tower_steps, 4),
           "pairs": st.pairs, "manifolds": st.manifolds, "contacts": st.contacts, "colors": st.colors,
           "roofline": roofline(stages, counts, "towers", (a.tower_steps, 10)), "stages": stage_table(stages, counts),
           "latency_ms": lat, "towers_standing": standing, "max_penetration": round(float(st.max_penetration), 5)}
    if cpu:
        threads = baseline_threads()
        v_cpu, dt_cpu = cpu_baseline_towers(scene, a.cpu_steps, threads)
        v_1t, dt_1t = cpu_baseline_towers(scene, max(2, a.cpu_steps // 4), 1)
        out["cpu_baseline"] = {
            "value": round(v_cpu, 1), "unit": "body-steps/s", "cores": threads, "kind": "port",
            "sample": "oracle (C++ restatement of the reference step; %d threads: parallel body stages, broadphase "
                      "and pair updates, colour-parallel solver), same 81,920-box towers scene, %d steps after 1 "
                      "untimed step, %.1f s" % (threads, a.cpu_steps, dt_cpu),
            "single_thread_value": round(v_1t, 1),
            "single_thread_sample": "%d steps, %.1f s" % (max(2, a.cpu_steps // 4), dt_1t), "host": host_cpu()}
        out["speedup_vs_cpu"] = round(value / v_cpu, 1)
    if parity:
        out["parity"] = parity_towers()
    return out


def served_leg(rate, duration, max_wall):
    """Served rollouts: Poisson arrivals at `rate`/s for `duration` s of wall
    time, each a random pile of N ~ logU[16,2048] bodies (boxes, spheres and
    16-point hulls in equal shares, own ground) stepped S ~ U[60,600] times at
    8 iterations; continuous batching on one world handle (admission against
    capacity and a 200 GB budget), then the queue drains."""
    from bepuphysics1_amd import scenes
    from bepuphysics1_amd.serve import Server, run_poisson
    srv = Server(capacity=4_000_000, settings=batched_settings(), hbm_budget=200e9, hull_shapes=scenes.hull_library())

    def make_request(rng):
        n = int(np.exp(rng.uniform(np.log(16), np.log(2048))))
        return scenes.random_pile_fast(n, rng, sphere_fraction=1 / 3, hull_fraction=1 / 3), int(rng.integers(60, 601))

    out = run_poisson(srv, rate, duration, make_request, seed=0, drain=True, max_wall=max_wall)
    srv.close()
    out["workload"] = ("Poisson %g requests/s for %g s, N ~ logU[16,2048] boxes + spheres + 16-point hulls, "
                       "S ~ U[60,600], 8 iterations, continuous batching" % (rate, duration))
    return out


def stress_leg(n, reps, cpu):
    """Broadphase + narrowphase stress: n bodies (50% boxes, 25% spheres, 25%
    16-point hulls) at random in a cube sized for a mean AABB degree of 4, one
    step from a fresh world (every pair new), timed after upload; the oracle's
    step on the same scene beside it."""
    from bepuphysics1_amd import Space, scenes
    sc = scenes.stress(n, seed=0, shapes=scenes.STRESS_MIX)
    best, stages, stats = None, None, None
    for _ in range(reps):
        sp = Space(capacity=n)
        sp.register_hulls(scenes.hull_library())
        sp.add_scene(sc)
        sp.motion_states()  # upload and synchronise
        sp.set_timing(True)
        t0 = time.perf_counter()
        sp.Update(1)
        sp.synchronize()
        dt = time.perf_counter() - t0
        if best is None or dt < best:
            best, stages, stats = dt, sp.stage_times(), sp.stats()
        sp.close()
    out = {"workload": "stress, %d bodies (50%% boxes, 25%% spheres, 25%% hulls), 1 step from a fresh world" % n,
           "ms_per_step": round(1e3 * best, 4), "pairs": stats.pairs, "manifolds": stats.manifolds,
           "contacts": stats.contacts, "convex_pairs_by_state": list(stats.convex_pairs),
           "pairs_per_s": round(stats.pairs / best, 1),
           "stages_ms": {k: round(v, 4) for k, v in stages.items()}}
    if cpu:
        from oracle import refcpu
        w = refcpu.World()
        w.add_scene(sc, oracle_hull_table())
        t0 = time.perf_counter()
        w.step(1)
        out["cpu_oracle_ms"] = round(1e3 * (time.perf_counter() - t0), 2)
        w.close()
    return out


def mesh_leg(n, warmup, steps, cpu, parity, kinds=("terrain", "static_mesh")):
    """Static meshes and terrains: n bodies of all four shapes dropped onto a
    bumpy terrain, and onto a bumpy static mesh (128 x 128 quads, 32,768
    triangles), stepped `warmup` steps (the bodies land) and then `steps`
    timed device-resident steps; stage times over 20 more steps.  Parity:
    the motion states after the warm-up against the oracle's, bit for bit.
    cpu: the oracle's next 10 steps on host threads."""
    from bepuphysics1_amd import Space, scenes
    out = {}
    threads = baseline_threads()
    for kind in kinds:
        if kind == "terrain":
            sc, meshes = scenes.terrain_pile(n, seed=1)
            t = meshes[0]
            desc = "%d bodies (boxes, spheres, 16-point hulls, capsules) on a %dx%d terrain (spacing 2, " \
                   "amplitude 1)" % (n, t.heights.shape[0], t.heights.shape[1])
        else:
            sc, meshes = scenes.mesh_pile(n, seed=1, cells=128)
            desc = "%d bodies (boxes, spheres, 16-point hulls, capsules) on a 128x128-quad static mesh " \
                   "(32,768 triangles, amplitude 1)" % n
        sp = Space(capacity=n + 1)
        sp.register_hulls(scenes.hull_library())
        sp.register_meshes(meshes)
        sp.add_scene(sc)
        sp.Update(warmup)
        sp.synchronize()
        st0 = sp.motion_states()
        sp.synchronize()
        t0 = time.perf_counter()
        sp.Update(steps)
        sp.synchronize()
        dt = (time.perf_counter() - t0) / steps
        sp.set_timing(True)
        sp.Update(20)
        stages = sp.stage_times()
        sp.set_timing(False)
        st = sp.stats()
        counts = counts_of(st, 10)
        mesh_ms = stages.get("mesh", 0.0)
        mesh_bytes = algorithmic_bytes("mesh", **counts)
        leg = {"workload": desc + ", dt 1/60, 10 iterations",
               "value": round(len(sc) / dt, 1), "unit": "body-steps/s", "ms_per_step": round(1e3 * dt, 4),
               "steps": steps, "warmup": warmup, "pairs": st.pairs, "manifolds": st.manifolds,
               "contacts": st.contacts, "mesh_pairs": st.mesh_pairs, "mesh_testers": st.mesh_testers,
               "convex_pairs_by_state": list(st.convex_pairs),
               "stages": stage_table(stages, counts),
               "roofline": {"bound": "latency (one pair per lane: triangle GJK/MPR chains)", "kernel": "k_mesh",
                            "kernel_ms": round(mesh_ms, 5), "alg_bytes_per_launch": int(mesh_bytes),
                            "achieved": round(mesh_bytes / (mesh_ms * 1e-3) / 1e9, 2) if mesh_ms > 0 else None,
                            "peak": HBM_PEAK_GBS, "unit": "GB/s",
                            "frac": round(mesh_bytes / (mesh_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 5)
                            if mesh_ms > 0 else None,
                            "traffic": measured_traffic("mesh", "mesh_" + kind)}}
        sp.close()
        if parity or cpu:
            from oracle import refcpu
            w = refcpu.World()
            w.set_threads(threads)
            w.add_scene(sc, oracle_hull_table(), meshes)
            w.step(warmup)
            if parity:
                leg["parity"] = {"steps": warmup, "bit_identical_states":
                                 bool(np.array_equal(f32bits(st0), f32bits(w.motion_states())))}
            if cpu:
                t0 = time.perf_counter()
                w.step(10)
                ct = (time.perf_counter() - t0) / 10
                leg["cpu_baseline"] = {"value": round(len(sc) / ct, 1), "unit": "body-steps/s", "cores": threads,
                                       "kind": "port", "sample": "the oracle's steps %d-%d of the same scene" %
                                       (warmup + 1, warmup + 10)}
                leg["speedup_vs_cpu"] = round(leg["value"] / leg["cpu_baseline"]["value"], 1)
            w.close()
        out[kind] = leg
    return out


def calibrate_copy(nbytes=1 << 30):
    """Plain device-to-device copy of a known size (the HIP runtime's copy
    kernel), run in the same process as the workload so that one profiler
    pass records both: the counter values it reads for `nbytes` read and
    `nbytes` written scale FETCH_SIZE / WRITE_SIZE to bytes."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    if hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(nbytes)) or hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(nbytes)):
        raise RuntimeError("hipMalloc failed")
    hip.hipMemset(a, 1, ctypes.c_size_t(nbytes))
    for _ in range(3):
        hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3)  # hipMemcpyDeviceToDevice
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    return nbytes


def _csv_rows(path):
    import csv
    with open(path, newline="") as f:
        return list(csv.DictReader(f))


def _window_index(rows, steps, warmup):
    """Per launch (rows in dispatch order): whether it belongs to a step the
    stage events time.  A step starts at its k_step_reset launch.  The
    profiled command (--legs batched --latency-steps 0 --autotune-steps 0
    --plain-sample 0 --no-parity --no-cpu-baseline --settled-steps 0) runs the
    batch for warmup + steps steps, then its replay for as many; the events
    time the replay's steps warmup + 1 .. warmup + steps."""
    run_len = warmup + steps
    total = sum(1 for r in rows if "k_step_reset" in r["Kernel_Name"])
    first = total - run_len + warmup  # steps before the replay's timed ones
    g = 0
    out = []
    for r in rows:
        if "k_step_reset" in r["Kernel_Name"]:
            g += 1
        out.append(g > first)
    return out


def summarize_trace(directory, steps, warmup, fetch_scale, write_scale, workload="batched"):
    """Per-kernel launch times (kernel trace) and HBM bytes (FETCH_SIZE /
    WRITE_SIZE counter passes, scaled) of a rocprofv3 --output-format csv run
    of one leg, over all launches and over the timed window's, and the bytes
    per launch `roofline` reads (profiles/traffic_<workload>.json): the
    window's for the batched leg (its replay's steps warmup + 1 .. warmup +
    steps), every launch's for the towers leg (a steady scene)."""
    import glob
    out = {"workload": workload, "window": [steps, warmup], "fetch_scale": fetch_scale, "write_scale": write_scale,
           "kernels": {}}
    traces = glob.glob(os.path.join(directory, "**", "*kernel_trace.csv"), recursive=True)
    counters = glob.glob(os.path.join(directory, "**", "*counter_collection.csv"), recursive=True)
    for path in traces:
        rows = sorted(_csv_rows(path), key=lambda r: int(r["Dispatch_Id"]))
        win = _window_index(rows, steps, warmup)
        for r, w in zip(rows, win):
            k = out["kernels"].setdefault(r["Kernel_Name"], {"calls": 0, "us": 0.0, "win_calls": 0, "win_us": 0.0})
            us = (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3
            k["calls"] += 1
            k["us"] += us
            if w:
                k["win_calls"] += 1
                k["win_us"] += us
    for path in counters:
        rows = _csv_rows(path)
        by_dispatch = {}
        for r in rows:
            d = by_dispatch.setdefault(int(r["Dispatch_Id"]), {"Kernel_Name": r["Kernel_Name"]})
            d[r["Counter_Name"]] = d.get(r["Counter_Name"], 0.0) + float(r["Counter_Value"])
        disp = [by_dispatch[i] for i in sorted(by_dispatch)]
        win = _window_index(disp, steps, warmup)
        for d, w in zip(disp, win):
            k = out["kernels"].setdefault(d["Kernel_Name"], {"calls": 0, "us": 0.0, "win_calls": 0, "win_us": 0.0})
            for name in ("FETCH_SIZE", "WRITE_SIZE"):
                if name in d:
                    k.setdefault(name + "_kib_all", []).append(d[name])
                    if w:
                        k.setdefault(name + "_kib_win", []).append(d[name])
    for name, k in out["kernels"].items():
        k["avg_us"] = k["us"] / k["calls"] if k["calls"] else None
        k["win_avg_us"] = k["win_us"] / k["win_calls"] if k["win_calls"] else None
        for c, scale in (("FETCH_SIZE", fetch_scale), ("WRITE_SIZE", write_scale)):
            for part in ("all", "win"):
                vals = k.pop("%s_kib_%s" % (c, part), None)
                if vals:
                    k["%s_kib_%s" % (c, part)] = float(np.mean(vals))
                    k["%s_bytes_%s" % (c, part)] = float(np.mean(vals)) * 1024.0 * scale
    part = "win" if workload == "batched" else "all"
    out["bytes_per_launch"] = {
        name: int(k["FETCH_SIZE_bytes_" + part] + k["WRITE_SIZE_bytes_" + part])
        for name, k in out["kernels"].items()
        if "FETCH_SIZE_bytes_" + part in k and "WRITE_SIZE_bytes_" + part in k}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300, help="timed batched steps")
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--worlds", type=int, default=4096)
    ap.add_argument("--autotune-steps", type=int, default=10,
                    help="steps per candidate and round of the launch autotuner (0: default configuration)")
    ap.add_argument("--autotune-after", type=int, default=160,
                    help="steps a copy of the batch runs before the autotuner times candidates on it")
    ap.add_argument("--latency-steps", type=int, default=50)
    ap.add_argument("--tuning", default="",
                    help="fixed launch configuration instead of the autotuner's, e.g. convex_lds=0,narrowphase_block=64")
    ap.add_argument("--settled-steps", type=int, default=50,
                    help="steps timed on the autotuner's copy of the batch once it has settled")
    ap.add_argument("--plain-sample", type=int, default=48,
                    help="worlds of the batch also stepped alone, for the batched/plain ratio (0: skip)")
    ap.add_argument("--cpu-worlds", type=int, default=4096, help="worlds in the oracle's sample")
    ap.add_argument("--cpu-steps", type=int, default=30)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-parity", action="store_true")
    ap.add_argument("--no-reference-defaults", action="store_true",
                    help="skip the batch under the reference's defaults (sleeping)")
    ap.add_argument("--calibrate-copy", action="store_true")
    ap.add_argument("--nx", type=int, default=64)
    ap.add_argument("--nz", type=int, default=64)
    ap.add_argument("--height", type=int, default=20)
    ap.add_argument("--tower-steps", type=int, default=200)
    ap.add_argument("--served-rate", type=float, default=2000.0)
    ap.add_argument("--served-seconds", type=float, default=4.0)
    ap.add_argument("--stress-bodies", type=int, default=32768)
    ap.add_argument("--mesh-bodies", type=int, default=16384)
    ap.add_argument("--mesh-steps", type=int, default=120)
    ap.add_argument("--mesh-kinds", default="terrain,static_mesh", help="grounds of the mesh leg (for profiling one)")
    ap.add_argument("--legs", default="batched,towers,served,stress,mesh",
                    help="comma list of workloads; without batched only the other legs' results are printed")
    ap.add_argument("--summarize-trace", metavar="DIR",
                    help="print per-kernel averages over the timed window of a rocprofv3 csv run in DIR and exit")
    ap.add_argument("--fetch-scale", type=float, default=2.0,
                    help="bytes per FETCH_SIZE KiB unit / 1024, from --calibrate-copy")
    ap.add_argument("--write-scale", type=float, default=1.0)
    a = ap.parse_args()
    legs = set(x.strip() for x in a.legs.split(",") if x.strip())
    if a.summarize_trace:
        workload = "batched" if "batched" in legs else "towers"
        steps, warmup = (a.steps, a.warmup) if workload == "batched" else (a.tower_steps, 10)
        print(json.dumps(summarize_trace(a.summarize_trace, steps, warmup, a.fetch_scale, a.write_scale, workload),
                         indent=1))
        return

    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: runs on one GPU only; --gpus %d is not supported "
              "(independent worlds shard across GPUs without communication)" % a.gpus,
              file=sys.stderr)
        sys.exit(2)

    from bepuphysics1_amd import scenes
    global MIX
    MIX = (scenes.SHAPE_BOX, scenes.SHAPE_SPHERE, scenes.SHAPE_CONVEX_HULL)
    cpu = not a.no_cpu_baseline
    copy_bytes = calibrate_copy() if a.calibrate_copy else None

    if "batched" not in legs:
        # Secondary workloads alone (kernel profiles of one workload at a time).
        out = {"legs": sorted(legs), "calibration_copy_bytes": copy_bytes}
        if "towers" in legs:
            out["towers"] = towers_leg(a, cpu, not a.no_parity)
        if "served" in legs:
            out["served"] = served_leg(a.served_rate, a.served_seconds, a.served_seconds + 20.0)
        if "stress" in legs:
            out["stress"] = stress_leg(a.stress_bodies, 3, cpu)
        if "mesh" in legs:
            out["mesh"] = mesh_leg(a.mesh_bodies, 60, a.mesh_steps, cpu, not a.no_parity, a.mesh_kinds.split(","))
        print(json.dumps(out))
        return

    worlds, b = batched_run(a)
    st, counts, stages = b["stats"], b["counts"], b["stages"]
    plain = None
    if a.plain_sample > 0:
        pv, ps = plain_path(worlds, a.plain_sample, 10)
        plain = {"value": round(pv, 1), "sample": ps, "ratio_batched_vs_plain": round(b["value"] / pv, 1)}
    cpu_line = None
    if cpu:
        threads = baseline_threads()
        # The oracle's sample sits in the middle of the GPU's timed window (the
        # piles settle over the run and later steps carry more contacts).
        cpu_line = cpu_baseline_batched(worlds, a.cpu_worlds, a.warmup + a.steps // 2, a.cpu_steps, threads)
    parity = None if a.no_parity else parity_batched(16, 60)
    step_alg = step_bytes(**counts)
    line = {
        "metric": "body-steps/sec @ <=1e-4 impulse err; p50/p99 step latency, batched worlds",
        "value": round(b["value"], 1),
        "unit": "body-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(b["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "batched many-world: %d worlds, N ~ logU[8,256] boxes + spheres + 16-point hulls "
                               "dropped onto own grounds, dt 1/60, %d iterations, one step of every world per "
                               "bepu_step" % (a.worlds, BATCHED_ITERATIONS),
                   "worlds": a.worlds, "bodies": b["bodies"], "pairs": st.pairs, "manifolds": st.manifolds,
                   "contacts": st.contacts, "colors": st.colors, "convex_pairs_by_state": list(st.convex_pairs),
                   "autotune": b["tuning"]},
        "roofline": roofline(stages, counts, "batched", (a.steps, a.warmup)),
        "settled": b["settled"],
        "reference_defaults": b["reference_defaults"],
        "step_roofline": {"alg_bytes_per_step": int(step_alg),
                          "achieved_GBps": round(step_alg / (b["ms_per_step"] * 1e-3) / 1e9, 2),
                          "frac": round(step_alg / (b["ms_per_step"] * 1e-3) / 1e9 / HBM_PEAK_GBS, 5)},
        "stages": stage_table(stages, counts),
        "latency_ms": b["latency"],
        "cpu_baseline": cpu_line,
        "speedup_vs_cpu": round(b["value"] / cpu_line["value"], 1) if cpu_line else None,
        "plain_path": plain,
        "parity": parity,
        "towers": towers_leg(a, cpu, not a.no_parity) if "towers" in legs else None,
        "served": served_leg(a.served_rate, a.served_seconds, a.served_seconds + 20.0) if "served" in legs else None,
        "stress": stress_leg(a.stress_bodies, 3, cpu) if "stress" in legs else None,
        "mesh": mesh_leg(a.mesh_bodies, 60, a.mesh_steps, cpu, not a.no_parity, a.mesh_kinds.split(",")) if "mesh" in legs else None,
        "calibration_copy_bytes": copy_bytes,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
