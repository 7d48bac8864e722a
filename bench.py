"""Benchmark of the gfx950 step path: body-steps/s on one MI355X.

Workload (N=1): the large stacking scene -- a 64 x 64 grid of 20-high box
towers (81,920 dynamic boxes, half extent 0.5, 0.01 seeded jitter) on a
static ground, dt 1/60, 10 solver iterations, FP32, warm starting on.
Inputs are resident on the device before the timed region; the timed region
is K full steps (force, AABB, broadphase, narrowphase, colouring, prestep,
warm start + iterations, integration) bracketed by device synchronisation.

The JSON line also carries
  roofline      the dominant kernel's algorithmic bytes per launch over its
                average duration (HIP events on the world's stream), against
                the MI355X HBM peak of 8 TB/s;
  cpu_baseline  the oracle (this repository's C++ restatement of the
                reference path) timed on up to 16 host threads, and on one,
                on a bounded sample of the same scene;
  parity        a small scene stepped on both sides: pair sets and impulses;
  latency       p50 / p99 of single synchronised steps;
  batched, served, stress
                the many-world batch, continuous batching of rollouts and the
                broadphase/narrowphase stress scene.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--legs towers,batched,served,stress]
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

# Algorithmic bytes per unit of work, per stage (FP32, each element counted
# once per stage and once per solver iteration, no cache-reuse credit):
#   body_pre      force/inertia 160 + AABB 80 per body
#   broadphase    Morton 32 + radix sort 64 + tree build 88 + traversal leaf read 24 per body
#   narrowphase   312 per broadphase pair
#   islands       union-find and island bucketing: 40 per body (parent init 4,
#                 compress 12, count 4, body list 8, local index 4, tier list 8)
#                 + 32 per manifold (union reads 16, count 8, constraint list 8)
#   island_solve  prestep 332 + 72 c, warm start 190 + 70 c and per iteration
#                 316 + 44 c, per manifold of an island solved in one workgroup
#   global_solve  the same per manifold of a large island, plus 64 for its
#                 colouring (pair read 24, incidence 16, colour/priority 16, order 8)
#   integrate     80 per body
# c = contacts of the manifold (the scene's mean is used per tier).
def manifold_bytes(m, contacts, iters):
    return (332 + 190) * m + (72 + 70) * contacts + iters * (316 * m + 44 * contacts)


def algorithmic_bytes(stage, n_body, n_pair, n_cons, n_contacts, iters, n_global=0):
    c_mean = n_contacts / n_cons if n_cons else 0.0
    m_isl = n_cons - n_global
    if stage == "body_pre":
        return 240 * n_body
    if stage == "broadphase":
        return 208 * n_body
    if stage == "narrowphase":
        return 312 * n_pair
    if stage == "islands":
        return 40 * n_body + 32 * n_cons
    if stage == "island_solve":
        return manifold_bytes(m_isl, c_mean * m_isl, iters)
    if stage == "global_solve":
        return 64 * n_global + manifold_bytes(n_global, c_mean * n_global, iters)
    if stage == "integrate":
        return 80 * n_body
    return 0


# Kernels of each timed stage, for the measured HBM traffic of the dominant one
# (profiles/traffic.json: FETCH_SIZE / WRITE_SIZE per launch from rocprofv3
# counter passes, scaled by a plain copy of known size).
STAGE_KERNELS = {
    "island_solve": ["bepu::k_island_pack(bepu::DeviceWorld)", "bepu::k_island_medium(bepu::DeviceWorld)"],
    "narrowphase": ["bepu::k_narrowphase(bepu::DeviceWorld)"],
    "broadphase": ["bepu::k_grid_insert(bepu::DeviceWorld)", "bepu::k_grid_query(bepu::DeviceWorld)"],
    "global_solve": ["bepu::k_global(bepu::DeviceWorld, unsigned long)"],
}


def measured_traffic(stage):
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if stage not in STAGE_KERNELS or not os.path.exists(path):
        return None
    t = json.load(open(path))["bytes_per_launch"]
    ks = [k for k in STAGE_KERNELS[stage] if k in t]
    return sum(t[k] for k in ks) if ks else None


def step_bytes(n_body, n_pair, n_cons, n_contacts, iters, n_global=0):
    """Whole-step algorithmic bytes: B_body*N + B_pair*P + sum over manifolds."""
    return 528 * n_body + 312 * n_pair + (332 + 190) * n_cons + (72 + 70) * n_contacts + \
        iters * (316 * n_cons + 44 * n_contacts)


def parity_check():
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
    err = 0.0
    same_keys = set(ig) == set(io)
    if same_keys:
        for k, ro in io.items():
            rg = ig[k]
            a = np.array(list(rg.normal) + list(rg.friction) + list(rg.sliding) + [rg.twist], np.float64)
            b = np.array(list(ro.normal) + list(ro.friction) + list(ro.sliding) + [ro.twist], np.float64)
            err = max(err, float((np.abs(a - b) / np.maximum(np.maximum(np.abs(a), np.abs(b)), 1e-3)).max()))
    motion = float(np.abs(sp.motion_states() - w.motion_states()).max())
    sp.close()
    return {"scene": "towers 8x8x20, 60 steps", "pairs_bit_exact": pairs_exact, "constraint_sets_equal": same_keys,
            "max_impulse_rel_err": err, "max_motion_abs_diff": motion}


def cpu_baseline(scene, steps, threads=1):
    """Oracle (C++ restatement of the reference step) on the same scene; one
    warm-up step excluded.  With threads > 1 the body stages, broadphase cells
    and pair updates run in parallel and the solver colour by colour (same
    results as one thread)."""
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


MIX = None  # boxes, spheres and 16-point hulls in equal shares (set in main)


def batched_leg(n_worlds, steps, warmup, sample_worlds, sample_steps):
    """4096 independent piles (N ~ logU[8,256]; boxes, spheres and 16-point
    hulls; own grounds; 8 iterations) stepped as sub-worlds of one batch,
    against the plain path: the same worlds stepped one at a time, each in its
    own world handle (sampled, then scaled to all worlds)."""
    from bepuphysics1_amd import Space, scenes
    from bepuphysics1_amd._lib import default_settings
    st = default_settings()
    st.iteration_limit = 8
    worlds = scenes.batched_worlds(n_worlds, seed=0, shapes=MIX)
    n_dyn = int(sum((w["mass"] > 0).sum() for w in worlds))
    sp = Space(capacity=sum(len(w) for w in worlds), settings=st)
    sp.register_hulls(scenes.hull_library())
    for w in worlds:
        sp.add_world(w)
    sp.Update(warmup)
    sp.stats()
    t0 = time.perf_counter()
    sp.Update(steps)
    stats = sp.stats()
    dt = time.perf_counter() - t0
    # Stage times in a separate run: per-stage events cost time of their own.
    sp.set_timing(True)
    sp.Update(min(steps, 20))
    stages = sp.stage_times()
    sp.close()
    batched = n_dyn * steps / dt
    out = {"workload": "%d worlds, N ~ logU[8,256] boxes + spheres + 16-point hulls on own grounds, dt 1/60, "
                       "8 iterations" % n_worlds,
           "bodies": n_dyn, "steps": steps, "value": round(batched, 1), "ms_per_step": round(1e3 * dt / steps, 4),
           "pairs": stats.pairs, "manifolds": stats.manifolds,
           "stages_ms": {k: round(v, 4) for k, v in stages.items()}}
    if sample_worlds <= 0:
        return out
    # Plain path: every world alone, one step after another.
    rng = np.random.default_rng(1)
    pick = rng.choice(n_worlds, size=min(sample_worlds, n_worlds), replace=False)
    t_sum, b_sum = 0.0, 0
    for k in pick:
        one = Space(capacity=len(worlds[k]), settings=st)
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
    plain = b_sum / t_sum  # body-steps/s when worlds are stepped one at a time
    out.update({"plain_value": round(plain, 1),
                "plain_sample": "%d worlds x %d single steps, one world per handle" % (len(pick), sample_steps),
                "ratio_vs_plain": round(batched / plain, 1)})
    return out


def served_leg(rate, duration, max_wall):
    """Served rollouts: Poisson arrivals at `rate`/s for `duration` s of wall
    time, each a random pile of N ~ logU[16,2048] bodies (boxes, spheres and
    16-point hulls in equal shares, own ground) stepped S ~ U[60,600] times at
    8 iterations; continuous batching on one world handle (admission against
    capacity and a 200 GB budget), then the queue drains."""
    from bepuphysics1_amd import scenes
    from bepuphysics1_amd._lib import default_settings
    from bepuphysics1_amd.serve import Server, run_poisson
    st = default_settings()
    st.iteration_limit = 8
    srv = Server(capacity=1_000_000, settings=st, hbm_budget=200e9, hull_shapes=scenes.hull_library())

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
           "contacts": stats.contacts, "pairs_per_s": round(stats.pairs / best, 1),
           "stages_ms": {k: round(v, 4) for k, v in stages.items()}}
    if cpu:
        from oracle import refcpu
        from bepuphysics1_amd import compute_hull
        w = refcpu.World()
        w.add_scene(sc, [compute_hull(p)[:2] for p in scenes.hull_library()])
        t0 = time.perf_counter()
        w.step(1)
        out["cpu_oracle_ms"] = round(1e3 * (time.perf_counter() - t0), 2)
        w.close()
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--nx", type=int, default=64)
    ap.add_argument("--nz", type=int, default=64)
    ap.add_argument("--height", type=int, default=20)
    ap.add_argument("--cpu-steps", type=int, default=20)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-parity", action="store_true")
    ap.add_argument("--latency-steps", type=int, default=50)
    ap.add_argument("--calibrate-copy", action="store_true")
    ap.add_argument("--batched-worlds", type=int, default=4096)
    ap.add_argument("--batched-steps", type=int, default=50)
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--plain-sample", type=int, default=48,
                    help="worlds of the batch also stepped alone, for the batched/plain ratio (0: skip)")
    ap.add_argument("--served-rate", type=float, default=2000.0)
    ap.add_argument("--served-seconds", type=float, default=3.0)
    ap.add_argument("--no-served", action="store_true")
    ap.add_argument("--stress-bodies", type=int, default=32768)
    ap.add_argument("--no-stress", action="store_true")
    ap.add_argument("--legs", default="towers,batched,served,stress",
                    help="comma list of workloads to run; without towers only the other legs' results are printed")
    a = ap.parse_args()
    legs = set(x.strip() for x in a.legs.split(",") if x.strip())
    for name, off in (("batched", a.no_batched), ("served", a.no_served), ("stress", a.no_stress)):
        if off:
            legs.discard(name)

    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: runs on one GPU only; --gpus %d is not supported "
              "(independent worlds shard across GPUs without communication)" % a.gpus,
              file=sys.stderr)
        sys.exit(2)

    from bepuphysics1_amd import Space, scenes
    global MIX
    MIX = (scenes.SHAPE_BOX, scenes.SHAPE_SPHERE, scenes.SHAPE_CONVEX_HULL)

    if "towers" not in legs:
        # Secondary workloads alone (kernel profiles of one workload at a time).
        out = {"legs": sorted(legs)}
        if "batched" in legs:
            out["batched"] = batched_leg(a.batched_worlds, a.batched_steps, 5, a.plain_sample, 10)
        if "served" in legs:
            out["served"] = served_leg(a.served_rate, a.served_seconds, a.served_seconds + 20.0)
        if "stress" in legs:
            out["stress"] = stress_leg(a.stress_bodies, 3, False)
        print(json.dumps(out))
        return

    scene = scenes.towers(nx=a.nx, nz=a.nz, height=a.height, seed=0)
    n_dyn = int((scene["mass"] > 0).sum())
    sp = Space(capacity=len(scene))
    sp.add_scene(scene)
    iters = sp.Solver.IterationLimit
    if a.warmup:
        sp.Update(a.warmup)
    sp.stats()  # synchronise

    t0 = time.perf_counter()
    sp.Update(a.steps)
    st = sp.stats()  # synchronises the world's stream
    elapsed = time.perf_counter() - t0
    # Stage times (HIP events between the stages) in a separate run, so that
    # the events' own cost stays out of `value`.
    sp.set_timing(True)
    sp.Update(min(a.steps, 50))
    stages = sp.stage_times()
    sp.set_timing(False)

    value = n_dyn * a.steps / elapsed
    ms_per_step = 1e3 * elapsed / a.steps

    counts = dict(n_body=st.bodies, n_pair=st.pairs, n_cons=st.manifolds, n_contacts=st.contacts, iters=iters,
                  n_global=st.global_manifolds)
    per_stage = {}
    for name, ms in stages.items():
        b = algorithmic_bytes(name, **counts)
        per_stage[name] = {"ms": round(ms, 5), "alg_bytes": int(b),
                           "GBps": round(b / (ms * 1e-3) / 1e9, 2) if ms > 0 else None}
    dominant = max(stages, key=lambda k: stages[k])
    dom_bytes = algorithmic_bytes(dominant, **counts)
    achieved = dom_bytes / (stages[dominant] * 1e-3) / 1e9
    step_alg = step_bytes(**counts)

    # latency of single synchronised steps
    lat = []
    for _ in range(a.latency_steps):
        t = time.perf_counter()
        sp.Update(1)
        sp.stats()
        lat.append(1e3 * (time.perf_counter() - t))
    ms_state = sp.motion_states()
    top = ms_state[1:].reshape(a.nx * a.nz, a.height, 13)[:, -1, :]
    start_top = scene["state"][1:].reshape(a.nx * a.nz, a.height, 13)[:, -1, :]
    # A tower stands while its top box has drifted less than half a box sideways
    # and sits within one box height of where it started (contacts are soft:
    # 20 stacked contacts compress by about 20 x allowed penetration).
    standing = int(((np.abs(top[:, 0:3:2] - start_top[:, 0:3:2]).max(axis=1) < 0.5) &
                    (np.abs(top[:, 1] - start_top[:, 1]) < 1.0)).sum())
    sp.close()

    cpu = None
    if not a.no_cpu_baseline:
        threads = max(1, min(16, os.cpu_count() or 1))
        v_cpu, dt_cpu = cpu_baseline(scene, a.cpu_steps, threads)
        v_1t, dt_1t = cpu_baseline(scene, max(2, a.cpu_steps // 4), 1)
        cpu = {"value": round(v_cpu, 1), "unit": "body-steps/s", "cores": threads, "kind": "port",
               "sample": "oracle (C++ restatement of the reference step; %d threads: parallel body stages, "
                         "broadphase and pair updates, colour-parallel solver), same 81,920-box towers scene, "
                         "%d steps after 1 untimed step, %.1f s" % (threads, a.cpu_steps, dt_cpu),
               "single_thread_value": round(v_1t, 1),
               "single_thread_sample": "%d steps, %.1f s" % (max(2, a.cpu_steps // 4), dt_1t)}

    parity = None if a.no_parity else parity_check()
    copy_bytes = calibrate_copy() if a.calibrate_copy else None
    batched = batched_leg(a.batched_worlds, a.batched_steps, 5, a.plain_sample, 10) if "batched" in legs else None
    served = served_leg(a.served_rate, a.served_seconds, a.served_seconds + 20.0) if "served" in legs else None
    stress = stress_leg(a.stress_bodies, 3, not a.no_cpu_baseline) if "stress" in legs else None

    line = {
        "metric": "body-steps/sec @ <=1e-4 impulse err; p50/p99 step latency",
        "value": round(value, 1),
        "unit": "body-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "towers %dx%dx%d boxes (plain single-world GPU path), dt 1/60, %d iterations"
                   % (a.nx, a.nz, a.height, iters),
                   "bodies": n_dyn, "pairs": st.pairs, "manifolds": st.manifolds, "contacts": st.contacts,
                   "colors": st.colors},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": measured_traffic(dominant),
                     "kernel": dominant,
                     "kernel_ms": round(stages[dominant], 5), "alg_bytes_per_launch": int(dom_bytes)},
        "step_roofline": {"alg_bytes_per_step": int(step_alg),
                          "achieved_GBps": round(step_alg / (ms_per_step * 1e-3) / 1e9, 2),
                          "frac": round(step_alg / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 5)},
        "stages": per_stage,
        "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 4), "p99": round(float(np.percentile(lat, 99)), 4)},
        "cpu_baseline": cpu,
        "speedup_vs_cpu": round(value / cpu["value"], 1) if cpu else None,
        "parity": parity,
        "batched": batched,
        "served": served,
        "stress": stress,
        "calibration_copy_bytes": copy_bytes,
        "towers_standing": standing,
        "max_penetration": round(float(st.max_penetration), 5),
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
