"""Continuous batching of simulation rollouts on one GPU world handle.

A request is a world (a scene array) to be stepped S times.  All resident
worlds are sub-worlds of one device world and advance together, one
`bepu_step` per tick.  Between ticks, finished worlds are read back and
removed (their body ranges and ids are freed) and waiting requests are
admitted while the body capacity and the HBM budget allow, so new work joins
a batch that is already running on the device.  Every world's result is
bit-identical to stepping it alone: sub-worlds never pair with each other,
their colouring uses world-local ids, and an inserted world never inherits
the contact history of a previous occupant of its ids.

The loop is pipelined: a tick queues its step, queues the readback of the
worlds that finish with it (and frees their ranges, in stream order after
that readback), and then collects the readback queued one tick earlier, so
the device always has the next step queued while the host reads results and
admits new requests.  Inserts copy through pinned memory without waiting
for the device either.

There is no reference counterpart: the reference steps one Space per
Update call (Space.cs:527-531); each request here would be its own Space.
"""
import collections
import time

import numpy as np

from .space import Space

# Device bytes per body slot of a world handle: body arrays (~250 B), and the
# pair, hash, island, solver and general-convex arrays sized 4 pairs per body
# (~4.4 KB).
BYTES_PER_BODY = 4700


class Request:
    __slots__ = ("rid", "scene", "steps", "arrival", "admitted", "finished", "done", "world", "first", "result",
                 "tick_latencies", "n_dyn", "failed")

    def __init__(self, rid, scene, steps, arrival):
        self.rid, self.scene, self.steps, self.arrival = rid, scene, int(steps), arrival
        self.n_dyn = int((scene["mass"] > 0).sum())
        self.admitted = self.finished = None
        self.done = 0
        self.world = self.first = None
        self.result = None
        self.failed = False
        self.tick_latencies = []


class Server:
    """Admission queue + resident batch.  `capacity` bodies are allocated on
    the device up front; `hbm_budget` (bytes) caps admission as well.
    `hull_shapes`: point clouds of the hull shapes requests refer to."""

    def __init__(self, capacity, settings=None, hbm_budget=200e9, clock=time.perf_counter, hull_shapes=None):
        self.capacity = int(min(capacity, hbm_budget // BYTES_PER_BODY))
        self.space = Space(capacity=self.capacity, settings=settings)
        if hull_shapes is not None:
            # Hull bodies of the requests name these shapes by index.
            self.space.register_hulls(hull_shapes)
        self.queue = collections.deque()
        self.pending = collections.deque()  # (readback request, finished requests) not yet collected
        self.resident = {}
        self.completed = []
        self.resident_bodies = 0
        self.clock = clock
        self.ticks = 0
        self.body_steps = 0
        self._next = 0

    def submit(self, scene, steps, arrival=None):
        r = Request(self._next, scene, steps, self.clock() if arrival is None else arrival)
        self._next += 1
        self.queue.append(r)
        return r

    def admit(self):
        """Admits queued requests in arrival order while they fit (one
        insertion call for all of them)."""
        batch, room = [], self.capacity - self.resident_bodies
        for r in self.queue:
            if len(r.scene) > room:
                break
            batch.append(r)
            room -= len(r.scene)
        if not batch:
            return
        placed = self.space.insert_worlds([r.scene for r in batch])
        now = self.clock()
        for r, (wid, first) in zip(batch, placed):
            self.queue.popleft()
            r.world, r.first = wid, first
            r.admitted = now
            self.resident[r.rid] = r
            self.resident_bodies += len(r.scene)

    def tick(self):
        """Queues one step of every resident world and the readback of those
        it finishes (one readback for all of them, also queued when none
        finishes: it marks the step's end), frees their ranges, then collects
        the readbacks of earlier ticks: all of them once nothing is resident."""
        if not self.resident and not self.pending:
            return 0.0
        t0 = self.clock()
        stepped = list(self.resident.values())
        if stepped:
            self.space.Update(1)
            self.ticks += 1
            done = []
            for r in stepped:
                r.done += 1
                self.body_steps += r.n_dyn
                if r.done >= r.steps:
                    done.append(r)
            req = self.space.request_motion_states_ranges([r.first for r in done], [len(r.scene) for r in done])
            if done:
                self.space.remove_worlds([r.world for r in done])
                for r in done:
                    del self.resident[r.rid]
                    self.resident_bodies -= len(r.scene)
            self.pending.append((req, done))
        while len(self.pending) > (1 if self.resident else 0):
            req, done = self.pending.popleft()
            states = self.space.collect_motion_states(req)
            now = self.clock()
            for r, ms in zip(done, states):
                r.result = ms
                r.finished = now
                # A world evicted by the non-finite guard is quarantined:
                # it completes with its error instead of a result.
                r.failed = not bool(np.isfinite(ms).all())
                self.completed.append(r)
        dt = self.clock() - t0
        for r in stepped:
            r.tick_latencies.append(dt)
        return dt

    def close(self):
        self.space.close()


def run_poisson(server, rate, duration, make_request, seed=0, drain=True, max_wall=None):
    """Arrivals of a Poisson process (`rate` per second of wall time) for
    `duration` seconds, served tick by tick.  `make_request(rng)` returns
    (scene, steps).  The clients build their requests before the clock
    starts (a request arrives with its scene); each is submitted at its
    arrival time.  Returns a summary dict."""
    rng = np.random.default_rng(seed)
    arrivals = []
    t = rng.exponential(1.0 / rate)
    while t <= duration:
        scene, steps = make_request(rng)
        arrivals.append((t, scene, steps))
        t += rng.exponential(1.0 / rate)
    offered = sum(int((sc["mass"] > 0).sum()) * st for _, sc, st in arrivals)
    t0 = time.perf_counter()
    k = 0
    waiting_max = 0
    trace = []  # (seconds since start, body-steps so far) after every tick
    while True:
        now = time.perf_counter() - t0
        while k < len(arrivals) and arrivals[k][0] <= now:
            at, scene, steps = arrivals[k]
            server.submit(scene, steps, arrival=t0 + at)
            k += 1
        waiting_max = max(waiting_max, len(server.queue))
        server.admit()
        if server.resident or server.pending:
            server.tick()
            trace.append((time.perf_counter() - t0, server.body_steps))
        if now > duration and (not drain or (not server.resident and not server.queue and not server.pending)):
            break
        if max_wall is not None and now > max_wall:
            break
    wall = time.perf_counter() - t0
    # The rate over the second half of the arrival window: the resident set
    # has built up by then (requests run up to 600 steps), and arrivals go on.
    half = [(t, b) for t, b in trace if duration / 2 <= t <= duration]
    steady = (half[-1][1] - half[0][1]) / (half[-1][0] - half[0][0]) if len(half) > 1 else None
    step_lat = np.array([x for r in server.completed for x in r.tick_latencies]) if server.completed else np.zeros(1)
    return {
        "wall_s": round(wall, 3),
        "requests_completed": len(server.completed),
        "requests_waiting": len(server.queue),
        "requests_waiting_max": waiting_max,
        "requests_offered": len(arrivals),
        "ticks": server.ticks,
        "body_steps": server.body_steps,
        "value": round(server.body_steps / wall, 1),
        "steady_value": round(steady, 1) if steady else None,
        "offered_body_steps_per_s": round(offered / duration, 1),
        "step_latency_ms": {"p50": round(1e3 * float(np.percentile(step_lat, 50)), 4),
                            "p99": round(1e3 * float(np.percentile(step_lat, 99)), 4)},
        "request_latency_s": {
            "p50": round(float(np.percentile([r.finished - r.arrival for r in server.completed], 50)), 4)
            if server.completed else None,
            "p99": round(float(np.percentile([r.finished - r.arrival for r in server.completed], 99)), 4)
            if server.completed else None},
    }
