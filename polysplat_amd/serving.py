"""Serving loop: resident scenes, camera requests of mixed resolutions,
memory-based admission and continuous batching over shared launches.

The reference renders one frame per Python call and blocks on the pair
count between preprocess and sort (pybind.cpp:1122-1125); it has no serving
loop.  Here every frame is one call of ``ops.forward_frame`` whose stages
pass the pair count on the device (no host round trip inside a frame), and
the loop keeps several groups of frames in flight on separate HIP streams:

* requests arrive on the host clock (Poisson process, orbit poses,
  resolutions drawn uniformly);
* a request is admitted when the bytes its keys, sort space, tile ranges,
  feature records and image need fit the key-memory budget next to the
  batches already on the GPU;
* whenever a stream slot is free, all admitted-and-waiting requests up to
  a frame cap and an estimated-GPU-time cap form the next group (first in,
  first out), so requests that arrive while others render start as soon
  as a slot frees instead of waiting for a fixed batch to fill;
* a frame whose pairs exceed its estimate is detected from its device
  count (read back asynchronously) and re-queued at the front with the
  measured count.

Latency is arrival -> image ready: a frame's completion is the GPU
timestamp of an event recorded after its own render (the pipelined
executor records one per frame, so a frame is not charged for the frames
queued behind it in its group), mapped onto the host clock through a start
event recorded on an idle device.

The scheduling policy (``Scheduler``) is independent of the GPU, so it is
tested on the CPU with a simulated executor.
"""
from __future__ import annotations

import collections
import math
import time
from dataclasses import dataclass, field

import numpy as np

from . import scenes as S

RESOLUTIONS = ((1280, 720), (1920, 1080), (3840, 2160), (7680, 4320))
TILE_ORDER = 0x200  # PS_FORWARD_TILE_ORDER: OR-ed into a forward variant
SCENE_SIZES = (1_000_000, 3_000_000, 10_000_000, 30_000_000)
BOX = 50.0
ORBIT_RADIUS = 1.5 * BOX
SORT_FIXED_BYTES = 2 * 1024 * 256 * 4 + 4096  # per-block digit counts of the radix sort


@dataclass
class Request:
    rid: int
    scene: int
    width: int
    height: int
    azimuth: float
    arrival_s: float
    bg: tuple = (0.0, 0.0, 0.0)
    # filled in by the loop
    batch: int = -1
    attempts: int = 0
    pairs: int = -1
    done_ms: float = math.nan

    @property
    def key(self):
        return (self.scene, self.width, self.height)

    @property
    def tiles(self):
        return ((self.width + 15) // 16) * ((self.height + 15) // 16)

    @property
    def latency_ms(self):
        return self.done_ms - 1e3 * self.arrival_s


def orbit_camera(width: int, height: int, azimuth: float, radius: float = ORBIT_RADIUS,
                 fov_deg: float = 60.0) -> S.CameraSpec:
    """Pinhole camera on a horizontal circle around the origin, looking at it.

    The rotation is the camera-to-world matrix given row by row, the layout
    of cameras.json that the reference passes to its bindings
    (pybind.cpp:140-143): its third column (the camera's +z) points from the
    camera to the origin.  Azimuth 0 is the benchmark camera at (0, 0, -r).
    """
    import torch

    s, c = math.sin(azimuth), math.cos(azimuth)
    pos = torch.tensor([radius * s, 0.0, -radius * c], dtype=torch.float32)
    # rotation about +y by -azimuth: maps (0, 0, 1) to (-sin a, 0, cos a)
    rot = torch.tensor([[c, 0.0, -s], [0.0, 1.0, 0.0], [s, 0.0, c]], dtype=torch.float32)
    fx = width / (2.0 * math.tan(math.radians(fov_deg) / 2.0))
    return S.CameraSpec(width, height, pos, rot, fx, fx)


def poisson_requests(n: int, rate: float, n_scenes: int, resolutions=RESOLUTIONS,
                     seed: int = 7) -> list[Request]:
    """n requests with exponential inter-arrival times (rate per second),
    uniform scene, uniform resolution and uniform orbit azimuth."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, n))
    sc = rng.integers(0, n_scenes, n)
    rs = rng.integers(0, len(resolutions), n)
    az = rng.uniform(0.0, 2.0 * math.pi, n)
    return [Request(i, int(sc[i]), resolutions[rs[i]][0], resolutions[rs[i]][1], float(az[i]),
                    float(t[i])) for i in range(n)]


def frame_bytes(P: int, width: int, height: int, pairs: int, feat_bytes: int = 24) -> int:
    """Device bytes one frame needs inside a batch: keys and values (unsorted
    and sorted), the sort's ping-pong copy, tile ranges, per-Gaussian
    feature records, the preprocess counts and the image."""
    tiles = ((width + 15) // 16) * ((height + 15) // 16)
    return (pairs * (12 + 12 + 12) + tiles * 8 + P * (feat_bytes + 4 + 8) + 3 * width * height
            + 4096)


class Estimates:
    """Pairs and GPU milliseconds per (scene, width, height), learned from
    completed frames; pairs keep the largest count seen."""

    def __init__(self, margin: float = 1.08, slack: int = 65536):
        self.pairs: dict = {}
        self.ms: dict = {}
        self.margin = margin
        self.slack = slack

    def observe(self, key, pairs: int | None = None, ms: float | None = None):
        if pairs is not None and pairs >= 0:
            self.pairs[key] = max(self.pairs.get(key, 0), int(pairs))
        if ms is not None:
            old = self.ms.get(key)
            self.ms[key] = ms if old is None else 0.8 * old + 0.2 * ms

    def capacity(self, key) -> int:
        return int(self.pairs.get(key, 0) * self.margin) + self.slack

    def cost_ms(self, key) -> float:
        return self.ms.get(key, 1.0)


class Scheduler:
    """Admission and batch formation (no GPU).

    ``budget`` bounds the bytes of all batches in flight.  ``form_batch``
    takes waiting requests in queue order while the batch stays within
    ``max_frames`` and ``max_batch_ms`` of estimated GPU time (a single
    request is always allowed) and the budget has room; the first request
    that does not fit ends the batch.  A request that could never fit the
    whole budget is rejected.

    Queue order: ``"fifo"`` is arrival order; ``"slack"`` takes the waiting
    requests by increasing estimated cost minus ``slack_k`` times the time
    they have waited (ms), so short frames go ahead of long ones that have
    not waited long.  Requests put back after an overflow stay first either
    way.
    """

    def __init__(self, scene_sizes, budget: int, est: Estimates, max_frames: int = 8,
                 max_batch_ms: float = 6.0, feat_bytes: int = 24, order: str = "fifo",
                 slack_k: float = 1.0):
        if order not in ("fifo", "slack"):
            raise ValueError(f"unknown queue order {order!r}")
        self.order = order
        self.slack_k = float(slack_k)
        self.P = list(scene_sizes)
        self.budget = int(budget)
        self.est = est
        self.max_frames = max_frames
        self.max_batch_ms = max_batch_ms
        self.feat_bytes = feat_bytes
        self.waiting: collections.deque = collections.deque()
        self.in_flight_bytes = 0
        self.rejected: list = []
        self.memory_waits = 0  # batches ended early because the budget was full

    def request_bytes(self, r: Request) -> int:
        return frame_bytes(self.P[r.scene], r.width, r.height, self.est.capacity(r.key),
                           self.feat_bytes)

    def arrive(self, r: Request):
        if self.request_bytes(r) + SORT_FIXED_BYTES > self.budget:
            self.rejected.append(r)
            return
        self.waiting.append(r)

    def requeue_front(self, reqs):
        for r in reversed(list(reqs)):
            r.attempts += 1
            self.waiting.appendleft(r)

    def _reorder(self, now_s: float):
        retried = [r for r in self.waiting if r.attempts > 0]
        fresh = [r for r in self.waiting if r.attempts == 0]
        fresh.sort(key=lambda r: (self.est.cost_ms(r.key)
                                  - self.slack_k * 1e3 * (now_s - r.arrival_s), r.arrival_s))
        self.waiting = collections.deque(retried + fresh)

    def form_batch(self, now_s: float | None = None):
        """Returns (requests, bytes) or ([], 0) when nothing can start now.
        ``now_s`` (the host clock of the arrivals) is needed by the slack
        order."""
        if self.order == "slack" and now_s is not None and len(self.waiting) > 1:
            self._reorder(now_s)
        batch, nbytes, ms = [], SORT_FIXED_BYTES, 0.0
        while self.waiting and len(batch) < self.max_frames:
            r = self.waiting[0]
            b = self.request_bytes(r)
            c = self.est.cost_ms(r.key)
            if batch and ms + c > self.max_batch_ms:
                break
            if self.in_flight_bytes + nbytes + b > self.budget:
                self.memory_waits += 1
                break
            self.waiting.popleft()
            batch.append(r)
            nbytes += b
            ms += c
        if not batch:
            return [], 0
        self.in_flight_bytes += nbytes
        return batch, nbytes

    def release(self, nbytes: int):
        self.in_flight_bytes -= nbytes
        assert self.in_flight_bytes >= 0


@dataclass
class InFlight:
    bid: int
    reqs: list
    nbytes: int
    capacities: list
    slot: int
    done: object = None            # torch.cuda.Event after the group's last frame
    host_counts: object = None     # pinned int32 [frames]: pairs each frame needed
    frame_events: list = field(default_factory=list)  # one event after each frame
    frame_ms: list = field(default_factory=list)      # each frame's completion, ms
    buffers: list = field(default_factory=list)
    outputs: list = field(default_factory=list)


class GpuExecutor:
    """(Not used by the benchmark or the tests: a GPU run of this executor
    faulted the device once and the cause has not been found; the
    benchmark serves with PipelinedExecutor.)

    Runs a group of admitted requests on one stream slot: each frame is
    one ``ops.forward_frame`` call (preprocess, sort with tile ranges,
    persistent render; the pair count stays on the device), launched back to
    back on the slot's HIP stream, so frames of different slots overlap on
    the device.  All streams keep the default priority.  A frame's pair
    count is copied to pinned host memory without waiting; a frame whose
    pairs exceeded its key buffers is reported for a retry."""

    def __init__(self, scenes_dev, device="cuda:0", slots: int = 2, feat_prec: str = "fp16",
                 radius: float = ORBIT_RADIUS):
        import torch

        from . import _C

        self.torch = torch
        self.ops = _C.ops
        self.dev = torch.device(device)
        self.scenes = scenes_dev  # list of dicts: pos, shs, opac, cov (device), P
        self.slots = [torch.cuda.Stream(device=self.dev) for _ in range(slots)]
        self.busy = [None] * slots
        self.feat_prec = {"fp32": 0, "fp16": 1, "bf16": 2}[feat_prec]
        self.radius = radius
        self.start = torch.cuda.Event(enable_timing=True)
        self.t0 = None
        self._bid = 0
        self._empty = torch.empty(0, dtype=torch.float32, device=self.dev)

    def begin(self):
        torch = self.torch
        torch.cuda.synchronize(self.dev)
        self.start.record(self.slots[0])
        torch.cuda.synchronize(self.dev)
        self.t0 = time.perf_counter()

    def now_s(self) -> float:
        return time.perf_counter() - self.t0

    def free_slot(self):
        for i, b in enumerate(self.busy):
            if b is None:
                return i
        return None

    def _frame(self, r, cap):
        """Allocates one frame's buffers on the current stream and enqueues it."""
        torch = self.torch
        dev = self.dev
        sc = self.scenes[r.scene]
        cap = max(int(cap), 1)
        ku = torch.empty(cap, dtype=torch.int64, device=dev)
        vu = torch.empty(cap, dtype=torch.int32, device=dev)
        ks = torch.empty(cap, dtype=torch.int64, device=dev)
        vs = torch.empty(cap, dtype=torch.int32, device=dev)
        space = torch.empty(self.ops.get_sort_buffer_size(cap), dtype=torch.int8, device=dev)
        curr = torch.empty(1, dtype=torch.int32, device=dev)
        ranges = torch.empty((r.tiles, 2), dtype=torch.int32, device=dev)
        ws = torch.empty(self.ops.forward_workspace_size(sc["P"]), dtype=torch.uint8, device=dev)
        ws[:2048].zero_()  # the 2 KiB tile-queue area (per-XCD counters) starts at zero
        rec = torch.empty(24 * max(sc["P"], 1), dtype=torch.uint8, device=dev)
        out = torch.empty((r.height, r.width, 3), dtype=torch.uint8, device=dev)
        cam = orbit_camera(r.width, r.height, r.azimuth, self.radius)
        e = self._empty
        self.ops.forward_frame(sc["pos"], sc["shs"], sc["opac"], sc["cov"], r.width, r.height,
                               cam.position, cam.rotation, cam.focal_x, cam.focal_y, cam.zFar,
                               cam.zNear, 1, self.feat_prec, curr, space, ku, vu, ks, vs, e, e, e,
                               rec, ranges, torch.tensor(r.bg, dtype=torch.float32), out, ws,
                               False)
        return curr, out, [ku, vu, ks, vs, space, curr, ranges, ws, rec]

    def launch(self, reqs, nbytes, capacities, slot) -> InFlight:
        torch = self.torch
        s = self.slots[slot]
        fl = InFlight(self._bid, reqs, nbytes, list(capacities), slot)
        self._bid += 1
        host = torch.empty(len(reqs), dtype=torch.int32, pin_memory=True)
        with torch.cuda.stream(s):
            for i, (r, cap) in enumerate(zip(reqs, capacities)):
                curr, out, bufs = self._frame(r, cap)
                host[i:i + 1].copy_(curr, non_blocking=True)
                fl.outputs.append(out)
                fl.buffers.append(bufs)
            fl.done = torch.cuda.Event(enable_timing=True)
            fl.done.record(s)
        fl.host_counts = host
        self.busy[slot] = fl
        return fl

    def poll(self):
        """Completed groups (their slots are freed): (group, done_ms, counts)."""
        out = []
        for i, fl in enumerate(self.busy):
            if fl is not None and fl.done.query():
                done_ms = self.start.elapsed_time(fl.done)
                self.busy[i] = None
                out.append((fl, done_ms, fl.host_counts.tolist()))
        return out

    def idle(self) -> bool:
        return all(b is None for b in self.busy)


class SequentialExecutor:
    """Runs admitted groups on one stream, frame after frame, each frame one
    synchronous ``ops.forward_frame`` call (its pair count is read back
    before the call returns).  Buffers are allocated zeroed and reused:
    key, value and sort buffers are shared by all frames and grow when a
    frame needs more pairs (the frame is then rendered again), the feature
    records and the forward workspace are kept per scene, tile ranges per
    resolution.  A group is complete when ``launch`` returns; ``poll`` hands
    it back with its completion time on the host clock."""

    slots_n = 1

    def __init__(self, scenes_dev, device="cuda:0", feat_prec: str = "fp16",
                 radius: float = ORBIT_RADIUS, margin: float = 1.05):
        import torch

        from . import _C

        self.torch = torch
        self.ops = _C.ops
        self.dev = torch.device(device)
        self.scenes = scenes_dev
        self.feat_prec = {"fp32": 0, "fp16": 1, "bf16": 2}[feat_prec]
        self.radius = radius
        self.margin = margin
        self.busy = [None]
        self.t0 = None
        self._bid = 0
        self.cap = 0
        self.keys = None
        self.per_scene = {}
        self.per_res = {}
        self.grows = 0
        self._empty = torch.empty(0, dtype=torch.float32, device=self.dev)

    def _zeros(self, n, dtype):
        return self.torch.zeros(n, dtype=dtype, device=self.dev)

    def _ensure_keys(self, cap: int):
        if cap <= self.cap:
            return
        torch = self.torch
        cap = int(cap)
        self.keys = None  # release the old buffers first
        self.keys = [self._zeros(cap, torch.int64), self._zeros(cap, torch.int32),
                     self._zeros(cap, torch.int64), self._zeros(cap, torch.int32),
                     self._zeros(self.ops.get_sort_buffer_size(cap), torch.int8),
                     self._zeros(1, torch.int32)]
        self.cap = cap

    def _scene_bufs(self, s: int):
        if s not in self.per_scene:
            torch = self.torch
            P = self.scenes[s]["P"]
            self.per_scene[s] = (self._zeros(24 * max(P, 1), torch.uint8),
                                 self._zeros(self.ops.forward_workspace_size(P), torch.uint8))
        return self.per_scene[s]

    def _ranges(self, r):
        key = (r.width, r.height)
        if key not in self.per_res:
            self.per_res[key] = self.torch.zeros((r.tiles, 2), dtype=self.torch.int32,
                                                 device=self.dev)
        return self.per_res[key]

    def begin(self):
        self.torch.cuda.synchronize(self.dev)
        self.t0 = time.perf_counter()

    def now_s(self) -> float:
        return time.perf_counter() - self.t0

    def free_slot(self):
        return 0 if self.busy[0] is None else None

    def render(self, r, cap: int):
        """One frame; returns (pairs, image tensor).  Grows the key buffers
        and renders again when the frame needs more than their capacity."""
        torch = self.torch
        self._ensure_keys(max(int(cap), 1))
        sc = self.scenes[r.scene]
        rec, ws = self._scene_bufs(r.scene)
        ranges = self._ranges(r)
        cam = orbit_camera(r.width, r.height, r.azimuth, self.radius)
        bg = torch.tensor(r.bg, dtype=torch.float32)
        e = self._empty
        while True:
            ku, vu, ks, vs, space, curr = self.keys
            out = torch.empty((r.height, r.width, 3), dtype=torch.uint8, device=self.dev)
            try:
                n = self.ops.forward_frame(sc["pos"], sc["shs"], sc["opac"], sc["cov"], r.width,
                                           r.height, cam.position, cam.rotation, cam.focal_x,
                                           cam.focal_y, cam.zFar, cam.zNear, 1, self.feat_prec,
                                           curr, space, ku, vu, ks, vs, e, e, e, rec, ranges, bg,
                                           out, ws, True)
                return n, out
            except RuntimeError as err:
                if "Too many k-v pairs" not in str(err):
                    raise
                n = int(curr.item())  # the count the frame needed
                self.grows += 1
                self._ensure_keys(int(n * self.margin) + 4096)

    def launch(self, reqs, nbytes, capacities, slot) -> InFlight:
        fl = InFlight(self._bid, reqs, nbytes, list(capacities), slot)
        self._bid += 1
        counts = []
        for i, r in enumerate(reqs):
            n, out = self.render(r, capacities[i])
            counts.append(n)
            fl.outputs.append(out)
            fl.capacities[i] = max(fl.capacities[i], n)  # rendered: never a retry
        fl.host_counts = counts
        fl.done = 1e3 * self.now_s()
        self.busy[0] = fl
        return fl

    def poll(self):
        fl = self.busy[0]
        if fl is None:
            return []
        self.busy[0] = None
        return [(fl, fl.done, list(fl.host_counts))]

    def idle(self) -> bool:
        return self.busy[0] is None


class PipelinedExecutor(SequentialExecutor):
    """Frames on one stream, enqueued back to back: each frame is one
    ``ops.forward_frame_post`` call, whose pair count is read back later
    (``ops.forward_wait``) from a pinned slot the library owns, so the GPU
    runs the next admitted frame while the host is still handling the last
    one, instead of idling between a frame's end and the next frame's
    launches.  Up to ``slots`` groups are in flight.  Buffers as in
    SequentialExecutor: frames on one stream run in order, so consecutive
    frames can share the key, sort and range buffers and a scene's records
    and workspace; only each frame's image is its own.  Key buffers grow
    before a frame whose estimate exceeds them; a frame whose pairs still
    did not fit is reported with its count for a retry.

    With ``graphs`` each (scene, resolution) pipeline is captured once as a
    HIP graph (``ops.FrameGraph``) over these shared buffers and replayed
    per request with its pose, background and image: two launches per frame
    instead of about twenty.  Growing the key buffers drops the graphs
    (they point at the old ones); they are captured again on next use."""

    def __init__(self, scenes_dev, device="cuda:0", feat_prec: str = "fp16",
                 radius: float = ORBIT_RADIUS, margin: float = 1.05, slots: int = 3,
                 graphs: bool = False, tile_order_pairs: int = 20_000_000,
                 tile_order_per_tile: int = 500):
        super().__init__(scenes_dev, device=device, feat_prec=feat_prec, radius=radius,
                         margin=margin)
        torch = self.torch
        self.slots_n = slots
        self.inflight: collections.deque = collections.deque()
        self.stream = torch.cuda.current_stream(self.dev)
        self.start = torch.cuda.Event(enable_timing=True)
        self.use_graphs = graphs
        self.graphs = {}
        self.captures = 0
        # Frames expected to hold at least this many pairs (and that many per
        # tile) render their tiles heaviest first (PS_FORWARD_TILE_ORDER): on
        # MI355X that takes 0.07 ms and saves about 0.4 ms on a 30M-Gaussian
        # 4K frame; tile_order_pairs 0: never.
        self.tile_order_pairs = tile_order_pairs
        self.tile_order_per_tile = tile_order_per_tile
        self.ordered_frames = 0

    def _variant(self, r, cap: int) -> int:
        if (self.tile_order_pairs > 0 and cap >= self.tile_order_pairs
                and cap >= self.tile_order_per_tile * r.tiles):
            self.ordered_frames += 1
            return 1 | TILE_ORDER
        return 1

    def _ensure_keys(self, cap: int):
        if cap > self.cap and getattr(self, "graphs", None):
            # the graphs point at the buffers about to be replaced; the
            # stream is ordered, so frames already enqueued finish first
            self.torch.cuda.current_stream(self.dev).synchronize()
            self.graphs.clear()
        super()._ensure_keys(cap)

    def _graph(self, r, cam, rec, ws, ranges, variant):
        key = (r.scene, r.width, r.height, variant)
        g = self.graphs.get(key)
        if g is None:
            sc = self.scenes[r.scene]
            ku, vu, ks, vs, space, curr = self.keys
            e = self._empty
            g = self.ops.FrameGraph(sc["pos"], sc["shs"], sc["opac"], sc["cov"], r.width,
                                    r.height, cam.position, cam.rotation, cam.focal_x,
                                    cam.focal_y, cam.zFar, cam.zNear, variant, self.feat_prec,
                                    curr, space, ku, vu, ks, vs, e, e, e, rec, ranges, ws)
            self.graphs[key] = g
            self.captures += 1
        return g

    def begin(self):
        torch = self.torch
        torch.cuda.synchronize(self.dev)
        self.start.record(self.stream)
        torch.cuda.synchronize(self.dev)
        self.t0 = time.perf_counter()

    def free_slot(self):
        return len(self.inflight) if len(self.inflight) < self.slots_n else None

    def _post(self, r, cap: int):
        torch = self.torch
        self._ensure_keys(max(int(cap), 1))
        sc = self.scenes[r.scene]
        rec, ws = self._scene_bufs(r.scene)
        ranges = self._ranges(r)
        cam = orbit_camera(r.width, r.height, r.azimuth, self.radius)
        bg = torch.tensor(r.bg, dtype=torch.float32)
        e = self._empty
        ku, vu, ks, vs, space, curr = self.keys
        out = torch.empty((r.height, r.width, 3), dtype=torch.uint8, device=self.dev)
        variant = self._variant(r, cap)
        if self.use_graphs:
            g = self._graph(r, cam, rec, ws, ranges, variant)
            t = g.launch(cam.position, cam.rotation, cam.focal_x, cam.focal_y, cam.zFar,
                         cam.zNear, bg, out, 2)
            return t, out
        t = self.ops.forward_frame_post(sc["pos"], sc["shs"], sc["opac"], sc["cov"], r.width,
                                        r.height, cam.position, cam.rotation, cam.focal_x,
                                        cam.focal_y, cam.zFar, cam.zNear, variant,
                                        self.feat_prec, curr,
                                        space, ku, vu, ks, vs, e, e, e, rec, ranges, bg, out, ws)
        return t, out

    def launch(self, reqs, nbytes, capacities, slot) -> InFlight:
        torch = self.torch
        fl = InFlight(self._bid, reqs, nbytes, list(capacities), slot)
        self._bid += 1
        tickets = []
        for i, r in enumerate(reqs):
            t, out = self._post(r, capacities[i])
            tickets.append(t)
            fl.outputs.append(out)
            fl.capacities[i] = self.cap  # the room the frame had
            # A request's image is ready when its own frame's kernels are
            # done, not when the rest of its group is.
            ev = torch.cuda.Event(enable_timing=True)
            ev.record(self.stream)
            fl.frame_events.append(ev)
        fl.host_counts = tickets
        fl.done = fl.frame_events[-1] if fl.frame_events else torch.cuda.Event(enable_timing=True)
        if not fl.frame_events:
            fl.done.record(self.stream)
        self.inflight.append(fl)
        return fl

    def poll(self):
        """Completed groups, oldest first: (group, done_ms, counts).  A frame
        that overflowed reports its count, above its capacity."""
        out = []
        while self.inflight and self.inflight[0].done.query():
            fl = self.inflight.popleft()
            counts = []
            for t, cap in zip(fl.host_counts, fl.capacities):
                n, overflowed = self.ops.forward_wait(t)
                counts.append(n if not overflowed else max(n, cap + 1))
                if overflowed:
                    self.grows += 1
                    self._ensure_keys(int(n * self.margin) + 4096)
            fl.frame_ms = [self.start.elapsed_time(ev) for ev in fl.frame_events]
            out.append((fl, self.start.elapsed_time(fl.done), counts))
        return out

    def idle(self) -> bool:
        return not self.inflight


class ServingLoop:
    """Drives requests through a Scheduler and an executor in real time."""

    def __init__(self, scheduler: Scheduler, executor, keep_outputs: bool = False):
        self.sched = scheduler
        self.ex = executor
        self.keep_outputs = keep_outputs
        self.outputs: dict = {}
        self.batches: list = []  # (bid, [rid...], launch_s, done_ms)
        self.overflows = 0

    def run(self, requests, spin_s: float = 2e-5):
        reqs = sorted(requests, key=lambda r: r.arrival_s)
        pending = collections.deque(reqs)
        done: list = []
        self.ex.begin()
        n_total = len(reqs)
        while len(done) + len(self.sched.rejected) < n_total:
            now = self.ex.now_s()
            while pending and pending[0].arrival_s <= now:
                self.sched.arrive(pending.popleft())
            progressed = False
            for fl, done_ms, counts in self.ex.poll():
                progressed = True
                self.sched.release(fl.nbytes)
                retry = []
                frame_ms = getattr(fl, "frame_ms", None) or [done_ms] * len(fl.reqs)
                for r, c, cap, img, f_ms in zip(fl.reqs, counts, fl.capacities, fl.outputs,
                                                frame_ms):
                    self.sched.est.observe(r.key, pairs=c)
                    if c > cap:
                        retry.append(r)  # not rendered: its pairs did not fit
                        continue
                    r.pairs = c
                    r.done_ms = f_ms
                    r.batch = fl.bid
                    if self.keep_outputs:
                        self.outputs[r.rid] = img
                    done.append(r)
                if retry:
                    self.overflows += 1
                    self.sched.requeue_front(retry)
                self.batches.append((fl.bid, [r.rid for r in fl.reqs if r not in retry],
                                     done_ms))
            while True:
                slot = self.ex.free_slot()
                if slot is None:
                    break
                batch, nbytes = self.sched.form_batch(self.ex.now_s())
                if not batch:
                    break
                caps = [self.sched.est.capacity(r.key) for r in batch]
                self.ex.launch(batch, nbytes, caps, slot)
                progressed = True
            if not progressed:
                if not pending and not self.sched.waiting and self.ex.idle():
                    break
                time.sleep(spin_s)
        return done


def summarize(done, rejected, wall_s: float, deadline_ms: float = 1000.0 / 60.0,
              batches=None) -> dict:
    lat = np.array([r.latency_ms for r in done], dtype=np.float64)
    res = {
        "requests": len(done) + len(rejected),
        "completed": len(done),
        "rejected": len(rejected),
        "p50_ms": float(np.percentile(lat, 50)) if len(lat) else math.nan,
        "p90_ms": float(np.percentile(lat, 90)) if len(lat) else math.nan,
        "p99_ms": float(np.percentile(lat, 99)) if len(lat) else math.nan,
        "max_ms": float(lat.max()) if len(lat) else math.nan,
        "deadline_ms": deadline_ms,
        "deadline_miss_frac": float((lat > deadline_ms).mean()) if len(lat) else math.nan,
        "frames_per_s": len(done) / wall_s if wall_s > 0 else math.nan,
        "wall_s": wall_s,
    }
    if batches:
        sizes = np.array([len(b[1]) for b in batches])
        res["batches"] = len(batches)
        res["mean_frames_per_batch"] = float(sizes.mean())
        res["max_frames_per_batch"] = int(sizes.max())
    return res


def _ops():
    from . import _C

    return _C.ops


def make_executor(scenes_dev, device="cuda:0", **kw):
    """The executor the benchmark and the GPU tests serve with."""
    return PipelinedExecutor(scenes_dev, device=device, **kw)


def make_resident_scenes(sizes=SCENE_SIZES, device="cuda:0", seed0: int = 101, box: float = BOX,
                         mix=None):
    """Scenes generated on the device with the C2 distributions (uniform
    means in a `box` cube, log-scales N(-4.5, 0.5)) and stored with FP16 SH."""
    import torch

    mix = mix if mix is not None else S.CONFIGS["c2"]["mix"]
    out = []
    for i, P in enumerate(sizes):
        sc = S.make_scene(P, box, mix, seed=seed0 + i, device=device)
        out.append({"P": P, "pos": sc.position, "shs": sc.shs.to(torch.float16).contiguous(),
                    "opac": sc.opacity, "cov": sc.cov3d})
        del sc
    return out


def calibrate(executor, est: Estimates, n_scenes: int, resolutions=RESOLUTIONS,
              azimuths=(0.0, math.pi / 4), repeats: int = 3):
    """Renders every (scene, resolution) alone at a few orbit poses: warms the
    kernels and seeds the pair and time estimates."""
    if isinstance(executor, SequentialExecutor):  # and PipelinedExecutor
        _calibrate_sequential(executor, est, n_scenes, resolutions, azimuths, repeats)
        if getattr(executor, "use_graphs", False):
            # capture every (scene, resolution) graph before serving starts
            torch = executor.torch
            for s in range(n_scenes):
                for (w, h) in resolutions:
                    r = Request(-1, s, w, h, azimuths[0], 0.0)
                    executor.launch([r], 0, [executor.cap], 0)
                    torch.cuda.synchronize(executor.dev)
                    executor.poll()
        return
    torch = executor.torch
    executor.begin()
    for s in range(n_scenes):
        for (w, h) in resolutions:
            for a in azimuths:
                r = Request(-1, s, w, h, a, 0.0)
                cap = 1 << 20
                for _ in range(4):
                    executor.launch([r], 0, [cap], 0)
                    torch.cuda.synchronize(executor.dev)
                    (_, _, counts), = executor.poll()
                    if counts[0] <= cap:
                        break
                    cap = int(counts[0] * 1.05) + 4096
                est.observe(r.key, pairs=counts[0])
                t = []
                for _ in range(repeats):
                    ev0 = torch.cuda.Event(enable_timing=True)
                    ev0.record(executor.slots[0])
                    executor.launch([r], 0, [cap], 0)
                    torch.cuda.synchronize(executor.dev)
                    (fl, _, _), = executor.poll()
                    t.append(ev0.elapsed_time(fl.done))
                est.observe(r.key, ms=float(np.median(t)))


def _calibrate_sequential(ex: SequentialExecutor, est: Estimates, n_scenes: int, resolutions,
                          azimuths, repeats: int):
    torch = ex.torch
    ex.begin()
    for s in range(n_scenes):
        for (w, h) in resolutions:
            for a in azimuths:
                r = Request(-1, s, w, h, a, 0.0)
                n, _ = ex.render(r, ex.cap)
                est.observe(r.key, pairs=n)
                t = []
                for _ in range(repeats):
                    ev0 = torch.cuda.Event(enable_timing=True)
                    ev1 = torch.cuda.Event(enable_timing=True)
                    ev0.record()
                    ex.render(r, ex.cap)
                    ev1.record()
                    torch.cuda.synchronize(ex.dev)
                    t.append(ev0.elapsed_time(ev1))
                est.observe(r.key, ms=float(np.median(t)))
