"""Serving policy on the CPU: admission by memory, continuous batching,
overflow retry and latency bookkeeping, driven by a simulated executor
(a single virtual GPU that runs batches back to back)."""
import math

import pytest

from polysplat_amd import serving as SV


class SimExecutor:
    """Virtual GPU: a batch starts when the device is free and takes the sum
    of its frames' costs; pair counts come from a fixed table."""

    def __init__(self, cost_ms, pairs, slots=2, clock=None):
        self.cost_ms = cost_ms
        self.pairs = pairs
        self.busy = [None] * slots
        self.gpu_free_ms = 0.0
        self.launched = []
        self.t = 0.0  # simulated host clock (s)

    def begin(self):
        self.t = 0.0

    def now_s(self):
        # every call advances the simulated clock a little: the loop polls
        self.t += 1e-4
        return self.t

    def free_slot(self):
        for i, b in enumerate(self.busy):
            if b is None:
                return i
        return None

    def launch(self, reqs, nbytes, capacities, slot):
        start = max(self.gpu_free_ms, 1e3 * self.t)
        end = start + sum(self.cost_ms[r.key] for r in reqs)
        self.gpu_free_ms = end
        fl = SV.InFlight(len(self.launched), reqs, nbytes, capacities, slot)
        fl.done = end
        fl.outputs = [f"img{r.rid}" for r in reqs]
        self.busy[slot] = fl
        self.launched.append((fl, self.t))
        return fl

    def poll(self):
        out = []
        for i, fl in enumerate(self.busy):
            if fl is not None and fl.done <= 1e3 * self.t:
                counts = [self.pairs[r.key] for r in fl.reqs]
                self.busy[i] = None
                out.append((fl, fl.done, counts))
        return out

    def idle(self):
        return all(b is None for b in self.busy)


def _setup(budget=1 << 34, max_frames=8, max_ms=6.0, est_pairs=None):
    sizes = (1000, 4000)
    res = ((64, 64), (128, 96))
    cost = {(s, w, h): 0.5 * (s + 1) * (w * h) / 4096 for s in range(2) for w, h in res}
    pairs = {(s, w, h): 700 * (s + 1) * (w * h) // 4096 for s in range(2) for w, h in res}
    est = SV.Estimates(margin=1.0, slack=0)
    for k in pairs:
        est.observe(k, pairs=pairs[k] if est_pairs is None else est_pairs, ms=cost[k])
    sched = SV.Scheduler(sizes, budget, est, max_frames=max_frames, max_batch_ms=max_ms)
    return sizes, res, cost, pairs, est, sched


def test_poisson_requests_are_seeded_and_ordered():
    a = SV.poisson_requests(500, 240.0, 4, seed=3)
    b = SV.poisson_requests(500, 240.0, 4, seed=3)
    assert [(r.scene, r.width, r.arrival_s) for r in a] == [(r.scene, r.width, r.arrival_s)
                                                            for r in b]
    t = [r.arrival_s for r in a]
    assert t == sorted(t)
    rate = len(t) / t[-1]
    assert 200 < rate < 290
    assert {r.scene for r in a} == {0, 1, 2, 3}
    assert {(r.width, r.height) for r in a} == set(SV.RESOLUTIONS)


def test_orbit_camera_looks_at_origin():
    import torch

    for az in (0.0, 0.7, math.pi, 4.0):
        cam = SV.orbit_camera(640, 480, az)
        c2w = cam.rotation  # rows of the camera-to-world matrix
        fwd = c2w[:, 2]  # camera +z in world coordinates
        to_origin = -cam.position / cam.position.norm()
        assert torch.allclose(fwd, to_origin, atol=1e-6)
        assert torch.allclose(c2w @ c2w.T, torch.eye(3), atol=1e-6)
        assert abs(float(torch.det(c2w)) - 1.0) < 1e-6
    base = SV.orbit_camera(800, 800, 0.0)
    assert base.position.tolist() == [0.0, 0.0, -75.0]
    assert torch.equal(base.rotation, torch.eye(3))


def test_admission_respects_the_memory_budget():
    sizes, res, cost, pairs, est, sched = _setup(max_ms=100.0)
    reqs = [SV.Request(i, 1, *res[1], 0.0, 0.0) for i in range(6)]
    one = sched.request_bytes(reqs[0])
    # room for exactly three requests' bytes plus the sort's fixed part
    sched.budget = SV.SORT_FIXED_BYTES + 3 * one
    for r in reqs:
        sched.arrive(r)
    b1, n1 = sched.form_batch()
    assert [r.rid for r in b1] == [0, 1, 2]
    assert sched.form_batch() == ([], 0)  # budget exhausted while b1 is in flight
    sched.release(n1)
    b2, n2 = sched.form_batch()
    assert [r.rid for r in b2] == [3, 4, 5]
    sched.release(n2)
    assert sched.in_flight_bytes == 0


def test_request_larger_than_budget_is_rejected():
    sizes, res, cost, pairs, est, sched = _setup(budget=10_000)
    r = SV.Request(0, 1, *res[1], 0.0, 0.0)
    sched.arrive(r)
    assert sched.rejected == [r] and not sched.waiting


def test_batches_are_capped_by_frames_and_estimated_time():
    sizes, res, cost, pairs, est, sched = _setup(max_frames=3, max_ms=100.0)
    for i in range(7):
        sched.arrive(SV.Request(i, 0, *res[0], 0.0, 0.0))
    assert [len(sched.form_batch()[0]) for _ in range(3)] == [3, 3, 1]
    sizes, res, cost, pairs, est, sched = _setup(max_frames=64, max_ms=2.0)
    for i in range(4):
        sched.arrive(SV.Request(i, 1, *res[1], 0.0, 0.0))  # 3 ms each: one per batch
    assert [len(sched.form_batch()[0]) for _ in range(4)] == [1, 1, 1, 1]


def test_slack_order_puts_short_frames_first_until_long_ones_have_waited():
    # costs: scene 1 at 128x96 is 3.0 ms, scene 0 at 64x64 is 0.5 ms
    sizes, res, cost, pairs, est, sched = _setup(max_frames=1)
    big = SV.Request(0, 1, *res[1], 0.0, arrival_s=0.0)
    small = SV.Request(1, 0, *res[0], 0.0, arrival_s=0.001)
    for order, now, first in (("fifo", 0.002, 0),
                              ("slack", 0.002, 1),    # 3.0 - 2 > 0.5 - 1
                              ("slack", 0.010, 1)):   # 3.0 - 10 > 0.5 - 9
        sizes, res, cost, pairs, est, sched = _setup(max_frames=1)
        sched.order = order
        sched.arrive(big)
        sched.arrive(small)
        assert sched.form_batch(now)[0][0].rid == first, (order, now)
    # the long frame goes first once it has waited its extra cost longer
    sizes, res, cost, pairs, est, sched = _setup(max_frames=1)
    sched.order = "slack"
    late_small = SV.Request(1, 0, *res[0], 0.0, arrival_s=0.005)
    sched.arrive(big)
    sched.arrive(late_small)
    assert sched.form_batch(0.006)[0][0].rid == 0  # 3.0 - 6 < 0.5 - 1
    # a request put back after an overflow stays first
    sizes, res, cost, pairs, est, sched = _setup(max_frames=1)
    sched.order = "slack"
    sched.arrive(small)
    retried = SV.Request(2, 1, *res[1], 0.0, arrival_s=0.0015)
    sched.requeue_front([retried])
    assert sched.form_batch(0.002)[0][0].rid == 2
    with pytest.raises(ValueError):
        SV.Scheduler(sizes, 1 << 30, est, order="lifo")


def test_loop_completes_everything_in_order_with_latencies():
    sizes, res, cost, pairs, est, sched = _setup()
    reqs = SV.poisson_requests(300, 400.0, 2, resolutions=res, seed=5)
    ex = SimExecutor(cost, pairs)
    loop = SV.ServingLoop(sched, ex, keep_outputs=True)
    done = loop.run(reqs, spin_s=0.0)
    assert len(done) == 300 and not sched.rejected
    assert sorted(r.rid for r in done) == list(range(300))
    assert all(r.latency_ms > 0 for r in done)
    assert loop.outputs[17] == "img17"
    # continuous batching: later batches pick up requests that arrived while
    # earlier ones ran, so some batches hold several frames
    assert max(len(b[1]) for b in loop.batches) > 1
    # arrival order is kept across groups
    order = [rid for b in sorted(loop.batches, key=lambda b: b[0]) for rid in b[1]]
    assert order == sorted(order)
    s = SV.summarize(done, sched.rejected, wall_s=ex.t, batches=loop.batches)
    assert s["completed"] == 300 and s["p99_ms"] >= s["p50_ms"] > 0
    assert 0.0 <= s["deadline_miss_frac"] <= 1.0


def test_overflowing_batch_is_retried_with_learned_counts():
    # estimates far too small: the first batches overflow, their requests
    # come back first with the measured counts, and everything completes
    sizes, res, cost, pairs, est, sched = _setup(est_pairs=10)
    reqs = SV.poisson_requests(40, 500.0, 2, resolutions=res, seed=9)
    ex = SimExecutor(cost, pairs)
    loop = SV.ServingLoop(sched, ex)
    done = loop.run(reqs, spin_s=0.0)
    assert len(done) == 40
    assert loop.overflows >= 1
    assert any(r.attempts > 0 for r in done)
    for k, v in pairs.items():
        assert est.pairs[k] == v


@pytest.mark.parametrize("P,w,h,pairs", [(1_000_000, 1280, 720, 1_500_000),
                                         (30_000_000, 7680, 4320, 40_000_000)])
def test_frame_bytes_covers_every_buffer(P, w, h, pairs):
    tiles = ((w + 15) // 16) * ((h + 15) // 16)
    need = pairs * (8 + 4) * 3 + tiles * 8 + 24 * P + 3 * w * h
    assert SV.frame_bytes(P, w, h, pairs) >= need
