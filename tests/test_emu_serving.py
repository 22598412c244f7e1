"""The serving loop's data path on the CPU, through the host emulation of the
kernels: two executor slots run their groups on two host threads (frames of
different groups overlap in time, as on two streams), every frame gets
freshly allocated buffers filled with random bytes, capacities come from the
loop's estimates (so some frames overflow and are retried), and the scenes
give tiles longer than the per-wave segment sort holds.  Each completed
request's image must equal the same request rendered alone with zeroed
buffers."""
import queue
import threading
import time

import numpy as np
import pytest

from polysplat_amd import scenes as S
from polysplat_amd import serving as SV
from tests import capi
from tests.emu import frames as HF

SIZES = (6_000, 24_000)
RES = ((256, 144), (160, 96))
BOX = 6.0
RADIUS = 1.5 * BOX
MIX = [(0.9, -3.0, 0.5), (0.1, -1.8, 0.4)]


def _scenes():
    return [S.make_scene(P, BOX, MIX, seed=300 + i) for i, P in enumerate(SIZES)]


class EmuExecutor:
    """Slots are host threads; a group's frames run back to back on its
    slot's thread through the emulated ps_forward."""

    def __init__(self, L, scenes, slots=2, seed=0):
        self.L = L
        self.scenes = scenes
        self.busy = [None] * slots
        self.jobs = [queue.Queue() for _ in range(slots)]
        self.seed = seed
        self.t0 = time.perf_counter()
        self.threads = [threading.Thread(target=self._worker, args=(i,), daemon=True)
                        for i in range(slots)]
        for t in self.threads:
            t.start()
        self.errors = []

    def _worker(self, slot):
        while True:
            fl = self.jobs[slot].get()
            if fl is None:
                return
            try:
                counts = []
                for r, cap in zip(fl.reqs, fl.capacities):
                    cam = SV.orbit_camera(r.width, r.height, r.azimuth, RADIUS)
                    f = HF.HostFrame(self.L, self.scenes[r.scene], cam, capacity=max(cap, 1),
                                     garbage=True, seed=self.seed + r.rid * 7 + r.attempts)
                    st, n, img = f.forward(r.bg, capi.V_LITE)
                    assert st in (capi.PS_OK, capi.PS_ERR_OVERFLOW), st
                    assert (st == capi.PS_ERR_OVERFLOW) == (n > cap), (st, n, cap)
                    counts.append(n)
                    fl.outputs.append(img)
                fl.host_counts = counts
            except Exception as e:  # reported by the test
                self.errors.append(e)
                fl.host_counts = [0] * len(fl.reqs)
            fl.done_ms = 1e3 * self.now_s()
            fl.done.set()

    def begin(self):
        self.t0 = time.perf_counter()

    def now_s(self):
        return time.perf_counter() - self.t0

    def free_slot(self):
        for i, b in enumerate(self.busy):
            if b is None:
                return i
        return None

    def launch(self, reqs, nbytes, capacities, slot):
        fl = SV.InFlight(0, reqs, nbytes, list(capacities), slot)
        fl.done = threading.Event()
        self.busy[slot] = fl
        self.jobs[slot].put(fl)
        return fl

    def poll(self):
        out = []
        for i, fl in enumerate(self.busy):
            if fl is not None and fl.done.is_set():
                self.busy[i] = None
                out.append((fl, fl.done_ms, fl.host_counts))
        return out

    def idle(self):
        return all(b is None for b in self.busy)

    def close(self):
        for q in self.jobs:
            q.put(None)
        for t in self.threads:
            t.join()


@pytest.fixture(scope="module")
def emu():
    import subprocess
    import sys

    subprocess.run([sys.executable, str(HF.HERE / "build.py")], check=True,
                   stdout=subprocess.DEVNULL)
    return HF.load()


def test_serving_frames_equal_frames_rendered_alone(emu):
    scs = _scenes()
    est = SV.Estimates(margin=1.0, slack=0)
    # Seed estimates below the true counts for one key, so its first frames
    # overflow and are retried with the measured count.
    for s in range(len(SIZES)):
        for (w, h) in RES:
            est.observe((s, w, h), ms=1.0)
    sched = SV.Scheduler(SIZES, 1 << 30, est, max_frames=3, max_batch_ms=50.0)
    reqs = SV.poisson_requests(8, 400.0, len(SIZES), resolutions=RES, seed=5)
    for i, r in enumerate(reqs):
        r.bg = ((i % 3) / 2.0, 0.25, 1.0 - (i % 2))
    ex = EmuExecutor(emu, scs)
    try:
        loop = SV.ServingLoop(sched, ex, keep_outputs=True)
        done = loop.run(reqs, spin_s=1e-3)
    finally:
        ex.close()
    assert not ex.errors, ex.errors
    assert len(done) == len(reqs) and not sched.rejected
    assert loop.overflows > 0  # the first estimates were too small
    assert max(len(b[1]) for b in loop.batches) > 1
    for r in done:
        cam = SV.orbit_camera(r.width, r.height, r.azimuth, RADIUS)
        f = HF.HostFrame(emu, scs[r.scene], cam, capacity=r.pairs + 16)
        st, n, img = f.forward(r.bg, capi.V_LITE)
        capi.check_with(emu, st)
        assert n == r.pairs
        np.testing.assert_array_equal(loop.outputs[r.rid], img, err_msg=f"request {r.rid}")
