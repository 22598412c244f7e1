"""Serving loop on the GPU: every request rendered in the mixed stream
(grouped with others, overlapping on two streams) is byte-identical to the
same request rendered alone through the single-frame entry point."""
import math

import numpy as np
import pytest
import torch

from polysplat_amd import _C
from polysplat_amd import serving as SV
from tests.gpu_helpers import say

pytestmark = [
    pytest.mark.gpu,
    pytest.mark.skip(reason="this test faulted the GPU in one run (two streams, orbit poses, "
                            "long segments) and the cause has not been found yet"),
]

RES = ((320, 180), (640, 360), (800, 600))
SIZES = (30_000, 80_000, 200_000)
BOX = 10.0


def _alone(sc, r, radius):
    dev = torch.device("cuda:0")
    cam = SV.orbit_camera(r.width, r.height, r.azimuth, radius)
    cap = r.pairs + 16
    ku = torch.empty(cap, dtype=torch.int64, device=dev)
    vu = torch.empty(cap, dtype=torch.int32, device=dev)
    ks = torch.empty(cap, dtype=torch.int64, device=dev)
    vs = torch.empty(cap, dtype=torch.int32, device=dev)
    space = torch.empty(_C.ops.get_sort_buffer_size(cap), dtype=torch.int8, device=dev)
    curr = torch.zeros(1, dtype=torch.int32, device=dev)
    P = sc["P"]
    e = torch.empty(0, dtype=torch.float32, device=dev)
    rec = torch.empty(24 * P, dtype=torch.uint8, device=dev)
    tiles = ((r.width + 15) // 16) * ((r.height + 15) // 16)
    ranges = torch.empty((tiles, 2), dtype=torch.int32, device=dev)
    out = torch.empty((r.height, r.width, 3), dtype=torch.uint8, device=dev)
    ws = torch.zeros(_C.ops.forward_workspace_size(P), dtype=torch.uint8, device=dev)
    n = _C.ops.forward_frame(sc["pos"], sc["shs"], sc["opac"], sc["cov"], r.width, r.height,
                             cam.position, cam.rotation, cam.focal_x, cam.focal_y, cam.zFar,
                             cam.zNear, 1, 1, curr, space, ku, vu, ks, vs, e, e, e, rec, ranges,
                             torch.tensor(r.bg, dtype=torch.float32), out, ws, True)
    return n, out.cpu().numpy()


def test_results_are_identical_alone_and_in_the_mix(gpu):
    radius = 1.5 * BOX
    scenes = SV.make_resident_scenes(SIZES, device="cuda:0", box=BOX,
                                     mix=[(0.9, -3.5, 0.5), (0.1, -2.0, 0.4)])
    ex = SV.GpuExecutor(scenes, slots=2, radius=radius)
    est = SV.Estimates()
    say("serving calibrate")
    SV.calibrate(ex, est, len(SIZES), resolutions=RES, azimuths=(0.0,), repeats=1)
    sched = SV.Scheduler(SIZES, 4 << 30, est, max_frames=6, max_batch_ms=50.0)
    reqs = SV.poisson_requests(48, 4000.0, len(SIZES), resolutions=RES, seed=11)
    for i, r in enumerate(reqs):
        r.bg = ((i % 3) / 2.0, 0.25, 1.0 - (i % 2))
    loop = SV.ServingLoop(sched, ex, keep_outputs=True)
    say("serving run", len(reqs))
    done = loop.run(reqs)
    torch.cuda.synchronize()
    assert len(done) == len(reqs) and not sched.rejected
    sizes = [len(b[1]) for b in loop.batches]
    say("batches", sizes, "overflows", loop.overflows)
    assert max(sizes) > 1  # requests were batched together
    summary = SV.summarize(done, sched.rejected, wall_s=ex.now_s(), batches=loop.batches)
    say("summary", summary)
    assert all(r.latency_ms > 0 and math.isfinite(r.latency_ms) for r in done)
    for r in done:
        n, img = _alone(scenes[r.scene], r, radius)
        assert n == r.pairs, (r.rid, n, r.pairs)
        mix = loop.outputs[r.rid].cpu().numpy()
        np.testing.assert_array_equal(mix, img, err_msg=f"request {r.rid} differs in the mix")
