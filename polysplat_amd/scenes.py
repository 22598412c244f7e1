"""Seeded synthetic scenes and cameras for the benchmark configurations.

The reference is benchmarked on trained scenes that are not available here;
these generators produce Gaussians with the distributions the benchmark
configurations specify:

* means uniform in a cube of side ``box``;
* per-axis log-scales from a normal (or a mixture of normals);
* rotations uniform on S^3 (normalised N(0,1) 4-vectors, w first);
* opacity = sigmoid(N(0, 1.5));
* 48 SH coefficients ~ N(0, 0.3), laid out [16][3] (coefficient, RGB).

cov3D follows the reference's scene preparation (pybind.cpp:22-56):
M = S * R, Sigma = M^T M, upper triangle (00, 01, 02, 11, 12, 22).

Cameras are pinholes with a 60 degree horizontal field of view, placed at
(0, 0, -1.5 * box) looking down +z with an identity rotation,
zNear = 0.01, zFar = 100 (example.py:33-44).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import torch

# name -> (P, box, width, height, log-scale mixture [(weight, mean, std)])
CONFIGS = {
    "c1": dict(P=100_000, box=10.0, width=800, height=800, mix=[(1.0, -4.5, 0.5)]),
    "c2": dict(P=3_000_000, box=50.0, width=1920, height=1080, mix=[(1.0, -4.5, 0.5)]),
    "c3": dict(P=6_000_000, box=50.0, width=3840, height=2160,
               mix=[(0.90, -5.0, 0.4), (0.09, -3.0, 0.4), (0.01, -1.0, 0.3)]),
    # City: 64x64 blocks on a 2 km plane, log-normal block heights, seen by an
    # elevated oblique camera (make_city_scene, config_camera).
    "c4": dict(P=56_000_000, box=2000.0, width=9216, height=6912, mix=[(1.0, -4.5, 0.5)],
               city=dict(blocks=64, extent=1000.0, footprint=0.7, height_median=25.0,
                         height_sigma=0.6, cam_pos=(0.0, -600.0, -1400.0),
                         cam_target=(0.0, 0.0, -100.0), z_far=5000.0)),
}


@dataclass
class SceneArrays:
    position: torch.Tensor  # f32 [P, 3]
    shs: torch.Tensor       # f32 [P, 48]
    opacity: torch.Tensor   # f32 [P]
    cov3d: torch.Tensor     # f32 [P, 6]
    extra: dict = field(default_factory=dict)

    @property
    def num_vertex(self) -> int:
        return int(self.opacity.shape[0])


def cov3d_from_scale_rot(scale: torch.Tensor, rot: torch.Tensor) -> torch.Tensor:
    """Sigma = (S R)^T (S R) with R from a w-first unit quaternion.

    Computed in float64 and rounded once to float32."""
    s = scale.double()
    q = rot.double()
    q = q / q.norm(dim=1, keepdim=True)
    r, x, y, z = q.unbind(1)
    # R[a][r] holds the nine values in the order the reference lists them in
    # its column-major matrix constructor, so R[a][r] is element (r, a) of the
    # mathematical rotation Rm, and Sigma = Rm^T diag(s^2) Rm becomes
    # Sigma[a][b] = sum_r R[a][r] s_r^2 R[b][r].
    R = torch.stack([
        torch.stack([1 - 2 * (y * y + z * z), 2 * (x * y - r * z), 2 * (x * z + r * y)], 1),
        torch.stack([2 * (x * y + r * z), 1 - 2 * (x * x + z * z), 2 * (y * z - r * x)], 1),
        torch.stack([2 * (x * z - r * y), 2 * (y * z + r * x), 1 - 2 * (x * x + y * y)], 1),
    ], 1)
    M = R * s[:, None, :]
    # Sigma[a][b] = sum_r M[a][r] M[b][r], summed elementwise: no batched
    # matrix product with a transposed operand on the device.
    sig = (M[:, :, None, :] * M[:, None, :, :]).sum(-1)
    out = torch.stack([sig[:, 0, 0], sig[:, 0, 1], sig[:, 0, 2],
                       sig[:, 1, 1], sig[:, 1, 2], sig[:, 2, 2]], 1)
    return out.float()


def make_scene(P: int, box: float, mix, seed: int = 1, device="cpu",
               chunk: int = 4_000_000) -> SceneArrays:
    """Generate P Gaussians on `device` in chunks (bounded host memory)."""
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    pos, shs, op, cov = [], [], [], []
    weights = torch.tensor([m[0] for m in mix], dtype=torch.float64)
    cum = torch.cumsum(weights / weights.sum(), 0)
    for c0 in range(0, P, chunk):
        n = min(chunk, P - c0)
        p = (torch.rand(n, 3, generator=g, device=device) - 0.5) * box
        # Mixture component per Gaussian, independent draws per axis.
        u = torch.rand(n, generator=g, device=device, dtype=torch.float64)
        comp = torch.searchsorted(cum.to(device), u).clamp_(max=len(mix) - 1)
        means = torch.tensor([m[1] for m in mix], device=device, dtype=torch.float32)[comp]
        stds = torch.tensor([m[2] for m in mix], device=device, dtype=torch.float32)[comp]
        ls = torch.randn(n, 3, generator=g, device=device) * stds[:, None] + means[:, None]
        rot = torch.randn(n, 4, generator=g, device=device)
        o = torch.sigmoid(torch.randn(n, generator=g, device=device) * 1.5)
        sh = torch.randn(n, 48, generator=g, device=device) * 0.3
        pos.append(p.float())
        shs.append(sh.float())
        op.append(o.float())
        cov.append(cov3d_from_scale_rot(torch.exp(ls), rot))
    cat = lambda xs: torch.cat(xs, 0).contiguous()  # noqa: E731
    return SceneArrays(cat(pos), cat(shs), cat(op), cat(cov))


def make_config_scene(name: str, seed: int = 1, device="cpu", P: int | None = None) -> SceneArrays:
    c = CONFIGS[name]
    if "city" in c:
        return make_city_scene(P if P is not None else c["P"], seed=seed, device=device,
                               **c["city"])
    return make_scene(P if P is not None else c["P"], c["box"], c["mix"], seed, device)


def make_city_scene(P: int, seed: int = 4, device="cpu", blocks: int = 64,
                    extent: float = 1000.0, footprint: float = 0.7,
                    height_median: float = 25.0, height_sigma: float = 0.6,
                    mix=((1.0, -4.5, 0.5),), chunk: int = 4_000_000, **_camera) -> SceneArrays:
    """P Gaussians in a blocks x blocks grid over x, z in [-extent, extent]:
    each Gaussian picks a block uniformly, lies in the central `footprint`
    fraction of it and at a height uniform in [0, h_block], with h_block
    log-normal per block (median height_median, log-sigma height_sigma).
    Up is -y (the camera convention of make_camera looks along +z with +y
    down the image).  Scales, rotations, opacities and SH as make_scene."""
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    pitch = 2.0 * extent / blocks
    hb = torch.exp(torch.randn(blocks * blocks, generator=g, device=device) * height_sigma
                   + math.log(height_median))
    mean, std = mix[0][1], mix[0][2]
    pos, shs, op, cov = [], [], [], []
    for c0 in range(0, P, chunk):
        n = min(chunk, P - c0)
        b = torch.randint(0, blocks * blocks, (n,), generator=g, device=device)
        bx = (b % blocks).float()
        bz = torch.div(b, blocks, rounding_mode="floor").float()
        u = torch.rand(n, 3, generator=g, device=device)
        x = -extent + (bx + 0.5) * pitch + (u[:, 0] - 0.5) * footprint * pitch
        z = -extent + (bz + 0.5) * pitch + (u[:, 1] - 0.5) * footprint * pitch
        y = -u[:, 2] * hb[b]
        ls = torch.randn(n, 3, generator=g, device=device) * std + mean
        rot = torch.randn(n, 4, generator=g, device=device)
        o = torch.sigmoid(torch.randn(n, generator=g, device=device) * 1.5)
        sh = torch.randn(n, 48, generator=g, device=device) * 0.3
        pos.append(torch.stack([x, y, z], 1).float())
        shs.append(sh.float())
        op.append(o.float())
        cov.append(cov3d_from_scale_rot(torch.exp(ls), rot))
    cat = lambda xs: torch.cat(xs, 0).contiguous()  # noqa: E731
    return SceneArrays(cat(pos), cat(shs), cat(op), cat(cov))


@dataclass
class CameraSpec:
    width: int
    height: int
    position: torch.Tensor  # CPU f32 [3]
    rotation: torch.Tensor  # CPU f32 [3, 3]
    focal_x: float
    focal_y: float
    zFar: float = 100.0
    zNear: float = 0.01
    id: int = 0
    img_name: str = "synthetic"


def make_camera(width: int, height: int, box: float, fov_deg: float = 60.0,
                distance: float | None = None) -> CameraSpec:
    fx = width / (2.0 * math.tan(math.radians(fov_deg) / 2.0))
    d = 1.5 * box if distance is None else distance
    return CameraSpec(width, height, torch.tensor([0.0, 0.0, -d], dtype=torch.float32),
                      torch.eye(3, dtype=torch.float32), fx, fx)


def look_at_camera(width: int, height: int, position, target, fov_deg: float = 60.0,
                   z_far: float = 100.0) -> CameraSpec:
    """Pinhole camera at `position` looking at `target`, with no roll about the
    viewing direction (camera x stays in the world x-z plane).  The rotation
    is the camera-to-world matrix given row by row; its third column is the
    viewing direction."""
    p = torch.tensor(position, dtype=torch.float64)
    f = torch.tensor(target, dtype=torch.float64) - p
    f = f / f.norm()
    xa = torch.tensor([f[2].item(), 0.0, -f[0].item()], dtype=torch.float64)
    xa = xa / xa.norm()
    ya = torch.linalg.cross(f, xa)
    rot = torch.stack([xa, ya, f], 1).float()  # columns: camera x, y, z in world
    fx = width / (2.0 * math.tan(math.radians(fov_deg) / 2.0))
    return CameraSpec(width, height, p.float(), rot, fx, fx, zFar=z_far)


def config_camera(name: str) -> CameraSpec:
    c = CONFIGS[name]
    if "city" in c:
        k = c["city"]
        return look_at_camera(c["width"], c["height"], k["cam_pos"], k["cam_target"],
                              z_far=k["z_far"])
    return make_camera(c["width"], c["height"], c["box"])
