"""PLY scene loader (reference pybind.cpp:58-120) on synthetic INRIA-layout
files: SH re-layout exactly, sigmoid opacity, exp scale, normalised w-first
quaternion, and Sigma = (S R)^T (S R) against a float64 evaluation."""
import math

import numpy as np
import pytest

import polysplat_amd as polysplat

PROPS = (["x", "y", "z", "nx", "ny", "nz", "f_dc_0", "f_dc_1", "f_dc_2"]
         + [f"f_rest_{i}" for i in range(45)] + ["opacity", "scale_0", "scale_1", "scale_2",
                                                 "rot_0", "rot_1", "rot_2", "rot_3"])


def write_ply(path, rec: np.ndarray):
    assert rec.shape[1] == 62
    head = ["ply", "format binary_little_endian 1.0", f"element vertex {rec.shape[0]}"]
    head += [f"property float {p}" for p in PROPS] + ["end_header"]
    with open(path, "wb") as f:
        f.write(("\n".join(head) + "\n").encode())
        f.write(rec.astype("<f4").tobytes())


def random_records(n, seed=0):
    rng = np.random.default_rng(seed)
    rec = rng.normal(size=(n, 62)).astype(np.float32)
    rec[:, 55:58] = rng.normal(-3, 1, size=(n, 3))
    return rec


def test_loadply_layout_and_transforms(tmp_path):
    n = 500
    rec = random_records(n)
    p = tmp_path / "s.ply"
    write_ply(p, rec)
    num, pos, shs, op, cov = polysplat.ops.loadPly(str(p))
    assert num == n
    np.testing.assert_array_equal(pos.numpy(), rec[:, :3])
    sh = shs.numpy().reshape(n, 16, 3)
    np.testing.assert_array_equal(sh[:, 0, :], rec[:, 6:9])
    rest = rec[:, 9:54]
    for j in range(1, 16):
        np.testing.assert_array_equal(sh[:, j, 0], rest[:, j - 1])
        np.testing.assert_array_equal(sh[:, j, 1], rest[:, j - 1 + 15])
        np.testing.assert_array_equal(sh[:, j, 2], rest[:, j - 1 + 30])
    np.testing.assert_allclose(op.numpy(), 1 / (1 + np.exp(-rec[:, 54].astype(np.float64))),
                               rtol=3e-7)
    # Sigma in float64: q normalised, R from (w, x, y, z), Sigma = R^T S^2 R
    for i in range(0, n, 17):
        s = np.exp(rec[i, 55:58].astype(np.float64))
        q = rec[i, 58:62].astype(np.float64)
        q /= np.linalg.norm(q)
        r, x, y, z = q
        # the nine values fill the columns of the matrix
        Rm = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - r * z), 2 * (x * z + r * y)],
                       [2 * (x * y + r * z), 1 - 2 * (x * x + z * z), 2 * (y * z - r * x)],
                       [2 * (x * z - r * y), 2 * (y * z + r * x), 1 - 2 * (x * x + y * y)]]).T
        Sig = Rm.T @ np.diag(s * s) @ Rm
        ref = [Sig[0, 0], Sig[0, 1], Sig[0, 2], Sig[1, 1], Sig[1, 2], Sig[2, 2]]
        np.testing.assert_allclose(cov.numpy()[i], ref, rtol=2e-5, atol=1e-6 * s.max() ** 2)


def test_loadply_identity_rotation_gives_diagonal_exactly(tmp_path):
    rec = random_records(8, seed=1)
    rec[:, 58:62] = [2.0, 0.0, 0.0, 0.0]  # unnormalised identity
    write_ply(tmp_path / "d.ply", rec)
    _, _, _, _, cov = polysplat.ops.loadPly(str(tmp_path / "d.ply"))
    c = cov.numpy()
    s = np.exp(rec[:, 55:58].astype(np.float32))
    np.testing.assert_allclose(c[:, [0, 3, 5]], s * s, rtol=6e-7)  # expf vs numpy exp: 1 ulp
    assert (c[:, [1, 2, 4]] == 0).all()


def test_loadply_errors(tmp_path):
    with pytest.raises(RuntimeError, match="Failed to open"):
        polysplat.ops.loadPly(str(tmp_path / "missing.ply"))
    (tmp_path / "bad.ply").write_bytes(b"ply\nformat ascii 1.0\nelement vertex 1\nend_header\n")
    with pytest.raises(RuntimeError):
        polysplat.ops.loadPly(str(tmp_path / "bad.ply"))
    rec = random_records(4)
    write_ply(tmp_path / "t.ply", rec)
    data = (tmp_path / "t.ply").read_bytes()
    (tmp_path / "trunc.ply").write_bytes(data[:-10])
    with pytest.raises(RuntimeError, match="truncated"):
        polysplat.ops.loadPly(str(tmp_path / "trunc.ply"))


def test_scene_loadply_on_cpu_device(tmp_path):
    write_ply(tmp_path / "s.ply", random_records(10))
    sc = polysplat.Scene("cpu")
    sc.loadPly(str(tmp_path / "s.ply"))
    assert sc.num_vertex == 10 and sc.shs.shape == (10, 48) and sc.shs_half.dtype.itemsize == 2
