"""Pins the CPU oracle against what the mathematics fixes.

The reference ships no unit tests or golden vectors for this path, so each
oracle function is checked here against closed forms, textbook values,
brute force over tiny inputs and invariants chosen so that a dropped term,
a wrong sign or index, or a transposed operand fails at least one check.
"""
import math

import numpy as np
import pytest
import torch

import oracle as O
from polysplat_amd import scenes
from polysplat_amd.metrics import diff_metrics

F32 = np.float32


def ulp(x):
    return float(np.spacing(np.float32(abs(x)) if x != 0 else np.float32(1e-38)))


# ------------------------------------------------------------------ log2
def test_log2_matches_libm_within_two_ulp():
    rng = np.random.default_rng(0)
    xs = np.concatenate([
        rng.uniform(1e-6, 1.0, 20000), rng.uniform(1.0, 1e6, 2000),
        10.0 ** rng.uniform(-37, 37, 2000),
        np.array([1.0, 0.5, 2.0, 0.25, 1.5, 0.75, 1 - 2**-24, 1 + 2**-23, 0.70710677, 1.4142135]),
    ]).astype(F32)
    worst = 0.0
    for x in xs[::7]:
        got = O.log2(float(x))
        ref = math.log2(float(x))
        err = abs(got - ref) / max(ulp(ref), 2**-24)
        worst = max(worst, err)
    assert worst <= 2.0, worst
    # Exact powers of two and special values.
    for e in range(-30, 30):
        assert O.log2(2.0**e) == e
    assert O.log2(0.0) == -math.inf
    assert math.isnan(O.log2(-1.0))
    sub = float(np.float32(1e-40))
    assert abs(O.log2(sub) - math.log2(sub)) < 1e-5


# ------------------------------------------------------------------ camera
def _cam(W=640, H=480, pos=(0.0, 0.0, 0.0), rot=None, fx=500.0, fy=520.0, zf=100.0, zn=0.01):
    rot = np.eye(3, dtype=F32) if rot is None else np.asarray(rot, F32)
    return O.camera(W, H, pos, rot, fx, fy, zf, zn)


def test_camera_matrices_closed_form():
    c = np.array([1.0, -2.0, 3.0], F32)
    cam = _cam(pos=c)
    view, proj, tf = O.camera_matrices(cam)
    V = view.reshape(4, 4).T  # row-major view of the column-major matrix
    np.testing.assert_array_equal(V[:3, :3], np.eye(3))
    np.testing.assert_array_equal(V[:3, 3], -c)
    Pm = proj.reshape(4, 4).T
    # x row: 2 fx / W, y row: 2 fy / H, w row picks view-space z.
    assert Pm[0, 0] == pytest.approx(2 * 500.0 / 640, rel=1e-6)
    assert Pm[1, 1] == pytest.approx(2 * 520.0 / 480, rel=1e-6)
    np.testing.assert_allclose(Pm[3, :3], [0, 0, 1])
    assert Pm[3, 3] == pytest.approx(-c[2])
    assert Pm[2, 2] == pytest.approx(100.0 / (100.0 - 0.01), rel=1e-6)
    np.testing.assert_allclose(tf, [640 / (2 * 500.0), 480 / (2 * 520.0)], rtol=1e-7)


def test_camera_rotation_is_read_as_columns():
    # The nine rotation values fill the columns of the view rotation: for a
    # 90 degree rotation about z given row by row, view(0,1) = r[3].
    r = np.array([[0, -1, 0], [1, 0, 0], [0, 0, 1]], F32)
    view, _, _ = O.camera_matrices(_cam(rot=r))
    V = view.reshape(4, 4).T
    np.testing.assert_array_equal(V[:3, :3], r.T)


# ------------------------------------------------------------------ projection
def test_cov2d_on_axis_isotropic_closed_form():
    cam = _cam(fx=800.0, fy=600.0)
    for z, sig in [(5.0, 0.1), (20.0, 0.7), (1.0, 0.01)]:
        cov3 = np.array([sig**2, 0, 0, sig**2, 0, sig**2], F32)
        out = O.cov2d(np.array([0, 0, z], F32), cov3, cam)
        np.testing.assert_allclose(out[0], (800 * sig / z) ** 2 + 0.3, rtol=1e-5)
        np.testing.assert_allclose(out[2], (600 * sig / z) ** 2 + 0.3, rtol=1e-5)
        assert abs(out[1]) < 1e-6


def test_cov2d_matches_ewa_jacobian_off_axis():
    # cov = J W Sigma W^T J^T + 0.3 I in double for a generic pose.
    rng = np.random.default_rng(1)
    ang = 0.3
    rot = np.array([[math.cos(ang), -math.sin(ang), 0], [math.sin(ang), math.cos(ang), 0],
                    [0, 0, 1]], F32)
    cam = _cam(pos=(0.5, -0.2, -4.0), rot=rot, fx=700.0, fy=650.0)
    view, _, tf = O.camera_matrices(cam)
    Vm = view.reshape(4, 4).T.astype(np.float64)
    for _ in range(50):
        p = rng.uniform(-1, 1, 3).astype(F32)
        A = rng.normal(size=(3, 3))
        S = (A @ A.T * 0.01).astype(F32)
        cov3 = np.array([S[0, 0], S[0, 1], S[0, 2], S[1, 1], S[1, 2], S[2, 2]], F32)
        t = Vm[:3, :3] @ p.astype(np.float64) + Vm[:3, 3]
        x, y, z = t
        lx, ly = 1.3 * tf[0], 1.3 * tf[1]
        x = min(lx, max(-lx, x / z)) * z
        y = min(ly, max(-ly, y / z)) * z
        J = np.array([[700.0 / z, 0, -700.0 * x / z**2], [0, 650.0 / z, -650.0 * y / z**2]])
        Wr = Vm[:3, :3]
        Sd = np.array([[cov3[0], cov3[1], cov3[2]], [cov3[1], cov3[3], cov3[4]],
                       [cov3[2], cov3[4], cov3[5]]], np.float64)
        ref = J @ Wr @ Sd @ Wr.T @ J.T + 0.3 * np.eye(2)
        got = O.cov2d(p, cov3, cam)
        np.testing.assert_allclose(got, [ref[0, 0], ref[0, 1], ref[1, 1]], rtol=2e-4, atol=1e-4)


def _one(pos, scale=0.05, opacity=0.9, shs=None, W=64, H=48, fx=60.0, cam_pos=(0, 0, 0)):
    pos = np.asarray(pos, F32).reshape(1, 3)
    cov = np.array([[scale**2, 0, 0, scale**2, 0, scale**2]], F32)
    sh = np.zeros((1, 48), F32) if shs is None else np.asarray(shs, F32).reshape(1, 48)
    cam = O.camera(W, H, cam_pos, np.eye(3, dtype=F32), fx, fx)
    return pos, sh, np.array([opacity], F32), cov, cam


def test_point_on_axis_projects_to_principal_point():
    pos, sh, op, cov, cam = _one([0, 0, 5.0])
    r = O.preprocess(pos, sh, op, cov, cam)
    assert r["M"] >= 1
    np.testing.assert_array_equal(r["xy"][0], [(64 - 1) / 2, (48 - 1) / 2])


def test_near_plane_cull_boundary():
    for z, expect in [(0.2, 0), (0.19, 0), (np.nextafter(F32(0.2), F32(1)), 1), (0.25, 1)]:
        pos, sh, op, cov, cam = _one([0, 0, float(z)], scale=0.001)
        r = O.preprocess(pos, sh, op, cov, cam)
        assert (r["M"] > 0) == bool(expect), z


def test_conic_is_inverse_covariance():
    rng = np.random.default_rng(2)
    cam = _cam(fx=300.0, fy=300.0)
    view, _, _ = O.camera_matrices(cam)
    for _ in range(30):
        p = np.array([rng.uniform(-0.5, 0.5), rng.uniform(-0.5, 0.5), rng.uniform(2, 6)], F32)
        A = rng.normal(size=(3, 3)) * 0.05
        S = (A @ A.T).astype(F32)
        cov3 = np.array([[S[0, 0], S[0, 1], S[0, 2], S[1, 1], S[1, 2], S[2, 2]]], F32)
        r = O.preprocess(p.reshape(1, 3), np.zeros((1, 48), F32), np.array([0.95], F32), cov3, cam)
        if r["M"] == 0:
            continue
        cov = O.cov2d(p, cov3[0], cam).astype(np.float64)
        k = -1.4426950216293334961
        a, b, c = r["conic"][0, 0] / (0.5 * k), r["conic"][0, 1] / k, r["conic"][0, 2] / (0.5 * k)
        Cm = np.array([[cov[0], cov[1]], [cov[1], cov[2]]])
        Q = np.array([[a, b], [b, c]])
        np.testing.assert_allclose(Q @ Cm, np.eye(2), atol=2e-5)
        assert r["conic"][0, 3] == pytest.approx(math.log2(0.95), abs=1e-6)


# ------------------------------------------------------------------ SH
def _y_real(l, m, d):
    """Real SH basis with the sign convention of the degree-3 expansion used by
    3DGS renderers (Condon-Shortley phase folded in)."""
    x, y, z = d
    c0 = 1 / (2 * math.sqrt(math.pi))
    c1 = math.sqrt(3 / (4 * math.pi))
    table = {
        (0, 0): c0,
        (1, -1): -c1 * y, (1, 0): c1 * z, (1, 1): -c1 * x,
        (2, -2): 0.5 * math.sqrt(15 / math.pi) * x * y,
        (2, -1): -0.5 * math.sqrt(15 / math.pi) * y * z,
        (2, 0): 0.25 * math.sqrt(5 / math.pi) * (2 * z * z - x * x - y * y),
        (2, 1): -0.5 * math.sqrt(15 / math.pi) * x * z,
        (2, 2): 0.25 * math.sqrt(15 / math.pi) * (x * x - y * y),
        (3, -3): -0.25 * math.sqrt(35 / (2 * math.pi)) * y * (3 * x * x - y * y),
        (3, -2): 0.5 * math.sqrt(105 / math.pi) * x * y * z,
        (3, -1): -0.25 * math.sqrt(21 / (2 * math.pi)) * y * (4 * z * z - x * x - y * y),
        (3, 0): 0.25 * math.sqrt(7 / math.pi) * z * (2 * z * z - 3 * x * x - 3 * y * y),
        (3, 1): -0.25 * math.sqrt(21 / (2 * math.pi)) * x * (4 * z * z - x * x - y * y),
        (3, 2): 0.25 * math.sqrt(105 / math.pi) * z * (x * x - y * y),
        (3, 3): -0.25 * math.sqrt(35 / (2 * math.pi)) * x * (x * x - 3 * y * y),
    }
    return table[(l, m)]


LM = [(0, 0), (1, -1), (1, 0), (1, 1)] + [(2, m) for m in range(-2, 3)] + [(3, m) for m in range(-3, 4)]


@pytest.mark.parametrize("k", range(16))
def test_sh_single_coefficient_known_answers(k):
    rng = np.random.default_rng(10 + k)
    dirs = [np.array(v, np.float64) for v in ([1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0],
                                               [0, 0, 1], [0, 0, -1])]
    dirs += [v / np.linalg.norm(v) for v in rng.normal(size=(6, 3))]
    for d in dirs:
        sh = np.zeros(48, F32)
        coef = np.array([0.7, -0.4, 1.3], F32)
        sh[3 * k:3 * k + 3] = coef
        p = (d * 3.0).astype(F32)
        got = O.sh_color(p, np.zeros(3, F32), sh)
        y = _y_real(*LM[k], p.astype(np.float64) / np.linalg.norm(p.astype(np.float64)))
        ref = np.maximum(0.5 + y * coef.astype(np.float64), 0.0)
        np.testing.assert_allclose(got, ref, atol=3e-6)


def _round_f32(q):
    """A rational rounded once to the nearest float32 (ties to even), in
    integer arithmetic: the exact result of one fused multiply-add."""
    from fractions import Fraction
    q = Fraction(q)
    if q == 0:
        return F32(0.0)
    sign, q = (-1 if q < 0 else 1), abs(q)
    e = q.numerator.bit_length() - q.denominator.bit_length()
    if Fraction(2) ** e > q:
        e -= 1
    e = max(e, -126)                     # subnormals share the 2^-149 grid
    scaled = q / Fraction(2) ** (e - 23)  # 24-bit significand grid
    n, rem = divmod(scaled.numerator, scaled.denominator)
    if 2 * rem > scaled.denominator or (2 * rem == scaled.denominator and n % 2 == 1):
        n += 1
    return F32(sign * float(Fraction(n) * Fraction(2) ** (e - 23)))


def _fma32(a, b, c):
    from fractions import Fraction
    return _round_f32(Fraction(float(a)) * Fraction(float(b)) + Fraction(float(c)))


def test_round_f32_known_answers():
    from fractions import Fraction
    assert _round_f32(Fraction(1, 3)) == F32(1) / F32(3)
    assert _round_f32(1 + Fraction(1, 2 ** 24)) == F32(1.0)           # tie -> even
    assert _round_f32(1 + Fraction(3, 2 ** 24)) == F32(1 + 2 ** -22)  # tie -> even (up)
    assert _round_f32(Fraction(2 ** -149)) == np.float32(2 ** -149)
    # the fused result differs from the two-step one where the product's
    # low bits matter
    a = F32(1 + 2 ** -12)  # a * a = 1 + 2^-11 + 2^-24
    assert _fma32(a, a, F32(-1.0)) == F32(2 ** -11 + 2 ** -24)
    assert a * a - F32(1.0) == F32(2 ** -11)


def test_sh_accumulates_by_fused_multiply_adds_in_term_order():
    """The colour sum 0.5 + sum_k basis_k * sh_k is one correctly rounded
    fused multiply-add per term, in term order (the reference's
    `r - b * sh + ...` contracted), and the basis products keep the
    reference's left-to-right order.  An unfused chain fails this check."""
    rng = np.random.default_rng(21)
    C0, C1 = F32(0.28209479177387814), F32(0.4886025119029199)
    C2 = [F32(v) for v in (1.0925484305920792, -1.0925484305920792, 0.31539156525252005,
                           -1.0925484305920792, 0.5462742152960396)]
    C3 = [F32(v) for v in (-0.5900435899266435, 2.890611442640554, -0.4570457994644658,
                           0.3731763325901154, -0.4570457994644658, 1.445305721320277,
                           -0.5900435899266435)]
    unfused_differs = 0
    for _ in range(300):
        p = rng.normal(size=3).astype(F32) * F32(4)
        sh = (rng.normal(size=48) * 0.6).astype(F32)
        l2 = p[0] * p[0] + p[1] * p[1] + p[2] * p[2]
        rs = F32(1.0) / np.sqrt(l2, dtype=F32)
        x, y, z = p[0] * rs, p[1] * rs, p[2] * rs
        xx, yy, zz, xy, yz, xz = x * x, y * y, z * z, x * y, y * z, x * z
        two, three, four = F32(2), F32(3), F32(4)
        b = [C0, -(C1 * y), C1 * z, -(C1 * x), C2[0] * xy, C2[1] * yz,
             C2[2] * (two * zz - xx - yy), C2[3] * xz, C2[4] * (xx - yy),
             C3[0] * y * (three * xx - yy), C3[1] * xy * z, C3[2] * y * (four * zz - xx - yy),
             C3[3] * z * (two * zz - three * xx - three * yy), C3[4] * x * (four * zz - xx - yy),
             C3[5] * z * (xx - yy), C3[6] * x * (xx - three * yy)]
        got = O.sh_color(p, np.zeros(3, F32), sh)
        for ch in range(3):
            r = _fma32(b[0], sh[ch], F32(0.5))
            u = b[0] * sh[ch] + F32(0.5)
            for k in range(1, 16):
                r = _fma32(b[k], sh[3 * k + ch], r)
                u = u + b[k] * sh[3 * k + ch]
            want = max(r, F32(0.0))
            assert got[ch].view(np.uint32) == want.view(np.uint32), (ch, got[ch], want)
            unfused_differs += int(max(u, F32(0.0)) != want)
    assert unfused_differs > 0


# ------------------------------------------------------------------ tile coverage
def _rect_min_q(cx, cy, a, b, c, x0, y0, x1, y1):
    """min over the pixel rectangle [x0,x1]x[y0,y1] of q(d) = a dx^2 + 2b dx dy + c dy^2,
    d = p - centre, in double (edges minimised in closed form)."""
    if x0 <= cx <= x1 and y0 <= cy <= y1:
        return 0.0
    best = math.inf
    for fixed_x in (x0, x1):  # vertical edges: q as a function of y
        dx = fixed_x - cx
        dy = -b * dx / c
        dy = min(max(dy, y0 - cy), y1 - cy)
        best = min(best, a * dx * dx + 2 * b * dx * dy + c * dy * dy)
    for fixed_y in (y0, y1):
        dy = fixed_y - cy
        dx = -b * dy / a
        dx = min(max(dx, x0 - cx), x1 - cx)
        best = min(best, a * dx * dx + 2 * b * dx * dy + c * dy * dy)
    return best


def test_tile_test_agrees_with_exact_geometry():
    rng = np.random.default_rng(3)
    mismatches, checked = 0, 0
    for _ in range(3000):
        cx, cy = rng.uniform(-20, 80, 2)
        # random SPD conic from a random 2x2 covariance
        s1, s2 = 10 ** rng.uniform(-0.5, 1.6, 2)
        th = rng.uniform(0, math.pi)
        R = np.array([[math.cos(th), -math.sin(th)], [math.sin(th), math.cos(th)]])
        Cv = R @ np.diag([s1 * s1 + 0.3, s2 * s2 + 0.3]) @ R.T
        Q = np.linalg.inv(Cv)
        a, b, c = (float(F32(Q[0, 0])), float(F32(Q[0, 1])), float(F32(Q[1, 1])))
        power = float(F32(rng.uniform(0.5, 5.5)))
        for tx in range(-1, 6):
            for ty in range(-1, 6):
                x0, y0 = 16 * tx, 16 * ty
                got = O.tile_hits(tx, ty, float(F32(cx)), float(F32(cy)), a, b, c, power)
                m = _rect_min_q(float(F32(cx)), float(F32(cy)), a, b, c, x0, y0, x0 + 15, y0 + 15)
                exact = m <= 2 * power
                checked += 1
                if got != exact:
                    mismatches += 1
                    # only allowed at the boundary of the ellipse
                    assert abs(m - 2 * power) <= 1e-4 * (2 * power + m), (m, 2 * power)
    assert checked > 50000 and mismatches < checked * 1e-3


def test_emission_invariants_small_scene():
    s = scenes.make_scene(4000, 6.0, [(0.8, -3.5, 0.5), (0.2, -1.5, 0.4)], seed=5)
    cs = scenes.make_camera(200, 136, 6.0)
    cam = O.camera_from_spec(cs)
    pos, sh, op, cov = s.position.numpy(), s.shs.numpy(), s.opacity.numpy(), s.cov3d.numpy()
    r = O.preprocess(pos, sh, op, cov, cam)
    keys, vals, counts = r["keys"], r["vals"], r["counts"]
    assert r["M"] == counts.sum() == len(keys)
    gx = (200 + 15) // 16
    tiles = (keys >> np.uint64(32)).astype(np.int64)
    # canonical order: Gaussian id ascending, then tile id ascending
    order = np.lexsort((tiles, vals))
    np.testing.assert_array_equal(order, np.arange(len(keys)))
    # per Gaussian: same depth bits on every key, the centre tile is emitted
    # when the centre is on screen, and every tile lies inside the rect the
    # radius allows.
    depth_bits = (keys & np.uint64(0xffffffff)).astype(np.uint32)
    xy = r["xy"]
    for g in np.unique(vals)[:500]:
        sel = vals == g
        assert len(np.unique(depth_bits[sel])) == 1
        t = tiles[sel]
        cx, cy = xy[g]
        if 0 <= cx < 200 and 0 <= cy < 136:
            assert int(cy // 16) * gx + int(cx // 16) in set(t.tolist())


def test_preprocess_bruteforce_tile_sets():
    # Every (Gaussian, tile) decision equals an independent FP64 test up to
    # boundary cases.
    s = scenes.make_scene(600, 4.0, [(0.5, -2.5, 0.4), (0.5, -1.2, 0.3)], seed=7)
    cs = scenes.make_camera(96, 80, 4.0)
    cam = O.camera_from_spec(cs)
    pos, sh, op, cov = s.position.numpy(), s.shs.numpy(), s.opacity.numpy(), s.cov3d.numpy()
    r = O.preprocess(pos, sh, op, cov, cam)
    gx, gy = 6, 5
    emitted = set(zip(r["vals"].tolist(), (r["keys"] >> np.uint64(32)).astype(np.int64).tolist()))
    k = -1.4426950216293334961
    bad = 0
    for g in np.unique(r["vals"]):
        cx, cy = map(float, r["xy"][g])
        a, b, c = r["conic"][g, 0] / (0.5 * k), r["conic"][g, 1] / k, r["conic"][g, 2] / (0.5 * k)
        power = math.log(256.0 * float(op[g]))
        for t in range(gx * gy):
            tx, ty = t % gx, t // gx
            m = _rect_min_q(cx, cy, a, b, c, 16 * tx, 16 * ty, 16 * tx + 15, 16 * ty + 15)
            exact = m <= 2 * power
            if exact != ((g, t) in emitted):
                # allowed: boundary cases, and tiles outside the conservative
                # rect the reference clips to
                if abs(m - 2 * power) > 1e-3 * (2 * power + m):
                    bad += 1
    assert bad <= 2, bad


# ------------------------------------------------------------------ sort / ranges
def test_sort_is_stable_by_key():
    rng = np.random.default_rng(4)
    for n in [0, 1, 2, 7, 1000, 65537]:
        keys = (rng.integers(0, 50, n).astype(np.uint64) << np.uint64(32)) | rng.integers(
            0, 4, n).astype(np.uint64)
        vals = rng.permutation(n).astype(np.uint32)
        ks, vs = O.sort_pairs(keys, vals)
        ref = sorted(range(n), key=lambda i: (int(keys[i]), i))
        np.testing.assert_array_equal(ks, keys[ref])
        np.testing.assert_array_equal(vs, vals[ref])


def test_tile_ranges_modes():
    keys = np.array([(t << 32) | d for t, d in [(0, 1), (0, 5), (2, 0), (2, 0), (2, 3), (5, 9)]],
                    np.uint64)
    r0 = O.tile_ranges(keys, 7, 0)
    r1 = O.tile_ranges(keys, 7, 1)
    np.testing.assert_array_equal(r0, [[0, 2], [2, 2], [2, 5], [5, 5], [5, 5], [5, 6], [6, 6]])
    np.testing.assert_array_equal(r1, [[0, 2], [0, 0], [2, 5], [0, 0], [0, 0], [5, 6], [0, 0]])
    assert (O.tile_ranges(np.zeros(0, np.uint64), 3, 0) == 0).all()


# ------------------------------------------------------------------ render
def _render_one(feats, W=32, H=32, bg=(0.1, 0.2, 0.3), check=8, dt=False):
    xy = np.array([f[0] for f in feats], F32)
    con = np.array([f[1] for f in feats], F32)
    rgb = np.array([list(f[2]) + [1.0] for f in feats], F32)
    keys = []
    gx = (W + 15) // 16
    # put every Gaussian in every tile, in list order
    for t in range((W + 15) // 16 * ((H + 15) // 16)):
        for i in range(len(feats)):
            keys.append((t, i))
    vals = np.array([i for _, i in keys], np.uint32)
    k = np.array([(t << 32) for t, _ in keys], np.uint64)
    ranges = O.tile_ranges(k, gx * ((H + 15) // 16), 0)
    return O.render_f32(W, H, ranges, vals, xy, rgb, con, bg, check, dt)


def test_render_empty_tiles_write_background():
    ranges = np.zeros((4, 2), np.int32)
    img, img_f, stop, need = O.render_f32(20, 20, ranges, np.zeros(1, np.uint32),
                                          np.zeros((1, 2), F32), np.zeros((1, 4), F32),
                                          np.zeros((1, 4), F32), (0.5, 0.25, 1.0), 8)
    assert (img[..., 0] == 127).all() and (img[..., 1] == 63).all() and (img[..., 2] == 255).all()


def test_render_single_gaussian_at_pixel_centre():
    o = 0.6
    feats = [((5.0, 7.0), (-0.5 * 1.4427 * 0.2, 0.0, -0.5 * 1.4427 * 0.2, math.log2(o)),
              (0.8, 0.4, 0.2))]
    img, img_f, _, _ = _render_one(feats, bg=(0.0, 0.0, 0.0))
    # at the centre alpha = o, T = 1 - o, C = rgb * o
    np.testing.assert_allclose(img_f[7, 5], np.array([0.8, 0.4, 0.2]) * o, rtol=1e-6)
    assert tuple(img[7, 5]) == tuple(int(v * o * 255) for v in (0.8, 0.4, 0.2))
    # one pixel away the Gaussian weight drops by 2^(-0.5*log2e*0.2) = exp(-0.1)
    assert img_f[7, 6, 0] == pytest.approx(0.8 * o * math.exp(-0.1), rel=1e-4)


def test_render_front_to_back_two_layers_and_saturation_rule():
    big = (-1e-9, 0.0, -1e-9, math.log2(0.999999))  # ~opaque everywhere
    feats = [((8.0, 8.0), big, (1.0, 0.0, 0.0)), ((8.0, 8.0), big, (0.0, 1.0, 0.0))]
    img, img_f, stop, need = _render_one(feats, bg=(0.0, 0.0, 1.0), check=1)
    # the first layer takes alpha ~= 0.999999, the second gets only T ~= 1e-6;
    # T < 1e-4 is then cleared to 0 so no background leaks in.
    assert img[0, 0, 0] == 254 or img[0, 0, 0] == 255
    assert img[0, 0, 2] == 0
    assert (stop == 1).all()  # saturation reached after the first Gaussian
    assert (need == 1).all()


def test_render_u8_truncates_and_clamps():
    feats = [((8.0, 8.0), (-1e-9, 0.0, -1e-9, math.log2(0.5)), (3.0, 0.5, -1.0))]
    img, img_f, _, _ = _render_one(feats, bg=(0.0, 0.0, 0.0))
    assert img[0, 0, 0] == 255  # clamp high
    assert img[0, 0, 1] == 63   # 0.25 * 255 = 63.75 -> 63 (truncation)
    assert img[0, 0, 2] == 0    # clamp low


def test_render_alpha_skip_threshold_is_log2_of_one_over_255():
    """The _dt renderers skip a pixel's blend when its power (log2 alpha) is
    below log2(1/255) (render.cu:352-362: kept when power >= -7.994353375f).
    A flat Gaussian (zero conic) puts the same power on every pixel, so the
    output is rgb * 2^w + (1 - 2^w) * bg when kept and bg when skipped."""
    thr = np.float32(-7.994353375)
    assert abs(float(thr) - math.log2(1.0 / 255.0)) < 1e-6
    below = np.nextafter(thr, np.float32(-np.inf))
    rgb, bg = (0.8, 0.4, 0.2), (0.5, 0.25, 0.125)
    for w in (thr, below, np.float32(-7.5), np.float32(-9.0)):
        feats = [((8.0, 8.0), (0.0, 0.0, 0.0, float(w)), rgb)]
        for dt in (False, True):
            img, img_f, _, _ = _render_one(feats, bg=bg, dt=dt)
            a = 2.0 ** float(w)
            kept = (not dt) or w >= thr
            want = [c * a + (1.0 - a) * b for c, b in zip(rgb, bg)] if kept else list(bg)
            np.testing.assert_allclose(img_f[3, 11], want, rtol=2e-6, err_msg=f"w={w} dt={dt}")
            assert (img_f == img_f[0, 0]).all()  # every pixel the same


def test_render_fp32_matches_fp64_c1():
    s = scenes.make_config_scene("c1", seed=1)
    cam = O.camera_from_spec(scenes.config_camera("c1"))
    r = O.frame(s.position.numpy(), s.shs.numpy(), s.opacity.numpy(), s.cov3d.numpy(), cam,
                (0, 0, 0), fp64=True)
    m = diff_metrics(r["image"], r["image64"])
    assert m["max_abs_diff"] <= 1 and m["psnr_db"] >= 50, m


# ------------------------------------------------------------------ half precision helpers
def test_half_conversions_round_to_nearest_even():
    rng = np.random.default_rng(6)
    xs = np.concatenate([rng.normal(0, 3, 5000), 10.0 ** rng.uniform(-9, 4.8, 5000)]).astype(F32)
    xs = np.concatenate([xs, -xs, np.array([0.0, 65504.0, 65520.0, 6e-8, 3e-8], F32)])
    L = O.lib()
    f16 = np.array([L.orc_f32_to_f16(float(x)) for x in xs], np.uint16)
    np.testing.assert_array_equal(f16, xs.astype(np.float16).view(np.uint16))
    bf = np.array([L.orc_f32_to_bf16(float(x)) for x in xs], np.uint16)
    ref = torch.from_numpy(xs).to(torch.bfloat16).view(torch.int16).numpy().view(np.uint16)
    np.testing.assert_array_equal(bf, ref)
