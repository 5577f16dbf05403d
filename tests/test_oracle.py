"""CPU tests of the oracle (test infrastructure) against known answers, closed forms and brute force.

Each check is chosen so a plausible restatement mistake fails it: a wrong RNG mapping fails the
Thrust / PCG known answers, a wrong intersection formula fails the closed forms, a leaky triangle
test fails the watertight sweep, a wrong PDF weight fails the normalisation and the furnace.
"""
import json
import math
import os

import numpy as np
import pytest

import oracle_py as O
from cortexraytracercuda_amd import Scene, scenes

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name):
    with open(os.path.join(G, name)) as f:
        return json.load(f)


# ------------------------------------------------------------------ RNG known answers
def test_minstd_thrust_documented_kat():
    # minstd_rand: the 10000th output of a default-constructed engine is 399268537
    # (thrust/random/linear_congruential_engine.h), the same as std::minstd_rand.
    out, _ = O.minstd(1, 10000)
    assert int(out[-1]) == 399268537
    assert _load("thrust_kat.json")["minstd_raw_10000th_default"] == 399268537


def test_minstd_uniform_float_matches_thrust():
    kat = _load("thrust_kat.json")["minstd_uniform_float"]
    for seed, bits in kat.items():
        _, u = O.minstd(int(seed), len(bits))
        assert np.array_equal(u.view(np.uint32), np.array(bits, dtype=np.uint32)), seed


def test_uniform_int_axis_matches_thrust():
    # bvh_node's random split axis: uniform_int_distribution<int>(0, 2) on minstd_rand(1984)
    kat = _load("thrust_kat.json")["uniform_int_0_2_seed1984"]
    assert list(O.minstd_uniform_int(1984, len(kat), 0, 2)) == kat


def test_get_int_quirk_always_min():
    # uniform_real_distribution<int> divides integers: always returns the lower bound
    # (core/src/misc/gpu_randomizer.cuh:29-36) -> light sampling always picks light 0
    assert set(_load("thrust_kat.json")["uniform_real_int_0_3_seed7"]) == {0}


def test_pcg32_matches_vendored_pcg():
    kat = _load("pcg32_kat.json")
    for c in kat["cases"]:
        out = O.pcg32(c["initstate"], c["initseq"], len(c["out"]))
        assert list(map(int, out)) == c["out"], c["initstate"]
    # the pcg-c-basic demo values for (42, 54)
    assert list(map(int, O.pcg32(42, 54, 3))) == [0xa15c02b7, 0x7b47f409, 0xba1d3330]


def test_keyed_streams_are_distinct_and_deterministic():
    a = O.keyed(1, 10, 0, 0, 8)
    b = O.keyed(1, 10, 0, 0, 8)
    c = O.keyed(1, 10, 0, 1, 8)
    d = O.keyed(1, 10, 1, 0, 8)
    e = O.keyed(2, 10, 0, 0, 8)
    assert np.array_equal(a, b)
    for x in (c, d, e):
        assert not np.array_equal(a, x)


# ------------------------------------------------------------------ tonemap
def test_tonemap_closed_forms():
    spp = 16
    acc = np.array([[0, 0, 0, 16], [16, 16, 16, 16], [np.nan, 4, 1e30, 16], [-1, 0.25 * 16, 1e-9, 16]], np.float32)
    rgb = O.tonemap(acc, spp)
    assert list(rgb[0]) == [0, 0, 0]
    assert list(rgb[1]) == [255, 255, 255]           # sqrt(1) clamps to 0.999 -> int(255.74)
    assert rgb[2, 0] == 0 and rgb[2, 2] == 255        # NaN -> 0, huge -> clamp
    assert rgb[2, 1] == int(255.99 * 0.5)             # 4/16 -> sqrt 0.5 -> 127
    assert rgb[3, 0] == 0                             # negative sqrt -> NaN is clamped low? (NaN < 0 false)
    assert rgb[3, 1] == 127


def test_tonemap_monotone_exhaustive():
    # every float32 in [0, 2] (step 2^-16) maps monotonically to [0, 255]
    x = np.arange(0, 2 * 65536 + 1, dtype=np.float64) / 65536.0
    acc = np.zeros((x.size, 4), np.float32)
    acc[:, 0] = x
    acc[:, 3] = 1
    rgb = O.tonemap(acc, 1)[:, 0].astype(int)
    assert rgb[0] == 0 and rgb[-1] == 255
    assert np.all(np.diff(rgb) >= 0)
    # each level k is entered where sqrt(x) * 255.99 crosses k
    for k in (1, 64, 128, 200, 254):
        first = np.argmax(rgb >= k)
        assert math.sqrt(x[first]) * 255.99 >= k - 1e-3


# ------------------------------------------------------------------ intersections
def _one_sphere(c=(0, 0, 0), r=1.0):
    s = Scene()
    m = s.lambertian("m", (0.5, 0.5, 0.5))
    s.sphere(c, r, m)
    return s


@pytest.mark.parametrize("prec", [32, 64])
def test_sphere_centre_ray_closed_form(prec):
    s = _one_sphere((0, 0, -5), 1.5)
    # ray through the centre: t = (|oc| - r) / |d| for an unnormalised direction
    d = np.array([0.0, 0.0, -2.0])
    rays = np.array([[0, 0, 0, 1e-5, *d, 1e30]])
    for ref_style in (True, False):
        prim, _, t, _ = O.trace(s, rays, precision=prec, ref_style=ref_style)
        assert prim[0] == 0
        assert abs(t[0] - (5 - 1.5) / 2.0) <= 1e-6
    # origin inside: the far root is taken (strict interval)
    rays = np.array([[0, 0, -5, 1e-5, 0, 1, 0, 1e30]])
    prim, _, t, _ = O.trace(s, rays, precision=prec)
    assert prim[0] == 0 and abs(t[0] - 1.5) <= 1e-6
    # tangent-plane miss and a behind-the-origin miss
    rays = np.array([[0, 1.6, 0, 1e-5, 0, 0, -1, 1e30], [0, 0, -10, 1e-5, 0, 0, -1, 1e30]])
    prim, _, _, _ = O.trace(s, rays, precision=prec)
    assert list(prim) == [-1, -1]


def test_triangle_barycentric_closed_form():
    s = Scene()
    m = s.lambertian("m", (0.5, 0.5, 0.5))
    v = np.array([[0, 0, -2], [2, 0, -2], [0, 2, -2]], np.float32)
    s.mesh(v, np.array([[0, 1, 2]]), m)
    rng = np.random.default_rng(3)
    for _ in range(50):
        b1, b2 = rng.uniform(0.05, 0.45, 2)
        p = (1 - b1 - b2) * v[0] + b1 * v[1] + b2 * v[2]
        o = np.array([0.3, -0.2, 1.0])
        d = (p - o) * 0.7
        for ref_style in (True, False):
            prim, _, t, uv = O.trace(s, np.array([[*o, 1e-5, *d, 1e30]]), precision=64, ref_style=ref_style)
            assert prim[0] == 0
            assert abs(t[0] - 1 / 0.7) < 1e-9
            assert abs(uv[0, 0] - b1) < 1e-9 and abs(uv[0, 1] - b2) < 1e-9


def test_quad_and_light_conventions():
    s = Scene()
    m = s.lambertian("m", (0.5, 0.5, 0.5))
    s.quad((0, 0, -3), (2, 0, 0), (0, 1, 0), m)            # corner form, alpha/beta in [0, 1]
    s.area_light((5, 0.5, -3), (2, 0, 0), (0, 1, 0), invisible=True)  # centred form, +-0.5
    # light is global id 0, quad id 1
    rays = np.array([[1.5, 0.25, 0, 1e-5, 0, 0, -1, 1e30],     # quad at alpha 0.75, beta 0.25
                     [5.5, 0.75, 0, 1e-5, 0, 0, -1, 1e30],     # light centre + (0.5u/2, 0.25v) -> uv .75 .75
                     [2.01, 0.5, 0, 1e-5, 0, 0, -1, 1e30]])    # just outside the quad
    prim, _, t, uv = O.trace(s, rays, camera=False, precision=64)
    assert list(prim) == [1, 0, -1]
    assert np.allclose(uv[0], [0.75, 0.25]) and np.allclose(uv[1], [0.75, 0.75])
    # camera rays do not see invisible lights (directional_light.cuh:106-109)
    prim, _, _, _ = O.trace(s, rays[1:2], camera=True, precision=64)
    assert prim[0] == -1


def test_watertight_closed_mesh_has_no_leaks():
    # rays from inside a closed icosphere towards its vertices/edges must all hit (FP32 oracle)
    v, f = scenes.icosphere(2)
    s = Scene()
    m = s.lambertian("m", (0.5, 0.5, 0.5))
    s.mesh(v.astype(np.float32), f.astype(np.uint32), m)
    e = np.concatenate([v, 0.5 * (v[f[:, 0]] + v[f[:, 1]]), (v[f[:, 0]] + v[f[:, 1]] + v[f[:, 2]]) / 3.0])
    o = np.zeros_like(e)
    rays = np.concatenate([o, np.full((e.shape[0], 1), 1e-5), e, np.full((e.shape[0], 1), 1e30)], 1)
    prim, _, _, _ = O.trace(s, rays, precision=32, ref_style=False)
    assert np.all(prim >= 0)


def test_bvh_equals_brute_force():
    s = scenes.large_scene(levels=3, n_spheres=60)
    rng = np.random.default_rng(5)
    o = rng.uniform(-12, 12, (4000, 3))
    d = rng.normal(size=(4000, 3))
    rays = np.concatenate([o, np.full((4000, 1), 1e-5), d, np.full((4000, 1), 1e30)], 1)
    pb, _, tb, _ = O.trace(s, rays, precision=64, accel=0)
    pf, _, tf, _ = O.trace(s, rays, precision=64, accel=1)
    same = pb == pf
    assert np.all(same | (np.abs(tb - tf) <= 1e-12 * np.abs(tf)))
    assert np.allclose(tb[same], tf[same])


# ------------------------------------------------------------------ estimator pieces
def test_light_pdf_integrates_to_one():
    s = Scene()
    s.area_light((0, 2, 0), (1, 0, 0), (0, 0, 1), invisible=True)   # normal -y, faces the origin
    rng = np.random.default_rng(1)
    n = 400000
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    pdf = O.light_pdf(s, np.zeros((n, 3)), d)
    est = pdf.mean() * 4 * math.pi   # E_uniform[pdf / (1/4pi)] = integral of pdf over directions
    assert abs(est - 1.0) < 0.02, est


def test_white_furnace_floor_under_sky():
    # Lambertian floor of albedo rho under a uniform sky of radiance 1, camera looking down, one
    # tiny far light of zero emission: pixel -> rho * (1 - fraction of cosine mass on the light)
    rho = 0.6
    s = Scene()
    s.set_camera((0, 1, 0), (0, 0, 0), (0, 0, -1), vfov=30.0, aspect_ratio=1.0, focus_dist=1.0)
    s.background = (1.0, 1.0, 1.0)
    m = s.lambertian("floor", (rho, rho, rho))
    s.quad((-50, 0, -50), (100, 0, 0), (0, 0, 100), m)
    s.area_light((0, 100, 0), (0.01, 0, 0), (0, 0, 0.01), rgb=(0, 0, 0), invisible=True)
    acc, _ = O.render(s, 16, 16, 64, 2, seed=3, rng="pcg", precision=64, threads=os.cpu_count() or 4)
    val = acc[..., :3].sum() / (16 * 16 * 64 * 3)
    assert abs(val - rho) < 0.01, val
    acc32, _ = O.render(s, 16, 16, 64, 2, seed=2580, rng="minstd", precision=32, threads=os.cpu_count() or 4)
    val32 = acc32[..., :3].sum() / (16 * 16 * 64 * 3)
    assert abs(val32 - rho) < 0.01, val32


def test_spp_normalisation_quirk():
    # floor(sqrt(spp))^2 samples are traced but the tonemap divides by spp (gpu.cu:776-792)
    s = Scene()
    s.background = (1.0, 1.0, 1.0)
    acc, _ = O.render(s, 4, 4, 5, 2, seed=1, rng="pcg", precision=64)
    assert np.all(acc[..., 3] == 4)
    assert np.allclose(acc[..., :3], 4.0)
    rgb = O.tonemap(acc, 5)
    assert np.all(rgb == int(255.99 * math.sqrt(4 / 5)))


def test_oracles_agree_statistically(box):
    a64, s64 = O.render(box, 48, 48, 16, 8, seed=1, rng="pcg", precision=64, threads=os.cpu_count() or 4)
    a32, s32 = O.render(box, 48, 48, 16, 8, seed=2580, rng="minstd", precision=32, threads=os.cpu_count() or 4)
    m64 = a64[..., :3].mean() / 16
    m32 = a32[..., :3].mean() / 16
    assert abs(m64 - m32) < 0.03 * m64
