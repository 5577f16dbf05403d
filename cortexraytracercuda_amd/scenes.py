"""Seeded synthetic scenes for the five workloads of the render hot path.

  box_scene       small box: 6 walls as 12 triangles, one area-light quad, 3 spheres
                  (Lambertian, metal fuzz 0.1, dielectric IOR 1.5), constant albedos
  large_scene     icosphere subdivided to 1,310,720 triangles displaced by 3-octave value noise
                  (amplitude 0.05) + 500 random spheres (60/30/10 % Lambertian/metal/dielectric)
  instanced_scene 4,096 instances of a 5,000-triangle displaced torus with procedural 2048^2
                  RGBA8 checker+noise textures
  sphere_mesh     displaced lat-long sphere with a requested triangle count (batched / serving)

All randomness comes from numpy's PCG64 generator seeded explicitly, and lattice noise values from
a 32-bit PCG hash, so every scene is a pure function of its arguments.
"""
import math

import numpy as np

from .scene import Scene


# ------------------------------------------------------------------ noise
def pcg_hash(v):
    """32-bit PCG hash (RXS-M-XS output of one LCG step) on uint32 arrays."""
    v = np.asarray(v, dtype=np.uint32)
    with np.errstate(over="ignore"):
        state = v * np.uint32(747796405) + np.uint32(2891336453)
        word = ((state >> ((state >> np.uint32(28)) + np.uint32(4))) ^ state) * np.uint32(277803737)
        return (word >> np.uint32(22)) ^ word


def _lattice(ix, iy, iz, seed):
    with np.errstate(over="ignore"):
        h = pcg_hash(iz.astype(np.uint32) + pcg_hash(np.uint32(seed)))
        h = pcg_hash(iy.astype(np.uint32) + h)
        h = pcg_hash(ix.astype(np.uint32) + h)
    return (h >> np.uint32(8)).astype(np.float64) / float(1 << 24)


def value_noise(p, seed):
    """Trilinear value noise in [0, 1) with smoothstep weights; p: (n, 3) float64."""
    fl = np.floor(p)
    f = p - fl
    i = fl.astype(np.int64)
    w = f * f * (3.0 - 2.0 * f)
    acc = np.zeros(p.shape[0])
    for dx in (0, 1):
        for dy in (0, 1):
            for dz in (0, 1):
                val = _lattice(i[:, 0] + dx, i[:, 1] + dy, i[:, 2] + dz, seed)
                wx = w[:, 0] if dx else 1.0 - w[:, 0]
                wy = w[:, 1] if dy else 1.0 - w[:, 1]
                wz = w[:, 2] if dz else 1.0 - w[:, 2]
                acc += wx * wy * wz * val
    return acc


def fbm3(p, seed, base_freq=4.0):
    """3 octaves (f, 2f, 4f) with weights 1, 1/2, 1/4, normalised to [0, 1)."""
    tot = np.zeros(p.shape[0])
    for k, wgt in enumerate((1.0, 0.5, 0.25)):
        tot += wgt * value_noise(p * (base_freq * (2 ** k)), seed + 7919 * k)
    return tot / 1.75


# ------------------------------------------------------------------ meshes
def icosphere(levels):
    """Unit icosphere: 20 * 4^levels triangles, shared vertices."""
    t = (1.0 + 5 ** 0.5) / 2.0
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                  [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], dtype=np.float64)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4], [11, 10, 2],
                  [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5], [2, 4, 11],
                  [6, 2, 10], [8, 6, 7], [9, 8, 1]], dtype=np.int64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    for _ in range(levels):
        e = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
        e.sort(axis=1)
        key = e[:, 0] * (v.shape[0] + 1) + e[:, 1]
        uniq, inv = np.unique(key, return_inverse=True)
        a = uniq // (v.shape[0] + 1)
        b = uniq % (v.shape[0] + 1)
        mid = v[a] + v[b]
        mid /= np.linalg.norm(mid, axis=1, keepdims=True)
        base = v.shape[0]
        v = np.concatenate([v, mid])
        n = f.shape[0]
        m01 = base + inv[:n]
        m12 = base + inv[n:2 * n]
        m20 = base + inv[2 * n:]
        f = np.concatenate([np.stack([f[:, 0], m01, m20], 1), np.stack([f[:, 1], m12, m01], 1),
                            np.stack([f[:, 2], m20, m12], 1), np.stack([m01, m12, m20], 1)])
    return v, f


def displace_sphere(v, radius, amplitude, seed):
    n = fbm3(v, seed)
    return (v * (radius + amplitude * (2.0 * n - 1.0))[:, None]).astype(np.float32)


def latlong_sphere(n_tri_target):
    """Lat-long sphere with about n_tri_target triangles (2 * n_phi * (n_theta - 1))."""
    n_theta = max(3, int(round(math.sqrt(n_tri_target / 4.0))) + 1)
    n_phi = max(3, int(round(n_tri_target / (2.0 * (n_theta - 1)))))
    th = np.linspace(0.0, math.pi, n_theta + 1)[1:-1]
    ph = np.linspace(0.0, 2 * math.pi, n_phi, endpoint=False)
    T, P = np.meshgrid(th, ph, indexing="ij")
    ring = np.stack([np.sin(T) * np.cos(P), np.cos(T), np.sin(T) * np.sin(P)], -1).reshape(-1, 3)
    v = np.concatenate([[[0, 1, 0]], ring, [[0, -1, 0]]]).astype(np.float64)
    nr = n_theta - 1
    faces = []
    j = np.arange(n_phi)
    jn = (j + 1) % n_phi
    faces.append(np.stack([np.zeros(n_phi, np.int64), 1 + jn, 1 + j], 1))
    for r in range(nr - 1):
        a = 1 + r * n_phi + j
        b = 1 + r * n_phi + jn
        c = 1 + (r + 1) * n_phi + j
        d = 1 + (r + 1) * n_phi + jn
        faces.append(np.stack([a, b, d], 1))
        faces.append(np.stack([a, d, c], 1))
    last = v.shape[0] - 1
    faces.append(np.stack([np.full(n_phi, last), 1 + (nr - 1) * n_phi + j, 1 + (nr - 1) * n_phi + jn], 1))
    return v, np.concatenate(faces)


def torus(R=1.0, r=0.3, nu=50, nv=50):
    """Torus grid: 2 * nu * nv triangles, positions, per-vertex uvs (u around R, v around r)."""
    u = np.arange(nu) / nu
    v = np.arange(nv) / nv
    U, Vv = np.meshgrid(u, v, indexing="ij")
    a, b = 2 * math.pi * U, 2 * math.pi * Vv
    pos = np.stack([(R + r * np.cos(b)) * np.cos(a), r * np.sin(b), (R + r * np.cos(b)) * np.sin(a)], -1).reshape(-1, 3)
    nrm = np.stack([np.cos(b) * np.cos(a), np.sin(b), np.cos(b) * np.sin(a)], -1).reshape(-1, 3)
    uv = np.stack([U, Vv], -1).reshape(-1, 2)
    i = np.arange(nu)[:, None]
    j = np.arange(nv)[None, :]
    p00 = (i * nv + j).reshape(-1)
    p10 = (((i + 1) % nu) * nv + j).reshape(-1)
    p01 = (i * nv + (j + 1) % nv).reshape(-1)
    p11 = (((i + 1) % nu) * nv + (j + 1) % nv).reshape(-1)
    f = np.concatenate([np.stack([p00, p10, p11], 1), np.stack([p00, p11, p01], 1)])
    return pos, nrm, uv.astype(np.float32), f


def random_rotation(rng):
    """Uniform random rotation (Shoemake's uniform quaternion)."""
    u1, u2, u3 = rng.random(3)
    q = np.array([math.sqrt(1 - u1) * math.sin(2 * math.pi * u2), math.sqrt(1 - u1) * math.cos(2 * math.pi * u2),
                  math.sqrt(u1) * math.sin(2 * math.pi * u3), math.sqrt(u1) * math.cos(2 * math.pi * u3)])
    x, y, z, w = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def checker_noise_texture(size, seed, cells=16):
    """size x size RGBA8: two-colour checker modulated by value noise."""
    rng = np.random.default_rng(seed)
    c0 = rng.uniform(0.2, 0.9, 3)
    c1 = rng.uniform(0.05, 0.5, 3)
    y, x = np.mgrid[0:size, 0:size]
    chk = ((x * cells // size) + (y * cells // size)) % 2
    p = np.stack([x.reshape(-1) / size * 8.0, y.reshape(-1) / size * 8.0, np.full(size * size, 0.5)], 1)
    n = value_noise(p, seed).reshape(size, size)
    base = np.where(chk[..., None] == 0, c0, c1) * (0.6 + 0.4 * n[..., None])
    img = np.empty((size, size, 4), np.uint8)
    img[..., :3] = np.clip(base * 255.0, 0, 255).astype(np.uint8)
    img[..., 3] = 255
    return img


# ------------------------------------------------------------------ workloads
def box_scene(coloured_walls=True):
    """Small box scene: closed box [-1,1]^3 as 12 triangles, area light centred at (0, 0.99, 0)
    facing down (u x v = -y), spheres of radius 0.3 resting on the floor; camera inside the box
    at (0, 0, 0.9) looking down -z, vfov 70, aspect 1."""
    s = Scene()
    s.set_camera((0.0, 0.0, 0.9), (0.0, 0.0, -1.0), (0.0, 1.0, 0.0), vfov=70.0, aspect_ratio=1.0, aperture=0.0,
                 focus_dist=10.0)
    s.background = (0.0, 0.0, 0.0)
    white = s.lambertian("white", (0.73, 0.73, 0.73))
    red = s.lambertian("red", (0.65, 0.05, 0.05)) if coloured_walls else white
    green = s.lambertian("green", (0.12, 0.45, 0.15)) if coloured_walls else white
    s.lambertian("diffuse_ball", (0.8, 0.3, 0.3))
    s.metal("metal_ball", (0.8, 0.85, 0.88), fuzz=0.1)
    s.dielectric("glass_ball", 1.5)
    c = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1], [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]],
                 dtype=np.float32)
    # two triangles per wall: floor, ceiling, back, front, left, right
    faces = np.array([[0, 1, 5], [0, 5, 4], [3, 7, 6], [3, 6, 2], [0, 3, 2], [0, 2, 1], [4, 5, 6], [4, 6, 7],
                      [0, 4, 7], [0, 7, 3], [1, 2, 6], [1, 6, 5]], dtype=np.uint32)
    mats = np.array([white, white, white, white, white, white, white, white, red, red, green, green], dtype=np.uint32)
    s.mesh(c, faces, white, tri_materials=mats)
    s.area_light((0.0, 0.99, 0.0), (-0.5, 0.0, 0.0), (0.0, 0.0, -0.5), rgb=(15.0, 15.0, 15.0), intensity=1.0,
                 invisible=True)
    s.sphere((-0.55, -0.7, -0.45), 0.3, "diffuse_ball")
    s.sphere((0.05, -0.7, -0.2), 0.3, "metal_ball")
    s.sphere((0.6, -0.7, -0.5), 0.3, "glass_ball")
    return s


def shapes_scene():
    """The box scene's walls and light with analytic shapes in place of the spheres: a checkered
    disk on the floor, an image-textured cylinder, a metal cone, a glass tube lying on its side
    (rotated about z, translated) and an elliptical disk on the back wall (rotated, translated and
    scaled non-uniformly)."""
    s = box_scene()
    s.spheres = []
    s.checker_texture("floor_check", 0.05, (0.9, 0.9, 0.9), (0.1, 0.2, 0.6))
    s.image_texture("noise", checker_noise_texture(64, seed=7, cells=8))
    s.lambertian("disk_mat", texture="floor_check")
    s.lambertian("tube_mat", texture="noise")
    s.metal("cone_mat", (0.85, 0.8, 0.7), fuzz=0.05)
    s.dielectric("glass_tube", 1.5)
    s.disk((-0.5, -0.99, -0.4), 0.35, 0.02, "disk_mat", mapping=(1, 1, 0, 0, 2, 2))
    s.cylinder((0.0, -1.0, -0.5), 0.25, 0.6, "tube_mat", mapping=(1, 1, 0, 0, 2, 1))
    s.cone((0.55, -1.0, -0.3), 0.3, 0.7, "cone_mat")
    s.cylinder((0.0, 0.0, 0.0), 0.12, 0.7, "glass_tube", rotate_deg=(0, 0, 90), translate=(0.3, -0.2, -0.1))
    s.disk((0.0, 0.0, 0.0), 0.2, 0.01, "red", rotate_deg=(90, 0, 0), translate=(0.2, -0.98, -0.3),
           scale=(1.5, 1.0, 1.0))
    return s


def large_scene(seed=1, levels=8, n_spheres=500, aspect=16.0 / 9.0):
    """Icosphere of radius 3 subdivided `levels` times (1,310,720 triangles at 8) displaced by
    3-octave value noise of amplitude 0.05, plus `n_spheres` spheres with centres U([-10,10]^3),
    radius U(0.1, 0.5), 60/30/10 % Lambertian/metal/dielectric. One area light above the scene and
    a sky background provide the illumination."""
    rng = np.random.default_rng(seed)
    s = Scene()
    s.set_camera((0.0, 4.0, 24.0), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), vfov=45.0, aspect_ratio=aspect, focus_dist=10.0)
    s.background = (0.7, 0.8, 1.0)
    body = s.lambertian("body", (0.62, 0.6, 0.58))
    v, f = icosphere(levels)
    s.mesh(displace_sphere(v, 3.0, 0.05, seed), f.astype(np.uint32), body)
    for k in range(n_spheres):
        c = rng.uniform(-10.0, 10.0, 3)
        r = rng.uniform(0.1, 0.5)
        u = rng.random()
        if u < 0.6:
            m = s.lambertian(None, tuple(rng.uniform(0.1, 0.9, 3)))
        elif u < 0.9:
            m = s.metal(None, tuple(rng.uniform(0.1, 0.9, 3)), fuzz=float(rng.uniform(0.0, 0.5)))
        else:
            m = s.dielectric(None, 1.5)
        s.sphere(tuple(c), float(r), m)
    s.area_light((0.0, 14.0, 0.0), (10.0, 0.0, 0.0), (0.0, 0.0, 10.0), rgb=(6.0, 6.0, 6.0), intensity=1.0, invisible=True)
    return s


def instanced_scene(seed=1, n_instances=4096, n_textures=4, tex_size=2048, nu=50, nv=50, extent=50.0,
                    aspect=16.0 / 9.0):
    """n_instances random rigid placements (uniform rotation, translation in [-extent, extent]^3)
    of a displaced torus (2*nu*nv triangles) with image textures."""
    rng = np.random.default_rng(seed)
    s = Scene()
    s.set_camera((0.0, 10.0, 120.0), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), vfov=60.0, aspect_ratio=aspect, focus_dist=10.0)
    s.background = (0.6, 0.7, 0.9)
    pos, nrm, uv, f = torus(1.0, 0.3, nu, nv)
    n = fbm3(pos, seed + 17)
    pos = pos + (0.05 * (2.0 * n - 1.0))[:, None] * nrm
    mats = []
    for k in range(n_textures):
        tex = s.image_texture("tex%d" % k, checker_noise_texture(tex_size, seed * 1000 + k))
        mats.append(s.lambertian("tor%d" % k, (1, 1, 1), texture=tex))
    mesh = s.mesh(pos.astype(np.float32), f.astype(np.uint32), mats[0], uvs=uv)
    for k in range(n_instances):
        R = random_rotation(rng) * rng.uniform(1.0, 2.5)
        t = rng.uniform(-extent, extent, 3)
        xf = np.concatenate([R, t[:, None]], 1)
        s.instance(mesh, xf, material_override=mats[k % n_textures])
    s.area_light((0.0, 80.0, 0.0), (60.0, 0.0, 0.0), (0.0, 0.0, 60.0), rgb=(4.0, 4.0, 4.0), intensity=1.0, invisible=True)
    return s


def sphere_mesh_scene(n_tris, seed=1, aspect=1.0):
    """Box-scene lighting around a displaced lat-long sphere of about n_tris triangles."""
    s = box_scene()
    s.set_camera((0.0, 0.0, 0.9), (0.0, 0.0, -1.0), (0.0, 1.0, 0.0), vfov=70.0, aspect_ratio=aspect, focus_dist=10.0)
    v, f = latlong_sphere(max(64, int(n_tris)))
    p = displace_sphere(v, 0.45, 0.05, seed) + np.array([0.0, -0.2, -0.45], np.float32)
    m = s.lambertian(None, (0.5, 0.6, 0.7))
    s.mesh(p, f.astype(np.uint32), m)
    return s
