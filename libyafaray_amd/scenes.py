"""Seeded synthetic scenes and ray batches of the benchmark workloads.

Every generator is deterministic in its seed (numpy PCG64) and returns plain
float32 / uint32 arrays, so the GPU library and the CPU oracle see identical
inputs.

Workloads
  cornell_random_tris  -- Cornell-style box of 5 unit walls (10 triangles)
                          plus n random small triangles inside it.
  heightfield_scene    -- a res x res fractal value-noise heightfield
                          (6 octaves, amplitude 0.2 over the unit square,
                          2 (res-1)^2 triangles) plus n random triangles
                          above it; the ray-cast microbenchmark geometry.
  camera_rays          -- pinhole camera rays, one per pixel centre, built as
                          PerspectiveCamera::shootRay does
                          (src/camera/camera.cc:89-110,
                          src/camera/camera_perspective.cc:97-111,165-180).
  incoherent_rays      -- origins uniform in a box, directions uniform on S^2.
"""
from __future__ import annotations

import functools

import numpy as np

F32 = np.float32
NO_VERTEX = 0xFFFFFFFF  # 4th index of a triangle in a face list


def _hash2(ix: np.ndarray, iy: np.ndarray, seed: int) -> np.ndarray:
    """Integer lattice hash -> [0, 1) (splitmix64 finaliser)."""
    with np.errstate(over="ignore"):
        h = (ix.astype(np.uint64) * np.uint64(0x9E3779B97F4A7C15)) ^ \
            (iy.astype(np.uint64) * np.uint64(0xC2B2AE3D27D4EB4F)) ^ np.uint64(seed * 0x165667B19E3779F9 & (2**64 - 1))
        h ^= h >> np.uint64(30)
        h *= np.uint64(0xBF58476D1CE4E5B9)
        h ^= h >> np.uint64(27)
        h *= np.uint64(0x94D049BB133111EB)
        h ^= h >> np.uint64(31)
    return (h >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))


def value_noise(x: np.ndarray, y: np.ndarray, seed: int) -> np.ndarray:
    """Bilinear value noise with a smoothstep fade over the integer lattice of
    _hash2.  The lattice values are hashed once per lattice point of the
    covered range and gathered, which gives the same numbers as hashing every
    corner of every sample."""
    ix = np.floor(x)
    iy = np.floor(y)
    fx = x - ix
    fy = y - iy
    ix = ix.astype(np.int64)
    iy = iy.astype(np.int64)
    sx = fx * fx * (3.0 - 2.0 * fx)
    sy = fy * fy * (3.0 - 2.0 * fy)
    x0, y0 = int(ix.min()), int(iy.min())
    lx, ly = np.meshgrid(np.arange(x0, int(ix.max()) + 2), np.arange(y0, int(iy.max()) + 2), indexing="xy")
    lattice = _hash2(lx, ly, seed)
    cx, cy = ix - x0, iy - y0
    if x.ndim == 2 and y.ndim == 2 and x.shape[0] == 1 and y.shape[1] == 1:
        # x along the columns, y along the rows (a separable grid): gather rows, then columns
        r0, r1 = lattice[cy[:, 0]], lattice[cy[:, 0] + 1]
        v00, v10 = r0[:, cx[0]], r0[:, cx[0] + 1]
        v01, v11 = r1[:, cx[0]], r1[:, cx[0] + 1]
    else:
        v00 = lattice[cy, cx]
        v10 = lattice[cy, cx + 1]
        v01 = lattice[cy + 1, cx]
        v11 = lattice[cy + 1, cx + 1]
    a = v00 + (v10 - v00) * sx
    b = v01 + (v11 - v01) * sx
    return a + (b - a) * sy


def fractal_heights(res: int, octaves: int = 6, amplitude: float = 0.2, base_freq: float = 4.0,
                    seed: int = 0) -> np.ndarray:
    t = np.linspace(0.0, 1.0, res)
    x, y = t[None, :], t[:, None]  # the "xy" meshgrid, broadcast
    h = np.zeros((res, res))
    amp = 1.0
    freq = base_freq
    for o in range(octaves):
        h += amp * value_noise(x * freq, y * freq, seed * 131 + o)
        amp *= 0.5
        freq *= 2.0
    h = (h - h.min()) / max(h.max() - h.min(), 1e-30) * amplitude
    return h


def grid_mesh(heights: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """Vertices over [0,1]^2 with z = heights; two triangles per cell."""
    res = heights.shape[0]
    t = np.linspace(0.0, 1.0, res)
    x, y = np.meshgrid(t, t, indexing="xy")
    verts = np.stack([x, y, heights], axis=-1).reshape(-1, 3).astype(F32)
    i, j = np.meshgrid(np.arange(res - 1), np.arange(res - 1), indexing="ij")
    v00 = (i * res + j).reshape(-1)
    v01 = v00 + 1
    v10 = v00 + res
    v11 = v10 + 1
    tris = np.empty((2 * v00.size, 3), dtype=np.uint32)
    tris[0::2] = np.stack([v00, v01, v11], axis=-1)
    tris[1::2] = np.stack([v00, v11, v10], axis=-1)
    return verts, tris


def grid_quads(heights: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """The heightfield of grid_mesh with one quad face per cell, (v00, v01,
    v11, v10) (non-planar where the cell's corners are), as faces of 4 indices."""
    res = heights.shape[0]
    verts, _ = grid_mesh(heights)
    i, j = np.meshgrid(np.arange(res - 1), np.arange(res - 1), indexing="ij")
    v00 = (i * res + j).reshape(-1)
    faces = np.stack([v00, v00 + 1, v00 + res + 1, v00 + res], axis=-1).astype(np.uint32)
    return verts, faces


def quad_face_scene(res: int = 64, n_random: int = 2000, seed: int = 0):
    """Faces of 4 indices (4th NO_VERTEX for a triangle): the heightfield as
    quads, random planar quads (parallelograms) and random slightly bent quads
    above it, and random triangles."""
    rng = np.random.default_rng(seed)
    hv, hf = grid_quads(fractal_heights(res, seed=seed))
    n = n_random
    c = rng.uniform((0.0, 0.0, 0.22), (1.0, 1.0, 0.5), size=(n, 3))
    a = unit_vectors(rng, n) * rng.uniform(0.01, 0.03, (n, 1))
    b = unit_vectors(rng, n) * rng.uniform(0.01, 0.03, (n, 1))
    q = np.stack([c - a - b, c + a - b, c + a + b, c - a + b], 1)
    q[1::2, 3] += unit_vectors(rng, (n + 1) // 2)[: q[1::2].shape[0]] * 0.004
    qv = q.reshape(-1, 3).astype(F32)
    qf = np.arange(4 * n, dtype=np.uint32).reshape(n, 4)
    qf[::4, 3] = NO_VERTEX  # a quarter of them triangles (v0, v1, v2)
    verts = np.concatenate([hv, qv]).astype(F32)
    faces = np.concatenate([hf, qf + np.where(qf == NO_VERTEX, 0, hv.shape[0]).astype(np.uint32)])
    return verts, faces.astype(np.uint32)


def unit_vectors(rng: np.random.Generator, n: int) -> np.ndarray:
    z = rng.uniform(-1.0, 1.0, n)
    phi = rng.uniform(0.0, 2.0 * np.pi, n)
    r = np.sqrt(np.maximum(0.0, 1.0 - z * z))
    return np.stack([r * np.cos(phi), r * np.sin(phi), z], axis=-1)


def random_triangles(rng: np.random.Generator, n: int, lo, hi, edge=(0.01, 0.03)) -> np.ndarray:
    """n triangles (n, 3 vertices, xyz): centroid uniform in the box [lo, hi],
    each vertex at a distance U(edge) from the centroid along a random unit
    direction."""
    c = np.column_stack([rng.uniform(lo[k], hi[k], n) for k in range(3)])
    out = np.empty((n, 3, 3))
    for k in range(3):
        length = rng.uniform(edge[0], edge[1], n)[:, None]
        out[:, k, :] = c + length * unit_vectors(rng, n)
    return out.astype(F32)


def merge(meshes: list[tuple[np.ndarray, np.ndarray]]) -> tuple[np.ndarray, np.ndarray]:
    verts, tris, off = [], [], 0
    for v, t in meshes:
        verts.append(v)
        tris.append(t.astype(np.int64) + off)
        off += v.shape[0]
    return np.concatenate(verts).astype(F32), np.concatenate(tris).astype(np.uint32)


def soup(tri_verts: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    n = tri_verts.shape[0]
    return tri_verts.reshape(-1, 3).astype(F32), np.arange(3 * n, dtype=np.uint32).reshape(n, 3)


def heightfield_scene(res: int = 1024, n_random: int = 100_000, seed: int = 0):
    """Ray-cast microbenchmark geometry: the heightfield plus n_random triangles
    with centroids in [0,1]^2 x [0.22, 0.5], above the surface."""
    rng = np.random.default_rng(seed)
    hf = grid_mesh(fractal_heights(res, seed=seed))
    if not n_random:
        return hf
    rnd = soup(random_triangles(rng, n_random, (0.0, 0.0, 0.22), (1.0, 1.0, 0.5)))
    return merge([hf, rnd])


def cornell_walls() -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    """5 axis-aligned unit walls of [0,1]^3 (open at y = 0), 2 triangles each.
    Returns verts, tris and a wall index per triangle (0 floor, 1 ceiling,
    2 back, 3 left, 4 right)."""
    quads = [
        [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0)],  # floor z=0
        [(0, 0, 1), (0, 1, 1), (1, 1, 1), (1, 0, 1)],  # ceiling z=1
        [(0, 1, 0), (1, 1, 0), (1, 1, 1), (0, 1, 1)],  # back y=1
        [(0, 0, 0), (0, 1, 0), (0, 1, 1), (0, 0, 1)],  # left x=0
        [(1, 0, 0), (1, 0, 1), (1, 1, 1), (1, 1, 0)],  # right x=1
    ]
    verts = np.array([p for q in quads for p in q], dtype=F32)
    tris = []
    wall = []
    for w in range(5):
        b = 4 * w
        tris += [[b, b + 1, b + 2], [b, b + 2, b + 3]]
        wall += [w, w]
    return verts, np.array(tris, dtype=np.uint32), np.array(wall, dtype=np.int32)


def cornell_random_tris(n_random: int = 10_000, seed: int = 0):
    rng = np.random.default_rng(seed)
    wv, wt, _ = cornell_walls()
    rnd = soup(random_triangles(rng, n_random, (0.2, 0.2, 0.2), (0.8, 0.8, 0.8)))
    return merge([(wv, wt), rnd])


def _normalize32(v: np.ndarray) -> np.ndarray:
    """Vec::normalize on float32 rows: v * (1 / sqrt(x^2 + y^2 + z^2))."""
    v = v.astype(F32)
    l2 = (v[..., 0] * v[..., 0] + v[..., 1] * v[..., 1]) + v[..., 2] * v[..., 2]
    inv = (F32(1.0) / np.sqrt(l2)).astype(F32)
    return (v * inv[..., None]).astype(F32)


def camera_rays(resx: int, resy: int, eye, target, up_point, focal: float) -> np.ndarray:
    """Pinhole rays through pixel centres (x + 0.5, y + 0.5), laid out row-major."""
    eye = np.asarray(eye, dtype=F32)
    cam_y = (np.asarray(up_point, dtype=F32) - eye).astype(F32)
    cam_z = (np.asarray(target, dtype=F32) - eye).astype(F32)
    cam_x = np.cross(cam_z, cam_y).astype(F32)
    cam_y = np.cross(cam_z, cam_x).astype(F32)
    cam_x, cam_y, cam_z = (_normalize32(a) for a in (cam_x, cam_y, cam_z))
    aspect = F32(resy) / F32(resx)
    vright = cam_x
    vup = (aspect * cam_y).astype(F32)
    vto = (cam_z * F32(focal) - F32(0.5) * (vup + vright)).astype(F32)
    vup = (vup / F32(resy)).astype(F32)
    vright = (vright / F32(resx)).astype(F32)
    px = (np.arange(resx, dtype=F32) + F32(0.5))
    py = (np.arange(resy, dtype=F32) + F32(0.5))
    PX, PY = np.meshgrid(px, py, indexing="xy")
    d = (vright[None, None, :] * PX[..., None] + vup[None, None, :] * PY[..., None]).astype(F32) + vto
    d = _normalize32(d.reshape(-1, 3))
    rays = np.zeros((d.shape[0], 8), dtype=F32)
    rays[:, 0:3] = eye
    rays[:, 3] = 0.0
    rays[:, 4:7] = d
    rays[:, 7] = -1.0
    return rays


def incoherent_rays(rng: np.random.Generator, n: int, lo, hi, tmax=None) -> np.ndarray:
    lo = np.asarray(lo, dtype=np.float64)
    hi = np.asarray(hi, dtype=np.float64)
    rays = np.zeros((n, 8), dtype=F32)
    rays[:, 0:3] = rng.uniform(lo, hi, size=(n, 3))
    rays[:, 4:7] = unit_vectors(rng, n)
    rays[:, 3] = 0.0
    rays[:, 7] = -1.0 if tmax is None else tmax
    return rays


def shadow_tmax(rng: np.random.Generator, n: int) -> np.ndarray:
    """tmax uniform in (0, 1]."""
    return (1.0 - rng.uniform(0.0, 1.0, n)).astype(F32)


# Ray-cast microbenchmark (16M rays on the ~2.2M-triangle heightfield scene).
RAYCAST_CAMERA = dict(eye=(0.5, -0.45, 0.75), target=(0.5, 0.55, 0.05), up_point=(0.5, -0.45, 1.75),
                      focal=1.0 / (2.0 * np.tan(np.radians(20.0))))


def raycast_workload(res: int = 1024, n_random: int = 100_000, cam_res=(4096, 2048), n_incoherent=None,
                     seed: int = 0):
    """Geometry plus coherent, incoherent and shadow ray batches."""
    verts, tris = heightfield_scene(res, n_random, seed)
    rng = np.random.default_rng(seed + 1)
    coherent = camera_rays(cam_res[0], cam_res[1], **RAYCAST_CAMERA)
    n_inc = coherent.shape[0] if n_incoherent is None else n_incoherent
    lo = verts.min(axis=0)
    hi = verts.max(axis=0)
    incoherent = incoherent_rays(rng, n_inc, lo, hi)
    shadow = incoherent.copy()
    shadow[:, 7] = shadow_tmax(rng, n_inc)
    return verts, tris, coherent, incoherent, shadow


# ---------------------------------------------------------------- path tracing
def cornell_path_scene(n_random: int = 10_000, seed: int = 0, res: int = 256, spp: int = 16):
    """Cornell-style box + random triangles lit by one 0.2 x 0.2 ceiling area light
    of radiance 10 (256 x 256, 16 spp, 5 bounces, Russian roulette from depth 3).

    Materials are shinydiffuse (Lambert): walls 0.7 grey, left wall red, right
    wall green, each random triangle its own grey albedo U[0.2, 0.8].  The area
    light emits colour * power * pi (src/light/light_area.h:78) and its pdf
    carries the quad area through the unnormalised normal (:76), so radiance L
    corresponds to colour * power = L / pi with the reference's pi-free Lambert.
    It is a pure light (no mesh), so camera and path rays do not see it.
    """
    rng = np.random.default_rng(seed)
    wv, wt, wall = cornell_walls()
    tv = random_triangles(rng, n_random, (0.2, 0.2, 0.2), (0.8, 0.8, 0.8))
    verts, tris = merge([(wv, wt), soup(tv)])
    wall_mat = np.array([0, 0, 0, 1, 2])  # floor, ceiling, back grey; left red; right green
    mat_id = np.concatenate([wall_mat[wall], 3 + np.arange(n_random)]).astype(np.uint32)
    albedo = rng.uniform(0.2, 0.8, n_random)
    materials = [dict(color=(0.7, 0.7, 0.7)), dict(color=(0.7, 0.1, 0.1)), dict(color=(0.1, 0.7, 0.1))]
    materials += [dict(color=(a, a, a)) for a in albedo]
    lights = [dict(type="area", corner=(0.4, 0.4, 0.999), point1=(0.4, 0.6, 0.999), point2=(0.6, 0.4, 0.999),
                   color=(1.0, 1.0, 1.0), power=10.0 / np.pi, samples=4)]
    camera = dict(from_=(0.5, -1.6, 0.5), to=(0.5, 0.5, 0.5), up=(0.5, -1.6, 1.5),
                  focal=float(1.0 / (2.0 * np.tan(np.radians(20.0)))), resx=res, resy=res)
    settings = dict(path_samples=1, bounces=5, rr_min_bounces=2, aa_samples=spp, aa_pixelwidth=1.5,
                    background=None, seed=0)
    return dict(verts=verts, tris=tris, mat_id=mat_id, materials=materials, lights=lights, camera=camera,
                settings=settings)


def quad_tris(c, e1, e2) -> tuple[np.ndarray, np.ndarray]:
    """Quad c, c + e1, c + e1 + e2, c + e2 as two triangles whose face normal is
    e1 x e2."""
    c, e1, e2 = (np.asarray(v, dtype=F32) for v in (c, e1, e2))
    verts = np.array([c, c + e1, c + e1 + e2, c + e2], dtype=F32)
    return verts, np.array([[0, 1, 2], [0, 2, 3]], dtype=np.uint32)


def box_tris(lo, hi) -> tuple[np.ndarray, np.ndarray]:
    """Axis-aligned box as 12 triangles with outward face normals."""
    lo = np.asarray(lo, dtype=F32)
    s = np.asarray(hi, dtype=F32) - lo
    x, y, z = (np.array(v, dtype=F32) for v in ((s[0], 0, 0), (0, s[1], 0), (0, 0, s[2])))
    faces = [(lo, y, x), (lo + z, x, y), (lo, z, y), (lo + x, y, z), (lo, x, z), (lo + y, z, x)]
    return merge([quad_tris(c, e1, e2) for c, e1, e2 in faces])


def cornell_materials_scene(n_random: int = 10_000, seed: int = 0, res: int = 256, spp: int = 16,
                            glossy_branches: float = 0.0, raydepth: int = 5, point_lights: int = 2):
    """The Cornell box of cornell_path_scene with every material of the path:
    random triangles drawn 40 % shinydiffuse / 30 % glossy (roughness r ~ U[0.05, 0.5],
    Blinn exponent 2 / r^2 - 2) / 15 % mirror / 15 % glass (IOR 1.5); a glass
    cube, a mirror panel before the back wall, a light_mat quad under the ceiling
    light, point lights and a constant background of 0.1.  A fraction
    `glossy_branches` of the glossy triangles has as_diffuse = false, so they
    recurse through 8 glossy branches instead of being sampled as diffuse."""
    rng = np.random.default_rng(seed)
    wv, wt, wall = cornell_walls()
    tv = random_triangles(rng, n_random, (0.2, 0.2, 0.2), (0.8, 0.8, 0.8))
    # the cube floats 1/512 above the floor: coplanar faces would make exact
    # ties in t, which the kd-tree and the BVH break differently
    cube = box_tris((0.55, 0.3, 1.0 / 512), (0.8, 0.55, 0.25))
    panel = quad_tris((0.1, 0.95, 0.1), (0.4, 0.0, 0.0), (0.0, 0.0, 0.6))
    lamp = quad_tris((0.45, 0.45, 0.98), (0.0, 0.1, 0.0), (0.1, 0.0, 0.0))
    verts, tris = merge([(wv, wt), soup(tv), cube, panel, lamp])
    materials = [dict(color=(0.7, 0.7, 0.7)), dict(color=(0.7, 0.1, 0.1)), dict(color=(0.1, 0.7, 0.1)),
                 dict(type="glass", IOR=1.5, filter_color=(0.9, 0.95, 1.0), transmit_filter=0.3),
                 dict(type="mirror", color=(0.95, 0.95, 0.9), reflect=0.9),
                 dict(type="light_mat", color=(1.0, 0.9, 0.8), power=3.0)]
    wall_mat = np.array([0, 0, 0, 1, 2])
    kind = rng.uniform(0.0, 1.0, n_random)
    for k in range(n_random):
        u = kind[k]
        a = float(rng.uniform(0.2, 0.8))
        if u < 0.4:
            materials.append(dict(color=(a, a, a)))
        elif u < 0.7:
            r = float(rng.uniform(0.05, 0.5))
            m = dict(type="glossy", color=(a, 0.9 * a, 0.8 * a), exponent=2.0 / (r * r) - 2.0,
                     diffuse_reflect=float(rng.uniform(0.0, 0.6)), diffuse_color=(0.8 * a, a, a),
                     glossy_reflect=float(rng.uniform(0.2, 1.0)))
            if rng.uniform() < glossy_branches:
                m["as_diffuse"] = False
            materials.append(m)
        elif u < 0.85:
            materials.append(dict(type="mirror", color=(a, a, 0.9 * a), reflect=float(rng.uniform(0.5, 1.0))))
        else:
            materials.append(dict(type="glass", IOR=1.5))
    mat_id = np.concatenate([wall_mat[wall], 6 + np.arange(n_random), np.full(12, 3), np.full(2, 4),
                             np.full(2, 5)]).astype(np.uint32)
    lights = [dict(type="area", corner=(0.4, 0.4, 0.999), point1=(0.4, 0.6, 0.999), point2=(0.6, 0.4, 0.999),
                   color=(1.0, 1.0, 1.0), power=10.0 / np.pi, samples=4)]
    for k in range(point_lights):
        p = rng.uniform(0.15, 0.85, 3)
        lights.append(dict(type="point", **{"from": (float(p[0]), float(p[1]), float(p[2]))},
                           color=(1.0, 0.95, 0.9), power=float(np.exp(rng.uniform(np.log(0.01), np.log(0.1))))))
    camera = dict(from_=(0.5, -1.6, 0.5), to=(0.5, 0.5, 0.5), up=(0.5, -1.6, 1.5),
                  focal=float(1.0 / (2.0 * np.tan(np.radians(20.0)))), resx=res, resy=res)
    settings = dict(path_samples=1, bounces=5, rr_min_bounces=2, aa_samples=spp, aa_pixelwidth=1.5,
                    background=(0.1, 0.1, 0.1), seed=0, raydepth=raydepth)
    return dict(verts=verts, tris=tris, mat_id=mat_id, materials=materials, lights=lights, camera=camera,
                settings=settings)


def grid_quad_tris(c, e1, e2, n1: int, n2: int, skip=None) -> tuple[np.ndarray, np.ndarray]:
    """Quad c + [0,1] e1 + [0,1] e2 tessellated into n1 x n2 cells (2 triangles
    each, face normal e1 x e2); skip(i, j) -> True leaves cell (i, j) out."""
    c, e1, e2 = (np.asarray(v, dtype=np.float64) for v in (c, e1, e2))
    i = np.arange(n1 + 1)[:, None, None] / n1
    j = np.arange(n2 + 1)[None, :, None] / n2
    verts = (c + i * e1 + j * e2).reshape(-1, 3).astype(F32)
    cells = [(a, b) for a in range(n1) for b in range(n2) if skip is None or not skip(a, b)]
    ii = np.array([a for a, _ in cells], dtype=np.int64)
    jj = np.array([b for _, b in cells], dtype=np.int64)
    v00 = ii * (n2 + 1) + jj
    v10 = (ii + 1) * (n2 + 1) + jj
    v11 = (ii + 1) * (n2 + 1) + jj + 1
    v01 = ii * (n2 + 1) + jj + 1
    tris = np.concatenate([np.stack([v00, v10, v11], -1), np.stack([v00, v11, v01], -1)]).astype(np.uint32)
    return verts, tris


def occluded_room_scene(res=(1280, 720), spp: int = 16, seed: int = 0, wall_cells: int = 180,
                        n_boxes: int = 150, bounces: int = 10, slit: float = 0.1):
    """Learned-guiding workload: a 4 x 3 x 3 closed room (walls tessellated, plus
    seeded box furniture, about 500k triangles) lit only through a `slit`-wide
    door slit from an adjacent 2 x 3 x 3 room holding a 1 x 1 ceiling area light
    of radiance 50.  All surfaces Lambertian; camera in the main room."""
    rng = np.random.default_rng(seed)
    parts, mats = [], []
    W, D, H = 4.0, 3.0, 3.0          # main room: x in [0, 4], y in [0, 3], z in [0, 3]
    n = wall_cells
    inward = [  # (corner, e1, e2) with e1 x e2 pointing into the room
        ((0, 0, 0), (0, D, 0), (W, 0, 0)),        # floor (normal +z)
        ((0, 0, H), (W, 0, 0), (0, D, 0)),        # ceiling (-z)
        ((0, 0, 0), (W, 0, 0), (0, 0, H)),        # y = 0 wall (+y)
        ((0, D, 0), (0, 0, H), (W, 0, 0)),        # y = D wall (-y)
        ((0, 0, 0), (0, 0, H), (0, D, 0)),        # x = 0 wall (+x)
    ]
    for c, e1, e2 in inward:
        parts.append(grid_quad_tris(c, e1, e2, n, n))
        mats.append(0)
    # dividing wall x = W with a door slit: y in [1.45, 1.45 + slit], z in [0, 2]
    y0, y1 = 1.5 - slit / 2, 1.5 + slit / 2
    nd = n
    def in_slit(i, j):
        ya, yb = i * D / nd, (i + 1) * D / nd
        za = j * H / nd
        return yb > y0 and ya < y1 and za < 2.0
    # the slit edges must fall on cell boundaries: cells of D / nd wide
    parts.append(grid_quad_tris((W, 0, 0), (0, D, 0), (0, 0, H), nd, nd, skip=in_slit))
    mats.append(0)
    # adjacent room x in [W, W + 2]
    W2 = 2.0
    m = max(8, (2 * n) // 3)
    adj = [
        ((W, 0, 0), (0, D, 0), (W2, 0, 0)),
        ((W, 0, H), (W2, 0, 0), (0, D, 0)),
        ((W, 0, 0), (W2, 0, 0), (0, 0, H)),
        ((W, D, 0), (0, 0, H), (W2, 0, 0)),
        ((W + W2, 0, 0), (0, D, 0), (0, 0, H)),
    ]
    for c, e1, e2 in adj:
        parts.append(grid_quad_tris(c, e1, e2, m, m))
        mats.append(0)
    # box furniture in the main room, clear of the camera: most on the floor, some shelves
    for k in range(n_boxes):
        sx, sy, sz = rng.uniform(0.1, 0.6, 3)
        x0 = rng.uniform(1.0, W - 0.3 - sx)
        y_ = rng.uniform(0.1, D - 0.1 - sy)
        z0 = 0.0 if rng.uniform() < 0.8 else rng.uniform(0.5, 2.0)
        parts.append(box_tris((x0, y_, z0 + 1.0 / 512), (x0 + sx, y_ + sy, z0 + sz)))
        mats.append(1 + k)
    verts, tris = merge(parts)
    mat_id = np.concatenate([np.full(len(t), mm, np.uint32) for (_, t), mm in zip(parts, mats)])
    albedo = rng.uniform(0.3, 0.8, n_boxes)
    materials = [dict(color=(0.7, 0.7, 0.7))] + [dict(color=(a, 0.9 * a, 0.8 * a)) for a in albedo]
    lights = [dict(type="area", corner=(W + 0.5, 1.0, H - 0.001), point1=(W + 0.5, 2.0, H - 0.001),
                   point2=(W + 1.5, 1.0, H - 0.001), color=(1.0, 1.0, 1.0), power=50.0 / np.pi, samples=1)]
    camera = dict(from_=(0.4, 0.6, 1.7), to=(4.0, 1.8, 0.9), up=(0.4, 0.6, 2.7),
                  focal=float(1.0 / (2.0 * np.tan(np.radians(35.0)))), resx=int(res[0]), resy=int(res[1]))
    settings = dict(path_samples=1, bounces=bounces, rr_min_bounces=3, aa_samples=spp, aa_pixelwidth=1.5,
                    background=None, seed=seed, raydepth=0)
    return dict(verts=verts, tris=tris, mat_id=mat_id, materials=materials, lights=lights, camera=camera,
                settings=settings)


def _icosahedron():
    """The 12 vertices (0, +-1, +-phi) and their cyclic permutations, scaled to
    the unit sphere, and the 20 faces: the vertex triples at mutual distance 2
    (the edge length), each ordered so its normal points outward."""
    phi = (1.0 + 5.0 ** 0.5) / 2.0
    raw = []
    for a in (-1.0, 1.0):
        for b in (-phi, phi):
            raw += [(0.0, a, b), (a, b, 0.0), (b, 0.0, a)]
    p = np.array(raw)
    d = np.linalg.norm(p[:, None] - p[None], axis=-1)
    edge = np.abs(d - 2.0) < 1e-9
    faces = []
    for i in range(12):
        for j in range(i + 1, 12):
            for k in range(j + 1, 12):
                if edge[i, j] and edge[j, k] and edge[i, k]:
                    n = np.cross(p[j] - p[i], p[k] - p[i])
                    faces.append((i, j, k) if n @ (p[i] + p[j] + p[k]) > 0 else (i, k, j))
    assert len(faces) == 20
    return [q / np.linalg.norm(q) for q in p], faces


def icosphere_tris(center, radius: float, subdiv: int = 2, base: str = "octahedron") -> tuple[np.ndarray, np.ndarray]:
    """Sphere mesh: a base solid with each face split into 4 `subdiv` times and
    its vertices pushed to the sphere, outward face normals.  base
    "octahedron": 8 * 4^subdiv triangles; "icosahedron": 20 * 4^subdiv (1280
    at subdiv 3)."""
    v, f = _unit_sphere(subdiv, base)
    verts = (v * radius + np.asarray(center, np.float64)).astype(F32)
    return verts, f.copy()


@functools.lru_cache(maxsize=None)
def _unit_sphere(subdiv: int, base: str):
    if base == "icosahedron":
        v, f = _icosahedron()
    else:
        v = [np.array(p, np.float64) for p in ((1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1))]
        f = [(0, 2, 4), (2, 1, 4), (1, 3, 4), (3, 0, 4), (2, 0, 5), (1, 2, 5), (3, 1, 5), (0, 3, 5)]
    for _ in range(subdiv):
        mid, nf = {}, []

        def m(a, b):
            k = (min(a, b), max(a, b))
            if k not in mid:
                p = v[a] + v[b]
                v.append(p / np.linalg.norm(p))
                mid[k] = len(v) - 1
            return mid[k]
        for a, b, c in f:
            ab, bc, ca = m(a, b), m(b, c), m(c, a)
            nf += [(a, ab, ca), (ab, b, bc), (ca, bc, c), (ab, bc, ca)]
        f = nf
    v, f = np.array(v), np.array(f, np.uint32)
    v.flags.writeable = False
    f.flags.writeable = False
    return v, f


def object_light_scene(n_random: int = 2000, seed: int = 0, res: int = 64, spp: int = 4,
                       visibility: str = "no_shadows", double_sided: bool = False, subdiv: int = 2):
    """Cornell box lit by two object lights (include/light/light_object_light.h):
    a light_mat sphere mesh (8 * 4^subdiv triangles) hanging in the box, and a
    two-triangle ceiling panel facing down; no area light.  Both light meshes
    carry the object `visibility` ("normal", "no_shadows", "shadow_only",
    "invisible"); the rest of the scene is "normal".  Random triangles are
    shinydiffuse and glossy, and one point light keeps the shadows of the
    panel's back side lit."""
    from .pt import VISIBILITY
    rng = np.random.default_rng(seed)
    wv, wt, wall = cornell_walls()
    tv = random_triangles(rng, n_random, (0.15, 0.15, 0.05), (0.85, 0.85, 0.6))
    sphere = icosphere_tris((0.35, 0.6, 0.75), 0.08, subdiv)
    panel = quad_tris((0.55, 0.3, 0.97), (0.0, 0.2, 0.0), (0.2, 0.0, 0.0))  # normal -z
    verts, tris = merge([(wv, wt), soup(tv), sphere, panel])
    n_wall, n_sph = len(wt), len(sphere[1])
    materials = [dict(color=(0.7, 0.7, 0.7)), dict(color=(0.7, 0.1, 0.1)), dict(color=(0.1, 0.7, 0.1)),
                 dict(type="light_mat", color=(1.0, 0.85, 0.6), power=4.0),
                 dict(type="light_mat", color=(0.7, 0.8, 1.0), power=2.0, double_sided=double_sided)]
    wall_mat = np.array([0, 0, 0, 1, 2])
    for k in range(n_random):
        a = float(rng.uniform(0.2, 0.8))
        if rng.uniform() < 0.6:
            materials.append(dict(color=(a, a, a)))
        else:
            r = float(rng.uniform(0.1, 0.5))
            materials.append(dict(type="glossy", color=(a, 0.9 * a, 0.8 * a), exponent=2.0 / (r * r) - 2.0,
                                  diffuse_reflect=float(rng.uniform(0.0, 0.6)), diffuse_color=(0.8 * a, a, a),
                                  glossy_reflect=float(rng.uniform(0.2, 1.0))))
    mat_id = np.concatenate([wall_mat[wall], 5 + np.arange(n_random), np.full(n_sph, 3),
                             np.full(2, 4)]).astype(np.uint32)
    first_sph = n_wall + n_random
    vis = np.full(len(tris), 3, np.uint8)
    vis[first_sph:] = VISIBILITY[visibility]
    lights = [dict(type="objectlight", first_tri=first_sph, n_tris=n_sph, color=(1.0, 0.85, 0.6), power=4.0,
                   samples=4),
              dict(type="objectlight", first_tri=first_sph + n_sph, n_tris=2, color=(0.7, 0.8, 1.0), power=2.0,
                   samples=2, double_sided=double_sided),
              dict(type="point", **{"from": (0.8, 0.2, 0.3)}, color=(1.0, 1.0, 1.0), power=0.02)]
    camera = dict(from_=(0.5, -1.6, 0.5), to=(0.5, 0.5, 0.5), up=(0.5, -1.6, 1.5),
                  focal=float(1.0 / (2.0 * np.tan(np.radians(20.0)))), resx=res, resy=res)
    settings = dict(path_samples=1, bounces=4, rr_min_bounces=2, aa_samples=spp, aa_pixelwidth=1.5,
                    background=None, seed=seed, raydepth=2)
    return dict(verts=verts, tris=tris, mat_id=mat_id, vis=vis, materials=materials, lights=lights,
                camera=camera, settings=settings)


def portal_scene(n_random: int = 2000, seed: int = 0, res: int = 64, spp: int = 4, visibility: str = "invisible",
                 cells: int = 6, power: float = 2.0, ibl_clamp_sampling: float = 0.0,
                 background=(0.9, 0.95, 1.0), as_object_light: bool = False):
    """Cornell box lit by a background portal light
    (include/light/light_background_portal.h): a window x, z in [0.25, 0.75]
    cut into the back wall (y = 1), covered by the portal's mesh of `cells` x
    `cells` quads (2 * cells^2 triangles, face normal -y into the box)
    carrying the object `visibility`, with the constant `background` behind it
    (and seen through the open side), the portal's `power` and
    "ibl_clamp_sampling".  Nothing lies behind the window, so the BSDF half of
    the light's MIS, whose shadow ray has no end, reaches the background.  Random triangles
    are shinydiffuse and glossy, the floor a sharp glossy (so the BSDF half
    of the light's MIS matters); one point light.  With `as_object_light` the
    window is instead a one-sided object light of colour `background` and
    power `power` / pi: the same light up to float rounding of its colour."""
    from .pt import VISIBILITY
    rng = np.random.default_rng(seed)
    wv, wt, wall = cornell_walls()
    keep = wall != 2
    back = grid_quad_tris((0.0, 1.0, 0.0), (1.0, 0.0, 0.0), (0.0, 0.0, 1.0), 4, 4,
                          skip=lambda a, b: a in (1, 2) and b in (1, 2))
    wall = np.concatenate([wall[keep], np.full(len(back[1]), 2)])
    tv = random_triangles(rng, n_random, (0.15, 0.15, 0.05), (0.85, 0.85, 0.6))
    portal = grid_quad_tris((0.25, 1.0, 0.25), (0.5, 0.0, 0.0), (0.0, 0.0, 0.5), cells, cells)
    verts, tris = merge([(wv, wt[keep]), back, soup(tv), portal])
    n_wall, n_portal = len(wall), len(portal[1])
    materials = [dict(color=(0.7, 0.7, 0.7)), dict(color=(0.7, 0.1, 0.1)), dict(color=(0.1, 0.7, 0.1)),
                 dict(color=(0.2, 0.2, 0.25)),
                 dict(type="glossy", color=(0.8, 0.8, 0.8), exponent=400.0, diffuse_reflect=0.2,
                      diffuse_color=(0.6, 0.6, 0.6), glossy_reflect=0.8)]
    wall_mat = np.array([4, 0, 0, 1, 2])
    for k in range(n_random):
        a = float(rng.uniform(0.2, 0.8))
        if rng.uniform() < 0.6:
            materials.append(dict(color=(a, a, a)))
        else:
            r = float(rng.uniform(0.1, 0.5))
            materials.append(dict(type="glossy", color=(a, 0.9 * a, 0.8 * a), exponent=2.0 / (r * r) - 2.0,
                                  diffuse_reflect=float(rng.uniform(0.0, 0.6)), diffuse_color=(0.8 * a, a, a),
                                  glossy_reflect=float(rng.uniform(0.2, 1.0))))
    mat_id = np.concatenate([wall_mat[wall], 5 + np.arange(n_random), np.full(n_portal, 3)]).astype(np.uint32)
    first = n_wall + n_random
    vis = np.full(len(tris), 3, np.uint8)
    vis[first:] = VISIBILITY[visibility]
    if as_object_light:
        light = dict(type="objectlight", first_tri=first, n_tris=n_portal, color=tuple(background),
                     power=float(power / np.pi), samples=16)
    else:
        light = dict(type="bgPortalLight", first_tri=first, n_tris=n_portal, power=power, samples=16,
                     ibl_clamp_sampling=ibl_clamp_sampling)
    lights = [light, dict(type="point", **{"from": (0.8, 0.6, 0.9)}, color=(1.0, 1.0, 1.0), power=0.05)]
    camera = dict(from_=(0.5, -1.6, 0.5), to=(0.5, 0.5, 0.5), up=(0.5, -1.6, 1.5),
                  focal=float(1.0 / (2.0 * np.tan(np.radians(20.0)))), resx=res, resy=res)
    settings = dict(path_samples=1, bounces=4, rr_min_bounces=2, aa_samples=spp, aa_pixelwidth=1.5,
                    background=tuple(background), seed=seed, raydepth=2)
    return dict(verts=verts, tris=tris, mat_id=mat_id, vis=vis, materials=materials, lights=lights,
                camera=camera, settings=settings)


def spot_room_scene(res: int = 64, spp: int = 16, seed: int = 0, h: float = 0.1, size: float = 0.1,
                    bounces: int = 2):
    """Guiding workload where the incoming light is concentrated: the five
    Cornell walls (albedo 0.7, open at y = 0) lit by one small area light
    (`size` square, power 200) facing down at height `h` above the floor
    centre, so it lights a small spot of the floor and the rest of the box
    sees mostly that spot.  No roulette; path depth `bounces`."""
    wv, wt, _ = cornell_walls()
    c = (0.5 - size / 2, 0.5 - size / 2, h)
    # point1 / point2 chosen so that the light faces -z (towards the floor)
    lights = [dict(type="area", corner=c, point1=(c[0], c[1] + size, h), point2=(c[0] + size, c[1], h),
                   color=(1.0, 1.0, 1.0), power=200.0, samples=1)]
    camera = dict(from_=(0.5, -1.6, 0.5), to=(0.5, 0.5, 0.5), up=(0.5, -1.6, 1.5),
                  focal=float(1.0 / (2.0 * np.tan(np.radians(20.0)))), resx=res, resy=res)
    settings = dict(path_samples=1, bounces=bounces, rr_min_bounces=10, aa_samples=spp, aa_pixelwidth=1.5,
                    background=None, seed=seed, raydepth=0)
    return dict(verts=wv, tris=wt, mat_id=np.zeros(len(wt), np.uint32), materials=[dict(color=(0.7, 0.7, 0.7))],
                lights=lights, camera=camera, settings=settings)


def terrain_spheres_scene(res=(1920, 1080), spp: int = 64, seed: int = 0, hf_res: int = 1024, n_spheres: int = 500,
                          subdiv: int = 3, n_point_lights: int = 64, depth: int = 8, area_lights=None):
    """Mixed-material path-tracing workload: the ray-cast heightfield
    (`hf_res`^2 vertices) with `n_spheres` icospheres above it (20 * 4^subdiv
    triangles each: 1280 at the default subdiv 3), sphere materials drawn 40 % shinydiffuse / 30 % glossy
    (roughness r ~ U[0.05, 0.5], Blinn exponent 2 / r^2 - 2) / 15 % mirror /
    15 % glass (IOR 1.5), `n_point_lights` point lights of power log-U[1, 100],
    one area light (or the `area_lights` given, as (corner, point1, point2,
    power) tuples) and a constant background of 0.1; path depth `depth`.  The
    defaults are the full frame; tests use a small one."""
    rng = np.random.default_rng(seed)
    hf = grid_mesh(fractal_heights(hf_res, seed=seed))
    parts, mats = [hf], [0]
    materials = [dict(color=(0.5, 0.45, 0.4))]
    for k in range(n_spheres):
        c = (float(rng.uniform(0.1, 0.9)), float(rng.uniform(0.1, 0.9)), float(rng.uniform(0.3, 0.5)))
        parts.append(icosphere_tris(c, float(rng.uniform(0.02, 0.05)), subdiv, base="icosahedron"))
        mats.append(len(materials))
        u, a = rng.uniform(), float(rng.uniform(0.2, 0.8))
        if u < 0.4:
            materials.append(dict(color=(a, a, a)))
        elif u < 0.7:
            r = float(rng.uniform(0.05, 0.5))
            materials.append(dict(type="glossy", color=(a, 0.9 * a, 0.8 * a), exponent=2.0 / (r * r) - 2.0,
                                  diffuse_reflect=float(rng.uniform(0.0, 0.6)), diffuse_color=(0.8 * a, a, a),
                                  glossy_reflect=float(rng.uniform(0.2, 1.0))))
        elif u < 0.85:
            materials.append(dict(type="mirror", color=(a, a, 0.9 * a), reflect=float(rng.uniform(0.5, 1.0))))
        else:
            materials.append(dict(type="glass", IOR=1.5))
    verts, tris = merge(parts)
    mat_id = np.concatenate([np.full(len(t), m, np.uint32) for (_, t), m in zip(parts, mats)])
    if area_lights is None:
        area_lights = [((0.4, 0.4, 1.2), (0.4, 0.6, 1.2), (0.6, 0.4, 1.2), 20.0 / np.pi)]
    lights = [dict(type="area", corner=c, point1=p1, point2=p2, color=(1.0, 1.0, 1.0), power=float(pw), samples=4)
              for c, p1, p2, pw in area_lights]
    for k in range(n_point_lights):
        p = (float(rng.uniform(0.0, 1.0)), float(rng.uniform(0.0, 1.0)), float(rng.uniform(0.6, 1.0)))
        lights.append(dict(type="point", **{"from": p}, color=(1.0, 0.95, 0.9),
                           power=float(np.exp(rng.uniform(0.0, np.log(100.0))))))
    camera = dict(from_=(0.5, -0.6, 0.9), to=(0.5, 0.55, 0.15), up=(0.5, -0.6, 1.9),
                  focal=float(1.0 / (2.0 * np.tan(np.radians(30.0)))), resx=int(res[0]), resy=int(res[1]))
    settings = dict(path_samples=1, bounces=depth, rr_min_bounces=3, aa_samples=spp, aa_pixelwidth=1.5,
                    background=(0.1, 0.1, 0.1), seed=seed, raydepth=depth)
    return dict(verts=verts, tris=tris, mat_id=mat_id, materials=materials, lights=lights, camera=camera,
                settings=settings)


def large_terrain_scene(res=(3840, 2160), spp: int = 256, seed: int = 0, hf_res: int = 4096, n_spheres: int = 2000,
                        n_point_lights: int = 256, depth: int = 12):
    """Large-scene workload: a `hf_res`^2 fractal heightfield (33.5M triangles
    at 4096) with `n_spheres` 1280-triangle icospheres (2.56M at 2000), the
    mixed material draw of terrain_spheres_scene, `n_point_lights` point
    lights, and two downward-facing 0.2 x 0.2 area lights hung low over the
    terrain (z = 0.12 and 0.16, below its highest ridges at 0.2), so that the
    terrain hides them from part of the scene; path depth `depth`."""
    area = [((0.2, 0.25, 0.12), (0.2, 0.45, 0.12), (0.4, 0.25, 0.12), 8.0 / np.pi),
            ((0.6, 0.6, 0.16), (0.6, 0.8, 0.16), (0.8, 0.6, 0.16), 8.0 / np.pi)]
    return terrain_spheres_scene(res=res, spp=spp, seed=seed, hf_res=hf_res, n_spheres=n_spheres, subdiv=3,
                                 n_point_lights=n_point_lights, depth=depth, area_lights=area)


def smooth_spheres_scene(res: int = 64, spp: int = 4, seed: int = 0, subdiv: int = 2, flat_every: int = 3):
    """Cornell box with icospheres of every material whose shading normals are
    the sphere's radial vertex normals (as yafaray_addNormal gives them), so a
    hit's shading normal n is the barycentric blend of its corners' normals
    while the geometric normal ng stays the facet's.  corner_normals is the
    (n_tris, 12) array of ymx_pt_set_normals: the walls are flat (w = -1);
    on every `flat_every`-th sphere one corner of each triangle has no normal
    (w = 0: the face normal stands in)."""
    rng = np.random.default_rng(seed)
    wv, wt, wall = cornell_walls()
    meshes, normals, mats = [(wv, wt)], [np.full((len(wt), 12), -1.0, F32)], []
    materials = [dict(color=(0.7, 0.7, 0.7)), dict(color=(0.7, 0.1, 0.1)), dict(color=(0.1, 0.7, 0.1))]
    kinds = [dict(color=(0.8, 0.8, 0.8)),
             dict(type="glossy", color=(0.8, 0.7, 0.6), exponent=60.0, diffuse_reflect=0.3,
                  diffuse_color=(0.6, 0.7, 0.8), glossy_reflect=0.8),
             dict(type="glossy", color=(0.7, 0.8, 0.6), exponent=20.0, diffuse_reflect=0.2, as_diffuse=False),
             dict(type="mirror", color=(0.9, 0.9, 0.85), reflect=0.9),
             dict(type="glass", IOR=1.5, filter_color=(0.95, 0.9, 1.0))]
    centers = [(0.3, 0.35, 0.18), (0.7, 0.35, 0.2), (0.5, 0.65, 0.22), (0.28, 0.7, 0.5), (0.72, 0.68, 0.55)]
    mat_id = [np.array([0, 0, 0, 1, 2])[wall]]
    for k, (c, kind) in enumerate(zip(centers, kinds)):
        r = 0.15 + 0.03 * rng.uniform()
        sv, st = icosphere_tris(c, r, subdiv)
        vn = (sv.astype(np.float64) - np.asarray(c)) / r
        vn = (vn / np.linalg.norm(vn, axis=1, keepdims=True)).astype(F32)
        cn = np.zeros((len(st), 3, 4), F32)
        cn[:, :, :3] = vn[st]
        cn[:, :, 3] = 1.0
        if flat_every and k % flat_every == 0:
            cn[:, 1, :] = 0.0
        meshes.append((sv, st))
        normals.append(cn.reshape(-1, 12))
        materials.append(kind)
        mat_id.append(np.full(len(st), len(materials) - 1))
    verts, tris = merge(meshes)
    lights = [dict(type="area", corner=(0.4, 0.4, 0.999), point1=(0.4, 0.6, 0.999), point2=(0.6, 0.4, 0.999),
                   color=(1.0, 1.0, 1.0), power=10.0 / np.pi, samples=4),
              dict(type="point", **{"from": (0.5, 0.2, 0.8)}, color=(1.0, 0.95, 0.9), power=0.05)]
    camera = dict(from_=(0.5, -1.6, 0.5), to=(0.5, 0.5, 0.5), up=(0.5, -1.6, 1.5),
                  focal=float(1.0 / (2.0 * np.tan(np.radians(20.0)))), resx=res, resy=res)
    settings = dict(path_samples=1, bounces=5, rr_min_bounces=2, aa_samples=spp, aa_pixelwidth=1.5,
                    background=(0.1, 0.1, 0.1), seed=0, raydepth=5)
    return dict(verts=verts, tris=tris, mat_id=np.concatenate(mat_id).astype(np.uint32), materials=materials,
                lights=lights, camera=camera, settings=settings, corner_normals=np.concatenate(normals))
