"""Synthetic workloads (voxel sets, meshes, cameras) for the benchmark
configurations.  All randomness comes from SplitMix64 so every input is a
pure function of its seed; the generators run on the host with numpy.

Voxel sets are returned as sorted unique 3-bit interleaved Morton codes with
x the most significant bit of each triple, so ``code & 7`` is the child slot
4x+2y+z of the finest octree level (octree.hpp:33-42).
"""
from __future__ import annotations

import math

import numpy as np

MASK64 = (1 << 64) - 1


# ---------------------------------------------------------------- RNG
def splitmix64(state: int):
    """Yield an endless SplitMix64 stream (Steele, Lea, Flood 2014)."""
    while True:
        state = (state + 0x9E3779B97F4A7C15) & MASK64
        z = state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        yield z ^ (z >> 31)


def uniforms(seed: int, n: int) -> np.ndarray:
    g = splitmix64(seed)
    return np.array([(next(g) >> 11) * (1.0 / (1 << 53)) for _ in range(n)], dtype=np.float64)


def _mix_u64(x: np.ndarray) -> np.ndarray:
    x = x.astype(np.uint64)
    x ^= x >> np.uint64(30)
    x *= np.uint64(0xBF58476D1CE4E5B9)
    x ^= x >> np.uint64(27)
    x *= np.uint64(0x94D049BB133111EB)
    x ^= x >> np.uint64(31)
    return x


# ---------------------------------------------------------------- Morton codes
def _spread3(v: np.ndarray) -> np.ndarray:
    x = v.astype(np.uint64) & np.uint64(0x1FFFFF)
    x = (x | (x << np.uint64(32))) & np.uint64(0x1F00000000FFFF)
    x = (x | (x << np.uint64(16))) & np.uint64(0x1F0000FF0000FF)
    x = (x | (x << np.uint64(8))) & np.uint64(0x100F00F00F00F00F)
    x = (x | (x << np.uint64(4))) & np.uint64(0x10C30C30C30C30C3)
    x = (x | (x << np.uint64(2))) & np.uint64(0x1249249249249249)
    return x


def _compact3(x: np.ndarray) -> np.ndarray:
    x = x.astype(np.uint64) & np.uint64(0x1249249249249249)
    x = (x ^ (x >> np.uint64(2))) & np.uint64(0x10C30C30C30C30C3)
    x = (x ^ (x >> np.uint64(4))) & np.uint64(0x100F00F00F00F00F)
    x = (x ^ (x >> np.uint64(8))) & np.uint64(0x1F0000FF0000FF)
    x = (x ^ (x >> np.uint64(16))) & np.uint64(0x1F00000000FFFF)
    x = (x ^ (x >> np.uint64(32))) & np.uint64(0x1FFFFF)
    return x


def morton_encode(x, y, z) -> np.ndarray:
    return (_spread3(x) << np.uint64(2)) | (_spread3(y) << np.uint64(1)) | _spread3(z)


def morton_decode(code: np.ndarray):
    c = code.astype(np.uint64)
    return _compact3(c >> np.uint64(2)), _compact3(c >> np.uint64(1)), _compact3(c)


def sorted_codes(x, y, z) -> np.ndarray:
    return np.unique(morton_encode(x, y, z))


# ---------------------------------------------------------------- value noise / fBm
def _lattice(ix: np.ndarray, iy: np.ndarray, iz: np.ndarray | None, seed: int) -> np.ndarray:
    h = ix.astype(np.int64).astype(np.uint64) * np.uint64(0x8DA6B343)
    h ^= iy.astype(np.int64).astype(np.uint64) * np.uint64(0xD8163841)
    if iz is not None:
        h ^= iz.astype(np.int64).astype(np.uint64) * np.uint64(0xCB1AB31F)
    h ^= np.uint64((seed * 0x9E3779B97F4A7C15) & MASK64)
    h = _mix_u64(h)
    return (h >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))


def _smooth(t):
    return t * t * (3.0 - 2.0 * t)


def value_noise2(x: np.ndarray, y: np.ndarray, seed: int) -> np.ndarray:
    ix, iy = np.floor(x), np.floor(y)
    fx, fy = _smooth(x - ix), _smooth(y - iy)
    ix, iy = ix.astype(np.int64), iy.astype(np.int64)
    v00 = _lattice(ix, iy, None, seed)
    v10 = _lattice(ix + 1, iy, None, seed)
    v01 = _lattice(ix, iy + 1, None, seed)
    v11 = _lattice(ix + 1, iy + 1, None, seed)
    a = v00 + (v10 - v00) * fx
    b = v01 + (v11 - v01) * fx
    return a + (b - a) * fy


def value_noise3(x, y, z, seed: int) -> np.ndarray:
    ix, iy, iz = np.floor(x), np.floor(y), np.floor(z)
    fx, fy, fz = _smooth(x - ix), _smooth(y - iy), _smooth(z - iz)
    ix, iy, iz = ix.astype(np.int64), iy.astype(np.int64), iz.astype(np.int64)
    out = 0.0
    for dz in (0, 1):
        wz = fz if dz else 1.0 - fz
        for dy in (0, 1):
            wy = fy if dy else 1.0 - fy
            for dx in (0, 1):
                wx = fx if dx else 1.0 - fx
                out = out + wx * wy * wz * _lattice(ix + dx, iy + dy, iz + dz, seed)
    return out


def fbm2(x, y, octaves: int, lacunarity: float, gain: float, seed: int) -> np.ndarray:
    s, amp, freq, norm = 0.0, 1.0, 1.0, 0.0
    for o in range(octaves):
        s = s + amp * value_noise2(x * freq, y * freq, seed * 1000 + o)
        norm += amp
        amp *= gain
        freq *= lacunarity
    return s / norm


def fbm3(x, y, z, octaves: int, lacunarity: float, gain: float, seed: int) -> np.ndarray:
    s, amp, freq, norm = 0.0, 1.0, 1.0, 0.0
    for o in range(octaves):
        s = s + amp * value_noise3(x * freq, y * freq, z * freq, seed * 1000 + o)
        norm += amp
        amp *= gain
        freq *= lacunarity
    return s / norm


# ---------------------------------------------------------------- terrain (configs 2, 5)
def terrain_heights(levels: int, seed: int = 1, octaves: int = 6, lacunarity: float = 2.0, gain: float = 0.5,
                    hmin: float = 0.2, hmax: float = 0.6, base_cells: float = 8.0) -> np.ndarray:
    """2D fBm heightfield of side 2^levels, heights in voxels (int64, [x, z] indexing)."""
    n = 1 << levels
    c = (np.arange(n, dtype=np.float64) + 0.5) * (base_cells / n)
    X, Z = np.meshgrid(c, c, indexing="ij")
    f = fbm2(X, Z, octaves, lacunarity, gain, seed)
    f = (f - f.min()) / max(f.max() - f.min(), 1e-12)     # stretch to the full range
    return np.floor((hmin + (hmax - hmin) * f) * n).astype(np.int64).clip(0, n - 1)


def terrain_voxels(levels: int, seed: int = 1, **kw):
    """Surface voxels: each column filled from the lowest of its 4-neighbour
    heights (and its own) up to its height."""
    h = terrain_heights(levels, seed, **kw)
    p = np.pad(h, 1, mode="edge")
    lo = np.minimum.reduce([h, p[:-2, 1:-1], p[2:, 1:-1], p[1:-1, :-2], p[1:-1, 2:]])
    cnt = (h - lo + 1).ravel()
    n = 1 << levels
    xs = np.repeat(np.arange(n, dtype=np.int64), n)
    zs = np.tile(np.arange(n, dtype=np.int64), n)
    X = np.repeat(xs, cnt)
    Z = np.repeat(zs, cnt)
    start = np.repeat(np.cumsum(cnt) - cnt, cnt)
    Y = np.repeat(lo.ravel(), cnt) + (np.arange(cnt.sum(), dtype=np.int64) - start)
    return X, Y, Z


def terrain_codes(levels: int, seed: int = 1, **kw) -> np.ndarray:
    x, y, z = terrain_voxels(levels, seed, **kw)
    return sorted_codes(x, y, z)


# ---------------------------------------------------------------- caves (configs 4, 5)
def cave_codes(levels: int, seed: int = 3, octaves: int = 5, threshold: float = 0.1, base_cells: float = 4.0,
               chunk: int = 64) -> np.ndarray:
    """Surface voxels (solid with an empty 6-neighbour; outside the grid is
    empty) of the solid {fBm3 - 0.5 > threshold}, evaluated in x-slabs."""
    n = 1 << levels
    c = (np.arange(n, dtype=np.float64) + 0.5) * (base_cells / n)
    out = []
    for x0 in range(0, n, chunk):
        x1 = min(n, x0 + chunk)
        lo, hi = max(0, x0 - 1), min(n, x1 + 1)
        X, Y, Z = np.meshgrid(c[lo:hi], c, c, indexing="ij")
        s = fbm3(X, Y, Z, octaves, 2.0, 0.5, seed) - 0.5 > threshold
        P = np.zeros((x1 - x0 + 2, n + 2, n + 2), dtype=bool)
        a = lo - (x0 - 1)
        P[a:a + (hi - lo), 1:-1, 1:-1] = s
        core = P[1:-1, 1:-1, 1:-1]
        full = (P[:-2, 1:-1, 1:-1] & P[2:, 1:-1, 1:-1] & P[1:-1, :-2, 1:-1] & P[1:-1, 2:, 1:-1]
                & P[1:-1, 1:-1, :-2] & P[1:-1, 1:-1, 2:])
        xi, yi, zi = np.nonzero(core & ~full)
        out.append(morton_encode(xi + x0, yi, zi))
    return np.unique(np.concatenate(out)) if out else np.zeros(0, np.uint64)


# ---------------------------------------------------------------- materials (config 5)
def material_attrs(codes: np.ndarray, levels: int, seed: int = 4, n_values: int = 16,
                   base_cells: float = 8.0) -> np.ndarray:
    """Per-voxel u8 material in the codes' order: 3D value noise at the voxel
    centre (seed 4), quantized to n_values spatially coherent levels."""
    x, y, z = morton_decode(codes)
    s = base_cells / (1 << levels)
    v = value_noise3((x.astype(np.float64) + 0.5) * s, (y.astype(np.float64) + 0.5) * s,
                     (z.astype(np.float64) + 0.5) * s, seed)
    return np.minimum(np.floor(v * n_values), n_values - 1).astype(np.uint8)


# ---------------------------------------------------------------- icosphere meshes (config 1)
def icosphere(subdiv: int):
    t = (1.0 + math.sqrt(5.0)) / 2.0
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
         (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11),
         (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    verts = [np.array(p, dtype=np.float64) / np.linalg.norm(p) for p in v]
    for _ in range(subdiv):
        cache = {}

        def mid(a, b):
            k = (min(a, b), max(a, b))
            if k not in cache:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                cache[k] = len(verts) - 1
            return cache[k]

        nf = []
        for a, b, c in f:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        f = nf
    return np.array(verts), np.array(f, dtype=np.int64)


def spheres_mesh(n_spheres: int = 64, subdiv: int = 3, seed: int = 0, rmin: float = 0.02, rmax: float = 0.15):
    """Union (as a triangle soup) of icospheres; returns (vertices f32 [V,3], faces i64 [F,3])."""
    u = uniforms(seed, 4 * n_spheres).reshape(n_spheres, 4)
    bv, bf = icosphere(subdiv)
    V, F = [], []
    for i in range(n_spheres):
        c = 0.1 + 0.8 * u[i, :3]
        r = rmin + (rmax - rmin) * u[i, 3]
        F.append(bf + i * len(bv))
        V.append(bv * r + c)
    return np.concatenate(V).astype(np.float32), np.concatenate(F)


def write_obj(path: str, verts: np.ndarray, faces: np.ndarray) -> None:
    with open(path, "w") as f:
        for p in verts:
            f.write("v %.9g %.9g %.9g\n" % (float(p[0]), float(p[1]), float(p[2])))
        for a, b, c in faces + 1:
            f.write("f %d %d %d\n" % (a, b, c))


def write_ply(path: str, verts: np.ndarray, faces: np.ndarray) -> None:
    with open(path, "w") as f:
        f.write("ply\nformat ascii 1.0\nelement vertex %d\nproperty float x\nproperty float y\n"
                "property float z\nelement face %d\nproperty list uchar int vertex_indices\nend_header\n"
                % (len(verts), len(faces)))
        for p in verts:
            f.write("%.9g %.9g %.9g\n" % (float(p[0]), float(p[1]), float(p[2])))
        for a, b, c in faces:
            f.write("3 %d %d %d\n" % (a, b, c))


def mesh_bbox(verts: np.ndarray):
    return verts.min(axis=0).astype(np.float32), verts.max(axis=0).astype(np.float32)


# ---------------------------------------------------------------- Menger sponge (config 3)
def menger_codes(levels: int, order: int) -> np.ndarray:
    """Menger sponge of the given order filling the 2^levels grid (surface and interior)."""
    n = 1 << levels
    c = np.arange(n)
    X, Y, Z = np.meshgrid(c, c, c, indexing="ij")
    solid = np.ones((n, n, n), dtype=bool)
    for k in range(order):
        s = 3 ** (k + 1)
        mx = ((X * s) // n) % 3 == 1
        my = ((Y * s) // n) % 3 == 1
        mz = ((Z * s) // n) % 3 == 1
        solid &= ~((mx & my) | (my & mz) | (mx & mz))
    x, y, z = np.nonzero(solid)
    return sorted_codes(x, y, z)


def _surface(solid: np.ndarray) -> np.ndarray:
    """Voxels of a boolean grid with an empty 6-neighbour (outside is empty)."""
    P = np.pad(solid, 1)
    full = (P[:-2, 1:-1, 1:-1] & P[2:, 1:-1, 1:-1] & P[1:-1, :-2, 1:-1] & P[1:-1, 2:, 1:-1]
            & P[1:-1, 1:-1, :-2] & P[1:-1, 1:-1, 2:])
    return solid & ~full


def menger_surface_codes(levels: int, order: int) -> np.ndarray:
    """Surface voxels of the Menger sponge of the given order filling the 2^levels grid."""
    n = 1 << levels
    c = np.arange(n)
    X, Y, Z = np.meshgrid(c, c, c, indexing="ij")
    solid = np.ones((n, n, n), dtype=bool)
    for k in range(order):
        s = 3 ** (k + 1)
        mx = ((X * s) // n) % 3 == 1
        my = ((Y * s) // n) % 3 == 1
        mz = ((Z * s) // n) % 3 == 1
        solid &= ~((mx & my) | (my & mz) | (mx & mz))
    return sorted_codes(*np.nonzero(_surface(solid)))


def sphere_balls(n_spheres: int, seed: int, rmin: float = 0.02, rmax: float = 0.15) -> np.ndarray:
    """(cx, cy, cz, r) of the icosphere scene's spheres (the same draws as spheres_mesh)."""
    u = uniforms(seed, 4 * n_spheres).reshape(n_spheres, 4)
    return np.column_stack([0.1 + 0.8 * u[:, :3], rmin + (rmax - rmin) * u[:, 3]]).astype(np.float64)


def balls_codes(levels: int, balls: np.ndarray) -> np.ndarray:
    """Surface voxels of a union of balls, voxel centres p = (i + 0.5) / 2^levels."""
    n = 1 << levels
    c = (np.arange(n, dtype=np.float64) + 0.5) / n
    X, Y, Z = np.meshgrid(c, c, c, indexing="ij")
    solid = np.zeros((n, n, n), dtype=bool)
    for cx, cy, cz, r in balls:
        dx, dy, dz = X - cx, Y - cy, Z - cz
        solid |= dx * dx + dy * dy + dz * dz <= r * r
    return sorted_codes(*np.nonzero(_surface(solid)))


def batch_specs(n_scenes: int = 256, seed: int = 0, depth=(6, 10)):
    """The multi-scene batch: depth uniform in depth[0]..depth[1]; the kind is
    drawn per scene: spheres (8-128 icospheres' balls), fBm terrain, or a
    Menger sponge of order 3-5 (surface voxels)."""
    u = uniforms(seed * 7919 + 17, 4 * n_scenes).reshape(n_scenes, 4)
    specs = []
    for i in range(n_scenes):
        levels = depth[0] + int(u[i, 0] * (depth[1] - depth[0] + 1))
        kind = ("spheres", "terrain", "menger")[int(u[i, 1] * 3)]
        sd = seed * 100003 + i
        if kind == "spheres":
            specs.append(dict(kind=kind, levels=levels, seed=sd, n_spheres=8 + int(u[i, 2] * 121)))
        elif kind == "terrain":
            specs.append(dict(kind=kind, levels=levels, seed=sd))
        else:
            specs.append(dict(kind=kind, levels=levels, order=3 + int(u[i, 2] * 3)))
    return specs


def spec_codes(spec) -> np.ndarray:
    """Host (NumPy) voxels of one batch scene (small depths only)."""
    if spec["kind"] == "spheres":
        return balls_codes(spec["levels"], sphere_balls(spec["n_spheres"], spec["seed"]))
    if spec["kind"] == "terrain":
        return terrain_codes(spec["levels"], seed=spec["seed"])
    return menger_surface_codes(spec["levels"], spec["order"])


# ---------------------------------------------------------------- cameras
def _look_at(eye, center, up):
    f = center - eye
    f = f / np.linalg.norm(f)
    s = np.cross(f, up)
    s = s / np.linalg.norm(s)
    u = np.cross(s, f)
    m = np.eye(4)
    m[0, :3], m[1, :3], m[2, :3] = s, u, -f
    m[0, 3], m[1, 3], m[2, 3] = -s @ eye, -u @ eye, f @ eye
    return m


def _perspective(fovy, aspect, near, far):
    t = 1.0 / math.tan(fovy / 2.0)
    m = np.zeros((4, 4))
    m[0, 0], m[1, 1] = t / aspect, t
    m[2, 2], m[2, 3] = (far + near) / (near - far), 2 * far * near / (near - far)
    m[3, 2] = -1.0
    return m


def projection_factor(height: int, fovy_deg: float = 60.0, pixel_tolerance: float = 1.0,
                      screen_divisor: float = 1.0) -> float:
    """OctreeDDARenderer::getProjectionFactor (octree_dda_renderer.cpp:503-507), in float32."""
    f32 = np.float32
    fovy = f32(fovy_deg) * f32(3.14159265359) / f32(180.0)
    inv = f32(1.0) / (f32(2.0) * f32(math.tan(float(f32(0.5) * fovy))))
    tol = f32(pixel_tolerance) / (f32(height) / f32(screen_divisor))
    return float(f32(inv / tol))


def orbit_camera(seed: int, width: int, height: int, center=(0.5, 0.5, 0.5), radius: float | None = None,
                 fovy_deg: float = 60.0, elev=(20.0, 50.0)):
    """Seeded orbit camera looking at `center` (GL conventions).  Returns
    (view_inv, proj_inv) as column-major float32[16] arrays."""
    u = uniforms(seed, 2)
    yaw = 2 * math.pi * u[0]
    pitch = math.radians(elev[0] + (elev[1] - elev[0]) * u[1])
    if radius is None:
        radius = 1.5 * math.sqrt(3.0) / 2.0   # 1.5 x the half diagonal of the unit scene box
    c = np.array(center, dtype=np.float64)
    eye = c + radius * np.array([math.cos(pitch) * math.cos(yaw), math.sin(pitch), math.cos(pitch) * math.sin(yaw)])
    view = _look_at(eye, c, np.array([0.0, 1.0, 0.0]))
    proj = _perspective(math.radians(fovy_deg), width / height, 0.01, 100.0)
    vi = np.linalg.inv(view).astype(np.float32)
    pi = np.linalg.inv(proj).astype(np.float32)
    # GL column-major: element (row r, col c) at index c*4 + r
    return vi.T.ravel().copy(), pi.T.ravel().copy()


def trace_params(width: int, height: int, seed: int, levels: int, light=(1.0, 1.0, 1.0), shadows=True,
                 max_iters: int = 300, radius: float | None = None, center=(0.5, 0.5, 0.5)):
    """Renderer defaults: maxIters 300, drawLevel = levels, pixel tolerance 1,
    light at the scene box max (octree_dda_renderer.cpp:47, 66, 222-223)."""
    from . import TraceParams  # noqa: circular-safe (package import)
    vi, pi = orbit_camera(seed, width, height, center=center, radius=radius)
    p = TraceParams()
    p.view_inv[:] = [float(x) for x in vi]
    p.proj_inv[:] = [float(x) for x in pi]
    p.width, p.height, p.max_iters, p.draw_level = width, height, max_iters, levels
    p.projection_factor = projection_factor(height)
    p.light_pos[:] = [float(x) for x in light]
    p.shadows = 1 if shadows else 0
    return p
