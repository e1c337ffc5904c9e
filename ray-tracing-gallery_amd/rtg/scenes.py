"""Seeded synthetic scenes for the trace step.

The reference renders a fixed, unseeded demo scene (src/scene.rs:34-159: a
ground plane, tori, a character, a portal, an alpha-masked fence and 100
random tori drawn with rand::thread_rng). The workloads here are seeded
procedural scenes with the same building blocks and the reference's
conventions:

  * a model = shared vertex streams (positions/normals vec3, uvs vec2) plus
    one u32 index buffer per material geometry (src/util_structs.rs:903-916,
    1113-1137);
  * materials are 1x1 RGBA32F fallback images for the base colour factor and
    for metallic-roughness [1, roughness, metallic, 1]
    (src/util_structs.rs:918-930, 1026-1029), or RGBA8 sRGB textures;
  * alpha-MASK geometry is non-opaque and runs any_hit_alpha_clip;
  * instances carry custom index = model id, mask 0xFF, hit group
    Textured/Mirror/Portal and CULL_DISABLE for double-sided geometry
    (src/gpu_structs.rs:27-53, src/scene.rs:95-156);
  * texture 2 is a 64x64 blue-noise mask, texture 3 the GGX LUT slot
    (src/main.rs:416-460, :595-596).

Random numbers come from a counter-based SplitMix64 stream. Scalar
transcendentals go through Python's math module (one libm call per value) so
that the generated floats do not depend on numpy's SIMD dispatch.
"""
import math

import numpy as np

from . import abi

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def _mix64(x):
    """SplitMix64 finaliser (Steele, Lea, Flood 2014) on a uint64 array."""
    x = (x + np.uint64(0x9E3779B97F4A7C15)) & _M64
    x = ((x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & _M64
    x = ((x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & _M64
    return x ^ (x >> np.uint64(31))


class Rng:
    """Counter-based stream: value k = mix64(key ^ mix64(k))."""

    def __init__(self, seed):
        self.key = int(_mix64(np.array([seed & 0xFFFFFFFFFFFFFFFF], dtype=np.uint64))[0])
        self.counter = 0

    def u64(self, n):
        c = np.arange(self.counter, self.counter + n, dtype=np.uint64)
        self.counter += n
        return _mix64(np.uint64(self.key) ^ _mix64(c))

    def uniform(self, n, lo=0.0, hi=1.0):
        u = (self.u64(n) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
        return lo + (hi - lo) * u

    def integers(self, n, lo, hi):
        return (lo + (self.u64(n) % np.uint64(hi - lo)).astype(np.int64)).astype(np.int64)

    def normal(self, n):
        """Box-Muller with libm log/cos."""
        u1 = self.uniform(n)
        u2 = self.uniform(n)
        r = np.array([math.sqrt(-2.0 * math.log(1.0 - a)) for a in u1])
        c = np.array([math.cos(2.0 * math.pi * b) for b in u2])
        return r * c


_vsin = np.frompyfunc(math.sin, 1, 1)
_vcos = np.frompyfunc(math.cos, 1, 1)
_vexp = np.frompyfunc(math.exp, 1, 1)


def dsin(x):
    return np.asarray(_vsin(np.asarray(x, dtype=np.float64)), dtype=np.float64)


def dcos(x):
    return np.asarray(_vcos(np.asarray(x, dtype=np.float64)), dtype=np.float64)


def dexp(x):
    return np.asarray(_vexp(np.asarray(x, dtype=np.float64)), dtype=np.float64)


# --------------------------------------------------------------------------
# Meshes
# --------------------------------------------------------------------------
class Mesh:
    """Shared vertex streams and per-material index buffers of one model
    (src/util_structs.rs:903-916)."""

    def __init__(self, name, positions, normals, uvs, geometries):
        self.name = name
        self.positions = np.ascontiguousarray(positions, dtype=np.float32)
        self.normals = np.ascontiguousarray(normals, dtype=np.float32)
        self.uvs = np.ascontiguousarray(uvs, dtype=np.float32)
        # list of (indices (T,3) uint32, opaque bool)
        self.geometries = [(np.ascontiguousarray(ix, dtype=np.uint32).reshape(-1, 3), bool(op))
                           for ix, op in geometries]

    @property
    def triangle_count(self):
        return sum(len(ix) for ix, _ in self.geometries)


def torus(seg_major=48, seg_minor=24, R=1.0, r=0.35, name="torus"):
    """(seg_major+1) x (seg_minor+1) vertices with a uv seam,
    2 * seg_major * seg_minor triangles (the tori of resources/tori.glb have
    48 x 12 segments; the workload asks for 48 x 24)."""
    i = np.arange(seg_major + 1, dtype=np.float64)
    j = np.arange(seg_minor + 1, dtype=np.float64)
    u = 2.0 * math.pi * (i % seg_major) / seg_major
    v = 2.0 * math.pi * (j % seg_minor) / seg_minor
    cu, su = dcos(u)[:, None], dsin(u)[:, None]
    cv, sv = dcos(v)[None, :], dsin(v)[None, :]
    px = (R + r * cv) * cu
    pz = (R + r * cv) * su
    py = np.broadcast_to(r * sv, px.shape)
    nx = cv * cu
    nz = cv * su
    ny = np.broadcast_to(sv, nx.shape)
    pos = np.stack([px, py, pz], -1).reshape(-1, 3)
    nrm = np.stack([nx, ny, nz], -1).reshape(-1, 3)
    uv = np.stack(np.meshgrid(i / seg_major, j / seg_minor, indexing="ij"), -1).reshape(-1, 2)
    a = (np.arange(seg_major)[:, None] * (seg_minor + 1) + np.arange(seg_minor)[None, :]).reshape(-1)
    b = a + (seg_minor + 1)
    tris = np.concatenate([np.stack([a, b, a + 1], -1), np.stack([a + 1, b, b + 1], -1)], 0)
    return Mesh(name, pos, nrm, uv, [(tris, True)])


def icosphere(level=4, radius=1.0, name="icosphere"):
    """Subdivided icosahedron: 20 * 4**level triangles (level 4: 5,120)."""
    t = (1.0 + math.sqrt(5.0)) / 2.0
    verts = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
             (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    verts = [tuple(c / math.sqrt(1.0 + t * t) for c in vv) for vv in verts]
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
             (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11),
             (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    for _ in range(level):
        cache = {}
        nf = []

        def mid(a, b):
            key = (a, b) if a < b else (b, a)
            if key not in cache:
                pa, pb = verts[a], verts[b]
                m = [(pa[k] + pb[k]) * 0.5 for k in range(3)]
                ln = math.sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2])
                verts.append(tuple(c / ln for c in m))
                cache[key] = len(verts) - 1
            return cache[key]

        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    p = np.array(verts, dtype=np.float64)
    uv = np.stack([0.5 + p[:, 0] * 0.5, 0.5 + p[:, 1] * 0.5], -1)
    return Mesh(name, p * radius, p, uv, [(np.array(faces), True)])


def quad(width, height, uv_scale=1.0, name="quad", opaque=True, vertical=False):
    """A two-triangle rectangle in the xz plane (y up normal), or in the xy
    plane (z normal) when vertical."""
    hw, hh = width * 0.5, height * 0.5
    if vertical:
        pos = [(-hw, 0.0, 0.0), (hw, 0.0, 0.0), (hw, height, 0.0), (-hw, height, 0.0)]
        nrm = [(0.0, 0.0, 1.0)] * 4
    else:
        pos = [(-hw, 0.0, -hh), (hw, 0.0, -hh), (hw, 0.0, hh), (-hw, 0.0, hh)]
        nrm = [(0.0, 1.0, 0.0)] * 4
    uv = [(0.0, 0.0), (uv_scale, 0.0), (uv_scale, uv_scale), (0.0, uv_scale)]
    return Mesh(name, pos, nrm, uv, [(np.array([[0, 1, 2], [0, 2, 3]]), opaque)])


def _random_unit_vectors(rng, n):
    v = rng.normal(3 * n).reshape(n, 3)
    return v / np.sqrt((v * v).sum(1, keepdims=True))


def soup(n_tris, median_edge, sigma=0.5, extent=1.0, rng=None, name="soup", cube=False):
    """Random triangle soup: centres uniform in [-extent, extent]^3 (or
    [0, extent]^3 if cube), random orientation, edge length lognormal with the
    given median. Each triangle has its own three vertices and face normal."""
    rng = rng or Rng(1)
    if cube:
        c = rng.uniform(3 * n_tris, 0.0, extent).reshape(n_tris, 3)
    else:
        c = rng.uniform(3 * n_tris, -extent, extent).reshape(n_tris, 3)
    edge = median_edge * dexp(sigma * rng.normal(n_tris))
    e1 = _random_unit_vectors(rng, n_tris)
    tmp = _random_unit_vectors(rng, n_tris)
    e2 = np.cross(e1, tmp)
    e2 /= np.sqrt((e2 * e2).sum(1, keepdims=True))
    e3 = np.cross(e1, e2)
    # equilateral-ish triangle of side `edge` around the centre
    s = edge[:, None]
    v0 = c + e1 * s * 0.57735
    v1 = c + (-0.5 * e1 * 0.57735 + 0.5 * e2) * s
    v2 = c + (-0.5 * e1 * 0.57735 - 0.5 * e2) * s
    pos = np.stack([v0, v1, v2], 1).reshape(-1, 3)
    nrm = np.repeat(e3, 3, axis=0)
    uv = np.tile(np.array([[0.5, 0.0], [1.0, 1.0], [0.0, 1.0]]), (n_tris, 1))
    idx = np.arange(3 * n_tris).reshape(-1, 3)
    return Mesh(name, pos, nrm, uv, [(idx, True)])


# --------------------------------------------------------------------------
# Textures
# --------------------------------------------------------------------------
class TextureData:
    def __init__(self, width, height, fmt, filt, data, name=""):
        self.width, self.height, self.format, self.filter = int(width), int(height), int(fmt), int(filt)
        self.data = np.ascontiguousarray(data)
        self.name = name
        expect = width * height * (16 if fmt == abi.RTG_FORMAT_RGBA32_FLOAT else 4)
        assert self.data.nbytes == expect, (name, self.data.nbytes, expect)


def colour_texture(rgba, name=""):
    """1x1 R32G32B32A32_SFLOAT (create_single_colour_image, src/util_functions.rs:216-237)."""
    return TextureData(1, 1, abi.RTG_FORMAT_RGBA32_FLOAT, abi.RTG_FILTER_NEAREST,
                       np.array(rgba, dtype=np.float32), name)


def checker_texture(size=256, period=32, name="checker"):
    """RGBA8 sRGB checker whose alpha alternates 0 / 255 (an alpha-MASK fence)."""
    y, x = np.mgrid[0:size, 0:size]
    on = ((x // period) + (y // period)) % 2 == 0
    img = np.zeros((size, size, 4), dtype=np.uint8)
    img[..., 0] = np.where(on, 200, 40)
    img[..., 1] = np.where(on, 140, 40)
    img[..., 2] = np.where(on, 60, 40)
    img[..., 3] = np.where(on, 255, 0)
    return TextureData(size, size, abi.RTG_FORMAT_RGBA8_SRGB, abi.RTG_FILTER_NEAREST, img, name)


def blue_noise(size=64, seed=7, sigma=1.5):
    """Void-and-cluster blue noise (Ulichney 1993) on a torus, seeded.
    Returns a size x size uint8 array of ranks scaled to 0..255."""
    n = size * size
    rng = Rng(seed)
    d = np.arange(size, dtype=np.float64)
    d = np.minimum(d, size - d)
    g1 = np.array([math.exp(-(v * v) / (2.0 * sigma * sigma)) for v in d])
    kernel = g1[:, None] * g1[None, :]  # separable Gaussian on the torus, centred at (0, 0)

    def kern_at(p):
        y, x = divmod(int(p), size)
        return np.roll(np.roll(kernel, y, 0), x, 1).reshape(-1)

    ones = n // 10
    order = np.argsort(rng.u64(n), kind="stable")
    pattern = np.zeros(n, dtype=bool)
    pattern[order[:ones]] = True
    energy = np.zeros(n)
    for p in np.nonzero(pattern)[0]:
        energy += kern_at(p)

    def tightest(pat, en):
        return int(np.argmax(np.where(pat, en, -np.inf)))

    def largest_void(pat, en):
        return int(np.argmin(np.where(pat, np.inf, en)))

    for _ in range(4 * n):
        c = tightest(pattern, energy)
        pattern[c] = False
        energy -= kern_at(c)
        v = largest_void(pattern, energy)
        if v == c:
            pattern[c] = True
            energy += kern_at(c)
            break
        pattern[v] = True
        energy += kern_at(v)
    rank = np.zeros(n, dtype=np.int64)
    # phase 1: rank the initial ones by removing tightest clusters
    pat = pattern.copy()
    en = energy.copy()
    k = int(pat.sum())
    for r in range(k - 1, -1, -1):
        c = tightest(pat, en)
        pat[c] = False
        en -= kern_at(c)
        rank[c] = r
    # phase 2 and 3: fill the largest voids
    pat = pattern.copy()
    en = energy.copy()
    for r in range(k, n):
        v = largest_void(pat, en)
        pat[v] = True
        en += kern_at(v)
        rank[v] = r
    return ((rank * 256) // n).astype(np.uint8).reshape(size, size)


def blue_noise_texture(size=64, seed=7):
    bn = blue_noise(size, seed)
    img = np.empty((size, size, 4), dtype=np.uint8)
    img[..., 0] = bn
    img[..., 1] = bn
    img[..., 2] = bn
    img[..., 3] = 255
    return TextureData(size, size, abi.RTG_FORMAT_RGBA8_UNORM, abi.RTG_FILTER_NEAREST, img, "blue noise")


# --------------------------------------------------------------------------
# Transforms and camera
# --------------------------------------------------------------------------
def rotation_from_uniform(u1, u2, u3):
    """Uniform random rotation from three U[0,1) numbers (Shoemake 1992)."""
    a, b = math.sqrt(1.0 - u1), math.sqrt(u1)
    x, y = a * math.sin(2 * math.pi * u2), a * math.cos(2 * math.pi * u2)
    z, w = b * math.sin(2 * math.pi * u3), b * math.cos(2 * math.pi * u3)
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
        [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
        [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)],
    ])


def rotation_y(angle):
    c, s = math.cos(angle), math.sin(angle)
    return np.array([[c, 0.0, s], [0.0, 1.0, 0.0], [-s, 0.0, c]])


def affine(translation, rotation=None, scale=1.0):
    """Row-major 3x4 object-to-world matrix T * R * S (the instance transform
    of src/gpu_structs.rs:55-58)."""
    m = np.zeros((3, 4))
    r = np.eye(3) if rotation is None else np.asarray(rotation, dtype=np.float64)
    m[:, :3] = r * scale
    m[:, 3] = translation
    return m.astype(np.float32)


def look_at_inverse(eye, target, up=(0.0, 1.0, 0.0)):
    """Camera-to-world matrix (the view_inverse uniform) of a right-handed
    camera looking down -z."""
    e = np.asarray(eye, dtype=np.float64)
    f = np.asarray(target, dtype=np.float64) - e
    f /= math.sqrt(float(f @ f))
    r = np.cross(f, np.asarray(up, dtype=np.float64))
    r /= math.sqrt(float(r @ r))
    u = np.cross(r, f)
    m = np.eye(4)
    m[:3, 0], m[:3, 1], m[:3, 2], m[:3, 3] = r, u, -f, e
    return m.astype(np.float32)


def proj_inverse_reversed_z(fov_y, aspect, near=0.1):
    """Inverse of a Vulkan (y down) reversed infinite-z perspective, the
    projection of src/main.rs:586-593: maps NDC (x, y, 1, 1) to a view-space
    point whose xyz is proportional to the view ray."""
    sy = 1.0 / math.tan(fov_y * 0.5)
    sx = sy / aspect
    m = np.zeros((4, 4))
    m[0, 0] = 1.0 / sx
    m[1, 1] = -1.0 / sy
    m[2, 3] = -1.0
    m[3, 2] = 1.0 / near
    return m.astype(np.float32)


def fps_view_inverse(eye, pitch, yaw):
    """Camera-to-world matrix of the reference's first-person camera
    (FirstPersonCamera::as_view_matrix, src/main.rs:1205-1231): the view
    matrix's rows are the right, up and back axes, so its inverse has them as
    columns, with the eye as translation. pitch > 0 looks up; yaw = pi looks
    along +z."""
    cy, sy, cp, sp = math.cos(yaw), math.sin(yaw), math.cos(pitch), math.sin(pitch)
    m = np.eye(4)
    m[:3, 0] = (cy, 0.0, -sy)
    m[:3, 1] = (sy * sp, cp, cy * sp)
    m[:3, 2] = (sy * cp, -sp, cp * cy)
    m[:3, 3] = eye
    return m.astype(np.float32)


def sun_direction(pitch, yaw):
    """Sun::as_normal (src/main.rs:1189-1197)."""
    return np.array([math.cos(pitch) * math.sin(yaw), math.sin(pitch), math.cos(pitch) * math.cos(yaw)],
                    dtype=np.float32)


# --------------------------------------------------------------------------
# Scene container
# --------------------------------------------------------------------------
class Model:
    """ModelInfo entry: a mesh's vertex streams and one GeometryInfo per
    geometry (diffuse, metallic-roughness, normal map image indices)."""

    def __init__(self, mesh_index, materials):
        self.mesh = int(mesh_index)
        self.materials = [tuple(int(v) for v in m) for m in materials]


class InstanceData:
    def __init__(self, transform, model, mesh, hit_group=abi.RTG_HIT_TEXTURED, flags=0, mask=0xFF):
        self.transform = np.asarray(transform, dtype=np.float32).reshape(3, 4)
        self.model = int(model)
        self.mesh = int(mesh)
        self.hit_group = int(hit_group)
        self.flags = int(flags)
        self.mask = int(mask)


class Scene:
    def __init__(self, name):
        self.name = name
        self.meshes = []
        self.models = []
        self.textures = []
        self.instances = []
        self.width = 512
        self.height = 512
        self.view_inverse = np.eye(4, dtype=np.float32)
        self.proj_inverse = np.eye(4, dtype=np.float32)
        self.sun_dir = np.array([0.0, 1.0, 0.0], dtype=np.float32)
        self.sun_radius = 0.0
        self.blue_noise_index = 2
        self.lut_index = 3
        self.frame_index = 0
        self.shadow_rays = 2
        self.max_traces = 3
        self.spp = 1
        self.seed = 0

    def add_texture(self, tex):
        self.textures.append(tex)
        return len(self.textures) - 1

    def add_mesh(self, mesh):
        self.meshes.append(mesh)
        return len(self.meshes) - 1

    def add_model(self, mesh_index, materials):
        self.models.append(Model(mesh_index, materials))
        return len(self.models) - 1

    def material(self, base, roughness, metallic=0.0):
        d = self.add_texture(colour_texture([base[0], base[1], base[2], 1.0], "base"))
        m = self.add_texture(colour_texture([1.0, roughness, metallic, 1.0], "mr"))
        return (d, m, -1)

    @property
    def instanced_triangles(self):
        return sum(self.meshes[i.mesh].triangle_count for i in self.instances)

    def instance_array(self, blas_addresses):
        """Packed 64-byte instance records (AccelerationStructureInstance::new,
        src/gpu_structs.rs:27-53)."""
        a = np.zeros(len(self.instances), dtype=abi.INSTANCE_DTYPE)
        for k, ins in enumerate(self.instances):
            a[k]["transform"] = ins.transform
            a[k]["custom_index_and_mask"] = abi.pack24_8(ins.model, ins.mask)
            a[k]["sbt_offset_and_flags"] = abi.pack24_8(ins.hit_group, ins.flags)
            a[k]["blas"] = blas_addresses[ins.mesh]
        return a

    def uniforms(self, frame_index=None):
        return abi.uniforms_from(self.view_inverse, self.proj_inverse, self.sun_dir, self.sun_radius,
                                 self.blue_noise_index, self.lut_index,
                                 self.frame_index if frame_index is None else frame_index)

    def trace_opts(self):
        o = abi.TraceOpts()
        o.shadow_rays = self.shadow_rays
        o.max_traces = self.max_traces
        o.spp = self.spp
        o.flags = 0
        return o

    # -- serialisation of every input (fixtures) --
    def to_arrays(self):
        out = {"meta": np.array([self.width, self.height, self.blue_noise_index, self.lut_index, self.frame_index,
                                 self.shadow_rays, self.max_traces, self.spp, self.seed], dtype=np.int64),
               "view_inverse": self.view_inverse, "proj_inverse": self.proj_inverse,
               "sun": np.array([*self.sun_dir, self.sun_radius], dtype=np.float32)}
        for k, m in enumerate(self.meshes):
            out[f"mesh{k}_pos"] = m.positions
            out[f"mesh{k}_nrm"] = m.normals
            out[f"mesh{k}_uv"] = m.uvs
            for g, (ix, op) in enumerate(m.geometries):
                out[f"mesh{k}_geom{g}_idx"] = ix
                out[f"mesh{k}_geom{g}_opaque"] = np.array([op])
        for k, mo in enumerate(self.models):
            out[f"model{k}"] = np.array([mo.mesh] + [v for mat in mo.materials for v in mat], dtype=np.int64)
        for k, t in enumerate(self.textures):
            out[f"tex{k}_meta"] = np.array([t.width, t.height, t.format, t.filter], dtype=np.int64)
            out[f"tex{k}_data"] = t.data
        out["inst_xf"] = np.array([i.transform for i in self.instances], dtype=np.float32).reshape(-1, 3, 4)
        out["inst_meta"] = np.array([[i.model, i.mesh, i.hit_group, i.flags, i.mask] for i in self.instances],
                                    dtype=np.int64).reshape(-1, 5)
        return out

    @staticmethod
    def from_arrays(a, name="fixture"):
        s = Scene(name)
        meta = a["meta"]
        (s.width, s.height, s.blue_noise_index, s.lut_index, s.frame_index, s.shadow_rays, s.max_traces, s.spp,
         s.seed) = [int(v) for v in meta]
        s.view_inverse = a["view_inverse"]
        s.proj_inverse = a["proj_inverse"]
        s.sun_dir = a["sun"][:3].copy()
        s.sun_radius = float(a["sun"][3])
        k = 0
        while f"mesh{k}_pos" in a:
            geoms = []
            g = 0
            while f"mesh{k}_geom{g}_idx" in a:
                geoms.append((a[f"mesh{k}_geom{g}_idx"], bool(a[f"mesh{k}_geom{g}_opaque"][0])))
                g += 1
            s.meshes.append(Mesh(f"mesh{k}", a[f"mesh{k}_pos"], a[f"mesh{k}_nrm"], a[f"mesh{k}_uv"], geoms))
            k += 1
        k = 0
        while f"model{k}" in a:
            v = [int(x) for x in a[f"model{k}"]]
            s.models.append(Model(v[0], [tuple(v[1 + 3 * i:4 + 3 * i]) for i in range((len(v) - 1) // 3)]))
            k += 1
        k = 0
        while f"tex{k}_meta" in a:
            w, h, f, fl = [int(x) for x in a[f"tex{k}_meta"]]
            s.textures.append(TextureData(w, h, f, fl, a[f"tex{k}_data"]))
            k += 1
        for xf, m in zip(a["inst_xf"], a["inst_meta"]):
            s.instances.append(InstanceData(xf, m[0], m[1], m[2], m[3], m[4]))
        return s


def _base_textures(scene, bn_seed):
    """Texture slots 0..3 as in src/main.rs:416-460: green, pink, blue noise,
    GGX LUT (the LUT is loaded but unused by the BRDF, pbr.glsl:107-158)."""
    scene.add_texture(colour_texture([0.0, 1.0, 0.0, 1.0], "green"))
    scene.add_texture(colour_texture([1.0, 0.0, 1.0, 1.0], "pink"))
    scene.add_texture(blue_noise_texture(64, bn_seed))
    scene.add_texture(colour_texture([1.0, 1.0, 1.0, 1.0], "ggx lut placeholder"))


# --------------------------------------------------------------------------
# Workloads
# --------------------------------------------------------------------------
def small_scene(seed=0x5EED0000, width=512, height=512, n_tori=16):
    """Small plain-method scene: n_tori tori (48 x 24 segments, R=1, r=0.35)
    at y=1 with x, z uniform in [-8, 8] and random rotations, a 40 x 40 ground
    quad, 60 degree pinhole camera, fixed sun, one hard shadow ray per pixel
    (sun_radius 0), roughness U[0.2, 0.8], base colour U[0, 1]^3."""
    rng = Rng(seed)
    s = Scene("small")
    s.seed = seed
    _base_textures(s, seed & 0xFFFF)
    t_mesh = s.add_mesh(torus(48, 24, 1.0, 0.35))
    g_mesh = s.add_mesh(quad(40.0, 40.0, 4.0, "ground"))
    ground_model = s.add_model(g_mesh, [s.material((0.55, 0.55, 0.5), 0.8)])
    s.instances.append(InstanceData(affine((0.0, 0.0, 0.0)), ground_model, g_mesh))
    pos = rng.uniform(2 * n_tori, -8.0, 8.0).reshape(n_tori, 2)
    rot = rng.uniform(3 * n_tori).reshape(n_tori, 3)
    cols = rng.uniform(3 * n_tori).reshape(n_tori, 3)
    rough = rng.uniform(n_tori, 0.2, 0.8)
    for k in range(n_tori):
        m = s.add_model(t_mesh, [s.material(cols[k], rough[k])])
        R = rotation_from_uniform(*rot[k])
        s.instances.append(InstanceData(affine((pos[k, 0], 1.0, pos[k, 1]), R), m, t_mesh))
    s.width, s.height = width, height
    s.view_inverse = look_at_inverse((0.0, 11.0, -17.0), (0.0, 0.5, 0.0))
    s.proj_inverse = proj_inverse_reversed_z(math.radians(60.0), width / height)
    s.sun_dir = sun_direction(0.9, 0.6)
    s.sun_radius = 0.0
    s.shadow_rays = 1
    s.max_traces = 3
    s.spp = 1
    return s


def default_scene(seed=0x5EED0005, width=320, height=180, n_tori=100, eye=(0.0, 3.5, -3.0), pitch=-0.75):
    """The reference's default scene (src/scene.rs:95-156), seeded: a 20 x 20
    ground plane (the 2 x 2 plane scaled by 10), a torus at (0, 1, 0), an
    icosphere where the reference loads its glTF character, the double-sided
    Portal plane at y = 1 (the 2 x 2 plane translated by (0, 1, 0), hit group
    Portal: rays that hit it continue from the hit point moved up by 5), the
    double-sided alpha-clipped fence at (2, 0, 2), and n_tori small tori at
    T(U[-10,10], U[0.5,2.5], U[-10,10]) Ry(U[0,100)) S(U[0.1,0.3]), half
    Textured and half Mirror. Camera: the reference's first-person camera
    (yaw 180 degrees, 59 degree field of view, src/main.rs:568-593), raised and
    pitched down so that the portal and the ground its rays reach are in view;
    sun pitch 0.5, yaw 1.0, radius 0.05 (src/main.rs:577-584); 2 shadow rays
    per hit as closest_hit_textured.glsl:195."""
    rng = Rng(seed)
    s = Scene("default")
    s.seed = seed
    _base_textures(s, seed & 0xFFFF)
    checker = s.add_texture(checker_texture(256, 32))
    plane = s.add_mesh(quad(2.0, 2.0, 1.0, "plane"))
    t_mesh = s.add_mesh(torus(48, 12, 1.0, 0.35))
    i_mesh = s.add_mesh(icosphere(3, 1.0))
    fence_mesh = s.add_mesh(quad(2.0, 2.0, 2.0, "fence", opaque=False, vertical=True))
    ground = s.add_model(plane, [s.material((0.55, 0.55, 0.5), 0.8)])
    s.instances.append(InstanceData(affine((0.0, 0.0, 0.0), None, 10.0), ground, plane))
    tm = s.add_model(t_mesh, [s.material((0.8, 0.3, 0.2), 0.4)])
    s.instances.append(InstanceData(affine((0.0, 1.0, 0.0)), tm, t_mesh))
    im = s.add_model(i_mesh, [s.material((0.7, 0.7, 0.75), 0.3, 1.0)])
    s.instances.append(InstanceData(affine((-2.0, 0.5, -1.0), rotation_y(math.radians(150.0)), 0.5), im, i_mesh))
    pm = s.add_model(plane, [s.material((1.0, 1.0, 1.0), 0.5)])
    s.instances.append(InstanceData(affine((0.0, 1.0, 0.0)), pm, plane, abi.RTG_HIT_PORTAL,
                                    abi.RTG_INSTANCE_TRIANGLE_FACING_CULL_DISABLE))
    fm = s.add_model(fence_mesh, [(checker, s.material((1, 1, 1), 0.6)[1], -1)])
    s.instances.append(InstanceData(affine((2.0, 0.0, 2.0)), fm, fence_mesh, abi.RTG_HIT_TEXTURED,
                                    abi.RTG_INSTANCE_TRIANGLE_FACING_CULL_DISABLE))
    pos = rng.uniform(3 * n_tori).reshape(n_tori, 3)
    ang = rng.uniform(n_tori, 0.0, 100.0)
    scl = rng.uniform(n_tori, 0.1, 0.3)
    mirror = rng.uniform(n_tori) < 0.5
    cols = rng.uniform(3 * n_tori).reshape(n_tori, 3)
    for k in range(n_tori):
        tr = (-10.0 + 20.0 * pos[k, 0], 0.5 + 2.0 * pos[k, 1], -10.0 + 20.0 * pos[k, 2])
        m = s.add_model(t_mesh, [s.material(cols[k], 0.5)])
        s.instances.append(InstanceData(affine(tr, rotation_y(ang[k]), scl[k]), m, t_mesh,
                                        abi.RTG_HIT_MIRROR if mirror[k] else abi.RTG_HIT_TEXTURED))
    s.width, s.height = width, height
    s.view_inverse = fps_view_inverse(eye, pitch, math.pi)
    s.proj_inverse = proj_inverse_reversed_z(math.radians(59.0), width / height)
    s.sun_dir = sun_direction(0.5, 1.0)
    s.sun_radius = 0.05
    s.shadow_rays = 2
    s.max_traces = 3
    s.spp = 1
    return s


class AnimatedScene(Scene):
    """A scene whose object rotations are re-randomised every frame
    (the per-frame transform write + TLAS UPDATE of src/scene.rs:167-204)."""

    def __init__(self, name):
        super().__init__(name)
        self.animated = []  # (instance index, translation, scale)

    def transforms_for_frame(self, frame):
        rng = Rng(self.seed * 1000003 + frame + 1)
        rot = rng.uniform(3 * len(self.animated)).reshape(-1, 3)
        out = [i.transform.copy() for i in self.instances]
        for k, (idx, tr, sc) in enumerate(self.animated):
            out[idx] = affine(tr, rotation_from_uniform(*rot[k]), sc)
        return out

    def set_frame(self, frame):
        for k, xf in enumerate(self.transforms_for_frame(frame)):
            self.instances[k].transform = xf
        self.frame_index = frame


def large_scene(seed=0x5EED0001, width=1920, height=1080, n_objects=1024, n_fences=64, n_mirrors=32,
                n_soup_meshes=16, soup_tris=1200, extent=36.0, p_torus=0.45, p_ico=0.10):
    """The 1080p ~2M-triangle animated scene: n_objects instances drawn from
    tori (48 x 24, probability p_torus), level-4 icospheres (5,120 tris,
    p_ico) and random triangle soups (lognormal edge length, median 0.05),
    n_fences alpha-clipped quads with a 256x256 RGBA8 checker alpha,
    n_mirrors mirror quads, a ground quad; objects spread over
    [-extent, extent]^2 with the camera scaled to match; 4 blue-noise
    soft-shadow rays per hit (64x64 mask, 2 degree cone); every frame
    re-randomises object rotations (TLAS refit)."""
    rng = Rng(seed)
    s = AnimatedScene("large")
    s.seed = seed
    _base_textures(s, seed & 0xFFFF)
    checker = s.add_texture(checker_texture(256, 32))
    t_mesh = s.add_mesh(torus(48, 24, 1.0, 0.35))
    i_mesh = s.add_mesh(icosphere(4, 1.0))
    soup_meshes = [s.add_mesh(soup(soup_tris, 0.05, 0.5, 1.0, Rng(seed + 17 + k), f"soup{k}"))
                   for k in range(n_soup_meshes)]
    fence_mesh = s.add_mesh(quad(1.6, 2.0, 2.0, "fence", opaque=False, vertical=True))
    mirror_mesh = s.add_mesh(quad(3.0, 2.5, 1.0, "mirror", vertical=True))
    ground_mesh = s.add_mesh(quad(240.0, 240.0, 24.0, "ground"))

    gm = s.add_model(ground_mesh, [s.material((0.5, 0.5, 0.48), 0.85)])
    s.instances.append(InstanceData(affine((0.0, 0.0, 0.0)), gm, ground_mesh))

    kind = rng.uniform(n_objects)
    pos = rng.uniform(2 * n_objects, -extent, extent).reshape(-1, 2)
    height_ = rng.uniform(n_objects, 0.6, 2.2)
    scale = rng.uniform(n_objects, 0.6, 1.4)
    cols = rng.uniform(3 * n_objects).reshape(-1, 3)
    rough = rng.uniform(n_objects, 0.2, 0.8)
    metal = (rng.uniform(n_objects) < 0.2).astype(np.float64)
    soup_pick = rng.integers(n_objects, 0, n_soup_meshes)
    for k in range(n_objects):
        if kind[k] < p_torus:
            mesh = t_mesh
        elif kind[k] < p_torus + p_ico:
            mesh = i_mesh
        else:
            mesh = soup_meshes[int(soup_pick[k])]
        m = s.add_model(mesh, [s.material(cols[k], rough[k], metal[k])])
        tr = (pos[k, 0], height_[k], pos[k, 1])
        s.animated.append((len(s.instances), tr, float(scale[k])))
        s.instances.append(InstanceData(affine(tr, None, scale[k]), m, mesh))
    inner = extent * (30.0 / 36.0)
    fpos = rng.uniform(2 * n_fences, -inner, inner).reshape(-1, 2)
    fang = rng.uniform(n_fences, 0.0, 2 * math.pi)
    fm = s.add_model(fence_mesh, [(checker, s.material((1, 1, 1), 0.6)[1], -1)])
    for k in range(n_fences):
        s.instances.append(InstanceData(affine((fpos[k, 0], 0.0, fpos[k, 1]), rotation_y(fang[k])), fm, fence_mesh,
                                        abi.RTG_HIT_TEXTURED, abi.RTG_INSTANCE_TRIANGLE_FACING_CULL_DISABLE))
    mpos = rng.uniform(2 * n_mirrors, -inner, inner).reshape(-1, 2)
    mang = rng.uniform(n_mirrors, 0.0, 2 * math.pi)
    mm = s.add_model(mirror_mesh, [s.material((0.9, 0.9, 0.9), 0.1)])
    for k in range(n_mirrors):
        s.instances.append(InstanceData(affine((mpos[k, 0], 0.0, mpos[k, 1]), rotation_y(mang[k])), mm,
                                        mirror_mesh, abi.RTG_HIT_MIRROR,
                                        abi.RTG_INSTANCE_TRIANGLE_FACING_CULL_DISABLE))
    s.width, s.height = width, height
    cam = extent / 36.0
    s.view_inverse = look_at_inverse((0.0, 16.0 * cam, -44.0 * cam), (0.0, 0.0, -2.0 * cam))
    s.proj_inverse = proj_inverse_reversed_z(math.radians(60.0), width / height)
    s.sun_dir = sun_direction(0.85, 2.4)
    s.sun_radius = math.tan(math.radians(1.0))  # 2 degree cone
    s.shadow_rays = 4
    s.max_traces = 3
    s.spp = 1
    s.set_frame(0)
    return s


BATCH_RESOLUTIONS = ((128, 128), (512, 512), (1280, 720), (1920, 1080))


def batch_scene(seed, target_tris, width, height, spp):
    """One scene of the batched workload with about target_tris instanced
    triangles: the generator of large_scene with as many objects as the
    target needs (one triangle soup of target_tris - 2 triangles when it is
    below a few thousand), fences and mirrors in proportion, and the camera
    scaled to the spread of the objects."""
    per_object = 0.45 * 2304 + 0.10 * 5120 + 0.45 * 1200
    if target_tris < 3 * 2304:
        s = large_scene(seed, width, height, n_objects=1, n_fences=0, n_mirrors=0, n_soup_meshes=1,
                        soup_tris=max(1, int(target_tris) - 2), extent=2.0, p_torus=0.0, p_ico=0.0)
    else:
        n = max(1, int(round((target_tris - 2) / per_object)))
        s = large_scene(seed, width, height, n_objects=n, n_fences=min(64, n // 16), n_mirrors=min(32, n // 32),
                        n_soup_meshes=min(16, n), extent=max(4.0, 36.0 * math.sqrt(n / 1024.0)))
    s.spp = int(spp)
    return s


def batch_workload(n_scenes=64, seed=0x5EED0003, tri_range=(1_000, 500_000), resolutions=BATCH_RESOLUTIONS,
                   spp_range=(1, 4)):
    """The batched launch: n_scenes independent scenes with triangle counts
    log-uniform in tri_range, resolutions drawn from `resolutions` and
    samples per pixel uniform in spp_range (inclusive)."""
    rng = Rng(seed)
    lo, hi = math.log(tri_range[0]), math.log(tri_range[1])
    tris = [math.exp(x) for x in rng.uniform(n_scenes, lo, hi)]
    res = rng.integers(n_scenes, 0, len(resolutions))
    spp = rng.integers(n_scenes, spp_range[0], spp_range[1] + 1)
    out = []
    for k in range(n_scenes):
        w, h = resolutions[int(res[k])]
        out.append(batch_scene(seed + 1 + k, float(tris[k]), w, h, int(spp[k])))
    return out


SERVE_RESOLUTIONS = ((256, 256), (1280, 720), (1920, 1080))


def served_stream(n_scenes=32, n_requests=160, seed=0x5EED0004, resolutions=SERVE_RESOLUTIONS, spp_range=(1, 8),
                  scene_size=(256, 256)):
    """The served stream: n_scenes scenes of the small-scene generator stay
    resident; each of n_requests requests names one of them, a resolution
    from `resolutions`, samples per pixel uniform in spp_range (inclusive) and
    new random rotations of every torus (a transform update and refit).
    Returns (scenes, requests), a request being a dict with scene, width,
    height, spp, frame and transforms (row-major 3x4 per instance)."""
    scenes_ = [small_scene(seed=seed + 1 + k, width=scene_size[0], height=scene_size[1]) for k in range(n_scenes)]
    rng = Rng(seed)
    reqs = []
    for i in range(n_requests):
        sid = int(rng.integers(1, 0, n_scenes)[0])
        w, h = resolutions[int(rng.integers(1, 0, len(resolutions))[0])]
        spp = int(rng.integers(1, spp_range[0], spp_range[1] + 1)[0])
        xf = [ins.transform.copy() for ins in scenes_[sid].instances]
        rot = rng.uniform(3 * len(xf)).reshape(-1, 3)
        for k in range(1, len(xf)):
            xf[k] = affine(xf[k][:, 3], rotation_from_uniform(*rot[k]), 1.0)
        reqs.append({"scene": sid, "width": w, "height": h, "spp": spp, "frame": i, "transforms": xf})
    return scenes_, reqs


def incoherent_rays(n_rays, lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0), seed=0x5EED0002):
    """Rays with origins uniform in the box and directions uniform on the
    sphere (z = 2u - 1, phi = 2 pi v), tmin = 0, tmax = +inf."""
    rng = Rng(seed)
    lo = np.asarray(lo, dtype=np.float64)
    hi = np.asarray(hi, dtype=np.float64)
    o = lo + (hi - lo) * rng.uniform(3 * n_rays).reshape(-1, 3)
    z = 2.0 * rng.uniform(n_rays) - 1.0
    phi = 2.0 * math.pi * rng.uniform(n_rays)
    rxy = np.sqrt(np.maximum(0.0, 1.0 - z * z))
    d = np.stack([rxy * dcos(phi), rxy * dsin(phi), z], -1)
    rays = np.zeros(n_rays, dtype=abi.RAY_DTYPE)
    rays["origin"] = o.astype(np.float32)
    rays["direction"] = d.astype(np.float32)
    rays["tmin"] = 0.0
    rays["tmax"] = np.inf
    return rays


def incoherent_scene(n_tris=1_000_000, median_edge=0.01, seed=0x5EED0002):
    """One BLAS of n_tris random triangles in the unit cube (centres U[0,1]^3,
    lognormal edge length with the given median), one identity instance."""
    s = Scene("incoherent")
    s.seed = seed
    _base_textures(s, seed & 0xFFFF)
    mesh = s.add_mesh(soup(n_tris, median_edge, 0.5, 1.0, Rng(seed + 1), "cube soup", cube=True))
    m = s.add_model(mesh, [s.material((0.8, 0.8, 0.8), 0.5)])
    s.instances.append(InstanceData(affine((0.0, 0.0, 0.0)), m, mesh))
    return s
