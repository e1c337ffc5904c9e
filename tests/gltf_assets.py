"""Synthetic .glb assets (written with rtg.gltf.encode_glb) and a scene built
from them, shared by the CPU and GPU glTF tests."""
import math

import numpy as np

from rtg import abi, gltf, image, scenes


def checker_png(size=32, period=8, alpha=True):
    y, x = np.mgrid[0:size, 0:size]
    on = ((x // period + y // period) % 2) == 0
    a = np.zeros((size, size, 4), np.uint8)
    a[..., 0] = np.where(on, 230, 40)
    a[..., 1] = np.where(on, 180, 90)
    a[..., 2] = 60
    a[..., 3] = np.where(on, 255, 0) if alpha else 255
    return image.encode_png(a)


def normal_png(size=16):
    a = np.zeros((size, size, 4), np.uint8)
    y, x = np.mgrid[0:size, 0:size]
    a[..., 0] = (128 + 40 * np.sin(x * 0.7)).astype(np.uint8)
    a[..., 1] = (128 + 40 * np.cos(y * 0.5)).astype(np.uint8)
    a[..., 2] = 230
    a[..., 3] = 255
    return image.encode_png(a)


def prim_of(mesh, material=None):
    ix, _ = mesh.geometries[0]
    p = {"positions": mesh.positions, "normals": mesh.normals, "uvs": mesh.uvs, "indices": ix.reshape(-1)}
    if material is not None:
        p["material"] = material
    return p


def fence_glb():
    """A vertical quad with an alpha-MASK checker (nearest) plus a torus with a
    factor-only material and a normal map: two materials, two primitives."""
    q = scenes.quad(2.0, 2.0, 1.0, "fence", vertical=True)
    t = scenes.torus(16, 8, 0.6, 0.2)
    mats = [{"base_texture": 0, "alpha_mode": "MASK", "nearest": True, "roughness": 0.7, "metallic": 0.0},
            {"base_colour": [0.3, 0.6, 0.9, 1.0], "roughness": 0.35, "metallic": 0.5, "normal_texture": 1}]
    return gltf.encode_glb([prim_of(q, 0), prim_of(t, 1)], mats, [checker_png(), normal_png()])


def plain_glb():
    """A mesh without materials (the fallback diffuse image is used)."""
    return gltf.encode_glb([prim_of(scenes.icosphere(1, 0.5))])


def gltf_scene(width=64, height=48, seed=0x5EED0200):
    """Ground + glTF models instanced a few times, looked at from the front."""
    s = scenes.Scene("gltf")
    s.seed = seed
    scenes._base_textures(s, seed & 0xFFFF)
    fence_mesh, fence_model = gltf.load_glb(fence_glb(), s, "fence")
    ball_mesh, ball_model = gltf.load_glb(plain_glb(), s, "ball", fallback_image_index=1)
    ground = s.add_mesh(scenes.quad(20.0, 20.0, 4.0, "ground"))
    gm = s.add_model(ground, [s.material((0.5, 0.5, 0.45), 0.8)])
    s.instances.append(scenes.InstanceData(scenes.affine((0, 0, 0)), gm, ground))
    rng = scenes.Rng(seed)
    for k in range(4):
        tr = (-3.0 + 2.0 * k, 1.0, 0.5 * (k % 2))
        s.instances.append(scenes.InstanceData(scenes.affine(tr, scenes.rotation_y(0.4 * k)), fence_model, fence_mesh,
                                               abi.RTG_HIT_TEXTURED, abi.RTG_INSTANCE_TRIANGLE_FACING_CULL_DISABLE))
        s.instances.append(scenes.InstanceData(
            scenes.affine((tr[0], 0.6, 2.0), scenes.rotation_from_uniform(*rng.uniform(3))), ball_model, ball_mesh))
    s.width, s.height = width, height
    s.view_inverse = scenes.look_at_inverse((0.0, 3.0, -8.0), (0.0, 0.8, 0.5))
    s.proj_inverse = scenes.proj_inverse_reversed_z(math.radians(60.0), width / height)
    s.sun_dir = scenes.sun_direction(0.7, 2.2)
    s.sun_radius = math.tan(math.radians(1.0))
    s.shadow_rays = 2
    return s
