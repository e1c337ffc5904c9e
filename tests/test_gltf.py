"""glTF ingestion (the reference's Model::load_gltf semantics): accessors,
material-to-geometry and texture-table order, fallbacks, the reference's own
assets when the reference tree is present, and a scene made of glTF models
traced by the oracle (BVH traversal equal to brute force)."""
import json
import os
import struct

import numpy as np
import pytest

import oracle_host as oh
from gltf_assets import fence_glb, gltf_scene, plain_glb
from rtg import abi, gltf, scenes

REF_RESOURCES = "/root/reference/resources"


def test_round_trip_materials_textures_and_geometries():
    s = scenes.Scene("t")
    base = len(s.textures)
    mesh_i, model_i = gltf.load_glb(fence_glb(), s, "fence")
    m = s.meshes[mesh_i]
    q = scenes.quad(2.0, 2.0, 1.0, "fence", vertical=True)
    t = scenes.torus(16, 8, 0.6, 0.2)
    # shared vertex streams: quad then torus; the torus indices are offset
    assert np.array_equal(m.positions, np.concatenate([q.positions, t.positions]))
    assert np.array_equal(m.uvs, np.concatenate([q.uvs, t.uvs]))
    (ix0, op0), (ix1, op1) = m.geometries
    assert np.array_equal(ix0, q.geometries[0][0]) and not op0  # alpha MASK: not opaque
    assert np.array_equal(ix1, t.geometries[0][0] + len(q.positions)) and op1
    # texture entries in push order: m0 diffuse, m0 MR, m1 diffuse, m1 MR, m1 normal
    tx = s.textures[base:]
    assert [(e.format, e.filter, e.width) for e in tx] == [
        (abi.RTG_FORMAT_RGBA8_SRGB, abi.RTG_FILTER_NEAREST, 32),
        (abi.RTG_FORMAT_RGBA32_FLOAT, abi.RTG_FILTER_NEAREST, 1),
        (abi.RTG_FORMAT_RGBA32_FLOAT, abi.RTG_FILTER_NEAREST, 1),
        (abi.RTG_FORMAT_RGBA32_FLOAT, abi.RTG_FILTER_NEAREST, 1),
        (abi.RTG_FORMAT_RGBA8_UNORM, abi.RTG_FILTER_LINEAR, 16)]
    assert np.array_equal(tx[1].data, np.float32([1.0, 0.7, 0.0, 1.0]))   # [1, roughness, metallic, 1]
    assert np.array_equal(tx[2].data, np.float32([0.3, 0.6, 0.9, 1.0]))   # baseColorFactor as the backup
    assert np.array_equal(tx[3].data, np.float32([1.0, 0.35, 0.5, 1.0]))
    assert s.models[model_i].materials == [(base, base + 1, -1), (base + 2, base + 3, base + 4)]


def test_no_materials_uses_the_fallback_image():
    s = scenes.Scene("t")
    _, model_i = gltf.load_glb(plain_glb(), s, "ball", fallback_image_index=7)
    (d, mr, nm), = s.models[model_i].materials
    assert d == 7 and nm == -1
    assert np.array_equal(s.textures[mr].data, np.float32([1.0, 1.0, 0.0, 1.0]))
    assert s.meshes[s.models[model_i].mesh].geometries[0][1] is True


def _glb(doc, blob):
    js = json.dumps(doc).encode()
    js += b" " * ((4 - len(js) % 4) % 4)
    blob = bytes(blob) + b"\0" * ((4 - len(blob) % 4) % 4)
    return (struct.pack("<III", gltf.GLB_MAGIC, 2, 28 + len(js) + len(blob)) + struct.pack("<II", len(js), 0x4E4F534A)
            + js + struct.pack("<II", len(blob), 0x004E4942) + blob)


def test_strided_and_normalised_accessors():
    # interleaved vec3 positions (stride 20: 12 B position + 2 x u16 normalised uv)
    pos = np.arange(12, dtype=np.float32).reshape(4, 3)
    uv = np.array([[0, 65535], [32768, 0], [65535, 65535], [1, 2]], np.uint16)
    inter = b"".join(pos[k].tobytes() + uv[k].tobytes() + b"\0\0\0\0" for k in range(4))
    doc = {"asset": {"version": "2.0"}, "buffers": [{"byteLength": len(inter)}],
           "bufferViews": [{"buffer": 0, "byteOffset": 0, "byteLength": len(inter), "byteStride": 20}],
           "accessors": [{"bufferView": 0, "componentType": 5126, "count": 4, "type": "VEC3"},
                         {"bufferView": 0, "byteOffset": 12, "componentType": 5123, "count": 4, "type": "VEC2",
                          "normalized": True}]}
    d, b = gltf.parse_glb(_glb(doc, inter))
    assert np.array_equal(gltf.read_accessor(d, b, 0), pos)
    got = gltf._as_f32(gltf.read_accessor(d, b, 1), True)
    assert np.array_equal(got, uv.astype(np.float32) / np.float32(65535))


def test_malformed_input_is_refused():
    with pytest.raises(gltf.GltfError):
        gltf.parse_glb(b"\0" * 40)
    with pytest.raises(gltf.GltfError):
        gltf.parse_glb(fence_glb()[:100])
    doc = {"asset": {"version": "2.0"}, "meshes": [{"primitives": [{"attributes": {}, "mode": 1}]}]}
    with pytest.raises(gltf.GltfError):
        gltf.load_glb(_glb(doc, b""), scenes.Scene("t"))


def test_glb_scene_traversal_equals_brute_force():
    s = gltf_scene(32, 24)
    hs = oh.HostScene(s)
    rays = scenes.incoherent_rays(4000, (-5, 0.0, -3), (5, 3, 3), seed=21)
    for mode in (abi.RTG_QUERY_CLOSEST, abi.RTG_QUERY_OCCLUSION):
        a, _ = hs.trace_rays(rays, mode=mode)
        b, _ = hs.trace_rays(rays, mode=mode, brute=True)
        assert np.array_equal(np.asarray(a).view(np.uint32), np.asarray(b).view(np.uint32))
    rgba, rgb, hits, cnt = hs.trace_image()
    assert cnt.anyhit_calls > 0  # the alpha-MASK fence runs the any-hit stage
    assert (hits["instance"] != abi.RTG_NO_HIT).mean() > 0.3


@pytest.mark.skipif(not os.path.isdir(REF_RESOURCES), reason="reference tree not present")
@pytest.mark.parametrize("name, tris, opaque", [("plane.glb", 2, [True]), ("tori.glb", None, None),
                                                ("fence.glb", None, None), ("lain.glb", 45448, None)])
def test_reference_assets(name, tris, opaque):
    data = open(os.path.join(REF_RESOURCES, name), "rb").read()
    s = scenes.Scene("ref")
    scenes._base_textures(s, 7)
    mesh_i, model_i = gltf.load_glb(data, s, name)
    m = s.meshes[mesh_i]
    if tris is not None:
        assert m.triangle_count == tris
    if opaque is not None:
        assert [op for _, op in m.geometries] == opaque
    if name == "fence.glb":
        assert not all(op for _, op in m.geometries)  # the alpha-MASK fence
    assert len(m.positions) == len(m.normals) == len(m.uvs)
    for ix, _ in m.geometries:
        assert len(ix) == 0 or ix.max() < len(m.positions)
    assert len(s.models[model_i].materials) == len(m.geometries)
    # the host build over the asset's triangles: every primitive exactly once
    blas = oh.build_blas(m)
    assert blas.info().prim_count == m.triangle_count
