"""glTF 2.0 binary (.glb) ingestion into a scene, as the reference loads its
models (Model::load_gltf, src/util_structs.rs:1052-1156), and a small .glb
writer for round trips.

What the reference does, and this module repeats:
* one geometry per material, in material order; a geometry is opaque iff the
  material's alphaMode is OPAQUE (src/util_structs.rs:992);
* per material, in this order, one texture-table entry each
  (ImageManager::push_image, :1330-1349): the base colour texture as
  RGBA8 sRGB, or a 1x1 RGBA32F image of baseColorFactor when there is none
  (the factor is not applied to a texture); the metallic-roughness texture,
  also as RGBA8 sRGB, or 1x1 [1, roughness, metallic, 1]; the normal texture
  as RGBA8 UNORM, or no entry and index -1 (:986-1026);
* a texture is filtered linearly unless its sampler's magFilter is NEAREST
  (:952-953); images come from buffer views holding PNG data (:955-967),
  decoded to RGBA8 (`to_rgba8`, src/util_functions.rs:239-265);
* without materials, one opaque geometry with the caller's fallback diffuse
  image and a 1x1 [1, 1, 0, 1] metallic-roughness image (:1087-1108);
* every primitive of every mesh, in order, appends its POSITION, NORMAL and
  TEXCOORD_0 to shared vertex streams and its indices, offset by the vertex
  count so far, to the geometry of its material (material 0 when unset)
  (:1110-1134); node transforms are not applied.

Deviation: primitives whose mode is not TRIANGLES (4) are refused, where
the reference would read their indices as triangles.
"""
import io
import json
import struct

import numpy as np

from . import abi
from .scenes import Mesh, TextureData

GLB_MAGIC = 0x46546C67  # 'glTF'
CHUNK_JSON = 0x4E4F534A
CHUNK_BIN = 0x004E4942
NEAREST = 9728

_COMPONENT = {5120: np.int8, 5121: np.uint8, 5122: np.int16, 5123: np.uint16, 5125: np.uint32, 5126: np.float32}
_COUNT = {"SCALAR": 1, "VEC2": 2, "VEC3": 3, "VEC4": 4, "MAT2": 4, "MAT3": 9, "MAT4": 16}


class GltfError(ValueError):
    pass


def parse_glb(data):
    """(document dict, BIN chunk bytes) of a .glb file."""
    data = bytes(data)
    if len(data) < 20:
        raise GltfError("too short for a .glb")
    magic, version, length = struct.unpack_from("<III", data, 0)
    if magic != GLB_MAGIC or version != 2:
        raise GltfError("not a glTF 2.0 binary")
    if length > len(data):
        raise GltfError("truncated .glb")
    pos, doc, blob = 12, None, b""
    while pos + 8 <= length:
        clen, ctype = struct.unpack_from("<II", data, pos)
        body = data[pos + 8:pos + 8 + clen]
        if len(body) != clen:
            raise GltfError("truncated chunk")
        if ctype == CHUNK_JSON and doc is None:
            doc = json.loads(body.decode("utf-8"))
        elif ctype == CHUNK_BIN and not blob:
            blob = body
        pos += 8 + clen
    if doc is None:
        raise GltfError("no JSON chunk")
    return doc, blob


def _view_bytes(doc, blob, view_index):
    v = doc["bufferViews"][view_index]
    if v.get("buffer", 0) != 0:
        raise GltfError("only the binary chunk buffer is supported")
    off = v.get("byteOffset", 0)
    return blob[off:off + v["byteLength"]], v.get("byteStride")


def read_accessor(doc, blob, index):
    """The accessor's elements as an (count, components) array of its
    component type (normalisation is left to the caller)."""
    a = doc["accessors"][index]
    if "sparse" in a:
        raise GltfError("sparse accessors are not supported")
    dt = np.dtype(_COMPONENT[a["componentType"]]).newbyteorder("<")
    nc = _COUNT[a["type"]]
    count = a["count"]
    if "bufferView" not in a:
        return np.zeros((count, nc), dt)
    raw, stride = _view_bytes(doc, blob, a["bufferView"])
    off = a.get("byteOffset", 0)
    elem = dt.itemsize * nc
    stride = stride or elem
    if count and off + (count - 1) * stride + elem > len(raw):
        raise GltfError(f"accessor {index} runs past its buffer view")
    out = np.empty((count, nc), dt)
    buf = np.frombuffer(raw, np.uint8)
    if stride == elem:
        out[:] = np.frombuffer(raw, dt, count=count * nc, offset=off).reshape(count, nc)
    else:
        rows = np.lib.stride_tricks.as_strided(buf[off:], shape=(count, elem), strides=(stride, 1))
        out[:] = np.ascontiguousarray(rows).view(dt).reshape(count, nc)
    return out


def _as_f32(a, normalized):
    """gltf `into_f32`: float stays, normalised integers map to [0, 1] or
    [-1, 1]."""
    if a.dtype == np.float32:
        return a.astype(np.float32)
    if not normalized:
        raise GltfError("integer texture coordinates must be normalised")
    info = np.iinfo(a.dtype)
    f = a.astype(np.float32) / np.float32(info.max)
    return np.maximum(f, np.float32(-1.0)) if info.min < 0 else f


def decode_png(data):
    """(width, height, RGBA8 bytes) of PNG data."""
    from PIL import Image
    with Image.open(io.BytesIO(bytes(data))) as im:
        if im.format != "PNG":
            raise GltfError("images must be PNG")
        rgba = np.asarray(im.convert("RGBA"), dtype=np.uint8)
    return rgba.shape[1], rgba.shape[0], rgba


def _texture_entry(doc, blob, tex_index, fmt, name):
    t = doc["textures"][tex_index]
    filt = abi.RTG_FILTER_LINEAR
    if "sampler" in t and doc["samplers"][t["sampler"]].get("magFilter") == NEAREST:
        filt = abi.RTG_FILTER_NEAREST
    img = doc["images"][t["source"]]
    if "bufferView" not in img:
        raise GltfError("image source is a uri, which is not supported")
    raw, _ = _view_bytes(doc, blob, img["bufferView"])
    w, h, rgba = decode_png(raw)
    return TextureData(w, h, fmt, filt, rgba.reshape(-1), name)


def _colour_entry(rgba, name):
    return TextureData(1, 1, abi.RTG_FORMAT_RGBA32_FLOAT, abi.RTG_FILTER_NEAREST,
                       np.asarray(rgba, np.float32), name)


def load_glb(data, scene, name="model", fallback_image_index=0):
    """Adds the .glb's textures, one mesh and one model to `scene`; returns
    (mesh index, model index)."""
    doc, blob = parse_glb(data)
    materials, geometries = [], []
    for i, m in enumerate(doc.get("materials", [])):
        pbr = m.get("pbrMetallicRoughness", {})
        if "baseColorTexture" in pbr:
            d = scene.add_texture(_texture_entry(doc, blob, pbr["baseColorTexture"]["index"],
                                                 abi.RTG_FORMAT_RGBA8_SRGB, f"{name} image {i} diffuse"))
        else:
            d = scene.add_texture(_colour_entry(pbr.get("baseColorFactor", [1.0, 1.0, 1.0, 1.0]),
                                                f"{name} image {i} diffuse"))
        if "metallicRoughnessTexture" in pbr:
            mr = scene.add_texture(_texture_entry(doc, blob, pbr["metallicRoughnessTexture"]["index"],
                                                  abi.RTG_FORMAT_RGBA8_SRGB, f"{name} image {i} metallic roughness"))
        else:
            mr = scene.add_texture(_colour_entry([1.0, pbr.get("roughnessFactor", 1.0), pbr.get("metallicFactor", 1.0),
                                                  1.0], f"{name} image {i} metallic roughness"))
        nm = -1
        if "normalTexture" in m:
            nm = scene.add_texture(_texture_entry(doc, blob, m["normalTexture"]["index"], abi.RTG_FORMAT_RGBA8_UNORM,
                                                  f"{name} image {i} normal map"))
        materials.append((d, mr, nm))
        geometries.append(([], m.get("alphaMode", "OPAQUE") == "OPAQUE"))
    if not materials:
        mr = scene.add_texture(_colour_entry([1.0, 1.0, 0.0, 1.0], f"{name} default metallic roughness"))
        materials.append((fallback_image_index, mr, -1))
        geometries.append(([], True))
    positions, normals, uvs = [], [], []
    nverts = 0
    for mesh in doc.get("meshes", []):
        for prim in mesh["primitives"]:
            if prim.get("mode", 4) != 4:
                raise GltfError("only TRIANGLES primitives are supported")
            if "indices" not in prim:
                raise GltfError("primitives without indices are not supported")
            attrs = prim["attributes"]
            for key in ("POSITION", "NORMAL", "TEXCOORD_0"):
                if key not in attrs:
                    raise GltfError(f"primitive without {key}")
            idx = read_accessor(doc, blob, prim["indices"]).reshape(-1).astype(np.uint32)
            p = _as_f32(read_accessor(doc, blob, attrs["POSITION"]), False)
            n = _as_f32(read_accessor(doc, blob, attrs["NORMAL"]), False)
            ta = doc["accessors"][attrs["TEXCOORD_0"]]
            t = _as_f32(read_accessor(doc, blob, attrs["TEXCOORD_0"]), ta.get("normalized", False))
            if not (len(p) == len(n) == len(t)):
                raise GltfError("attribute counts differ")
            if len(idx) % 3:
                raise GltfError("index count is not a multiple of 3")
            geometries[prim.get("material", 0)][0].append(idx + np.uint32(nverts))
            positions.append(p)
            normals.append(n)
            uvs.append(t)
            nverts += len(p)
    cat = lambda xs, w: np.concatenate(xs).astype(np.float32) if xs else np.zeros((0, w), np.float32)
    geoms = [((np.concatenate(ix) if ix else np.zeros(0, np.uint32)).reshape(-1, 3), op) for ix, op in geometries]
    mesh_index = scene.add_mesh(Mesh(name, cat(positions, 3), cat(normals, 3), cat(uvs, 2), geoms))
    model_index = scene.add_model(mesh_index, materials)
    return mesh_index, model_index


# --------------------------------------------------------------------------
# Writer (round trips and synthetic test assets)
# --------------------------------------------------------------------------
def encode_glb(primitives, materials=(), images=()):
    """A .glb holding one mesh.

    primitives: dicts with positions (N,3), normals (N,3), uvs (N,2) float32,
        indices (M,) u32, optional material index.
    materials: dicts with optional base_colour (4), metallic, roughness,
        alpha_mode, base_texture / mr_texture / normal_texture (image index),
        nearest (bool: magFilter NEAREST for its textures).
    images: PNG bytes.
    """
    blob = bytearray()
    views, accessors = [], []

    def add_view(data):
        while len(blob) % 4:
            blob.append(0)
        views.append({"buffer": 0, "byteOffset": len(blob), "byteLength": len(data)})
        blob.extend(data)
        return len(views) - 1

    def add_accessor(arr, ctype, typ):
        v = add_view(np.ascontiguousarray(arr).tobytes())
        acc = {"bufferView": v, "componentType": ctype, "count": int(arr.shape[0]), "type": typ}
        if typ == "VEC3" and ctype == 5126 and len(arr):
            acc["min"] = [float(x) for x in arr.min(0)]
            acc["max"] = [float(x) for x in arr.max(0)]
        accessors.append(acc)
        return len(accessors) - 1

    prims = []
    for p in primitives:
        d = {"attributes": {"POSITION": add_accessor(np.asarray(p["positions"], np.float32), 5126, "VEC3"),
                            "NORMAL": add_accessor(np.asarray(p["normals"], np.float32), 5126, "VEC3"),
                            "TEXCOORD_0": add_accessor(np.asarray(p["uvs"], np.float32), 5126, "VEC2")},
             "indices": add_accessor(np.asarray(p["indices"], np.uint32).reshape(-1), 5125, "SCALAR")}
        if "material" in p:
            d["material"] = int(p["material"])
        prims.append(d)
    img_docs = [{"bufferView": add_view(png), "mimeType": "image/png"} for png in images]
    samplers = [{"magFilter": NEAREST, "minFilter": NEAREST}, {"magFilter": 9729, "minFilter": 9729}]
    textures, mats = [], []

    def tex(img, nearest):
        textures.append({"source": int(img), "sampler": 0 if nearest else 1})
        return {"index": len(textures) - 1}

    for m in materials:
        pbr = {"baseColorFactor": [float(x) for x in m.get("base_colour", [1, 1, 1, 1])],
               "metallicFactor": float(m.get("metallic", 1.0)), "roughnessFactor": float(m.get("roughness", 1.0))}
        if "base_texture" in m:
            pbr["baseColorTexture"] = tex(m["base_texture"], m.get("nearest", False))
        if "mr_texture" in m:
            pbr["metallicRoughnessTexture"] = tex(m["mr_texture"], m.get("nearest", False))
        md = {"pbrMetallicRoughness": pbr, "alphaMode": m.get("alpha_mode", "OPAQUE")}
        if "normal_texture" in m:
            md["normalTexture"] = tex(m["normal_texture"], m.get("nearest", False))
        mats.append(md)
    doc = {"asset": {"version": "2.0"}, "buffers": [{"byteLength": len(blob)}], "bufferViews": views,
           "accessors": accessors, "meshes": [{"primitives": prims}], "scenes": [{"nodes": [0]}],
           "nodes": [{"mesh": 0}], "scene": 0}
    if mats:
        doc["materials"] = mats
    if img_docs:
        doc["images"] = img_docs
        doc["samplers"] = samplers
        doc["textures"] = textures
    js = json.dumps(doc, separators=(",", ":")).encode()
    js += b" " * ((4 - len(js) % 4) % 4)
    while len(blob) % 4:
        blob.append(0)
    total = 12 + 8 + len(js) + 8 + len(blob)
    return (struct.pack("<III", GLB_MAGIC, 2, total) + struct.pack("<II", len(js), CHUNK_JSON) + js +
            struct.pack("<II", len(blob), CHUNK_BIN) + bytes(blob))
