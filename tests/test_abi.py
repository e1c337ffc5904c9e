"""The C ABI: struct layouts of include/rtg.h and the exported symbols of
librtg.so (no compute calls; runs without a GPU)."""
import ctypes as C
import os
import re
import subprocess

import pytest

from conftest import PKG, REPO
from rtg import abi, lib

HEADER = os.path.join(REPO, "include", "rtg.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(rtg_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_binding_list():
    assert declared_functions() == sorted(lib.EXPORTS)


def test_library_exports_every_declared_symbol():
    if not os.path.exists(lib.LIB_PATH):
        subprocess.run(["make", "-C", os.path.join(PKG, "csrc")], check=True, capture_output=True)
    L = C.CDLL(lib.LIB_PATH)
    missing = [f for f in declared_functions() if not hasattr(L, f)]
    assert not missing, missing
    assert lib.load().rtg_abi_version() == 1


def test_struct_layouts_match_the_reference_abi():
    # vk::AccelerationStructureInstanceKHR (src/gpu_structs.rs:17-25)
    assert C.sizeof(abi.Instance) == 64
    assert abi.Instance.custom_index_and_mask.offset == 48
    assert abi.Instance.sbt_offset_and_flags.offset == 52
    assert abi.Instance.acceleration_structure_reference.offset == 56
    # Uniforms (shared-structs/src/lib.rs:7-19): Vec3A sun_dir at 128, 176 B
    assert abi.Uniforms.sun_dir.offset == 128
    assert abi.Uniforms.sun_radius.offset == 144
    assert abi.Uniforms.blue_noise_texture_index.offset == 148
    assert abi.Uniforms.ggx_lut_texture_index.offset == 152
    assert abi.Uniforms.frame_index.offset == 156
    assert abi.Uniforms.show_heatmap.offset == 160
    assert C.sizeof(abi.Uniforms) == 176
    # ModelInfo / GeometryInfo (shared-structs/src/lib.rs:21-47)
    assert C.sizeof(abi.ModelInfo) == 32
    assert C.sizeof(abi.GeometryInfo) == 24
    assert abi.GeometryInfo.normal_map_image_index.offset == 16


def test_packed24_8():
    # vk::Packed24_8::new(low24, high8) = low24 | high8 << 24
    assert abi.pack24_8(5, 0xFF) == 0xFF000005
    assert abi.pack24_8(0xFFFFFF, 0) == 0x00FFFFFF
    assert abi.pack24_8(0x1000001, 1) == 0x01000001  # low 24 bits only


def test_context_without_gpu_fails_loudly():
    import torch
    if torch.cuda.is_available():
        pytest.skip("GPU present")
    with pytest.raises(lib.RtgError):
        lib.Context(0)


def test_python_constants_equal_the_header_defines():
    text = open(HEADER).read()
    defines = {m.group(1): int(m.group(2), 0) for m in
               re.finditer(r"#define\s+(RTG_[A-Z0-9_]+)\s+(0x[0-9a-fA-F]+|\d+)u?\b", text)}
    checked = 0
    for name, value in defines.items():
        if hasattr(abi, name):
            assert getattr(abi, name) == value, name
            checked += 1
    assert checked >= 20
    kinds = sorted((v, k) for k, v in defines.items() if k.startswith("RTG_KERNEL_") and k != "RTG_KERNEL_COUNT")
    assert len(abi.RTG_KERNEL_NAMES) == defines["RTG_KERNEL_COUNT"] == len(kinds)
    for v, k in kinds:
        assert getattr(abi, k) == v


def test_tuning_layout():
    # rtg_tuning: seven u32 settings + one reserved word
    assert C.sizeof(abi.Tuning) == 32
    assert abi.Tuning.path_order.offset == 16
    assert abi.Tuning.bounce_blocks.offset == 20
    assert abi.Tuning.stack_lds_entries.offset == 24
    t = abi.Tuning.make()
    assert t.as_dict() == {"blocks_per_cu": 4, "feed_chunk": 64, "refill_closest": 64, "refill_occlusion": 64,
                           "path_order": 1, "bounce_blocks": 192, "stack_lds_entries": 16}


MIRRORS = {
    "rtg_instance": abi.Instance, "rtg_uniforms": abi.Uniforms, "rtg_model_info": abi.ModelInfo,
    "rtg_geometry_info": abi.GeometryInfo, "rtg_texture": abi.Texture, "rtg_geometry": abi.Geometry,
    "rtg_build_sizes": abi.BuildSizes, "rtg_as_info": abi.AsInfo, "rtg_ray": abi.Ray, "rtg_hit": abi.Hit,
    "rtg_scene_bindings": abi.SceneBindings, "rtg_trace_opts": abi.TraceOpts, "rtg_image_job": abi.ImageJob,
    "rtg_trace_counters": abi.TraceCounters, "rtg_tuning": abi.Tuning, "rtg_server_config": abi.ServerConfig,
    "rtg_request": abi.Request, "rtg_completion": abi.Completion, "rtg_server_stats": abi.ServerStats,
}


def c_fields(cname):
    """Field names of a struct in rtg.h, in declaration order."""
    text = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    body = re.search(r"typedef struct %s \{(.*?)\} %s;" % (cname, cname), text, re.S).group(1)
    return [re.search(r"(\w+)\s*(\[[^\]]*\])*\s*$", d.strip()).group(1) for d in body.split(";") if d.strip()]


def test_ctypes_mirrors_match_the_c_compiler(tmp_path):
    """Every ctypes mirror has the field names, offsets and size that gcc
    gives the struct in include/rtg.h."""
    lines = ["#include <stddef.h>", "#include <stdio.h>", '#include "rtg.h"', "int main(void) {"]
    for cname in MIRRORS:
        lines.append('printf("%s %%zu\\n", sizeof(%s));' % (cname, cname))
        for f in c_fields(cname):
            lines.append('printf("%s.%s %%zu\\n", offsetof(%s, %s));' % (cname, f, cname, f))
    lines += ["return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(REPO, "include"), "-o", str(exe), str(src)], check=True)
    got = dict(line.split() for line in subprocess.run([str(exe)], check=True, capture_output=True,
                                                        text=True).stdout.splitlines())
    for cname, py in MIRRORS.items():
        names = [n for n, _ in py._fields_]
        assert names == c_fields(cname), cname
        assert C.sizeof(py) == int(got[cname]), cname
        for n in names:
            assert getattr(py, n).offset == int(got[f"{cname}.{n}"]), (cname, n)
