"""CPU-side checks of the C ABI: the library loads, exports exactly what include/restir_abi.h
declares, the wire formats have the reference's sizes and offsets, argument errors come back as
error codes, and the host-side builders equal the oracle bit for bit. No kernel is launched."""
import ctypes as C
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

import oracle
from restir_amd import abi, scenes

ROOT = Path(__file__).resolve().parent.parent
HEADER = ROOT / "include" / "restir_abi.h"


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return set(re.findall(r"^\s*(?:const\s+)?(?:int|void|char)\s*\*?\s*(restir_\w+)\s*\(", text, flags=re.M))


def test_header_and_binding_declare_the_same_entry_points():
    assert declared_functions() == set(abi.SIGNATURES)


def test_header_and_binding_flags_agree():
    text = HEADER.read_text()

    def value(name):
        m = re.search(rf"#define {name} \(?(?:1 << (\d+)|(\d+)u?)\)?", text)
        assert m, name
        return 1 << int(m.group(1)) if m.group(1) else int(m.group(2))

    for name, v in (("RESTIR_VISIBILITY_REUSE_FLAG", abi.VISIBILITY_REUSE_FLAG),
                    ("RESTIR_TEMPORAL_REUSE_FLAG", abi.TEMPORAL_REUSE_FLAG),
                    ("RESTIR_STANDARD_WEIGHTS_FLAG", abi.STANDARD_WEIGHTS_FLAG),
                    ("RESTIR_INPUT_VISIBLE_FLAG", abi.INPUT_VISIBLE_FLAG),
                    ("RESTIR_SHADE_VISIBILITY_FLAG", abi.SHADE_VISIBILITY_FLAG)):
        assert value(name) == v, name
    uniform_flags = [abi.VISIBILITY_REUSE_FLAG, abi.TEMPORAL_REUSE_FLAG, abi.STANDARD_WEIGHTS_FLAG,
                     abi.INPUT_VISIBLE_FLAG]
    assert len(set(uniform_flags)) == len(uniform_flags)


def test_library_exports_every_declared_symbol(product_lib):
    out = subprocess.run(["nm", "-D", "--defined-only", str(abi.LIB_PATH)], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r"\bT (restir_\w+)", out))
    missing = declared_functions() - exported
    assert not missing, missing
    for name in declared_functions():
        assert hasattr(product_lib, name)


def test_header_static_asserts_hold(tmp_path):
    src = tmp_path / "t.cpp"
    src.write_text('#include "restir_abi.h"\nint main() { return 0; }\n')
    subprocess.run(["g++", "-std=c++17", "-fsyntax-only", f"-I{HEADER.parent}", str(src)], check=True)
    src_c = tmp_path / "t.c"
    src_c.write_text('#include "restir_abi.h"\nint main(void) { return 0; }\n')
    subprocess.run(["gcc", "-std=c11", "-fsyntax-only", f"-I{HEADER.parent}", str(src_c)], check=True)


def test_wire_format_offsets():
    R = abi.RESERVOIR_DTYPE
    assert R.fields["lightIndex"][1] == 32 and R.fields["pHat"][1] == 36
    assert R.fields["sumWeights"][1] == 40 and R.fields["w"][1] == 44 and R.fields["M"][1] == 48
    N = abi.NODE_DTYPE
    assert N.fields["left"][1] == 64 and N.fields["right"][1] == 68 and N.itemsize == 80
    U = abi.Uniforms
    assert U.cameraPos.offset == 64 and U.screenSize.offset == 80 and U.frame.offset == 88
    assert U.initialLightSampleCount.offset == 92 and U.temporalSampleCountMultiplier.offset == 96
    assert U.spatialPosThreshold.offset == 100 and U.spatialNormalThreshold.offset == 104
    assert U.spatialNeighbors.offset == 108 and U.spatialRadius.offset == 112 and U.flags.offset == 116


def test_argument_errors_without_device(product_lib):
    assert product_lib.restir_create(None, 0) == -1
    tri = np.zeros((1, 9), dtype=np.float32)
    nodes = np.zeros(1, dtype=abi.NODE_DTYPE)
    tris = np.zeros(1, dtype=abi.TRIANGLE_DTYPE)
    assert product_lib.restir_build_aabb_tree(tri.ctypes.data, 1, nodes.ctypes.data,
                                              tris.ctypes.data, None) == -1  # needs >= 2 triangles
    assert product_lib.restir_set_lights(None, None, 0, None, 0, None, 0) == -1
    assert product_lib.restir_set_bvh(None, None, 0, None, 0) == -1
    assert product_lib.restir_camera_projection_view(None, 1.0, None) == -1
    assert product_lib.restir_version() == abi.ABI_VERSION
    assert f"#define RESTIR_ABI_VERSION {abi.ABI_VERSION}" in HEADER.read_text()


def test_create_reports_missing_device(product_lib):
    import torch
    if torch.cuda.device_count() > 0:
        pytest.skip("a HIP device is present")
    h = C.c_void_p()
    assert product_lib.restir_create(C.byref(h), 0) == -7


@pytest.mark.parametrize("which", ["cornell", "terrain", "random"])
def test_bvh_builder_matches_oracle(which):
    if which == "cornell":
        tri = scenes.cornell_box(1).tri_xyz
    elif which == "terrain":
        tri = scenes.terrain(2).tri_xyz
    else:
        tri = np.random.default_rng(3).normal(size=(5000, 9)).astype(np.float32)
    n1, t1, d1 = abi.build_aabb_tree(tri)
    n2, t2, d2 = oracle.build_aabb_tree(tri, abi.NODE_DTYPE, abi.TRIANGLE_DTYPE)
    assert n1.tobytes() == n2.tobytes() and t1.tobytes() == t2.tobytes() and d1 == d2


@pytest.mark.parametrize("n", [1, 64, 10000])
def test_alias_builder_matches_oracle(n):
    rng = np.random.default_rng(n)
    pts = scenes.make_point_lights(rng.uniform(-1, 1, (n, 3)),
                                   np.exp(rng.uniform(-4.6, 2.3, (n, 3))))
    a = abi.build_alias_table(pts, None)
    b = oracle.build_alias_table(pts, None, abi.ALIAS_DTYPE)
    assert a.tobytes() == b.tobytes()


def test_alias_builder_triangle_lights_match_oracle():
    rng = np.random.default_rng(9)
    tl = scenes.make_tri_lights(rng.uniform(-1, 1, (300, 9)).astype(np.float32),
                                np.exp(rng.uniform(-2, 2, (300, 3))))
    a = abi.build_alias_table(None, tl)
    b = oracle.build_alias_table(None, tl, abi.ALIAS_DTYPE)
    assert a.tobytes() == b.tobytes()


def test_projection_view_matches_oracle():
    sc = scenes.cornell_box(1)
    for aspect in (1.0, 16 / 9):
        a = abi.projection_view(sc.camera, aspect)
        b = oracle.projection_view(abi.camera_struct(sc.camera), aspect)
        np.testing.assert_array_equal(a, b)
        # the camera position projects to w = 0; the look-at point to the screen centre
        pv = a.reshape(4, 4).T
        la = np.array([*sc.camera["lookAt"], 1.0], dtype=np.float32)
        clip = pv @ la
        assert abs(clip[0] / clip[3]) < 1e-6 and abs(clip[1] / clip[3]) < 1e-6


def test_light_buffer_packing():
    pts = scenes.cornell_box(1).point_lights
    buf = abi.pack_buffer(pts, abi.POINT_LIGHT_DTYPE)
    assert len(buf) == 16 + 64 * 32
    assert np.frombuffer(buf[:4].tobytes(), dtype=np.int32)[0] == 64
    assert buf[16:].tobytes() == pts.tobytes()
