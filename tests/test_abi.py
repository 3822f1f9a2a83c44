"""CPU checks of the C-ABI library: it loads, exports every function include/nrc.h
declares, and its host-only entry points behave (no GPU calls here)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import vknrc_amd as v
from vknrc_amd import workloads as wl

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "nrc.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(nrc_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_the_boundary():
    names = declared_functions()
    for f in ("nrc_cache_create", "nrc_cache_reset", "nrc_cache_destroy", "nrc_infer", "nrc_train", "nrc_frame",
              "nrc_get_state", "nrc_set_state", "nrc_set_use_ema", "nrc_encode", "nrc_mlp_forward_f16",
              "nrc_mlp_gradient_f16"):
        assert f in names


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(v.nrc.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_defaults_are_the_reference_constants():
    d = v.default_config()
    assert d.hidden_layers == 5 and d.train_batch_size == 16384           # VkNRCState.hpp:23
    assert abs(d.learning_rate - 0.002) < 1e-9 and abs(d.epsilon - 1e-8) < 1e-15  # nrc_optimize.comp:10-11
    assert abs(d.beta1 - 0.9) < 1e-7 and abs(d.beta2 - 0.999) < 1e-7 and abs(d.ema_alpha - 0.99) < 1e-7
    assert d.loss_scale == 1.0 and not d.dynamic_loss_scale and not d.use_ema_weights
    assert d.encoding == v.ENCODING_TRI12 and d.loss == v.LOSS_RELATIVE_L2_LUMINANCE


def test_weight_count():
    lib = v.load_library()
    assert lib.nrc_weight_count(5) == 20672 == v.weight_count(5)   # VkNRCState.hpp:25
    assert lib.nrc_weight_count(4) == 16576
    assert lib.nrc_weight_count(3) == 0 and lib.nrc_weight_count(6) == 0


def test_init_weights_seeded_kaiming():
    a = v.init_weights(v.NRCConfig(seed=7))
    b = v.init_weights(v.NRCConfig(seed=7))
    c = v.init_weights(v.NRCConfig(seed=8))
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    assert a.size == 20672
    assert abs(a.std() - np.sqrt(2 / 64)) < 0.005 and abs(a.mean()) < 0.005    # VkNRCState.cpp:41


def test_init_weights_xavier_uniform_bounds():
    w = v.init_weights(v.NRCConfig(seed=3, hidden_layers=4, init=v.INIT_XAVIER_UNIFORM))
    hid, out = w[:4 * 4096], w[4 * 4096:]
    assert np.abs(hid).max() <= np.sqrt(6 / 128) and np.abs(hid).max() > 0.95 * np.sqrt(6 / 128)
    assert np.abs(out).max() <= np.sqrt(6 / 67)


@pytest.mark.parametrize("field,value", [("hidden_layers", 3), ("encoding", 7), ("loss", 9), ("train_batch_size", 0),
                                         ("train_batch_size", 70000), ("learning_rate", -1.0), ("beta2", 1.0),
                                         ("loss_scale", 0.0)])
def test_invalid_desc_is_rejected_with_a_message(field, value):
    cfg = v.NRCConfig()
    setattr(cfg, field, value)
    with pytest.raises(v.NRCError) as e:
        v.init_weights(cfg)
    assert e.value.status == 1 and str(e.value)


def test_null_arguments_return_errors_not_aborts():
    lib = v.load_library()
    assert lib.nrc_infer(None, None, None, 10, None, 0, None) == 1
    assert lib.nrc_train(None, None, None, 10, 0, 0, None) == 1
    assert lib.nrc_infer(None, None, None, 0, None, 0, None) == 1       # NULL cache is checked first
    assert lib.nrc_encode(5, None, 3, None, None) == 1
    assert lib.nrc_mlp_forward_f16(3, None, None, 1, None, None) == 1
    assert lib.nrc_last_error().decode()
    assert lib.nrc_status_string(4).decode() == "HIP error"


def test_struct_layouts_match_the_header():
    assert C.sizeof(v.nrc.OptimizerState) == 40
    assert C.sizeof(v.nrc.CacheDesc) == 64
    assert C.sizeof(v.nrc.GroupItem) == 64


def test_group_calls_validate_before_touching_a_device():
    L = v.load_library()
    item = v.nrc.GroupItem()  # NULL cache
    assert L.nrc_group_infer(C.byref(item), 1, v.PATH_FUSED_F16, None) != 0
    assert b"NULL cache" in L.nrc_last_error() or b"hip" in L.nrc_last_error().lower()
    g = C.c_void_p()
    assert L.nrc_group_frame_graph(None, 0, v.PATH_FUSED_F16, None) != 0
    assert L.nrc_graph_launch(None, None) != 0
    L.nrc_graph_destroy(None)


def test_workload_generators():
    rng = np.random.default_rng(0)
    q = wl.make_queries(1000, rng)
    assert q.dtype == np.float32 and q.shape == (1000, 14)
    assert np.all(np.abs(q[:, 0:3]) <= 1) and np.all((q[:, 3:7] >= 0) & (q[:, 3:7] <= 1))
    lights = wl.GaussianLights(5)
    r = wl.make_train_records(2000, rng, lights)
    assert r.shape == (2000, 17) and np.all(r[:, 14:] >= 0) and np.isfinite(r).all()
    img = wl.make_image(1)
    assert img.shape == (256, 256, 3) and img.min() >= 0 and img.max() <= 1
    ir = wl.make_image_records(100, rng, img)
    # bilinear sampling at texel centres returns the texel
    uv = np.array([[0.5 / 256, 0.5 / 256], [100.5 / 256, 7.5 / 256]])
    assert np.allclose(wl.sample_bilinear(img, uv), img[[0, 7], [0, 100]], atol=1e-6)
    assert ir.shape == (100, 17)


def _header_layout():
    """sizeof and offsetof of the ABI structs as a C99 compiler lays out include/nrc.h."""
    import os
    import subprocess
    import tempfile
    inc = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include")
    fields = {"nrc_cache_desc": [], "nrc_optimizer_state": [], "nrc_optimizer_entry": [], "nrc_group_item": [],
              "nrc_screen": [], "nrc_texture": [], "nrc_material": ["diffuse_texture_id", "specular_texture_id",
                                                                     "roughness"],
              "nrc_packed_input": [], "nrc_eval_record": ["packed_input"],
              "nrc_wire_train_record": ["factor", "packed_input"], "nrc_traced_path": [],
              "nrc_scene": ["d_textures", "texture_count", "d_primitives"],
              "nrc_scene_primitive": ["texcoords", "material_id"], "nrc_query": [], "nrc_train_record": []}
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "nrc.h"', "int main(void) {"]
    for t, fs in fields.items():
        lines.append(f'  printf("{t} %zu\\n", sizeof({t}));')
        for f in fs:
            lines.append(f'  printf("{t}.{f} %zu\\n", offsetof({t}, {f}));')
    lines += ["  return 0;", "}"]
    with tempfile.TemporaryDirectory() as d:
        src, exe = os.path.join(d, "layout.c"), os.path.join(d, "layout")
        open(src, "w").write("\n".join(lines) + "\n")
        subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-I", inc, src, "-o", exe], check=True,
                       capture_output=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    return {k: int(n) for k, n in (l.split() for l in out.splitlines())}


def test_python_mirrors_match_the_c_layout_of_the_header():
    h = _header_layout()
    n = v.nrc
    assert C.sizeof(n.CacheDesc) == h["nrc_cache_desc"]
    assert C.sizeof(n.OptimizerState) == h["nrc_optimizer_state"]
    assert h["nrc_optimizer_entry"] == 16
    assert C.sizeof(n.GroupItem) == h["nrc_group_item"]
    assert C.sizeof(n.Screen) == h["nrc_screen"]
    assert C.sizeof(n.Texture) == h["nrc_texture"]
    assert C.sizeof(n.Scene) == h["nrc_scene"]
    for f in ("d_textures", "texture_count", "d_primitives"):
        assert getattr(n.Scene, f).offset == h[f"nrc_scene.{f}"], f
    assert h["nrc_scene_primitive"] == 64 and h["nrc_scene_primitive.texcoords"] == 36
    assert h["nrc_scene_primitive.material_id"] == 60
    assert n.PACKED_INPUT.itemsize == h["nrc_packed_input"] == 16
    assert n.EVAL_RECORD.itemsize == h["nrc_eval_record"] == 20
    assert n.EVAL_RECORD.fields["packed_input"][1] == h["nrc_eval_record.packed_input"]
    assert n.WIRE_TRAIN_RECORD.itemsize == h["nrc_wire_train_record"] == 40
    for f in ("factor", "packed_input"):
        assert n.WIRE_TRAIN_RECORD.fields[f][1] == h[f"nrc_wire_train_record.{f}"], f
    assert n.MATERIAL.itemsize == h["nrc_material"]
    for f in ("diffuse_texture_id", "specular_texture_id", "roughness"):
        assert n.MATERIAL.fields[f][1] == h[f"nrc_material.{f}"], f
    assert n.TRACED_PATH.itemsize == h["nrc_traced_path"]
    assert h["nrc_query"] == 56 and h["nrc_train_record"] == 68


def test_scene_calls_validate_before_touching_a_device():
    L = v.load_library()
    s = v.nrc.Scene()                                   # every buffer NULL
    assert L.nrc_scene_pack_primitives(C.byref(s), 10, C.c_void_p(4096), None) != 0
    assert b"scene buffer NULL" in L.nrc_last_error()
    assert L.nrc_scene_pack_primitives(None, 10, C.c_void_p(4096), None) != 0
    assert L.nrc_scene_pack_primitives(C.byref(s), 0, None, None) == 0     # nothing to pack
    fake = [C.c_void_p(256 * (i + 1)) for i in range(7)]
    s2 = v.nrc.Scene(*[f.value for f in fake], None, 0, None)
    assert L.nrc_scene_pack_primitives(C.byref(s2), 10, None, None) != 0
    assert b"NULL buffer" in L.nrc_last_error()
    assert L.nrc_scene_pack_primitives(C.byref(s2), 10, C.c_void_p(4096 + 16), None) != 0
    assert b"64-byte aligned" in L.nrc_last_error()
    s2.d_primitives = 4096 + 4                          # the unpack needs 16-byte aligned records
    assert L.nrc_unpack_eval_records(C.byref(s2), C.c_void_p(4096), None, 10, C.c_void_p(8192), None) != 0
    assert b"16-byte aligned" in L.nrc_last_error()


def test_counts_beyond_the_query_limit_are_refused_before_touching_a_device():
    """Every call that takes a query or record count refuses counts above NRC_MAX_QUERIES with
    NRC_ERROR_CAPACITY (5) before any launch: the launches round counts up to whole tiles and blocks
    in 32 bits.  The buffers here are never dereferenced."""
    import os
    import re
    hdr = open(os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "nrc.h")).read()
    limit = int(re.search(r"#define NRC_MAX_QUERIES (0x[0-9A-Fa-f]+)u", hdr).group(1), 16)
    assert limit == v.nrc.MAX_QUERIES
    L = v.load_library()
    big = limit + 1
    p = [C.c_void_p(4096 * (i + 1)) for i in range(8)]
    assert L.nrc_composite(p[0], p[1], None, big, None, None, None, 0, None) == 5
    assert L.nrc_encode(v.ENCODING_SINCOS6, p[0], big, p[1], None) == 5
    assert L.nrc_sph_encode(p[0], big, p[1], None) == 5
    assert L.nrc_mlp_forward_f16(5, p[0], p[1], big, p[2], None) == 5
    assert L.nrc_composite_records(p[0], p[1], None, big, None, None, 0, None) == 5
    assert L.nrc_build_train_records(p[0], big, p[1], p[2], 16384, p[3], p[4], 1, None) == 5
    assert L.nrc_build_train_records(p[0], 1, p[1], p[2], 16384, p[3], p[4], big, None) == 5
    scene = v.nrc.Scene(*[4096 * (i + 1) for i in range(7)], None, 0, None)
    assert L.nrc_unpack_eval_records(C.byref(scene), p[0], None, big, p[1], None) == 5
    assert L.nrc_unpack_train_records(C.byref(scene), p[0], None, big, p[1], None) == 5
    assert b"NRC_MAX_QUERIES" in L.nrc_last_error()
