"""Pins the record front-end oracle (oracle/frontend.py) with closed forms on
hand-built scenes: UnpackNRCInput (NRCRecord.glsl:98-125), the scene accessors
(Scene.glsl:50-69), the texture sampler, the wire-record epilogue
(nrc_inference.comp:50-72), the path suffix sums (path_tracer.comp:346-350) and the
screen pass (screen.frag:10-31)."""
import numpy as np
import pytest

import vknrc_amd.nrc as n
from oracle import epilogue as ep
from oracle import frontend as fe

NT = n.NO_TEXTURE


def one_triangle(v0=(0, 0, 0), v1=(1, 0, 0), v2=(0, 1, 0), transform=None, material=None, textures=(),
                 texcoords=((0, 0), (1, 0), (0, 1))):
    m = np.zeros(1, n.MATERIAL)
    m["diffuse"] = (0.25, 0.5, 0.75)
    m["specular"] = (0.1, 0.2, 0.3)
    m["roughness"] = 0.6
    m["diffuse_texture_id"] = NT
    m["specular_texture_id"] = NT
    if material:
        for k, val in material.items():
            m[k] = val
    T = np.zeros((1, 3, 4))
    T[0, :, :3] = np.eye(3)
    if transform is not None:
        T[0] = transform
    return {"vertices": np.array([v0, v1, v2], np.float32), "vertex_indices": np.array([0, 1, 2], np.uint32),
            "texcoords": np.array(texcoords, np.float32), "texcoord_indices": np.array([0, 1, 2], np.uint32),
            "materials": m, "material_ids": np.array([0], np.uint32), "transforms": T.astype(np.float32),
            "textures": list(textures)}


def packed(b1, b2, d=(0, 0), flip=False, prim=0, inst=0):
    p = np.zeros(1, n.PACKED_INPUT)
    p["primitive_id"] = prim
    p["flip_bit_instance_id"] = inst | (int(flip) << 31)
    p["barycentric_2x16u"] = int(b1) | (int(b2) << 16)
    p["scattered_dir_2x16u"] = int(d[0]) | (int(d[1]) << 16)
    return p


def test_unorm2x16():
    assert np.array_equal(fe.unorm2x16(0), [0.0, 0.0])
    assert np.array_equal(fe.unorm2x16(0xFFFFFFFF), [1.0, 1.0])
    assert np.allclose(fe.unorm2x16(0x8000 | (1 << 16)), [32768 / 65535, 1 / 65535], rtol=0, atol=0)


def test_corners_normal_and_flip():
    s = one_triangle()
    q = fe.unpack(s, packed(0, 0))[0]
    assert np.allclose(q[0:3], [0, 0, 0]) and np.allclose(q[5:7], [0.5, 0.0])        # +z: v = acos(1) / pi = 0
    assert np.allclose(fe.unpack(s, packed(65535, 0))[0][0:3], [1, 0, 0])
    assert np.allclose(fe.unpack(s, packed(0, 65535))[0][0:3], [0, 1, 0])
    assert np.allclose(fe.unpack(s, packed(0, 0, flip=True))[0][5:7], [0.5, 1.0])   # -z
    # the centroid from unorm16 thirds
    c = fe.unpack(s, packed(21845, 21845))[0]
    assert np.allclose(c[0:3], [21845 / 65535, 21845 / 65535, 0], atol=1e-15)
    # a triangle in the xz-plane: cross((1,0,0), (0,0,1)) = (0,-1,0): u = 0.5 - 1/4, v = 1/2
    s2 = one_triangle(v2=(0, 0, 1))
    assert np.allclose(fe.unpack(s2, packed(0, 0))[0][5:7], [0.25, 0.5])
    # the scattered direction passes through as unorm16
    assert np.allclose(fe.unpack(s, packed(0, 0, d=(65535, 13107)))[0][3:5], [1.0, 0.2])


def test_instance_transform_rows():
    R = np.array([[0, -1, 0], [1, 0, 0], [0, 0, 1]], np.float64)       # 90 degrees about z
    T = np.concatenate([2 * R, np.array([[5.0], [-1.0], [0.5]])], 1)   # scale 2, then translate
    s = one_triangle(transform=T)
    q = fe.unpack(s, packed(65535, 0))[0]                              # vertex 1 = (1, 0, 0)
    assert np.allclose(q[0:3], [5.0, 1.0, 0.5])
    assert np.allclose(q[5:7], [0.5, 0.0])                             # the normal stays +z


def test_material_constants_and_roughness():
    q = fe.unpack(one_triangle(), packed(100, 200))[0]
    assert np.allclose(q[7], 0.6) and np.allclose(q[8:11], [0.25, 0.5, 0.75]) and np.allclose(q[11:14], [0.1, 0.2, 0.3])


def rgba(r, g, b, a=255):
    return np.uint32(r | (g << 8) | (b << 16) | (a << 24))


def test_srgb_decode_endpoints():
    assert fe.srgb_decode(0) == 0.0 and fe.srgb_decode(255) == 1.0
    assert np.isclose(fe.srgb_decode(10), 10 / 255 / 12.92)                 # linear segment below 0.04045
    assert np.isclose(fe.srgb_decode(128), ((128 / 255 + 0.055) / 1.055) ** 2.4)


def test_texture_bilinear_repeat():
    img = np.array([[rgba(255, 0, 0), rgba(0, 255, 0)], [rgba(0, 0, 255), rgba(255, 255, 255)]], np.uint32)
    # texel centres return the texel
    assert np.allclose(fe.sample_texture(img, np.array([0.25]), np.array([0.25]))[0], [1, 0, 0])
    assert np.allclose(fe.sample_texture(img, np.array([0.75]), np.array([0.75]))[0], [1, 1, 1])
    # half-way between the two top texels
    assert np.allclose(fe.sample_texture(img, np.array([0.5]), np.array([0.25]))[0], [0.5, 0.5, 0])
    # repeat: u + 1 samples the same point; u = 0 blends the last and the first column
    assert np.allclose(fe.sample_texture(img, np.array([1.25]), np.array([-0.75]))[0], [1, 0, 0])
    assert np.allclose(fe.sample_texture(img, np.array([0.0]), np.array([0.25]))[0], [0.5, 0.5, 0])


def test_textured_material_uses_interpolated_texcoords():
    img = np.full((4, 4), rgba(255, 128, 0), np.uint32)
    s = one_triangle(material={"diffuse_texture_id": 0}, textures=[img])
    q = fe.unpack(s, packed(30000, 20000))[0]
    assert np.allclose(q[8:11], [1.0, fe.srgb_decode(128), 0.0])          # a constant texture
    assert np.allclose(q[11:14], [0.1, 0.2, 0.3])                          # specular stays constant
    # an out-of-range texture id falls back to the constant colour
    s2 = one_triangle(material={"diffuse_texture_id": 7}, textures=[img])
    assert np.allclose(fe.unpack(s2, packed(1, 1))[0][8:11], [0.25, 0.5, 0.75])
    # a gradient texture along u: the texcoord is the barycentric blend of the corners'
    grad = np.array([[rgba(0, 0, 0), rgba(255, 0, 0)]] * 2, np.uint32)
    s3 = one_triangle(material={"diffuse_texture_id": 0}, textures=[grad], texcoords=((0.25, 0.25), (0.75, 0.25), (0.25, 0.75)))
    assert np.allclose(fe.unpack(s3, packed(0, 0))[0][8], 0.0)           # u = 0.25: texel 0
    assert np.allclose(fe.unpack(s3, packed(65535, 0))[0][8], 1.0)       # u = 0.75: texel 1
    assert np.allclose(fe.unpack(s3, packed(32768, 0))[0][8], 32768 / 65535, atol=1e-12)


def test_composite_records_closed_form():
    rec = np.zeros(3, n.EVAL_RECORD)
    rec["dst"] = [ep.dst_screen(1, 0), ep.dst_train(2, 1, 2), 0xFFFFFFFF]
    rgb = np.array([[1, 2, 3], [0.5, 0.25, 2], [9, 9, 9]], np.float32)
    bf = np.zeros((1, 2, 4), np.float32)
    bf[0, 1] = (1, 1, 1, 2)
    fgb = np.zeros((1, 2, 2), np.float32)
    fgb[0, 1] = (3, 4)
    batches = np.zeros(4 * 4, n.WIRE_TRAIN_RECORD)
    batches["bias"] = 1.0
    batches["factor"] = 2.0
    fe.composite_records(rgb, rec, 3, bf, fgb, 2, 1, 2, batches, 4)
    assert np.array_equal(bf[0, 1], [1 + 2 * 1, 1 + 3 * 2, 1 + 4 * 3, 0])
    assert np.array_equal(bf[0, 0], [0, 0, 0, 0])
    want = np.array([1 + 2 * 0.5, 1 + 2 * 0.25, 1 + 2 * 2], np.float32)
    for j in range(16):
        b, i = divmod(j, 4)
        assert np.array_equal(batches["bias"][j], want if (b == 2 and i in (1, 2)) else [1, 1, 1])


def test_path_suffix_closed_form():
    l = np.array([[1, 0, 0], [0, 2, 0], [0, 0, 4]], np.float32)
    c = np.array([[0.5, 0.5, 0.5], [0.25, 1, 2], [1, 1, 1]], np.float32)
    bias, factor = fe.path_suffix(l, c, 3)
    # bias_2 = l2, bias_1 = l1 + c1 l2, bias_0 = l0 + c0 (l1 + c1 l2); factor_i = prod c_i..c_2
    assert np.array_equal(bias[2], [0, 0, 4]) and np.array_equal(bias[1], [0, 2, 8])
    assert np.array_equal(bias[0], [1, 1, 4]) and np.array_equal(factor[0], [0.125, 0.5, 1])
    b1, f1 = fe.path_suffix(l, c, 1)
    assert np.array_equal(b1[0], l[0]) and np.array_equal(f1[0], c[0])


def test_screen_tonemap_closed_form_and_accumulation():
    # the white point maps to 1 exactly; black to 0; the curve is increasing
    assert np.isclose(fe.filmic(3.2), 1.0)
    h = np.linspace(0.01, 10, 200)
    assert np.all(np.diff(fe.filmic(h)) > 0)
    a = 1.425 * 1.0 + 0.05
    f1 = (a + 0.004) / (a + 0.55 + 0.0491) - 0.0821
    aw = 1.425 * 3.2 + 0.05
    fw = (3.2 * aw + 0.004) / (3.2 * (aw + 0.55) + 0.0491) - 0.0821
    out, _ = fe.screen(np.array([[[1.0, 3.2, 0.0, 7]]]))
    assert np.allclose(out[0, 0], [(f1 / fw) ** (1 / 2.2), 1.0, 0.0, 1.0])
    # accumulating frames x_1..x_N with counts 0..N-1 gives their mean
    rng = np.random.default_rng(3)
    frames = rng.uniform(0, 2, (6, 2, 2, 4))
    acc = None
    for k, fr in enumerate(frames):
        _, acc = fe.screen(fr, acc, accumulate=True, count=k)
    assert np.allclose(acc[..., :3], frames[..., :3].mean(0)) and np.all(acc[..., 3] == 0)
