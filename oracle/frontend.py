"""CPU restatement of the record front-end, the wire-record epilogue, the path
suffix sums and the screen pass (numpy).

TEST INFRASTRUCTURE ONLY: imported by tests/ and never by the product package.

  unpack               UnpackNRCInput, shader/src/NRCRecord.glsl:98-125, with the
                       scene accessors of Scene.glsl:50-69 (float64)
  sample_texture       texture() of an RGBA8 sRGB image with linear filtering and
                       repeat addressing (VkScene.cpp:193, NRCRenderGraph.cpp:132)
  composite_records    nrc_inference.comp:50-72 over 20-byte eval records and
                       40-byte train records (float32, each product and sum rounded)
  path_suffix          path_tracer.comp:346-350 (float32, rounded as the kernel)
  screen               screen.frag:10-31 (float64)

A scene is a dict with vertices [V][3], vertex_indices [3P], texcoords [T][2],
texcoord_indices [3P], materials (a MATERIAL structured array), material_ids [P],
transforms [I][3][4] and textures (a list of [h][w] uint32 RGBA8 images).
"""
from __future__ import annotations

import numpy as np

NO_TEXTURE = 0xFFFFFFFF


def unorm2x16(v) -> np.ndarray:
    """unpackUnorm2x16: (low half, high half) / 65535."""
    v = np.asarray(v, np.uint64)
    return np.stack([(v & 0xFFFF) / 65535.0, (v >> 16) / 65535.0], -1)


def sph_encode(d) -> np.ndarray:
    """NRCSphEncode (NRCRecord.glsl:47-49)."""
    d = np.asarray(d, np.float64)
    x, y, z = d[..., 0], d[..., 1], d[..., 2]
    u = np.where((x == 0) & (y == 0), 0.5, 0.5 + np.arctan2(y, x) / (2 * np.pi))
    return np.stack([u, np.arccos(np.clip(z, -1, 1)) / np.pi], -1)


def srgb_decode(c8) -> np.ndarray:
    c = np.asarray(c8, np.float64) / 255.0
    return np.where(c <= 0.04045, c / 12.92, ((c + 0.055) / 1.055) ** 2.4)


def sample_texture(img, u, v) -> np.ndarray:
    """Bilinear RGB of an RGBA8 sRGB image [h][w] at (u, v), repeat addressing; texels are
    decoded before the blend."""
    img = np.asarray(img, np.uint32)
    h, w = img.shape
    x = np.asarray(u, np.float64) * w - 0.5
    y = np.asarray(v, np.float64) * h - 0.5
    x0, y0 = np.floor(x), np.floor(y)
    ax, ay = x - x0, y - y0
    x0, y0 = x0.astype(np.int64), y0.astype(np.int64)
    xs = (np.mod(x0, w), np.mod(x0 + 1, w))
    ys = (np.mod(y0, h), np.mod(y0 + 1, h))

    def texel(yy, xx):
        p = img[yy, xx].astype(np.uint64)
        return np.stack([srgb_decode((p >> (8 * c)) & 255) for c in range(3)], -1)

    top = texel(ys[0], xs[0]) * (1 - ax)[..., None] + texel(ys[0], xs[1]) * ax[..., None]
    bot = texel(ys[1], xs[0]) * (1 - ax)[..., None] + texel(ys[1], xs[1]) * ax[..., None]
    return top * (1 - ay)[..., None] + bot * ay[..., None]


def _color(scene, const, tex, u, v):
    out = np.array(const, np.float64)
    textures = scene.get("textures", [])
    for t in np.unique(tex):
        if t == NO_TEXTURE or t >= len(textures):
            continue
        m = tex == t
        out[m] = sample_texture(textures[t], u[m], v[m])
    return out


def unpack(scene: dict, packed: np.ndarray) -> np.ndarray:
    """[n][14] float64 queries of PACKED_INPUT rows (UnpackNRCInput, NRCRecord.glsl:98-125)."""
    prim = packed["primitive_id"].astype(np.int64)
    inst = (packed["flip_bit_instance_id"] & 0x7FFFFFFF).astype(np.int64)
    flip = (packed["flip_bit_instance_id"] >> 31) != 0
    V = np.asarray(scene["vertices"], np.float64)
    VI = np.asarray(scene["vertex_indices"], np.int64)
    T = np.asarray(scene["transforms"], np.float64)
    verts = []
    for k in range(3):
        p4 = np.concatenate([V[VI[prim * 3 + k]], np.ones((len(prim), 1))], 1)
        verts.append(np.einsum("nj,nij->ni", p4, T[inst]))      # row i of the 3x4 transform . (x, y, z, 1)
    v0, v1, v2 = verts
    n = np.cross(v1 - v0, v2 - v0)
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    n[flip] *= -1
    b12 = unorm2x16(packed["barycentric_2x16u"])
    b = np.concatenate([1.0 - b12[:, :1] - b12[:, 1:], b12], 1)
    q = np.zeros((len(prim), 14))
    q[:, 0:3] = v0 * b[:, :1] + v1 * b[:, 1:2] + v2 * b[:, 2:3]
    q[:, 3:5] = unorm2x16(packed["scattered_dir_2x16u"])
    q[:, 5:7] = sph_encode(n)
    mats = scene["materials"][np.asarray(scene["material_ids"], np.int64)[prim]]
    q[:, 7] = mats["roughness"]
    TC = np.asarray(scene["texcoords"], np.float64)
    TI = np.asarray(scene["texcoord_indices"], np.int64)
    uv = sum(TC[TI[prim * 3 + k]] * b[:, k:k + 1] for k in range(3))
    q[:, 8:11] = _color(scene, mats["diffuse"], mats["diffuse_texture_id"], uv[:, 0], uv[:, 1])
    q[:, 11:14] = _color(scene, mats["specular"], mats["specular_texture_id"], uv[:, 0], uv[:, 1])
    return q


def _mad(b, f, p):
    """b + f * p in float32 with the product and the sum rounded separately."""
    return (np.float32(b) + np.float32(np.float32(f) * np.float32(p))).astype(np.float32)


def composite_records(rgb, records, count, bias_factor_r=None, factor_gb=None, width=0, height=0, pitch=None,
                      batches=None, batch_stride=0):
    """nrc_inference.comp:50-72 on host arrays, in place: bias_factor_r [h][pitch][4] and factor_gb
    [h][pitch][2] float32, batches a WIRE_TRAIN_RECORD array of NRC_TRAIN_BATCH_COUNT * batch_stride."""
    pitch = width if pitch is None else pitch
    rgb = np.asarray(rgb, np.float32)
    for i in range(count):
        e = int(records["dst"][i])
        if e == 0xFFFFFFFF:
            continue
        p = rgb[i]
        v = e >> 1
        if e & 1 == 0:
            x, y = v & 0x7FFF, v >> 15
            if bias_factor_r is None or x >= width or y >= height:
                continue
            b = bias_factor_r[y, x]
            f = (b[3], factor_gb[y, x, 0], factor_gb[y, x, 1])
            bias_factor_r[y, x] = [_mad(b[c], f[c], p[c]) for c in range(3)] + [0.0]
        else:
            bb, l, r = v & 3, (v >> 2) & 0x3FFF, v >> 16
            if batches is None:
                continue
            for j in range(l, min(r, batch_stride - 1) + 1):
                rec = batches[bb * batch_stride + j]
                rec["bias"] = [_mad(rec["bias"][c], rec["factor"][c], p[c]) for c in range(3)]
                batches[bb * batch_stride + j] = rec


def path_suffix(light, color, bounce: int):
    """path_tracer.comp:346-350: for i = bounce - 2 .. 0, light[i] += color[i] * light[i + 1] and
    color[i] *= color[i + 1], in float32 as the kernel rounds them.  Returns (bias, factor) of the
    bounce vertices."""
    L = np.array(light[:bounce], np.float32)
    Cc = np.array(color[:bounce], np.float32)
    for i in range(bounce - 2, -1, -1):
        L[i] = _mad(L[i], Cc[i], L[i + 1])
        Cc[i] = (Cc[i] * Cc[i + 1]).astype(np.float32)
    return L, Cc


def filmic(h, white: float = 3.2):
    """ToneMapFilmic_Hejl2015 (screen.frag:10-15) per channel, over the curve's value at the white point."""
    def f(x):
        a = 1.425 * x + 0.05
        return (x * a + 0.004) / (x * (a + 0.55) + 0.0491) - 0.0821
    return f(np.asarray(h, np.float64)) / f(white)


def screen(color, acc=None, accumulate=False, count=0):
    """screen.frag:17-30 on [h][w][4] arrays: returns (tone-mapped RGBA, new accumulation image)."""
    c = np.asarray(color, np.float64)[..., :3].copy()
    new_acc = None
    if accumulate:
        if count != 0:
            c = (c + np.asarray(acc, np.float64)[..., :3] * count) / (count + 1)
        new_acc = np.concatenate([c, np.zeros(c.shape[:-1] + (1,))], -1)
    out = np.concatenate([np.maximum(filmic(c), 0.0) ** (1 / 2.2), np.ones(c.shape[:-1] + (1,))], -1)
    return out, new_acc
