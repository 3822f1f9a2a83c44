"""Seeded synthetic workloads for the NRC hot path (queries, training records, targets).

The reference's records come from its path tracer (path_tracer.comp, out of
scope here); these generators produce records of the same layout
(UnpackedNRCInput, NRCRecord.glsl:40-45 -- see include/nrc.h) with analytic
targets, identical for the CPU oracle and the GPU paths.

  queries          position U[-1,1]^3; scattered direction and normal uniform on
                   the sphere, sphere-encoded to [0,1]^2 (NRCSphEncode); roughness
                   U(0,1); diffuse and specular albedo U(0,1)^3
  Gaussian lights  L(x) = sum_j I_j exp(-1/2 (x-mu_j)^T S_j^-1 (x-mu_j)),
                   I_j ~ LogNormal(0, 1.5), anisotropic S_j = R diag(s^2) R^T;
                   target = L(x) * (diffuse + specular) / 2
  image            256x256 RGB: squash of 32 random 2-D sinusoids with
                   frequencies U(1,32); records carry uv in position[0:2] and the
                   bilinear (clamp-to-edge) sample as target
"""
from __future__ import annotations

import numpy as np

QUERY_FLOATS = 14
TRAIN_FLOATS = 17
MASTER_SEED = 0x4E52430000000000


# the workloads whose seeds derive from MASTER_SEED
IMAGE_MLP = 0      # 256x256 image, 64 -> 4x64 -> 3, UV_SINCOS16, L2 loss
RADIANCE = 1       # one radiance cache, 2,073,600 queries + 4 x 16,384 training records a frame
RADIANCE_FP8 = 2   # the same cache data, FP8 inference over 500 frames
CACHES16 = 3       # 16 caches with uneven query and batch sizes
SERVED = 4         # Poisson render sessions


def workload_seed(workload: int, sub: int = 0) -> int:
    """The seed of stream `sub` of a workload (one of the constants above)."""
    return (MASTER_SEED + (workload << 16) + sub) & 0xFFFFFFFFFFFFFFFF


def _sphere(rng, n):
    v = rng.normal(size=(n, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    return v


def sph_encode(d: np.ndarray) -> np.ndarray:
    """NRCSphEncode (NRCRecord.glsl:47-49) in numpy."""
    x, y, z = d[:, 0], d[:, 1], d[:, 2]
    u = np.where((x == 0) & (y == 0), 0.5, 0.5 + np.arctan2(y, x) / (2 * np.pi))
    v = np.arccos(np.clip(z, -1, 1)) / np.pi
    return np.stack([u, v], 1)


def make_queries(n: int, rng: np.random.Generator) -> np.ndarray:
    q = np.empty((n, QUERY_FLOATS), np.float32)
    q[:, 0:3] = rng.uniform(-1.0, 1.0, (n, 3))
    q[:, 3:5] = sph_encode(_sphere(rng, n))
    q[:, 5:7] = sph_encode(_sphere(rng, n))
    q[:, 7] = rng.uniform(0.0, 1.0, n)
    q[:, 8:14] = rng.uniform(0.0, 1.0, (n, 6))
    return q


class GaussianLights:
    def __init__(self, seed: int, count: int = 16):
        rng = np.random.default_rng(seed)
        self.mu = rng.uniform(-1.0, 1.0, (count, 3))
        self.intensity = rng.lognormal(0.0, 1.5, count)
        inv = []
        for _ in range(count):
            a = rng.normal(size=(3, 3))
            qm, _ = np.linalg.qr(a)
            s = rng.uniform(0.15, 0.6, 3)
            inv.append(qm @ np.diag(1.0 / (s * s)) @ qm.T)
        self.inv_cov = np.array(inv)

    def radiance(self, pos: np.ndarray, chunk: int = 1 << 18) -> np.ndarray:
        out = np.empty(pos.shape[0], np.float64)
        for i in range(0, pos.shape[0], chunk):
            p = pos[i:i + chunk].astype(np.float64)
            d = p[:, None, :] - self.mu[None, :, :]
            q = np.einsum("njk,jkl,njl->nj", d, self.inv_cov, d)
            out[i:i + chunk] = (self.intensity[None, :] * np.exp(-0.5 * q)).sum(1)
        return out

    def targets(self, q: np.ndarray) -> np.ndarray:
        L = self.radiance(q[:, 0:3])
        return (L[:, None] * 0.5 * (q[:, 8:11] + q[:, 11:14])).astype(np.float32)


def make_train_records(n: int, rng: np.random.Generator, lights: GaussianLights) -> np.ndarray:
    q = make_queries(n, rng)
    r = np.empty((n, TRAIN_FLOATS), np.float32)
    r[:, :QUERY_FLOATS] = q
    r[:, QUERY_FLOATS:] = lights.targets(q)
    return r


def make_image(seed: int, size: int = 256, waves: int = 32) -> np.ndarray:
    rng = np.random.default_rng(seed)
    fx = rng.uniform(1, 32, (3, waves))
    fy = rng.uniform(1, 32, (3, waves))
    amp = rng.uniform(0, 1, (3, waves))
    ph = rng.uniform(0, 2 * np.pi, (3, waves))
    t = (np.arange(size) + 0.5) / size
    u, v = np.meshgrid(t, t, indexing="xy")
    img = np.empty((size, size, 3), np.float64)
    for c in range(3):
        s = np.zeros_like(u)
        for k in range(waves):
            s += amp[c, k] * np.sin(2 * np.pi * (fx[c, k] * u + fy[c, k] * v) + ph[c, k])
        img[:, :, c] = 0.5 + 0.5 * s / amp[c].sum()
    return img.astype(np.float32)


def sample_bilinear(img: np.ndarray, uv: np.ndarray) -> np.ndarray:
    h, w, _ = img.shape
    x = np.clip(uv[:, 0] * w - 0.5, 0, w - 1)
    y = np.clip(uv[:, 1] * h - 0.5, 0, h - 1)
    x0 = np.floor(x).astype(np.int64)
    y0 = np.floor(y).astype(np.int64)
    x1 = np.minimum(x0 + 1, w - 1)
    y1 = np.minimum(y0 + 1, h - 1)
    fx = (x - x0)[:, None]
    fy = (y - y0)[:, None]
    top = img[y0, x0] * (1 - fx) + img[y0, x1] * fx
    bot = img[y1, x0] * (1 - fx) + img[y1, x1] * fx
    return (top * (1 - fy) + bot * fy).astype(np.float32)


def make_image_records(n: int, rng: np.random.Generator, img: np.ndarray) -> np.ndarray:
    r = np.zeros((n, TRAIN_FLOATS), np.float32)
    uv = rng.uniform(0, 1, (n, 2)).astype(np.float32)
    r[:, 0:2] = uv
    r[:, QUERY_FLOATS:] = sample_bilinear(img, uv.astype(np.float64))
    return r


def caches16(n: int = 16) -> list[dict]:
    """The 16-cache workload: per cache, its query count (log-uniform in [65,536, 2,073,600]), its
    training batch (uniform in {4096, 8192, 16384}), its weight seed and its 4 batches of training
    records (Gaussian lights seeded per cache).  Every cache reads the first `queries` records of a
    shared query buffer."""
    rng = np.random.default_rng(workload_seed(CACHES16))
    nq = np.exp(rng.uniform(np.log(65536), np.log(2073600), n)).astype(np.int64)
    nb = rng.choice([4096, 8192, 16384], n)
    out = []
    for i in range(n):
        recs = make_train_records(4 * int(nb[i]), rng, GaussianLights(workload_seed(CACHES16, 200 + i)))
        out.append({"queries": int(nq[i]), "batch": int(nb[i]), "seed": workload_seed(CACHES16, 100 + i),
                    "records": recs})
    return out



SCENE = 5          # synthetic scenes and wire records for the record front-end


def make_scene(seed: int, primitives: int = 4096, vertices: int = 3000, instances: int = 8, materials: int = 16,
               textures: int = 4, texture_size: int = 64) -> dict:
    """A random triangle scene in the layout of the reference's scene buffers (Scene.glsl:26-47):
    shared vertices and texcoords indexed per primitive corner, instances with random rotations,
    scales and translations (3x4 rows), materials of which about half sample one of the RGBA8 sRGB
    textures for diffuse and a quarter for specular."""
    import vknrc_amd.nrc as _n
    rng = np.random.default_rng(seed)
    s = {"vertices": rng.uniform(-1, 1, (vertices, 3)).astype(np.float32),
         "vertex_indices": rng.integers(0, vertices, primitives * 3).astype(np.uint32),
         "texcoords": rng.uniform(-0.5, 1.5, (vertices, 2)).astype(np.float32),
         "texcoord_indices": rng.integers(0, vertices, primitives * 3).astype(np.uint32),
         "material_ids": rng.integers(0, materials, primitives).astype(np.uint32)}
    # degenerate triangles have no normal: redraw corners until every triangle has area
    vi = s["vertex_indices"].reshape(-1, 3)
    for _ in range(100):
        v = s["vertices"][vi].astype(np.float64)
        area = np.linalg.norm(np.cross(v[:, 1] - v[:, 0], v[:, 2] - v[:, 0]), axis=1)
        bad = area < 1e-3
        if not bad.any():
            break
        vi[bad] = rng.integers(0, vertices, (int(bad.sum()), 3))
    s["vertex_indices"] = vi.reshape(-1).astype(np.uint32)
    T = np.zeros((instances, 3, 4), np.float32)
    for i in range(instances):
        q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
        T[i, :, :3] = q * rng.uniform(0.5, 2.0)
        T[i, :, 3] = rng.uniform(-3, 3, 3)
    s["transforms"] = T
    m = np.zeros(materials, _n.MATERIAL)
    m["diffuse"] = rng.uniform(0, 1, (materials, 3))
    m["specular"] = rng.uniform(0, 1, (materials, 3))
    m["emission"] = 0.0
    m["roughness"] = rng.uniform(0, 1, materials)
    m["ior"] = 1.5
    m["diffuse_texture_id"] = np.where(rng.uniform(size=materials) < 0.5,
                                       rng.integers(0, max(textures, 1), materials), _n.NO_TEXTURE) if textures \
        else _n.NO_TEXTURE
    m["specular_texture_id"] = np.where(rng.uniform(size=materials) < 0.25,
                                        rng.integers(0, max(textures, 1), materials), _n.NO_TEXTURE) if textures \
        else _n.NO_TEXTURE
    m["emission_texture_id"] = _n.NO_TEXTURE
    s["materials"] = m
    s["textures"] = [rng.integers(0, 2 ** 32, (texture_size, texture_size), dtype=np.uint64).astype(np.uint32)
                     for _ in range(textures)]
    return s


def make_packed_inputs(scene: dict, n: int, rng: np.random.Generator) -> np.ndarray:
    """n packed path vertices (MakeNRCInput, path_tracer.comp:159-166) on random primitives and
    instances: unorm16 barycentrics inside the triangle and a unorm16 sphere-encoded direction."""
    import vknrc_amd.nrc as _n
    p = np.zeros(n, _n.PACKED_INPUT)
    prims = len(scene["material_ids"])
    p["primitive_id"] = rng.integers(0, prims, n)
    flip = rng.integers(0, 2, n).astype(np.uint64)
    p["flip_bit_instance_id"] = (rng.integers(0, len(scene["transforms"]), n).astype(np.uint64) | (flip << 31))
    b = rng.uniform(0, 1, (n, 2))
    out = b.sum(1) > 1
    b[out] = 1 - b[out][:, ::-1]
    bq = np.round(b * 65535).astype(np.uint32)
    p["barycentric_2x16u"] = bq[:, 0] | (bq[:, 1] << 16)
    d = sph_encode(_sphere(rng, n))
    dq = np.round(np.clip(d, 0, 1) * 65535).astype(np.uint32)
    p["scattered_dir_2x16u"] = dq[:, 0] | (dq[:, 1] << 16)
    return p


def make_traced_paths(scene: dict, n: int, rng: np.random.Generator) -> np.ndarray:
    """n traced paths (nrc_traced_path): 1..8 vertices with random packed inputs, emissions and BRDF
    colours, random train batches, a quarter ending at a light."""
    import vknrc_amd.nrc as _n
    P = np.zeros(n, _n.TRACED_PATH)
    P["bounce"] = rng.integers(1, _n.MAX_BOUNCE + 1, n)
    P["batch"] = rng.integers(0, 4, n)
    P["light_terminate"] = rng.uniform(size=n) < 0.25
    P["light"] = rng.lognormal(-1.0, 1.0, (n, _n.MAX_BOUNCE, 3))
    P["color"] = rng.uniform(0, 1, (n, _n.MAX_BOUNCE, 3))
    P["input"] = make_packed_inputs(scene, n * _n.MAX_BOUNCE, rng).reshape(n, _n.MAX_BOUNCE)
    return P
