"""Python mirror of the NRC state interface over the C ABI in include/nrc.h.

``NRCState`` follows the reference's ``VkNRCState`` (src/VkNRCState.hpp:17-90):
construction builds and initialises the cache, ``reset_mlp_buffers`` is
ResetMLPBuffers, ``set_use_ema_weights`` is SetUseEMAWeights, and the static
sizes match GetWeightCount / GetTrainBatchCount / GetTrainBatchSize.  ``infer``
replaces the NNDispatch<NNInference> pass and ``train`` the NNTrain pass group.

Device buffers are anything with ``data_ptr()`` (torch tensors on the HIP device)
or raw integer addresses.  The library is loaded from this directory; a missing
library raises instead of falling back to anything else.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libvknrc.so")

WIDTH = 64
OUTPUT_WIDTH = 3
QUERY_FLOATS = 14
TRAIN_RECORD_FLOATS = 17
TRAIN_BATCH_COUNT = 4
DEFAULT_TRAIN_BATCH_SIZE = 16384
MAX_TRAIN_BATCH_SIZE = 65536
MAX_QUERIES = 0x80000000  # most queries or records one call takes (NRC_MAX_QUERIES)

PATH_PLAIN_F32, PATH_FUSED_F16, PATH_FUSED_F8 = 0, 1, 2
ENCODING_TRI12, ENCODING_SINCOS6, ENCODING_UV_SINCOS16 = 0, 1, 2
LOSS_RELATIVE_L2_LUMINANCE, LOSS_L2 = 0, 1
INIT_KAIMING_NORMAL, INIT_XAVIER_UNIFORM = 0, 1
TRAIN_WRITE_USE_WEIGHTS = 1

STATUS = {0: "ok", 1: "invalid argument", 2: "unsupported", 3: "out of memory", 4: "HIP error",
          5: "capacity exceeded"}

# Wire records of the renderer (include/nrc.h): numpy layouts of nrc_packed_input,
# nrc_eval_record, nrc_wire_train_record, nrc_material and nrc_traced_path.
MAX_BOUNCE = 8
NO_TEXTURE = 0xFFFFFFFF
PACKED_INPUT = np.dtype([("primitive_id", "<u4"), ("flip_bit_instance_id", "<u4"), ("barycentric_2x16u", "<u4"),
                         ("scattered_dir_2x16u", "<u4")])
EVAL_RECORD = np.dtype([("dst", "<u4"), ("packed_input", PACKED_INPUT)])
WIRE_TRAIN_RECORD = np.dtype([("bias", "<f4", 3), ("factor", "<f4", 3), ("packed_input", PACKED_INPUT)])
MATERIAL = np.dtype([("diffuse", "<f4", 3), ("diffuse_texture_id", "<u4"), ("specular", "<f4", 3),
                     ("specular_texture_id", "<u4"), ("emission", "<f4", 3), ("emission_texture_id", "<u4"),
                     ("metallic", "<f4"), ("roughness", "<f4"), ("ior", "<f4"), ("pad", "<f4")])
TRACED_PATH = np.dtype([("bounce", "<u4"), ("batch", "<u4"), ("light_terminate", "<u4"), ("pad", "<u4"),
                        ("light", "<f4", (MAX_BOUNCE, 3)), ("color", "<f4", (MAX_BOUNCE, 3)),
                        ("input", PACKED_INPUT, MAX_BOUNCE)])
assert (PACKED_INPUT.itemsize, EVAL_RECORD.itemsize, WIRE_TRAIN_RECORD.itemsize, MATERIAL.itemsize,
        TRACED_PATH.itemsize) == (16, 20, 40, 64, 336)


class NRCError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"{STATUS.get(status, status)}: {message}")
        self.status = status


class CacheDesc(C.Structure):
    _fields_ = [("hidden_layers", C.c_uint32), ("encoding", C.c_int), ("loss", C.c_int), ("init", C.c_int),
                ("seed", C.c_uint64), ("train_batch_size", C.c_uint32), ("learning_rate", C.c_float),
                ("beta1", C.c_float), ("beta2", C.c_float), ("epsilon", C.c_float), ("ema_alpha", C.c_float),
                ("loss_scale", C.c_float), ("dynamic_loss_scale", C.c_int32),
                ("loss_scale_growth_interval", C.c_uint32), ("use_ema_weights", C.c_int32)]


class OptimizerState(C.Structure):
    _fields_ = [("t", C.c_uint32), ("beta1_t", C.c_float), ("beta2_t", C.c_float), ("alpha_t", C.c_float),
                ("alpha_t_1", C.c_float), ("loss_scale", C.c_float), ("clean_steps", C.c_uint32),
                ("skipped_steps", C.c_uint32), ("last_loss", C.c_float), ("last_count", C.c_uint32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Screen(C.Structure):
    """nrc_screen: RGBA32F (bias rgb, factor r) and RG32F (factor gb) images [height][pitch]."""
    _fields_ = [("d_bias_factor_r", C.c_void_p), ("d_factor_gb", C.c_void_p), ("width", C.c_uint32),
                ("height", C.c_uint32), ("pitch", C.c_uint32)]


class Texture(C.Structure):
    """nrc_texture: an RGBA8 sRGB image on the device."""
    _fields_ = [("d_rgba8", C.c_void_p), ("width", C.c_uint32), ("height", C.c_uint32)]


class Scene(C.Structure):
    """nrc_scene: device pointers of the scene buffers (Scene.glsl:26-47)."""
    _fields_ = [("d_vertices", C.c_void_p), ("d_vertex_indices", C.c_void_p), ("d_texcoords", C.c_void_p),
                ("d_texcoord_indices", C.c_void_p), ("d_materials", C.c_void_p), ("d_material_ids", C.c_void_p),
                ("d_transforms", C.c_void_p), ("d_textures", C.c_void_p), ("texture_count", C.c_uint32),
                ("d_primitives", C.c_void_p)]


class GroupItem(C.Structure):
    """nrc_group_item: one cache's inputs and outputs in a grouped call."""
    _fields_ = [("cache", C.c_void_p), ("d_queries", C.c_void_p), ("d_query_count", C.c_void_p),
                ("max_queries", C.c_uint32), ("d_rgb", C.c_void_p), ("d_batches", C.c_void_p),
                ("d_batch_counts", C.c_void_p), ("max_batch_records", C.c_uint32)]


_lib = None


def load_library():
    """Loads libvknrc.so (raises when it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built; run vknrc_amd.build.build() first")
    L = C.CDLL(LIB_PATH)
    P, U32, I32 = C.c_void_p, C.c_uint32, C.c_int32
    L.nrc_status_string.restype = C.c_char_p
    L.nrc_status_string.argtypes = [C.c_int]
    L.nrc_last_error.restype = C.c_char_p
    L.nrc_cache_desc_default.argtypes = [C.POINTER(CacheDesc)]
    L.nrc_cache_desc_default.restype = None
    L.nrc_weight_count.restype = U32
    L.nrc_weight_count.argtypes = [U32]
    L.nrc_init_weights.argtypes = [C.POINTER(CacheDesc), C.c_uint64, P]
    L.nrc_cache_create.argtypes = [C.POINTER(CacheDesc), C.POINTER(P)]
    L.nrc_cache_reset.argtypes = [P, C.c_uint64]
    L.nrc_cache_destroy.argtypes = [P]
    L.nrc_cache_destroy.restype = None
    L.nrc_cache_get_desc.argtypes = [P, C.POINTER(CacheDesc)]
    L.nrc_set_use_ema.argtypes = [P, I32]
    L.nrc_infer.argtypes = [P, P, P, U32, P, C.c_int, P]
    L.nrc_train.argtypes = [P, P, P, U32, C.c_int, U32, P]
    L.nrc_frame.argtypes = [P, P, P, U32, P, P, P, U32, C.c_int, P]
    L.nrc_get_state.argtypes = [P, P, C.POINTER(OptimizerState), P, P]
    L.nrc_set_state.argtypes = [P, P, C.POINTER(OptimizerState)]
    L.nrc_get_fp8_scales.argtypes = [P, P, P, P]
    L.nrc_cache_footprint.argtypes = [C.POINTER(CacheDesc), C.c_int, P]
    L.nrc_cache_set_grad_chunks.argtypes = [P, U32]
    L.nrc_composite.argtypes = [P, P, P, U32, C.POINTER(Screen), P, P, U32, P]
    L.nrc_dst_screen.argtypes = [U32, U32]
    L.nrc_dst_screen.restype = U32
    L.nrc_dst_train.argtypes = [U32, U32, U32]
    L.nrc_dst_train.restype = U32
    L.nrc_set_infer_workgroups_per_cu.argtypes = [U32]
    L.nrc_encode.argtypes = [C.c_int, P, U32, P, P]
    L.nrc_sph_encode.argtypes = [P, U32, P, P]
    L.nrc_mlp_forward_f16.argtypes = [U32, P, P, U32, P, P]
    SC = C.POINTER(Scene)
    L.nrc_scene_pack_primitives.argtypes = [SC, U32, P, P]
    L.nrc_unpack_eval_records.argtypes = [SC, P, P, U32, P, P]
    L.nrc_unpack_train_records.argtypes = [SC, P, P, U32, P, P]
    L.nrc_composite_records.argtypes = [P, P, P, U32, C.POINTER(Screen), P, U32, P]
    L.nrc_build_train_records.argtypes = [P, U32, P, P, U32, P, P, U32, P]
    L.nrc_screen_tonemap.argtypes = [P, P, U32, U32, U32, I32, U32, P, P]
    L.nrc_mlp_gradient_f16.argtypes = [U32, P, P, P, U32, P, P]
    L.nrc_mlp_gradient_f16_noalloc.argtypes = [U32, P, P, P, U32, P, P]
    GI = C.POINTER(GroupItem)
    L.nrc_group_infer.argtypes = [GI, U32, C.c_int, P]
    L.nrc_group_train.argtypes = [GI, U32, U32, C.c_int, U32, P]
    L.nrc_group_frame.argtypes = [GI, U32, C.c_int, P]
    L.nrc_group_frame_graph.argtypes = [GI, U32, C.c_int, C.POINTER(P)]
    L.nrc_graph_launch.argtypes = [P, P]
    L.nrc_graph_destroy.argtypes = [P]
    L.nrc_graph_destroy.restype = None
    for f in ("nrc_init_weights", "nrc_cache_create", "nrc_cache_reset", "nrc_cache_get_desc", "nrc_set_use_ema",
              "nrc_infer", "nrc_train", "nrc_frame", "nrc_get_state", "nrc_set_state", "nrc_encode",
              "nrc_sph_encode", "nrc_mlp_forward_f16", "nrc_mlp_gradient_f16", "nrc_mlp_gradient_f16_noalloc",
              "nrc_group_infer", "nrc_scene_pack_primitives", "nrc_unpack_eval_records", "nrc_unpack_train_records",
              "nrc_composite_records",
              "nrc_build_train_records", "nrc_screen_tonemap",
              "nrc_group_train", "nrc_group_frame", "nrc_group_frame_graph", "nrc_graph_launch"):
        getattr(L, f).restype = C.c_int
    _lib = L
    return L


def _check(status: int) -> None:
    if status != 0:
        raise NRCError(status, load_library().nrc_last_error().decode())


def _ptr(x):
    if x is None:
        return None
    if isinstance(x, int):
        return C.c_void_p(x)
    return C.c_void_p(x.data_ptr())


def _stream(stream):
    if stream is None:
        try:
            import torch
            if torch.cuda.is_available():
                return C.c_void_p(torch.cuda.current_stream().cuda_stream)
        except ImportError:
            pass
        return None
    if isinstance(stream, int):
        return C.c_void_p(stream)
    return C.c_void_p(stream.cuda_stream)


def weight_count(hidden_layers: int = 5) -> int:
    return hidden_layers * WIDTH * WIDTH + OUTPUT_WIDTH * WIDTH


@dataclass
class NRCConfig:
    """Cache description; defaults are the reference constants (Constant.glsl:5-13,
    nrc_optimize.comp:10-11, VkNRCState.hpp:22-25)."""
    hidden_layers: int = 5
    encoding: int = ENCODING_TRI12
    loss: int = LOSS_RELATIVE_L2_LUMINANCE
    init: int = INIT_KAIMING_NORMAL
    seed: int = 0
    train_batch_size: int = DEFAULT_TRAIN_BATCH_SIZE
    learning_rate: float = 0.002
    beta1: float = 0.9
    beta2: float = 0.999
    epsilon: float = 1e-8
    ema_alpha: float = 0.99
    loss_scale: float = 1.0
    dynamic_loss_scale: bool = False
    loss_scale_growth_interval: int = 2000
    use_ema_weights: bool = False

    def to_c(self) -> CacheDesc:
        d = CacheDesc()
        for k, _ in CacheDesc._fields_:
            setattr(d, k, int(getattr(self, k)) if isinstance(getattr(self, k), bool) else getattr(self, k))
        return d


def default_config() -> NRCConfig:
    """The C library's defaults, read back through nrc_cache_desc_default."""
    d = CacheDesc()
    load_library().nrc_cache_desc_default(C.byref(d))
    kw = {k: getattr(d, k) for k, _ in CacheDesc._fields_}
    kw["dynamic_loss_scale"] = bool(kw["dynamic_loss_scale"])
    kw["use_ema_weights"] = bool(kw["use_ema_weights"])
    return NRCConfig(**kw)


def init_weights(config: NRCConfig, seed: int | None = None) -> np.ndarray:
    """Host-only: the FP32 initial weights of a cache built from ``config``."""
    w = np.zeros(weight_count(config.hidden_layers), np.float32)
    d = config.to_c()
    _check(load_library().nrc_init_weights(C.byref(d), int(config.seed if seed is None else seed),
                                           w.ctypes.data_as(C.c_void_p)))
    return w


def set_infer_workgroups_per_cu(workgroups: int) -> None:
    """Persistent inference workgroups per CU for later launches (results do not depend on it)."""
    _check(load_library().nrc_set_infer_workgroups_per_cu(int(workgroups)))


def cache_footprint(config: NRCConfig, path: int = PATH_FUSED_F16) -> int:
    """Device bytes a cache built from ``config`` allocates (host-only; nrc_cache_footprint)."""
    d = config.to_c()
    b = C.c_uint64(0)
    _check(load_library().nrc_cache_footprint(C.byref(d), int(path), C.byref(b)))
    return int(b.value)


class NRCState:
    """One neural radiance cache: weights, use-weights and optimizer state on the device."""

    TRAIN_BATCH_COUNT = TRAIN_BATCH_COUNT

    def __init__(self, config: NRCConfig | None = None):
        self.config = config or NRCConfig()
        self._lib = load_library()
        self._desc = self.config.to_c()
        h = C.c_void_p()
        _check(self._lib.nrc_cache_create(C.byref(self._desc), C.byref(h)))
        self._h = h

    def close(self) -> None:
        if getattr(self, "_h", None) and self._h.value:
            self._lib.nrc_cache_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- VkNRCState interface
    @staticmethod
    def get_weight_count(hidden_layers: int = 5) -> int:
        return weight_count(hidden_layers)

    @staticmethod
    def get_train_batch_count() -> int:
        return TRAIN_BATCH_COUNT

    def get_train_batch_size(self) -> int:
        return self.config.train_batch_size

    @property
    def weight_count(self) -> int:
        return weight_count(self.config.hidden_layers)

    def reset_mlp_buffers(self, seed: int | None = None) -> None:
        _check(self._lib.nrc_cache_reset(self._h, int(self.config.seed if seed is None else seed)))

    def set_use_ema_weights(self, use_ema: bool) -> None:
        self.config.use_ema_weights = bool(use_ema)
        _check(self._lib.nrc_set_use_ema(self._h, int(bool(use_ema))))

    def is_use_ema_weights(self) -> bool:
        return self.config.use_ema_weights

    # ---- NN passes
    def infer(self, queries, out, count=None, max_count: int | None = None, path: int = PATH_FUSED_F16,
              stream=None) -> None:
        n = int(max_count if max_count is not None else out.shape[0])
        _check(self._lib.nrc_infer(self._h, _ptr(queries), _ptr(count), n, _ptr(out), int(path), _stream(stream)))

    def train(self, records, count=None, max_count: int | None = None, path: int = PATH_FUSED_F16,
              write_use_weights: bool = True, stream=None) -> None:
        n = int(max_count if max_count is not None else records.shape[0])
        flags = TRAIN_WRITE_USE_WEIGHTS if write_use_weights else 0
        _check(self._lib.nrc_train(self._h, _ptr(records), _ptr(count), n, int(path), flags, _stream(stream)))

    def frame(self, queries, out, batches, query_count=None, batch_counts=None, max_queries=None,
              max_batch_records=None, path: int = PATH_FUSED_F16, stream=None) -> None:
        nq = int(max_queries if max_queries is not None else out.shape[0])
        nb = int(max_batch_records if max_batch_records is not None else batches.shape[0] // TRAIN_BATCH_COUNT)
        _check(self._lib.nrc_frame(self._h, _ptr(queries), _ptr(query_count), nq, _ptr(out), _ptr(batches),
                                   _ptr(batch_counts), nb, int(path), _stream(stream)))

    # ---- state export / import (synchronous)
    def get_state(self) -> dict:
        wc = self.weight_count
        entries = np.zeros((wc, 4), np.float32)
        w16 = np.zeros(wc, np.uint16)
        u16 = np.zeros(wc, np.uint16)
        st = OptimizerState()
        _check(self._lib.nrc_get_state(self._h, entries.ctypes.data_as(C.c_void_p), C.byref(st),
                                       w16.ctypes.data_as(C.c_void_p), u16.ctypes.data_as(C.c_void_p)))
        return {"m": entries[:, 0].copy(), "v": entries[:, 1].copy(), "weight": entries[:, 2].copy(),
                "ema_weight": entries[:, 3].copy(), "weights_f16": w16.view(np.float16).copy(),
                "use_weights_f16": u16.view(np.float16).copy(), "state": st.as_dict()}

    def set_grad_chunks(self, chunks: int) -> None:
        """64-record chunks per gradient workgroup of this cache's fused training steps (0: automatic)."""
        _check(self._lib.nrc_cache_set_grad_chunks(self._h, int(chunks)))
        self.grad_chunks = int(chunks)

    def get_fp8_scales(self) -> dict:
        """The power-of-two scales of the last FP8 inference (synchronous): layer_scales[l] = C_l,
        l = 0..H (activations enter layer l as C_l a_l, weights as (C_{l+1}/C_l) W_l), out_scale,
        and the number of FP8 inferences in the activation history."""
        sc = np.zeros(self.config.hidden_layers + 1, np.float32)
        o = C.c_float(0.0)
        f = C.c_uint32(0)
        _check(self._lib.nrc_get_fp8_scales(self._h, sc.ctypes.data_as(C.c_void_p), C.byref(o), C.byref(f)))
        return {"layer_scales": sc, "out_scale": float(o.value), "frames": int(f.value)}

    def get_optimizer_state(self) -> dict:
        """The scalar optimizer state only (t, beta^t, alpha^t, loss scale, last loss); synchronous."""
        st = OptimizerState()
        _check(self._lib.nrc_get_state(self._h, None, C.byref(st), None, None))
        return st.as_dict()

    def set_state(self, weight=None, m=None, v=None, ema_weight=None, state: dict | None = None) -> None:
        wc = self.weight_count
        ent = None
        if weight is not None:
            ent = np.zeros((wc, 4), np.float32)
            ent[:, 2] = weight
            ent[:, 0] = 0 if m is None else m
            ent[:, 1] = 0 if v is None else v
            ent[:, 3] = weight if ema_weight is None else ema_weight
        st = None
        if state is not None:
            st = OptimizerState()
            for k, _ in OptimizerState._fields_:
                if k in state:
                    setattr(st, k, state[k])
        _check(self._lib.nrc_set_state(self._h, None if ent is None else ent.ctypes.data_as(C.c_void_p),
                                       None if st is None else C.byref(st)))


# ---- stateless entry points
def encode(queries, n: int, features, encoding: int = ENCODING_TRI12, stream=None) -> None:
    _check(load_library().nrc_encode(int(encoding), _ptr(queries), int(n), _ptr(features), _stream(stream)))


DST_INVALID = 0xFFFFFFFF


def dst_screen(x, y):
    """Destination code of screen pixel (x, y) (EncodeNRCEvalDstScreen, NRCRecord.glsl:19)."""
    return ((np.asarray(x, np.uint64) & 0x7FFF | ((np.asarray(y, np.uint64) & 0x7FFF) << 15)) << 1).astype(np.uint32)


def dst_train(b, l, r):
    """Destination code of training records l..r of batch b (EncodeNRCEvalDstTrain, NRCRecord.glsl:20-22)."""
    v = (np.asarray(b, np.uint64) & 3) | ((np.asarray(l, np.uint64) & 0x3FFF) << 2) | \
        ((np.asarray(r, np.uint64) & 0x3FFF) << 16)
    return ((v << 1) | 1).astype(np.uint32)


def composite(rgb, dst, max_count: int, count=None, bias_factor_r=None, factor_gb=None, width: int = 0,
              height: int = 0, pitch: int | None = None, batches=None, train_factors=None, batch_stride: int = 0,
              stream=None) -> None:
    """The inference epilogue (nrc_composite): screen composite into bias_factor_r and back-fill of the
    training targets of batches, from the inference output rgb and per-query destination codes."""
    scr = None
    if bias_factor_r is not None:
        scr = Screen(bias_factor_r.data_ptr(), factor_gb.data_ptr(), int(width), int(height),
                     int(pitch if pitch is not None else width))
    _check(load_library().nrc_composite(_ptr(rgb), _ptr(dst), _ptr(count), int(max_count),
                                        None if scr is None else C.byref(scr), _ptr(batches),
                                        _ptr(train_factors), int(batch_stride), _stream(stream)))


class DeviceScene:
    """A scene uploaded to the device (torch tensors) with its nrc_scene descriptor.

    vertices [V][3] f32, vertex_indices [3P] u32, texcoords [T][2] f32, texcoord_indices [3P] u32,
    materials: numpy array of MATERIAL, material_ids [P] u32, transforms [I][3][4] f32,
    textures: list of [h][w] uint32 RGBA8 arrays (R in the low byte)."""

    def __init__(self, torch, vertices, vertex_indices, texcoords, texcoord_indices, materials, material_ids,
                 transforms, textures=()):
        dev = "cuda"

        def put(a, dt):
            a = np.ascontiguousarray(np.asarray(a, dt))
            return torch.from_numpy(a.view(np.int32) if a.dtype == np.uint32 else a).to(dev)

        self.vertices = put(vertices, np.float32)
        self.vertex_indices = put(vertex_indices, np.uint32)
        self.texcoords = put(texcoords, np.float32)
        self.texcoord_indices = put(texcoord_indices, np.uint32)
        mats = np.ascontiguousarray(np.asarray(materials, MATERIAL))
        self.materials = torch.from_numpy(mats.view(np.uint8).copy()).to(dev)
        self.material_ids = put(material_ids, np.uint32)
        self.transforms = put(transforms, np.float32)
        self.texture_images = [put(t, np.uint32) for t in textures]
        desc = np.zeros(max(1, len(textures)), np.dtype([("p", "<u8"), ("w", "<u4"), ("h", "<u4")]))
        for i, (img, host) in enumerate(zip(self.texture_images, textures)):
            desc[i] = (img.data_ptr(), np.asarray(host).shape[1], np.asarray(host).shape[0])
        self.texture_table = torch.from_numpy(desc.view(np.uint8).copy()).to(dev)
        self.c = Scene(self.vertices.data_ptr(), self.vertex_indices.data_ptr(), self.texcoords.data_ptr(),
                       self.texcoord_indices.data_ptr(), self.materials.data_ptr(), self.material_ids.data_ptr(),
                       self.transforms.data_ptr(), self.texture_table.data_ptr() if textures else None,
                       len(textures), None)
        self.torch = torch
        self.primitive_count = int(np.asarray(material_ids).shape[0])
        self.primitives = None

    def pack_primitives(self, stream=None) -> None:
        """Builds the 64-byte packed primitive records (nrc_scene_pack_primitives); later unpacks read
        them."""
        self.primitives = self.torch.empty(max(1, self.primitive_count) * 64, dtype=self.torch.uint8,
                                           device=self.vertices.device)
        _check(load_library().nrc_scene_pack_primitives(C.byref(self.c), self.primitive_count,
                                                        self.primitives.data_ptr(), _stream(stream)))
        self.c.d_primitives = self.primitives.data_ptr()

    def use_packed(self, on: bool) -> None:
        """Switches the unpack between the packed records (after pack_primitives) and the index arrays."""
        if on and self.primitives is None:
            raise RuntimeError("pack_primitives() first")
        self.c.d_primitives = self.primitives.data_ptr() if on else None


def unpack_eval_records(scene: DeviceScene, records, max_count: int, queries, count=None, stream=None) -> None:
    """UnpackNRCInput of nrc_eval_record rows into nrc_query rows (nrc_unpack_eval_records)."""
    _check(load_library().nrc_unpack_eval_records(C.byref(scene.c), _ptr(records), _ptr(count), int(max_count),
                                                  _ptr(queries), _stream(stream)))


def unpack_train_records(scene: DeviceScene, records, max_count: int, out, count=None, stream=None) -> None:
    """nrc_wire_train_record rows -> nrc_train_record rows (query + target = bias)."""
    _check(load_library().nrc_unpack_train_records(C.byref(scene.c), _ptr(records), _ptr(count), int(max_count),
                                                   _ptr(out), _stream(stream)))


def composite_records(rgb, records, max_count: int, count=None, bias_factor_r=None, factor_gb=None, width: int = 0,
                      height: int = 0, pitch: int | None = None, batches=None, batch_stride: int = 0,
                      stream=None) -> None:
    """The inference epilogue over wire records (nrc_composite_records)."""
    scr = None
    if bias_factor_r is not None:
        scr = Screen(bias_factor_r.data_ptr(), factor_gb.data_ptr(), int(width), int(height),
                     int(pitch if pitch is not None else width))
    _check(load_library().nrc_composite_records(_ptr(rgb), _ptr(records), _ptr(count), int(max_count),
                                                None if scr is None else C.byref(scr), _ptr(batches),
                                                int(batch_stride), _stream(stream)))


def build_train_records(paths, n_paths: int, batches, batch_counts, batch_capacity: int, eval_records, eval_count,
                        eval_capacity: int, stream=None) -> None:
    """The path suffix sums and record appends of ExtendedPathTrace (nrc_build_train_records)."""
    _check(load_library().nrc_build_train_records(_ptr(paths), int(n_paths), _ptr(batches), _ptr(batch_counts),
                                                  int(batch_capacity), _ptr(eval_records), _ptr(eval_count),
                                                  int(eval_capacity), _stream(stream)))


def screen_tonemap(color, width: int, height: int, out, accumulate_img=None, accumulate: bool = False,
                   accumulate_count: int = 0, pitch: int | None = None, stream=None) -> None:
    """The screen pass (nrc_screen_tonemap): optional accumulation, filmic tone map and gamma."""
    _check(load_library().nrc_screen_tonemap(_ptr(color), _ptr(accumulate_img), int(width), int(height),
                                             int(pitch if pitch is not None else width), int(bool(accumulate)),
                                             int(accumulate_count), _ptr(out), _stream(stream)))


def sph_encode(dirs, n: int, uv, stream=None) -> None:
    _check(load_library().nrc_sph_encode(_ptr(dirs), int(n), _ptr(uv), _stream(stream)))


def mlp_forward_f16(weights_f16, inputs_f16, n: int, out, hidden_layers: int = 5, stream=None) -> None:
    _check(load_library().nrc_mlp_forward_f16(int(hidden_layers), _ptr(weights_f16), _ptr(inputs_f16), int(n),
                                              _ptr(out), _stream(stream)))


def mlp_gradient_f16(weights_f16, inputs_f16, targets, n: int, dw, hidden_layers: int = 5, stream=None) -> None:
    _check(load_library().nrc_mlp_gradient_f16(int(hidden_layers), _ptr(weights_f16), _ptr(inputs_f16),
                                               _ptr(targets), int(n), _ptr(dw), _stream(stream)))


def mlp_gradient_f16_noalloc(weights_f16, inputs_f16, targets, n: int, dw, hidden_layers: int = 5,
                             stream=None) -> None:
    """mlp_gradient_f16 with the library's fixed device scratch (no allocation); n <= 16384."""
    _check(load_library().nrc_mlp_gradient_f16_noalloc(int(hidden_layers), _ptr(weights_f16), _ptr(inputs_f16),
                                                       _ptr(targets), int(n), _ptr(dw), _stream(stream)))


# ---- many caches at once
def _items(entries):
    """Builds the nrc_group_item array from dicts with keys cache (NRCState), queries, query_count,
    max_queries, rgb, batches, batch_counts, max_batch_records (all optional but cache)."""
    arr = (GroupItem * max(1, len(entries)))()
    keep = []
    for i, e in enumerate(entries):
        st = e["cache"]
        q, rgb, b = e.get("queries"), e.get("rgb"), e.get("batches")
        mq = e.get("max_queries", 0 if q is None else (rgb.shape[0] if rgb is not None else q.shape[0]))
        mb = e.get("max_batch_records", 0 if b is None else b.shape[0] // TRAIN_BATCH_COUNT)
        addr = lambda x: None if x is None else (x if isinstance(x, int) else x.data_ptr())  # noqa: E731
        arr[i] = GroupItem(st._h.value, addr(q), addr(e.get("query_count")), int(mq), addr(rgb), addr(b),
                           addr(e.get("batch_counts")), int(mb))
        keep.append(e)
    return arr, keep


def group_infer(entries, path: int = PATH_FUSED_F16, stream=None) -> None:
    arr, _ = _items(entries)
    _check(load_library().nrc_group_infer(arr, len(entries), int(path), _stream(stream)))


def group_train(entries, batch: int, path: int = PATH_FUSED_F16, write_use_weights: bool = False,
                stream=None) -> None:
    arr, _ = _items(entries)
    flags = TRAIN_WRITE_USE_WEIGHTS if write_use_weights else 0
    _check(load_library().nrc_group_train(arr, len(entries), int(batch), int(path), flags, _stream(stream)))


def group_frame(entries, path: int = PATH_FUSED_F16, stream=None) -> None:
    arr, _ = _items(entries)
    _check(load_library().nrc_group_frame(arr, len(entries), int(path), _stream(stream)))


class GroupGraph:
    """A group frame captured once as a HIP graph (nrc_group_frame_graph) and replayed per frame."""

    def __init__(self, entries, path: int = PATH_FUSED_F16):
        self._lib = load_library()
        arr, self._keep = _items(entries)
        h = C.c_void_p()
        _check(self._lib.nrc_group_frame_graph(arr, len(entries), int(path), C.byref(h)))
        self._h = h

    def launch(self, stream=None) -> None:
        _check(self._lib.nrc_graph_launch(self._h, _stream(stream)))

    def close(self) -> None:
        if getattr(self, "_h", None) and self._h.value:
            self._lib.nrc_graph_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
