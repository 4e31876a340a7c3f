"""Encoder for the rhip_stream.h command stream.

A stream records a sequence of calls on SoftGLRender's abstract Renderer
interface (src/Render/Renderer.h:24-59) so one seeded scene can be replayed,
call for call, through RendererHIP and through the CPU checkers. Enum values
are the reference's own (src/Render/Texture.h:78-135, RenderStates.h:13-99).
"""
from __future__ import annotations

import struct
from dataclasses import dataclass, field

import numpy as np

MAGIC = 0x31534852
VERSION = 1

OP_TEXTURE, OP_TEX_INIT, OP_TEX_DATA, OP_TEX_LEVEL = 1, 2, 3, 4
OP_VAO, OP_FBO, OP_BEGIN, OP_VIEWPORT, OP_DRAW, OP_END, OP_READ, OP_FRAME = 5, 6, 7, 8, 9, 10, 11, 12

# Texture.h enums
TEX_2D, TEX_CUBE = 0, 1
FMT_RGBA8, FMT_FLOAT32 = 0, 1
USAGE_SAMPLER, USAGE_UPLOAD, USAGE_COLOR, USAGE_DEPTH, USAGE_OUTPUT = 1, 2, 4, 8, 16
(FILTER_NEAREST, FILTER_LINEAR, FILTER_NEAREST_MIPMAP_NEAREST, FILTER_LINEAR_MIPMAP_NEAREST,
 FILTER_NEAREST_MIPMAP_LINEAR, FILTER_LINEAR_MIPMAP_LINEAR) = range(6)
WRAP_REPEAT, WRAP_MIRRORED_REPEAT, WRAP_CLAMP_TO_EDGE, WRAP_CLAMP_TO_BORDER = range(4)
BORDER_BLACK, BORDER_WHITE = 0, 1

# RenderStates.h enums
(DEPTH_NEVER, DEPTH_LESS, DEPTH_EQUAL, DEPTH_LEQUAL, DEPTH_GREATER, DEPTH_NOTEQUAL,
 DEPTH_GEQUAL, DEPTH_ALWAYS) = range(8)
(BF_ZERO, BF_ONE, BF_SRC_COLOR, BF_SRC_ALPHA, BF_DST_COLOR, BF_DST_ALPHA, BF_ONE_MINUS_SRC_COLOR,
 BF_ONE_MINUS_SRC_ALPHA, BF_ONE_MINUS_DST_COLOR, BF_ONE_MINUS_DST_ALPHA) = range(10)
BLEND_ADD, BLEND_SUBTRACT, BLEND_REVERSE_SUBTRACT, BLEND_MIN, BLEND_MAX = range(5)
PRIM_POINT, PRIM_LINE, PRIM_TRIANGLE = range(3)
POLY_POINT, POLY_LINE, POLY_FILL = range(3)

# Shading models and defines (rhip_stream.h)
SHADING_BASIC, SHADING_BLINN_PHONG, SHADING_PBR, SHADING_BLINN_PHONG_DIR, SHADING_VERTEX_COLOR, \
    SHADING_NEURAL_PBR, SHADING_SKYBOX, SHADING_IBL_IRRADIANCE, SHADING_IBL_PREFILTER, SHADING_FXAA = range(10)
DEF_ALBEDO, DEF_NORMAL, DEF_EMISSIVE, DEF_AO, DEF_METALROUGH, DEF_EQUIRECT = 1, 2, 4, 8, 16, 32
SLOT_ALBEDO, SLOT_NORMAL, SLOT_EMISSIVE, SLOT_AO, SLOT_SHADOW, SLOT_METALROUGH, SLOT_IRRADIANCE, \
    SLOT_PREFILTER = range(8)
# the skybox, IBL and FXAA shaders' samplers (rhip_stream.h RHS_SLOT_EQUIRECT / _SCREEN / _CUBE)
SLOT_EQUIRECT = SLOT_SCREEN = SLOT_ALBEDO
SLOT_CUBE = SLOT_IRRADIANCE

MODEL_BYTES, SCENE_BYTES, MATERIAL_BYTES = 256, 64, 48


@dataclass
class RenderStates:
    """RenderStates (src/Render/RenderStates.h:52-63)."""
    blend: bool = False
    blend_func_rgb: int = BLEND_ADD
    blend_src_rgb: int = BF_ONE
    blend_dst_rgb: int = BF_ZERO
    blend_func_alpha: int = BLEND_ADD
    blend_src_alpha: int = BF_ONE
    blend_dst_alpha: int = BF_ZERO
    depth_test: bool = False
    depth_mask: bool = True
    depth_func: int = DEPTH_LESS
    cull_face: bool = False
    primitive_type: int = PRIM_TRIANGLE
    polygon_mode: int = POLY_FILL
    line_width: float = 1.0

    def set_blend_factor(self, src, dst):
        self.blend_src_rgb = self.blend_src_alpha = src
        self.blend_dst_rgb = self.blend_dst_alpha = dst

    def words(self) -> bytes:
        return struct.pack("<13if", int(self.blend), self.blend_func_rgb, self.blend_src_rgb, self.blend_dst_rgb,
                           self.blend_func_alpha, self.blend_src_alpha, self.blend_dst_alpha,
                           int(self.depth_test), int(self.depth_mask), self.depth_func, int(self.cull_face),
                           self.primitive_type, self.polygon_mode, self.line_width)


def _mat4_bytes(m) -> bytes:
    # column-major like glm::mat4
    return np.asarray(m, dtype=np.float32).T.reshape(-1).tobytes()


def _mat3_bytes(m) -> bytes:
    # glm::mat3 with GLM_FORCE_ALIGNED: three vec3 columns padded to 16 bytes
    m = np.asarray(m, dtype=np.float32)
    out = np.zeros((3, 4), np.float32)
    out[:, :3] = m.T
    return out.tobytes()


def uniforms_model(model, mvp, shadow_mvp=None, reverse_z=False) -> bytes:
    """UniformsModel (src/Viewer/Material.h:70-76): 256 bytes."""
    model = np.asarray(model, np.float64)
    inv_t = np.linalg.inv(model[:3, :3]).T if abs(np.linalg.det(model[:3, :3])) > 0 else np.eye(3)
    b = bytearray(MODEL_BYTES)
    struct.pack_into("<I", b, 0, 1 if reverse_z else 0)
    b[16:80] = _mat4_bytes(model)
    b[80:144] = _mat4_bytes(mvp)
    b[144:192] = _mat3_bytes(inv_t)
    b[192:256] = _mat4_bytes(np.eye(4) if shadow_mvp is None else shadow_mvp)
    return bytes(b)


def uniforms_scene(ambient, camera_pos, light_pos, light_color) -> bytes:
    """UniformsScene (src/Viewer/Material.h:63-68): 64 bytes."""
    b = bytearray(SCENE_BYTES)
    for i, v in enumerate((ambient, camera_pos, light_pos, light_color)):
        struct.pack_into("<3f", b, 16 * i, *[float(x) for x in v])
    return bytes(b)


def uniforms_material(enable_light=1, enable_ibl=0, enable_shadow=0, point_size=1.0, k_specular=0.5,
                      base_color=(1, 1, 1, 1)) -> bytes:
    """UniformsMaterial (src/Viewer/Material.h:78-86): 48 bytes."""
    b = bytearray(MATERIAL_BYTES)
    struct.pack_into("<3I2f", b, 0, enable_light, enable_ibl, enable_shadow, point_size, k_specular)
    struct.pack_into("<4f", b, 32, *[float(x) for x in base_color])
    return bytes(b)


def _pad4(data: bytes) -> bytes:
    r = len(data) % 4
    return data if r == 0 else data + b"\0" * (4 - r)


@dataclass
class Stream:
    chunks: list = field(default_factory=list)
    reads: list = field(default_factory=list)  # (tex, layer, level, which, nbytes, dtype, shape)
    textures: dict = field(default_factory=dict)

    def __post_init__(self):
        self.chunks.append(struct.pack("<II", MAGIC, VERSION))

    def _cmd(self, op: int, payload: bytes):
        payload = _pad4(payload)
        self.chunks.append(struct.pack("<II", op, len(payload) // 4))
        self.chunks.append(payload)

    def texture(self, tid, width, height, type=TEX_2D, format=FMT_RGBA8, usage=USAGE_SAMPLER, mipmaps=False,
                msaa=False, filter_min=FILTER_NEAREST, filter_mag=FILTER_NEAREST, wrap=WRAP_CLAMP_TO_EDGE,
                border=BORDER_BLACK):
        self.textures[tid] = dict(w=width, h=height, type=type, format=format, msaa=msaa, mipmaps=mipmaps)
        self._cmd(OP_TEXTURE, struct.pack("<14i", tid, type, format, width, height, usage, int(mipmaps), int(msaa),
                                          filter_min, filter_mag, wrap, wrap, wrap, border))

    def tex_init(self, tid):
        self._cmd(OP_TEX_INIT, struct.pack("<i", tid))

    def tex_data(self, tid, layers):
        """layers: list of arrays (h, w, 4) uint8 or (h, w) float32, one per layer."""
        data = b"".join(np.ascontiguousarray(l).tobytes() for l in layers)
        self._cmd(OP_TEX_DATA, struct.pack("<2i", tid, len(layers)) + data)

    def tex_level(self, tid, layer, level, data):
        self._cmd(OP_TEX_LEVEL, struct.pack("<3i", tid, layer, level) + np.ascontiguousarray(data).tobytes())

    def vao(self, vid, vertices: np.ndarray, indices: np.ndarray):
        """vertices: structured/float array (nverts, stride/4) float32; indices int32."""
        v = np.ascontiguousarray(vertices, dtype=np.float32)
        idx = np.ascontiguousarray(indices, dtype=np.int32).reshape(-1)
        stride = v.shape[1] * 4
        self._cmd(OP_VAO, struct.pack("<4i", vid, stride, v.shape[0], idx.size) + v.tobytes() + idx.tobytes())

    def fbo(self, fid, color=-1, color_layer=0, color_level=0, depth=-1):
        self._cmd(OP_FBO, struct.pack("<5i", fid, color, color_layer, color_level, depth))

    def begin(self, fid, color_flag=True, depth_flag=True, clear_color=(0, 0, 0, 0), clear_depth=1.0):
        self._cmd(OP_BEGIN, struct.pack("<3i5f", fid, int(color_flag), int(depth_flag), *clear_color, clear_depth))

    def viewport(self, x, y, w, h):
        self._cmd(OP_VIEWPORT, struct.pack("<4i", x, y, w, h))

    def draw(self, vao, shading, defines, states: RenderStates, model: bytes, scene: bytes, material: bytes,
             samplers=()):
        assert len(model) == MODEL_BYTES and len(scene) == SCENE_BYTES and len(material) == MATERIAL_BYTES
        s = b"".join(struct.pack("<2i", slot, tex) for slot, tex in samplers)
        self._cmd(OP_DRAW, struct.pack("<iiI", vao, shading, defines) + states.words() + model + scene + material +
                  struct.pack("<i", len(samplers)) + s)

    def end(self):
        self._cmd(OP_END, b"")

    def frame(self):
        self._cmd(OP_FRAME, b"")

    def read(self, tid, layer=0, level=0, which=0):
        t = self.textures[tid]
        w, h = max(1, t["w"] >> level), max(1, t["h"] >> level)
        spp = 4 if which == 1 else 1
        if t["format"] == FMT_RGBA8:
            shape, dtype = (h, w, spp, 4) if spp > 1 else (h, w, 4), np.uint8
        else:
            shape, dtype = (h, w, spp) if spp > 1 else (h, w), np.float32
        nbytes = int(np.prod(shape)) * np.dtype(dtype).itemsize
        self.reads.append((tid, layer, level, which, nbytes, dtype, shape))
        self._cmd(OP_READ, struct.pack("<4i", tid, layer, level, which))

    def bytes(self) -> bytes:
        return b"".join(self.chunks)

    def output_bytes(self) -> int:
        return sum(r[4] for r in self.reads)

    def split_outputs(self, buf: bytes):
        """Cuts the concatenated RHS_READ output into arrays, in read order."""
        out, off = [], 0
        for (_, _, _, _, nbytes, dtype, shape) in self.reads:
            out.append(np.frombuffer(buf[off:off + nbytes], dtype=dtype).reshape(shape).copy())
            off += nbytes
        return out
