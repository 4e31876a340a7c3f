"""RendererHIP for Python: SoftGLRender's abstract Renderer interface
(src/Render/Renderer.h:24-59) and its resource classes, implemented on librhip.

Method names, argument meaning and error behaviour follow the reference:
create* return a resource object (or None after logging, like RendererSoft's
nullptr returns); draw() renders the bound VAO with the bound program,
resources and pipeline states into the bound framebuffer.
"""
from __future__ import annotations

import ctypes
import logging
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from . import stream as S

log = logging.getLogger("softglrender_amd")

RENDERER_HIP = "HIP"  # next to Renderer_SOFT / _OPENGL / _Vulkan (Renderer.h:18-22)

_SLOT_BY_NAME = {"u_albedoMap": S.SLOT_ALBEDO, "u_normalMap": S.SLOT_NORMAL, "u_emissiveMap": S.SLOT_EMISSIVE,
                 "u_aoMap": S.SLOT_AO, "u_shadowMap": S.SLOT_SHADOW, "u_metalRoughnessMap": S.SLOT_METALROUGH,
                 "u_irradianceMap": S.SLOT_IRRADIANCE, "u_prefilterMap": S.SLOT_PREFILTER,
                 "u_equirectangularMap": S.SLOT_EQUIRECT, "u_screenTexture": S.SLOT_SCREEN, "u_cubeMap": S.SLOT_CUBE}
_BLOCK_SIZES = {"UniformsModel": S.MODEL_BYTES, "UniformsScene": S.SCENE_BYTES, "UniformsMaterial": S.MATERIAL_BYTES}
# the IBL prefilter's and the quad filter's blocks take the material block's place (Material.h:88-95)
_BLOCK_ALIASES = {"UniformsPrefilter": "UniformsMaterial", "UniformsQuadFilter": "UniformsMaterial"}
_DEFINES = {"ALBEDO_MAP": S.DEF_ALBEDO, "NORMAL_MAP": S.DEF_NORMAL, "EMISSIVE_MAP": S.DEF_EMISSIVE, "AO_MAP": S.DEF_AO,
            "METALROUGHNESS_MAP": S.DEF_METALROUGH, "EQUIRECTANGULAR_MAP": S.DEF_EQUIRECT}


@dataclass
class TextureDesc:
    """TextureDesc (Texture.h:62-71) plus the texel layout (per texture here)."""
    width: int = 0
    height: int = 0
    type: int = S.TEX_2D
    format: int = S.FMT_RGBA8
    usage: int = S.USAGE_SAMPLER
    useMipmaps: bool = False
    multiSample: bool = False
    tag: str = ""
    layout: int = 0


@dataclass
class SamplerDesc:
    """SamplerDesc (Texture.h:33-41)."""
    filterMin: int = S.FILTER_NEAREST
    filterMag: int = S.FILTER_NEAREST
    wrapS: int = S.WRAP_CLAMP_TO_EDGE
    wrapT: int = S.WRAP_CLAMP_TO_EDGE
    wrapR: int = S.WRAP_CLAMP_TO_EDGE
    borderColor: int = S.BORDER_BLACK


@dataclass
class ClearStates:
    """ClearStates (RenderStates.h:65-70)."""
    depthFlag: bool = False
    colorFlag: bool = False
    clearColor: tuple = (0.0, 0.0, 0.0, 0.0)
    clearDepth: float = 1.0


@dataclass
class VertexArray:
    """VertexArray (Vertex.h:26-35): float attributes, int32 indices."""
    vertexSize: int = 0
    vertexes: np.ndarray = None   # (n, stride/4) float32
    indices: np.ndarray = None    # int32


class TextureHIP(TextureDesc):
    def __init__(self, r: "RendererHIP", handle: int, desc: TextureDesc):
        super().__init__(**desc.__dict__)
        self._r, self.handle = r, handle

    def getId(self):
        return self.handle

    def getLevelWidth(self, level):
        return max(1, self.width >> level)

    def getLevelHeight(self, level):
        return max(1, self.height >> level)

    def setSamplerDesc(self, s: SamplerDesc):
        sd = _lib.SamplerDesc(s.filterMin, s.filterMag, s.wrapS, s.wrapT, s.wrapR, s.borderColor)
        _lib.check(self._r.lib.rhip_texture_set_sampler(self._r.dev, self.handle, ctypes.byref(sd)))

    def initImageData(self):
        _lib.check(self._r.lib.rhip_texture_init(self._r.dev, self.handle))

    def setImageData(self, buffers):
        """Level-0 images, one per layer: (h, w, 4) uint8 or (h, w) float32, rows from y = 0."""
        data = b"".join(np.ascontiguousarray(b).tobytes() for b in buffers)
        _lib.check(self._r.lib.rhip_texture_upload(self._r.dev, self.handle, len(buffers), data, len(data)))

    def setLevelData(self, layer, level, image):
        data = np.ascontiguousarray(image).tobytes()
        _lib.check(self._r.lib.rhip_texture_upload_level(self._r.dev, self.handle, layer, level, data, len(data)))

    def readImage(self, layer=0, level=0, samples=False):
        """Reads one level (row-major); samples=True returns the raw 4x samples."""
        w, h = self.getLevelWidth(level), self.getLevelHeight(level)
        spp = 4 if samples else 1
        if self.format == S.FMT_RGBA8:
            out = np.empty((h, w, spp, 4) if samples else (h, w, 4), np.uint8)
        else:
            out = np.empty((h, w, spp) if samples else (h, w), np.float32)
        _lib.check(self._r.lib.rhip_texture_read(self._r.dev, self.handle, layer, level, 1 if samples else 0,
                                                 out.ctypes.data, out.nbytes))
        return out

    def dumpImage(self, path, layer, level):
        np.save(path, self.readImage(layer, level))

    def __del__(self):
        try:
            if self._r.dev:
                self._r.lib.rhip_texture_destroy(self._r.dev, self.handle)
        except Exception:
            pass


class VertexArrayObjectHIP:
    def __init__(self, r, handle, nbytes):
        self._r, self.handle, self.nbytes = r, handle, nbytes

    def getId(self):
        return self.handle

    def updateVertexData(self, data: np.ndarray):
        b = np.ascontiguousarray(data, np.float32).tobytes()
        _lib.check(self._r.lib.rhip_vao_update(self._r.dev, self.handle, b, len(b)))


class FrameBufferHIP:
    """FrameBuffer (Framebuffer.h:20-91)."""

    def __init__(self, r, handle, offscreen):
        self._r, self.handle, self.offscreen = r, handle, offscreen
        self.color = None
        self.depth = None

    def getId(self):
        return self.handle

    def isValid(self):
        return self.color is not None or self.depth is not None

    def setColorAttachment(self, tex: TextureHIP, level=0, face=0):
        _lib.check(self._r.lib.rhip_fbo_attach_color(self._r.dev, self.handle, tex.handle, face, level))
        self.color = tex

    def setDepthAttachment(self, tex: TextureHIP):
        _lib.check(self._r.lib.rhip_fbo_attach_depth(self._r.dev, self.handle, tex.handle))
        self.depth = tex


class ShaderProgramHIP:
    """ShaderProgram (ShaderProgram.h:15-58): defines + a shading model that picks the kernels."""

    def __init__(self, r):
        self._r = r
        self.defines = 0
        self.shading = None
        self.handle = None

    def getId(self):
        return self.handle

    def addDefine(self, name):
        self.defines |= _DEFINES[name]

    def addDefines(self, names):
        for n in names:
            self.addDefine(n)

    def setShading(self, shading: int) -> bool:
        """Counterpart of ShaderProgramSoft::SetShaders (ShaderProgramSoft.h:28-57)."""
        h = ctypes.c_int()
        rc = self._r.lib.rhip_program_create(self._r.dev, shading, self.defines, ctypes.byref(h))
        if rc != 0:
            log.error("setShading failed: %s", self._r.lib.rhip_last_error().decode())
            return False
        self.shading, self.handle = shading, h.value
        return True


class PipelineStates:
    """PipelineStates (PipelineStates.h:13-22)."""

    def __init__(self, renderStates: S.RenderStates):
        self.renderStates = renderStates


class UniformBlockHIP:
    """UniformBlock (Uniform.h:38-47): bytes are snapshotted at draw()."""

    def __init__(self, name, size):
        self.name, self.blockSize = name, size
        self.buffer = bytearray(size)

    def setSubData(self, data: bytes, offset=0):
        self.buffer[offset:offset + len(data)] = data

    def setData(self, data: bytes):
        self.setSubData(data, 0)


class UniformSamplerHIP:
    """UniformSampler (Uniform.h:49-58)."""

    def __init__(self, name, desc):
        self.name, self.desc = name, desc
        self.texture = None

    def setTexture(self, tex: TextureHIP):
        self.texture = tex


@dataclass
class ShaderResources:
    blocks: dict = field(default_factory=dict)
    samplers: dict = field(default_factory=dict)


class RendererHIP:
    """Renderer (Renderer.h:24-59) on one MI355X through librhip."""

    def __init__(self, device: int = 0):
        self.lib = _lib.load()
        self.device = device
        self.dev = None
        self._vao = self._prog = self._res = self._ps = None
        self._batch = None  # views being recorded by beginBatch / endBatch
        self._pass = None
        self._vp = (0, 0, 0, 0)
        # draw() through rhip_draw_noalloc (the arena in the library's static scratch) instead of rhip_draw
        self.no_alloc = False

    def type(self):
        return RENDERER_HIP

    def create(self) -> bool:
        d = ctypes.c_void_p()
        _lib.check(self.lib.rhip_device_create(self.device, ctypes.byref(d)))
        self.dev = d
        return True

    def destroy(self):
        if self.dev:
            self.lib.rhip_device_destroy(self.dev)
            self.dev = None

    # framebuffer
    def createFrameBuffer(self, offscreen=True):
        h = ctypes.c_int()
        _lib.check(self.lib.rhip_fbo_create(self.dev, ctypes.byref(h)))
        return FrameBufferHIP(self, h.value, offscreen)

    # texture
    def createTexture(self, desc: TextureDesc):
        td = _lib.TextureDesc(desc.width, desc.height, desc.type, desc.format, desc.usage, int(desc.useMipmaps),
                              int(desc.multiSample), desc.layout)
        h = ctypes.c_int()
        rc = self.lib.rhip_texture_create(self.dev, ctypes.byref(td), ctypes.byref(h))
        if rc != 0:
            log.error("createTexture failed: %s", self.lib.rhip_last_error().decode())
            return None
        return TextureHIP(self, h.value, desc)

    # vertex
    def createVertexArrayObject(self, va: VertexArray):
        v = np.ascontiguousarray(va.vertexes, np.float32)
        idx = np.ascontiguousarray(va.indices, np.int32).reshape(-1)
        stride = va.vertexSize or v.shape[1] * 4
        h = ctypes.c_int()
        _lib.check(self.lib.rhip_vao_create(self.dev, v.ctypes.data, stride, v.shape[0], idx.ctypes.data, idx.size,
                                            ctypes.byref(h)))
        return VertexArrayObjectHIP(self, h.value, v.nbytes)

    # shader program
    def createShaderProgram(self):
        return ShaderProgramHIP(self)

    # pipeline states
    def createPipelineStates(self, renderStates: S.RenderStates):
        return PipelineStates(renderStates)

    # uniform
    def createUniformBlock(self, name, size):
        return UniformBlockHIP(name, size)

    def createUniformSampler(self, name, desc: TextureDesc):
        return UniformSamplerHIP(name, desc)

    # pipeline
    def beginRenderPass(self, frameBuffer: FrameBufferHIP, states: ClearStates):
        cs = _lib.ClearStates(int(states.depthFlag), int(states.colorFlag),
                              (ctypes.c_float * 4)(*states.clearColor), states.clearDepth)
        if self._batch is not None:
            self._pass = [frameBuffer.handle, cs]
            return
        _lib.check(self.lib.rhip_begin_render_pass(self.dev, frameBuffer.handle, ctypes.byref(cs)))

    def setViewPort(self, x, y, width, height):
        self._vp = (x, y, width, height)
        if self._batch is None:
            _lib.check(self.lib.rhip_set_viewport(self.dev, x, y, width, height))

    def setVertexArrayObject(self, vao):
        self._vao = vao

    def setShaderProgram(self, program):
        self._prog = program

    def setShaderResources(self, resources: ShaderResources):
        if resources is not None:
            self._res = resources

    def setPipelineStates(self, states: PipelineStates):
        self._ps = states

    def draw(self):
        if self._vao is None or self._prog is None or self._prog.handle is None or self._ps is None:
            return
        dd, keep = self._draw_desc()
        if self._batch is not None:
            if self._pass is None:
                raise RuntimeError("batched draw outside a render pass")
            fbo, cs = self._pass
            self._batch.append((_lib.BatchView(fbo, (ctypes.c_int * 4)(*self._vp), cs, dd), keep))
            self._pass[1] = _lib.ClearStates()  # later draws of the pass clear nothing
            return
        fn = self.lib.rhip_draw_noalloc if self.no_alloc else self.lib.rhip_draw
        _lib.check(fn(self.dev, ctypes.byref(dd)))

    def _draw_desc(self):
        rs = self._ps.renderStates
        st = _lib.RhsStates(int(rs.blend), rs.blend_func_rgb, rs.blend_src_rgb, rs.blend_dst_rgb,
                            rs.blend_func_alpha, rs.blend_src_alpha, rs.blend_dst_alpha, int(rs.depth_test),
                            int(rs.depth_mask), rs.depth_func, int(rs.cull_face), rs.primitive_type,
                            rs.polygon_mode, rs.line_width)
        blocks = {}
        samplers = [0] * 8
        if self._res is not None:
            for b in self._res.blocks.values():
                blocks[_BLOCK_ALIASES.get(b.name, b.name)] = bytes(b.buffer)
            for s in self._res.samplers.values():
                if s.texture is not None and s.name in _SLOT_BY_NAME:
                    samplers[_SLOT_BY_NAME[s.name]] = s.texture.handle
        keep = [ctypes.create_string_buffer(blocks.get(n, bytes(sz)), sz) for n, sz in _BLOCK_SIZES.items()]
        dd = _lib.DrawDesc(self._vao.handle, self._prog.handle, st, ctypes.cast(keep[0], ctypes.c_void_p),
                           ctypes.cast(keep[1], ctypes.c_void_p), ctypes.cast(keep[2], ctypes.c_void_p),
                           (ctypes.c_int * 8)(*samplers))
        return dd, keep

    def endRenderPass(self):
        if self._batch is not None:
            self._pass = None
            return
        _lib.check(self.lib.rhip_end_render_pass(self.dev))

    # ------------------------------------------------------------- batched views (no reference counterpart)

    def beginBatch(self):
        """Render passes and draws up to endBatch() are recorded, not run: each draw becomes one view of a batch."""
        self._batch = []
        self._pass = None

    def endBatch(self, tile=None, waves=None) -> "BatchHIP":
        """Creates the batch of the recorded views (one graph of launches over all of them)."""
        views, self._batch, self._pass = self._batch, None, None
        if not views:
            raise RuntimeError("no draws recorded")
        arr = (_lib.BatchView * len(views))(*[v for v, _ in views])
        cfg = ctypes.byref(_lib.BatchConfig(tile, waves)) if tile is not None else None
        h = ctypes.c_int()
        _lib.check(self.lib.rhip_batch_create(self.dev, arr, len(views), cfg, ctypes.byref(h)))
        return BatchHIP(self, h.value, len(views))

    def waitIdle(self):
        _lib.check(self.lib.rhip_wait_idle(self.dev))

    def setHierarchicalZ(self, enabled: bool):
        """Depth-bound culling on (default) or off for later draws (rhip_set_hierarchical_z)."""
        _lib.check(self.lib.rhip_set_hierarchical_z(self.dev, 1 if enabled else 0))

    # measurement helpers (not part of the reference interface)
    def stats(self) -> dict:
        s = _lib.Stats()
        _lib.check(self.lib.rhip_stats_get(self.dev, ctypes.byref(s)))
        return {n: getattr(s, n) for n, _ in _lib.Stats._fields_}

    def resetStats(self):
        _lib.check(self.lib.rhip_stats_reset(self.dev))

    # ------------------------------------------------------------- learned PBR shader (no reference counterpart)

    NN_RECORD_FLOATS = 32

    def neural_init(self, seed: int = 1):
        _lib.check(self.lib.rhip_neural_init(self.dev, seed))

    def neural_set_weights(self, w):
        w = np.ascontiguousarray(w, dtype=np.float32)
        _lib.check(self.lib.rhip_neural_set_weights(self.dev, w.ctypes.data, w.size))

    def neural_get_weights(self):
        w = np.zeros(self.lib.rhip_neural_param_count(), np.float32)
        _lib.check(self.lib.rhip_neural_get_weights(self.dev, w.ctypes.data, w.size))
        return w

    def neural_records(self, capacity: int):
        _lib.check(self.lib.rhip_neural_records(self.dev, capacity))

    def neural_capture(self, enable: bool, shift: int = 0, material: int = 0):
        _lib.check(self.lib.rhip_neural_capture(self.dev, int(enable), shift, material))

    def neural_material(self, material: int):
        """Material (0..63) of the draws that follow, for the per-material fallback."""
        _lib.check(self.lib.rhip_neural_material(self.dev, material))

    def neural_fallback(self, materials):
        """Materials whose learned-shader draws use the plain PBR shader instead."""
        arr = (ctypes.c_int * max(len(materials), 1))(*materials)
        _lib.check(self.lib.rhip_neural_fallback(self.dev, arr, len(materials)))

    def neural_record_count(self) -> int:
        n = ctypes.c_longlong()
        _lib.check(self.lib.rhip_neural_record_count(self.dev, ctypes.byref(n)))
        return n.value

    def neural_read_records(self, first: int, count: int):
        out = np.zeros((count, self.NN_RECORD_FLOATS), np.float32)
        _lib.check(self.lib.rhip_neural_read_records(self.dev, first, count, out.ctypes.data))
        return out

    def neural_write_records(self, first: int, rec):
        rec = np.ascontiguousarray(rec, dtype=np.float32).reshape(-1, self.NN_RECORD_FLOATS)
        _lib.check(self.lib.rhip_neural_write_records(self.dev, first, rec.shape[0], rec.ctypes.data))

    def neural_gradient(self, first: int, count: int):
        g = np.zeros(self.lib.rhip_neural_param_count(), np.float32)
        loss = ctypes.c_double()
        _lib.check(self.lib.rhip_neural_gradient(self.dev, first, count, g.ctypes.data, ctypes.byref(loss)))
        return g, loss.value

    def neural_train(self, first: int, count: int, steps: int, batch: int = 16384, lr_start: float = 3e-3,
                     lr_end: float = 1e-4, seed: int = 1):
        losses = np.zeros(max(steps, 1), np.float64)
        _lib.check(self.lib.rhip_neural_train(self.dev, first, count, steps, batch, lr_start, lr_end, seed,
                                              losses.ctypes.data))
        return losses[:steps]

    def neural_evaluate(self, first: int, count: int, material: int = -1, outputs: bool = False):
        a, b = _lib.NnError(), _lib.NnError()
        out = np.zeros((count, 3), np.float32) if outputs else None
        _lib.check(self.lib.rhip_neural_evaluate(self.dev, first, count, material, ctypes.byref(a), ctypes.byref(b),
                                                 out.ctypes.data if outputs else None))
        err = lambda e: {"mean": e.mean, "p99": e.p99, "max": e.max, "samples": e.samples}
        return err(a), err(b), out

    def profile(self, on: bool = True):
        """Per-kernel HIP-event timing on the renderer's stream (clears previous records)."""
        _lib.check(self.lib.rhip_profile_enable(self.dev, int(on)))

    def profileResults(self) -> dict:
        """{kernel name: (total ms, launches)} recorded since profile(True)."""
        out, i = {}, 0
        name = ctypes.create_string_buffer(128)
        ms, cnt = ctypes.c_double(), ctypes.c_longlong()
        while self.lib.rhip_profile_query(self.dev, i, name, 128, ctypes.byref(ms), ctypes.byref(cnt)) == 0:
            out[name.value.decode()] = (ms.value, cnt.value)
            i += 1
        return out

    def stream(self) -> int:
        return self.lib.rhip_device_stream(self.dev)

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class BatchHIP:
    """A captured batch of views: run() replays it; tune() is the autotuner."""

    def __init__(self, r: RendererHIP, handle: int, n_views: int):
        self.r, self.handle, self.n_views = r, handle, n_views

    def run(self):
        _lib.check(self.r.lib.rhip_batch_run(self.r.dev, self.handle))

    def config(self):
        c = _lib.BatchConfig()
        _lib.check(self.r.lib.rhip_batch_get_config(self.r.dev, self.handle, ctypes.byref(c)))
        return c.tile, c.waves

    def set_config(self, tile, waves):
        _lib.check(self.r.lib.rhip_batch_set_config(self.r.dev, self.handle, ctypes.byref(_lib.BatchConfig(tile, waves))))

    def tune(self, reps=3):
        """Times every (tile, waves) configuration and keeps the fastest; returns
        ((tile, waves), [((tile, waves), ms per batch), ...])."""
        cap = 64
        arr = (_lib.BatchTiming * cap)()
        n = ctypes.c_int()
        best = _lib.BatchConfig()
        _lib.check(self.r.lib.rhip_batch_tune(self.r.dev, self.handle, reps, arr, cap, ctypes.byref(n),
                                              ctypes.byref(best)))
        return (best.tile, best.waves), [((t.config.tile, t.config.waves), t.ms) for t in arr[:min(n.value, cap)]]

    def view_times(self):
        """Per-view completion times (ms from the start) of the last replay."""
        out = np.zeros(self.n_views, np.float64)
        _lib.check(self.r.lib.rhip_batch_view_times(self.r.dev, self.handle, out.ctypes.data, self.n_views))
        return out

    def destroy(self):
        if self.handle and self.r.dev:
            _lib.check(self.r.lib.rhip_batch_destroy(self.r.dev, self.handle))
        self.handle = 0

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass
