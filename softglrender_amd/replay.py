"""Replays an rhip_stream.h command stream through RendererHIP's Renderer
interface, the same calls SoftGLRender's Viewer makes (Viewer.cpp:100-168,
388-396, 497-540)."""
from __future__ import annotations

import struct
import time

import numpy as np

from . import stream as S
from .renderer import (ClearStates, RendererHIP, SamplerDesc, ShaderResources, TextureDesc, VertexArray)

_SLOT_NAMES = ["u_albedoMap", "u_normalMap", "u_emissiveMap", "u_aoMap", "u_shadowMap", "u_metalRoughnessMap",
               "u_irradianceMap", "u_prefilterMap"]
_DEF_NAMES = ["ALBEDO_MAP", "NORMAL_MAP", "EMISSIVE_MAP", "AO_MAP", "METALROUGHNESS_MAP", "EQUIRECTANGULAR_MAP"]


def _slot_name(shading, slot):
    """The sampler name a shader declares for a slot (SkyboxSoft.h, IBL*Soft.h and FxaaSoft.h name theirs)."""
    if shading == S.SHADING_FXAA and slot == S.SLOT_SCREEN:
        return "u_screenTexture"
    if shading in (S.SHADING_SKYBOX, S.SHADING_IBL_IRRADIANCE, S.SHADING_IBL_PREFILTER):
        if slot == S.SLOT_CUBE:
            return "u_cubeMap"
        if slot == S.SLOT_EQUIRECT:
            return "u_equirectangularMap"
    return _SLOT_NAMES[slot]


def parse(data: bytes):
    """Yields (opcode, payload bytes)."""
    magic, ver = struct.unpack_from("<II", data, 0)
    if magic != S.MAGIC or ver != S.VERSION:
        raise ValueError("not an rhip stream")
    pos = 8
    while pos < len(data):
        op, n = struct.unpack_from("<II", data, pos)
        yield op, data[pos + 8:pos + 8 + 4 * n]
        pos += 8 + 4 * n


class Replayer:
    def __init__(self, renderer: RendererHIP, layouts: dict | None = None):
        self.r = renderer
        self.tex, self.vao, self.fbo, self.prog = {}, {}, {}, {}
        self.layouts = layouts or {}
        self.ub = {n: renderer.createUniformBlock(n, sz) for n, sz in
                   (("UniformsModel", S.MODEL_BYTES), ("UniformsScene", S.SCENE_BYTES),
                    ("UniformsMaterial", S.MATERIAL_BYTES))}
        self.samplers = {}

    def exec(self, op, p, reads):
        r = self.r
        if op == S.OP_TEXTURE:
            v = struct.unpack_from("<14i", p)
            tid = v[0]
            desc = TextureDesc(width=v[3], height=v[4], type=v[1], format=v[2], usage=v[5], useMipmaps=bool(v[6]),
                               multiSample=bool(v[7]), layout=self.layouts.get(tid, 0))
            t = r.createTexture(desc)
            if t is None:
                raise RuntimeError("createTexture failed")
            t.setSamplerDesc(SamplerDesc(v[8], v[9], v[10], v[11], v[12], v[13]))
            self.tex[tid] = t
        elif op == S.OP_TEX_INIT:
            self.tex[struct.unpack_from("<i", p)[0]].initImageData()
        elif op == S.OP_TEX_DATA:
            tid, layers = struct.unpack_from("<2i", p)
            t = self.tex[tid]
            per = t.width * t.height * 4
            imgs = []
            for l in range(layers):
                raw = np.frombuffer(p, np.uint8, per, 8 + per * l)
                imgs.append(raw.reshape(t.height, t.width, 4) if t.format == S.FMT_RGBA8 else
                            raw.view(np.float32).reshape(t.height, t.width))
            t.setImageData(imgs)
        elif op == S.OP_TEX_LEVEL:
            tid, layer, level = struct.unpack_from("<3i", p)
            t = self.tex[tid]
            w, h = t.getLevelWidth(level), t.getLevelHeight(level)
            t.setLevelData(layer, level, np.frombuffer(p, np.uint8, w * h * 4, 12))
        elif op == S.OP_VAO:
            vid, stride, nv, ni = struct.unpack_from("<4i", p)
            vb = (nv * stride + 3) // 4 * 4
            verts = np.frombuffer(p, np.float32, nv * stride // 4, 16).reshape(nv, stride // 4)
            idx = np.frombuffer(p, np.int32, ni, 16 + vb)
            self.vao[vid] = r.createVertexArrayObject(VertexArray(stride, verts, idx))
        elif op == S.OP_FBO:
            fid, ct, cl, clev, dt = struct.unpack_from("<5i", p)
            f = r.createFrameBuffer(True)
            if ct >= 0:
                f.setColorAttachment(self.tex[ct], clev, cl)
            if dt >= 0:
                f.setDepthAttachment(self.tex[dt])
            self.fbo[fid] = f
        elif op == S.OP_BEGIN:
            fid, cf, df = struct.unpack_from("<3i", p)
            cc = struct.unpack_from("<5f", p, 12)
            r.beginRenderPass(self.fbo[fid], ClearStates(bool(df), bool(cf), cc[:4], cc[4]))
        elif op == S.OP_VIEWPORT:
            r.setViewPort(*struct.unpack_from("<4i", p))
        elif op == S.OP_DRAW:
            vao, shading, defines = struct.unpack_from("<iiI", p)
            sv = struct.unpack_from("<13if", p, 12)
            rs = S.RenderStates(bool(sv[0]), sv[1], sv[2], sv[3], sv[4], sv[5], sv[6], bool(sv[7]), bool(sv[8]),
                                sv[9], bool(sv[10]), sv[11], sv[12], sv[13])
            key = (shading, defines)
            if key not in self.prog:
                pg = r.createShaderProgram()
                pg.addDefines([n for i, n in enumerate(_DEF_NAMES) if defines & (1 << i)])
                if not pg.setShading(shading):
                    raise RuntimeError(f"unsupported shading {shading}")
                self.prog[key] = pg
            self.ub["UniformsModel"].setData(p[68:68 + 256])
            self.ub["UniformsScene"].setData(p[324:324 + 64])
            self.ub["UniformsMaterial"].setData(p[388:388 + 48])
            ns = struct.unpack_from("<i", p, 436)[0]
            res = ShaderResources(blocks=dict(self.ub))
            for i in range(ns):
                slot, tid = struct.unpack_from("<2i", p, 440 + 8 * i)
                name = _slot_name(shading, slot)
                if name not in self.samplers:
                    self.samplers[name] = r.createUniformSampler(name, TextureDesc())
                self.samplers[name].setTexture(self.tex[tid])
                res.samplers[name] = self.samplers[name]
            # a sampler not named by this draw stays unbound
            res.samplers = {k: v for k, v in res.samplers.items()}
            r.setVertexArrayObject(self.vao[vao])
            r.setShaderProgram(self.prog[key])
            r.setShaderResources(res)
            r.setPipelineStates(r.createPipelineStates(rs))
            r.draw()
        elif op == S.OP_END:
            r.endRenderPass()
        elif op == S.OP_READ:
            tid, layer, level, which = struct.unpack_from("<4i", p)
            reads.append(self.tex[tid].readImage(layer, level, samples=bool(which)))
        elif op == S.OP_FRAME:
            pass
        else:
            raise ValueError(f"unknown opcode {op}")


def replay(renderer: RendererHIP, data: bytes, repeats: int = 1, layouts: dict | None = None):
    """Setup part once, frame part `repeats` times, reads at the end.
    Returns (list of read arrays, list of per-frame wall ms incl. a final waitIdle)."""
    rp = Replayer(renderer, layouts)
    setup, frame, reads_cmds = [], [], []
    in_frame = False
    for op, p in parse(data):
        if op == S.OP_FRAME:
            in_frame = True
        elif op == S.OP_READ:
            reads_cmds.append((op, p))
        elif in_frame:
            frame.append((op, p))
        else:
            setup.append((op, p))
    outs = []
    for op, p in setup:
        rp.exec(op, p, outs)
    renderer.waitIdle()
    times = []
    for _ in range(max(1, repeats)):
        t0 = time.perf_counter()
        for op, p in frame:
            rp.exec(op, p, outs)
        renderer.waitIdle()
        times.append((time.perf_counter() - t0) * 1e3)
    for op, p in reads_cmds:
        rp.exec(op, p, outs)
    return outs, times


def split(data: bytes):
    """(setup, frame, reads) command lists of a stream."""
    setup, frame, reads = [], [], []
    in_frame = False
    for op, p in parse(data):
        if op == S.OP_FRAME:
            in_frame = True
        elif op == S.OP_READ:
            reads.append((op, p))
        elif in_frame:
            frame.append((op, p))
        else:
            setup.append((op, p))
    return setup, frame, reads


def replay_batched(renderer: RendererHIP, data: bytes, layouts: dict | None = None, tile=None, waves=None):
    """Runs the setup part, then records the frame part as one batch (each draw
    one view). Returns (replayer, batch, read commands); the batch has not run yet."""
    rp = Replayer(renderer, layouts)
    setup, frame, reads = split(data)
    sink = []
    for op, p in setup:
        rp.exec(op, p, sink)
    renderer.waitIdle()
    renderer.beginBatch()
    try:
        for op, p in frame:
            rp.exec(op, p, sink)
    except BaseException:
        renderer._batch = None
        raise
    batch = renderer.endBatch(tile, waves)
    batch.keep = rp  # the resources the batch uses live as long as it does
    return rp, batch, reads


def run_reads(rp: Replayer, reads):
    """Executes read commands; returns the arrays they read."""
    outs = []
    for op, p in reads:
        rp.exec(op, p, outs)
    return outs
