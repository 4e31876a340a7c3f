/*
 * replay.cpp -- drives RendererHIP through the abstract Renderer interface the
 * way SoftGLRender's Viewer does (Viewer.cpp:100-168 frame, :388-396
 * pipelineDraw, :497-540 texture setup), from an rhip_stream.h command stream.
 *
 * rhip_host_replay() is the C entry point tests and benchmarks use to exercise
 * the C++ class end to end; it needs nothing of SoftGLRender beyond the
 * interface RendererHIP implements.
 */
#include <chrono>
#include <cstring>
#include <map>
#include <vector>

#include "RendererHIP.h"
#include "rhip_stream.h"

using namespace SoftGL;

namespace {

const char *kSlotNames[RHS_SLOT_COUNT] = {"u_albedoMap", "u_normalMap", "u_emissiveMap", "u_aoMap",
                                          "u_shadowMap", "u_metalRoughnessMap", "u_irradianceMap", "u_prefilterMap"};

struct Replayer {
  RendererHIP &r;
  std::map<int, std::shared_ptr<Texture>> tex;
  std::map<int, std::shared_ptr<VertexArrayObject>> vao;
  std::map<int, std::shared_ptr<FrameBuffer>> fbo;
  std::map<uint64_t, std::shared_ptr<ShaderProgram>> programs;
  std::shared_ptr<UniformBlock> ub[3];
  std::shared_ptr<UniformSampler> samplers[RHS_SLOT_COUNT];
  std::map<int, int> layouts;

  explicit Replayer(RendererHIP &rr) : r(rr) {
    ub[0] = r.createUniformBlock("UniformsModel", RHS_MODEL_BYTES);
    ub[1] = r.createUniformBlock("UniformsScene", RHS_SCENE_BYTES);
    ub[2] = r.createUniformBlock("UniformsMaterial", RHS_MATERIAL_BYTES);
  }

  template <typename T>
  static std::vector<std::shared_ptr<Buffer<T>>> images(int w, int h, int layers, const uint8_t *src) {
    std::vector<std::shared_ptr<Buffer<T>>> v;
    for (int l = 0; l < layers; l++) {
      auto b = Buffer<T>::makeDefault(w, h);
      for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
          T t;
          memcpy(&t, src + ((size_t)l * w * h + (size_t)y * w + x) * 4, 4);
          b->set(x, y, t);
        }
      v.push_back(b);
    }
    return v;
  }

  int exec(int op, const uint32_t *p, uint32_t n, std::vector<uint8_t> &out) {
    switch (op) {
      case RHS_TEXTURE: {
        TextureDesc d;
        d.type = (TextureType)p[1];
        d.format = (TextureFormat)p[2];
        d.width = (int)p[3];
        d.height = (int)p[4];
        d.usage = p[5];
        d.useMipmaps = p[6] != 0;
        d.multiSample = p[7] != 0;
        auto it = layouts.find((int)p[0]);
        r.setTextureLayout(it == layouts.end() ? RHIP_LAYOUT_LINEAR : it->second);
        auto t = r.createTexture(d);
        if (!t) return -1;
        SamplerDesc s;
        s.filterMin = (FilterMode)p[8];
        s.filterMag = (FilterMode)p[9];
        s.wrapS = (WrapMode)p[10];
        s.wrapT = (WrapMode)p[11];
        s.wrapR = (WrapMode)p[12];
        s.borderColor = (BorderColor)p[13];
        t->setSamplerDesc(s);
        tex[(int)p[0]] = t;
        return 0;
      }
      case RHS_TEX_INIT:
        tex.at((int)p[0])->initImageData();
        return 0;
      case RHS_TEX_DATA: {
        auto &t = tex.at((int)p[0]);
        if (t->format == TextureFormat_RGBA8)
          t->setImageData(images<RGBA>(t->width, t->height, (int)p[1], (const uint8_t *)(p + 2)));
        else
          t->setImageData(images<float>(t->width, t->height, (int)p[1], (const uint8_t *)(p + 2)));
        return 0;
      }
      case RHS_TEX_LEVEL: {
        auto *t = dynamic_cast<TextureHIP *>(tex.at((int)p[0]).get());
        size_t bytes = (size_t)t->getLevelWidth(p[2]) * t->getLevelHeight(p[2]) * 4;
        return t->setLevelData(p[1], p[2], p + 3, bytes) ? 0 : -1;
      }
      case RHS_VAO: {
        VertexArray va;
        size_t stride = p[1], nv = p[2], ni = p[3];
        va.vertexSize = stride;
        va.vertexesDesc.push_back({stride / 4, stride, 0});
        va.vertexesBuffer = (uint8_t *)(p + 4);
        va.vertexesBufferLength = nv * stride;
        va.indexBuffer = (int32_t *)(p + 4 + (nv * stride + 3) / 4);
        va.indexBufferLength = ni * 4;
        auto v = r.createVertexArrayObject(va);
        if (!v) return -1;
        vao[(int)p[0]] = v;
        return 0;
      }
      case RHS_FBO: {
        auto f = r.createFrameBuffer(true);
        int ct = (int)p[1], dt = (int)p[4];
        if (ct >= 0) {
          auto &t = tex.at(ct);
          if (t->type == TextureType_CUBE)
            f->setColorAttachment(t, (CubeMapFace)p[2], (int)p[3]);
          else
            f->setColorAttachment(t, (int)p[3]);
        }
        if (dt >= 0) f->setDepthAttachment(tex.at(dt));
        fbo[(int)p[0]] = f;
        return 0;
      }
      case RHS_BEGIN: {
        ClearStates cs;
        cs.colorFlag = p[1] != 0;
        cs.depthFlag = p[2] != 0;
        cs.clearColor = glm::vec4(rhs_f32(p[3]), rhs_f32(p[4]), rhs_f32(p[5]), rhs_f32(p[6]));
        cs.clearDepth = rhs_f32(p[7]);
        r.beginRenderPass(fbo.at((int)p[0]), cs);
        return 0;
      }
      case RHS_VIEWPORT:
        r.setViewPort((int)p[0], (int)p[1], (int)p[2], (int)p[3]);
        return 0;
      case RHS_DRAW: {
        RhsDraw d;
        if (rhs_decode_draw(p, n, &d)) return -1;
        uint64_t key = ((uint64_t)d.shading << 32) | d.defines;
        auto it = programs.find(key);
        if (it == programs.end()) {
          auto prog = r.createShaderProgram();
          static const char *defs[5] = {"ALBEDO_MAP", "NORMAL_MAP", "EMISSIVE_MAP", "AO_MAP", "METALROUGHNESS_MAP"};
          for (int i = 0; i < 5; i++)
            if (d.defines & (1u << i)) prog->addDefine(defs[i]);
          /* the Viewer's loadShaders step (ViewerSoftware.h:54-69) */
          if (!dynamic_cast<ShaderProgramHIP *>(prog.get())->setShading(d.shading)) return -1;
          it = programs.emplace(key, prog).first;
        }
        ub[0]->setData(d.model, RHS_MODEL_BYTES);
        ub[1]->setData(d.scene, RHS_SCENE_BYTES);
        ub[2]->setData(d.material, RHS_MATERIAL_BYTES);
        auto res = std::make_shared<ShaderResources>();
        for (int i = 0; i < 3; i++) res->blocks[i] = ub[i];
        for (int i = 0; i < d.nsamplers; i++) {
          int slot = d.slot[i];
          if (slot < 0 || slot >= RHS_SLOT_COUNT) return -1;
          auto &t = tex.at(d.tex[i]);
          if (!samplers[slot]) samplers[slot] = r.createUniformSampler(kSlotNames[slot], *t);
          samplers[slot]->setTexture(t);
          res->samplers[slot] = samplers[slot];
        }
        RenderStates rs;
        rs.blend = d.states.blend != 0;
        rs.blendParams.blendFuncRgb = (BlendFunction)d.states.blendFuncRgb;
        rs.blendParams.blendSrcRgb = (BlendFactor)d.states.blendSrcRgb;
        rs.blendParams.blendDstRgb = (BlendFactor)d.states.blendDstRgb;
        rs.blendParams.blendFuncAlpha = (BlendFunction)d.states.blendFuncAlpha;
        rs.blendParams.blendSrcAlpha = (BlendFactor)d.states.blendSrcAlpha;
        rs.blendParams.blendDstAlpha = (BlendFactor)d.states.blendDstAlpha;
        rs.depthTest = d.states.depthTest != 0;
        rs.depthMask = d.states.depthMask != 0;
        rs.depthFunc = (DepthFunction)d.states.depthFunc;
        rs.cullFace = d.states.cullFace != 0;
        rs.primitiveType = (PrimitiveType)d.states.primitiveType;
        rs.polygonMode = (PolygonMode)d.states.polygonMode;
        rs.lineWidth = d.states.lineWidth;
        auto ps = r.createPipelineStates(rs);
        r.setVertexArrayObject(vao.at(d.vao));
        r.setShaderProgram(it->second);
        r.setShaderResources(res);
        r.setPipelineStates(ps);
        r.draw();
        return 0;
      }
      case RHS_END:
        r.endRenderPass();
        return 0;
      case RHS_READ: {
        auto *t = dynamic_cast<TextureHIP *>(tex.at((int)p[0]).get());
        size_t bytes = (size_t)t->getLevelWidth(p[2]) * t->getLevelHeight(p[2]) * 4 * (p[3] == 1 ? 4 : 1);
        size_t off = out.size();
        out.resize(off + bytes);
        return t->readPixels(p[1], p[2], p[3] == 1, out.data() + off, bytes) ? 0 : -1;
      }
      case RHS_FRAME:
        return 0;
      default:
        return -1;
    }
  }
};

}  // namespace

extern "C" {

/* Replays `stream` through RendererHIP on HIP device `device`: the part before
 * RHS_FRAME once, the frame part `repeats` times (frame_ms[i] = wall time of
 * repeat i including a final waitIdle), then the RHS_READ commands. `layouts`
 * holds (texture id, RHIP_LAYOUT_*) pairs. Returns output bytes or a negative
 * error (-2: out_cap too small). */
long long rhip_host_replay(const void *stream, size_t bytes, void *out, size_t out_cap, int repeats,
                           double *frame_ms, int device, const int *layouts, int nlayouts) {
  RhsReader rd;
  if (rhs_open(&rd, stream, bytes)) return -1;
  RendererHIP r(device);
  if (!r.create()) return -1;
  long long result = -1;
  {
    Replayer rp(r);
    for (int i = 0; i < nlayouts; i++) rp.layouts[layouts[2 * i]] = layouts[2 * i + 1];
    std::vector<uint8_t> outv;
    struct Cmd {
      int op;
      const uint32_t *p;
      uint32_t n;
    };
    std::vector<Cmd> frame, reads;
    bool inFrame = false, failed = false;
    const uint32_t *p;
    uint32_t n;
    int op;
    while ((op = rhs_next(&rd, &p, &n)) > 0 && !failed) {
      if (op == RHS_FRAME) inFrame = true;
      else if (op == RHS_READ) reads.push_back({op, p, n});
      else if (inFrame) frame.push_back({op, p, n});
      else failed = rp.exec(op, p, n, outv) != 0;
    }
    if (!failed && op >= 0) {
      r.waitIdle();
      for (int i = 0; i < (repeats < 1 ? 1 : repeats) && !failed; i++) {
        auto t0 = std::chrono::steady_clock::now();
        for (auto &c : frame)
          if (rp.exec(c.op, c.p, c.n, outv)) failed = true;
        r.waitIdle();
        auto t1 = std::chrono::steady_clock::now();
        if (frame_ms) frame_ms[i] = std::chrono::duration<double, std::milli>(t1 - t0).count();
      }
      for (auto &c : reads)
        if (!failed && rp.exec(c.op, c.p, c.n, outv)) failed = true;
      if (!failed) {
        if (outv.size() > out_cap) result = -2;
        else {
          memcpy(out, outv.data(), outv.size());
          result = (long long)outv.size();
        }
      }
    }
  }
  r.destroy();
  return result;
}
}
