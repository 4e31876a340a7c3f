/*
 * ref_harness.cpp -- TEST INFRASTRUCTURE ONLY (never linked into the product).
 *
 * Replays an rhip_stream.h command stream through SoftGLRender's own
 * RendererSoft, compiled from the reference sources where they lie (see
 * oracle/Makefile), so the reference's output on a seeded scene can be used to
 * pin the CPU restatement (oracle/cpu_ref.cpp) and timed as the CPU baseline.
 *
 * Only the reference's public API is used: Renderer (Renderer.h:24-59),
 * TextureSoft::getImage for read-back (TextureSoft.h:154-156), and the shader
 * extension point ShaderSoft + CREATE_SHADER_OVERRIDE (ShaderSoft.h:51-196) for
 * the two workloads that have no reference shader (a directional-light
 * Blinn-Phong and flat per-vertex colour).
 */
#include <chrono>
#include <map>
#include <algorithm>
#include <vector>

#include "Render/Software/RendererSoft.h"
#include "Render/Software/TextureSoft.h"
#include "Render/Software/UniformSoft.h"
#include "Viewer/Shader/Software/BasicSoft.h"
#include "Viewer/Shader/Software/BlinnPhongSoft.h"
#include "Viewer/Shader/Software/PbrSoft.h"
#include "Viewer/Shader/Software/SkyboxSoft.h"
#include "Viewer/Shader/Software/IBLIrradianceSoft.h"
#include "Viewer/Shader/Software/IBLPrefilterSoft.h"
#include "Viewer/Shader/Software/FxaaSoft.h"

#include "rhip_stream.h"

using namespace SoftGL;

namespace HarnessShaders {

/* Blinn-Phong with a directional light: the reference VS/FS with the light
 * vector replaced by a constant direction (u_pointLightPosition holds the
 * direction towards the light) and no range attenuation. */
class VSDir : public ShaderBlinnPhong::VS {
 public:
  CREATE_SHADER_CLONE(VSDir)
  void shaderMain() override {
    SoftGL::ShaderBlinnPhong::VS::shaderMain();
    v->v_lightDirection = u->u_pointLightPosition;
  }
};

class FSDir : public ShaderBlinnPhong::FS {
 public:
  CREATE_SHADER_CLONE(FSDir)
  void shaderMain() override {
    const float specExp = 128.f;
    glm::vec4 base = def->ALBEDO_MAP ? texture(u->u_albedoMap, v->v_texCoord) : u->u_baseColor;
    glm::vec3 N = GetNormalFromMap();
    float ao = def->AO_MAP ? texture(u->u_aoMap, v->v_texCoord).r : 1.f;
    glm::vec3 amb = glm::vec3(base) * u->u_ambientColor * ao;
    glm::vec3 dif(0.f), spe(0.f), emi(0.f);
    if (u->u_enableLight) {
      glm::vec3 L = glm::normalize(v->v_lightDirection);
      float nl = glm::max(glm::dot(N, L), 0.0f);
      dif = u->u_pointLightColor * glm::vec3(base) * nl;
      glm::vec3 C = glm::normalize(v->v_cameraDirection);
      glm::vec3 H = glm::normalize(L + C);
      float nh = glm::max(glm::dot(N, H), 0.0f);
      spe = u->u_kSpecular * glm::vec3(glm::pow(nh, specExp));
      if (u->u_enableShadow) {
        float s = 1.0f - ShadowCalculation(v->v_shadowFragPos, N);
        dif *= s;
        spe *= s;
      }
    }
    if (def->EMISSIVE_MAP) emi = texture(u->u_emissiveMap, v->v_texCoord);
    gl->FragColor = glm::vec4(amb + dif + spe + emi, base.a);
  }
};

/* Flat per-vertex colour: position + RGBA32F colour, colour passed through. */
namespace VColor {
struct ShaderDefines {};
struct ShaderAttributes {
  glm::vec3 a_position;
  glm::vec4 a_color;
};
struct ShaderUniforms {
  glm::int32_t u_reverseZ;
  glm::mat4 u_modelMatrix;
  glm::mat4 u_modelViewProjectionMatrix;
  glm::mat3 u_inverseTransposeModelMatrix;
  glm::mat4 u_shadowMVPMatrix;
};
struct ShaderVaryings {
  glm::vec4 v_color;
};
class Base : public ShaderSoft {
 public:
  CREATE_SHADER_OVERRIDE
  std::vector<std::string> &getDefines() override {
    static std::vector<std::string> d;
    return d;
  }
  std::vector<UniformDesc> &getUniformsDesc() override {
    static std::vector<UniformDesc> d = {{"UniformsModel", offsetof(ShaderUniforms, u_reverseZ)}};
    return d;
  }
};
class VS : public Base {
 public:
  CREATE_SHADER_CLONE(VS)
  void shaderMain() override {
    gl->Position = u->u_modelViewProjectionMatrix * glm::vec4(a->a_position, 1.0);
    v->v_color = a->a_color;
  }
};
class FS : public Base {
 public:
  CREATE_SHADER_CLONE(FS)
  void shaderMain() override { gl->FragColor = v->v_color; }
};
}  // namespace VColor

}  // namespace HarnessShaders

namespace {

const char *kSlotNames[RHS_SLOT_COUNT] = {"u_albedoMap", "u_normalMap", "u_emissiveMap", "u_aoMap",
                                          "u_shadowMap", "u_metalRoughnessMap", "u_irradianceMap", "u_prefilterMap"};

/* The sampler name a shading model's shader declares for a slot. */
const char *slotName(int shading, int slot) {
  if (shading == RHS_SHADING_FXAA && slot == RHS_SLOT_SCREEN) return "u_screenTexture";
  if (shading == RHS_SHADING_SKYBOX || shading == RHS_SHADING_IBL_IRRADIANCE || shading == RHS_SHADING_IBL_PREFILTER) {
    if (slot == RHS_SLOT_CUBE) return "u_cubeMap";
    if (slot == RHS_SLOT_EQUIRECT) return "u_equirectangularMap";
  }
  return kSlotNames[slot];
}

struct Harness {
  RendererSoft r;
  std::map<int, std::shared_ptr<Texture>> tex;
  std::map<int, std::shared_ptr<VertexArrayObject>> vao;
  std::map<int, std::shared_ptr<FrameBuffer>> fbo;
  std::map<uint64_t, std::shared_ptr<ShaderProgram>> programs;
  std::shared_ptr<UniformBlock> ubModel, ubScene, ubMaterial, ubPrefilter, ubQuadFilter;
  std::map<std::string, std::shared_ptr<UniformSampler>> samplers;
  std::vector<std::shared_ptr<PipelineStates>> keepStates;
  std::string err;

  Harness() {
    ubModel = r.createUniformBlock("UniformsModel", RHS_MODEL_BYTES);
    ubScene = r.createUniformBlock("UniformsScene", RHS_SCENE_BYTES);
    ubMaterial = r.createUniformBlock("UniformsMaterial", RHS_MATERIAL_BYTES);
    /* the IBL prefilter's and the quad filter's blocks (Material.h:88-95) come from the stream's material bytes */
    ubPrefilter = r.createUniformBlock("UniformsPrefilter", 8);
    ubQuadFilter = r.createUniformBlock("UniformsQuadFilter", 8);
  }

  std::shared_ptr<ShaderProgram> program(int shading, uint32_t defines) {
    uint64_t key = ((uint64_t)shading << 32) | defines;
    auto it = programs.find(key);
    if (it != programs.end()) return it->second;
    auto prog = r.createShaderProgram();
    static const char *defNames[6] = {"ALBEDO_MAP", "NORMAL_MAP",         "EMISSIVE_MAP",
                                      "AO_MAP",     "METALROUGHNESS_MAP", "EQUIRECTANGULAR_MAP"};
    for (int i = 0; i < 6; i++)
      if (defines & (1u << i)) prog->addDefine(defNames[i]);
    auto *ps = dynamic_cast<ShaderProgramSoft *>(prog.get());
    bool ok = false;
    switch (shading) {
      case RHS_SHADING_BASIC:
        ok = ps->SetShaders(std::make_shared<ShaderBasic::VS>(), std::make_shared<ShaderBasic::FS>());
        break;
      case RHS_SHADING_BLINN_PHONG:
        ok = ps->SetShaders(std::make_shared<ShaderBlinnPhong::VS>(), std::make_shared<ShaderBlinnPhong::FS>());
        break;
      case RHS_SHADING_PBR:
        ok = ps->SetShaders(std::make_shared<ShaderPbrIBL::VS>(), std::make_shared<ShaderPbrIBL::FS>());
        break;
      case RHS_SHADING_BLINN_PHONG_DIR:
        ok = ps->SetShaders(std::make_shared<HarnessShaders::VSDir>(), std::make_shared<HarnessShaders::FSDir>());
        break;
      case RHS_SHADING_VERTEX_COLOR:
        ok = ps->SetShaders(std::make_shared<HarnessShaders::VColor::VS>(),
                            std::make_shared<HarnessShaders::VColor::FS>());
        break;
      case RHS_SHADING_SKYBOX:
        ok = ps->SetShaders(std::make_shared<ShaderSkybox::VS>(), std::make_shared<ShaderSkybox::FS>());
        break;
      case RHS_SHADING_IBL_IRRADIANCE:
        ok = ps->SetShaders(std::make_shared<ShaderIBLIrradiance::VS>(), std::make_shared<ShaderIBLIrradiance::FS>());
        break;
      case RHS_SHADING_IBL_PREFILTER:
        ok = ps->SetShaders(std::make_shared<ShaderIBLPrefilter::VS>(), std::make_shared<ShaderIBLPrefilter::FS>());
        break;
      case RHS_SHADING_FXAA:
        ok = ps->SetShaders(std::make_shared<ShaderFXAA::VS>(), std::make_shared<ShaderFXAA::FS>());
        break;
      default:
        break;
    }
    if (!ok) return nullptr;
    programs[key] = prog;
    return prog;
  }

  template <typename T>
  static std::shared_ptr<Buffer<T>> fillBuffer(int w, int h, const uint8_t *src) {
    auto b = Buffer<T>::makeDefault(w, h);
    for (int y = 0; y < h; y++)
      for (int x = 0; x < w; x++) {
        T v;
        memcpy(&v, src + ((size_t)y * w + x) * sizeof(T), sizeof(T));
        b->set(x, y, v);
      }
    return b;
  }

  template <typename T>
  static void dumpBuffer(Buffer<T> *b, std::vector<uint8_t> &out) {
    size_t w = b->getWidth(), h = b->getHeight();
    size_t off = out.size();
    out.resize(off + w * h * sizeof(T));
    for (size_t y = 0; y < h; y++)
      for (size_t x = 0; x < w; x++) memcpy(out.data() + off + (y * w + x) * sizeof(T), b->get(x, y), sizeof(T));
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
        auto t = r.createTexture(d);
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
        int layers = (int)p[1];
        const uint8_t *src = (const uint8_t *)(p + 2);
        if (t->format == TextureFormat_RGBA8) {
          std::vector<std::shared_ptr<Buffer<RGBA>>> bufs;
          for (int l = 0; l < layers; l++)
            bufs.push_back(fillBuffer<RGBA>(t->width, t->height, src + (size_t)l * t->width * t->height * 4));
          t->setImageData(bufs);
        } else {
          std::vector<std::shared_ptr<Buffer<float>>> bufs;
          for (int l = 0; l < layers; l++)
            bufs.push_back(fillBuffer<float>(t->width, t->height, src + (size_t)l * t->width * t->height * 4));
          t->setImageData(bufs);
        }
        return 0;
      }
      case RHS_TEX_LEVEL: {
        auto &t = tex.at((int)p[0]);
        int layer = (int)p[1], level = (int)p[2];
        const uint8_t *src = (const uint8_t *)(p + 3);
        if (t->format == TextureFormat_RGBA8) {
          auto *ts = dynamic_cast<TextureSoft<RGBA> *>(t.get());
          auto &img = ts->getImage(layer).getBuffer(level);
          int w = img->width, h = img->height;
          for (int y = 0; y < h; y++)
            for (int x = 0; x < w; x++) {
              RGBA v;
              memcpy(&v, src + ((size_t)y * w + x) * 4, 4);
              img->buffer->set(x, y, v);
            }
        } else {
          auto *ts = dynamic_cast<TextureSoft<float> *>(t.get());
          auto &img = ts->getImage(layer).getBuffer(level);
          int w = img->width, h = img->height;
          for (int y = 0; y < h; y++)
            for (int x = 0; x < w; x++) {
              float v;
              memcpy(&v, src + ((size_t)y * w + x) * 4, 4);
              img->buffer->set(x, y, v);
            }
        }
        return 0;
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
        vao[(int)p[0]] = r.createVertexArrayObject(va);
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
        auto prog = program(d.shading, d.defines);
        if (!prog) {
          err = "unsupported shading";
          return -1;
        }
        ubModel->setData(d.model, RHS_MODEL_BYTES);
        ubScene->setData(d.scene, RHS_SCENE_BYTES);
        ubMaterial->setData(d.material, RHS_MATERIAL_BYTES);
        auto res = std::make_shared<ShaderResources>();
        res->blocks[0] = ubModel;
        res->blocks[1] = ubScene;
        res->blocks[2] = ubMaterial;
        if (d.shading == RHS_SHADING_IBL_PREFILTER) {
          ubPrefilter->setData(d.material, 8);
          res->blocks[3] = ubPrefilter;
        } else if (d.shading == RHS_SHADING_FXAA) {
          ubQuadFilter->setData(d.material, 8);
          res->blocks[4] = ubQuadFilter;
        }
        for (int i = 0; i < d.nsamplers; i++) {
          int slot = d.slot[i];
          auto &t = tex.at(d.tex[i]);
          const std::string name = slotName(d.shading, slot);
          auto &smp = samplers[name];
          if (!smp) {
            TextureDesc td;
            td.type = t->type;
            td.format = t->format;
            smp = r.createUniformSampler(name, td);
          }
          smp->setTexture(t);
          res->samplers[slot] = smp;
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
        keepStates.push_back(ps);
        if (keepStates.size() > 64) keepStates.erase(keepStates.begin());
        r.setVertexArrayObject(vao.at(d.vao));
        r.setShaderProgram(prog);
        r.setShaderResources(res);
        r.setPipelineStates(ps);
        r.draw();
        return 0;
      }
      case RHS_END:
        r.endRenderPass();
        return 0;
      case RHS_READ: {
        auto &t = tex.at((int)p[0]);
        int layer = (int)p[1], level = (int)p[2], which = (int)p[3];
        if (t->format == TextureFormat_RGBA8) {
          auto &img = dynamic_cast<TextureSoft<RGBA> *>(t.get())->getImage(layer).getBuffer(level);
          if (which == 1) dumpBuffer(img->bufferMs4x.get(), out);
          else dumpBuffer(img->buffer.get(), out);
        } else {
          auto &img = dynamic_cast<TextureSoft<float> *>(t.get())->getImage(layer).getBuffer(level);
          if (which == 1) dumpBuffer(img->bufferMs4x.get(), out);
          else dumpBuffer(img->buffer.get(), out);
        }
        return 0;
      }
      case RHS_FRAME:
        return 0;
      default:
        err = "unknown opcode";
        return -1;
    }
  }
};

}  // namespace

extern "C" {

/* Replays `stream`. The part before RHS_FRAME runs once; the part after it
 * (minus RHS_READ) runs `repeats` times and each run's wall time is written to
 * frame_ms[i]; the RHS_READ commands then run once. Returns the number of output
 * bytes, or -1 on error, or -2 if `out_cap` is too small. */
long long ref_replay(const void *stream, size_t bytes, void *out, size_t out_cap, int repeats, double *frame_ms) {
  RhsReader rd;
  if (rhs_open(&rd, stream, bytes)) return -1;
  Harness h;
  std::vector<uint8_t> result;
  struct Cmd {
    int op;
    const uint32_t *p;
    uint32_t n;
  };
  std::vector<Cmd> frame, reads;
  bool inFrame = false;
  const uint32_t *p;
  uint32_t n;
  int op;
  while ((op = rhs_next(&rd, &p, &n)) > 0) {
    if (op == RHS_FRAME) {
      inFrame = true;
      continue;
    }
    if (op == RHS_READ) {
      reads.push_back({op, p, n});
      continue;
    }
    if (inFrame) {
      frame.push_back({op, p, n});
    } else if (h.exec(op, p, n, result)) {
      return -1;
    }
  }
  if (op < 0) return -1;
  if (repeats < 1) repeats = 1;
  for (int i = 0; i < repeats; i++) {
    auto t0 = std::chrono::steady_clock::now();
    for (auto &c : frame)
      if (h.exec(c.op, c.p, c.n, result)) return -1;
    auto t1 = std::chrono::steady_clock::now();
    if (frame_ms) frame_ms[i] = std::chrono::duration<double, std::milli>(t1 - t0).count();
  }
  for (auto &c : reads)
    if (h.exec(c.op, c.p, c.n, result)) return -1;
  if (result.size() > out_cap) return -2;
  memcpy(out, result.data(), result.size());
  return (long long)result.size();
}

}
