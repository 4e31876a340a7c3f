/*
 * rhip_stream.h -- a flat "command stream" that records a sequence of calls
 * against SoftGLRender's abstract Renderer interface (src/Render/Renderer.h:24-59)
 * so the same scene can be replayed, call for call, through
 *   - RendererHIP (this library, via the C-ABI in rhip.h or the C++ class),
 *   - the CPU restatement under oracle/ (test infrastructure only), and
 *   - SoftGLRender's own RendererSoft (oracle/_ref, built from the reference).
 *
 * Layout: a header {magic, version} followed by commands
 * {u32 opcode, u32 payload_words, payload...}. Every field is a 32-bit
 * little-endian word (u32 / i32 / f32); byte payloads are padded to 4 bytes.
 *
 * The enums carried in the stream use the numeric values of the reference's
 * enums (Texture.h:78-135, RenderStates.h:13-99) so a stream recorded against
 * the reference API replays unchanged.
 */
#ifndef RHIP_STREAM_H
#define RHIP_STREAM_H

#include <stdint.h>
#include <string.h>

#define RHS_MAGIC 0x31534852u /* "RHS1" */
#define RHS_VERSION 1u

enum RhsOp {
  RHS_TEXTURE = 1,   /* id type format w h usage useMipmaps multiSample filterMin filterMag wrapS wrapT wrapR border */
  RHS_TEX_INIT = 2,  /* id                        -> Texture::initImageData()                                  */
  RHS_TEX_DATA = 3,  /* id nlayers bytes...        -> Texture::setImageData(level-0 buffers, one per layer)    */
  RHS_TEX_LEVEL = 4, /* id layer level bytes...    -> overwrite one mip level (pre-filtered IBL chains)        */
  RHS_VAO = 5,       /* id stride nverts nidx vertex-bytes int32-indices                                       */
  RHS_FBO = 6,       /* id colorTex colorLayer colorLevel depthTex   (-1 = no attachment)                      */
  RHS_BEGIN = 7,     /* fbo colorFlag depthFlag clearR clearG clearB clearA clearDepth                          */
  RHS_VIEWPORT = 8,  /* x y w h                                                                                 */
  RHS_DRAW = 9,      /* vao shading defines states[14] model[64] scene[16] material[12] nsamplers (slot tex)*  */
  RHS_END = 10,      /* endRenderPass                                                                           */
  RHS_READ = 11,     /* tex layer level which(0 = single-sample / resolved, 1 = raw 4x samples)                 */
  RHS_FRAME = 12,    /* marks where the per-frame part of the stream starts                                      */
};

/* Shading models. 0..2 are the reference's Shading_BaseColor / BlinnPhong / PBR
 * (Material.h:24-33); 3..5 are workloads with no reference shader:
 * a Blinn-Phong directional light, flat per-vertex colour, and the learned PBR;
 * 6..9 are the reference's Shading_Skybox, Shading_IBL_Irradiance,
 * Shading_IBL_Prefilter and Shading_FXAA (SkyboxSoft.h, IBLIrradianceSoft.h,
 * IBLPrefilterSoft.h, FxaaSoft.h), drawn into targets without depth writes. */
enum RhsShading {
  RHS_SHADING_BASIC = 0,
  RHS_SHADING_BLINN_PHONG = 1,
  RHS_SHADING_PBR = 2,
  RHS_SHADING_BLINN_PHONG_DIR = 3,
  RHS_SHADING_VERTEX_COLOR = 4,
  RHS_SHADING_NEURAL_PBR = 5,
  RHS_SHADING_SKYBOX = 6,         /* u_cubeMap, or u_equirectangularMap with RHS_DEF_EQUIRECTANGULAR_MAP */
  RHS_SHADING_IBL_IRRADIANCE = 7, /* u_cubeMap */
  RHS_SHADING_IBL_PREFILTER = 8,  /* u_cubeMap; material block = UniformsIBLPrefilter {srcResolution, roughness} */
  RHS_SHADING_FXAA = 9,           /* u_screenTexture; material block = UniformsQuadFilter {screenSize} */
};

/* Shader define bits (BlinnPhongSoft.h:14-19, PbrSoft.h:14-20). */
enum RhsDefine {
  RHS_DEF_ALBEDO_MAP = 1,
  RHS_DEF_NORMAL_MAP = 2,
  RHS_DEF_EMISSIVE_MAP = 4,
  RHS_DEF_AO_MAP = 8,
  RHS_DEF_METALROUGHNESS_MAP = 16,
  RHS_DEF_EQUIRECTANGULAR_MAP = 32, /* SkyboxSoft.h:14-16 */
};

/* Sampler slots, in the order the reference shaders declare them
 * (BlinnPhongSoft.h:51-55, PbrSoft.h:52-60). */
enum RhsSlot {
  RHS_SLOT_ALBEDO = 0,
  RHS_SLOT_NORMAL = 1,
  RHS_SLOT_EMISSIVE = 2,
  RHS_SLOT_AO = 3,
  RHS_SLOT_SHADOW = 4,
  RHS_SLOT_METAL_ROUGHNESS = 5,
  RHS_SLOT_IRRADIANCE = 6,
  RHS_SLOT_PREFILTER = 7,
  RHS_SLOT_COUNT = 8,
  /* the skybox, IBL and FXAA shaders' samplers share the slots above:
   * u_equirectangularMap and u_screenTexture in slot 0, u_cubeMap in slot 6 */
  RHS_SLOT_EQUIRECT = RHS_SLOT_ALBEDO,
  RHS_SLOT_SCREEN = RHS_SLOT_ALBEDO,
  RHS_SLOT_CUBE = RHS_SLOT_IRRADIANCE,
};

/* RenderStates (RenderStates.h:52-63) flattened to 14 words. */
typedef struct RhsStates {
  int32_t blend;
  int32_t blendFuncRgb, blendSrcRgb, blendDstRgb;
  int32_t blendFuncAlpha, blendSrcAlpha, blendDstAlpha;
  int32_t depthTest, depthMask, depthFunc;
  int32_t cullFace, primitiveType, polygonMode;
  float lineWidth;
} RhsStates;

/* Uniform block sizes of the reference's std140-like C++ structs with
 * GLM_FORCE_ALIGNED (Material.h:63-95): Model 256 B, Scene 64 B, Material 48 B. */
#define RHS_MODEL_BYTES 256
#define RHS_SCENE_BYTES 64
#define RHS_MATERIAL_BYTES 48

typedef struct RhsDraw {
  int32_t vao;
  int32_t shading;
  uint32_t defines;
  RhsStates states;
  uint8_t model[RHS_MODEL_BYTES];
  uint8_t scene[RHS_SCENE_BYTES];
  uint8_t material[RHS_MATERIAL_BYTES];
  int32_t nsamplers;
  int32_t slot[RHS_SLOT_COUNT];
  int32_t tex[RHS_SLOT_COUNT];
} RhsDraw;

/* A minimal cursor over a stream held in memory. */
typedef struct RhsReader {
  const uint32_t *w;
  size_t n;   /* total words */
  size_t pos; /* current word */
} RhsReader;

static inline int rhs_open(RhsReader *r, const void *data, size_t bytes) {
  r->w = (const uint32_t *)data;
  r->n = bytes / 4;
  r->pos = 0;
  if (r->n < 2 || r->w[0] != RHS_MAGIC || r->w[1] != RHS_VERSION) return -1;
  r->pos = 2;
  return 0;
}

/* Returns the opcode and points *payload / *nwords at the payload, or 0 at the end
 * and -1 on a truncated stream. */
static inline int rhs_next(RhsReader *r, const uint32_t **payload, uint32_t *nwords) {
  if (r->pos >= r->n) return 0;
  if (r->pos + 2 > r->n) return -1;
  uint32_t op = r->w[r->pos], len = r->w[r->pos + 1];
  if (r->pos + 2 + (size_t)len > r->n) return -1;
  *payload = r->w + r->pos + 2;
  *nwords = len;
  r->pos += 2 + (size_t)len;
  return (int)op;
}

static inline float rhs_f32(uint32_t w) {
  float f;
  memcpy(&f, &w, 4);
  return f;
}

/* Decodes a RHS_DRAW payload. Returns 0 on success. */
static inline int rhs_decode_draw(const uint32_t *p, uint32_t n, RhsDraw *d) {
  const uint32_t fixed = 3 + 14 + 64 + 16 + 12 + 1;
  if (n < fixed) return -1;
  d->vao = (int32_t)p[0];
  d->shading = (int32_t)p[1];
  d->defines = p[2];
  memcpy(&d->states, p + 3, 14 * 4);
  memcpy(d->model, p + 17, RHS_MODEL_BYTES);
  memcpy(d->scene, p + 81, RHS_SCENE_BYTES);
  memcpy(d->material, p + 97, RHS_MATERIAL_BYTES);
  d->nsamplers = (int32_t)p[109];
  if (d->nsamplers < 0 || d->nsamplers > RHS_SLOT_COUNT || n < fixed + 2u * (uint32_t)d->nsamplers) return -1;
  for (int i = 0; i < d->nsamplers; i++) {
    d->slot[i] = (int32_t)p[110 + 2 * i];
    d->tex[i] = (int32_t)p[111 + 2 * i];
  }
  return 0;
}

#endif /* RHIP_STREAM_H */
