/*
 * RendererHIP.cpp -- the Renderer interface on librhip's C-ABI. Each method
 * forwards to one rhip_* call; errors are logged and reported the way the
 * software renderer reports them (nullptr / no-op; RendererSoft.cpp:28-34).
 */
#include "RendererHIP.h"

#include <cstdio>
#include <cstring>

namespace SoftGL {

namespace {

void logError(const char *what) { fprintf(stderr, "[RendererHIP] %s failed: %s\n", what, rhip_last_error()); }

bool ok(int rc, const char *what) {
  if (rc == RHIP_OK) return true;
  logError(what);
  return false;
}

/* Block names the software shaders declare (BlinnPhongSoft.h:84-96, PbrSoft.h:89-103). */
int blockLocation(const std::string &name) {
  if (name == "UniformsModel") return 0;
  if (name == "UniformsScene") return 1;
  if (name == "UniformsMaterial") return 2;
  /* the IBL prefilter's and the quad filter's blocks take the material block's place (Material.h:88-95) */
  if (name == "UniformsPrefilter" || name == "UniformsQuadFilter") return 2;
  return -1;
}

int samplerSlot(const std::string &name) {
  static const char *names[RHS_SLOT_COUNT] = {"u_albedoMap", "u_normalMap", "u_emissiveMap", "u_aoMap",
                                              "u_shadowMap", "u_metalRoughnessMap", "u_irradianceMap",
                                              "u_prefilterMap"};
  for (int i = 0; i < RHS_SLOT_COUNT; i++)
    if (name == names[i]) return i;
  /* the skybox, IBL and FXAA shaders (SkyboxSoft.h, IBLIrradianceSoft.h, FxaaSoft.h) */
  if (name == "u_equirectangularMap") return RHS_SLOT_EQUIRECT;
  if (name == "u_screenTexture") return RHS_SLOT_SCREEN;
  if (name == "u_cubeMap") return RHS_SLOT_CUBE;
  return -1;
}

/* Copies one layer image (any Buffer layout) into row-major texels. */
template <typename T>
void packLinear(Buffer<T> &b, std::vector<uint8_t> &out) {
  size_t w = b.getWidth(), h = b.getHeight(), off = out.size();
  out.resize(off + w * h * 4);
  for (size_t y = 0; y < h; y++)
    for (size_t x = 0; x < w; x++) memcpy(out.data() + off + (y * w + x) * 4, b.get(x, y), 4);
}

}  // namespace

RhipDevice::~RhipDevice() {
  if (dev) rhip_device_destroy(dev);
}

/* ------------------------------------------------------------------ textures */

TextureHIP::TextureHIP(std::shared_ptr<RhipDevice> dev, const TextureDesc &desc, int handle)
    : dev_(std::move(dev)), handle_(handle) {
  width = desc.width;
  height = desc.height;
  type = desc.type;
  format = desc.format;
  usage = desc.usage;
  useMipmaps = desc.useMipmaps;
  multiSample = desc.multiSample;
  tag = desc.tag;
}

TextureHIP::~TextureHIP() { rhip_texture_destroy(dev_->dev, handle_); }

void TextureHIP::setSamplerDesc(SamplerDesc &s) {
  rhip_sampler_desc d{(int)s.filterMin, (int)s.filterMag, (int)s.wrapS, (int)s.wrapT, (int)s.wrapR,
                      (int)s.borderColor};
  ok(rhip_texture_set_sampler(dev_->dev, handle_, &d), "setSamplerDesc");
}

void TextureHIP::initImageData() { ok(rhip_texture_init(dev_->dev, handle_), "initImageData"); }

void TextureHIP::setImageData(const std::vector<std::shared_ptr<Buffer<RGBA>>> &buffers) {
  if (format != TextureFormat_RGBA8) {
    fprintf(stderr, "[RendererHIP] setImageData: RGBA data for a FLOAT32 texture\n");
    return;
  }
  std::vector<uint8_t> data;
  for (auto &b : buffers) packLinear(*b, data);
  ok(rhip_texture_upload(dev_->dev, handle_, (int)buffers.size(), data.data(), data.size()), "setImageData");
}

void TextureHIP::setImageData(const std::vector<std::shared_ptr<Buffer<float>>> &buffers) {
  if (format != TextureFormat_FLOAT32) {
    fprintf(stderr, "[RendererHIP] setImageData: float data for an RGBA8 texture\n");
    return;
  }
  std::vector<uint8_t> data;
  for (auto &b : buffers) packLinear(*b, data);
  ok(rhip_texture_upload(dev_->dev, handle_, (int)buffers.size(), data.data(), data.size()), "setImageData");
}

bool TextureHIP::readPixels(uint32_t layer, uint32_t level, bool samples, void *out, size_t bytes) {
  return ok(rhip_texture_read(dev_->dev, handle_, (int)layer, (int)level, samples ? 1 : 0, out, bytes), "readPixels");
}

bool TextureHIP::setLevelData(uint32_t layer, uint32_t level, const void *data, size_t bytes) {
  return ok(rhip_texture_upload_level(dev_->dev, handle_, (int)layer, (int)level, data, bytes), "setLevelData");
}

void TextureHIP::dumpImage(const char *path, uint32_t layer, uint32_t level) {
  if (multiSample) return; /* like TextureSoft::dumpImageSoft (TextureSoft.h:230-233) */
  uint32_t w = getLevelWidth(level), h = getLevelHeight(level);
  std::vector<uint8_t> px((size_t)w * h * 4);
  if (!readPixels(layer, level, false, px.data(), px.size())) return;
  FILE *f = fopen(path, "wb");
  if (!f) {
    fprintf(stderr, "[RendererHIP] dumpImage: cannot open %s\n", path);
    return;
  }
  if (format == TextureFormat_RGBA8) {
    fprintf(f, "P6\n%u %u\n255\n", w, h);
    for (int y = (int)h - 1; y >= 0; y--) /* rows bottom-up, y up (ImageUtils::writeImage flips) */
      for (uint32_t x = 0; x < w; x++) fwrite(&px[((size_t)y * w + x) * 4], 1, 3, f);
  } else {
    fprintf(f, "Pf\n%u %u\n-1.0\n", w, h);
    fwrite(px.data(), 4, (size_t)w * h, f);
  }
  fclose(f);
}

/* ------------------------------------------------------------------ vertex arrays / framebuffers */

VertexArrayObjectHIP::~VertexArrayObjectHIP() { rhip_vao_destroy(dev_->dev, handle_); }

void VertexArrayObjectHIP::updateVertexData(void *data, size_t length) {
  ok(rhip_vao_update(dev_->dev, handle_, data, length), "updateVertexData");
}

FrameBufferHIP::~FrameBufferHIP() { rhip_fbo_destroy(dev_->dev, handle_); }

void FrameBufferHIP::setColorAttachment(std::shared_ptr<Texture> &color, int level) {
  auto *t = dynamic_cast<TextureHIP *>(color.get());
  if (!t || !ok(rhip_fbo_attach_color(dev_->dev, handle_, t->handle(), 0, level), "setColorAttachment")) return;
  FrameBuffer::setColorAttachment(color, level);
}

void FrameBufferHIP::setColorAttachment(std::shared_ptr<Texture> &color, CubeMapFace face, int level) {
  auto *t = dynamic_cast<TextureHIP *>(color.get());
  if (!t || !ok(rhip_fbo_attach_color(dev_->dev, handle_, t->handle(), (int)face, level), "setColorAttachment"))
    return;
  FrameBuffer::setColorAttachment(color, face, level);
}

void FrameBufferHIP::setDepthAttachment(std::shared_ptr<Texture> &depth) {
  auto *t = dynamic_cast<TextureHIP *>(depth.get());
  if (!t || !ok(rhip_fbo_attach_depth(dev_->dev, handle_, t->handle()), "setDepthAttachment")) return;
  FrameBuffer::setDepthAttachment(depth);
}

/* ------------------------------------------------------------------ programs and uniforms */

void ShaderProgramHIP::addDefine(const std::string &def) {
  static const char *names[6] = {"ALBEDO_MAP", "NORMAL_MAP",         "EMISSIVE_MAP",
                                 "AO_MAP",     "METALROUGHNESS_MAP", "EQUIRECTANGULAR_MAP"};
  for (int i = 0; i < 6; i++)
    if (def == names[i]) defines_ |= 1u << i;
}

bool ShaderProgramHIP::setShading(int shading) {
  int h = 0;
  if (!ok(rhip_program_create(dev_->dev, shading, defines_, &h), "setShading")) return false;
  handle_ = h;
  return true;
}

void ShaderProgramHIP::setBlock(int location, const std::vector<uint8_t> &bytes) {
  if (location >= 0 && location < 3) blocks_[location] = bytes;
}

void ShaderProgramHIP::setSampler(int slot, int texture) {
  if (slot >= 0 && slot < RHS_SLOT_COUNT) samplers_[slot] = texture;
}

int UniformBlockHIP::getLocation(ShaderProgram &) { return blockLocation(name); }

void UniformBlockHIP::bindProgram(ShaderProgram &program, int location) {
  if (auto *p = dynamic_cast<ShaderProgramHIP *>(&program)) p->setBlock(location, buffer_);
}

void UniformBlockHIP::setSubData(void *data, int len, int offset) {
  if (offset < 0 || len < 0 || (size_t)(offset + len) > buffer_.size()) {
    fprintf(stderr, "[RendererHIP] uniform block %s: %d bytes at %d exceed %zu\n", name.c_str(), len, offset,
            buffer_.size());
    return;
  }
  memcpy(buffer_.data() + offset, data, len);
}

int UniformSamplerHIP::getLocation(ShaderProgram &) { return samplerSlot(name); }

void UniformSamplerHIP::bindProgram(ShaderProgram &program, int location) {
  auto *p = dynamic_cast<ShaderProgramHIP *>(&program);
  auto *t = dynamic_cast<TextureHIP *>(tex_.get());
  if (p) p->setSampler(location, t ? t->handle() : 0);
}

/* ------------------------------------------------------------------ renderer */

bool RendererHIP::create() {
  rhip_device d = nullptr;
  if (!ok(rhip_device_create(hipDevice_, &d), "create")) return false;
  dev_ = std::make_shared<RhipDevice>();
  dev_->dev = d;
  return true;
}

void RendererHIP::destroy() {
  vao_ = nullptr;
  program_ = nullptr;
  states_ = nullptr;
  statesHold_.reset();
  dev_.reset(); /* the device goes with the last resource that still holds it */
}

std::shared_ptr<FrameBuffer> RendererHIP::createFrameBuffer(bool offscreen) {
  int h = 0;
  if (!dev_ || !ok(rhip_fbo_create(dev_->dev, &h), "createFrameBuffer")) return nullptr;
  return std::make_shared<FrameBufferHIP>(dev_, h, offscreen);
}

std::shared_ptr<Texture> RendererHIP::createTexture(const TextureDesc &desc) {
  rhip_texture_desc d{desc.width, desc.height, (int)desc.type, (int)desc.format, desc.usage,
                      desc.useMipmaps ? 1 : 0, desc.multiSample ? 1 : 0, layout_};
  int h = 0;
  if (!dev_ || !ok(rhip_texture_create(dev_->dev, &d, &h), "createTexture")) return nullptr;
  return std::make_shared<TextureHIP>(dev_, desc, h);
}

std::shared_ptr<VertexArrayObject> RendererHIP::createVertexArrayObject(const VertexArray &va) {
  if (!dev_ || va.vertexesDesc.empty()) return nullptr;
  int stride = (int)va.vertexesDesc[0].stride;
  int h = 0;
  if (!ok(rhip_vao_create(dev_->dev, va.vertexesBuffer, stride, (int)(va.vertexesBufferLength / stride),
                          va.indexBuffer, (int)(va.indexBufferLength / sizeof(int32_t)), &h),
          "createVertexArrayObject"))
    return nullptr;
  return std::make_shared<VertexArrayObjectHIP>(dev_, h);
}

std::shared_ptr<ShaderProgram> RendererHIP::createShaderProgram() {
  if (!dev_) return nullptr;
  return std::make_shared<ShaderProgramHIP>(dev_);
}

std::shared_ptr<PipelineStates> RendererHIP::createPipelineStates(const RenderStates &renderStates) {
  return std::make_shared<PipelineStates>(renderStates);
}

std::shared_ptr<UniformBlock> RendererHIP::createUniformBlock(const std::string &name, int size) {
  return std::make_shared<UniformBlockHIP>(name, size);
}

std::shared_ptr<UniformSampler> RendererHIP::createUniformSampler(const std::string &name, const TextureDesc &desc) {
  return std::make_shared<UniformSamplerHIP>(name, desc.type, desc.format);
}

void RendererHIP::beginRenderPass(std::shared_ptr<FrameBuffer> &frameBuffer, const ClearStates &states) {
  auto *f = dynamic_cast<FrameBufferHIP *>(frameBuffer.get());
  if (!f || !dev_) return;
  rhip_clear_states cs{states.depthFlag ? 1 : 0, states.colorFlag ? 1 : 0,
                       {states.clearColor.r, states.clearColor.g, states.clearColor.b, states.clearColor.a},
                       states.clearDepth};
  ok(rhip_begin_render_pass(dev_->dev, f->handle(), &cs), "beginRenderPass");
}

void RendererHIP::setViewPort(int x, int y, int width, int height) {
  if (dev_) ok(rhip_set_viewport(dev_->dev, x, y, width, height), "setViewPort");
}

void RendererHIP::setVertexArrayObject(std::shared_ptr<VertexArrayObject> &vao) {
  vao_ = dynamic_cast<VertexArrayObjectHIP *>(vao.get());
}

void RendererHIP::setShaderProgram(std::shared_ptr<ShaderProgram> &program) {
  program_ = dynamic_cast<ShaderProgramHIP *>(program.get());
}

void RendererHIP::setShaderResources(std::shared_ptr<ShaderResources> &resources) {
  if (resources && program_) program_->bindResources(*resources);
}

void RendererHIP::setPipelineStates(std::shared_ptr<PipelineStates> &states) {
  statesHold_ = states;
  states_ = states ? &states->renderStates : nullptr;
}

void RendererHIP::draw() {
  if (!dev_ || !vao_ || !program_ || !states_) return; /* RendererSoft::draw (RendererSoft.cpp:137-139) */
  const RenderStates &r = *states_;
  rhip_draw_desc d{};
  d.vao = vao_->handle();
  d.program = program_->handle();
  d.states = RhsStates{r.blend ? 1 : 0,
                       (int)r.blendParams.blendFuncRgb,
                       (int)r.blendParams.blendSrcRgb,
                       (int)r.blendParams.blendDstRgb,
                       (int)r.blendParams.blendFuncAlpha,
                       (int)r.blendParams.blendSrcAlpha,
                       (int)r.blendParams.blendDstAlpha,
                       r.depthTest ? 1 : 0,
                       r.depthMask ? 1 : 0,
                       (int)r.depthFunc,
                       r.cullFace ? 1 : 0,
                       (int)r.primitiveType,
                       (int)r.polygonMode,
                       r.lineWidth};
  uint8_t model[RHS_MODEL_BYTES] = {}, scene[RHS_SCENE_BYTES] = {}, material[RHS_MATERIAL_BYTES] = {};
  const std::vector<uint8_t> &m = program_->block(0), &s = program_->block(1), &t = program_->block(2);
  memcpy(model, m.data(), std::min(m.size(), sizeof model));
  memcpy(scene, s.data(), std::min(s.size(), sizeof scene));
  memcpy(material, t.data(), std::min(t.size(), sizeof material));
  d.uniform_model = model;
  d.uniform_scene = scene;
  d.uniform_material = material;
  for (int i = 0; i < RHS_SLOT_COUNT; i++) d.sampler_tex[i] = program_->sampler(i);
  ok(rhip_draw(dev_->dev, &d), "draw");
}

void RendererHIP::endRenderPass() {
  if (dev_) ok(rhip_end_render_pass(dev_->dev), "endRenderPass");
}

void RendererHIP::waitIdle() {
  if (dev_) ok(rhip_wait_idle(dev_->dev), "waitIdle");
}

}  // namespace SoftGL
