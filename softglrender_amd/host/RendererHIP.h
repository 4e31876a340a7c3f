/*
 * RendererHIP -- SoftGLRender's abstract Renderer (src/Render/Renderer.h:24-59)
 * implemented on librhip, the MI355X rasteriser (include/rhip.h).
 *
 * Written against the reference's interface headers: it compiles inside a
 * SoftGLRender checkout (next to src/Render/Software) and, here, against the
 * API-compatible restatement in softglrender_amd/host/compat. A Viewer picks it
 * like the Software / OpenGL / Vulkan renderers; its loadShaders maps a
 * ShadingModel to ShaderProgramHIP::setShading.
 */
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "Base/UUID.h"
#include "Render/Renderer.h"
#include "rhip.h"

namespace SoftGL {

#ifdef SOFTGL_HAS_RENDERER_HIP
constexpr RendererType kRendererTypeHIP = Renderer_HIP;
#else
/* the value a Renderer_HIP entry after Renderer_Vulkan takes (Renderer.h:18-22) */
constexpr RendererType kRendererTypeHIP = static_cast<RendererType>(3);
#endif

/* The device a renderer and all its resources share; freed with the last of them. */
struct RhipDevice {
  rhip_device dev = nullptr;
  ~RhipDevice();
};

class TextureHIP : public Texture {
 public:
  TextureHIP(std::shared_ptr<RhipDevice> dev, const TextureDesc &desc, int handle);
  ~TextureHIP() override;
  int getId() const override { return uuid_.get(); }
  void setSamplerDesc(SamplerDesc &sampler) override;
  void initImageData() override;
  void setImageData(const std::vector<std::shared_ptr<Buffer<RGBA>>> &buffers) override;
  void setImageData(const std::vector<std::shared_ptr<Buffer<float>>> &buffers) override;
  /* Writes the level as binary PPM (RGBA8, alpha dropped) or PFM (FLOAT32). */
  void dumpImage(const char *path, uint32_t layer, uint32_t level) override;

  /* Row-major read-back (the host copy ViewerSoftware::swapBuffer uploads to GL). */
  bool readPixels(uint32_t layer, uint32_t level, bool samples, void *out, size_t bytes);
  /* Overwrites one mip level (IBL chains computed elsewhere). */
  bool setLevelData(uint32_t layer, uint32_t level, const void *data, size_t bytes);
  int handle() const { return handle_; }

 private:
  std::shared_ptr<RhipDevice> dev_;
  int handle_;
  UUID<TextureHIP> uuid_;
};

class VertexArrayObjectHIP : public VertexArrayObject {
 public:
  VertexArrayObjectHIP(std::shared_ptr<RhipDevice> dev, int handle) : dev_(std::move(dev)), handle_(handle) {}
  ~VertexArrayObjectHIP();
  int getId() const override { return uuid_.get(); }
  void updateVertexData(void *data, size_t length) override;
  int handle() const { return handle_; }

 private:
  std::shared_ptr<RhipDevice> dev_;
  int handle_;
  UUID<VertexArrayObjectHIP> uuid_;
};

class FrameBufferHIP : public FrameBuffer {
 public:
  FrameBufferHIP(std::shared_ptr<RhipDevice> dev, int handle, bool offscreen)
      : FrameBuffer(offscreen), dev_(std::move(dev)), handle_(handle) {}
  ~FrameBufferHIP();
  int getId() const override { return uuid_.get(); }
  bool isValid() override { return colorReady_ || depthReady_; }
  void setColorAttachment(std::shared_ptr<Texture> &color, int level) override;
  void setColorAttachment(std::shared_ptr<Texture> &color, CubeMapFace face, int level) override;
  void setDepthAttachment(std::shared_ptr<Texture> &depth) override;
  int handle() const { return handle_; }

 private:
  std::shared_ptr<RhipDevice> dev_;
  int handle_;
  UUID<FrameBufferHIP> uuid_;
};

/* A program is a shading model plus defines; uniform blocks are copied into it
 * when resources are bound (ShaderProgram::bindResources, like the software
 * renderer's memcpy at bind time, ShaderProgramSoft.h:61-70). */
class ShaderProgramHIP : public ShaderProgram {
 public:
  explicit ShaderProgramHIP(std::shared_ptr<RhipDevice> dev) : dev_(std::move(dev)) {}
  int getId() const override { return uuid_.get(); }
  void addDefine(const std::string &def) override;
  /* Selects the kernels (counterpart of ShaderProgramSoft::SetShaders); RhsShading value. */
  bool setShading(int shading);
  int handle() const { return handle_; }

  void setBlock(int location, const std::vector<uint8_t> &bytes);
  void setSampler(int slot, int texture);
  const std::vector<uint8_t> &block(int location) const { return blocks_[location]; }
  int sampler(int slot) const { return samplers_[slot]; }

 private:
  std::shared_ptr<RhipDevice> dev_;
  uint32_t defines_ = 0;
  int handle_ = 0;
  std::vector<uint8_t> blocks_[3];
  int samplers_[RHS_SLOT_COUNT] = {};
  UUID<ShaderProgramHIP> uuid_;
};

class UniformBlockHIP : public UniformBlock {
 public:
  UniformBlockHIP(const std::string &name, int size) : UniformBlock(name, size), buffer_(size) {}
  int getLocation(ShaderProgram &program) override;
  void bindProgram(ShaderProgram &program, int location) override;
  void setSubData(void *data, int len, int offset) override;
  void setData(void *data, int len) override { setSubData(data, len, 0); }

 private:
  std::vector<uint8_t> buffer_;
};

class UniformSamplerHIP : public UniformSampler {
 public:
  UniformSamplerHIP(const std::string &name, TextureType type, TextureFormat format)
      : UniformSampler(name, type, format) {}
  int getLocation(ShaderProgram &program) override;
  void bindProgram(ShaderProgram &program, int location) override;
  void setTexture(const std::shared_ptr<Texture> &tex) override { tex_ = tex; }

 private:
  std::shared_ptr<Texture> tex_;
};

class RendererHIP : public Renderer {
 public:
  explicit RendererHIP(int hipDevice = 0) : hipDevice_(hipDevice) {}
  RendererType type() override { return kRendererTypeHIP; }
  bool create() override;
  void destroy() override;

  std::shared_ptr<FrameBuffer> createFrameBuffer(bool offscreen) override;
  std::shared_ptr<Texture> createTexture(const TextureDesc &desc) override;
  std::shared_ptr<VertexArrayObject> createVertexArrayObject(const VertexArray &vertexArray) override;
  std::shared_ptr<ShaderProgram> createShaderProgram() override;
  std::shared_ptr<PipelineStates> createPipelineStates(const RenderStates &renderStates) override;
  std::shared_ptr<UniformBlock> createUniformBlock(const std::string &name, int size) override;
  std::shared_ptr<UniformSampler> createUniformSampler(const std::string &name, const TextureDesc &desc) override;

  void beginRenderPass(std::shared_ptr<FrameBuffer> &frameBuffer, const ClearStates &states) override;
  void setViewPort(int x, int y, int width, int height) override;
  void setVertexArrayObject(std::shared_ptr<VertexArrayObject> &vao) override;
  void setShaderProgram(std::shared_ptr<ShaderProgram> &program) override;
  void setShaderResources(std::shared_ptr<ShaderResources> &resources) override;
  void setPipelineStates(std::shared_ptr<PipelineStates> &states) override;
  void draw() override;
  void endRenderPass() override;
  void waitIdle() override;

  /* Texel layout given to textures created from now on (RHIP_LAYOUT_*); the
   * reference fixes one layout for all buffers at compile time (Buffer.h:215-227). */
  void setTextureLayout(int layout) { layout_ = layout; }
  rhip_device device() const { return dev_ ? dev_->dev : nullptr; }

 private:
  int hipDevice_;
  int layout_ = RHIP_LAYOUT_LINEAR;
  std::shared_ptr<RhipDevice> dev_;
  VertexArrayObjectHIP *vao_ = nullptr;
  ShaderProgramHIP *program_ = nullptr;
  const RenderStates *states_ = nullptr;
  std::shared_ptr<PipelineStates> statesHold_;
};

}  // namespace SoftGL
