#pragma once
#include <memory>
#include "Texture.h"
namespace SoftGL {
struct FrameBufferAttachment {
  std::shared_ptr<Texture> tex = nullptr;
  uint32_t layer = 0;
  uint32_t level = 0;
};
class FrameBuffer {
 public:
  explicit FrameBuffer(bool offscreen) : offscreen_(offscreen) {}
  virtual ~FrameBuffer() = default;
  virtual int getId() const = 0;
  virtual bool isValid() = 0;
  virtual void setColorAttachment(std::shared_ptr<Texture> &color, int level) {
    colorAttachment_ = {color, 0, (uint32_t)level};
    colorReady_ = true;
  }
  virtual void setColorAttachment(std::shared_ptr<Texture> &color, CubeMapFace face, int level) {
    colorAttachment_ = {color, (uint32_t)face, (uint32_t)level};
    colorReady_ = true;
  }
  virtual void setDepthAttachment(std::shared_ptr<Texture> &depth) {
    depthAttachment_ = {depth, 0, 0};
    depthReady_ = true;
  }
  const FrameBufferAttachment &getColorAttachment() const { return colorAttachment_; }
  const FrameBufferAttachment &getDepthAttachment() const { return depthAttachment_; }
  bool isColorReady() const { return colorReady_; }
  bool isDepthReady() const { return depthReady_; }
  bool isMultiSample() const {
    if (colorReady_) return colorAttachment_.tex->multiSample;
    if (depthReady_) return depthAttachment_.tex->multiSample;
    return false;
  }
  bool isOffscreen() const { return offscreen_; }
  void setOffscreen(bool offscreen) { offscreen_ = offscreen; }
 protected:
  bool offscreen_ = false;
  bool colorReady_ = false;
  bool depthReady_ = false;
  FrameBufferAttachment colorAttachment_{};
  FrameBufferAttachment depthAttachment_{};
};
}  // namespace SoftGL
