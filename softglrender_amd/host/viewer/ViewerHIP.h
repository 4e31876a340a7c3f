/*
 * ViewerHIP -- the SoftGLRender Viewer backend that renders with RendererHIP
 * (librhip on an MI355X), beside ViewerSoftware, ViewerOpenGL and ViewerVulkan
 * (src/Viewer/ViewerSoftware.h:21-70 is the backend it stands next to).
 *
 * Installed as src/Viewer/ViewerHIP.h of a SoftGLRender checkout, with
 * softglrender_amd/host and include/ on the include path; viewer.patch adds
 * the Renderer_HIP enum value, registers the viewer in ViewerManager and adds
 * the "HIP" button to the config panel.
 *
 * The frame is rendered on the GPU; swapBuffer reads the resolved colour
 * attachment back (row-major RGBA8) and uploads it to the window's GL texture,
 * as the software viewer uploads its colour buffer. Every shading model the
 * Viewer asks for runs on the GPU, including the skybox, FXAA and the IBL
 * precompute passes. The IBL `.tex` cache stays the software renderer's
 * (Environment.cpp:206-240 only consults it for Renderer_SOFT): the GPU
 * recomputes the maps.
 */
#pragma once

#include <cstdint>
#include <memory>
#include <vector>

#include "Viewer.h"
#include "Render/OpenGL/OpenGLUtils.h"
#include "RendererHIP.h"

namespace SoftGL {
namespace View {

class ViewerHIP : public Viewer {
 public:
  ViewerHIP(Config &config, Camera &camera) : Viewer(config, camera) {}

  void configRenderer() override {
    camera_->setReverseZ(config_.reverseZ);
    cameraDepth_->setReverseZ(config_.reverseZ);
  }

  int swapBuffer() override {
    auto *tex = dynamic_cast<TextureHIP *>(texColorMain_.get());
    if (tex == nullptr) {
      return outTexId_;
    }
    const size_t bytes = (size_t) tex->width * tex->height * 4;
    if (pixels_.size() != bytes) {
      pixels_.resize(bytes);
    }
    // the resolved image (one sample per pixel), waiting for the frame's kernels
    if (!tex->readPixels(0, 0, false, pixels_.data(), bytes)) {
      return outTexId_;
    }
    GL_CHECK(glBindTexture(GL_TEXTURE_2D, outTexId_));
    GL_CHECK(glTexSubImage2D(GL_TEXTURE_2D, 0, 0, 0, (int) tex->width, (int) tex->height, GL_RGBA,
                             GL_UNSIGNED_BYTE, pixels_.data()));
    return outTexId_;
  }

  std::shared_ptr<Renderer> createRenderer() override {
    auto renderer = std::make_shared<RendererHIP>(device_);
    if (!renderer->create()) {
      return nullptr;
    }
    return renderer;
  }

  bool loadShaders(ShaderProgram &program, ShadingModel shading) override {
    auto *programHIP = dynamic_cast<ShaderProgramHIP *>(&program);
    if (programHIP == nullptr) {
      return false;
    }
    // the defines the Viewer added (ALBEDO_MAP ... EQUIRECTANGULAR_MAP) select the kernel specialisation
    switch (shading) {
      case Shading_BaseColor:
        return programHIP->setShading(RHS_SHADING_BASIC);
      case Shading_BlinnPhong:
        return programHIP->setShading(RHS_SHADING_BLINN_PHONG);
      case Shading_PBR:
        return programHIP->setShading(RHS_SHADING_PBR);
      case Shading_Skybox:
        return programHIP->setShading(RHS_SHADING_SKYBOX);
      case Shading_FXAA:
        return programHIP->setShading(RHS_SHADING_FXAA);
      case Shading_IBL_Irradiance:
        return programHIP->setShading(RHS_SHADING_IBL_IRRADIANCE);
      case Shading_IBL_Prefilter:
        return programHIP->setShading(RHS_SHADING_IBL_PREFILTER);
      default:
        break;
    }
    return false;
  }

 private:
  int device_ = 0; /* HIP device ordinal */
  std::vector<uint8_t> pixels_;
};

}
}
