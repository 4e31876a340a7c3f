#pragma once
#include "RenderStates.h"
namespace SoftGL {
class PipelineStates {
 public:
  explicit PipelineStates(const RenderStates &states) : renderStates(states) {}
  virtual ~PipelineStates() = default;
  RenderStates renderStates;
};
}  // namespace SoftGL
