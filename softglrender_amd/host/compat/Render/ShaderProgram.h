#pragma once
#include <set>
#include <string>
#include <unordered_map>
#include "Uniform.h"
namespace SoftGL {
class ShaderProgram {
 public:
  virtual ~ShaderProgram() = default;
  virtual int getId() const = 0;
  virtual void addDefine(const std::string &def) = 0;
  virtual void addDefines(const std::set<std::string> &defs) {
    for (auto &d : defs) addDefine(d);
  }
  virtual void bindResources(ShaderResources &resources) {
    for (auto &kv : resources.blocks) bindUniform(*kv.second);
    for (auto &kv : resources.samplers) bindUniform(*kv.second);
  }
 protected:
  /* a uniform's location is looked up once per program and cached by uniform id */
  virtual bool bindUniform(Uniform &uniform) {
    int hash = uniform.getHash(), location;
    auto it = uniformLocations_.find(hash);
    if (it == uniformLocations_.end()) location = uniformLocations_[hash] = uniform.getLocation(*this);
    else location = it->second;
    if (location < 0) return false;
    uniform.bindProgram(*this, location);
    return true;
  }
  std::unordered_map<int, int> uniformLocations_;
};
}  // namespace SoftGL
