/* A 2D image buffer with the interface of the reference's Buffer<T>
 * (row-major storage here; callers address texels through get / set). */
#pragma once
#include <cstddef>
#include <cstring>
#include <memory>
#include <vector>
#include "GLMInc.h"
namespace SoftGL {
enum BufferLayout { Layout_Linear, Layout_Tiled, Layout_Morton };
template <typename T>
class Buffer {
 public:
  static std::shared_ptr<Buffer<T>> makeDefault(size_t w, size_t h) {
    auto b = std::make_shared<Buffer<T>>();
    b->create(w, h);
    return b;
  }
  virtual ~Buffer() = default;
  void create(size_t w, size_t h, const uint8_t *data = nullptr) {
    width_ = w;
    height_ = h;
    data_.assign(w * h, T());
    if (data) memcpy(data_.data(), data, w * h * sizeof(T));
  }
  virtual BufferLayout getLayout() const { return Layout_Linear; }
  size_t getWidth() const { return width_; }
  size_t getHeight() const { return height_; }
  T *getRawDataPtr() { return data_.data(); }
  size_t getRawDataSize() const { return data_.size(); }
  size_t getRawDataBytesSize() const { return data_.size() * sizeof(T); }
  bool empty() const { return data_.empty(); }
  T *get(size_t x, size_t y) { return (x < width_ && y < height_) ? &data_[x + y * width_] : nullptr; }
  void set(size_t x, size_t y, const T &v) {
    if (x < width_ && y < height_) data_[x + y * width_] = v;
  }
  void setAll(T v) {
    for (auto &e : data_) e = v;
  }
 protected:
  size_t width_ = 0, height_ = 0;
  std::vector<T> data_;
};
}  // namespace SoftGL
