/* The few glm vector types the render interface names (glm's vec4 / u8vec4 /
 * tvec4 with the member names the interface uses). */
#pragma once
#include <cstdint>
namespace glm {
template <typename T>
struct tvec4 {
  union { T x, r; };
  union { T y, g; };
  union { T z, b; };
  union { T w, a; };
  tvec4() : x(T(0)), y(T(0)), z(T(0)), w(T(0)) {}
  explicit tvec4(T s) : x(s), y(s), z(s), w(s) {}
  tvec4(T x_, T y_, T z_, T w_) : x(x_), y(y_), z(z_), w(w_) {}
};
using vec4 = tvec4<float>;
using u8vec4 = tvec4<uint8_t>;
}  // namespace glm
using RGBA = glm::u8vec4;
