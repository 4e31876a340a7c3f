#pragma once
namespace SoftGL {
/* Per-type increasing object ids (getId() of every resource). */
template <typename T>
class UUID {
 public:
  UUID() : uuid_(counter()++) {}
  int get() const { return uuid_; }
 private:
  static int &counter() { static int c = 0; return c; }
  int uuid_;
};
}  // namespace SoftGL
