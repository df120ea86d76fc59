// weldform/matvec.h -- the small vector/matrix types the SPH::Domain facade
// exposes in its public members, with the conventions of the reference's
// External/matvec.h:540-880 (Vec3_t / Mat3_t, row-major (i,j) access, Mult
// as left/right products, Dyad, Trans, OrthoSys::I) and the container
// Array<T> (External/array.h:36).  No blitz++ dependency.
#pragma once
#include <cmath>
#include <cstddef>
#include <stdexcept>
#include <string>
#include <vector>

struct Vec3_t {
  double c[3];
  Vec3_t() : c{0.0, 0.0, 0.0} {}
  Vec3_t(double v) : c{v, v, v} {}  // NOLINT: scalar assignment as in blitz
  Vec3_t(double x, double y, double z) : c{x, y, z} {}
  double &operator()(int i) { return c[i]; }
  double operator()(int i) const { return c[i]; }
  double &operator[](int i) { return c[i]; }
  double operator[](int i) const { return c[i]; }
  Vec3_t &operator+=(const Vec3_t &o) {
    for (int k = 0; k < 3; ++k) c[k] += o.c[k];
    return *this;
  }
  Vec3_t &operator-=(const Vec3_t &o) {
    for (int k = 0; k < 3; ++k) c[k] -= o.c[k];
    return *this;
  }
  Vec3_t &operator*=(double s) {
    for (int k = 0; k < 3; ++k) c[k] *= s;
    return *this;
  }
  Vec3_t &operator/=(double s) {
    for (int k = 0; k < 3; ++k) c[k] /= s;
    return *this;
  }
};
inline Vec3_t operator+(Vec3_t a, const Vec3_t &b) { return a += b; }
inline Vec3_t operator-(Vec3_t a, const Vec3_t &b) { return a -= b; }
inline Vec3_t operator-(Vec3_t a) { return a *= -1.0; }
inline Vec3_t operator*(Vec3_t a, double s) { return a *= s; }
inline Vec3_t operator*(double s, Vec3_t a) { return a *= s; }
inline Vec3_t operator/(Vec3_t a, double s) { return a /= s; }
inline double dot(const Vec3_t &a, const Vec3_t &b) { return a(0) * b(0) + a(1) * b(1) + a(2) * b(2); }
inline double norm(const Vec3_t &a) { return std::sqrt(dot(a, a)); }
inline Vec3_t cross(const Vec3_t &a, const Vec3_t &b) {
  return Vec3_t(a(1) * b(2) - a(2) * b(1), a(2) * b(0) - a(0) * b(2), a(0) * b(1) - a(1) * b(0));
}

struct Mat3_t {
  double m[9];
  Mat3_t() : m{0, 0, 0, 0, 0, 0, 0, 0, 0} {}
  Mat3_t(double v) { for (double &e : m) e = v; }  // NOLINT
  double &operator()(int i, int j) { return m[3 * i + j]; }
  double operator()(int i, int j) const { return m[3 * i + j]; }
};
// element-wise helpers as defined by External/matvec.h:549-575
inline Mat3_t operator*(double a, const Mat3_t &A) {
  Mat3_t M;
  for (int k = 0; k < 9; ++k) M.m[k] = A.m[k] * a;
  return M;
}
inline Mat3_t operator+(const Mat3_t &A, const Mat3_t &B) {
  Mat3_t M;
  for (int k = 0; k < 9; ++k) M.m[k] = A.m[k] + B.m[k];
  return M;
}
inline Mat3_t operator-(const Mat3_t &A, const Mat3_t &B) {
  Mat3_t M;
  for (int k = 0; k < 9; ++k) M.m[k] = A.m[k] - B.m[k];
  return M;
}
inline void set_to_zero(Vec3_t &v) { v = Vec3_t(); }
inline void set_to_zero(Mat3_t &M) { M = Mat3_t(); }
inline void Trans(const Mat3_t &M, Mat3_t &Mt) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) Mt(i, j) = M(j, i);
}
// {B} = {A}^T [M] (left) and {B} = [M]{A} (right), External/matvec.h:829-842
inline void Mult(const Vec3_t &A, const Mat3_t &M, Vec3_t &B) {
  for (int j = 0; j < 3; ++j) B(j) = A(0) * M(0, j) + A(1) * M(1, j) + A(2) * M(2, j);
}
inline void Mult(const Mat3_t &M, const Vec3_t &A, Vec3_t &B) {
  for (int i = 0; i < 3; ++i) B(i) = A(0) * M(i, 0) + A(1) * M(i, 1) + A(2) * M(i, 2);
}
inline void Mult(const Mat3_t &A, const Mat3_t &B, Mat3_t &M) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) M(i, j) = A(i, 2) * B(2, j) + A(i, 1) * B(1, j) + A(i, 0) * B(0, j);
}
inline void Dyad(const Vec3_t &A, const Vec3_t &B, Mat3_t &M) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) M(i, j) = A(i) * B(j);
}

namespace OrthoSys {
inline const Mat3_t &identity() {
  static Mat3_t I = [] {
    Mat3_t M;
    M(0, 0) = M(1, 1) = M(2, 2) = 1.0;
    return M;
  }();
  return I;
}
static const Mat3_t &I = identity();
}  // namespace OrthoSys

// growable array with the reference's Array<T> interface
template <class T>
class Array {
 public:
  size_t Size() const { return v_.size(); }
  void Push(const T &x) { v_.push_back(x); }
  void Resize(size_t n) { v_.resize(n); }
  void Clear() { v_.clear(); }
  T &operator[](size_t i) { return v_[i]; }
  const T &operator[](size_t i) const { return v_[i]; }
  T &Last() { return v_.back(); }
  void DelItem(size_t i) { v_.erase(v_.begin() + (long)i); }
  typename std::vector<T>::iterator begin() { return v_.begin(); }
  typename std::vector<T>::iterator end() { return v_.end(); }

 private:
  std::vector<T> v_;
};

// errors are thrown as Fatal*, like External/fatal.h
class Fatal {
 public:
  explicit Fatal(const std::string &m) : msg_(m) {}
  const std::string &Msg() const { return msg_; }

 private:
  std::string msg_;
};
