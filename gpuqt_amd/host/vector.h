// vector.h -- class Vector with the reference's interface (src/vector.h:23-47),
// holding a complex vector on the GPU through the C ABI (include/lsqt_hip.h).
#pragma once

#include "common.h"

struct lsqt_vec;
struct lsqt_ctx;

namespace lsqt_host {
// the process-wide device context (device 0, or $LSQT_DEVICE)
lsqt_ctx* default_context();
void release_default_context();
}  // namespace lsqt_host

class Vector
{
public:
  Vector(int n);             // zero-filled (src/vector.cu:61-71)
  Vector(Vector& original);  // deep copy (src/vector.cu:92-105)
  ~Vector();

  void add(Vector& other);                                     // this += other
  void copy(Vector& other);                                    // this = other
  void apply_sz(Vector& other);                                // this = (-1)^n other
  void copy_from_host(real* other_real, real* other_imag);     // H2D
  void copy_to_host(real* target_real, real* target_imag);     // D2H (synchronises)
  void swap(Vector& other);                                    // exchanges storage
  // per 512-site block b: target[offset + b] = sum_n this_n conj(other_n)
  // (src/vector.cu:331-344)
  void inner_product_1(int number_of_atoms, Vector& other, Vector& target, int offset);
  // target[m] = sum of moment m's block partials held in this vector
  // (src/vector.cu:425-436)
  void inner_product_2(int number_of_atoms, int number_of_moments, Vector& target);
  // The device storage: n interleaved (re, im) pairs.  Replaces the public
  // split arrays real_part / imag_part of the reference (src/vector.h:39-40);
  // like them it changes owner on swap.
  real* device_data() const;

  int size() const { return n; }
  lsqt_vec* handle() const { return v; }

private:
  Vector& operator=(const Vector&) = delete;
  lsqt_vec* v = nullptr;
  int n = 0;
};
