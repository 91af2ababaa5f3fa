// vector.cpp -- Vector over the C ABI; one process-wide device context.
#include "vector.h"

#include <cstdio>
#include <cstdlib>
#include <mutex>

#include "lsqt_hip.h"

namespace lsqt_host {

void check_status(int status, const char* call, const char* file, int line)
{
  if (status == LSQT_OK)
    return;
  // the reference's CHECK report (src/common.h:31-42), then exit(1)
  printf("HIP Error:\n");
  printf("    File:       %s\n", file);
  printf("    Line:       %d\n", line);
  printf("    Error code: %d\n", status);
  printf("    Error text: %s: %s (%s)\n", lsqt_strerror(status), lsqt_last_error(), call);
  fflush(stdout);
  exit(1);
}

static lsqt_ctx* g_ctx = nullptr;
static std::once_flag g_once;

lsqt_ctx* default_context()
{
  std::call_once(g_once, [] {
    const char* dev = getenv("LSQT_DEVICE");
    int device = dev ? atoi(dev) : 0;
    LSQT_CHECK(lsqt_ctx_create(device, &g_ctx));
  });
  return g_ctx;
}

void release_default_context()
{
  if (g_ctx) {
    lsqt_ctx_destroy(g_ctx);
    g_ctx = nullptr;
  }
}

}  // namespace lsqt_host

Vector::Vector(int n) : n(n)
{
  LSQT_CHECK(lsqt_vec_create(lsqt_host::default_context(), n, &v));
}

Vector::Vector(Vector& original) : n(original.n)
{
  LSQT_CHECK(lsqt_vec_clone(original.v, &v));
}

Vector::~Vector()
{
  lsqt_vec_destroy(v);
}

void Vector::add(Vector& other)
{
  LSQT_CHECK(lsqt_vec_add(v, other.v));
}

void Vector::copy(Vector& other)
{
  LSQT_CHECK(lsqt_vec_copy(v, other.v));
}

void Vector::apply_sz(Vector& other)
{
  LSQT_CHECK(lsqt_vec_apply_sz(v, other.v));
}

void Vector::copy_from_host(real* other_real, real* other_imag)
{
  LSQT_CHECK(lsqt_vec_from_host(v, other_real, other_imag));
}

void Vector::copy_to_host(real* target_real, real* target_imag)
{
  LSQT_CHECK(lsqt_vec_to_host(v, target_real, target_imag));
}

void Vector::swap(Vector& other)
{
  LSQT_CHECK(lsqt_vec_swap(v, other.v));
}

void Vector::inner_product_1(int number_of_atoms, Vector& other, Vector& target, int offset)
{
  LSQT_CHECK(lsqt_vec_inner_product_1(v, number_of_atoms, other.v, target.v, offset));
}

void Vector::inner_product_2(int number_of_atoms, int number_of_moments, Vector& target)
{
  LSQT_CHECK(lsqt_vec_inner_product_2(v, number_of_atoms, number_of_moments, target.v));
}

real* Vector::device_data() const
{
  void* p = nullptr;
  LSQT_CHECK(lsqt_vec_device_data(v, &p));
  return static_cast<real*>(p);
}
