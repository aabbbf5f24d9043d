// Device helpers shared by the kernel translation units of libcpml_gpu.
#pragma once

#include <hip/hip_runtime.h>

namespace cpml {

constexpr int BX = 64;  // points per wavefront row along x

// C-PML correction of a derivative: psi <- b*psi + a*d ; d <- d/K + psi
// (e.g. seismic_CPML_2D_isotropic_second_order.f90:567-571).
template <typename T>
__device__ __forceinline__ T cpml(T d, T *psi, T a, T b, T K) {
  const T m = b * *psi + a * d;
  *psi = m;
  return d / K + m;
}

// Same recursion with the old value read from one array and the new one
// written to another (or not written when psi_new is null).
template <typename T>
__device__ __forceinline__ T cpml2(T d, const T *psi_old, T *psi_new, T a, T b, T K) {
  const T m = b * *psi_old + a * d;
  if (psi_new) *psi_new = m;
  return d / K + m;
}

// Raw buffer resource over `bytes` bytes at p (gfx9 family word 3).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const void *p, unsigned bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, (int)bytes,
                                           0x00020000);
}

template <typename T>
struct BufOp;

template <>
struct BufOp<float> {
  __device__ static float ld(__amdgpu_buffer_rsrc_t r, unsigned v, unsigned s) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, v, s, 0));
  }
  __device__ static void st(float x, __amdgpu_buffer_rsrc_t r, unsigned v, unsigned s) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, x), r, v, s, 0);
  }
};

template <>
struct BufOp<double> {
  __device__ static double ld(__amdgpu_buffer_rsrc_t r, unsigned v, unsigned s) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, v, s, 0));
  }
  __device__ static void st(double x, __amdgpu_buffer_rsrc_t r, unsigned v, unsigned s) {
    typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, x), r, v, s, 0);
  }
};

}  // namespace cpml
