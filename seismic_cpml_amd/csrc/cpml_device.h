// Device helpers shared by the kernel translation units of libcpml_gpu.
#pragma once

#include <hip/hip_runtime.h>

namespace cpml {

constexpr int BX = 64;  // points per wavefront row along x

// Position of a tile's z chunk in its XCD band -> z chunk.  When the chunks
// divide into 8 equal groups, band b holds the chunks c with c % 8 == b, so
// every XCD gets the same share of the (slower) absorbing z-slab planes
// instead of the first and last bands holding all of them.
__device__ __forceinline__ int zchunk_of(int band_pos, int ntz) {
  if (ntz & 7) return band_pos;
  const int m = ntz >> 3;
  return band_pos / m + 8 * (band_pos % m);
}

// d / K.  The reference profiles use K_MAX_PML = 1 (2D2:215), so K is 1 at
// every point and x / 1 == x exactly in IEEE arithmetic: the correctly
// rounded division (a dozen instructions) is only executed where K != 1.
template <typename T>
__device__ __forceinline__ T over_k(T d, T K) {
  return K == T(1) ? d : d / K;
}

// C-PML correction of a derivative: psi <- b*psi + a*d ; d <- d/K + psi
// (e.g. seismic_CPML_2D_isotropic_second_order.f90:567-571).
template <typename T>
__device__ __forceinline__ T cpml(T d, T *psi, T a, T b, T K) {
  const T m = b * *psi + a * d;
  *psi = m;
  return over_k(d, K) + m;
}

// Same recursion with the old value read from one array and the new one
// written to another (or not written when psi_new is null).
template <typename T>
__device__ __forceinline__ T cpml2(T d, const T *psi_old, T *psi_new, T a, T b, T K) {
  const T m = b * *psi_old + a * d;
  if (psi_new) *psi_new = m;
  return over_k(d, K) + m;
}

// The kernel's only argument, read through the kernel-argument segment with a
// pointer the compiler cannot prove loop-invariant.  Kernels whose argument
// block holds more pointers than fit in scalar registers re-read the ones
// they need inside their z loop (scalar loads, cached) instead of keeping
// them live across it (which spills scalar registers into vector lanes).
template <typename A>
__device__ __forceinline__ const __attribute__((address_space(4))) A *kernarg() {
  auto q = (const __attribute__((address_space(4))) A *)__builtin_amdgcn_kernarg_segment_ptr();
  asm volatile("" : "+s"(q));
  return q;
}

// Raw buffer resource over `bytes` bytes at p (gfx9 family word 3).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const void *p, unsigned bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, (int)bytes,
                                           0x00020000);
}

template <typename T>
struct BufOp;

// Cache policy of accesses whose data no other thread touches (a point's own
// old stress, coefficients, results): when the state is much larger than the
// caches, the non-temporal hint (bit 1 of the policy operand on gfx950)
// streams them past the L2, which keeps its capacity for the halo rows and
// planes that neighbouring tiles re-read.  Grids whose state stays resident
// in the caches from one step to the next keep plain accesses (NT = false).
constexpr int kNonTemporal = 2;

template <bool NT, typename T>
__device__ __forceinline__ T ld_once(const T *p) {
  if constexpr (NT) return __builtin_nontemporal_load(p);
  else return *p;
}
template <bool NT, typename T>
__device__ __forceinline__ void st_once(T *p, T x) {
  if constexpr (NT) __builtin_nontemporal_store(x, p);
  else *p = x;
}

template <>
struct BufOp<float> {
  __device__ static float ld(__amdgpu_buffer_rsrc_t r, unsigned v, unsigned s) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, v, s, 0));
  }
  __device__ static void st(float x, __amdgpu_buffer_rsrc_t r, unsigned v, unsigned s) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, x), r, v, s, 0);
  }
  template <int POL>
  __device__ static float ldp(__amdgpu_buffer_rsrc_t r, unsigned v, unsigned s) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, v, s, POL));
  }
  template <int POL>
  __device__ static void stp(float x, __amdgpu_buffer_rsrc_t r, unsigned v, unsigned s) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, x), r, v, s, POL);
  }
};

template <>
struct BufOp<double> {
  typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
  __device__ static double ld(__amdgpu_buffer_rsrc_t r, unsigned v, unsigned s) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, v, s, 0));
  }
  __device__ static void st(double x, __amdgpu_buffer_rsrc_t r, unsigned v, unsigned s) {
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, x), r, v, s, 0);
  }
  template <int POL>
  __device__ static double ldp(__amdgpu_buffer_rsrc_t r, unsigned v, unsigned s) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, v, s, POL));
  }
  template <int POL>
  __device__ static void stp(double x, __amdgpu_buffer_rsrc_t r, unsigned v, unsigned s) {
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, x), r, v, s, POL);
  }
};

}  // namespace cpml
