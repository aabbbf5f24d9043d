// Learned thin absorbing boundary for the 3D solver (FP32 fields).
//
// The grid is the interior plus a 4-cell strip on every face, updated by the
// staggered-grid step either without C-PML or with a 4-cell C-PML of its
// own.  After each step a per-face corrector adds a correction to the strip:
//
//   x  = the 9 fields (vx vy vz sxx syy szz sxy sxz syz), each divided by a
//        fixed scale, in a 4-cell-deep box along the face;
//   h1 = W1 * x,  h2 = W2 * h1,  y = W3 * h2     (3x3x3 convolutions, 16
//        channels, zero padding outside the box and outside the grid);
//   u += y * scale on the outermost cells the face owns (each boundary cell
//        belongs to its nearest face, ties to the lowest face index
//        x-,x+,y-,y+,z-,z+).
// The inner strip cells are updated by the plain step alone: their stencils
// stay inside the grid, and what the wide C-PML does differently there (weak
// damping) is four orders of magnitude below the correction the outermost
// cell needs, whose stencil reaches past the grid edge.
//
// The corrector is linear (no activation, no bias): the elastic system is
// linear, so an absorbing boundary that does not depend on the wave amplitude
// or its sign is a linear map of the local state.  Only the outputs needed at
// depth 0 are computed: h1 at d < 3, h2 at d < 2, y at d = 0.
//
// Training (FP32 kernels below) regresses the correction on pairs produced by
// the solver itself: a wide-C-PML "teacher" run gives the state S_n; the
// student grid (teacher cells 16 inward, i.e. the inner 4 cells of a 20-cell
// C-PML become the strip) is loaded with the window of S_n (fields; a thin
// C-PML of the student keeps its own memory variables) and stepped once,
// giving P(S_n); the target is T_{n+1} - P(S_n) on the owned
// outermost cells, with T_{n+1} the teacher's next state.  The loss of a face
// over a mini-batch is sum (y - target)^2 / max(sum target^2, 1% of the
// input energy).
// Inference runs the same network with BF16 weights and activations and FP32
// accumulation on the MFMA units (v_mfma_f32_16x16x32_bf16: two taps of 16
// channels per instruction).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/cpml_gpu.h"
#include "cpml_internal.h"

// Training and inference need no bit-exact association: let the compiler
// contract multiply-adds here.
#pragma clang fp contract(fast)

namespace cpml {
namespace abc {

constexpr int S = 4;            // strip depth (cells)
constexpr int SO = 1;           // depths the corrector writes (the outermost cell)
constexpr int D = SO + 3;       // box depth: corrected cells + receptive field of 3 layers
constexpr int C = 16;           // channel stride (9 fields, 16 hidden)
constexpr int NF = 9;           // fields
constexpr int TAPS = 27;
constexpr int WL = TAPS * C * C;  // weights of one layer of one face: [27][16 in][16 out]
constexpr int WF = 3 * WL;        // one face
constexpr int WN = 6 * WF;        // all faces
constexpr int DMAX1 = SO + 2, DMAX2 = SO + 1, DMAX3 = SO;
constexpr int PAIRS = (TAPS + 1) / 2;  // MFMA K = 32: taps (2p, 2p + 1); tap 27 is zero padding
constexpr int WBN = 6 * 3 * PAIRS * 64 * 8;  // BF16 B-operand fragments of all faces and layers
constexpr int NCHUNK = 64;        // cell chunks of the weight-gradient reduction
constexpr int TB = 256;

struct Face {
  int ax, side, ta, tb, na, nb;
  long off7, off4;  // first cell of this face in the 7-deep / 4-deep buffers
  int tiles, tile0, nta;  // inference tiles (8 x 8 columns) of this face
};

struct Geo {
  int n[3];
  long ld, ps;
  Face f[6];
  long cells7, cells4;
  int max7, max4;  // largest face box (cells), largest strip
  int tiles;
};

struct Fields {
  float *f[NF];
};

struct Scales {
  float s[NF];
};

Geo make_geo(const SolverView &v) {
  Geo g{};
  g.n[0] = v.nx;
  g.n[1] = v.ny;
  g.n[2] = v.nz;
  g.ld = v.ld;
  g.ps = v.ps;
  long o7 = 0, o4 = 0;
  int t0 = 0;
  for (int q = 0; q < 6; ++q) {
    Face &F = g.f[q];
    F.ax = q >> 1;
    F.side = q & 1;
    F.ta = F.ax == 0 ? 1 : 0;
    F.tb = F.ax == 2 ? 1 : 2;
    F.na = g.n[F.ta];
    F.nb = g.n[F.tb];
    F.off7 = o7;
    F.off4 = o4;
    o7 += (long)D * F.na * F.nb;
    o4 += (long)S * F.na * F.nb;
    g.max7 = std::max(g.max7, D * F.na * F.nb);
    g.max4 = std::max(g.max4, S * F.na * F.nb);
    F.nta = (F.na + 7) / 8;
    F.tiles = F.nta * ((F.nb + 7) / 8);
    F.tile0 = t0;
    t0 += F.tiles;
  }
  g.cells7 = o7;
  g.cells4 = o4;
  g.tiles = t0;
  return g;
}

// grid coordinates of face cell (d, a, b)
__device__ __forceinline__ void gcoord(const Geo &g, const Face &F, int d, int a, int b, int c[3]) {
  c[F.ax] = F.side ? g.n[F.ax] - 1 - d : d;
  c[F.ta] = a;
  c[F.tb] = b;
}

__device__ __forceinline__ int owner(const Geo &g, const int c[3]) {
  int best = 0, bd = c[0];
  const int dist[6] = {c[0], g.n[0] - 1 - c[0], c[1], g.n[1] - 1 - c[1], c[2], g.n[2] - 1 - c[2]};
#pragma unroll
  for (int q = 1; q < 6; ++q)
    if (dist[q] < bd) {
      bd = dist[q];
      best = q;
    }
  return best;
}

__device__ __forceinline__ unsigned short to_bf16(float x) {
  unsigned u = __float_as_uint(x);
  u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even (finite inputs)
  return (unsigned short)(u >> 16);
}

__device__ __forceinline__ double block_sum(double v) {
  __shared__ double red[TB / 64];
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0;
  if (threadIdx.x == 0)
    for (int w = 0; w < TB / 64; ++w) s += red[w];
  return s;  // valid in thread 0
}

// ------------------------------------------------------------ training --

// Student box x (normalised), target on the strip and ownership mask, and the
// per-face normaliser sum(x^2).
__global__ __launch_bounds__(TB) void k_gather(Geo g, Fields s, Fields t, long tld, long tps,
                                               long toff, Scales inv, float *X, float *T,
                                               float *M, double *norm, double *tnorm) {
  const int f = blockIdx.y;
  const Face &F = g.f[f];
  const int q = blockIdx.x * TB + threadIdx.x;
  double acc = 0, tacc = 0;
  if (q < D * F.na * F.nb) {
    const int a = q % F.na, b = (q / F.na) % F.nb, d = q / (F.na * F.nb);
    int c[3];
    gcoord(g, F, d, a, b, c);
    const long o = c[0] + c[1] * g.ld + c[2] * g.ps;
    float *x = X + (F.off7 + q) * C;
    float u[NF];
#pragma unroll
    for (int k = 0; k < NF; ++k) {
      u[k] = s.f[k][o];
      x[k] = u[k] * inv.s[k];
      acc += (double)x[k] * x[k];
    }
#pragma unroll
    for (int k = NF; k < C; ++k) x[k] = 0.f;
    if (d < S) {
      const long to = c[0] + c[1] * tld + c[2] * tps + toff;
      float *tt = T + (F.off4 + q) * C;
      const bool own = owner(g, c) == f && d < SO;
#pragma unroll
      for (int k = 0; k < NF; ++k) {
        tt[k] = (t.f[k][to] - u[k]) * inv.s[k];
        if (own) tacc += (double)tt[k] * tt[k];
      }
#pragma unroll
      for (int k = NF; k < C; ++k) tt[k] = 0.f;
      M[F.off4 + q] = own ? 1.f : 0.f;
    }
  }
  const double bs = block_sum(acc);
  __syncthreads();
  const double bt = block_sum(tacc);
  if (threadIdx.x == 0) {
    atomicAdd(norm + f, bs);
    atomicAdd(tnorm + f, bt);
  }
}

// 3x3x3 convolution over every face box, outputs at depth d < dmax (zero
// beyond).  TRANS: the input-gradient of a layer (weights transposed and
// flipped), reading the output-gradient as input.
template <bool TRANS>
__global__ __launch_bounds__(TB) void k_conv(Geo g, const float *in, float *out, const float *W,
                                             int dmax) {
  __shared__ float w[WL];
  const int f = blockIdx.y;
  const Face &F = g.f[f];
  for (int e = threadIdx.x; e < WL; e += TB) w[e] = W[(long)f * WF + e];
  __syncthreads();
  const int q = blockIdx.x * TB + threadIdx.x;
  if (q >= D * F.na * F.nb) return;
  const int a = q % F.na, b = (q / F.na) % F.nb, d = q / (F.na * F.nb);
  float acc[C];
#pragma unroll
  for (int k = 0; k < C; ++k) acc[k] = 0.f;
  if (d < dmax) {
    for (int t = 0; t < TAPS; ++t) {
      const int d2 = d + t / 9 - 1, a2 = a + (t / 3) % 3 - 1, b2 = b + t % 3 - 1;
      if (d2 < 0 || d2 >= D || a2 < 0 || a2 >= F.na || b2 < 0 || b2 >= F.nb) continue;
      const float4 *xp = reinterpret_cast<const float4 *>(
          in + (F.off7 + ((long)d2 * F.nb + b2) * F.na + a2) * C);
      float x[C];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const float4 v = xp[k];
        x[4 * k] = v.x;
        x[4 * k + 1] = v.y;
        x[4 * k + 2] = v.z;
        x[4 * k + 3] = v.w;
      }
      const float *wt = w + (TRANS ? (TAPS - 1 - t) : t) * C * C;
#pragma unroll
      for (int ci = 0; ci < C; ++ci)
#pragma unroll
        for (int co = 0; co < C; ++co)
          acc[co] += x[ci] * (TRANS ? wt[co * C + ci] : wt[ci * C + co]);
    }
  }
  float4 *op = reinterpret_cast<float4 *>(out + (F.off7 + q) * C);
#pragma unroll
  for (int k = 0; k < 4; ++k)
    op[k] = make_float4(acc[4 * k], acc[4 * k + 1], acc[4 * k + 2], acc[4 * k + 3]);
}

// Partial weight gradients: sum over a chunk of output cells (d < dmax) of
// gout[cell][co] * in[cell + tap][ci].  Block (chunk, tap, face), thread (ci, co).
__global__ __launch_bounds__(TB) void k_wgrad(Geo g, const float *in, const float *gout,
                                              float *part, int dmax) {
  __shared__ float sg[64 * C], sx[64 * C];
  const int chunk = blockIdx.x, t = blockIdx.y, f = blockIdx.z;
  const Face &F = g.f[f];
  const long ncell = (long)dmax * F.na * F.nb;
  const long c0 = ncell * chunk / NCHUNK, c1 = ncell * (chunk + 1) / NCHUNK;
  const int td = t / 9 - 1, ta = (t / 3) % 3 - 1, tb = t % 3 - 1;
  const int ci = threadIdx.x >> 4, co = threadIdx.x & 15;
  float acc = 0.f;
  for (long base = c0; base < c1; base += 64) {
    for (int e = threadIdx.x; e < 64 * C; e += TB) {
      const long cell = base + e / C;
      const int ch = e % C;
      float gv = 0.f, xv = 0.f;
      if (cell < c1) {
        gv = gout[(F.off7 + cell) * C + ch];
        const int a = (int)(cell % F.na), b = (int)((cell / F.na) % F.nb),
                  d = (int)(cell / ((long)F.na * F.nb));
        const int d2 = d + td, a2 = a + ta, b2 = b + tb;
        if (d2 >= 0 && d2 < D && a2 >= 0 && a2 < F.na && b2 >= 0 && b2 < F.nb)
          xv = in[(F.off7 + ((long)d2 * F.nb + b2) * F.na + a2) * C + ch];
      }
      sg[e] = gv;
      sx[e] = xv;
    }
    __syncthreads();
#pragma unroll 8
    for (int r = 0; r < 64; ++r) acc += sg[r * C + co] * sx[r * C + ci];
    __syncthreads();
  }
  part[(((long)f * TAPS + t) * NCHUNK + chunk) * (C * C) + threadIdx.x] = acc;
}

// G (+)= sum of the chunk partials; accumulate over the samples of a mini-batch
__global__ __launch_bounds__(TB) void k_wreduce(const float *part, float *G, int layer,
                                                int accumulate) {
  const int ft = blockIdx.x;  // face * 27 + tap
  const int f = ft / TAPS, t = ft % TAPS;
  float s = 0.f;
  for (int k = 0; k < NCHUNK; ++k) s += part[((long)ft * NCHUNK + k) * (C * C) + threadIdx.x];
  float *gp = G + (long)f * WF + layer * WL + t * C * C + threadIdx.x;
  *gp = accumulate ? *gp + s : s;
}

// Squared error of every face over its owned strip cells and the gradient of
// the last layer's output (both unnormalised; the host divides by the target
// energy of the mini-batch).
__global__ __launch_bounds__(TB) void k_loss(Geo g, const float *Y, const float *T, const float *M,
                                             float *G3, double *err) {
  const int f = blockIdx.y;
  const Face &F = g.f[f];
  const int q = blockIdx.x * TB + threadIdx.x;
  double l = 0;
  if (q < D * F.na * F.nb) {
    const int d = q / (F.na * F.nb);
    float *gp = G3 + (F.off7 + q) * C;
    if (d < S) {
      const float m = M[F.off4 + q];
      const float *y = Y + (F.off7 + q) * C;
      const float *tt = T + (F.off4 + q) * C;
#pragma unroll
      for (int k = 0; k < C; ++k) {
        const float diff = k < NF ? (y[k] - tt[k]) * m : 0.f;
        gp[k] = 2.f * diff;
        l += (double)diff * diff;
      }
    } else {
#pragma unroll
      for (int k = 0; k < C; ++k) gp[k] = 0.f;
    }
  }
  const double bs = block_sum(l);
  if (threadIdx.x == 0) atomicAdd(err + f, bs);
}

// Input noise for training: x += nu * rms(box of the face) * N(0, 1) on the 9
// field channels (hash-based normal deviates, Irwin-Hall of four uniforms).
// With the target unchanged this adds nu^2 * rms^2 * |W|^2 to the expected
// loss, so the corrector learns not to amplify what the data never shows.
__device__ __forceinline__ float hash_unit(unsigned x) {
  x ^= x >> 16;
  x *= 0x7feb352dU;
  x ^= x >> 15;
  x *= 0x846ca68bU;
  x ^= x >> 16;
  return (x >> 8) * (1.0f / 16777216.0f);
}

__global__ __launch_bounds__(TB) void k_noise(Geo g, float *X, const double *norm, float nu,
                                              unsigned seed) {
  const int f = blockIdx.y;
  const Face &F = g.f[f];
  const int q = blockIdx.x * TB + threadIdx.x;
  const int n7 = D * F.na * F.nb;
  if (q >= n7) return;
  const float rms = (float)sqrt(norm[f] / (9.0 * n7));
  float *x = X + (F.off7 + q) * C;
  const unsigned base = (seed * 6u + f) * 0x9e3779b9U + (unsigned)q * 9u;
#pragma unroll
  for (int k = 0; k < NF; ++k) {
    const unsigned h = (base + k) * 4u;
    const float z = (hash_unit(h) + hash_unit(h + 1) + hash_unit(h + 2) + hash_unit(h + 3) - 2.f) *
                    1.7320508f;  // unit variance
    x[k] += nu * rms * z;
  }
}

__global__ void k_adam(float *W, const float *G, Scales gscale, float *m, float *v, int n,
                       float lr, float b1, float b2, float eps, float bc1, float bc2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gr = G[i] * gscale.s[i / WF];
  const float mi = b1 * m[i] + (1.f - b1) * gr;
  const float vi = b2 * v[i] + (1.f - b2) * gr * gr;
  m[i] = mi;
  v[i] = vi;
  W[i] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
}

// FP32 master weights -> BF16 MFMA B-operand fragments of the 16x16x32
// instruction, one per tap pair: lane l holds output channel n = l%16 and
// the 8 K values of its lane group g = l/16, i.e. input channels
// 8*(g%2) .. +7 of tap 2p + g/2 (zero for the padding tap 27):
// Wb[face][layer][p][lane][j] = W[face][layer][2p + g/2][8*(g%2) + j][n]
__global__ void k_to_frag(const float *W, unsigned short *Wb) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= WBN) return;
  const int j = e & 7, lane = (e >> 3) & 63, rest = e >> 9;
  const int p = rest % PAIRS, fl = rest / PAIRS;  // fl = face * 3 + layer
  const int g = lane >> 4, tap = 2 * p + (g >> 1), k = 8 * (g & 1) + j, n = lane & 15;
  Wb[e] = tap < TAPS ? to_bf16(W[((long)fl * TAPS + tap) * C * C + k * C + n])
                     : (unsigned short)0;
}

// 7-deep layer-3 output -> planar correction [face][field][4-deep cell]
__global__ __launch_bounds__(TB) void k_y_to_corr(Geo g, const float *Y, float *corr) {
  const int f = blockIdx.y;
  const Face &F = g.f[f];
  const int q = blockIdx.x * TB + threadIdx.x;
  const int n4 = S * F.na * F.nb;
  if (q >= n4) return;
#pragma unroll
  for (int k = 0; k < NF; ++k) corr[NF * F.off4 + (long)k * n4 + q] = Y[(F.off7 + q) * C + k];
}

// ----------------------------------------------------------- inference --

typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int TT = 8;                          // tile: 8 x 8 columns of the face
constexpr int TBI = 512;                       // threads of the inference kernel
constexpr int XA = TT + 6, XN = XA * XA * D;   // input box 14 x 14 x D
constexpr int H1A = TT + 4, H1N = H1A * H1A * (D - 1);
constexpr int H2A = TT + 2, H2N = H2A * H2A * (D - 2);
static_assert(H2N + 1 <= XN + 1, "h2 reuses the input buffer");

// One layer on the MFMA units.  Output region OA x OA x OD (a fastest, then b,
// then d) of the tile, centred like the input region IA = OA + 2; input cells
// at depth -1 (and the padding tap) read the zero cell `izero`.  One
// 16x16x32 MFMA covers a tap pair: lane l of a wave supplies row m = l%16 (an
// output cell) of A with input channels 8*(g%2)..+7 of tap 2p + g/2
// (g = l/16), one 16-byte LDS read, and holds D[m = 4*g + r][n = l%16].
// Each wave keeps GPW groups of 16 output cells in flight, so GPW
// independent MFMA chains hide the MFMA latency.
template <int OA, int OD, bool LAST>
__device__ __forceinline__ void mfma_layer(const unsigned short *in, int izero,
                                           unsigned short *out, const bf8 *__restrict__ wf,
                                           const Face &F, int a0, int b0, float *corr) {
  constexpr int IA = OA + 2;
  constexpr int ON = OA * OA * OD, NG = (ON + 15) / 16;
  constexpr int HALO = (OA - TT) / 2;
  constexpr int NW = TBI / 64, GPW = 4;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  bf8 bw[PAIRS];
#pragma unroll
  for (int p = 0; p < PAIRS; ++p) bw[p] = wf[p * 64 + lane];
  const int hi = lane >> 5;             // second tap of the pair
  const int kq = ((lane >> 4) & 1) * 8;  // channel half
  for (int g0 = wave * GPW; g0 < NG; g0 += NW * GPW) {
    // GPW independent 16-cell groups; a group past the layer's last one
    // recomputes the last group (no branch inside the tap loop) and is not
    // stored
    int base[GPW];
    bool top[GPW];
    f4 acc[GPW];
#pragma unroll
    for (int u = 0; u < GPW; ++u) {
      const int gu = min(g0 + u, NG - 1);
      const int oc = min(gu * 16 + (lane & 15), ON - 1);
      const int al = oc % OA, bl = (oc / OA) % OA, d = oc / (OA * OA);
      base[u] = (d * IA + bl + 1) * IA + al + 1;
      top[u] = d == 0;
      acc[u] = f4{0.f, 0.f, 0.f, 0.f};
    }
    auto a_of = [&](int u, int p) -> bf8 {
      const int t = 2 * p + hi;
      const int off = (t / 9 - 1) * IA * IA + (t % 3 - 1) * IA + ((t / 3) % 3 - 1);
      const int ic = (t >= TAPS || (t < 9 && top[u])) ? izero : base[u] + off;
      return *reinterpret_cast<const bf8 *>(in + ic * C + kq);
    };
    // A operands of pair p + 1 are read from LDS while the MFMAs of pair p run
    bf8 cur[GPW], nxt[GPW];
#pragma unroll
    for (int u = 0; u < GPW; ++u) cur[u] = a_of(u, 0);
#pragma unroll
    for (int p = 0; p < PAIRS; ++p) {
      if (p + 1 < PAIRS) {
#pragma unroll
        for (int u = 0; u < GPW; ++u) nxt[u] = a_of(u, p + 1);
      }
#pragma unroll
      for (int u = 0; u < GPW; ++u)
        acc[u] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(cur[u], bw[p], acc[u], 0, 0, 0);
      if (p + 1 < PAIRS) {
#pragma unroll
        for (int u = 0; u < GPW; ++u) cur[u] = nxt[u];
      }
    }
    const int n = lane & 15;
#pragma unroll
    for (int u = 0; u < GPW; ++u) {
      if (g0 + u >= NG) break;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = (g0 + u) * 16 + 4 * (lane >> 4) + r;
        if (m >= ON) continue;
        const int ml = m % OA, mb = (m / OA) % OA, md = m / (OA * OA);
        const int a = a0 - HALO + ml, b = b0 - HALO + mb;
        const bool inside = a >= 0 && a < F.na && b >= 0 && b < F.nb;
        if (!LAST) {
          out[m * C + n] = inside ? to_bf16(acc[u][r]) : (unsigned short)0;
        } else if (inside && n < NF) {
          const long n4 = (long)S * F.na * F.nb;
          corr[NF * F.off4 + n * n4 + ((long)md * F.nb + b) * F.na + a] = acc[u][r];
        }
      }
    }
  }
}

__global__ __launch_bounds__(TBI) void k_infer(Geo g, Fields u, Scales inv,
                                              const unsigned short *__restrict__ Wb,
                                              float *corr) {
  __shared__ __attribute__((aligned(16))) unsigned short bx[(XN + 1) * C];
  __shared__ __attribute__((aligned(16))) unsigned short bh1[(H1N + 1) * C];
  int f = 0;
#pragma unroll
  for (int q = 1; q < 6; ++q)
    if ((int)blockIdx.x >= g.f[q].tile0) f = q;
  const Face &F = g.f[f];
  const int tile = blockIdx.x - F.tile0;
  const int a0 = (tile % F.nta) * TT, b0 = (tile / F.nta) * TT;

  // input box: normalised fields in BF16, channels 9..15 zero
  // x faces: depth runs along x, the contiguous direction, so the lanes walk
  // the box depth first; y and z faces walk a (= x) first
  const bool xface = F.ax == 0;
  for (int q = threadIdx.x; q < XN; q += TBI) {
    const int al = xface ? (q / D) % XA : q % XA;
    const int bl = xface ? q / (D * XA) : (q / XA) % XA;
    const int d = xface ? q % D : q / (XA * XA);
    const int a = a0 - 3 + al, b = b0 - 3 + bl;
    unsigned pk[C / 2] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (a >= 0 && a < F.na && b >= 0 && b < F.nb) {
      int c[3];
      gcoord(g, F, d, a, b, c);
      const long o = c[0] + c[1] * g.ld + c[2] * g.ps;
#pragma unroll
      for (int k = 0; k < NF; ++k) {
        const unsigned h = to_bf16(u.f[k][o] * inv.s[k]);
        pk[k >> 1] |= (k & 1) ? h << 16 : h;
      }
    }
    uint4 *dst = reinterpret_cast<uint4 *>(bx + ((d * XA + bl) * XA + al) * C);
    dst[0] = make_uint4(pk[0], pk[1], pk[2], pk[3]);
    dst[1] = make_uint4(pk[4], pk[5], pk[6], pk[7]);
  }
  if (threadIdx.x < C) {
    bx[XN * C + threadIdx.x] = 0;
    bh1[H1N * C + threadIdx.x] = 0;
  }
  __syncthreads();
  const bf8 *wf = reinterpret_cast<const bf8 *>(Wb) + (long)f * 3 * PAIRS * 64;
  mfma_layer<H1A, D - 1, false>(bx, XN, bh1, wf, F, a0, b0, nullptr);
  __syncthreads();
  unsigned short *bh2 = bx;  // the input box is dead now
  if (threadIdx.x < C) bh2[H2N * C + threadIdx.x] = 0;
  mfma_layer<H2A, D - 2, false>(bh1, H1N, bh2, wf + PAIRS * 64, F, a0, b0, nullptr);
  __syncthreads();
  mfma_layer<TT, SO, true>(bh2, H2N, nullptr, wf + 2 * PAIRS * 64, F, a0, b0, corr);
}

// u += correction * scale on the strip cells each face owns
__global__ __launch_bounds__(TB) void k_apply(Geo g, Fields u, Scales sc, const float *corr) {
  const int f = blockIdx.y;
  const Face &F = g.f[f];
  const int q = blockIdx.x * TB + threadIdx.x;
  const int n4 = S * F.na * F.nb;
  if (q >= SO * F.na * F.nb) return;
  // x faces walk the (contiguous) depth first
  const bool xface = F.ax == 0;
  const int a = xface ? (q / SO) % F.na : q % F.na;
  const int b = xface ? q / (SO * F.na) : (q / F.na) % F.nb;
  const int d = xface ? q % SO : q / (F.na * F.nb);
  const int cell = (d * F.nb + b) * F.na + a;
  int c[3];
  gcoord(g, F, d, a, b, c);
  if (owner(g, c) != f) return;
  const long o = c[0] + c[1] * g.ld + c[2] * g.ps;
#pragma unroll
  for (int k = 0; k < NF; ++k) u.f[k][o] += corr[NF * F.off4 + (long)k * n4 + cell] * sc.s[k];
}

// window copy: dst interior (i,j,k) <- src (i+i0, j+j0, k+k0), all 9 fields
__global__ __launch_bounds__(TB) void k_window(Fields dst, int nx, int ny, int nz, long ld,
                                               long ps, Fields src, long sld, long sps,
                                               long soff) {
  const long q = (long)blockIdx.x * TB + threadIdx.x;
  if (q >= (long)nx * ny * nz) return;
  const int i = (int)(q % nx), j = (int)((q / nx) % ny), k = (int)(q / ((long)nx * ny));
  const long o = i + j * ld + k * ps, so = i + j * sld + k * sps + soff;
#pragma unroll
  for (int c = 0; c < NF; ++c) dst.f[c][o] = src.f[c][so];
}

Fields fields_of(const SolverView &v) {
  Fields F{};
  for (int k = 0; k < NF; ++k) F.f[k] = static_cast<float *>(v.f[k]);
  return F;
}

}  // namespace abc
}  // namespace cpml

using namespace cpml::abc;

struct cpml_abc_net {
  int device = 0;
  float *W = nullptr;             // FP32 master weights [6][3][27][16][16]
  unsigned short *Wb = nullptr;   // BF16 fragments
  bool wb_dirty = true;
  Scales scale{}, inv{};
  // training state
  float *G = nullptr, *m = nullptr, *v = nullptr;
  int t = 0;
  double lr = 1e-3, b1 = 0.9, b2 = 0.999, eps = 1e-8;
  float noise = 0.f;  // relative input noise during training
  unsigned samples = 0;
  // work buffers, sized for the largest geometry seen
  long cap7 = 0, cap4 = 0;
  float *X = nullptr, *H1 = nullptr, *H2 = nullptr, *Y = nullptr, *G1 = nullptr, *G2 = nullptr,
        *G3 = nullptr, *T = nullptr, *M = nullptr, *part = nullptr, *corr = nullptr;
  double *red = nullptr;  // per face: box energy [0,6), squared error [6,12), target energy [12,18)
  int batch = 1, acc_count = 0;      // mini-batch: samples per Adam step, samples so far
  double acc_tn[6] = {0, 0, 0, 0, 0, 0};  // per face: target energy of the batch
  double acc_xn[6] = {0, 0, 0, 0, 0, 0};  // per face: input (box) energy of the batch
  std::vector<void *> allocs;
  std::string err;
};

namespace {

thread_local std::string g_abc_error;

int afail(cpml_abc_net *n, int code, const std::string &m) {
  if (n)
    n->err = m;
  else
    g_abc_error = m;
  return code;
}

#define ACHK(n, call)                                                                     \
  do {                                                                                    \
    hipError_t e_ = (call);                                                               \
    if (e_ != hipSuccess)                                                                 \
      return afail((n), CPML_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

int aalloc(cpml_abc_net *n, void **p, size_t bytes) {
  hipError_t e = hipMalloc(p, std::max<size_t>(bytes, 256));
  if (e != hipSuccess) {
    *p = nullptr;
    return afail(n, e == hipErrorOutOfMemory ? CPML_ERR_OOM : CPML_ERR_HIP,
                 std::string("hipMalloc: ") + hipGetErrorString(e));
  }
  n->allocs.push_back(*p);
  ACHK(n, hipMemset(*p, 0, std::max<size_t>(bytes, 256)));
  return CPML_OK;
}

void afree(cpml_abc_net *n, void *p) {
  if (!p) return;
  auto it = std::find(n->allocs.begin(), n->allocs.end(), p);
  if (it != n->allocs.end()) n->allocs.erase(it);
  (void)hipFree(p);
}

int ensure_buffers(cpml_abc_net *n, const Geo &g) {
  if (g.cells7 <= n->cap7 && g.cells4 <= n->cap4) return CPML_OK;
  float **b7[7] = {&n->X, &n->H1, &n->H2, &n->Y, &n->G1, &n->G2, &n->G3};
  for (auto p : b7) {
    afree(n, *p);
    *p = nullptr;
  }
  for (float **p : {&n->T, &n->M, &n->corr}) {
    afree(n, *p);
    *p = nullptr;
  }
  int rc;
  for (auto p : b7)
    if ((rc = aalloc(n, (void **)p, (size_t)g.cells7 * C * sizeof(float)))) return rc;
  if ((rc = aalloc(n, (void **)&n->T, (size_t)g.cells4 * C * sizeof(float)))) return rc;
  if ((rc = aalloc(n, (void **)&n->M, (size_t)g.cells4 * sizeof(float)))) return rc;
  if ((rc = aalloc(n, (void **)&n->corr, (size_t)g.cells4 * NF * sizeof(float)))) return rc;
  n->cap7 = g.cells7;
  n->cap4 = g.cells4;
  return CPML_OK;
}

int check_view(cpml_abc_net *n, const cpml::SolverView &v, const char *who) {
  if (v.ndim != 3 || v.prec != 4 || v.order != 2)
    return afail(n, CPML_ERR_ARG, std::string(who) +
                                      ": the learned boundary needs a 3D FP32 2nd-order handle");
  if (std::min(v.nx, std::min(v.ny, v.nz)) < 2 * D)
    return afail(n, CPML_ERR_STENCIL, std::string(who) + ": grid smaller than two strip boxes");
  return CPML_OK;
}

int sync_bf16(cpml_abc_net *n, hipStream_t s) {
  if (!n->wb_dirty) return CPML_OK;
  hipLaunchKernelGGL(k_to_frag, dim3((WBN + 255) / 256), dim3(256), 0, s, n->W, n->Wb);
  ACHK(n, hipGetLastError());
  n->wb_dirty = false;
  return CPML_OK;
}

// Forward pass (FP32) of the box already in n->X.
hipError_t forward32(cpml_abc_net *n, const Geo &g, hipStream_t s) {
  const dim3 grid((g.max7 + TB - 1) / TB, 6);
  hipLaunchKernelGGL(k_conv<false>, grid, dim3(TB), 0, s, g, n->X, n->H1, n->W, DMAX1);
  hipLaunchKernelGGL(k_conv<false>, grid, dim3(TB), 0, s, g, n->H1, n->H2, n->W + WL, DMAX2);
  hipLaunchKernelGGL(k_conv<false>, grid, dim3(TB), 0, s, g, n->H2, n->Y, n->W + 2 * WL, DMAX3);
  return hipGetLastError();
}

// 1 / loss denominator: target energy, floored at 1% of the input energy;
// 0 when both vanish (no wave in the box yet).
double loss_scale(double target_energy, double input_energy) {
  const double d = std::max(target_energy, 1e-2 * input_energy);
  return d > 1e-30 ? 1.0 / d : 0.0;
}

// gather + forward + loss + backward of one sample into n->G (added to it when
// accumulate); per-face squared error and owned target energy into err6, tn6
int grad32(cpml_abc_net *n, const cpml::SolverView &st, const cpml::SolverView &te, int i0,
           int j0, int k0, int accumulate, double *err6, double *tn6, double *xn6 = nullptr,
           float noise = 0.f) {
  const Geo g = make_geo(st);
  int rc = ensure_buffers(n, g);
  if (rc) return rc;
  if (!n->G) return afail(n, CPML_ERR_STATE, "call cpml_abc_train_init first");
  if (i0 < 0 || j0 < 0 || k0 < 0 || i0 + st.nx > te.nx || j0 + st.ny > te.ny ||
      k0 + st.nz > te.nz)
    return afail(n, CPML_ERR_ARG, "student window outside the teacher grid");
  hipStream_t s = st.stream;
  ACHK(n, hipMemsetAsync(n->red, 0, 18 * sizeof(double), s));
  const dim3 grid((g.max7 + TB - 1) / TB, 6);
  const long toff = i0 + j0 * te.ld + k0 * te.ps;
  hipLaunchKernelGGL(k_gather, grid, dim3(TB), 0, s, g, fields_of(st), fields_of(te), te.ld,
                     te.ps, toff, n->inv, n->X, n->T, n->M, n->red, n->red + 12);
  if (noise > 0)
    hipLaunchKernelGGL(k_noise, grid, dim3(TB), 0, s, g, n->X, n->red, noise, n->samples++);
  ACHK(n, forward32(n, g, s));
  hipLaunchKernelGGL(k_loss, grid, dim3(TB), 0, s, g, n->Y, n->T, n->M, n->G3, n->red + 6);
  const dim3 wg(NCHUNK, TAPS, 6);
  hipLaunchKernelGGL(k_wgrad, wg, dim3(TB), 0, s, g, n->H2, n->G3, n->part, DMAX3);
  hipLaunchKernelGGL(k_wreduce, dim3(6 * TAPS), dim3(TB), 0, s, n->part, n->G, 2, accumulate);
  hipLaunchKernelGGL(k_conv<true>, grid, dim3(TB), 0, s, g, n->G3, n->G2, n->W + 2 * WL, DMAX2);
  hipLaunchKernelGGL(k_wgrad, wg, dim3(TB), 0, s, g, n->H1, n->G2, n->part, DMAX2);
  hipLaunchKernelGGL(k_wreduce, dim3(6 * TAPS), dim3(TB), 0, s, n->part, n->G, 1, accumulate);
  hipLaunchKernelGGL(k_conv<true>, grid, dim3(TB), 0, s, g, n->G2, n->G1, n->W + WL, DMAX1);
  hipLaunchKernelGGL(k_wgrad, wg, dim3(TB), 0, s, g, n->X, n->G1, n->part, DMAX1);
  hipLaunchKernelGGL(k_wreduce, dim3(6 * TAPS), dim3(TB), 0, s, n->part, n->G, 0, accumulate);
  ACHK(n, hipGetLastError());
  double red[18];
  ACHK(n, hipMemcpyAsync(red, n->red, sizeof(red), hipMemcpyDeviceToHost, s));
  ACHK(n, hipStreamSynchronize(s));
  for (int f = 0; f < 6; ++f) {
    err6[f] = red[6 + f];
    tn6[f] = red[12 + f];
    if (xn6) xn6[f] = red[f];
  }
  return CPML_OK;
}

int views(cpml_abc_net *n, cpml_solver *h, cpml::SolverView *v, const char *who) {
  if (!h) return afail(n, CPML_ERR_ARG, std::string(who) + ": null handle");
  if (cpml::solver_view(h, v) != CPML_OK) return afail(n, CPML_ERR_STATE, "solver view failed");
  if (v->device != n->device) return afail(n, CPML_ERR_ARG, "handle on another device");
  ACHK(n, hipSetDevice(n->device));
  return check_view(n, *v, who);
}

}  // namespace

namespace cpml {

long abc_corr_floats(const SolverView &v) { return make_geo(v).cells4 * NF; }

int abc_prepare(cpml_abc_net *n, const SolverView &v, std::string *err) {
  int rc = check_view(n, v, "set_boundary");
  if (!rc) rc = sync_bf16(n, v.stream);
  if (rc && err) *err = n->err;
  return rc;
}

hipError_t abc_enqueue(cpml_abc_net *n, float *corr, const SolverView &v) {
  const Geo g = make_geo(v);
  const Fields u = fields_of(v);
  hipLaunchKernelGGL(k_infer, dim3(g.tiles), dim3(TBI), 0, v.stream, g, u, n->inv, n->Wb, corr);
  hipLaunchKernelGGL(k_apply, dim3((g.max4 + TB - 1) / TB, 6),
                     dim3(TB), 0, v.stream, g, u, n->scale, corr);
  return hipGetLastError();
}

}  // namespace cpml

extern "C" {

int cpml_abc_net_create(cpml_abc_net **np, const float *weights, const float *scales) {
  if (!np) return afail(nullptr, CPML_ERR_ARG, "null net pointer");
  *np = nullptr;
  if (!scales) return afail(nullptr, CPML_ERR_ARG, "null scales");
  for (int k = 0; k < NF; ++k)
    if (!(scales[k] > 0) || !std::isfinite(scales[k]))
      return afail(nullptr, CPML_ERR_ARG, "scales must be positive and finite");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    return afail(nullptr, CPML_ERR_NODEVICE, "no HIP device visible");
  auto *n = new cpml_abc_net();
  auto bail = [&](int rc) {
    g_abc_error = n->err;
    cpml_abc_net_destroy(n);
    return rc;
  };
  if (hipGetDevice(&n->device) != hipSuccess) {
    n->err = "hipGetDevice failed";
    return bail(CPML_ERR_HIP);
  }
  for (int k = 0; k < NF; ++k) {
    n->scale.s[k] = scales[k];
    n->inv.s[k] = 1.f / scales[k];
  }
  int rc = aalloc(n, (void **)&n->W, WN * sizeof(float));
  if (!rc) rc = aalloc(n, (void **)&n->Wb, WBN * sizeof(unsigned short));
  if (!rc) rc = aalloc(n, (void **)&n->red, 18 * sizeof(double));
  if (!rc) rc = aalloc(n, (void **)&n->part, (size_t)6 * TAPS * NCHUNK * C * C * sizeof(float));
  if (!rc && weights && hipMemcpy(n->W, weights, WN * sizeof(float), hipMemcpyHostToDevice))
    rc = afail(n, CPML_ERR_HIP, "weight upload failed");
  if (rc) return bail(rc);
  *np = n;
  return CPML_OK;
}

int cpml_abc_net_destroy(cpml_abc_net *n) {
  if (!n) return CPML_OK;
  (void)hipSetDevice(n->device);
  (void)hipDeviceSynchronize();
  for (void *p : n->allocs) (void)hipFree(p);
  delete n;
  return CPML_OK;
}

int cpml_abc_net_get_weights(cpml_abc_net *n, float *weights) {
  if (!n || !weights) return afail(n, CPML_ERR_ARG, "null argument");
  ACHK(n, hipSetDevice(n->device));
  ACHK(n, hipDeviceSynchronize());
  ACHK(n, hipMemcpy(weights, n->W, WN * sizeof(float), hipMemcpyDeviceToHost));
  return CPML_OK;
}

int cpml_abc_net_set_weights(cpml_abc_net *n, const float *weights) {
  if (!n || !weights) return afail(n, CPML_ERR_ARG, "null argument");
  ACHK(n, hipSetDevice(n->device));
  ACHK(n, hipDeviceSynchronize());
  ACHK(n, hipMemcpy(n->W, weights, WN * sizeof(float), hipMemcpyHostToDevice));
  n->wb_dirty = true;
  return CPML_OK;
}

int cpml_abc_train_init(cpml_abc_net *n, double lr, double beta1, double beta2, double eps,
                        int batch) {
  if (!n) return afail(n, CPML_ERR_ARG, "null net");
  if (!(lr > 0) || !(beta1 >= 0 && beta1 < 1) || !(beta2 >= 0 && beta2 < 1) || !(eps > 0) ||
      batch < 1 || batch > 4096)
    return afail(n, CPML_ERR_ARG, "bad optimiser parameters");
  ACHK(n, hipSetDevice(n->device));
  int rc;
  for (float **p : {&n->G, &n->m, &n->v}) {
    if (!*p && (rc = aalloc(n, (void **)p, WN * sizeof(float)))) return rc;
    ACHK(n, hipMemset(*p, 0, WN * sizeof(float)));
  }
  n->t = 0;
  n->batch = batch;
  n->acc_count = 0;
  for (double &v : n->acc_tn) v = 0;
  for (double &v : n->acc_xn) v = 0;
  n->lr = lr;
  n->b1 = beta1;
  n->b2 = beta2;
  n->eps = eps;
  return CPML_OK;
}

int cpml_abc_train_noise(cpml_abc_net *n, double rel_noise) {
  if (!n) return afail(n, CPML_ERR_ARG, "null net");
  if (!(rel_noise >= 0 && rel_noise <= 10)) return afail(n, CPML_ERR_ARG, "noise out of range");
  n->noise = (float)rel_noise;
  return CPML_OK;
}

int cpml_abc_grad(cpml_abc_net *n, cpml_solver *student, cpml_solver *teacher, int i0, int j0,
                  int k0, double *loss6, float *grad) {
  if (!n) return afail(n, CPML_ERR_ARG, "null net");
  cpml::SolverView st{}, te{};
  double err[6], tn[6], xn[6];
  int rc = views(n, student, &st, "student");
  if (!rc) rc = views(n, teacher, &te, "teacher");
  if (!rc) rc = grad32(n, st, te, i0, j0, k0, 0, err, tn, xn);
  if (rc) return rc;
  n->acc_count = 0;  // the gradient buffer no longer holds a partial mini-batch
  for (double &v : n->acc_tn) v = 0;
  for (double &v : n->acc_xn) v = 0;
  if (loss6)
    for (int f = 0; f < 6; ++f) loss6[f] = loss_scale(tn[f], xn[f]) * err[f];
  if (grad) {
    ACHK(n, hipMemcpy(grad, n->G, WN * sizeof(float), hipMemcpyDeviceToHost));
    for (int f = 0; f < 6; ++f) {
      const float sc = (float)loss_scale(tn[f], xn[f]);
      for (int e = 0; e < WF; ++e) grad[f * WF + e] *= sc;
    }
  }
  return CPML_OK;
}

int cpml_abc_train_step(cpml_abc_net *n, cpml_solver *student, cpml_solver *teacher, int i0,
                        int j0, int k0, int it, double *loss6) {
  if (!n) return afail(n, CPML_ERR_ARG, "null net");
  cpml::SolverView st{}, te{};
  int rc = views(n, student, &st, "student");
  if (!rc) rc = views(n, teacher, &te, "teacher");
  if (rc) return rc;
  if (st.learned) return afail(n, CPML_ERR_ARG, "the student must run without a corrector");
  if (i0 < 0 || j0 < 0 || k0 < 0 || i0 + st.nx > te.nx || j0 + st.ny > te.ny ||
      k0 + st.nz > te.nz)
    return afail(n, CPML_ERR_ARG, "student window outside the teacher grid");
  // S_n -> student, one plain step: P(S_n); teacher one step: T_{n+1}
  const long cells = (long)st.nx * st.ny * st.nz;
  hipLaunchKernelGGL(k_window, dim3((unsigned)((cells + TB - 1) / TB)), dim3(TB), 0, st.stream,
                     fields_of(st), st.nx, st.ny, st.nz, st.ld, st.ps, fields_of(te), te.ld, te.ps,
                     i0 + j0 * te.ld + k0 * te.ps);
  ACHK(n, hipGetLastError());
  ACHK(n, hipStreamSynchronize(st.stream));
  if ((rc = cpml_gpu_run(student, 1, 1)))
    return afail(n, rc, "student step failed");
  if ((rc = cpml_gpu_run(teacher, it, it)))
    return afail(n, rc, "teacher step failed");
  // the runs may have switched state buffers: take fresh views
  if ((rc = views(n, student, &st, "student")) || (rc = views(n, teacher, &te, "teacher")))
    return rc;
  double err[6], tn[6], xn[6];
  if ((rc = grad32(n, st, te, i0, j0, k0, n->acc_count > 0, err, tn, xn, n->noise))) return rc;
  for (int f = 0; f < 6; ++f) {
    n->acc_tn[f] += tn[f];
    n->acc_xn[f] += xn[f];
  }
  if (loss6)
    for (int f = 0; f < 6; ++f) loss6[f] = loss_scale(tn[f], xn[f]) * err[f];
  if (++n->acc_count < n->batch) return CPML_OK;
  // Adam step on the mini-batch: gradient of sum(err) / sum(target energy), per
  // face, the denominator floored at 1% of the batch's input energy so that
  // batches where the wave has not reached the strip yet do not dominate.
  Scales gs{};
  for (int f = 0; f < 6; ++f) gs.s[f] = (float)loss_scale(n->acc_tn[f], n->acc_xn[f]);
  n->t += 1;
  const float bc1 = (float)(1.0 - std::pow(n->b1, n->t));
  const float bc2 = (float)(1.0 - std::pow(n->b2, n->t));
  hipLaunchKernelGGL(k_adam, dim3((WN + 255) / 256), dim3(256), 0, st.stream, n->W, n->G, gs,
                     n->m, n->v, WN, (float)n->lr, (float)n->b1, (float)n->b2, (float)n->eps,
                     bc1, bc2);
  ACHK(n, hipGetLastError());
  n->wb_dirty = true;
  n->acc_count = 0;
  for (double &v : n->acc_tn) v = 0;
  for (double &v : n->acc_xn) v = 0;
  ACHK(n, hipStreamSynchronize(st.stream));
  return CPML_OK;
}

int cpml_abc_correction(cpml_abc_net *n, cpml_solver *h, int use_mfma, float *out) {
  if (!n || !out) return afail(n, CPML_ERR_ARG, "null argument");
  cpml::SolverView v{};
  int rc = views(n, h, &v, "handle");
  if (rc) return rc;
  const Geo g = make_geo(v);
  if ((rc = ensure_buffers(n, g))) return rc;
  hipStream_t s = v.stream;
  if (use_mfma) {
    if ((rc = sync_bf16(n, s))) return rc;
    ACHK(n, hipMemsetAsync(n->corr, 0, (size_t)g.cells4 * NF * sizeof(float), s));
    hipLaunchKernelGGL(k_infer, dim3(g.tiles), dim3(TBI), 0, s, g, fields_of(v), n->inv, n->Wb,
                       n->corr);
  } else {
    ACHK(n, hipMemsetAsync(n->red, 0, 18 * sizeof(double), s));
    const dim3 grid((g.max7 + TB - 1) / TB, 6);
    hipLaunchKernelGGL(k_gather, grid, dim3(TB), 0, s, g, fields_of(v), fields_of(v), v.ld, v.ps,
                       0L, n->inv, n->X, n->T, n->M, n->red, n->red + 12);
    ACHK(n, forward32(n, g, s));
    hipLaunchKernelGGL(k_y_to_corr, grid, dim3(TB), 0, s, g, n->Y, n->corr);
  }
  ACHK(n, hipGetLastError());
  ACHK(n, hipMemcpyAsync(out, n->corr, (size_t)g.cells4 * NF * sizeof(float),
                         hipMemcpyDeviceToHost, s));
  ACHK(n, hipStreamSynchronize(s));
  return CPML_OK;
}

int cpml_abc_last_error(cpml_abc_net *n, char *buf, int len) {
  if (!buf || len <= 0) return CPML_ERR_ARG;
  const std::string &m = n ? n->err : g_abc_error;
  const size_t k = std::min(m.size(), (size_t)len - 1);
  std::memcpy(buf, m.data(), k);
  buf[k] = '\0';
  return CPML_OK;
}

}  // extern "C"
