// Internal declarations shared by the HIP translation units of libcpml_gpu.
//
// Device layout (all arrays x-fastest, like the reference's Fortran arrays):
//
//  * wavefields and media: one padded array each.  Interior point (i,j,k)
//    (0-based; Fortran index i+1, j+1, k+1) lives at  base + i + j*ld + k*ps.
//    A halo of R = order/2 cells on every side is allocated and kept at zero
//    (the reference's ghost points, 2D4:205 0:NX+1, ADE8:272 -4:NX+4), so the
//    stencils read neighbours without bounds tests.  The x halo is XPAD
//    elements wide so that interior rows start 256-byte aligned.
//  * C-PML memory variables: only inside the absorbing slabs.  Along x the
//    slab is the columns [0, wl) U [nx-wr, nx) stored compactly as s in
//    [0, wl+wr); along y and z likewise for rows and planes.  Outside the slab
//    the reference's memory variable stays exactly 0 (a = 0) and K = 1, so
//    the update reduces to the plain derivative there.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "../../include/cpml_gpu.h"

namespace cpml {

constexpr int kMaxRecBlock = 256;

// Holberg (1987) 8th-order staggered coefficients as printed in
// seismic_ADEPML_2D_elastic_RK4_eighth_order.f90:263-266.
constexpr double kHolberg1 = 1.231666;
constexpr double kHolberg2 = -1.041182e-1;
constexpr double kHolberg3 = 2.063707e-2;
constexpr double kHolberg4 = -3.570998e-3;

// State size (bytes) above which the launchers pick the kernels with
// non-temporal own-point accesses: the 256 MiB last-level (infinity) cache of
// the MI355X.  Smaller states stay cache-resident from one step to the next.
constexpr double kStreamingBytes = 256.0 * 1024 * 1024;

struct Slab {
  int wl = 0, wr = 0, w = 0;  // widths at the min and max side, total
  int n = 0;                  // grid extent along the axis
  __host__ __device__ bool in(int i) const {
    return i >= 0 && i < n && (i < wl || i >= n - wr);
  }
  __host__ __device__ int s(int i) const { return i < wl ? i : i - (n - wr) + wl; }
};

// Slab widths from one axis's profiles: a point belongs to the slab when any
// of a, a_half differs from 0 or K, K_half from 1; the widths are the extents
// of such points from the min side (first half) and the max side.
template <typename T>
inline void slab_from_profile(const T *a, const T *ah, const T *K, const T *Kh, int n, Slab &s) {
  const int half = (n + 1) / 2;
  int wl = 0, wr = 0;
  for (int i = 0; i < n; ++i) {
    const bool nz = a[i] != T(0) || ah[i] != T(0) || K[i] != T(1) || Kh[i] != T(1);
    if (!nz) continue;
    if (i < half) wl = i + 1 > wl ? i + 1 : wl;
    else wr = n - i > wr ? n - i : wr;
  }
  s.wl = wl;
  s.wr = wr;
  s.w = wl + wr;
  s.n = n;
}

// Profiles along one axis, 6 consecutive arrays of length n:
// a, b, K, a_half, b_half, K_half (the reference's a_x, b_x, K_x, ...).
template <typename T>
struct Prof {
  const T *a, *b, *K, *ah, *bh, *Kh;
};

template <typename T>
struct Step2D {
  T *vx, *vy, *sxx, *syy, *sxy;  // interior origin (halo at negative offsets)
  const T *lam, *mu, *rho;
  // x-slab memories: [0] dvx_dx [1] dvy_dx [2] dsxx_dx [3] dsxy_dx, index s + wx*j
  T *mx[4];
  // y-slab memories: [0] dvy_dy [1] dvx_dy [2] dsxy_dy [3] dsyy_dy, index i + mld*s
  T *my[4];
  Prof<T> px, py;
  Slab sx, sy;
  int nx, ny;
  long ld;   // row stride of fields
  long mld;  // row stride of y-slab memories
  T den_x, den_y, dt;
  // source (0-based, isrc < 0: none) and step counter
  int isrc, jsrc;
  const T *fx, *fy;
  const int *step;  // current step it (1-based), advanced by the record kernel
  int tile_y;       // rows per block (4, 8 or 16)
  int streaming;    // state larger than the caches: non-temporal own-point accesses
};

template <typename T>
struct Step3D {
  T *vx, *vy, *vz, *sxx, *syy, *szz, *sxy, *sxz, *syz;
  const T *clam, *cmu, *cbuoy;  // DELTAT*lambda, DELTAT*mu, DELTAT/rho
  const T *rho;                 // raw density (source term, 3D:1081-1082)
  const T *lam, *mu;            // raw Lame parameters (4th-order step)
  // x-slab: [0] dvx_dx [1] dvy_dx [2] dvz_dx [3] dsxx_dx [4] dsxy_dx [5] dsxz_dx
  T *mx[6];
  // y-slab: [0] dvy_dy [1] dvx_dy [2] dvz_dy [3] dsxy_dy [4] dsyy_dy [5] dsyz_dy
  T *my[6];
  // z-slab: [0] dvz_dz [1] dvx_dz [2] dvy_dz [3] dsxz_dz [4] dsyz_dz [5] dszz_dz
  T *mz[6];
  Prof<T> px, py, pz;
  Slab sx, sy, sz;
  int nx, ny, nz;
  long ld, ps;    // field strides
  long base;      // offset of interior point (0,0,0) from the allocation start
  long xpad;      // x halo width (interior rows start 256-byte aligned)
  int R;          // halo rows / planes on each side
  long mld, mps;  // memory strides: x-slab (w, w*ny), y-slab (mld, mld*wy), z (mld, mld*ny)
  T odx, ody, odz;
  T dt;
  int isrc, jsrc, ksrc;
  const T *fx, *fy;
  const int *step;
  int kchunk;  // planes per block along z
  int tile_y;  // rows per block (4, 8 or 16)
  int streaming;  // state larger than the caches: non-temporal own-point accesses
};

// Fused 3D step: reads the state in `a`, writes the next state here.
template <typename T>
struct Fused3D {
  Step3D<T> a;
  T *vx, *vy, *vz, *sxx, *syy, *szz, *sxy, *sxz, *syz;  // next state, interior origin
  T *mx[6], *my[6], *mz[6];                             // next memory variables
  // Tiles whose points and ring touch no absorbing slab form a box in tile
  // coordinates [box[0], box[1]) x [box[2], box[3]) x [box[4], box[5]); set by
  // the launcher.
  int box[6];
  // z chunks: planes [0, z0), [z0, z1), [z1, nz) cut into zn[0], zn[1], zn[2]
  // chunks of at most kchunk planes each; set by the launcher.
  int z0, z1, zn[3];
};

// Receivers / step counter: one block.
template <typename T>
struct Record {
  const T *vx, *vy, *vz;  // interior origin
  long ld, ps;
  int nrec;
  int avg4;        // 1: vx as the mean of (i,j), (i+1,j), (i,j+1), (i+1,j+1) (ADE8:879-882)
  const int *idx;  // nrec offsets (i + j*ld + k*ps)
  T *sx, *sy, *sz;  // [nstep][nrec]
  int *step;
  // the state is that of step it + skip (a launch that advanced skip + 1 steps
  // and recorded the first skip itself): recorded as step it + skip, and the
  // counter advances by skip + 1
  int skip;
};

// Time-blocked 3D step: nt leapfrog steps per launch, state a -> the arrays
// below.  A workgroup is a tile of `ow` columns x (its wavefronts - 2 nt)
// rows marching through a z chunk; tiles whose points and ring avoid the x
// and y slabs form the box [box[0], box[1]) x [box[2], box[3]) of tile
// coordinates, and z chunks are cut per region as in Fused3D (z0, z1, zn).
template <typename T>
struct TB3D {
  Step3D<T> a;                                          // state in
  T *vx, *vy, *vz, *sxx, *syy, *szz, *sxy, *sxz, *syz;  // state out, interior origin
  T *mx[6], *my[6], *mz[6];                             // memory variables out
  int ow, ty;          // owned columns and rows of a tile
  int box[4];
  int z0, z1, zn[3];
  // receivers of each x-y tile (t = ty * ntx + tx) sorted by plane: entries
  // [rec_start[t], rec_start[t+1]) of rec_i, rec_j, rec_k (0-based) and rec_slot
  const int *rec_start, *rec_i, *rec_j, *rec_k, *rec_slot;
  T *sisx, *sisy, *sisz;  // [nstep][nrec]
  int nrec;
};

// Time-blocked 2D step: nt leapfrog steps per launch, state a -> the arrays
// below.  Work items are (strip, chunk): 64-lane wavefronts owning `ow`
// columns of a strip and the rows [jb, jb + jc) of a chunk.  Items whose
// columns and rows (with their rings) avoid the absorbing slabs form the box
// of strips [sx0, sx1) x chunks [cy0, cy1); the others run in the launch built
// with the C-PML code.
template <typename T>
struct TB2D {
  Step2D<T> a;                   // state in (fields, memory variables, coefficients)
  T *vx, *vy, *sxx, *syy, *sxy;  // state out, interior origin
  T *mx[4], *my[4];              // memory variables out (layout of a.mx, a.my)
  int ow, jc, nstrips, nchunks;
  int sx0, sx1, cy0, cy1;
  // receivers of each strip sorted by row: strip t owns entries
  // [rec_start[t], rec_start[t+1]) of rec_i, rec_j (0-based) and rec_slot
  // (the receiver's index in the seismogram rows)
  const int *rec_start, *rec_i, *rec_j, *rec_slot;
  T *sisx, *sisy;  // [nstep][nrec]
  int nrec;
  // energy of every level (single grid, energy on): each item's sums over
  // its owned points inside the window ew (0-based i0, i1, j0, j1), written
  // to epart[(level * eslots + item) * 2 + {kinetic, potential}]
  double *epart;
  int ew[4];
  int eslots;
};

// Energy over a window, per step, accumulated in double.
template <typename T>
struct Energy {
  const T *vx, *vy, *vz, *sxx, *syy, *szz, *sxy, *sxz, *syz;
  const T *lam, *mu, *rho;  // raw media (field layout)
  long ld, ps;
  int i0, i1, j0, j1, k0, k1;  // 0-based inclusive
  int ndim, corrected3d, cadence;
  double *partial;  // [nblocks][2]
  double *out_kin, *out_pot;  // [nstep]
  const int *step;
  int offset;  // the state is that of step *step + offset (time blocks: NT - 1)
};

// What the learned-boundary code needs of a solver handle: sizes, strides and
// the current state (interior origin of each field), its stream and device.
struct SolverView {
  int ndim, order, prec, nx, ny, nz;
  long ld, ps;
  void *f[9];  // vx vy vz sxx syy szz sxy sxz syz
  int slab_w[6];
  int learned;  // a boundary corrector is attached
  hipStream_t stream;
  int device;
};

// defined in cpml_api.hip
int solver_view(cpml_solver *h, SolverView *v);
// message returned by cpml_gpu_last_error(NULL, ...) on this thread
void set_thread_error(const std::string &msg);
// defined in cpml_abc.hip: correction buffer size, weight upload, per-step launch
long abc_corr_floats(const SolverView &v);
int abc_prepare(cpml_abc_net *n, const SolverView &v, std::string *err);
hipError_t abc_enqueue(cpml_abc_net *n, float *corr, const SolverView &v);

// Kernel launchers (defined in the .hip files).
template <typename T>
hipError_t launch_stress2d(int order, const Step2D<T> &p, hipStream_t s);
template <typename T>
hipError_t launch_velocity2d(int order, const Step2D<T> &p, hipStream_t s);
template <typename T>
hipError_t launch_stress3d(const Step3D<T> &p, hipStream_t s);
template <typename T>
hipError_t launch_velocity3d(const Step3D<T> &p, hipStream_t s);
// ADE-PML RK4 step of the 8th-order program: the 2D fields (s), three stage
// copies of vx vy sxx syy sxy (d[level][field], padded like the fields), the
// eight memory arrays (padded), the stage tables a, b ([4][n]: x, x_half, y,
// y_half) and the force at each stage time ([NSTEP][4]).
template <typename T>
struct Ade8 {
  Step2D<T> s;         // X (vx vy sxx syy sxy) at the step's start: X0, stage 1's state
  T *d[2][5];          // the stage states of stages 2..4, alternating
  T *xn[5];            // X at the step's end (the other state buffer)
  T *mem[8];  // dvx_dx dvx_dy dvy_dx dvy_dy dsxx_dx dsyy_dy dsxy_dx dsxy_dy
  const T *ax, *bx, *axh, *bxh, *ay, *by, *ayh, *byh;
  const T *f4x, *f4y;
  T dx, dy;
  int npml;
};
template <typename T>
hipError_t launch_ade8_step(const Ade8<T> &a, hipStream_t s);

// 2D viscoelastic step: the 2D step plus the attenuation memory variables
// (dense nx*ny per SLS plane) and their per-SLS coefficients.
constexpr int kMaxSls = 8;
template <typename T>
struct Visco2D {
  Step2D<T> s;
  T *re1, *re11, *re13;
  long nn;   // nx * ny
  int nsls;
  T P1[kMaxSls], P2[kMaxSls];  // DELTAT_phi_nu1, _nu2
  T H1[kMaxSls], H2[kMaxSls];  // HALF_DELTAT_over_tau_sigma_nu1, _nu2
  T M1[kMaxSls], M2[kMaxSls];  // multiplication_factor_tau_sigma_nu1, _nu2
  T odx, ody;                  // ONE_OVER_DELTAX, ONE_OVER_DELTAY
  T n8x, n8y, o24x, o24y;      // NINE_OVER_8_DELTA*, ONE_OVER_24_DELTA* (order 4)
  T *sisp;                     // pressure seismograms [NSTEP][NREC]
  const int *recidx;
  int nrec;
};
template <typename T>
hipError_t launch_stress2dve(int order, const Visco2D<T> &v, hipStream_t s);
template <typename T>
hipError_t launch_velocity2dve(int order, const Visco2D<T> &v, hipStream_t s);
// 4th-order 3D step (elastic part of the 3D viscoelastic program)
template <typename T>
hipError_t launch_stress3d4(const Step3D<T> &p, hipStream_t s);
template <typename T>
hipError_t launch_velocity3d4(const Step3D<T> &p, hipStream_t s);
template <typename T>
hipError_t launch_fused3d(const Fused3D<T> &p, hipStream_t s);
template <typename T>
hipError_t launch_record(const Record<T> &p, hipStream_t s);
// nt steps per launch (3D, second order); tiles of ow columns x (nw - 2 nt)
// rows marching through z chunks of at most kchunk planes
// (the launches of tiles that touch a slab go to `side` when given, so they
// run beside the slab-free tiles; the caller joins the two streams)
template <typename T>
hipError_t launch_tb3d(int nt, const TB3D<T> &p, hipStream_t s, hipStream_t side);
int tb3d_max_nt(int prec);
// nt steps per launch (2D, plain elastic); clean = items without C-PML code
template <typename T>
hipError_t launch_tb2d(int order, int nt, const TB2D<T> &p, hipStream_t s, hipStream_t side);
// The energy of the NT levels of one time-blocked launch from its items' sums
// (TB2D::epart), into e.out_kin / e.out_pot at steps *e.step .. + NT - 1.
template <typename T>
hipError_t launch_tb2d_energy(const Energy<T> &e, const double *epart, int eslots, int nt,
                              hipStream_t s);
// largest nt launch_tb2d has a kernel for, per order and precision
int tb2d_max_nt(int order, int prec);
// many grids: items (grid, strip, chunk) of one class (cls bit 0: the x
// slabs' C-PML code, bit 1: the y slabs') from the descriptor table `grids`,
// on `blocks` workgroups of 4 wavefronts
template <typename T>
hipError_t launch_tb2d_multi(int order, int nt, int cls, const TB2D<T> *grids,
                             const int4 *items, int nitems, int blocks, hipStream_t s);
// strips [box[0], box[1]) x chunks [box[2], box[3]) of one grid need no C-PML code
void tb2d_clean_box(int order, int nt, int nx, int ny, int ow, int jc, const Slab &sx,
                    const Slab &sy, int box[4]);
template <typename T>
hipError_t launch_energy(const Energy<T> &p, int nblocks, hipStream_t s);
template <typename T>
int energy_blocks(const Energy<T> &p);
// PNM snapshot parameters (Fortran indices): source, PML thickness, the
// layers in use (x min, x max, y min, y max), ghost border (1: the image is
// the ghost-padded array of the 4th-order program).
struct Image {
  int isrc, jsrc, npml, use[4], ghost;
  long ps;  // plane stride of the receiver offsets
};
template <typename T>
hipError_t launch_color_image(const T *f, int nx, int ny, long ld, unsigned long long *amax,
                              const Image &g, const int *recidx, int nrec, unsigned char *rgb,
                              hipStream_t s);
template <typename T>
hipError_t launch_vmax(const T *vx, const T *vy, const T *vz, int nx, int ny, int nz, long ld,
                       long ps, unsigned long long *out, hipStream_t s);

}  // namespace cpml
