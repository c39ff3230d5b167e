// Shared device definitions of the D3Q19 PSC-IMB kernels (gfx950).
//
// Lattice constants follow src/lbm_dem/psc3d.py:17-43 of the reference:
// direction 0 is rest, 1-6 the axis directions (+x,-x,+y,-y,+z,-z), 7-18
// the face diagonals in (+,+),(-,-),(+,-),(-,+) pairs over xy, xz, yz; the
// opposite of direction i is its pair partner.  Weights 1/3, 1/18, 1/36,
// cs^2 = 1/3.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace psc {

constexpr int Q = 19;

__host__ __device__ constexpr int cx_of(int i) {
  constexpr int t[Q] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
  return t[i];
}
__host__ __device__ constexpr int cy_of(int i) {
  constexpr int t[Q] = {0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1};
  return t[i];
}
__host__ __device__ constexpr int cz_of(int i) {
  constexpr int t[Q] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1};
  return t[i];
}
__host__ __device__ constexpr int c_of(int i, int axis) {
  return axis == 0 ? cx_of(i) : (axis == 1 ? cy_of(i) : cz_of(i));
}
__host__ __device__ constexpr int opp_of(int i) { return i == 0 ? 0 : ((i & 1) ? i + 1 : i - 1); }

// Weights as the reference's Python floats 1/3, 1/18, 1/36 (correctly rounded).
__host__ __device__ constexpr double w_of(int i) {
  return i == 0 ? 1.0 / 3.0 : (i <= 6 ? 1.0 / 18.0 : 1.0 / 36.0);
}
// q = n / d, r = n % d for n < 2^24 from an FP32 reciprocal inv ~ 1/d,
// corrected to the exact quotient (a few instructions instead of the
// integer division sequence).
__device__ __forceinline__ void divmod_f(uint32_t n, uint32_t d, float inv, uint32_t& q, uint32_t& r) {
  q = (uint32_t)((float)n * inv);
  int rr = (int)n - (int)(q * d);
  while (rr < 0) { --q; rr += (int)d; }
  while (rr >= (int)d) { ++q; rr -= (int)d; }
  r = (uint32_t)rr;
}

// Exact n / d for every n < 2^31 by a multiply and a shift (Granlund and
// Montgomery, "Division by invariant integers using multiplication", 1994,
// theorem 4.2 with N = 31): with l = ceil(log2 d) and m = ceil(2^(31+l) / d),
// 2^(31+l) <= m d < 2^(31+l) + 2^l, so floor(n m / 2^(31+l)) = floor(n / d);
// m < 2^32 because d > 2^(l-1).
struct Divider {
  uint32_t m, sh, d;
};
inline Divider make_divider(uint32_t d) {
  uint32_t l = 0;
  while ((1ull << l) < d) ++l;
  const unsigned long long p = 1ull << (31 + l);
  return Divider{(uint32_t)((p + d - 1) / d), 31 + l, d};
}
__device__ __forceinline__ void divmod_exact(uint32_t n, const Divider& v, uint32_t& q, uint32_t& r) {
  q = (uint32_t)(((unsigned long long)n * v.m) >> v.sh);
  r = n - q * v.d;
}

constexpr double kCs2 = 1.0 / 3.0;
constexpr double kCs2Sq = kCs2 * kCs2;

// Periodic neighbour coordinate: d in {-1, 0, 1}.
__device__ __forceinline__ int wrap_step(int v, int d, int n) {
  int r = v + d;
  r = (r < 0) ? r + n : r;
  r = (r >= n) ? r - n : r;
  return r;
}

// Wrap an arbitrary (possibly negative) coordinate into [0, n).
__device__ __forceinline__ int wrap_any(int v, int n) {
  int r = v % n;
  return r < 0 ? r + n : r;
}

// Wrap a coordinate v in [-n, 2n) into [0, n): the box of every sphere
// narrower than the lattice stays within one period of it.
__device__ __forceinline__ int wrap_near(int v, int n) { return v < 0 ? v + n : (v >= n ? v - n : v); }

// Owner table of the FP32 moving step: owner[c] is the particle whose sphere
// alone covers cell c, or kSharedCell where several spheres cover it.
constexpr uint32_t kSharedCell = 0xffffffffu;

// Solid-velocity inputs of a moving FP32 step.  Without an owner table the
// numerator planes hold sum_p alpha_p (U + omega x r)_p at every covered cell
// (plane stride N); with one, they hold it at shared cells only, and a cell
// with a single owner takes that particle's U + omega x r from its motion
// record: box origin, centre, U and omega (FP64), so that the kernel forms
// the voxelizer's lever arm (box coordinate + 1/2 - centre) and velocity in
// FP64 and rounds once per cell.  Rounding U, omega or the centre to FP32
// once per particle would give every cell of a sphere the same error, which
// adds up in its load instead of averaging out.
struct alignas(16) MotionRecord {
  int32_t x0, y0, z0, pad;
  double cx, cy, cz, vx, vy, vz, wx, wy, wz, pad2;
};
static_assert(sizeof(MotionRecord) == 96, "motion record layout");
struct SolidMotion {
  const float* us_num;
  const uint32_t* owner;
  const MotionRecord* motion;
  uint32_t nmotion;             // records in motion; an owner id at or past it reads the numerator planes
  unsigned long long us_plane;  // cells per numerator plane (nx*ny*nz, whatever the state's plane size)
};

// Cells per direction plane of a single lattice's FP32 state: nx*ny*nz, or,
// with 64-cell-aligned planes, that rounded up to 64.
__host__ __device__ inline unsigned long long state_plane(unsigned long long ncell, bool aligned64) {
  return aligned64 ? (ncell + 63ull) / 64ull * 64ull : ncell;
}

// Sum of K doubles per thread over the workgroup: wave64 butterfly then LDS across waves.
// Result valid in thread 0.  blockDim.x must be a multiple of 64, <= 1024.
template <int K>
__device__ __forceinline__ void block_sum(double (&v)[K], double* lds /* K*16 doubles */) {
#pragma unroll
  for (int k = 0; k < K; ++k) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v[k] += __shfl_xor(v[k], off, 64);
  }
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int nwave = blockDim.x >> 6;
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) lds[wave * K + k] = v[k];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) {
      double acc = lds[k];
      for (int w = 1; w < nwave; ++w) acc += lds[w * K + k];
      v[k] = acc;
    }
  }
}

}  // namespace psc

// Error reporting shared by the translation units (defined in psc_abi.hip).
namespace psc {
int set_error(int code, const char* fmt, ...);
int check_launch(const char* what);
}  // namespace psc
