// FP64 kernels with the reference's operation order: the plain FP64 step,
// sub-sample voxelization, particle loads, moments/readouts and the f
// import/export.  This translation unit is compiled with -ffp-contract=off so
// that every multiply and add rounds separately, as NumPy does; the FP64 step
// then reproduces src/lbm_dem/psc3d.py bit for bit on cells owned by one
// particle.
#include <mutex>
#include <vector>

#include "psc_common.h"
#include "psc_div.h"
#include "psc_loads.h"
#include "../../include/psc_hip.h"

namespace psc {
namespace {

// ((cx*ux + cy*uy) + cz*uz) for cx, cy, cz in {-1, 0, 1}, as NumPy forms it
// (equilibrium, psc3d.py:101).  Products with 0 only contribute signed zeros,
// so the sum of the non-zero terms in axis order rounds identically.
template <int I>
__device__ __forceinline__ double cdot(double ux, double uy, double uz) {
  constexpr int a = cx_of(I), b = cy_of(I), c = cz_of(I);
  double acc = 0.0;
  bool first = true;
  if (a != 0) { acc = a > 0 ? ux : -ux; first = false; }
  if (b != 0) { const double t = b > 0 ? uy : -uy; acc = first ? t : acc + t; first = false; }
  if (c != 0) { const double t = c > 0 ? uz : -uz; acc = first ? t : acc + t; first = false; }
  return acc;
}

// sum_i v_i c_{i,AX} accumulated in direction order (np.sum(..., axis=0)).
template <int AX>
__device__ __forceinline__ double moment(const double (&v)[Q]) {
  double acc = 0.0;
  bool first = true;
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    const int c = c_of(i, AX);
    if (c == 0) continue;
    const double t = c > 0 ? v[i] : -v[i];
    acc = first ? t : acc + t;
    first = false;
  }
  return acc;
}

// Reciprocals of the constant divisors, correctly rounded (compile-time IEEE
// division), for the divisions of psc_div.h: every quotient below is IEEE
// division's, bit for bit.
constexpr double kInvCs2 = 1.0 / kCs2;
constexpr double kInvCs2Sq = 1.0 / kCs2Sq;

template <int I>
__device__ __forceinline__ double feq_term(double rho, double ux, double uy, double uz, double uu) {
  const double cu = cdot<I>(ux, uy, uz);
  return (w_of(I) * rho) * (((1.0 + div_pos(cu, kCs2, kInvCs2)) + div_pos((0.5 * cu) * cu, kCs2Sq, kInvCs2Sq)) -
                            div_pos(0.5 * uu, kCs2, kInvCs2));
}

template <int... Is>
struct Seq {};

template <int N, int... Is>
struct MakeSeq : MakeSeq<N - 1, N - 1, Is...> {};
template <int... Is>
struct MakeSeq<0, Is...> {
  using type = Seq<Is...>;
};
using Dirs = MakeSeq<Q>::type;

template <int... Is>
__device__ __forceinline__ void equilibrium_all(Seq<Is...>, double rho, double ux, double uy, double uz,
                                                double (&out)[Q]) {
  const double uu = (ux * ux + uy * uy) + uz * uz;
  ((out[Is] = feq_term<Is>(rho, ux, uy, uz, uu)), ...);
}

template <int I>
__device__ __forceinline__ double guo_term(double pref, double uf, double ux, double uy, double uz,
                                           double fx, double fy, double fz) {
  const double cu = cdot<I>(ux, uy, uz);
  const double cf = cdot<I>(fx, fy, fz);
  return (pref * w_of(I)) * (div_pos(cf - uf, kCs2, kInvCs2) + div_pos(cu * cf, kCs2Sq, kInvCs2Sq));
}

template <int... Is>
__device__ __forceinline__ void guo_all(Seq<Is...>, double pref, double ux, double uy, double uz, double fx,
                                        double fy, double fz, double (&out)[Q]) {
  const double uf = (ux * fx + uy * fy) + uz * fz;
  ((out[Is] = guo_term<Is>(pref, uf, ux, uy, uz, fx, fy, fz)), ...);
}

struct Dims {
  int nx, ny, nz;
  int pz;                 // row pitch along z (cells); >= nz, padding cells are never used
  unsigned long long ps;  // cells per direction plane (>= nx*ny*pz)
};

// Plane size of a lattice with an explicit row pitch: nx*ny*pz rounded up to
// 64 cells, so planes start 256-byte aligned (the batched-window layout).
__host__ __device__ inline unsigned long long plane_cells(int nx, int ny, int pz) {
  const unsigned long long n = (unsigned long long)nx * ny * pz;
  return (n + 63ull) / 64ull * 64ull;
}

// Cell of the thread in a (z-tiles, y-tiles, x) grid.
__device__ __forceinline__ bool thread_cell(const Dims d, int& ix, int& iy, int& iz) {
  iz = blockIdx.x * blockDim.x + threadIdx.x;
  iy = blockIdx.y * blockDim.y + threadIdx.y;
  ix = blockIdx.z;
  return iz < d.nz && iy < d.ny;
}

__device__ __forceinline__ uint32_t cell_id(const Dims d, int x, int y, int z) {
  return ((uint32_t)x * (uint32_t)d.ny + (uint32_t)y) * (uint32_t)d.pz + (uint32_t)z;
}

// Address of direction I's value for cell (ix,iy,iz) shifted by s*c_I.
template <int I>
__device__ __forceinline__ uint32_t shifted(const Dims d, int ix, int iy, int iz, int s) {
  return cell_id(d, wrap_step(ix, s * cx_of(I), d.nx), wrap_step(iy, s * cy_of(I), d.ny),
                 wrap_step(iz, s * cz_of(I), d.nz));
}

// NT: non-temporal loads and stores (the state streams through once per step).
template <bool NT>
__device__ __forceinline__ double ld64(const double* p) {
  return NT ? __builtin_nontemporal_load(p) : *p;
}
template <bool NT>
__device__ __forceinline__ void st64(double* p, double v) {
  if (NT) __builtin_nontemporal_store(v, p);
  else *p = v;
}

template <bool PULL, bool NT = false, int... Is>
__device__ __forceinline__ void load_all(Seq<Is...>, const double* __restrict__ in, size_t N, const Dims d,
                                         int ix, int iy, int iz, double (&f)[Q]) {
  ((f[Is] = ld64<NT>(in + (size_t)Is * N + (PULL ? shifted<Is>(d, ix, iy, iz, -1) : cell_id(d, ix, iy, iz)))), ...);
}

template <bool PULL, bool NT = false, int... Is>
__device__ __forceinline__ void store_all(Seq<Is...>, double* __restrict__ out, size_t N, const Dims d, int ix,
                                          int iy, int iz, const double (&f)[Q]) {
  (st64<NT>(out + (size_t)Is * N + (PULL ? cell_id(d, ix, iy, iz) : shifted<Is>(d, ix, iy, iz, 1)), f[Is]), ...);
}

// ---------------------------------------------------------------------------
// K2 (FP64 plain): psc3d.py:211-242 in the reference's operation order.
// ---------------------------------------------------------------------------
template <bool PULL, bool MOVING, bool FORCE, bool IMB, bool NT>
__global__ __launch_bounds__(256) void k_step_f64(const double* __restrict__ fin, double* __restrict__ fout,
                                                  const uint32_t* __restrict__ csum,
                                                  const double* __restrict__ usn, double* __restrict__ imb,
                                                  Dims d, double s3, double ys3, double tau, double ytau,
                                                  double pref, double fx, double fy, double fz) {
  int ix, iy, iz;
  if (!thread_cell(d, ix, iy, iz)) return;
  const size_t N = (size_t)d.ps;
  const uint32_t cell = cell_id(d, ix, iy, iz);
  double f[Q];
  load_all<PULL, NT>(Dirs{}, fin, N, d, ix, iy, iz, f);

  // rho, u with the half-force shift (psc3d.py:211-214)
  double rho = f[0];
#pragma unroll
  for (int i = 1; i < Q; ++i) rho = rho + f[i];
  // every division is IEEE's, bit for bit (psc_div.h): by rho through one
  // division for its reciprocal, by tau, s^3 and the constants through theirs
  const double yrho = 1.0 / rho;
  const double ux = div_any(moment<0>(f) + 0.5 * fx, rho, yrho);
  const double uy = div_any(moment<1>(f) + 0.5 * fy, rho, yrho);
  const double uz = div_any(moment<2>(f) + 0.5 * fz, rho, yrho);
  // second half-force shift (psc3d.py:217-219)
  const double uxe = ux + div_any(0.5 * fx, rho, yrho);
  const double uye = uy + div_any(0.5 * fy, rho, yrho);
  const double uze = uz + div_any(0.5 * fz, rho, yrho);

  double feqf[Q];
  equilibrium_all(Dirs{}, rho, uxe, uye, uze, feqf);

  // solid fraction and B (psc3d.py:174-175, 228-229); s3 == 0 selects a
  // given FP64 solid-fraction field in place of the sub-sample counts
  const bool field = s3 == 0.0;
  const uint32_t cnt = field ? 0u : csum[cell];
  const double asum = field ? ((const double*)csum)[cell] : div_pos((double)cnt, s3, ys3);
  const double solid = asum < 1.0 ? asum : 1.0;
  const double b = solid / ((solid + tau * (1.0 - solid)) + 1.0e-14);

  double fnew[Q];
  double gterm[Q];
  if (FORCE) guo_all(Dirs{}, pref, uxe, uye, uze, fx, fy, fz, gterm);

  if (field ? solid == 0.0 : cnt == 0) {
    // B = 0: collision = 1*Omega_bgk + 0*Omega_s
#pragma unroll
    for (int i = 0; i < Q; ++i) {
      const double obgk = -div_pos(f[i] - feqf[i], tau, ytau);
      fnew[i] = FORCE ? f[i] + (obgk + gterm[i]) : f[i] + obgk;
    }
  } else {
    double feqs[Q];
    if (MOVING) {
      // solid velocity: numerator / alpha_sum (psc3d.py:202-205)
      double usx = usn[cell], usy = usn[N + cell], usz = usn[2 * N + cell];
      if (asum > 1.0e-14) {
        const double yas = 1.0 / asum;
        usx = div_pos(usx, asum, yas);
        usy = div_pos(usy, asum, yas);
        usz = div_pos(usz, asum, yas);
      }
      equilibrium_all(Dirs{}, rho, usx, usy, usz, feqs);
    } else {
#pragma unroll
      for (int i = 0; i < Q; ++i) feqs[i] = w_of(i) * rho;
    }
    double bos[Q];
    const double omb = 1.0 - b;
#pragma unroll
    for (int i = 0; i < Q; ++i) {
      const int o = opp_of(i);
      const double obgk = -div_pos(f[i] - feqf[i], tau, ytau);
      const double os = ((f[o] - f[i]) + feqs[i]) - feqf[o];
      bos[i] = b * os;
      const double coll = omb * obgk + bos[i];
      fnew[i] = FORCE ? f[i] + (coll + gterm[i]) : f[i] + coll;
    }
    if (IMB) {  // interleaved: imb[3 cell + axis]
      imb[3 * (size_t)cell] = moment<0>(bos);
      imb[3 * (size_t)cell + 1] = moment<1>(bos);
      imb[3 * (size_t)cell + 2] = moment<2>(bos);
    }
  }
  store_all<PULL, NT>(Dirs{}, fout, N, d, ix, iy, iz, fnew);
}

// ---------------------------------------------------------------------------
// Plain D3Q19 BGK with Guo forcing and bounce-back at a binary solid mask
// (scripts/run_3d_window_lbm.py:95-146), pull form: the state holds the
// post-collision values; a step gathers the streamed populations, applies the
// solid-node loop f_i <- f_opp(i) in direction order (each odd member takes
// its even partner's value, the even member keeps it: the loop runs in place),
// then collides with u = 0 in solid cells:
//   u = (sum f c + F/2)/rho (the half shift only on components with F != 0),
//   f' = (f - (f - feq)/tau) + Guo(u).
// ---------------------------------------------------------------------------
// NT: non-temporal state loads and stores (the same bits).
template <bool FORCE, bool NT>
__global__ __launch_bounds__(256) void k_bb_step_f64(const double* __restrict__ fin, double* __restrict__ fout,
                                                     const uint8_t* __restrict__ solid, Dims d, double tau,
                                                     double ytau, double pref, double fx, double fy, double fz) {
  int ix, iy, iz;
  if (!thread_cell(d, ix, iy, iz)) return;
  const size_t N = (size_t)d.ps;
  const uint32_t cell = cell_id(d, ix, iy, iz);
  double f[Q];
  load_all<true, NT>(Dirs{}, fin, N, d, ix, iy, iz, f);
  const bool s = solid[cell] != 0;
  if (s) {
#pragma unroll
    for (int i = 1; i < Q; i += 2) f[i] = f[i + 1];
  }
  double rho = f[0];
#pragma unroll
  for (int i = 1; i < Q; ++i) rho = rho + f[i];
  const double yrho = 1.0 / rho;  // IEEE divisions through the reciprocal (psc_div.h)
  double ux = div_any(fx != 0.0 ? moment<0>(f) + 0.5 * fx : moment<0>(f), rho, yrho);
  double uy = div_any(fy != 0.0 ? moment<1>(f) + 0.5 * fy : moment<1>(f), rho, yrho);
  double uz = div_any(fz != 0.0 ? moment<2>(f) + 0.5 * fz : moment<2>(f), rho, yrho);
  if (s) ux = uy = uz = 0.0;
  double feq[Q];
  equilibrium_all(Dirs{}, rho, ux, uy, uz, feq);
  double g[Q];
  if (FORCE) guo_all(Dirs{}, pref, ux, uy, uz, fx, fy, fz, g);
  double out[Q];
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    const double v = f[i] - div_pos(f[i] - feq[i], tau, ytau);
    out[i] = FORCE ? v + g[i] : v;
  }
  store_all<true, NT>(Dirs{}, fout, N, d, ix, iy, iz, out);
}

// Binary voxelization (scripts/run_3d_window_lbm.py:73-92,
// run_local_fines_lbm_permeability.py:39-59): one workgroup per sphere over
// its inclusive cell box; a cell is solid when
// ((ax - x)^2 + (ay - y)^2) + (az - z)^2 <= r^2 with the cell-centre
// coordinates ax, ay, az given per axis.  Every writer stores 1: no atomics.
__global__ __launch_bounds__(256) void k_voxelize_binary(const double* __restrict__ spheres,
                                                         const int32_t* __restrict__ boxes,
                                                         const double* __restrict__ axes, int nx, int ny, int nz,
                                                         uint8_t* __restrict__ solid) {
  const double* q = spheres + 4 * (size_t)blockIdx.x;
  const int32_t* b = boxes + 6 * (size_t)blockIdx.x;
  const int i0 = b[0], j0 = b[2], k0 = b[4];
  const int ei = b[1] - i0 + 1, ej = b[3] - j0 + 1, ek = b[5] - k0 + 1;
  if (ei <= 0 || ej <= 0 || ek <= 0) return;
  const double x = q[0], y = q[1], z = q[2], r = q[3];
  const double r2 = r * r;
  const double* ax = axes;
  const double* ay = axes + nx;
  const double* az = axes + nx + ny;
  const int vol = ei * ej * ek;
  for (int e = threadIdx.x; e < vol; e += blockDim.x) {
    const int lk = e % ek;
    const int t = e / ek;
    const int lj = t % ej;
    const int li = t / ej;
    const int i = i0 + li, j = j0 + lj, k = k0 + lk;
    const double dx = ax[i] - x, dy = ay[j] - y, dz = az[k] - z;
    if ((dx * dx + dy * dy) + dz * dz <= r2) solid[((size_t)i * ny + j) * nz + k] = 1;
  }
}

// ---------------------------------------------------------------------------
// K1: sub-sample voxelization (psc3d.py:60-94 / 151-173) and the solid
// velocity numerator (psc3d.py:190-201 / 352-367).
// ---------------------------------------------------------------------------
struct Particle {
  double x, y, z, r, vx, vy, vz, wx, wy, wz;
};

__device__ __forceinline__ Particle load_particle(const double* __restrict__ p, int i) {
  const double* q = p + 10 * (size_t)i;
  return Particle{q[0], q[1], q[2], q[3], q[4], q[5], q[6], q[7], q[8], q[9]};
}

// Count of samples (gx+o_a, gy+o_b, gz+o_c), o_k = (k+0.5)/s, with
// (dx*dx + dy*dy) + dz*dz <= r*r.  Cells whose extreme samples decide the
// answer for every sample are resolved without the s^3 loop: correctly
// rounded subtraction, squaring and addition are monotonic, so the farthest
// (nearest) sample bound computed with the same operations is exact.
template <int S>
__device__ __forceinline__ int sample_count_t(int gx, int gy, int gz, const Particle& p, double r2) {
  double dx2[S], dy2[S], dz2[S];
  double dx0 = 0, dx1 = 0, dy0 = 0, dy1 = 0, dz0 = 0, dz1 = 0;
#pragma unroll
  for (int k = 0; k < S; ++k) {
    const double o = ((double)k + 0.5) / (double)S;
    const double dx = ((double)gx + o) - p.x;
    const double dy = ((double)gy + o) - p.y;
    const double dz = ((double)gz + o) - p.z;
    dx2[k] = dx * dx;
    dy2[k] = dy * dy;
    dz2[k] = dz * dz;
    if (k == 0) { dx0 = dx; dy0 = dy; dz0 = dz; }
    if (k == S - 1) { dx1 = dx; dy1 = dy; dz1 = dz; }
  }
  const double mx = fmax(dx2[0], dx2[S - 1]), my = fmax(dy2[0], dy2[S - 1]), mz = fmax(dz2[0], dz2[S - 1]);
  if ((mx + my) + mz <= r2) return S * S * S;
  const double nxm = (dx0 <= 0.0 && dx1 >= 0.0) ? 0.0 : fmin(dx2[0], dx2[S - 1]);
  const double nym = (dy0 <= 0.0 && dy1 >= 0.0) ? 0.0 : fmin(dy2[0], dy2[S - 1]);
  const double nzm = (dz0 <= 0.0 && dz1 >= 0.0) ? 0.0 : fmin(dz2[0], dz2[S - 1]);
  if ((nxm + nym) + nzm > r2) return 0;
  int count = 0;
#pragma unroll
  for (int a = 0; a < S; ++a)
#pragma unroll
    for (int bb = 0; bb < S; ++bb) {
      const double sxy = dx2[a] + dy2[bb];
#pragma unroll
      for (int c = 0; c < S; ++c) count += (sxy + dz2[c] <= r2) ? 1 : 0;
    }
  return count;
}

__device__ int sample_count_any(int gx, int gy, int gz, const Particle& p, double r2, int s) {
  const double ds = (double)s;
  const double o0 = 0.5 / ds, o1 = ((double)(s - 1) + 0.5) / ds;
  const double dx0 = ((double)gx + o0) - p.x, dx1 = ((double)gx + o1) - p.x;
  const double dy0 = ((double)gy + o0) - p.y, dy1 = ((double)gy + o1) - p.y;
  const double dz0 = ((double)gz + o0) - p.z, dz1 = ((double)gz + o1) - p.z;
  const double mx = fmax(dx0 * dx0, dx1 * dx1), my = fmax(dy0 * dy0, dy1 * dy1), mz = fmax(dz0 * dz0, dz1 * dz1);
  if ((mx + my) + mz <= r2) return s * s * s;
  const double nxm = (dx0 <= 0.0 && dx1 >= 0.0) ? 0.0 : fmin(dx0 * dx0, dx1 * dx1);
  const double nym = (dy0 <= 0.0 && dy1 >= 0.0) ? 0.0 : fmin(dy0 * dy0, dy1 * dy1);
  const double nzm = (dz0 <= 0.0 && dz1 >= 0.0) ? 0.0 : fmin(dz0 * dz0, dz1 * dz1);
  if ((nxm + nym) + nzm > r2) return 0;
  int count = 0;
  for (int a = 0; a < s; ++a) {
    const double dx = ((double)gx + ((double)a + 0.5) / ds) - p.x;
    for (int bb = 0; bb < s; ++bb) {
      const double dy = ((double)gy + ((double)bb + 0.5) / ds) - p.y;
      const double sxy = dx * dx + dy * dy;
      for (int c = 0; c < s; ++c) {
        const double dz = ((double)gz + ((double)c + 0.5) / ds) - p.z;
        count += (sxy + dz * dz <= r2) ? 1 : 0;
      }
    }
  }
  return count;
}

// Sample count with the sub-sample count fixed at compile time (S = 0: s at
// run time), so that a kernel holds the registers of its own S only.
template <int S>
__device__ __forceinline__ int sample_count(int gx, int gy, int gz, const Particle& p, double r2, int s) {
  if constexpr (S == 0) return sample_count_any(gx, gy, gz, p, r2, s);
  else return sample_count_t<S>(gx, gy, gz, p, r2);
}

struct Box {
  int x0, y0, z0, ex, ey, ez, off;
};
__device__ __forceinline__ Box load_box(const int32_t* __restrict__ b, int p) {
  const int32_t* q = b + 8 * (size_t)p;
  return Box{q[0], q[1], q[2], q[3], q[4], q[5], q[6]};
}

// alpha = count / s^3 for counts up to kAlphaTable - 1, made by the same FP64
// division as the direct form (identical bits), looked up instead of divided.
constexpr int kAlphaTable = 513;

// Solid count of one box cell into count_sum and, by USP: 0 nothing else;
// 1 / 2 its solid-velocity numerator contribution into the FP32 / FP64
// planes us_num; 3 (owner table) particle pid into the owner slot usn[cell]
// (plain store: a cell covered by one sphere has one writer; a shared cell's
// slot is overwritten by k_shared_owners).  NEAR: the box lies within one
// period of the lattice on every axis (wrap by compare and add).
template <int USP, bool NEAR>
__device__ __forceinline__ uint32_t voxel_deposit(int gx, int gy, int gz, int cnt, const Particle& p, int pid,
                                                  const Dims d, double s3, const double* alpha, size_t N,
                                                  uint32_t* __restrict__ csum, void* __restrict__ usn) {
  const uint32_t cell = NEAR ? cell_id(d, wrap_near(gx, d.nx), wrap_near(gy, d.ny), wrap_near(gz, d.nz))
                             : cell_id(d, wrap_any(gx, d.nx), wrap_any(gy, d.ny), wrap_any(gz, d.nz));
  atomicAdd(csum + cell, (uint32_t)cnt);
  if (USP == 3) {
    ((uint32_t*)usn)[cell] = (uint32_t)pid;
  } else if (USP != 0) {
    const double a = alpha != nullptr ? alpha[cnt] : (double)cnt / s3;
    const double rx = ((double)gx + 0.5) - p.x;
    const double ry = ((double)gy + 0.5) - p.y;
    const double rz = ((double)gz + 0.5) - p.z;
    const double vx = (p.vx + p.wy * rz) - p.wz * ry;
    const double vy = (p.vy + p.wz * rx) - p.wx * rz;
    const double vz = (p.vz + p.wx * ry) - p.wy * rx;
    if (USP == 2) {
      double* u = (double*)usn;
      unsafeAtomicAdd(u + cell, a * vx);
      unsafeAtomicAdd(u + N + cell, a * vy);
      unsafeAtomicAdd(u + 2 * N + cell, a * vz);
    } else {
      float* u = (float*)usn;
      unsafeAtomicAdd(u + cell, (float)(a * vx));
      unsafeAtomicAdd(u + N + cell, (float)(a * vy));
      unsafeAtomicAdd(u + 2 * N + cell, (float)(a * vz));
    }
  }
  return cell;
}

// One wavefront per chunk: the chunk's cells in rounds of 64 (box order), the
// entries of each round compacted with a ballot, so entry order = box order
// (deterministic) and no workgroup barrier is needed.
constexpr int kVoxThreads = 64;

template <int USP, int S>
__global__ __launch_bounds__(kVoxThreads) void k_voxelize(const double* __restrict__ parts,
                                                          const int32_t* __restrict__ boxes,
                                                          const int32_t* __restrict__ chunks, int chunk_cells, Dims d,
                                                          int s, uint32_t* __restrict__ ent_cell,
                                                          uint32_t* __restrict__ ent_loc, uint16_t* __restrict__ ent_cnt,
                                                          int32_t* __restrict__ chunk_nnz, uint32_t* __restrict__ csum,
                                                          void* __restrict__ usn, uint32_t cell_base) {
  __shared__ double alpha_lds[(USP == 1 || USP == 2) ? kAlphaTable : 1];
  const int lane = (int)threadIdx.x;
  const int pid = chunks[2 * blockIdx.x];
  const int e0 = chunks[2 * blockIdx.x + 1];
  const Box bx = load_box(boxes, pid);
  const Particle p = load_particle(parts, pid);
  const int vol = bx.ex * bx.ey * bx.ez;
  const int e1 = min(e0 + chunk_cells, vol);
  const double r2 = p.r * p.r;
  const int s3i = s * s * s;
  const double s3 = (double)s3i;
  const size_t N = (size_t)d.ps;
  const size_t out0 = (size_t)blockIdx.x * chunk_cells;
  const double* alpha = nullptr;
  if ((USP == 1 || USP == 2) && s3i < kAlphaTable) {
    for (int i = lane; i <= s3i; i += kVoxThreads) alpha_lds[i] = (double)i / s3;
    __syncthreads();
    alpha = alpha_lds;
  }
  const bool near = bx.x0 >= -d.nx && bx.x0 + bx.ex <= 2 * d.nx && bx.y0 >= -d.ny && bx.y0 + bx.ey <= 2 * d.ny &&
                    bx.z0 >= -d.nz && bx.z0 + bx.ez <= 2 * d.nz;
  int written = 0;
  // box coordinates of entry e: FP32-reciprocal division (exact after
  // correction) while the box is below 2^24 cells, integer division beyond
  const bool fast = vol < (1 << 24);
  const float inv_ez = 1.0f / (float)bx.ez, inv_ey = 1.0f / (float)bx.ey;
  for (int r = e0; r < e1; r += kVoxThreads) {
    const int e = r + lane;
    int cnt = 0, lx = 0, ly = 0, lz = 0;
    uint32_t cell = 0;
    if (e < e1) {
      if (fast) {
        uint32_t t, qz, qy, qx;
        divmod_f((uint32_t)e, (uint32_t)bx.ez, inv_ez, t, qz);
        divmod_f(t, (uint32_t)bx.ey, inv_ey, qx, qy);
        lz = (int)qz;
        ly = (int)qy;
        lx = (int)qx;
      } else {
        lz = e % bx.ez;
        const int t = e / bx.ez;
        ly = t % bx.ey;
        lx = t / bx.ey;
      }
      const int gx = bx.x0 + lx, gy = bx.y0 + ly, gz = bx.z0 + lz;
      cnt = sample_count<S>(gx, gy, gz, p, r2, s);
      if (cnt) {
        cell = near ? voxel_deposit<USP, true>(gx, gy, gz, cnt, p, pid, d, s3, alpha, N, csum, usn)
                    : voxel_deposit<USP, false>(gx, gy, gz, cnt, p, pid, d, s3, alpha, N, csum, usn);
      }
    }
    const unsigned long long m = __ballot(cnt != 0);
    if (cnt) {
      const size_t o = out0 + written + __popcll(m & ((1ull << lane) - 1ull));
      ent_cell[o] = cell + cell_base;
      ent_loc[o] = (uint32_t)lx | ((uint32_t)ly << kLocBits) | ((uint32_t)lz << (2 * kLocBits));
      ent_cnt[o] = (uint16_t)cnt;
    }
    written += __popcll(m);
  }
  if (lane == 0) chunk_nnz[blockIdx.x] = written;
}

constexpr int kSharedUnroll = 4;  // entries per thread per round of the owner-table passes

// Owner table, second pass over the entries: a cell whose count sum differs
// from the count of one of its entries is covered by more than one sphere.
// Its owner slot is marked shared and each of its spheres adds
// alpha (U + omega x r) into the FP32 numerator planes there, exactly as the
// plane form (USP 1) does at every cell; cells with one owner are left alone.
__global__ __launch_bounds__(256) void k_shared_owners(const double* __restrict__ parts,
                                                       const int32_t* __restrict__ boxes,
                                                       const int32_t* __restrict__ chunks, int chunk_cells, Dims d,
                                                       int s, const uint32_t* __restrict__ ent_cell,
                                                       const uint32_t* __restrict__ ent_loc,
                                                       const uint16_t* __restrict__ ent_cnt,
                                                       const int32_t* __restrict__ chunk_nnz,
                                                       const uint32_t* __restrict__ csum, uint32_t* __restrict__ owner,
                                                       float* __restrict__ us_num) {
  const int c = blockIdx.x;
  const int nnz = chunk_nnz[c];
  const size_t out0 = (size_t)c * chunk_cells;
  const size_t N = (size_t)d.ps;
  const double s3 = (double)(s * s * s);
  const int pid = chunks[2 * c];
  const Box bx = load_box(boxes, pid);  // one sphere per chunk
  const Particle p = load_particle(parts, pid);
  constexpr uint32_t kLocMask = (1u << kLocBits) - 1u;
  // kSharedUnroll entries per thread per round, their loads issued together
  for (int k0 = threadIdx.x; k0 < nnz; k0 += kSharedUnroll * blockDim.x) {
    uint32_t cell[kSharedUnroll], tot[kSharedUnroll];
    int cnt[kSharedUnroll];
#pragma unroll
    for (int j = 0; j < kSharedUnroll; ++j) {
      const int k = k0 + j * (int)blockDim.x;
      cell[j] = k < nnz ? ent_cell[out0 + k] : 0u;
      cnt[j] = k < nnz ? (int)ent_cnt[out0 + k] : 0;
    }
#pragma unroll
    for (int j = 0; j < kSharedUnroll; ++j) tot[j] = cnt[j] ? csum[cell[j]] : 0u;
#pragma unroll
    for (int j = 0; j < kSharedUnroll; ++j) {
      if (cnt[j] == 0 || tot[j] == (uint32_t)cnt[j]) continue;  // past the entries, or one owner
      const int k = k0 + j * (int)blockDim.x;
      owner[cell[j]] = kSharedCell;
      const uint32_t loc = ent_loc[out0 + k];
      const int gx = bx.x0 + (int)(loc & kLocMask);
      const int gy = bx.y0 + (int)((loc >> kLocBits) & kLocMask);
      const int gz = bx.z0 + (int)((loc >> (2 * kLocBits)) & kLocMask);
      const double a = (double)cnt[j] / s3;
      const double rx = ((double)gx + 0.5) - p.x;
      const double ry = ((double)gy + 0.5) - p.y;
      const double rz = ((double)gz + 0.5) - p.z;
      const double vx = (p.vx + p.wy * rz) - p.wz * ry;
      const double vy = (p.vy + p.wz * rx) - p.wx * rz;
      const double vz = (p.vz + p.wx * ry) - p.wy * rx;
      unsafeAtomicAdd(us_num + cell[j], (float)(a * vx));
      unsafeAtomicAdd(us_num + N + cell[j], (float)(a * vy));
      unsafeAtomicAdd(us_num + 2 * N + cell[j], (float)(a * vz));
    }
  }
}

// Owner table: zero count_sum at the cells of a previous voxelization's
// entries, and the numerator planes at those of them that were shared (count
// sum != the entry's count; a thread that finds the sum already zeroed by
// another entry of the same cell zeroes the planes too).  The owner slots
// need no clearing: they are read only where count_sum > 0, and every
// voxelization rewrites them there.
__global__ __launch_bounds__(256) void k_clear_owners(const uint32_t* __restrict__ ent_cell,
                                                      const uint16_t* __restrict__ ent_cnt,
                                                      const int32_t* __restrict__ chunk_nnz, int chunk_cells, size_t N,
                                                      uint32_t* __restrict__ csum, float* __restrict__ us_num) {
  const int c = blockIdx.x;
  const int nnz = chunk_nnz[c];
  const size_t out0 = (size_t)c * chunk_cells;
  for (int k0 = threadIdx.x; k0 < nnz; k0 += kSharedUnroll * blockDim.x) {
    uint32_t cell[kSharedUnroll], tot[kSharedUnroll];
    int cnt[kSharedUnroll];
#pragma unroll
    for (int j = 0; j < kSharedUnroll; ++j) {
      const int k = k0 + j * (int)blockDim.x;
      cell[j] = k < nnz ? ent_cell[out0 + k] : 0u;
      cnt[j] = k < nnz ? (int)ent_cnt[out0 + k] : 0;
    }
#pragma unroll
    for (int j = 0; j < kSharedUnroll; ++j) tot[j] = cnt[j] ? csum[cell[j]] : 0u;
#pragma unroll
    for (int j = 0; j < kSharedUnroll; ++j) {
      if (cnt[j] == 0) continue;  // past the entries
      if (tot[j] != (uint32_t)cnt[j]) {
        us_num[cell[j]] = 0.0f;
        us_num[N + cell[j]] = 0.0f;
        us_num[2 * N + cell[j]] = 0.0f;
      }
      csum[cell[j]] = 0u;
    }
  }
}

// Zero count_sum (and the solid-velocity numerator) at the cells listed in
// a previous voxelization's entries: the incremental alternative to clearing
// the whole lattice before re-voxelizing moving spheres.
template <int USP>
__global__ __launch_bounds__(256) void k_clear_entries(const uint32_t* __restrict__ ent_cell,
                                                       const int32_t* __restrict__ chunk_nnz, int chunk_cells,
                                                       size_t N, uint32_t* __restrict__ csum,
                                                       void* __restrict__ usn) {
  const int c = blockIdx.x;
  const int nnz = chunk_nnz[c];
  const size_t out0 = (size_t)c * chunk_cells;
  for (int k = threadIdx.x; k < nnz; k += blockDim.x) {
    const uint32_t cell = ent_cell[out0 + k];
    csum[cell] = 0u;
    if (USP == 1) {
      float* u = (float*)usn;
      u[cell] = 0.0f;
      u[N + cell] = 0.0f;
      u[2 * N + cell] = 0.0f;
    } else if (USP == 2) {
      double* u = (double*)usn;
      u[cell] = 0.0;
      u[N + cell] = 0.0;
      u[2 * N + cell] = 0.0;
    }
  }
}

// ---------------------------------------------------------------------------
// K3: node-summed force/torque per particle (psc3d.py:255-281, 375-407).
// Pass 1: one workgroup per chunk sums its compacted entries in FP64.  Pass
// 2: one wavefront per particle adds its chunk partials.  Fixed thread
// assignment and reduction trees: deterministic.  Bodies in psc_loads.h.
// ---------------------------------------------------------------------------
// Threads per chunk workgroup (64, 128 or 256: the same bits, see psc_loads.h).
#ifndef PSC_LOADS_THREADS
#define PSC_LOADS_THREADS 256
#endif

// Workgroups of the chunk pass: 0 = one per chunk; otherwise at most this
// many, each summing chunks blockIdx.x, blockIdx.x + gridDim.x, ... (the
// per-chunk sums and their bits do not depend on the grid).
#ifndef PSC_LOADS_GRID
#define PSC_LOADS_GRID 0
#endif

// Entries per thread per round of the chunk pass (a register batch: the sums
// and their bits do not depend on it).
#ifndef PSC_LOADS_KB
#define PSC_LOADS_KB kLoadsKB
#endif

template <typename TI>
__global__ __launch_bounds__(PSC_LOADS_THREADS) void k_loads_chunk(LoadsArgs a, const TI* __restrict__ imb, Dims d,
                                                                   double* __restrict__ partials) {
  __shared__ double lds[6 * 4];
  for (int c = blockIdx.x; c < a.nchunks; c += gridDim.x)
    loads_chunk_body<TI, PSC_LOADS_KB>(a, imb, d.nx, d.ny, d.nz, c, threadIdx.x, blockDim.x, partials, lds);
}

// The loads still pending at the end of a fused run, in one launch:
// workgroup b < nchunks sums chunk b of the last step's imb into part_last,
// workgroup nchunks + p adds up particle p's partials of the step before
// (part_prev) into loads.  The two halves touch different buffers; the last
// step's per-particle pass follows as k_loads_final.
template <typename TI>
__global__ __launch_bounds__(PSC_LOADS_THREADS) void k_loads_flush(LoadsArgs a, const TI* __restrict__ imb, Dims d,
                                                                   double* __restrict__ part_last,
                                                                   const double* __restrict__ part_prev,
                                                                   double* __restrict__ loads) {
  __shared__ double lds[6 * 4];
  const int b = blockIdx.x;
  if (b < a.nchunks)
    loads_chunk_body<TI, PSC_LOADS_KB>(a, imb, d.nx, d.ny, d.nz, b, threadIdx.x, blockDim.x, part_last, lds);
  else
    loads_final_body(a.chunk_first, b - a.nchunks, threadIdx.x, part_prev, loads);
}

inline int loads_grid(int nchunks) {
  return (PSC_LOADS_GRID > 0 && PSC_LOADS_GRID < nchunks) ? PSC_LOADS_GRID : nchunks;
}

__global__ __launch_bounds__(64) void k_loads_final(const int32_t* __restrict__ chunk_first,
                                                    const double* __restrict__ partials,
                                                    double* __restrict__ loads) {
  loads_final_body(chunk_first, blockIdx.x, threadIdx.x, partials, loads);
}

// ---------------------------------------------------------------------------
// K4: post-stream moments and FP64 readout sums (psc3d.py:243-253).
// ---------------------------------------------------------------------------
constexpr int kMomThreads = 256;
constexpr int64_t kMomMaxBlocks = 2048;

template <typename TS, bool DEV, bool PULL, bool HAVE_STATE, typename TI>
__global__ __launch_bounds__(kMomThreads) void k_moments(const TS* __restrict__ st, const uint32_t* __restrict__ csum,
                                                         const TI* __restrict__ imb, Dims d, double s3, int afield,
                                                         double hfx, double hfy, double hfz,
                                                         double* __restrict__ rho_u, double* __restrict__ partials) {
  __shared__ double lds[PSC_NSUM * 16];
  const size_t N = (size_t)d.ps;                            // plane stride
  const size_t ncell = (size_t)d.nx * d.ny * d.pz;          // cells incl. row padding
  double acc[PSC_NSUM];
#pragma unroll
  for (int k = 0; k < PSC_NSUM; ++k) acc[k] = 0.0;
  for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < ncell; c += (size_t)gridDim.x * blockDim.x) {
    if (d.pz != d.nz && (int)(c % d.pz) >= d.nz) continue;  // row padding
    // afield: 0 sub-sample counts, 1 FP32 / 2 FP64 solid-fraction field
    double solid;
    if (afield == 0) {
      const double asum = (double)csum[c] / s3;
      solid = asum < 1.0 ? asum : 1.0;
    } else {
      const double a = afield == 1 ? (double)((const float*)csum)[c] : ((const double*)csum)[c];
      solid = a < 1.0 ? a : 1.0;
    }
    const double fl = 1.0 - solid;
    acc[0] += solid;
    acc[1] += fl;
    if (imb != nullptr && solid != 0.0) {
      acc[8] += (double)imb[3 * c];
      acc[9] += (double)imb[3 * c + 1];
      acc[10] += (double)imb[3 * c + 2];
    }
    if (HAVE_STATE) {
      const int iz = (int)(c % d.pz);
      const size_t t = c / d.pz;
      const int iy = (int)(t % d.ny);
      const int ix = (int)(t / d.ny);
      double f[Q];
#pragma unroll
      for (int i = 0; i < Q; ++i) {
        const uint32_t src = PULL ? cell_id(d, wrap_step(ix, -cx_of(i), d.nx), wrap_step(iy, -cy_of(i), d.ny),
                                            wrap_step(iz, -cz_of(i), d.nz))
                                  : (uint32_t)c;
        f[i] = (double)st[(size_t)i * N + src];
      }
      double rho;
      if (DEV) {
        // rho = 1 + sum g
        double s = f[0];
#pragma unroll
        for (int i = 1; i < Q; ++i) s = s + f[i];
        rho = 1.0 + s;
      } else {
        rho = f[0];
#pragma unroll
        for (int i = 1; i < Q; ++i) rho = rho + f[i];
      }
      const double ux = (moment<0>(f) + hfx) / rho;
      const double uy = (moment<1>(f) + hfy) / rho;
      const double uz = (moment<2>(f) + hfz) / rho;
      if (rho_u != nullptr) {  // fields: planes of ncell (no row padding here)
        rho_u[c] = rho;
        rho_u[ncell + c] = ux;
        rho_u[2 * ncell + c] = uy;
        rho_u[3 * ncell + c] = uz;
      }
      const bool finite = isfinite(rho) && isfinite(ux) && isfinite(uy) && isfinite(uz);
      acc[2] += ux;
      acc[3] += uy;
      acc[4] += uz;
      acc[5] += fl * ux;
      acc[6] += fl * uy;
      acc[7] += fl * uz;
      acc[11] += rho;
      acc[12] += finite ? 0.0 : 1.0;
    }
  }
  block_sum<PSC_NSUM>(acc, lds);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < PSC_NSUM; ++k) partials[(size_t)PSC_NSUM * blockIdx.x + k] = acc[k];
  }
}

__global__ __launch_bounds__(256) void k_sum_partials(const double* __restrict__ partials, int nblk,
                                                      double* __restrict__ out) {
  __shared__ double lds[PSC_NSUM * 16];
  double acc[PSC_NSUM];
#pragma unroll
  for (int k = 0; k < PSC_NSUM; ++k) acc[k] = 0.0;
  for (int b = threadIdx.x; b < nblk; b += blockDim.x)
#pragma unroll
    for (int k = 0; k < PSC_NSUM; ++k) acc[k] += partials[(size_t)PSC_NSUM * b + k];
  block_sum<PSC_NSUM>(acc, lds);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < PSC_NSUM; ++k) out[k] = acc[k];
  }
}

// ---------------------------------------------------------------------------
// f export / import and state init
// ---------------------------------------------------------------------------
template <typename TS, bool DEV, bool PULL>
__global__ void k_export(const TS* __restrict__ st, Dims d, double* __restrict__ fout) {
  const size_t N = (size_t)d.ps;
  for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < N; c += (size_t)gridDim.x * blockDim.x) {
    const int iz = (int)(c % d.pz);
    const size_t t = c / d.pz;
    const int iy = (int)(t % d.ny);
    const int ix = (int)(t / d.ny);
#pragma unroll
    for (int i = 0; i < Q; ++i) {
      const uint32_t src = PULL ? cell_id(d, wrap_step(ix, -cx_of(i), d.nx), wrap_step(iy, -cy_of(i), d.ny),
                                          wrap_step(iz, -cz_of(i), d.nz))
                                : (uint32_t)c;
      const double v = (double)st[(size_t)i * N + src];
      fout[(size_t)i * N + c] = DEV ? w_of(i) + v : v;
    }
  }
}

template <typename TS, bool DEV, bool PULL>
__global__ void k_import(const double* __restrict__ fin, Dims d, TS* __restrict__ st) {
  const size_t N = (size_t)d.ps;
  for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < N; c += (size_t)gridDim.x * blockDim.x) {
    const int iz = (int)(c % d.pz);
    const size_t t = c / d.pz;
    const int iy = (int)(t % d.ny);
    const int ix = (int)(t / d.ny);
#pragma unroll
    for (int i = 0; i < Q; ++i) {
      // the state at y holds the value that export reads back at y + c_i
      const uint32_t src = PULL ? cell_id(d, wrap_step(ix, cx_of(i), d.nx), wrap_step(iy, cy_of(i), d.ny),
                                          wrap_step(iz, cz_of(i), d.nz))
                                : (uint32_t)c;
      const double v = fin[(size_t)i * N + src];
      st[(size_t)i * N + c] = (TS)(DEV ? v - w_of(i) : v);
    }
  }
}

__global__ void k_init_plain(double* __restrict__ f, size_t N) {
  for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < N; c += (size_t)gridDim.x * blockDim.x) {
#pragma unroll
    for (int i = 0; i < Q; ++i) f[(size_t)i * N + c] = w_of(i) * 1.0;
  }
}

// Module functions equilibrium / guo_force (psc3d.py:97-114) over n points:
// out[i * n + c], the reference's (19, *shape) layout, in its operation order.
__global__ void k_equilibrium_f64(const double* __restrict__ rho, const double* __restrict__ ux,
                                  const double* __restrict__ uy, const double* __restrict__ uz, size_t n,
                                  double* __restrict__ out) {
  for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (size_t)gridDim.x * blockDim.x) {
    double f[Q];
    equilibrium_all(Dirs{}, rho[c], ux[c], uy[c], uz[c], f);
#pragma unroll
    for (int i = 0; i < Q; ++i) out[(size_t)i * n + c] = f[i];
  }
}

__global__ void k_guo_f64(double fx, double fy, double fz, const double* __restrict__ ux,
                          const double* __restrict__ uy, const double* __restrict__ uz, size_t n, double pref,
                          double* __restrict__ out) {
  for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (size_t)gridDim.x * blockDim.x) {
    double g[Q];
    guo_all(Dirs{}, pref, ux[c], uy[c], uz[c], fx, fy, fz, g);
#pragma unroll
    for (int i = 0; i < Q; ++i) out[(size_t)i * n + c] = g[i];
  }
}

int grid_for(size_t n, int threads, int cap = 8192) {
  size_t g = (n + threads - 1) / threads;
  if (g > (size_t)cap) g = cap;
  if (g < 1) g = 1;
  return (int)g;
}

}  // namespace

// ------------------------------- launchers ---------------------------------
int launch_bb_step_f64(const double* in, double* out, const uint8_t* solid, int nx, int ny, int nz, double tau,
                       double fx, double fy, double fz, int bz, int by, int flags, hipStream_t stream) {
  const Dims d{nx, ny, nz, nz, (unsigned long long)nx * ny * nz};
  const dim3 block(bz, by, 1);
  const dim3 grid((nz + bz - 1) / bz, (ny + by - 1) / by, nx);
  const double pref = 1.0 - 0.5 / tau;
  const double ytau = 1.0 / tau;  // RN(1/tau): host IEEE division
  const bool force = fx != 0.0 || fy != 0.0 || fz != 0.0;
  const bool nt = (flags & PSC_STEP_NONTEMPORAL) != 0;  // the only launch flag of the FP64 step
  if (force && nt) k_bb_step_f64<true, true><<<grid, block, 0, stream>>>(in, out, solid, d, tau, ytau, pref, fx, fy, fz);
  else if (force) k_bb_step_f64<true, false><<<grid, block, 0, stream>>>(in, out, solid, d, tau, ytau, pref, fx, fy, fz);
  else if (nt) k_bb_step_f64<false, true><<<grid, block, 0, stream>>>(in, out, solid, d, tau, ytau, pref, fx, fy, fz);
  else k_bb_step_f64<false, false><<<grid, block, 0, stream>>>(in, out, solid, d, tau, ytau, pref, fx, fy, fz);
  return check_launch("psc_bb_step(f64)");
}

int launch_voxelize_binary(const double* spheres, const int32_t* boxes, int nspheres, const double* axes, int nx,
                           int ny, int nz, uint8_t* solid, hipStream_t stream) {
  if (nspheres == 0) return PSC_OK;
  k_voxelize_binary<<<nspheres, 256, 0, stream>>>(spheres, boxes, axes, nx, ny, nz, solid);
  return check_launch("psc_voxelize_binary");
}

int launch_step_f64(const double* in, double* out, const uint32_t* csum, const double* usn, double* imb,
                    int layout, int nx, int ny, int nz, int s, double tau, double fx, double fy, double fz,
                    int bz, int by, int flags, hipStream_t stream) {
  const Dims d{nx, ny, nz, nz, (unsigned long long)nx * ny * nz};
  const dim3 block(bz, by, 1);
  const dim3 grid((nz + bz - 1) / bz, (ny + by - 1) / by, nx);
  const double s3 = (double)(s * s * s);  // 0: csum is an FP64 solid-fraction field
  const double ys3 = 1.0 / s3, ytau = 1.0 / tau;  // RN(1/s^3), RN(1/tau): host IEEE divisions
  const double pref = 1.0 - 0.5 / tau;
  const bool force = (fx != 0.0 || fy != 0.0 || fz != 0.0);
  const bool moving = usn != nullptr;
  const bool pull = layout == PSC_PULL;
  const bool wimb = imb != nullptr;
  const bool nt = (flags & PSC_STEP_NONTEMPORAL) != 0;  // the only launch flag of the FP64 step
#define PSC_F64_CASE(P, M, F, I)                                                                             \
  if (pull == P && moving == M && force == F && wimb == I) {                                               \
    if (nt)                                                                                                \
      k_step_f64<P, M, F, I, true><<<grid, block, 0, stream>>>(in, out, csum, usn, imb, d, s3, ys3, tau, ytau, \
                                                               pref, fx, fy, fz);                           \
    else                                                                                                   \
      k_step_f64<P, M, F, I, false><<<grid, block, 0, stream>>>(in, out, csum, usn, imb, d, s3, ys3, tau,     \
                                                                ytau, pref, fx, fy, fz);                    \
    return check_launch("psc_step(f64)");                                                                  \
  }
  PSC_F64_CASE(false, false, false, false)
  PSC_F64_CASE(false, false, false, true)
  PSC_F64_CASE(false, false, true, false)
  PSC_F64_CASE(false, false, true, true)
  PSC_F64_CASE(false, true, false, false)
  PSC_F64_CASE(false, true, false, true)
  PSC_F64_CASE(false, true, true, false)
  PSC_F64_CASE(false, true, true, true)
  PSC_F64_CASE(true, false, false, false)
  PSC_F64_CASE(true, false, false, true)
  PSC_F64_CASE(true, false, true, false)
  PSC_F64_CASE(true, false, true, true)
  PSC_F64_CASE(true, true, false, false)
  PSC_F64_CASE(true, true, false, true)
  PSC_F64_CASE(true, true, true, false)
  PSC_F64_CASE(true, true, true, true)
#undef PSC_F64_CASE
  return set_error(PSC_EINVAL, "psc_step(f64): no kernel variant");
}

int launch_voxelize(const double* parts, const int32_t* boxes, const int32_t* chunks, int nchunks, int chunk_cells,
                    int nx, int ny, int nz, int pz, int s, uint32_t* ent_cell, uint32_t* ent_loc, uint16_t* ent_cnt,
                    int32_t* chunk_nnz, uint32_t* csum, void* usn, int usp, uint32_t cell_base,
                    hipStream_t stream) {
  if (nchunks == 0) return PSC_OK;
  // us_num: three planes of nx*ny*pz cells, the planes the step kernels and
  // the incremental clear read (count_sum is a single plane)
  const Dims d{nx, ny, nz, pz, (unsigned long long)nx * ny * pz};
#define PSC_VOX(U, S)                                                                                       \
  k_voxelize<U, S><<<nchunks, kVoxThreads, 0, stream>>>(parts, boxes, chunks, chunk_cells, d, s, ent_cell, ent_loc, \
                                                        ent_cnt, chunk_nnz, csum, U ? usn : nullptr, cell_base)
#define PSC_VOX_S(U)             \
  switch (s) {                   \
    case 1: PSC_VOX(U, 1); break; \
    case 2: PSC_VOX(U, 2); break; \
    case 3: PSC_VOX(U, 3); break; \
    case 4: PSC_VOX(U, 4); break; \
    case 5: PSC_VOX(U, 5); break; \
    case 6: PSC_VOX(U, 6); break; \
    default: PSC_VOX(U, 0); break; \
  }
  if (usn == nullptr || usp == 0) {
    PSC_VOX_S(0)
  } else if (usp == 1) {
    PSC_VOX_S(1)
  } else {
    PSC_VOX_S(2)
  }
#undef PSC_VOX_S
#undef PSC_VOX
  return check_launch("psc_voxelize");
}

int launch_loads(const double* parts, const int32_t* boxes, const int32_t* chunks, const int32_t* chunk_first, int np,
                 int nchunks, int chunk_cells, const uint32_t* ent_cell, const uint32_t* ent_loc,
                 const uint16_t* ent_cnt, const int32_t* chunk_nnz, const uint32_t* csum, const void* imb, int imbp,
                 int nx, int ny, int nz, double* partials, double* loads, hipStream_t stream) {
  if (np == 0) return PSC_OK;
  const Dims d{nx, ny, nz, nz, (unsigned long long)nx * ny * nz};
  const LoadsArgs la{parts, boxes, chunks, chunk_first, np, nchunks, chunk_cells, ent_cell, ent_loc, ent_cnt,
                     chunk_nnz, csum};
  if (nchunks > 0) {
    if (imbp == 2)
      k_loads_chunk<double><<<loads_grid(nchunks), PSC_LOADS_THREADS, 0, stream>>>(la, (const double*)imb, d, partials);
    else
      k_loads_chunk<float><<<loads_grid(nchunks), PSC_LOADS_THREADS, 0, stream>>>(la, (const float*)imb, d, partials);
    const int rc = check_launch("psc_particle_loads(chunks)");
    if (rc) return rc;
  }
  k_loads_final<<<np, 64, 0, stream>>>(chunk_first, partials, loads);
  return check_launch("psc_particle_loads(final)");
}

int launch_loads_chunks(const LoadsArgs& la, const void* imb, int imbp, int nx, int ny, int nz, double* partials,
                        hipStream_t stream) {
  if (la.nchunks == 0) return PSC_OK;
  const Dims d{nx, ny, nz, nz, (unsigned long long)nx * ny * nz};
  if (imbp == 2) k_loads_chunk<double><<<loads_grid(la.nchunks), PSC_LOADS_THREADS, 0, stream>>>(la, (const double*)imb, d, partials);
  else k_loads_chunk<float><<<loads_grid(la.nchunks), PSC_LOADS_THREADS, 0, stream>>>(la, (const float*)imb, d, partials);
  return check_launch("psc_particle_loads(chunks)");
}

int launch_loads_flush(const LoadsArgs& la, const void* imb, int imbp, int nx, int ny, int nz, double* part_last,
                       const double* part_prev, double* loads, hipStream_t stream) {
  const long long nb = (long long)la.nchunks + la.np;
  if (nb == 0) return PSC_OK;
  if (nb > 0x7fffffffll) return set_error(PSC_EINVAL, "psc_run_steps: too many chunks and particles");
  const Dims d{nx, ny, nz, nz, (unsigned long long)nx * ny * nz};
  if (imbp == 2)
    k_loads_flush<double><<<(int)nb, PSC_LOADS_THREADS, 0, stream>>>(la, (const double*)imb, d, part_last, part_prev,
                                                                     loads);
  else
    k_loads_flush<float><<<(int)nb, PSC_LOADS_THREADS, 0, stream>>>(la, (const float*)imb, d, part_last, part_prev,
                                                                    loads);
  return check_launch("psc_run_steps(pending loads)");
}

int launch_loads_final(const int32_t* chunk_first, int np, const double* partials, double* loads, hipStream_t stream) {
  if (np == 0) return PSC_OK;
  k_loads_final<<<np, 64, 0, stream>>>(chunk_first, partials, loads);
  return check_launch("psc_particle_loads(final)");
}

// Motion records of the owner table (MotionRecord): one thread per particle,
// from the particle and box tables of the voxelization.
__global__ __launch_bounds__(256) void k_motion_table(const double* __restrict__ parts,
                                                      const int32_t* __restrict__ boxes, int np,
                                                      MotionRecord* __restrict__ motion) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  const Particle p = load_particle(parts, i);
  const Box b = load_box(boxes, i);
  MotionRecord m;
  m.x0 = b.x0;
  m.y0 = b.y0;
  m.z0 = b.z0;
  m.pad = 0;
  m.cx = p.x;
  m.cy = p.y;
  m.cz = p.z;
  m.vx = p.vx;
  m.vy = p.vy;
  m.vz = p.vz;
  m.wx = p.wx;
  m.wy = p.wy;
  m.wz = p.wz;
  m.pad2 = 0.0;
  motion[i] = m;
}

int launch_voxelize_owners(const double* parts, const int32_t* boxes, int np, const int32_t* chunks, int nchunks,
                           int chunk_cells, int nx, int ny, int nz, int s, uint32_t* ent_cell, uint32_t* ent_loc,
                           uint16_t* ent_cnt, int32_t* chunk_nnz, uint32_t* csum, uint32_t* owner, float* us_num,
                           void* motion, hipStream_t stream) {
  if (np > 0) {
    k_motion_table<<<(np + 255) / 256, 256, 0, stream>>>(parts, boxes, np, (MotionRecord*)motion);
    const int rc = check_launch("psc_voxelize_owners(motion table)");
    if (rc) return rc;
  }
  if (nchunks == 0) return PSC_OK;
  const Dims d{nx, ny, nz, nz, (unsigned long long)nx * ny * nz};
  void* own = owner;
#define PSC_VOX_OWN(S)                                                                                            \
  k_voxelize<3, S><<<nchunks, kVoxThreads, 0, stream>>>(parts, boxes, chunks, chunk_cells, d, s, ent_cell, ent_loc, \
                                                        ent_cnt, chunk_nnz, csum, own, 0u)
  switch (s) {
    case 1: PSC_VOX_OWN(1); break;
    case 2: PSC_VOX_OWN(2); break;
    case 3: PSC_VOX_OWN(3); break;
    case 4: PSC_VOX_OWN(4); break;
    case 5: PSC_VOX_OWN(5); break;
    case 6: PSC_VOX_OWN(6); break;
    default: PSC_VOX_OWN(0); break;
  }
#undef PSC_VOX_OWN
  const int rc = check_launch("psc_voxelize_owners");
  if (rc) return rc;
  k_shared_owners<<<nchunks, 256, 0, stream>>>(parts, boxes, chunks, chunk_cells, d, s, ent_cell, ent_loc, ent_cnt,
                                               chunk_nnz, csum, owner, us_num);
  return check_launch("psc_voxelize_owners(shared cells)");
}

int launch_clear_owners(const uint32_t* ent_cell, const uint16_t* ent_cnt, const int32_t* chunk_nnz, int nchunks,
                        int chunk_cells, int nx, int ny, int nz, uint32_t* csum, float* us_num, hipStream_t stream) {
  if (nchunks == 0) return PSC_OK;
  k_clear_owners<<<nchunks, 256, 0, stream>>>(ent_cell, ent_cnt, chunk_nnz, chunk_cells, (size_t)nx * ny * nz, csum,
                                              us_num);
  return check_launch("psc_clear_entries_owners");
}

int launch_clear_entries(const uint32_t* ent_cell, const int32_t* chunk_nnz, int nchunks, int chunk_cells, int nx,
                         int ny, int nz, uint32_t* csum, void* usn, int usp, hipStream_t stream) {
  if (nchunks == 0) return PSC_OK;
  const size_t N = (size_t)nx * ny * nz;
  if (usn == nullptr || usp == 0)
    k_clear_entries<0><<<nchunks, 256, 0, stream>>>(ent_cell, chunk_nnz, chunk_cells, N, csum, nullptr);
  else if (usp == 1)
    k_clear_entries<1><<<nchunks, 256, 0, stream>>>(ent_cell, chunk_nnz, chunk_cells, N, csum, usn);
  else
    k_clear_entries<2><<<nchunks, 256, 0, stream>>>(ent_cell, chunk_nnz, chunk_cells, N, csum, usn);
  return check_launch("psc_clear_entries");
}

int64_t moments_blocks(int64_t n) {
  int64_t b = (n + (int64_t)kMomThreads * 8 - 1) / ((int64_t)kMomThreads * 8);
  if (b > kMomMaxBlocks) b = kMomMaxBlocks;
  if (b < 1) b = 1;
  return b;
}

int launch_moments(const void* st, int storage, int layout, const uint32_t* csum, const void* imb, int nx, int ny,
                   int nz, int pz, unsigned long long ps, int s, double fx, double fy, double fz, double* rho_u,
                   double* partials, double* sums, hipStream_t stream) {
  const Dims d{nx, ny, nz, pz, ps};
  const size_t N = (size_t)nx * ny * pz;
  const int nblk = (int)moments_blocks((int64_t)N);
  // s == 0: csum is a solid-fraction field in the storage precision
  const double s3 = (double)(s * s * s);
  const int af = s == 0 ? (storage == PSC_F32_DEV ? 1 : 2) : 0;
  const double hx = 0.5 * fx, hy = 0.5 * fy, hz = 0.5 * fz;
  const bool pull = layout == PSC_PULL;
  if (storage == PSC_F32_DEV) {
    const float* g = (const float*)st;
    const float* im = (const float*)imb;
    if (g == nullptr)
      k_moments<float, true, false, false, float><<<nblk, kMomThreads, 0, stream>>>(g, csum, im, d, s3, af, hx, hy, hz,
                                                                                    rho_u, partials);
    else if (pull)
      k_moments<float, true, true, true, float><<<nblk, kMomThreads, 0, stream>>>(g, csum, im, d, s3, af, hx, hy, hz,
                                                                                  rho_u, partials);
    else
      k_moments<float, true, false, true, float><<<nblk, kMomThreads, 0, stream>>>(g, csum, im, d, s3, af, hx, hy, hz,
                                                                                   rho_u, partials);
  } else {
    const double* f = (const double*)st;
    const double* im = (const double*)imb;
    if (f == nullptr)
      k_moments<double, false, false, false, double><<<nblk, kMomThreads, 0, stream>>>(f, csum, im, d, s3, af, hx, hy,
                                                                                       hz, rho_u, partials);
    else if (pull)
      k_moments<double, false, true, true, double><<<nblk, kMomThreads, 0, stream>>>(f, csum, im, d, s3, af, hx, hy, hz,
                                                                                     rho_u, partials);
    else
      k_moments<double, false, false, true, double><<<nblk, kMomThreads, 0, stream>>>(f, csum, im, d, s3, af, hx, hy,
                                                                                      hz, rho_u, partials);
  }
  int rc = check_launch("psc_moments");
  if (rc) return rc;
  k_sum_partials<<<1, 256, 0, stream>>>(partials, nblk, sums);
  return check_launch("psc_moments(final)");
}

int launch_export(const void* st, int storage, int layout, int nx, int ny, int nz, double* f, hipStream_t stream) {
  const Dims d{nx, ny, nz, nz, (unsigned long long)nx * ny * nz};
  const size_t N = (size_t)nx * ny * nz;
  const int g = grid_for(N, 256);
  const bool pull = layout == PSC_PULL;
  if (storage == PSC_F32_DEV) {
    if (pull) k_export<float, true, true><<<g, 256, 0, stream>>>((const float*)st, d, f);
    else k_export<float, true, false><<<g, 256, 0, stream>>>((const float*)st, d, f);
  } else {
    if (pull) k_export<double, false, true><<<g, 256, 0, stream>>>((const double*)st, d, f);
    else k_export<double, false, false><<<g, 256, 0, stream>>>((const double*)st, d, f);
  }
  return check_launch("psc_export_f64");
}

int launch_import(const double* f, int storage, int layout, int nx, int ny, int nz, void* st, hipStream_t stream) {
  const Dims d{nx, ny, nz, nz, (unsigned long long)nx * ny * nz};
  const size_t N = (size_t)nx * ny * nz;
  const int g = grid_for(N, 256);
  const bool pull = layout == PSC_PULL;
  if (storage == PSC_F32_DEV) {
    if (pull) k_import<float, true, true><<<g, 256, 0, stream>>>(f, d, (float*)st);
    else k_import<float, true, false><<<g, 256, 0, stream>>>(f, d, (float*)st);
  } else {
    if (pull) k_import<double, false, true><<<g, 256, 0, stream>>>(f, d, (double*)st);
    else k_import<double, false, false><<<g, 256, 0, stream>>>(f, d, (double*)st);
  }
  return check_launch("psc_import_f64");
}

int launch_equilibrium_f64(const double* rho, const double* ux, const double* uy, const double* uz, size_t n,
                           double* out, hipStream_t stream) {
  k_equilibrium_f64<<<grid_for(n, 256), 256, 0, stream>>>(rho, ux, uy, uz, n, out);
  return check_launch("psc_equilibrium");
}

int launch_guo_f64(double fx, double fy, double fz, const double* ux, const double* uy, const double* uz, size_t n,
                   double tau, double* out, hipStream_t stream) {
  // prefactor 1 - 0.5/tau as psc3d.py:107 forms it
  k_guo_f64<<<grid_for(n, 256), 256, 0, stream>>>(fx, fy, fz, ux, uy, uz, n, 1.0 - 0.5 / tau, out);
  return check_launch("psc_guo_force");
}

int launch_init_plain(double* f, size_t N, hipStream_t stream) {
  k_init_plain<<<grid_for(N, 256), 256, 0, stream>>>(f, N);
  return check_launch("psc_init_state");
}

}  // namespace psc

// Scratch memory of psc_moments_noalloc: the FP64 block partials of
// k_moments, kMomMaxBlocks blocks x PSC_NSUM sums (208 KiB), which every
// lattice fits (the block count is capped).  Each GPU holds its own copy
// until the process ends; host code reaches it only through the device
// address that moments_scratch looks up.
static __device__ __align__(256) double scratch[psc::kMomMaxBlocks * PSC_NSUM];

namespace psc {
namespace {
std::mutex g_scratch_mu;
std::vector<double*> g_scratch_dev;  // device address of `scratch` per device (nullptr: not looked up yet)
}  // namespace

// Device address and capacity (doubles) of the scratch array on the current
// device, looked up once per device with that device current.
int moments_scratch(double** out, int64_t* capacity) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return set_error((int)e, "psc_moments_noalloc: hipGetDevice: %s", hipGetErrorString(e));
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if ((int)g_scratch_dev.size() <= dev) g_scratch_dev.resize(dev + 1, nullptr);
  if (g_scratch_dev[dev] == nullptr) {
    void* ptr = nullptr;
    e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || ptr == nullptr)
      return set_error(e != hipSuccess ? (int)e : PSC_EINVAL, "psc_moments_noalloc: hipGetSymbolAddress: %s",
                       hipGetErrorString(e));
    g_scratch_dev[dev] = (double*)ptr;
  }
  *out = g_scratch_dev[dev];
  *capacity = kMomMaxBlocks * PSC_NSUM;
  return PSC_OK;
}

}  // namespace psc
