// FP32 deviation-form PSC step (the bandwidth-bound production kernel).
//
// Storage is g_i = f_i - w_i in FP32, structure of arrays.  One thread owns
// one lattice cell; a workgroup is a (block_z x block_y) tile of one x-plane
// or, with the flat mapping, a run of consecutive cells of an x plane, so a
// wavefront reads and writes 64 consecutive cells of each direction plane
// (coalesced).  No MFMA, and the cell update uses no LDS: the step is
// HBM-bound at 19*4 B read + 19*4 B written + 4 B of solid count per cell
// (+ 12 B of imb at cells with solid).  The step kernels that also carry
// the particle-load reduction of earlier steps (k_step_f32_loads, its
// in-place pair form k_step_f32_aa and the LOADS form of k_step_f32_batched)
// use a few hundred bytes of LDS in those load workgroups only.  The in-place
// pairs (step_cell, AA = 1 and 2) update one state buffer, so a 128^3 state
// (159 MB) stays largely in the 256 MB Infinity Cache.
//
// The arithmetic restates src/lbm_dem/psc3d.py:211-236 in deviation form:
//   rho = 1 + sum g,  j = sum g c,  u_eff = (j + F)/rho   (both half shifts)
//   geq_i = w_i (drho + rho (3 cu + 4.5 cu^2 - 1.5 u^2))
//   Omega_bgk = (geq - g)/tau,  Omega_s = g_opp - g + geq_s - geq_opp
//   g' = g + (1-B) Omega_bgk + B Omega_s + Guo(u_eff)
//   imb = sum_i B Omega_s_i c_i
// Because sum_i w_i = 1 and sum_i w_i c_i = 0, the w_i of f cancel in every
// difference above and only the deviations are ever stored or summed.
//
// Momentum-exact FP32 coefficients.  The linear equilibrium term 3 w_i rho
// (c_i.u) is formed as k_i (c_i.(j + F)) with k = kAxis on the 6 axis links
// and kDiag on the 12 diagonals, chosen so that 2 kAxis + 8 kDiag = 1 exactly
// in FP32 (kAxis = 0x3e2aaaaa ~ 1/6, kDiag = (1 - 2 kAxis)/8 ~ 1/12; both
// are exact and the sum is exact).  With the FP32-rounded 3*w_i instead,
// sum_i 3 w_i c_ix^2 = 6 fl(1/18) != 1 and every BGK relaxation would inject
// a systematic ~1e-8 j of momentum per step; over thousands of steps that
// biases the permeability by ~1e-5.  Forming the term from j + F rather than
// rho * ((j + F)/rho) also removes the rho*(1/rho) rounding from it.  The
// remaining (quadratic and mass) terms are even in c_i and carry no momentum
// in any arithmetic, because a direction and its opposite compute them
// identically.  The Guo term uses the same k_i for its linear part.
#include "psc_common.h"
#include "psc_loads.h"
#include "../../include/psc_hip.h"

namespace psc {
namespace {

__device__ __forceinline__ float k_lin(int i) {
  // 0x3e2aaaaa and (1 - 2*0x3e2aaaaa)/8, see the header comment
  return i == 0 ? 0.0f : (i <= 6 ? 0.16666665673255920410f : 0.08333333581686019897f);
}

struct Dims {
  int nx, ny, nz;
  int pz;              // row pitch along z (cells); >= nz, padding cells are never used
  unsigned long long ps;  // cells per direction plane (>= nx*ny*pz)
};

// The flat cell mapping: a workgroup takes consecutive cells of the whole
// lattice's flat index c = (x ny + y) nz + z, so every workgroup's range
// starts on a 1 KB boundary (256 cells) whatever the row length, and the
// 64-byte granules at its ends are written by one workgroup only.
struct FlatMap {
  Divider nz, ny;
};
__device__ __forceinline__ void flat_cell(uint32_t c, const FlatMap& f, int& ix, int& iy, int& iz) {
  uint32_t q, r, x, y;
  divmod_exact(c, f.nz, q, r);
  divmod_exact(q, f.ny, x, y);
  ix = (int)x;
  iy = (int)y;
  iz = (int)r;
}

__device__ __forceinline__ uint32_t cell_id(const Dims d, int x, int y, int z) {
  return ((uint32_t)x * (uint32_t)d.ny + (uint32_t)y) * (uint32_t)d.pz + (uint32_t)z;
}

// Load / store of one FP32 value at byte offset `off` of a direction plane,
// in the saddr form of global_load/store: the plane base (uniform) in an SGPR
// pair, the offset a 32-bit VGPR, no 64-bit address arithmetic per access.
// Left to itself the compiler folds an offset that several planes share (a
// cell's own offset: every store of a pull step, every load and store of the
// local step of an in-place pair) into one 64-bit VGPR address and adds each
// plane's base to it with 64-bit VALU operations, 19 to 38 per cell, and
// holds those addresses in registers (the local in-place step: 109 VGPRs
// instead of 70).  The empty asm statements only pin the values into an SGPR
// pair (the base, uniform across the workgroup wherever these are called)
// and a VGPR (the offset), which hides the shared sum from that folding; they
// touch no memory.  The pointers are cast back to the global address space so
// the accesses stay global_load/store.  NT selects non-temporal hints.
typedef __attribute__((address_space(1))) char gchar_t;
typedef __attribute__((address_space(1))) float gfloat_t;
__device__ __forceinline__ const gchar_t* sgpr_base(const char* p) {
  asm("" : "+s"(p));
  return (const gchar_t*)p;
}
__device__ __forceinline__ gchar_t* sgpr_base(char* p) {
  asm("" : "+s"(p));
  return (gchar_t*)p;
}
template <bool NT>
__device__ __forceinline__ float ld(const char* __restrict__ base, uint32_t off) {
  const gfloat_t* p = (const gfloat_t*)(sgpr_base(base) + off);
  return NT ? __builtin_nontemporal_load(p) : *p;
}
template <bool NT>
__device__ __forceinline__ void st(char* __restrict__ base, uint32_t off, float v) {
  asm("" : "+v"(off));
  gfloat_t* p = (gfloat_t*)(sgpr_base(base) + off);
  if (NT) __builtin_nontemporal_store(v, p);
  else *p = v;
}

struct StepConsts {
  float inv_s3, tau, itau, omt, pref, Fx, Fy, Fz;
  int field;  // 1: the 4-byte solid input holds alpha as an FP32 field instead of a sub-sample count
};

// The collision of one cell from its gathered populations g (deviation
// form) and solid count cnt: post-collision values gn and, at a solid cell,
// its imb (mx, my, mz; 0 in fluid).  sm/cell/N/d/ix..iz: the solid-velocity
// inputs, the cell's index and coordinates, read only by MOVING solid cells.
// Every step kernel calls this one function, so every launch form computes
// the same bits.
template <bool MOVING>
__device__ __forceinline__ void collide_cell(const float (&g)[Q], uint32_t cnt, const StepConsts& k,
                                             const SolidMotion& sm, uint32_t cell, size_t N, const Dims& d, int ix,
                                             int iy, int iz, float (&gn)[Q], float& mx, float& my, float& mz) {
  // No implicit contraction: the fused operations are the explicit fmaf
  // calls, so the bits depend on this source only, not on how the compiler
  // fuses the code of each kernel and branch that inlines it (the fluid branch
  // and the solid branch at B = 0 then agree exactly, and so do all launch
  // forms, moving ones included).
#pragma clang fp contract(off)
  const float drho = ((g[0] + (g[1] + g[2])) + ((g[3] + g[4]) + (g[5] + g[6]))) +
                     (((g[7] + g[8]) + (g[9] + g[10])) + ((g[11] + g[12]) + (g[13] + g[14]))) +
                     ((g[15] + g[16]) + (g[17] + g[18]));
  const float jx = (g[1] - g[2]) + ((g[7] - g[8]) + (g[9] - g[10])) + ((g[11] - g[12]) + (g[13] - g[14]));
  const float jy = (g[3] - g[4]) + ((g[7] - g[8]) - (g[9] - g[10])) + ((g[15] - g[16]) + (g[17] - g[18]));
  const float jz = (g[5] - g[6]) + ((g[11] - g[12]) - (g[13] - g[14])) + ((g[15] - g[16]) - (g[17] - g[18]));
  const float rho = 1.0f + drho;
  // hardware reciprocal (1 ulp): 1/rho only scales the quadratic terms, the
  // momentum-carrying linear terms are formed from j + F directly
  const float ir = __builtin_amdgcn_rcpf(rho);
  const float jFx = jx + k.Fx, jFy = jy + k.Fy, jFz = jz + k.Fz;  // rho * u_eff
  const float ux = jFx * ir, uy = jFy * ir, uz = jFz * ir;
  const float usq15 = 1.5f * (ux * ux + uy * uy + uz * uz);
  const float uF3 = 3.0f * (ux * k.Fx + uy * k.Fy + uz * k.Fz);

  // per weight class: w*drho, w*rho, pref*w
  auto wdr = [&](int i) { return (float)w_of(i) * drho; };
  auto wrh = [&](int i) { return (float)w_of(i) * rho; };
  auto pw = [&](int i) { return k.pref * (float)w_of(i); };

  mx = 0.f;
  my = 0.f;
  mz = 0.f;
  // The branch is taken per wavefront: only a wavefront of fluid cells takes
  // the fluid form; any solid cell sends the whole wavefront through the
  // solid form, whose B = 0 case reproduces the fluid form bit for bit (the
  // relaxation factors reduce exactly to 1 - 1/tau and 1/tau and the solid
  // term to an exact 0), so the bits of a cell do not depend on its
  // neighbours in the wavefront and a mixed wavefront runs one form, not both.
  if (!__any(cnt != 0)) {
    // fluid cell: g' = (1 - 1/tau) g + geq/tau + Guo, pairwise even/odd parts
    {
      const float geq0 = wdr(0) - wrh(0) * usq15;
      const float gu0 = -pw(0) * uF3;
      gn[0] = fmaf(g[0], k.omt, fmaf(geq0, k.itau, gu0));
    }
#pragma unroll
    for (int i = 1; i < Q; i += 2) {
      const int o = i + 1;
      const float cu = (float)cx_of(i) * ux + (float)cy_of(i) * uy + (float)cz_of(i) * uz;
      const float cj = (float)cx_of(i) * jFx + (float)cy_of(i) * jFy + (float)cz_of(i) * jFz;
      const float cf = (float)cx_of(i) * k.Fx + (float)cy_of(i) * k.Fy + (float)cz_of(i) * k.Fz;
      const float S = fmaf(wrh(i), fmaf(4.5f * cu, cu, -usq15), wdr(i));  // even part of geq
      const float A = k_lin(i) * cj;                                      // odd part of geq
      const float Gs = pw(i) * fmaf(9.0f * cu, cf, -uF3);                 // even part of Guo
      const float Ga = k.pref * (k_lin(i) * cf);                          // odd part of Guo
      const float P = fmaf(S, k.itau, Gs);
      const float M = fmaf(A, k.itau, Ga);
      gn[i] = fmaf(g[i], k.omt, P + M);
      gn[o] = fmaf(g[o], k.omt, P - M);
    }
  } else {
    // B = 0 in fluid cells (a = 0): then om1 = 1/tau, om0 = 1 - 1/tau and
    // every B * (...) is 0, and the expressions below are the fluid form's
    const float a = k.field ? __uint_as_float(cnt) : fminf((float)cnt * k.inv_s3, 1.0f);
    const float B = a / ((a + k.tau * (1.0f - a)) + 1.0e-14f);
    const float om1 = (1.0f - B) * k.itau;  // (1 - B)/tau: weight of geq
    const float om0 = fmaf(B, k.itau, k.omt);  // 1 - (1 - B)/tau: weight of g
    float sx = 0.f, sy = 0.f, sz = 0.f, ssq15 = 0.f;
    if (MOVING && cnt != 0) {
      const uint32_t c = cell;
      const uint32_t o = sm.owner != nullptr ? sm.owner[c] : kSharedCell;
      if (o < sm.nmotion) {  // kSharedCell, and any id past the motion records, read the planes
        // the one sphere covering the cell: U + omega x r in FP64 with the
        // voxelizer's lever arm (the cell's box coordinate, found from the
        // box origin; a box spans at most one period), rounded to FP32.
        // Unfused, as the voxelizer computes it.
        const MotionRecord& m = sm.motion[o];
        const double rx = ((double)(m.x0 + wrap_near(ix - m.x0, d.nx)) + 0.5) - m.cx;
        const double ry = ((double)(m.y0 + wrap_near(iy - m.y0, d.ny)) + 0.5) - m.cy;
        const double rz = ((double)(m.z0 + wrap_near(iz - m.z0, d.nz)) + 0.5) - m.cz;
        sx = (float)((m.vx + m.wy * rz) - m.wz * ry);
        sy = (float)((m.vy + m.wz * rx) - m.wx * rz);
        sz = (float)((m.vz + m.wx * ry) - m.wy * rx);
      } else {
        const float ia = 1.0f / ((float)cnt * k.inv_s3);
        sx = sm.us_num[c] * ia;
        sy = sm.us_num[N + c] * ia;
        sz = sm.us_num[2 * N + c] * ia;
      }
      ssq15 = 1.5f * (sx * sx + sy * sy + sz * sz);
    }
    {
      const float geq0 = wdr(0) - wrh(0) * usq15;
      const float geqs0 = MOVING ? wdr(0) - wrh(0) * ssq15 : wdr(0);
      const float gu0 = -pw(0) * uF3;
      const float os0 = B * (geqs0 - geq0);
      gn[0] = fmaf(g[0], om0, fmaf(geq0, om1, gu0)) + os0;
    }
#pragma unroll
    for (int i = 1; i < Q; i += 2) {
      const int o = i + 1;
      const float cu = (float)cx_of(i) * ux + (float)cy_of(i) * uy + (float)cz_of(i) * uz;
      const float cj = (float)cx_of(i) * jFx + (float)cy_of(i) * jFy + (float)cz_of(i) * jFz;
      const float cf = (float)cx_of(i) * k.Fx + (float)cy_of(i) * k.Fy + (float)cz_of(i) * k.Fz;
      const float S = fmaf(wrh(i), fmaf(4.5f * cu, cu, -usq15), wdr(i));
      const float A = k_lin(i) * cj;
      const float Gs = pw(i) * fmaf(9.0f * cu, cf, -uF3);
      const float Ga = k.pref * (k_lin(i) * cf);
      float Ss = wdr(i), As = 0.0f;  // solid equilibrium, even and odd parts
      if (MOVING) {
        const float cs = (float)cx_of(i) * sx + (float)cy_of(i) * sy + (float)cz_of(i) * sz;
        Ss = fmaf(wrh(i), fmaf(4.5f * cs, cs, -ssq15), wdr(i));
        As = k_lin(i) * (rho * cs);
      }
      const float geq_i = S + A, geq_o = S - A;
      const float os_i = B * ((g[o] - g[i]) + ((Ss + As) - geq_o));
      const float os_o = B * ((g[i] - g[o]) + ((Ss - As) - geq_i));
      const float P = fmaf(S, om1, Gs);
      const float M = fmaf(A, om1, Ga);
      gn[i] = fmaf(g[i], om0, P + M) + os_i;
      gn[o] = fmaf(g[o], om0, P - M) + os_o;
      const float dm = os_i - os_o;  // c_i os_i + c_o os_o
      if (cx_of(i)) mx += (float)cx_of(i) * dm;
      if (cy_of(i)) my += (float)cy_of(i) * dm;
      if (cz_of(i)) mz += (float)cz_of(i) * dm;
    }
  }
}

// One cell: gather, collide, store.  imb is interleaved: the three
// components of a cell are consecutive floats (imb[3 c + axis]), so a cell's
// imb is one 12-byte store here and one 12-byte read in the load reduction.
//
// AA (in place, one state buffer; gin == gout), for a pair of steps that
// starts and ends in the pull form S[i][x] = post-collision g_i(x):
//   AA = 1 (gather-scatter): read S[i][x - c_i] (the pull gather), collide,
//          write g_i(x) to S[opp(i)][x + c_i], the location this thread read
//          as S[opp(i)][x - c_opp(i)].  Afterwards S[opp(i)][y] holds the
//          post-stream population f_i(y).
//   AA = 2 (local): read S[opp(i)][x] = f_i(x), collide, write S[i][x]: the
//          pull form again.
// Every location is read and written by one thread only, so the in-place
// update has no races, and the collision sees the inputs of the pull step
// in the same order: results are bit-identical to two pull steps.
template <bool PULL, bool MOVING, bool IMB, bool NT, bool IMBNT = false, int AA = 0>
__device__ __forceinline__ void step_cell(const float* gin, float* gout,
                                          const uint32_t* __restrict__ csum, const SolidMotion& sm,
                                          float* __restrict__ imb, const Dims d, const StepConsts& k, int ix,
                                          int iy, int iz) {
  const size_t plane = (size_t)d.ps * sizeof(float);  // bytes per direction plane
  const int xm = wrap_step(ix, -1, d.nx), xp = wrap_step(ix, 1, d.nx);
  const int ym = wrap_step(iy, -1, d.ny), yp = wrap_step(iy, 1, d.ny);
  const int zm = wrap_step(iz, -1, d.nz), zp = wrap_step(iz, 1, d.nz);
  auto X = [&](int s) { return s < 0 ? xm : (s > 0 ? xp : ix); };
  auto Y = [&](int s) { return s < 0 ? ym : (s > 0 ? yp : iy); };
  auto Z = [&](int s) { return s < 0 ? zm : (s > 0 ? zp : iz); };
  // byte offset of the cell displaced by (sx, sy, sz)
  auto off = [&](int sx, int sy, int sz) -> uint32_t { return cell_id(d, X(sx), Y(sy), Z(sz)) << 2; };
  const uint32_t here = cell_id(d, ix, iy, iz) << 2;
  const char* gi = (const char*)gin;
  char* go = (char*)gout;

  float g[Q];
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    if (AA == 2) g[i] = ld<NT>(gi + (size_t)opp_of(i) * plane, here);
    else g[i] = ld<NT>(gi + (size_t)i * plane, PULL ? off(-cx_of(i), -cy_of(i), -cz_of(i)) : here);
  }

  const uint32_t cnt = csum[cell_id(d, ix, iy, iz)];
  // destination offset of direction i
  auto dst = [&](int i) -> uint32_t { return PULL ? here : off(cx_of(i), cy_of(i), cz_of(i)); };

  // Post-collision values; stored by one store per direction after the
  // collision, so every store keeps the uniform-base + 32-bit-offset form.
  float gn[Q];
  float mx, my, mz;  // imb of a solid cell (0 in fluid)
  collide_cell<MOVING>(g, cnt, k, sm, cell_id(d, ix, iy, iz), (size_t)sm.us_plane, d, ix, iy, iz, gn, mx, my, mz);
  // imb is defined at solid cells only.  Lanes write it in groups of 16: a
  // group that holds a solid cell writes all 16 lanes (zeros at fluid cells),
  // so that 16 consecutive cells starting at a multiple of 16 fill three
  // whole 64-byte granules (a partly written granule costs HBM a
  // read-modify-write), and groups of fluid cells write nothing.
  const unsigned long long solid_lanes = __ballot(cnt != 0);
  if (IMB && ((solid_lanes >> (__lane_id() & 48u)) & 0xFFFFull) != 0) {
    float* q = imb + 3 * (size_t)cell_id(d, ix, iy, iz);
    if (IMBNT) {  // streaming stores: imb much larger than the caches
      __builtin_nontemporal_store(mx, q);
      __builtin_nontemporal_store(my, q + 1);
      __builtin_nontemporal_store(mz, q + 2);
    } else {
      q[0] = mx;
      q[1] = my;
      q[2] = mz;
    }
  }
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    if (AA == 1) st<NT>(go + (size_t)opp_of(i) * plane, off(cx_of(i), cy_of(i), cz_of(i)), gn[i]);
    else if (AA == 2) st<NT>(go + (size_t)i * plane, here, gn[i]);
    else st<NT>(go + (size_t)i * plane, dst(i), gn[i]);
  }
}


// Grid: (z tiles, y tiles, x).  A thread updates CPT cells at consecutive y.
// The dynamic LDS allocation is unused storage that caps the number of
// resident workgroups per CU (an occupancy limit chosen by the tuner); it
// leaves kLdsHeadroom bytes free so the small load-reduction workgroups of
// the previous step, running on another stream, can share the CU.
//
// LIN (flat mapping; grid (workgroups per x plane, x)): a workgroup takes
// blockDim.x*blockDim.y consecutive cells of an x plane's flat (y, z) index,
// so rows that are not a whole number of wavefronts leave no idle lanes
// (136-cell rows keep 136 of 192 lanes busy with one row per z tile).
template <bool PULL, bool MOVING, bool IMB, bool NT, int CPT>
__global__ __launch_bounds__(256) void k_step_f32(const float* __restrict__ gin, float* __restrict__ gout,
                                                  const uint32_t* __restrict__ csum,
                                                  SolidMotion sm, float* __restrict__ imb, Dims d,
                                                  StepConsts k, FlatMap) {
  const int iz = blockIdx.x * blockDim.x + threadIdx.x;
  const int iy0 = (blockIdx.y * blockDim.y + threadIdx.y) * CPT;
  const int ix = blockIdx.z;
  if (iz >= d.nz) return;
#pragma unroll
  for (int c = 0; c < CPT; ++c) {
    const int iy = iy0 + c;
    if (iy < d.ny)
      step_cell<PULL, MOVING, IMB, NT>(gin, gout, csum, sm, imb, d, k, ix, iy, iz);
  }
}

// The flat mapping.  Its register budget is held to six wavefronts per SIMD:
// the static forms need fewer registers anyway, and the moving form then
// fits without spilling (77 VGPRs instead of 93 unbounded, five wavefronts).
template <bool PULL, bool MOVING, bool IMB, bool NT>
__global__ __launch_bounds__(256, 6) void k_step_f32_flat(const float* __restrict__ gin, float* __restrict__ gout,
                                                          const uint32_t* __restrict__ csum, SolidMotion sm,
                                                          float* __restrict__ imb, Dims d, StepConsts k,
                                                          FlatMap fm) {
  const uint32_t c = blockIdx.x * (blockDim.x * blockDim.y) + threadIdx.y * blockDim.x + threadIdx.x;
  if (c >= (uint32_t)d.nx * (uint32_t)d.ny * (uint32_t)d.nz) return;
  int ix, iy, iz;
  flat_cell(c, fm, ix, iy, iz);
  step_cell<PULL, MOVING, IMB, NT>(gin, gout, csum, sm, imb, d, k, ix, iy, iz);
}

// Static pull step with the particle-load reduction of earlier steps in the
// same launch.  Besides the cell workgroups (the (z tile, y tile, x) order of
// k_step_f32) the grid holds nchunk chunk workgroups, which sum the chunks of
// the previous step's imb into part_out, and nfinal final workgroups, which
// add up the per-particle partials that the previous launch wrote into
// part_prev (the loads of the step before the previous one).  Each launch
// only reads what earlier launches wrote, so kernel boundaries order all of
// it, with no device-scope fences.  The extra workgroups come first, one
// after every period - 1 cell workgroups (period 1: all of them ahead of the
// cells), so their latency-bound gathers start at once and overlap the cell
// workgroups' streaming while they hold only part of the CU slots.
struct FusedLoads {
  LoadsArgs a;
  const float* imb_prev;     // imb written by the previous step
  double* part_out;          // 6 per chunk: the previous step's chunk sums
  const double* part_prev;   // 6 per chunk, written by the previous launch
  double* loads;             // 6 per particle
  int nchunk, nfinal;        // extra workgroups of each kind (0: none)
  int period;                // extra workgroups lead, one per period workgroups (>= 1)
};

// The load workgroups of a fused step launch: with nextra = nchunk + nfinal
// of them, workgroup b < nextra * period with b % period == period - 1 runs
// extra item b / period (a chunk pass on imb_prev, then per-particle final
// passes) and returns true; every other b is mapped to its index among the
// cell workgroups.
__device__ __forceinline__ bool fused_loads_workgroup(const FusedLoads& fl, const Dims& d, int& b) {
  const int nextra = fl.nchunk + fl.nfinal;
  if (nextra == 0) return false;
  if (b >= nextra * fl.period) {
    b -= nextra;
    return false;
  }
  const int e = b / fl.period, r = b - e * fl.period;
  if (r != fl.period - 1) {
    b = e * (fl.period - 1) + r;
    return false;
  }
  const int tid = threadIdx.y * blockDim.x + threadIdx.x;
  if (e < fl.nchunk) {
    __shared__ double lds[6 * 4];
    loads_chunk_body<float>(fl.a, fl.imb_prev, d.nx, d.ny, d.nz, e, tid, blockDim.x * blockDim.y, fl.part_out,
                               lds);
  } else {
    loads_final_body(fl.a.chunk_first, e - fl.nchunk, tid, fl.part_prev, fl.loads);
  }
  return true;
}

// LIN: the flat mapping of k_step_f32_flat (cell workgroup b takes cells
// b * blockDim.x * blockDim.y onwards); gz is then the number of cell
// workgroups and gy is 1.
template <bool NT, int CPT, bool LIN>
__global__ __launch_bounds__(256) void k_step_f32_loads(const float* __restrict__ gin, float* __restrict__ gout,
                                                        const uint32_t* __restrict__ csum, float* __restrict__ imb,
                                                        Dims d, StepConsts k, FusedLoads fl, int gz, int gy,
                                                        FlatMap fm) {
  int b = blockIdx.x;
  if (fused_loads_workgroup(fl, d, b)) return;
  if (LIN) {
    if (b >= gz) return;
    const uint32_t c = (uint32_t)b * (blockDim.x * blockDim.y) + threadIdx.y * blockDim.x + threadIdx.x;
    if (c >= (uint32_t)d.nx * (uint32_t)d.ny * (uint32_t)d.nz) return;
    int ix, iy, iz;
    flat_cell(c, fm, ix, iy, iz);
    step_cell<true, false, true, NT>(gin, gout, csum, SolidMotion{}, imb, d, k, ix, iy, iz);
    return;
  }
  const int t = b / gz;
  const int ix = t / gy;
  if (ix >= d.nx) return;
  const int iz = (b - t * gz) * blockDim.x + threadIdx.x;
  const int iy0 = ((t - ix * gy) * blockDim.y + threadIdx.y) * CPT;
  if (iz >= d.nz) return;
#pragma unroll
  for (int c = 0; c < CPT; ++c) {
    const int iy = iy0 + c;
    if (iy < d.ny)
      step_cell<true, false, true, NT>(gin, gout, csum, SolidMotion{}, imb, d, k, ix, iy, iz);
  }
}

// In-place (AA) form of k_step_f32_loads: one state buffer g, AA = 1 or 2 as
// in step_cell, one cell per thread, the same load workgroups.  g carries no
// __restrict__: it is read and written.
template <bool NT, bool LIN, int AA>
__global__ __launch_bounds__(256) void k_step_f32_aa(float* g, const uint32_t* __restrict__ csum,
                                                     float* __restrict__ imb, Dims d, StepConsts k, FusedLoads fl,
                                                     int gz, int gy, FlatMap fm) {
  int b = blockIdx.x;
  if (fused_loads_workgroup(fl, d, b)) return;
  int ix, iy, iz;
  if (LIN) {
    if (b >= gz) return;
    const uint32_t c = (uint32_t)b * (blockDim.x * blockDim.y) + threadIdx.y * blockDim.x + threadIdx.x;
    if (c >= (uint32_t)d.nx * (uint32_t)d.ny * (uint32_t)d.nz) return;
    flat_cell(c, fm, ix, iy, iz);
  } else {
    const int t = b / gz;
    ix = t / gy;
    if (ix >= d.nx) return;
    iz = (b - t * gz) * blockDim.x + threadIdx.x;
    iy = (t - ix * gy) * blockDim.y + threadIdx.y;
    if (iz >= d.nz || iy >= d.ny) return;
  }
  step_cell<true, false, true, NT, false, AA>(g, g, csum, SolidMotion{}, imb, d, k, ix, iy, iz);
}

// ---------------------------------------------------------------------------
// Plain D3Q19 BGK + Guo + bounce-back at a binary solid mask
// (scripts/run_3d_window_lbm.py:95-146) in deviation form, pull layout: the
// streamed populations are gathered, solid cells apply the in-place
// reflection loop of the reference (odd members take their even partner's
// value), then the cell collides with u = (j + F/2)/rho, zero in solids:
//   g' = (1 - 1/tau) g + geq/tau + Guo(u).
// The linear equilibrium and Guo terms use the momentum-exact coefficients
// k_i and j + F/2, as in step_cell.
// ---------------------------------------------------------------------------
template <bool NT>
__device__ __forceinline__ void bb_cell(const float* __restrict__ gin, float* __restrict__ gout,
                                        const uint8_t* __restrict__ solid, const Dims d, const StepConsts& k,
                                        int ix, int iy, int iz) {
  const size_t plane = (size_t)d.ps * sizeof(float);
  const int xm = wrap_step(ix, -1, d.nx), xp = wrap_step(ix, 1, d.nx);
  const int ym = wrap_step(iy, -1, d.ny), yp = wrap_step(iy, 1, d.ny);
  const int zm = wrap_step(iz, -1, d.nz), zp = wrap_step(iz, 1, d.nz);
  auto X = [&](int s) { return s < 0 ? xm : (s > 0 ? xp : ix); };
  auto Y = [&](int s) { return s < 0 ? ym : (s > 0 ? yp : iy); };
  auto Z = [&](int s) { return s < 0 ? zm : (s > 0 ? zp : iz); };
  auto off = [&](int sx, int sy, int sz) -> uint32_t { return cell_id(d, X(sx), Y(sy), Z(sz)) << 2; };
  const uint32_t here = cell_id(d, ix, iy, iz) << 2;
  const char* gi = (const char*)gin;
  char* go = (char*)gout;
  float g[Q];
#pragma unroll
  for (int i = 0; i < Q; ++i) g[i] = ld<NT>(gi + (size_t)i * plane, off(-cx_of(i), -cy_of(i), -cz_of(i)));
  const bool s = solid[cell_id(d, ix, iy, iz)] != 0;
  if (s) {
#pragma unroll
    for (int i = 1; i < Q; i += 2) g[i] = g[i + 1];
  }
  const float drho = ((g[0] + (g[1] + g[2])) + ((g[3] + g[4]) + (g[5] + g[6]))) +
                     (((g[7] + g[8]) + (g[9] + g[10])) + ((g[11] + g[12]) + (g[13] + g[14]))) +
                     ((g[15] + g[16]) + (g[17] + g[18]));
  const float jx = (g[1] - g[2]) + ((g[7] - g[8]) + (g[9] - g[10])) + ((g[11] - g[12]) + (g[13] - g[14]));
  const float jy = (g[3] - g[4]) + ((g[7] - g[8]) - (g[9] - g[10])) + ((g[15] - g[16]) + (g[17] - g[18]));
  const float jz = (g[5] - g[6]) + ((g[11] - g[12]) - (g[13] - g[14])) + ((g[15] - g[16]) - (g[17] - g[18]));
  const float rho = 1.0f + drho;
  const float ir = __builtin_amdgcn_rcpf(rho);
  // rho * u with the single half-force shift; zero in solid cells
  const float jFx = s ? 0.0f : jx + 0.5f * k.Fx, jFy = s ? 0.0f : jy + 0.5f * k.Fy, jFz = s ? 0.0f : jz + 0.5f * k.Fz;
  const float ux = jFx * ir, uy = jFy * ir, uz = jFz * ir;
  const float usq15 = 1.5f * (ux * ux + uy * uy + uz * uz);
  const float uF3 = 3.0f * (ux * k.Fx + uy * k.Fy + uz * k.Fz);
  auto wdr = [&](int i) { return (float)w_of(i) * drho; };
  auto wrh = [&](int i) { return (float)w_of(i) * rho; };
  auto pw = [&](int i) { return k.pref * (float)w_of(i); };
  float gn[Q];
  gn[0] = fmaf(g[0], k.omt, fmaf(wdr(0) - wrh(0) * usq15, k.itau, -pw(0) * uF3));
#pragma unroll
  for (int i = 1; i < Q; i += 2) {
    const int o = i + 1;
    const float cu = (float)cx_of(i) * ux + (float)cy_of(i) * uy + (float)cz_of(i) * uz;
    const float cj = (float)cx_of(i) * jFx + (float)cy_of(i) * jFy + (float)cz_of(i) * jFz;
    const float cf = (float)cx_of(i) * k.Fx + (float)cy_of(i) * k.Fy + (float)cz_of(i) * k.Fz;
    const float S = fmaf(wrh(i), fmaf(4.5f * cu, cu, -usq15), wdr(i));
    const float A = k_lin(i) * cj;
    const float Gs = pw(i) * fmaf(9.0f * cu, cf, -uF3);
    const float Ga = k.pref * (k_lin(i) * cf);
    const float P = fmaf(S, k.itau, Gs);
    const float M = fmaf(A, k.itau, Ga);
    gn[i] = fmaf(g[i], k.omt, P + M);
    gn[o] = fmaf(g[o], k.omt, P - M);
  }
#pragma unroll
  for (int i = 0; i < Q; ++i) st<NT>(go + (size_t)i * plane, here, gn[i]);
}

template <bool NT>
__global__ __launch_bounds__(256) void k_bb_step_f32(const float* __restrict__ gin, float* __restrict__ gout,
                                                     const uint8_t* __restrict__ solid, Dims d, StepConsts k) {
  const int iz = blockIdx.x * blockDim.x + threadIdx.x;
  const int iy = blockIdx.y * blockDim.y + threadIdx.y;
  const int ix = blockIdx.z;
  if (iz >= d.nz || iy >= d.ny) return;
  bb_cell<NT>(gin, gout, solid, d, k, ix, iy, iz);
}

// ---------------------------------------------------------------------------
// Batched windows: one persistent launch steps many independent periodic
// lattices.  Window w's state is a contiguous block of 19 planes of N_w
// cells at element offset state_off; its count_sum and imb cells sit at
// cell_off in arrays shared by all windows (imb plane stride = total cells).
// Work items are spans of up to 256 consecutive cells of one window, so the
// window parameters are uniform across a workgroup; workgroups loop over the
// items with a grid stride.
// ---------------------------------------------------------------------------
struct BatchWindow {
  long long state_off;  // elements
  long long cell_off;   // cells
  int nx, ny, nz, pz;   // extents and z row pitch (pz >= nz)
  float inv_pz, inv_nypz;  // reciprocals for the index decomposition
  float fx, fy, fz;
  int span;             // cells per work item (<= workgroup size; <= 0: the workgroup size)
  long long ps;         // cells per direction plane (multiple of 64: 256-byte aligned planes)
};

// With LOADS, the item sequence also holds the load work of earlier steps
// (fl: as for k_step_f32_loads, over the flattened cell index of all
// windows): nchunk + nfinal extra items lead it, one in every fl.period.
template <bool IMB, bool NT, int BS, bool LOADS, bool IMBNT>
__global__ __launch_bounds__(BS) void k_step_f32_batched(const float* __restrict__ gin, float* __restrict__ gout,
                                                         const uint32_t* __restrict__ csum,
                                                         float* __restrict__ imb, size_t imb_stride,
                                                         const BatchWindow* __restrict__ wins,
                                                         const int2* __restrict__ items, int nitems,
                                                         StepConsts k0, FusedLoads fl, int total) {
  const int tid = threadIdx.x;
  const int nextra = LOADS ? fl.nchunk + fl.nfinal : 0;
  for (int it0 = blockIdx.x; it0 < (LOADS ? total : nitems); it0 += gridDim.x) {
    int it = it0;
    if (LOADS && nextra > 0) {
      if (it0 < nextra * fl.period) {
        const int e = it0 / fl.period, r = it0 - e * fl.period;
        if (r == fl.period - 1) {
          if (e < fl.nchunk) {
            __shared__ double lds[6 * 4];
            loads_chunk_body<float>(fl.a, fl.imb_prev, 1, 1, (int)imb_stride, e, tid, BS, fl.part_out, lds);
          } else {
            loads_final_body(fl.a.chunk_first, e - fl.nchunk, tid, fl.part_prev, fl.loads);
          }
          continue;
        }
        it = e * (fl.period - 1) + r;
      } else {
        it = it0 - nextra;
      }
      if (it >= nitems) continue;
    }
    const int2 item = items[it];  // (window, first cell)
    const BatchWindow w = wins[item.x];
    const int span = w.span > 0 ? w.span : BS;
    const uint32_t c = (uint32_t)(item.y + tid);
    const uint32_t Nw = (uint32_t)w.nx * w.ny * w.pz;
    if (tid >= span || c >= Nw) continue;
    uint32_t ix, rem, iy, iz;
    divmod_f(c, (uint32_t)(w.ny * w.pz), w.inv_nypz, ix, rem);
    divmod_f(rem, (uint32_t)w.pz, w.inv_pz, iy, iz);
    if (iz >= (uint32_t)w.nz) continue;  // row padding
    StepConsts k = k0;
    k.Fx = w.fx;
    k.Fy = w.fy;
    k.Fz = w.fz;
    const Dims d{w.nx, w.ny, w.nz, w.pz, (unsigned long long)w.ps};
    step_cell<true, false, IMB, NT, IMBNT>(gin + w.state_off, gout + w.state_off, csum + w.cell_off, SolidMotion{},
                                           IMB ? imb + 3 * w.cell_off : nullptr, d, k, (int)ix, (int)iy, (int)iz);
  }
}

}  // namespace

constexpr size_t kLdsPerCU = 163840;
constexpr size_t kLdsHeadroom = 8192;

// owner/motion/nmotion (optional, with usn): the owner table of a moving step.
int launch_step_f32(const float* in, float* out, const uint32_t* csum, const float* usn, float* imb, int layout,
                    int nx, int ny, int nz, int s, double tau, double fx, double fy, double fz, int bz, int by,
                    int nt, hipStream_t stream, const uint32_t* owner, const void* motion, int nmotion) {
  const unsigned long long N = (unsigned long long)nx * ny * nz;
  const Dims d{nx, ny, nz, nz, state_plane(N, (nt & PSC_STEP_PLANES64) != 0)};
  const SolidMotion sm{usn, owner, (const MotionRecord*)motion, owner != nullptr ? (uint32_t)nmotion : 0u, N};
  const int cpt = (nt >> 16) & 3 ? (nt >> 16) & 3 : 1;
  const int bpc = (nt >> 8) & 0xff;
  size_t lds = 0;
  if (bpc > 0) {
    lds = (kLdsPerCU - kLdsHeadroom) / (size_t)bpc;
    if (lds > 65536) lds = 65536;
  }
  const bool lin = (nt & PSC_STEP_LINEAR) != 0;
  const dim3 block(bz, by, 1);
  const int rows = (ny + cpt - 1) / cpt;
  const size_t ncell = (size_t)nx * ny * nz;
  const unsigned nflat = (unsigned)((ncell + (size_t)bz * by - 1) / ((size_t)bz * by));
  const dim3 grid = lin ? dim3(nflat, 1, 1) : dim3((nz + bz - 1) / bz, (rows + by - 1) / by, nx);
  const FlatMap fm{make_divider((uint32_t)nz), make_divider((uint32_t)ny)};
  if (cpt != 1 && cpt != 2) return set_error(PSC_EINVAL, "psc_step(f32): cells per thread %d", cpt);
  if (ncell > ((size_t)1 << 30))
    return set_error(PSC_EINVAL, "psc_step(f32): at most 2^30 cells (32-bit byte offsets in a plane)");
  if (lin && cpt != 1) return set_error(PSC_EINVAL, "psc_step(f32): the flat mapping takes one cell per thread");
  StepConsts k;
  k.field = s == 0;
  k.inv_s3 = k.field ? 0.0f : (float)(1.0 / (double)(s * s * s));
  k.tau = (float)tau;
  k.itau = (float)(1.0 / tau);
  k.omt = (float)(1.0 - 1.0 / tau);
  k.pref = (float)(1.0 - 0.5 / tau);
  k.Fx = (float)fx;
  k.Fy = (float)fy;
  k.Fz = (float)fz;
  const bool pull = layout == PSC_PULL;
  const bool moving = usn != nullptr;
  const bool wimb = imb != nullptr;
  const bool ntb = (nt & 1) != 0;
#define PSC_F32_CASE(P, M, I, T)                                                                  \
  if (pull == P && moving == M && wimb == I && ntb == T) {                                       \
    if (lin)                                                                                     \
      k_step_f32_flat<P, M, I, T><<<grid, block, lds, stream>>>(in, out, csum, sm, imb, d, k,      \
                                                                    fm);                         \
    else if (cpt == 2)                                                                           \
      k_step_f32<P, M, I, T, 2><<<grid, block, lds, stream>>>(in, out, csum, sm, imb, d, k,         \
                                                                     fm);                        \
    else                                                                                         \
      k_step_f32<P, M, I, T, 1><<<grid, block, lds, stream>>>(in, out, csum, sm, imb, d, k,         \
                                                                     fm);                        \
    return check_launch("psc_step(f32)");                                                        \
  }
#define PSC_F32_CASES(T)                 \
  PSC_F32_CASE(false, false, false, T)   \
  PSC_F32_CASE(false, false, true, T)    \
  PSC_F32_CASE(false, true, false, T)    \
  PSC_F32_CASE(false, true, true, T)     \
  PSC_F32_CASE(true, false, false, T)    \
  PSC_F32_CASE(true, false, true, T)     \
  PSC_F32_CASE(true, true, false, T)     \
  PSC_F32_CASE(true, true, true, T)
  PSC_F32_CASES(false)
  PSC_F32_CASES(true)
#undef PSC_F32_CASES
#undef PSC_F32_CASE
  return set_error(PSC_EINVAL, "psc_step(f32): no kernel variant");
}

// Static FP32 pull step writing imb.  With imb_prev set, the launch also sums
// the chunks of imb_prev (the previous step's imb) into part_out; with
// part_prev set, it also adds up part_prev's per-particle partials into loads.
// period: the extra workgroups lead, one in every `period` workgroups.  Same
// flags as launch_step_f32.  aa: 0 = pull step from in to out; 1 or 2 = the
// gather-scatter or local step of an in-place pair (step_cell), in == out.
int launch_step_f32_loads(const float* in, float* out, const uint32_t* csum, float* imb, const float* imb_prev,
                          const LoadsArgs& la, double* part_out, const double* part_prev, double* loads, int period,
                          int nx, int ny, int nz, int s, double tau, double fx, double fy, double fz, int bz, int by,
                          int nt, hipStream_t stream, int aa) {
  const Dims d{nx, ny, nz, nz, state_plane((unsigned long long)nx * ny * nz, (nt & PSC_STEP_PLANES64) != 0)};
  const int cpt = (nt >> 16) & 3 ? (nt >> 16) & 3 : 1;
  const int bpc = (nt >> 8) & 0xff;
  size_t lds = 0;
  if (bpc > 0) {
    lds = (kLdsPerCU - kLdsHeadroom) / (size_t)bpc;
    if (lds > 65536) lds = 65536;
  }
  if (cpt != 1 && cpt != 2) return set_error(PSC_EINVAL, "psc_step(f32+loads): cells per thread %d", cpt);
  if (aa < 0 || aa > 2 || (aa > 0 && (cpt != 1 || in != out)))
    return set_error(PSC_EINVAL, "psc_step(f32+loads): in-place form %d needs one buffer and one cell per thread", aa);
  if (aa == 0 && in == out) return set_error(PSC_EINVAL, "psc_step(f32+loads): input and output alias");
  if ((bz * by) % 64 != 0 || bz * by > 256) return set_error(PSC_ELAUNCH, "psc_step(f32+loads): block %dx%d", bz, by);
  if ((size_t)nx * ny * nz > ((size_t)1 << 30))
    return set_error(PSC_EINVAL, "psc_step(f32+loads): at most 2^30 cells (32-bit byte offsets in a plane)");
  const bool lin = (nt & PSC_STEP_LINEAR) != 0;
  if (lin && cpt != 1) return set_error(PSC_EINVAL, "psc_step(f32+loads): the flat mapping takes one cell per thread");
  const int rows = (ny + cpt - 1) / cpt;
  const int gz = lin ? (int)(((size_t)nx * ny * nz + (size_t)bz * by - 1) / ((size_t)bz * by)) : (nz + bz - 1) / bz;
  const int gy = lin ? 1 : (rows + by - 1) / by;
  const FlatMap fm{make_divider((uint32_t)nz), make_divider((uint32_t)ny)};
  const long long ncell = lin ? (long long)gz : (long long)gz * gy * nx;
  FusedLoads fl{la, imb_prev, part_out, part_prev, loads, 0, 0, 1};
  if (imb_prev != nullptr) {
    if (part_out == nullptr) return set_error(PSC_EINVAL, "psc_step(f32+loads): null partials");
    fl.nchunk = la.nchunks;
  }
  if (part_prev != nullptr) {
    if (loads == nullptr) return set_error(PSC_EINVAL, "psc_step(f32+loads): null loads");
    fl.nfinal = la.np;
  }
  if (period < 1) return set_error(PSC_EINVAL, "psc_step(f32+loads): period %d", period);
  fl.period = period;
  long long grid = ncell;
  const long long extra = (long long)fl.nchunk + fl.nfinal;
  if (extra > 0) {
    grid = ncell + extra;
    if (grid < extra * period) grid = extra * period;  // every extra workgroup inside the interleaved head
  }
  if (grid > 0x7fffffffll) return set_error(PSC_EINVAL, "psc_step(f32+loads): grid too large");
  StepConsts k;
  k.field = s == 0;
  k.inv_s3 = k.field ? 0.0f : (float)(1.0 / (double)(s * s * s));
  k.tau = (float)tau;
  k.itau = (float)(1.0 / tau);
  k.omt = (float)(1.0 - 1.0 / tau);
  k.pref = (float)(1.0 - 0.5 / tau);
  k.Fx = (float)fx;
  k.Fy = (float)fy;
  k.Fz = (float)fz;
  const dim3 block(bz, by, 1);
  const bool ntb = (nt & 1) != 0;
  const unsigned g = (unsigned)grid;
  if (aa > 0) {
#define PSC_F32AA_CASE(T, L, A)                                                                        \
    if (ntb == T && lin == L && aa == A) {                                                             \
      k_step_f32_aa<T, L, A><<<g, block, lds, stream>>>(out, csum, imb, d, k, fl, gz, gy, fm);         \
      return check_launch("psc_step(f32+loads, in place)");                                            \
    }
    PSC_F32AA_CASE(true, false, 1) PSC_F32AA_CASE(true, false, 2) PSC_F32AA_CASE(false, false, 1)
    PSC_F32AA_CASE(false, false, 2) PSC_F32AA_CASE(true, true, 1) PSC_F32AA_CASE(true, true, 2)
    PSC_F32AA_CASE(false, true, 1) PSC_F32AA_CASE(false, true, 2)
#undef PSC_F32AA_CASE
    return set_error(PSC_EINVAL, "psc_step(f32+loads): no in-place kernel variant");
  }
#define PSC_F32L_ARGS in, out, csum, imb, d, k, fl, gz, gy, fm
  if (ntb && lin) k_step_f32_loads<true, 1, true><<<g, block, lds, stream>>>(PSC_F32L_ARGS);
  else if (lin) k_step_f32_loads<false, 1, true><<<g, block, lds, stream>>>(PSC_F32L_ARGS);
  else if (ntb && cpt == 2) k_step_f32_loads<true, 2, false><<<g, block, lds, stream>>>(PSC_F32L_ARGS);
  else if (ntb) k_step_f32_loads<true, 1, false><<<g, block, lds, stream>>>(PSC_F32L_ARGS);
  else if (cpt == 2) k_step_f32_loads<false, 2, false><<<g, block, lds, stream>>>(PSC_F32L_ARGS);
  else k_step_f32_loads<false, 1, false><<<g, block, lds, stream>>>(PSC_F32L_ARGS);
#undef PSC_F32L_ARGS
  return check_launch("psc_step(f32+loads)");
}

int launch_bb_step_f32(const float* in, float* out, const uint8_t* solid, int nx, int ny, int nz, double tau,
                       double fx, double fy, double fz, int bz, int by, int flags, hipStream_t stream) {
  const Dims d{nx, ny, nz, nz, (unsigned long long)nx * ny * nz};
  const int bpc = (flags >> 8) & 0xff;
  size_t lds = 0;
  if (bpc > 0) {
    lds = (kLdsPerCU - kLdsHeadroom) / (size_t)bpc;
    if (lds > 65536) lds = 65536;
  }
  if ((size_t)nx * ny * nz > ((size_t)1 << 30))
    return set_error(PSC_EINVAL, "psc_bb_step(f32): at most 2^30 cells (32-bit byte offsets in a plane)");
  StepConsts k;
  k.field = 0;
  k.inv_s3 = 0.0f;
  k.tau = (float)tau;
  k.itau = (float)(1.0 / tau);
  k.omt = (float)(1.0 - 1.0 / tau);
  k.pref = (float)(1.0 - 0.5 / tau);
  k.Fx = (float)fx;
  k.Fy = (float)fy;
  k.Fz = (float)fz;
  const dim3 block(bz, by, 1);
  const dim3 grid((nz + bz - 1) / bz, (ny + by - 1) / by, nx);
  if (flags & 1) k_bb_step_f32<true><<<grid, block, lds, stream>>>(in, out, solid, d, k);
  else k_bb_step_f32<false><<<grid, block, lds, stream>>>(in, out, solid, d, k);
  return check_launch("psc_bb_step(f32)");
}

// Batched launch: win/items are device tables (see BatchWindow); grid_blocks
// = 0 sizes the persistent grid to the item count.  With `la` set, the launch
// also carries the load work of earlier steps (see k_step_f32_batched):
// chunk sums of imb_prev into part_out when imb_prev is set, per-particle
// sums of part_prev into loads when part_prev is set, the extra items leading
// one in every `period`.
int launch_step_f32_batched(const float* in, float* out, const uint32_t* csum, float* imb, size_t imb_stride,
                            const void* wins, const void* items, int nitems, int s, double tau, int flags,
                            int grid_blocks, hipStream_t stream, const LoadsArgs* la = nullptr,
                            const float* imb_prev = nullptr, double* part_out = nullptr,
                            const double* part_prev = nullptr, double* loads = nullptr, int period = 1) {
  FusedLoads fl{};
  fl.period = period;
  long long total = nitems;
  if (la != nullptr) {
    if (period < 1) return set_error(PSC_EINVAL, "psc_step_batched_loads: period %d", period);
    if (imb == nullptr) return set_error(PSC_EINVAL, "psc_step_batched_loads: the step must write imb");
    fl.a = *la;
    fl.imb_prev = imb_prev;
    fl.part_out = part_out;
    fl.part_prev = part_prev;
    fl.loads = loads;
    if (imb_prev != nullptr) {
      if (part_out == nullptr) return set_error(PSC_EINVAL, "psc_step_batched_loads: null partials");
      fl.nchunk = la->nchunks;
    }
    if (part_prev != nullptr) {
      if (loads == nullptr) return set_error(PSC_EINVAL, "psc_step_batched_loads: null loads");
      fl.nfinal = la->np;
    }
    const long long nextra = (long long)fl.nchunk + fl.nfinal;
    if (nextra > 0) {
      total = nitems + nextra;
      if (total < nextra * period) total = nextra * period;
    }
    if (total > 0x7fffffffll) return set_error(PSC_EINVAL, "psc_step_batched_loads: too many items");
  }
  if (total <= 0) return PSC_OK;
  const int bpc = (flags >> 8) & 0xff;
  size_t lds = 0;
  if (bpc > 0) {
    lds = (kLdsPerCU - kLdsHeadroom) / (size_t)bpc;
    if (lds > 65536) lds = 65536;
  }
  StepConsts k;
  k.field = 0;
  k.inv_s3 = (float)(1.0 / (double)(s * s * s));
  k.tau = (float)tau;
  k.itau = (float)(1.0 / tau);
  k.omt = (float)(1.0 - 1.0 / tau);
  k.pref = (float)(1.0 - 0.5 / tau);
  k.Fx = k.Fy = k.Fz = 0.0f;
  const int tot = (int)total;
  const int grid = (grid_blocks > 0 && grid_blocks < tot) ? grid_blocks : tot;
  const BatchWindow* w = (const BatchWindow*)wins;
  const int2* it = (const int2*)items;
  const bool nt = (flags & 1) != 0;
  const bool wide = (flags & PSC_BATCH_WG512) != 0;
  const bool ld = la != nullptr;
  const bool inb = imb != nullptr && (flags & PSC_BATCH_IMB_NT) != 0;
#define PSC_BATCH_CASE(I, T, L, J)                                                                             \
  if ((imb != nullptr) == I && nt == T && ld == L && inb == J) {                                              \
    if (wide)                                                                                                 \
      k_step_f32_batched<I, T, 512, L, J><<<grid, 512, lds, stream>>>(in, out, csum, imb, imb_stride, w, it,   \
                                                                      nitems, k, fl, tot);                    \
    else                                                                                                      \
      k_step_f32_batched<I, T, 256, L, J><<<grid, 256, lds, stream>>>(in, out, csum, imb, imb_stride, w, it,   \
                                                                      nitems, k, fl, tot);                    \
    return check_launch("psc_step_batched");                                                                  \
  }
  PSC_BATCH_CASE(true, true, false, false)
  PSC_BATCH_CASE(true, false, false, false)
  PSC_BATCH_CASE(false, true, false, false)
  PSC_BATCH_CASE(false, false, false, false)
  PSC_BATCH_CASE(true, true, true, false)
  PSC_BATCH_CASE(true, false, true, false)
  PSC_BATCH_CASE(true, true, false, true)
  PSC_BATCH_CASE(true, false, false, true)
  PSC_BATCH_CASE(true, true, true, true)
  PSC_BATCH_CASE(true, false, true, true)
#undef PSC_BATCH_CASE
  return set_error(PSC_EINVAL, "psc_step_batched: no kernel variant");
}

}  // namespace psc
