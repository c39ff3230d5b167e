// FP32 deviation-form PSC step (the bandwidth-bound production kernel).
//
// Storage is g_i = f_i - w_i in FP32, structure of arrays.  One thread owns
// one lattice cell; a workgroup is a (block_z x block_y) tile of one x-plane,
// so a wavefront reads and writes 64 consecutive z cells of each direction
// plane (coalesced).  No LDS and no MFMA: the step is HBM-bound at
// 19*4 B read + 19*4 B written + 4 B of solid count per cell.
//
// The arithmetic restates src/lbm_dem/psc3d.py:211-236 in deviation form:
//   rho = 1 + sum g,  j = sum g c,  u_eff = (j + F)/rho   (both half shifts)
//   geq_i = w_i (drho + rho (3 cu + 4.5 cu^2 - 1.5 u^2))
//   Omega_bgk = (geq - g)/tau,  Omega_s = g_opp - g + geq_s - geq_opp
//   g' = g + (1-B) Omega_bgk + B Omega_s + Guo(u_eff)
//   imb = sum_i B Omega_s_i c_i
// Because sum_i w_i = 1 and sum_i w_i c_i = 0, the w_i of f cancel in every
// difference above and only the deviations are ever stored or summed.
#include "psc_common.h"
#include "../../include/psc_hip.h"

namespace psc {
namespace {

struct Dims {
  int nx, ny, nz;
};

__device__ __forceinline__ uint32_t cell_id(const Dims d, int x, int y, int z) {
  return ((uint32_t)x * (uint32_t)d.ny + (uint32_t)y) * (uint32_t)d.nz + (uint32_t)z;
}

template <bool PULL, bool MOVING, bool IMB>
__global__ __launch_bounds__(256) void k_step_f32(const float* __restrict__ gin, float* __restrict__ gout,
                                                  const uint32_t* __restrict__ csum,
                                                  const float* __restrict__ usn, float* __restrict__ imb, Dims d,
                                                  float inv_s3, float tau, float inv_tau, float pref, float Fx,
                                                  float Fy, float Fz) {
  const int iz = blockIdx.x * blockDim.x + threadIdx.x;
  const int iy = blockIdx.y * blockDim.y + threadIdx.y;
  const int ix = blockIdx.z;
  if (iz >= d.nz || iy >= d.ny) return;
  const size_t N = (size_t)d.nx * d.ny * d.nz;
  const int xm = wrap_step(ix, -1, d.nx), xp = wrap_step(ix, 1, d.nx);
  const int ym = wrap_step(iy, -1, d.ny), yp = wrap_step(iy, 1, d.ny);
  const int zm = wrap_step(iz, -1, d.nz), zp = wrap_step(iz, 1, d.nz);
  const uint32_t cell = cell_id(d, ix, iy, iz);

  // neighbour coordinate for a displacement s in {-1,0,1}
  auto X = [&](int s) { return s < 0 ? xm : (s > 0 ? xp : ix); };
  auto Y = [&](int s) { return s < 0 ? ym : (s > 0 ? yp : iy); };
  auto Z = [&](int s) { return s < 0 ? zm : (s > 0 ? zp : iz); };

  float g[Q];
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    const uint32_t src = PULL ? cell_id(d, X(-cx_of(i)), Y(-cy_of(i)), Z(-cz_of(i))) : cell;
    g[i] = gin[(size_t)i * N + src];
  }

  const float drho = ((g[0] + (g[1] + g[2])) + ((g[3] + g[4]) + (g[5] + g[6]))) +
                     (((g[7] + g[8]) + (g[9] + g[10])) + ((g[11] + g[12]) + (g[13] + g[14]))) +
                     ((g[15] + g[16]) + (g[17] + g[18]));
  const float jx = (g[1] - g[2]) + ((g[7] - g[8]) + (g[9] - g[10])) + ((g[11] - g[12]) + (g[13] - g[14]));
  const float jy = (g[3] - g[4]) + ((g[7] - g[8]) - (g[9] - g[10])) + ((g[15] - g[16]) + (g[17] - g[18]));
  const float jz = (g[5] - g[6]) + ((g[11] - g[12]) - (g[13] - g[14])) + ((g[15] - g[16]) - (g[17] - g[18]));
  const float rho = 1.0f + drho;
  const float ir = 1.0f / rho;
  const float ux = (jx + Fx) * ir;
  const float uy = (jy + Fy) * ir;
  const float uz = (jz + Fz) * ir;
  const float usq15 = 1.5f * (ux * ux + uy * uy + uz * uz);
  const float uF3 = 3.0f * (ux * Fx + uy * Fy + uz * Fz);

  float geq[Q];
  float guo[Q];
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    const float w = (float)w_of(i);
    const float cu = (float)cx_of(i) * ux + (float)cy_of(i) * uy + (float)cz_of(i) * uz;
    const float cf = (float)cx_of(i) * Fx + (float)cy_of(i) * Fy + (float)cz_of(i) * Fz;
    geq[i] = w * drho + (w * rho) * (cu * (3.0f + 4.5f * cu) - usq15);
    guo[i] = (pref * w) * ((3.0f * cf - uF3) + 9.0f * cu * cf);
  }

  const uint32_t cnt = csum[cell];
  float gn[Q];
  if (cnt == 0) {
#pragma unroll
    for (int i = 0; i < Q; ++i) gn[i] = g[i] + (geq[i] - g[i]) * inv_tau + guo[i];
  } else {
    const float a = fminf((float)cnt * inv_s3, 1.0f);
    const float B = a / ((a + tau * (1.0f - a)) + 1.0e-14f);
    const float omb = (1.0f - B) * inv_tau;
    float geqs[Q];
    if (MOVING) {
      const float asum = (float)cnt * inv_s3;
      const float ia = 1.0f / asum;
      const float sx = usn[cell] * ia, sy = usn[N + cell] * ia, sz = usn[2 * N + cell] * ia;
      const float ssq15 = 1.5f * (sx * sx + sy * sy + sz * sz);
#pragma unroll
      for (int i = 0; i < Q; ++i) {
        const float w = (float)w_of(i);
        const float cu = (float)cx_of(i) * sx + (float)cy_of(i) * sy + (float)cz_of(i) * sz;
        geqs[i] = w * drho + (w * rho) * (cu * (3.0f + 4.5f * cu) - ssq15);
      }
    } else {
#pragma unroll
      for (int i = 0; i < Q; ++i) geqs[i] = (float)w_of(i) * drho;
    }
    float mx = 0.f, my = 0.f, mz = 0.f;
#pragma unroll
    for (int i = 0; i < Q; ++i) {
      const int o = opp_of(i);
      const float os = B * ((g[o] - g[i]) + (geqs[i] - geq[o]));
      gn[i] = g[i] + (geq[i] - g[i]) * omb + os + guo[i];
      mx += (float)cx_of(i) * os;
      my += (float)cy_of(i) * os;
      mz += (float)cz_of(i) * os;
    }
    if (IMB) {
      imb[cell] = mx;
      imb[N + cell] = my;
      imb[2 * N + cell] = mz;
    }
  }

#pragma unroll
  for (int i = 0; i < Q; ++i) {
    const uint32_t dst = PULL ? cell : cell_id(d, X(cx_of(i)), Y(cy_of(i)), Z(cz_of(i)));
    gout[(size_t)i * N + dst] = gn[i];
  }
}

}  // namespace

int launch_step_f32(const float* in, float* out, const uint32_t* csum, const float* usn, float* imb, int layout,
                    int nx, int ny, int nz, int s, double tau, double fx, double fy, double fz, int bz, int by,
                    hipStream_t stream) {
  const Dims d{nx, ny, nz};
  const dim3 block(bz, by, 1);
  const dim3 grid((nz + bz - 1) / bz, (ny + by - 1) / by, nx);
  const float inv_s3 = (float)(1.0 / (double)(s * s * s));
  const float ftau = (float)tau;
  const float inv_tau = (float)(1.0 / tau);
  const float pref = (float)(1.0 - 0.5 / tau);
  const bool pull = layout == PSC_PULL;
  const bool moving = usn != nullptr;
  const bool wimb = imb != nullptr;
#define PSC_F32_CASE(P, M, I)                                                                                   \
  if (pull == P && moving == M && wimb == I) {                                                                 \
    k_step_f32<P, M, I><<<grid, block, 0, stream>>>(in, out, csum, usn, imb, d, inv_s3, ftau, inv_tau, pref,     \
                                                    (float)fx, (float)fy, (float)fz);                          \
    return check_launch("psc_step(f32)");                                                                      \
  }
  PSC_F32_CASE(false, false, false)
  PSC_F32_CASE(false, false, true)
  PSC_F32_CASE(false, true, false)
  PSC_F32_CASE(false, true, true)
  PSC_F32_CASE(true, false, false)
  PSC_F32_CASE(true, false, true)
  PSC_F32_CASE(true, true, false)
  PSC_F32_CASE(true, true, true)
#undef PSC_F32_CASE
  return set_error(PSC_EINVAL, "psc_step(f32): no kernel variant");
}

}  // namespace psc
