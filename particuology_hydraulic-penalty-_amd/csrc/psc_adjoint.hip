// Reverse mode of the PSC step for permeability gradients (gfx950).
//
// Forward model (pull layout, stationary solids, solid fraction alpha given
// per cell): at cell x the step gathers f_i = s_t[i](x - c_i), collides
// (psc3d.py:211-232 with F a vector and both half-force shifts) and stores
// s_{t+1}[i](x) = f_i + (1-B) Obgk_i + B Os_i + Guo_i with
//   rho = sum f, u = (j + F)/rho, feq_i(rho, u) = w_i rho (1 + 3cu + 4.5cu^2 - 1.5u^2),
//   Obgk_i = (feq_i - f_i)/tau,  Os_i = f_opp(i) - f_i + w_i rho - feq_opp(i),
//   Guo_i = (1 - 1/(2tau)) w_i (3 (c_i - u).F + 9 (c_i.u)(c_i.F)),
//   B = alpha/(alpha + tau (1-alpha) + 1e-14).
// Writing a = 1 - b - B, b = (1-B)/tau, the post-collision value is
//   post_i = a f_i + B f_opp(i) + b feq_i - B feq_opp(i) + B w_i rho + Guo_i,
// so with lam = dL/d post and Lam_i = b lam_i - B lam_opp(i):
//   dL/df_k = a lam_k + B lam_opp(k) + g_rho + g_u.(c_k - u)/rho
//   g_rho = sum_i Lam_i feq_i/rho + B sum_i lam_i w_i
//   g_u   = rho sum_i Lam_i w_i (3c_i + 9(c_i.u)c_i - 3u)
//           + (1 - 1/(2tau)) sum_i lam_i w_i (9 (c_i.F) c_i - 3F)
//   dL/dB = sum_i lam_i (Os_i - Obgk_i),  dB/dalpha = (tau + 1e-14)/den^2.
// dL/df_k belongs to s_t[k](x - c_k): the reverse step scatters it there,
// one write per (cell, direction), a permutation, so no atomics.
//
// Smooth solid fraction: alpha_p(x) = S(t), t = r + 1/2 - |x - c|, S the
// quintic smoothstep 6t^5 - 15t^4 + 10t^3 on [0, 1] (C2 continuous), and
// alpha = 1 - prod_p (1 - alpha_p): equal to alpha_p where one particle
// covers the cell, within [0, 1], and smooth in every centre and radius.
#include "psc_common.h"
#include "../../include/psc_hip.h"

namespace psc {
namespace {

struct ADims {
  int nx, ny, nz;
};

__device__ __forceinline__ uint32_t acell(const ADims d, int x, int y, int z) {
  return ((uint32_t)x * (uint32_t)d.ny + (uint32_t)y) * (uint32_t)d.nz + (uint32_t)z;
}

// Quintic smoothstep and its derivative.
__device__ __forceinline__ double smooth5(double t, double& ds) {
  if (t <= 0.0) {
    ds = 0.0;
    return 0.0;
  }
  if (t >= 1.0) {
    ds = 0.0;
    return 1.0;
  }
  const double om = 1.0 - t;
  ds = 30.0 * t * t * om * om;
  return t * t * t * (10.0 + t * (-15.0 + 6.0 * t));
}

// One box entry of a chunk: the particle, its unwrapped cell and the
// distance terms.  Returns false past the end of the box.
struct Entry {
  double dx, dy, dz, dist, a, ds;
  uint32_t cell;
};

__device__ __forceinline__ bool box_entry(const double* __restrict__ q, const int32_t* __restrict__ b, int e,
                                          const ADims d, Entry& out) {
  const int ex = b[3], ey = b[4], ez = b[5];
  if (e >= ex * ey * ez) return false;
  const int lz = e % ez;
  const int t = e / ez;
  const int ly = t % ey;
  const int lx = t / ey;
  const int gx = b[0] + lx, gy = b[1] + ly, gz = b[2] + lz;
  out.dx = ((double)gx + 0.5) - q[0];
  out.dy = ((double)gy + 0.5) - q[1];
  out.dz = ((double)gz + 0.5) - q[2];
  out.dist = sqrt(out.dx * out.dx + out.dy * out.dy + out.dz * out.dz);
  out.a = smooth5(q[3] + 0.5 - out.dist, out.ds);
  out.cell = acell(d, wrap_any(gx, d.nx), wrap_any(gy, d.ny), wrap_any(gz, d.nz));
  return true;
}

__global__ __launch_bounds__(256) void k_fill_u32(uint32_t* __restrict__ p, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = 0u;
}
__global__ __launch_bounds__(256) void k_fill_u64(unsigned long long* __restrict__ p, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    p[i] = 0ull;
}

// log(1 - alpha_p) is summed per cell in fixed point (units of 2^-50) with
// integer atomics, so the sum and alpha are independent of the order in
// which particles reach a cell.  |log(1 - a)| < 37 for every double a < 1,
// and 2^63 units hold the sum of over 200 such terms.
constexpr double kLogScale = 1125899906842624.0;  // 2^50
constexpr double kLogUnit = 1.0 / kLogScale;

__device__ __forceinline__ double log_sum(const unsigned long long* __restrict__ logsum, uint32_t c) {
  return (double)(long long)logsum[c] * kLogUnit;
}

// One workgroup per chunk of one particle's box entries.
__global__ __launch_bounds__(256) void k_smooth_scatter(const double* __restrict__ parts,
                                                        const int32_t* __restrict__ boxes,
                                                        const int32_t* __restrict__ chunks, int chunk_cells,
                                                        ADims d, uint32_t* __restrict__ nfull,
                                                        unsigned long long* __restrict__ logsum) {
  const int p = chunks[2 * blockIdx.x];
  const int e0 = chunks[2 * blockIdx.x + 1];
  const double* q = parts + 10 * (size_t)p;
  const int32_t* b = boxes + 8 * (size_t)p;
  for (int k = threadIdx.x; k < chunk_cells; k += blockDim.x) {
    Entry en;
    if (!box_entry(q, b, e0 + k, d, en)) break;
    if (en.a <= 0.0) continue;
    if (en.a >= 1.0) {
      atomicAdd(&nfull[en.cell], 1u);
    } else {
      const long long v = llrint(log1p(-en.a) * kLogScale);
      atomicAdd(&logsum[en.cell], (unsigned long long)v);
    }
  }
}

__global__ __launch_bounds__(256) void k_smooth_finish(const uint32_t* __restrict__ nfull,
                                                       const unsigned long long* __restrict__ logsum, size_t n,
                                                       double* __restrict__ a64, float* __restrict__ a32) {
  for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (size_t)gridDim.x * blockDim.x) {
    const double ls = log_sum(logsum, (uint32_t)c);
    // +0 exactly where no particle reaches the cell (the fluid test of the step)
    const double a = nfull[c] != 0u ? 1.0 : (ls == 0.0 ? 0.0 : -expm1(ls));
    if (a64 != nullptr) a64[c] = a;
    if (a32 != nullptr) a32[c] = (float)a;
  }
}

// dL/d(centre, radius): d alpha/d alpha_p = prod_{q != p} (1 - alpha_q)
// = exp(logsum - log(1 - alpha_p)) when no particle saturates the cell (0
// otherwise), d alpha_p/dc = S'(t) (x - c)/|x - c|, d alpha_p/dr = S'(t).
// Pass 1 writes one partial per chunk, pass 2 adds a particle's chunks in
// order: the gradient is deterministic.
__global__ __launch_bounds__(256) void k_smooth_grad(const double* __restrict__ parts,
                                                     const int32_t* __restrict__ boxes,
                                                     const int32_t* __restrict__ chunks, int chunk_cells, ADims d,
                                                     const uint32_t* __restrict__ nfull,
                                                     const unsigned long long* __restrict__ logsum,
                                                     const double* __restrict__ galpha,
                                                     double* __restrict__ partials) {
  __shared__ double lds[4 * 16];
  const int p = chunks[2 * blockIdx.x];
  const int e0 = chunks[2 * blockIdx.x + 1];
  const double* q = parts + 10 * (size_t)p;
  const int32_t* b = boxes + 8 * (size_t)p;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  for (int k = threadIdx.x; k < chunk_cells; k += blockDim.x) {
    Entry en;
    if (!box_entry(q, b, e0 + k, d, en)) break;
    if (!(en.a > 0.0 && en.a < 1.0) || nfull[en.cell] != 0u || !(en.dist > 0.0)) continue;
    const double others = exp(log_sum(logsum, en.cell) - log1p(-en.a));
    const double g = galpha[en.cell] * others * en.ds;
    const double gi = g / en.dist;
    acc[0] += gi * en.dx;
    acc[1] += gi * en.dy;
    acc[2] += gi * en.dz;
    acc[3] += g;
  }
  block_sum<4>(acc, lds);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < 4; ++k) partials[4 * (size_t)blockIdx.x + k] = acc[k];
  }
}

__global__ __launch_bounds__(256) void k_smooth_grad_final(const int32_t* __restrict__ chunk_first, int np,
                                                           const double* __restrict__ partials,
                                                           double* __restrict__ grad) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 4 * np) return;
  const int p = t >> 2, k = t & 3;
  double s = 0.0;
  for (int c = chunk_first[p]; c < chunk_first[p + 1]; ++c) s += partials[4 * (size_t)c + k];
  grad[t] = s;
}

// ---------------------------------------------------------------------------
// Adjoint step.  R = float with deviation storage (g = f - w) or double with
// plain storage.  The Jacobian needs only rho, u, B and, for dL/dB, the
// deviations, all formed here from the gathered state.
// ---------------------------------------------------------------------------
template <typename R>
struct AdjConsts {
  R tau, itau, pref, Fx, Fy, Fz;
};

// Load / store of one value at byte offset `off` of a direction plane whose
// base is uniform across the wavefront, in the saddr form (SGPR base + 32-bit
// VGPR offset).  The empty asm statements pin the base into an SGPR pair and
// the offset into a VGPR, which keeps the compiler from folding an offset
// shared by several planes into 64-bit VGPR addresses (64-bit VALU
// arithmetic per access, and registers to hold them); they touch no memory.
// The pointers are cast back to the global address space.
template <typename R, bool NT>
__device__ __forceinline__ R ldo(const char* __restrict__ base, uint32_t off) {
  typedef __attribute__((address_space(1))) const char gc_t;
  typedef __attribute__((address_space(1))) const R gr_t;
  asm("" : "+s"(base));
  asm("" : "+v"(off));
  gr_t* p = (gr_t*)((gc_t*)base + off);
  return NT ? __builtin_nontemporal_load(p) : *p;
}
template <typename R, bool NT>
__device__ __forceinline__ void sto(char* __restrict__ base, uint32_t off, R v) {
  typedef __attribute__((address_space(1))) char gc_t;
  typedef __attribute__((address_space(1))) R gr_t;
  asm("" : "+s"(base));
  asm("" : "+v"(off));
  gr_t* p = (gr_t*)((gc_t*)base + off);
  if (NT) __builtin_nontemporal_store(v, p);
  else *p = v;
}

// The populations enter dL/dB only through
//   sum_i lam_i ((g_opp(i) - g_i) + g_i/tau) = sum_i g_i (lam_opp(i) - (1 - 1/tau) lam_i),
// which is accumulated while the state is gathered, so the 19 populations
// are never held at once: the adjoint lam_i and a few sums are all that
// stay live.
template <typename R, bool DEV, bool NT>
__global__ __launch_bounds__(256) void k_adjoint_step(const R* __restrict__ st, const R* __restrict__ lin,
                                                      R* __restrict__ lout, const R* __restrict__ alpha,
                                                      double* __restrict__ galpha, ADims d, AdjConsts<R> k) {
  const int iz = blockIdx.x * blockDim.x + threadIdx.x;
  const int iy = blockIdx.y * blockDim.y + threadIdx.y;
  const int ix = blockIdx.z;
  if (iz >= d.nz || iy >= d.ny) return;
  const size_t plane = (size_t)d.nx * d.ny * d.nz * sizeof(R);  // bytes per direction plane
  const uint32_t here = acell(d, ix, iy, iz);
  const int xm = wrap_step(ix, -1, d.nx), xp = wrap_step(ix, 1, d.nx);
  const int ym = wrap_step(iy, -1, d.ny), yp = wrap_step(iy, 1, d.ny);
  const int zm = wrap_step(iz, -1, d.nz), zp = wrap_step(iz, 1, d.nz);
  auto X = [&](int s) { return s < 0 ? xm : (s > 0 ? xp : ix); };
  auto Y = [&](int s) { return s < 0 ? ym : (s > 0 ? yp : iy); };
  auto Z = [&](int s) { return s < 0 ? zm : (s > 0 ? zp : iz); };
  // byte offset of the cell x - c_i
  auto nb = [&](int i) -> uint32_t {
    return acell(d, X(-cx_of(i)), Y(-cy_of(i)), Z(-cz_of(i))) * (uint32_t)sizeof(R);
  };
  const uint32_t hoff = here * (uint32_t)sizeof(R);
  const char* sb = (const char*)st;
  const char* lb = (const char*)lin;
  char* ob = (char*)lout;

  R lam[Q];
#pragma unroll
  for (int i = 0; i < Q; ++i) lam[i] = ldo<R, NT>(lb + (size_t)i * plane, hoff);
  const R omt = (R)1 - k.itau;
  R drho = 0, jx = 0, jy = 0, jz = 0, Gl = 0;
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    R g = ldo<R, NT>(sb + (size_t)i * plane, nb(i));
    if (!DEV) g -= (R)w_of(i);
    drho += g;
    if (cx_of(i)) jx += (R)cx_of(i) * g;
    if (cy_of(i)) jy += (R)cy_of(i) * g;
    if (cz_of(i)) jz += (R)cz_of(i) * g;
    Gl += g * (lam[opp_of(i)] - omt * lam[i]);
  }
  const R rho = (R)1 + drho;
  const R ir = (R)1 / rho;
  const R ux = (jx + k.Fx) * ir, uy = (jy + k.Fy) * ir, uz = (jz + k.Fz) * ir;
  const R usq = ux * ux + uy * uy + uz * uz;
  const R a = alpha[here];
  const R den = (a + k.tau * ((R)1 - a)) + (R)1.0e-14;
  const R B = a / den;
  const R bb = ((R)1 - B) * k.itau;
  const R A = ((R)1 - bb) - B;

  R S0 = 0, S1 = 0, S2 = 0, Lw = 0;
  R V1x = 0, V1y = 0, V1z = 0, V2x = 0, V2y = 0, V2z = 0, LFx = 0, LFy = 0, LFz = 0;
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    const R w = (R)w_of(i);
    const R Lam = bb * lam[i] - B * lam[opp_of(i)];
    const R cu = (R)cx_of(i) * ux + (R)cy_of(i) * uy + (R)cz_of(i) * uz;
    const R cF = (R)cx_of(i) * k.Fx + (R)cy_of(i) * k.Fy + (R)cz_of(i) * k.Fz;
    const R Lw_i = Lam * w;
    S0 += Lw_i;
    S1 += Lw_i * cu;
    S2 += Lw_i * cu * cu;
    const R lw = lam[i] * w;
    Lw += lw;
    const R t2 = Lw_i * cu, tF = lw * cF;
    if (cx_of(i)) { V1x += (R)cx_of(i) * Lw_i; V2x += (R)cx_of(i) * t2; LFx += (R)cx_of(i) * tF; }
    if (cy_of(i)) { V1y += (R)cy_of(i) * Lw_i; V2y += (R)cy_of(i) * t2; LFy += (R)cy_of(i) * tF; }
    if (cz_of(i)) { V1z += (R)cz_of(i) * Lw_i; V2z += (R)cz_of(i) * t2; LFz += (R)cz_of(i) * tF; }
  }
  const R grho = (S0 + (R)3 * S1 + (R)4.5 * S2 - (R)1.5 * usq * S0) + B * Lw;
  const R gux = rho * ((R)3 * V1x + (R)9 * V2x - (R)3 * ux * S0) + k.pref * ((R)9 * LFx - (R)3 * k.Fx * Lw);
  const R guy = rho * ((R)3 * V1y + (R)9 * V2y - (R)3 * uy * S0) + k.pref * ((R)9 * LFy - (R)3 * k.Fy * Lw);
  const R guz = rho * ((R)3 * V1z + (R)9 * V2z - (R)3 * uz * S0) + k.pref * ((R)9 * LFz - (R)3 * k.Fz * Lw);
  const R gu_u = gux * ux + guy * uy + guz * uz;
  if (a > (R)0 && a < (R)1) {
    // dL/dB = sum_i lam_i (Os_i - Obgk_i) = Gl + sum_i lam_i (geqs_i - geq_opp(i) - geq_i/tau)
    // with geq_i = w_i (drho + rho (3cu + 4.5cu^2 - 1.5u^2)), geqs_i = w_i drho.
    const R w0 = (R)w_of(0);
    const R geq0 = w0 * (drho - rho * (R)1.5 * usq);
    R dLdB = Gl + lam[0] * ((w0 * drho - geq0) - geq0 * k.itau);
#pragma unroll
    for (int i = 1; i < Q; i += 2) {
      const int o = i + 1;
      const R w = (R)w_of(i);
      const R cu = (R)cx_of(i) * ux + (R)cy_of(i) * uy + (R)cz_of(i) * uz;
      const R even = w * (drho + rho * ((R)4.5 * cu * cu - (R)1.5 * usq));
      const R odd = w * rho * (R)3 * cu;
      const R geq_i = even + odd, geq_o = even - odd;
      const R geqs = w * drho;
      dLdB += lam[i] * ((geqs - geq_o) - geq_i * k.itau) + lam[o] * ((geqs - geq_i) - geq_o * k.itau);
    }
    const double dBda = ((double)k.tau + 1.0e-14) / ((double)den * (double)den);
    galpha[here] += (double)dLdB * dBda;
  }
#pragma unroll
  for (int kk = 0; kk < Q; ++kk) {
    const R gc = (R)cx_of(kk) * gux + (R)cy_of(kk) * guy + (R)cz_of(kk) * guz;
    const R mu = (A * lam[kk] + B * lam[opp_of(kk)]) + (grho + (gc - gu_u) * ir);
    sto<R, NT>(ob + (size_t)kk * plane, nb(kk), mu);
  }
}

// Seed: lam_T(x - c_k) = weight*coef (1 - alpha(x)) (e.c_k - e.u(x))/rho(x),
// galpha(x) = -weight*coef e.u(x), u = (j + F/2)/rho post-stream.
template <typename R, bool DEV>
__global__ __launch_bounds__(256) void k_adjoint_seed(const R* __restrict__ st, const R* __restrict__ alpha,
                                                      R* __restrict__ lam, double* __restrict__ galpha, ADims d,
                                                      double hfx, double hfy, double hfz, double ex, double ey,
                                                      double ez, double coef) {
  const int iz = blockIdx.x * blockDim.x + threadIdx.x;
  const int iy = blockIdx.y * blockDim.y + threadIdx.y;
  const int ix = blockIdx.z;
  if (iz >= d.nz || iy >= d.ny) return;
  const size_t N = (size_t)d.nx * d.ny * d.nz;
  const uint32_t here = acell(d, ix, iy, iz);
  uint32_t src[Q];
  double drho = 0.0, jx = 0.0, jy = 0.0, jz = 0.0;
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    src[i] = acell(d, wrap_step(ix, -cx_of(i), d.nx), wrap_step(iy, -cy_of(i), d.ny), wrap_step(iz, -cz_of(i), d.nz));
    double g = (double)st[(size_t)i * N + src[i]];
    if (!DEV) g -= w_of(i);
    drho += g;
    jx += (double)cx_of(i) * g;
    jy += (double)cy_of(i) * g;
    jz += (double)cz_of(i) * g;
  }
  const double rho = 1.0 + drho;
  const double eu = (ex * (jx + hfx) + ey * (jy + hfy) + ez * (jz + hfz)) / rho;
  const double a = (double)alpha[here];
  const double wgt = coef * (1.0 - a) / rho;
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    const double ec = ex * cx_of(i) + ey * cy_of(i) + ez * cz_of(i);
    lam[(size_t)i * N + src[i]] = (R)(wgt * (ec - eu));
  }
  galpha[here] = -coef * eu;
}

// 64 x 4 tiles, or 32 x 8 when rows are a multiple of 32 but not of 64
// (e.g. 96 cells: no half-empty wavefronts at the row end).
dim3 cell_grid(const ADims d, dim3& block) {
  const int bz = (d.nz % 64 != 0 && d.nz % 32 == 0) ? 32 : 64;
  block = dim3(bz, 256 / bz, 1);
  return dim3((d.nz + bz - 1) / bz, (d.ny + block.y - 1) / block.y, d.nx);
}

}  // namespace

int launch_smooth_alpha(const double* parts, const int32_t* boxes, const int32_t* chunks, int nchunks,
                        int chunk_cells, int nx, int ny, int nz, uint32_t* nfull, unsigned long long* logsum,
                        double* a64, float* a32, hipStream_t stream) {
  const ADims d{nx, ny, nz};
  const size_t n = (size_t)nx * ny * nz;
  const int fill_blocks = (int)((n + 255) / 256 < 8192 ? (n + 255) / 256 : 8192);
  k_fill_u32<<<fill_blocks, 256, 0, stream>>>(nfull, n);
  k_fill_u64<<<fill_blocks, 256, 0, stream>>>(logsum, n);
  int rc = check_launch("psc_smooth_alpha(clear)");
  if (rc) return rc;
  if (nchunks > 0) {
    k_smooth_scatter<<<nchunks, 256, 0, stream>>>(parts, boxes, chunks, chunk_cells, d, nfull, logsum);
    rc = check_launch("psc_smooth_alpha(scatter)");
    if (rc) return rc;
  }
  k_smooth_finish<<<fill_blocks, 256, 0, stream>>>(nfull, logsum, n, a64, a32);
  return check_launch("psc_smooth_alpha(finish)");
}

int launch_smooth_alpha_grad(const double* parts, const int32_t* boxes, const int32_t* chunks,
                             const int32_t* chunk_first, int nchunks, int chunk_cells, int np, int nx, int ny, int nz,
                             const uint32_t* nfull, const unsigned long long* logsum, const double* galpha,
                             double* partials, double* grad, hipStream_t stream) {
  const ADims d{nx, ny, nz};
  if (nchunks > 0) {
    k_smooth_grad<<<nchunks, 256, 0, stream>>>(parts, boxes, chunks, chunk_cells, d, nfull, logsum, galpha,
                                               partials);
    const int rc = check_launch("psc_smooth_alpha_grad");
    if (rc) return rc;
  }
  if (np == 0) return PSC_OK;
  k_smooth_grad_final<<<(4 * np + 255) / 256, 256, 0, stream>>>(chunk_first, np, partials, grad);
  return check_launch("psc_smooth_alpha_grad(final)");
}

int launch_adjoint_seed(const void* st, const void* alpha, void* lam, double* galpha, int storage, int nx, int ny,
                        int nz, double tau, double fx, double fy, double fz, double weight, hipStream_t stream) {
  const ADims d{nx, ny, nz};
  const double fm = sqrt(fx * fx + fy * fy + fz * fz);
  const double nu = (tau - 0.5) / 3.0;
  const double coef = weight * nu / (fm * (double)nx * (double)ny * (double)nz);
  dim3 block;
  const dim3 grid = cell_grid(d, block);
  if (storage == PSC_F32_DEV)
    k_adjoint_seed<float, true><<<grid, block, 0, stream>>>((const float*)st, (const float*)alpha, (float*)lam,
                                                            galpha, d, 0.5 * fx, 0.5 * fy, 0.5 * fz, fx / fm,
                                                            fy / fm, fz / fm, coef);
  else
    k_adjoint_seed<double, false><<<grid, block, 0, stream>>>((const double*)st, (const double*)alpha,
                                                              (double*)lam, galpha, d, 0.5 * fx, 0.5 * fy,
                                                              0.5 * fz, fx / fm, fy / fm, fz / fm, coef);
  return check_launch("psc_adjoint_seed");
}

int launch_adjoint_step(const void* st, const void* lin, void* lout, const void* alpha, double* galpha, int storage,
                        int nx, int ny, int nz, double tau, double fx, double fy, double fz, int bz, int by,
                        int flags, hipStream_t stream) {
  const ADims d{nx, ny, nz};
  const size_t esz = storage == PSC_F32_DEV ? 4 : 8;
  if ((size_t)nx * ny * nz * esz > 0xffffffffull)
    return set_error(PSC_EINVAL, "psc_adjoint_step: a direction plane must stay below 4 GiB");
  dim3 block;
  dim3 grid = cell_grid(d, block);
  if (bz > 0 && by > 0) {
    block = dim3(bz, by, 1);
    grid = dim3((nz + bz - 1) / bz, (ny + by - 1) / by, nx);
  }
  // bits 8-15: cap on resident workgroups per CU through an LDS reservation
  const int bpc = (flags >> 8) & 0xff;
  size_t lds = 0;
  if (bpc > 0) {
    lds = (163840 - 8192) / (size_t)bpc;
    if (lds > 65536) lds = 65536;
  }
  const bool nt = (flags & PSC_STEP_NONTEMPORAL) != 0;
  if (storage == PSC_F32_DEV) {
    const AdjConsts<float> k{(float)tau, (float)(1.0 / tau), (float)(1.0 - 0.5 / tau), (float)fx, (float)fy,
                             (float)fz};
    if (nt)
      k_adjoint_step<float, true, true><<<grid, block, lds, stream>>>((const float*)st, (const float*)lin,
                                                                    (float*)lout, (const float*)alpha, galpha, d, k);
    else
      k_adjoint_step<float, true, false><<<grid, block, lds, stream>>>((const float*)st, (const float*)lin,
                                                                     (float*)lout, (const float*)alpha, galpha, d, k);
  } else {
    const AdjConsts<double> k{tau, 1.0 / tau, 1.0 - 0.5 / tau, fx, fy, fz};
    k_adjoint_step<double, false, false><<<grid, block, lds, stream>>>((const double*)st, (const double*)lin,
                                                                     (double*)lout, (const double*)alpha, galpha,
                                                                     d, k);
  }
  return check_launch("psc_adjoint_step");
}

}  // namespace psc
