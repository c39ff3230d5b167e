// Particle-load reduction bodies (psc3d.py:255-281, 375-407) of the K3
// kernels.  FP contraction is switched off inside, so the bits do not depend
// on the translation unit that instantiates them.
#pragma once

#include "psc_common.h"

namespace psc {

constexpr int kLocBits = 10;  // box-local coordinate bits of an entry (extent <= 1024)
// Entries per thread per round (a register batch; it does not change the bits).
// Two keep the chunk pass at 5 waves per SIMD: with eight it needed 144 VGPRs
// (3 waves) and took 313 instead of 185 us at 512^3 (18.7 M entries).
constexpr int kLoadsKB = 2;

struct LoadsArgs {
  const double* parts;       // P x 10
  const int32_t* boxes;      // P x 8
  const int32_t* chunks;     // C x 2
  const int32_t* chunk_first;  // P + 1
  int np, nchunks, chunk_cells;
  const uint32_t* ent_cell;
  const uint32_t* ent_loc;
  const uint16_t* ent_cnt;
  const int32_t* chunk_nnz;
  const uint32_t* csum;
};

// Sum over the entries of chunk c of -owner*imb and its torque, into
// partials[6c..6c+5].  The summation order is fixed by a 256-lane virtual
// layout (entry k goes to virtual lane k % 256, lanes summed by a wave
// butterfly, the four virtual waves in order): physical wave p of a
// workgroup of nt threads (nt a multiple of 64, at most 256) runs virtual
// waves p, p + nt/64, ..., so every such nt produces the same bits, and so
// does every register batch KB.  lds needs 4*6 doubles.
constexpr int kLoadsVL = 256;

template <typename TI, int KB = kLoadsKB>
__device__ __forceinline__ void loads_chunk_body(const LoadsArgs& a, const TI* __restrict__ imb, int nx, int ny,
                                                 int nz, int c, int tid, int nt, double* __restrict__ partials,
                                                 double* lds) {
#pragma clang fp contract(off)
  const int pid = a.chunks[2 * c];
  const int nnz = a.chunk_nnz[c];
  const int32_t* b = a.boxes + 8 * (size_t)pid;
  const int x0 = b[0], y0 = b[1], z0 = b[2];
  const double* q = a.parts + 10 * (size_t)pid;
  const double px = q[0], py = q[1], pz = q[2];
  const size_t out0 = (size_t)c * a.chunk_cells;
  constexpr uint32_t kMask = (1u << kLocBits) - 1u;
  const int lane = tid & 63, pwave = tid >> 6, nwave = nt >> 6;
  for (int vw = pwave; vw < kLoadsVL / 64; vw += nwave) {
    const int v = vw * 64 + lane;  // virtual lane
    double acc[6] = {0, 0, 0, 0, 0, 0};
    for (int k0 = 0; k0 < nnz; k0 += KB * kLoadsVL) {
      uint32_t cell[KB], loc[KB], cnt[KB], tot[KB];
      TI ix_[KB], iy_[KB], iz_[KB];
#pragma unroll
      for (int j = 0; j < KB; ++j) {
        const int k = k0 + j * kLoadsVL + v;
        const bool ok = k < nnz;
        cell[j] = ok ? a.ent_cell[out0 + k] : 0u;
        loc[j] = ok ? a.ent_loc[out0 + k] : 0u;
        cnt[j] = ok ? (uint32_t)a.ent_cnt[out0 + k] : 0u;
      }
#pragma unroll
      for (int j = 0; j < KB; ++j) {
        const bool ok = cnt[j] != 0u;
        tot[j] = ok ? a.csum[cell[j]] : 1u;
        ix_[j] = ok ? imb[3 * (size_t)cell[j]] : (TI)0;
        iy_[j] = ok ? imb[3 * (size_t)cell[j] + 1] : (TI)0;
        iz_[j] = ok ? imb[3 * (size_t)cell[j] + 2] : (TI)0;
      }
#pragma unroll
      for (int j = 0; j < KB; ++j) {
        if (cnt[j] == 0u) continue;
        // owner = alpha_p / sum alpha = cnt / count_sum (exactly 1 for one owner)
        const double owner = (cnt[j] == tot[j]) ? 1.0 : (double)cnt[j] / (double)tot[j];
        const double nfx = -owner * (double)ix_[j];
        const double nfy = -owner * (double)iy_[j];
        const double nfz = -owner * (double)iz_[j];
        const double rx = ((double)(x0 + (int)(loc[j] & kMask)) + 0.5) - px;
        const double ry = ((double)(y0 + (int)((loc[j] >> kLocBits) & kMask)) + 0.5) - py;
        const double rz = ((double)(z0 + (int)((loc[j] >> (2 * kLocBits)) & kMask)) + 0.5) - pz;
        acc[0] += nfx;
        acc[1] += nfy;
        acc[2] += nfz;
        acc[3] += ry * nfz - rz * nfy;
        acc[4] += rz * nfx - rx * nfz;
        acc[5] += rx * nfy - ry * nfx;
      }
    }
    // wave butterfly over this virtual wave, sum kept per virtual wave
#pragma unroll
    for (int j = 0; j < 6; ++j)
#pragma unroll
      for (int off = 32; off >= 1; off >>= 1) acc[j] += __shfl_xor(acc[j], off, 64);
    if (lane == 0) {
#pragma unroll
      for (int j = 0; j < 6; ++j) lds[vw * 6 + j] = acc[j];
    }
  }
  __syncthreads();
  if (tid == 0) {
#pragma unroll
    for (int j = 0; j < 6; ++j) {
      double s = lds[j];
      for (int w = 1; w < kLoadsVL / 64; ++w) s += lds[w * 6 + j];
      partials[6 * (size_t)c + j] = s;
    }
  }
  __syncthreads();  // lds is reused when a workgroup sums several chunks
}

// Per-particle sum of the chunk partials (first wavefront of the group).
__device__ __forceinline__ void loads_final_body(const int32_t* __restrict__ chunk_first, int pid, int tid,
                                                 const double* __restrict__ partials, double* __restrict__ loads) {
#pragma clang fp contract(off)
  if (tid >= 64) return;
  const int c0 = chunk_first[pid], c1 = chunk_first[pid + 1];
  double acc[6] = {0, 0, 0, 0, 0, 0};
  for (int cc = c0 + tid; cc < c1; cc += 64)
#pragma unroll
    for (int j = 0; j < 6; ++j) acc[j] += partials[6 * (size_t)cc + j];
#pragma unroll
  for (int j = 0; j < 6; ++j)
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) acc[j] += __shfl_xor(acc[j], off, 64);
  if (tid == 0)
#pragma unroll
    for (int j = 0; j < 6; ++j) loads[6 * (size_t)pid + j] = acc[j];
}

}  // namespace psc
