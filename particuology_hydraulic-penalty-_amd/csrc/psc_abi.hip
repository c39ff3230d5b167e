// extern "C" entry points of libpsc_hip.so: argument checking, error
// reporting and dispatch to the kernel launchers of psc_exact.hip and
// psc_fast.hip.  The declarations and their reference citations are in
// include/psc_hip.h.
#include <stdarg.h>
#include <stdio.h>

#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "psc_common.h"
#include "psc_loads.h"
#include "../../include/psc_hip.h"

namespace psc {

namespace {
thread_local std::string g_last_error;
}

int set_error(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}

int check_launch(const char* what) {
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_error((int)e, "%s: %s", what, hipGetErrorString(e));
  return PSC_OK;
}

// launchers defined in the kernel translation units
int launch_step_f64(const double*, double*, const uint32_t*, const double*, double*, int, int, int, int, int, double,
                    double, double, double, int, int, int, hipStream_t);
int launch_step_f32(const float*, float*, const uint32_t*, const float*, float*, int, int, int, int, int, double,
                    double, double, double, int, int, int, hipStream_t, const uint32_t* = nullptr,
                    const void* = nullptr, int = 0);
int launch_voxelize_owners(const double*, const int32_t*, int, const int32_t*, int, int, int, int, int, int, uint32_t*,
                           uint32_t*, uint16_t*, int32_t*, uint32_t*, uint32_t*, float*, void*, hipStream_t);
int launch_clear_owners(const uint32_t*, const uint16_t*, const int32_t*, int, int, int, int, int, uint32_t*, float*,
                        hipStream_t);
int launch_voxelize(const double*, const int32_t*, const int32_t*, int, int, int, int, int, int, int, uint32_t*,
                    uint32_t*, uint16_t*, int32_t*, uint32_t*, void*, int, uint32_t, hipStream_t);
int launch_step_f32_batched(const float*, float*, const uint32_t*, float*, size_t, const void*, const void*, int, int,
                            double, int, int, hipStream_t, const LoadsArgs*, const float*, double*, const double*,
                            double*, int);
int launch_loads(const double*, const int32_t*, const int32_t*, const int32_t*, int, int, int, const uint32_t*,
                 const uint32_t*, const uint16_t*, const int32_t*, const uint32_t*, const void*, int, int, int, int,
                 double*, double*, hipStream_t);
int launch_moments(const void*, int, int, const uint32_t*, const void*, int, int, int, int, unsigned long long, int,
                   double, double, double, double*, double*, double*, hipStream_t);
int64_t moments_blocks(int64_t);
int moments_scratch(double**, int64_t*);
int launch_clear_entries(const uint32_t*, const int32_t*, int, int, int, int, int, uint32_t*, void*, int, hipStream_t);
int launch_loads_chunks(const LoadsArgs&, const void*, int, int, int, int, double*, hipStream_t);
int launch_loads_final(const int32_t*, int, const double*, double*, hipStream_t);
int launch_loads_flush(const LoadsArgs&, const void*, int, int, int, int, double*, const double*, double*,
                       hipStream_t);
int launch_step_f32_loads(const float*, float*, const uint32_t*, float*, const float*, const LoadsArgs&, double*,
                          const double*, double*, int, int, int, int, int, double, double, double, double, int, int,
                          int, hipStream_t, int);

int launch_smooth_alpha(const double*, const int32_t*, const int32_t*, int, int, int, int, int, uint32_t*,
                        unsigned long long*, double*, float*, hipStream_t);
int launch_smooth_alpha_grad(const double*, const int32_t*, const int32_t*, const int32_t*, int, int, int, int, int,
                             int, const uint32_t*, const unsigned long long*, const double*, double*, double*,
                             hipStream_t);
int launch_adjoint_seed(const void*, const void*, void*, double*, int, int, int, int, double, double, double, double,
                        double, hipStream_t);
int launch_adjoint_step(const void*, const void*, void*, const void*, double*, int, int, int, int, double, double,
                        double, double, int, int, int, hipStream_t);

int launch_bb_step_f64(const double*, double*, const uint8_t*, int, int, int, double, double, double, double, int, int,
                       int, hipStream_t);
int launch_bb_step_f32(const float*, float*, const uint8_t*, int, int, int, double, double, double, double, int, int,
                       int, hipStream_t);
int launch_voxelize_binary(const double*, const int32_t*, int, const double*, int, int, int, uint8_t*, hipStream_t);

int launch_export(const void*, int, int, int, int, int, double*, hipStream_t);
int launch_import(const double*, int, int, int, int, int, void*, hipStream_t);
int launch_init_plain(double*, size_t, hipStream_t);
int launch_equilibrium_f64(const double*, const double*, const double*, const double*, size_t, double*, hipStream_t);
int launch_guo_f64(double, double, double, const double*, const double*, const double*, size_t, double, double*,
                   hipStream_t);

namespace {

__global__ void k_clear16(uint4* __restrict__ p, size_t n16) {
  const uint4 z = make_uint4(0u, 0u, 0u, 0u);
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += (size_t)gridDim.x * blockDim.x)
    p[i] = z;
}

__global__ void k_clear1(unsigned char* __restrict__ p, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = 0;
}

// Streaming copy: 4 independent 16-byte non-temporal loads in flight per
// thread, then the 4 stores; grid-stride over 64-byte groups.
typedef float f32x4 __attribute__((ext_vector_type(4)));
__global__ __launch_bounds__(256) void k_copy4(const f32x4* __restrict__ s, f32x4* __restrict__ d, size_t n4) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (; i + 3 * stride < n4; i += 4 * stride) {
    f32x4 v[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) v[u] = __builtin_nontemporal_load(s + i + u * stride);
#pragma unroll
    for (int u = 0; u < 4; ++u) __builtin_nontemporal_store(v[u], d + i + u * stride);
  }
  for (; i < n4; i += stride) d[i] = s[i];
}

__global__ void k_copy1(const float* __restrict__ s, float* __restrict__ d, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    d[i] = s[i];
}

int grid_cap(size_t n, int threads) {
  size_t g = (n + threads - 1) / threads;
  if (g > 16384) g = 16384;
  return g < 1 ? 1 : (int)g;
}

// Largest FP32 lattice: the kernels address a direction plane with 32-bit
// byte offsets (the last cell of a 2^30-cell plane is at 2^32 - 4).
constexpr unsigned long long kMaxF32Cells = 1ull << 30;

bool dims_ok(int nx, int ny, int nz) {
  if (nx < 1 || ny < 1 || nz < 1) return false;
  const unsigned long long n = (unsigned long long)nx * ny * nz;
  return n <= 0x7fffffffULL;
}

}  // namespace
}  // namespace psc

using namespace psc;

namespace psc {
// Argument checks of psc_step; resolves the default block shape into *bz, *by.
int step_args(const void* d_in, const void* d_out, const uint32_t* d_count_sum, const void* d_us_num, int storage,
              int layout, int nx, int ny, int nz, int subsamples, double tau, int flags, int* bz, int* by) {
  if (!dims_ok(nx, ny, nz)) return set_error(PSC_EINVAL, "psc_step: bad lattice %dx%dx%d", nx, ny, nz);
  // the FP32 kernels address a direction plane with 32-bit byte offsets
  if (storage == PSC_F32_DEV && (unsigned long long)nx * ny * nz > kMaxF32Cells)
    return set_error(PSC_EINVAL, "psc_step: FP32 lattice %dx%dx%d exceeds 2^30 cells (32-bit byte offsets in a plane)", nx, ny,
                     nz);
  if (d_in == nullptr || d_out == nullptr || d_count_sum == nullptr) return set_error(PSC_EINVAL, "psc_step: null");
  if (d_in == d_out) return set_error(PSC_EINVAL, "psc_step: input and output alias");
  if (layout != PSC_PUSH && layout != PSC_PULL) return set_error(PSC_EINVAL, "psc_step: layout %d", layout);
  if (subsamples < 0 || subsamples * subsamples * subsamples > 65535)
    return set_error(PSC_EINVAL, "psc_step: subsamples %d", subsamples);
  if (subsamples == PSC_ALPHA_FIELD && d_us_num != nullptr)
    return set_error(PSC_EINVAL, "psc_step: a solid-fraction field (subsamples 0) takes no solid velocity");
  if (!(tau > 0.5)) return set_error(PSC_EINVAL, "psc_step: tau %g must exceed 1/2", tau);
  // default: 64-cell row tiles, 32 when rows are a multiple of 32 but not of 64
  if (*bz <= 0) *bz = (nz >= 64 && !(nz % 64 != 0 && nz % 32 == 0)) ? 64 : 32;
  if (*by <= 0) *by = 256 / *bz > 0 ? 256 / *bz : 1;
  if ((flags & ~(PSC_STEP_NONTEMPORAL | PSC_STEP_BPC_MASK | PSC_STEP_CPT_MASK | PSC_STEP_LINEAR | PSC_STEP_PLANES64)) !=
      0)
    return set_error(PSC_EINVAL, "psc_step: unknown flags 0x%x", flags);
  if ((flags & PSC_STEP_PLANES64) && storage != PSC_F32_DEV)
    return set_error(PSC_EINVAL, "psc_step: 64-cell-aligned planes (PSC_STEP_PLANES64) need FP32 storage");
  if (((flags & PSC_STEP_CPT_MASK) >> 16) > 2) return set_error(PSC_EINVAL, "psc_step: cells per thread > 2");
  if ((flags & PSC_STEP_LINEAR) && ((flags & PSC_STEP_CPT_MASK) >> 16) > 1)
    return set_error(PSC_EINVAL, "psc_step: the flat mapping takes one cell per thread");
  if (*bz * *by > 256 || *bz * *by < 1 || ((*bz * *by) % 64) != 0)
    return set_error(PSC_ELAUNCH, "psc_step: block %dx%d must hold a multiple of 64 threads, at most 256", *bz, *by);
  return PSC_OK;
}
}  // namespace psc

extern "C" {

int psc_version(void) { return 10000; }

const char* psc_last_error(void) { return g_last_error.c_str(); }

int psc_lattice_constants(int32_t* c, double* w, int32_t* opp) {
  if (c == nullptr || w == nullptr || opp == nullptr) return set_error(PSC_EINVAL, "psc_lattice_constants: null");
  for (int i = 0; i < Q; ++i) {
    c[3 * i] = cx_of(i);
    c[3 * i + 1] = cy_of(i);
    c[3 * i + 2] = cz_of(i);
    w[i] = w_of(i);
    opp[i] = opp_of(i);
  }
  return PSC_OK;
}

int psc_clear(void* d_ptr, int64_t nbytes, void* stream) {
  if (nbytes < 0) return set_error(PSC_EINVAL, "psc_clear: negative size");
  if (nbytes == 0) return PSC_OK;
  if (d_ptr == nullptr) return set_error(PSC_EINVAL, "psc_clear: null pointer");
  hipStream_t s = (hipStream_t)stream;
  const uintptr_t a = (uintptr_t)d_ptr;
  if ((a & 15u) == 0) {
    const size_t n16 = (size_t)nbytes / 16;
    if (n16) k_clear16<<<grid_cap(n16, 256), 256, 0, s>>>((uint4*)d_ptr, n16);
    const size_t tail = (size_t)nbytes - n16 * 16;
    if (tail) k_clear1<<<1, 64, 0, s>>>((unsigned char*)d_ptr + n16 * 16, tail);
  } else {
    k_clear1<<<grid_cap((size_t)nbytes, 256), 256, 0, s>>>((unsigned char*)d_ptr, (size_t)nbytes);
  }
  return check_launch("psc_clear");
}

int psc_init_state(void* d_state, int32_t storage, int64_t ncell, void* stream) {
  if (d_state == nullptr || ncell <= 0) return set_error(PSC_EINVAL, "psc_init_state: bad state or size");
  if (storage == PSC_F32_DEV) return psc_clear(d_state, (int64_t)Q * ncell * 4, stream);
  if (storage == PSC_F64_PLAIN) return launch_init_plain((double*)d_state, (size_t)ncell, (hipStream_t)stream);
  return set_error(PSC_EINVAL, "psc_init_state: storage %d", storage);
}

int psc_equilibrium(const double* d_rho, const double* d_ux, const double* d_uy, const double* d_uz, int64_t n,
                    double* d_out, void* stream) {
  if (n < 0) return set_error(PSC_EINVAL, "psc_equilibrium: negative size");
  if (n == 0) return PSC_OK;
  if (d_rho == nullptr || d_ux == nullptr || d_uy == nullptr || d_uz == nullptr || d_out == nullptr)
    return set_error(PSC_EINVAL, "psc_equilibrium: null pointer");
  return launch_equilibrium_f64(d_rho, d_ux, d_uy, d_uz, (size_t)n, d_out, (hipStream_t)stream);
}

int psc_guo_force(double fx, double fy, double fz, const double* d_ux, const double* d_uy, const double* d_uz,
                  int64_t n, double tau, double* d_out, void* stream) {
  if (n < 0) return set_error(PSC_EINVAL, "psc_guo_force: negative size");
  if (!(tau > 0.0)) return set_error(PSC_EINVAL, "psc_guo_force: tau %g must be positive", tau);
  if (n == 0) return PSC_OK;
  if (d_ux == nullptr || d_uy == nullptr || d_uz == nullptr || d_out == nullptr)
    return set_error(PSC_EINVAL, "psc_guo_force: null pointer");
  return launch_guo_f64(fx, fy, fz, d_ux, d_uy, d_uz, (size_t)n, tau, d_out, (hipStream_t)stream);
}

int psc_voxelize(const double* d_particles, const int32_t* d_boxes, const int32_t* d_chunks, int32_t nchunks,
                 int32_t chunk_cells, int32_t nx, int32_t ny, int32_t nz, int32_t pitch_z, int32_t subsamples,
                 uint32_t* d_ent_cell,
                 uint32_t* d_ent_loc, uint16_t* d_ent_cnt, int32_t* d_chunk_nnz, uint32_t* d_count_sum,
                 void* d_us_num, int32_t us_precision, uint32_t cell_base, void* stream) {
  if (!dims_ok(nx, ny, nz)) return set_error(PSC_EINVAL, "psc_voxelize: bad lattice %dx%dx%d", nx, ny, nz);
  if (subsamples < 1 || subsamples * subsamples * subsamples > 65535)
    return set_error(PSC_EINVAL, "psc_voxelize: subsamples %d outside [1, 40]", subsamples);
  if (nchunks < 0 || chunk_cells < 1) return set_error(PSC_EINVAL, "psc_voxelize: bad chunking");
  if (nchunks > 0 && (d_particles == nullptr || d_boxes == nullptr || d_chunks == nullptr || d_ent_cell == nullptr ||
                      d_ent_loc == nullptr || d_ent_cnt == nullptr || d_chunk_nnz == nullptr ||
                      d_count_sum == nullptr))
    return set_error(PSC_EINVAL, "psc_voxelize: null pointer");
  if (d_us_num != nullptr && us_precision != 1 && us_precision != 2)
    return set_error(PSC_EINVAL, "psc_voxelize: us_precision %d", us_precision);
  const int pz = pitch_z > 0 ? pitch_z : nz;
  if (pz < nz || !dims_ok(nx, ny, pz)) return set_error(PSC_EINVAL, "psc_voxelize: pitch_z %d < nz %d", pitch_z, nz);
  return launch_voxelize(d_particles, d_boxes, d_chunks, nchunks, chunk_cells, nx, ny, nz, pz, subsamples, d_ent_cell,
                         d_ent_loc, d_ent_cnt, d_chunk_nnz, d_count_sum, d_us_num, d_us_num ? us_precision : 0,
                         cell_base, (hipStream_t)stream);
}

int psc_clear_entries(const uint32_t* d_ent_cell, const int32_t* d_chunk_nnz, int32_t nchunks, int32_t chunk_cells,
                      int32_t nx, int32_t ny, int32_t nz, uint32_t* d_count_sum, void* d_us_num,
                      int32_t us_precision, void* stream) {
  if (!dims_ok(nx, ny, nz)) return set_error(PSC_EINVAL, "psc_clear_entries: bad lattice");
  if (nchunks < 0 || chunk_cells < 1) return set_error(PSC_EINVAL, "psc_clear_entries: sizes");
  if (nchunks > 0 && (d_ent_cell == nullptr || d_chunk_nnz == nullptr || d_count_sum == nullptr))
    return set_error(PSC_EINVAL, "psc_clear_entries: null pointer");
  if (d_us_num != nullptr && us_precision != 1 && us_precision != 2)
    return set_error(PSC_EINVAL, "psc_clear_entries: us_precision %d", us_precision);
  return launch_clear_entries(d_ent_cell, d_chunk_nnz, nchunks, chunk_cells, nx, ny, nz, d_count_sum, d_us_num,
                              d_us_num ? us_precision : 0, (hipStream_t)stream);
}

int psc_voxelize_owners(const double* d_particles, const int32_t* d_boxes, int32_t nparticles,
                        const int32_t* d_chunks, int32_t nchunks, int32_t chunk_cells, int32_t nx, int32_t ny,
                        int32_t nz, int32_t subsamples, uint32_t* d_ent_cell, uint32_t* d_ent_loc,
                        uint16_t* d_ent_cnt, int32_t* d_chunk_nnz, uint32_t* d_count_sum, uint32_t* d_owner,
                        float* d_us_shared, void* d_motion, void* stream) {
  if (!dims_ok(nx, ny, nz)) return set_error(PSC_EINVAL, "psc_voxelize_owners: bad lattice %dx%dx%d", nx, ny, nz);
  if (subsamples < 1 || subsamples * subsamples * subsamples > 65535)
    return set_error(PSC_EINVAL, "psc_voxelize_owners: subsamples %d outside [1, 40]", subsamples);
  if (nchunks < 0 || chunk_cells < 1 || nparticles < 0) return set_error(PSC_EINVAL, "psc_voxelize_owners: sizes");
  if ((nparticles > 0 || nchunks > 0) && (d_particles == nullptr || d_boxes == nullptr || d_motion == nullptr))
    return set_error(PSC_EINVAL, "psc_voxelize_owners: null particle, box or motion table");
  if (nchunks > 0 && (d_chunks == nullptr || d_ent_cell == nullptr || d_ent_loc == nullptr || d_ent_cnt == nullptr ||
                      d_chunk_nnz == nullptr || d_count_sum == nullptr || d_owner == nullptr ||
                      d_us_shared == nullptr))
    return set_error(PSC_EINVAL, "psc_voxelize_owners: null pointer");
  return launch_voxelize_owners(d_particles, d_boxes, nparticles, d_chunks, nchunks, chunk_cells, nx, ny, nz,
                                subsamples, d_ent_cell, d_ent_loc, d_ent_cnt, d_chunk_nnz, d_count_sum, d_owner,
                                d_us_shared, d_motion, (hipStream_t)stream);
}

int psc_clear_entries_owners(const uint32_t* d_ent_cell, const uint16_t* d_ent_cnt, const int32_t* d_chunk_nnz,
                             int32_t nchunks, int32_t chunk_cells, int32_t nx, int32_t ny, int32_t nz,
                             uint32_t* d_count_sum, float* d_us_shared, void* stream) {
  if (!dims_ok(nx, ny, nz)) return set_error(PSC_EINVAL, "psc_clear_entries_owners: bad lattice");
  if (nchunks < 0 || chunk_cells < 1) return set_error(PSC_EINVAL, "psc_clear_entries_owners: sizes");
  if (nchunks > 0 && (d_ent_cell == nullptr || d_ent_cnt == nullptr || d_chunk_nnz == nullptr ||
                      d_count_sum == nullptr || d_us_shared == nullptr))
    return set_error(PSC_EINVAL, "psc_clear_entries_owners: null pointer");
  return launch_clear_owners(d_ent_cell, d_ent_cnt, d_chunk_nnz, nchunks, chunk_cells, nx, ny, nz, d_count_sum,
                             d_us_shared, (hipStream_t)stream);
}

int psc_step_owners(const float* d_in, float* d_out, const uint32_t* d_count_sum, const uint32_t* d_owner,
                    const float* d_us_shared, const void* d_motion, int32_t nparticles, float* d_imb, int32_t nx,
                    int32_t ny, int32_t nz, int32_t subsamples, double tau, double fx, double fy, double fz,
                    int32_t block_z, int32_t block_y, int32_t flags, void* stream) {
  if (d_owner == nullptr || d_us_shared == nullptr || (nparticles > 0 && d_motion == nullptr))
    return set_error(PSC_EINVAL, "psc_step_owners: null owner table, numerators or motion table");
  if (nparticles < 0) return set_error(PSC_EINVAL, "psc_step_owners: nparticles %d", nparticles);
  if (((uintptr_t)d_motion & 15) != 0) return set_error(PSC_EINVAL, "psc_step_owners: motion table not 16-byte aligned");
  int bz = block_z, by = block_y;
  const int rc = psc::step_args(d_in, d_out, d_count_sum, d_us_shared, PSC_F32_DEV, PSC_PULL, nx, ny, nz, subsamples,
                                tau, flags, &bz, &by);
  if (rc) return rc;
  return launch_step_f32(d_in, d_out, d_count_sum, d_us_shared, d_imb, PSC_PULL, nx, ny, nz, subsamples, tau, fx, fy,
                         fz, bz, by, flags, (hipStream_t)stream, d_owner, d_motion, nparticles);
}

int psc_step(const void* d_in, void* d_out, const uint32_t* d_count_sum, const void* d_us_num, void* d_imb,
             int32_t storage, int32_t layout, int32_t nx, int32_t ny, int32_t nz, int32_t subsamples, double tau,
             double fx, double fy, double fz, int32_t block_z, int32_t block_y, int32_t flags, void* stream) {
  int bz = block_z, by = block_y;
  const int rc = psc::step_args(d_in, d_out, d_count_sum, d_us_num, storage, layout, nx, ny, nz, subsamples, tau,
                                flags, &bz, &by);
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  if (storage == PSC_F32_DEV)
    return launch_step_f32((const float*)d_in, (float*)d_out, d_count_sum, (const float*)d_us_num, (float*)d_imb,
                           layout, nx, ny, nz, subsamples, tau, fx, fy, fz, bz, by, flags, s);
  if (storage == PSC_F64_PLAIN)
    return launch_step_f64((const double*)d_in, (double*)d_out, d_count_sum, (const double*)d_us_num,
                           (double*)d_imb, layout, nx, ny, nz, subsamples, tau, fx, fy, fz, bz, by, flags, s);
  return set_error(PSC_EINVAL, "psc_step: storage %d", storage);
}

int psc_step_batched(const float* d_in, float* d_out, const uint32_t* d_count_sum, float* d_imb, int64_t imb_stride,
                     const PscWindow* d_windows, const int32_t* d_items, int32_t nitems, int32_t subsamples,
                     double tau, int32_t flags, int32_t grid_blocks, void* stream) {
  if (d_in == nullptr || d_out == nullptr || d_count_sum == nullptr || d_windows == nullptr ||
      (nitems > 0 && d_items == nullptr))
    return set_error(PSC_EINVAL, "psc_step_batched: null pointer");
  if (d_in == d_out) return set_error(PSC_EINVAL, "psc_step_batched: input and output alias");
  if (nitems < 0 || grid_blocks < 0) return set_error(PSC_EINVAL, "psc_step_batched: sizes");
  if (d_imb != nullptr && imb_stride <= 0) return set_error(PSC_EINVAL, "psc_step_batched: imb stride");
  if (subsamples < 1 || subsamples * subsamples * subsamples > 65535)
    return set_error(PSC_EINVAL, "psc_step_batched: subsamples %d", subsamples);
  if (!(tau > 0.5)) return set_error(PSC_EINVAL, "psc_step_batched: tau %g must exceed 1/2", tau);
  if ((flags & ~(PSC_STEP_NONTEMPORAL | PSC_STEP_BPC_MASK | PSC_STEP_CPT_MASK | PSC_BATCH_WG512 | PSC_BATCH_IMB_NT)) != 0)
    return set_error(PSC_EINVAL, "psc_step_batched: unknown flags 0x%x", flags);
  return launch_step_f32_batched(d_in, d_out, d_count_sum, d_imb, (size_t)imb_stride, d_windows, d_items, nitems,
                                 subsamples, tau, flags, grid_blocks, (hipStream_t)stream, nullptr, nullptr, nullptr,
                                 nullptr, nullptr, 1);
}

int psc_step_batched_loads(const float* d_in, float* d_out, const uint32_t* d_count_sum, float* d_imb,
                           int64_t imb_stride, const PscWindow* d_windows, const int32_t* d_items, int32_t nitems,
                           int32_t subsamples, double tau, int32_t flags, int32_t grid_blocks,
                           const PscLoadsTables* t, const float* d_imb_prev, double* d_partials_out,
                           const double* d_partials_prev, double* d_loads, int32_t period, void* stream) {
  if (d_in == nullptr || d_out == nullptr || d_count_sum == nullptr || d_windows == nullptr || d_imb == nullptr ||
      t == nullptr || (nitems > 0 && d_items == nullptr))
    return set_error(PSC_EINVAL, "psc_step_batched_loads: null pointer");
  if (d_in == d_out) return set_error(PSC_EINVAL, "psc_step_batched_loads: input and output alias");
  if (nitems < 0 || grid_blocks < 0 || imb_stride <= 0 || imb_stride > 0x7fffffffll || period < 1)
    return set_error(PSC_EINVAL, "psc_step_batched_loads: sizes");
  if (subsamples < 1 || subsamples * subsamples * subsamples > 65535)
    return set_error(PSC_EINVAL, "psc_step_batched_loads: subsamples %d", subsamples);
  if (!(tau > 0.5)) return set_error(PSC_EINVAL, "psc_step_batched_loads: tau %g must exceed 1/2", tau);
  if ((flags & ~(PSC_STEP_NONTEMPORAL | PSC_STEP_BPC_MASK | PSC_STEP_CPT_MASK | PSC_BATCH_WG512 | PSC_BATCH_IMB_NT)) != 0)
    return set_error(PSC_EINVAL, "psc_step_batched_loads: unknown flags 0x%x", flags);
  if (t->nparticles < 0 || t->nchunks < 0 || t->chunk_cells < 1)
    return set_error(PSC_EINVAL, "psc_step_batched_loads: table sizes");
  const bool tables_ok = t->particles != nullptr && t->boxes != nullptr && t->chunks != nullptr &&
                         t->ent_cell != nullptr && t->ent_loc != nullptr && t->ent_cnt != nullptr &&
                         t->chunk_nnz != nullptr && t->count_sum != nullptr;
  if (d_imb_prev != nullptr && (d_imb_prev == d_imb || (t->nchunks > 0 && !tables_ok)))
    return set_error(PSC_EINVAL, "psc_step_batched_loads: chunk tables / imb buffers");
  if (d_partials_prev != nullptr && (t->chunk_first == nullptr || d_partials_prev == d_partials_out))
    return set_error(PSC_EINVAL, "psc_step_batched_loads: partial-sum buffers");
  const LoadsArgs la{t->particles, t->boxes, t->chunks, t->chunk_first, t->nparticles, t->nchunks, t->chunk_cells,
                     t->ent_cell, t->ent_loc, t->ent_cnt, t->chunk_nnz, t->count_sum};
  return launch_step_f32_batched(d_in, d_out, d_count_sum, d_imb, (size_t)imb_stride, d_windows, d_items, nitems,
                                 subsamples, tau, flags, grid_blocks, (hipStream_t)stream, &la, d_imb_prev,
                                 d_partials_out, d_partials_prev, d_loads, period);
}

int psc_particle_loads(const double* d_particles, const int32_t* d_boxes, const int32_t* d_chunks,
                       const int32_t* d_chunk_first, int32_t nparticles, int32_t nchunks, int32_t chunk_cells,
                       const uint32_t* d_ent_cell, const uint32_t* d_ent_loc, const uint16_t* d_ent_cnt,
                       const int32_t* d_chunk_nnz, const uint32_t* d_count_sum, const void* d_imb,
                       int32_t imb_precision, int32_t nx, int32_t ny, int32_t nz, double* d_partials,
                       double* d_loads, void* stream) {
  if (!dims_ok(nx, ny, nz)) return set_error(PSC_EINVAL, "psc_particle_loads: bad lattice");
  if (nparticles < 0 || nchunks < 0 || chunk_cells < 1) return set_error(PSC_EINVAL, "psc_particle_loads: sizes");
  if (imb_precision != 1 && imb_precision != 2) return set_error(PSC_EINVAL, "psc_particle_loads: imb precision");
  if (nparticles > 0 && (d_chunk_first == nullptr || d_loads == nullptr || d_partials == nullptr))
    return set_error(PSC_EINVAL, "psc_particle_loads: null output");
  if (nchunks > 0 && (d_particles == nullptr || d_boxes == nullptr || d_chunks == nullptr || d_ent_cell == nullptr ||
                      d_ent_loc == nullptr || d_ent_cnt == nullptr || d_chunk_nnz == nullptr ||
                      d_count_sum == nullptr || d_imb == nullptr))
    return set_error(PSC_EINVAL, "psc_particle_loads: null input");
  return launch_loads(d_particles, d_boxes, d_chunks, d_chunk_first, nparticles, nchunks, chunk_cells, d_ent_cell,
                      d_ent_loc, d_ent_cnt, d_chunk_nnz, d_count_sum, d_imb, imb_precision, nx, ny, nz, d_partials,
                      d_loads, (hipStream_t)stream);
}

int64_t psc_moments_blocks(int64_t ncell) { return moments_blocks(ncell); }

}  // extern "C"

namespace psc {
namespace {
// Argument checks shared by psc_moments and psc_moments_noalloc; resolves the
// row pitch and plane size.
int moments_args(const char* who, const void* d_state, int32_t storage, int32_t layout, const uint32_t* d_count_sum,
                 int32_t nx, int32_t ny, int32_t nz, int32_t pitch_z, int32_t subsamples, const double* d_rho_u,
                 const double* d_sums, int* pz, unsigned long long* ps) {
  if (!dims_ok(nx, ny, nz)) return set_error(PSC_EINVAL, "%s: bad lattice", who);
  if (storage != PSC_F32_DEV && storage != PSC_F64_PLAIN) return set_error(PSC_EINVAL, "%s: storage", who);
  if (layout != PSC_PUSH && layout != PSC_PULL) return set_error(PSC_EINVAL, "%s: layout", who);
  if (d_count_sum == nullptr || d_sums == nullptr) return set_error(PSC_EINVAL, "%s: null", who);
  if (d_state == nullptr && d_rho_u != nullptr) return set_error(PSC_EINVAL, "%s: fields need a state", who);
  if (subsamples < 0 || subsamples * subsamples * subsamples > 65535)
    return set_error(PSC_EINVAL, "%s: subsamples %d", who, subsamples);
  *pz = pitch_z > 0 ? pitch_z : nz;
  if (*pz < nz || !dims_ok(nx, ny, *pz)) return set_error(PSC_EINVAL, "%s: pitch_z %d < nz %d", who, pitch_z, nz);
  if (*pz > nz && d_rho_u != nullptr) return set_error(PSC_EINVAL, "%s: fields need unpadded rows", who);
  // explicit pitch: planes of nx*ny*pitch_z cells rounded up to 64 (batched-window layout)
  const unsigned long long n = (unsigned long long)nx * ny * *pz;
  *ps = pitch_z > 0 ? (n + 63ull) / 64ull * 64ull : n;
  return PSC_OK;
}
}  // namespace
}  // namespace psc

extern "C" {

int psc_moments(const void* d_state, int32_t storage, int32_t layout, const uint32_t* d_count_sum, const void* d_imb,
                int32_t nx, int32_t ny, int32_t nz, int32_t pitch_z, int32_t subsamples, double fx, double fy, double fz,
                double* d_rho_u, double* d_partials, double* d_sums, void* stream) {
  int pz = 0;
  unsigned long long ps = 0;
  int rc = moments_args("psc_moments", d_state, storage, layout, d_count_sum, nx, ny, nz, pitch_z, subsamples,
                        d_rho_u, d_sums, &pz, &ps);
  if (rc) return rc;
  if (d_partials == nullptr) return set_error(PSC_EINVAL, "psc_moments: null");
  return launch_moments(d_state, storage, layout, d_count_sum, d_imb, nx, ny, nz, pz, ps, subsamples, fx, fy, fz,
                        d_rho_u, d_partials, d_sums, (hipStream_t)stream);
}

int psc_moments_noalloc(const void* d_state, int32_t storage, int32_t layout, const uint32_t* d_count_sum,
                        const void* d_imb, int32_t nx, int32_t ny, int32_t nz, int32_t pitch_z, int32_t subsamples,
                        double fx, double fy, double fz, double* d_rho_u, double* d_sums, void* stream) {
  int pz = 0;
  unsigned long long ps = 0;
  int rc = moments_args("psc_moments_noalloc", d_state, storage, layout, d_count_sum, nx, ny, nz, pitch_z, subsamples,
                        d_rho_u, d_sums, &pz, &ps);
  if (rc) return rc;
  double* partials = nullptr;
  int64_t capacity = 0;
  rc = moments_scratch(&partials, &capacity);
  if (rc) return rc;
  const int64_t need = moments_blocks((int64_t)((unsigned long long)nx * ny * pz)) * PSC_NSUM;
  if (need > capacity)
    return set_error(PSC_EINVAL, "psc_moments_noalloc: %lld partial sums exceed the %lld of the scratch array",
                     (long long)need, (long long)capacity);
  return launch_moments(d_state, storage, layout, d_count_sum, d_imb, nx, ny, nz, pz, ps, subsamples, fx, fy, fz,
                        d_rho_u, partials, d_sums, (hipStream_t)stream);
}

int psc_export_f64(const void* d_state, int32_t storage, int32_t layout, int32_t nx, int32_t ny, int32_t nz,
                   double* d_f, void* stream) {
  if (!dims_ok(nx, ny, nz) || d_state == nullptr || d_f == nullptr)
    return set_error(PSC_EINVAL, "psc_export_f64: bad arguments");
  if ((storage != PSC_F32_DEV && storage != PSC_F64_PLAIN) || (layout != PSC_PUSH && layout != PSC_PULL))
    return set_error(PSC_EINVAL, "psc_export_f64: storage/layout");
  if ((const void*)d_f == d_state) return set_error(PSC_EINVAL, "psc_export_f64: aliasing");
  return launch_export(d_state, storage, layout, nx, ny, nz, d_f, (hipStream_t)stream);
}

int psc_import_f64(const double* d_f, int32_t storage, int32_t layout, int32_t nx, int32_t ny, int32_t nz,
                   void* d_state, void* stream) {
  if (!dims_ok(nx, ny, nz) || d_state == nullptr || d_f == nullptr)
    return set_error(PSC_EINVAL, "psc_import_f64: bad arguments");
  if ((storage != PSC_F32_DEV && storage != PSC_F64_PLAIN) || (layout != PSC_PUSH && layout != PSC_PULL))
    return set_error(PSC_EINVAL, "psc_import_f64: storage/layout");
  if ((const void*)d_f == d_state) return set_error(PSC_EINVAL, "psc_import_f64: aliasing");
  return launch_import(d_f, storage, layout, nx, ny, nz, d_state, (hipStream_t)stream);
}

int psc_copy_f32(const float* d_src, float* d_dst, int64_t n, void* stream) {
  if (n < 0 || (n > 0 && (d_src == nullptr || d_dst == nullptr))) return set_error(PSC_EINVAL, "psc_copy_f32: args");
  if (n == 0) return PSC_OK;
  hipStream_t s = (hipStream_t)stream;
  if ((((uintptr_t)d_src | (uintptr_t)d_dst) & 15u) == 0 && (n % 4) == 0)
    k_copy4<<<grid_cap((size_t)n / 16, 256), 256, 0, s>>>((const f32x4*)d_src, (f32x4*)d_dst, (size_t)n / 4);
  else
    k_copy1<<<grid_cap((size_t)n, 256), 256, 0, s>>>(d_src, d_dst, (size_t)n);
  return check_launch("psc_copy_f32");
}

int psc_voxelize_binary(const double* d_spheres, const int32_t* d_boxes, int32_t nspheres, const double* d_axes,
                        int32_t nx, int32_t ny, int32_t nz, uint8_t* d_solid, void* stream) {
  if (!dims_ok(nx, ny, nz)) return set_error(PSC_EINVAL, "psc_voxelize_binary: bad lattice");
  if (nspheres < 0) return set_error(PSC_EINVAL, "psc_voxelize_binary: sphere count %d", nspheres);
  if (d_solid == nullptr || d_axes == nullptr || (nspheres > 0 && (d_spheres == nullptr || d_boxes == nullptr)))
    return set_error(PSC_EINVAL, "psc_voxelize_binary: null");
  return launch_voxelize_binary(d_spheres, d_boxes, nspheres, d_axes, nx, ny, nz, d_solid, (hipStream_t)stream);
}

int psc_bb_step(const void* d_in, void* d_out, const uint8_t* d_solid, int32_t storage, int32_t nx, int32_t ny,
                int32_t nz, double tau, double fx, double fy, double fz, int32_t block_z, int32_t block_y,
                int32_t flags, void* stream) {
  if (!dims_ok(nx, ny, nz)) return set_error(PSC_EINVAL, "psc_bb_step: bad lattice %dx%dx%d", nx, ny, nz);
  if (d_in == nullptr || d_out == nullptr || d_solid == nullptr) return set_error(PSC_EINVAL, "psc_bb_step: null");
  if (d_in == d_out) return set_error(PSC_EINVAL, "psc_bb_step: input and output alias");
  if (!(tau > 0.5)) return set_error(PSC_EINVAL, "psc_bb_step: tau %g must exceed 1/2", tau);
  if ((flags & ~(PSC_STEP_NONTEMPORAL | PSC_STEP_BPC_MASK)) != 0)
    return set_error(PSC_EINVAL, "psc_bb_step: unknown flags 0x%x", flags);
  int bz = block_z, by = block_y;
  if (bz <= 0) bz = (nz >= 64 && !(nz % 64 != 0 && nz % 32 == 0)) ? 64 : 32;
  if (by <= 0) by = 256 / bz > 0 ? 256 / bz : 1;
  if (bz * by > 256 || bz * by < 1 || ((bz * by) % 64) != 0)
    return set_error(PSC_ELAUNCH, "psc_bb_step: block %dx%d must hold a multiple of 64 threads, at most 256", bz, by);
  hipStream_t s = (hipStream_t)stream;
  if (storage == PSC_F32_DEV)
    return launch_bb_step_f32((const float*)d_in, (float*)d_out, d_solid, nx, ny, nz, tau, fx, fy, fz, bz, by, flags,
                              s);
  if (storage == PSC_F64_PLAIN)
    return launch_bb_step_f64((const double*)d_in, (double*)d_out, d_solid, nx, ny, nz, tau, fx, fy, fz, bz, by, flags,
                              s);
  return set_error(PSC_EINVAL, "psc_bb_step: storage %d", storage);
}

int psc_smooth_alpha(const double* d_particles, const int32_t* d_boxes, const int32_t* d_chunks, int32_t nchunks,
                     int32_t chunk_cells, int32_t nx, int32_t ny, int32_t nz, uint32_t* d_nfull, int64_t* d_logsum,
                     double* d_alpha64, float* d_alpha32, void* stream) {
  if (!dims_ok(nx, ny, nz)) return set_error(PSC_EINVAL, "psc_smooth_alpha: bad lattice");
  if (nchunks < 0 || chunk_cells < 1) return set_error(PSC_EINVAL, "psc_smooth_alpha: sizes");
  if (d_nfull == nullptr || d_logsum == nullptr) return set_error(PSC_EINVAL, "psc_smooth_alpha: null scratch");
  if (nchunks > 0 && (d_particles == nullptr || d_boxes == nullptr || d_chunks == nullptr))
    return set_error(PSC_EINVAL, "psc_smooth_alpha: null geometry");
  return launch_smooth_alpha(d_particles, d_boxes, d_chunks, nchunks, chunk_cells, nx, ny, nz, d_nfull,
                             (unsigned long long*)d_logsum, d_alpha64, d_alpha32, (hipStream_t)stream);
}

int psc_smooth_alpha_grad(const double* d_particles, const int32_t* d_boxes, const int32_t* d_chunks,
                          const int32_t* d_chunk_first, int32_t nchunks, int32_t chunk_cells, int32_t nparticles,
                          int32_t nx, int32_t ny, int32_t nz, const uint32_t* d_nfull, const int64_t* d_logsum,
                          const double* d_galpha, double* d_partials, double* d_grad, void* stream) {
  if (!dims_ok(nx, ny, nz)) return set_error(PSC_EINVAL, "psc_smooth_alpha_grad: bad lattice");
  if (nchunks < 0 || chunk_cells < 1 || nparticles < 0)
    return set_error(PSC_EINVAL, "psc_smooth_alpha_grad: sizes");
  if (nparticles > 0 && (d_grad == nullptr || d_chunk_first == nullptr))
    return set_error(PSC_EINVAL, "psc_smooth_alpha_grad: null output or chunk_first");
  if (nchunks > 0 && (d_particles == nullptr || d_boxes == nullptr || d_chunks == nullptr || d_nfull == nullptr ||
                      d_logsum == nullptr || d_galpha == nullptr || d_partials == nullptr))
    return set_error(PSC_EINVAL, "psc_smooth_alpha_grad: null input");
  return launch_smooth_alpha_grad(d_particles, d_boxes, d_chunks, d_chunk_first, nchunks, chunk_cells, nparticles,
                                  nx, ny, nz, d_nfull, (const unsigned long long*)d_logsum, d_galpha, d_partials,
                                  d_grad, (hipStream_t)stream);
}

int psc_adjoint_seed(const void* d_state, const void* d_alpha, void* d_lam, double* d_galpha, int32_t storage,
                     int32_t nx, int32_t ny, int32_t nz, double tau, double fx, double fy, double fz, double weight,
                     void* stream) {
  if (!dims_ok(nx, ny, nz)) return set_error(PSC_EINVAL, "psc_adjoint_seed: bad lattice");
  if (storage != PSC_F32_DEV && storage != PSC_F64_PLAIN) return set_error(PSC_EINVAL, "psc_adjoint_seed: storage");
  if (d_state == nullptr || d_alpha == nullptr || d_lam == nullptr || d_galpha == nullptr)
    return set_error(PSC_EINVAL, "psc_adjoint_seed: null");
  if (d_state == d_lam) return set_error(PSC_EINVAL, "psc_adjoint_seed: state and adjoint alias");
  if (!(tau > 0.5)) return set_error(PSC_EINVAL, "psc_adjoint_seed: tau %g must exceed 1/2", tau);
  if (!(fx * fx + fy * fy + fz * fz > 0.0))
    return set_error(PSC_EINVAL, "psc_adjoint_seed: the permeability needs a non-zero body force");
  return launch_adjoint_seed(d_state, d_alpha, d_lam, d_galpha, storage, nx, ny, nz, tau, fx, fy, fz, weight,
                             (hipStream_t)stream);
}

int psc_adjoint_step(const void* d_state, const void* d_lam_in, void* d_lam_out, const void* d_alpha,
                     double* d_galpha, int32_t storage, int32_t nx, int32_t ny, int32_t nz, double tau, double fx,
                     double fy, double fz, int32_t block_z, int32_t block_y, int32_t flags, void* stream) {
  if (!dims_ok(nx, ny, nz)) return set_error(PSC_EINVAL, "psc_adjoint_step: bad lattice");
  if (storage != PSC_F32_DEV && storage != PSC_F64_PLAIN) return set_error(PSC_EINVAL, "psc_adjoint_step: storage");
  if (d_state == nullptr || d_lam_in == nullptr || d_lam_out == nullptr || d_alpha == nullptr || d_galpha == nullptr)
    return set_error(PSC_EINVAL, "psc_adjoint_step: null");
  if (d_lam_in == d_lam_out || d_state == d_lam_out)
    return set_error(PSC_EINVAL, "psc_adjoint_step: the output adjoint aliases an input");
  if (!(tau > 0.5)) return set_error(PSC_EINVAL, "psc_adjoint_step: tau %g must exceed 1/2", tau);
  if ((flags & ~(PSC_STEP_NONTEMPORAL | PSC_STEP_BPC_MASK)) != 0)
    return set_error(PSC_EINVAL, "psc_adjoint_step: unknown flags 0x%x", flags);
  if ((block_z > 0) != (block_y > 0) ||
      (block_z > 0 && (block_z * block_y > 256 || (block_z * block_y) % 64 != 0)))
    return set_error(PSC_ELAUNCH, "psc_adjoint_step: block %dx%d must hold a multiple of 64 threads, at most 256",
                     block_z, block_y);
  return launch_adjoint_step(d_state, d_lam_in, d_lam_out, d_alpha, d_galpha, storage, nx, ny, nz, tau, fx, fy, fz,
                             block_z, block_y, flags, (hipStream_t)stream);
}

}  // extern "C"

// ------------------------------------------------------------------------
// psc_run_steps: native step loop with hipGraph replay
// ------------------------------------------------------------------------
namespace psc {
namespace {

struct GraphEntry {
  PscStepPlan plan;
  int steps;
  int start_parity;
  int kind;  // one of SegmentKind
  hipGraphExec_t exec;
};
std::mutex g_graph_mu;
std::vector<GraphEntry> g_graphs;
constexpr size_t kMaxGraphs = 64;

// Copy of a plan with every padding byte zero, so plans compare with memcmp.
PscStepPlan normalized(const PscStepPlan& p) {
  PscStepPlan k;
  std::memset(&k, 0, sizeof(k));
  k.state_a = p.state_a;
  k.state_b = p.state_b;
  k.count_sum = p.count_sum;
  k.us_num = p.us_num;
  k.imb = p.imb;
  k.imb_alt = p.imb_alt;
  k.storage = p.storage;
  k.layout = p.layout;
  k.nx = p.nx;
  k.ny = p.ny;
  k.nz = p.nz;
  k.subsamples = p.subsamples;
  k.tau = p.tau;
  k.fx = p.fx;
  k.fy = p.fy;
  k.fz = p.fz;
  k.block_z = p.block_z;
  k.block_y = p.block_y;
  k.flags = p.flags;
  k.particles = p.particles;
  k.boxes = p.boxes;
  k.chunks = p.chunks;
  k.chunk_first = p.chunk_first;
  k.nparticles = p.nparticles;
  k.nchunks = p.nchunks;
  k.chunk_cells = p.chunk_cells;
  k.ent_cell = p.ent_cell;
  k.ent_loc = p.ent_loc;
  k.ent_cnt = p.ent_cnt;
  k.chunk_nnz = p.chunk_nnz;
  k.partials = p.partials;
  k.partials_alt = p.partials_alt;
  k.loads = p.loads;
  k.fuse_loads = p.fuse_loads;
  return k;
}

int one_step(const PscStepPlan& p, int parity, hipStream_t s) {
  void* in = parity ? p.state_b : p.state_a;
  void* out = parity ? p.state_a : p.state_b;
  // with a second imb buffer, the input parity selects it (as in the pipeline)
  void* imb = (parity && p.imb_alt != nullptr) ? p.imb_alt : p.imb;
  double* partials = (parity && p.partials_alt != nullptr) ? p.partials_alt : p.partials;
  int rc = psc_step(in, out, p.count_sum, p.us_num, imb, p.storage, p.layout, p.nx, p.ny, p.nz, p.subsamples,
                    p.tau, p.fx, p.fy, p.fz, p.block_z, p.block_y, p.flags, s);
  if (rc) return rc;
  if (p.nparticles > 0) {
    rc = psc_particle_loads(p.particles, p.boxes, p.chunks, p.chunk_first, p.nparticles, p.nchunks, p.chunk_cells,
                            p.ent_cell, p.ent_loc, p.ent_cnt, p.chunk_nnz, p.count_sum, imb,
                            p.storage == PSC_F32_DEV ? 1 : 2, p.nx, p.ny, p.nz, partials, p.loads, s);
  }
  return rc;
}

// Library-owned stream per device for graph capture: the caller's stream may
// be the legacy default stream, which cannot be captured.  Captured work does
// not execute, so capturing on a side stream orders nothing; the graph is
// launched on the caller's stream.
std::vector<hipStream_t> g_capture_streams;

int capture_stream(hipStream_t* out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return set_error((int)e, "psc_run_steps: hipGetDevice: %s", hipGetErrorString(e));
  if ((int)g_capture_streams.size() <= dev) g_capture_streams.resize(dev + 1, nullptr);
  if (g_capture_streams[dev] == nullptr) {
    e = hipStreamCreateWithFlags(&g_capture_streams[dev], hipStreamNonBlocking);
    if (e != hipSuccess) return set_error((int)e, "psc_run_steps: capture stream: %s", hipGetErrorString(e));
  }
  *out = g_capture_streams[dev];
  return PSC_OK;
}

// Side stream and events of the overlapped loop (per device, library-owned,
// default priority).
struct SideRes {
  hipStream_t side = nullptr;
  hipEvent_t ev_step = nullptr;
  hipEvent_t ev_loads[2] = {nullptr, nullptr};
};
std::vector<SideRes> g_side;

int side_resources(SideRes** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return set_error((int)e, "psc_run_steps: hipGetDevice: %s", hipGetErrorString(e));
  if ((int)g_side.size() <= dev) g_side.resize(dev + 1);
  SideRes& r = g_side[dev];
  if (r.side == nullptr) {
    e = hipStreamCreateWithFlags(&r.side, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&r.ev_step, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&r.ev_loads[0], hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&r.ev_loads[1], hipEventDisableTiming);
    if (e != hipSuccess) return set_error((int)e, "psc_run_steps: side stream: %s", hipGetErrorString(e));
  }
  *out = &r;
  return PSC_OK;
}

#define PSC_HIP_TRY(x)                                                                          \
  do {                                                                                          \
    const hipError_t e_ = (x);                                                                  \
    if (e_ != hipSuccess) return set_error((int)e_, "psc_run_steps: %s", hipGetErrorString(e_)); \
  } while (0)

// n steps from parity `par`.  With imb_alt and partials_alt set and
// particles present, the particle-load reduction of step k runs on a side
// stream while step k+1 streams the lattice on `s` (step k+2 waits for it,
// since it rewrites the imb buffer step k's loads read); the side stream is
// joined back into `s` at the end, so on return every step's loads have been
// computed and d_loads holds the last step's.  Otherwise steps and loads run
// in order on `s`.
// Fused form (fuse_loads > 0, FP32 pull storage, no solid velocity, particles,
// imb_alt and partials_alt set): a step with input parity q writes imb[q]
// and, in the same launch, sums the chunks of the previous step's imb[q^1]
// into partials[q^1] and adds up partials[q] (the step before that) into
// d_loads, so a step is one launch.  *pc / *pf say whether a previous step's
// chunks / partials are pending; a sequence starts with neither.
// With PSC_RUN_INPLACE in the plan's flags, steps go in pairs that update the
// state buffer of parity *par in place (a gather-scatter step, then a local
// step; see step_cell): the imb and partials buffers still alternate every
// step, and after a pair both parities agree again.  An odd count starts with
// an ordinary pull step from one buffer to the other, so the pairs (and the
// runs that follow) stay on the buffer the run ends in, which the Infinity
// Cache then already holds.
int run_fused(const PscStepPlan& p, int n, int* par, bool* pc, bool* pf, hipStream_t s) {
  const LoadsArgs la{p.particles, p.boxes, p.chunks, p.chunk_first, p.nparticles, p.nchunks, p.chunk_cells,
                     p.ent_cell, p.ent_loc, p.ent_cnt, p.chunk_nnz, p.count_sum};
  void* imbs[2] = {p.imb, p.imb_alt};
  double* prt[2] = {p.partials, p.partials_alt};
  int bz = p.block_z, by = p.block_y;
  if (bz <= 0) bz = (p.nz >= 64 && !(p.nz % 64 != 0 && p.nz % 32 == 0)) ? 64 : 32;
  if (by <= 0) by = 256 / bz > 0 ? 256 / bz : 1;
  const bool inplace = (p.flags & PSC_RUN_INPLACE) != 0;
  const int flags = p.flags & ~PSC_RUN_INPLACE;
  for (int k = 0; k < n; ++k) {
    const int q = *par;
    // in-place pairs: steps j and j+1 for even j = k - (n odd); the odd
    // count's first step (j = -1) is an ordinary one
    const int j = k - (n & 1);
    const int aa = (!inplace || j < 0) ? 0 : ((j & 1) ? 2 : 1);
    const int sb = aa == 2 ? q ^ 1 : q;  // the pair's state buffer (the parity at its start)
    float* st_in = (float*)(sb ? p.state_b : p.state_a);
    float* st_out = aa ? st_in : (float*)(sb ? p.state_a : p.state_b);
    const int rc = launch_step_f32_loads(st_in, st_out, p.count_sum, (float*)imbs[q],
                                         *pc ? (const float*)imbs[q ^ 1] : nullptr, la, prt[q ^ 1],
                                         *pf ? prt[q] : nullptr, p.loads, p.fuse_loads, p.nx, p.ny, p.nz,
                                         p.subsamples, p.tau, p.fx, p.fy, p.fz, bz, by, flags, s, aa);
    if (rc) return rc;
    *pf = *pc;
    *pc = true;
    *par ^= 1;
  }
  return PSC_OK;
}

// The pending loads after run_fused, oldest first, so d_loads ends with the
// last step's (the final pass also writes zeros for particles without chunks).
int flush_fused(const PscStepPlan& p, int par, bool pc, bool pf, hipStream_t s) {
  const LoadsArgs la{p.particles, p.boxes, p.chunks, p.chunk_first, p.nparticles, p.nchunks, p.chunk_cells,
                     p.ent_cell, p.ent_loc, p.ent_cnt, p.chunk_nnz, p.count_sum};
  void* imbs[2] = {p.imb, p.imb_alt};
  double* prt[2] = {p.partials, p.partials_alt};
  int rc = PSC_OK;
  if (pf && pc) {
    // the previous step's per-particle pass and the last step's chunk pass
    // share a launch, then the last step's per-particle pass
    rc = launch_loads_flush(la, imbs[par ^ 1], 1, p.nx, p.ny, p.nz, prt[par ^ 1], prt[par], p.loads, s);
    return rc ? rc : launch_loads_final(p.chunk_first, p.nparticles, prt[par ^ 1], p.loads, s);
  }
  if (pf) rc = launch_loads_final(p.chunk_first, p.nparticles, prt[par], p.loads, s);
  if (rc == PSC_OK && pc) {
    rc = launch_loads_chunks(la, imbs[par ^ 1], 1, p.nx, p.ny, p.nz, prt[par ^ 1], s);
    if (rc == PSC_OK) rc = launch_loads_final(p.chunk_first, p.nparticles, prt[par ^ 1], p.loads, s);
  }
  return rc;
}

bool fused_plan(const PscStepPlan& p) {
  return p.fuse_loads > 0 && p.nparticles > 0 && p.storage == PSC_F32_DEV && p.layout == PSC_PULL &&
         p.us_num == nullptr && p.imb != nullptr && p.imb_alt != nullptr && p.partials != nullptr &&
         p.partials_alt != nullptr;
}

int run_sequence(const PscStepPlan& p, int n, int* par, hipStream_t s) {
  if (fused_plan(p)) {
    bool pc = false, pf = false;
    const int rc = run_fused(p, n, par, &pc, &pf, s);
    return rc ? rc : flush_fused(p, *par, pc, pf, s);
  }
  const bool overlap = p.nparticles > 0 && p.imb != nullptr && p.imb_alt != nullptr && p.partials != nullptr &&
                       p.partials_alt != nullptr;
  if (!overlap) {
    for (int k = 0; k < n; ++k) {
      const int rc = one_step(p, *par, s);
      if (rc) return rc;
      *par ^= 1;
    }
    return PSC_OK;
  }
  if (n == 0) return PSC_OK;
  SideRes* r = nullptr;
  int rc = side_resources(&r);
  if (rc) return rc;
  const LoadsArgs la{p.particles, p.boxes, p.chunks, p.chunk_first, p.nparticles, p.nchunks, p.chunk_cells,
                     p.ent_cell, p.ent_loc, p.ent_cnt, p.chunk_nnz, p.count_sum};
  void* imbs[2] = {p.imb, p.imb_alt};
  double* prt[2] = {p.partials, p.partials_alt};
  const int imbp = p.storage == PSC_F32_DEV ? 1 : 2;
  // the side stream starts after the work already queued on s
  PSC_HIP_TRY(hipEventRecord(r->ev_step, s));
  PSC_HIP_TRY(hipStreamWaitEvent(r->side, r->ev_step, 0));
  for (int k = 0; k < n; ++k) {
    const int pk = *par;
    if (k >= 2) PSC_HIP_TRY(hipStreamWaitEvent(s, r->ev_loads[k & 1], 0));
    rc = psc_step(pk ? p.state_b : p.state_a, pk ? p.state_a : p.state_b, p.count_sum, p.us_num, imbs[pk],
                  p.storage, p.layout, p.nx, p.ny, p.nz, p.subsamples, p.tau, p.fx, p.fy, p.fz, p.block_z,
                  p.block_y, p.flags, s);
    if (rc) return rc;
    PSC_HIP_TRY(hipEventRecord(r->ev_step, s));
    PSC_HIP_TRY(hipStreamWaitEvent(r->side, r->ev_step, 0));
    rc = launch_loads_chunks(la, imbs[pk], imbp, p.nx, p.ny, p.nz, prt[pk], r->side);
    if (rc) return rc;
    rc = launch_loads_final(p.chunk_first, p.nparticles, prt[pk], p.loads, r->side);
    if (rc) return rc;
    PSC_HIP_TRY(hipEventRecord(r->ev_loads[k & 1], r->side));
    *par ^= 1;
  }
  PSC_HIP_TRY(hipStreamWaitEvent(s, r->ev_loads[(n - 1) & 1], 0));
  return PSC_OK;
}

// The pieces a run is made of.  The fused kinds differ in whether the loads
// of earlier steps are pending at the start (a fresh start has none) and in
// whether the pending loads are reduced at the end (flush).
enum SegmentKind {
  kSeq = 0,          // run_sequence: unfused steps (loads in line or on the side stream)
  kFusedFresh = 1,   // fused steps from a fresh start, no flush
  kFusedCont = 2,    // fused steps continuing (both loads stages pending), no flush
  kFusedTail = 3,    // fused steps continuing, then the pending loads
  kFusedWhole = 4,   // fused steps from a fresh start, then the pending loads
};

// Run one segment on stream s (also what a captured graph of it holds).
int exec_segment(const PscStepPlan& p, int kind, int steps, int* par, hipStream_t s) {
  if (kind == kSeq) return run_sequence(p, steps, par, s);
  const bool cont = kind == kFusedCont || kind == kFusedTail;
  bool pc = cont, pf = cont;
  const int rc = run_fused(p, steps, par, &pc, &pf, s);
  if (rc || kind == kFusedFresh || kind == kFusedCont) return rc;
  return flush_fused(p, *par, pc, pf, s);
}

int graph_for(const PscStepPlan& p, int steps, int parity, int kind, hipGraphExec_t* out) {
  std::lock_guard<std::mutex> lock(g_graph_mu);
  for (const GraphEntry& e : g_graphs) {
    if (e.steps == steps && e.start_parity == parity && e.kind == kind &&
        std::memcmp(&e.plan, &p, sizeof(p)) == 0) {
      *out = e.exec;
      return PSC_OK;
    }
  }
  hipStream_t s = nullptr;
  const int rcs = capture_stream(&s);
  if (rcs) return rcs;
  hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) return set_error((int)e, "psc_run_steps: begin capture: %s", hipGetErrorString(e));
  int par = parity;
  const int rc = exec_segment(p, kind, steps, &par, s);
  hipGraph_t graph = nullptr;
  e = hipStreamEndCapture(s, &graph);
  if (rc) {
    if (graph) (void)hipGraphDestroy(graph);
    return rc;
  }
  if (e != hipSuccess) return set_error((int)e, "psc_run_steps: end capture: %s", hipGetErrorString(e));
  hipGraphExec_t exec = nullptr;
  e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (e != hipSuccess) return set_error((int)e, "psc_run_steps: instantiate: %s", hipGetErrorString(e));
  // Upload the executable graph now (on the library's capture stream, waited
  // for here), so that its first launch -- often the first graph of a timed
  // run, right after two direct steps -- does not carry the upload: without
  // it the GPU idled ~17 us between those steps and the graph's first kernel.
  e = hipGraphUpload(exec, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) {
    (void)hipGraphExecDestroy(exec);
    return set_error((int)e, "psc_run_steps: graph upload: %s", hipGetErrorString(e));
  }
  if (g_graphs.size() >= kMaxGraphs) {
    (void)hipGraphExecDestroy(g_graphs.front().exec);
    g_graphs.erase(g_graphs.begin());
  }
  g_graphs.push_back(GraphEntry{p, steps, parity, kind, exec});
  *out = exec;
  return PSC_OK;
}

}  // namespace
}  // namespace psc

namespace psc {
namespace {

// A run as segments: `times` executions of `steps` steps of one kind, as
// replays of a cached graph or as direct launches.  With graphs, a fused run
// starts with two direct steps (the GPU starts on them at once, while the
// host submits the first graph), then replays a continuing graph of
// graph_steps steps as often as it fits and ends with one graph of the
// remaining 1 .. graph_steps steps that also reduces the pending loads, so
// no graph is followed by a direct launch.  Consecutive graph launches leave
// the GPU idle for a few microseconds between them, so a run has as few as
// its length allows.  An unfused run is replays of graph_steps steps, then
// one graph of the rest.
struct Segment {
  int kind, steps, times;
  bool direct;
};

std::vector<Segment> run_segments(const PscStepPlan& key, int nsteps, bool use_graph, int graph_steps) {
  std::vector<Segment> out;
  if (nsteps <= 0) return out;
  const bool fused = fused_plan(key);
  // fused: two direct steps, three for an odd count (its ordinary step
  // first), so that every graph holds an even number of steps
  const int head = 2 + (nsteps & 1);
  if (!use_graph || (fused && nsteps <= head)) {
    out.push_back({fused ? kFusedWhole : kSeq, nsteps, 1, true});
    return out;
  }
  const int G = graph_steps;
  if (fused) {
    out.push_back({kFusedFresh, head, 1, true});
    int reps = (nsteps - head) / G, tail = (nsteps - head) % G;
    if (tail == 0) {
      --reps;
      tail = G;
    }
    if (reps > 0) out.push_back({kFusedCont, G, reps, false});
    out.push_back({kFusedTail, tail, 1, false});
    return out;
  }
  if (nsteps / G > 0) out.push_back({kSeq, G, nsteps / G, false});
  if (nsteps % G) out.push_back({kSeq, nsteps % G, 1, false});
  return out;
}

int check_run_plan(const PscStepPlan& key) {
  if ((key.flags & PSC_RUN_INPLACE) && !fused_plan(key))
    return set_error(PSC_EINVAL, "psc_run_steps: in-place steps need the fused form (FP32 pull, no solid velocity, "
                     "fuse_loads > 0, particles, alternate imb and partials)");
  if (fused_plan(key)) {
    int bz = key.block_z, by = key.block_y;
    int rc = step_args(key.state_a, key.state_b, key.count_sum, key.us_num, key.storage, key.layout, key.nx, key.ny,
                       key.nz, key.subsamples, key.tau, key.flags & ~PSC_RUN_INPLACE, &bz, &by);
    if (rc) return rc;
    if ((key.flags & PSC_RUN_INPLACE) && (key.flags & PSC_STEP_CPT_MASK) > (1 << 16))
      return set_error(PSC_EINVAL, "psc_run_steps: in-place steps take one cell per thread");
    if (key.imb == key.imb_alt || key.partials == key.partials_alt)
      return set_error(PSC_EINVAL, "psc_run_steps: imb/partials buffers alias");
  }
  return PSC_OK;
}

int check_run_args(const PscStepPlan* plan, int32_t nsteps, int32_t use_graph, int32_t graph_steps,
                   const int32_t* parity) {
  if (plan == nullptr || parity == nullptr || nsteps < 0) return set_error(PSC_EINVAL, "psc_run_steps: arguments");
  if (*parity != 0 && *parity != 1) return set_error(PSC_EINVAL, "psc_run_steps: parity %d", *parity);
  if (use_graph && (graph_steps < 2 || (graph_steps & 1)))
    return set_error(PSC_EINVAL, "psc_run_steps: graph_steps %d must be even and >= 2", graph_steps);
  return PSC_OK;
}

int replay(const PscStepPlan& key, int steps, int parity, int kind, int times, hipStream_t s) {
  hipGraphExec_t exec = nullptr;
  const int rc = graph_for(key, steps, parity, kind, &exec);
  if (rc) return rc;
  for (int t = 0; t < times; ++t) {
    const hipError_t e = hipGraphLaunch(exec, s);
    if (e != hipSuccess) return set_error((int)e, "psc_run_steps: graph launch: %s", hipGetErrorString(e));
  }
  return PSC_OK;
}

}  // namespace
}  // namespace psc

extern "C" int psc_run_steps(const PscStepPlan* plan, int32_t nsteps, int32_t use_graph, int32_t graph_steps,
                             int32_t* parity, void* stream) {
  int rc = check_run_args(plan, nsteps, use_graph, graph_steps, parity);
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  const PscStepPlan key = normalized(*plan);
  rc = check_run_plan(key);
  if (rc) return rc;
  int par = *parity;
  for (const Segment& g : run_segments(key, nsteps, use_graph != 0, graph_steps)) {
    if (g.direct) {
      rc = exec_segment(key, g.kind, g.steps, &par, s);
    } else {
      rc = replay(key, g.steps, par, g.kind, g.times, s);
      if (g.steps & g.times & 1) par ^= 1;
    }
    if (rc) return rc;
  }
  *parity = par;
  return PSC_OK;
}

extern "C" int psc_prepare_steps(const PscStepPlan* plan, int32_t nsteps, int32_t graph_steps,
                                 const int32_t* parity) {
  int rc = check_run_args(plan, nsteps, 1, graph_steps, parity);
  if (rc) return rc;
  const PscStepPlan key = normalized(*plan);
  rc = check_run_plan(key);
  if (rc) return rc;
  int par = *parity;
  for (const Segment& g : run_segments(key, nsteps, true, graph_steps)) {
    if (!g.direct) {
      hipGraphExec_t exec = nullptr;
      rc = graph_for(key, g.steps, par, g.kind, &exec);
      if (rc) return rc;
    }
    if (g.steps & g.times & 1) par ^= 1;
  }
  return PSC_OK;
}

namespace psc {
namespace {

int check_deferred(const char* who, const PscStepPlan* plan, int32_t parity, const int32_t* pending,
                   PscStepPlan* key) {
  if (plan == nullptr || pending == nullptr) return set_error(PSC_EINVAL, "%s: arguments", who);
  if (parity != 0 && parity != 1) return set_error(PSC_EINVAL, "%s: parity %d", who, parity);
  if (*pending & ~(PSC_PENDING_CHUNKS | PSC_PENDING_FINAL))
    return set_error(PSC_EINVAL, "%s: pending %d", who, *pending);
  *key = normalized(*plan);
  if (!fused_plan(*key))
    return set_error(PSC_EINVAL, "%s: needs a fused plan (FP32 pull, no solid velocity, fuse_loads > 0, particles, "
                     "alternate imb and partials)", who);
  return check_run_plan(*key);
}

}  // namespace
}  // namespace psc

extern "C" int psc_run_steps_deferred(const PscStepPlan* plan, int32_t nsteps, int32_t* parity, int32_t* pending,
                                      void* stream) {
  if (parity == nullptr || nsteps < 0) return set_error(PSC_EINVAL, "psc_run_steps_deferred: arguments");
  PscStepPlan key;
  int rc = check_deferred("psc_run_steps_deferred", plan, *parity, pending, &key);
  if (rc) return rc;
  int par = *parity;
  bool pc = (*pending & PSC_PENDING_CHUNKS) != 0, pf = (*pending & PSC_PENDING_FINAL) != 0;
  rc = run_fused(key, nsteps, &par, &pc, &pf, (hipStream_t)stream);
  if (rc) return rc;
  *parity = par;
  *pending = (pc ? PSC_PENDING_CHUNKS : 0) | (pf ? PSC_PENDING_FINAL : 0);
  return PSC_OK;
}

extern "C" int psc_flush_loads(const PscStepPlan* plan, int32_t parity, int32_t* pending, void* stream) {
  PscStepPlan key;
  int rc = check_deferred("psc_flush_loads", plan, parity, pending, &key);
  if (rc) return rc;
  rc = flush_fused(key, parity, (*pending & PSC_PENDING_CHUNKS) != 0, (*pending & PSC_PENDING_FINAL) != 0,
                   (hipStream_t)stream);
  if (rc) return rc;
  *pending = 0;
  return PSC_OK;
}

extern "C" int psc_graph_cache_clear(void) {
  std::lock_guard<std::mutex> lock(g_graph_mu);
  for (GraphEntry& e : g_graphs) (void)hipGraphExecDestroy(e.exec);
  g_graphs.clear();
  return PSC_OK;
}
