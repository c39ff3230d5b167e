/*
 * psc_hip.h -- C ABI of the MI355X (gfx950) D3Q19 partially-saturated-cell
 * immersed-boundary (PSC-IMB) lattice-Boltzmann library, libpsc_hip.so.
 *
 * The reference (wangjianfttt/particuology_Hydraulic-penalty-) exposes this
 * path as the Python class API of src/lbm_dem/psc3d.py; it has no foreign
 * function interface of its own.  Each entry point below replaces one piece
 * of that module (cited per function) and is what the Python facade
 * (package module psc3d.py) binds through ctypes.
 *
 * Conventions
 *   - Every function returns int: 0 on success, a negative PSC_E* code on a
 *     bad argument, or a positive hipError_t from the HIP runtime.
 *     psc_last_error() gives the message of the last failure on this thread.
 *   - Pointers named d_* are device addresses (hipMalloc / torch tensors);
 *     nothing here allocates, frees or synchronises, so every launch is
 *     asynchronous on `stream` (a hipStream_t, NULL = default stream).
 *   - Lattice fields are structure-of-arrays: direction-major planes of
 *     N = nx*ny*nz cells, cell id = (ix*ny + iy)*nz + iz (iz fastest), the
 *     C order of the reference's (19, n, n, n) arrays.  The one exception is
 *     imb, whose 3 components of a cell are consecutive (imb[3*cell + axis]):
 *     it is written and read one cell at a time.
 *   - Distribution storage: PSC_F32_DEV stores g_i = f_i - w_i in FP32,
 *     PSC_F64_PLAIN stores f_i in FP64 with the reference's exact operation
 *     order.  Stream layout PSC_PUSH keeps the post-stream state (the
 *     reference's `f`); PSC_PULL keeps the post-collision state and gathers
 *     on the next step.
 *   - Particle table: P rows of 10 doubles (x, y, z, r, vx, vy, vz, wx, wy,
 *     wz), lattice units, the fields of Particle3D (psc3d.py:46-57).
 *   - Box table: P rows of 8 int32 (x0, y0, z0, ex, ey, ez, entry_offset, 0):
 *     the cell bounding box [x0, x0+ex) ... of psc3d.py:73-78 and the offset
 *     of the particle's first per-cell count in the count array.
 *   - Chunk table: rows of 2 int32 (particle, first box entry); each chunk
 *     covers up to `chunk_cells` consecutive box entries (ix, iy, iz
 *     lexicographic) of one particle.
 */
#ifndef PSC_HIP_H
#define PSC_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PSC_OK 0
#define PSC_EINVAL (-1) /* bad argument (null pointer, size, range)          */
#define PSC_ELAUNCH (-2) /* launch configuration outside the kernel's limits  */

#define PSC_F32_DEV 1   /* FP32 deviation storage g = f - w                  */
#define PSC_F64_PLAIN 2 /* FP64 plain storage f, reference operation order   */

#define PSC_PUSH 0
#define PSC_PULL 1

/* psc_step flags */
#define PSC_STEP_NONTEMPORAL 1      /* bit 0: non-temporal state loads/stores; the  */
                                    /* only flag the FP64 step uses                 */
#define PSC_STEP_BPC_MASK 0xff00    /* bits 8-15: cap on resident workgroups per CU */
#define PSC_STEP_CPT_MASK 0x30000   /* bits 16-17: cells per thread along y (1, 2)  */
#define PSC_BATCH_WG512 0x40000     /* psc_step_batched: 512-thread workgroups      */
#define PSC_BATCH_IMB_NT 0x80000    /* psc_step_batched: non-temporal imb stores    */
#define PSC_RUN_INPLACE 0x200000    /* bit 21, PscStepPlan.flags of psc_run_steps   */
                                    /* in the fused form only: steps go in pairs    */
                                    /* updating one state buffer in place (AA       */
                                    /* pattern), half the footprint, same bits      */
#define PSC_STEP_PLANES64 0x400000  /* bit 22: FP32 steps; each direction plane of  */
                                    /* the state holds nx*ny*nz rounded up to 64    */
                                    /* cells (256-byte aligned planes; allocate     */
                                    /* 19 such planes; psc_moments reads them with  */
                                    /* pitch_z = nz); count_sum, imb and us_num     */
                                    /* keep nx*ny*nz cells                          */
#define PSC_STEP_LINEAR 0x100000    /* bit 20: FP32 steps; a workgroup takes        */
                                    /* block_z*block_y consecutive cells of the     */
                                    /* lattice's flat (x, y, z) index instead of a  */
                                    /* (z, y) tile; one cell per thread; no idle    */
                                    /* lanes at row ends, whole granules written    */

#define PSC_NSUM 13 /* number of FP64 sums produced by psc_moments           */

/* subsamples value of psc_step / psc_moments / psc_run_steps selecting a
 * given solid-fraction field: d_count_sum then points to alpha in [0, 1],
 * one value per cell in the storage precision (float for PSC_F32_DEV,
 * double for PSC_F64_PLAIN), e.g. the output of psc_smooth_alpha. */
#define PSC_ALPHA_FIELD 0

/* Library version (major*10000 + minor*100 + patch). */
int psc_version(void);

/* Message of the last failing call on this host thread ("" if none). */
const char* psc_last_error(void);

/* D3Q19 velocity set, weights and opposite table (psc3d.py:17-43), host
 * memory: c[19*3] int32, w[19] double, opp[19] int32. Host-only, no GPU. */
int psc_lattice_constants(int32_t* c, double* w, int32_t* opp);

/* Zero nbytes of device memory at d_ptr with a fill kernel (graph-safe).   */
int psc_clear(void* d_ptr, int64_t nbytes, void* stream);

/* Initial state rho = 1, u = 0 (psc3d.py:136-140 / :312-316): f = w in
 * PSC_F64_PLAIN, g = 0 in PSC_F32_DEV.  ncell = N. */
int psc_init_state(void* d_state, int32_t storage, int64_t ncell, void* stream);

/* Module function equilibrium(rho, ux, uy, uz) (psc3d.py:97-103) over n
 * points, FP64, in the reference's operation order: d_out[i*n + c] for
 * direction i (the reference's (19, *shape) array). */
int psc_equilibrium(const double* d_rho, const double* d_ux, const double* d_uy, const double* d_uz, int64_t n,
                    double* d_out, void* stream);

/* Module function guo_force(fx, fy, fz, ux, uy, uz, tau) (psc3d.py:106-114)
 * over n points, FP64, in the reference's operation order; layout as
 * psc_equilibrium. */
int psc_guo_force(double fx, double fy, double fz, const double* d_ux, const double* d_uy, const double* d_uz,
                  int64_t n, double tau, double* d_out, void* stream);

/* K1 -- sphere-to-lattice solid fraction by s^3 sub-sample counting.
 * Replaces _sparse_particle_alpha_task (psc3d.py:60-94) with the sum of
 * SparsePscD3Q19.rebuild_geometry (psc3d.py:328-345) / the dense
 * PscD3Q19.rebuild_geometry (psc3d.py:151-179), and the numerator of
 * solid_velocity (psc3d.py:181-206, :347-373).  One workgroup per chunk:
 *   entries of chunk c (the cells of its box range with count > 0, in box
 *   order) are written at [c*chunk_cells, c*chunk_cells + d_chunk_nnz[c]):
 *     d_ent_cell  wrapped cell id + cell_base (uint32; cell_base places a
 *                 window's cells in arrays shared by batched windows, 0
 *                 otherwise)
 *     d_ent_loc   box-local coordinates lx | ly << 10 | lz << 20 (box
 *                 extents must not exceed 1024 cells)
 *     d_ent_cnt   number of sub-samples inside (uint16; alpha = cnt/s^3)
 *   d_count_sum[cell] += count              (must be cleared before the call)
 *   d_us_num[a*N+cell] += (count/s^3)*(U + omega x r)_a   if d_us_num != NULL
 *                     (FP32 if us_precision == 1, FP64 if 2; must be cleared)
 * Box cells are wrapped into the lattice, so clamped boxes (the reference,
 * psc3d.py:73-78) and unclamped periodic boxes use the same call; lever
 * arms use the unwrapped cell position.  pitch_z (0 = nz) is the row pitch
 * of count_sum/us_num when rows are padded (batched windows); N = nx*ny*pitch_z
 * is the plane size of d_us_num, the planes psc_step and psc_clear_entries
 * read.  Deterministic apart from the
 * order of the FP atomics into d_us_num on cells shared by two spheres. */
int psc_voxelize(const double* d_particles, const int32_t* d_boxes, const int32_t* d_chunks,
                 int32_t nchunks, int32_t chunk_cells, int32_t nx, int32_t ny, int32_t nz,
                 int32_t pitch_z, int32_t subsamples, uint32_t* d_ent_cell, uint32_t* d_ent_loc, uint16_t* d_ent_cnt,
                 int32_t* d_chunk_nnz, uint32_t* d_count_sum, void* d_us_num, int32_t us_precision,
                 uint32_t cell_base, void* stream);

/* Incremental re-voxelization support: zero d_count_sum (and d_us_num, when
 * not NULL) at every cell listed in the entries of a previous psc_voxelize
 * call (same tables).  Replaces a full-lattice clear when moving spheres are
 * re-voxelized every step (psc3d.py:209-210 with static_geometry=False). */
int psc_clear_entries(const uint32_t* d_ent_cell, const int32_t* d_chunk_nnz, int32_t nchunks,
                      int32_t chunk_cells, int32_t nx, int32_t ny, int32_t nz, uint32_t* d_count_sum,
                      void* d_us_num, int32_t us_precision, void* stream);

/* Owner-table voxelization for FP32 moving steps: the same count_sum and
 * entries as psc_voxelize (pitch_z = nz, cell_base 0; psc3d.py:151-179,
 * :328-345), and the solid velocity of psc3d.py:181-206 / :347-373 in a
 * second form.  d_owner[cell] (N uint32) receives the particle whose sphere
 * alone covers the cell, or 0xffffffff where several spheres cover it; at
 * those shared cells only, d_us_shared (three FP32 planes of N cells,
 * zero elsewhere on entry) accumulates sum_p alpha_p (U + omega x r)_p as
 * psc_voxelize's d_us_num does everywhere.  d_motion (96 bytes per
 * particle, 16-byte aligned) receives each particle's motion record: box
 * origin (3 int32 + 4 bytes of padding), then centre, U and omega (9 doubles
 * + 8 bytes of padding).  psc_step_owners forms the single-owner velocity
 * from it, so the three solid-velocity atomics per covered cell are gone.  d_count_sum must be cleared before the call; the owner table needs
 * no clearing.  Sphere boxes must span at most one period (as the periodic
 * table builder enforces). */
int psc_voxelize_owners(const double* d_particles, const int32_t* d_boxes, int32_t nparticles,
                        const int32_t* d_chunks, int32_t nchunks, int32_t chunk_cells, int32_t nx, int32_t ny,
                        int32_t nz, int32_t subsamples, uint32_t* d_ent_cell, uint32_t* d_ent_loc,
                        uint16_t* d_ent_cnt, int32_t* d_chunk_nnz, uint32_t* d_count_sum, uint32_t* d_owner,
                        float* d_us_shared, void* d_motion, void* stream);

/* The incremental clear of the owner-table form (psc3d.py:209-210 with
 * static_geometry=False): zero d_count_sum at every cell of a previous
 * psc_voxelize_owners call's entries and d_us_shared at those that were
 * shared, which leaves d_us_shared zero everywhere. */
int psc_clear_entries_owners(const uint32_t* d_ent_cell, const uint16_t* d_ent_cnt, const int32_t* d_chunk_nnz,
                             int32_t nchunks, int32_t chunk_cells, int32_t nx, int32_t ny, int32_t nz,
                             uint32_t* d_count_sum, float* d_us_shared, void* stream);

/* psc_step for FP32 deviation storage, pull layout, moving spheres voxelized
 * by psc_voxelize_owners (psc3d.py:208-246 with the solid velocity of
 * psc3d.py:181-206): at a cell with one owner the solid velocity is that
 * particle's U + omega x r, formed from its motion record in FP64 with the
 * voxelizer's lever arm (box coordinate + 1/2 - centre) and rounded to FP32;
 * at a shared cell it is d_us_shared/(count_sum/s^3) as in psc_step.  d_motion: the
 * nparticles motion records of that voxelization; an owner id at or past
 * nparticles (an owner table that does not belong to them) reads d_us_shared
 * instead of the records.  Other arguments as psc_step. */
int psc_step_owners(const float* d_in, float* d_out, const uint32_t* d_count_sum, const uint32_t* d_owner,
                    const float* d_us_shared, const void* d_motion, int32_t nparticles, float* d_imb, int32_t nx,
                    int32_t ny, int32_t nz, int32_t subsamples, double tau, double fx, double fy, double fz,
                    int32_t block_z, int32_t block_y, int32_t flags, void* stream);

/* K2 -- fused PSC collision + streaming, one lattice step.
 * Replaces PscD3Q19.collide_stream_psc (psc3d.py:208-246) with equilibrium
 * (psc3d.py:97-103) and guo_force (psc3d.py:106-114) inlined.  The solid
 * fraction of a cell is min(count_sum/s^3, 1); B = a/(a + tau(1-a) + 1e-14);
 * the solid velocity is us_num/(count_sum/s^3) (zero if d_us_num is NULL).
 * The body force (fx, fy, fz) generalises body_force_z; (0, 0, Fz) is the
 * reference.  d_imb (optional; 3 values per cell, interleaved: d_imb[3*cell
 * + axis]) receives sum_i (B Omega_s)_i c_i
 * (psc3d.py:234-236) in every cell with count_sum > 0; other cells are left
 * untouched or set to 0 (FP32 storage writes every lane of a wavefront that
 * holds a solid cell), so readers mask imb by count_sum > 0.  d_in and d_out
 * must not alias.  block_z/block_y select the
 * workgroup shape (0 = default); flags: PSC_STEP_NONTEMPORAL streams the
 * distributions with non-temporal loads/stores (both storages); with FP32
 * storage bits 8-15 also cap the resident workgroups per CU (through an LDS
 * reservation, 0 = no cap) and bits 16-17 set the cells per thread (1 or 2);
 * the FP64 step ignores those;
 * storage/layout as above.  Precision of d_in/d_out/d_us_num/d_imb follows
 * `storage`.  subsamples = PSC_ALPHA_FIELD reads the solid fraction from a
 * field instead (no solid velocity: d_us_num must be NULL); a cell is fluid
 * where the field is +0. */
int psc_step(const void* d_in, void* d_out, const uint32_t* d_count_sum, const void* d_us_num,
             void* d_imb, int32_t storage, int32_t layout, int32_t nx, int32_t ny, int32_t nz,
             int32_t subsamples, double tau, double fx, double fy, double fz, int32_t block_z,
             int32_t block_y, int32_t flags, void* stream);

/* One lattice window of a batch (device table entry, 64 bytes). */
typedef struct PscWindow {
  int64_t state_off;  /* element offset of the window's 19 planes in the state buffers */
  int64_t cell_off;   /* cell offset in count_sum and imb (shared by all windows)      */
  int32_t nx, ny, nz; /* window extents                                               */
  int32_t pz;         /* z row pitch >= nz (rows padded for aligned access)           */
  float inv_pz, inv_nypz; /* 1/pz and 1/(ny*pz) in FP32 (index decomposition)        */
  float fx, fy, fz;   /* body force of the window                                    */
  int32_t span;       /* cells per work item, at most the workgroup size (<= 0: the   */
                      /* workgroup size); items of whole rows keep rows unsplit       */
  int64_t ps;         /* cells per direction plane: nx*ny*pz rounded up to 64 (256 B) */
} PscWindow;

/* K5 -- batched FP32 deviation step over many independent periodic windows
 * in one persistent launch (pull layout, stationary solids).  Window w's
 * state is 19 planes of its N_w = nx*ny*pz cells (rows padded to the pitch
 * pz; padding cells are never read or written), each plane ps cells long,
 * at d_in + state_off; its
 * count_sum/imb cells start at cell_off (imb: 3 interleaved values per cell,
 * imb_stride cells in all).  d_items lists (window, first cell) pairs; an item covers the
 * window's `span` cells from there (the workgroup size when span <= 0).  Each
 * workgroup (256 threads, 512 with PSC_BATCH_WG512) steps one item at a time,
 * looping over the items with a stride of grid_blocks (0 = one workgroup per
 * item).
 * Each window's step is collide_stream_psc (psc3d.py:208-246), the same
 * arithmetic as psc_step with PSC_F32_DEV.  N_w must stay below 2^22 cells.
 * flags: PSC_STEP_NONTEMPORAL, the workgroups-per-CU cap of psc_step,
 * PSC_BATCH_WG512 and PSC_BATCH_IMB_NT. */
int psc_step_batched(const float* d_in, float* d_out, const uint32_t* d_count_sum, float* d_imb,
                     int64_t imb_stride, const PscWindow* d_windows, const int32_t* d_items,
                     int32_t nitems, int32_t subsamples, double tau, int32_t flags,
                     int32_t grid_blocks, void* stream);

/* Tables of the particle-load reduction (the arguments of
 * psc_particle_loads that do not change from step to step). */
typedef struct PscLoadsTables {
  const double* particles;
  const int32_t* boxes;
  const int32_t* chunks;
  const int32_t* chunk_first;
  int32_t nparticles, nchunks, chunk_cells;
  const uint32_t* ent_cell;
  const uint32_t* ent_loc;
  const uint16_t* ent_cnt;
  const int32_t* chunk_nnz;
  const uint32_t* count_sum;
} PscLoadsTables;

/* psc_step_batched with the particle-load reduction of earlier steps in the
 * same launch (the batched form of psc_run_steps' fused loads; it replaces
 * the pair psc_step_batched + psc_particle_loads of the drivers' per-step
 * loop, psc3d.py:208-281, for many windows).  The tables describe every
 * window's particles over the flattened cells of all windows (nx = ny = 1,
 * nz = imb_stride in psc_particle_loads' terms); FP32 imb.  With d_imb_prev
 * set the launch also sums its chunks (the previous step's imb) into
 * d_partials_out; with d_partials_prev set it also adds those partials up
 * per particle into d_loads (the loads of the step before the previous
 * one).  The extra work items lead the item sequence, one in every `period`
 * (>= 1).  d_imb must be set, and differ from d_imb_prev, d_partials_out from
 * d_partials_prev.  The loads have the bits of psc_particle_loads. */
int psc_step_batched_loads(const float* d_in, float* d_out, const uint32_t* d_count_sum, float* d_imb,
                           int64_t imb_stride, const PscWindow* d_windows, const int32_t* d_items,
                           int32_t nitems, int32_t subsamples, double tau, int32_t flags,
                           int32_t grid_blocks, const PscLoadsTables* tables, const float* d_imb_prev,
                           double* d_partials_out, const double* d_partials_prev, double* d_loads,
                           int32_t period, void* stream);

/* K3 -- node-summed hydrodynamic force and torque per particle.
 * Replaces accumulate_particle_loads (psc3d.py:255-281 dense, :375-407
 * sparse): owner = alpha_p / sum alpha = cnt / count_sum, node force
 * -owner*imb, torque about the centre with lever arm (cell + 1/2) - centre.
 * Pass 1 sums the entries psc_voxelize wrote for each chunk in FP64 into
 * d_partials (nchunks*6 doubles); pass 2 adds the partials of chunks
 * d_chunk_first[p] .. d_chunk_first[p+1]-1 into d_loads (P*6 doubles: Fx Fy
 * Fz Tx Ty Tz).  Fixed thread assignment and reduction trees: deterministic.
 * d_imb: 3 interleaved values per cell (d_imb[3*cell + axis]);
 * imb_precision: 1 = FP32 imb, 2 = FP64 imb. */
int psc_particle_loads(const double* d_particles, const int32_t* d_boxes, const int32_t* d_chunks,
                       const int32_t* d_chunk_first, int32_t nparticles, int32_t nchunks,
                       int32_t chunk_cells, const uint32_t* d_ent_cell, const uint32_t* d_ent_loc,
                       const uint16_t* d_ent_cnt, const int32_t* d_chunk_nnz,
                       const uint32_t* d_count_sum, const void* d_imb, int32_t imb_precision,
                       int32_t nx, int32_t ny, int32_t nz, double* d_partials, double* d_loads,
                       void* stream);

/* K4 -- post-stream moments and FP64 readout sums.
 * Replaces the tail of collide_stream_psc (psc3d.py:243-253) and the
 * validation readouts (run_precision_validation.py:142-147,438;
 * run_psc_permeability_resolution.py:144-160).  For every cell it forms
 * rho = sum f, u = (sum f c + F/2)/rho of the post-stream state; if d_rho_u
 * is not NULL it stores rho, ux, uy, uz as 4 FP64 planes.  d_sums receives
 * PSC_NSUM FP64 values in this order:
 *   0 sum solid, 1 sum (1-solid), 2-4 sum u, 5-7 sum (1-solid) u,
 *   8-10 sum imb over cells with count_sum > 0, 11 sum rho,
 *   12 number of cells with a non-finite moment.
 * solid = min(count_sum/s^3, 1), or the field value for subsamples =
 * PSC_ALPHA_FIELD.  d_partials needs psc_moments_blocks(N)*
 * PSC_NSUM doubles.  Deterministic (fixed-order two-pass reduction).
 * d_imb may be NULL (sums 8-10 are then 0).  pitch_z (0 = nz): row pitch of a
 * padded lattice (padding cells are skipped; fields need pitch_z = 0 or nz,
 * and are written as 4 planes of nx*ny*nz); with pitch_z > 0 each direction
 * plane holds nx*ny*pitch_z cells rounded up to a multiple of 64 (the
 * batched-window layout; pitch_z = nz reads a PSC_STEP_PLANES64 state). */
int psc_moments(const void* d_state, int32_t storage, int32_t layout, const uint32_t* d_count_sum,
                const void* d_imb, int32_t nx, int32_t ny, int32_t nz, int32_t pitch_z,
                int32_t subsamples, double fx,
                double fy, double fz, double* d_rho_u, double* d_partials, double* d_sums,
                void* stream);

/* Number of partial-sum blocks psc_moments uses for N cells. */
int64_t psc_moments_blocks(int64_t ncell);

/* psc_moments without the d_partials workspace: it allocates nothing (no
 * hipMalloc, no library allocation) and keeps the block partials in a
 * library-owned device array of 2048 x PSC_NSUM doubles (208 KiB per GPU,
 * held until the process ends), which every lattice fits (the block count
 * is capped at 2048); a larger need is rejected with PSC_EINVAL.  No
 * counterpart in the reference: it is for callers that cannot allocate
 * (code under graph capture, serving loops, harnesses that forbid it).
 * Every call on a GPU uses that GPU's one array and writes every partial
 * it reads, so nothing has to start at zero; calls that overlap on one
 * GPU (different streams or threads) race.  Same arguments, results (bit
 * for bit) and errors as psc_moments otherwise. */
int psc_moments_noalloc(const void* d_state, int32_t storage, int32_t layout, const uint32_t* d_count_sum,
                        const void* d_imb, int32_t nx, int32_t ny, int32_t nz, int32_t pitch_z,
                        int32_t subsamples, double fx, double fy, double fz, double* d_rho_u, double* d_sums,
                        void* stream);

/* Post-stream distributions f as FP64 (19 planes), the reference's `f`
 * attribute (psc3d.py:239-242): w + g for PSC_F32_DEV, gathered along c_i
 * for PSC_PULL. */
int psc_export_f64(const void* d_state, int32_t storage, int32_t layout, int32_t nx, int32_t ny,
                   int32_t nz, double* d_f, void* stream);

/* Inverse of psc_export_f64: set the state from post-stream FP64 f. */
int psc_import_f64(const double* d_f, int32_t storage, int32_t layout, int32_t nx, int32_t ny,
                   int32_t nz, void* d_state, void* stream);

/* Everything one static-geometry step needs: the K2 arguments of psc_step
 * and, when nparticles > 0, the K3 arguments of psc_particle_loads.  Plain
 * pointers and sizes; the state is ping-ponged between state_a and state_b. */
typedef struct PscStepPlan {
  void* state_a;
  void* state_b;
  const uint32_t* count_sum;
  const void* us_num;
  void* imb;
  void* imb_alt;        /* optional second imb buffer (enables pipelining)   */
  int32_t storage, layout, nx, ny, nz, subsamples;
  double tau, fx, fy, fz;
  int32_t block_z, block_y, flags;
  const double* particles;
  const int32_t* boxes;
  const int32_t* chunks;
  const int32_t* chunk_first;
  int32_t nparticles, nchunks, chunk_cells;
  const uint32_t* ent_cell;
  const uint32_t* ent_loc;
  const uint16_t* ent_cnt;
  const int32_t* chunk_nnz;
  double* partials;
  double* partials_alt; /* optional second partials buffer (pipelining)      */
  double* loads;
  int32_t fuse_loads;   /* k > 0: fuse the load reduction into the step kernels */
} PscStepPlan;

/* nsteps repetitions of collide_stream_psc with static geometry
 * (psc3d.py:208-253 in a loop, as the drivers run it, e.g.
 * run_psc_permeability_resolution.py:137-138): K2 from the current buffer
 * into the other one, then K3.  The step starts from state_a when *parity is
 * 0 (state_b when 1); on return *parity names the buffer holding the state.
 * A step writes imb, or imb_alt when it is set and the step's input has
 * parity 1 (then the last step's imb is the one of parity 1 - *parity on
 * return).  With particles and imb_alt/partials_alt set, the load reduction
 * of step k runs on a library-owned side stream concurrently with step k+1
 * (fork/join through events, captured into the graph as parallel branches);
 * every step's loads are computed and d_loads ends with the last step's.
 * With fuse_loads = k > 0, FP32 pull storage, no solid velocity, particles
 * and imb_alt/partials_alt set, the loads are fused instead: the kernel of
 * step k also sums step k-1's chunks (extra workgroups, one per chunk) and
 * adds up step k-2's chunk partials per particle (one per particle), so a
 * step is one launch; the extra workgroups lead the grid, one in every k
 * workgroups (k = 1: all ahead of the cell workgroups).  The loads still
 * pending after the last step run after the loop.  Every step's loads are
 * computed, with the bits of the separate reduction, and d_loads ends with
 * the last step's.  PSC_RUN_INPLACE in flags (fused form only) runs the steps
 * in pairs on one state buffer, in place: a gather-scatter step that leaves
 * the post-stream populations of direction i in plane opp(i), then a local
 * step back to the pull form.  An odd count starts with an ordinary step to
 * the other buffer and pairs on that one, so on return *parity and the state
 * have their usual meaning, and the next run continues on the buffer the
 * cache holds; the bits equal those of ordinary steps.  At
 * 128^3 one 160 MB state buffer stays resident in the 256 MB Infinity Cache
 * where two do not.
 * use_graph = 1 runs the steps as captured hipGraphs (graph_steps even,
 * >= 2): a fused run takes two direct steps (the GPU starts at once while
 * the host submits the first graph), replays a graph of graph_steps steps as
 * often as it fits and ends with a graph of the remaining 1 .. graph_steps
 * steps that also reduces the pending loads.  An unfused run is replays of a
 * graph_steps graph and one graph of the rest.
 * Consecutive graph launches leave the GPU idle for a few microseconds, so
 * longer graphs cost less.  Graphs are cached per plan contents, kind, step
 * count and starting parity until psc_graph_cache_clear(); use_graph = 0
 * launches the same kernels directly.  A graph is uploaded to the GPU when it
 * is created (the host waits for that upload on the library's own capture
 * stream, never on `stream`), so its first launch costs no more than later
 * ones. */
int psc_run_steps(const PscStepPlan* plan, int32_t nsteps, int32_t use_graph, int32_t graph_steps,
                  int32_t* parity, void* stream);

/* Capture, without launching anything, the graphs that psc_run_steps(plan,
 * nsteps, 1, graph_steps, parity, ...) replays, and upload them, so that
 * the run itself starts with no capture and no upload (e.g. a timed run after
 * a warm-up of another length).  Same checks and errors as psc_run_steps; the state is not
 * touched.  No counterpart in the reference, whose loop has no launches. */
int psc_prepare_steps(const PscStepPlan* plan, int32_t nsteps, int32_t graph_steps, const int32_t* parity);

/* Destroy every cached hipGraph of psc_run_steps. */
int psc_graph_cache_clear(void);

/* Bits of *pending in psc_run_steps_deferred / psc_flush_loads. */
#define PSC_PENDING_CHUNKS 1  /* the chunk pass of the last step's imb is owed        */
#define PSC_PENDING_FINAL 2   /* the per-particle pass of the step before it is owed  */

/* nsteps collide_stream_psc steps (psc3d.py:208-253) of a fused plan (the
 * fused form of psc_run_steps: fuse_loads > 0, FP32 pull storage, no solid
 * velocity, particles, imb_alt and partials_alt set; PSC_RUN_INPLACE allowed)
 * as direct launches whose particle loads (psc3d.py:255-281) are left
 * pending: each launch steps the state, writes the imb buffer of its input
 * parity and reduces, in the same launch, the loads that the steps before it
 * still owe; *pending says what is owed on entry and on return.  The steps
 * are those of psc_run_steps (with PSC_RUN_INPLACE: pairs in place, an odd
 * count's first step an ordinary one), so on return *parity names the buffer
 * holding the state.  Start a sequence with *pending = 0; do not interleave
 * other steps on the same buffers (psc_flush_loads, or a reduction of the
 * last step's loads with psc_particle_loads, ends it).
 * No counterpart in the reference, whose step reduces the loads before it
 * returns: for callers stepping one call at a time that read the loads only
 * now and then, as run_psc_permeability_resolution.py:137-143 does. */
int psc_run_steps_deferred(const PscStepPlan* plan, int32_t nsteps, int32_t* parity, int32_t* pending,
                           void* stream);

/* Reduce the loads still pending after psc_run_steps_deferred (the same plan
 * and the parity it returned), oldest first, so that d_loads then holds the
 * last step's loads with the bits of psc_particle_loads after that step; sets
 * *pending to 0.  No counterpart in the reference (see
 * psc_run_steps_deferred). */
int psc_flush_loads(const PscStepPlan* plan, int32_t parity, int32_t* pending, void* stream);

/* Plain device copy of n floats, the bandwidth reference for the roofline
 * and for scaling HBM byte counters (no counterpart in the reference). */
int psc_copy_f32(const float* d_src, float* d_dst, int64_t n, void* stream);

/* ------------------------------------------------------------------------
 * Plain bounce-back D3Q19 window solver of the drivers (the drag proxy and
 * the fines permeability probes).
 * ---------------------------------------------------------------------- */

/* Binary voxelization.  Replaces voxelize (scripts/run_3d_window_lbm.py:
 * 73-92) and voxelize_variable (scripts/run_local_fines_lbm_permeability.py:
 * 39-59): cell (i, j, k) of the nx*ny*nz mask d_solid (uint8, C order) is set
 * to 1 when ((ax[i]-x)^2 + (ay[j]-y)^2) + (az[k]-z)^2 <= r^2 for a sphere
 * whose inclusive cell box (d_boxes: nspheres rows of int32 i0, i1, j0, j1,
 * k0, k1, within the lattice) contains it.  d_spheres: nspheres rows of
 * double x, y, z, r; d_axes: the cell-centre coordinates ax (nx), ay (ny),
 * az (nz), concatenated, in the spheres' units.  Cells are only ever set:
 * clear d_solid first.  FP64 with the reference's operation order. */
int psc_voxelize_binary(const double* d_spheres, const int32_t* d_boxes, int32_t nspheres,
                        const double* d_axes, int32_t nx, int32_t ny, int32_t nz, uint8_t* d_solid,
                        void* stream);

/* One step of run_lbm (scripts/run_3d_window_lbm.py:95-146) in the pull
 * layout: BGK with the Guo force (fx, fy, fz); (0, 0, force_z) is the
 * reference), velocity u = (sum f c + F/2)/rho zeroed in solid cells
 * (d_solid != 0), streaming, and at solid cells the reference's in-place
 * reflection loop f_i <- f_opp(i) in direction order, after which both
 * members of each opposite pair hold the even member's incoming value.
 * The state holds post-collision values: gather + reflection of it is the
 * reference's f.  PSC_F64_PLAIN reproduces the reference bit for bit;
 * PSC_F32_DEV stores g = f - w in FP32.  Start from psc_init_state.
 * block/flags as psc_step (no cells-per-thread bits). */
int psc_bb_step(const void* d_in, void* d_out, const uint8_t* d_solid, int32_t storage, int32_t nx,
                int32_t ny, int32_t nz, double tau, double fx, double fy, double fz, int32_t block_z,
                int32_t block_y, int32_t flags, void* stream);

/* ------------------------------------------------------------------------
 * Differentiation: permeability gradients with respect to the particles.
 * No counterpart in the reference (its solid fraction is a piecewise-
 * constant sub-sample count, psc3d.py:83-89, with zero derivative almost
 * everywhere).  The smooth solid fraction below replaces it for gradient
 * work; the adjoint entry points are the reverse mode of psc_step in the
 * pull layout with a PSC_ALPHA_FIELD solid fraction and stationary solids.
 * ---------------------------------------------------------------------- */

/* K1s -- smooth solid fraction of a periodic lattice:
 *   alpha(x) = 1 - prod_p (1 - S(r_p + 1/2 - |x - c_p|)),
 * x the cell centre (ix + 1/2, ...), S the quintic smoothstep on [0, 1]
 * (S = 0 below 0, 1 above 1).  Boxes/chunks as for psc_voxelize, built
 * without clamping (periodic); each box must contain the support
 * |x - c| < r + 1/2.  Per cell, d_nfull (uint32) counts the particles with
 * S = 1 and d_logsum (int64) sums log(1 - S) of the others in fixed point
 * (units of 2^-50, integer atomics: the result does not depend on the order
 * of the particles and is bitwise reproducible); both are cleared here and
 * kept for psc_smooth_alpha_grad.  alpha is written as FP64 (d_alpha64)
 * and/or FP32 (d_alpha32); either may be NULL. */
int psc_smooth_alpha(const double* d_particles, const int32_t* d_boxes, const int32_t* d_chunks,
                     int32_t nchunks, int32_t chunk_cells, int32_t nx, int32_t ny, int32_t nz,
                     uint32_t* d_nfull, int64_t* d_logsum, double* d_alpha64, float* d_alpha32,
                     void* stream);

/* Chain rule through psc_smooth_alpha: given dL/dalpha per cell (d_galpha,
 * FP64) and the d_nfull/d_logsum of the same call, writes d_grad (P rows of
 * 4 doubles: dL/dx, dL/dy, dL/dz, dL/dr).  Pass 1 writes one partial per
 * chunk into d_partials (nchunks*4 doubles), pass 2 adds the chunks
 * d_chunk_first[p] .. d_chunk_first[p+1]-1 of each particle in order
 * (deterministic). */
int psc_smooth_alpha_grad(const double* d_particles, const int32_t* d_boxes, const int32_t* d_chunks,
                          const int32_t* d_chunk_first, int32_t nchunks, int32_t chunk_cells,
                          int32_t nparticles, int32_t nx, int32_t ny, int32_t nz,
                          const uint32_t* d_nfull, const int64_t* d_logsum, const double* d_galpha,
                          double* d_partials, double* d_grad, void* stream);

/* Adjoint seed of the permeability along the body force after the last
 * step, k = nu sum_x (1 - alpha) (e . u) / (N |F|), nu = (tau - 1/2)/3,
 * e = F/|F|, u = (j + F/2)/rho of the post-stream state (the readout of
 * run_psc_permeability_resolution.py:144-150 with a field alpha).  Writes
 * d_lam = weight * dk/d(state) (19 planes, pull layout, storage precision)
 * and sets d_galpha = weight * dk/dalpha (explicit part, FP64 per cell). */
int psc_adjoint_seed(const void* d_state, const void* d_alpha, void* d_lam, double* d_galpha,
                     int32_t storage, int32_t nx, int32_t ny, int32_t nz, double tau, double fx,
                     double fy, double fz, double weight, void* stream);

/* K6 -- one reverse step.  d_state is the forward state the step started
 * from (pull layout), d_lam_in the adjoint of the state it produced; writes
 * d_lam_out, the adjoint of d_state (transposed collision Jacobian, then
 * the reversed streaming as a scatter), and adds dL/dalpha of this step to
 * d_galpha in every cell with 0 < alpha < 1 (the only cells where a smooth
 * solid fraction depends on the particles).  Storage precision as for
 * psc_step (FP32: deviations, adjoint in FP32; FP64: plain, adjoint in
 * FP64); d_alpha as for PSC_ALPHA_FIELD.  block_z/block_y: workgroup
 * shape (0, 0 = 64 x 4, or 32 x 8 for rows that are a multiple of 32 but
 * not of 64); flags: PSC_STEP_NONTEMPORAL and the workgroups-per-CU cap
 * (bits 8-15) of psc_step. */
int psc_adjoint_step(const void* d_state, const void* d_lam_in, void* d_lam_out, const void* d_alpha,
                     double* d_galpha, int32_t storage, int32_t nx, int32_t ny, int32_t nz,
                     double tau, double fx, double fy, double fz, int32_t block_z, int32_t block_y,
                     int32_t flags, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* PSC_HIP_H */
