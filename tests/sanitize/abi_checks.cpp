// Host-side checks of the C ABI under AddressSanitizer and UndefinedBehaviorSanitizer
// (built by tests/test_abi_sanitize.py with host-only instrumentation; no GPU is used).
// Every entry point is called with arguments it must reject -- null pointers, bad
// lattices (including extents whose products overflow 32 bits), bad flags, sizes and
// plans -- and the calls that complete without a device (constants, size queries,
// zero-sized work) are called with valid ones.  The program prints one line per check
// and exits non-zero on the first unexpected result.
#include <climits>
#include <cstdio>
#include <cstring>

#include "psc_hip.h"

static int g_fail = 0;

static void expect(bool ok, const char* what) {
  std::printf("%s %s | %s\n", ok ? "ok  " : "FAIL", what, psc_last_error());
  if (!ok) g_fail = 1;
}

static bool msg_has(const char* s) { return std::strstr(psc_last_error(), s) != nullptr; }

int main() {
  void* p16 = reinterpret_cast<void*>(16);
  void* p32 = reinterpret_cast<void*>(32);
  void* p64 = reinterpret_cast<void*>(64);
  const uint32_t* cs = reinterpret_cast<const uint32_t*>(64);

  expect(psc_version() >= 10000, "version");
  int32_t c[57];
  double w[19];
  int32_t opp[19];
  expect(psc_lattice_constants(c, w, opp) == 0 && opp[1] == 2 && c[3] == 1, "lattice constants");
  expect(psc_lattice_constants(nullptr, w, opp) != 0 && msg_has("null"), "lattice constants null");
  expect(psc_moments_blocks(0) >= 1 && psc_moments_blocks(1LL << 40) >= 1, "moments blocks bounds");
  expect(psc_clear(nullptr, 0, nullptr) == 0, "clear of zero bytes");
  expect(psc_clear(nullptr, -1, nullptr) != 0, "clear negative");
  expect(psc_clear(nullptr, 16, nullptr) != 0 && msg_has("null"), "clear null");
  expect(psc_init_state(nullptr, 1, 8, nullptr) != 0, "init null");
  expect(psc_init_state(p16, 7, 8, nullptr) != 0 && msg_has("storage"), "init storage");
  expect(psc_equilibrium(nullptr, nullptr, nullptr, nullptr, 0, nullptr, nullptr) == 0, "equilibrium n=0");
  expect(psc_equilibrium(nullptr, nullptr, nullptr, nullptr, -3, nullptr, nullptr) != 0, "equilibrium negative");
  expect(psc_equilibrium(nullptr, nullptr, nullptr, nullptr, 5, nullptr, nullptr) != 0 && msg_has("null"),
         "equilibrium null");
  expect(psc_guo_force(0, 0, 1e-6, nullptr, nullptr, nullptr, 5, 0.0, nullptr, nullptr) != 0 && msg_has("tau"),
         "guo tau");

  // lattices: empty, negative, products beyond 2^31, FP32 beyond 2^30
  const int bad[][3] = {{0, 8, 8}, {-1, 8, 8}, {INT_MAX, INT_MAX, INT_MAX}, {65536, 65536, 2}, {1 << 20, 1 << 20, 1 << 20}};
  for (const auto& d : bad) {
    expect(psc_step(p16, p32, cs, nullptr, nullptr, 1, 1, d[0], d[1], d[2], 2, 0.8, 0, 0, 0, 0, 0, 0, nullptr) != 0,
           "step bad lattice");
    expect(psc_moments(p16, 1, 1, cs, nullptr, d[0], d[1], d[2], 0, 2, 0, 0, 0, nullptr, static_cast<double*>(p64),
                       static_cast<double*>(p32), nullptr) != 0,
           "moments bad lattice");
    expect(psc_voxelize(nullptr, nullptr, nullptr, 0, 2048, d[0], d[1], d[2], 0, 2, nullptr, nullptr, nullptr, nullptr,
                        nullptr, nullptr, 0, 0, nullptr) != 0,
           "voxelize bad lattice");
    expect(psc_export_f64(p16, 1, 1, d[0], d[1], d[2], static_cast<double*>(p32), nullptr) != 0, "export bad lattice");
    expect(psc_bb_step(p16, p32, reinterpret_cast<const uint8_t*>(64), 1, d[0], d[1], d[2], 0.8, 0, 0, 0, 0, 0, 0,
                       nullptr) != 0,
           "bounce-back bad lattice");
  }
  expect(psc_step(p16, p32, cs, nullptr, nullptr, 1, 1, 1025, 1024, 1024, 2, 0.8, 0, 0, 0, 0, 0, 0, nullptr) != 0 &&
             msg_has("2^30"),
         "step fp32 2^30");
  expect(psc_step(p16, p16, cs, nullptr, nullptr, 1, 1, 8, 8, 8, 2, 0.8, 0, 0, 0, 0, 0, 0, nullptr) != 0 &&
             msg_has("alias"),
         "step alias");
  expect(psc_step(p16, p32, cs, nullptr, nullptr, 1, 1, 8, 8, 8, 2, 0.8, 0, 0, 0, 48, 3, 0, nullptr) != 0,
         "step block");
  expect(psc_step(p16, p32, cs, nullptr, nullptr, 1, 1, 8, 8, 8, 2, 0.8, 0, 0, 0, 64, 1, 1 << 30, nullptr) != 0 &&
             msg_has("flags"),
         "step flags");
  expect(psc_step(p16, p32, cs, nullptr, nullptr, 1, 1, 8, 8, 8, 41, 0.8, 0, 0, 0, 0, 0, 0, nullptr) != 0 &&
             msg_has("subsamples"),
         "step subsamples");
  expect(psc_step(p16, p32, cs, nullptr, nullptr, 9, 1, 8, 8, 8, 2, 0.8, 0, 0, 0, 0, 0, 0, nullptr) != 0,
         "step storage");

  // batched windows, loads, readouts, adjoint, bounce-back
  expect(psc_step_batched(nullptr, nullptr, nullptr, nullptr, 0, nullptr, nullptr, 0, 2, 0.8, 0, 0, nullptr) != 0,
         "batched null");
  PscLoadsTables t;
  std::memset(&t, 0, sizeof(t));
  expect(psc_step_batched_loads(static_cast<float*>(p16), static_cast<float*>(p32), cs, static_cast<float*>(p64), 10,
                                reinterpret_cast<const PscWindow*>(p64), reinterpret_cast<const int32_t*>(p64), 1, 3,
                                0.8, 0, 0, &t, nullptr, nullptr, nullptr, nullptr, 0, nullptr) != 0,
         "batched loads period");
  expect(psc_particle_loads(nullptr, nullptr, nullptr, nullptr, 3, 2, 2048, nullptr, nullptr, nullptr, nullptr, nullptr,
                            nullptr, 1, 8, 8, 8, nullptr, nullptr, nullptr) != 0,
         "particle loads null");
  expect(psc_particle_loads(nullptr, nullptr, nullptr, nullptr, 0, 0, 2048, nullptr, nullptr, nullptr, nullptr, nullptr,
                            nullptr, 3, 8, 8, 8, nullptr, nullptr, nullptr) != 0 &&
             msg_has("precision"),
         "particle loads precision");
  expect(psc_moments_noalloc(p16, 1, 1, nullptr, nullptr, 8, 8, 8, 0, 2, 0, 0, 0, nullptr, static_cast<double*>(p64),
                             nullptr) != 0 &&
             msg_has("null"),
         "moments noalloc null");
  // valid arguments: without a device the scratch lookup must fail cleanly
  expect(psc_moments_noalloc(p16, 1, 1, cs, nullptr, 8, 8, 8, 0, 2, 0, 0, 0, nullptr, static_cast<double*>(p64),
                             nullptr) != 0,
         "moments noalloc without a device");
  expect(psc_import_f64(static_cast<const double*>(p16), 1, 1, 8, 8, 8, p16, nullptr) != 0 && msg_has("alias"),
         "import alias");
  expect(psc_copy_f32(nullptr, nullptr, 0, nullptr) == 0, "copy of zero");
  expect(psc_copy_f32(nullptr, nullptr, 5, nullptr) != 0, "copy null");
  expect(psc_clear_entries(nullptr, nullptr, 3, 2048, 8, 8, 8, nullptr, nullptr, 0, nullptr) != 0, "clear entries");
  // owner table
  expect(psc_voxelize_owners(nullptr, nullptr, 4, nullptr, 0, 2048, 8, 8, 8, 3, nullptr, nullptr, nullptr, nullptr,
                             nullptr, nullptr, nullptr, nullptr, nullptr) != 0 &&
             msg_has("motion"),
         "owner voxelize tables");
  expect(psc_voxelize_owners(static_cast<const double*>(p64), static_cast<const int32_t*>(p32), 4, nullptr, 3, 2048,
                             8, 8, 8, 3, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, p16,
                             nullptr) != 0 &&
             msg_has("null"),
         "owner voxelize null");
  expect(psc_voxelize_owners(nullptr, nullptr, -1, nullptr, 0, 2048, 8, 8, 8, 3, nullptr, nullptr, nullptr, nullptr,
                             nullptr, nullptr, nullptr, nullptr, nullptr) != 0,
         "owner voxelize sizes");
  expect(psc_voxelize_owners(nullptr, nullptr, 0, nullptr, 0, 2048, 8, 8, 8, 41, nullptr, nullptr, nullptr, nullptr,
                             nullptr, nullptr, nullptr, nullptr, nullptr) != 0 &&
             msg_has("subsamples"),
         "owner voxelize subsamples");
  expect(psc_clear_entries_owners(nullptr, nullptr, nullptr, 3, 2048, 8, 8, 8, nullptr, nullptr, nullptr) != 0 &&
             msg_has("null"),
         "owner clear null");
  expect(psc_clear_entries_owners(nullptr, nullptr, nullptr, 0, 2048, 8, 8, 8, nullptr, nullptr, nullptr) == 0,
         "owner clear of nothing");
  expect(psc_step_owners(static_cast<const float*>(p16), static_cast<float*>(p32), cs, nullptr, nullptr, nullptr,
                         4, nullptr, 8, 8, 8, 3, 0.8, 0, 0, 0, 64, 1, 0, nullptr) != 0 &&
             msg_has("owner"),
         "owner step tables");
  expect(psc_step_owners(static_cast<const float*>(p16), static_cast<float*>(p32), cs,
                         static_cast<const uint32_t*>(p32), static_cast<const float*>(p16), p64, -1, nullptr, 8, 8, 8,
                         3, 0.8, 0, 0, 0, 64, 1, 0, nullptr) != 0 &&
             msg_has("nparticles"),
         "owner step particle count");
  expect(psc_step_owners(static_cast<const float*>(p16), static_cast<float*>(p32), cs,
                         static_cast<const uint32_t*>(p32), static_cast<const float*>(p16),
                         static_cast<const char*>(p16) + 4, 4, nullptr, 8, 8, 8, 3, 0.8, 0, 0, 0, 64, 1, 0,
                         nullptr) != 0 &&
             msg_has("aligned"),
         "owner step motion alignment");
  expect(psc_step_owners(static_cast<const float*>(p16), static_cast<float*>(p16), cs,
                         static_cast<const uint32_t*>(p32), static_cast<const float*>(p16), p64, 4, nullptr, 8, 8, 8,
                         3, 0.8, 0, 0, 0, 64, 1, 0, nullptr) != 0 &&
             msg_has("alias"),
         "owner step alias");
  expect(psc_voxelize_binary(nullptr, nullptr, -1, nullptr, 8, 8, 8, nullptr, nullptr) != 0, "binary voxelize");
  expect(psc_smooth_alpha(nullptr, nullptr, nullptr, 0, 64, 8, 8, 8, nullptr, nullptr, nullptr, nullptr, nullptr) != 0,
         "smooth alpha scratch");
  expect(psc_smooth_alpha_grad(nullptr, nullptr, nullptr, nullptr, 2, 64, 1, 8, 8, 8, nullptr, nullptr, nullptr,
                               nullptr, nullptr, nullptr) != 0,
         "smooth alpha grad");
  expect(psc_adjoint_seed(p16, p32, p64, static_cast<double*>(p64), 1, 8, 8, 8, 0.8, 0, 0, 0, 1.0, nullptr) != 0 &&
             msg_has("force"),
         "adjoint seed force");
  expect(psc_adjoint_step(p16, p32, p32, p64, static_cast<double*>(p64), 1, 8, 8, 8, 0.8, 0, 0, 1e-6, 0, 0, 0,
                          nullptr) != 0 &&
             msg_has("alias"),
         "adjoint step alias");

  // native loop plans
  PscStepPlan plan;
  std::memset(&plan, 0, sizeof(plan));
  int32_t par = 0;
  expect(psc_run_steps(nullptr, 4, 0, 16, &par, nullptr) != 0, "run null plan");
  plan.state_a = p16;
  plan.state_b = p32;
  plan.count_sum = cs;
  plan.storage = 1;
  plan.layout = 1;
  plan.nx = plan.ny = plan.nz = 8;
  plan.subsamples = 2;
  plan.tau = 0.8;
  expect(psc_run_steps(&plan, 4, 1, 7, &par, nullptr) != 0 && msg_has("graph_steps"), "run odd graph size");
  par = 2;
  expect(psc_run_steps(&plan, 4, 1, 8, &par, nullptr) != 0 && msg_has("parity"), "run parity");
  par = 0;
  expect(psc_prepare_steps(&plan, -1, 8, &par) != 0, "prepare negative");
  plan.flags = PSC_RUN_INPLACE;
  expect(psc_run_steps(&plan, 4, 0, 8, &par, nullptr) != 0 && msg_has("fused form"), "run in place unfused");
  plan.flags = 0;
  plan.imb = p64;
  plan.imb_alt = p64;
  plan.partials = static_cast<double*>(p64);
  plan.partials_alt = static_cast<double*>(p32);
  plan.nparticles = 2;
  plan.fuse_loads = 5;
  expect(psc_run_steps(&plan, 4, 0, 8, &par, nullptr) != 0 && msg_has("alias"), "run fused alias");
  plan.imb_alt = p32;
  plan.nx = INT_MAX;
  expect(psc_prepare_steps(&plan, 4, 8, &par) != 0, "prepare bad lattice");
  expect(psc_graph_cache_clear() == 0, "graph cache clear");

  // a long message is truncated, not overrun
  char longname[1024];
  std::memset(longname, 'x', sizeof(longname) - 1);
  longname[sizeof(longname) - 1] = 0;
  expect(psc_step(p16, p32, cs, nullptr, nullptr, 1, 7, 8, 8, 8, 2, 0.8, 0, 0, 0, 0, 0, 0, nullptr) != 0,
         "step layout");
  expect(std::strlen(psc_last_error()) < 512, "error text bounded");
  std::printf("%s\n", g_fail ? "FAILED" : "ALL OK");
  return g_fail;
}
