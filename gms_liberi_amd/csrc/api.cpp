// C ABI of the MI355X Fock build: handle lifecycle, setup, build.
//
// Host side of the drop-in for src/interface/liberi_interface.F90 (handle
// API), src/liberi_shell_pair.F90 (setup) and src/liberi_driver.F90
// (compute_integrals) of gms_libERI.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/liberi_hip.h"
#include "liberi_internal.hpp"

namespace liberi {

const ClassDef kClasses[kNumClasses] = {
    {"int0000", P_SS, P_SS, CMP_GE, PS_SELF},     {"int0001", P_SS, P_SP, CMP_GT, PS_CROSS},
    {"int0011", P_SS, P_PP, CMP_GT, PS_CROSS},    {"int0111", P_SP, P_PP, CMP_GT, PS_CROSS},
    {"int0101", P_SP, P_SP, CMP_GE, PS_SELF},     {"int1111", P_PP, P_PP, CMP_GE, PS_SELF},
    {"int0002", P_SS, P_SD, CMP_GT, PS_CROSS},    {"int0012", P_SS, P_PD, CMP_GT, PS_CROSS},
    {"int0102", P_SP, P_SD, CMP_GT, PS_CROSS},    {"int0112", P_SP, P_PD, CMP_GT, PS_CROSS},
    {"int0022", P_SS, P_DD, CMP_GT, PS_CROSS},    {"int0122", P_SP, P_DD, CMP_GT, PS_CROSS},
    {"int1102", P_PP, P_SD, CMP_GT, PS_CROSS},    {"int0202", P_SD, P_SD, CMP_GT, PS_SELF},
    {"int2211", P_DD, P_PP, CMP_GT, PS_SELF},     {"int2121", P_PD, P_PD, CMP_GT, PS_SELF},
    {"int2120", P_PD, P_SD, CMP_GT, PS_SELF},     {"int2111", P_PD, P_PP, CMP_GT, PS_SELF},
    {"int2220", P_DD, P_SD, CMP_GT, PS_SELF},     {"int2221", P_DD, P_PD, CMP_GT, PS_SELF},
    {"int2222gen", P_DD, P_DD, CMP_NONE, PS_SELF}, {"int3333", P_FF, P_FF, CMP_NONE, PS_SELF},
    {"int0030", P_SS, P_SF, CMP_GT, PS_SELF},     {"int0031", P_SS, P_PF, CMP_GT, PS_SELF},
    {"int0032", P_SS, P_DF, CMP_GT, PS_SELF},     {"int1030", P_SP, P_SF, CMP_GT, PS_SELF},
    {"int3330", P_FF, P_SF, CMP_GT, PS_SELF},     {"int3011", P_SF, P_PP, CMP_GT, PS_SELF},
    {"int3020", P_SF, P_SD, CMP_GT, PS_SELF},     {"int2130", P_PD, P_SF, CMP_GT, PS_SELF},
    {"int2230", P_DD, P_SF, CMP_GT, PS_SELF},     {"int3030", P_SF, P_SF, CMP_GT, PS_SELF},
    {"int3110", P_PF, P_SP, CMP_GT, PS_SELF},     {"int3111", P_PF, P_PP, CMP_GT, PS_SELF},
    {"int3120", P_PF, P_SD, CMP_GT, PS_SELF},     {"int3121", P_PF, P_PD, CMP_GT, PS_SELF},
    {"int2231", P_DD, P_PF, CMP_GT, PS_SELF},     {"int3130", P_PF, P_SF, CMP_GT, PS_SELF},
    {"int3131", P_PF, P_PF, CMP_GT, PS_SELF},     {"int3210", P_DF, P_SP, CMP_GT, PS_SELF},
    {"int3331", P_FF, P_PF, CMP_GT, PS_SELF},     {"int3211", P_DF, P_PP, CMP_GT, PS_SELF},
    {"int3220", P_DF, P_SD, CMP_GT, PS_SELF},     {"int3221", P_DF, P_PD, CMP_GT, PS_SELF},
    {"int3222", P_DF, P_DD, CMP_GT, PS_SELF},     {"int3230", P_DF, P_SF, CMP_GT, PS_SELF},
    {"int3231", P_DF, P_PF, CMP_GT, PS_SELF},     {"int3232", P_DF, P_DF, CMP_GT, PS_SELF},
    {"int3300", P_FF, P_SS, CMP_GT, PS_SELF},     {"int3310", P_FF, P_SP, CMP_GT, PS_SELF},
    {"int3311", P_FF, P_PP, CMP_GT, PS_SELF},     {"int3320", P_FF, P_SD, CMP_GT, PS_SELF},
    {"int3321", P_FF, P_PD, CMP_GT, PS_SELF},     {"int3322", P_FF, P_DD, CMP_GT, PS_SELF},
    {"int3332", P_FF, P_DF, CMP_GT, PS_SELF},
};

}  // namespace liberi

using namespace liberi;

// ------------------------------------------------------------------ errors
static thread_local std::string g_last_error;

static int fail(int code, const std::string &msg) {
  g_last_error = msg;
  return code;
}

#define HIP_TRY(expr)                                                                                   \
  do {                                                                                                  \
    hipError_t e_ = (expr);                                                                             \
    if (e_ != hipSuccess)                                                                               \
      return fail(e_ == hipErrorOutOfMemory ? LIBERI_ENOMEM : LIBERI_EDEVICE,                           \
                  std::string(#expr) + ": " + hipGetErrorString(e_));                                   \
  } while (0)

static void forget_tuned_choices();

// --------------------------------------------------------------- handle
struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
};

struct ClassState {
  bool active = false;
  bool row_ok = false;  // a row-tiled kernel exists and fits in LDS
  int variant = -1;     // form chosen for the class: 0 direct, 1 row-tiled, 3 cooperative (-1: not yet)
  int fp32 = -1;        // mixed form: 1 FP32 allowed, 0 not (-1: not yet chosen for the current tau)
  float ms = -1.f;      // device time of the class's launch, measured at the first build
  int64_t nquart = 0, nprimq = 0;
  int32_t nseg = 0;
  int tile = -1;        // row-tile length chosen for the class, index into kTileLens (-1: not yet)
  DevBuf ket_list, seg_start, seg_bra, seg_ket0;
  // the row tiles of this rank's share, for each tile length
  DevBuf tile_q[kNumTileLens], tile_seg[kNumTileLens];
  int32_t t_begin[kNumTileLens] = {}, t_end[kNumTileLens] = {};
  ClassDev dev{};
};

// Point the class's launch arguments at its row tiles of length kTileLens[k].
static void use_tiles(ClassState &cs_, int k) {
  cs_.dev.tile_q = (const int64_t *)cs_.tile_q[k].p;
  cs_.dev.tile_seg = (const int32_t *)cs_.tile_seg[k].p;
  cs_.dev.t_begin = cs_.t_begin[k];
  cs_.dev.t_end = cs_.t_end[k];
}

struct liberi_ctx {
  // options
  double schwarz_cut = kRefCutoff, value_cut = kRefCutoff, fp32_tau = 0.0, exact_large_t = 0.0;
  double prim_cut = 1e-18;
  int row_tiles = 1;  // 0 never, 1 per class (autotuned or default), 2 wherever they fit
  bool autotune = true;
  double max_device_bytes = 0.0;  // HBM budget of the handle (0: none)
  int rank_split = 0;             // 0 shares balanced by estimated cost, 1 equal quartet counts
  // basis
  bool initialized = false;
  int32_t nsh = 0, natoms = 0, num_bas = 0, mxgtot = 0, rank = 0, nprocs = 1;
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  std::vector<DevBuf> owned;
  DevBuf pairs, prims, dens, fock, scratch;
  int64_t scratch_per_wg = 0;
  DevBuf col_bf[4];     // compact columns per angular momentum -> basis function
  int32_t ncol[4] = {};
  DevBuf aux;           // per build: max |D|, FP32 quartet count
  // fused group launch: work items, every class's ClassDev, the queue head
  DevBuf fused_items, fused_cds, fused_head;
  int64_t fused_n = 0;
  int fused_nwg = 0;
  bool fused = false;      // LIBERI_OPT_FUSED
  int row_orient = 0;      // LIBERI_OPT_ROW_ORIENT
  int tile_len = -1;       // LIBERI_OPT_TILE_LEN as an index into kTileLens; -1 chosen per class
  bool last_fused = false; // the last build ran the fused launch
  DevBuf tune_F;        // autotuner's scratch Fock matrix
  DevBuf dens_t, dens_b, fock_b;  // UHF: Da + Db, and the beta density / Fock of the host entry point
  hipEvent_t tev[2] = {};
  // batched form: the build's class launches spread over kSide streams and
  // captured once into a graph, replayed by later builds on the same D and F
  bool use_graph = true;
  hipStream_t side[4] = {};
  hipEvent_t fork_ev = nullptr, join_ev[4] = {};
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  const double *g_dD = nullptr;
  double *g_dF = nullptr;
  bool fp32_pending = false;
  ClassState cls[kNumClasses];
  liberi_stats stats{};
  // diagnostics
  bool class_timing = false, timing_pending = false, build_pending = false;
  int only_class = -1;
  hipEvent_t cev0[kNumClasses] = {}, cev1[kNumClasses] = {};
};

// fold the events of the last build into the statistics (waits for it)
static void collect_timing(liberi_ctx *h) {
  if (h->build_pending) {
    if (hipEventSynchronize(h->ev1) == hipSuccess) {
      float ms = 0.f;
      if (hipEventElapsedTime(&ms, h->ev0, h->ev1) == hipSuccess) h->stats.last_build_ms = ms;
    }
    h->build_pending = false;
    if (h->fp32_pending) {
      uint64_t n = 0;
      if (hipMemcpy(&n, (uint64_t *)h->aux.p + 1, sizeof(n), hipMemcpyDeviceToHost) == hipSuccess)
        h->stats.fp32_quartets = (int64_t)n;
      h->fp32_pending = false;
    }
  }
  if (!h->timing_pending) return;
  for (int c = 0; c < kNumClasses; ++c) {
    h->stats.class_ms[c] = 0.0;
    if (!h->cls[c].active || !h->cev0[c]) continue;
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, h->cev0[c], h->cev1[c]) == hipSuccess) h->stats.class_ms[c] = ms;
  }
  h->timing_pending = false;
}

static int dev_alloc(liberi_ctx *h, DevBuf &b, size_t bytes) {
  b.bytes = bytes;
  if (bytes == 0) {
    b.p = nullptr;
    return LIBERI_OK;
  }
  if (h->max_device_bytes > 0.0 && (double)h->stats.device_bytes + (double)bytes > h->max_device_bytes) {
    b.bytes = 0;
    return fail(LIBERI_ENOMEM, "device budget of the handle (LIBERI_OPT_MAX_DEVICE_BYTES = " +
                                   std::to_string((int64_t)h->max_device_bytes) + ") exceeded by an allocation of " +
                                   std::to_string(bytes) + " bytes");
  }
  hipError_t e = hipMalloc(&b.p, bytes);
  if (e != hipSuccess)
    return fail(e == hipErrorOutOfMemory ? LIBERI_ENOMEM : LIBERI_EDEVICE,
                "hipMalloc(" + std::to_string(bytes) + "): " + hipGetErrorString(e));
  h->stats.device_bytes += (int64_t)bytes;
  return LIBERI_OK;
}

template <class T>
static int dev_upload(liberi_ctx *h, DevBuf &b, const std::vector<T> &v) {
  int rc = dev_alloc(h, b, v.size() * sizeof(T));
  if (rc) return rc;
  if (!v.empty()) HIP_TRY(hipMemcpy(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  return LIBERI_OK;
}

static void dev_free(DevBuf &b) {
  if (b.p) (void)hipFree(b.p);
  b.p = nullptr;
  b.bytes = 0;
}

static void drop_graph(liberi_ctx *h) {
  if (h->gexec) (void)hipGraphExecDestroy(h->gexec);
  if (h->graph) (void)hipGraphDestroy(h->graph);
  h->gexec = nullptr;
  h->graph = nullptr;
  h->g_dD = nullptr;
  h->g_dF = nullptr;
}

static void release_device(liberi_ctx *h) {
  drop_graph(h);
  for (auto &c : h->cls) {
    dev_free(c.ket_list);
    dev_free(c.seg_start);
    dev_free(c.seg_bra);
    dev_free(c.seg_ket0);
    for (int k = 0; k < kNumTileLens; ++k) dev_free(c.tile_q[k]), dev_free(c.tile_seg[k]);
    c = ClassState{};
  }
  dev_free(h->pairs);
  dev_free(h->prims);
  dev_free(h->dens);
  dev_free(h->fock);
  dev_free(h->scratch);
  dev_free(h->aux);
  dev_free(h->fused_items);
  dev_free(h->fused_cds);
  dev_free(h->fused_head);
  h->fused_n = 0;
  dev_free(h->tune_F);
  dev_free(h->dens_t);
  dev_free(h->dens_b);
  dev_free(h->fock_b);
  for (auto &b : h->col_bf) dev_free(b);
  for (auto &n : h->ncol) n = 0;
  h->stats = liberi_stats{};
  h->initialized = false;
}

extern "C" {

const char *liberi_c_last_error(void) { return g_last_error.c_str(); }

const char *liberi_c_class_name(int32_t c) { return (c >= 0 && c < kNumClasses) ? kClasses[c].name : ""; }

// Kernel attributes of every class, once per device in the process, at
// handle creation (never inside a stream capture)
static int prepare_kernels(int device) {
  static std::mutex mu;
  static std::vector<int> done;
  std::lock_guard<std::mutex> lk(mu);
  if (std::find(done.begin(), done.end(), device) != done.end()) return LIBERI_OK;
  for (int c = 0; c < kNumClasses; ++c)
    if (int e = prepare_class(c))
      return fail(LIBERI_EDEVICE, std::string("kernel attributes of ") + kClasses[c].name + ": " +
                                      hipGetErrorString((hipError_t)e));
  done.push_back(device);
  return LIBERI_OK;
}

int liberi_c_create(liberi_ctx **ctx) {
  if (!ctx) return fail(LIBERI_EARG, "liberi_c_create: NULL handle pointer");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
    return fail(LIBERI_EDEVICE, "liberi_c_create: no HIP device visible");
  auto *h = new (std::nothrow) liberi_ctx();
  if (!h) return fail(LIBERI_ENOMEM, "liberi_c_create: out of host memory");
  HIP_TRY(hipGetDevice(&h->device));
  hipDeviceProp_t prop;
  HIP_TRY(hipGetDeviceProperties(&prop, h->device));
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
    delete h;
    return fail(LIBERI_EDEVICE, std::string("liberi_c_create: device is ") + prop.gcnArchName +
                                    ", this library is built for gfx950 (MI355X)");
  }
  if (int rc = prepare_kernels(h->device)) {
    delete h;
    return rc;
  }
  HIP_TRY(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
  HIP_TRY(hipEventCreate(&h->ev0));
  HIP_TRY(hipEventCreate(&h->ev1));
  *ctx = h;
  return LIBERI_OK;
}

int liberi_c_cleanup(liberi_ctx *h) {
  if (!h) return fail(LIBERI_EARG, "liberi_c_cleanup: NULL handle");
  // not set up: nothing to do (src/interface/liberi_interface.F90:235), but
  // device data of a setup that failed part of the way is released as well
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  release_device(h);
  return LIBERI_OK;
}

int liberi_c_destroy(liberi_ctx **ctx) {
  if (!ctx) return fail(LIBERI_EARG, "liberi_c_destroy: NULL handle pointer");
  liberi_ctx *h = *ctx;
  if (!h) return LIBERI_OK;
  liberi_c_cleanup(h);
  if (h->ev0) (void)hipEventDestroy(h->ev0);
  if (h->ev1) (void)hipEventDestroy(h->ev1);
  for (hipEvent_t e : h->tev)
    if (e) (void)hipEventDestroy(e);
  for (hipEvent_t e : h->join_ev)
    if (e) (void)hipEventDestroy(e);
  if (h->fork_ev) (void)hipEventDestroy(h->fork_ev);
  for (hipStream_t s2 : h->side)
    if (s2) (void)hipStreamDestroy(s2);
  for (int c = 0; c < kNumClasses; ++c) {
    if (h->cev0[c]) (void)hipEventDestroy(h->cev0[c]);
    if (h->cev1[c]) (void)hipEventDestroy(h->cev1[c]);
  }
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
  *ctx = nullptr;
  return LIBERI_OK;
}

int liberi_c_set_option(liberi_ctx *h, int32_t opt, double v) {
  if (!h) return fail(LIBERI_EARG, "liberi_c_set_option: NULL handle");
  drop_graph(h);  // a captured build holds the options it was captured with
  if (opt == LIBERI_OPT_ONLY_CLASS) {
    if (!(v >= -1.0 && v < kNumClasses) || v != std::floor(v))
      return fail(LIBERI_EARG, "liberi_c_set_option: ONLY_CLASS must be -1 or a class index 0..54");
    h->only_class = (int)v;
    return LIBERI_OK;
  }
  if (!(v >= 0.0) || !std::isfinite(v)) return fail(LIBERI_EARG, "liberi_c_set_option: value must be finite, >= 0");
  switch (opt) {
    case LIBERI_OPT_CLASS_TIMING: h->class_timing = v != 0.0; break;
    case LIBERI_OPT_SCHWARZ_CUTOFF: h->schwarz_cut = v; break;
    case LIBERI_OPT_VALUE_CUTOFF: h->value_cut = v; break;
    case LIBERI_OPT_FP32_TAU:
      if (v != h->fp32_tau)
        for (auto &c : h->cls) c.fp32 = -1;  // chosen again at the next build
      h->fp32_tau = v;
      break;
    case LIBERI_OPT_EXACT_LARGE_T: h->exact_large_t = v != 0.0; break;
    case LIBERI_OPT_PRIM_CUTOFF: h->prim_cut = v; break;
    case LIBERI_OPT_ROW_TILES:
      if (v != 0.0 && v != 1.0 && v != 2.0) return fail(LIBERI_EARG, "liberi_c_set_option: ROW_TILES must be 0, 1 or 2");
      h->row_tiles = (int)v;
      break;
    case LIBERI_OPT_GRAPH: h->use_graph = v != 0.0; break;
    case LIBERI_OPT_MAX_DEVICE_BYTES: h->max_device_bytes = v; break;
    case LIBERI_OPT_RANK_SPLIT:
      if (v != 0.0 && v != 1.0) return fail(LIBERI_EARG, "liberi_c_set_option: RANK_SPLIT must be 0 or 1");
      h->rank_split = (int)v;
      break;
    case LIBERI_OPT_FUSED: h->fused = v != 0.0; break;
    case LIBERI_OPT_ROW_ORIENT:
      if (v != 0.0 && v != 1.0 && v != 2.0) return fail(LIBERI_EARG, "liberi_c_set_option: ROW_ORIENT must be 0, 1 or 2");
      h->row_orient = (int)v;
      break;
    case LIBERI_OPT_TILE_LEN: {
      int k = -1;
      for (int i = 0; i < kNumTileLens; ++i)
        if (v == (double)kTileLens[i]) k = i;
      if (v != 0.0 && k < 0)
        return fail(LIBERI_EARG, "liberi_c_set_option: TILE_LEN must be 0, 512, 1024, 4096 or 16384");
      h->tile_len = v == 0.0 ? -1 : k;
      for (auto &c : h->cls) c.tile = -1;  // chosen again at the next build
      break;
    }
    case LIBERI_OPT_AUTOTUNE:
      h->autotune = v != 0.0;
      // decide (and time) again at the next build, without earlier handles' choices
      for (auto &c : h->cls) c.variant = -1, c.fp32 = -1, c.tile = -1, c.ms = -1.f;
      forget_tuned_choices();
      break;
    default: return fail(LIBERI_EARG, "liberi_c_set_option: unknown option " + std::to_string(opt));
  }
  return LIBERI_OK;
}

int liberi_c_get_option(liberi_ctx *h, int32_t opt, double *v) {
  if (!h || !v) return fail(LIBERI_EARG, "liberi_c_get_option: NULL argument");
  switch (opt) {
    case LIBERI_OPT_SCHWARZ_CUTOFF: *v = h->schwarz_cut; break;
    case LIBERI_OPT_VALUE_CUTOFF: *v = h->value_cut; break;
    case LIBERI_OPT_FP32_TAU: *v = h->fp32_tau; break;
    case LIBERI_OPT_EXACT_LARGE_T: *v = h->exact_large_t; break;
    case LIBERI_OPT_PRIM_CUTOFF: *v = h->prim_cut; break;
    case LIBERI_OPT_CLASS_TIMING: *v = h->class_timing; break;
    case LIBERI_OPT_ONLY_CLASS: *v = h->only_class; break;
    case LIBERI_OPT_ROW_TILES: *v = h->row_tiles; break;
    case LIBERI_OPT_AUTOTUNE: *v = h->autotune; break;
    case LIBERI_OPT_GRAPH: *v = h->use_graph; break;
    case LIBERI_OPT_MAX_DEVICE_BYTES: *v = h->max_device_bytes; break;
    case LIBERI_OPT_RANK_SPLIT: *v = h->rank_split; break;
    case LIBERI_OPT_FUSED: *v = h->fused; break;
    case LIBERI_OPT_ROW_ORIENT: *v = h->row_orient; break;
    case LIBERI_OPT_TILE_LEN: *v = h->tile_len < 0 ? 0.0 : (double)kTileLens[h->tile_len]; break;
    default: return fail(LIBERI_EARG, "liberi_c_get_option: unknown option " + std::to_string(opt));
  }
  return LIBERI_OK;
}

int liberi_c_get_stats(liberi_ctx *h, liberi_stats *out) {
  if (!h || !out) return fail(LIBERI_EARG, "liberi_c_get_stats: NULL argument");
  collect_timing(h);
  *out = h->stats;
  for (int c = 0; c < kNumClasses; ++c) {
    const ClassState &cs_ = h->cls[c];
    const bool row = cs_.row_ok && (h->row_tiles == 2 || (h->row_tiles == 1 && cs_.variant == 1));
    out->class_variant[c] = !cs_.active || cs_.variant < 0       ? -1
                            : h->last_fused && class_fused_nq(c) > 0 ? 4
                            : row                                 ? 1 + cs_.dev.transposed
                            : cs_.variant == 3 || class_kernel_kind(c) == 2 ? 3
                                                                  : 0;
    out->class_fp32[c] = cs_.active && h->fp32_tau > 0.0 && cs_.fp32 == 1;
    out->class_tile[c] = cs_.active && row && cs_.tile >= 0 && !(h->last_fused && class_fused_nq(c) > 0)
                             ? (int32_t)kTileLens[cs_.tile]
                             : 0;
  }
  return LIBERI_OK;
}

// ------------------------------------------------------------------ setup
// Where the reference stops using its tabulated Boys function / Rys roots.
// Rotated-axis and ERIC kernels (classes 0-13 and 22-25) take the asymptotic
// Boys chain F_m = F_{m-1} (2m-1)/2T for T > t_max = 25
// (src/liberi_boys.F90:22, src/rhf/rot_axis/int1111.F90:389-396,
// src/rhf/eric/int0030_ericgen.F90:300-311); the Rys kernels take the
// asymptotic roots for X >= 37, 43, 49, 55, 60, 65 with 2..7 roots
// (e.g. src/rhf/rot_axis/int2222_rysgen.F90:642).  Switching the Rys rule at
// the same points reproduces those values.
static double reference_switch(int c) {
  if (c <= 13 || (c >= 22 && c <= 25)) return std::nextafter(25.0, INFINITY);  // table iff T <= 25
  static const double thr[8] = {0, 0, 37.0, 43.0, 49.0, 55.0, 60.0, 65.0};       // table iff X < thr
  return thr[class_nroots(c)];
}

struct HostShell {
  int l, k, bf0;
  double c[3];
  const double *e, *w;
};

// Shells from the reference's setup arrays (src/liberi_shell_pair.F90:108-135),
// with the argument checks the reference leaves to the caller.
static int read_shells(const char *who, int32_t nsh, int32_t natoms, int32_t num_bas, int32_t mxgtot,
                       const int32_t *ang_mom, const int32_t *contr_num, const int32_t *sh_loc,
                       const int32_t *atom_num, const int32_t *atom_loc, const double *exponents, const double *cs,
                       const double *cp, const double *cdc, const double *cf, const double *coords,
                       std::vector<HostShell> &sh) {
  const std::string w0 = std::string(who) + ": ";
  if (nsh <= 0 || natoms <= 0 || num_bas <= 0 || mxgtot <= 0)
    return fail(LIBERI_EARG, w0 + "nsh, natoms, num_bas, mxgtot must be positive");
  // the packed matrices hold num_bas(num_bas+1)/2 elements, a default
  // integer in the reference's interface (liberi_interface.F90:202-217);
  // the kernels index them in 32 bits
  if (num_bas > kMaxBasis)
    return fail(LIBERI_EARG, w0 + "num_bas above " + std::to_string(kMaxBasis) + " (packed size exceeds 32 bits)");
  if (!ang_mom || !contr_num || !sh_loc || !atom_num || !atom_loc || !exponents || !cs || !cp || !cdc || !cf ||
      !coords)
    return fail(LIBERI_EARG, w0 + "NULL array");
  sh.assign(nsh, HostShell{});
  const int64_t coef_len = (int64_t)mxgtot * nsh;
  for (int s = 0; s < nsh; ++s) {
    const int l = ang_mom[s] - 1, k = contr_num[s], loc = sh_loc[s] - 1, a = atom_num[s] - 1;
    if (l < 0 || l > 3) return fail(LIBERI_EARG, w0 + "ang_mom must be 1..4 (shell " + std::to_string(s + 1) + ")");
    if (k < 1 || loc < 0 || loc + k > coef_len)
      return fail(LIBERI_EARG, w0 + "contr_num/sh_loc out of range (shell " + std::to_string(s + 1) + ")");
    if (a < 0 || a >= natoms) return fail(LIBERI_EARG, w0 + "atom_num out of range");
    const int bf0 = atom_loc[s] - 1;
    if (bf0 < 0 || bf0 + (l + 1) * (l + 2) / 2 > num_bas) return fail(LIBERI_EARG, w0 + "atom_loc out of range");
    const double *w = l == 0 ? cs : l == 1 ? cp : l == 2 ? cdc : cf;
    for (int i = 0; i < k; ++i)
      if (!(exponents[loc + i] > 0.0) || !std::isfinite(exponents[loc + i]) || !std::isfinite(w[loc + i]))
        return fail(LIBERI_EARG, w0 + "exponents must be finite and > 0, coefficients finite (shell " +
                                     std::to_string(s + 1) + ")");
    sh[s] = HostShell{l, k, bf0, {coords[3 * a], coords[3 * a + 1], coords[3 * a + 2]}, exponents + loc, w + loc};
  }
  return LIBERI_OK;
}

static int ncart_host(int l) { return (l + 1) * (l + 2) / 2; }

// Shell pair (sA, sB) and its primitive pairs appended to the device records
// (the per-primitive-pair quantities of src/liberi_shell_pair.F90:416-459);
// rank[s] = position of shell s among the shells of its angular momentum.
static void append_pair(std::vector<PairRec> &pairs, std::vector<PrimRec> &prims, const std::vector<HostShell> &sh,
                        const std::vector<int> *rank, int sA, int sB) {
  const HostShell &A = sh[sA], &B = sh[sB];
  PairRec r{};
  r.prim_off = (int32_t)prims.size();
  r.nprim = A.k * B.k;
  r.bfA = A.bf0;
  r.bfB = B.bf0;
  r.ABx = A.c[0] - B.c[0];
  r.ABy = A.c[1] - B.c[1];
  r.ABz = A.c[2] - B.c[2];
  r.sameAB = sA == sB;
  r.shA = sA;
  r.cA = rank ? (*rank)[sA] * ncart_host(A.l) : 0;
  r.cB = rank ? (*rank)[sB] * ncart_host(B.l) : 0;
  const double ab2 = r.ABx * r.ABx + r.ABy * r.ABy + r.ABz * r.ABz;
  for (int i = 0; i < A.k; ++i)
    for (int j = 0; j < B.k; ++j) {
      const double a = A.e[i], b = B.e[j], p = a + b;
      PrimRec pr;
      pr.p = p;
      pr.Px = (a * A.c[0] + b * B.c[0]) / p;
      pr.Py = (a * A.c[1] + b * B.c[1]) / p;
      pr.Pz = (a * A.c[2] + b * B.c[2]) / p;
      pr.PAx = pr.Px - A.c[0];
      pr.PAy = pr.Py - A.c[1];
      pr.PAz = pr.Pz - A.c[2];
      pr.K = A.w[i] * B.w[j] * std::exp(-a * b / p * ab2) / p;
      prims.push_back(pr);
    }
  pairs.push_back(r);
}

// Work items of the fused group launch: every active class that has a body
// in it, in runs of kFusedRounds rounds of its quartets per item; the classes
// in order of their primitive quartets (largest first), the items of a class
// dealt out over its whole quartet range (item k takes run k x stride mod n,
// stride coprime with n) so that the workgroups running at the same time
// meet few common Fock rows.
constexpr int64_t kFusedRounds = 32;

static int setup_fused(liberi_ctx *h) {
  std::vector<int> cls;
  for (int c = 0; c < kNumClasses; ++c)
    if (h->cls[c].active && class_fused_nq(c) > 0 && h->cls[c].dev.q_end > h->cls[c].dev.q_begin) cls.push_back(c);
  h->fused_n = 0;
  if (cls.empty()) return LIBERI_OK;
  std::stable_sort(cls.begin(), cls.end(), [&](int a, int b) { return h->cls[a].nprimq > h->cls[b].nprimq; });
  std::vector<FusedItem> items;
  for (int c : cls) {
    const ClassDev &d = h->cls[c].dev;
    const int64_t nq = d.q_end - d.q_begin, run = kFusedRounds * class_fused_nq(c);
    const int64_t n = (nq + run - 1) / run, stride = chunk_stride(n);
    for (int64_t k = 0; k < n; ++k) {
      const int64_t j = k * stride % n;
      items.push_back(FusedItem{c, 0, j * run, std::min(nq, (j + 1) * run)});
    }
  }
  std::vector<ClassDev> cds(kNumClasses);
  for (int c = 0; c < kNumClasses; ++c) cds[c] = h->cls[c].dev;
  int rc;
  if ((rc = dev_upload(h, h->fused_items, items)) || (rc = dev_upload(h, h->fused_cds, cds)) ||
      (rc = dev_alloc(h, h->fused_head, sizeof(double))))
    return rc;
  int ncu = 0;
  if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, h->device) != hipSuccess || ncu <= 0) ncu = 256;
  h->fused_nwg = 12 * ncu;  // three one-wave workgroups per SIMD
  h->fused_n = (int64_t)items.size();
  return LIBERI_OK;
}

// Setup proper, on a handle whose device data is released; the caller
// releases whatever it leaves behind when it fails part of the way through.
static int setup_basis(liberi_ctx *h, int32_t nsh, int32_t natoms, int32_t num_bas, int32_t mxgtot,
                       const std::vector<HostShell> &sh, const double *schwrz, int32_t my_rank, int32_t num_procs) {
  h->nsh = nsh;
  h->natoms = natoms;
  h->num_bas = num_bas;
  h->mxgtot = mxgtot;
  h->rank = my_rank;
  h->nprocs = num_procs;

  // ---- shell lists per L, ascending index (src/liberi_shell_pair.F90:137-344)
  std::vector<int> lists[4];
  for (int s = 0; s < nsh; ++s) lists[sh[s].l].push_back(s);
  // compact columns: the functions of each angular momentum in shell order
  std::vector<int> rank(nsh);
  {
    std::vector<int32_t> cols[4];
    for (int l = 0; l < 4; ++l) {
      for (size_t k = 0; k < lists[l].size(); ++k) {
        rank[lists[l][k]] = (int)k;
        for (int m = 0; m < ncart_host(l); ++m) cols[l].push_back(sh[lists[l][k]].bf0 + m);
      }
      h->ncol[l] = (int32_t)cols[l].size();
      if (int rc1 = dev_upload(h, h->col_bf[l], cols[l])) return rc1;
    }
  }

  // ---- pair classes: records, Schwarz values (compute_*_pairs, :348-1682)
  std::vector<PairRec> pairs;
  std::vector<PrimRec> prims;
  std::vector<double> pair_x;
  std::vector<std::vector<int32_t>> class_pairs(kNumPairClasses);
  auto add_pair = [&](int pc, int sA, int sB) {
    const int hi = std::max(sA, sB), lo = std::min(sA, sB);
    pair_x.push_back(schwrz[(int64_t)hi * (hi + 1) / 2 + lo]);
    class_pairs[pc].push_back((int32_t)pairs.size());
    append_pair(pairs, prims, sh, &rank, sA, sB);
    pairs.back().x = pair_x.back();
  };
  for (int pc = 0; pc < kNumPairClasses; ++pc) {
    const int l1 = kPairL[pc][0], l2 = kPairL[pc][1];
    if (l1 == l2) {
      for (size_t ii = 0; ii < lists[l1].size(); ++ii)
        for (size_t jj = 0; jj <= ii; ++jj) add_pair(pc, lists[l1][ii], lists[l1][jj]);
    } else {
      for (int x : lists[l1])
        for (int y : lists[l2]) add_pair(pc, y, x);  // A = the higher-L shell
    }
  }
  if (prims.size() > (size_t)INT32_MAX || pairs.size() > (size_t)INT32_MAX)
    return fail(LIBERI_EARG, "liberi_c_setup: basis too large for 32-bit pair indices");

  // ---- primitive-pair screen.  Bound of one primitive pair's share of any
  // integral: B = 2pi^(5/2)^(1/2) |K| (1 + |P-A| + 1/sqrt(p))^(la+lb) (the
  // Rys 2-D factors are polynomials in P-A, P-Q and 1/2p of that degree, the
  // Boys values are <= 1); a primitive quartet contributes at most about
  // B_ab B_cd.  A primitive pair of shell pair ab only ever meets the
  // primitive pairs of ab's partners: the shell pairs cd with
  // x_ab x_cd >= cutoff (every quartet test compares that product), and every
  // pair of the other list in the classes without a quartet test.  With M_ab
  // the largest B among those partners, primitive pairs with B M_ab <
  // prim_cut change no integral by more than prim_cut per primitive quartet,
  // far below the value cutoff, and are dropped.  Quartet lists and counts are
  // untouched (a pair may end up with no primitives and then contributes
  // nothing).
  if (h->prim_cut > 0.0) {
    const size_t np = pairs.size();
    std::vector<double> bound(prims.size()), pmax(np, 0.0);
    std::vector<int> pair_pc(np);
    for (int pc = 0; pc < kNumPairClasses; ++pc)
      for (int32_t p : class_pairs[pc]) pair_pc[p] = pc;
    for (size_t ip = 0; ip < np; ++ip) {
      const PairRec &r = pairs[ip];
      const int lsum = kPairL[pair_pc[ip]][0] + kPairL[pair_pc[ip]][1];
      for (int k = 0; k < r.nprim; ++k) {
        const PrimRec &q = prims[r.prim_off + k];
        const double pa = std::sqrt(q.PAx * q.PAx + q.PAy * q.PAy + q.PAz * q.PAz);
        const double b = std::sqrt(kTwoPi52) * std::fabs(q.K) * std::pow(1.0 + pa + 1.0 / std::sqrt(q.p), lsum);
        bound[r.prim_off + k] = b;
        pmax[ip] = std::max(pmax[ip], b);
      }
    }
    // partners through the quartet test: pairs by Schwarz value descending,
    // running maximum of their bounds
    std::vector<int32_t> byx(np);
    for (size_t ip = 0; ip < np; ++ip) byx[ip] = (int32_t)ip;
    std::sort(byx.begin(), byx.end(), [&](int32_t a, int32_t b) { return pair_x[a] > pair_x[b]; });
    std::vector<double> xdesc(np), runmax(np);
    for (size_t k = 0; k < np; ++k) {
      xdesc[k] = pair_x[byx[k]];
      runmax[k] = std::max(k ? runmax[k - 1] : 0.0, pmax[byx[k]]);
    }
    // partners in the classes without a quartet test: the other list's pairs
    std::vector<double> class_max(kNumPairClasses, 0.0), untested(kNumPairClasses, 0.0);
    for (size_t ip = 0; ip < np; ++ip) class_max[pair_pc[ip]] = std::max(class_max[pair_pc[ip]], pmax[ip]);
    for (int c = 0; c < kNumClasses; ++c)
      if (kClasses[c].cmp == CMP_NONE) {
        untested[kClasses[c].bra] = std::max(untested[kClasses[c].bra], class_max[kClasses[c].ket]);
        untested[kClasses[c].ket] = std::max(untested[kClasses[c].ket], class_max[kClasses[c].bra]);
      }
    std::vector<PrimRec> kept;
    kept.reserve(prims.size());
    for (size_t ip = 0; ip < np; ++ip) {
      PairRec &r = pairs[ip];
      double m = untested[pair_pc[ip]];
      if (pair_x[ip] > 0.0) {
        // x_cd >= cutoff / x_ab, with a margin for the rounding of the quotient
        const double t = h->schwarz_cut / pair_x[ip] * (1.0 - 1e-12);
        const size_t n = (size_t)(std::upper_bound(xdesc.begin(), xdesc.end(), t, std::greater<double>()) - xdesc.begin());
        if (n > 0) m = std::max(m, runmax[n - 1]);
      }
      const int32_t off = (int32_t)kept.size();
      for (int k = 0; k < r.nprim; ++k)
        if (bound[r.prim_off + k] * m >= h->prim_cut) kept.push_back(prims[r.prim_off + k]);
      r.prim_off = off;
      r.nprim = (int32_t)kept.size() - off;
    }
    h->stats.prim_pairs_dropped = (int64_t)(prims.size() - kept.size());
    prims.swap(kept);
  }
  h->stats.prim_pairs = (int64_t)prims.size();

  int rc;
  if ((rc = dev_upload(h, h->pairs, pairs))) return rc;
  if ((rc = dev_upload(h, h->prims, prims))) return rc;
  const int64_t ntri = (int64_t)num_bas * (num_bas + 1) / 2;
  if ((rc = dev_alloc(h, h->dens, ntri * sizeof(double)))) return rc;
  if ((rc = dev_alloc(h, h->aux, 2 * sizeof(double)))) return rc;
  if (int e = launch_fill_zero((double *)h->aux.p, 2, h->stream))
    return fail(LIBERI_EDEVICE, std::string("zero aux: ") + hipGetErrorString((hipError_t)e));
  if ((rc = dev_alloc(h, h->fock, ntri * sizeof(double)))) return rc;

  // ---- per class: pair pre-screen (host), quartet segments (device)
  std::vector<DevBuf> tmp;
  auto cleanup_tmp = [&]() {
    for (auto &b : tmp) dev_free(b);
    tmp.clear();
  };
  int64_t max_scratch = 0;
  for (int c = 0; c < kNumClasses; ++c) {
    const ClassDef &def = kClasses[c];
    ClassState &cs_ = h->cls[c];
    const auto &lb = class_pairs[def.bra], &lk = class_pairs[def.ket];
    int lmax = std::max(std::max(kPairL[def.bra][0], kPairL[def.bra][1]), std::max(kPairL[def.ket][0], kPairL[def.ket][1]));
    // the driver calls the class only if its highest shell type exists
    // (src/liberi_driver.F90:69,72,80,105); an empty pair list has no quartets
    if (lists[lmax].empty() || lb.empty() || lk.empty()) continue;
    auto maxval = [&](const std::vector<int32_t> &l) {
      double m = -INFINITY;
      for (int32_t p : l) m = std::max(m, pair_x[p]);
      return m;
    };
    const double mb = maxval(lb), mk = maxval(lk);
    const double cut = h->schwarz_cut;
    const double sb = def.ps == PS_CROSS ? cut / mk : cut / mb;
    const double sk = def.ps == PS_CROSS ? cut / mb : cut / mk;
    const bool diag = def.bra == def.ket;
    auto group_key = [&](int32_t p) { return (int64_t)pairs[p].nprim; };
    auto keep = [&](const std::vector<int32_t> &l, double thr) {
      std::vector<int32_t> out;
      for (int32_t p : l)
        if (pair_x[p] >= thr) out.push_back(p);
      // sort by (group key, Schwarz value descending, pair index)
      std::sort(out.begin(), out.end(), [&](int32_t a, int32_t b) {
        if (group_key(a) != group_key(b)) return group_key(a) < group_key(b);
        if (pair_x[a] != pair_x[b]) return pair_x[a] > pair_x[b];
        return a < b;
      });
      return out;
    };
    std::vector<int32_t> kb = keep(lb, sb);
    std::vector<int32_t> kk = diag ? kb : keep(lk, sk);
    if (kb.empty() || kk.empty()) continue;
    // Orientation of the row tiles: the fixed pair of a tile keeps its
    // exchange blocks on chip, the other pair's Coulomb block goes to F per
    // quartet.  Fix the pair whose Coulomb block is larger, among the
    // orientations whose accumulators fit in LDS.  The quartet set (and the
    // count) is the same either way: the test is on the product x_bra*x_ket.
    bool transposed = false;
    {
      const int nbra = ncart_host(kPairL[def.bra][0]) * ncart_host(kPairL[def.bra][1]);
      const int nket = ncart_host(kPairL[def.ket][0]) * ncart_host(kPairL[def.ket][1]);
      const double xsw = h->exact_large_t ? class_rys_xmax(c) : reference_switch(c);
      const int64_t l0 = class_row_lds(c, 0, h->ncol, xsw), l1 = diag ? -1 : class_row_lds(c, 1, h->ncol, xsw);
      const bool fit0 = l0 >= 0, fit1 = l1 >= 0;  // (-1: no row form, or its LDS would not fit)
      transposed = fit1 && (!fit0 || nket > nbra);
      if (h->row_orient == 1 && fit0) transposed = false;
      if (h->row_orient == 2 && fit1) transposed = true;
      cs_.row_ok = fit0 || fit1;
    }
    if (transposed) kb.swap(kk);
    // ket groups (equal group key)
    std::vector<int32_t> gbeg, gend, kgroup(kk.size());
    for (size_t j = 0; j < kk.size(); ++j) {
      if (j == 0 || group_key(kk[j]) != group_key(kk[j - 1])) {
        if (j) gend.push_back((int32_t)j);
        gbeg.push_back((int32_t)j);
      }
      kgroup[j] = (int32_t)gbeg.size() - 1;
    }
    gend.push_back((int32_t)kk.size());
    const int ng = (int)gbeg.size();
    std::vector<double> bx(kb.size()), kx(kk.size());
    std::vector<int32_t> bgroup(kb.size()), bpos(kb.size());
    for (size_t i = 0; i < kb.size(); ++i) {
      bx[i] = pair_x[kb[i]];
      bgroup[i] = diag ? kgroup[i] : 0;
      bpos[i] = (int32_t)i;
    }
    for (size_t j = 0; j < kk.size(); ++j) kx[j] = pair_x[kk[j]];
    DevBuf d_bx, d_bg, d_bp, d_kx, d_gb, d_ge, d_len;
    if ((rc = dev_upload(h, d_bx, bx)) || (rc = dev_upload(h, d_bg, bgroup)) || (rc = dev_upload(h, d_bp, bpos)) ||
        (rc = dev_upload(h, d_kx, kx)) || (rc = dev_upload(h, d_gb, gbeg)) || (rc = dev_upload(h, d_ge, gend))) {
      cleanup_tmp();
      return rc;
    }
    tmp = {d_bx, d_bg, d_bp, d_kx, d_gb, d_ge};
    const int64_t nsegc = (int64_t)kb.size() * ng;
    if ((rc = dev_alloc(h, d_len, nsegc * sizeof(int64_t)))) {
      cleanup_tmp();
      return rc;
    }
    tmp.push_back(d_len);
    int e = launch_segment_lengths((const double *)d_bx.p, (const int32_t *)d_bg.p, (const int32_t *)d_bp.p,
                                   (int32_t)kb.size(), (const double *)d_kx.p, (const int32_t *)d_gb.p,
                                   (const int32_t *)d_ge.p, ng, diag, def.cmp, cut, (int64_t *)d_len.p, h->stream);
    std::vector<int64_t> len(nsegc);
    if (e == hipSuccess)
      e = hipMemcpyAsync(len.data(), d_len.p, nsegc * sizeof(int64_t), hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    for (auto &b : tmp) h->stats.device_bytes -= (int64_t)b.bytes;
    cleanup_tmp();
    if (e != hipSuccess)
      return fail(LIBERI_EDEVICE, std::string("segment generation: ") + hipGetErrorString((hipError_t)e));
    std::vector<int64_t> sstart, tq[kNumTileLens];
    std::vector<int32_t> sbra, sket0, tseg[kNumTileLens];
    int64_t total = 0, primq = 0;
    for (size_t i = 0; i < kb.size(); ++i) {
      const int64_t row0 = total;
      const size_t sfirst = sstart.size();
      for (int g = 0; g < ng; ++g) {
        const int64_t L = len[i * ng + g];
        if (L <= 0) continue;
        sstart.push_back(total);
        sbra.push_back(kb[i]);
        sket0.push_back(gbeg[g]);
        total += L;
        primq += L * (int64_t)pairs[kb[i]].nprim * pairs[kk[gbeg[g]]].nprim;
      }
      // row tiles: the row's quartets in runs of at most kTileLens[k]
      for (int k = 0; k < kNumTileLens; ++k) {
        size_t sg = sfirst;
        for (int64_t t0 = row0; t0 < total; t0 += kTileLens[k]) {
          while (sg + 1 < sstart.size() && sstart[sg + 1] <= t0) ++sg;
          tq[k].push_back(t0);
          tseg[k].push_back((int32_t)sg);
        }
      }
    }
    sstart.push_back(total);
    for (auto &t : tq) t.push_back(total);
    cs_.nquart = total;
    cs_.nprimq = primq;
    if (total == 0) continue;
    if (sbra.size() >= (size_t)INT32_MAX || tq[0].size() >= (size_t)INT32_MAX)
      return fail(LIBERI_ESIZE, std::string("class ") + def.name + ": too many segments");
    cs_.nseg = (int32_t)sbra.size();
    if ((rc = dev_upload(h, cs_.ket_list, kk)) || (rc = dev_upload(h, cs_.seg_start, sstart)) ||
        (rc = dev_upload(h, cs_.seg_bra, sbra)) || (rc = dev_upload(h, cs_.seg_ket0, sket0)))
      return rc;
    for (int k = 0; k < kNumTileLens; ++k)
      if ((rc = dev_upload(h, cs_.tile_q[k], tq[k])) || (rc = dev_upload(h, cs_.tile_seg[k], tseg[k]))) return rc;
    cs_.active = true;
    ClassDev &d = cs_.dev;
    d.bra_list = nullptr;
    d.ket_list = (const int32_t *)cs_.ket_list.p;
    d.seg_start = (const int64_t *)cs_.seg_start.p;
    d.seg_bra = (const int32_t *)cs_.seg_bra.p;
    d.seg_ket0 = (const int32_t *)cs_.seg_ket0.p;
    d.nseg = cs_.nseg;
    d.diag = diag;
    // this rank's contiguous share of the class.  The reference splits each
    // class's index range into equal counts (e.g.
    // src/rhf/rot_axis/int0001.F90:100-104; LIBERI_OPT_RANK_SPLIT = 1);
    // by default the shares are equal in estimated cost instead: a quartet
    // costs its primitive quartets plus kQuartetCost for its HRR and
    // digestion, and the kept lists run from few to many primitives, so equal
    // counts leave the last ranks the most work (C60 in 8 shares, time of the
    // slowest over the mean: equal counts 1.39; kQuartetCost 4: 1.26,
    // 8: 1.08, 16: 1.08, 32: 1.16).
    if (h->rank_split == 1 || num_procs == 1) {
      d.q_begin = total * my_rank / num_procs;
      d.q_end = total * (my_rank + 1) / num_procs;
    } else {
      constexpr double kQuartetCost = 12.0;
      const size_t ns = sbra.size();
      std::vector<double> cum(ns + 1, 0.0), unit(ns);
      for (size_t k = 0; k < ns; ++k) {
        unit[k] = (double)pairs[sbra[k]].nprim * pairs[kk[sket0[k]]].nprim + kQuartetCost;
        cum[k + 1] = cum[k] + unit[k] * (double)(sstart[k + 1] - sstart[k]);
      }
      // first quartet at or past cost b
      auto q_at = [&](double b) -> int64_t {
        if (b <= 0.0) return 0;
        if (b >= cum[ns]) return total;
        const size_t k = (size_t)(std::upper_bound(cum.begin(), cum.end(), b) - cum.begin()) - 1;
        const int64_t q = sstart[k] + (int64_t)std::ceil((b - cum[k]) / unit[k]);
        return std::min(q, sstart[k + 1]);
      };
      d.q_begin = q_at(cum[ns] * my_rank / num_procs);
      d.q_end = my_rank + 1 == num_procs ? total : q_at(cum[ns] * (my_rank + 1) / num_procs);
    }
    // this share's primitive quartets
    {
      int64_t pq = 0;
      for (size_t k = 0; k < sbra.size(); ++k) {
        const int64_t a = std::max(sstart[k], d.q_begin), b = std::min(sstart[k + 1], d.q_end);
        if (b > a) pq += (b - a) * (int64_t)pairs[sbra[k]].nprim * pairs[kk[sket0[k]]].nprim;
      }
      primq = pq;
    }
    // tiles overlapping [q_begin, q_end): tq is strictly increasing, tq.back() = total
    for (int k = 0; k < kNumTileLens; ++k) {
      const std::vector<int64_t> &t = tq[k];
      const int32_t ntile = (int32_t)t.size() - 1;
      int32_t b = (int32_t)(std::upper_bound(t.begin(), t.end(), d.q_begin) - t.begin()) - 1;
      int32_t e = (int32_t)(std::lower_bound(t.begin(), t.end(), d.q_end) - t.begin());
      cs_.t_begin[k] = b = std::max<int32_t>(0, std::min(b, ntile));
      cs_.t_end[k] = std::max(b, std::min(e, ntile));
    }
    use_tiles(cs_, kDefaultTile);
    d.row_mode = 0;
    d.transposed = transposed;
    d.xswitch = h->exact_large_t ? class_rys_xmax(c) : reference_switch(c);
    h->stats.quartets[c] = d.q_end - d.q_begin;
    h->stats.prim_quartets[c] = primq;
    h->stats.total_quartets += d.q_end - d.q_begin;
    max_scratch = std::max(max_scratch, class_scratch_doubles(c));
  }
  h->scratch_per_wg = max_scratch;
  if (max_scratch > 0)
    if ((rc = dev_alloc(h, h->scratch, (size_t)max_scratch * kCoopWorkgroups * sizeof(double)))) return rc;
  if ((rc = setup_fused(h))) return rc;
  HIP_TRY(hipStreamSynchronize(h->stream));
  h->initialized = true;
  return LIBERI_OK;
}

int liberi_c_setup(liberi_ctx *h, int32_t nsh, int32_t natoms, int32_t num_bas, int32_t mxgtot,
                   const int32_t *ang_mom, const int32_t *contr_num, const int32_t *sh_loc, const int32_t *atom_num,
                   const int32_t *atom_loc, const int32_t *start_bas, const int32_t *end_bas,
                   const double *exponents, const double *cs, const double *cp, const double *cdc, const double *cf,
                   const double *coords, const double *schwrz, int32_t my_rank, int32_t num_procs) {
  (void)start_bas;  // stored but never read by the reference's hot path either
  (void)end_bas;
  if (!h) return fail(LIBERI_EARG, "liberi_c_setup: NULL handle");
  if (!schwrz) return fail(LIBERI_EARG, "liberi_c_setup: NULL array");
  if (num_procs < 1 || my_rank < 0 || my_rank >= num_procs)
    return fail(LIBERI_EARG, "liberi_c_setup: need 0 <= my_rank < num_procs");
  std::vector<HostShell> sh;
  if (int rc0 = read_shells("liberi_c_setup", nsh, natoms, num_bas, mxgtot, ang_mom, contr_num, sh_loc, atom_num,
                            atom_loc, exponents, cs, cp, cdc, cf, coords, sh))
    return rc0;
  // a set-up handle is cleaned up first (src/interface/liberi_interface.F90:133-136);
  // so is whatever an earlier setup that failed part of the way left behind
  liberi_c_cleanup(h);
  const int rc = setup_basis(h, nsh, natoms, num_bas, mxgtot, sh, schwrz, my_rank, num_procs);
  if (rc != LIBERI_OK) {
    const std::string msg = g_last_error;
    (void)hipStreamSynchronize(h->stream);
    release_device(h);
    g_last_error = msg;
  }
  return rc;
}

// ------------------------------------------------------------------ build
// Autotuner: for every class that has both a direct and a row-tiled kernel,
// time each on this build's density (into a scratch Fock matrix, best of two
// runs each) and keep the faster.  Runs once per setup, at the first build;
// classes too small to time reliably, and every class when tuning is off, take
// the default: row tiles for the lane-per-quartet classes (whose cost is the
// atomics), direct digestion for the workgroup classes (whose occupancy the
// LDS accumulators would cut).
constexpr int64_t kTuneMinQuartets = 20000;
constexpr int kSideStreams = 4;

// Choices made so far in this process: later handles (a request loop sets up
// one per molecule) reuse them instead of timing again, but only for a class
// of about the same size -- the key is the class and its quartet count to
// within a factor of 8 -- and, for the FP32 choice, the same switch tau.
static std::mutex g_tune_mu;
static std::map<std::tuple<int, int, double>, int> g_tuned;  // (class, size, tau or -1) -> choice

// classes within a factor of 8 in quartet count share a choice
static int size_bucket(int64_t nq) {
  int b = 0;
  while (nq > 1) nq >>= 1, ++b;
  return b / 3;
}

static int tuned_choice(int c, int64_t nq, double tau) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  auto it = g_tuned.find({c, size_bucket(nq), tau});
  return it == g_tuned.end() ? -1 : it->second;
}

static void set_tuned_choice(int c, int64_t nq, double tau, int v) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_tuned[{c, size_bucket(nq), tau}] = v;
}

static void forget_tuned_choices() {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_tuned.clear();
}

// Whether class c runs its row-tiled kernel in this build.
static bool class_row_mode(const liberi_ctx *h, int c) {
  const ClassState &cs_ = h->cls[c];
  return cs_.row_ok && cs_.dev.t_end > cs_.dev.t_begin &&
         (h->row_tiles == 2 || (h->row_tiles == 1 && cs_.variant == 1));
}

// One launch of class c (row-tiled or direct) into the scratch Fock matrix,
// timed with events on st.
static int time_class(liberi_ctx *h, const BuildArgs &ba, hipStream_t st, int c, bool row_mode, bool fp32,
                      float &ms) {
  const int64_t ntri = (int64_t)h->num_bas * (h->num_bas + 1) / 2;
  if (!h->tune_F.p)
    if (int rc = dev_alloc(h, h->tune_F, ntri * sizeof(double))) return rc;
  for (auto &e : h->tev)
    if (!e) HIP_TRY(hipEventCreate(&e));
  BuildArgs bt = ba;
  bt.F = (double *)h->tune_F.p;
  ClassDev cd = h->cls[c].dev;
  cd.row_mode = row_mode;
  cd.fp32 = fp32;
  HIP_TRY(hipEventRecord(h->tev[0], st));
  if (int e = launch_class(c, cd, bt, st, kCoopWorkgroups))
    return fail(LIBERI_EDEVICE, std::string("timing ") + kClasses[c].name + ": " + hipGetErrorString((hipError_t)e));
  HIP_TRY(hipEventRecord(h->tev[1], st));
  HIP_TRY(hipEventSynchronize(h->tev[1]));
  HIP_TRY(hipEventElapsedTime(&ms, h->tev[0], h->tev[1]));
  return LIBERI_OK;
}

static int choose_variants(liberi_ctx *h, const BuildArgs &ba, hipStream_t st) {
  for (int c = 0; c < kNumClasses; ++c) {
    ClassState &cs_ = h->cls[c];
    if (!cs_.active || cs_.variant >= 0) continue;
    const int64_t nq = cs_.dev.q_end - cs_.dev.q_begin;
    // untimed choice: row tiles for the lane classes, and for the group
    // classes with a wave-split form unless the class is too small to time
    const int dflt =
        (cs_.row_ok && (class_kernel_kind(c) == 0 || (class_has_split(c) && (!h->autotune || nq >= kTuneMinQuartets))))
            ? 1
            : 0;
    if (class_has_coop(c)) {
      // group class with a cooperative form (variant 3), and with a
      // wave-split row form (variant 1) where it has one: one launch of each
      // is cheap at any size, so they are timed without a size threshold
      const bool row = cs_.row_ok && cs_.dev.t_end > cs_.dev.t_begin && h->row_tiles != 0;
      cs_.variant = (dflt == 1 && h->row_tiles != 0) || h->row_tiles == 2 ? 1 : 0;
      cs_.dev.coop = 0;
      if (!h->autotune || h->row_tiles == 2 || (h->only_class >= 0 && c != h->only_class)) continue;
      if (const int prev = tuned_choice(c, nq, -3.0); prev >= 0 && (prev != 1 || row)) {
        cs_.variant = prev;
        cs_.dev.coop = prev == 3;
        continue;
      }
      const int forms[3] = {0, 3, 1};  // group, cooperative, wave-split rows
      float best[3] = {1e30f, 1e30f, 1e30f};
      for (int rep = 0; rep < 6; ++rep) {
        const int k = rep % 3;
        if (forms[k] == 1 && !row) continue;
        float ms = 0.f;
        cs_.dev.coop = forms[k] == 3;
        if (int rc = time_class(h, ba, st, c, forms[k] == 1, false, ms)) return rc;
        best[k] = std::min(best[k], ms);
      }
      const int kbest = best[1] < best[0] ? (best[2] < best[1] ? 2 : 1) : (best[2] < best[0] ? 2 : 0);
      cs_.variant = forms[kbest];
      cs_.dev.coop = cs_.variant == 3;
      cs_.ms = best[kbest];
      set_tuned_choice(c, nq, -3.0, cs_.variant);
      continue;
    }
    if (!h->autotune || h->row_tiles != 1 || !cs_.row_ok || nq < kTuneMinQuartets || cs_.dev.t_end <= cs_.dev.t_begin ||
        (h->only_class >= 0 && c != h->only_class)) {
      cs_.variant = dflt;
      continue;
    }
    if (const int prev = tuned_choice(c, nq, -1.0); prev >= 0) {
      cs_.variant = prev;
      continue;
    }
    float best[2] = {1e30f, 1e30f};
    for (int rep = 0; rep < 4; ++rep) {
      const int v = rep & 1;
      float ms = 0.f;
      if (int rc = time_class(h, ba, st, c, v == 1, false, ms)) return rc;
      best[v] = std::min(best[v], ms);
    }
    cs_.variant = best[1] < best[0] ? 1 : 0;
    cs_.ms = std::min(best[0], best[1]);
    set_tuned_choice(c, nq, -1.0, cs_.variant);
  }
  // row-tile length of the row-tiled classes: shorter tiles give more
  // workgroups and more flushes of the exchange rows, longer ones fewer
  // workgroups (rows longer than a tile are split)
  for (int c = 0; c < kNumClasses; ++c) {
    ClassState &cs_ = h->cls[c];
    if (!cs_.active || cs_.tile >= 0) continue;
    const int64_t nq = cs_.dev.q_end - cs_.dev.q_begin;
    cs_.tile = h->tile_len >= 0 ? h->tile_len : kDefaultTile;
    if (h->tile_len >= 0 || !h->autotune || !class_row_mode(h, c) || nq < kTuneMinQuartets ||
        (h->only_class >= 0 && c != h->only_class)) {
      use_tiles(cs_, cs_.tile);
      continue;
    }
    if (const int prev = tuned_choice(c, nq, -2.0); prev >= 0) {
      cs_.tile = prev;
      use_tiles(cs_, cs_.tile);
      continue;
    }
    float best[kNumTileLens];
    for (int k = 0; k < kNumTileLens; ++k) best[k] = 1e30f;
    for (int rep = 0; rep < 2; ++rep)
      for (int k = 0; k < kNumTileLens; ++k) {
        if (cs_.t_end[k] <= cs_.t_begin[k]) continue;
        use_tiles(cs_, k);
        float ms = 0.f;
        if (int rc = time_class(h, ba, st, c, true, false, ms)) return rc;
        best[k] = std::min(best[k], ms);
      }
    for (int k = 0; k < kNumTileLens; ++k)
      if (best[k] < best[cs_.tile]) cs_.tile = k;
    use_tiles(cs_, cs_.tile);
    cs_.ms = best[cs_.tile];
    set_tuned_choice(c, nq, -2.0, cs_.tile);
  }
  // mixed form: FP32 stays on only in the classes where it is measured
  // faster than FP64 by at least kFp32Gain (without packed FP32 an FP32
  // instruction costs what an FP64 one does, so most classes lose by the
  // conversions); the lane-per-quartet classes have an FP32 path
  constexpr float kFp32Gain = 0.02f;
  for (int c = 0; c < kNumClasses; ++c) {
    ClassState &cs_ = h->cls[c];
    if (!cs_.active || cs_.fp32 >= 0) continue;
    const int64_t nq = cs_.dev.q_end - cs_.dev.q_begin;
    if (!(h->fp32_tau > 0.0) || class_kernel_kind(c) != 0) {
      cs_.fp32 = h->fp32_tau > 0.0 ? 0 : -1;
      continue;
    }
    if (!h->autotune || nq < kTuneMinQuartets || (h->only_class >= 0 && c != h->only_class)) {
      cs_.fp32 = 1;  // untimed: the form as asked for
      continue;
    }
    if (const int prev = tuned_choice(c, nq, h->fp32_tau); prev >= 0) {
      cs_.fp32 = prev;
      continue;
    }
    float best[2] = {1e30f, 1e30f};
    for (int rep = 0; rep < 4; ++rep) {
      const int v = rep & 1;
      float ms = 0.f;
      if (int rc = time_class(h, ba, st, c, class_row_mode(h, c), v == 1, ms)) return rc;
      best[v] = std::min(best[v], ms);
    }
    cs_.fp32 = best[1] < (1.f - kFp32Gain) * best[0] ? 1 : 0;
    set_tuned_choice(c, nq, h->fp32_tau, cs_.fp32);
  }
  // one timed launch of every class not timed above: the measured times
  // order the classes over the side streams
  for (int c = 0; c < kNumClasses; ++c) {
    ClassState &cs_ = h->cls[c];
    if (!cs_.active || cs_.ms >= 0.f || (h->only_class >= 0 && c != h->only_class)) continue;
    if (int rc = time_class(h, ba, st, c, class_row_mode(h, c), cs_.fp32 == 1, cs_.ms)) return rc;
  }
  return LIBERI_OK;
}

// max |D| of the build into aux[0] and a zero FP32 quartet count into aux[1]
// (the mixed form's switch reads them)
static int enqueue_dmax(liberi_ctx *h, const double *dD, hipStream_t st) {
  const int64_t ntri = (int64_t)h->num_bas * (h->num_bas + 1) / 2;
  int e = launch_fill_zero((double *)h->aux.p, 2, st);
  if (e == hipSuccess) e = launch_abs_max(dD, ntri, (double *)h->aux.p, st);
  if (e != hipSuccess) return fail(LIBERI_EDEVICE, std::string("max |D|: ") + hipGetErrorString((hipError_t)e));
  return LIBERI_OK;
}

// Zero F (and find max |D| for the mixed form), then every class's kernel.
// With ns > 1 the classes are spread over the side streams, longest first
// onto the least-loaded stream, forked from and joined back into st.
static int enqueue_build(liberi_ctx *h, const BuildArgs &ba, const double *dD, double *dF, hipStream_t st, int ns) {
  const int64_t ntri = (int64_t)h->num_bas * (h->num_bas + 1) / 2;
  // fock = 0 (src/liberi_driver.F90:64)
  if (int e = launch_fill_zero(dF, ntri, st))
    return fail(LIBERI_EDEVICE, std::string("zero F: ") + hipGetErrorString((hipError_t)e));
  if (h->fp32_tau > 0.0)
    if (int rc = enqueue_dmax(h, dD, st)) return rc;
  // the fused group launch, where on: its classes leave the per-class launches
  const bool fused = h->fused && h->fused_n > 0 && !h->class_timing && h->only_class < 0;
  h->last_fused = fused;
  std::vector<int> order;
  double fused_cost = 0.0;
  for (int c = 0; c < kNumClasses; ++c)
    if (h->cls[c].active && (h->only_class < 0 || c == h->only_class)) {
      if (fused && class_fused_nq(c) > 0) continue;
      order.push_back(c);
    }
  std::vector<double> load(ns, 0.0);
  // cost: the class's measured time (first build), else a proxy scaled to
  // milliseconds: primitive quartets x Rys roots x sqrt(cartesian components)
  auto cost = [&](int c) {
    if (h->cls[c].ms >= 0.f) return (double)h->cls[c].ms;
    const int la = kPairL[kClasses[c].bra][1], lb = kPairL[kClasses[c].bra][0];
    const int lc = kPairL[kClasses[c].ket][1], ld = kPairL[kClasses[c].ket][0];
    const double ncomp = (double)ncart_host(la) * ncart_host(lb) * ncart_host(lc) * ncart_host(ld);
    return 1e-9 * (double)h->cls[c].nprimq * class_nroots(c) * std::sqrt(ncomp);
  };
  if (ns > 1) {
    std::sort(order.begin(), order.end(), [&](int a, int b) { return cost(a) > cost(b); });
    HIP_TRY(hipEventRecord(h->fork_ev, st));
    for (int k = 0; k < ns; ++k) HIP_TRY(hipStreamWaitEvent(h->side[k], h->fork_ev, 0));
  }
  if (fused) {
    // one persistent launch for the fused classes, first onto the least-loaded
    // stream (the cooperative classes keep side stream 0 to themselves)
    for (int c = 0; c < kNumClasses; ++c)
      if (h->cls[c].active && class_fused_nq(c) > 0) fused_cost += cost(c);
    int k = 0;
    if (ns > 1) {
      k = ns > 1 ? 1 : 0;
      load[k] += fused_cost;
    }
    hipStream_t fs = ns > 1 ? h->side[k] : st;
    int e = launch_fill_zero((double *)h->fused_head.p, 1, fs);
    if (e == hipSuccess)
      e = launch_group_fused((const ClassDev *)h->fused_cds.p, ba, (const FusedItem *)h->fused_items.p, h->fused_n,
                             (unsigned long long *)h->fused_head.p, h->fused_nwg, fs);
    if (e != hipSuccess)
      return fail(LIBERI_EDEVICE, std::string("fused group launch: ") + hipGetErrorString((hipError_t)e));
  }
  for (int c : order) {
    hipStream_t cs = st;
    if (ns > 1) {
      // the cooperative kernels share one per-workgroup scratch buffer: their
      // classes all go to side stream 0, which runs them one after another
      const int k = class_scratch_doubles(c) > 0 ? 0 : (int)(std::min_element(load.begin(), load.end()) - load.begin());
      load[k] += cost(c);
      cs = h->side[k];
    }
    if (h->class_timing && ns == 1) {
      if (!h->cev0[c]) HIP_TRY(hipEventCreate(&h->cev0[c]));
      if (!h->cev1[c]) HIP_TRY(hipEventCreate(&h->cev1[c]));
      HIP_TRY(hipEventRecord(h->cev0[c], cs));
    }
    ClassDev cd = h->cls[c].dev;
    cd.row_mode = class_row_mode(h, c);
    cd.fp32 = h->cls[c].fp32 == 1;
    int e = launch_class(c, cd, ba, cs, kCoopWorkgroups);
    if (e != hipSuccess)
      return fail(LIBERI_EDEVICE, std::string("launch ") + kClasses[c].name + ": " + hipGetErrorString((hipError_t)e));
    if (h->class_timing && ns == 1) HIP_TRY(hipEventRecord(h->cev1[c], cs));
  }
  if (ns > 1)
    for (int k = 0; k < ns; ++k) {
      HIP_TRY(hipEventRecord(h->join_ev[k], h->side[k]));
      HIP_TRY(hipStreamWaitEvent(st, h->join_ev[k], 0));
    }
  return LIBERI_OK;
}

// RHF: DK = D, k_scale = 1.  A UHF spin component: D = Da + Db, DK = that
// spin's density, k_scale = 2 (F_s = J[Da+Db] - K[Ds]).
static int build_on_device(liberi_ctx *h, const double *dD, double *dF, hipStream_t st, const double *dK = nullptr,
                           double k_scale = 1.0) {
  BuildArgs ba;
  ba.pairs = (const PairRec *)h->pairs.p;
  ba.prims = (const PrimRec *)h->prims.p;
  ba.D = dD;
  ba.DK = dK ? dK : dD;
  ba.k_scale = k_scale;
  ba.F = dF;
  ba.value_cut = h->value_cut;
  ba.scratch = (double *)h->scratch.p;
  ba.scratch_per_wg = h->scratch_per_wg;
  ba.nbf = h->num_bas;
  ba.fp32_tau = h->fp32_tau;
  ba.dmax = (const double *)h->aux.p;
  ba.fp32_count = (uint64_t *)h->aux.p + 1;
  for (int l = 0; l < 4; ++l) {
    ba.col_bf[l] = (const int32_t *)h->col_bf[l].p;
    ba.ncol[l] = h->ncol[l];
  }
  // the autotuner's timed launches see this build's max |D| (mixed form)
  if (h->fp32_tau > 0.0)
    if (int rc = enqueue_dmax(h, dD, st)) return rc;
  if (int rc = choose_variants(h, ba, st)) return rc;
  // the batched form: capture once (per D, F), replay afterwards.  Per-class
  // timing and single-class diagnostics run eagerly on st.
  const bool graph = h->use_graph && !h->class_timing && h->only_class < 0 && ba.DK == dD && k_scale == 1.0;
  if (graph && (!h->gexec || h->g_dD != dD || h->g_dF != dF)) {
    drop_graph(h);
    for (int k = 0; k < kSideStreams; ++k) {
      if (!h->side[k]) HIP_TRY(hipStreamCreateWithFlags(&h->side[k], hipStreamNonBlocking));
      if (!h->join_ev[k]) HIP_TRY(hipEventCreateWithFlags(&h->join_ev[k], hipEventDisableTiming));
    }
    if (!h->fork_ev) HIP_TRY(hipEventCreateWithFlags(&h->fork_ev, hipEventDisableTiming));
    HIP_TRY(hipStreamSynchronize(h->stream));
    HIP_TRY(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    int rc = enqueue_build(h, ba, dD, dF, h->stream, kSideStreams);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(h->stream, &g);
    if (rc) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    if (e != hipSuccess) return fail(LIBERI_EDEVICE, std::string("graph capture: ") + hipGetErrorString(e));
    h->graph = g;
    HIP_TRY(hipGraphInstantiate(&h->gexec, h->graph, nullptr, nullptr, 0));
    h->g_dD = dD;
    h->g_dF = dF;
  }
  HIP_TRY(hipEventRecord(h->ev0, st));
  if (graph) {
    HIP_TRY(hipGraphLaunch(h->gexec, st));
  } else if (int rc = enqueue_build(h, ba, dD, dF, st, 1)) {
    return rc;
  }
  HIP_TRY(hipEventRecord(h->ev1, st));
  h->build_pending = true;
  h->fp32_pending = h->fp32_tau > 0.0;
  if (!h->fp32_pending) h->stats.fp32_quartets = 0;
  h->timing_pending = h->class_timing && !graph;
  if (!h->timing_pending)
    for (double &v : h->stats.class_ms) v = 0.0;
  return LIBERI_OK;
}

int liberi_c_fock_build(liberi_ctx *h, const double *density, double *fock, int64_t n) {
  if (!h) return fail(LIBERI_EARG, "liberi_c_fock_build: NULL handle");
  if (!h->initialized)
    return fail(LIBERI_ESTATE, "liberi_fock_build called before liberi_setup");  // liberi_interface.F90:219-222
  const int64_t ntri = (int64_t)h->num_bas * (h->num_bas + 1) / 2;
  if (n != ntri)
    return fail(LIBERI_ESIZE, "liberi_c_fock_build: n=" + std::to_string(n) + ", expected " + std::to_string(ntri));
  if (!density || !fock) return fail(LIBERI_EARG, "liberi_c_fock_build: NULL array");
  HIP_TRY(hipMemcpyAsync(h->dens.p, density, ntri * sizeof(double), hipMemcpyHostToDevice, h->stream));
  int rc = build_on_device(h, (const double *)h->dens.p, (double *)h->fock.p, h->stream);
  if (rc) return rc;
  HIP_TRY(hipMemcpyAsync(fock, h->fock.p, ntri * sizeof(double), hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(hipStreamSynchronize(h->stream));
  collect_timing(h);
  return LIBERI_OK;
}

int liberi_c_fock_build_device(liberi_ctx *h, const double *dD, double *dF, int64_t n, void *stream) {
  if (!h) return fail(LIBERI_EARG, "liberi_c_fock_build_device: NULL handle");
  if (!h->initialized) return fail(LIBERI_ESTATE, "liberi_fock_build called before liberi_setup");
  const int64_t ntri = (int64_t)h->num_bas * (h->num_bas + 1) / 2;
  if (n != ntri)
    return fail(LIBERI_ESIZE, "liberi_c_fock_build_device: n=" + std::to_string(n) + ", expected " + std::to_string(ntri));
  if (!dD || !dF) return fail(LIBERI_EARG, "liberi_c_fock_build_device: NULL array");
  return build_on_device(h, dD, dF, stream ? (hipStream_t)stream : h->stream);
}

// ------------------------------------------------------------------- UHF
int liberi_c_fock_build_uhf_device(liberi_ctx *h, const double *dDa, const double *dDb, double *dFa, double *dFb,
                                   int64_t n, void *stream) {
  if (!h) return fail(LIBERI_EARG, "liberi_c_fock_build_uhf_device: NULL handle");
  if (!h->initialized) return fail(LIBERI_ESTATE, "liberi_fock_build called before liberi_setup");
  const int64_t ntri = (int64_t)h->num_bas * (h->num_bas + 1) / 2;
  if (n != ntri)
    return fail(LIBERI_ESIZE, "liberi_c_fock_build_uhf_device: n=" + std::to_string(n) + ", expected " + std::to_string(ntri));
  if (!dDa || !dDb || !dFa || !dFb) return fail(LIBERI_EARG, "liberi_c_fock_build_uhf_device: NULL array");
  hipStream_t st = stream ? (hipStream_t)stream : h->stream;
  if (!h->dens_t.p)
    if (int rc = dev_alloc(h, h->dens_t, ntri * sizeof(double))) return rc;
  double *dt = (double *)h->dens_t.p;
  if (int e = launch_add(dDa, dDb, dt, ntri, st))
    return fail(LIBERI_EDEVICE, std::string("Da + Db: ") + hipGetErrorString((hipError_t)e));
  if (int rc = build_on_device(h, dt, dFa, st, dDa, 2.0)) return rc;
  return build_on_device(h, dt, dFb, st, dDb, 2.0);
}

int liberi_c_fock_build_uhf(liberi_ctx *h, const double *Da, const double *Db, double *Fa, double *Fb, int64_t n) {
  if (!h) return fail(LIBERI_EARG, "liberi_c_fock_build_uhf: NULL handle");
  if (!h->initialized) return fail(LIBERI_ESTATE, "liberi_fock_build called before liberi_setup");
  const int64_t ntri = (int64_t)h->num_bas * (h->num_bas + 1) / 2;
  if (n != ntri)
    return fail(LIBERI_ESIZE, "liberi_c_fock_build_uhf: n=" + std::to_string(n) + ", expected " + std::to_string(ntri));
  if (!Da || !Db || !Fa || !Fb) return fail(LIBERI_EARG, "liberi_c_fock_build_uhf: NULL array");
  if (!h->dens_b.p)
    if (int rc = dev_alloc(h, h->dens_b, ntri * sizeof(double))) return rc;
  if (!h->fock_b.p)
    if (int rc = dev_alloc(h, h->fock_b, ntri * sizeof(double))) return rc;
  const size_t bytes = ntri * sizeof(double);
  HIP_TRY(hipMemcpyAsync(h->dens.p, Da, bytes, hipMemcpyHostToDevice, h->stream));
  HIP_TRY(hipMemcpyAsync(h->dens_b.p, Db, bytes, hipMemcpyHostToDevice, h->stream));
  if (int rc = liberi_c_fock_build_uhf_device(h, (const double *)h->dens.p, (const double *)h->dens_b.p,
                                              (double *)h->fock.p, (double *)h->fock_b.p, n, h->stream))
    return rc;
  HIP_TRY(hipMemcpyAsync(Fa, h->fock.p, bytes, hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(hipMemcpyAsync(Fb, h->fock_b.p, bytes, hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(hipStreamSynchronize(h->stream));
  collect_timing(h);
  return LIBERI_OK;
}

// ------------------------------------------- GAMESS direct-link entry points
static liberi_ctx *g_handle = nullptr;  // gms_liberi_wrapper.F90:23-25 (module `save` handle)
static int64_t g_matrix_size = 0;

int liberi_c_init(int32_t natoms, int32_t num_bas, int32_t nsh, int32_t mxgtot, const int32_t *ang_mom,
                  const int32_t *contr_num, const int32_t *sh_loc, const int32_t *atom_num, const int32_t *atom_loc,
                  const int32_t *start_bas, const int32_t *end_bas, const double *exponents, const double *cs,
                  const double *cp, const double *cd, const double *cf, const double *coords, const double *schwrz,
                  int32_t my_rank, int32_t num_procs) {
  if (g_handle) liberi_c_destroy(&g_handle);  // gms_liberi_wrapper.F90:80-83
  int rc = liberi_c_create(&g_handle);
  if (rc) return rc;
  rc = liberi_c_setup(g_handle, nsh, natoms, num_bas, mxgtot, ang_mom, contr_num, sh_loc, atom_num, atom_loc,
                      start_bas, end_bas, exponents, cs, cp, cd, cf, coords, schwrz, my_rank, num_procs);
  if (rc) {
    std::string msg = g_last_error;
    liberi_c_destroy(&g_handle);
    return fail(rc, msg);
  }
  g_matrix_size = (int64_t)num_bas * (num_bas + 1) / 2;
  return LIBERI_OK;
}

int liberi_c_fock(const double *density, double *fock, int64_t n) {
  if (!g_handle) return fail(LIBERI_ESTATE, "liberi_fock called before liberi_init");  // :117-120
  if (n != g_matrix_size)
    return fail(LIBERI_ESIZE, "matrix size mismatch in liberi_fock: expected " + std::to_string(g_matrix_size) +
                                  ", got " + std::to_string(n));  // :122-127
  return liberi_c_fock_build(g_handle, density, fock, n);
}

int liberi_c_finalize(void) {
  if (g_handle) liberi_c_destroy(&g_handle);
  g_matrix_size = 0;
  return LIBERI_OK;
}

// ------------------------------------------------------ Schwarz integrals
int liberi_c_schwarz(int32_t nsh, int32_t natoms, int32_t num_bas, int32_t mxgtot, const int32_t *ang_mom,
                     const int32_t *contr_num, const int32_t *sh_loc, const int32_t *atom_num,
                     const int32_t *atom_loc, const double *exponents, const double *cs, const double *cp,
                     const double *cd, const double *cf, const double *coords, double *schwrz_int) {
  if (!schwrz_int) return fail(LIBERI_EARG, "liberi_c_schwarz: NULL output");
  SchwarzStage s;
  if (int rc = schwarz_stage("liberi_c_schwarz", nsh, natoms, num_bas, mxgtot, ang_mom, contr_num, sh_loc,
                             atom_num, atom_loc, exponents, cs, cp, cd, cf, coords, s))
    return rc;
  if (int rc = schwarz_check_device("liberi_c_schwarz")) return rc;
  void *dp = nullptr, *dq = nullptr, *dl = nullptr, *dout = nullptr;
  hipStream_t st = nullptr;
  auto release = [&]() {
    if (st) (void)hipStreamDestroy(st);
    for (void *p : {dp, dq, dl, dout})
      if (p) (void)hipFree(p);
  };
  std::vector<double> out(s.list.size());
  hipError_t e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipMalloc(&dp, s.pairs.size() * sizeof(PairRec));
  if (e == hipSuccess) e = hipMalloc(&dq, s.prims.size() * sizeof(PrimRec));
  if (e == hipSuccess) e = hipMalloc(&dl, s.list.size() * sizeof(int32_t));
  if (e == hipSuccess) e = hipMalloc(&dout, s.list.size() * sizeof(double));
  if (e == hipSuccess)
    e = (hipError_t)schwarz_enqueue(s, (PairRec *)dp, (PrimRec *)dq, (int32_t *)dl, (double *)dout, st);
  if (e == hipSuccess) e = hipMemcpyAsync(out.data(), dout, out.size() * sizeof(double), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  release();
  if (e != hipSuccess)
    return fail(e == hipErrorOutOfMemory ? LIBERI_ENOMEM : LIBERI_EDEVICE,
                std::string("liberi_c_schwarz: ") + hipGetErrorString(e));
  schwarz_scatter(s, out.data(), schwrz_int);
  return LIBERI_OK;
}

}  // extern "C"

// Host staging shared by the two Schwarz entry points (liberi_c_schwarz here,
// liberi_c_schwarz_noalloc in noalloc.hip).
namespace liberi {

int schwarz_stage(const char *who, int32_t nsh, int32_t natoms, int32_t num_bas, int32_t mxgtot,
                  const int32_t *ang_mom, const int32_t *contr_num, const int32_t *sh_loc, const int32_t *atom_num,
                  const int32_t *atom_loc, const double *exponents, const double *cs, const double *cp,
                  const double *cd, const double *cf, const double *coords, SchwarzStage &s) {
  std::vector<HostShell> sh;
  if (int rc = read_shells(who, nsh, natoms, num_bas, mxgtot, ang_mom, contr_num, sh_loc, atom_num, atom_loc,
                           exponents, cs, cp, cd, cf, coords, sh))
    return rc;
  // every shell pair i >= j, A = the shell of higher angular momentum, grouped by (la, lb)
  std::vector<int32_t> by_l[16];
  s.pairs.clear();
  s.prims.clear();
  s.where.clear();
  s.list.clear();
  for (int i = 0; i < nsh; ++i)
    for (int j = 0; j <= i; ++j) {
      const int sA = sh[i].l >= sh[j].l ? i : j, sB = sA == i ? j : i;
      by_l[sh[sA].l * 4 + sh[sB].l].push_back((int32_t)s.pairs.size());
      s.where.push_back((int64_t)i * (i + 1) / 2 + j);
      append_pair(s.pairs, s.prims, sh, nullptr, sA, sB);
    }
  if (s.prims.size() > (size_t)INT32_MAX || s.pairs.size() > (size_t)INT32_MAX)
    return fail(LIBERI_EARG, std::string(who) + ": basis too large for 32-bit pair indices");
  for (int k = 0; k < 16; ++k) {
    s.off[k] = s.list.size();
    s.list.insert(s.list.end(), by_l[k].begin(), by_l[k].end());
  }
  s.off[16] = s.list.size();
  return LIBERI_OK;
}

int schwarz_check_device(const char *who) {
  int dev = 0;
  hipDeviceProp_t prop;
  if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess)
    return fail(LIBERI_EDEVICE, std::string(who) + ": no HIP device visible");
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    return fail(LIBERI_EDEVICE, std::string(who) + ": device is " + prop.gcnArchName +
                                    ", this library is built for gfx950 (MI355X)");
  return LIBERI_OK;
}

int schwarz_enqueue(const SchwarzStage &s, PairRec *dp, PrimRec *dq, int32_t *dl, double *dout, void *stream) {
  hipStream_t st = (hipStream_t)stream;
  hipError_t e = hipMemcpyAsync(dp, s.pairs.data(), s.pairs.size() * sizeof(PairRec), hipMemcpyHostToDevice, st);
  if (e == hipSuccess)
    e = hipMemcpyAsync(dq, s.prims.data(), s.prims.size() * sizeof(PrimRec), hipMemcpyHostToDevice, st);
  if (e == hipSuccess)
    e = hipMemcpyAsync(dl, s.list.data(), s.list.size() * sizeof(int32_t), hipMemcpyHostToDevice, st);
  for (int k = 0; k < 16 && e == hipSuccess; ++k) {
    const int32_t n = (int32_t)(s.off[k + 1] - s.off[k]);
    if (n > 0) e = (hipError_t)launch_schwarz(k / 4, k % 4, dp, dq, dl + s.off[k], n, dout + s.off[k], st);
  }
  return (int)e;
}

void schwarz_scatter(const SchwarzStage &s, const double *out, double *schwrz_int) {
  for (size_t k = 0; k < s.list.size(); ++k) schwrz_int[s.where[s.list[k]]] = out[k];
}

int set_last_error(int code, const std::string &msg) { return fail(code, msg); }

}  // namespace liberi
