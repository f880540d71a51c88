// capi.cpp -- the C ABI of libccr_hip.so (include/ccr_hip.h): contexts, plan compile/upload,
// kernel execution with HIP-event timing, read-back into caller-owned ccr_exposure_out arrays.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <functional>
#include <memory>
#include <mutex>
#include <random>
#include <string>
#include <vector>

#include "../../include/ccr_hip.h"
#include "launch.hpp"
#include "mlp_host.hpp"
#include "draws.hpp"
#include "plan.hpp"


using namespace ccrhip;

namespace {

// One replayed normal stream in FP64 on the device (the reference's draws for a seed, first n),
// shared by every plan of the context that replays a prefix of it.
struct StreamBuf {
  uint64_t seed = 0;
  int64_t n = 0;
  double* d = nullptr;
  double draws_ms = 0.0, upload_ms = 0.0;
  ~StreamBuf() {
    if (d) (void)hipFree(d);
  }
};

}  // namespace

struct ccr_hip_ctx {
  int device = 0;
  int precision = CCR_FP64;
  int block = 0;  // path-kernel workgroup size; 0: chosen per plan by occupancy
  hipStream_t stream = nullptr;
  hipEvent_t ev[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};  // start, AMC, path, reduce, PFE
  std::string err;
  int fault = CCR_FAULT_NONE;
  bool timing_pending = false;  // events recorded by a launch not yet read into `timing`
  double upload_ms = 0.0, draws_ms = 0.0;
  ccr_hip_timing timing{};
  // Normal streams recently generated on this context (at most two kept; a plan holds its own
  // reference, so an evicted stream lives as long as the plans replaying it).
  std::vector<std::shared_ptr<StreamBuf>> streams;
  // Shared workspace: plans compiled while share_ws is set keep their per-execute buffers
  // (exposure rows, wave partials, AMC states and values) here, sized to the largest of them.
  bool share_ws = false;
  void* ws = nullptr;
  size_t ws_bytes = 0;
  // Initial-margin proxy for plans compiled while im_quantile > 0.
  double im_quantile = 0.0, im_mpor = 10.0 / 250.0;
  // Per-trade statistics for plans compiled while set.
  bool per_trade = false;
  // Nested conditional initial margin for plans compiled while nested_inner > 0.
  int nested_inner = 0;
  double nested_q = 0.99, nested_mpor = 10.0 / 250.0;
  // Book mode (common grids, pathwise identical scenarios) for plans compiled while set, and the
  // grids of the whole book when it is compiled as several plans (empty: each plan's own union).
  bool book = false;
  std::vector<double> book_rep, book_sim;
  // Book plans: netting sets valued per workgroup by the book kernel (1: one set per workgroup).
  // Four, the unrolled set loop with its state in registers: the market step is shared by four
  // sets (500-set portfolio, 200,000 paths, 256 threads: path kernel 362 ms at 4, 400 ms at 2).
  int book_group = 4;
  // Bermudan regressions of plans compiled while set: -1 the reference's state basis on the
  // shared per-date designs; CCR_BASIS_STATE / CCR_BASIS_SWAP_RATE_ANNUITY on the regression
  // engine, its Grams in FP64 or on FP32 matrix cores.
  int basis = -1;
  int gram32 = 0;
  // Surrogate inputs of the last surrogate execute (grown as needed).
  float* sur_feats = nullptr;
  size_t sur_feats_bytes = 0;
  // Executes replay one captured HIP graph per plan while set.
  bool use_graph = false;
  // Launches are being captured into a graph (no timing events recorded), and whether the last
  // launch recorded the per-stage events (graph replays record the total only).
  bool capturing = false, stage_events = true;
};

namespace {

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
};

}  // namespace

struct ccr_hip_plan {
  HostPlan host;
  int device = 0;
  std::vector<DevBuf> bufs;
  GlobalDev* d_global = nullptr;
  SetDev* d_sets = nullptr;
  int32_t* d_block_set = nullptr;
  void* d_stream = nullptr;   // the stream in the plan's precision (FP64: the shared buffer itself)
  std::shared_ptr<StreamBuf> stream64;
  void* d_dom = nullptr;     // domestic normals of the AMC states pass
  void* d_steps = nullptr;
  DateRec* d_dates = nullptr;
  void* d_date_reals = nullptr;
  void* d_bonds = nullptr;
  void* d_fixed = nullptr;
  int32_t* d_fixstore = nullptr;
  void* d_poly = nullptr;     // bond-polynomial coefficients in the plan's precision
  void* d_opts = nullptr;
  void* d_rows = nullptr;
  double* d_partials = nullptr;
  double* d_stats = nullptr;
  double* d_pfe = nullptr;
  // American Monte Carlo (null when no set holds a Bermudan swaption).
  BermDev* d_berms = nullptr;
  BermDate* d_bdates = nullptr;
  double* d_cont = nullptr;
  int16_t* d_exidx = nullptr;
  double* d_states = nullptr;
  double* d_bermv = nullptr;
  double* d_ctimes = nullptr;
  AmcArgs amc{};
  double draws_ms = 0.0, upload_ms = 0.0;
  // Per-execute buffers placed in the context's shared workspace: (field, byte offset).
  std::vector<std::pair<void**, size_t>> ws_slots;
  size_t ws_need = 0;
  // Initial-margin proxy (null / 0 when off): |dV| rows, per-wave deflator sums, their quantiles
  // and sums per (set, simulation date), and the last execute's read-back.
  double im_quantile = 0.0, im_mpor = 0.0;
  void* d_im_rows = nullptr;
  double* d_im_part = nullptr;
  double* d_im_q = nullptr;
  double* d_im_dsum = nullptr;
  std::vector<double> im_q, im_dsum;
  bool im_ready = false;
  // Per-trade statistics (tstride = 0 when off): per-(date, trade) records and reals, per-wave
  // partials, the reduced statistics [set][R_max][tstride] and the last execute's read-back.
  int32_t tstride = 0;
  DateRec* d_tdates = nullptr;
  void* d_tdreals = nullptr;
  double* d_tpart = nullptr;
  double* d_tstats = nullptr;
  std::vector<double> tstats;
  bool tstats_ready = false;
  // Nested conditional initial margin (null when off): per set date the target date, IM rows and
  // per-wave partials (laid out as the exposure rows and statistics), their reductions
  // [set][R_max][kStats] (slots 0: E[D IM], 1: E[IM]) and the 5 / 50 / 95% quantiles [3][set][R_max].
  int32_t* d_ntarget = nullptr;
  void* d_nrows = nullptr;
  double* d_npart = nullptr;
  double* d_nstats = nullptr;
  double* d_nbands = nullptr;
  int32_t nested_inner = 0;
  double nested_q = 0.0;
  uint64_t seed = 0;
  // The plan's execute captured as a HIP graph (graph mode), with the workspace it was bound to.
  hipGraphExec_t graph_exec = nullptr;
  void* graph_ws = nullptr;
  ~ccr_hip_plan() {
    if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
    for (auto& b : bufs)
      if (b.p) (void)hipFree(b.p);
  }
};

namespace {

thread_local std::string gCreateError;

void check(hipError_t e, const char* what) {
  if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation)
    fail(CCR_ERR_OUT_OF_MEMORY, std::string(what) + ": " + hipGetErrorString(e));
  if (e != hipSuccess) fail(CCR_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

int32_t guarded(ccr_hip_ctx* ctx, const std::function<void()>& f) {
  try {
    f();
    if (ctx) ctx->err.clear();
    return CCR_OK;
  } catch (const Error& e) {
    if (ctx) ctx->err = e.what();
    return e.code;
  } catch (const std::bad_alloc&) {
    if (ctx) ctx->err = "host allocation failed";
    return CCR_ERR_OUT_OF_MEMORY;
  } catch (const std::exception& e) {
    if (ctx) ctx->err = e.what();
    return CCR_ERR_RUNTIME;
  }
}

void* devAlloc(ccr_hip_ctx* ctx, ccr_hip_plan* plan, size_t bytes);

// A per-execute buffer: the plan's own allocation, or a slot of the context's shared workspace
// (bound to an address by bindWorkspace before every execute).
template <typename T>
void execAlloc(ccr_hip_ctx* ctx, ccr_hip_plan* plan, T*& field, size_t bytes) {
  if (!ctx->share_ws) {
    field = static_cast<T*>(devAlloc(ctx, plan, bytes));
    return;
  }
  plan->ws_slots.push_back({reinterpret_cast<void**>(&field), plan->ws_need});
  plan->ws_need += (bytes + 255) / 256 * 256;
  field = nullptr;
}

// Points the plan's per-execute buffers (workspace slots) into the memory at `base`.
void bindSlots(ccr_hip_plan* plan, void* base) {
  for (auto& s : plan->ws_slots) *s.first = static_cast<char*>(base) + s.second;
  AmcArgs& a = plan->amc;
  a.states = plan->d_states;
  a.cont = plan->d_cont;
  a.exidx = plan->d_exidx;
  plan->d_bermv = a.bermv;
}

void bindWorkspace(ccr_hip_ctx* ctx, ccr_hip_plan* plan) {
  if (plan->ws_slots.empty()) return;
  if (ctx->ws_bytes < plan->ws_need) {
    check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    if (ctx->ws) check(hipFree(ctx->ws), "hipFree");
    ctx->ws = nullptr;
    ctx->ws_bytes = 0;
    check(hipMalloc(&ctx->ws, plan->ws_need), "hipMalloc (shared workspace)");
    ctx->ws_bytes = plan->ws_need;
  }
  bindSlots(plan, ctx->ws);
}

void* devAlloc(ccr_hip_ctx* ctx, ccr_hip_plan* plan, size_t bytes) {
  if (ctx->fault == CCR_FAULT_ALLOC) {
    ctx->fault = CCR_FAULT_NONE;
    fail(CCR_ERR_OUT_OF_MEMORY, "hipMalloc: out of memory (forced by ccr_hip_set_fault)");
  }
  void* p = nullptr;
  check(hipMalloc(&p, bytes < 16 ? 16 : bytes), "hipMalloc");
  plan->bufs.push_back({p, bytes});
  return p;
}

template <typename T>
T* upload(ccr_hip_ctx* ctx, ccr_hip_plan* plan, const T* src, size_t n) {
  T* d = static_cast<T*>(devAlloc(ctx, plan, n * sizeof(T)));
  if (n) check(hipMemcpyAsync(d, src, n * sizeof(T), hipMemcpyHostToDevice, ctx->stream), "hipMemcpyAsync");
  return d;
}

void* uploadReals(ccr_hip_ctx* ctx, ccr_hip_plan* plan, const std::vector<double>& v, bool fp32) {
  if (!fp32) return upload(ctx, plan, v.data(), v.size());
  std::vector<float> f(v.begin(), v.end());
  void* d = upload(ctx, plan, f.data(), f.size());
  check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
  return d;
}

// Generates the first n normals of the stream for `seed` (on the host's cores, bit-identical to
// the sequential stream) in chunks into two pinned staging buffers, overlapping the generation of
// one chunk with the upload of the other.
std::shared_ptr<StreamBuf> generateStream64(ccr_hip_ctx* ctx, uint64_t seed, int64_t n) {
  auto buf = std::make_shared<StreamBuf>();
  buf->seed = seed;
  buf->n = n;
  if (ctx->fault == CCR_FAULT_ALLOC) {
    ctx->fault = CCR_FAULT_NONE;
    fail(CCR_ERR_OUT_OF_MEMORY, "hipMalloc: out of memory (forced by ccr_hip_set_fault)");
  }
  check(hipMalloc(reinterpret_cast<void**>(&buf->d), static_cast<size_t>(n > 0 ? n : 1) * sizeof(double)), "hipMalloc");
  if (n == 0) return buf;
  const int64_t chunk = 1 << 22;
  void* stage[2] = {nullptr, nullptr};
  hipEvent_t done[2] = {nullptr, nullptr};
  check(hipHostMalloc(&stage[0], static_cast<size_t>(chunk) * sizeof(double), hipHostMallocDefault), "hipHostMalloc");
  check(hipHostMalloc(&stage[1], static_cast<size_t>(chunk) * sizeof(double), hipHostMallocDefault), "hipHostMalloc");
  check(hipEventCreate(&done[0]), "hipEventCreate");
  check(hipEventCreate(&done[1]), "hipEventCreate");
  NormalStream normals(seed);
  double genMs = 0.0;
  const auto t0 = std::chrono::steady_clock::now();
  int which = 0;
  bool pending[2] = {false, false};
  for (int64_t off = 0; off < n; off += chunk, which ^= 1) {
    const int64_t m = std::min(chunk, n - off);
    if (pending[which]) check(hipEventSynchronize(done[which]), "hipEventSynchronize");
    const auto g0 = std::chrono::steady_clock::now();
    normals.fill(static_cast<double*>(stage[which]), m);
    genMs += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - g0).count();
    check(hipMemcpyAsync(buf->d + off, stage[which], static_cast<size_t>(m) * sizeof(double), hipMemcpyHostToDevice,
                         ctx->stream),
          "hipMemcpyAsync");
    check(hipEventRecord(done[which], ctx->stream), "hipEventRecord");
    pending[which] = true;
  }
  check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
  const double total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  (void)hipEventDestroy(done[0]);
  (void)hipEventDestroy(done[1]);
  (void)hipHostFree(stage[0]);
  (void)hipHostFree(stage[1]);
  buf->draws_ms = genMs;
  buf->upload_ms = total - genMs;
  return buf;
}

// The plan's replayed stream: a prefix of a stream this context already holds for the seed, or
// a newly generated one; FP32 plans narrow it on the device (round to nearest, as the host cast).
void uploadStream(ccr_hip_ctx* ctx, ccr_hip_plan* plan, uint64_t seed, bool fp32) {
  const int64_t n = plan->host.stream_len;
  std::shared_ptr<StreamBuf> src;
  for (const auto& s : ctx->streams)
    if (s->seed == seed && s->n >= n) src = s;
  plan->draws_ms = 0.0;
  if (!src) {
    src = generateStream64(ctx, seed, n);
    ctx->streams.push_back(src);
    if (ctx->streams.size() > 2) ctx->streams.erase(ctx->streams.begin());
    plan->draws_ms = src->draws_ms;
    plan->upload_ms += src->upload_ms;
  }
  plan->stream64 = src;
  if (!fp32) {
    plan->d_stream = src->d;
    return;
  }
  plan->d_stream = devAlloc(ctx, plan, static_cast<size_t>(n) * sizeof(float));
  if (n > 0) check(launchNarrow(src->d, static_cast<float*>(plan->d_stream), n, ctx->stream), "stream narrowing");
  check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
}

// Workgroup size of the path kernel with the most waves resident per CU for this plan (its LDS
// rows per path and its kernel specialisation's registers); ties go to the larger workgroup.
int occupancyBlock(const HostPlan& h) {
  const int ldsRows = h.group > 1 ? h.g.max_slots + h.group * (h.g.max_ring + h.price_rows)
                                  : h.g.max_ring + h.price_rows + h.g.max_slots + h.max_groups;
  const bool trades = h.max_groups > 0;
  int best = kMaxBlock, bestWaves = -1;
  for (int block = kMaxBlock; block >= 64; block -= 64) {
    int wg = 0;
    const hipError_t e = h.precision == CCR_FP32
                             ? pathOccupancy<float>(h.g.nrates, h.g.nassets, h.g.ncredits, !h.opts.empty(), ldsRows,
                                                    block, &wg, trades, h.group)
                             : pathOccupancy<double>(h.g.nrates, h.g.nassets, h.g.ncredits, !h.opts.empty(), ldsRows,
                                                     block, &wg, trades, h.group);
    check(e, "hipOccupancyMaxActiveBlocksPerMultiprocessor");
    const int waves = wg * (block / 64);
    if (waves > bestWaves) {
      bestWaves = waves;
      best = block;
    }
  }
  return best;
}

template <typename Real>
PathArgs<Real> pathArgs(const ccr_hip_plan* plan);

// Timing event i of an execute (0 start, 1 after the AMC stage, 2 after the path kernel, 3 after
// the reductions, 4 end); not recorded while the launches are captured into a graph.
void markEvent(ccr_hip_ctx* ctx, int i) {
  if (!ctx->capturing) check(hipEventRecord(ctx->ev[i], ctx->stream), "hipEventRecord");
}

void compilePlan(ccr_hip_ctx* ctx, const ccr_scenario_generator* generator, const ccr_simulation_config* config,
                 const double* base_grid, int32_t base_size, const ccr_netting_set* sets, int32_t num_sets,
                 double pfe_quantile, int32_t precision, ccr_hip_plan** out) {
  if (!ctx || !generator || !config || !out) invalid("null argument");
  if (!base_grid && base_size > 0) invalid("null base grid");
  if (precision != CCR_FP64 && precision != CCR_FP32) invalid("precision must be CCR_FP64 or CCR_FP32");
  if (ctx->fault == CCR_FAULT_PLAN) {
    ctx->fault = CCR_FAULT_NONE;
    invalid("invalid plan (forced by ccr_hip_set_fault)");
  }
  check(hipSetDevice(ctx->device), "hipSetDevice");
  auto plan = std::make_unique<ccr_hip_plan>();
  plan->device = ctx->device;
  const bool bookGrids = ctx->book && !ctx->book_rep.empty();
  plan->host = buildPlan(*generator, *config, base_grid, base_size, sets, num_sets, pfe_quantile, precision,
                         ctx->block > 0 ? ctx->block : kMaxBlock, ctx->per_trade, ctx->book,
                         ctx->im_quantile > 0.0, bookGrids ? &ctx->book_rep : nullptr,
                         bookGrids ? &ctx->book_sim : nullptr, ctx->basis, ctx->nested_inner, ctx->nested_mpor);
  HostPlan& h = plan->host;
  if (ctx->book && ctx->book_group > 1 && !ctx->per_trade && ctx->im_quantile <= 0.0 && h.opts.empty() &&
      h.sets.size() > 1)
    h.group = ctx->book_group;
  if (ctx->block == 0) assignBlocks(h, occupancyBlock(h));
  const bool fp32 = precision == CCR_FP32;
  const auto t0 = std::chrono::steady_clock::now();
  plan->d_global = upload(ctx, plan.get(), &h.g, 1);
  plan->d_sets = upload(ctx, plan.get(), h.sets.data(), h.sets.size());
  plan->d_block_set = upload(ctx, plan.get(), h.block_set.data(), h.block_set.size());
  plan->d_steps = uploadReals(ctx, plan.get(), h.steps, fp32);
  plan->d_dates = upload(ctx, plan.get(), h.dates.data(), h.dates.size());
  plan->d_date_reals = uploadReals(ctx, plan.get(), h.date_reals, fp32);
  plan->d_bonds = uploadReals(ctx, plan.get(), h.bonds, fp32);
  plan->d_fixed = uploadReals(ctx, plan.get(), h.fixed, fp32);
  plan->d_fixstore = upload(ctx, plan.get(), h.fixstore.data(), h.fixstore.size());
  plan->d_poly = uploadReals(ctx, plan.get(), h.poly, fp32);
  plan->d_opts = h.opts.empty() ? nullptr : uploadReals(ctx, plan.get(), h.opts, fp32);
  const size_t elem = fp32 ? sizeof(float) : sizeof(double);
  const int64_t P = h.g.num_paths;
  execAlloc(ctx, plan.get(), plan->d_rows, static_cast<size_t>(num_sets) * h.max_rep * P * elem);
  execAlloc(ctx, plan.get(), plan->d_partials, static_cast<size_t>(h.total_waves) * h.max_rep * kStats * sizeof(double));
  plan->seed = config->seed;
  if (h.nested_inner > 0) {
    plan->nested_inner = h.nested_inner;
    plan->nested_q = ctx->nested_q;
    plan->d_ntarget = upload(ctx, plan.get(), h.nested_target.data(), h.nested_target.size());
    execAlloc(ctx, plan.get(), plan->d_nrows, static_cast<size_t>(num_sets) * h.max_rep * P * elem);
    execAlloc(ctx, plan.get(), plan->d_npart, static_cast<size_t>(h.total_waves) * h.max_rep * kStats * sizeof(double));
    plan->d_nstats = static_cast<double*>(
        devAlloc(ctx, plan.get(), static_cast<size_t>(num_sets) * h.max_rep * kStats * sizeof(double)));
    plan->d_nbands = static_cast<double*>(devAlloc(ctx, plan.get(), static_cast<size_t>(3) * num_sets * h.max_rep * sizeof(double)));
  }
  if (ctx->im_quantile > 0.0) {
    // The proxy needs the netted value on every simulation date (the plan values every date
    // when it is on: book plans simulate dates of other sets that their own valuation skips).
    for (const DateRec& d : h.dates)
      if (!(d.flags & kValued)) invalid("initial margin: a simulation date is not valued");
    plan->im_quantile = ctx->im_quantile;
    plan->im_mpor = ctx->im_mpor;
    const int64_t dates = h.max_dates, wavesPerSet = h.total_waves / num_sets;
    for (int s = 0; s < num_sets; ++s) {
      h.sets[static_cast<size_t>(s)].im_off = static_cast<int64_t>(s) * (dates - 1) * P;
      h.sets[static_cast<size_t>(s)].im_part_off = static_cast<int64_t>(s) * wavesPerSet * dates;
    }
    execAlloc(ctx, plan.get(), plan->d_im_rows, static_cast<size_t>(num_sets) * (dates > 1 ? dates - 1 : 1) * P * elem);
    execAlloc(ctx, plan.get(), plan->d_im_part, static_cast<size_t>(h.total_waves) * dates * sizeof(double));
    plan->d_im_q = static_cast<double*>(devAlloc(ctx, plan.get(), static_cast<size_t>(num_sets) * dates * sizeof(double)));
    plan->d_im_dsum = static_cast<double*>(devAlloc(ctx, plan.get(), static_cast<size_t>(num_sets) * dates * sizeof(double)));
    check(hipMemcpyAsync(plan->d_sets, h.sets.data(), h.sets.size() * sizeof(SetDev), hipMemcpyHostToDevice,
                         ctx->stream),
          "hipMemcpyAsync");
  }
  if (h.max_groups > 0) {
    plan->tstride = kTradeStats * h.max_groups;
    const int64_t per = static_cast<int64_t>(h.max_rep) * plan->tstride;
    for (SetDev& sd : h.sets) sd.tpart_off = static_cast<int64_t>(sd.wave_begin) * per;
    plan->d_tdates = upload(ctx, plan.get(), h.tdates.data(), h.tdates.size());
    plan->d_tdreals = uploadReals(ctx, plan.get(), h.tdreals, fp32);
    execAlloc(ctx, plan.get(), plan->d_tpart, static_cast<size_t>(h.total_waves) * per * sizeof(double));
    plan->d_tstats = static_cast<double*>(devAlloc(ctx, plan.get(), static_cast<size_t>(num_sets) * per * sizeof(double)));
    check(hipMemcpyAsync(plan->d_sets, h.sets.data(), h.sets.size() * sizeof(SetDev), hipMemcpyHostToDevice,
                         ctx->stream),
          "hipMemcpyAsync");
  }
  plan->d_stats = static_cast<double*>(devAlloc(ctx, plan.get(), static_cast<size_t>(num_sets) * h.max_rep * kStats * sizeof(double)));
  plan->d_pfe = static_cast<double*>(devAlloc(ctx, plan.get(), static_cast<size_t>(num_sets) * h.max_rep * sizeof(double)));
  if (!h.berms.empty()) {
    auto dalloc = [&](int64_t n) { return static_cast<double*>(devAlloc(ctx, plan.get(), static_cast<size_t>(n) * sizeof(double))); };
    const int nB = static_cast<int>(h.berms.size());
    const int ncd = static_cast<int>(h.cdates.size() / 2);
    const int ncols = static_cast<int>(h.exlist.size());
    const int nchunks = static_cast<int>((P + kAmcChunk - 1) / kAmcChunk);
    plan->d_berms = upload(ctx, plan.get(), h.berms.data(), h.berms.size());
    plan->d_bdates = upload(ctx, plan.get(), h.bdates.data(), h.bdates.size());
    execAlloc(ctx, plan.get(), plan->d_cont, static_cast<size_t>(h.cont_len) * sizeof(double));
    execAlloc(ctx, plan.get(), plan->d_exidx, static_cast<size_t>(nB) * P * sizeof(int16_t));
    execAlloc(ctx, plan.get(), plan->d_states, static_cast<size_t>(2 * h.state_len) * sizeof(double));
    AmcArgs& a = plan->amc;
    a.sets = plan->d_sets;
    a.berms = plan->d_berms;
    a.bdates = plan->d_bdates;
    a.evterms = upload(ctx, plan.get(), h.evterms.data(), h.evterms.size());
    a.eidx = upload(ctx, plan.get(), h.eidx.data(), h.eidx.size());
    a.exlist = upload(ctx, plan.get(), h.exlist.data(), h.exlist.size());
    a.cdates = ncd > 0 ? upload(ctx, plan.get(), h.cdates.data(), h.cdates.size()) : nullptr;
    a.bdate_berm = upload(ctx, plan.get(), h.bdate_berm.data(), h.bdate_berm.size());
    a.bdate_cd = upload(ctx, plan.get(), h.bdate_cd.data(), h.bdate_cd.size());
    a.clive = upload(ctx, plan.get(), h.clive.data(), h.clive.size());
    a.clive_rng = upload(ctx, plan.get(), h.clive_rng.data(), h.clive_rng.size());
    a.poly = fp32 ? nullptr : static_cast<const double*>(plan->d_poly);
    a.polyf = fp32 ? static_cast<const float*>(plan->d_poly) : nullptr;
    a.states = plan->d_states;
    execAlloc(ctx, plan.get(), a.ev, static_cast<size_t>(h.ev_len) * sizeof(double));
    execAlloc(ctx, plan.get(), a.strategy, static_cast<size_t>(h.ev_len) * sizeof(double));
    execAlloc(ctx, plan.get(), a.part,
              static_cast<size_t>(nchunks) * std::max(nB * kRegMom2, ncd * 10) * sizeof(double));
    a.regp = dalloc(static_cast<int64_t>(nB) * 8);
    a.coef = dalloc(static_cast<int64_t>(nB) * 4);
    a.csum = dalloc(std::max(1, ncd) * 3);
    a.gram = dalloc(std::max(1, ncd) * 10);
    const int nkc = static_cast<int>((P + kGemmK - 1) / kGemmK);
    a.tiles = h.tiles.empty() ? nullptr : upload(ctx, plan.get(), h.tiles.data(), h.tiles.size());
    a.narrow = h.ntiles.empty() ? nullptr : upload(ctx, plan.get(), h.ntiles.data(), h.ntiles.size());
    a.gcols = h.gcols.empty() ? nullptr : upload(ctx, plan.get(), h.gcols.data(), h.gcols.size());
    a.gberm = h.gberm.empty() ? nullptr : upload(ctx, plan.get(), h.gberm.data(), h.gberm.size());
    a.tcd = h.tcd.empty() ? nullptr : upload(ctx, plan.get(), h.tcd.data(), h.tcd.size());
    a.cmap = ncd > 0 ? upload(ctx, plan.get(), h.cmap.data(), h.cmap.size()) : nullptr;
    a.cuniq = ncd > 0 ? upload(ctx, plan.get(), h.cuniq.data(), h.cuniq.size()) : nullptr;
    a.ncuniq = static_cast<int32_t>(h.cuniq.size());
    execAlloc(ctx, plan.get(), a.gpart, static_cast<size_t>(nkc) * h.bdates.size() * 4 * sizeof(double));
    a.cont = plan->d_cont;
    a.exidx = plan->d_exidx;
    if (ncd > 0) execAlloc(ctx, plan.get(), a.bermv, static_cast<size_t>(ncd) * P * sizeof(double));
    plan->d_bermv = a.bermv;
    plan->d_ctimes = ncd > 0 ? upload(ctx, plan.get(), h.ctimes.data(), h.ctimes.size()) : nullptr;
    a.paths = P;
    a.state_len = h.state_len;
    a.nberms = nB;
    a.ncols = ncols;
    a.ncdates = ncd;
    a.nbdates = static_cast<int>(h.bdates.size());
    a.nchunks = nchunks;
    a.ncchunks = (nchunks + kContGroup - 1) / kContGroup;
    a.ntiles = static_cast<int>(h.tiles.size());
    a.nnarrow = static_cast<int>(h.ntiles.size());
    a.nkc = nkc;
    a.basis = -1;
    if (h.basis >= 0) {
      const int nrecs = static_cast<int>(h.rrecs.size());
      const int nbc = static_cast<int>((P + kBasisChunk - 1) / kBasisChunk);
      a.basis = h.basis;
      a.gram32 = ctx->gram32;
      a.nrecs = nrecs;
      a.nex_recs = h.nex_recs;
      a.nbc = nbc;
      a.max_exercises = std::max(1, h.max_exercises);
      a.rrecs = upload(ctx, plan.get(), h.rrecs.data(), h.rrecs.size());
      a.rex = upload(ctx, plan.get(), h.rex.data(), h.rex.size());
      a.rcont = upload(ctx, plan.get(), h.rcont.data(), h.rcont.size());
      a.saterms = upload(ctx, plan.get(), h.saterms.data(), h.saterms.size());
      execAlloc(ctx, plan.get(), a.rpart,
                static_cast<size_t>(std::max(1, nrecs)) * nbc * kBasisGram * sizeof(double));
      execAlloc(ctx, plan.get(), a.rfit, static_cast<size_t>(std::max(1, nrecs)) * kBasisStride * sizeof(double));
    }
  }
  check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
  plan->upload_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  uploadStream(ctx, plan.get(), config->seed, fp32);
  if (h.dom_len > 0) {
    plan->d_dom = devAlloc(ctx, plan.get(), static_cast<size_t>(h.dom_len) * elem);
    const int n = static_cast<int>(h.sets.size());
    const hipError_t e = fp32 ? launchDomesticNormals<float>(pathArgs<float>(plan.get()), h.sets.data(), n,
                                                             static_cast<float*>(plan->d_dom), ctx->stream)
                              : launchDomesticNormals<double>(pathArgs<double>(plan.get()), h.sets.data(), n,
                                                              static_cast<double*>(plan->d_dom), ctx->stream);
    check(e, "domestic normals gather");
    check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
  }
  *out = plan.release();
}

template <typename Real>
PathArgs<Real> pathArgs(const ccr_hip_plan* plan) {
  const HostPlan& h = plan->host;
  PathArgs<Real> a{};
  a.g = plan->d_global;
  a.sets = plan->d_sets;
  a.block_set = plan->d_block_set;
  a.stream = static_cast<const Real*>(plan->d_stream);
  a.dom = static_cast<const Real*>(plan->d_dom);
  a.steps = static_cast<const Real*>(plan->d_steps);
  a.dates = plan->d_dates;
  a.date_reals = static_cast<const Real*>(plan->d_date_reals);
  a.bonds = static_cast<const Real*>(plan->d_bonds);
  a.fixed = static_cast<const Real*>(plan->d_fixed);
  a.fixstore = plan->d_fixstore;
  a.poly = static_cast<const Real*>(plan->d_poly);
  a.opts = static_cast<const Real*>(plan->d_opts);
  a.bermv = plan->d_bermv;
  a.states = nullptr;
  a.states_set = -1;
  a.rows = static_cast<Real*>(plan->d_rows);
  a.partials = plan->d_partials;
  a.max_rep = h.max_rep;
  a.ring = h.g.max_ring;
  a.price_rows = h.price_rows;
  a.slots = h.g.max_slots;
  a.nrates = h.g.nrates;
  a.nassets = h.g.nassets;
  a.ncredits = h.g.ncredits;
  a.has_options = h.opts.empty() ? 0 : 1;
  a.group = h.group;
  a.blocks_per_set = h.sets.empty() ? 0 : h.sets[0].num_blocks;
  a.nsets = static_cast<int32_t>(h.sets.size());
  a.im_rows = static_cast<Real*>(plan->d_im_rows);
  a.im_part = plan->d_im_part;
  a.im_dates = h.max_dates;
  a.tstride = plan->tstride;
  a.tdates = plan->d_tdates;
  a.tdreals = static_cast<const Real*>(plan->d_tdreals);
  a.tpart = plan->d_tpart;
  a.state_len = h.state_len;
  return a;
}

// Levels of the nested initial margin's bands (FusedOptions::bandLevels, fused_exposure.hpp:323).
constexpr double kNestedBands[3] = {0.05, 0.5, 0.95};

int setRows(const HostPlan& h, int s) {
  int rows = 0;
  for (size_t i = 0; i < h.cdates.size() / 2; ++i)
    if (h.cdates[2 * i] == s) ++rows;
  return rows;
}

// Rows (continuation dates, in order) of set s up to its last one on which some swaption can be
// nonzero; past it every value is an exact 0.
int liveRows(const HostPlan& h, int s) {
  int row = 0, live = 0;
  for (size_t i = 0; i < h.cdates.size() / 2; ++i) {
    if (h.cdates[2 * i] != s) continue;
    ++row;
    if (h.clive_rng[2 * i + 1] > h.clive_rng[2 * i]) live = row;
  }
  return live;
}

// The distinct exercise dates (grid indices, ascending, at most kSurrogateMaxExercises) of the
// Bermudans of the given sets.
SurrogateExercises exercisesOf(const HostPlan& h, const std::vector<int>& sets) {
  std::vector<int32_t> d;
  for (int s : sets) {
    const SetDev& S = h.sets[static_cast<size_t>(s)];
    for (int b = S.berm_beg; b < S.berm_end; ++b) {
      const BermDev& bm = h.berms[static_cast<size_t>(b)];
      for (int k = 0; k < bm.ne; ++k) d.push_back(h.eidx[static_cast<size_t>(bm.eidx_off + k)]);
    }
  }
  std::sort(d.begin(), d.end());
  d.erase(std::unique(d.begin(), d.end()), d.end());
  SurrogateExercises ex{};
  ex.count = static_cast<int32_t>(std::min<size_t>(d.size(), kSurrogateMaxExercises));
  for (int m = 0; m < ex.count; ++m) ex.date[m] = d[static_cast<size_t>(m)];
  return ex;
}

// The valued dates (grid indices, in row order) of set s's continuation rows.
std::vector<int32_t> valuedDates(const HostPlan& h, int s) {
  std::vector<int32_t> d;
  for (size_t i = 0; i < h.cdates.size() / 2; ++i)
    if (h.cdates[2 * i] == s) d.push_back(h.cdates[2 * i + 1]);
  return d;
}

// Outputs per surrogate network: the netting sets of one group (sets with Bermudans that share
// recorded states) are split into networks of at most this many outputs.
constexpr int kSurrogateMaxOutputs = 128;

}  // namespace

struct ccr_hip_surrogate {
  // One network output: the Bermudan value of netting set `set` on its valued dates.
  struct Output {
    int set = -1;
    int nberms = 0;
    bool enabled = true;
    double ee_error = 0.0;
    double ymean = 0.0, ysd = 1.0;
    int32_t cd_base = 0, rows = 0, live = 0;  // its continuation rows in the plan; rows >= live are 0
    std::vector<int32_t> dates;               // its valued dates (grid indices), checked against a plan
  };
  // One network over the rows of a group of sets: the union of their valued dates up to the last
  // one with a live swaption (urows rows), every output a set.
  struct Net {
    std::vector<Output> outs;
    std::unique_ptr<MlpRun> run;  // trained weights (inference only)
    int inputs = kSurrogateBase;
    double xmean[kSurrogateBase + kSurrogateMaxExercises], xsd[kSurrogateBase + kSurrogateMaxExercises];
    SurrogateExercises ex{};
    int num_dates = 0, urows = 0;
    std::vector<int32_t> urow, map;  // urow [num_dates]: row of each date (-1: none); map [output][urows]
    std::vector<double> utimes;      // [urows]
    int32_t* d_urow = nullptr;
    double* d_utimes = nullptr;
    int32_t* d_map = nullptr;        // the enabled outputs' entries (the others -1)
    InferOut* d_outs = nullptr;
  };
  int device = 0;
  int num_sets = 0;
  std::vector<Net> nets;
  std::vector<void*> bufs;
  ~ccr_hip_surrogate() {
    for (void* p : bufs) (void)hipFree(p);
  }
  template <typename T>
  T* upload(const std::vector<T>& v, hipStream_t stream) {
    void* p = nullptr;
    check(hipMalloc(&p, std::max<size_t>(v.size() * sizeof(T), 256)), "hipMalloc");
    bufs.push_back(p);
    if (!v.empty()) check(hipMemcpyAsync(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, stream), "hipMemcpyAsync");
    check(hipStreamSynchronize(stream), "hipStreamSynchronize");
    return static_cast<T*>(p);
  }
};

namespace {

// The networks of a plan (no weights yet): groups of sets with Bermudans sharing recorded states,
// split into at most kSurrogateMaxOutputs outputs each, with their row tables.
std::vector<ccr_hip_surrogate::Net> surrogateNets(const HostPlan& h) {
  std::vector<std::vector<int>> groups;
  std::vector<int64_t> offs;
  for (int s = 0; s < static_cast<int>(h.sets.size()); ++s) {
    const SetDev& S = h.sets[static_cast<size_t>(s)];
    if (S.berm_end == S.berm_beg) continue;
    const auto it = std::find(offs.begin(), offs.end(), S.state_off);
    if (it == offs.end()) {
      offs.push_back(S.state_off);
      groups.push_back({s});
    } else {
      groups[static_cast<size_t>(it - offs.begin())].push_back(s);
    }
  }
  std::vector<ccr_hip_surrogate::Net> nets;
  for (const auto& g : groups) {
    const size_t parts = (g.size() + kSurrogateMaxOutputs - 1) / kSurrogateMaxOutputs;
    for (size_t q = 0; q < parts; ++q) {
      const size_t b0 = g.size() * q / parts, b1 = g.size() * (q + 1) / parts;
      ccr_hip_surrogate::Net net;
      std::vector<int> sets(g.begin() + static_cast<std::ptrdiff_t>(b0), g.begin() + static_cast<std::ptrdiff_t>(b1));
      net.num_dates = h.sets[static_cast<size_t>(sets[0])].num_dates;
      std::vector<int32_t> uni;
      std::vector<double> tOf(static_cast<size_t>(net.num_dates), 0.0);
      for (int s : sets) {
        const SetDev& S = h.sets[static_cast<size_t>(s)];
        ccr_hip_surrogate::Output o;
        o.set = s;
        o.nberms = S.berm_end - S.berm_beg;
        o.cd_base = S.cd_base;
        o.dates = valuedDates(h, s);
        o.rows = static_cast<int32_t>(o.dates.size());
        o.live = liveRows(h, s);
        for (int r = 0; r < o.live; ++r) {
          uni.push_back(o.dates[static_cast<size_t>(r)]);
          tOf[static_cast<size_t>(o.dates[static_cast<size_t>(r)])] = h.ctimes[static_cast<size_t>(o.cd_base + r)];
        }
        net.outs.push_back(std::move(o));
      }
      std::sort(uni.begin(), uni.end());
      uni.erase(std::unique(uni.begin(), uni.end()), uni.end());
      net.urows = static_cast<int>(uni.size());
      net.urow.assign(static_cast<size_t>(net.num_dates), -1);
      for (int u = 0; u < net.urows; ++u) {
        net.urow[static_cast<size_t>(uni[static_cast<size_t>(u)])] = u;
        net.utimes.push_back(tOf[static_cast<size_t>(uni[static_cast<size_t>(u)])]);
      }
      net.map.assign(net.outs.size() * static_cast<size_t>(net.urows), -1);
      for (size_t o = 0; o < net.outs.size(); ++o) {
        const auto& out = net.outs[o];
        for (int r = 0; r < out.live; ++r)
          net.map[o * static_cast<size_t>(net.urows) + static_cast<size_t>(net.urow[static_cast<size_t>(out.dates[static_cast<size_t>(r)])])] = r;
      }
      net.ex = exercisesOf(h, sets);
      net.inputs = kSurrogateBase + net.ex.count;
      nets.push_back(std::move(net));
    }
  }
  return nets;
}

// Whether plan h has the sets, valued dates and continuation rows the surrogate was trained on.
bool surrogateFits(const HostPlan& h, const ccr_hip_surrogate* sur) {
  if (static_cast<int>(h.sets.size()) != sur->num_sets) return false;
  for (const auto& net : sur->nets)
    for (const auto& o : net.outs) {
      const SetDev& S = h.sets[static_cast<size_t>(o.set)];
      if (S.berm_end - S.berm_beg != o.nberms || S.cd_base != o.cd_base || S.num_dates != net.num_dates ||
          valuedDates(h, o.set) != o.dates || liveRows(h, o.set) != o.live)
        return false;
    }
  return true;
}

// Surrogate inputs (the context's feature buffer) of every network with an enabled output,
// inference straight into the Bermudan value rows of its enabled sets, exact zeros past each
// enabled set's last live row.
void surrogateRows(ccr_hip_ctx* ctx, ccr_hip_plan* plan, const ccr_hip_surrogate* sur) {
  const HostPlan& h = plan->host;
  const int64_t P = h.g.num_paths;
  for (const auto& net : sur->nets) {
    bool any = false;
    for (const auto& o : net.outs) any = any || o.enabled;
    if (!any || net.urows == 0) continue;
    if (static_cast<size_t>(net.urows) * P * net.inputs * sizeof(float) > ctx->sur_feats_bytes)
      invalid("surrogate: feature buffer not sized for this plan");
    const SetDev& S = h.sets[static_cast<size_t>(net.outs[0].set)];
    check(launchSurrogateFeatures(plan->d_states, S.state_off, P, net.d_urow, net.d_utimes, net.num_dates, net.ex,
                                  ctx->sur_feats, ctx->stream),
          "surrogate features launch");
    check(mlpInferRows(net.run->device(), ctx->sur_feats, static_cast<int64_t>(net.urows) * P, net.xmean, net.xsd, P,
                       net.urows, net.d_map, net.d_outs, plan->d_bermv, ctx->stream),
          "surrogate inference launch");
    for (const auto& o : net.outs)
      if (o.enabled && o.rows > o.live)
        check(launchZeroRows(plan->d_bermv + (static_cast<int64_t>(o.cd_base) + o.live) * P,
                             static_cast<int64_t>(o.rows - o.live) * P, ctx->stream),
              "surrogate zero rows launch");
  }
}

// Grows the context's surrogate feature buffer to what the networks need on `plan`.
void reserveSurrogateFeatures(ccr_hip_ctx* ctx, const ccr_hip_plan* plan, const ccr_hip_surrogate* sur) {
  size_t need = 0;
  for (const auto& net : sur->nets)
    need = std::max(need, static_cast<size_t>(net.urows) * static_cast<size_t>(plan->host.g.num_paths) *
                              static_cast<size_t>(net.inputs) * sizeof(float));
  if (need <= ctx->sur_feats_bytes) return;
  check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
  if (ctx->sur_feats) (void)hipFree(ctx->sur_feats);
  ctx->sur_feats = nullptr;
  ctx->sur_feats_bytes = 0;
  check(hipMalloc(reinterpret_cast<void**>(&ctx->sur_feats), need), "hipMalloc");
  ctx->sur_feats_bytes = need;
}

// States pass and regressions (sets with Bermudan swaptions) or the surrogate, then the path
// kernel; events 1 and 2 close the two stages.
template <typename Real>
void launchPathStages(ccr_hip_ctx* ctx, ccr_hip_plan* plan, int blocks, const ccr_hip_surrogate* sur) {
  const HostPlan& h = plan->host;
  PathArgs<Real> a = pathArgs<Real>(plan);
  if (!h.berms.empty()) {
    PathArgs<Real> s = a;
    s.states = plan->d_states;
    s.states_set = h.states_set;
    const int sblocks = h.states_set >= 0 ? h.sets[static_cast<size_t>(h.states_set)].num_blocks : blocks;
    check(launchStates<Real>(s, sblocks, h.block, ctx->stream), "states kernel launch");
    bool allSurrogate = sur != nullptr;
    for (int si = 0; sur && si < static_cast<int>(h.sets.size()); ++si) {
      if (h.sets[static_cast<size_t>(si)].berm_end == h.sets[static_cast<size_t>(si)].berm_beg) continue;
      bool on = false;
      for (const auto& net : sur->nets)
        for (const auto& o : net.outs) on = on || (o.set == si && o.enabled);
      allSurrogate = allSurrogate && on;
    }
    if (!allSurrogate)
      check(launchAmc<Real>(plan->amc, a.bonds, a.fixed, h.max_exercises, ctx->stream), "regression kernel launch");
    if (sur) surrogateRows(ctx, plan, sur);
  }
  markEvent(ctx, 1);
  check(launchPath<Real>(a, blocks, h.block, ctx->stream), "path kernel launch");
  markEvent(ctx, 2);
}

// wsBase: the memory the plan's per-execute buffers are bound to for this execute (null: the
// context's shared workspace, grown if needed).
void launchAll(ccr_hip_ctx* ctx, ccr_hip_plan* plan, const ccr_hip_surrogate* sur = nullptr, void* wsBase = nullptr) {
  if (ctx->fault == CCR_FAULT_LAUNCH) {
    ctx->fault = CCR_FAULT_NONE;
    fail(CCR_ERR_DEVICE, "kernel launch failed (forced by ccr_hip_set_fault)");
  }
  check(hipSetDevice(ctx->device), "hipSetDevice");
  if (wsBase) bindSlots(plan, wsBase);
  else bindWorkspace(ctx, plan);
  const HostPlan& h = plan->host;
  const int nsets = static_cast<int>(h.sets.size());
  const int blocks = static_cast<int>(h.block_set.size());
  markEvent(ctx, 0);
  if (h.precision == CCR_FP32) launchPathStages<float>(ctx, plan, blocks, sur);
  else launchPathStages<double>(ctx, plan, blocks, sur);
  check(launchReduce(plan->d_sets, plan->d_partials, nsets, h.max_rep, h.total_waves / nsets,
                     1.0 / static_cast<double>(h.g.num_paths), plan->d_stats, ctx->stream),
        "reduce kernel launch");
  if (plan->tstride > 0) {
    check(launchTradeReduce(plan->d_sets, plan->d_tpart, nsets, h.max_rep, plan->tstride, h.total_waves / nsets,
                            h.g.num_paths, plan->d_tstats, ctx->stream),
          "per-trade reduce kernel launch");
    plan->tstats_ready = false;
  }
  markEvent(ctx, 3);
  if (h.precision == CCR_FP32)
    check(launchPfe<float>(plan->d_sets, static_cast<const float*>(plan->d_rows), nsets, h.max_rep, h.g.num_paths,
                           h.pfe_q, plan->d_pfe, ctx->stream),
          "pfe kernel launch");
  else
    check(launchPfe<double>(plan->d_sets, static_cast<const double*>(plan->d_rows), nsets, h.max_rep, h.g.num_paths,
                            h.pfe_q, plan->d_pfe, ctx->stream),
          "pfe kernel launch");
  if (plan->d_im_q) {
    const int dates = h.max_dates;
    if (dates > 1) {
      if (h.precision == CCR_FP32)
        check(launchPfe<float>(plan->d_sets, static_cast<const float*>(plan->d_im_rows), nsets, dates - 1,
                               h.g.num_paths, plan->im_quantile, plan->d_im_q, ctx->stream, 1),
              "initial-margin select launch");
      else
        check(launchPfe<double>(plan->d_sets, static_cast<const double*>(plan->d_im_rows), nsets, dates - 1,
                                h.g.num_paths, plan->im_quantile, plan->d_im_q, ctx->stream, 1),
              "initial-margin select launch");
    }
    check(launchImReduce(plan->d_sets, plan->d_im_part, nsets, dates, h.total_waves / nsets, plan->d_im_dsum,
                         ctx->stream),
          "initial-margin reduce launch");
    plan->im_ready = false;
  }
  if (plan->d_ntarget) {
    NestedArgs n{};
    n.target = plan->d_ntarget;
    n.inner = plan->nested_inner;
    n.level = plan->nested_q;
    n.seed = plan->seed;
    n.rows_raw = plan->d_nrows;
    n.partials = plan->d_npart;
    const int64_t per = static_cast<int64_t>(nsets) * h.max_rep;
    if (h.precision == CCR_FP32) {
      check(launchNested<float>(pathArgs<float>(plan), n, blocks, h.block, ctx->stream), "nested initial margin launch");
      for (int k = 0; k < 3; ++k)
        check(launchPfe<float>(plan->d_sets, static_cast<const float*>(plan->d_nrows), nsets, h.max_rep, h.g.num_paths,
                               kNestedBands[k], plan->d_nbands + k * per, ctx->stream),
              "nested initial margin select launch");
    } else {
      check(launchNested<double>(pathArgs<double>(plan), n, blocks, h.block, ctx->stream), "nested initial margin launch");
      for (int k = 0; k < 3; ++k)
        check(launchPfe<double>(plan->d_sets, static_cast<const double*>(plan->d_nrows), nsets, h.max_rep,
                                h.g.num_paths, kNestedBands[k], plan->d_nbands + k * per, ctx->stream),
              "nested initial margin select launch");
    }
    check(launchReduce(plan->d_sets, plan->d_npart, nsets, h.max_rep, h.total_waves / nsets,
                       1.0 / static_cast<double>(h.g.num_paths), plan->d_nstats, ctx->stream),
          "nested initial margin reduce launch");
  }
  markEvent(ctx, 4);
  ctx->timing_pending = !ctx->capturing;
  ctx->stage_events = true;
  ctx->upload_ms = plan->upload_ms;
  ctx->draws_ms = plan->draws_ms;
}

// Graph mode: the plan's whole execute (every launch of launchAll and its timing events) captured
// once into a HIP graph and replayed; captured again when the shared workspace it was bound to
// has moved. The workspace is bound (and grown) before the capture, outside it.
void launchGraph(ccr_hip_ctx* ctx, ccr_hip_plan* plan) {
  check(hipSetDevice(ctx->device), "hipSetDevice");
  bindWorkspace(ctx, plan);
  if (!plan->graph_exec || plan->graph_ws != ctx->ws) {
    if (plan->graph_exec) {
      check(hipGraphExecDestroy(plan->graph_exec), "hipGraphExecDestroy");
      plan->graph_exec = nullptr;
    }
    hipGraph_t graph = nullptr;
    check(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    ctx->capturing = true;
    try {
      launchAll(ctx, plan);
    } catch (...) {
      ctx->capturing = false;
      (void)hipStreamEndCapture(ctx->stream, &graph);
      if (graph) (void)hipGraphDestroy(graph);
      throw;
    }
    ctx->capturing = false;
    check(hipStreamEndCapture(ctx->stream, &graph), "hipStreamEndCapture");
    const hipError_t e = hipGraphInstantiate(&plan->graph_exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) plan->graph_exec = nullptr;
    check(e, "hipGraphInstantiate");
    plan->graph_ws = ctx->ws;
  }
  markEvent(ctx, 0);
  check(hipGraphLaunch(plan->graph_exec, ctx->stream), "hipGraphLaunch");
  markEvent(ctx, 4);
  if (plan->d_im_q) plan->im_ready = false;
  if (plan->tstride > 0) plan->tstats_ready = false;
  ctx->timing_pending = true;
  ctx->stage_events = false;
  ctx->upload_ms = plan->upload_ms;
  ctx->draws_ms = plan->draws_ms;
}

void launchExecute(ccr_hip_ctx* ctx, ccr_hip_plan* plan) {
  if (ctx->use_graph && ctx->fault == CCR_FAULT_NONE) launchGraph(ctx, plan);
  else launchAll(ctx, plan);
}

// Reads the event timings of the last launch (waits for it to finish).
void readTimings(ccr_hip_ctx* ctx) {
  if (!ctx->timing_pending) return;
  ctx->timing_pending = false;
  check(hipEventSynchronize(ctx->ev[4]), "hipEventSynchronize");
  float m = 0, a = 0, b = 0, c = 0, t = 0;
  if (ctx->stage_events) {
    check(hipEventElapsedTime(&m, ctx->ev[0], ctx->ev[1]), "hipEventElapsedTime");
    check(hipEventElapsedTime(&a, ctx->ev[1], ctx->ev[2]), "hipEventElapsedTime");
    check(hipEventElapsedTime(&b, ctx->ev[2], ctx->ev[3]), "hipEventElapsedTime");
    check(hipEventElapsedTime(&c, ctx->ev[3], ctx->ev[4]), "hipEventElapsedTime");
  }
  check(hipEventElapsedTime(&t, ctx->ev[0], ctx->ev[4]), "hipEventElapsedTime");
  ctx->timing.amc_ms = m;
  ctx->timing.path_ms = a;
  ctx->timing.reduce_ms = b;
  ctx->timing.pfe_ms = c;
  ctx->timing.total_ms = t;
  ctx->timing.upload_ms = ctx->upload_ms;
  ctx->timing.draws_ms = ctx->draws_ms;
}

void readBack(ccr_hip_ctx* ctx, ccr_hip_plan* plan, ccr_exposure_out* outs) {
  const HostPlan& h = plan->host;
  const size_t nsets = h.sets.size();
  std::vector<double> stats(nsets * static_cast<size_t>(h.max_rep) * kStats), pfe(nsets * static_cast<size_t>(h.max_rep));
  check(hipMemcpyAsync(stats.data(), plan->d_stats, stats.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream),
        "hipMemcpyAsync");
  check(hipMemcpyAsync(pfe.data(), plan->d_pfe, pfe.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream),
        "hipMemcpyAsync");
  if (plan->d_im_q) {
    const size_t n = nsets * static_cast<size_t>(h.max_dates);
    plan->im_q.resize(n);
    plan->im_dsum.resize(n);
    check(hipMemcpyAsync(plan->im_q.data(), plan->d_im_q, n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream),
          "hipMemcpyAsync");
    check(hipMemcpyAsync(plan->im_dsum.data(), plan->d_im_dsum, n * sizeof(double), hipMemcpyDeviceToHost,
                         ctx->stream),
          "hipMemcpyAsync");
  }
  if (plan->tstride > 0) {
    plan->tstats.resize(nsets * static_cast<size_t>(h.max_rep) * static_cast<size_t>(plan->tstride));
    check(hipMemcpyAsync(plan->tstats.data(), plan->d_tstats, plan->tstats.size() * sizeof(double),
                         hipMemcpyDeviceToHost, ctx->stream),
          "hipMemcpyAsync");
  }
  check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
  plan->im_ready = plan->d_im_q != nullptr;
  plan->tstats_ready = plan->tstride > 0;
  readTimings(ctx);
  const int64_t P = h.g.num_paths;
  for (size_t s = 0; s < nsets; ++s) {
    const int R = h.sets[s].num_rep;
    if (outs[s].capacity < R) invalid("output capacity smaller than the reporting grid");
  }
  for (size_t s = 0; s < nsets; ++s) {
    ccr_exposure_out& o = outs[s];
    const int R = h.sets[s].num_rep;
    o.num_dates = R;
    for (int r = 0; r < R; ++r) {
      const double* st = stats.data() + (s * static_cast<size_t>(h.max_rep) + static_cast<size_t>(r)) * kStats;
      if (o.times) o.times[r] = h.rep_times[s][static_cast<size_t>(r)];
      if (o.expected_value) o.expected_value[r] = st[0];
      if (o.expected_exposure) o.expected_exposure[r] = st[1];
      if (o.expected_negative_exposure) o.expected_negative_exposure[r] = st[2];
      if (o.discounted_expected_exposure) o.discounted_expected_exposure[r] = st[3];
      if (o.discounted_expected_negative_exposure) o.discounted_expected_negative_exposure[r] = st[4];
      if (o.mean_deflator) o.mean_deflator[r] = st[5];
      if (o.mean_survival) o.mean_survival[r] = st[6];
      if (o.cva_increments) o.cva_increments[r] = st[7];
      if (o.dva_increments) o.dva_increments[r] = st[8];
      if (o.potential_future_exposure) o.potential_future_exposure[r] = pfe[s * static_cast<size_t>(h.max_rep) + static_cast<size_t>(r)];
    }
    if (o.exposure_rows) {
      const size_t n = static_cast<size_t>(R) * static_cast<size_t>(P);
      const size_t off = static_cast<size_t>(h.sets[s].rows_off);
      if (h.precision == CCR_FP32) {
        std::vector<float> f(n);
        check(hipMemcpy(f.data(), static_cast<float*>(plan->d_rows) + off, n * sizeof(float), hipMemcpyDeviceToHost), "hipMemcpy");
        for (size_t i = 0; i < n; ++i) o.exposure_rows[i] = f[i];
      } else {
        check(hipMemcpy(o.exposure_rows, static_cast<double*>(plan->d_rows) + off, n * sizeof(double), hipMemcpyDeviceToHost),
              "hipMemcpy");
      }
    }
  }
}

// Expected exposure per set and reporting date of the last launch on `plan`.
std::vector<std::vector<double>> expectedExposure(ccr_hip_ctx* ctx, ccr_hip_plan* plan) {
  const HostPlan& h = plan->host;
  const size_t ns = h.sets.size();
  std::vector<std::vector<double>> ee(ns, std::vector<double>(static_cast<size_t>(h.max_rep), 0.0));
  std::vector<ccr_exposure_out> outs(ns);
  for (size_t s = 0; s < ns; ++s) {
    outs[s] = ccr_exposure_out{};
    outs[s].capacity = h.max_rep;
    outs[s].expected_exposure = ee[s].data();
  }
  readBack(ctx, plan, outs.data());
  for (size_t s = 0; s < ns; ++s) ee[s].resize(static_cast<size_t>(outs[s].num_dates));
  return ee;
}

int32_t writeGrid(const TimeGrid& g, double* out, int32_t cap, int32_t* size) {
  if (size) *size = static_cast<int32_t>(g.size());
  if (out)
    for (size_t i = 0; i < g.size() && static_cast<int32_t>(i) < cap; ++i) out[i] = g[i];
  return CCR_OK;
}

// ---------------------------------------------------------------- allocation-free execution
// ccr_hip_execute_noalloc binds a plan's per-execute buffers (exposure rows, wave partials,
// American Monte Carlo states, exercise values, regressions and valuation rows) to this array
// instead of allocating. Its name is a host-side stand-in: host code uses the device address
// hipGetSymbolAddress gives for it, taken once per device in ccr_hip_create (with that device
// current) and kept in gScratchPtr. Every kernel writes each element of these buffers before
// reading it in the same execute, so nothing needs clearing between calls.
constexpr size_t kScratchBytes = static_cast<size_t>(64) << 20;
constexpr int kScratchDevices = 64;
std::mutex gScratchMutex;
void* gScratchPtr[kScratchDevices] = {};

}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

// The device address of `scratch` on `device` (current), looked up on the first call per device.
void scratchInit(int device) {
  if (device < 0 || device >= kScratchDevices) fail(CCR_ERR_UNSUPPORTED, "device ordinal beyond the scratch table");
  std::lock_guard<std::mutex> lock(gScratchMutex);
  if (gScratchPtr[device]) return;
  void* p = nullptr;
  check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress (scratch)");
  gScratchPtr[device] = p;
}

// Strided copy of one statistic of one set (element k of every kStats-wide date record) into
// the caller's array: a 2-D copy, no staging buffer.
void copyStat(double* dst, const double* stats, int k, int R, hipStream_t stream) {
  if (!dst || R <= 0) return;
  check(hipMemcpy2DAsync(dst, sizeof(double), stats + k, kStats * sizeof(double), sizeof(double),
                         static_cast<size_t>(R), hipMemcpyDeviceToHost, stream),
        "hipMemcpy2DAsync");
}

}  // namespace

extern "C" {

int32_t ccr_hip_abi_version(void) { return CCR_HIP_ABI_VERSION; }

int32_t ccr_hip_create(const ccr_hip_options* options, ccr_hip_ctx** out) {
  if (!out) return CCR_ERR_INVALID_ARGUMENT;
  *out = nullptr;
  auto ctx = std::make_unique<ccr_hip_ctx>();
  if (options) {
    ctx->device = options->device;
    ctx->precision = options->precision;
    ctx->block = options->block_size;
  }
  if (ctx->block < 0 || ctx->block % 64 != 0 || ctx->block > kMaxBlock) {
    gCreateError = "block_size must be a multiple of 64 up to 256 (the path kernel's launch bound)";
    return CCR_ERR_INVALID_ARGUMENT;
  }
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n == 0) {
    gCreateError = std::string("no HIP device: ") + hipGetErrorString(e);
    return CCR_ERR_DEVICE;
  }
  if (ctx->device < 0 || ctx->device >= n) {
    gCreateError = "device ordinal out of range";
    return CCR_ERR_OUT_OF_RANGE;
  }
  e = hipSetDevice(ctx->device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
  for (int i = 0; i < 5 && e == hipSuccess; ++i) e = hipEventCreate(&ctx->ev[i]);
  if (e != hipSuccess) {
    gCreateError = std::string("HIP initialisation failed: ") + hipGetErrorString(e);
    return CCR_ERR_DEVICE;
  }
  // The allocation-free entry point's scratch on this device (its address, once per device).
  try {
    scratchInit(ctx->device);
  } catch (const std::exception& x) {
    gCreateError = x.what();
    return CCR_ERR_DEVICE;
  }
  gCreateError.clear();
  *out = ctx.release();
  return CCR_OK;
}

void ccr_hip_destroy(ccr_hip_ctx* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  for (auto& e : ctx->ev)
    if (e) (void)hipEventDestroy(e);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  if (ctx->ws) (void)hipFree(ctx->ws);
  if (ctx->sur_feats) (void)hipFree(ctx->sur_feats);
  delete ctx;
}

const char* ccr_hip_last_error(const ccr_hip_ctx* ctx) { return ctx ? ctx->err.c_str() : gCreateError.c_str(); }

int32_t ccr_hip_set_nested_initial_margin(ccr_hip_ctx* ctx, int32_t inner_paths, double quantile,
                                          double margin_period_of_risk) {
  if (!ctx) return CCR_ERR_INVALID_ARGUMENT;
  return guarded(ctx, [&] {
    if (inner_paths == 0) {
      ctx->nested_inner = 0;
      return;
    }
    // The checks of gpu::compile (src/gpu/fused_exposure.cpp:220-246).
    if (inner_paths < 2 || inner_paths > 1024) invalid("nested initial margin: 2 to 1024 inner paths");
    if (!(quantile >= 0.0 && quantile <= 1.0)) invalid("imQuantile must be in [0, 1]");
    const double rank = quantile * static_cast<double>(inner_paths - 1);
    if (static_cast<double>(inner_paths - 1) - std::floor(rank) >= static_cast<double>(kNestedTop - 1))
      invalid("nested initial margin: the quantile needs more inner paths (keeps the 16 largest)");
    if (!(margin_period_of_risk > 0.0)) invalid("nested initial margin: the margin period must be positive");
    ctx->nested_inner = inner_paths;
    ctx->nested_q = quantile;
    ctx->nested_mpor = margin_period_of_risk;
  });
}

int32_t ccr_hip_nested_initial_margin(ccr_hip_ctx* ctx, const ccr_hip_plan* plan, int32_t set, ccr_nested_im_out* out) {
  if (!ctx) return CCR_ERR_INVALID_ARGUMENT;
  return guarded(ctx, [&] {
    if (!plan || !out) invalid("null argument");
    if (!plan->d_ntarget) invalid("nested initial margin: the plan was compiled without ccr_hip_set_nested_initial_margin");
    const HostPlan& h = plan->host;
    const int nsets = static_cast<int>(h.sets.size());
    if (set < 0 || set >= nsets) fail(CCR_ERR_OUT_OF_RANGE, "netting set index out of range");
    const int R = h.sets[static_cast<size_t>(set)].num_rep;
    if (out->capacity < R) fail(CCR_ERR_OUT_OF_RANGE, "nested initial margin: output capacity too small");
    check(hipSetDevice(ctx->device), "hipSetDevice");
    const size_t per = static_cast<size_t>(nsets) * h.max_rep;
    std::vector<double> st(per * kStats), bands(3 * per);
    check(hipMemcpyAsync(st.data(), plan->d_nstats, st.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream),
          "hipMemcpyAsync");
    check(hipMemcpyAsync(bands.data(), plan->d_nbands, bands.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream),
          "hipMemcpyAsync");
    check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    double* dst[5] = {out->discounted, out->mean, out->band05, out->band50, out->band95};
    for (int r = 0; r < R; ++r) {
      const size_t i = static_cast<size_t>(set) * h.max_rep + static_cast<size_t>(r);
      const double v[5] = {st[i * kStats], st[i * kStats + 1], bands[i], bands[per + i], bands[2 * per + i]};
      for (int k = 0; k < 5; ++k)
        if (dst[k]) dst[k][r] = v[k];
    }
    out->num_dates = R;
  });
}

int32_t ccr_hip_set_initial_margin(ccr_hip_ctx* ctx, double quantile, double margin_period_of_risk) {
  if (!ctx) return CCR_ERR_INVALID_ARGUMENT;
  if (quantile > 1.0 || (quantile > 0.0 && !(margin_period_of_risk > 0.0))) {
    ctx->err = "initial margin: quantile in (0, 1] and a positive margin period of risk";
    return CCR_ERR_INVALID_ARGUMENT;
  }
  ctx->im_quantile = quantile > 0.0 ? quantile : 0.0;
  ctx->im_mpor = margin_period_of_risk;
  return CCR_OK;
}

int32_t ccr_hip_initial_margin(ccr_hip_ctx* ctx, const ccr_hip_plan* plan, int32_t set, double* margin,
                               int32_t capacity, int32_t* size) {
  if (!ctx) return CCR_ERR_INVALID_ARGUMENT;
  return guarded(ctx, [&] {
    if (!plan || !margin) invalid("null argument");
    if (!plan->d_im_q) invalid("initial margin: the plan was compiled without ccr_hip_set_initial_margin");
    if (!plan->im_ready) invalid("initial margin: execute the plan (ccr_hip_execute) first");
    const HostPlan& h = plan->host;
    if (set < 0 || set >= static_cast<int32_t>(h.sets.size())) fail(CCR_ERR_OUT_OF_RANGE, "netting set index out of range");
    const std::vector<double>& t = h.sim_times[static_cast<size_t>(set)];
    const int32_t n = static_cast<int32_t>(t.size());
    if (size) *size = n;
    if (capacity < n) invalid("initial margin: capacity smaller than the simulation grid");
    // initialMarginProfile (src/xva/xva.cpp:59-76): q-quantile of |V(t_k) - V(t_k-1)| over the
    // paths x sqrt(mpor / dt) x the sum of the deflators / P; 0 at t = 0.
    // Quantiles: one per date after the first, (max_dates - 1) per set; deflator sums: one per
    // date, max_dates per set.
    const size_t qBase = static_cast<size_t>(set) * static_cast<size_t>(h.max_dates - 1);
    const size_t dBase = static_cast<size_t>(set) * static_cast<size_t>(h.max_dates);
    const double paths = static_cast<double>(h.g.num_paths);
    margin[0] = 0.0;
    for (int32_t k = 1; k < n; ++k) {
      const double scale = std::sqrt(plan->im_mpor / (t[static_cast<size_t>(k)] - t[static_cast<size_t>(k - 1)]));
      margin[k] = plan->im_q[qBase + static_cast<size_t>(k - 1)] * scale * plan->im_dsum[dBase + static_cast<size_t>(k)] / paths;
    }
  });
}

int32_t ccr_hip_set_graph(ccr_hip_ctx* ctx, int32_t enable) {
  if (!ctx) return CCR_ERR_INVALID_ARGUMENT;
  ctx->use_graph = enable != 0;
  return CCR_OK;
}

int32_t ccr_hip_set_book(ccr_hip_ctx* ctx, int32_t enable) {
  if (!ctx) return CCR_ERR_INVALID_ARGUMENT;
  ctx->book = enable != 0;
  return CCR_OK;
}

int32_t ccr_hip_set_book_group(ccr_hip_ctx* ctx, int32_t sets_per_workgroup) {
  if (!ctx) return CCR_ERR_INVALID_ARGUMENT;
  return guarded(ctx, [&] {
    if (sets_per_workgroup > 1 && !groupSupported(sets_per_workgroup))
      fail(CCR_ERR_UNSUPPORTED, "book kernel: 1, 2, 4 or 8 netting sets per workgroup");
    ctx->book_group = sets_per_workgroup;
  });
}

int32_t ccr_hip_set_regression(ccr_hip_ctx* ctx, int32_t basis, int32_t gram_precision) {
  if (!ctx) return CCR_ERR_INVALID_ARGUMENT;
  return guarded(ctx, [&] {
    if (basis != -1 && basis != CCR_BASIS_STATE && basis != CCR_BASIS_SWAP_RATE_ANNUITY)
      invalid("regression basis must be -1, CCR_BASIS_STATE or CCR_BASIS_SWAP_RATE_ANNUITY");
    if (gram_precision != CCR_FP64 && gram_precision != CCR_FP32)
      invalid("Gram precision must be CCR_FP64 or CCR_FP32");
    ctx->basis = basis;
    ctx->gram32 = gram_precision == CCR_FP32 ? 1 : 0;
  });
}

int32_t ccr_hip_set_book_grids(ccr_hip_ctx* ctx, const double* reporting, int32_t reporting_size,
                               const double* simulation, int32_t simulation_size) {
  if (!ctx) return CCR_ERR_INVALID_ARGUMENT;
  return guarded(ctx, [&] {
    if (reporting_size <= 0 || simulation_size <= 0 || !reporting || !simulation) {
      ctx->book_rep.clear();
      ctx->book_sim.clear();
      return;
    }
    ctx->book_rep.assign(reporting, reporting + reporting_size);
    ctx->book_sim.assign(simulation, simulation + simulation_size);
  });
}

int32_t ccr_hip_book_grids(const double* base_grid, int32_t base_size, const ccr_netting_set* sets, int32_t num_sets,
                           double* reporting, int32_t reporting_capacity, int32_t* reporting_size, double* simulation,
                           int32_t simulation_capacity, int32_t* simulation_size) {
  return guarded(nullptr, [&] {
    if (!sets || num_sets <= 0 || (!base_grid && base_size > 0)) invalid("null argument");
    const TimeGrid base(std::vector<double>(base_grid, base_grid + std::max(0, base_size)));
    const TimeGrid rep = bookReportingGrid(base, sets, num_sets);
    writeGrid(rep, reporting, reporting_capacity, reporting_size);
    writeGrid(bookSimulationGrid(rep, sets, num_sets), simulation, simulation_capacity, simulation_size);
  });
}

int32_t ccr_book_totals(const ccr_exposure_out* outs, const ccr_netting_set* sets, int32_t num_sets,
                        int32_t counterparty, ccr_exposure_out* total) {
  if (!outs || !sets || !total || num_sets <= 0) return CCR_ERR_INVALID_ARGUMENT;
  // sumProfiles (src/app/analyses.cpp:1769-1784) plus the pathwise CVA / DVA increments, which are
  // linear in the exposure on common scenarios; PFE does not add across netting sets.
  int32_t n = -1;
  for (int32_t s = 0; s < num_sets; ++s) {
    if (counterparty >= 0 && sets[s].counterparty_credit != counterparty) continue;
    if (n < 0) n = outs[s].num_dates;
    if (outs[s].num_dates != n) return CCR_ERR_INVALID_ARGUMENT;  // not one book: grids differ
    if (outs[s].times && outs[0].times)
      for (int32_t k = 0; k < n; ++k)
        if (outs[s].times[k] != outs[0].times[k]) return CCR_ERR_INVALID_ARGUMENT;
  }
  if (n < 0) return CCR_ERR_OUT_OF_RANGE;  // no netting set of this counterparty
  if (total->capacity < n) return CCR_ERR_INVALID_ARGUMENT;
  total->num_dates = n;
  double* const dst[] = {total->expected_value, total->expected_exposure, total->expected_negative_exposure,
                         total->discounted_expected_exposure, total->discounted_expected_negative_exposure,
                         total->cva_increments, total->dva_increments};
  for (int m = 0; m < 7; ++m)
    if (dst[m])
      for (int32_t k = 0; k < n; ++k) dst[m][k] = 0.0;
  bool first = true;
  for (int32_t s = 0; s < num_sets; ++s) {
    if (counterparty >= 0 && sets[s].counterparty_credit != counterparty) continue;
    const ccr_exposure_out& o = outs[s];
    const double* const src[] = {o.expected_value, o.expected_exposure, o.expected_negative_exposure,
                                 o.discounted_expected_exposure, o.discounted_expected_negative_exposure,
                                 o.cva_increments, o.dva_increments};
    for (int m = 0; m < 7; ++m)
      if (dst[m] && src[m])
        for (int32_t k = 0; k < n; ++k) dst[m][k] += src[m][k];
    if (first) {
      for (int32_t k = 0; k < n; ++k) {
        if (total->times && o.times) total->times[k] = o.times[k];
        if (total->mean_deflator && o.mean_deflator) total->mean_deflator[k] = o.mean_deflator[k];
        if (total->mean_survival && o.mean_survival) total->mean_survival[k] = o.mean_survival[k];
        if (total->potential_future_exposure) total->potential_future_exposure[k] = 0.0;
      }
      first = false;
    }
  }
  return CCR_OK;
}

int32_t ccr_hip_set_trade_statistics(ccr_hip_ctx* ctx, int32_t enable) {
  if (!ctx) return CCR_ERR_INVALID_ARGUMENT;
  ctx->per_trade = enable != 0;
  return CCR_OK;
}

int32_t ccr_hip_trade_statistics(ccr_hip_ctx* ctx, const ccr_hip_plan* plan, int32_t set, int32_t trade,
                                 ccr_trade_statistics* out) {
  if (!ctx) return CCR_ERR_INVALID_ARGUMENT;
  return guarded(ctx, [&] {
    if (!plan || !out) invalid("null argument");
    if (plan->tstride == 0) invalid("per-trade statistics: the plan was compiled without ccr_hip_set_trade_statistics");
    if (!plan->tstats_ready) invalid("per-trade statistics: execute the plan (ccr_hip_execute) first");
    const HostPlan& h = plan->host;
    if (set < 0 || set >= static_cast<int32_t>(h.sets.size())) fail(CCR_ERR_OUT_OF_RANGE, "netting set index out of range");
    const SetDev& S = h.sets[static_cast<size_t>(set)];
    if (trade < 0 || trade >= S.num_groups) fail(CCR_ERR_OUT_OF_RANGE, "trade index out of range");
    if (out->capacity < S.num_rep) invalid("output capacity smaller than the reporting grid");
    out->num_dates = S.num_rep;
    double* dst[kTradeStats] = {out->marginal_discounted, out->marginal, out->incremental_discounted,
                                out->standalone_discounted, out->standalone};
    const size_t base = static_cast<size_t>(set) * static_cast<size_t>(h.max_rep) * static_cast<size_t>(plan->tstride);
    for (int m = 0; m < kTradeStats; ++m) {
      if (!dst[m]) continue;
      for (int r = 0; r < S.num_rep; ++r)
        dst[m][r] = plan->tstats[base + static_cast<size_t>(r) * static_cast<size_t>(plan->tstride) +
                                 static_cast<size_t>(trade) * kTradeStats + static_cast<size_t>(m)];
    }
  });
}

int32_t ccr_hip_set_shared_workspace(ccr_hip_ctx* ctx, int32_t enable) {
  if (!ctx) return CCR_ERR_INVALID_ARGUMENT;
  ctx->share_ws = enable != 0;
  return CCR_OK;
}

int32_t ccr_hip_set_fault(ccr_hip_ctx* ctx, int32_t fault) {
  if (!ctx) return CCR_ERR_INVALID_ARGUMENT;
  if (fault < CCR_FAULT_NONE || fault > CCR_FAULT_PLAN) return CCR_ERR_INVALID_ARGUMENT;
  ctx->fault = fault;
  return CCR_OK;
}

int32_t ccr_hip_normals(uint64_t seed, int64_t n, double* out) {
  return guarded(nullptr, [&] {
    if (n < 0 || (n > 0 && !out)) invalid("ccr_hip_normals: bad arguments");
    parallelNormals(seed, n, out);
  });
}

int32_t ccr_hip_reporting_grid(const double* base_grid, int32_t base_size, const ccr_netting_set* set, double* times,
                               int32_t capacity, int32_t* size) {
  return guarded(nullptr, [&] {
    if (!set) invalid("null netting set");
    writeGrid(reportingGrid(TimeGrid(std::vector<double>(base_grid, base_grid + std::max(0, base_size))), *set), times,
              capacity, size);
  });
}

int32_t ccr_hip_simulation_grid(const double* base_grid, int32_t base_size, const ccr_netting_set* set, double* times,
                                int32_t capacity, int32_t* size) {
  return guarded(nullptr, [&] {
    if (!set) invalid("null netting set");
    writeGrid(simulationGrid(TimeGrid(std::vector<double>(base_grid, base_grid + std::max(0, base_size))), *set), times,
              capacity, size);
  });
}

int32_t ccr_hip_compile(ccr_hip_ctx* ctx, const ccr_scenario_generator* generator, const ccr_simulation_config* config,
                        const double* base_grid, int32_t base_size, const ccr_netting_set* sets, int32_t num_sets,
                        double pfe_quantile, int32_t precision, ccr_hip_plan** out) {
  if (!ctx) return CCR_ERR_INVALID_ARGUMENT;
  return guarded(ctx, [&] {
    compilePlan(ctx, generator, config, base_grid, base_size, sets, num_sets, pfe_quantile, precision, out);
  });
}

int32_t ccr_hip_plan_info_get(const ccr_hip_plan* plan, ccr_hip_plan_info* info) {
  if (!plan || !info) return CCR_ERR_INVALID_ARGUMENT;
  const HostPlan& h = plan->host;
  info->num_sets = static_cast<int32_t>(h.sets.size());
  info->num_paths = h.g.num_paths;
  info->max_sim_dates = h.max_dates;
  info->max_reporting_dates = h.max_rep;
  info->num_normals = h.g.nz;
  info->total_terms = h.total_terms;
  info->stream_normals = h.stream_len;
  info->revaluations = h.revaluations;
  info->term_evaluations = h.term_evaluations;
  const int64_t elem = h.precision == CCR_FP32 ? 4 : 8;
  int64_t rows = 0, normals = 0;
  const int64_t blocks = h.g.antithetic ? (static_cast<int64_t>(h.g.num_paths) + 1) / 2 : h.g.num_paths;
  for (const auto& s : h.sets) {
    rows += static_cast<int64_t>(s.num_rep) * h.g.num_paths;
    normals += blocks * s.stream_stride;  // each set replays its own slice of the stream
  }
  // Normals read by the path kernel, exposure rows written once and read by the PFE select.
  info->normals_read = normals;
  info->hbm_bytes = normals * elem + 2 * rows * elem;
  info->precision = h.precision;
  info->block_size = h.block;
  info->num_bermudans = static_cast<int32_t>(h.berms.size());
  info->num_exercises = static_cast<int32_t>(h.exlist.size());
  info->num_continuation_dates = static_cast<int32_t>(h.cdates.size() / 2);
  info->max_exercises = h.max_exercises;
  return CCR_OK;
}

int32_t ccr_hip_execute_async(ccr_hip_ctx* ctx, ccr_hip_plan* plan) {
  if (!ctx) return CCR_ERR_INVALID_ARGUMENT;
  return guarded(ctx, [&] {
    if (!plan) invalid("null plan");
    launchExecute(ctx, plan);
  });
}

int32_t ccr_hip_synchronize(ccr_hip_ctx* ctx) {
  if (!ctx) return CCR_ERR_INVALID_ARGUMENT;
  return guarded(ctx, [&] {
    check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    check(hipDeviceSynchronize(), "hipDeviceSynchronize");
  });
}

int32_t ccr_hip_execute(ccr_hip_ctx* ctx, ccr_hip_plan* plan, ccr_exposure_out* outs) {
  if (!ctx) return CCR_ERR_INVALID_ARGUMENT;
  return guarded(ctx, [&] {
    if (!plan || !outs) invalid("null argument");
    launchExecute(ctx, plan);
    readBack(ctx, plan, outs);
  });
}

int32_t ccr_hip_execute_noalloc(ccr_hip_ctx* ctx, ccr_hip_plan* plan, ccr_exposure_out* outs) {
  if (!ctx) return CCR_ERR_INVALID_ARGUMENT;
  // No std::function wrapper here: the call allocates nothing on the device and nothing on the
  // host unless it fails (the error message).
  try {
    if (!plan || !outs) invalid("null argument");
    if (plan->ws_slots.empty())
      invalid("ccr_hip_execute_noalloc: compile the plan after ccr_hip_set_shared_workspace(ctx, 1)");
    if (plan->ws_need > kScratchBytes)
      fail(CCR_ERR_OUT_OF_RANGE, "ccr_hip_execute_noalloc: the plan needs " + std::to_string(plan->ws_need) +
                                     " bytes of per-execute buffers; the scratch holds at most " +
                                     std::to_string(kScratchBytes) + " (64 MiB)");
    if (plan->d_im_q || plan->tstride > 0)
      fail(CCR_ERR_UNSUPPORTED, "ccr_hip_execute_noalloc: plans with the initial-margin proxy or per-trade "
                                "statistics need ccr_hip_execute");
    const HostPlan& h = plan->host;
    const size_t nsets = h.sets.size();
    const int64_t P = h.g.num_paths;
    for (size_t s = 0; s < nsets; ++s) {
      if (outs[s].capacity < h.sets[s].num_rep) invalid("output capacity smaller than the reporting grid");
      if (outs[s].exposure_rows && h.precision == CCR_FP32)
        fail(CCR_ERR_UNSUPPORTED, "ccr_hip_execute_noalloc: FP32 exposure rows need ccr_hip_execute (widening)");
    }
    void* base = ctx->device >= 0 && ctx->device < kScratchDevices ? gScratchPtr[ctx->device] : nullptr;
    if (!base) fail(CCR_ERR_DEVICE, "ccr_hip_execute_noalloc: no scratch address for this device");
    launchAll(ctx, plan, nullptr, base);
    for (size_t s = 0; s < nsets; ++s) {
      ccr_exposure_out& o = outs[s];
      const int R = h.sets[s].num_rep;
      const double* st = plan->d_stats + s * static_cast<size_t>(h.max_rep) * kStats;
      if (o.times) std::copy(h.rep_times[s].begin(), h.rep_times[s].begin() + R, o.times);
      copyStat(o.expected_value, st, 0, R, ctx->stream);
      copyStat(o.expected_exposure, st, 1, R, ctx->stream);
      copyStat(o.expected_negative_exposure, st, 2, R, ctx->stream);
      copyStat(o.discounted_expected_exposure, st, 3, R, ctx->stream);
      copyStat(o.discounted_expected_negative_exposure, st, 4, R, ctx->stream);
      copyStat(o.mean_deflator, st, 5, R, ctx->stream);
      copyStat(o.mean_survival, st, 6, R, ctx->stream);
      copyStat(o.cva_increments, st, 7, R, ctx->stream);
      copyStat(o.dva_increments, st, 8, R, ctx->stream);
      if (o.potential_future_exposure)
        check(hipMemcpyAsync(o.potential_future_exposure, plan->d_pfe + s * static_cast<size_t>(h.max_rep),
                             static_cast<size_t>(R) * sizeof(double), hipMemcpyDeviceToHost, ctx->stream),
              "hipMemcpyAsync");
      if (o.exposure_rows)
        check(hipMemcpyAsync(o.exposure_rows, static_cast<const double*>(plan->d_rows) + h.sets[s].rows_off,
                             static_cast<size_t>(R) * static_cast<size_t>(P) * sizeof(double), hipMemcpyDeviceToHost,
                             ctx->stream),
              "hipMemcpyAsync");
    }
    check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    for (size_t s = 0; s < nsets; ++s) outs[s].num_dates = h.sets[s].num_rep;
    readTimings(ctx);
    ctx->err.clear();
    return CCR_OK;
  } catch (const Error& e) {
    ctx->err = e.what();
    return e.code;
  } catch (const std::exception& e) {
    ctx->err = e.what();
    return CCR_ERR_RUNTIME;
  }
}

int32_t ccr_hip_last_timing(const ccr_hip_ctx* ctx, ccr_hip_timing* timing) {
  if (!ctx || !timing) return CCR_ERR_INVALID_ARGUMENT;
  ccr_hip_ctx* c = const_cast<ccr_hip_ctx*>(ctx);
  const int32_t rc = guarded(c, [&] { readTimings(c); });
  *timing = ctx->timing;
  return rc;
}

void ccr_hip_plan_destroy(ccr_hip_plan* plan) {
  if (!plan) return;
  (void)hipSetDevice(plan->device);
  delete plan;
}

int32_t ccr_hip_run_exposure(ccr_hip_ctx* ctx, const ccr_scenario_generator* generator, const ccr_simulation_config* config,
                             const double* base_grid, int32_t base_size, const ccr_netting_set* sets, int32_t num_sets,
                             double pfe_quantile, int32_t precision, ccr_exposure_out* outs) {
  if (!ctx) return CCR_ERR_INVALID_ARGUMENT;
  return guarded(ctx, [&] {
    if (!outs) invalid("null outputs");
    ccr_hip_plan* raw = nullptr;
    compilePlan(ctx, generator, config, base_grid, base_size, sets, num_sets, pfe_quantile, precision, &raw);
    std::unique_ptr<ccr_hip_plan, void (*)(ccr_hip_plan*)> plan(raw, ccr_hip_plan_destroy);
    launchAll(ctx, plan.get());
    readBack(ctx, plan.get(), outs);
  });
}

int32_t ccr_hip_autotune(ccr_hip_ctx* ctx, const ccr_scenario_generator* generator, const ccr_simulation_config* config,
                         const double* base_grid, int32_t base_size, const ccr_netting_set* sets, int32_t num_sets,
                         int32_t precision, int64_t sample_paths, int32_t* block_size, double* times_ms) {
  if (!ctx) return CCR_ERR_INVALID_ARGUMENT;
  const int keep = ctx->block;
  const int32_t rc = guarded(ctx, [&] {
    if (!config || sample_paths <= 0) invalid("ccr_hip_autotune: invalid arguments");
    ccr_simulation_config sample = *config;
    sample.num_paths = std::min<int64_t>(config->num_paths, sample_paths);
    const int candidates[4] = {64, 128, 192, 256};
    int best = keep;
    double bestMs = 1e300;
    for (int i = 0; i < 4; ++i) {
      ctx->block = candidates[i];
      ccr_hip_plan* raw = nullptr;
      compilePlan(ctx, generator, &sample, base_grid, base_size, sets, num_sets, 0.95, precision, &raw);
      std::unique_ptr<ccr_hip_plan, void (*)(ccr_hip_plan*)> plan(raw, ccr_hip_plan_destroy);
      double ms = 1e300;
      for (int rep = 0; rep < 4; ++rep) {
        launchAll(ctx, plan.get());
        readTimings(ctx);
        if (rep > 0) ms = std::min(ms, ctx->timing.path_ms);  // the first execute warms up
      }
      if (times_ms) times_ms[i] = ms;
      if (ms < bestMs) {
        bestMs = ms;
        best = candidates[i];
      }
    }
    ctx->block = best;
    if (block_size) *block_size = best;
  });
  if (rc != CCR_OK) ctx->block = keep;
  return rc;
}

int32_t ccr_hip_autotune_book(ccr_hip_ctx* ctx, const ccr_scenario_generator* generator,
                              const ccr_simulation_config* config, const double* base_grid, int32_t base_size,
                              const ccr_netting_set* sets, int32_t num_sets, int32_t precision, int64_t sample_paths,
                              int32_t* block_size, int32_t* sets_per_workgroup, double* times_ms) {
  if (!ctx) return CCR_ERR_INVALID_ARGUMENT;
  const int keepBlock = ctx->block, keepGroup = ctx->book_group;
  const int32_t rc = guarded(ctx, [&] {
    if (!config || sample_paths <= 0) invalid("ccr_hip_autotune_book: invalid arguments");
    if (!ctx->book) invalid("ccr_hip_autotune_book: the context is not in book mode");
    ccr_simulation_config sample = *config;
    sample.num_paths = std::min<int64_t>(config->num_paths, sample_paths);
    const int blocks[4] = {64, 128, 192, 256}, groups[4] = {1, 2, 4, 8};
    int bestBlock = keepBlock, bestGroup = keepGroup;
    double bestMs = 1e300;
    for (int i = 0; i < 4; ++i)
      for (int k = 0; k < 4; ++k) {
        if (times_ms) times_ms[4 * i + k] = 0.0;
        ctx->block = blocks[i];
        ctx->book_group = groups[k];
        ccr_hip_plan* raw = nullptr;
        compilePlan(ctx, generator, &sample, base_grid, base_size, sets, num_sets, 0.95, precision, &raw);
        std::unique_ptr<ccr_hip_plan, void (*)(ccr_hip_plan*)> plan(raw, ccr_hip_plan_destroy);
        // A group the plan cannot take (one set, options, the IM proxy) runs as group 1: only
        // the first such candidate is timed.
        const int effective = plan->host.group > 1 ? plan->host.group : 1;
        if (effective != groups[k]) continue;
        // A shape whose LDS does not fit a workgroup is skipped.
        int wg = 0;
        const int ldsRows = plan->host.g.max_slots + effective * (plan->host.g.max_ring + plan->host.price_rows);
        const hipError_t occ = precision == CCR_FP32
                                   ? pathOccupancy<float>(plan->host.g.nrates, plan->host.g.nassets, plan->host.g.ncredits,
                                                          !plan->host.opts.empty(), ldsRows, blocks[i], &wg, false,
                                                          effective)
                                   : pathOccupancy<double>(plan->host.g.nrates, plan->host.g.nassets, plan->host.g.ncredits,
                                                           !plan->host.opts.empty(), ldsRows, blocks[i], &wg, false,
                                                           effective);
        if (occ != hipSuccess || wg <= 0) continue;
        double ms = 1e300;
        for (int rep = 0; rep < 4; ++rep) {
          launchAll(ctx, plan.get());
          readTimings(ctx);
          if (rep > 0) ms = std::min(ms, ctx->timing.path_ms);  // the first execute warms up
        }
        if (times_ms) times_ms[4 * i + k] = ms;
        if (ms < bestMs) {
          bestMs = ms;
          bestBlock = blocks[i];
          bestGroup = groups[k];
        }
      }
    ctx->block = bestBlock;
    ctx->book_group = bestGroup;
    if (block_size) *block_size = bestBlock;
    if (sets_per_workgroup) *sets_per_workgroup = bestGroup;
  });
  if (rc != CCR_OK) {
    ctx->block = keepBlock;
    ctx->book_group = keepGroup;
  }
  return rc;
}

int32_t ccr_hip_bermudan(ccr_hip_ctx* ctx, const ccr_scenario_generator* generator, const ccr_simulation_config* config,
                         const double* grid, int32_t grid_size, const ccr_trade* trade, int32_t precision, double* price,
                         int32_t* exercise_index) {
  if (!ctx) return CCR_ERR_INVALID_ARGUMENT;
  return guarded(ctx, [&] {
    if (!trade || !price) invalid("null argument");
    if (trade->kind != CCR_TRADE_BERMUDAN_SWAPTION) invalid("ccr_hip_bermudan: the trade is not a Bermudan swaption");
    ccr_netting_set ns{};
    ns.num_trades = 1;
    ns.trades = trade;
    ns.counterparty_credit = -1;
    ns.own_credit = -1;
    ccr_hip_plan* raw = nullptr;
    compilePlan(ctx, generator, config, grid, grid_size, &ns, 1, 0.95, precision, &raw);
    std::unique_ptr<ccr_hip_plan, void (*)(ccr_hip_plan*)> plan(raw, ccr_hip_plan_destroy);
    launchAll(ctx, plan.get());
    const HostPlan& h = plan->host;
    const int64_t P = h.g.num_paths;
    // Date 0 is always valued (the grid starts at 0): its row is the first continuation row.
    std::vector<double> v0(static_cast<size_t>(P));
    check(hipMemcpyAsync(v0.data(), plan->d_bermv, v0.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream),
          "hipMemcpyAsync");
    std::vector<int16_t> ex(static_cast<size_t>(P));
    check(hipMemcpyAsync(ex.data(), plan->d_exidx, ex.size() * sizeof(int16_t), hipMemcpyDeviceToHost, ctx->stream),
          "hipMemcpyAsync");
    check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    double sum = 0.0;
    for (double v : v0) sum += v;
    *price = sum / static_cast<double>(P);
    if (exercise_index) {
      const int ne = h.berms[0].ne;
      for (int64_t p = 0; p < P; ++p) exercise_index[p] = ex[static_cast<size_t>(p)] >= ne ? -1 : ex[static_cast<size_t>(p)];
    }
  });
}

double ccr_expected_positive_exposure(const double* times, const double* ee, int32_t size, double horizon) {
  // ExposureProfile::timeAverage: src/exposure/exposure_profile.cpp:52-61.
  double integral = 0.0, end = 0.0;
  for (int32_t j = 1; j < size && times[j] <= horizon + 1e-10; ++j) {
    integral += ee[j] * (times[j] - times[j - 1]);
    end = times[j];
  }
  return end > 0.0 ? integral / end : (size > 0 ? ee[0] : 0.0);
}

void ccr_effective_expected_exposure(const double* ee, int32_t size, double* eff) {
  double running = 0.0;
  for (int32_t j = 0; j < size; ++j) {
    running = std::max(running, ee[j]);
    eff[j] = running;
  }
}

double ccr_effective_expected_positive_exposure(const double* times, const double* ee, int32_t size, double horizon) {
  std::vector<double> eff(static_cast<size_t>(std::max(size, 0)));
  ccr_effective_expected_exposure(ee, size, eff.data());
  return ccr_expected_positive_exposure(times, eff.data(), size, horizon);
}

double ccr_unilateral_cva(const double* times, const double* discounted_ee, int32_t size,
                          const ccr_credit_curve* counterparty, double recovery) {
  // unilateralCva: src/cva/cva.cpp:11-33 (trapezoid in EE* against the default probabilities).
  const Credit c(*counterparty);
  double cva = 0.0;
  for (int32_t i = 1; i < size; ++i) {
    const double ee = 0.5 * (discounted_ee[i - 1] + discounted_ee[i]);
    cva += (1.0 - recovery) * ee * (c.survival(times[i - 1]) - c.survival(times[i]));
  }
  return cva;
}

void ccr_bilateral_cva(const double* t, const double* dee, const double* dene, int32_t size,
                       const ccr_credit_curve* counterparty, double rc, const ccr_credit_curve* own, double ro,
                       double* cva, double* dva) {
  // bilateralCva: src/cva/cva.cpp:34-48 (first-to-default via the other party's survival at midpoints).
  const Credit c(*counterparty), o(*own);
  double a = 0.0, b = 0.0;
  for (int32_t i = 1; i < size; ++i) {
    const double mid = 0.5 * (t[i - 1] + t[i]);
    a += (1.0 - rc) * 0.5 * (dee[i - 1] + dee[i]) * o.survival(mid) * (c.survival(t[i - 1]) - c.survival(t[i]));
    b += (1.0 - ro) * 0.5 * (-dene[i - 1] - dene[i]) * c.survival(mid) * (o.survival(t[i - 1]) - o.survival(t[i]));
  }
  if (cva) *cva = a;
  if (dva) *dva = b;
}

void ccr_xva_defaults(ccr_xva_parameters* p) {
  if (!p) return;
  *p = ccr_xva_parameters{0.0100, 0.0, 0.0100, 0.10, 0.08, 1.0, 1.4};
}

namespace {

// int f(t) Q_c(t) Q_own(t) dt: trapezoid in f, both survivals at the interval mid-points
// (xva.cpp:18-27).
double survivalWeighted(const double* t, const double* f, int32_t n, const Credit& c, const Credit& o, double sign) {
  double total = 0.0;
  for (int32_t i = 1; i < n; ++i) {
    const double mid = 0.5 * (t[i - 1] + t[i]);
    total += 0.5 * (sign * f[i - 1] + sign * f[i]) * (t[i] - t[i - 1]) * c.survival(mid) * o.survival(mid);
  }
  return total;
}

}  // namespace

int32_t ccr_exposure_xva(const double* t, const double* dee, int32_t n, const ccr_credit_curve* counterparty,
                         double rc, const ccr_credit_curve* own, const ccr_xva_parameters* p, double* out) {
  if (!t || !dee || !counterparty || !own || !p || !out || n < 0) return CCR_ERR_INVALID_ARGUMENT;
  const Credit c(*counterparty), o(*own);
  double cva = 0.0;
  for (int32_t i = 1; i < n; ++i)
    cva += (1.0 - rc) * 0.5 * (dee[i - 1] + dee[i]) * o.survival(0.5 * (t[i - 1] + t[i])) *
           (c.survival(t[i - 1]) - c.survival(t[i]));
  const double weighted = survivalWeighted(t, dee, n, c, o, 1.0);
  out[0] = cva;
  out[1] = p->funding_spread_borrow * weighted;
  out[2] = p->hurdle_rate * p->capital_ratio * p->risk_weight * p->alpha * weighted;
  return CCR_OK;
}

int32_t ccr_allocate_exposure_xva(const double* t, const double* contributions, int32_t num_trades, int32_t n,
                                  const ccr_credit_curve* counterparty, double rc, const ccr_credit_curve* own,
                                  const ccr_xva_parameters* p, double* out) {
  if (!contributions || num_trades < 0) return CCR_ERR_INVALID_ARGUMENT;
  for (int32_t i = 0; i < num_trades; ++i) {
    const int32_t rcode = ccr_exposure_xva(t, contributions + static_cast<int64_t>(i) * n, n, counterparty, rc, own, p,
                                           out + 3 * static_cast<int64_t>(i));
    if (rcode != CCR_OK) return rcode;
  }
  return CCR_OK;
}

double ccr_funding_benefit(const double* t, const double* dene, int32_t n, const ccr_credit_curve* counterparty,
                           const ccr_credit_curve* own, double spread) {
  if (!t || !dene || !counterparty || !own) return 0.0;
  return spread * survivalWeighted(t, dene, n, Credit(*counterparty), Credit(*own), -1.0);
}

double ccr_margin_valuation_adjustment(const double* t, const double* dim, int32_t n,
                                       const ccr_credit_curve* counterparty, const ccr_credit_curve* own,
                                       double spread) {
  if (!t || !dim || !counterparty || !own) return 0.0;
  return spread * survivalWeighted(t, dim, n, Credit(*counterparty), Credit(*own), 1.0);
}

int32_t ccr_xva(const double* t, const double* dee, const double* dene, int32_t n, const double* dim,
                const ccr_credit_curve* counterparty, double rc, const ccr_credit_curve* own, double ro,
                const ccr_xva_parameters* p, ccr_xva_result* out) {
  if (!dene || !out) return CCR_ERR_INVALID_ARGUMENT;
  double e[3];
  const int32_t rc2 = ccr_exposure_xva(t, dee, n, counterparty, rc, own, p, e);
  if (rc2 != CCR_OK) return rc2;
  const Credit c(*counterparty), o(*own);
  ccr_xva_result r{};
  r.cva = e[0];
  r.fca = e[1];
  r.kva = e[2];
  for (int32_t i = 1; i < n; ++i)
    r.dva += (1.0 - ro) * 0.5 * (-dene[i - 1] - dene[i]) * c.survival(0.5 * (t[i - 1] + t[i])) *
             (o.survival(t[i - 1]) - o.survival(t[i]));
  r.fba = ccr_funding_benefit(t, dene, n, counterparty, own, p->funding_spread_lend);
  if (dim) r.mva = ccr_margin_valuation_adjustment(t, dim, n, counterparty, own, p->margin_funding_spread);
  *out = r;
  return CCR_OK;
}

int32_t ccr_hip_mlp_train(ccr_hip_ctx* ctx, const ccr_mlp_desc* desc, const ccr_mlp_train_options* o, const double* X,
                          const double* Y, int64_t n, double* params, double* stdz, double* loss, double* train_ms) {
  if (!ctx) return CCR_ERR_INVALID_ARGUMENT;
  return guarded(ctx, [&] {
    if (!desc || !o || !X || !Y || !params || !stdz) invalid("null argument");
    if (o->epochs < 0) invalid("Mlp: negative epochs");
    check(hipSetDevice(ctx->device), "hipSetDevice");
    const int din = desc->inputs, dout = desc->outputs;
    MlpRun run(*desc, o->precision, n, ctx->stream);
    double* xm = stdz;
    double* xsd = stdz + din;
    double* ym = stdz + 2 * din;
    double* ysd = stdz + 2 * din + dout;
    if (!o->warm) {
      standardiserFit(X, n, din, xm, xsd);
      standardiserFit(Y, n, dout, ym, ysd);
      mlpInit(run.device().sh, o->seed, params);
    }
    std::vector<double> xs(static_cast<size_t>(n * din)), ys(static_cast<size_t>(n * dout));
    for (int64_t r = 0; r < n; ++r) {
      for (int j = 0; j < din; ++j) xs[static_cast<size_t>(r * din + j)] = (X[r * din + j] - xm[j]) / xsd[j];
      for (int k = 0; k < dout; ++k) ys[static_cast<size_t>(r * dout + k)] = (Y[r * dout + k] - ym[k]) / ysd[k];
    }
    run.setParams(params);
    run.setInputs(xs.data());
    run.setTargets(ys.data());
    const auto t0 = std::chrono::steady_clock::now();
    run.train(o->epochs, o->learning_rate, o->weight_decay, o->batch);
    if (train_ms) *train_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (loss) *loss = run.loss();
    run.getParams(params);
  });
}

int32_t ccr_hip_surrogate_train(ccr_hip_ctx* ctx, ccr_hip_plan* train, ccr_hip_plan* holdout,
                                const ccr_surrogate_options* o, ccr_hip_surrogate** out, ccr_surrogate_report* report) {
  if (!ctx) return CCR_ERR_INVALID_ARGUMENT;
  return guarded(ctx, [&] {
    if (!train || !holdout || !o || !out) invalid("null argument");
    if (train == holdout) invalid("surrogate: the hold-out plan must differ from the training plan");
    if (!train->ws_slots.empty() || !holdout->ws_slots.empty())
      invalid("surrogate: plans compiled with the shared workspace cannot train a surrogate");
    const HostPlan& ht = train->host;
    const HostPlan& hh = holdout->host;
    if (ht.sets.size() != hh.sets.size()) invalid("surrogate: the plans hold different netting sets");
    for (size_t s = 0; s < ht.sets.size(); ++s)
      if (ht.sets[s].berm_end - ht.sets[s].berm_beg != hh.sets[s].berm_end - hh.sets[s].berm_beg)
        invalid("surrogate: the plans hold different netting sets");
    if (ht.berms.empty()) invalid("surrogate: no netting set holds a Bermudan swaption");
    if (o->epochs < 0 || o->gate_threshold < 0.0) invalid("surrogate: invalid options");
    check(hipSetDevice(ctx->device), "hipSetDevice");
    auto sur = std::make_unique<ccr_hip_surrogate>();
    sur->device = ctx->device;
    sur->num_sets = static_cast<int>(ht.sets.size());
    sur->nets = surrogateNets(ht);
    if (!surrogateFits(hh, sur.get())) invalid("surrogate: the plans hold different netting sets");
    ccr_surrogate_report rep{};
    // Method-generated pathwise values: the training plan executed with the regressions.
    launchAll(ctx, train);
    check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    const int64_t P = ht.g.num_paths;
    const int64_t nTrain = o->train_paths > 0 ? std::min<int64_t>(o->train_paths, P) : P;
    for (auto& net : sur->nets) {
      net.d_urow = sur->upload(net.urow, ctx->stream);
      net.d_utimes = sur->upload(net.utimes, ctx->stream);
      const int in = net.inputs, nout = static_cast<int>(net.outs.size());
      const int U = net.urows;
      // Inputs of the training rows (u, p < nTrain) and targets: each output's value on the dates
      // its set values (NaN elsewhere: a missing target the loss skips).
      std::vector<float> f(static_cast<size_t>(U) * P * in);
      if (U > 0) {
        float* feats = nullptr;
        check(hipMalloc(reinterpret_cast<void**>(&feats), f.size() * sizeof(float)), "hipMalloc");
        std::unique_ptr<float, hipError_t (*)(void*)> featsOwner(feats, hipFree);
        const SetDev& S = ht.sets[static_cast<size_t>(net.outs[0].set)];
        check(launchSurrogateFeatures(train->d_states, S.state_off, P, net.d_urow, net.d_utimes, net.num_dates, net.ex,
                                      feats, ctx->stream),
              "surrogate features launch");
        check(hipMemcpyAsync(f.data(), feats, f.size() * sizeof(float), hipMemcpyDeviceToHost, ctx->stream),
              "hipMemcpyAsync");
        check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
      }
      const int64_t n = static_cast<int64_t>(U) * nTrain;
      if (n == 0) invalid("surrogate: no live Bermudan rows to train on");
      std::vector<double> X(static_cast<size_t>(n * in)), Y(static_cast<size_t>(n * nout), std::nan(""));
      for (int u = 0; u < U; ++u)
        for (int64_t p = 0; p < nTrain; ++p)
          for (int i = 0; i < in; ++i)
            X[static_cast<size_t>((u * nTrain + p) * in + i)] = f[static_cast<size_t>((u * P + p) * in + i)];
      for (int k = 0; k < nout; ++k) {
        auto& out = net.outs[static_cast<size_t>(k)];
        std::vector<double> v(static_cast<size_t>(out.rows) * P);
        check(hipMemcpyAsync(v.data(), train->d_bermv + static_cast<int64_t>(out.cd_base) * P, v.size() * sizeof(double),
                             hipMemcpyDeviceToHost, ctx->stream),
              "hipMemcpyAsync");
        check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
        double sum = 0.0, cnt = 0.0;
        for (int u = 0; u < U; ++u) {
          const int32_t r = net.map[static_cast<size_t>(k) * U + u];
          if (r < 0) continue;
          for (int64_t p = 0; p < nTrain; ++p) {
            const double y = v[static_cast<size_t>(r * P + p)];
            Y[static_cast<size_t>((u * nTrain + p) * nout + k)] = y;
            sum += y;
            cnt += 1.0;
          }
        }
        // Standardiser of this output over its present targets (mean, population sd; 0 -> 1).
        out.ymean = cnt > 0.0 ? sum / cnt : 0.0;
        double ss = 0.0;
        for (int64_t i = 0; i < n; ++i) {
          const double y = Y[static_cast<size_t>(i * nout + k)];
          if (y == y) ss += (y - out.ymean) * (y - out.ymean);
        }
        out.ysd = cnt > 0.0 && ss > 0.0 ? std::sqrt(ss / cnt) : 1.0;
        for (int64_t i = 0; i < n; ++i) {
          double& y = Y[static_cast<size_t>(i * nout + k)];
          if (y == y) y = (y - out.ymean) / out.ysd;
        }
      }
      standardiserFit(X.data(), n, in, net.xmean, net.xsd);
      for (int64_t i = 0; i < n; ++i)
        for (int j = 0; j < in; ++j)
          X[static_cast<size_t>(i * in + j)] = (X[static_cast<size_t>(i * in + j)] - net.xmean[j]) / net.xsd[j];
      if (o->batch > 0) {  // mini-batches of consecutive rows: shuffle the rows once
        std::vector<int64_t> perm(static_cast<size_t>(n));
        for (int64_t i = 0; i < n; ++i) perm[static_cast<size_t>(i)] = i;
        std::mt19937_64 rng(o->seed);
        std::shuffle(perm.begin(), perm.end(), rng);
        std::vector<double> X2(X.size()), Y2(Y.size());
        for (int64_t i = 0; i < n; ++i) {
          const int64_t q = perm[static_cast<size_t>(i)];
          for (int j = 0; j < in; ++j) X2[static_cast<size_t>(i * in + j)] = X[static_cast<size_t>(q * in + j)];
          for (int k = 0; k < nout; ++k) Y2[static_cast<size_t>(i * nout + k)] = Y[static_cast<size_t>(q * nout + k)];
        }
        X.swap(X2);
        Y.swap(Y2);
      }
      const ccr_mlp_desc desc{in, o->hidden, o->hidden_layers, nout, CCR_ACT_SILU};
      std::vector<double> params(static_cast<size_t>(mlpShape(desc).count));
      {
        MlpRun trainer(desc, CCR_BF16, n, ctx->stream);
        mlpInit(trainer.device().sh, o->seed, params.data());
        trainer.setParams(params.data());
        trainer.setInputs(X.data());
        trainer.setTargets(Y.data());
        const auto t0 = std::chrono::steady_clock::now();
        trainer.train(o->epochs, o->learning_rate, o->weight_decay, o->batch);
        rep.train_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        rep.loss = std::max(rep.loss, trainer.loss());
        trainer.getParams(params.data());
      }
      net.run = std::make_unique<MlpRun>(desc, CCR_BF16, kMlpRows, ctx->stream);
      net.run->setParams(params.data());
      std::vector<InferOut> io;
      for (const auto& out : net.outs) io.push_back(InferOut{out.cd_base, out.ymean, out.ysd});
      net.d_outs = sur->upload(io, ctx->stream);
      net.d_map = sur->upload(net.map, ctx->stream);
      rep.samples += n;
      rep.candidate_sets += nout;
      rep.networks += 1;
    }
    // The gate: hold-out EE with the surrogate against the regression method, per netting set.
    launchAll(ctx, holdout);
    const auto plain = expectedExposure(ctx, holdout);
    reserveSurrogateFeatures(ctx, holdout, sur.get());
    launchAll(ctx, holdout, sur.get());
    const auto approx = expectedExposure(ctx, holdout);
    for (auto& net : sur->nets) {
      std::vector<int32_t> map = net.map;
      for (size_t k = 0; k < net.outs.size(); ++k) {
        auto& out = net.outs[k];
        const auto& a = plain[static_cast<size_t>(out.set)];
        const auto& b = approx[static_cast<size_t>(out.set)];
        double peak = 0.0, err = 0.0;
        for (size_t r = 0; r < a.size(); ++r) {
          peak = std::max(peak, std::fabs(a[r]));
          err = std::max(err, std::fabs(b[r] - a[r]));
        }
        out.ee_error = peak > 0.0 ? err / peak : err;
        out.enabled = out.ee_error <= o->gate_threshold;
        rep.max_ee_error = std::max(rep.max_ee_error, out.ee_error);
        rep.enabled_sets += out.enabled ? 1 : 0;
        if (!out.enabled)
          for (int u = 0; u < net.urows; ++u) map[k * static_cast<size_t>(net.urows) + static_cast<size_t>(u)] = -1;
      }
      check(hipMemcpyAsync(net.d_map, map.data(), map.size() * sizeof(int32_t), hipMemcpyHostToDevice, ctx->stream),
            "hipMemcpyAsync");
      check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    }
    if (report) *report = rep;
    *out = sur.release();
  });
}

int32_t ccr_hip_surrogate_set_info(const ccr_hip_surrogate* sur, int32_t set, int32_t* enabled, double* ee_error) {
  if (!sur || set < 0 || set >= sur->num_sets) return CCR_ERR_INVALID_ARGUMENT;
  if (enabled) *enabled = -1;
  if (ee_error) *ee_error = 0.0;
  for (const auto& net : sur->nets)
    for (const auto& o : net.outs)
      if (o.set == set) {
        if (enabled) *enabled = o.enabled ? 1 : 0;
        if (ee_error) *ee_error = o.ee_error;
      }
  return CCR_OK;
}

int32_t ccr_hip_execute_surrogate_async(ccr_hip_ctx* ctx, ccr_hip_plan* plan, const ccr_hip_surrogate* sur) {
  if (!ctx) return CCR_ERR_INVALID_ARGUMENT;
  return guarded(ctx, [&] {
    if (!plan || !sur) invalid("null argument");
    if (sur->device != ctx->device) invalid("surrogate: trained on another device");
    if (!surrogateFits(plan->host, sur)) invalid("surrogate: trained for other netting sets");
    reserveSurrogateFeatures(ctx, plan, sur);
    launchAll(ctx, plan, sur);
  });
}

int32_t ccr_hip_execute_surrogate(ccr_hip_ctx* ctx, ccr_hip_plan* plan, const ccr_hip_surrogate* sur,
                                  ccr_exposure_out* outs) {
  const int32_t rc = ccr_hip_execute_surrogate_async(ctx, plan, sur);
  if (rc != CCR_OK) return rc;
  return guarded(ctx, [&] {
    if (!outs) invalid("null outputs");
    readBack(ctx, plan, outs);
  });
}

void ccr_hip_surrogate_destroy(ccr_hip_surrogate* sur) {
  if (!sur) return;
  (void)hipSetDevice(sur->device);
  delete sur;
}

int32_t ccr_hip_mlp_gradient(ccr_hip_ctx* ctx, const ccr_mlp_desc* desc, int32_t precision, const double* params,
                             const double* xs, const double* ys, int64_t n, double* grad, double* loss) {
  if (!ctx) return CCR_ERR_INVALID_ARGUMENT;
  return guarded(ctx, [&] {
    if (!desc || !params || !xs || !ys || !grad) invalid("null argument");
    check(hipSetDevice(ctx->device), "hipSetDevice");
    MlpRun run(*desc, precision, n, ctx->stream);
    run.setParams(params);
    run.setInputs(xs);
    run.setTargets(ys);
    const double l = run.lossAndGradient(grad);
    if (loss) *loss = l;
  });
}

int32_t ccr_hip_mlp_predict(ccr_hip_ctx* ctx, const ccr_mlp_desc* desc, int32_t precision, const double* params,
                            const double* stdz, const double* X, int64_t n, double* Y) {
  if (!ctx) return CCR_ERR_INVALID_ARGUMENT;
  return guarded(ctx, [&] {
    if (!desc || !params || !stdz || !X || !Y) invalid("null argument");
    check(hipSetDevice(ctx->device), "hipSetDevice");
    const int din = desc->inputs, dout = desc->outputs;
    if (precision == CCR_BF16) {
      // The fused inference kernel (weights in LDS, no per-sample buffers) on raw rows.
      MlpRun run(*desc, CCR_BF16, kMlpRows, ctx->stream);
      run.setParams(params);
      std::vector<float> xf(static_cast<size_t>(n * din));
      for (size_t i = 0; i < xf.size(); ++i) xf[i] = static_cast<float>(X[i]);
      float *dx = nullptr, *dy = nullptr;
      check(hipMalloc(reinterpret_cast<void**>(&dx), xf.size() * sizeof(float)), "hipMalloc");
      std::unique_ptr<float, hipError_t (*)(void*)> ownX(dx, hipFree);
      check(hipMalloc(reinterpret_cast<void**>(&dy), static_cast<size_t>(n * dout) * sizeof(float)), "hipMalloc");
      std::unique_ptr<float, hipError_t (*)(void*)> ownY(dy, hipFree);
      check(hipMemcpyAsync(dx, xf.data(), xf.size() * sizeof(float), hipMemcpyHostToDevice, ctx->stream), "hipMemcpyAsync");
      check(mlpInfer(run.device(), dx, n, stdz, stdz + din, dy, ctx->stream), "MLP inference launch");
      std::vector<float> yf(static_cast<size_t>(n * dout));
      check(hipMemcpyAsync(yf.data(), dy, yf.size() * sizeof(float), hipMemcpyDeviceToHost, ctx->stream), "hipMemcpyAsync");
      check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
      for (int64_t r = 0; r < n; ++r)
        for (int k = 0; k < dout; ++k)
          Y[r * dout + k] = yf[static_cast<size_t>(r * dout + k)] * stdz[2 * din + dout + k] + stdz[2 * din + k];
      return;
    }
    MlpRun run(*desc, precision, n, ctx->stream);
    std::vector<double> xs(static_cast<size_t>(n * din));
    for (int64_t r = 0; r < n; ++r)
      for (int j = 0; j < din; ++j) xs[static_cast<size_t>(r * din + j)] = (X[r * din + j] - stdz[j]) / stdz[din + j];
    run.setParams(params);
    run.setInputs(xs.data());
    run.forward();
    std::vector<double> z(static_cast<size_t>(n * dout));
    run.outputs(z.data());
    for (int64_t r = 0; r < n; ++r)
      for (int k = 0; k < dout; ++k)
        Y[r * dout + k] = z[static_cast<size_t>(r * dout + k)] * stdz[2 * din + dout + k] + stdz[2 * din + k];
  });
}

}  // extern "C"
