// C ABI of relaxmap_amd (include/relaxmap.h) and the host-side driver that
// restates the reference's super-step loop around the device kernels.
#include "../../include/relaxmap.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <thread>
#include <vector>

#include "context.h"
#include "netio.h"

using namespace rmx;

namespace rmx {
void batch_partition(Ctx &c, const rmx_partition_opts &o, rmx_partition_result &r);   // batch.hip
}
static_assert(RMX_BUCKET_CUTS == kNumBuckets - 1, "one cut-off per bucket below the hubs");

struct rmx_ctx {
  Ctx c;
  std::vector<std::string> names;
  std::string last_error;
};

namespace {

template <typename F>
int guarded(rmx_ctx *ctx, F &&f) {
  if (!ctx) return RMX_EINVAL;
  try {
    ctx->last_error.clear();
    return f();
  } catch (const HipError &e) {
    ctx->last_error = e.what();
    return RMX_EDEVICE;
  } catch (const std::bad_alloc &) {
    ctx->last_error = "host allocation failed";
    return RMX_EDEVICE;
  } catch (const std::exception &e) {
    ctx->last_error = e.what();
    return RMX_EINVAL;
  }
}

int fail(rmx_ctx *ctx, int code, const std::string &msg) {
  ctx->last_error = msg;
  return code;
}

void free_ctx_device(Ctx &c) {
  c.L0.free_all();
  c.L1.free_all();
  c.L2.free_all();
  c.Lsub.free_all();
  c.node_w.free(); c.raw_w.free(); c.dangling.free(); c.raw_in.free(); c.node_sp.free();
  c.ms.free(); c.qv.free(); c.mNodes.free(); c.mUnits.free(); c.mStale.free(); c.mOne.free(); c.mMulti.free();
  c.b_off.free(); c.b_node.free(); c.b_seed.free(); c.b_S.free(); c.b_L.free(); c.b_C.free();
  c.b_Lold.free(); c.b_Lstep.free(); c.b_Louter.free(); c.b_nmod.free(); c.b_D.free(); c.b_R.free();
  c.b_iters.free(); c.b_step.free(); c.b_outer.free(); c.b_inner.free();
  c.b_ebeg.free(); c.b_ecnt.free(); c.b_ecur.free(); c.b_count.free();
  c.b_cbeg.free(); c.b_cend.free(); c.b_ichunk.free(); c.b_cpart.free(); c.n_chunks = 0;
  c.act.free(); c.act_next.free(); c.wmark.free(); c.target.free(); c.newmod.free(); c.stamp.free(); c.movers.free(); c.movers_st.free();
  c.empty_list.free(); c.want_empty.free(); c.rank_tmp.free(); c.agg_buf.free();
  c.batch_keys.free(); c.batch_list.free();
  c.hub_keys.free(); c.hub_out.free(); c.hub_in.free(); c.hub_plan.free(); c.hub_bm.free(); c.hub_bd.free();
  c.counters.free(); c.scal.free(); c.partial.free(); c.x.free(); c.y.free(); c.xf.free(); c.node_q.free();
  c.scratch.release();
  c.scratch2.release();
  c.cache.release();
}

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

extern "C" {

int rmx_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

int rmx_create(rmx_ctx **out, int device) {
  if (!out) return RMX_EINVAL;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return RMX_EDEVICE;
  if (device < 0 || device >= n) return RMX_EINVAL;
  rmx_ctx *ctx = new rmx_ctx();
  int rc = guarded(ctx, [&] {
    ctx->c.device = device;
    RMX_HIP(hipSetDevice(device));
    RMX_HIP(hipStreamCreateWithFlags(&ctx->c.stream, hipStreamNonBlocking));
    RMX_HIP(hipHostMalloc((void **)&ctx->c.h_scal, sizeof(double) * 16, hipHostMallocDefault));
    RMX_HIP(hipHostMalloc((void **)&ctx->c.h_cnt, sizeof(uint32_t) * 16, hipHostMallocDefault));
    return RMX_OK;
  });
  if (rc != RMX_OK) {
    delete ctx;
    return rc;
  }
  *out = ctx;
  return RMX_OK;
}

void rmx_destroy(rmx_ctx *ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->c.device);
  if (ctx->c.stream) (void)hipStreamSynchronize(ctx->c.stream);
  free_ctx_device(ctx->c);
  if (ctx->c.h_scal) (void)hipHostFree(ctx->c.h_scal);
  if (ctx->c.h_cnt) (void)hipHostFree(ctx->c.h_cnt);
  if (ctx->c.stream) (void)hipStreamDestroy(ctx->c.stream);
  delete ctx;
}

const char *rmx_last_error(const rmx_ctx *ctx) { return ctx ? ctx->last_error.c_str() : "null context"; }

int rmx_set_precision(rmx_ctx *ctx, int precision) {
  return guarded(ctx, [&] {
    if (precision != RMX_FP64 && precision != RMX_MIXED) return fail(ctx, RMX_EINVAL, "precision must be 0 or 1");
    // the arc weights each precision reads are made by rmx_initiate
    if (precision != ctx->c.precision) ctx->c.initiated = false;
    ctx->c.precision = precision;
    return RMX_OK;
  });
}

int rmx_set_batches(rmx_ctx *ctx, int batches) {
  return guarded(ctx, [&] {
    if (batches < 0 || batches > rmx::kMaxBatches)
      return fail(ctx, RMX_EINVAL, "batches must be in 0.." + std::to_string(rmx::kMaxBatches) + " (0: per level)");
    ctx->c.batches = batches;
    return RMX_OK;
  });
}

int rmx_set_fp64_flow(rmx_ctx *ctx, int blocked) {
  return guarded(ctx, [&] {
    if (blocked != 0 && blocked != 1) return fail(ctx, RMX_EINVAL, "FP64 flow path must be 0 or 1");
    ctx->c.fp64_blocked = blocked;
    ctx->c.initiated = false;   // the flows of the other path need a new initiate
    return RMX_OK;
  });
}

int rmx_set_hub_path(rmx_ctx *ctx, int global_tables) {
  return guarded(ctx, [&] {
    if (global_tables < 0 || global_tables > 2) return fail(ctx, RMX_EINVAL, "hub path must be 0, 1 or 2");
    ctx->c.hub_global = global_tables;
    return RMX_OK;
  });
}

int rmx_set_attempt_mode(rmx_ctx *ctx, int sequential) {
  return guarded(ctx, [&] {
    if (sequential != 0 && sequential != 1) return fail(ctx, RMX_EINVAL, "attempt mode must be 0 or 1");
    ctx->c.attempts_sequential = sequential;
    return RMX_OK;
  });
}

int rmx_set_graph_capture(rmx_ctx *ctx, int on) {
  return guarded(ctx, [&] {
    if (on != 0 && on != 1) return fail(ctx, RMX_EINVAL, "graph capture must be 0 or 1");
    ctx->c.use_graph = on;
    return RMX_OK;
  });
}

int rmx_set_candidate_bound(rmx_ctx *ctx, int bucket, int mode) {
  return guarded(ctx, [&] {
    if (bucket < -1 || bucket > 7) return fail(ctx, RMX_EINVAL, "candidate bound: bucket must be -1..7");
    if (mode != 0 && mode != 1) return fail(ctx, RMX_EINVAL, "candidate bound: mode must be 0 or 1");
    for (int b = 0; b < 8; ++b)
      if (bucket < 0 || bucket == b) ctx->c.bound_mode[b] = mode;
    return RMX_OK;
  });
}

}  // extern "C"

namespace {

// f(first_row, end_row) on row ranges of about equal arc counts, one host
// thread each (out_ptr monotone); returns the first range's non-empty message
template <typename F>
const char *rows_in_parallel(int32_t n, int64_t m, const int64_t *out_ptr, F f) {
  const int T = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  if (T == 1 || m < (1 << 20)) return f(0, n);
  std::vector<int32_t> cut(T + 1);
  for (int t = 0; t <= T; ++t)
    cut[t] = t == T ? n : (int32_t)(std::lower_bound(out_ptr, out_ptr + n + 1, m * t / T) - out_ptr);
  std::vector<const char *> msg(T, "");
  std::vector<std::thread> th;
  for (int t = 0; t < T; ++t) th.emplace_back([&, t] { msg[t] = f(cut[t], std::max(cut[t], cut[t + 1])); });
  for (auto &x : th) x.join();
  for (int t = 0; t < T; ++t)
    if (*msg[t]) return msg[t];
  return "";
}

// the out-CSR rules of rmx_graph_upload; "" when they hold
const char *check_csr(int32_t n, int64_t m, const int64_t *out_ptr, const int32_t *out_idx, const double *w) {
  if (n <= 0 || m < 0 || !out_ptr || (m > 0 && (!out_idx || !w))) return "empty network or null arrays";
  if (m > 0x7fffffffLL) return "more than 2^31 - 1 arcs (arc offsets on the device are 32-bit)";
  if (out_ptr[0] != 0 || out_ptr[n] != m) return "out_ptr does not span the arcs";
  for (int32_t i = 0; i < n; ++i)
    if (out_ptr[i + 1] < out_ptr[i]) return "out_ptr not monotone";
  return rows_in_parallel(n, m, out_ptr, [&](int32_t i0, int32_t i1) -> const char * {
    for (int32_t i = i0; i < i1; ++i) {
      for (int64_t k = out_ptr[i]; k < out_ptr[i + 1]; ++k) {
        int32_t d = out_idx[k];
        if (d < 0 || d >= n) return "arc head out of range";
        if (d == i) return "self-loop in arc list";
        if (k > out_ptr[i] && out_idx[k - 1] >= d) return "arcs not sorted or duplicated";
        if (!(w[k] > 0.0) || !std::isfinite(w[k])) return "arc weight not positive";
      }
    }
    return "";
  });
}

void install_batch(Ctx &c, int32_t n_inst, const int32_t *node_off, const uint64_t *seeds, const uint32_t *inst);

// default names "1".."n", made when first needed (an uploaded network has
// none of its own)
void set_names(rmx_ctx *ctx, int32_t n) {
  if ((int32_t)ctx->names.size() != n) {
    ctx->names.resize(n);
    for (int32_t i = 0; i < n; ++i) ctx->names[i] = std::to_string(i + 1);
  }
}

}  // namespace

extern "C" {

int rmx_graph_upload(rmx_ctx *ctx, int32_t n, int64_t m, const int64_t *out_ptr, const int32_t *out_idx,
                     const double *w, const double *node_w) {
  return guarded(ctx, [&] {
    const char *bad = check_csr(n, m, out_ptr, out_idx, w);
    if (*bad) return fail(ctx, RMX_EINVAL, bad);
    RMX_HIP(hipSetDevice(ctx->c.device));
    ctx->c.nb = 0;
    upload_graph(ctx->c, n, m, out_ptr, out_idx, w, node_w);
    ctx->names.clear();
    return RMX_OK;
  });
}

int rmx_batch_upload(rmx_ctx *ctx, int32_t n_inst, const int32_t *node_off, const uint64_t *seeds, int32_t n,
                     int64_t m, const int64_t *out_ptr, const int32_t *out_idx, const double *w,
                     const double *node_w) {
  return guarded(ctx, [&] {
    if (n_inst < 1 || !node_off || !seeds) return fail(ctx, RMX_EINVAL, "no instances or null arrays");
    const char *bad = check_csr(n, m, out_ptr, out_idx, w);
    if (*bad) return fail(ctx, RMX_EINVAL, bad);
    if (node_off[0] != 0 || node_off[n_inst] != n) return fail(ctx, RMX_EINVAL, "node_off does not span the nodes");
    for (int32_t I = 0; I < n_inst; ++I)
      if (node_off[I + 1] < node_off[I]) return fail(ctx, RMX_EINVAL, "node_off not monotone");
    std::vector<uint32_t> inst((size_t)n);
    bad = rows_in_parallel(n, m, out_ptr, [&](int32_t i0, int32_t i1) -> const char * {
      if (i0 >= i1) return "";
      // the instance of row i0: the last I with node_off[I] <= i0
      int32_t I = (int32_t)(std::upper_bound(node_off, node_off + n_inst + 1, i0) - node_off) - 1;
      for (int32_t i = i0; i < i1; ++i) {
        while (i >= node_off[I + 1]) ++I;
        inst[i] = (uint32_t)I;
        for (int64_t k = out_ptr[i]; k < out_ptr[i + 1]; ++k)
          if (out_idx[k] < node_off[I] || out_idx[k] >= node_off[I + 1]) return "an arc joins two instances";
      }
      return "";
    });
    if (*bad) return fail(ctx, RMX_EINVAL, bad);
    Ctx &c = ctx->c;
    RMX_HIP(hipSetDevice(c.device));
    c.nb = 0;
    upload_graph(c, n, m, out_ptr, out_idx, w, node_w);
    install_batch(c, n_inst, node_off, seeds, inst.data());
    ctx->names.clear();
    return RMX_OK;
  });
}

int rmx_batch_upload_seeds(rmx_ctx *ctx, int32_t n_graph, const int32_t *graph_off, int32_t n_seed,
                           const uint64_t *seeds, int32_t n, int64_t m, const int64_t *out_ptr,
                           const int32_t *out_idx, const double *w, const double *node_w) {
  return guarded(ctx, [&] {
    if (n_graph < 1 || n_seed < 1 || !graph_off || !seeds) return fail(ctx, RMX_EINVAL, "no networks, no seeds or null arrays");
    const char *bad = check_csr(n, m, out_ptr, out_idx, w);
    if (*bad) return fail(ctx, RMX_EINVAL, bad);
    if (graph_off[0] != 0 || graph_off[n_graph] != n) return fail(ctx, RMX_EINVAL, "graph_off does not span the nodes");
    for (int32_t g = 0; g < n_graph; ++g)
      if (graph_off[g + 1] < graph_off[g]) return fail(ctx, RMX_EINVAL, "graph_off not monotone");
    if ((int64_t)n * n_seed > 0x7fffffffLL || m * n_seed > 0x7fffffffLL)
      return fail(ctx, RMX_EINVAL, "the copies exceed 2^31 - 1 nodes or arcs");
    bad = rows_in_parallel(n, m, out_ptr, [&](int32_t i0, int32_t i1) -> const char * {
      if (i0 >= i1) return "";
      int32_t g = (int32_t)(std::upper_bound(graph_off, graph_off + n_graph + 1, i0) - graph_off) - 1;
      for (int32_t i = i0; i < i1; ++i) {
        while (i >= graph_off[g + 1]) ++g;
        for (int64_t k = out_ptr[i]; k < out_ptr[i + 1]; ++k)
          if (out_idx[k] < graph_off[g] || out_idx[k] >= graph_off[g + 1]) return "an arc joins two networks";
      }
      return "";
    });
    if (*bad) return fail(ctx, RMX_EINVAL, bad);
    Ctx &c = ctx->c;
    RMX_HIP(hipSetDevice(c.device));
    c.nb = 0;
    upload_graph_replicated(c, n_graph, graph_off, n_seed, n, m, out_ptr, out_idx, w, node_w);
    const int32_t n_inst = n_graph * n_seed;
    std::vector<int32_t> node_off((size_t)n_inst + 1);
    int64_t at = 0;
    for (int32_t g = 0; g < n_graph; ++g)
      for (int32_t r = 0; r < n_seed; ++r) {
        node_off[(size_t)g * n_seed + r] = (int32_t)at;
        at += graph_off[g + 1] - graph_off[g];
      }
    node_off[n_inst] = (int32_t)at;
    install_batch(c, n_inst, node_off.data(), seeds, nullptr);
    ctx->names.clear();
    return RMX_OK;
  });
}

}  // extern "C"

namespace {

// The instances of a batch on level 0 (node_off[n_inst + 1], one seed each):
// per-instance chunks of module ids, node offsets, the instance of every node
// (inst on the host, or null when L0.uinst already holds it) and the seeds.
void install_batch(Ctx &c, int32_t n_inst, const int32_t *node_off, const uint64_t *seeds, const uint32_t *inst) {
  const int32_t n = c.N;
  c.nb = n_inst;
  c.h_boff.assign(node_off, node_off + n_inst + 1);
  std::vector<uint64_t> sd(n_inst);
  for (int32_t I = 0; I < n_inst; ++I) sd[I] = mix64(seeds[I] ^ 0x6a09e667f3bcc909ull);
  std::vector<uint32_t> cbeg, cend, ichunk(1, 0u);
  for (int32_t I = 0; I < n_inst; ++I) {
    for (int64_t a = node_off[I]; a < node_off[I + 1]; a += kTotChunk) {
      cbeg.push_back((uint32_t)a);
      cend.push_back((uint32_t)std::min<int64_t>(a + kTotChunk, node_off[I + 1]));
    }
    ichunk.push_back((uint32_t)cbeg.size());
  }
  c.n_chunks = (int32_t)cbeg.size();
  c.b_cbeg.alloc(std::max<int64_t>(1, c.n_chunks));
  c.b_cend.alloc(std::max<int64_t>(1, c.n_chunks));
  c.b_ichunk.alloc((int64_t)n_inst + 1);
  c.b_cpart.alloc(4 * std::max<int64_t>(1, c.n_chunks));
  if (c.n_chunks > 0) {
    RMX_HIP(hipMemcpyAsync(c.b_cbeg.p, cbeg.data(), sizeof(uint32_t) * c.n_chunks, hipMemcpyHostToDevice, c.stream));
    RMX_HIP(hipMemcpyAsync(c.b_cend.p, cend.data(), sizeof(uint32_t) * c.n_chunks, hipMemcpyHostToDevice, c.stream));
  }
  RMX_HIP(hipMemcpyAsync(c.b_ichunk.p, ichunk.data(), sizeof(uint32_t) * (n_inst + 1), hipMemcpyHostToDevice,
                         c.stream));
  c.b_off.alloc((int64_t)n_inst + 1);
  c.b_node.alloc(n);
  c.b_seed.alloc(n_inst);
  c.L0.uoff.alloc((int64_t)n_inst + 1);
  c.L0.uinst.alloc(n);   // already filled when inst is null
  RMX_HIP(hipMemcpyAsync(c.b_off.p, c.h_boff.data(), sizeof(uint32_t) * (n_inst + 1), hipMemcpyHostToDevice,
                         c.stream));
  RMX_HIP(hipMemcpyAsync(c.L0.uoff.p, c.h_boff.data(), sizeof(uint32_t) * (n_inst + 1), hipMemcpyHostToDevice,
                         c.stream));
  if (inst) RMX_HIP(hipMemcpyAsync(c.L0.uinst.p, inst, sizeof(uint32_t) * n, hipMemcpyHostToDevice, c.stream));
  RMX_HIP(hipMemcpyAsync(c.b_node.p, c.L0.uinst.p, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, c.stream));
  RMX_HIP(hipMemcpyAsync(c.b_seed.p, sd.data(), sizeof(uint64_t) * n_inst, hipMemcpyHostToDevice, c.stream));
  c.sync();   // host staging vectors go out of scope
}

}  // namespace

namespace {

void install_cuts(Ctx &c, const BucketCuts &cuts) {
  c.cuts = cuts;
  if (c.L0.n > 0) build_buckets(c, c.L0);
  if (c.level == 1 && c.L1.n > 0) build_buckets(c, c.L1);
  c.sync();
}

}  // namespace

extern "C" {

int rmx_set_bucket_cutoffs(rmx_ctx *ctx, const int32_t *max_deg) {
  return guarded(ctx, [&] {
    Ctx &c = ctx->c;
    if (!max_deg) return fail(ctx, RMX_EINVAL, "null cut-offs");
    BucketCuts cuts = c.cuts;
    for (int k = 0; k < kNumBuckets - 1; ++k) {
      if (max_deg[k] < 0 || max_deg[k] > kBucketMaxDeg[k])
        return fail(ctx, RMX_EINVAL, "cut-off " + std::to_string(k) + " outside 0.." + std::to_string(kBucketMaxDeg[k]));
      if (k > 0 && max_deg[k] < max_deg[k - 1]) return fail(ctx, RMX_EINVAL, "cut-offs not ascending");
      cuts.v[k] = max_deg[k];
    }
    RMX_HIP(hipSetDevice(c.device));
    install_cuts(c, cuts);
    return RMX_OK;
  });
}

int rmx_get_bucket_cutoffs(rmx_ctx *ctx, int32_t *max_deg) {
  return guarded(ctx, [&] {
    if (!max_deg) return fail(ctx, RMX_EINVAL, "null cut-offs");
    for (int k = 0; k < kNumBuckets - 1; ++k) max_deg[k] = ctx->c.cuts.v[k];
    return RMX_OK;
  });
}

int rmx_autotune(rmx_ctx *ctx, int32_t *max_deg) {
  return guarded(ctx, [&] {
    Ctx &c = ctx->c;
    if (!c.initiated) return fail(ctx, RMX_ESTATE, "rmx_initiate first");
    RMX_HIP(hipSetDevice(c.device));
    install_cuts(c, autotune_buckets(c));
    if (max_deg)
      for (int k = 0; k < kNumBuckets - 1; ++k) max_deg[k] = c.cuts.v[k];
    return RMX_OK;
  });
}

int rmx_batch_results(rmx_ctx *ctx, double *codelength, int32_t *n_module, int32_t *iterations) {
  return guarded(ctx, [&] {
    Ctx &c = ctx->c;
    if (c.nb <= 0) return fail(ctx, RMX_ESTATE, "no batch uploaded");
    if (!c.initiated) return fail(ctx, RMX_ESTATE, "rmx_initiate first");
    module_totals(c, &c.sum_exit, &c.codelength);   // per instance, current partition
    if (codelength)
      RMX_HIP(hipMemcpyAsync(codelength, c.b_L.p, sizeof(double) * c.nb, hipMemcpyDeviceToHost, c.stream));
    if (n_module)
      RMX_HIP(hipMemcpyAsync(n_module, c.b_nmod.p, sizeof(int32_t) * c.nb, hipMemcpyDeviceToHost, c.stream));
    if (iterations)
      RMX_HIP(hipMemcpyAsync(iterations, c.b_iters.p, sizeof(int32_t) * c.nb, hipMemcpyDeviceToHost, c.stream));
    c.sync();
    return RMX_OK;
  });
}

struct rmx_net {
  rmx::NetFile f;
};

int rmx_net_read(const char *path, int threads, rmx_net **net, char *err, int err_len) {
  if (!net || !path) return RMX_EINVAL;
  *net = nullptr;
  try {
    set_parse_threads(threads > 0 ? (unsigned)threads : 0u);
    auto *h = new rmx_net;
    std::string e;
    const int rc = read_network(path, h->f, e);
    set_parse_threads(0);
    if (rc != 0) {
      if (err && err_len > 0) std::snprintf(err, (size_t)err_len, "%s", e.c_str());
      delete h;
      return rc == -2 ? RMX_EIO : RMX_EINVAL;
    }
    *net = h;
    return RMX_OK;
  } catch (const std::exception &x) {
    set_parse_threads(0);
    if (err && err_len > 0) std::snprintf(err, (size_t)err_len, "%s", x.what());
    return RMX_EINVAL;
  }
}

void rmx_net_free(rmx_net *net) { delete net; }

int rmx_net_info(const rmx_net *net, int32_t *n, int64_t *m, int32_t *n_self_links, int32_t *n_double_links) {
  if (!net) return RMX_EINVAL;
  if (n) *n = net->f.n;
  if (m) *m = (int64_t)net->f.idx.size();
  if (n_self_links) *n_self_links = net->f.n_self;
  if (n_double_links) *n_double_links = net->f.n_double;
  return RMX_OK;
}

int rmx_net_arrays(const rmx_net *net, const int64_t **out_ptr, const int32_t **out_idx, const double **w,
                   const double **node_w) {
  if (!net) return RMX_EINVAL;
  if (out_ptr) *out_ptr = net->f.ptr.data();
  if (out_idx) *out_idx = net->f.idx.data();
  if (w) *w = net->f.w.data();
  if (node_w) *node_w = net->f.node_w.data();
  return RMX_OK;
}

const char *rmx_net_name(const rmx_net *net, int32_t i) {
  if (!net || i < 0 || i >= net->f.n) return nullptr;
  return net->f.names[(size_t)i].c_str();
}

int rmx_net_write(const rmx_net *net, const char *path) {
  if (!net || !path) return RMX_EINVAL;
  try {
    std::string e;
    return write_binary(path, net->f, e) == 0 ? RMX_OK : RMX_EIO;
  } catch (...) {
    return RMX_EIO;
  }
}

int rmx_net_upload(rmx_ctx *ctx, const rmx_net *net) {
  return guarded(ctx, [&] {
    if (!net) return fail(ctx, RMX_EINVAL, "null network");
    const rmx::NetFile &f = net->f;
    int r = rmx_graph_upload(ctx, f.n, (int64_t)f.idx.size(), f.ptr.data(), f.idx.data(), f.w.data(),
                             f.node_w.data());
    if (r != RMX_OK) return r;
    ctx->names = f.names;
    return RMX_OK;
  });
}

int rmx_load_network(rmx_ctx *ctx, const char *path, int32_t *n, int64_t *m, int32_t *n_self, int32_t *n_double) {
  return guarded(ctx, [&] {
    if (!path) return fail(ctx, RMX_EINVAL, "null path");
    NetFile net;
    std::string err;
    int rc = read_network(path, net, err);
    if (rc != 0) return fail(ctx, rc == -2 ? RMX_EIO : RMX_EINVAL, err);
    int r = rmx_graph_upload(ctx, net.n, (int64_t)net.idx.size(), net.ptr.data(), net.idx.data(), net.w.data(),
                             net.node_w.data());
    if (r != RMX_OK) return r;
    ctx->names = std::move(net.names);   // after the upload, which resets the names to the default ones
    if (n) *n = net.n;
    if (m) *m = (int64_t)net.idx.size();
    if (n_self) *n_self = net.n_self;
    if (n_double) *n_double = net.n_double;
    return RMX_OK;
  });
}

int rmx_initiate(rmx_ctx *ctx, int32_t *iterations, double *codelength, int64_t *n_dangling) {
  return guarded(ctx, [&] {
    Ctx &c = ctx->c;
    if (c.N <= 0) return fail(ctx, RMX_ESTATE, "no network uploaded");
    RMX_HIP(hipSetDevice(c.device));
    c.level = 0;
    c.sweep_counter = 0;   // a re-initiated network repeats its sweeps exactly
    steady_state(c, 200, 50, 1.0e-15);
    init_modules_singletons(c);
    module_totals(c, &c.sum_exit, &c.codelength);
    if (c.nb > 0) c.initial_codelength = c.codelength;   // a batch: the sum over its instances
    ensure_sweep_buffers(c);
    reset_active(c);
    c.initiated = true;
    if (iterations) *iterations = c.iterations;
    if (codelength) *codelength = c.initial_codelength;
    if (n_dangling) *n_dangling = c.n_dangling;
    return RMX_OK;
  });
}

int rmx_get_flows(rmx_ctx *ctx, double *p) {
  return guarded(ctx, [&] {
    Ctx &c = ctx->c;
    if (!c.initiated) return fail(ctx, RMX_ESTATE, "rmx_initiate first");
    RMX_HIP(hipMemcpyAsync(p, c.L0.p.p, sizeof(double) * c.N, hipMemcpyDeviceToHost, c.stream));
    c.sync();
    return RMX_OK;
  });
}

int rmx_get_node_exit(rmx_ctx *ctx, double *q) {
  return guarded(ctx, [&] {
    Ctx &c = ctx->c;
    if (!c.initiated) return fail(ctx, RMX_ESTATE, "rmx_initiate first");
    RMX_HIP(hipMemcpyAsync(q, c.node_q.p, sizeof(double) * c.N, hipMemcpyDeviceToHost, c.stream));
    c.sync();
    return RMX_OK;
  });
}

int rmx_codelength(rmx_ctx *ctx, double *codelength, double *sum_exit, int32_t *n_module) {
  return guarded(ctx, [&] {
    Ctx &c = ctx->c;
    if (!c.initiated) return fail(ctx, RMX_ESTATE, "rmx_initiate first");
    if (codelength) *codelength = c.codelength;
    if (sum_exit) *sum_exit = c.sum_exit;
    if (n_module) *n_module = c.n_module;
    return RMX_OK;
  });
}

int rmx_calculate_codelength(rmx_ctx *ctx, double *codelength) {
  return guarded(ctx, [&] {
    Ctx &c = ctx->c;
    if (c.nb > 0) return fail(ctx, RMX_ESTATE, "a batch has one codelength per instance: rmx_batch_results");
    if (!c.initiated) return fail(ctx, RMX_ESTATE, "rmx_initiate first");
    propagate_labels(c);
    *codelength = calculate_codelength(c);
    return RMX_OK;
  });
}

int rmx_set_partition(rmx_ctx *ctx, const int32_t *part) {
  return guarded(ctx, [&] {
    Ctx &c = ctx->c;
    if (!c.initiated) return fail(ctx, RMX_ESTATE, "rmx_initiate first");
    for (int32_t i = 0; i < c.N; ++i)
      if (part[i] < 0 || part[i] >= c.N) return fail(ctx, RMX_EINVAL, "module index out of range");
    for (int32_t I = 0; I < c.nb; ++I)   // a batch: every module inside its instance's node ids
      for (uint32_t i = c.h_boff[I]; i < c.h_boff[I + 1]; ++i)
        if ((uint32_t)part[i] < c.h_boff[I] || (uint32_t)part[i] >= c.h_boff[I + 1])
          return fail(ctx, RMX_EINVAL, "a module joins two instances");
    c.level = 0;
    c.sweep_counter = 0;   // runs from an installed partition repeat exactly
    DevBuf<uint32_t> d(c.cache);
    d.alloc(c.N);
    RMX_HIP(hipMemcpyAsync(d.p, part, sizeof(uint32_t) * c.N, hipMemcpyHostToDevice, c.stream));
    set_partition(c, d.p);
    reset_active(c);
    c.sync();
    return RMX_OK;
  });
}

int rmx_get_partition(rmx_ctx *ctx, int32_t *part) {
  return guarded(ctx, [&] {
    Ctx &c = ctx->c;
    if (!c.initiated) return fail(ctx, RMX_ESTATE, "rmx_initiate first");
    propagate_labels(c);
    RMX_HIP(hipMemcpyAsync(part, c.L0.mod.p, sizeof(uint32_t) * c.N, hipMemcpyDeviceToHost, c.stream));
    c.sync();
    return RMX_OK;
  });
}

int rmx_reset_active(rmx_ctx *ctx) {
  return guarded(ctx, [&] {
    if (!ctx->c.initiated) return fail(ctx, RMX_ESTATE, "rmx_initiate first");
    reset_active(ctx->c);
    return RMX_OK;
  });
}

int rmx_move_sweep(rmx_ctx *ctx, int prioritized, double vthresh, uint64_t seed, int64_t *moved, double *codelength,
                   int64_t *edge_visits) {
  return guarded(ctx, [&] {
    Ctx &c = ctx->c;
    if (!c.initiated) return fail(ctx, RMX_ESTATE, "rmx_initiate first");
    int64_t mv = move_sweep(c, prioritized != 0, -vthresh, seed, edge_visits);
    if (moved) *moved = mv;
    if (codelength) *codelength = c.codelength;
    return RMX_OK;
  });
}

int rmx_evaluate_moves(rmx_ctx *ctx, int32_t capacity, int32_t *n_units, int32_t *best_module,
                       double *best_delta) {
  return guarded(ctx, [&] {
    Ctx &c = ctx->c;
    if (!c.initiated) return fail(ctx, RMX_ESTATE, "rmx_initiate first");
    if (c.nb > 0) return fail(ctx, RMX_ESTATE, "not on a batch of instances");
    const int32_t n = c.cur().n;
    if (n_units) *n_units = n;
    if (!best_module || !best_delta) return fail(ctx, RMX_EINVAL, "NULL output");
    if (capacity < n) return fail(ctx, RMX_EINVAL, "output arrays hold fewer entries than the level's units");
    DevBuf<uint32_t> bm(c.cache);
    DevBuf<double> bd(c.cache);
    bm.alloc(n);
    bd.alloc(n);
    evaluate_moves(c, bm.p, bd.p);
    std::vector<uint32_t> hm(n);
    RMX_HIP(hipMemcpyAsync(hm.data(), bm.p, sizeof(uint32_t) * n, hipMemcpyDeviceToHost, c.stream));
    RMX_HIP(hipMemcpyAsync(best_delta, bd.p, sizeof(double) * n, hipMemcpyDeviceToHost, c.stream));
    c.sync();
    for (int32_t u = 0; u < n; ++u) {
      // no candidate with dL < 0 (the reference keeps bestResult.diffCodeLen = 0.0, Module.cpp:1451)
      if (hm[u] == 0xffffffffu || !(best_delta[u] < 0.0)) {
        best_module[u] = -1;
        best_delta[u] = 0.0;
      } else {
        best_module[u] = (int32_t)hm[u];
      }
    }
    return RMX_OK;
  });
}

int rmx_update_members(rmx_ctx *ctx) {
  return guarded(ctx, [&] {
    if (!ctx->c.initiated) return fail(ctx, RMX_ESTATE, "rmx_initiate first");
    update_members(ctx->c);
    return RMX_OK;
  });
}

int rmx_contract(rmx_ctx *ctx, int32_t *n_super, int64_t *m_super) {
  return guarded(ctx, [&] {
    Ctx &c = ctx->c;
    if (!c.initiated) return fail(ctx, RMX_ESTATE, "rmx_initiate first");
    contract(c);
    reset_active(c);
    if (n_super) *n_super = c.L1.n;
    if (m_super) *m_super = c.L1.m;
    return RMX_OK;
  });
}

int rmx_get_super(rmx_ctx *ctx, double *size, double *exit, double *tp, double *dangling, int32_t *mod,
                  int32_t *n_members) {
  return guarded(ctx, [&] {
    Ctx &c = ctx->c;
    if (c.level != 1) return fail(ctx, RMX_ESTATE, "no super nodes (rmx_contract first)");
    const int32_t n = c.L1.n;
    std::vector<Stat4> us(n);
    std::vector<int32_t> nm(n);
    std::vector<uint32_t> md(n);
    RMX_HIP(hipMemcpyAsync(us.data(), c.L1.us.p, sizeof(Stat4) * n, hipMemcpyDeviceToHost, c.stream));
    RMX_HIP(hipMemcpyAsync(nm.data(), c.L1.nmem.p, sizeof(int32_t) * n, hipMemcpyDeviceToHost, c.stream));
    RMX_HIP(hipMemcpyAsync(md.data(), c.L1.mod.p, sizeof(uint32_t) * n, hipMemcpyDeviceToHost, c.stream));
    c.sync();
    for (int32_t i = 0; i < n; ++i) {
      const double P = fx_to(us[i].p), T = fx_to(us[i].t), D = fx_to(us[i].d), X = fx_to(us[i].x);
      if (size) size[i] = P;
      if (tp) tp[i] = T;
      if (dangling) dangling[i] = D;
      if (exit) exit[i] = (kAlpha * P + kBeta * D) * (1.0 - T) + kBeta * X;
      if (mod) mod[i] = (int32_t)md[i];
      if (n_members) n_members[i] = nm[i];
    }
    return RMX_OK;
  });
}

int rmx_get_super_arcs(rmx_ctx *ctx, int64_t *out_ptr, int32_t *out_idx, double *out_w, int64_t *in_ptr,
                       int32_t *in_idx, double *in_w) {
  return guarded(ctx, [&] {
    Ctx &c = ctx->c;
    if (c.level != 1) return fail(ctx, RMX_ESTATE, "no super nodes (rmx_contract first)");
    const int32_t n = c.L1.n;
    const int64_t m = c.L1.m;
    std::vector<uint32_t> op(n + 1), ip(n + 1);
    RMX_HIP(hipMemcpyAsync(op.data(), c.L1.out_ptr.p, sizeof(uint32_t) * (n + 1), hipMemcpyDeviceToHost, c.stream));
    RMX_HIP(hipMemcpyAsync(ip.data(), c.L1.in_ptr.p, sizeof(uint32_t) * (n + 1), hipMemcpyDeviceToHost, c.stream));
    if (m > 0) {
      if (out_idx) RMX_HIP(hipMemcpyAsync(out_idx, c.L1.out_idx.p, sizeof(uint32_t) * m, hipMemcpyDeviceToHost, c.stream));
      if (out_w) RMX_HIP(hipMemcpyAsync(out_w, c.L1.out_w.p, sizeof(double) * m, hipMemcpyDeviceToHost, c.stream));
      if (in_idx) RMX_HIP(hipMemcpyAsync(in_idx, c.L1.in_idx.p, sizeof(uint32_t) * m, hipMemcpyDeviceToHost, c.stream));
      if (in_w) {   // derived from the out-arcs on request (the sweeps never read them)
        DevBuf<double> tmp(c.cache);
        tmp.alloc(m);
        super_in_weights(c, c.L1, tmp.p);
        RMX_HIP(hipMemcpyAsync(in_w, tmp.p, sizeof(double) * m, hipMemcpyDeviceToHost, c.stream));
        c.sync();
      }
    }
    c.sync();
    for (int32_t i = 0; i <= n; ++i) {
      if (out_ptr) out_ptr[i] = op[i];
      if (in_ptr) in_ptr[i] = ip[i];
    }
    return RMX_OK;
  });
}

int rmx_to_node_level(rmx_ctx *ctx) {
  return guarded(ctx, [&] {
    Ctx &c = ctx->c;
    if (!c.initiated) return fail(ctx, RMX_ESTATE, "rmx_initiate first");
    to_node_level(c);
    reset_active(c);
    return RMX_OK;
  });
}

}  // extern "C"

// ------------------------------------------------------------------ driver
namespace {

struct Driver {
  Ctx &c;
  const rmx_partition_opts &o;
  rmx_partition_result &r;
  double t0;
  double t_sweeps = 0, t_contract = 0, t_coarse = 0;
  uint64_t sweep_no = 0;
  const double ln2 = std::log(2.0);

  int64_t sweep() {
    double a = now_s();
    int64_t vis = 0;
    int64_t moved = move_sweep(c, o.prior != 0, -o.vthresh, o.seed ^ (0x5851f42d4c957f2dull * ++sweep_no), &vis);
    t_sweeps += now_s() - a;
    r.sweeps++;
    r.edge_visits += vis;
    return moved;
  }

  void log_iter(const char *tag, int outer, int it, double t_it, int64_t moved) {
    if (!o.verbose) return;
    if (outer > 0)
      std::printf("%s %d-%d: code length =\t%g\t, elapsed time:\t%g\t(sec), accumulated time:\t%g\t(sec)\tsumExitPr = %g\tnumMoved:\t%lld\n",
                  tag, outer, it, c.codelength / ln2, t_it, now_s() - t0, c.sum_exit, (long long)moved);
    else
      std::printf("%s %d: code length =\t%g\t, elapsed time:\t%g\t(sec), accumulated time:\t%g\t(sec)\tsumExitPr = %g\tnumMoved:\t%lld\n",
                  tag, it, c.codelength / ln2, t_it, now_s() - t0, c.sum_exit, (long long)moved);
  }

  // stochastic_greedy_partition (OmpRelaxmap.cpp:263-429)
  void greedy(bool fine) {
    const double thr = o.threshold;
    if (fine) to_node_level(c);
    else {
      // coarse tune (main(), OmpRelaxmap.cpp:212-217): the sub-modules of
      // every module become the moving units, starting in their modules
      double a = now_s();
      generate_sub_modules(c, thr, o.max_iter, o.seed ^ (0x9e3779b97f4a7c15ull * ++sweep_no));
      contract_sub_modules(c);
      t_coarse += now_s() - a;
    }
    reset_active(c);
    if (o.verbose) std::printf("%d, ", c.cur().n);
    int iter = 0;
    bool stop = false;
    while (!stop && iter < o.max_iter) {
      double a = now_s();
      double old = c.codelength;
      int64_t moved = sweep();
      ++iter;
      if (old - c.codelength >= 0 && (old - c.codelength) / ln2 < thr) stop = true;
      log_iter("Iteration", 0, iter, now_s() - a, moved);
    }
    update_members(c);
    int outer = 1;
    double old;
    do {
      old = c.codelength;
      stop = false;
      double a = now_s();
      contract(c);
      t_contract += now_s() - a;
      reset_active(c);
      if (o.verbose) std::printf("%d, ", c.L1.n);
      int sp_iter = 0;
      while (!stop && sp_iter < o.max_iter) {
        double b = now_s();
        double inner = c.codelength;
        int64_t moved = sweep();
        ++sp_iter;
        if (inner - c.codelength >= 0.0 && (inner - c.codelength) / ln2 < thr) stop = true;
        log_iter("SuperIteration", outer, sp_iter, now_s() - b, moved);
      }
      update_members(c);
      ++outer;
    } while ((old - c.codelength) / ln2 > thr);
    propagate_labels(c);
  }
};

}  // namespace

void rmx::to_node_level(Ctx &c) {
  if (c.level == 1) {
    propagate_labels(c);
    c.level = 0;
  }
  // at the node level every node is a unit
  RMX_HIP(hipMemcpyAsync(c.mUnits.p, c.mNodes.p, sizeof(int32_t) * c.N, hipMemcpyDeviceToDevice, c.stream));
}

extern "C" int rmx_generate_sub_modules(rmx_ctx *ctx, double threshold, int32_t max_iter, uint64_t seed,
                                        int32_t *n_sub) {
  return guarded(ctx, [&] {
    Ctx &c = ctx->c;
    if (!c.initiated) return fail(ctx, RMX_ESTATE, "rmx_initiate first");
    if (max_iter < 0) return fail(ctx, RMX_EINVAL, "max_iter < 0");
    RMX_HIP(hipSetDevice(c.device));
    const int32_t k = rmx::generate_sub_modules(c, threshold, max_iter, seed);
    if (n_sub) *n_sub = k;
    return RMX_OK;
  });
}

extern "C" int rmx_get_sub_modules(rmx_ctx *ctx, int32_t *sub) {
  return guarded(ctx, [&] {
    Ctx &c = ctx->c;
    if (c.n_sub <= 0) return fail(ctx, RMX_ESTATE, "rmx_generate_sub_modules first");
    if (!sub) return fail(ctx, RMX_EINVAL, "null output");
    RMX_HIP(hipMemcpyAsync(sub, c.Lsub.mod.p, sizeof(uint32_t) * c.N, hipMemcpyDeviceToHost, c.stream));
    c.sync();
    return RMX_OK;
  });
}

extern "C" int rmx_contract_sub_modules(rmx_ctx *ctx, int32_t *n_super, int64_t *m_super) {
  return guarded(ctx, [&] {
    Ctx &c = ctx->c;
    if (c.n_sub <= 0) return fail(ctx, RMX_ESTATE, "rmx_generate_sub_modules first");
    RMX_HIP(hipSetDevice(c.device));
    rmx::contract_sub_modules(c);
    if (n_super) *n_super = c.L1.n;
    if (m_super) *m_super = c.L1.m;
    return RMX_OK;
  });
}

namespace {

// One run of main()'s super-step loop (OmpRelaxmap.cpp:181-219) from the
// current partition.
void run_super_steps(Ctx &c, const rmx_partition_opts &o, rmx_partition_result &r, Driver &d) {
  const double ln2 = std::log(2.0);
  bool fine = true;
  int step = 1;
  double old = c.codelength;
  d.greedy(fine);
  if (o.verbose) std::printf("SuperStep [%d] - codeLength = %g in %d modules.\n", step, c.codelength / ln2, c.n_module);
  bool next = !((old - c.codelength) / ln2 < o.threshold);
  while (next) {
    old = c.codelength;
    d.greedy(fine);
    ++step;
    if (o.verbose)
      std::printf("SuperStep [%d] - codeLength = %g in %d modules.\n", step, c.codelength / ln2, c.n_module);
    if ((old - c.codelength) / ln2 < o.threshold) next = false;
    fine = !fine;
  }
  r.super_steps += step;
}

// back to singleton modules on the node level (a fresh attempt)
void reset_to_singletons(Ctx &c) {
  propagate_labels(c);
  c.sweep_counter = 0;   // as after rmx_initiate
  init_modules_singletons(c);
  module_totals(c, &c.sum_exit, &c.codelength);
  reset_active(c);
}

}  // namespace

namespace {

// Attempts > 1 on one network as one batch of instances: a temporary context
// holds a copy of the network per attempt (seeds seed, seed + 1, ...), runs
// them together, and the copy with the lowest codelength (the first, on a
// tie) gives this context its partition.  Returns a status.
int batched_attempts(rmx_ctx *ctx, const rmx_partition_opts &o, rmx_partition_result &r) {
  Ctx &c = ctx->c;
  const int A = o.attempts;
  const int32_t n = c.N;
  const int64_t m = c.E;
  const double t_start = now_s();
  std::vector<uint32_t> p32((size_t)n + 1), idx((size_t)m);
  std::vector<double> w((size_t)m), nw((size_t)n);
  RMX_HIP(hipMemcpyAsync(p32.data(), c.L0.out_ptr.p, sizeof(uint32_t) * (n + 1), hipMemcpyDeviceToHost, c.stream));
  if (m > 0) {
    RMX_HIP(hipMemcpyAsync(idx.data(), c.L0.out_idx.p, sizeof(uint32_t) * m, hipMemcpyDeviceToHost, c.stream));
    RMX_HIP(hipMemcpyAsync(w.data(), c.raw_w.p, sizeof(double) * m, hipMemcpyDeviceToHost, c.stream));
  }
  RMX_HIP(hipMemcpyAsync(nw.data(), c.node_w.p, sizeof(double) * n, hipMemcpyDeviceToHost, c.stream));
  c.sync();
  const int64_t N = (int64_t)A * n;
  std::vector<int64_t> ptr((size_t)n + 1);
  for (int32_t i = 0; i <= n; ++i) ptr[i] = p32[i];
  const int32_t goff[2] = {0, n};
  std::vector<uint64_t> seeds((size_t)A);
  for (int a = 0; a < A; ++a) seeds[a] = o.seed + (uint64_t)a;
  rmx_ctx *t = nullptr;
  if (rmx_create(&t, c.device) != RMX_OK) return fail(ctx, RMX_EDEVICE, "attempts: no second context");
  struct Drop {
    rmx_ctx *t;
    ~Drop() { rmx_destroy(t); }
  } drop{t};
  t->c.precision = c.precision;
  t->c.batches = c.batches;
  t->c.cuts = c.cuts;
  t->c.hub_global = c.hub_global;
  t->c.use_graph = c.use_graph;
  for (int b = 0; b < 8; ++b) t->c.bound_mode[b] = c.bound_mode[b];
  auto check = [&](int rc, const char *what) {
    if (rc != RMX_OK) throw HipError(std::string("attempts: ") + what + ": " + rmx_last_error(t));
  };
  check(rmx_batch_upload_seeds(t, 1, goff, A, seeds.data(), n, m, ptr.data(), (const int32_t *)idx.data(), w.data(),
                               nw.data()),
        "upload");
  check(rmx_initiate(t, nullptr, nullptr, nullptr), "initiate");
  rmx_partition_opts o1 = o;
  o1.attempts = 1;
  o1.verbose = 0;
  rmx_partition_result rt;
  check(rmx_partition(t, &o1, &rt), "partition");
  std::vector<double> L((size_t)A);
  check(rmx_batch_results(t, L.data(), nullptr, nullptr), "results");
  int best = 0;
  for (int a = 1; a < A; ++a)
    if (L[a] < L[best]) best = a;
  if (o.verbose)
    for (int a = 0; a < A; ++a) std::printf("Attempt %d: Computed Code Length = %g\n", a + 1, L[a] / std::log(2.0));
  std::vector<int32_t> part((size_t)N);
  check(rmx_get_partition(t, part.data()), "partition");
  std::vector<uint32_t> bp((size_t)n);
  const int64_t b0 = (int64_t)best * n;
  for (int32_t i = 0; i < n; ++i) bp[i] = (uint32_t)(part[b0 + i] - b0);   // the copy's modules lie in its range
  DevBuf<uint32_t> d(c.cache);
  d.alloc(n);
  RMX_HIP(hipMemcpyAsync(d.p, bp.data(), sizeof(uint32_t) * n, hipMemcpyHostToDevice, c.stream));
  rmx::to_node_level(c);
  set_partition(c, d.p);
  c.sync();
  r.super_steps = rt.super_steps;
  r.sweeps = rt.sweeps;
  r.edge_visits = rt.edge_visits;
  r.time_sweeps_s = rt.time_sweeps_s;
  r.time_contract_s = rt.time_contract_s;
  r.time_coarse_s = rt.time_coarse_s;
  r.best_attempt = best;
  r.time_partition_s = now_s() - t_start;
  r.codelength = c.codelength;
  r.n_module = c.n_module;
  r.computed_codelength = calculate_codelength(c);
  c.timing.sweep_ms = rt.time_sweeps_s * 1e3;
  c.timing.contract_ms = rt.time_contract_s * 1e3;
  if (o.verbose) {
    std::printf("Time for partitioning : %g (sec)\n", r.time_partition_s);
    std::printf("DONE: Code Length = %g in %d modules, with %d nodes (attempt %d of %d).\n\n",
                c.codelength / std::log(2.0), c.n_module, c.N, best + 1, A);
  }
  return RMX_OK;
}

}  // namespace

extern "C" int rmx_partition(rmx_ctx *ctx, const rmx_partition_opts *opts, rmx_partition_result *res) {
  return guarded(ctx, [&] {
    Ctx &c = ctx->c;
    if (!c.initiated) return fail(ctx, RMX_ESTATE, "rmx_initiate first");
    if (!opts || !res) return fail(ctx, RMX_EINVAL, "null options or result");
    if (opts->max_iter < 0) return fail(ctx, RMX_EINVAL, "max_iter < 0");
    std::memset(res, 0, sizeof(*res));
    RMX_HIP(hipSetDevice(c.device));
    if (c.nb > 0) {   // a batch of instances: its seeds are the attempts
      if (opts->attempts > 1) return fail(ctx, RMX_EINVAL, "attempts > 1 on a batch: add instances instead");
      batch_partition(c, *opts, *res);
      return RMX_OK;
    }
    const double ln2 = std::log(2.0);
    const double t_start = now_s();
    const int attempts = opts->attempts > 1 ? opts->attempts : 1;
    if (attempts > 1 && !c.attempts_sequential && (int64_t)attempts * c.E <= ((int64_t)1 << 28) &&
        (int64_t)attempts * c.N <= ((int64_t)1 << 27))
      return batched_attempts(ctx, *opts, *res);
    DevBuf<uint32_t> best(c.cache);
    double best_L = 0.0;
    Driver d{c, *opts, *res, now_s()};
    for (int a = 0; a < attempts; ++a) {
      rmx_partition_opts oa = *opts;
      oa.seed = opts->seed + (uint64_t)a;
      Driver da{c, oa, *res, now_s()};
      if (a > 0) reset_to_singletons(c);
      run_super_steps(c, oa, *res, da);
      d.t_sweeps += da.t_sweeps;
      d.t_contract += da.t_contract;
      d.t_coarse += da.t_coarse;
      if (attempts > 1) {
        propagate_labels(c);
        const double L = calculate_codelength(c);
        if (opts->verbose) std::printf("Attempt %d: Computed Code Length = %g\n", a + 1, L / ln2);
        if (a == 0 || L < best_L) {
          best_L = L;
          res->best_attempt = a;
          best.alloc(c.N);
          RMX_HIP(hipMemcpyAsync(best.p, c.L0.mod.p, sizeof(uint32_t) * c.N, hipMemcpyDeviceToDevice, c.stream));
        }
      }
    }
    if (attempts > 1 && res->best_attempt != attempts - 1) {
      rmx::to_node_level(c);
      set_partition(c, best.p);   // the kept attempt, with fresh module statistics
    }
    c.sync();
    res->time_partition_s = now_s() - t_start;
    res->time_sweeps_s = d.t_sweeps;
    res->time_contract_s = d.t_contract;
    res->time_coarse_s = d.t_coarse;
    res->codelength = c.codelength;
    res->n_module = c.n_module;
    res->computed_codelength = calculate_codelength(c);
    c.timing.sweep_ms = d.t_sweeps * 1e3;
    c.timing.contract_ms = d.t_contract * 1e3;
    if (opts->verbose) {
      std::printf("Time for partitioning : %g (sec)\n", res->time_partition_s);
      std::printf("DONE: Code Length = %g in %d modules, with %d nodes.\n\n", c.codelength / ln2, c.n_module, c.N);
    }
    return RMX_OK;
  });
}

extern "C" int rmx_write_tree(rmx_ctx *ctx, const char *path) {
  return guarded(ctx, [&] {
    Ctx &c = ctx->c;
    if (!c.initiated) return fail(ctx, RMX_ESTATE, "rmx_initiate first");
    propagate_labels(c);
    std::vector<uint32_t> mod(c.N);
    std::vector<double> p(c.N);
    RMX_HIP(hipMemcpyAsync(mod.data(), c.L0.mod.p, sizeof(uint32_t) * c.N, hipMemcpyDeviceToHost, c.stream));
    RMX_HIP(hipMemcpyAsync(p.data(), c.L0.p.p, sizeof(double) * c.N, hipMemcpyDeviceToHost, c.stream));
    c.sync();
    std::string err;
    set_names(ctx, c.N);
    if (write_tree(path, c.codelength, mod, p, ctx->names, err) != 0) return fail(ctx, RMX_EIO, err);
    return RMX_OK;
  });
}

extern "C" int rmx_write_clu(rmx_ctx *ctx, const char *path) {
  return guarded(ctx, [&] {
    Ctx &c = ctx->c;
    if (!c.initiated) return fail(ctx, RMX_ESTATE, "rmx_initiate first");
    propagate_labels(c);
    std::vector<uint32_t> mod(c.N);
    RMX_HIP(hipMemcpyAsync(mod.data(), c.L0.mod.p, sizeof(uint32_t) * c.N, hipMemcpyDeviceToHost, c.stream));
    c.sync();
    std::string err;
    if (write_clu(path, mod, err) != 0) return fail(ctx, RMX_EIO, err);
    return RMX_OK;
  });
}

extern "C" int rmx_set_profiling(rmx_ctx *ctx, int on) {
  return guarded(ctx, [&] {
    ctx->c.prof.on = on != 0;
    if (on) ctx->c.prof.stats.clear();
    return RMX_OK;
  });
}

extern "C" int rmx_kernel_stats(rmx_ctx *ctx, int i, char *name, int name_len, int64_t *launches, double *ms,
                                double *bytes) {
  return guarded(ctx, [&] {
    const auto &st = ctx->c.prof.stats;
    if (i < 0) return (int)st.size();
    if (i >= (int)st.size()) return fail(ctx, RMX_EINVAL, "no such kernel entry");
    if (name && name_len > 0) std::snprintf(name, (size_t)name_len, "%s", st[i].name.c_str());
    if (launches) *launches = st[i].launches;
    if (ms) *ms = st[i].ms;
    if (bytes) *bytes = st[i].bytes;
    return RMX_OK;
  });
}

extern "C" int rmx_timing(rmx_ctx *ctx, double *flow_ms, double *sweep_ms, double *contract_ms, int64_t *edge_visits,
                          int64_t *sweeps) {
  return guarded(ctx, [&] {
    const Timing &t = ctx->c.timing;
    if (flow_ms) *flow_ms = t.flow_ms;
    if (sweep_ms) *sweep_ms = t.sweep_ms;
    if (contract_ms) *contract_ms = t.contract_ms;
    if (edge_visits) *edge_visits = t.edge_visits;
    if (sweeps) *sweeps = t.sweeps;
    return RMX_OK;
  });
}
