// A batch of independent networks partitioned together: main()'s super-step
// loop (OmpRelaxmap.cpp:181-219) around stochastic_greedy_partition
// (OmpRelaxmap.cpp:263-429), run for every instance of the batch at once.
//
// The instances share every launch: a sweep, a contraction or a coarse tune
// covers all of them.  Every decision of the reference's driver (leave the
// sweep loop, leave the outer contraction loop, stop the super steps) is
// taken per instance from its own codelength; an instance that has left a
// loop keeps its units inactive until the loop ends for all, so it follows
// the schedule it would follow alone (its batch draws come from its own
// seed).  A contraction or a coarse tune that reaches an instance after it
// has left the loop turns its modules into super nodes without moving
// anything, which leaves its partition as it is.
#include "devutil.h"

#include <chrono>
#include <cmath>

#include "relaxmap.h"

namespace rmx {

namespace {

// sweep loop (OmpRelaxmap.cpp:328-329, 400-401): an instance leaves it
// after a sweep that gains at least 0 and less than the threshold
__global__ void k_b_inner(int32_t nb, const double *__restrict__ Lold, const double *__restrict__ L, double thr,
                          uint8_t *__restrict__ flag, uint32_t *__restrict__ count) {
  __shared__ uint32_t lds[256 / kWave];
  uint32_t k = 0;
  for (int32_t I = threadIdx.x; I < nb; I += 256) {
    if (!flag[I]) continue;
    const double gain = Lold[I] - L[I];
    if (gain >= 0.0 && gain < thr) flag[I] = 0;
    else ++k;
  }
  k = block_sum<uint32_t, 256>(k, lds);
  if (threadIdx.x == 0) *count = k;
}

// outer contraction loop (the do-while of OmpRelaxmap.cpp:355-422): goes on
// while a round gains more than the threshold
__global__ void k_b_outer(int32_t nb, const double *__restrict__ Lold, const double *__restrict__ L, double thr,
                          uint8_t *__restrict__ flag, uint32_t *__restrict__ count) {
  __shared__ uint32_t lds[256 / kWave];
  uint32_t k = 0;
  for (int32_t I = threadIdx.x; I < nb; I += 256) {
    if (!flag[I]) continue;
    if (Lold[I] - L[I] > thr) ++k;
    else flag[I] = 0;
  }
  k = block_sum<uint32_t, 256>(k, lds);
  if (threadIdx.x == 0) *count = k;
}

// super steps (OmpRelaxmap.cpp:196-219): stop once a step gains less than
// the threshold
__global__ void k_b_step(int32_t nb, const double *__restrict__ Lold, const double *__restrict__ L, double thr,
                         uint8_t *__restrict__ flag, uint32_t *__restrict__ count) {
  __shared__ uint32_t lds[256 / kWave];
  uint32_t k = 0;
  for (int32_t I = threadIdx.x; I < nb; I += 256) {
    if (!flag[I]) continue;
    if (Lold[I] - L[I] < thr) flag[I] = 0;
    else ++k;
  }
  k = block_sum<uint32_t, 256>(k, lds);
  if (threadIdx.x == 0) *count = k;
}

// units of instances that have left the loop stay inactive
__global__ void k_b_mask(int32_t n, const uint32_t *__restrict__ uinst, const uint8_t *__restrict__ live,
                         uint8_t *__restrict__ act) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && !live[uinst[i]]) act[i] = 0;
}

__global__ void k_b_copy(int32_t nb, const uint8_t *__restrict__ a, uint8_t *__restrict__ b) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nb) b[i] = a[i];
}

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct BatchDriver {
  Ctx &c;
  const rmx_partition_opts &o;
  rmx_partition_result &r;
  const double thr;   // nats
  uint64_t sweep_no = 0;
  double t_sweeps = 0, t_contract = 0, t_coarse = 0;

  uint32_t read_count() {
    uint32_t k = 0;
    RMX_HIP(hipMemcpyAsync(&k, c.b_count.p, sizeof(uint32_t), hipMemcpyDeviceToHost, c.stream));
    c.sync();
    return k;
  }
  void keep_L(DevBuf<double> &dst) {
    RMX_HIP(hipMemcpyAsync(dst.p, c.b_L.p, sizeof(double) * c.nb, hipMemcpyDeviceToDevice, c.stream));
  }
  void copy_flags(const DevBuf<uint8_t> &a, DevBuf<uint8_t> &b) {
    k_b_copy<<<grid_for(c.nb, 256), 256, 0, c.stream>>>(c.nb, a.p, b.p);
    RMX_LAUNCH_CHECK();
  }
  void mask_active() {
    const Level &L = c.cur();
    if (L.n <= 0) return;
    k_b_mask<<<grid_for(L.n, 256), 256, 0, c.stream>>>(L.n, L.uinst.p, c.b_inner.p, c.act.p);
    RMX_LAUNCH_CHECK();
  }
  void start_loop(const DevBuf<uint8_t> &from) {
    reset_active(c);
    copy_flags(from, c.b_inner);
    mask_active();
  }
  void sweep() {
    const double a = now_s();
    int64_t vis = 0;
    move_sweep(c, o.prior != 0, -o.vthresh, o.seed ^ (0x5851f42d4c957f2dull * ++sweep_no), &vis);
    t_sweeps += now_s() - a;
    r.sweeps++;
    r.edge_visits += vis;
  }
  // sweeps until every instance in b_inner has left the loop, or max_iter
  void sweep_loop() {
    for (int it = 0; it < o.max_iter; ++it) {
      keep_L(c.b_Lold);
      sweep();
      k_b_inner<<<1, 256, 0, c.stream>>>(c.nb, c.b_Lold.p, c.b_L.p, thr, c.b_inner.p, c.b_count.p);
      RMX_LAUNCH_CHECK();
      mask_active();
      if (read_count() == 0) break;
    }
  }
  // stochastic_greedy_partition for the instances in b_step
  void greedy(bool fine) {
    if (fine) {
      to_node_level(c);
    } else {
      const double a = now_s();
      generate_sub_modules(c, o.threshold, o.max_iter, o.seed ^ (0x9e3779b97f4a7c15ull * ++sweep_no));
      contract_sub_modules(c);
      t_coarse += now_s() - a;
    }
    start_loop(c.b_step);
    sweep_loop();
    update_members(c);
    copy_flags(c.b_step, c.b_outer);
    while (true) {
      keep_L(c.b_Louter);
      const double a = now_s();
      contract(c);
      t_contract += now_s() - a;
      start_loop(c.b_outer);
      sweep_loop();
      update_members(c);
      k_b_outer<<<1, 256, 0, c.stream>>>(c.nb, c.b_Louter.p, c.b_L.p, thr, c.b_outer.p, c.b_count.p);
      RMX_LAUNCH_CHECK();
      if (read_count() == 0) break;
    }
    propagate_labels(c);
  }
};

}  // namespace

void batch_partition(Ctx &c, const rmx_partition_opts &o, rmx_partition_result &r) {
  const int32_t nb = c.nb;
  const double t0 = now_s();
  c.b_step.alloc(nb);
  c.b_outer.alloc(nb);
  c.b_inner.alloc(nb);
  c.b_Lold.alloc(nb);
  c.b_Lstep.alloc(nb);
  c.b_Louter.alloc(nb);
  c.b_count.alloc(4);
  fill_u8(c.b_step.p, 1, nb, c.stream);
  BatchDriver d{c, o, r, o.threshold * std::log(2.0)};
  // main(): F, F, C, F, C, ... until a super step gains less than the threshold
  bool fine = true;
  int steps = 0;
  while (true) {
    d.keep_L(c.b_Lstep);
    d.greedy(fine);
    ++steps;
    k_b_step<<<1, 256, 0, c.stream>>>(nb, c.b_Lstep.p, c.b_L.p, d.thr, c.b_step.p, c.b_count.p);
    RMX_LAUNCH_CHECK();
    if (d.read_count() == 0) break;
    if (steps > 1) fine = !fine;   // the first two steps are both fine
  }
  to_node_level(c);
  reset_active(c);
  module_totals(c, &c.sum_exit, &c.codelength);
  r.super_steps = steps;
  r.time_partition_s = now_s() - t0;
  r.time_sweeps_s = d.t_sweeps;
  r.time_contract_s = d.t_contract;
  r.time_coarse_s = d.t_coarse;
  r.codelength = c.codelength;
  r.computed_codelength = c.codelength;
  r.n_module = c.n_module;
  c.timing.sweep_ms = d.t_sweeps * 1e3;
  c.timing.contract_ms = d.t_contract * 1e3;
}

}  // namespace rmx
