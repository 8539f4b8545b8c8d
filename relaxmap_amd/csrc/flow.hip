// Steady-state flow: PageRank with teleportation and dangling mass as a CSR
// pull SpMV (Network::calculateSteadyState, Module.cpp:340-469), then arc
// flows, singleton modules and the initial codelength (Network::initiate,
// Module.cpp:210-333, and calibrate, :474-501).
//
// Per iteration: p'_i = (alpha + beta D) tau_i + beta sum_{j->i} w_ji p_j,
// p = p' / sum p', r = |p - p_prev|_1, D = sum of p over dangling nodes.
// Stop after iteration k when k >= 200, or k >= 50 and r <= 1e-15.  Mixed
// precision has an L1 budget of 1e-6 against the FP64 flows.  The iteration
// contracts by beta = 0.85 in L1, so once r <= (1 - beta) / beta * 5e-7
// (8.8e-8) its iterate is within half that budget of the fixed point; the
// FP32 products add about 2e-8.  It also stops when ten iterations no longer
// halve r, at the FP32 noise floor.
//
// FP64: the pull is bucketed by in-degree: a group of G in {4..64} lanes per
// row streams the row's in-arcs (coalesced in_idx / in_w) and gathers x;
// rows with more than 256 in-arcs get a 256-lane workgroup.  It runs on a
// relabelled copy of the in-CSR: nodes in order of descending out-degree, so
// the sources read most often share a compact, cache-resident part of x
// (the gathers of x dominate the traffic).  Every row keeps its arcs in
// their original order, so each row sum is the same as on the original
// labels; p is mapped back after the last iteration.
//
// Mixed precision: a blocked push / pull on the original labels, FP32
// weights and an FP32 copy of x, exact fixed-point row sums (below); the
// residual, the normalisation and D stay FP64.
#include "context.h"

#include <algorithm>
#include <vector>

namespace rmx {

namespace {

constexpr int kRedGrid = 1024;
constexpr int kRedBlock = 256;
constexpr double kMixedFlowL1 = 1e-6;   // mixed-precision flows against FP64, L1
// L1 residual that ends the mixed-precision iteration: the a-priori distance
// beta / (1 - beta) r to the fixed point is then half of kMixedFlowL1
constexpr double kMixedTol = 0.5 * kMixedFlowL1 * (1.0 - kBeta) / kBeta;
enum { kS = 0, kR = 1, kD = 2, kL = 3, kSE = 4, kSS = 5, kSP = 6 };

// grid of a grid-stride loop over n elements
inline unsigned stride_grid(int64_t n) { return (unsigned)std::max<int64_t>(1, std::min<int64_t>(16384, (n + 255) / 256)); }

constexpr int kNumRowBins = 6;
constexpr int kRowBinMax[kNumRowBins] = {4, 8, 16, 32, 256, 0x7fffffff};

template <int G>
__device__ inline double group_sum(double v) {
#pragma unroll
  for (int off = G / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, G);
  return v;
}
template <int G>
__device__ inline long long group_sum_ll(long long v) {
#pragma unroll
  for (int off = G / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, G);
  return v;
}

// --- row-normalised out weights (Module.cpp:237-247) --------------------
// Also the row sums, from which the in-arcs get the same quotients; mixed
// precision keeps an FP32 copy in out-CSR order (the push).
template <int G, bool MIXED>
__global__ void k_normalize(const uint32_t *__restrict__ optr, int32_t n, const double *__restrict__ raw,
                            double *__restrict__ out_w, float *__restrict__ out_wf, double *__restrict__ rowsum,
                            uint8_t *__restrict__ dangling) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t r = t / G;
  const int lane = (int)(t % G);
  if (r >= n) return;
  const uint32_t b = optr[r], e = optr[r + 1];
  double s = 0.0;
  for (uint32_t k0 = b + lane; k0 < e; k0 += kRowAhead * G) {
    double v[kRowAhead];
#pragma unroll
    for (int j = 0; j < kRowAhead; ++j) v[j] = k0 + j * G < e ? raw[k0 + j * G] : 0.0;
#pragma unroll
    for (int j = 0; j < kRowAhead; ++j)
      if (k0 + j * G < e) s += v[j];
  }
  s = group_sum<G>(s);
  if (lane == 0) {
    dangling[r] = (b == e) ? 1 : 0;
    rowsum[r] = s;
  }
  for (uint32_t k0 = b + lane; k0 < e; k0 += kRowAhead * G) {
    double v[kRowAhead];
#pragma unroll
    for (int j = 0; j < kRowAhead; ++j) v[j] = k0 + j * G < e ? raw[k0 + j * G] : 0.0;
#pragma unroll
    for (int j = 0; j < kRowAhead; ++j) {
      const uint32_t k = k0 + j * G;
      if (k < e) {
        const double q = v[j] / s;
        out_w[k] = q;
        if (MIXED) out_wf[k] = (float)q;
      }
    }
  }
}

// The in-arcs, in in-CSR order, each from its tail: the normalised weight
// raw / s (FP64 pull), or the flow p * (raw / s) in FP64 (in_w) or, in mixed
// precision, rounded to FP32 (in_wf): the sweeps of each precision read only
// their own array.  The operations are those of the out-arc, so both copies
// are identical.  The tail's p and row sum are read as one 16-byte pair.
__global__ void k_in_norm(const uint32_t *__restrict__ iidx, const double *__restrict__ raw_in, int64_t m,
                          const double *__restrict__ rowsum, double *__restrict__ in_w) {
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < m; j += (int64_t)gridDim.x * blockDim.x)
    in_w[j] = raw_in[j] / rowsum[iidx[j]];
}

__global__ void k_pack_ps(const double *__restrict__ p, const double *__restrict__ rowsum, int32_t n,
                          double2 *__restrict__ ps) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) ps[i] = make_double2(p[i], rowsum[i]);
}

__global__ void k_in_flows(const uint32_t *__restrict__ iidx, const double *__restrict__ raw_in, int64_t m,
                           const double2 *__restrict__ ps, int mixed, double *__restrict__ in_w,
                           float *__restrict__ in_wf, long long *__restrict__ in_wx) {
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < m; j += (int64_t)gridDim.x * blockDim.x) {
    const double2 t = ps[iidx[j]];
    const double f = t.x * (raw_in[j] / t.y);
    if (in_w) in_w[j] = f;
    if (in_wf) in_wf[j] = (float)f;
    in_wx[j] = fx_of(mixed ? (double)(float)f : f);
  }
}

__global__ __launch_bounds__(256) void k_row_bin(const uint32_t *__restrict__ iptr, int32_t n,
                                                 uint32_t *__restrict__ key, uint32_t *__restrict__ cnt,
                                                 unsigned long long *__restrict__ arcs) {
  __shared__ uint32_t s_cnt[kNumRowBins];
  __shared__ unsigned long long s_arcs[kNumRowBins];
  if (threadIdx.x < kNumRowBins) {
    s_cnt[threadIdx.x] = 0;
    s_arcs[threadIdx.x] = 0;
  }
  __syncthreads();
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    uint32_t d = iptr[i + 1] - iptr[i];
    uint32_t b = 0;
    while (b + 1 < (uint32_t)kNumRowBins && d > (uint32_t)kRowBinMax[b]) ++b;
    key[i] = b;
    atomicAdd(&s_cnt[b], 1u);
    atomicAdd(&s_arcs[b], (unsigned long long)d);
  }
  __syncthreads();
  if (threadIdx.x < kNumRowBins && s_cnt[threadIdx.x]) {
    atomicAdd(&cnt[threadIdx.x], s_cnt[threadIdx.x]);
    atomicAdd(&arcs[threadIdx.x], s_arcs[threadIdx.x]);
  }
}

// --- pull SpMV (FP64) ---------------------------------------------------
template <int G>
__global__ __launch_bounds__(256) void k_spmv(const uint32_t *__restrict__ rows, int32_t nrows,
                                              const uint32_t *__restrict__ iptr, const uint32_t *__restrict__ iidx,
                                              const double *__restrict__ iw, const double *__restrict__ x,
                                              const double *__restrict__ tau, const double *__restrict__ Dv,
                                              const uint32_t *__restrict__ rinst, double *__restrict__ y) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t g = t / G;
  const int lane = (int)(t % G);
  if (g >= nrows) return;
  const uint32_t r = rows[g];
  const uint32_t b = iptr[r], e = iptr[r + 1];
  double acc = 0.0;
  for (uint32_t k = b + lane; k < e; k += G) acc = fma(iw[k], x[iidx[k]], acc);
  acc = group_sum<G>(acc);
  if (lane == 0) y[r] = (kAlpha + kBeta * Dv[rinst ? rinst[r] : 0]) * tau[r] + kBeta * acc;
}

__global__ __launch_bounds__(256) void k_spmv_block(const uint32_t *__restrict__ rows, int32_t nrows,
                                                    const uint32_t *__restrict__ iptr,
                                                    const uint32_t *__restrict__ iidx, const double *__restrict__ iw,
                                                    const double *__restrict__ x, const double *__restrict__ tau,
                                                    const double *__restrict__ Dv,
                                                    const uint32_t *__restrict__ rinst, double *__restrict__ y) {
  __shared__ double lds[4];
  for (int64_t g = blockIdx.x; g < nrows; g += gridDim.x) {
    const uint32_t r = rows[g];
    const uint32_t b = iptr[r], e = iptr[r + 1];
    double acc = 0.0;
    for (uint32_t k = b + threadIdx.x; k < e; k += 256) acc = fma(iw[k], x[iidx[k]], acc);
    acc = block_sum<double, 256>(acc, lds);
    if (threadIdx.x == 0) y[r] = (kAlpha + kBeta * Dv[rinst ? rinst[r] : 0]) * tau[r] + kBeta * acc;
  }
}

// --- mixed precision: blocked push / pull ---------------------------------
// The heads are cut into bins of R consecutive rows.  Every out-arc has a
// fixed slot in a bin-major layout whose arcs keep their out-CSR (tail)
// order inside each bin.  Per iteration the push streams the out-CSR and
// writes x_tail * w into its slot: successive tails fill every bin's slots
// in increasing order, so the writes land in a few lines per bin.  The pull
// sums each bin's slots into its R rows in LDS.  The sums are exact
// fixed-point (2^62) sums of the FP32 products, independent of their order.
constexpr int kPushTile = 2048;             // arcs per push workgroup: 256 lanes x 8 consecutive arcs
// out_ptr entries the push keeps in LDS.  The 40 KB also hold the push to 4
// workgroups per CU: fewer tiles in flight narrow the range of slots each
// bin is written at, so L2 completes its lines before evicting them
// (3.4 ms instead of 4.5 ms at 8 workgroups per CU on 529M arcs).
constexpr int kPushRows = 10240;
constexpr int kPullRowsMax = 8192;          // rows per bin, at most (64 KB of LDS accumulators)
constexpr int kPullBlock = 512;
constexpr double kPullInv = 1.0 / 4611686018427387904.0;   // 2^-62
constexpr float kPullScale = 4611686018427387904.0f;       // 2^62

// tile t starts at arc t * kPushTile; trow[t] = the row that holds arc
// min(t * kPushTile, m - 1), the largest r with optr[r] <= that arc
__global__ void k_pb_tile_rows(const uint32_t *__restrict__ optr, int32_t n, int64_t m, int64_t ntiles,
                               uint32_t *__restrict__ trow) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t > ntiles) return;
  const int64_t k = std::min<int64_t>(t * kPushTile, m - 1);
  int64_t lo = 0, hi = n;   // optr[lo] <= k < optr[hi]
  while (hi - lo > 1) {
    const int64_t mid = (lo + hi) >> 1;
    if ((int64_t)optr[mid] <= k) lo = mid;
    else hi = mid;
  }
  trow[t] = (uint32_t)lo;
}

__global__ void k_pb_keys(const uint32_t *__restrict__ oidx, int64_t m, int rshift, uint32_t *__restrict__ key) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < m) key[k] = oidx[k] >> rshift;
}

// bptr[b] = first position of bin b in the sorted bin keys
__global__ void k_pb_bounds(const uint32_t *__restrict__ skey, int64_t m, int32_t nbins,
                            uint32_t *__restrict__ bptr) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b > nbins) return;
  int64_t lo = 0, hi = m;   // first position with key >= b
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if ((int64_t)skey[mid] < b) lo = mid + 1;
    else hi = mid;
  }
  bptr[b] = (uint32_t)lo;
}

// slot of every out-arc, and the head of every slot relative to its bin
__global__ void k_pb_layout(const uint32_t *__restrict__ sarc, const uint32_t *__restrict__ oidx, int64_t m,
                            int rshift, uint32_t *__restrict__ slot, uint16_t *__restrict__ bdst) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= m) return;
  const uint32_t a = sarc[j];
  slot[a] = (uint32_t)j;
  bdst[j] = (uint16_t)(oidx[a] & ((1u << rshift) - 1u));
}

// Slots without a sort (bins that fit a workgroup's LDS counters).  The arcs
// are cut into layout tiles; slots of bin b are handed out tile by tile:
// tile t owns the range [base(t, b), base(t, b) + count(t, b)) with
//   base(t, b) = bptr[b] + sum over t' < t of count(t', b),
// and bptr[b] = the arcs whose head lies below bin b, read off the in-CSR.
// Inside a tile the order of one bin's arcs follows LDS atomic order; the
// pull's fixed-point sums do not depend on it.  Each tile's arcs of a bin
// sit in one short run, so the push still writes a bin's lines from one
// XCD in nearly sequential order, and slot[] is written in arc order.
constexpr int kLayoutMaxBins = 16384;   // 64 KB of LDS counters
constexpr int kLayoutBlock = 1024;
constexpr int kLayoutChunk = 64;        // tiles per partial sum of the column scan

__global__ void k_pb_bptr(const uint32_t *__restrict__ iptr, int32_t n, int32_t nbins, int rshift,
                          uint32_t *__restrict__ bptr) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b <= nbins) bptr[b] = iptr[std::min<int64_t>(b << rshift, n)];
}

// count(t, b) into C[t * nbins + b]
__global__ __launch_bounds__(kLayoutBlock) void k_pb_count(const uint32_t *__restrict__ oidx, int64_t m,
                                                           int64_t tile_arcs, int32_t nbins, int rshift,
                                                           uint32_t *__restrict__ C) {
  __shared__ uint32_t cnt[kLayoutMaxBins];
  for (int b = threadIdx.x; b < nbins; b += kLayoutBlock) cnt[b] = 0u;
  __syncthreads();
  const int64_t a0 = (int64_t)blockIdx.x * tile_arcs, a1 = std::min<int64_t>(m, a0 + tile_arcs);
  for (int64_t k = a0 + threadIdx.x; k < a1; k += kLayoutBlock) atomicAdd(&cnt[oidx[k] >> rshift], 1u);
  __syncthreads();
  uint32_t *row = C + (int64_t)blockIdx.x * nbins;
  for (int b = threadIdx.x; b < nbins; b += kLayoutBlock) row[b] = cnt[b];
}

// P[q * nbins + b] = sum of C[t][b] over the tiles t of chunk q
__global__ void k_pb_colsum(const uint32_t *__restrict__ C, int64_t ntl, int32_t nbins, int64_t nq,
                            uint32_t *__restrict__ P) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nq * nbins) return;
  const int64_t q = i / nbins, b = i - q * nbins;
  const int64_t t1 = std::min<int64_t>(ntl, (q + 1) * kLayoutChunk);
  uint32_t s = 0;
  for (int64_t t = q * kLayoutChunk; t < t1; ++t) s += C[t * nbins + b];
  P[i] = s;
}

// P[q][b] = bptr[b] + the arcs of bin b in the chunks before q
__global__ void k_pb_colscan(const uint32_t *__restrict__ bptr, int32_t nbins, int64_t nq, uint32_t *__restrict__ P) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nbins) return;
  uint32_t run = bptr[b];
  for (int64_t q = 0; q < nq; ++q) {
    const uint32_t v = P[q * nbins + b];
    P[q * nbins + b] = run;
    run += v;
  }
}

// C[t][b] = base(t, b)
__global__ void k_pb_colbase(const uint32_t *__restrict__ P, int64_t ntl, int32_t nbins, int64_t nq,
                             uint32_t *__restrict__ C) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nq * nbins) return;
  const int64_t q = i / nbins, b = i - q * nbins;
  const int64_t t1 = std::min<int64_t>(ntl, (q + 1) * kLayoutChunk);
  uint32_t run = P[i];
  for (int64_t t = q * kLayoutChunk; t < t1; ++t) {
    const uint32_t v = C[t * nbins + b];
    C[t * nbins + b] = run;
    run += v;
  }
}

// The slot of every arc of a tile, and the head of every slot relative to
// its bin.  Neighbouring tiles write neighbouring runs of a bin, often in
// one cache line, so tiles go to the XCDs in contiguous eighths (as in the
// push): a line is then completed in one L2.
__global__ __launch_bounds__(kLayoutBlock) void k_pb_place(const uint32_t *__restrict__ oidx, int64_t m,
                                                           int64_t tile_arcs, int64_t ntl, int64_t per_xcd,
                                                           int32_t nbins, int rshift, const uint32_t *__restrict__ C,
                                                           uint32_t *__restrict__ slot, uint16_t *__restrict__ bdst) {
  __shared__ uint32_t base[kLayoutMaxBins];
  const int64_t tile = (int64_t)(blockIdx.x % 8) * per_xcd + blockIdx.x / 8;
  if (tile >= ntl) return;
  const uint32_t *row = C + tile * nbins;
  for (int b = threadIdx.x; b < nbins; b += kLayoutBlock) base[b] = row[b];
  __syncthreads();
  const uint32_t mask = (1u << rshift) - 1u;
  const int64_t a0 = tile * tile_arcs, a1 = std::min<int64_t>(m, a0 + tile_arcs);
  // four arcs per lane in flight
  for (int64_t k0 = a0 + threadIdx.x; k0 < a1; k0 += 4 * kLayoutBlock) {
    uint32_t d[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int64_t k = k0 + j * kLayoutBlock;
      d[j] = k < a1 ? oidx[k] : 0u;
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int64_t k = k0 + j * kLayoutBlock;
      if (k < a1) {
        const uint32_t r = atomicAdd(&base[d[j] >> rshift], 1u);
        slot[k] = r;
        bdst[r] = (uint16_t)(d[j] & mask);
      }
    }
  }
}

// Push: workgroups are spread so that each of the 8 XCDs (workgroup i runs
// on XCD i mod 8) takes one contiguous eighth of the tiles; every bin then
// receives each XCD's writes in one contiguous range.  The streamed slots
// and weights are read non-temporally, so that L2 keeps the partly written
// lines of the bins until they are complete.
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
__device__ inline u32x4 stream_u4(const uint32_t *p) { return __builtin_nontemporal_load((const u32x4 *)p); }
__device__ inline f32x4 stream_f4(const float *p) { return __builtin_nontemporal_load((const f32x4 *)p); }

__global__ __launch_bounds__(256) void k_pb_push(const uint32_t *__restrict__ optr, const uint32_t *__restrict__ trow,
                                                 int64_t ntiles, int64_t per_xcd, int64_t m,
                                                 const uint32_t *__restrict__ slot, const float *__restrict__ wf,
                                                 const float *__restrict__ xf, float *__restrict__ val) {
  __shared__ uint32_t sp[kPushRows];
  const int64_t tile = (int64_t)(blockIdx.x % 8) * per_xcd + blockIdx.x / 8;
  if (tile >= ntiles) return;
  const int64_t a0 = tile * kPushTile;
  const int64_t a1 = std::min<int64_t>(m, a0 + kPushTile);
  const uint32_t r0 = trow[tile], r1 = trow[tile + 1];
  const int64_t nr = (int64_t)r1 - r0 + 1;   // rows r0..r1 cover the tile; their bounds are optr[r0..r1+1]
  const bool in_lds = nr + 1 <= kPushRows;
  if (in_lds)
    for (int64_t i = threadIdx.x; i <= nr; i += 256) sp[i] = optr[r0 + i];
  __syncthreads();
  const int64_t k0 = a0 + 8 * (int64_t)threadIdx.x;
  if (k0 >= a1) return;
  const uint32_t *bounds = in_lds ? sp : optr + r0;
  // the row of k0: the largest i < nr with bounds[i] <= k0 (bounds[0] <= a0)
  int64_t lo = 0, hi = nr;
  while (hi - lo > 1) {
    const int64_t mid = (lo + hi) >> 1;
    if ((int64_t)bounds[mid] <= k0) lo = mid;
    else hi = mid;
  }
  int64_t row = lo;
  uint32_t nxt = bounds[row + 1];
  float xr = xf[r0 + row];
  if (k0 + 8 <= a1) {
    const u32x4 s0 = stream_u4(slot + k0), s1 = stream_u4(slot + k0 + 4);
    const f32x4 w0 = stream_f4(wf + k0), w1 = stream_f4(wf + k0 + 4);
    const uint32_t sl[8] = {s0.x, s0.y, s0.z, s0.w, s1.x, s1.y, s1.z, s1.w};
    const float ww[8] = {w0.x, w0.y, w0.z, w0.w, w1.x, w1.y, w1.z, w1.w};
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const uint32_t k = (uint32_t)(k0 + j);
      if (k >= nxt) {
        do {
          ++row;
          nxt = bounds[row + 1];
        } while (k >= nxt);
        xr = xf[r0 + row];
      }
      val[sl[j]] = xr * ww[j];
    }
  } else {
    for (int64_t k = k0; k < a1; ++k) {
      if ((uint32_t)k >= nxt) {
        do {
          ++row;
          nxt = bounds[row + 1];
        } while ((uint32_t)k >= nxt);
        xr = xf[r0 + row];
      }
      val[slot[k]] = xr * wf[k];
    }
  }
}

// The staged push's layout: tiles of kStageTile arcs.  Inside a tile the
// arcs are given positions in bin-major order (the tile's arcs of bin b take
// positions lbase(b) .. lbase(b) + count(t, b) - 1, lbase the tile's
// exclusive scan of its counts), and position p of tile t is tied to slot
// gslot[t kStageTile + p] = base(t, b) + (p - lbase(b)): the tile's run of
// slots in bin b.  Which arc of a run takes which position follows LDS
// atomic order; lpos[k] and gslot agree on it, and the pull's fixed-point
// sums do not depend on it.
constexpr int kStageTile = 16384;   // arcs per staged tile (64 KB of FP32 products in LDS)
constexpr int kStageBlock = 1024;

__global__ __launch_bounds__(kLayoutBlock) void k_pb_place_staged(const uint32_t *__restrict__ oidx, int64_t m,
                                                                  int32_t nbins, int rshift,
                                                                  const uint32_t *__restrict__ C,
                                                                  const uint32_t *__restrict__ Ccnt,
                                                                  uint16_t *__restrict__ lpos,
                                                                  uint32_t *__restrict__ gslot,
                                                                  uint16_t *__restrict__ bdst) {
  __shared__ uint32_t lcur[kLayoutMaxBins];   // next position of each bin in the tile
  __shared__ uint32_t goff[kLayoutMaxBins];   // slot of position p in bin b: goff[b] + p
  __shared__ uint32_t s_scan[kLayoutBlock / kWave];
  const int64_t tile = blockIdx.x;
  const uint32_t *cnt = Ccnt + tile * nbins, *base = C + tile * nbins;
  // exclusive scan of the tile's counts, kLayoutMaxBins / kLayoutBlock bins per lane
  constexpr int PER = kLayoutMaxBins / kLayoutBlock;
  uint32_t loc = 0;
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    const int b = threadIdx.x * PER + q;
    loc += b < nbins ? cnt[b] : 0u;
  }
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  uint32_t x = loc;
#pragma unroll
  for (int off = 1; off < kWave; off <<= 1) {
    const uint32_t y = __shfl_up(x, off, kWave);
    if (lane >= off) x += y;
  }
  if (lane == kWave - 1) s_scan[w] = x;
  __syncthreads();
  uint32_t run = x - loc;
  for (int i = 0; i < w; ++i) run += s_scan[i];
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    const int b = threadIdx.x * PER + q;
    if (b < nbins) {
      lcur[b] = run;
      goff[b] = base[b] - run;   // modulo 2^32, as the slots
      run += cnt[b];
    }
  }
  __syncthreads();
  const uint32_t mask = (1u << rshift) - 1u;
  const int64_t a0 = tile * kStageTile, a1 = std::min<int64_t>(m, a0 + kStageTile);
  for (int64_t k0 = a0 + threadIdx.x; k0 < a1; k0 += 4 * kLayoutBlock) {
    uint32_t d[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int64_t k = k0 + j * kLayoutBlock;
      d[j] = k < a1 ? oidx[k] : 0u;
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int64_t k = k0 + j * kLayoutBlock;
      if (k < a1) {
        const uint32_t b = d[j] >> rshift;
        const uint32_t p = atomicAdd(&lcur[b], 1u);
        const uint32_t sl = goff[b] + p;
        lpos[k] = (uint16_t)p;
        gslot[a0 + p] = sl;
        bdst[sl] = (uint16_t)(d[j] & mask);
      }
    }
  }
}

// grow[g] = the row of arc 8 g (one wavefront per row)
__global__ void k_pb_group_rows(const uint32_t *__restrict__ optr, int32_t n, uint32_t *__restrict__ grow) {
  const int64_t r = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const int lane = threadIdx.x & (kWave - 1);
  if (r >= n) return;
  const uint32_t g0 = (optr[r] + 7) / 8, g1 = (optr[r + 1] + 7) / 8;
  for (uint32_t g = g0 + lane; g < g1; g += kWave) grow[g] = (uint32_t)r;
}

// Staged push (FP32 products): one workgroup per tile of kStageTile arcs
// computes every product x_tail w into the tile's bin-major order in LDS,
// then writes the tile's runs of slots with consecutive lanes, so that the
// stores of a run leave as whole segments instead of one request per arc.
// The products are those of the plain push, so the pull's sums are the same
// bit for bit.  Tiles go to the XCDs in contiguous eighths.
__global__ __launch_bounds__(kStageBlock) void k_pb_push_staged(const uint32_t *__restrict__ optr,
                                                                const uint32_t *__restrict__ grow, int64_t ntl,
                                                                int64_t per_xcd, int64_t m,
                                                                const uint16_t *__restrict__ lpos,
                                                                const uint32_t *__restrict__ gslot,
                                                                const float *__restrict__ wf,
                                                                const float *__restrict__ xf,
                                                                float *__restrict__ val) {
  __shared__ float stage[kStageTile];
  const int64_t tile = (int64_t)(blockIdx.x % 8) * per_xcd + blockIdx.x / 8;
  if (tile >= ntl) return;
  const int64_t a0 = tile * kStageTile;
  const int64_t a1 = std::min<int64_t>(m, a0 + kStageTile);
  for (int64_t k0 = a0 + 8 * (int64_t)threadIdx.x; k0 < a1; k0 += 8 * kStageBlock) {
    uint32_t row = grow[k0 / 8];   // k0 is a multiple of 8
    uint32_t nxt = optr[row + 1];
    float xr = xf[row];
    if (k0 + 8 <= a1) {
      const u32x4 lp = stream_u4((const uint32_t *)(lpos + k0));   // 8 positions of 16 bits
      const f32x4 w0 = stream_f4(wf + k0), w1 = stream_f4(wf + k0 + 4);
      const uint32_t pw[4] = {lp.x, lp.y, lp.z, lp.w};
      const float ww[8] = {w0.x, w0.y, w0.z, w0.w, w1.x, w1.y, w1.z, w1.w};
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const uint32_t k = (uint32_t)(k0 + j);
        if (k >= nxt) {
          do {
            ++row;
            nxt = optr[row + 1];
          } while (k >= nxt);
          xr = xf[row];
        }
        stage[(pw[j >> 1] >> (16 * (j & 1))) & 0xffffu] = xr * ww[j];
      }
    } else {
      for (int64_t k = k0; k < a1; ++k) {
        if ((uint32_t)k >= nxt) {
          do {
            ++row;
            nxt = optr[row + 1];
          } while ((uint32_t)k >= nxt);
          xr = xf[row];
        }
        stage[lpos[k]] = xr * wf[k];
      }
    }
  }
  __syncthreads();
  const int n_t = (int)(a1 - a0);
  for (int p = threadIdx.x; p < n_t; p += kStageBlock) val[gslot[a0 + p]] = stage[p];
}

// Pull: one workgroup per bin of R rows
__global__ __launch_bounds__(kPullBlock) void k_pb_pull(const uint32_t *__restrict__ bptr,
                                                        const uint16_t *__restrict__ bdst,
                                                        const float *__restrict__ val, int32_t R, int32_t n,
                                                        const double *__restrict__ tau,
                                                        const double *__restrict__ Dv,
                                                        const uint32_t *__restrict__ rinst, double *__restrict__ y) {
  __shared__ unsigned long long acc[kPullRowsMax];
  for (int i = threadIdx.x; i < R; i += kPullBlock) acc[i] = 0ull;
  __syncthreads();
  const uint32_t b = bptr[blockIdx.x], e = bptr[blockIdx.x + 1];
  uint32_t j = b + threadIdx.x;
  for (; (int64_t)j + 3 * kPullBlock < (int64_t)e; j += 4 * kPullBlock) {
    const uint16_t d0 = bdst[j], d1 = bdst[j + kPullBlock], d2 = bdst[j + 2 * kPullBlock],
                   d3 = bdst[j + 3 * kPullBlock];
    const float v0 = val[j], v1 = val[j + kPullBlock], v2 = val[j + 2 * kPullBlock], v3 = val[j + 3 * kPullBlock];
    atomicAdd(&acc[d0], (unsigned long long)(v0 * kPullScale));
    atomicAdd(&acc[d1], (unsigned long long)(v1 * kPullScale));
    atomicAdd(&acc[d2], (unsigned long long)(v2 * kPullScale));
    atomicAdd(&acc[d3], (unsigned long long)(v3 * kPullScale));
  }
  for (; j < e; j += kPullBlock) atomicAdd(&acc[bdst[j]], (unsigned long long)(val[j] * kPullScale));
  __syncthreads();
  const int64_t row0 = (int64_t)blockIdx.x * R;
  for (int i = threadIdx.x; i < R && row0 + i < n; i += kPullBlock) {
    const int64_t r = row0 + i;
    y[r] = (kAlpha + kBeta * Dv[rinst ? rinst[r] : 0]) * tau[r] + kBeta * ((double)acc[i] * kPullInv);
  }
}

// FP64 blocked push / pull: the same layout with FP64 products, summed
// exactly.  Each product v (0 <= v <= 1) is split into two integers,
// hi = rint(v 2^62) and lo = rint((v 2^62 - hi) 2^40) (the difference is
// exact), accumulated in two 64-bit LDS words per row: the row's sum is then
// known to 2^-102, independent of the order of the products, and converted
// to FP64 once.  A bin holds at most kPullRowsMax64 rows (two 32 KB tables).
constexpr int kPullRowsMax64 = 4096;
// Row bounds the FP64 push keeps in LDS: 52 KB, 3 workgroups per CU.  At
// 529M arcs one FP64 steady state took 1.013 s at 3 (and at 2) workgroups
// per CU, 1.071 s at 4 (the FP32 push's 40 KB) and 1.254 s at 8; bins of
// 2048 and 8192 rows instead of 4096 took 1.17 and 1.06 s.
constexpr int kPushRows64 = 13312;
constexpr double kHiScale = 4611686018427387904.0;            // 2^62
constexpr double kLoScale = 1099511627776.0;                  // 2^40
constexpr double kHiInv = 1.0 / 4611686018427387904.0;        // 2^-62
constexpr double kLoInv = 1.0 / 5070602400912917605986812821504.0;   // 2^-102

typedef double f64x4 __attribute__((ext_vector_type(4)));
__device__ inline f64x4 stream_d4(const double *p) { return __builtin_nontemporal_load((const f64x4 *)p); }

template <int ROWS>
__global__ __launch_bounds__(256) void k_pb_push64(const uint32_t *__restrict__ optr,
                                                   const uint32_t *__restrict__ trow, int64_t ntiles, int64_t per_xcd,
                                                   int64_t m, const uint32_t *__restrict__ slot,
                                                   const double *__restrict__ w, const double *__restrict__ x,
                                                   double *__restrict__ val) {
  __shared__ uint32_t sp[ROWS];
  const int64_t tile = (int64_t)(blockIdx.x % 8) * per_xcd + blockIdx.x / 8;
  if (tile >= ntiles) return;
  const int64_t a0 = tile * kPushTile;
  const int64_t a1 = std::min<int64_t>(m, a0 + kPushTile);
  const uint32_t r0 = trow[tile], r1 = trow[tile + 1];
  const int64_t nr = (int64_t)r1 - r0 + 1;
  const bool in_lds = nr + 1 <= ROWS;
  if (in_lds)
    for (int64_t i = threadIdx.x; i <= nr; i += 256) sp[i] = optr[r0 + i];
  __syncthreads();
  const int64_t k0 = a0 + 8 * (int64_t)threadIdx.x;
  if (k0 >= a1) return;
  const uint32_t *bounds = in_lds ? sp : optr + r0;
  int64_t lo = 0, hi = nr;
  while (hi - lo > 1) {
    const int64_t mid = (lo + hi) >> 1;
    if ((int64_t)bounds[mid] <= k0) lo = mid;
    else hi = mid;
  }
  int64_t row = lo;
  uint32_t nxt = bounds[row + 1];
  double xr = x[r0 + row];
  if (k0 + 8 <= a1) {
    const u32x4 s0 = stream_u4(slot + k0), s1 = stream_u4(slot + k0 + 4);
    const f64x4 w0 = stream_d4(w + k0), w1 = stream_d4(w + k0 + 4);
    const uint32_t sl[8] = {s0.x, s0.y, s0.z, s0.w, s1.x, s1.y, s1.z, s1.w};
    const double ww[8] = {w0.x, w0.y, w0.z, w0.w, w1.x, w1.y, w1.z, w1.w};
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const uint32_t k = (uint32_t)(k0 + j);
      if (k >= nxt) {
        do {
          ++row;
          nxt = bounds[row + 1];
        } while (k >= nxt);
        xr = x[r0 + row];
      }
      val[sl[j]] = ww[j] * xr;
    }
  } else {
    for (int64_t k = k0; k < a1; ++k) {
      if ((uint32_t)k >= nxt) {
        do {
          ++row;
          nxt = bounds[row + 1];
        } while ((uint32_t)k >= nxt);
        xr = x[r0 + row];
      }
      val[slot[k]] = w[k] * xr;
    }
  }
}

__device__ inline void add_split(unsigned long long *hi, unsigned long long *lo, double v) {
  const double t = v * kHiScale;
  const double h = rint(t);
  atomicAdd(hi, (unsigned long long)(long long)h);
  atomicAdd(lo, (unsigned long long)(long long)rint((t - h) * kLoScale));
}

__global__ __launch_bounds__(kPullBlock) void k_pb_pull64(const uint32_t *__restrict__ bptr,
                                                          const uint16_t *__restrict__ bdst,
                                                          const double *__restrict__ val, int32_t R, int32_t n,
                                                          const double *__restrict__ tau,
                                                          const double *__restrict__ Dv,
                                                          const uint32_t *__restrict__ rinst,
                                                          double *__restrict__ y) {
  __shared__ unsigned long long acc_hi[kPullRowsMax64];
  __shared__ unsigned long long acc_lo[kPullRowsMax64];
  for (int i = threadIdx.x; i < R; i += kPullBlock) {
    acc_hi[i] = 0ull;
    acc_lo[i] = 0ull;
  }
  __syncthreads();
  const uint32_t b = bptr[blockIdx.x], e = bptr[blockIdx.x + 1];
  uint32_t j = b + threadIdx.x;
  for (; (int64_t)j + 3 * kPullBlock < (int64_t)e; j += 4 * kPullBlock) {
    const uint16_t d0 = bdst[j], d1 = bdst[j + kPullBlock], d2 = bdst[j + 2 * kPullBlock],
                   d3 = bdst[j + 3 * kPullBlock];
    const double v0 = val[j], v1 = val[j + kPullBlock], v2 = val[j + 2 * kPullBlock], v3 = val[j + 3 * kPullBlock];
    add_split(&acc_hi[d0], &acc_lo[d0], v0);
    add_split(&acc_hi[d1], &acc_lo[d1], v1);
    add_split(&acc_hi[d2], &acc_lo[d2], v2);
    add_split(&acc_hi[d3], &acc_lo[d3], v3);
  }
  for (; j < e; j += kPullBlock) add_split(&acc_hi[bdst[j]], &acc_lo[bdst[j]], val[j]);
  __syncthreads();
  const int64_t row0 = (int64_t)blockIdx.x * R;
  for (int i = threadIdx.x; i < R && row0 + i < n; i += kPullBlock) {
    const int64_t r = row0 + i;
    const double sum = (double)(long long)acc_hi[i] * kHiInv + (double)(long long)acc_lo[i] * kLoInv;
    y[r] = (kAlpha + kBeta * Dv[rinst ? rinst[r] : 0]) * tau[r] + kBeta * sum;
  }
}

// --- reductions in a fixed launch shape (deterministic) ---------------------
__global__ __launch_bounds__(kRedBlock) void k_sum_partials(const double *__restrict__ v, int32_t n,
                                                            double *__restrict__ part) {
  __shared__ double lds[kRedBlock / kWave];
  double acc = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kRedBlock + threadIdx.x; i < n; i += (int64_t)kRedGrid * kRedBlock)
    acc += v[i];
  acc = block_sum<double, kRedBlock>(acc, lds);
  if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

__device__ inline double reduce_partials(const double *__restrict__ part, double *lds) {
  double a = 0.0;
  for (int i = threadIdx.x; i < kRedGrid; i += kRedBlock) a += part[i];
  a = block_sum<double, kRedBlock>(a, lds);
  if (threadIdx.x == 0) lds[0] = a;
  __syncthreads();
  double s = lds[0];
  __syncthreads();
  return s;
}

// x <- y / sum(y); partials of |x_new - x_old| and of dangling mass
__global__ __launch_bounds__(kRedBlock) void k_norm_step(const double *__restrict__ y, double *__restrict__ x,
                                                         float *__restrict__ xf, const uint8_t *__restrict__ dang,
                                                         int32_t n, const double *__restrict__ part_sum,
                                                         double *__restrict__ part_r, double *__restrict__ part_d) {
  __shared__ double lds[kRedBlock / kWave];
  const double s = reduce_partials(part_sum, lds);
  double r = 0.0, dm = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kRedBlock + threadIdx.x; i < n; i += (int64_t)kRedGrid * kRedBlock) {
    double pn = y[i] / s;
    r += fabs(pn - x[i]);
    if (dang[i]) dm += pn;
    x[i] = pn;
    xf[i] = (float)pn;
  }
  r = block_sum<double, kRedBlock>(r, lds);
  if (threadIdx.x == 0) part_r[blockIdx.x] = r;
  dm = block_sum<double, kRedBlock>(dm, lds);
  if (threadIdx.x == 0) part_d[blockIdx.x] = dm;
}

__global__ __launch_bounds__(kRedBlock) void k_final_rd(const double *__restrict__ part_r,
                                                        const double *__restrict__ part_d, double *__restrict__ scal) {
  __shared__ double lds[kRedBlock / kWave];
  double r = reduce_partials(part_r, lds);
  double d = reduce_partials(part_d, lds);
  if (threadIdx.x == 0) {
    scal[kR] = r;
    scal[kD] = d;
  }
}

__global__ __launch_bounds__(kRedBlock) void k_init_x(double *__restrict__ x, float *__restrict__ xf,
                                                      const uint8_t *__restrict__ dang, int32_t n,
                                                      double *__restrict__ part_d, double *__restrict__ part_r) {
  __shared__ double lds[kRedBlock / kWave];
  const double v = 1.0 / n;
  double dm = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kRedBlock + threadIdx.x; i < n; i += (int64_t)kRedGrid * kRedBlock) {
    x[i] = v;
    xf[i] = (float)v;
    if (dang[i]) dm += v;
  }
  dm = block_sum<double, kRedBlock>(dm, lds);
  if (threadIdx.x == 0) {
    part_d[blockIdx.x] = dm;
    part_r[blockIdx.x] = 0.0;
  }
}

// --- a batch of instances: one workgroup per instance over its node range --
__device__ inline double block_total(double v, double *lds, double *out) {
  v = block_sum<double, kRedBlock>(v, lds);
  if (threadIdx.x == 0) *out = v;
  __syncthreads();
  return *out;
}

// teleport weights: tau = w / (sum of w over the node's instance)
__global__ __launch_bounds__(kRedBlock) void k_btau(const uint32_t *__restrict__ off, const double *__restrict__ nw,
                                                    double *__restrict__ tau) {
  __shared__ double lds[kRedBlock / kWave];
  __shared__ double tot;
  const uint32_t b = off[blockIdx.x], e = off[blockIdx.x + 1];
  double s = 0.0;
  for (uint32_t i = b + threadIdx.x; i < e; i += kRedBlock) s += nw[i];
  s = block_total(s, lds, &tot);
  for (uint32_t i = b + threadIdx.x; i < e; i += kRedBlock) tau[i] = nw[i] / s;
}

// p = 1/n on every instance, its dangling mass; nothing converged yet
__global__ __launch_bounds__(kRedBlock) void k_binit(const uint32_t *__restrict__ off,
                                                     const uint8_t *__restrict__ dang, double *__restrict__ x,
                                                     float *__restrict__ xf, double *__restrict__ D,
                                                     double *__restrict__ R, uint8_t *__restrict__ done) {
  __shared__ double lds[kRedBlock / kWave];
  __shared__ double tot;
  const uint32_t b = off[blockIdx.x], e = off[blockIdx.x + 1];
  const double v = e > b ? 1.0 / (double)(e - b) : 0.0;
  double dm = 0.0;
  for (uint32_t i = b + threadIdx.x; i < e; i += kRedBlock) {
    x[i] = v;
    xf[i] = (float)v;
    if (dang[i]) dm += v;
  }
  dm = block_total(dm, lds, &tot);
  if (threadIdx.x == 0) {
    D[blockIdx.x] = dm;
    R[blockIdx.x] = 0.0;
    done[blockIdx.x] = e > b ? 0 : 1;
  }
}

// p <- y / sum(y) over the instance, its L1 change and dangling mass; an
// instance that has stopped keeps its p
__global__ __launch_bounds__(kRedBlock) void k_bnorm(const uint32_t *__restrict__ off, const double *__restrict__ y,
                                                     const uint8_t *__restrict__ dang,
                                                     const uint8_t *__restrict__ done, double *__restrict__ x,
                                                     float *__restrict__ xf, double *__restrict__ D,
                                                     double *__restrict__ R) {
  __shared__ double lds[kRedBlock / kWave];
  __shared__ double tot;
  if (done[blockIdx.x]) return;
  const uint32_t b = off[blockIdx.x], e = off[blockIdx.x + 1];
  double s = 0.0;
  for (uint32_t i = b + threadIdx.x; i < e; i += kRedBlock) s += y[i];
  s = block_total(s, lds, &tot);
  double r = 0.0, dm = 0.0;
  for (uint32_t i = b + threadIdx.x; i < e; i += kRedBlock) {
    const double pn = y[i] / s;
    r += fabs(pn - x[i]);
    if (dang[i]) dm += pn;
    x[i] = pn;
    xf[i] = (float)pn;
  }
  __syncthreads();
  r = block_sum<double, kRedBlock>(r, lds);
  if (threadIdx.x == 0) R[blockIdx.x] = r;
  dm = block_sum<double, kRedBlock>(dm, lds);
  if (threadIdx.x == 0) D[blockIdx.x] = dm;
}

// the stop rule after iteration `it`, per instance; *running counts the
// instances that go on
__global__ __launch_bounds__(kRedBlock) void k_bcheck(int32_t nb, int it, int min_iter, int max_iter, double tol,
                                                      const double *__restrict__ R, uint8_t *__restrict__ done,
                                                      uint32_t *__restrict__ iters, uint32_t *__restrict__ running) {
  __shared__ uint32_t lds[kRedBlock / kWave];
  uint32_t k = 0;
  for (int32_t I = threadIdx.x; I < nb; I += kRedBlock) {
    if (done[I]) continue;
    if (it >= max_iter || (it >= min_iter && !(R[I] > tol))) {
      done[I] = 1;
      iters[I] = (uint32_t)it;
    } else {
      ++k;
    }
  }
  k = block_sum<uint32_t, kRedBlock>(k, lds);
  if (threadIdx.x == 0) *running = k;
}

// sum of phi(p) over every instance's nodes (constant during the partition)
__global__ __launch_bounds__(kRedBlock) void k_bconst(const uint32_t *__restrict__ off, const double *__restrict__ p,
                                                      double *__restrict__ C) {
  __shared__ double lds[kRedBlock / kWave];
  __shared__ double tot;
  const uint32_t b = off[blockIdx.x], e = off[blockIdx.x + 1];
  double a = 0.0;
  for (uint32_t i = b + threadIdx.x; i < e; i += kRedBlock) a += plogp(p[i]);
  a = block_total(a, lds, &tot);
  if (threadIdx.x == 0) C[blockIdx.x] = a;
}

// --- flows on arcs and node statistics (Module.cpp:267-325) ---------------
template <int G>
__global__ void k_flows(const uint32_t *__restrict__ optr, int32_t n, double *__restrict__ out_w,
                        float *__restrict__ out_wf, long long *__restrict__ out_wx, const double *__restrict__ x,
                        const double *__restrict__ tau, const uint8_t *__restrict__ dang, int mixed,
                        double *__restrict__ p_out, Stat4 *__restrict__ us, double *__restrict__ qlit) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t r = t / G;
  const int lane = (int)(t % G);
  if (r >= n) return;
  const uint32_t b = optr[r], e = optr[r + 1];
  const double pr = x[r];
  double sf = 0.0;
  long long sx = 0;
  for (uint32_t k0 = b + lane; k0 < e; k0 += kRowAhead * G) {
    double w[kRowAhead];
#pragma unroll
    for (int j = 0; j < kRowAhead; ++j) w[j] = k0 + j * G < e ? out_w[k0 + j * G] : 0.0;
#pragma unroll
    for (int j = 0; j < kRowAhead; ++j) {
      const uint32_t k = k0 + j * G;
      if (k >= e) break;
      const double f = pr * w[j];
      const float ff = (float)f;
      out_w[k] = f;
      if (mixed) out_wf[k] = ff;   // the contraction's weights; the sweeps read out_wx
      sf += f;
      const long long fx = fx_of(mixed ? (double)ff : f);
      out_wx[k] = fx;
      sx += fx;
    }
  }
  sf = group_sum<G>(sf);
  sx = group_sum_ll<G>(sx);
  if (lane == 0) {
    const double tp = tau[r];
    const bool dg = dang[r] != 0;
    p_out[r] = pr;
    Stat4 st;
    st.p = fx_of(pr);
    st.t = fx_of(tp);
    st.d = dg ? fx_of(pr) : 0;
    st.x = sx;
    us[r] = st;
    qlit[r] = dg ? (1.0 - tp) * pr : kAlpha * (1.0 - tp) * pr + kBeta * sf;
  }
}

__global__ void k_tau(const double *__restrict__ nw, int32_t n, const double *__restrict__ tot,
                      double *__restrict__ tau) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) tau[i] = nw[i] / tot[0];
}

// calibrate over singleton modules with the node exits of Module.cpp:86-95
__global__ __launch_bounds__(kRedBlock) void k_calib(const double *__restrict__ qlit, const double *__restrict__ p,
                                                     int32_t n, double *__restrict__ part) {
  __shared__ double lds[kRedBlock / kWave];
  double s = 0, se = 0, ss = 0, sp = 0;
  for (int64_t i = (int64_t)blockIdx.x * kRedBlock + threadIdx.x; i < n; i += (int64_t)kRedGrid * kRedBlock) {
    double q = qlit[i], pi = p[i];
    s += q;
    se += plogp(q);
    ss += plogp(q + pi);
    sp += plogp(pi);
  }
  s = block_sum<double, kRedBlock>(s, lds);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
  se = block_sum<double, kRedBlock>(se, lds);
  if (threadIdx.x == 0) part[kRedGrid + blockIdx.x] = se;
  ss = block_sum<double, kRedBlock>(ss, lds);
  if (threadIdx.x == 0) part[2 * kRedGrid + blockIdx.x] = ss;
  sp = block_sum<double, kRedBlock>(sp, lds);
  if (threadIdx.x == 0) part[3 * kRedGrid + blockIdx.x] = sp;
}

__global__ __launch_bounds__(kRedBlock) void k_calib_final(const double *__restrict__ part, double *__restrict__ scal) {
  __shared__ double lds[kRedBlock / kWave];
  double s = reduce_partials(part, lds);
  double se = reduce_partials(part + kRedGrid, lds);
  double ss = reduce_partials(part + 2 * kRedGrid, lds);
  double sp = reduce_partials(part + 3 * kRedGrid, lds);
  if (threadIdx.x == 0) {
    scal[kS] = s;
    scal[kSP] = sp;
    scal[kL] = plogp(s) - 2.0 * se + ss - sp;
  }
}

template <typename T>
__global__ void k_copy(const T *__restrict__ a, T *__restrict__ b, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = a[i];
}

__global__ void k_singletons(int32_t n, uint32_t *__restrict__ mod, int32_t *__restrict__ nmem,
                             int32_t *__restrict__ mNodes, int32_t *__restrict__ mUnits,
                             int32_t *__restrict__ mStale) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  mod[i] = (uint32_t)i;
  nmem[i] = 1;
  mNodes[i] = 1;
  mUnits[i] = 1;
  mStale[i] = 1;
}

// The pull side of the iteration: in-CSR, weights, teleport, dangling flags.
struct PullCSR {
  int32_t n = 0;
  const uint32_t *ptr = nullptr, *idx = nullptr;
  const double *w = nullptr;
  const double *tau = nullptr;
  const uint8_t *dangling = nullptr;
  const double *Dv = nullptr;        // dangling mass: one value, or one per instance (batch)
  const uint32_t *rinst = nullptr;   // instance of every row (batch), else null
};

__global__ void k_out_degree_key(const uint32_t *__restrict__ optr, int32_t n, uint32_t *__restrict__ key) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) key[i] = ~(optr[i + 1] - optr[i]);   // ascending key = descending degree
}

__global__ void k_invert(const uint32_t *__restrict__ inv, int32_t n, uint32_t *__restrict__ perm) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r < n) perm[inv[r]] = (uint32_t)r;
}

// per new row: in-degree, teleport weight and dangling flag of its node
__global__ void k_perm_rows(const uint32_t *__restrict__ inv, int32_t n, const uint32_t *__restrict__ iptr,
                            const double *__restrict__ tau, const uint8_t *__restrict__ dang,
                            uint32_t *__restrict__ deg, double *__restrict__ ptau, uint8_t *__restrict__ pdang) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const uint32_t v = inv[r];
  deg[r] = iptr[v + 1] - iptr[v];
  ptau[r] = tau[v];
  pdang[r] = dang[v];
}

// G lanes per new row: the node's in-arcs in their order, tails relabelled
template <int G>
__global__ void k_perm_arcs(const uint32_t *__restrict__ inv, const uint32_t *__restrict__ perm, int32_t n,
                            const uint32_t *__restrict__ iptr, const uint32_t *__restrict__ iidx,
                            const double *__restrict__ iw, const uint32_t *__restrict__ pptr,
                            uint32_t *__restrict__ pidx, double *__restrict__ pw) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t r = t / G;
  const int lane = (int)(t % G);
  if (r >= n) return;
  const uint32_t v = inv[r];
  const uint32_t b = iptr[v], e = iptr[v + 1], o = pptr[r];
  for (uint32_t k = b + lane; k < e; k += G) {
    const uint32_t j = o + (k - b);
    pidx[j] = perm[iidx[k]];
    pw[j] = iw[k];
  }
}

// p on the original labels
__global__ void k_unpermute(const double *__restrict__ xp, const uint32_t *__restrict__ perm, int32_t n,
                            double *__restrict__ x) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v < n) x[v] = xp[perm[v]];
}

// Row bins of the pull SpMV (in-degree classes), stable in row order.
struct RowBins {
  DevBuf<uint32_t> rows;
  int32_t off[kNumRowBins + 1] = {0};
  int64_t arcs[kNumRowBins] = {0};
};

// Algorithmic bytes of one bin launch: per in-arc the source id, the weight
// and the gathered flow (4 + w + x); per row the row id, its CSR bounds, tau
// and the written result (4 + 8 + 8 + 8).
template <int G>
void launch_spmv_bin(int bin, const RowBins &rb, const PullCSR &L, Ctx &c, double *y) {
  // profiling entries carry the kernels' own names
  static const char *names[kNumRowBins] = {"k_spmv<4>", "k_spmv<8>", "k_spmv<16>", "k_spmv<32>", "k_spmv<64>",
                                           "k_spmv_block"};
  const int32_t cnt = rb.off[bin + 1] - rb.off[bin];
  if (cnt <= 0) return;
  const uint32_t *rows = rb.rows.p + rb.off[bin];
  if (c.prof.on) c.prof.begin(c.stream);
  if (G > 0) {
    constexpr int GG = G > 0 ? G : 1;
    k_spmv<GG><<<grid_for((int64_t)cnt * GG, 256), 256, 0, c.stream>>>(rows, cnt, L.ptr, L.idx, L.w, c.x.p,
                                                                        L.tau, L.Dv, L.rinst, y);
  } else {
    unsigned g = cnt < 4096 ? (unsigned)cnt : 4096u;
    k_spmv_block<<<g, 256, 0, c.stream>>>(rows, cnt, L.ptr, L.idx, L.w, c.x.p, L.tau, L.Dv, L.rinst, y);
  }
  RMX_LAUNCH_CHECK();
  if (c.prof.on) c.prof.end(c.stream, names[bin], (double)rb.arcs[bin] * 20.0 + (double)cnt * 28.0);
}

void spmv_all(Ctx &c, const PullCSR &L, const RowBins &rb, double *y) {
  launch_spmv_bin<4>(0, rb, L, c, y);
  launch_spmv_bin<8>(1, rb, L, c, y);
  launch_spmv_bin<16>(2, rb, L, c, y);
  launch_spmv_bin<32>(3, rb, L, c, y);
  launch_spmv_bin<64>(4, rb, L, c, y);
  launch_spmv_bin<0>(5, rb, L, c, y);
}

// Rows per bin: a power of two from 256 to kPullRowsMax (kPullRowsMax64 for
// FP64 products), about n / 2048, so that a large network gives thousands of
// bins (workgroups of the pull).  The slots come from the per-tile counts
// above when the bins fit the LDS counters and the level has its in-CSR
// (FP32 products: in the staged push's tiles), else from a stable sort of
// the out-arcs by bin.  A network without arcs has no push tiles.
void build_push_bins(Ctx &c, const Level &L, PushBins &pb) {
  hipStream_t s = c.stream;
  const int32_t n = L.n;
  const int64_t m = L.m;
  pb.staged = false;
  pb.stiles = 0;
  pb.ntiles = 0;
  const int rows_max = pb.fp64 ? kPullRowsMax64 : kPullRowsMax;
  int rs = 8;
  while ((1 << rs) < rows_max && ((int64_t)1 << rs) * 2048 < n) ++rs;
  pb.rshift = rs;
  pb.R = 1 << rs;
  pb.nbins = (int32_t)(((int64_t)n + pb.R - 1) >> rs);
  pb.bptr.alloc((int64_t)pb.nbins + 1);
  if (m == 0) {
    fill_u32(pb.bptr.p, 0u, (int64_t)pb.nbins + 1, s);
    return;
  }
  if (!pb.fp64 && pb.nbins <= kLayoutMaxBins && L.in_ptr.n > n) {
    // FP32 products: the staged push's layout (tiles of kStageTile arcs)
    const int64_t ntl = (m + kStageTile - 1) / kStageTile;
    const int64_t nq = (ntl + kLayoutChunk - 1) / kLayoutChunk;
    k_pb_bptr<<<grid_for((int64_t)pb.nbins + 1, 256), 256, 0, s>>>(L.in_ptr.p, n, pb.nbins, rs, pb.bptr.p);
    RMX_LAUNCH_CHECK();
    DevBuf<uint32_t> C(c.cache), Ccnt(c.cache), P(c.cache);
    C.alloc(ntl * pb.nbins);
    Ccnt.alloc(ntl * pb.nbins);
    P.alloc(nq * pb.nbins);
    k_pb_count<<<(unsigned)ntl, kLayoutBlock, 0, s>>>(L.out_idx.p, m, kStageTile, pb.nbins, rs, C.p);
    RMX_LAUNCH_CHECK();
    RMX_HIP(hipMemcpyAsync(Ccnt.p, C.p, sizeof(uint32_t) * ntl * pb.nbins, hipMemcpyDeviceToDevice, s));
    k_pb_colsum<<<grid_for(nq * pb.nbins, 256), 256, 0, s>>>(C.p, ntl, pb.nbins, nq, P.p);
    RMX_LAUNCH_CHECK();
    k_pb_colscan<<<grid_for(pb.nbins, 256), 256, 0, s>>>(pb.bptr.p, pb.nbins, nq, P.p);
    RMX_LAUNCH_CHECK();
    k_pb_colbase<<<grid_for(nq * pb.nbins, 256), 256, 0, s>>>(P.p, ntl, pb.nbins, nq, C.p);
    RMX_LAUNCH_CHECK();
    pb.lpos.alloc(m);
    pb.gslot.alloc(m);
    pb.bdst.alloc(m);
    pb.grow.alloc((m + 7) / 8);
    pb.alloc_values(m);
    k_pb_place_staged<<<(unsigned)ntl, kLayoutBlock, 0, s>>>(L.out_idx.p, m, pb.nbins, rs, C.p, Ccnt.p, pb.lpos.p,
                                                             pb.gslot.p, pb.bdst.p);
    RMX_LAUNCH_CHECK();
    k_pb_group_rows<<<grid_for((int64_t)n * kWave, 256), 256, 0, s>>>(L.out_ptr.p, n, pb.grow.p);
    RMX_LAUNCH_CHECK();
    pb.staged = true;
    pb.stiles = ntl;
    return;   // C, Ccnt and P go back to the cache: later work on this stream runs after the layout
  }
  pb.ntiles = (m + kPushTile - 1) / kPushTile;
  pb.trow.alloc(pb.ntiles + 1);
  k_pb_tile_rows<<<grid_for(pb.ntiles + 1, 256), 256, 0, s>>>(L.out_ptr.p, n, m, pb.ntiles, pb.trow.p);
  RMX_LAUNCH_CHECK();
  if (pb.nbins <= kLayoutMaxBins && L.in_ptr.n > n) {
    // at most 8192 layout tiles of at least 64k arcs
    const int64_t tile_arcs = std::max<int64_t>(65536, (m + 8191) / 8192);
    const int64_t ntl = (m + tile_arcs - 1) / tile_arcs;
    const int64_t nq = (ntl + kLayoutChunk - 1) / kLayoutChunk;
    k_pb_bptr<<<grid_for((int64_t)pb.nbins + 1, 256), 256, 0, s>>>(L.in_ptr.p, n, pb.nbins, rs, pb.bptr.p);
    RMX_LAUNCH_CHECK();
    DevBuf<uint32_t> C(c.cache), P(c.cache);
    C.alloc(ntl * pb.nbins);
    P.alloc(nq * pb.nbins);
    k_pb_count<<<(unsigned)ntl, kLayoutBlock, 0, s>>>(L.out_idx.p, m, tile_arcs, pb.nbins, rs, C.p);
    RMX_LAUNCH_CHECK();
    k_pb_colsum<<<grid_for(nq * pb.nbins, 256), 256, 0, s>>>(C.p, ntl, pb.nbins, nq, P.p);
    RMX_LAUNCH_CHECK();
    k_pb_colscan<<<grid_for(pb.nbins, 256), 256, 0, s>>>(pb.bptr.p, pb.nbins, nq, P.p);
    RMX_LAUNCH_CHECK();
    k_pb_colbase<<<grid_for(nq * pb.nbins, 256), 256, 0, s>>>(P.p, ntl, pb.nbins, nq, C.p);
    RMX_LAUNCH_CHECK();
    pb.slot.alloc(m);
    pb.bdst.alloc(m);
    pb.alloc_values(m);
    const int64_t per_xcd = (ntl + 7) / 8;
    k_pb_place<<<(unsigned)(8 * per_xcd), kLayoutBlock, 0, s>>>(L.out_idx.p, m, tile_arcs, ntl, per_xcd, pb.nbins,
                                                                rs, C.p, pb.slot.p, pb.bdst.p);
    RMX_LAUNCH_CHECK();
    return;   // C and P go back to the cache: later work on this stream runs after the layout
  }
  DevBuf<uint32_t> key(c.cache), arc(c.cache);
  key.alloc(m);
  arc.alloc(m);
  k_pb_keys<<<grid_for(m, 256), 256, 0, s>>>(L.out_idx.p, m, rs, key.p);
  RMX_LAUNCH_CHECK();
  iota_u32(arc.p, m, s);
  if (pb.nbins > 1) {
    int bits = 1;
    while (((int64_t)1 << bits) < pb.nbins) ++bits;
    radix_sort_pairs_u32(key.p, arc.p, m, 0, bits, c.scratch, s);
  }
  k_pb_bounds<<<grid_for((int64_t)pb.nbins + 1, 256), 256, 0, s>>>(key.p, m, pb.nbins, pb.bptr.p);
  RMX_LAUNCH_CHECK();
  pb.slot.alloc(m);
  pb.bdst.alloc(m);
  pb.alloc_values(m);
  k_pb_layout<<<grid_for(m, 256), 256, 0, s>>>(arc.p, L.out_idx.p, m, rs, pb.slot.p, pb.bdst.p);
  RMX_LAUNCH_CHECK();
}

// y = one iteration's unnormalised p' from xf.  Algorithmic bytes: the push
// reads per arc its slot and FP32 weight and writes its product (12), per
// row its CSR bound and x (8); the pull reads per arc the product and its
// 16-bit head (6), per row tau and writes y (16, and 4 for the instance).
void push_pull(Ctx &c, const PushBins &pb, const Level &L, const PullCSR &F, double *y) {
  hipStream_t s = c.stream;
  if (pb.fp64) {   // FP64 products: per arc slot, FP64 weight and product (20 B); pull 10 B per arc
    if (pb.ntiles > 0) {
      const int64_t per = (pb.ntiles + 7) / 8;
      if (c.prof.on) c.prof.begin(s);
      k_pb_push64<kPushRows64><<<(unsigned)(8 * per), 256, 0, s>>>(L.out_ptr.p, pb.trow.p, pb.ntiles, per, L.m,
                                                                       pb.slot.p, L.out_w.p, c.x.p, pb.val64.p);
      RMX_LAUNCH_CHECK();
      if (c.prof.on)
        c.prof.end(s, "k_pb_push64<" + std::to_string(kPushRows64) + ">", (double)L.m * 20.0 + (double)L.n * 12.0);
    }
    if (c.prof.on) c.prof.begin(s);
    k_pb_pull64<<<(unsigned)pb.nbins, kPullBlock, 0, s>>>(pb.bptr.p, pb.bdst.p, pb.val64.p, pb.R, L.n, F.tau, F.Dv,
                                                          F.rinst, y);
    RMX_LAUNCH_CHECK();
    if (c.prof.on) c.prof.end(s, "k_pb_pull64", (double)L.m * 10.0 + (double)L.n * (F.rinst ? 20.0 : 16.0));
    return;
  }
  if (pb.staged && pb.stiles > 0) {
    // per arc its position, FP32 weight, the position's slot and the product (14 B),
    // per 8 arcs a row (0.5 B), per row its bound and x (8)
    const int64_t per = (pb.stiles + 7) / 8;
    if (c.prof.on) c.prof.begin(s);
    k_pb_push_staged<<<(unsigned)(8 * per), kStageBlock, 0, s>>>(L.out_ptr.p, pb.grow.p, pb.stiles, per, L.m,
                                                                  pb.lpos.p, pb.gslot.p, L.out_wf.p, c.xf.p,
                                                                  pb.val.p);
    RMX_LAUNCH_CHECK();
    if (c.prof.on) c.prof.end(s, "k_pb_push_staged", (double)L.m * 14.5 + (double)L.n * 8.0);
  } else if (pb.ntiles > 0) {
    const int64_t per = (pb.ntiles + 7) / 8;
    if (c.prof.on) c.prof.begin(s);
    k_pb_push<<<(unsigned)(8 * per), 256, 0, s>>>(L.out_ptr.p, pb.trow.p, pb.ntiles, per, L.m, pb.slot.p,
                                                    L.out_wf.p, c.xf.p, pb.val.p);
    RMX_LAUNCH_CHECK();
    if (c.prof.on) c.prof.end(s, "k_pb_push", (double)L.m * 12.0 + (double)L.n * 8.0);
  }
  if (c.prof.on) c.prof.begin(s);
  k_pb_pull<<<(unsigned)pb.nbins, kPullBlock, 0, s>>>(pb.bptr.p, pb.bdst.p, pb.val.p, pb.R, L.n, F.tau, F.Dv,
                                                      F.rinst, y);
  RMX_LAUNCH_CHECK();
  if (c.prof.on) c.prof.end(s, "k_pb_pull", (double)L.m * 6.0 + (double)L.n * (F.rinst ? 20.0 : 16.0));
}

}  // namespace

static void make_row_bins(Ctx &c, const PullCSR &L, RowBins &rb) {
  hipStream_t s = c.stream;
  DevBuf<uint32_t> key(c.cache), cnt(c.cache);
  key.alloc(L.n);
  cnt.alloc(32);
  fill_u32(cnt.p, 0u, 32, s);
  k_row_bin<<<(unsigned)std::min<int64_t>(2048, (L.n + 255) / 256), 256, 0, s>>>(L.ptr, L.n, key.p, cnt.p,
                                               (unsigned long long *)(cnt.p + 16));
  RMX_LAUNCH_CHECK();
  rb.rows.alloc(L.n);
  iota_u32(rb.rows.p, L.n, s);
  radix_sort_pairs_u32(key.p, rb.rows.p, L.n, 0, 8, c.scratch, s);
  uint32_t h[32];
  RMX_HIP(hipMemcpyAsync(h, cnt.p, sizeof(h), hipMemcpyDeviceToHost, s));
  c.sync();
  rb.off[0] = 0;
  for (int b = 0; b < kNumRowBins; ++b) {
    rb.off[b + 1] = rb.off[b] + (int32_t)h[b];
    rb.arcs[b] = (int64_t)((uint64_t)h[16 + 2 * b] | ((uint64_t)h[17 + 2 * b] << 32));
  }
}

void steady_state(Ctx &c, int max_iter, int min_iter, double tol) {
  hipStream_t s = c.stream;
  Level &L = c.L0;
  const int32_t n = L.n;
  const int64_t m = L.m;
  const bool mixed = c.precision == 1;
  const bool blocked = mixed || c.fp64_blocked;   // push / pull bins (else the FP64 pull)
  L.out_w.alloc(m);
  L.in_w.alloc(m);
  L.out_wf.alloc(m);
  L.in_wf.alloc(m);
  L.p.alloc(n);
  L.tau.alloc(n);
  L.us.alloc(n);
  L.nmem.alloc(n);
  L.mod.alloc(n);
  c.dangling.alloc(n);
  c.x.alloc(n);
  c.y.alloc(n);
  c.xf.alloc(n);
  c.node_q.alloc(n);
  c.scal.alloc(16);
  c.partial.alloc(8 * kRedGrid);
  double *part = c.partial.p;

  const bool batch = c.nb > 0;
  if (batch) {
    c.b_D.alloc(c.nb);
    c.b_R.alloc(c.nb);
    c.b_iters.alloc(c.nb);
    c.b_count.alloc(4);
  }
  // teleport weights tau = w / sum w (OmpRelaxmap.cpp:113-115), per instance in a batch
  if (batch) {
    k_btau<<<c.nb, kRedBlock, 0, s>>>(c.b_off.p, c.node_w.p, L.tau.p);
    RMX_LAUNCH_CHECK();
  } else {
    k_sum_partials<<<kRedGrid, kRedBlock, 0, s>>>(c.node_w.p, n, part);
    RMX_LAUNCH_CHECK();
    k_final_rd<<<1, kRedBlock, 0, s>>>(part, part, c.scal.p);   // scal[kR] = scal[kD] = sum w
    RMX_LAUNCH_CHECK();
    k_tau<<<grid_for(n, 256), 256, 0, s>>>(c.node_w.p, n, c.scal.p + kR, L.tau.p);
    RMX_LAUNCH_CHECK();
  }

  DevBuf<double> rowsum(c.cache);
  rowsum.alloc(n);
  if (mixed) {
    k_normalize<16, true><<<grid_for((int64_t)n * 16, 256), 256, 0, s>>>(L.out_ptr.p, n, c.raw_w.p, L.out_w.p,
                                                                        L.out_wf.p, rowsum.p, c.dangling.p);
    RMX_LAUNCH_CHECK();
  } else {
    k_normalize<16, false><<<grid_for((int64_t)n * 16, 256), 256, 0, s>>>(L.out_ptr.p, n, c.raw_w.p, L.out_w.p,
                                                                         L.out_wf.p, rowsum.p, c.dangling.p);
    RMX_LAUNCH_CHECK();
    if (m > 0 && !blocked) {   // the pull's in-arc weights
      k_in_norm<<<stride_grid(m), 256, 0, s>>>(L.in_idx.p, c.raw_in.p, m, rowsum.p, L.in_w.p);
      RMX_LAUNCH_CHECK();
    }
  }
  // FP64: the relabelled pull copy (descending out-degree), built on the
  // device; mixed: the push / pull bins on the original labels
  DevBuf<uint32_t> inv(c.cache), perm(c.cache), pptr(c.cache), pidx(c.cache), pdeg(c.cache);
  DevBuf<double> pw(c.cache), ptau(c.cache);
  DevBuf<float> pwf(c.cache);
  DevBuf<uint8_t> pdang(c.cache);
  PushBins &pb = c.pb;
  PullCSR F;
  F.n = n;
  const bool relabel = !batch && !blocked;
  if (!relabel) {   // instances keep their node ranges
    F.ptr = L.in_ptr.p;
    F.idx = L.in_idx.p;
    F.w = L.in_w.p;
    F.tau = L.tau.p;
    F.dangling = c.dangling.p;
    F.Dv = batch ? c.b_D.p : c.scal.p + kD;
    F.rinst = batch ? c.b_node.p : nullptr;
    // the layout depends on the network's structure and the product type
    // only: built by the first steady state after an upload, then reused
    if (blocked && (c.pb_graph != c.graph_version || pb.fp64 != !mixed)) {
      pb.fp64 = !mixed;
      build_push_bins(c, L, pb);
      c.pb_graph = c.graph_version;
    }
  } else {
    F.Dv = c.scal.p + kD;
    inv.alloc(n);
    perm.alloc(n);
    pptr.alloc((int64_t)n + 1);
    pdeg.alloc((int64_t)n + 1);
    ptau.alloc(n);
    pdang.alloc(n);
    pidx.alloc(m);
    pw.alloc(m);
    DevBuf<uint32_t> key(c.cache);
    key.alloc(n);
    k_out_degree_key<<<grid_for(n, 256), 256, 0, s>>>(L.out_ptr.p, n, key.p);
    RMX_LAUNCH_CHECK();
    iota_u32(inv.p, n, s);
    radix_sort_pairs_u32(key.p, inv.p, n, 0, 32, c.scratch, s);   // stable: ties by node id
    k_invert<<<grid_for(n, 256), 256, 0, s>>>(inv.p, n, perm.p);
    RMX_LAUNCH_CHECK();
    fill_u32(pdeg.p + n, 0u, 1, s);
    k_perm_rows<<<grid_for(n, 256), 256, 0, s>>>(inv.p, n, L.in_ptr.p, L.tau.p, c.dangling.p, pdeg.p, ptau.p,
                                                 pdang.p);
    RMX_LAUNCH_CHECK();
    exclusive_scan_u32(pdeg.p, pptr.p, (int64_t)n + 1, nullptr, c.scratch, s);
    if (m > 0) {
      k_perm_arcs<16><<<grid_for((int64_t)n * 16, 256), 256, 0, s>>>(inv.p, perm.p, n, L.in_ptr.p, L.in_idx.p,
                                                                      L.in_w.p, pptr.p, pidx.p, pw.p);
      RMX_LAUNCH_CHECK();
    }
    F.ptr = pptr.p;
    F.idx = pidx.p;
    F.w = pw.p;
    F.tau = ptau.p;
    F.dangling = pdang.p;
  }
  RowBins rb;
  if (!blocked) make_row_bins(c, F, rb);

  hipEvent_t e0, e1;
  RMX_HIP(hipEventCreate(&e0));
  RMX_HIP(hipEventCreate(&e1));
  RMX_HIP(hipEventRecord(e0, s));
  int it = 0;
  double r = 1.0;
  if (batch) {
    // every instance iterates until its own stop rule holds, then keeps its p
    const int32_t nb = c.nb;
    DevBuf<uint8_t> done(c.cache);
    done.alloc(nb);
    fill_u32(c.b_iters.p, 0u, nb, s);
    k_binit<<<nb, kRedBlock, 0, s>>>(c.b_off.p, c.dangling.p, c.x.p, c.xf.p, c.b_D.p, c.b_R.p, done.p);
    RMX_LAUNCH_CHECK();
    const int b_min = mixed ? 1 : min_iter;
    const double b_tol = mixed ? kMixedTol : tol;
    uint32_t running = 1;
    while (running > 0) {
      if (blocked) push_pull(c, pb, L, F, c.y.p);
      else spmv_all(c, F, rb, c.y.p);
      k_bnorm<<<nb, kRedBlock, 0, s>>>(c.b_off.p, c.y.p, c.dangling.p, done.p, c.x.p, c.xf.p, c.b_D.p, c.b_R.p);
      RMX_LAUNCH_CHECK();
      ++it;
      k_bcheck<<<1, kRedBlock, 0, s>>>(nb, it, b_min, max_iter, b_tol, c.b_R.p, done.p, c.b_iters.p, c.b_count.p);
      RMX_LAUNCH_CHECK();
      RMX_HIP(hipMemcpyAsync(&running, c.b_count.p, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
      c.sync();
    }
    std::vector<double> hr(nb);
    RMX_HIP(hipMemcpyAsync(hr.data(), c.b_R.p, sizeof(double) * nb, hipMemcpyDeviceToHost, s));
    c.sync();
    r = 0.0;
    for (double v : hr) r = v > r ? v : r;   // the largest residual of the batch
  } else {
  k_init_x<<<kRedGrid, kRedBlock, 0, s>>>(c.x.p, c.xf.p, F.dangling, n, part + 2 * kRedGrid, part + kRedGrid);
  RMX_LAUNCH_CHECK();
  k_final_rd<<<1, kRedBlock, 0, s>>>(part + kRedGrid, part + 2 * kRedGrid, c.scal.p);
  RMX_LAUNCH_CHECK();
  std::vector<double> hist;
  const int first_check = mixed ? 1 : min_iter;
  while (it < max_iter) {
    if (blocked) push_pull(c, pb, L, F, c.y.p);
    else spmv_all(c, F, rb, c.y.p);
    if (c.prof.on) c.prof.begin(s);
    k_sum_partials<<<kRedGrid, kRedBlock, 0, s>>>(c.y.p, n, part);
    RMX_LAUNCH_CHECK();
    k_norm_step<<<kRedGrid, kRedBlock, 0, s>>>(c.y.p, c.x.p, c.xf.p, F.dangling, n, part, part + kRedGrid,
                                               part + 2 * kRedGrid);
    RMX_LAUNCH_CHECK();
    k_final_rd<<<1, kRedBlock, 0, s>>>(part + kRedGrid, part + 2 * kRedGrid, c.scal.p);
    RMX_LAUNCH_CHECK();
    // y read twice, x read and written, FP32 copy and dangling flag
    if (c.prof.on) c.prof.end(s, "pr_normalize", (double)n * 37.0);
    ++it;
    if (it >= first_check) {
      RMX_HIP(hipMemcpyAsync(c.h_scal, c.scal.p, sizeof(double) * 8, hipMemcpyDeviceToHost, s));
      c.sync();
      r = c.h_scal[kR];
      if (!(r > (mixed ? kMixedTol : tol))) break;
      hist.push_back(r);
      if (mixed && hist.size() > 10 && r > 0.5 * hist[hist.size() - 11]) break;
    }
  }
  }
  RMX_HIP(hipEventRecord(e1, s));
  RMX_HIP(hipEventSynchronize(e1));
  float ms = 0;
  RMX_HIP(hipEventElapsedTime(&ms, e0, e1));
  c.timing.flow_ms = ms;
  RMX_HIP(hipEventDestroy(e0));
  RMX_HIP(hipEventDestroy(e1));
  c.iterations = it;
  c.resid = r;
  if (relabel) {   // p back on the original labels
    k_unpermute<<<grid_for(n, 256), 256, 0, s>>>(c.x.p, perm.p, n, c.y.p);
    RMX_LAUNCH_CHECK();
    c.x.swap(c.y);
  }
  c.sync();   // the relabelled copy goes back to the cache

  // flows, node statistics and the literal node exits
  // the sweeps' fixed-point arc weights: the same conversion fxw made of
  // out_wf / in_wf (mixed) or out_w / in_w (FP64), done once here
  L.out_wx.alloc(m);
  L.in_wx.alloc(m);
  k_flows<16><<<grid_for((int64_t)n * 16, 256), 256, 0, s>>>(L.out_ptr.p, n, L.out_w.p, L.out_wf.p, L.out_wx.p, c.x.p,
                                                              L.tau.p, c.dangling.p, mixed ? 1 : 0, L.p.p, L.us.p,
                                                              c.node_q.p);
  RMX_LAUNCH_CHECK();
  if (m > 0) {
    DevBuf<double2> ps(c.cache);
    ps.alloc(n);
    k_pack_ps<<<grid_for(n, 256), 256, 0, s>>>(c.x.p, rowsum.p, n, ps.p);
    RMX_LAUNCH_CHECK();
    k_in_flows<<<stride_grid(m), 256, 0, s>>>(L.in_idx.p, c.raw_in.p, m, ps.p, mixed ? 1 : 0,
                                               mixed ? nullptr : L.in_w.p, mixed ? L.in_wf.p : nullptr, L.in_wx.p);
    RMX_LAUNCH_CHECK();
    c.sync();   // ps goes back to the cache
  }
  k_calib<<<kRedGrid, kRedBlock, 0, s>>>(c.node_q.p, L.p.p, n, part);
  RMX_LAUNCH_CHECK();
  k_calib_final<<<1, kRedBlock, 0, s>>>(part, c.scal.p);
  RMX_LAUNCH_CHECK();
  RMX_HIP(hipMemcpyAsync(c.h_scal, c.scal.p, sizeof(double) * 8, hipMemcpyDeviceToHost, s));
  c.sync();
  c.initial_codelength = c.h_scal[kL];
  c.initial_sum_exit = c.h_scal[kS];
  c.all_log_all = c.h_scal[kSP];
  if (batch) {   // sum of phi(p) per instance; the codelengths come from the singleton modules
    c.b_C.alloc(c.nb);
    k_bconst<<<c.nb, kRedBlock, 0, s>>>(c.b_off.p, L.p.p, c.b_C.p);
    RMX_LAUNCH_CHECK();
  }
  {
    // number of dangling nodes
    std::vector<uint8_t> dg(n);
    RMX_HIP(hipMemcpyAsync(dg.data(), c.dangling.p, n, hipMemcpyDeviceToHost, s));
    c.sync();
    int64_t nd = 0;
    for (int32_t i = 0; i < n; ++i) nd += dg[i];
    c.n_dangling = nd;
  }
  c.x.free();
  c.xf.free();
  c.y.free();
}

void init_modules_singletons(Ctx &c) {
  hipStream_t s = c.stream;
  Level &L = c.L0;
  const int32_t n = L.n;
  c.ms.alloc(n);
  c.mNodes.alloc(n);
  c.mUnits.alloc(n);
  c.mStale.alloc(n);
  unsigned g = grid_for(n, 256);
  k_copy<Stat4><<<g, 256, 0, s>>>(L.us.p, c.ms.p, n);
  k_singletons<<<g, 256, 0, s>>>(n, L.mod.p, L.nmem.p, c.mNodes.p, c.mUnits.p, c.mStale.p);
  RMX_LAUNCH_CHECK();
  c.level = 0;
  c.n_module = n;
}

}  // namespace rmx
