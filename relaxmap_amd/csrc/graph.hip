// Graph upload, in-CSR construction and degree bucketing.
//
// The host hands over the out-CSR the reference builds in
// OmpRelaxmap.cpp:117-135 (arcs sorted by (src, dst), self-loops and w <= 0
// removed).  In-links are built on the device by a stable radix sort of the
// arcs by destination, so every in-list is in ascending source order as the
// reference's push_back loop produces it (OmpRelaxmap.cpp:158-162).
#include "context.h"

#include <algorithm>
#include <vector>

namespace rmx {

void Level::free_all() {
  out_ptr.free(); out_idx.free(); in_ptr.free(); in_idx.free();
  out_w.free(); in_w.free(); out_wf.free(); in_wf.free(); out_wx.free(); in_wx.free();
  p.free(); tau.free(); us.free();
  nmem.free(); mod.free(); bucket.free(); hub_off.free();
  uinst.free(); uoff.free();
  hub_pool = 0;
  hub_arcs = 0;
  hub_parts = 0;
  n = 0; m = 0;
}

void Level::swap(Level &o) {
  std::swap(n, o.n);
  std::swap(m, o.m);
  out_ptr.swap(o.out_ptr); out_idx.swap(o.out_idx); in_ptr.swap(o.in_ptr); in_idx.swap(o.in_idx);
  out_w.swap(o.out_w); in_w.swap(o.in_w); out_wf.swap(o.out_wf); in_wf.swap(o.in_wf);
  out_wx.swap(o.out_wx); in_wx.swap(o.in_wx);
  p.swap(o.p); tau.swap(o.tau); us.swap(o.us);
  nmem.swap(o.nmem); mod.swap(o.mod); bucket.swap(o.bucket); hub_off.swap(o.hub_off);
  for (int b = 0; b <= kNumBuckets; ++b) std::swap(bucket_off[b], o.bucket_off[b]);
  std::swap(hub_pool, o.hub_pool);
  std::swap(hub_arcs, o.hub_arcs);
  std::swap(hub_parts, o.hub_parts);
  std::swap(max_deg, o.max_deg);
  uinst.swap(o.uinst);
  uoff.swap(o.uoff);
}

namespace {

// src of every arc: one wave per row strides over the row's arcs
__global__ void k_arc_src(const uint32_t *__restrict__ ptr, int32_t n, uint32_t *__restrict__ src) {
  const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const int lane = threadIdx.x & (kWave - 1);
  if (wave >= n) return;
  for (uint32_t e = ptr[wave] + lane; e < ptr[wave + 1]; e += kWave) src[e] = (uint32_t)wave;
}

// FP64 weight of every in-arc (v <- u): the weight of u's out-arc to v,
// found by binary search in u's out-list (sorted by head)
__global__ void k_in_w_lookup(const uint32_t *__restrict__ iptr, const uint32_t *__restrict__ iidx, int32_t n,
                              const uint32_t *__restrict__ optr, const uint32_t *__restrict__ oidx,
                              const double *__restrict__ ow, double *__restrict__ iw) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  for (uint32_t k = iptr[v]; k < iptr[v + 1]; ++k) {
    const uint32_t u = iidx[k];
    uint32_t lo = optr[u], hi = optr[u + 1];
    while (lo < hi) {
      const uint32_t mid = (lo + hi) >> 1;
      if (oidx[mid] < (uint32_t)v) lo = mid + 1;
      else hi = mid;
    }
    iw[k] = (lo < optr[u + 1] && oidx[lo] == (uint32_t)v) ? ow[lo] : 0.0;
  }
}

// in_ptr from the destinations sorted ascending: in_ptr[d] is the first
// position whose destination is >= d, so position i fills every d in
// (key[i-1], key[i]] (and position m fills the tail up to n)
__global__ void k_ptr_from_sorted(const uint32_t *__restrict__ key, int64_t m, int32_t n,
                                  uint32_t *__restrict__ ptr) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i <= m; i += stride) {
    const int64_t lo = i == 0 ? 0 : (int64_t)key[i - 1] + 1;
    const int64_t hi = i == m ? (int64_t)n : (int64_t)key[i];
    for (int64_t d = lo; d <= hi; ++d) ptr[d] = (uint32_t)i;
  }
}

// after the stable sort: in_idx[k] = src[e_k], raw_in[k] = raw[e_k]
__global__ void k_in_from_sorted(const uint32_t *__restrict__ sorted_e, const uint32_t *__restrict__ src,
                                 const double *__restrict__ raw, int64_t m, uint32_t *__restrict__ in_idx,
                                 double *__restrict__ raw_in) {
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; k < m; k += stride) {
    uint32_t e = sorted_e[k];
    in_idx[k] = src[e];
    if (raw_in) raw_in[k] = raw[e];
  }
}

__global__ void k_bucket_key(const uint32_t *__restrict__ optr, const uint32_t *__restrict__ iptr, int32_t n,
                             BucketCuts cuts, uint32_t *__restrict__ key, uint32_t *__restrict__ maxdeg) {
  // grid-stride with a bounded grid and one atomic per workgroup: atomics on
  // a single address serialise across the device
  __shared__ uint32_t s_max[256 / kWave];
  uint32_t mx = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t deg = (optr[i + 1] - optr[i]) + (iptr[i + 1] - iptr[i]);
    uint32_t b = 0;
    while (b + 1 < (uint32_t)kNumBuckets && deg > (uint32_t)cuts.v[b]) ++b;
    key[i] = b;
    mx = deg > mx ? deg : mx;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const uint32_t o = __shfl_xor(mx, off, kWave);
    mx = o > mx ? o : mx;
  }
  if ((threadIdx.x & (kWave - 1)) == 0) s_max[threadIdx.x / kWave] = mx;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < 256 / kWave; ++w) mx = s_max[w] > mx ? s_max[w] : mx;
    if (mx) atomicMax(maxdeg, mx);
  }
}

__global__ void k_small_hist(const uint32_t *__restrict__ key, int64_t n, uint32_t *__restrict__ cnt) {
  __shared__ uint32_t h[16];
  if (threadIdx.x < 16) h[threadIdx.x] = 0;
  __syncthreads();
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) atomicAdd(&h[key[i] & 15u], 1u);
  __syncthreads();
  if (threadIdx.x < 16 && h[threadIdx.x]) atomicAdd(&cnt[threadIdx.x], h[threadIdx.x]);
}

__global__ void k_list_degree(const uint32_t *__restrict__ list, int32_t k, const uint32_t *__restrict__ optr,
                              const uint32_t *__restrict__ iptr, uint32_t *__restrict__ deg) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= k) return;
  const uint32_t u = list[i];
  deg[i] = (optr[u + 1] - optr[u]) + (iptr[u + 1] - iptr[u]);
}

unsigned gs_grid(int64_t n) {
  int64_t g = (n + 255) / 256;
  if (g > 16384) g = 16384;
  return (unsigned)(g < 1 ? 1 : g);
}

}  // namespace

namespace {

// the out-CSR, raw weights and node weights to the device, and the in-CSR
void upload_csr(Ctx &c, int32_t n, int64_t m, const int64_t *out_ptr, const int32_t *out_idx, const double *w,
                const double *node_w) {
  if (n <= 0) throw HipError("network has no nodes");
  if (m >= (int64_t)0xffffffffLL) throw HipError("more than 2^32-2 arcs");
  hipStream_t s = c.stream;
  Level &L = c.L0;
  L.free_all();
  L.n = n;
  L.m = m;
  c.N = n;
  c.E = m;
  std::vector<uint32_t> p32((size_t)n + 1);
  for (int32_t i = 0; i <= n; ++i) p32[i] = (uint32_t)out_ptr[i];
  L.out_ptr.alloc(n + 1);
  L.out_idx.alloc(m);
  c.raw_w.alloc(m);
  c.node_w.alloc(n);
  RMX_HIP(hipMemcpyAsync(L.out_ptr.p, p32.data(), sizeof(uint32_t) * (n + 1), hipMemcpyHostToDevice, s));
  if (m) {
    RMX_HIP(hipMemcpyAsync(L.out_idx.p, out_idx, sizeof(uint32_t) * m, hipMemcpyHostToDevice, s));
    RMX_HIP(hipMemcpyAsync(c.raw_w.p, w, sizeof(double) * m, hipMemcpyHostToDevice, s));
  }
  std::vector<double> nw((size_t)n, 1.0);
  if (node_w)
    for (int32_t i = 0; i < n; ++i) nw[i] = node_w[i] > 0.0 ? node_w[i] : 1.0;
  RMX_HIP(hipMemcpyAsync(c.node_w.p, nw.data(), sizeof(double) * n, hipMemcpyHostToDevice, s));
  c.sync();   // host staging vectors go out of scope
  build_in_csr(c, L, c.raw_w.p, &c.raw_in);
}

// Copies of the networks, one workgroup row per instance k (gridDim.y
// strides over the instances): instance k = g * R + r is network g, its
// nodes at NO[k], its arcs at AO[k].  Network g's rows are GO[g].. and its
// arcs GA[g].. in the uploaded union; every arc of it stays inside it, and
// its in-arcs occupy the same range GA[g].. of the in-CSR.
__global__ void k_rep_nodes(int32_t K, int32_t R, const uint32_t *__restrict__ NO, const uint32_t *__restrict__ AO,
                            const uint32_t *__restrict__ GO, const uint32_t *__restrict__ GA,
                            const uint32_t *__restrict__ gop, const uint32_t *__restrict__ gip,
                            const double *__restrict__ gnw, uint32_t *__restrict__ op, uint32_t *__restrict__ ip,
                            double *__restrict__ nw, uint32_t *__restrict__ inst) {
  for (int32_t k = blockIdx.y; k < K; k += gridDim.y) {
    const int32_t g = k / R;
    const uint32_t n0 = NO[k], nk = NO[k + 1] - n0, g0 = GO[g];
    const uint32_t base = AO[k] - GA[g];   // arc offset of the copy (mod 2^32)
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nk; i += gridDim.x * blockDim.x) {
      op[n0 + i] = base + gop[g0 + i];
      ip[n0 + i] = base + gip[g0 + i];
      nw[n0 + i] = gnw[g0 + i];
      inst[n0 + i] = (uint32_t)k;
    }
  }
}

__global__ void k_rep_arcs(int32_t K, int32_t R, const uint32_t *__restrict__ NO, const uint32_t *__restrict__ AO,
                           const uint32_t *__restrict__ GO, const uint32_t *__restrict__ GA,
                           const uint32_t *__restrict__ goi, const uint32_t *__restrict__ gii,
                           const double *__restrict__ gw, const double *__restrict__ gwi, uint32_t *__restrict__ oi,
                           uint32_t *__restrict__ ii, double *__restrict__ w, double *__restrict__ wi) {
  for (int32_t k = blockIdx.y; k < K; k += gridDim.y) {
    const int32_t g = k / R;
    const uint32_t a0 = AO[k], mk = AO[k + 1] - a0, ga = GA[g];
    const uint32_t shift = NO[k] - GO[g];   // node offset of the copy (mod 2^32)
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < mk; j += gridDim.x * blockDim.x) {
      oi[a0 + j] = goi[ga + j] + shift;
      ii[a0 + j] = gii[ga + j] + shift;
      w[a0 + j] = gw[ga + j];
      wi[a0 + j] = gwi[ga + j];
    }
  }
}

__global__ void k_rep_end(uint32_t N, uint32_t M, uint32_t *__restrict__ op, uint32_t *__restrict__ ip) {
  op[N] = M;
  ip[N] = M;
}

}  // namespace

void upload_graph(Ctx &c, int32_t n, int64_t m, const int64_t *out_ptr, const int32_t *out_idx,
                  const double *w, const double *node_w) {
  upload_csr(c, n, m, out_ptr, out_idx, w, node_w);
  build_buckets(c, c.L0);
  c.initiated = false;
  ++c.graph_version;
}

void upload_graph_replicated(Ctx &c, int32_t G, const int32_t *goff, int32_t R, int32_t n, int64_t m,
                             const int64_t *out_ptr, const int32_t *out_idx, const double *w,
                             const double *node_w) {
  const int64_t K = (int64_t)G * R, N = (int64_t)n * R, M = m * R;
  if (G < 1 || R < 1) throw HipError("no networks or no seeds");
  if (N > 0x7fffffffLL || M > 0x7fffffffLL) throw HipError("the copies exceed 2^31 - 1 nodes or arcs");
  upload_csr(c, n, m, out_ptr, out_idx, w, node_w);   // the networks once, with their in-CSR
  hipStream_t s = c.stream;
  Level &L = c.L0;
  // instance tables: NO, AO (K + 1 each), then GO, GA (G + 1 each)
  std::vector<uint32_t> tab((size_t)(2 * (K + 1) + 2 * (G + 1)));
  uint32_t *NO = tab.data(), *AO = NO + (K + 1), *GO = AO + (K + 1), *GA = GO + (G + 1);
  int64_t nn = 0, mm = 0;
  for (int32_t g = 0; g < G; ++g) {
    GO[g] = (uint32_t)goff[g];
    GA[g] = (uint32_t)out_ptr[goff[g]];
    const int64_t ng = goff[g + 1] - goff[g], mg = out_ptr[goff[g + 1]] - out_ptr[goff[g]];
    for (int32_t r = 0; r < R; ++r) {
      NO[(int64_t)g * R + r] = (uint32_t)nn;
      AO[(int64_t)g * R + r] = (uint32_t)mm;
      nn += ng;
      mm += mg;
    }
  }
  NO[K] = (uint32_t)N;
  AO[K] = (uint32_t)M;
  GO[G] = (uint32_t)n;
  GA[G] = (uint32_t)m;
  DevBuf<uint32_t> dtab(c.cache), op(c.cache), ip(c.cache), oi(c.cache), ii(c.cache), inst(c.cache);
  DevBuf<double> rw(c.cache), rwi(c.cache), nw(c.cache);
  dtab.alloc((int64_t)tab.size());
  RMX_HIP(hipMemcpyAsync(dtab.p, tab.data(), sizeof(uint32_t) * tab.size(), hipMemcpyHostToDevice, s));
  const uint32_t *dNO = dtab.p, *dAO = dNO + (K + 1), *dGO = dAO + (K + 1), *dGA = dGO + (G + 1);
  op.alloc(N + 1);
  ip.alloc(N + 1);
  oi.alloc(M);
  ii.alloc(M);
  rw.alloc(M);
  rwi.alloc(M);
  nw.alloc(N);
  inst.alloc(N);
  const int64_t nmax = [&] {
    int64_t x = 1;
    for (int32_t g = 0; g < G; ++g) x = std::max<int64_t>(x, goff[g + 1] - goff[g]);
    return x;
  }();
  const int64_t mmax = [&] {
    int64_t x = 1;
    for (int32_t g = 0; g < G; ++g) x = std::max<int64_t>(x, out_ptr[goff[g + 1]] - out_ptr[goff[g]]);
    return x;
  }();
  const unsigned gy = (unsigned)std::min<int64_t>(K, 65535);
  const dim3 gn((unsigned)std::min<int64_t>((nmax + 255) / 256, 64), gy);
  k_rep_nodes<<<gn, 256, 0, s>>>((int32_t)K, R, dNO, dAO, dGO, dGA, L.out_ptr.p, L.in_ptr.p, c.node_w.p, op.p, ip.p,
                                 nw.p, inst.p);
  RMX_LAUNCH_CHECK();
  k_rep_end<<<1, 1, 0, s>>>((uint32_t)N, (uint32_t)M, op.p, ip.p);
  RMX_LAUNCH_CHECK();
  if (M > 0) {
    const dim3 ga((unsigned)std::min<int64_t>((mmax + 255) / 256, 256), gy);
    k_rep_arcs<<<ga, 256, 0, s>>>((int32_t)K, R, dNO, dAO, dGO, dGA, L.out_idx.p, L.in_idx.p, c.raw_w.p,
                                  c.raw_in.p, oi.p, ii.p, rw.p, rwi.p);
    RMX_LAUNCH_CHECK();
  }
  L.out_ptr.swap(op);
  L.in_ptr.swap(ip);
  L.out_idx.swap(oi);
  L.in_idx.swap(ii);
  c.raw_w.swap(rw);
  c.raw_in.swap(rwi);
  c.node_w.swap(nw);
  L.uinst.swap(inst);
  L.n = (int32_t)N;
  L.m = M;
  c.N = (int32_t)N;
  c.E = M;
  c.sync();   // the networks' buffers and the host tables go out of scope
  build_buckets(c, L);
  c.initiated = false;
  ++c.graph_version;
}

namespace {

// raw_in (with raw): the raw weights of the in-arcs, in in-CSR order
void transpose(Ctx &c, Level &L, const double *raw, DevBuf<double> *raw_in) {
  hipStream_t s = c.stream;
  const int64_t m = L.m;
  const int32_t n = L.n;
  L.in_ptr.alloc(n + 1);
  L.in_idx.alloc(m);
  if (raw_in) raw_in->alloc(m);
  if (m == 0) {
    fill_u32(L.in_ptr.p, 0u, n + 1, s);
    c.sync();
    return;
  }
  {
    DevBuf<uint32_t> src(c.cache), keys(c.cache), vals(c.cache);
    src.alloc(m);
    keys.alloc(m);
    vals.alloc(m);
    k_arc_src<<<grid_for((int64_t)n * kWave, 256), 256, 0, s>>>(L.out_ptr.p, n, src.p);
    RMX_LAUNCH_CHECK();
    RMX_HIP(hipMemcpyAsync(keys.p, L.out_idx.p, sizeof(uint32_t) * m, hipMemcpyDeviceToDevice, s));
    iota_u32(vals.p, m, s);
    int bits = 1;
    while (bits < 32 && ((int64_t)1 << bits) < (int64_t)n) ++bits;
    radix_sort_pairs_u32(keys.p, vals.p, m, 0, bits, c.scratch, s);
    k_in_from_sorted<<<gs_grid(m), 256, 0, s>>>(vals.p, src.p, raw, m, L.in_idx.p, raw_in ? raw_in->p : nullptr);
    RMX_LAUNCH_CHECK();
    k_ptr_from_sorted<<<gs_grid(m + 1), 256, 0, s>>>(keys.p, m, n, L.in_ptr.p);
    RMX_LAUNCH_CHECK();
    c.sync();
  }
}

}  // namespace

void build_in_csr(Ctx &c, Level &L, const double *raw, DevBuf<double> *raw_in) { transpose(c, L, raw, raw_in); }

void super_in_weights(Ctx &c, const Level &L, double *in_w) {
  if (L.m == 0) return;
  k_in_w_lookup<<<grid_for(L.n, 256), 256, 0, c.stream>>>(L.in_ptr.p, L.in_idx.p, L.n, L.out_ptr.p, L.out_idx.p,
                                                          L.out_w.p, in_w);
  RMX_LAUNCH_CHECK();
}

void build_buckets(Ctx &c, Level &L) {
  hipStream_t s = c.stream;
  const int32_t n = L.n;
  DevBuf<uint32_t> keys(c.cache), cnt(c.cache);
  keys.alloc(n);
  cnt.alloc(16 + 1);
  fill_u32(cnt.p, 0u, 17, s);
  L.bucket.alloc(n);
  k_bucket_key<<<std::min<unsigned>(grid_for(n, 256), 2048u), 256, 0, s>>>(L.out_ptr.p, L.in_ptr.p, n, c.cuts,
                                                                          keys.p, cnt.p + 16);
  RMX_LAUNCH_CHECK();
  k_small_hist<<<gs_grid(n), 256, 0, s>>>(keys.p, n, cnt.p);
  RMX_LAUNCH_CHECK();
  iota_u32(L.bucket.p, n, s);
  radix_sort_pairs_u32(keys.p, L.bucket.p, n, 0, 8, c.scratch, s);
  uint32_t h[17];
  RMX_HIP(hipMemcpyAsync(h, cnt.p, sizeof(h), hipMemcpyDeviceToHost, s));
  c.sync();
  L.bucket_off[0] = 0;
  for (int b = 0; b < kNumBuckets; ++b) L.bucket_off[b + 1] = L.bucket_off[b] + (int32_t)h[b];
  L.max_deg = (int32_t)h[16];
  // hash-table regions of the hubs
  const int32_t nh = L.bucket_off[kNumBuckets] - L.bucket_off[kNumBuckets - 1];
  L.hub_off.alloc(nh + 1);
  L.hub_pool = 0;
  L.hub_arcs = 0;
  L.hub_parts = 0;
  if (nh > 0) {
    DevBuf<uint32_t> deg(c.cache);
    deg.alloc(nh);
    k_list_degree<<<grid_for(nh, 256), 256, 0, s>>>(L.bucket.p + L.bucket_off[kNumBuckets - 1], nh, L.out_ptr.p,
                                                    L.in_ptr.p, deg.p);
    RMX_LAUNCH_CHECK();
    std::vector<uint32_t> hd(nh);
    RMX_HIP(hipMemcpyAsync(hd.data(), deg.p, sizeof(uint32_t) * nh, hipMemcpyDeviceToHost, s));
    c.sync();
    std::vector<long long> off(nh + 1, 0);
    for (int32_t i = 0; i < nh; ++i) {
      long long cap = 1;
      while (cap < 2LL * hd[i] + 2) cap <<= 1;
      off[i + 1] = off[i] + cap;
      L.hub_arcs += hd[i];
      L.hub_parts += ((int64_t)hd[i] + kHubPart - 1) / kHubPart;
    }
    L.hub_pool = off[nh];
    RMX_HIP(hipMemcpyAsync(L.hub_off.p, off.data(), sizeof(long long) * (nh + 1), hipMemcpyHostToDevice, s));
    c.sync();
  }
}

}  // namespace rmx
