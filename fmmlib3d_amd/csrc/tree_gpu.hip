// Octree construction with the points in HBM.  The bounding box, the octant
// digit codes, their per-box histograms and the stable reordering of the
// points run on the device; the box logic (which sub-boxes exist and split,
// their breadth-first numbering) and the interaction lists run on the host
// from the histograms with the same code as build_octree, so the tree is the
// one build_octree makes, box for box (src/d3tstrcr.f:24-259), and the
// sorted-to-original permutations stay on the device.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <rocprim/device/device_radix_sort.hpp>

#include "tree_gpu.hpp"

namespace fmm3d {

namespace {

constexpr int kBBoxBlocks = 1024;

__global__ void k_bbox_partial(const double* __restrict__ p, int n, double* __restrict__ part) {
  __shared__ double red[6][256];
  double v[6] = {1e300, 1e300, 1e300, -1e300, -1e300, -1e300};
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    for (int d = 0; d < 3; d++) {
      const double x = p[3 * (size_t)i + d];
      v[d] = fmin(v[d], x);
      v[3 + d] = fmax(v[3 + d], x);
    }
  for (int d = 0; d < 6; d++) red[d][threadIdx.x] = v[d];
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int d = 0; d < 6; d++)
        red[d][threadIdx.x] = d < 3 ? fmin(red[d][threadIdx.x], red[d][threadIdx.x + w])
                                    : fmax(red[d][threadIdx.x], red[d][threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x < 6) part[6 * blockIdx.x + threadIdx.x] = red[threadIdx.x][0];
}

__global__ void k_iota(int* __restrict__ a, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) a[i] = i;
}

// frontier box of compact position t: fcum[f] <= t < fcum[f+1]
__device__ __forceinline__ int seg_of(const int* fcum, int nf, int t) {
  int lo = 0, hi = nf;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (fcum[mid] <= t) lo = mid;
    else hi = mid;
  }
  return lo;
}

// D octant digits of every point of the frontier boxes (strict ">" against
// the host-computed centres) and the per-box histogram of the codes
__global__ void k_codes(int nf, const int* __restrict__ fseg, const int* __restrict__ fcum, int total,
                        const double* __restrict__ pts, const int* __restrict__ idx, int D, int CT,
                        const double* __restrict__ ctab, int NB, uint32_t* __restrict__ code,
                        int* __restrict__ hist) {
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < total; t += gridDim.x * blockDim.x) {
    const int f = seg_of(fcum, nf, t);
    const int i = idx[fseg[f] + (t - fcum[f])];
    const double x = pts[3 * (size_t)i], y = pts[3 * (size_t)i + 1], z = pts[3 * (size_t)i + 2];
    const double* cx = ctab + (((size_t)f * 3 + 0) * D) * CT;
    const double* cy = ctab + (((size_t)f * 3 + 1) * D) * CT;
    const double* cz = ctab + (((size_t)f * 3 + 2) * D) * CT;
    uint32_t c = 0;
    int ii = 0, jj = 0, kk = 0;
    for (int d = 0; d < D; d++) {
      const int o = (x > cx[d * CT + ii]) | ((y > cy[d * CT + jj]) << 1) | ((z > cz[d * CT + kk]) << 2);
      c = (c << 3) | (uint32_t)o;
      ii = 2 * ii + (o & 1);
      jj = 2 * jj + ((o >> 1) & 1);
      kk = 2 * kk + ((o >> 2) & 1);
    }
    code[t] = c;
    atomicAdd(hist + (size_t)f * (NB + 1) + c + 1, 1);
  }
}

struct EffArgs {
  int Dk, Dfull, NS, NBk;
  int posd[8];
};

// sort key: (frontier box, code truncated where the point's box stops splitting)
__global__ void k_keys(int nf, const int* __restrict__ fseg, const int* __restrict__ fcum, int total,
                       const uint32_t* __restrict__ code, const char* __restrict__ splits, EffArgs ea,
                       const int* __restrict__ idx, uint32_t* __restrict__ key, int* __restrict__ val) {
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < total; t += gridDim.x * blockDim.x) {
    const int f = seg_of(fcum, nf, t);
    const uint32_t c = code[t] >> (3 * (ea.Dfull - ea.Dk));
    const char* sp = splits + (size_t)f * ea.NS;
    uint32_t e = c;
    for (int d = 1; d < ea.Dk; d++) {
      const uint32_t pre = c >> (3 * (ea.Dk - d));
      if (!sp[ea.posd[d] + pre]) { e = pre << (3 * (ea.Dk - d)); break; }
    }
    key[t] = (uint32_t)f * (uint32_t)ea.NBk + e;
    val[t] = idx[fseg[f] + (t - fcum[f])];
  }
}

__global__ void k_put_idx(int nf, const int* __restrict__ fseg, const int* __restrict__ fcum, int total,
                          const int* __restrict__ sval, int* __restrict__ idx) {
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < total; t += gridDim.x * blockDim.x) {
    const int f = seg_of(fcum, nf, t);
    idx[fseg[f] + (t - fcum[f])] = sval[t];
  }
}

unsigned grid_for(size_t n) { return (unsigned)std::max<size_t>(1, std::min<size_t>((n + 255) / 256, 8192)); }

// one point set (sources or targets) on the device
struct PointSet {
  const char* tag;
  const double* pts;
  int n;
  int* idx;
};

}  // namespace

bool device_tree_enabled() {
  const char* v = std::getenv("FMM3D_TREE");
  return !(v && std::strcmp(v, "host") == 0);
}

int build_octree_device(Context& ctx, hipStream_t s, const double* src, int ns, const double* trg, int nt,
                        int nbox, Octree& t, const int** sperm, const int** tperm, bool lists) {
  t.reset();
  if (ns < 1) return 128;
  if (nt < 0) nt = 0;
  Arena& A = ctx.arena;
  // ---- bounding box: per-block partial min/max, final reduction on the host
  double* part = (double*)A.get("tree_bbox", sizeof(double) * 6 * kBBoxBlocks * 2);
  double* hpart = (double*)A.get_host("tree_bbox_h", sizeof(double) * 6 * kBBoxBlocks * 2);
  const unsigned gs = std::min<unsigned>(kBBoxBlocks, grid_for(ns)), gt = nt > 0 ? std::min<unsigned>(kBBoxBlocks, grid_for(nt)) : 0;
  hipLaunchKernelGGL(k_bbox_partial, dim3(gs), dim3(256), 0, s, src, ns, part);
  if (nt > 0) hipLaunchKernelGGL(k_bbox_partial, dim3(gt), dim3(256), 0, s, trg, nt, part + 6 * kBBoxBlocks);
  FMM_HIP(hipMemcpyAsync(hpart, part, sizeof(double) * 6 * kBBoxBlocks * 2, hipMemcpyDeviceToHost, s));
  FMM_HIP(hipStreamSynchronize(s));
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  auto fold = [&](const double* pp, unsigned nblk) {
    for (unsigned b = 0; b < nblk; b++)
      for (int d = 0; d < 3; d++) {
        lo[d] = std::min(lo[d], pp[6 * b + d]);
        hi[d] = std::max(hi[d], pp[6 * b + 3 + d]);
      }
  };
  fold(hpart, gs);
  if (nt > 0) fold(hpart + 6 * kBBoxBlocks, gt);
  std::vector<int> frontier, next;
  tree_root(t, lo, hi, ns, nt, nbox, frontier);

  PointSet P[2] = {{"s", src, ns, (int*)A.get("tree_sidx", sizeof(int) * (size_t)ns)},
                   {"t", trg, nt, nt > 0 ? (int*)A.get("tree_tidx", sizeof(int) * (size_t)nt) : nullptr}};
  for (auto& q : P)
    if (q.n > 0) hipLaunchKernelGGL(k_iota, dim3(grid_for(q.n)), dim3(256), 0, s, q.idx, q.n);
  const int nmax = std::max(ns, nt);
  uint32_t* scode = (uint32_t*)A.get("tree_scode", sizeof(uint32_t) * (size_t)ns);
  uint32_t* tcode = nt > 0 ? (uint32_t*)A.get("tree_tcode", sizeof(uint32_t) * (size_t)nt) : nullptr;
  uint32_t* key0 = (uint32_t*)A.get("tree_key0", sizeof(uint32_t) * (size_t)nmax);
  uint32_t* key1 = (uint32_t*)A.get("tree_key1", sizeof(uint32_t) * (size_t)nmax);
  int* val0 = (int*)A.get("tree_val0", sizeof(int) * (size_t)nmax);
  int* val1 = (int*)A.get("tree_val1", sizeof(int) * (size_t)nmax);

  TreePass ps;
  int L = 0;
  while (!frontier.empty() && L < kTreeMaxLevel) {
    // the previous pass's sort has finished before its buffers are reused or regrown
    if (L > 0) FMM_HIP(hipStreamSynchronize(s));
    ps.frontier = frontier;
    tree_pass_setup(t, L, ps);
    const int nf = (int)frontier.size(), NB = ps.NB;
    // frontier segments (sorted positions) and their compact offsets
    std::vector<int> seg(4 * (size_t)(nf + 1), 0);  // sseg | scum | tseg | tcum
    int* sseg = seg.data();
    int* scum = sseg + (nf + 1);
    int* tseg = scum + (nf + 1);
    int* tcum = tseg + (nf + 1);
    for (int f = 0; f < nf; f++) {
      const int b = frontier[f];
      sseg[f] = t.isrc[b];
      scum[f + 1] = scum[f] + t.nsrc[b];
      tseg[f] = t.itrg[b];
      tcum[f + 1] = tcum[f] + t.ntrg[b];
    }
    const size_t hsz = (size_t)nf * (NB + 1);
    int* hseg = (int*)A.get_host("tree_seg_h", sizeof(int) * seg.size());
    double* hctab = (double*)A.get_host("tree_ctab_h", sizeof(double) * ps.ctab.size());
    std::memcpy(hseg, seg.data(), sizeof(int) * seg.size());
    std::memcpy(hctab, ps.ctab.data(), sizeof(double) * ps.ctab.size());
    int* dseg = (int*)A.get("tree_seg", sizeof(int) * seg.size());
    double* dctab = (double*)A.get("tree_ctab", sizeof(double) * ps.ctab.size());
    int* dhist = (int*)A.get("tree_hist", sizeof(int) * 2 * hsz);
    int* hhist = (int*)A.get_host("tree_hist_h", sizeof(int) * 2 * hsz);
    FMM_HIP(hipMemcpyAsync(dseg, hseg, sizeof(int) * seg.size(), hipMemcpyHostToDevice, s));
    FMM_HIP(hipMemcpyAsync(dctab, hctab, sizeof(double) * ps.ctab.size(), hipMemcpyHostToDevice, s));
    FMM_HIP(hipMemsetAsync(dhist, 0, sizeof(int) * 2 * hsz, s));
    const int* dsegs[2] = {dseg, dseg + 2 * (nf + 1)};
    const int* dcums[2] = {dseg + (nf + 1), dseg + 3 * (nf + 1)};
    const int totals[2] = {scum[nf], tcum[nf]};
    // the codes are kept per point set between the histogram and the sort
    uint32_t* codes[2] = {scode, tcode};
    for (int k = 0; k < 2; k++)
      if (totals[k] > 0)
        hipLaunchKernelGGL(k_codes, dim3(grid_for(totals[k])), dim3(256), 0, s, nf, dsegs[k], dcums[k], totals[k],
                           P[k].pts, P[k].idx, ps.D, ps.CT, dctab, NB, codes[k], dhist + k * hsz);
    FMM_HIP(hipMemcpyAsync(hhist, dhist, sizeof(int) * 2 * hsz, hipMemcpyDeviceToHost, s));
    FMM_HIP(hipStreamSynchronize(s));
    for (int k = 0; k < 2; k++) {
      int* h = (k ? ps.ht : ps.hs).data();
      const int* src_h = hhist + k * hsz;
      for (int f = 0; f < nf; f++) {
        const int* a = src_h + (size_t)f * (NB + 1);
        int* o = h + (size_t)f * (NB + 1);
        o[0] = 0;
        for (int c = 0; c < NB; c++) o[c + 1] = o[c] + a[c + 1];
      }
    }
    tree_pass_expand(t, nbox, ps, next);
    // stable sort of each frontier box's points by (box, effective code)
    char* hspl = (char*)A.get_host("tree_split_h", ps.splits.size());
    std::memcpy(hspl, ps.splits.data(), ps.splits.size());
    char* dspl = (char*)A.get("tree_split", ps.splits.size());
    FMM_HIP(hipMemcpyAsync(dspl, hspl, ps.splits.size(), hipMemcpyHostToDevice, s));
    EffArgs ea;
    ea.Dk = ps.Dk; ea.Dfull = ps.Dfull; ea.NS = ps.NS; ea.NBk = 1 << (3 * ps.Dk);
    for (int d = 0; d < 8; d++) ea.posd[d] = ps.posd[d];
    unsigned bits = 1;
    while (bits < 32 && ((uint64_t)nf * (uint64_t)ea.NBk) > (1ull << bits)) bits++;
    for (int k = 0; k < 2; k++) {
      const int n = totals[k];
      if (n == 0) continue;
      hipLaunchKernelGGL(k_keys, dim3(grid_for(n)), dim3(256), 0, s, nf, dsegs[k], dcums[k], n, codes[k], dspl, ea,
                         P[k].idx, key0, val0);
      size_t tmp_bytes = 0;
      FMM_HIP(rocprim::radix_sort_pairs(nullptr, tmp_bytes, key0, key1, val0, val1, (size_t)n, 0u, bits, s));
      void* tmp = A.get("tree_sort_tmp", std::max<size_t>(tmp_bytes, 256));
      FMM_HIP(rocprim::radix_sort_pairs(tmp, tmp_bytes, key0, key1, val0, val1, (size_t)n, 0u, bits, s));
      hipLaunchKernelGGL(k_put_idx, dim3(grid_for(n)), dim3(256), 0, s, nf, dsegs[k], dcums[k], n, val1, P[k].idx);
    }
    frontier.swap(next);
    L += ps.Dk;
  }
  t.nboxes = (int)t.level.size();
  FMM_HIP(hipGetLastError());
  *sperm = P[0].idx;
  *tperm = P[1].idx;
  if (lists) tree_lists(t);
  return 0;
}

}  // namespace fmm3d
