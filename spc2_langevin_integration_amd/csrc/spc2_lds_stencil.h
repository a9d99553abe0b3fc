// LDS staging of the 3 x 3 cell stencil and the pass-1 / pass-2 bodies of one
// pair workgroup, shared by the single-system kernels (spc2_pairs.hip) and the
// batched kernels (spc2_batch.hip).
//
// Work decomposition: one 256-lane workgroup per row chunk (a run of
// consecutive slots of one cell row, which spans W cells) or, in sparse
// systems, per row segment of W cells.  It stages the 3 x (W + 2) halo cells'
// tiles in LDS with coalesced reads (cells of a row are contiguous in the
// cell-sorted arrays), then each lane takes one centre particle.  In LDS the
// three stencil cells of one stencil row are adjacent, so a particle visits its
// 9 cells as 3 contiguous ranges, in the reference order (cells (-1,dy),
// (0,dy), (1,dy), ascending particle index inside a cell).  A workgroup whose halo does not fit the LDS capacity
// falls back to reading the same tiles from global memory with the same
// arithmetic and order.  Compiled with -ffp-contract=off.
#pragma once

#include <type_traits>

#include "spc2_pair_accumulators.h"

namespace spc2 {
namespace {

constexpr int kMaxSegW = 64;  // 3 (kMaxSegW + 2) halo cells, one per lane of the setup scan
constexpr int kMaxSlots = 3 * (kMaxSegW + 2);
constexpr double kChunkPerCell = 8.0;  // mean particles per cell from which rows are cut into chunks
// x offset that puts a candidate beyond every cutoff (r^2 about 1e12)
constexpr float kSelfOff = 1.0e6f;


// =============================================================================
// Row-segment setup and LDS staging
// =============================================================================

__device__ __forceinline__ int block_inclusive_scan(int v, int *wtot) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    int u = __shfl_up(v, d, 64);
    if (lane >= d) v += u;
  }
  if (lane == 63) wtot[w] = v;
  __syncthreads();
  int add = 0;
#pragma unroll
  for (int q = 0; q < kPairBlock / 64; ++q) add += (q < w) ? wtot[q] : 0;
  __syncthreads();
  return v + add;
}

struct SegShared {
  int off[kMaxSlots + 1];  // LDS offset of halo cell (r, q), flattened r * (W + 2) + q; off[3(W+2)] = total
  int gst[kMaxSlots];      // its first global slot
  int wtot[kPairBlock / 64];
};

struct Seg {
  int cy, cx0, W, ncol, ns, total;
};

// Stages the halo table of centre columns cx0 .. cx0 + W - 1 of row cy (W <=
// kMaxSegW): the 3 x (W + 2) cells' first slots and LDS offsets.
__device__ __forceinline__ Seg halo_setup(const KParams &P, const int *start, int cy, int cx0, int W, SegShared &sh) {
  Seg g;
  g.cy = cy;
  g.cx0 = cx0;
  g.W = W;
  g.ncol = g.W + 2;
  g.ns = 3 * g.ncol;
  int cnt = 0;
  if ((int)threadIdx.x < g.ns) {
    int r = threadIdx.x / g.ncol, q = threadIdx.x - r * g.ncol;
    int cyr = adjust_cell(g.cy + r - 1, P.ny);
    int cxq = g.cx0 - 1 + q;
    if (cxq < 0) cxq += P.nx;
    else if (cxq >= P.nx) cxq -= P.nx;
    int cell = cyr * P.nx + cxq;
    int g0 = start[cell];
    cnt = start[cell + 1] - g0;
    sh.gst[threadIdx.x] = g0;
  }
  int incl = block_inclusive_scan(cnt, sh.wtot);
  if ((int)threadIdx.x < g.ns) sh.off[threadIdx.x] = incl - cnt;
  if ((int)threadIdx.x == g.ns - 1) sh.off[g.ns] = incl;
  __syncthreads();
  g.total = sh.off[g.ns];
  return g;
}

// The centre slots [k0, k1) of pair workgroup `seg`, their row cy and the
// columns cx0 .. cx0 + W - 1 they occupy.  Row chunks (P.chunk > 0): slots
// j * chunk .. of the row, the row's last workgroup (j = segs - 1) taking the
// rest; false if the chunk starts past the row.  Row segments: seg_w cells.
struct Block {
  int k0, k1, cy, cx0, W;
};
__device__ __forceinline__ bool block_range(const KParams &P, const int *start, const Tiles &t, int seg, Block &b) {
  b.cy = seg / P.segs;
  const int j = seg - b.cy * P.segs, row = b.cy * P.nx;
  if (P.chunk == 0) {
    b.cx0 = j * P.seg_w;
    b.W = min(P.seg_w, P.nx - b.cx0);
    b.k0 = start[row + b.cx0];
    b.k1 = start[row + b.cx0 + b.W];
    return true;
  }
  const int re = start[row + P.nx];
  b.k0 = start[row] + j * P.chunk;
  if (b.k0 >= re) return false;
  b.k1 = j == P.segs - 1 ? re : min(re, b.k0 + P.chunk);
  b.cx0 = t.cell[b.k0] - row;
  b.W = t.cell[b.k1 - 1] - row - b.cx0 + 1;
  return true;
}

// largest halo slot s with off[s] <= e
__device__ __forceinline__ int find_slot(const SegShared &sh, int ns, int e) {
  int lo = 0, hi = ns - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (sh.off[mid] <= e) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// LDS element arrays of one precision and pass, carved from dynamic LDS.
template <int PREC, int PASS>
struct LdsTile;

template <int PASS>
struct LdsTile<P32K, PASS> {
  float4 *f4;
  int *slot;
  float *aj;
  static constexpr int bytes = 20 + (PASS == 2 ? 4 : 0);
  __device__ __forceinline__ LdsTile(void *base, int cap) {
    f4 = (float4 *)base;
    slot = (int *)(f4 + cap);
    aj = (float *)(slot + cap);
  }
  __device__ __forceinline__ void load(int e, int g, const Tiles &t, const EpiCtx &ec) {
    f4[e] = t.f4[g];
    slot[e] = g;
    if (PASS == 2) aj[e] = ec.aphi_slot_f[g];
  }
  __device__ __forceinline__ Src32 src() const { return Src32{f4, slot, aj}; }
};

template <int PREC, int PASS>
struct LdsTile {  // P64 / P64F
  double *x, *y, *a, *b, *aj;
  static constexpr int ncomp = (PREC == P64 ? 3 : 4) + (PASS == 2 ? 1 : 0);
  static constexpr int bytes = 8 * ncomp;
  __device__ __forceinline__ LdsTile(void *base, int cap) {
    x = (double *)base;
    y = x + cap;
    a = y + cap;
    b = a + cap;
    aj = (PREC == P64 ? a : b) + cap;
  }
  __device__ __forceinline__ void load(int e, int g, const Tiles &t, const EpiCtx &ec) {
    x[e] = t.x[g];
    y[e] = t.y[g];
    if (PREC == P64) {
      a[e] = t.p[g];
    } else {
      a[e] = t.c[g];
      b[e] = t.s[g];
    }
    if (PASS == 2) aj[e] = ec.aphi_slot[g];
  }
  __device__ __forceinline__ Src64 src() const {
    return PREC == P64 ? Src64{x, y, a, nullptr, nullptr, aj} : Src64{x, y, nullptr, a, b, aj};
  }
};

// global-memory source over the cell-sorted tiles (overflow fallback)
template <int PREC>
__device__ __forceinline__ typename SrcOf<PREC>::type global_src(const Tiles &t, const EpiCtx &e, int pass) {
  if constexpr (PREC == P32K) {
    return Src32{t.f4, nullptr, pass == 2 ? e.aphi_slot_f : nullptr};
  } else {
    return Src64{t.x, t.y, t.p, t.c, t.s, pass == 2 ? e.aphi_slot : nullptr};
  }
}

// Accumulators with a two-candidate step for the FP32 cell walks.
template <class Acc>
struct HasPair2 : std::false_type {};
template <>
struct HasPair2<ChimAcc<P32K>> : std::true_type {};
template <>
struct HasPair2<Chim2Acc<P32K>> : std::true_type {};
template <>
struct HasPair2<VortAcc<P32K>> : std::true_type {};

// The FP32 walk of one stencil cell, LDS indices [a, b), column offset ox, in
// FP32 partials added to the compensated sums after the cell.  SELF: the cell
// holds the particle itself (LDS index self), which is moved kSelfOff away,
// beyond every cutoff, so it adds the +0 terms of any candidate outside rho
// (partials start at +0 and never become -0, so the sums equal those that skip
// it).  One loop over the cell, not two split at the particle: a wavefront's
// lanes split at different places, so two loops would run the cell twice.
template <bool SELF, class Acc>
__device__ __forceinline__ void cell_run(Acc &acc, const Src32 &src, int a, int b, float ox, int self,
                                         const KParams &P) {
  acc.cell_begin();
  if constexpr (HasPair2<Acc>::value) {
    // the pairs walked by tile address, so that the loop keeps one induction
    // variable, each pair's tiles requested a pair ahead; then the odd last
    // candidate.  The read ahead of the last pair stays within the workgroup's
    // LDS allocation (at most two tiles past the staged halo, in the arrays
    // that follow it) and is not used.
    const float4 *t2 = src.f4 + a, *const last = src.f4 + (b - 1), *const own = src.f4 + self;
    if (t2 < last) {
      // the pair at t2 from the tiles ta, tb; whether another pair follows
      auto step = [&](const float4 &ta, const float4 &tb, int m) {
        const float oxa = SELF && t2 == own ? kSelfOff : ox, oxb = SELF && t2 + 1 == own ? kSelfOff : ox;
        if constexpr (std::is_same<Acc, Chim2Acc<P32K>>::value)  // reads aj[m]: the index too
          acc.pair2_x(src, t2, m, ta, tb, oxa, oxb, P);
        else
          acc.pair2_x(src, t2, ta, tb, oxa, oxb, P);
        t2 += 2;
        return t2 < last;
      };
      // two sets of tile registers taken in turn: each pair's tiles are read
      // into the set the previous pair has just used, with no copies
      float4 a0 = t2[0], b0 = t2[1];
      for (int m = a;; m += 4) {
        const float4 a1 = t2[2], b1 = t2[3];
        if (!step(a0, b0, m)) break;
        a0 = t2[2];
        b0 = t2[3];
        if (!step(a1, b1, m + 2)) break;
      }
    }
    if ((b - a) & 1) acc.template pair_x<SUM_GROUP>(src, b - 1, SELF && b - 1 == self ? kSelfOff : ox, P);
    acc.flush_cell();
  } else {
#pragma unroll 2
    for (int m = a; m < b; ++m) acc.template pair_x<SUM_GROUP>(src, m, SELF && m == self ? kSelfOff : ox, P);
    acc.flush();
  }
}

// Visits the candidates of the centre particle at LDS index ei (halo column qi)
// in reference order; returns the number of candidates.
template <class Acc, class S>
__device__ __forceinline__ int visit_lds(Acc &acc, const S &src, const SegShared &sh, int ncol, int qi, int ei,
                                         const KParams &P) {
  int cand = 0;
  if constexpr (std::is_same<S, Src32>::value) {
    // FP32 form: one loop per stencil cell, its column offset computed once
    // (the value the per-candidate selection gives inside that cell), its
    // terms summed in FP32 partials that are added to the compensated sums
    // after the cell
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      const int base = r * ncol + qi;
      const int lo = sh.off[base - 1], b1 = sh.off[base], b2 = sh.off[base + 1], hi = sh.off[base + 2];
      acc.row(r - 1, b1, b2, P);
      cand += hi - lo;
      const float ol = acc.co.col(-1), oc = acc.co.col(0), orr = acc.co.col(1);
      cell_run<false>(acc, src, lo, b1, ol, ei, P);
      if (r == 1) cell_run<true>(acc, src, b1, b2, oc, ei, P);
      else cell_run<false>(acc, src, b1, b2, oc, ei, P);
      cell_run<false>(acc, src, b2, hi, orr, ei, P);
    }
    return cand - 1;
  }
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    const int base = r * ncol + qi;
    const int lo = sh.off[base - 1], b1 = sh.off[base], b2 = sh.off[base + 1], hi = sh.off[base + 2];
    acc.row(r - 1, b1, b2, P);
    cand += hi - lo;
    if (r == 1) {
#pragma unroll 2
      for (int m = lo; m < ei; ++m) acc.pair(src, m, P);
#pragma unroll 2
      for (int m = ei + 1; m < hi; ++m) acc.pair(src, m, P);
    } else {
#pragma unroll 2
      for (int m = lo; m < hi; ++m) acc.pair(src, m, P);
    }
  }
  return cand - 1;
}

// Same visit over the global tiles, cell by cell (slot k in cell c).
template <class Acc, class S>
__device__ __forceinline__ int visit_global(Acc &acc, const S &src, const int *start, int c, int k, const KParams &P) {
  const int cx = c % P.nx, cy = c / P.nx;
  int cand = 0;
  for (int dyo = -1; dyo <= 1; ++dyo) {
    const int row = adjust_cell(cy + dyo, P.ny) * P.nx;
    for (int dxo = -1; dxo <= 1; ++dxo) {
      const int jc = row + adjust_cell(cx + dxo, P.nx);
      const int m0 = start[jc], m1 = start[jc + 1];
      // column offsets from the boundary trick: all of [m0, m1) is one column
      if (dxo < 0) acc.row(dyo, m1, m1, P);
      else if (dxo == 0) acc.row(dyo, m0, m1, P);
      else acc.row(dyo, m0, m0, P);
      cand += m1 - m0;
      if constexpr (std::is_same<S, Src32>::value) {
        // per cell FP32 partials, as visit_lds sums them
        acc.cell_begin();
        for (int m = m0; m < m1; ++m)
          if (m != k) acc.template pair<SUM_GROUP>(src, m, P);
        acc.flush();
      } else {
        for (int m = m0; m < m1; ++m)
          if (m != k) acc.pair(src, m, P);
      }
    }
  }
  return cand - 1;
}

// =============================================================================
// Pass-1 and pass-2 kernels over linked cells
// =============================================================================

template <class Acc, int PREC, int EPI>
__device__ __forceinline__ void pass1_segment(const KParams &P, const int *__restrict__ start, const Tiles &t,
                                              const int *__restrict__ slot_h, const int *__restrict__ slot_id,
                                              const double *gp, const EpiCtx &e, unsigned long long *counts, int seg,
                                              void *dyn_lds) {
  __shared__ SegShared sh;
  Block b;
  if (!block_range(P, start, t, seg, b)) return;  // the whole workgroup
  const bool fits = b.W <= kMaxSegW;
  Seg g{};
  if (fits) g = halo_setup(P, start, b.cy, b.cx0, b.W, sh);
  const bool staged = fits && g.total <= P.lds_cap;
  LdsTile<PREC, 1> lt(dyn_lds, P.lds_cap);
  if (staged) {
    for (int el = threadIdx.x; el < g.total; el += kPairBlock) {
      int s = find_slot(sh, g.ns, el);
      lt.load(el, sh.gst[s] + (el - sh.off[s]), t, e);
    }
  }
  __syncthreads();
  // the centre slots are contiguous from the first centre cell's, k0 its slot
  // (k0 - gst) of the cell
  const int e0 = staged ? sh.off[g.ncol + 1] + (b.k0 - sh.gst[g.ncol + 1]) : 0;
  const int ncen = b.k1 - b.k0;
  const int k0 = b.k0;
  long long cand_sum = 0, n_sum = 0, n2_sum = 0;
  for (int p = threadIdx.x; p < ncen; p += kPairBlock) {
    const int k = k0 + p, ei = e0 + p;
    Acc acc;
    int cand;
    if (staged) {
      int lo = 1, hi = g.W;  // centre column of ei
      while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (sh.off[g.ncol + mid] <= ei) lo = mid;
        else hi = mid - 1;
      }
      acc.init(lt.src(), ei, k, t, P);
      cand = visit_lds(acc, lt.src(), sh, g.ncol, lo, ei, P);
    } else {
      auto gs = global_src<PREC>(t, e, 1);
      acc.init(gs, k, k, t, P);
      cand = visit_global(acc, gs, start, t.cell[k], k, P);
    }
    const int h = slot_h[k], id = slot_id[k];
    Pass1Out o{};
    acc.finish(P, PREC == P32K ? gp[h] : 0.0, o);
    epilogue1<EPI>(P, e, h, id, k, o);
    cand_sum += cand;
    n_sum += o.n;
    n2_sum += o.n2;
  }
  if (counts) count_add(counts, cand_sum, n_sum, n2_sum);
}

template <class Acc, int PREC, int EPI>
__device__ __forceinline__ void pass2_segment(const KParams &P, const int *__restrict__ start, const Tiles &t,
                                              const int *__restrict__ slot_h, const int *__restrict__ slot_id,
                                              const double *gp, const EpiCtx &e, int seg, void *dyn_lds) {
  __shared__ SegShared sh;
  Block b;
  if (!block_range(P, start, t, seg, b)) return;
  const bool fits = b.W <= kMaxSegW;
  Seg g{};
  if (fits) g = halo_setup(P, start, b.cy, b.cx0, b.W, sh);
  const bool staged = fits && g.total <= P.lds_cap;
  LdsTile<PREC, 2> lt(dyn_lds, P.lds_cap);
  if (staged) {
    for (int el = threadIdx.x; el < g.total; el += kPairBlock) {
      int s = find_slot(sh, g.ns, el);
      lt.load(el, sh.gst[s] + (el - sh.off[s]), t, e);
    }
  }
  __syncthreads();
  const int e0 = staged ? sh.off[g.ncol + 1] + (b.k0 - sh.gst[g.ncol + 1]) : 0;
  const int ncen = b.k1 - b.k0;
  const int k0 = b.k0;
  for (int p = threadIdx.x; p < ncen; p += kPairBlock) {
    const int k = k0 + p, ei = e0 + p;
    const int h = slot_h[k], id = slot_id[k];
    Pass2In in{e.aphi_slot[k], e.n_slot[k], e.n2_slot[k], 0.0, 0.0};
    if (EPI == EPI_ITO2) {
      in.in0 = e.a0[h];
      in.in1 = e.a1[h];
    }
    Acc acc;
    if (staged) {
      int lo = 1, hi = g.W;
      while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (sh.off[g.ncol + mid] <= ei) lo = mid;
        else hi = mid - 1;
      }
      acc.init(lt.src(), ei, k, t, P, in);
      visit_lds(acc, lt.src(), sh, g.ncol, lo, ei, P);
    } else {
      auto gs = global_src<PREC>(t, e, 2);
      acc.init(gs, k, k, t, P, in);
      visit_global(acc, gs, start, t.cell[k], k, P);
    }
    Pass2Out o;
    acc.finish(P, PREC == P32K ? gp[h] : 0.0, o);
    epilogue2<EPI>(P, e, h, id, o);
  }
}

}  // namespace
}  // namespace spc2
