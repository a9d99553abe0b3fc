// Periodic 3x3-stencil pair sums on gfx950 (chimera and vortex arrays), the
// fused stepper epilogues, and the all-pairs path.  Compiled with
// -ffp-contract=off (see spc2_device.h).
//
// Work decomposition (linked cells): one 256-lane workgroup per row segment of
// W cells.  It stages the 3 x (W + 2) halo cells' tiles in LDS with coalesced
// reads (cells of a row segment are contiguous in the cell-sorted arrays), then
// each lane takes one centre particle.  In LDS the three stencil cells of one
// stencil row are adjacent, so a particle visits its 9 cells as 3 contiguous
// ranges, in the reference order (cells (-1,dy), (0,dy), (1,dy), ascending
// particle index inside a cell).  A segment whose halo does not fit the LDS
// capacity falls back to reading the same tiles from global memory with the
// same arithmetic and order.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>

#include "spc2_device.h"
#include "spc2_internal.h"
#include "spc2_pair_accumulators.h"

namespace spc2 {

namespace {

constexpr int kBlock = 256;
constexpr int kMaxSegW = 64;
constexpr int kMaxSlots = 3 * (kMaxSegW + 2);

inline int blocks_for(long long n, int b = kBlock) { return (int)((n + b - 1) / b); }

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
  for (int q = 0; q < kBlock / 64; ++q) add += (q < w) ? wtot[q] : 0;
  __syncthreads();
  return v + add;
}

struct SegShared {
  int off[kMaxSlots + 1];  // LDS offset of halo cell (r, q), flattened r * (W + 2) + q; off[3(W+2)] = total
  int gst[kMaxSlots];      // its first global slot
  int wtot[kBlock / 64];
};

struct Seg {
  int cy, cx0, W, ncol, ns, total;
};

__device__ __forceinline__ Seg seg_setup(const KParams &P, const int *start, SegShared &sh) {
  Seg g;
  const int seg = xcd_segment(blockIdx.x, gridDim.x);
  g.cy = seg / P.segs;
  g.cx0 = (seg - g.cy * P.segs) * P.seg_w;
  g.W = min(P.seg_w, P.nx - g.cx0);
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

// Visits the candidates of the centre particle at LDS index ei (halo column qi)
// in reference order; returns the number of candidates.
template <class Acc, class S>
__device__ __forceinline__ int visit_lds(Acc &acc, const S &src, const SegShared &sh, int ncol, int qi, int ei,
                                         const KParams &P) {
  int cand = 0;
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
      for (int m = m0; m < m1; ++m)
        if (m != k) acc.pair(src, m, P);
    }
  }
  return cand - 1;
}

// =============================================================================
// Pass-1 and pass-2 kernels over linked cells
// =============================================================================

template <class Acc, int PREC, int EPI>
__global__ __launch_bounds__(kBlock) void k_pass1_lds(KParams P, const int *__restrict__ start, Tiles t,
                                                      const int *__restrict__ slot_h, const int *__restrict__ slot_id,
                                                      const double *gp, EpiCtx e, unsigned long long *counts) {
  __shared__ SegShared sh;
  extern __shared__ float4 dyn_lds[];
  Seg g = seg_setup(P, start, sh);
  const bool staged = g.total <= P.lds_cap;
  LdsTile<PREC, 1> lt(dyn_lds, P.lds_cap);
  if (staged) {
    for (int el = threadIdx.x; el < g.total; el += kBlock) {
      int s = find_slot(sh, g.ns, el);
      lt.load(el, sh.gst[s] + (el - sh.off[s]), t, e);
    }
  }
  __syncthreads();
  const int e0 = sh.off[g.ncol + 1];
  const int ncen = sh.off[g.ncol + g.W + 1] - e0;
  const int k0 = sh.gst[g.ncol + 1];
  long long cand_sum = 0, n_sum = 0, n2_sum = 0;
  for (int p = threadIdx.x; p < ncen; p += kBlock) {
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
__global__ __launch_bounds__(kBlock) void k_pass2_lds(KParams P, const int *__restrict__ start, Tiles t,
                                                      const int *__restrict__ slot_h, const int *__restrict__ slot_id,
                                                      const double *gp, EpiCtx e) {
  __shared__ SegShared sh;
  extern __shared__ float4 dyn_lds[];
  Seg g = seg_setup(P, start, sh);
  const bool staged = g.total <= P.lds_cap;
  LdsTile<PREC, 2> lt(dyn_lds, P.lds_cap);
  if (staged) {
    for (int el = threadIdx.x; el < g.total; el += kBlock) {
      int s = find_slot(sh, g.ns, el);
      lt.load(el, sh.gst[s] + (el - sh.off[s]), t, e);
    }
  }
  __syncthreads();
  const int e0 = sh.off[g.ncol + 1];
  const int ncen = sh.off[g.ncol + g.W + 1] - e0;
  const int k0 = sh.gst[g.ncol + 1];
  for (int p = threadIdx.x; p < ncen; p += kBlock) {
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

// =============================================================================
// All pairs (rho > L/4): thread per particle, j = 0..N-1, stage positions not
// wrapped (ChimeraSystem.cpp:131-167, :184-254)
// =============================================================================

template <int PREC, int SYS>
__device__ __forceinline__ void allpairs_pass1(const KParams &P, int i, const double *gx, const double *gy,
                                               const double *gp, Pass1Out &o) {
  const double xi = gx[i], yi = gy[i], pi = gp[i];
  if (SYS == SPC2_CHIMERA) {
    if (PREC == P64) {
      double A = 0.0;
      int n = 0;
      for (int j = 0; j < P.n; ++j) {
        if (j == i) continue;
        double dx = min_image(gx[j] - xi, P.Lx, P.rLx);
        double dy = min_image(gy[j] - yi, P.Ly, P.rLy);
        double r2 = dx * dx + dy * dy;
        if (r2 <= P.rho2) {
          A += sin(gp[j] - pi - P.alpha);
          ++n;
        }
      }
      o.n = n;
      o.written = n != 0;
      o.aphi = n != 0 ? P.sigma * A / (double)n : 0.0;
    } else {
      // factored form; the FP32 variant accumulates in Kahan FP32 over FP64 geometry
      double S = 0.0, C = 0.0;
      Kahan Sf, Cf;
      Sf.init();
      Cf.init();
      int n = 0;
      for (int j = 0; j < P.n; ++j) {
        if (j == i) continue;
        double dx = min_image(gx[j] - xi, P.Lx, P.rLx);
        double dy = min_image(gy[j] - yi, P.Ly, P.rLy);
        double r2 = dx * dx + dy * dy;
        if (r2 <= P.rho2) {
          double sj, cj;
          sincos(gp[j], &sj, &cj);
          if (PREC == P32K) {
            Sf.add((float)sj);
            Cf.add((float)cj);
          } else {
            S += sj;
            C += cj;
          }
          ++n;
        }
      }
      if (PREC == P32K) {
        S = Sf.value();
        C = Cf.value();
      }
      double sa, ca;
      sincos(pi + P.alpha, &sa, &ca);
      o.n = n;
      o.written = n != 0;
      o.aphi = n != 0 ? P.sigma * (ca * S - sa * C) / (double)n : 0.0;
    }
    o.n2 = 0;
    o.vortex_add = 0;
  } else {
    double A = 0.0, B = 0.0, Cc = 0.0;
    int n0 = 0, n1 = 0, n2 = 0;
    for (int j = 0; j < P.n; ++j) {
      if (j == i) continue;
      double dx = min_image(gx[j] - xi, P.Lx, P.rLx);
      double dy = min_image(gy[j] - yi, P.Ly, P.rLy);
      double dr = sqrt(dx * dx + dy * dy);
      if (dr <= P.xi_a) {
        A += vortex_w(P, dr) * sin(gp[j] - pi);
        ++n0;
      } else if (dr <= P.rho) {
        B += vortex_w(P, dr) * sin(gp[j] - pi);
        ++n1;
      }
      if (dr <= P.xi_r) {
        Cc += -P.kappa * sin(atan2(dy, dx) - pi);
        ++n2;
      }
    }
    o.n = n0 + n1;
    o.n2 = n2;
    if (n0 + n1 > 0) {
      o.aphi = (A + B) / ((double)n0 + (double)n1);
      if (n2 > 0) o.aphi += Cc / (double)n2;
      o.written = 1;
      o.vortex_add = 0;
    } else {
      o.aphi = 0.0;
      o.written = 0;
      o.vortex_add = n2 > 0;
      o.coh = n2 > 0 ? Cc / (double)n2 : 0.0;
    }
  }
  o.phase = pi;
}

template <int PREC, int SYS, int EPI>
__global__ __launch_bounds__(kBlock) void k_pass1_allpairs(KParams P, const double *gx, const double *gy,
                                                           const double *gp, EpiCtx e, unsigned long long *counts) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool active = i < P.n;
  Pass1Out o{};
  if (active) {
    allpairs_pass1<PREC, SYS>(P, i, gx, gy, gp, o);
    epilogue1<EPI>(P, e, i, i, i, o);  // all pairs: home order is the id order
  }
  if (counts) count_add(counts, active ? P.n - 1 : 0, active ? o.n : 0, active ? o.n2 : 0);
}

template <int SYS, int EPI>
__global__ __launch_bounds__(kBlock) void k_pass2_allpairs(KParams P, const double *gx, const double *gy,
                                                           const double *gp, EpiCtx e) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P.n) return;
  const double xi = gx[i], yi = gy[i], pi = gp[i];
  const double ai = e.aphi_slot[i], n = e.n_slot[i];
  double a0 = 0.0, a1 = 0.0;
  if (EPI == EPI_ITO2) {
    a0 = e.a0[i];
    a1 = e.a1[i];
  }
  if (SYS == SPC2_VORTEX) {
    // VortexArraysSystemPtr.cpp:305-353 (FP64 form for every precision)
    const double n2 = e.n2_slot[i];
    const double vxi = P.v0 * cos(pi), vyi = P.v0 * sin(pi);
    for (int j = 0; j < P.n; ++j) {
      if (j == i) continue;
      double dx = min_image(gx[j] - xi, P.Lx, P.rLx);
      double dy = min_image(gy[j] - yi, P.Ly, P.rLy);
      double dr = sqrt(dx * dx + dy * dy);
      const bool al = (dr <= P.rho) && (n > 0.0), co = (dr <= P.xi_r) && (n2 > 0.0);
      if (!al && !co) continue;
      const double pj = gp[j];
      const double vxj = P.v0 * cos(pj), vyj = P.v0 * sin(pj);
      if (al) {
        double cph = cos(pj - pi);
        double sph = sin(pj - pi);
        double w = vortex_w(P, dr);
        double dw = vortex_dw(P, dr);
        a0 += -w * cph / n;
        a1 += ((vxj - vxi) * dw * dx / dr + (vyj - vyi) * dw * dy / dr) * sph / n + (e.aphi_slot[j] - ai) * w * cph / n -
              2.0 * P.D * sph * w / n;
      }
      if (co) {
        double th = atan2(dy, dx);
        double cpp = cos(th - pi);
        double spp = sin(th - pi);
        a0 += P.kappa * cpp / n2;
        a1 += -P.kappa * ((vxj - vxi) * (-dy) / dr / dr + (vyj - vyi) * dx / dr / dr - ai) * cpp / n2 +
              P.kappa * P.D * spp / n2;
      }
    }
    Pass2Out o;
    o.phase = pi;
    o.ai = ai;
    o.a0 = a0 * sqrt(2.0 * P.D);
    o.a1 = a1;
    epilogue2<EPI>(P, e, i, i, o);
    return;
  }
  // ChimeraSystem.cpp:223-254 (FP64 form for every precision: small-N path)
  for (int j = 0; j < P.n; ++j) {
    if (j == i) continue;
    double dx = min_image(gx[j] - xi, P.Lx, P.rLx);
    double dy = min_image(gy[j] - yi, P.Ly, P.rLy);
    double r2 = dx * dx + dy * dy;
    if ((r2 <= P.rho2) && (n != 0)) {
      double cph = cos(gp[j] - pi - P.alpha);
      a0 += cph / n;
      a1 += (e.aphi_slot[j] - ai) * cph / n;
    }
  }
  Pass2Out o;
  o.phase = pi;
  o.ai = ai;
  o.a0 = a0 * -sqrt(2.0 * P.D);
  o.a1 = a1 + -2.0 * P.D * ai;
  epilogue2<EPI>(P, e, i, i, o);
}

// =============================================================================
// dispatch
// =============================================================================

template <int PREC, int PASS>
size_t lds_bytes(const Sim &s) {
  return (size_t)s.kp.lds_cap * LdsTile<PREC, PASS>::bytes;
}

template <int PREC, int SYS, int EPI>
void pass1_typed(Sim &s, const EpiCtx &e) {
  unsigned long long *cnt = s.counting ? s.counts : nullptr;
  if (s.linked) {
    const dim3 grid(s.ny * s.kp.segs), blk(kBlock);
    const size_t sh = lds_bytes<PREC, 1>(s);
    if (SYS == SPC2_CHIMERA)
      launch_pair(s, k_pass1_lds<ChimAcc<PREC>, PREC, EPI>, grid, blk, sh, s.kp, (const int *)s.start, s.t,
                  (const int *)s.slot_h, (const int *)s.slot_id, (const double *)s.gp, e, cnt);
    else
      launch_pair(s, k_pass1_lds<VortAcc<PREC>, PREC, EPI>, grid, blk, sh, s.kp, (const int *)s.start, s.t,
                  (const int *)s.slot_h, (const int *)s.slot_id, (const double *)s.gp, e, cnt);
  } else {
    launch_pair(s, k_pass1_allpairs<PREC, SYS, EPI>, dim3(blocks_for(s.n)), dim3(kBlock), 0, s.kp,
                (const double *)s.gx, (const double *)s.gy, (const double *)s.gp, e, cnt);
  }
}

template <int PREC, int SYS>
void pass1_epi(Sim &s, int epi, const EpiCtx &e) {
  switch (epi) {
    case EPI_K: pass1_typed<PREC, SYS, EPI_K>(s, e); break;
    case EPI_ITO1: pass1_typed<PREC, SYS, EPI_ITO1>(s, e); break;
    case EPI_SE: pass1_typed<PREC, SYS, EPI_SE>(s, e); break;
    case EPI_RK4_FINAL: pass1_typed<PREC, SYS, EPI_RK4_FINAL>(s, e); break;
  }
}

template <int PREC>
void pass1_sys(Sim &s, int epi, const EpiCtx &e) {
  if (s.P.system == SPC2_CHIMERA)
    pass1_epi<PREC, SPC2_CHIMERA>(s, epi, e);
  else
    pass1_epi<PREC, SPC2_VORTEX>(s, epi, e);
}

template <int PREC, int SYS, int EPI>
void pass2_typed(Sim &s, const EpiCtx &e) {
  if (s.linked) {
    const size_t sh = lds_bytes<PREC, 2>(s);
    const dim3 grid(s.ny * s.kp.segs), blk(kBlock);
    if (SYS == SPC2_CHIMERA)
      launch_pair(s, k_pass2_lds<Chim2Acc<PREC>, PREC, EPI>, grid, blk, sh, s.kp, (const int *)s.start, s.t,
                  (const int *)s.slot_h, (const int *)s.slot_id, (const double *)s.gp, e);
    else
      launch_pair(s, k_pass2_lds<Vort2Acc<PREC>, PREC, EPI>, grid, blk, sh, s.kp, (const int *)s.start, s.t,
                  (const int *)s.slot_h, (const int *)s.slot_id, (const double *)s.gp, e);
  } else {
    launch_pair(s, k_pass2_allpairs<SYS, EPI>, dim3(blocks_for(s.n)), dim3(kBlock), 0, s.kp, (const double *)s.gx,
                (const double *)s.gy, (const double *)s.gp, e);
  }
}

template <int PREC, int SYS>
void pass2_epi(Sim &s, int epi, const EpiCtx &e) {
  if (epi == EPI_ITO2)
    pass2_typed<PREC, SYS, EPI_ITO2>(s, e);
  else
    pass2_typed<PREC, SYS, EPI_SRK>(s, e);
}

template <int PREC>
void pass2_sys(Sim &s, int epi, const EpiCtx &e) {
  if (s.P.system == SPC2_CHIMERA)
    pass2_epi<PREC, SPC2_CHIMERA>(s, epi, e);
  else
    pass2_epi<PREC, SPC2_VORTEX>(s, epi, e);
}

}  // namespace

void configure_segments(Sim &s) {
  KParams &k = s.kp;
  if (!s.linked) {
    k.seg_w = k.segs = 1;
    k.lds_cap = 0;
    return;
  }
  const double per_cell = (double)s.n / (double)s.nc;
  int w = (int)std::floor(218.0 / std::max(per_cell, 1e-9));
  w = std::max(1, std::min({w, kMaxSegW, s.nx}));
  k.seg_w = w;
  k.segs = (s.nx + w - 1) / w;
  const double halo = 3.0 * (w + 2) * per_cell;
  int cap = (int)std::ceil((1.6 * halo + 160.0) / 64.0) * 64;
  int elem = 0;
  switch (s.prec) {
    case P32K: elem = LdsTile<P32K, 2>::bytes; break;
    case P64: elem = LdsTile<P64, 2>::bytes; break;
    default: elem = LdsTile<P64F, 2>::bytes;
  }
  cap = std::min(cap, (60 * 1024) / elem);
  if (const char *env = std::getenv("SPC2_LDS_CAP")) cap = std::max(0, std::min(cap, std::atoi(env)));
  k.lds_cap = cap;
}

void launch_pass1(Sim &s, int epi, double *ox, double *oy, double *op, const StepArgs &sa, double *const k1[3],
                  double *const k2[3], double *const k3[3]) {
  EpiCtx e = make_ctx(s, ox, oy, op, sa, k1, k2, k3);
  switch (s.prec) {
    case P64: pass1_sys<P64>(s, epi, e); break;
    case P64F: pass1_sys<P64F>(s, epi, e); break;
    default: pass1_sys<P32K>(s, epi, e);
  }
}

void launch_pass2(Sim &s, int epi, double *a0, double *a1, const StepArgs &sa) {
  EpiCtx e = make_ctx(s, nullptr, nullptr, nullptr, sa, nullptr, nullptr, nullptr);
  e.a0 = a0;
  e.a1 = a1;
  switch (s.prec) {
    case P64: pass2_sys<P64>(s, epi, e); break;
    case P64F: pass2_sys<P64F>(s, epi, e); break;
    default: pass2_sys<P32K>(s, epi, e);
  }
}

}  // namespace spc2
