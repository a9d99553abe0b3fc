// Pair accumulators, stepper epilogues and launch helpers shared by the
// cell-stencil kernels (spc2_pairs.hip) and the Verlet-list kernels
// (spc2_verlet.hip).  Every accumulator has the same interface:
//   init(src, e, k, tiles, params[, pass-2 inputs])  centre particle at tile
//       index e (slot k for the FP64 re-checks),
//   row(dy, b1, b2, params)  the stencil row and column bounds of the next
//       candidates (FP32 cell-local forms only),
//   pair(src, m, params)     candidate at tile index m,
//   finish(params, phase, out).
// The FP32 forms also take Verlet-list entries: pair_c<MODE>(src, m, fx, fy,
// tile, aj, params) with the partner's tile loaded (its y in the Verlet
// passes' frame, ColOffsets::frame) and its stencil column and row as floats,
// and flush() at the end of each list group (SumMode).
// Compiled with -ffp-contract=off.
#pragma once

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include "spc2_device.h"
#include "spc2_internal.h"

namespace spc2 {
namespace {

constexpr int kPairBlock = 256;

// ---- tile sources -------------------------------------------------------------
// Src32 reads the FP32 tile (cell-local x, y and cos, sin of the phase).  The
// global slot of tile index m: slot[m] when `slot` is set, else, when the tiles
// are a staged halo of `rn` runs (run r's tiles from index roff[r] hold slots
// rgst[r], rgst[r] + 1, ...), found from the runs, else m itself.
struct Src32 {
  const float4 *f4;
  const int *slot;
  const float *aj;
  const int *rgst = nullptr, *roff = nullptr;
  int rn = 0;
  __device__ __forceinline__ int gslot(int m) const {
    if (slot) return slot[m];
    if (rgst) {
      int k = 0;
      while (k + 1 < rn && roff[k + 1] <= m) ++k;
      return rgst[k] + (m - roff[k]);
    }
    return m;
  }
};
struct Src64 {
  const double *x, *y, *p, *c, *s, *aj;
};

template <int PREC>
struct SrcOf {
  using type = Src64;
};
template <>
struct SrcOf<P32K> {
  using type = Src32;
};

// AlignmentStrength, VortexArraysSystemPtr.cpp:737-749
__device__ __forceinline__ double vortex_w(const KParams &P, double r) {
  if (r < P.xi_a) return P.mu_plus * (1.0 - r * r / (P.xi_a * P.xi_a));
  if (r < P.rho) return -P.mu_minus * 4.0 * (r - P.xi_a) * (P.rho - r) / (P.rho - P.xi_a) / (P.rho - P.xi_a);
  return 0.0;
}

// AlignmentStrengthDerivative, VortexArraysSystemPtr.cpp:751-763
__device__ __forceinline__ double vortex_dw(const KParams &P, double r) {
  if (r < P.xi_a) return P.mu_plus * (-2.0 * r / P.xi_a / P.xi_a);
  if (r < P.rho) return -P.mu_minus * 4.0 * (-2.0 * r + P.rho + P.xi_a) / (P.rho - P.xi_a) / (P.rho - P.xi_a);
  return 0.0;
}

// FP64 distance^2 of slots k and m with the reference min-image: decides the
// rare FP32 pairs that fall inside a cutoff band.
__device__ __forceinline__ double exact_r2(const double *tx, const double *ty, int k, int m, double Lx, double rLx,
                                           double Ly, double rLy) {
  double dx = min_image(tx[m] - tx[k], Lx, rLx);
  double dy = min_image(ty[m] - ty[k], Ly, rLy);
  return dx * dx + dy * dy;
}

// FP64 min-image separation of slots k and m, for cohesion pairs closer than
// a hundredth of a cell: there the 1/r and 1/r^2 factors would amplify the
// FP32 rounding of the cell-local coordinates (about 6e-8 of a cell width)
// beyond 1e-6 of the term.
__device__ __forceinline__ void exact_d(const double *tx, const double *ty, int k, int m, const KParams &P, double &dx,
                                        double &dy) {
  dx = min_image(tx[m] - tx[k], P.Lx, P.rLx);
  dy = min_image(ty[m] - ty[k], P.Ly, P.rLy);
}

// Per-particle result of pass 1.
struct Pass1Out {
  double phase;    // FP64 stage phase of i
  double aphi;     // phi derivative when `written`
  int written;     // 0: phi entry left as is (chimera n_i = 0; vortex n0+n1 = 0 and n2 = 0)
  int vortex_add;  // vortex: 1 if only the cohesion term applies (added to the old value)
  double coh;      // vortex cohesion / n2 for vortex_add
  int n;           // partners within rho
  int n2;          // cohesion partners
  int has_cs;      // cs, sn hold cos, sin of phase (computed by finish)
  double cs, sn;
};

// FP32 x offset of the candidate's cell column relative to i's: m < b1 is the
// left stencil column, b1 <= m < b2 the middle one, m >= b2 the right one.
// Written as -x_i + sel * h with sel in {-1, 0, 1} (exact), so that the
// compiler keeps every member in registers (a select between members can be
// turned into a select between their addresses, which pushes the whole
// accumulator to scratch memory).
struct ColOffsets {
  float oM, hxf, oy;
  int b1, b2;
  float kx, ky;  // -h - xl, -h - yl: the offsets of stencil column / row d are (d + 1) h + k
  float yl0, yf;  // the centre's cell-local y, and its y in the Verlet passes' frame
  __device__ __forceinline__ void set(float xl, float h) {
    oM = -xl;
    hxf = h;
  }
  __device__ __forceinline__ void set_codes(float xl, float yl, float hx, float hy) {
    kx = -hx - xl;
    ky = -hy - yl;
    yl0 = yl;
    yf = yl;
  }
  // The Verlet passes' y frame: the centre's y is yl + rows h_y (rows in {0, 1},
  // exact) and a list entry's tile arrives with its y in the same frame, so
  // that the separation's y is v.y - yf.
  __device__ __forceinline__ void frame(float rows_hy) { yf = yl0 + rows_hy; }
  // x / y offsets from a list entry's stencil column dcx + 1 and row dcy + 1
  // (each 0, 1 or 2, as floats)
  __device__ __forceinline__ float cdx(float fx, float hx) const { return __builtin_fmaf(fx, hx, kx); }
  __device__ __forceinline__ float cdy(float fy, float hy) const { return __builtin_fmaf(fy, hy, ky); }
  __device__ __forceinline__ float ox(int m) const {
    float sel = (float)((m >= b2) - (m < b1));
    return __builtin_fmaf(sel, hxf, oM);
  }
  // ox of stencil column d in {-1, 0, 1} (the value ox gives inside that column)
  __device__ __forceinline__ float col(int d) const { return __builtin_fmaf((float)d, hxf, oM); }
};

// How an FP32 pair term is summed. The Verlet walks sum every eight list
// entries in plain FP32 partials (the group's first entry starts them) and add
// the partials to the Kahan sums once per group (flush()): a compensated sum
// over groups, whose error bound is the Kahan bound plus 7 ulp of each group's
// absolute sum.  The cell walks do the same per stencil cell (cell_begin() zeroes
// the partials, SUM_GROUP adds, flush() after the cell): the Kahan bound plus
// (m - 1) ulp of each cell's absolute sum, m its candidates.  SUM_KAHAN adds
// each term to the Kahan sums.
enum SumMode { SUM_KAHAN = 0, SUM_GROUP_FIRST = 1, SUM_GROUP = 2 };

template <int MODE>
__device__ __forceinline__ void group_add(float &p, float x) {
  p = MODE == SUM_GROUP_FIRST ? x : p + x;
}
template <int MODE>
__device__ __forceinline__ void group_add_prod(float &p, float a, float b) {
  p = MODE == SUM_GROUP_FIRST ? a * b : __builtin_fmaf(a, b, p);
}

// A kernel-argument constant copied into a vector register once: an FP32
// instruction reads at most one scalar register, so keeping a constant that
// meets another one in the same instruction in a vector register saves a move
// per use.
__device__ __forceinline__ float in_vgpr(float x) {
  float r;
  asm("v_mov_b32 %0, %1" : "=v"(r) : "v"(x));
  return r;
}

// =============================================================================
// Chimera pass 1 (ChimeraSystem.cpp:391-424)
// =============================================================================
template <int PREC>
struct ChimAcc;

// FP64 reference form: A += sin(phi_j - phi_i - alpha) over |r_ij| <= rho
template <>
struct ChimAcc<P64> {
  double xi, yi, pi, A;
  int n;
  __device__ __forceinline__ void init(const Src64 &s, int e, int, const Tiles &, const KParams &) {
    xi = s.x[e];
    yi = s.y[e];
    pi = s.p[e];
    A = 0.0;
    n = 0;
  }
  __device__ __forceinline__ void row(int, int, int, const KParams &) {}
  __device__ __forceinline__ void pair(const Src64 &s, int m, const KParams &P) {
    double dx = min_image(s.x[m] - xi, P.Lx, P.rLx);
    double dy = min_image(s.y[m] - yi, P.Ly, P.rLy);
    double r2 = dx * dx + dy * dy;
    if (r2 <= P.rho2) {
      A += sin(s.p[m] - pi - P.alpha);
      ++n;
    }
  }
  __device__ __forceinline__ void finish(const KParams &P, double, Pass1Out &o) const {
    o.phase = pi;
    o.n = n;
    o.n2 = 0;
    o.written = n != 0;
    o.vortex_add = 0;
    o.aphi = n != 0 ? P.sigma * A / (double)n : 0.0;
  }
};

// FP64 factored: sum sin(phi_j - phi_i - alpha)
//   = cos(phi_i + alpha) sum sin phi_j - sin(phi_i + alpha) sum cos phi_j
template <>
struct ChimAcc<P64F> {
  double xi, yi, pi, S, C;
  int n;
  __device__ __forceinline__ void init(const Src64 &s, int e, int k, const Tiles &t, const KParams &) {
    xi = s.x[e];
    yi = s.y[e];
    pi = t.p[k];
    S = 0.0;
    C = 0.0;
    n = 0;
  }
  __device__ __forceinline__ void row(int, int, int, const KParams &) {}
  __device__ __forceinline__ void pair(const Src64 &s, int m, const KParams &P) {
    double dx = min_image(s.x[m] - xi, P.Lx, P.rLx);
    double dy = min_image(s.y[m] - yi, P.Ly, P.rLy);
    double r2 = dx * dx + dy * dy;
    if (r2 <= P.rho2) {
      S += s.s[m];
      C += s.c[m];
      ++n;
    }
  }
  __device__ __forceinline__ void finish(const KParams &P, double, Pass1Out &o) const {
    double sa, ca;
    sincos(pi + P.alpha, &sa, &ca);
    double A = ca * S - sa * C;
    o.phase = pi;
    o.n = n;
    o.n2 = 0;
    o.written = n != 0;
    o.vortex_add = 0;
    o.aphi = n != 0 ? P.sigma * A / (double)n : 0.0;
  }
};

// FP32 + Kahan over cell-local coordinates; cutoff band re-decided in FP64
template <>
struct ChimAcc<P32K> {
  float xl, yl;
  ColOffsets co;
  Kahan S, C;
  float ps, pc;  // the current list group's partial sums
  float pn;      // pair2_x's in-range count in the current cell
  int n, k;
  const double *tx, *ty;
  __device__ __forceinline__ void init(const Src32 &s, int e, int kk, const Tiles &t, const KParams &P) {
    float4 v = s.f4[e];
    xl = v.x;
    yl = v.y;
    co.set(xl, P.hxf);
    co.set_codes(xl, yl, P.hxf, P.hyf);
    S.init();
    C.init();
    ps = pc = 0.0f;
    pn = 0.0f;
    n = 0;
    k = kk;
    tx = t.x;
    ty = t.y;
  }
  __device__ __forceinline__ void row(int dyo, int b1, int b2, const KParams &P) {
    co.oy = (float)dyo * P.hyf - yl;
    co.b1 = b1;
    co.b2 = b2;
  }
  template <int MODE = SUM_KAHAN>
  __device__ __forceinline__ void pair(const Src32 &s, int m, const KParams &P) {
    const float4 v = s.f4[m];
    pair_d<MODE>(s, m, v, v.x + co.ox(m), v.y + co.oy, P);
  }
  // pair() with the column offset known (a run of candidates in one stencil column)
  template <int MODE = SUM_KAHAN>
  __device__ __forceinline__ void pair_x(const Src32 &s, int m, float ox, const KParams &P) {
    const float4 v = s.f4[m];
    pair_d<MODE>(s, m, v, v.x + ox, v.y + co.oy, P);
  }
  // the cell walks' partial sums over one stencil cell start at zero
  __device__ __forceinline__ void cell_begin() { ps = pc = pn = 0.0f; }
  // a list entry with the partner's tile already loaded (aj: pass-2 input,
  // unused here); returns whether the partner is within rho
  template <int MODE>
  __device__ __forceinline__ bool pair_c(const Src32 &s, int m, float fx, float fy, const float4 &v, float,
                                         const KParams &P) {
    return pair_d<MODE>(s, m, v, v.x + co.cdx(fx, P.hxf), v.y - co.yf, P);
  }
  __device__ __forceinline__ void flush() {
    S.add(ps);
    C.add(pc);
  }
  // flush() after a cell walked with pair2_x, with its count
  __device__ __forceinline__ void flush_cell() {
    flush();
    n += (int)pn;
  }
  // Two consecutive cell candidates, the staged tiles a, b read from t2 (LDS
  // indices m = t2 - s.f4 and m + 1; column offsets oxa, oxb), into the cell's partials,
  // with one branch for both band re-checks.  An in-range
  // flag f in {0, 1} enters as fma(f, x, p): p + x when in range and p + 0
  // otherwise, the same values as pair_x<SUM_GROUP> for each in turn.
  __device__ __forceinline__ void pair2_x(const Src32 &s, const float4 *t2, const float4 &a, const float4 &b,
                                          float oxa, float oxb, const KParams &P) {
    const float dxa = a.x + oxa, dya = a.y + co.oy, dxb = b.x + oxb, dyb = b.y + co.oy;
    const float r2a = __builtin_fmaf(dxa, dxa, dya * dya), r2b = __builtin_fmaf(dxb, dxb, dyb * dyb);
    bool ina = r2a <= P.rho2f_hi, inb = r2b <= P.rho2f_hi;
    const bool ba = (r2a >= P.rho2f_lo) & ina, bb = (r2b >= P.rho2f_lo) & inb;
    if (ba | bb) {
      const int m = (int)(t2 - s.f4);
      if (ba) ina = exact_r2(tx, ty, k, s.gslot(m), P.Lx, P.rLx, P.Ly, P.rLy) <= P.rho2;
      if (bb) inb = exact_r2(tx, ty, k, s.gslot(m + 1), P.Lx, P.rLx, P.Ly, P.rLy) <= P.rho2;
    }
    const float fa = ina ? 1.0f : 0.0f, fb = inb ? 1.0f : 0.0f;
    ps = __builtin_fmaf(fa, a.w, ps);
    pc = __builtin_fmaf(fa, a.z, pc);
    ps = __builtin_fmaf(fb, b.w, ps);
    pc = __builtin_fmaf(fb, b.z, pc);
    pn += fa + fb;  // the cell's count so far, exact in FP32
  }
  template <int MODE = SUM_KAHAN>
  __device__ __forceinline__ bool pair_d(const Src32 &s, int m, const float4 &v, float dx, float dy,
                                         const KParams &P) {
    float r2 = __builtin_fmaf(dx, dx, dy * dy);
    bool in = r2 <= P.rho2f_hi;  // outside the band, the same as r2 <= rho2f
    if ((r2 >= P.rho2f_lo) & in) in = exact_r2(tx, ty, k, s.gslot(m), P.Lx, P.rLx, P.Ly, P.rLy) <= P.rho2;
    if constexpr (MODE == SUM_KAHAN) {
      S.add(in ? v.w : 0.0f);
      C.add(in ? v.z : 0.0f);
    } else {
      group_add<MODE>(ps, in ? v.w : 0.0f);
      group_add<MODE>(pc, in ? v.z : 0.0f);
    }
    n += in;
    return in;
  }
  __device__ __forceinline__ void finish(const KParams &P, double phase, Pass1Out &o) const {
    double sa, ca;
    sincos(phase + P.alpha, &sa, &ca);
    double A = ca * S.value() - sa * C.value();
    o.phase = phase;
    o.n = n;
    o.n2 = 0;
    o.written = n != 0;
    o.vortex_add = 0;
    o.aphi = n != 0 ? P.sigma * A / (double)n : 0.0;
  }
};

// =============================================================================
// Vortex arrays pass 1 (VortexArraysSystemPtr.cpp:459-504)
// =============================================================================
template <int PREC>
struct VortAcc;

template <>
struct VortAcc<P64> {
  double xi, yi, pi, A, B, Cc;
  int n0, n1, n2;
  __device__ __forceinline__ void init(const Src64 &s, int e, int, const Tiles &, const KParams &) {
    xi = s.x[e];
    yi = s.y[e];
    pi = s.p[e];
    A = B = Cc = 0.0;
    n0 = n1 = n2 = 0;
  }
  __device__ __forceinline__ void row(int, int, int, const KParams &) {}
  __device__ __forceinline__ void pair(const Src64 &s, int m, const KParams &P) {
    double dx = min_image(s.x[m] - xi, P.Lx, P.rLx);
    double dy = min_image(s.y[m] - yi, P.Ly, P.rLy);
    double r2 = dx * dx + dy * dy;
    if (r2 > P.outer2 * 1.000000000001) return;  // cannot pass the sqrt tests below
    double dr = sqrt(r2);
    double pj = s.p[m];
    if (dr <= P.xi_a) {
      A += vortex_w(P, dr) * sin(pj - pi);
      ++n0;
    } else if (dr <= P.rho) {
      B += vortex_w(P, dr) * sin(pj - pi);
      ++n1;
    }
    if (dr <= P.xi_r) {
      Cc += -P.kappa * sin(atan2(dy, dx) - pi);
      ++n2;
    }
  }
  __device__ __forceinline__ void finish(const KParams &, double, Pass1Out &o) const {
    o.phase = pi;
    o.n = n0 + n1;
    o.n2 = n2;
    double n01 = (double)n0 + (double)n1;
    if (n0 + n1 > 0) {
      o.aphi = (A + B) / n01;
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
};

// factored: sum W sin(phi_j - phi_i) = cos phi_i sum W sin phi_j
//   - sin phi_i sum W cos phi_j;  sin(atan2(dy, dx) - phi_i) = (dy cos phi_i - dx sin phi_i) / r
template <>
struct VortAcc<P64F> {
  double xi, yi, pi, SW, CW, Tx, Ty;
  int n01, n2;
  __device__ __forceinline__ void init(const Src64 &s, int e, int k, const Tiles &t, const KParams &) {
    xi = s.x[e];
    yi = s.y[e];
    pi = t.p[k];
    SW = CW = Tx = Ty = 0.0;
    n01 = n2 = 0;
  }
  __device__ __forceinline__ void row(int, int, int, const KParams &) {}
  __device__ __forceinline__ void pair(const Src64 &s, int m, const KParams &P) {
    double dx = min_image(s.x[m] - xi, P.Lx, P.rLx);
    double dy = min_image(s.y[m] - yi, P.Ly, P.rLy);
    double r2 = dx * dx + dy * dy;
    if (r2 > P.outer2 * 1.000000000001) return;
    double dr = sqrt(r2);
    if (dr <= P.rho) {
      double w = vortex_w(P, dr);
      SW += w * s.s[m];
      CW += w * s.c[m];
      ++n01;
    }
    if (dr <= P.xi_r) {
      if (dr > 0.0) {
        Tx += dx / dr;
        Ty += dy / dr;
      } else {
        Tx += 1.0;  // atan2(+0, +0) = 0
      }
      ++n2;
    }
  }
  __device__ __forceinline__ void finish(const KParams &P, double, Pass1Out &o) const {
    double sp, cp;
    sincos(pi, &sp, &cp);
    o.phase = pi;
    o.n = n01;
    o.n2 = n2;
    double coh = n2 > 0 ? -P.kappa * (cp * Ty - sp * Tx) / (double)n2 : 0.0;
    if (n01 > 0) {
      o.aphi = (cp * SW - sp * CW) / (double)n01 + coh;
      o.written = 1;
      o.vortex_add = 0;
    } else {
      o.aphi = 0.0;
      o.written = 0;
      o.vortex_add = n2 > 0;
      o.coh = coh;
    }
  }
};

template <>
struct VortAcc<P32K> {
  float xl, yl;
  ColOffsets co;
  Kahan SW, CW, Tx, Ty;
  float psw, pcw;    // the current list group's partial sums
  float mu_pv, wav;  // mu_pf and w_af in vector registers
  int n01, n2, k;
  const double *tx, *ty;
  __device__ __forceinline__ void init(const Src32 &s, int e, int kk, const Tiles &t, const KParams &P) {
    float4 v = s.f4[e];
    xl = v.x;
    yl = v.y;
    co.set(xl, P.hxf);
    co.set_codes(xl, yl, P.hxf, P.hyf);
    SW.init();
    CW.init();
    Tx.init();
    Ty.init();
    psw = pcw = 0.0f;
    mu_pv = in_vgpr(P.mu_pf);
    wav = in_vgpr(P.w_af);
    n01 = n2 = 0;
    k = kk;
    tx = t.x;
    ty = t.y;
  }
  __device__ __forceinline__ void row(int dyo, int b1, int b2, const KParams &P) {
    co.oy = (float)dyo * P.hyf - yl;
    co.b1 = b1;
    co.b2 = b2;
  }
  // Branch-free except for the rare cases (cutoff band, cohesion range): every
  // lane evaluates W and adds W sin, W cos masked to zero outside rho.
  template <int MODE = SUM_KAHAN>
  __device__ __forceinline__ void pair(const Src32 &s, int m, const KParams &P) {
    const float4 v = s.f4[m];
    pair_d<MODE>(s, m, v, v.x + co.ox(m), v.y + co.oy, P);
  }
  // pair() with the column offset known (a run of candidates in one stencil column)
  template <int MODE = SUM_KAHAN>
  __device__ __forceinline__ void pair_x(const Src32 &s, int m, float ox, const KParams &P) {
    const float4 v = s.f4[m];
    pair_d<MODE>(s, m, v, v.x + ox, v.y + co.oy, P);
  }
  // the cell walks' partial sums over one stencil cell start at zero
  __device__ __forceinline__ void cell_begin() { psw = pcw = 0.0f; }
  // a list entry with the partner's tile already loaded (aj: pass-2 input,
  // unused here).  DEFER: the cohesion term is left to the caller, who gets
  // whether the partner lies within xi_r's band and calls cohesion() later, with
  // the same separation, in list order.
  template <int MODE, bool DEFER = false>
  __device__ __forceinline__ bool pair_c(const Src32 &s, int m, float fx, float fy, const float4 &v, float,
                                         const KParams &P) {
    return pair_d<MODE, DEFER>(s, m, v, v.x + co.cdx(fx, P.hxf), v.y - co.yf, P);
  }
  __device__ __forceinline__ void flush() {
    SW.add(psw);
    CW.add(pcw);
  }
  __device__ __forceinline__ void flush_cell() { flush(); }
  // W(r) of an FP32 r^2, masked to zero outside rho
  __device__ __forceinline__ float weight(float r2, bool inr, const KParams &P) const {
    float r = __builtin_amdgcn_sqrtf(r2);
    float w_in = __builtin_fmaf(-P.mu_p_over_xa2f, r2, mu_pv);
    float w_an = __builtin_fmaf(r, __builtin_fmaf(-P.c4f, r, wav), P.w_bf);
    float w = r2 <= P.xa2f ? w_in : w_an;
    return inr ? w : 0.0f;
  }
  // An entry's stages for walks that test two entries before summing either:
  // the FP32 rho test (returns whether the FP64 re-check rho_exact decides),
  // then term() with the decision.  Together, the values of pair_d<MODE, true>.
  __device__ __forceinline__ bool rho_test(float r2, bool &inr, const KParams &P) const {
    inr = r2 <= P.rho2f_hi;
    return (r2 >= P.rho2f_lo) & inr;
  }
  __device__ __forceinline__ bool rho_exact(const Src32 &s, int m, const KParams &P) const {
    return sqrt(exact_r2(tx, ty, k, s.gslot(m), P.Lx, P.rLx, P.Ly, P.rLy)) <= P.rho;
  }
  template <int MODE>
  __device__ __forceinline__ void term(const float4 &v, float r2, bool inr, const KParams &P) {
    const float w = weight(r2, inr, P);
    group_add_prod<MODE>(psw, w, v.w);
    group_add_prod<MODE>(pcw, w, v.z);
    n01 += inr;
  }
  // Two consecutive cell candidates, the staged tiles a, b read from t2 (LDS
  // indices m = t2 - s.f4 and m + 1; column offsets oxa, oxb), with one branch for both rho
  // band re-checks and one for both cohesion tests.
  // The terms, counts and cohesion sums are those of pair_x<SUM_GROUP> for
  // each in turn (the W partials and the cohesion sums are separate chains).
  __device__ __forceinline__ void pair2_x(const Src32 &s, const float4 *t2, const float4 &a, const float4 &b,
                                          float oxa, float oxb, const KParams &P) {
    const float dxa = a.x + oxa, dya = a.y + co.oy, dxb = b.x + oxb, dyb = b.y + co.oy;
    const float r2a = __builtin_fmaf(dxa, dxa, dya * dya), r2b = __builtin_fmaf(dxb, dxb, dyb * dyb);
    bool ina = r2a <= P.rho2f_hi, inb = r2b <= P.rho2f_hi;
    const bool ba = (r2a >= P.rho2f_lo) & ina, bb = (r2b >= P.rho2f_lo) & inb;
    if (ba | bb) {
      const int m = (int)(t2 - s.f4);
      if (ba) ina = sqrt(exact_r2(tx, ty, k, s.gslot(m), P.Lx, P.rLx, P.Ly, P.rLy)) <= P.rho;
      if (bb) inb = sqrt(exact_r2(tx, ty, k, s.gslot(m + 1), P.Lx, P.rLx, P.Ly, P.rLy)) <= P.rho;
    }
    const float wa = weight(r2a, ina, P), wb = weight(r2b, inb, P);
    psw = __builtin_fmaf(wa, a.w, psw);
    pcw = __builtin_fmaf(wa, a.z, pcw);
    psw = __builtin_fmaf(wb, b.w, psw);
    pcw = __builtin_fmaf(wb, b.z, pcw);
    n01 += (int)ina + (int)inb;
    const bool ca = r2a <= P.xr2f_hi, cb = r2b <= P.xr2f_hi;
    if (ca | cb) {
      const int m = (int)(t2 - s.f4);
      if (ca) cohesion(s, m, dxa, dya, r2a, P);
      if (cb) cohesion(s, m + 1, dxb, dyb, r2b, P);
    }
  }
  template <int MODE = SUM_KAHAN, bool DEFER = false>
  __device__ __forceinline__ bool pair_d(const Src32 &s, int m, const float4 &v, float dx, float dy,
                                         const KParams &P) {
    float r2 = __builtin_fmaf(dx, dx, dy * dy);
    // Outside the rho band the FP32 test decides (r2 <= rho2 + band there means
    // r2 <= rho2 - band); inside it the reference's FP64 test sqrt(r^2) <= rho
    // does (VortexArraysSystemPtr.cpp:473-488).
    bool inr = r2 <= P.rho2f_hi;
    if ((r2 >= P.rho2f_lo) & inr) inr = sqrt(exact_r2(tx, ty, k, s.gslot(m), P.Lx, P.rLx, P.Ly, P.rLy)) <= P.rho;
    float r = __builtin_amdgcn_sqrtf(r2);
    float w_in = __builtin_fmaf(-P.mu_p_over_xa2f, r2, mu_pv);
    // c4 (r - xi_a)(rho - r) in Horner form (two FMAs; the same FP32 error bound)
    float w_an = __builtin_fmaf(r, __builtin_fmaf(-P.c4f, r, wav), P.w_bf);
    float w = r2 <= P.xa2f ? w_in : w_an;
    w = inr ? w : 0.0f;
    if constexpr (MODE == SUM_KAHAN) {
      SW.add_prod(w, v.w);
      CW.add_prod(w, v.z);
    } else {
      group_add_prod<MODE>(psw, w, v.w);
      group_add_prod<MODE>(pcw, w, v.z);
    }
    n01 += inr;
    if constexpr (DEFER) {
      return r2 <= P.xr2f_hi;
    } else {
      if (r2 <= P.xr2f_hi) cohesion(s, m, dx, dy, r2, P);
      return false;
    }
  }
  // The cohesion term of a partner within xi_r or its band (r2 <= xr2f_hi):
  // decided as for rho.
  __device__ __forceinline__ void cohesion(const Src32 &s, int m, float dx, float dy, float r2, const KParams &P) {
    bool inc = r2 <= P.xr2f;
    if (r2 >= P.xr2f - P.band_xr2f) inc = sqrt(exact_r2(tx, ty, k, s.gslot(m), P.Lx, P.rLx, P.Ly, P.rLy)) <= P.xi_r;
    if (inc) {
      if (r2 >= P.close2f) {
        float ir = __builtin_amdgcn_rsqf(r2);
        Tx.add_prod(dx, ir);
        Ty.add_prod(dy, ir);
      } else {
        double ddx, ddy;
        exact_d(tx, ty, k, s.gslot(m), P, ddx, ddy);
        double dr = sqrt(ddx * ddx + ddy * ddy);
        if (dr > 0.0) {
          Tx.add((float)(ddx / dr));
          Ty.add((float)(ddy / dr));
        } else {
          Tx.add(1.0f);  // atan2(+0, +0) = 0
        }
      }
      ++n2;
    }
  }
  __device__ __forceinline__ void finish(const KParams &P, double phase, Pass1Out &o) const {
    double sp, cp;
    sincos(phase, &sp, &cp);
    o.phase = phase;
    o.has_cs = 1;  // the epilogue's velocity uses the same cos, sin
    o.cs = cp;
    o.sn = sp;
    o.n = n01;
    o.n2 = n2;
    double coh = n2 > 0 ? -P.kappa * (cp * Ty.value() - sp * Tx.value()) / (double)n2 : 0.0;
    if (n01 > 0) {
      o.aphi = (cp * SW.value() - sp * CW.value()) / (double)n01 + coh;
      o.written = 1;
      o.vortex_add = 0;
    } else {
      o.aphi = 0.0;
      o.written = 0;
      o.vortex_add = n2 > 0;
      o.coh = coh;
    }
  }
};

// =============================================================================
// Chimera pass 2 (Ito form, ChimeraSystem.cpp:589-644): over the same partners,
// cos(phi_j - phi_i - alpha) / n_i and (a_j - a_i) cos(...) / n_i (no sigma in
// these terms, as in the reference)
// =============================================================================
struct Pass2Out {
  double phase, a0, a1, ai;
};

// Per-particle pass-2 inputs: pass-1 phi derivative, partner counts (within
// rho; cohesion), and the additional-derivative values before the call.
struct Pass2In {
  double ai, n, n2, in0, in1;
};

template <int PREC>
struct Chim2Acc;

template <>
struct Chim2Acc<P64> {
  double xi, yi, pi, ai, n, a0, a1;
  __device__ __forceinline__ void init(const Src64 &s, int e, int, const Tiles &, const KParams &, const Pass2In &in) {
    xi = s.x[e];
    yi = s.y[e];
    pi = s.p[e];
    ai = in.ai;
    n = in.n;
    a0 = in.in0;
    a1 = in.in1;
  }
  __device__ __forceinline__ void row(int, int, int, const KParams &) {}
  __device__ __forceinline__ void pair(const Src64 &s, int m, const KParams &P) {
    double dx = min_image(s.x[m] - xi, P.Lx, P.rLx);
    double dy = min_image(s.y[m] - yi, P.Ly, P.rLy);
    double r2 = dx * dx + dy * dy;
    if ((r2 <= P.rho2) && (n != 0)) {
      double cph = cos(s.p[m] - pi - P.alpha);
      a0 += cph / n;
      a1 += (s.aj[m] - ai) * cph / n;
    }
  }
  __device__ __forceinline__ void finish(const KParams &P, double, Pass2Out &o) const {
    o.phase = pi;
    o.ai = ai;
    o.a0 = a0 * -sqrt(2.0 * P.D);
    o.a1 = a1 + -2.0 * P.D * ai;
  }
};

template <>
struct Chim2Acc<P64F> {
  double xi, yi, pi, ai, n, in0, in1, C, S, Pc, Qs;
  __device__ __forceinline__ void init(const Src64 &s, int e, int k, const Tiles &t, const KParams &, const Pass2In &in) {
    xi = s.x[e];
    yi = s.y[e];
    pi = t.p[k];
    ai = in.ai;
    n = in.n;
    in0 = in.in0;
    in1 = in.in1;
    C = S = Pc = Qs = 0.0;
  }
  __device__ __forceinline__ void row(int, int, int, const KParams &) {}
  __device__ __forceinline__ void pair(const Src64 &s, int m, const KParams &P) {
    double dx = min_image(s.x[m] - xi, P.Lx, P.rLx);
    double dy = min_image(s.y[m] - yi, P.Ly, P.rLy);
    double r2 = dx * dx + dy * dy;
    if (r2 <= P.rho2) {
      double aj = s.aj[m];
      C += s.c[m];
      S += s.s[m];
      Pc += aj * s.c[m];
      Qs += aj * s.s[m];
    }
  }
  __device__ __forceinline__ void finish(const KParams &P, double, Pass2Out &o) const {
    double sa, ca;
    sincos(pi + P.alpha, &sa, &ca);
    double a0 = in0, a1 = in1;
    if (n != 0) {
      double sc = ca * C + sa * S;
      a0 += sc / n;
      a1 += (ca * Pc + sa * Qs - ai * sc) / n;
    }
    o.phase = pi;
    o.ai = ai;
    o.a0 = a0 * -sqrt(2.0 * P.D);
    o.a1 = a1 + -2.0 * P.D * ai;
  }
};

template <>
struct Chim2Acc<P32K> {
  float xl, yl;
  ColOffsets co;
  double ai, n, in0, in1;
  Kahan C, S, Pc, Qs;
  float pc, ps, ppc, pqs;  // the current list group's partial sums
  int k;
  const double *tx, *ty;
  __device__ __forceinline__ void init(const Src32 &s, int e, int kk, const Tiles &t, const KParams &P, const Pass2In &in) {
    float4 v = s.f4[e];
    xl = v.x;
    yl = v.y;
    co.set(xl, P.hxf);
    co.set_codes(xl, yl, P.hxf, P.hyf);
    ai = in.ai;
    n = in.n;
    in0 = in.in0;
    in1 = in.in1;
    C.init();
    S.init();
    Pc.init();
    Qs.init();
    pc = ps = ppc = pqs = 0.0f;
    k = kk;
    tx = t.x;
    ty = t.y;
  }
  __device__ __forceinline__ void row(int dyo, int b1, int b2, const KParams &P) {
    co.oy = (float)dyo * P.hyf - yl;
    co.b1 = b1;
    co.b2 = b2;
  }
  template <int MODE = SUM_KAHAN>
  __device__ __forceinline__ void pair(const Src32 &s, int m, const KParams &P) {
    pair_x<MODE>(s, m, co.ox(m), P);
  }
  // pair() with the column offset known (a run of candidates in one stencil column)
  template <int MODE = SUM_KAHAN>
  __device__ __forceinline__ void pair_x(const Src32 &s, int m, float ox, const KParams &P) {
    float4 v = s.f4[m];
    float dx = v.x + ox, dy = v.y + co.oy;
    float r2 = __builtin_fmaf(dx, dx, dy * dy);
    bool in = r2 <= P.rho2f_hi;  // outside the band, the same as r2 <= rho2f
    if ((r2 >= P.rho2f_lo) & in) in = exact_r2(tx, ty, k, s.gslot(m), P.Lx, P.rLx, P.Ly, P.rLy) <= P.rho2;
    sum<MODE>(v, in, in ? s.aj[m] : 0.0f);
  }
  // the cell walks' partial sums over one stencil cell start at zero
  __device__ __forceinline__ void cell_begin() { pc = ps = ppc = pqs = 0.0f; }
  // a list entry with the partner's tile and pass-1 derivative already loaded
  template <int MODE>
  __device__ __forceinline__ void pair_c(const Src32 &s, int m, float fx, float fy, const float4 &v, float ajv,
                                         const KParams &P) {
    float dx = v.x + co.cdx(fx, P.hxf), dy = v.y - co.yf;
    float r2 = __builtin_fmaf(dx, dx, dy * dy);
    bool in = r2 <= P.rho2f_hi;  // outside the band, the same as r2 <= rho2f
    if ((r2 >= P.rho2f_lo) & in) in = exact_r2(tx, ty, k, s.gslot(m), P.Lx, P.rLx, P.Ly, P.rLy) <= P.rho2;
    sum<MODE>(v, in, ajv);  // aj multiplies cos, sin terms that are zero outside rho
  }
  // a list entry with pass 1's rho decision: v = (cos, sin, aj, 0) of the partner
  template <int MODE>
  __device__ __forceinline__ void pair_m(const float4 &v, bool in) {
    const float cz = in ? v.x : 0.0f, sw = in ? v.y : 0.0f;
    group_add<MODE>(pc, cz);
    group_add<MODE>(ps, sw);
    group_add_prod<MODE>(ppc, v.z, cz);
    group_add_prod<MODE>(pqs, v.z, sw);
  }
  __device__ __forceinline__ void flush() {
    C.add(pc);
    S.add(ps);
    Pc.add(ppc);
    Qs.add(pqs);
  }
  __device__ __forceinline__ void flush_cell() { flush(); }
  // Two consecutive cell candidates (tiles a, b read from t2, LDS index m =
  // t2 - s.f4), as
  // ChimAcc<P32K>::pair2_x: the in-range
  // flag f and f * aj enter as fma factors, the same values as pair_x<SUM_GROUP>
  // for each in turn (the halo's pass-1 derivatives are finite).
  __device__ __forceinline__ void pair2_x(const Src32 &s, const float4 *t2, int m, const float4 &a, const float4 &b,
                                          float oxa, float oxb, const KParams &P) {
    const float dxa = a.x + oxa, dya = a.y + co.oy, dxb = b.x + oxb, dyb = b.y + co.oy;
    const float r2a = __builtin_fmaf(dxa, dxa, dya * dya), r2b = __builtin_fmaf(dxb, dxb, dyb * dyb);
    bool ina = r2a <= P.rho2f_hi, inb = r2b <= P.rho2f_hi;
    const bool ba = (r2a >= P.rho2f_lo) & ina, bb = (r2b >= P.rho2f_lo) & inb;
    if (ba | bb) {
      if (ba) ina = exact_r2(tx, ty, k, s.gslot(m), P.Lx, P.rLx, P.Ly, P.rLy) <= P.rho2;
      if (bb) inb = exact_r2(tx, ty, k, s.gslot(m + 1), P.Lx, P.rLx, P.Ly, P.rLy) <= P.rho2;
    }
    const float fa = ina ? 1.0f : 0.0f, fb = inb ? 1.0f : 0.0f;
    const float ga = fa * s.aj[m], gb = fb * s.aj[m + 1];
    pc = __builtin_fmaf(fa, a.z, pc);
    ps = __builtin_fmaf(fa, a.w, ps);
    ppc = __builtin_fmaf(ga, a.z, ppc);
    pqs = __builtin_fmaf(ga, a.w, pqs);
    pc = __builtin_fmaf(fb, b.z, pc);
    ps = __builtin_fmaf(fb, b.w, ps);
    ppc = __builtin_fmaf(gb, b.z, ppc);
    pqs = __builtin_fmaf(gb, b.w, pqs);
  }
  template <int MODE = SUM_KAHAN>
  __device__ __forceinline__ void sum(const float4 &v, bool in, float aj) {
    float cz = in ? v.z : 0.0f, sw = in ? v.w : 0.0f;
    if constexpr (MODE == SUM_KAHAN) {
      C.add(cz);
      S.add(sw);
      Pc.add(aj * cz);
      Qs.add(aj * sw);
    } else {
      group_add<MODE>(pc, cz);
      group_add<MODE>(ps, sw);
      group_add_prod<MODE>(ppc, aj, cz);
      group_add_prod<MODE>(pqs, aj, sw);
    }
  }
  __device__ __forceinline__ void finish(const KParams &P, double phase, Pass2Out &o) const {
    double sa, ca;
    sincos(phase + P.alpha, &sa, &ca);
    double a0 = in0, a1 = in1;
    if (n != 0) {
      double sc = ca * C.value() + sa * S.value();
      a0 += sc / n;
      a1 += (ca * Pc.value() + sa * Qs.value() - ai * sc) / n;
    }
    o.phase = phase;
    o.ai = ai;
    o.a0 = a0 * -sqrt(2.0 * P.D);
    o.a1 = a1 + -2.0 * P.D * ai;
  }
};

// =============================================================================
// Vortex-arrays pass 2 (Ito form, VortexArraysSystemPtr.cpp:662-733): over the
// partners within rho, -W cos(dphi) / n01 and the drift derivative of
// W(r) sin(dphi); over the partners within xi_r, the cohesion counterparts;
// additional[0] is scaled by sqrt(2 D) at the end.  Velocities are
// v (cos phi, sin phi) of the stage phases, as the pass-1 derivative holds them.
// =============================================================================
template <int PREC>
struct Vort2Acc;

// FP64 reference form
template <>
struct Vort2Acc<P64> {
  double xi, yi, pi, vxi, vyi, ai, n1, n2, a0, a1;
  __device__ __forceinline__ void init(const Src64 &s, int e, int, const Tiles &, const KParams &P, const Pass2In &in) {
    xi = s.x[e];
    yi = s.y[e];
    pi = s.p[e];
    vxi = P.v0 * cos(pi);
    vyi = P.v0 * sin(pi);
    ai = in.ai;
    n1 = in.n;
    n2 = in.n2;
    a0 = in.in0;
    a1 = in.in1;
  }
  __device__ __forceinline__ void row(int, int, int, const KParams &) {}
  __device__ __forceinline__ void pair(const Src64 &s, int m, const KParams &P) {
    double dx = min_image(s.x[m] - xi, P.Lx, P.rLx);
    double dy = min_image(s.y[m] - yi, P.Ly, P.rLy);
    double r2 = dx * dx + dy * dy;
    if (r2 > P.outer2 * 1.000000000001) return;
    double dr = sqrt(r2);
    const bool al = (dr <= P.rho) && (n1 > 0.0), co = (dr <= P.xi_r) && (n2 > 0.0);
    if (!al && !co) return;
    const double pj = s.p[m];
    const double vxj = P.v0 * cos(pj), vyj = P.v0 * sin(pj);
    if (al) {
      double cph = cos(pj - pi);
      double sph = sin(pj - pi);
      double w = vortex_w(P, dr);
      double dw = vortex_dw(P, dr);
      a0 += -w * cph / n1;
      a1 += ((vxj - vxi) * dw * dx / dr + (vyj - vyi) * dw * dy / dr) * sph / n1 + (s.aj[m] - ai) * w * cph / n1 -
            2.0 * P.D * sph * w / n1;
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
  __device__ __forceinline__ void finish(const KParams &P, double, Pass2Out &o) const {
    o.phase = pi;
    o.ai = ai;
    o.a0 = a0 * sqrt(2.0 * P.D);
    o.a1 = a1;
  }
};

// Factored FP64 form: cos, sin of (phi_j - phi_i) and of (atan2(dy, dx) - phi_i)
// from the tile's cos phi, sin phi and dx / r, dy / r, no per-pair trig; the
// sums are divided by the counts once.
template <>
struct Vort2Acc<P64F> {
  double xi, yi, pi, ci, si, ai, n1, n2, in0, in1, A0, A1, C0, C1;
  __device__ __forceinline__ void init(const Src64 &s, int e, int k, const Tiles &t, const KParams &, const Pass2In &in) {
    xi = s.x[e];
    yi = s.y[e];
    pi = t.p[k];
    ci = s.c[e];
    si = s.s[e];
    ai = in.ai;
    n1 = in.n;
    n2 = in.n2;
    in0 = in.in0;
    in1 = in.in1;
    A0 = A1 = C0 = C1 = 0.0;
  }
  __device__ __forceinline__ void row(int, int, int, const KParams &) {}
  __device__ __forceinline__ void pair(const Src64 &s, int m, const KParams &P) {
    double dx = min_image(s.x[m] - xi, P.Lx, P.rLx);
    double dy = min_image(s.y[m] - yi, P.Ly, P.rLy);
    double r2 = dx * dx + dy * dy;
    if (r2 > P.outer2 * 1.000000000001) return;
    double dr = sqrt(r2);
    const bool al = dr <= P.rho, co = dr <= P.xi_r;
    if (!al && !co) return;
    const double cj = s.c[m], sj = s.s[m];
    const double dvx = P.v0 * cj - P.v0 * ci, dvy = P.v0 * sj - P.v0 * si;
    if (al) {
      double cph = cj * ci + sj * si;
      double sph = sj * ci - cj * si;
      double w = vortex_w(P, dr);
      double g = vortex_dw(P, dr) / dr;
      A0 += w * cph;
      A1 += (dvx * dx + dvy * dy) * g * sph + (s.aj[m] - ai) * w * cph - 2.0 * P.D * sph * w;
    }
    if (co) {
      // atan2(+0, +0) = 0 for a coincident partner
      double cpp = dr > 0.0 ? (dx * ci + dy * si) / dr : ci;
      double spp = dr > 0.0 ? (dy * ci - dx * si) / dr : -si;
      C0 += cpp;
      C1 += -P.kappa * ((dvy * dx - dvx * dy) / r2 - ai) * cpp + P.kappa * P.D * spp;
    }
  }
  __device__ __forceinline__ void finish(const KParams &P, double, Pass2Out &o) const {
    double a0 = in0, a1 = in1;
    if (n1 > 0.0) {
      a0 += -A0 / n1;
      a1 += A1 / n1;
    }
    if (n2 > 0.0) {
      a0 += P.kappa * C0 / n2;
      a1 += C1 / n2;
    }
    o.phase = pi;
    o.ai = ai;
    o.a0 = a0 * sqrt(2.0 * P.D);
    o.a1 = a1;
  }
};

// FP32 form over cell-local coordinates with Kahan sums; cutoff decisions
// (rho, xi_r, and the xi_a branch of W') within the band are taken in FP64.
template <>
struct Vort2Acc<P32K> {
  float xl, yl, cif, sif, aif;
  ColOffsets co;
  double ai, n1, n2, in0, in1, C0, C1;
  Kahan A0, A1;
  int k;
  const double *tx, *ty;
  __device__ __forceinline__ void init(const Src32 &s, int e, int kk, const Tiles &t, const KParams &P, const Pass2In &in) {
    float4 v = s.f4[e];
    xl = v.x;
    yl = v.y;
    cif = v.z;
    sif = v.w;
    co.set(xl, P.hxf);
    co.set_codes(xl, yl, P.hxf, P.hyf);
    ai = in.ai;
    aif = (float)in.ai;
    n1 = in.n;
    n2 = in.n2;
    in0 = in.in0;
    in1 = in.in1;
    A0.init();
    A1.init();
    C0 = C1 = 0.0;
    k = kk;
    tx = t.x;
    ty = t.y;
  }
  __device__ __forceinline__ void row(int dyo, int b1, int b2, const KParams &P) {
    co.oy = (float)dyo * P.hyf - yl;
    co.b1 = b1;
    co.b2 = b2;
  }
  // (few candidates reach the sums, which stay per term: MODE, cell_begin()
  // and flush() have nothing to do)
  template <int MODE = SUM_KAHAN>
  __device__ __forceinline__ void pair(const Src32 &s, int m, const KParams &P) { pair_x(s, m, co.ox(m), P); }
  // pair() with the column offset known (a run of candidates in one stencil column)
  template <int MODE = SUM_KAHAN>
  __device__ __forceinline__ void pair_x(const Src32 &s, int m, float ox, const KParams &P) {
    const float4 v = s.f4[m];
    pair_t(s, m, v, v.x + ox, v.y + co.oy, 0.0f, false, P);
  }
  __device__ __forceinline__ void cell_begin() {}
  // a list entry with the partner's tile and pass-1 derivative already loaded
  // (few entries reach the sums, which stay per entry: flush() has nothing to add)
  template <int MODE>
  __device__ __forceinline__ void pair_c(const Src32 &s, int m, float fx, float fy, const float4 &v, float ajv,
                                         const KParams &P) {
    pair_t(s, m, v, v.x + co.cdx(fx, P.hxf), v.y - co.yf, ajv, true, P);
  }
  __device__ __forceinline__ void flush() {}
  __device__ __forceinline__ void pair_t(const Src32 &s, int m, const float4 &v, float dx, float dy, float ajv,
                                         bool have_aj, const KParams &P) {
    float r2 = __builtin_fmaf(dx, dx, dy * dy);
    bool inr = r2 <= P.rho2f, inc = r2 <= P.xr2f, inner = r2 < P.xa2f;
    if ((__builtin_fabsf(r2 - P.rho2f) <= P.band_rho2f) | (__builtin_fabsf(r2 - P.xr2f) <= P.band_xr2f) |
        (__builtin_fabsf(r2 - P.xa2f) <= P.band_xa2f)) {
      double d = sqrt(exact_r2(tx, ty, k, s.gslot(m), P.Lx, P.rLx, P.Ly, P.rLy));
      inr = d <= P.rho;
      inc = d <= P.xi_r;
      inner = d < P.xi_a;
    }
    if (!inr && !inc) return;
    const float cj = v.z, sj = v.w;
    const float dvx = P.v0f * cj - P.v0f * cif, dvy = P.v0f * sj - P.v0f * sif;
    const float ir = r2 > 0.0f ? __builtin_amdgcn_rsqf(r2) : 0.0f;
    if (inr) {
      float r = r2 * ir;
      float cph = cj * cif + sj * sif;
      float sph = sj * cif - cj * sif;
      float w = inner ? __builtin_fmaf(-P.mu_p_over_xa2f, r2, P.mu_pf) : P.c4f * (r - P.xaf) * (P.rhof - r);
      // W'(r) / r
      float g = inner ? -2.0f * P.mu_p_over_xa2f : P.c4f * (P.rhof + P.xaf - 2.0f * r) * ir;
      float aj = have_aj ? ajv : s.aj[m];
      A0.add_prod(w, cph);
      A1.add((dvx * dx + dvy * dy) * g * sph + (aj - aif) * w * cph - 2.0f * P.Df * sph * w);
    }
    if (inc) {
      double cpp, spp, q;
      if (r2 >= P.close2f) {
        cpp = (dx * cif + dy * sif) * ir;
        spp = (dy * cif - dx * sif) * ir;
        q = (dvy * dx - dvx * dy) * ir * ir;
      } else {
        double ddx, ddy;
        exact_d(tx, ty, k, s.gslot(m), P, ddx, ddy);
        const double dr2 = ddx * ddx + ddy * ddy, dr = sqrt(dr2);
        const double cid = cif, sid = sif;
        cpp = dr > 0.0 ? (ddx * cid + ddy * sid) / dr : cid;
        spp = dr > 0.0 ? (ddy * cid - ddx * sid) / dr : -sid;
        q = ((double)dvy * ddx - (double)dvx * ddy) / dr2;
      }
      C0 += cpp;
      C1 += -P.kappa * (q - ai) * cpp + P.kappa * P.D * spp;
    }
  }
  __device__ __forceinline__ void finish(const KParams &P, double phase, Pass2Out &o) const {
    double a0 = in0, a1 = in1;
    if (n1 > 0.0) {
      a0 += -A0.value() / n1;
      a1 += A1.value() / n1;
    }
    if (n2 > 0.0) {
      a0 += P.kappa * C0 / n2;
      a1 += C1 / n2;
    }
    o.phase = phase;
    o.ai = ai;
    o.a0 = a0 * sqrt(2.0 * P.D);
    o.a1 = a1;
  }
};

// =============================================================================
// Epilogues
// =============================================================================
struct EpiCtx {
  double *ox, *oy, *op;           // k_next / derivative / k_4 by particle (phi in/out)
  const double *k1x, *k1y, *k1p;  // RK4 combine inputs by particle
  const double *k2x, *k2y, *k2p;
  const double *k3x, *k3y, *k3p;
  double *sx, *sy, *sp;           // master state by particle
  double *aphi_slot, *n_slot;     // pass-1 results by slot (pass-2 inputs)
  double *n2_slot;                // vortex cohesion partner counts by slot
  float *aphi_slot_f;
  double *a0, *a1;                // additional by particle (in/out)
  StepArgs sa;
  // Verlet passes: the next stage's tiles, written by the pass-1 epilogue
  // (next = 1: s + kcdt_n k of this launch; 2: the state this launch wrote)
  Tiles tn;
  double *gpn;
  double kcdt_n;
  int next;
  // Batch steps: the state-writing epilogue (SE, SRK) also stages and bins the
  // new state for the next step (stage arrays bgx/bgy/bgp by particle, its
  // cell, its rank among the cell's count); null bcell: a separate launch does
  // that at the next step.  bcoff: the replica's first cell.
  double *bgx, *bgy, *bgp;
  int *bcell, *bcount, *blocal;
  int bcoff;
};

// What a pass-1 epilogue computed for its particle: the velocity, the phi
// derivative and, for the state updates, the new state.
struct EpiState {
  double ax, ay, ap, x, y, p;
};

__device__ __forceinline__ void wrap_state(const KParams &P, double &x, double &y) {
  // the array form of the wrap, applied to every particle (PeriodicBoundaryConditions.cpp:133-141)
  x = x - floor(x * P.rLx) * P.Lx;
  y = y - floor(y * P.rLy) * P.Ly;
}

// The new state (x, y, p) of particle h staged for the next step as the
// batch's stage-and-bin launch stages it: the conditional wrap, the cell of the
// wrapped position (the replica's one cell on the all-pairs path) and a rank
// among that cell's particles from a wave-aggregated atomic on its count.
__device__ __forceinline__ void bin_next(const KParams &P, const EpiCtx &e, int h, double x, double y, double p) {
  int lc = 0;
  if (P.linked) {
    const Box b{P.Lx, P.Ly, P.rLx, P.rLy};
    cond_wrap(b, x, y);
    lc = cell_coord(y, b.Ly, P.ny) * P.nx + cell_coord(x, b.Lx, P.nx);
  }
  const int c = e.bcoff + lc;
  e.bcell[h] = c;
  e.bgx[h] = x;
  e.bgy[h] = y;
  e.bgp[h] = p;
  e.blocal[h] = wave_aggregated_rank(e.bcount, c, true);
}

// Final phi derivative given the in/out value (k_next or derivative entry
// before the call, *old, or 0 when old is null): the reference writes it only
// when partners exist, so the value is read only for a particle without them.
__device__ __forceinline__ double phi_out(const Pass1Out &o, const double *old) {
  if (o.written) return o.aphi;
  const double v = old ? *old : 0.0;
  if (o.vortex_add) return v + o.coh;
  return v;
}

template <int EPI>
__device__ __forceinline__ EpiState epilogue1(const KParams &P, const EpiCtx &e, int h, int id, int slot,
                                              const Pass1Out &o) {
  double sn, cs;
  if (o.has_cs) {
    sn = o.sn;
    cs = o.cs;
  } else {
    sincos(o.phase, &sn, &cs);
  }
  double ax = P.v0 * cs, ay = P.v0 * sn;
  EpiState r{ax, ay, 0.0, 0.0, 0.0, 0.0};
  if (EPI == EPI_K || EPI == EPI_ITO1 || EPI == EPI_RK4_FINAL) {
    double ap = phi_out(o, e.sa.zero_in || !e.op ? nullptr : e.op + h);
    r.ap = ap;
    // Null outputs are not written: k_4's x / y components (the combine below
    // consumes them), a fused RK4 step's k_1..k_3 x / y (rebuilt from the stage
    // phases by the combine) and an SRK step's derivative (read by nothing).
    if (EPI != EPI_RK4_FINAL && e.ox) {
      e.ox[h] = ax;
      e.oy[h] = ay;
    }
    if (e.op) e.op[h] = ap;
    if (EPI == EPI_ITO1) {
      e.aphi_slot[slot] = ap;
      e.aphi_slot_f[slot] = (float)ap;
      e.n_slot[slot] = (double)o.n;
      e.n2_slot[slot] = (double)o.n2;
    }
    if (EPI == EPI_RK4_FINAL) {
      // RungeKutta4StepperPtr.cpp:72-77: (k1 + 2 k2 + 2 k3 + k4) * dt / 6 for
      // x, y and (...) / 6 * dt for phi, evaluated left to right
      const double dt = e.sa.dt;
      double x = e.sx[h], y = e.sy[h], p = e.sp[h];
      const double k1p = e.k1p[h], k2p = e.k2p[h], k3p = e.k3p[h];
      double k1x, k1y, k2x, k2y, k3x, k3y;
      if (e.k1x) {
        k1x = e.k1x[h], k1y = e.k1y[h], k2x = e.k2x[h], k2y = e.k2y[h], k3x = e.k3x[h], k3y = e.k3y[h];
      } else {
        // k_i = v0 (cos, sin) of stage i's phase: the state's for stage 1,
        // p + (dt / 2) k_1 and p + (dt / 2) k_2 for stages 2 and 3, formed with
        // the operations and operands of the epilogues that wrote those stages
        // (contraction off), so that each k_i is the same bit for bit
        const double hdt = 0.5 * dt;
        double sn_i, cs_i;
        sincos(p, &sn_i, &cs_i);
        k1x = P.v0 * cs_i, k1y = P.v0 * sn_i;
        sincos(p + hdt * k1p, &sn_i, &cs_i);
        k2x = P.v0 * cs_i, k2y = P.v0 * sn_i;
        sincos(p + hdt * k2p, &sn_i, &cs_i);
        k3x = P.v0 * cs_i, k3y = P.v0 * sn_i;
      }
      x += (k1x + 2.0 * k2x + 2.0 * k3x + ax) * dt / 6.0;
      y += (k1y + 2.0 * k2y + 2.0 * k3y + ay) * dt / 6.0;
      p += (k1p + 2.0 * k2p + 2.0 * k3p + ap) / 6.0 * dt;
      if (e.sa.wrap) wrap_state(P, x, y);
      e.sx[h] = x;
      e.sy[h] = y;
      e.sp[h] = p;
      r.x = x;
      r.y = y;
      r.p = p;
    }
  } else if (EPI == EPI_SE) {
    // StochasticEulerStepperPtr.hpp:44-53 with a zero-filled derivative and the
    // natural frequency in the drift as (omega + a_phi) dt
    double ap = phi_out(o, nullptr);
    double xi;
    if (e.sa.noise) {
      xi = e.sa.noise[id];  // materialised noise is indexed by particle id
    } else {
      double z1;
      normal_pair(e.sa.seed, e.sa.step, kStreamStep, (uint32_t)id, xi, z1);
    }
    const double dt = e.sa.dt;
    double dW = e.sa.sqdt * xi;
    double w = e.sa.omega ? e.sa.omega[h] : 0.0;
    double x = e.sx[h], y = e.sy[h], p = e.sp[h];
    x += ax * dt;
    y += ay * dt;
    p += (w + ap) * dt + e.sa.sq2D * dW;
    if (e.sa.wrap) wrap_state(P, x, y);
    e.sx[h] = x;
    e.sy[h] = y;
    e.sp[h] = p;
    if (e.bcell) bin_next(P, e, h, x, y, p);
    r.ap = ap;
    r.x = x;
    r.y = y;
    r.p = p;
  }
  return r;
}

template <int EPI>
__device__ __forceinline__ EpiState epilogue2(const KParams &P, const EpiCtx &e, int h, int id, const Pass2Out &o) {
  EpiState r{0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  if (EPI == EPI_ITO2) {
    e.a0[h] = o.a0;
    e.a1[h] = o.a1;
  } else {
    // StochasticRungeKuttaStepperPtr.hpp:52-66
    double sn, cs;
    sincos(o.phase, &sn, &cs);
    const double ax = P.v0 * cs, ay = P.v0 * sn, ap = o.ai;
    double U_1, U_2;
    if (e.sa.noise) {
      U_1 = e.sa.noise[2 * id];
      U_2 = e.sa.noise[2 * id + 1];
    } else {
      normal_pair(e.sa.seed, e.sa.step, kStreamStep, (uint32_t)id, U_1, U_2);
    }
    const double dt = e.sa.dt, D = e.sa.D, sq2D = e.sa.sq2D;
    double delta_W = U_1 * e.sa.sqdt;
    double delta_Z = e.sa.h1 * (U_1 + U_2 / e.sa.sqrt3);
    double w = e.sa.omega ? e.sa.omega[h] : 0.0;
    double x = e.sx[h], y = e.sy[h], p = e.sp[h];
    x += ax * dt - sq2D * ay * delta_Z - 0.5 * (ap * ay + D * ax) * dt * dt;
    y += ay * dt + sq2D * ax * delta_Z + 0.5 * (ap * ax - D * ay) * dt * dt;
    p += (w + ap) * dt + sq2D * delta_W + o.a0 * delta_Z + 0.5 * o.a1 * dt * dt;
    if (e.sa.wrap) wrap_state(P, x, y);
    e.sx[h] = x;
    e.sy[h] = y;
    e.sp[h] = p;
    if (e.bcell) bin_next(P, e, h, x, y, p);
    r.x = x;
    r.y = y;
    r.p = p;
  }
  return r;
}

// Pair counters: workgroup totals added to one of kCountSlots slot triples
// (blockIdx modulo kCountSlots), so that the atomics of a launch spread over
// many addresses; the host sums the slots.  Called by every thread.
// count_add_lds keeps the workgroup's partials in LDS the caller provides;
// count_add declares its own.
__device__ __forceinline__ void count_add_lds(long long (*part)[kPairBlock / 64], unsigned long long *counts,
                                              long long a, long long b, long long c) {
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) {
    a += __shfl_xor(a, d, 64);
    b += __shfl_xor(b, d, 64);
    c += __shfl_xor(c, d, 64);
  }
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    part[0][w] = a;
    part[1][w] = b;
    part[2][w] = c;
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    long long v = 0;
#pragma unroll
    for (int q = 0; q < kPairBlock / 64; ++q) v += part[threadIdx.x][q];
    atomicAdd(&counts[3 * (blockIdx.x % kCountSlots) + threadIdx.x], (unsigned long long)v);
  }
}

__device__ __forceinline__ void count_add(unsigned long long *counts, long long a, long long b, long long c) {
  __shared__ long long part[3][kPairBlock / 64];
  count_add_lds(part, counts, a, b, c);
}


// Blocks are dealt to the 8 XCDs round robin; map them so that each XCD takes
// a contiguous run of row segments and its L2 sees neighbouring halo rows.
__device__ __forceinline__ int xcd_segment(int b, int nb) {
  const int per = nb / 8, rem = nb % 8, xcd = b % 8, idx = b / 8;
  return xcd < rem ? xcd * (per + 1) + idx : rem * (per + 1) + (xcd - rem) * per + idx;
}


inline EpiCtx make_ctx(Sim &s, double *ox, double *oy, double *op, const StepArgs &sa, double *const k1[3],
                double *const k2[3], double *const k3[3]) {
  EpiCtx e{};
  e.ox = ox;
  e.oy = oy;
  e.op = op;
  if (k1) {
    e.k1x = k1[0];
    e.k1y = k1[1];
    e.k1p = k1[2];
    e.k2x = k2[0];
    e.k2y = k2[1];
    e.k2p = k2[2];
    e.k3x = k3[0];
    e.k3y = k3[1];
    e.k3p = k3[2];
  }
  e.sx = s.sx;
  e.sy = s.sy;
  e.sp = s.sp;
  e.aphi_slot = s.aphi_slot;
  e.n_slot = s.n_slot;
  e.n2_slot = s.n2_slot;
  e.aphi_slot_f = s.aphi_slot_f;
  e.a0 = s.add[0];
  e.a1 = s.add[1];
  e.sa = sa;
  return e;
}

// Launches a pair kernel on the handle's stream; with profiling on, the HIP
// events are bound to the kernel's own dispatch (hipExtLaunchKernelGGL), so
// their elapsed time is the kernel's execution time.
template <typename... KArgs, typename... Args>
void launch_pair(Sim &s, void (*kern)(KArgs...), dim3 grid, dim3 blk, size_t sh, Args... args) {
  if (s.profiling) {
    hipEvent_t b, e;
    s.prof_take(b, e);
    hipExtLaunchKernelGGL(kern, grid, blk, sh, s.stream, b, e, 0, args...);
    s.prof_pending.emplace_back(b, e);
  } else {
    hipLaunchKernelGGL(kern, grid, blk, sh, s.stream, args...);
  }
}

// Lanes per particle of the FP32 chimera all-pairs pass (allpairs_group_lanes):
// a group of kApLanes consecutive lanes shares one particle, lane g summing
// partners j = g, g + kApLanes, ...; small all-pairs systems have too few
// particles to fill the device one lane each.
constexpr int kApLanes = 16;
template <int PREC, int SYS>
constexpr int allpairs_lanes() { return PREC == P32K && SYS == SPC2_CHIMERA ? kApLanes : 1; }
__host__ __device__ inline int allpairs_lanes(int prec, int sys) {
  return prec == P32K && sys == SPC2_CHIMERA ? kApLanes : 1;
}

// The FP32 chimera all-pairs pass 1 of particle i by its group's lane g (every
// lane of the group calls it): each lane's Kahan sums over its partners, in
// ascending j, then the lanes' FP64 values and counts added across the group in
// a fixed butterfly (the same total in every lane).  Geometry and the cos, sin
// of phi_j as allpairs_pass1 forms them.
__device__ __forceinline__ void allpairs_pass1_group(const KParams &P, int i, int g, const double *gx,
                                                     const double *gy, const double *gp, const float4 *tiles,
                                                     Pass1Out &o) {
  const double xi = gx[i], yi = gy[i], pi = gp[i];
  Kahan Sf, Cf;
  Sf.init();
  Cf.init();
  int n = 0;
  for (int j = g; j < P.n; j += kApLanes) {
    if (j == i) continue;
    double dx = min_image(gx[j] - xi, P.Lx, P.rLx);
    double dy = min_image(gy[j] - yi, P.Ly, P.rLy);
    double r2 = dx * dx + dy * dy;
    if (r2 <= P.rho2) {
      const float4 tj = tiles[j];
      Sf.add(tj.w);
      Cf.add(tj.z);
      ++n;
    }
  }
  double S = Sf.value(), C = Cf.value();
#pragma unroll
  for (int d = 1; d < kApLanes; d <<= 1) {
    S += __shfl_xor(S, d, kApLanes);
    C += __shfl_xor(C, d, kApLanes);
    n += __shfl_xor(n, d, kApLanes);
  }
  double sa, ca;
  sincos(pi + P.alpha, &sa, &ca);
  o.n = n;
  o.written = n != 0;
  o.aphi = n != 0 ? P.sigma * (ca * S - sa * C) / (double)n : 0.0;
  o.n2 = 0;
  o.vortex_add = 0;
  o.phase = pi;
}

// All pairs (rho > L/4): thread per particle, j = 0..N-1 ascending, stage
// positions not wrapped (ChimeraSystem.cpp:131-167, :184-254); arrays by
// particle id.
// tiles (FP32 form, optional): FP32 tiles by particle id whose cos and sin are
// (float) of the FP64 sincos of gp, read instead of forming them per pair (the
// same values).
template <int PREC, int SYS>
__device__ __forceinline__ void allpairs_pass1(const KParams &P, int i, const double *gx, const double *gy,
                                               const double *gp, Pass1Out &o, const float4 *tiles = nullptr) {
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
          if (PREC == P32K && tiles) {
            const float4 tj = tiles[j];
            Sf.add(tj.w);
            Cf.add(tj.z);
            ++n;
            continue;
          }
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


// All-pairs pass 2 (rho > L/4) of particle i, arrays by particle id (FP64 terms
// for every precision: the small-N path; the FP32 chimera form spreads them over
// a lane group, allpairs_pass2_group); in0, in1: the additional values before
// the call (zero for the SRK step).
template <int SYS>
__device__ __forceinline__ void allpairs_pass2(const KParams &P, int i, const double *gx, const double *gy,
                                               const double *gp, const double *aphi, const double *nslot,
                                               const double *n2slot, double in0, double in1, Pass2Out &o) {
  const double xi = gx[i], yi = gy[i], pi = gp[i];
  const double ai = aphi[i], n = nslot[i];
  double a0 = in0, a1 = in1;
  if (SYS == SPC2_VORTEX) {
    // VortexArraysSystemPtr.cpp:305-353
    const double n2 = n2slot[i];
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
        a1 += ((vxj - vxi) * dw * dx / dr + (vyj - vyi) * dw * dy / dr) * sph / n + (aphi[j] - ai) * w * cph / n -
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
    o.phase = pi;
    o.ai = ai;
    o.a0 = a0 * sqrt(2.0 * P.D);
    o.a1 = a1;
    return;
  }
  // ChimeraSystem.cpp:223-254
  for (int j = 0; j < P.n; ++j) {
    if (j == i) continue;
    double dx = min_image(gx[j] - xi, P.Lx, P.rLx);
    double dy = min_image(gy[j] - yi, P.Ly, P.rLy);
    double r2 = dx * dx + dy * dy;
    if ((r2 <= P.rho2) && (n != 0)) {
      double cph = cos(gp[j] - pi - P.alpha);
      a0 += cph / n;
      a1 += (aphi[j] - ai) * cph / n;
    }
  }
  o.phase = pi;
  o.ai = ai;
  o.a0 = a0 * -sqrt(2.0 * P.D);
  o.a1 = a1 + -2.0 * P.D * ai;
}

// Chimera all-pairs pass 2 of particle i by lane g of its group (the FP32
// form, allpairs_lanes lanes per particle; every lane of the group calls it):
// allpairs_pass2's FP64 terms, each lane's over its partners j = g, g + G, ...,
// added across the group in a fixed butterfly and then to in0, in1.
__device__ __forceinline__ void allpairs_pass2_group(const KParams &P, int i, int g, const double *gx,
                                                     const double *gy, const double *gp, const double *aphi,
                                                     const double *nslot, double in0, double in1, Pass2Out &o) {
  const double xi = gx[i], yi = gy[i], pi = gp[i];
  const double ai = aphi[i], n = nslot[i];
  double s0 = 0.0, s1 = 0.0;
  for (int j = g; j < P.n; j += kApLanes) {
    if (j == i) continue;
    double dx = min_image(gx[j] - xi, P.Lx, P.rLx);
    double dy = min_image(gy[j] - yi, P.Ly, P.rLy);
    double r2 = dx * dx + dy * dy;
    if ((r2 <= P.rho2) && (n != 0)) {
      double cph = cos(gp[j] - pi - P.alpha);
      s0 += cph / n;
      s1 += (aphi[j] - ai) * cph / n;
    }
  }
#pragma unroll
  for (int d = 1; d < kApLanes; d <<= 1) {
    s0 += __shfl_xor(s0, d, kApLanes);
    s1 += __shfl_xor(s1, d, kApLanes);
  }
  o.phase = pi;
  o.ai = ai;
  o.a0 = (in0 + s0) * -sqrt(2.0 * P.D);
  o.a1 = (in1 + s1) + -2.0 * P.D * ai;
}

}  // namespace
}  // namespace spc2
