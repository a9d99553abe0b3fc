// spc2 gfx950 kernels: linked-cell build (binning, wavefront scan, stable
// counting sort), periodic 3x3-stencil pair sums, fused steppers and the
// order-parameter reduction.  Compiled with -ffp-contract=off (see
// spc2_device.h).
#include <hip/hip_runtime.h>

#include "spc2_device.h"
#include "spc2_internal.h"

namespace spc2 {

namespace {

constexpr int kBlock = 256;
constexpr double kTwoPi = 6.283185307179586;  // 2.0 * M_PI

inline int blocks_for(long long n, int b = kBlock) { return (int)((n + b - 1) / b); }

// =============================================================================
// K1: stage state + binning + histogram
// =============================================================================
// Stage s + (k_coef * dt) * k (ChimeraSystem.cpp:274-276, evaluated as
// (k_coef * dt) * k_prev like the reference); on the linked-cell path the
// position is wrapped if outside the box and binned
// (ChimeraSystem.cpp:291-309).  `local` is the arrival rank inside the cell,
// used only to place the particle before the stable re-ranking.
__global__ __launch_bounds__(kBlock) void k_stage_bin(int n, int linked, Box b, int nx, int ny, const double *sx,
                                                      const double *sy, const double *sp, const double *kx,
                                                      const double *ky, const double *kp, double kcdt, double *gx,
                                                      double *gy, double *gp, int *cell_of, int *count, int *local) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double x = sx[i], y = sy[i], p = sp[i];
  if (kx) {
    x = x + kcdt * kx[i];
    y = y + kcdt * ky[i];
    p = p + kcdt * kp[i];
  }
  if (linked) {
    cond_wrap(b, x, y);
    int cx = cell_coord(x, b.Lx, nx);
    int cy = cell_coord(y, b.Ly, ny);
    int c = cy * nx + cx;
    cell_of[i] = c;
    local[i] = atomicAdd(&count[c], 1);
  }
  gx[i] = x;
  gy[i] = y;
  gp[i] = p;
}

// =============================================================================
// K2: exclusive scan of the cell histogram (one workgroup, wavefront scans)
// =============================================================================
// Each of the 1024 threads owns a contiguous run of cells; run totals are
// scanned inside each 64-lane wavefront with shuffles, wavefront totals across
// the 16 wavefronts through LDS.  The histogram is cleared for the next build.
__global__ __launch_bounds__(1024) void k_scan(int *count, int *start, int nc) {
  __shared__ int wtot[16];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int chunk = (nc + 1023) / 1024;
  const int lo = min(nc, t * chunk), hi = min(nc, lo + chunk);
  int s = 0;
  for (int c = lo; c < hi; ++c) s += count[c];
  int v = s;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    int u = __shfl_up(v, d, 64);
    if (lane >= d) v += u;
  }
  if (lane == 63) wtot[w] = v;
  __syncthreads();
  if (w == 0) {
    int x = lane < 16 ? wtot[lane] : 0;
#pragma unroll
    for (int d = 1; d < 16; d <<= 1) {
      int u = __shfl_up(x, d, 64);
      if (lane >= d) x += u;
    }
    if (lane < 16) wtot[lane] = x;  // inclusive over wavefronts
  }
  __syncthreads();
  int run = (v - s) + (w > 0 ? wtot[w - 1] : 0);
  for (int c = lo; c < hi; ++c) {
    int k = count[c];
    start[c] = run;
    run += k;
    count[c] = 0;
  }
  if (t == 1023) start[nc] = run;
}

// K3a: place each particle at cell start + arrival rank
__global__ __launch_bounds__(kBlock) void k_place(int n, const int *cell_of, const int *local, const int *start,
                                                  int *tmp) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  tmp[start[cell_of[i]] + local[i]] = i;
}

// K3b: stable counting sort.  The final position of particle i inside its cell
// is the number of cell members with a smaller index, which reproduces the
// reference's ascending push_back order (ChimeraSystem.cpp:335-338).  The same
// thread gathers i's stage data into the cell-sorted tile arrays.
template <int PREC>
__global__ __launch_bounds__(kBlock) void k_rank_tiles(int n, int nx, double hx, double hy, const int *tmp,
                                                       const int *cell_of, const int *start, const double *gx,
                                                       const double *gy, const double *gp, int *sorted, Tiles t) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  int i = tmp[k];
  int c = cell_of[i];
  int lo = start[c], hi = start[c + 1];
  int r = 0;
  for (int m = lo; m < hi; ++m) r += tmp[m] < i;
  int dst = lo + r;
  sorted[dst] = i;
  t.cell[dst] = c;
  double x = gx[i], y = gy[i], p = gp[i];
  if (PREC == P32K) {
    int cx = c % nx, cy = c / nx;
    double sn, cs;
    sincos(p, &sn, &cs);
    t.f4[dst] = make_float4((float)(x - (double)cx * hx), (float)(y - (double)cy * hy), (float)cs, (float)sn);
    t.x[dst] = x;
    t.y[dst] = y;
  } else {
    t.x[dst] = x;
    t.y[dst] = y;
    t.p[dst] = p;
    if (PREC == P64F) {
      double sn, cs;
      sincos(p, &sn, &cs);
      t.c[dst] = cs;
      t.s[dst] = sn;
    }
  }
}

// =============================================================================
// Pair accumulators.  Each visits i's candidates j != i in the reference order
// (stencil cells (-1,-1), (0,-1), ..., (1,1), ascending index inside a cell).
// =============================================================================

// AlignmentStrength, VortexArraysSystemPtr.cpp:737-749
__device__ __forceinline__ double vortex_w(const KParams &P, double r) {
  if (r < P.xi_a) return P.mu_plus * (1.0 - r * r / (P.xi_a * P.xi_a));
  if (r < P.rho) return -P.mu_minus * 4.0 * (r - P.xi_a) * (P.rho - r) / (P.rho - P.xi_a) / (P.rho - P.xi_a);
  return 0.0;
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
};

// ---- chimera, FP64 reference form (ChimeraSystem.cpp:391-424) ----------------
template <int PREC>
struct ChimAcc;

template <>
struct ChimAcc<P64> {
  double xi, yi, pi, A;
  int n;
  __device__ void init(const Tiles &t, int k, const KParams &) {
    xi = t.x[k];
    yi = t.y[k];
    pi = t.p[k];
    A = 0.0;
    n = 0;
  }
  __device__ void cell(int, int, const KParams &) {}
  __device__ void pair(const Tiles &t, int m, const KParams &P) {
    double dx = min_image(t.x[m] - xi, P.Lx, P.rLx);
    double dy = min_image(t.y[m] - yi, P.Ly, P.rLy);
    double r2 = dx * dx + dy * dy;
    if (r2 <= P.rho2) {
      A += sin(t.p[m] - pi - P.alpha);
      ++n;
    }
  }
  __device__ void finish(const KParams &P, double, Pass1Out &o) const {
    o.phase = pi;
    o.n = n;
    o.n2 = 0;
    o.written = n != 0;
    o.vortex_add = 0;
    o.aphi = n != 0 ? P.sigma * A / (double)n : 0.0;
  }
};

// ---- chimera, FP64 factored: sum sin(phi_j - phi_i - alpha)
//      = cos(phi_i + alpha) sum sin phi_j - sin(phi_i + alpha) sum cos phi_j ----
template <>
struct ChimAcc<P64F> {
  double xi, yi, pi, S, C;
  int n;
  __device__ void init(const Tiles &t, int k, const KParams &) {
    xi = t.x[k];
    yi = t.y[k];
    pi = t.p[k];
    S = 0.0;
    C = 0.0;
    n = 0;
  }
  __device__ void cell(int, int, const KParams &) {}
  __device__ void pair(const Tiles &t, int m, const KParams &P) {
    double dx = min_image(t.x[m] - xi, P.Lx, P.rLx);
    double dy = min_image(t.y[m] - yi, P.Ly, P.rLy);
    double r2 = dx * dx + dy * dy;
    if (r2 <= P.rho2) {
      S += t.s[m];
      C += t.c[m];
      ++n;
    }
  }
  __device__ void finish(const KParams &P, double, Pass1Out &o) const {
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

// FP64 distance^2 of slots k and m with the reference min-image: decides the
// rare FP32 pairs that fall inside the cutoff band.
__device__ __noinline__ double exact_r2(const Tiles &t, int k, int m, const KParams &P) {
  double dx = min_image(t.x[m] - t.x[k], P.Lx, P.rLx);
  double dy = min_image(t.y[m] - t.y[k], P.Ly, P.rLy);
  return dx * dx + dy * dy;
}

// ---- chimera, FP32 + Kahan over cell-local coordinates -------------------------
template <>
struct ChimAcc<P32K> {
  float xl, yl, offx, offy;
  Kahan S, C;
  int n, k;
  __device__ void init(const Tiles &t, int kk, const KParams &) {
    float4 v = t.f4[kk];
    xl = v.x;
    yl = v.y;
    S.init();
    C.init();
    n = 0;
    k = kk;
  }
  __device__ void cell(int dxo, int dyo, const KParams &P) {
    offx = (float)dxo * P.hxf - xl;
    offy = (float)dyo * P.hyf - yl;
  }
  __device__ void pair(const Tiles &t, int m, const KParams &P) {
    float4 v = t.f4[m];
    float dx = v.x + offx, dy = v.y + offy;
    float r2 = __builtin_fmaf(dx, dx, dy * dy);
    bool in = r2 <= P.rho2f;
    if (__builtin_fabsf(r2 - P.rho2f) <= P.band_rho2f) in = exact_r2(t, k, m, P) <= P.rho2;
    if (in) {
      S.add(v.w);
      C.add(v.z);
      ++n;
    }
  }
  __device__ void finish(const KParams &P, double phase, Pass1Out &o) const {
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

// ---- vortex arrays, FP64 reference form (VortexArraysSystemPtr.cpp:459-504) ----
template <int PREC>
struct VortAcc;

template <>
struct VortAcc<P64> {
  double xi, yi, pi, A, B, Cc;
  int n0, n1, n2;
  __device__ void init(const Tiles &t, int k, const KParams &) {
    xi = t.x[k];
    yi = t.y[k];
    pi = t.p[k];
    A = B = Cc = 0.0;
    n0 = n1 = n2 = 0;
  }
  __device__ void cell(int, int, const KParams &) {}
  __device__ void pair(const Tiles &t, int m, const KParams &P) {
    double dx = min_image(t.x[m] - xi, P.Lx, P.rLx);
    double dy = min_image(t.y[m] - yi, P.Ly, P.rLy);
    double r2 = dx * dx + dy * dy;
    if (r2 > P.outer2 * 1.000000000001) return;  // cannot pass the sqrt tests below
    double dr = sqrt(r2);
    double pj = t.p[m];
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
  __device__ void finish(const KParams &, double, Pass1Out &o) const {
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

// ---- vortex arrays, factored: sum W sin(phi_j - phi_i) = cos phi_i sum W sin phi_j
//      - sin phi_i sum W cos phi_j; sin(atan2(dy, dx) - phi_i)
//      = (dy cos phi_i - dx sin phi_i) / r ------------------------------------------
template <>
struct VortAcc<P64F> {
  double xi, yi, pi, SW, CW, Tx, Ty;
  int n01, n2;
  __device__ void init(const Tiles &t, int k, const KParams &) {
    xi = t.x[k];
    yi = t.y[k];
    pi = t.p[k];
    SW = CW = Tx = Ty = 0.0;
    n01 = n2 = 0;
  }
  __device__ void cell(int, int, const KParams &) {}
  __device__ void pair(const Tiles &t, int m, const KParams &P) {
    double dx = min_image(t.x[m] - xi, P.Lx, P.rLx);
    double dy = min_image(t.y[m] - yi, P.Ly, P.rLy);
    double r2 = dx * dx + dy * dy;
    if (r2 > P.outer2 * 1.000000000001) return;
    double dr = sqrt(r2);
    if (dr <= P.rho) {
      double w = vortex_w(P, dr);
      SW += w * t.s[m];
      CW += w * t.c[m];
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
  __device__ void finish(const KParams &P, double, Pass1Out &o) const {
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
  float xl, yl, offx, offy;
  Kahan SW, CW, Tx, Ty;
  int n01, n2, k;
  __device__ void init(const Tiles &t, int kk, const KParams &) {
    float4 v = t.f4[kk];
    xl = v.x;
    yl = v.y;
    SW.init();
    CW.init();
    Tx.init();
    Ty.init();
    n01 = n2 = 0;
    k = kk;
  }
  __device__ void cell(int dxo, int dyo, const KParams &P) {
    offx = (float)dxo * P.hxf - xl;
    offy = (float)dyo * P.hyf - yl;
  }
  __device__ void pair(const Tiles &t, int m, const KParams &P) {
    float4 v = t.f4[m];
    float dx = v.x + offx, dy = v.y + offy;
    float r2 = __builtin_fmaf(dx, dx, dy * dy);
    if (r2 <= P.outer2f * 1.00002f) {
      bool inr = r2 <= P.rho2f, inc = r2 <= P.xr2f;
      if (__builtin_fabsf(r2 - P.rho2f) <= P.band_rho2f || __builtin_fabsf(r2 - P.xr2f) <= P.band_xr2f) {
        // the reference tests sqrt(r^2) <= cutoff in FP64 (VortexArraysSystemPtr.cpp:473-488)
        double d = sqrt(exact_r2(t, k, m, P));
        inr = d <= P.rho;
        inc = d <= P.xi_r;
      }
      if (inr) {
        float w;
        if (r2 <= P.xa2f) {
          w = __builtin_fmaf(-P.mu_p_over_xa2f, r2, P.mu_pf);
        } else {
          float r = __builtin_sqrtf(r2);
          w = P.c4f * (r - P.xaf) * (P.rhof - r);
        }
        SW.add(w * v.w);
        CW.add(w * v.z);
        ++n01;
      }
      if (inc) {
        if (r2 > 0.0f) {
          float ir = 1.0f / __builtin_sqrtf(r2);
          Tx.add(dx * ir);
          Ty.add(dy * ir);
        } else {
          Tx.add(1.0f);
        }
        ++n2;
      }
    }
  }
  __device__ void finish(const KParams &P, double phase, Pass1Out &o) const {
    double sp, cp;
    sincos(phase, &sp, &cp);
    o.phase = phase;
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
// Epilogues
// =============================================================================

struct EpiCtx {
  double *ox, *oy, *op;          // k_next / derivative / k_4 by particle (phi in/out)
  const double *k1x, *k1y, *k1p;  // RK4 combine inputs by particle
  const double *k2x, *k2y, *k2p;
  const double *k3x, *k3y, *k3p;
  double *sx, *sy, *sp;          // master state by particle
  double *aphi_slot, *n_slot;    // pass-1 results by slot (pass 2 inputs)
  float *aphi_slot_f;
  double *a0, *a1;               // additional by particle (in/out)
  StepArgs sa;
};

__device__ __forceinline__ void wrap_state(const KParams &P, double &x, double &y) {
  // array form of the wrap, applied to every particle (PeriodicBoundaryConditions.cpp:133-141)
  x = x - floor(x * P.rLx) * P.Lx;
  y = y - floor(y * P.rLy) * P.Ly;
}

// Final phi derivative given the in/out value `old` (k_next or derivative
// entry before the call): the reference writes it only when partners exist.
__device__ __forceinline__ double phi_out(const Pass1Out &o, double old) {
  if (o.written) return o.aphi;
  if (o.vortex_add) return old + o.coh;
  return old;
}

template <int EPI>
__device__ __forceinline__ void epilogue1(const KParams &P, const EpiCtx &e, int id, int slot, const Pass1Out &o) {
  double sn, cs;
  sincos(o.phase, &sn, &cs);
  double ax = P.v0 * cs, ay = P.v0 * sn;
  if (EPI == EPI_K || EPI == EPI_ITO1 || EPI == EPI_RK4_FINAL) {
    double ap = phi_out(o, e.sa.zero_in ? 0.0 : e.op[id]);
    e.ox[id] = ax;
    e.oy[id] = ay;
    e.op[id] = ap;
    if (EPI == EPI_ITO1) {
      e.aphi_slot[slot] = ap;
      e.aphi_slot_f[slot] = (float)ap;
      e.n_slot[slot] = (double)o.n;
    }
    if (EPI == EPI_RK4_FINAL) {
      // RungeKutta4StepperPtr.cpp:72-77: (k1 + 2 k2 + 2 k3 + k4) * dt / 6 for
      // x, y and (...) / 6 * dt for phi, evaluated left to right
      const double dt = e.sa.dt;
      double x = e.sx[id], y = e.sy[id], p = e.sp[id];
      x += (e.k1x[id] + 2.0 * e.k2x[id] + 2.0 * e.k3x[id] + ax) * dt / 6.0;
      y += (e.k1y[id] + 2.0 * e.k2y[id] + 2.0 * e.k3y[id] + ay) * dt / 6.0;
      p += (e.k1p[id] + 2.0 * e.k2p[id] + 2.0 * e.k3p[id] + ap) / 6.0 * dt;
      wrap_state(P, x, y);
      e.sx[id] = x;
      e.sy[id] = y;
      e.sp[id] = p;
    }
  } else if (EPI == EPI_SE) {
    // StochasticEulerStepperPtr.hpp:44-53 with a zero-filled derivative and the
    // natural frequency in the drift as (omega + a_phi) dt
    double ap = phi_out(o, 0.0);
    double xi;
    if (e.sa.noise) {
      xi = e.sa.noise[id];
    } else {
      double z1;
      normal_pair(e.sa.seed, e.sa.step, kStreamStep, (uint32_t)id, xi, z1);
    }
    const double dt = e.sa.dt;
    double dW = e.sa.sqdt * xi;
    double w = e.sa.omega ? e.sa.omega[id] : 0.0;
    double x = e.sx[id], y = e.sy[id], p = e.sp[id];
    x += ax * dt;
    y += ay * dt;
    p += (w + ap) * dt + e.sa.sq2D * dW;
    wrap_state(P, x, y);
    e.sx[id] = x;
    e.sy[id] = y;
    e.sp[id] = p;
  }
}

// =============================================================================
// Pass 1 over the cell list: thread per slot (cell-sorted order)
// =============================================================================

__device__ __forceinline__ void count_add(unsigned long long *counts, long long a, long long b, long long c) {
  // wavefront reduction, one atomic per wavefront
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) {
    a += __shfl_xor(a, d, 64);
    b += __shfl_xor(b, d, 64);
    c += __shfl_xor(c, d, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    atomicAdd(&counts[0], (unsigned long long)a);
    atomicAdd(&counts[1], (unsigned long long)b);
    atomicAdd(&counts[2], (unsigned long long)c);
  }
}

template <class Acc, int PREC, int EPI>
__global__ __launch_bounds__(kBlock) void k_pass1_cells(KParams P, const int *__restrict__ start, Tiles t,
                                                        const int *__restrict__ sorted, const double *gp, EpiCtx e,
                                                        unsigned long long *counts) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const bool active = k < P.n;
  long long cand = 0;
  Pass1Out o{};
  if (active) {
    Acc acc;
    acc.init(t, k, P);
    const int c = t.cell[k];
    const int cx = c % P.nx, cy = c / P.nx;
    for (int dyo = -1; dyo <= 1; ++dyo) {
      const int row = adjust_cell(cy + dyo, P.ny) * P.nx;
      for (int dxo = -1; dxo <= 1; ++dxo) {
        const int jc = row + adjust_cell(cx + dxo, P.nx);
        const int m0 = start[jc], m1 = start[jc + 1];
        acc.cell(dxo, dyo, P);
        cand += m1 - m0;
        for (int m = m0; m < m1; ++m)
          if (m != k) acc.pair(t, m, P);
      }
    }
    cand -= 1;
    const int id = sorted[k];
    acc.finish(P, PREC == P32K ? gp[id] : 0.0, o);
    epilogue1<EPI>(P, e, id, k, o);
  }
  if (counts) count_add(counts, active ? cand : 0, active ? o.n : 0, active ? o.n2 : 0);
}

// =============================================================================
// Pass 1 over all pairs (rho > L/4): thread per particle, j = 0..N-1, stage
// positions not wrapped (ChimeraSystem.cpp:131-167)
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
      // factored form; FP32 variant accumulates in Kahan FP32 over FP64 geometry
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
    epilogue1<EPI>(P, e, i, i, o);
  }
  if (counts) count_add(counts, active ? P.n - 1 : 0, active ? o.n : 0, active ? o.n2 : 0);
}

// =============================================================================
// Pass 2 (chimera Ito form, ChimeraSystem.cpp:589-644): sums over the same
// partners of cos(phi_j - phi_i - alpha) / n_i and (a_j - a_i) cos(...) / n_i
// (no sigma in these terms, as in the reference)
// =============================================================================

struct Pass2Out {
  double phase, a0, a1, ai;
};

template <int PREC>
struct Chim2Acc;

template <>
struct Chim2Acc<P64> {
  double xi, yi, pi, ai, n, a0, a1;
  __device__ void init(const Tiles &t, int k, const EpiCtx &e, double in0, double in1) {
    xi = t.x[k];
    yi = t.y[k];
    pi = t.p[k];
    ai = e.aphi_slot[k];
    n = e.n_slot[k];
    a0 = in0;
    a1 = in1;
  }
  __device__ void cell(int, int, const KParams &) {}
  __device__ void pair(const Tiles &t, const EpiCtx &e, int m, const KParams &P) {
    double dx = min_image(t.x[m] - xi, P.Lx, P.rLx);
    double dy = min_image(t.y[m] - yi, P.Ly, P.rLy);
    double r2 = dx * dx + dy * dy;
    if ((r2 <= P.rho2) && (n != 0)) {
      double cph = cos(t.p[m] - pi - P.alpha);
      a0 += cph / n;
      a1 += (e.aphi_slot[m] - ai) * cph / n;
    }
  }
  __device__ void finish(const KParams &P, double, Pass2Out &o) const {
    o.phase = pi;
    o.ai = ai;
    o.a0 = a0 * -sqrt(2.0 * P.D);
    o.a1 = a1 + -2.0 * P.D * ai;
  }
};

template <>
struct Chim2Acc<P64F> {
  double xi, yi, pi, ai, n, in0, in1, C, S, Pc, Qs;
  __device__ void init(const Tiles &t, int k, const EpiCtx &e, double i0, double i1) {
    xi = t.x[k];
    yi = t.y[k];
    pi = t.p[k];
    ai = e.aphi_slot[k];
    n = e.n_slot[k];
    in0 = i0;
    in1 = i1;
    C = S = Pc = Qs = 0.0;
  }
  __device__ void cell(int, int, const KParams &) {}
  __device__ void pair(const Tiles &t, const EpiCtx &e, int m, const KParams &P) {
    double dx = min_image(t.x[m] - xi, P.Lx, P.rLx);
    double dy = min_image(t.y[m] - yi, P.Ly, P.rLy);
    double r2 = dx * dx + dy * dy;
    if (r2 <= P.rho2) {
      double aj = e.aphi_slot[m];
      C += t.c[m];
      S += t.s[m];
      Pc += aj * t.c[m];
      Qs += aj * t.s[m];
    }
  }
  __device__ void finish(const KParams &P, double, Pass2Out &o) const {
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
  float xl, yl, offx, offy;
  double ai, n, in0, in1;
  Kahan C, S, Pc, Qs;
  int k;
  __device__ void init(const Tiles &t, int kk, const EpiCtx &e, double i0, double i1) {
    float4 v = t.f4[kk];
    xl = v.x;
    yl = v.y;
    ai = e.aphi_slot[kk];
    n = e.n_slot[kk];
    in0 = i0;
    in1 = i1;
    C.init();
    S.init();
    Pc.init();
    Qs.init();
    k = kk;
  }
  __device__ void cell(int dxo, int dyo, const KParams &P) {
    offx = (float)dxo * P.hxf - xl;
    offy = (float)dyo * P.hyf - yl;
  }
  __device__ void pair(const Tiles &t, const EpiCtx &e, int m, const KParams &P) {
    float4 v = t.f4[m];
    float dx = v.x + offx, dy = v.y + offy;
    float r2 = __builtin_fmaf(dx, dx, dy * dy);
    bool in = r2 <= P.rho2f;
    if (__builtin_fabsf(r2 - P.rho2f) <= P.band_rho2f) in = exact_r2(t, k, m, P) <= P.rho2;
    if (in) {
      float aj = e.aphi_slot_f[m];
      C.add(v.z);
      S.add(v.w);
      Pc.add(aj * v.z);
      Qs.add(aj * v.w);
    }
  }
  __device__ void finish(const KParams &P, double phase, Pass2Out &o) const {
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

template <int EPI>
__device__ __forceinline__ void epilogue2(const KParams &P, const EpiCtx &e, int id, const Pass2Out &o) {
  if (EPI == EPI_ITO2) {
    e.a0[id] = o.a0;
    e.a1[id] = o.a1;
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
    double w = e.sa.omega ? e.sa.omega[id] : 0.0;
    double x = e.sx[id], y = e.sy[id], p = e.sp[id];
    x += ax * dt - sq2D * ay * delta_Z - 0.5 * (ap * ay + D * ax) * dt * dt;
    y += ay * dt + sq2D * ax * delta_Z + 0.5 * (ap * ax - D * ay) * dt * dt;
    p += (w + ap) * dt + sq2D * delta_W + o.a0 * delta_Z + 0.5 * o.a1 * dt * dt;
    wrap_state(P, x, y);
    e.sx[id] = x;
    e.sy[id] = y;
    e.sp[id] = p;
  }
}

template <int PREC, int EPI>
__global__ __launch_bounds__(kBlock) void k_pass2_cells(KParams P, const int *__restrict__ start, Tiles t,
                                                        const int *__restrict__ sorted, const double *gp, EpiCtx e) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= P.n) return;
  const int id = sorted[k];
  double in0 = 0.0, in1 = 0.0;
  if (EPI == EPI_ITO2) {
    in0 = e.a0[id];
    in1 = e.a1[id];
  }
  Chim2Acc<PREC> acc;
  acc.init(t, k, e, in0, in1);
  const int c = t.cell[k];
  const int cx = c % P.nx, cy = c / P.nx;
  for (int dyo = -1; dyo <= 1; ++dyo) {
    const int row = adjust_cell(cy + dyo, P.ny) * P.nx;
    for (int dxo = -1; dxo <= 1; ++dxo) {
      const int jc = row + adjust_cell(cx + dxo, P.nx);
      const int m0 = start[jc], m1 = start[jc + 1];
      acc.cell(dxo, dyo, P);
      for (int m = m0; m < m1; ++m)
        if (m != k) acc.pair(t, e, m, P);
    }
  }
  Pass2Out o;
  acc.finish(P, PREC == P32K ? gp[id] : 0.0, o);
  epilogue2<EPI>(P, e, id, o);
}

template <int PREC, int EPI>
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
  epilogue2<EPI>(P, e, i, o);
}

// =============================================================================
// Order parameter (two-level FP64 reduction, deterministic)
// =============================================================================

constexpr int kRedBlocks = 1024;

__global__ __launch_bounds__(kBlock) void k_order_partial(int n, const double *sp, double *red) {
  __shared__ double sc[kBlock], ss[kBlock];
  double c = 0.0, s = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    double sn, cs;
    sincos(sp[i], &sn, &cs);
    c += cs;
    s += sn;
  }
  sc[threadIdx.x] = c;
  ss[threadIdx.x] = s;
  __syncthreads();
  for (int d = kBlock / 2; d > 0; d >>= 1) {
    if ((int)threadIdx.x < d) {
      sc[threadIdx.x] += sc[threadIdx.x + d];
      ss[threadIdx.x] += ss[threadIdx.x + d];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    red[2 * blockIdx.x] = sc[0];
    red[2 * blockIdx.x + 1] = ss[0];
  }
}

__global__ __launch_bounds__(kBlock) void k_order_final(int nb, const double *red, double *out2) {
  __shared__ double sc[kBlock], ss[kBlock];
  double c = 0.0, s = 0.0;
  for (int b = threadIdx.x; b < nb; b += blockDim.x) {
    c += red[2 * b];
    s += red[2 * b + 1];
  }
  sc[threadIdx.x] = c;
  ss[threadIdx.x] = s;
  __syncthreads();
  for (int d = kBlock / 2; d > 0; d >>= 1) {
    if ((int)threadIdx.x < d) {
      sc[threadIdx.x] += sc[threadIdx.x + d];
      ss[threadIdx.x] += ss[threadIdx.x + d];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out2[0] = sc[0];
    out2[1] = ss[0];
  }
}

// =============================================================================
// Initial state, materialised noise, fill
// =============================================================================

__global__ __launch_bounds__(kBlock) void k_init_state(int n, double Lx, double Ly, uint64_t seed, double omega_std,
                                                       double *sx, double *sy, double *sp, double *omega) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double ux, uy, up, unused;
  uniform_pair(seed, 0, kStreamInitPos, (uint32_t)i, ux, uy);
  uniform_pair(seed, 0, kStreamInitPhase, (uint32_t)i, up, unused);
  sx[i] = ux * Lx;
  sy[i] = uy * Ly;
  sp[i] = up * kTwoPi;
  double z0, z1;
  normal_pair(seed, 0, kStreamOmega, (uint32_t)i, z0, z1);
  omega[i] = z0 * omega_std;
}

__global__ __launch_bounds__(kBlock) void k_noise(int n, int per_particle, uint64_t seed, uint64_t step,
                                                  double *out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double z0, z1;
  normal_pair(seed, step, kStreamStep, (uint32_t)i, z0, z1);
  if (per_particle == 1) {
    out[i] = z0;
  } else {
    out[2 * i] = z0;
    out[2 * i + 1] = z1;
  }
}

__global__ __launch_bounds__(kBlock) void k_fill(double *p, double v, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// =============================================================================
// dispatch helpers
// =============================================================================

template <int PREC>
struct ChimSel {
  using type = ChimAcc<PREC>;
};

EpiCtx make_ctx(Sim &s, double *ox, double *oy, double *op, const StepArgs &sa, double *const k1[3],
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
  e.aphi_slot_f = s.aphi_slot_f;
  e.a0 = s.add[0];
  e.a1 = s.add[1];
  e.sa = sa;
  return e;
}

template <int PREC, int SYS, int EPI>
void pass1_typed(Sim &s, const EpiCtx &e) {
  unsigned long long *cnt = s.counting ? s.counts : nullptr;
  if (s.linked) {
    if (SYS == SPC2_CHIMERA)
      hipLaunchKernelGGL((k_pass1_cells<ChimAcc<PREC>, PREC, EPI>), dim3(blocks_for(s.n)), dim3(kBlock), 0, s.stream,
                         s.kp, s.start, s.t, s.sorted, s.gp, e, cnt);
    else
      hipLaunchKernelGGL((k_pass1_cells<VortAcc<PREC>, PREC, EPI>), dim3(blocks_for(s.n)), dim3(kBlock), 0, s.stream,
                         s.kp, s.start, s.t, s.sorted, s.gp, e, cnt);
  } else {
    hipLaunchKernelGGL((k_pass1_allpairs<PREC, SYS, EPI>), dim3(blocks_for(s.n)), dim3(kBlock), 0, s.stream, s.kp,
                       s.gx, s.gy, s.gp, e, cnt);
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

template <int PREC, int EPI>
void pass2_typed(Sim &s, const EpiCtx &e) {
  if (s.linked)
    hipLaunchKernelGGL((k_pass2_cells<PREC, EPI>), dim3(blocks_for(s.n)), dim3(kBlock), 0, s.stream, s.kp, s.start,
                       s.t, s.sorted, s.gp, e);
  else
    hipLaunchKernelGGL((k_pass2_allpairs<PREC, EPI>), dim3(blocks_for(s.n)), dim3(kBlock), 0, s.stream, s.kp, s.gx,
                       s.gy, s.gp, e);
}

template <int PREC>
void pass2_epi(Sim &s, int epi, const EpiCtx &e) {
  if (epi == EPI_ITO2)
    pass2_typed<PREC, EPI_ITO2>(s, e);
  else
    pass2_typed<PREC, EPI_SRK>(s, e);
}

struct ProfScope {
  Sim &s;
  explicit ProfScope(Sim &sim) : s(sim) {
    if (s.profiling) (void)hipEventRecord(s.ev0, s.stream);
  }
  ~ProfScope() {
    if (s.profiling) {
      (void)hipEventRecord(s.ev1, s.stream);
      (void)hipEventSynchronize(s.ev1);
      float ms = 0.0f;
      (void)hipEventElapsedTime(&ms, s.ev0, s.ev1);
      s.prof_ms += ms;
      s.prof_launches += 1;
    }
  }
};

}  // namespace

// =============================================================================
// launchers
// =============================================================================

void launch_stage_bin(Sim &s, const double *kx, const double *ky, const double *kp, double kcdt) {
  Box b{s.kp.Lx, s.kp.Ly, s.kp.rLx, s.kp.rLy};
  hipLaunchKernelGGL(k_stage_bin, dim3(blocks_for(s.n)), dim3(kBlock), 0, s.stream, s.n, (int)s.linked, b, s.nx,
                     s.ny, s.sx, s.sy, s.sp, kx, ky, kp, kcdt, s.gx, s.gy, s.gp, s.cell_of, s.count, s.local);
}

void launch_scan(Sim &s) { hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, s.stream, s.count, s.start, s.nc); }

void launch_place(Sim &s) {
  hipLaunchKernelGGL(k_place, dim3(blocks_for(s.n)), dim3(kBlock), 0, s.stream, s.n, s.cell_of, s.local, s.start,
                     s.tmp);
}

void launch_rank_tiles(Sim &s) {
  dim3 g(blocks_for(s.n)), b(kBlock);
  switch (s.prec) {
    case P64:
      hipLaunchKernelGGL(k_rank_tiles<P64>, g, b, 0, s.stream, s.n, s.nx, s.kp.hx, s.kp.hy, s.tmp, s.cell_of, s.start,
                         s.gx, s.gy, s.gp, s.sorted, s.t);
      break;
    case P64F:
      hipLaunchKernelGGL(k_rank_tiles<P64F>, g, b, 0, s.stream, s.n, s.nx, s.kp.hx, s.kp.hy, s.tmp, s.cell_of,
                         s.start, s.gx, s.gy, s.gp, s.sorted, s.t);
      break;
    default:
      hipLaunchKernelGGL(k_rank_tiles<P32K>, g, b, 0, s.stream, s.n, s.nx, s.kp.hx, s.kp.hy, s.tmp, s.cell_of,
                         s.start, s.gx, s.gy, s.gp, s.sorted, s.t);
  }
}

void launch_pass1(Sim &s, int epi, double *ox, double *oy, double *op, const StepArgs &sa, double *const k1[3],
                  double *const k2[3], double *const k3[3]) {
  EpiCtx e = make_ctx(s, ox, oy, op, sa, k1, k2, k3);
  ProfScope prof(s);
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
  ProfScope prof(s);
  switch (s.prec) {
    case P64: pass2_epi<P64>(s, epi, e); break;
    case P64F: pass2_epi<P64F>(s, epi, e); break;
    default: pass2_epi<P32K>(s, epi, e);
  }
}

void launch_order_param(Sim &s, double *out2) {
  int nb = std::min(kRedBlocks, std::max(1, blocks_for(s.n)));
  hipLaunchKernelGGL(k_order_partial, dim3(nb), dim3(kBlock), 0, s.stream, s.n, s.sp, s.red);
  hipLaunchKernelGGL(k_order_final, dim3(1), dim3(kBlock), 0, s.stream, nb, s.red, out2);
}

void launch_init_state(Sim &s, double omega_std) {
  hipLaunchKernelGGL(k_init_state, dim3(blocks_for(s.n)), dim3(kBlock), 0, s.stream, s.n, s.P.Lx, s.P.Ly, s.P.seed,
                     omega_std, s.sx, s.sy, s.sp, s.omega);
}

void launch_noise(Sim &s, int per_particle, uint64_t step, double *out) {
  hipLaunchKernelGGL(k_noise, dim3(blocks_for(s.n)), dim3(kBlock), 0, s.stream, s.n, per_particle, s.P.seed, step,
                     out);
}

void launch_fill(Sim &s, double *p, double v, size_t n) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_fill, dim3((unsigned)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, s.stream, p, v, n);
}

}  // namespace spc2
