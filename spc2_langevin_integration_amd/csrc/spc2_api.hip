// spc2 C ABI: handle lifetime, host <-> device copies and the per-method
// pipelines that chain the kernels of spc2_cells.hip, spc2_pairs.hip and
// spc2_misc.hip on the handle's stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <mutex>
#include <new>
#include <string>
#include <utility>
#include <vector>

#include "spc2_internal.h"

using namespace spc2;

struct spc2_sim {
  Sim s;
};

namespace {

thread_local std::string g_err;

spc2_status fail(spc2_status st, const std::string &msg) {
  g_err = msg;
  return st;
}

#define HIP_TRY(expr)                                                                           \
  do {                                                                                          \
    hipError_t e_ = (expr);                                                                     \
    if (e_ != hipSuccess)                                                                       \
      return fail(SPC2_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));             \
  } while (0)

// Checks the launches queued since the last check.
spc2_status check_launch(const char *what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(SPC2_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
  return SPC2_OK;
}

template <class T>
hipError_t dalloc(T **p, size_t count) {
  return hipMalloc((void **)p, sizeof(T) * std::max<size_t>(count, 1));
}

void free_all(Sim &s) {
  auto f = [](void *p) {
    if (p) (void)hipFree(p);
  };
  f(s.sx); f(s.sy); f(s.sp); f(s.omega);
  for (auto &kk : s.k)
    for (auto &p : kk) f(p);
  for (auto &p : s.der) f(p);
  for (auto &p : s.add) f(p);
  f(s.gx); f(s.gy); f(s.gp);
  f(s.cell_of); f(s.count); f(s.local); f(s.start); f(s.tmp); f(s.tmpid); f(s.slot_h); f(s.slot_id);
  f(s.hid); f(s.hpos); f(s.spare_i);
  for (auto &p : s.spare) f(p);
  for (auto &p : s.io) f(p);
  f(s.t.x); f(s.t.y); f(s.t.p); f(s.t.c); f(s.t.s); f(s.t.f4); f(s.t.cell);
  f(s.t2.x); f(s.t2.y); f(s.t2.p); f(s.t2.c); f(s.t2.s); f(s.t2.f4); f(s.gp2);
  f(s.aphi_slot); f(s.n_slot); f(s.n2_slot); f(s.aphi_slot_f);
  f(s.noise); f(s.counts); f(s.red); f(s.bsum); f(s.snap);
  f(s.vcnt); f(s.wmax); f(s.wrow); f(s.vlist); f(s.xb); f(s.yb); f(s.vmask);
  f(s.bh_nrun); f(s.bh_run); f(s.bh_off); f(s.bh_max); f(s.vslot); f(s.wmax_s); f(s.wrow_s);
  f(s.wmax8); f(s.wrow8); f(s.vrho);
  for (auto &kk : s.kev)
    for (auto &p : kk) f(p);
  if (s.snap_ev) (void)hipEventDestroy(s.snap_ev);
}

}  // namespace

namespace spc2 {

spc2_status set_error(spc2_status st, const std::string &msg) { return fail(st, msg); }

int precision_code(int32_t p) { return p == SPC2_FP64 ? P64 : (p == SPC2_FP32_KAHAN ? P32K : P64F); }

spc2_status cell_grid(const spc2_params &p, bool &linked, int &nx, int &ny) {
  linked = p.rho <= 0.25 * p.Lx;  // ChimeraSystem.cpp:90, VortexArraysSystemPtr.cpp:113
  nx = ny = 1;
  if (!linked) return SPC2_OK;
  double fx = p.Lx / p.rho, fy = p.Ly / p.rho;
  if (fx > 2.0e9 || fy > 2.0e9 || fx * fy > 1.0e9) return fail(SPC2_ERR_INVALID, "too many cells for the linked-cell path");
  nx = (int)(p.Lx / p.rho);  // num_subcells_, ChimeraSystem.cpp:32-33
  ny = (int)(p.Ly / p.rho);
  if (nx < 3 || ny < 3) return fail(SPC2_ERR_INVALID, "linked cells need at least 3 cells per side (rho <= L/3)");
  return SPC2_OK;
}

void fill_kparams_from(KParams &k, const spc2_params &P, int n, int nx, int ny, bool linked) {
  k.n = n;
  k.nx = nx;
  k.ny = ny;
  k.linked = linked;
  k.Lx = P.Lx;
  k.Ly = P.Ly;
  k.rLx = 1.0 / P.Lx;
  k.rLy = 1.0 / P.Ly;
  k.hx = nx > 0 ? P.Lx / nx : P.Lx;
  k.hy = ny > 0 ? P.Ly / ny : P.Ly;
  k.hxf = (float)k.hx;
  k.hyf = (float)k.hy;
  k.rho = P.rho;
  k.rho2 = P.rho * P.rho;  // rho_squared_, ChimeraSystem.cpp:24
  k.rho2f = (float)k.rho2;
  k.v0 = P.v0;
  k.sigma = P.sigma;
  k.alpha = P.alpha;
  k.D = P.D;
  k.mu_plus = P.mu_plus;
  k.mu_minus = P.mu_minus;
  k.xi_a = P.xi_a;
  k.xi_r = P.xi_r;
  k.kappa = P.kappa;
  double outer = P.system == SPC2_VORTEX ? std::max(P.rho, P.xi_r) : P.rho;
  k.outer2 = outer * outer;
  k.xa2f = (float)(P.xi_a * P.xi_a);
  k.xr2f = (float)(P.xi_r * P.xi_r);
  k.outer2f = (float)k.outer2;
  k.xaf = (float)P.xi_a;
  k.rhof = (float)P.rho;
  k.mu_pf = (float)P.mu_plus;
  k.mu_p_over_xa2f = P.xi_a > 0 ? (float)(P.mu_plus / (P.xi_a * P.xi_a)) : 0.0f;
  k.c4f = P.rho > P.xi_a ? (float)(-P.mu_minus * 4.0 / ((P.rho - P.xi_a) * (P.rho - P.xi_a))) : 0.0f;
  {
    const double c4 = P.rho > P.xi_a ? -P.mu_minus * 4.0 / ((P.rho - P.xi_a) * (P.rho - P.xi_a)) : 0.0;
    k.w_af = (float)(c4 * (P.xi_a + P.rho));
    k.w_bf = (float)(-c4 * P.xi_a * P.rho);
  }
  k.kappaf = (float)P.kappa;
  k.v0f = (float)P.v0;
  k.Df = (float)P.D;
  // FP32 r^2 from cell-local coordinates is within a few 1e-7 relative of the
  // FP64 value; pairs closer than 1e-5 relative to a cutoff take the FP64 test.
  k.band_rho2f = (float)(1e-5 * k.rho2);
  k.band_xr2f = (float)(1e-5 * P.xi_r * P.xi_r);
  k.band_xa2f = (float)(1e-5 * P.xi_a * P.xi_a);
  k.rho2f_lo = k.rho2f - k.band_rho2f;
  k.rho2f_hi = k.rho2f + k.band_rho2f;
  k.xr2f_hi = k.xr2f + k.band_xr2f;
  const double close = 1e-2 * std::min(k.hx, k.hy);
  k.close2f = (float)(close * close);
}

}  // namespace spc2

namespace {

void fill_kparams(Sim &s) { fill_kparams_from(s.kp, s.P, s.n, s.nx, s.ny, s.linked); }

spc2_status alloc_all(Sim &s) {
  const size_t n = (size_t)s.n, nc = (size_t)std::max(s.nc, 1);
#define A(ptr, cnt) HIP_TRY(dalloc(&(ptr), (cnt)))
  A(s.sx, n); A(s.sy, n); A(s.sp, n); A(s.omega, n);
  for (auto &kk : s.k)
    for (auto &p : kk) A(p, n);
  for (auto &p : s.der) A(p, n);
  for (auto &p : s.add) A(p, n);
  A(s.gx, n); A(s.gy, n); A(s.gp, n);
  A(s.cell_of, n); A(s.count, nc); A(s.local, n); A(s.start, nc + 1); A(s.tmp, n); A(s.tmpid, n);
  A(s.slot_h, n); A(s.slot_id, n); A(s.hid, n); A(s.hpos, n); A(s.spare_i, n);
  for (auto &p : s.io) A(p, n);
  if (s.linked)
    for (auto &p : s.spare) A(p, n);
  A(s.t.cell, n);
  for (Tiles *tt : {&s.t, &s.t2}) {
    if (tt == &s.t2 && !s.verlet) break;
    if (s.prec == P32K) {
      A(tt->f4, n);
      A(tt->x, n); A(tt->y, n);  // FP64 positions for the cutoff-band re-check
    } else {
      A(tt->x, n); A(tt->y, n); A(tt->p, n);
      if (s.prec == P64F) {
        A(tt->c, n); A(tt->s, n);
      }
    }
  }
  if (s.verlet) {
    A(s.gp2, n);
    s.t2.cell = s.t.cell;
  }
  A(s.aphi_slot, n); A(s.n_slot, n); A(s.n2_slot, n); A(s.aphi_slot_f, n);
  A(s.counts, 3 * kCountSlots); A(s.red, 2 * 1024 + 8);
  A(s.bsum, (size_t)scan_blocks(std::max({s.nc, s.nw, 1})) + 1);
  if (s.verlet) {
    A(s.vcnt, n); A(s.wmax, (size_t)s.nw + 1); A(s.wrow, (size_t)s.nw + 1); A(s.xb, n); A(s.yb, n);
    {  // room for 1.5x the mean candidate count (9 build cells), 1..16 words
      const double mean = 9.0 * (double)s.n / (double)std::max(s.nc, 1);
      s.vmask_words = (int)std::min(16.0, std::max(1.0, std::ceil(1.5 * mean / 32.0)));
      A(s.vmask, (size_t)s.vmask_words * n);
    }
    {  // the LDS layout's halo runs and lane slots (every precision form)
      s.nblk = (s.n + 255) / 256;
      A(s.bh_nrun, (size_t)s.nblk); A(s.bh_run, (size_t)s.nblk * 2 * kVerletRunMax);
      A(s.bh_off, (size_t)s.nblk * (kVerletRunMax + 1)); A(s.bh_max, 1);
      A(s.vslot, n); A(s.wmax_s, (size_t)s.nw + 1); A(s.wrow_s, (size_t)s.nw + 1);
      A(s.wmax8, (size_t)s.nw + 1); A(s.wrow8, (size_t)s.nw + 1);
    }
  }
#undef A
  // Cleared on the handle's stream, so that the first kernels on that
  // (non-blocking) stream see zeros: the cell histogram must start at zero,
  // since the first binning's arrival ranks index the placement.
  HIP_TRY(hipMemsetAsync(s.count, 0, sizeof(int) * nc, s.stream));
  HIP_TRY(hipMemsetAsync(s.omega, 0, sizeof(double) * n, s.stream));
  for (auto &kk : s.k)
    for (auto &p : kk) HIP_TRY(hipMemsetAsync(p, 0, sizeof(double) * n, s.stream));
  for (auto &p : s.der) HIP_TRY(hipMemsetAsync(p, 0, sizeof(double) * n, s.stream));
  for (auto &p : s.add) HIP_TRY(hipMemsetAsync(p, 0, sizeof(double) * n, s.stream));
  HIP_TRY(hipMemsetAsync(s.counts, 0, sizeof(unsigned long long) * 3 * kCountSlots, s.stream));
  HIP_TRY(hipMemsetAsync(s.sx, 0, sizeof(double) * n, s.stream));
  HIP_TRY(hipMemsetAsync(s.sy, 0, sizeof(double) * n, s.stream));
  HIP_TRY(hipMemsetAsync(s.sp, 0, sizeof(double) * n, s.stream));
  HIP_TRY(hipStreamSynchronize(s.stream));
  return SPC2_OK;
}

// AoS host array by particle id (stride `stride`, `comps` components) <->
// home-ordered device arrays, through the id-ordered transfer buffers.
spc2_status upload_aos(Sim &s, const double *aos, int stride, double *const *dst, int comps) {
  std::vector<double> buf((size_t)s.n);
  for (int c = 0; c < comps; ++c) {
    for (int i = 0; i < s.n; ++i) buf[(size_t)i] = aos[(size_t)i * stride + c];
    HIP_TRY(hipMemcpyAsync(s.io[c], buf.data(), sizeof(double) * (size_t)s.n, hipMemcpyHostToDevice, s.stream));
    launch_to_home(s, s.io[c], dst[c]);
    HIP_TRY(hipStreamSynchronize(s.stream));
  }
  return SPC2_OK;
}

spc2_status download_aos(Sim &s, double *aos, int stride, double *const *src, int comps) {
  std::vector<double> buf((size_t)s.n);
  for (int c = 0; c < comps; ++c) {
    launch_from_home(s, src[c], s.io[c]);
    HIP_TRY(hipMemcpyAsync(buf.data(), s.io[c], sizeof(double) * (size_t)s.n, hipMemcpyDeviceToHost, s.stream));
    HIP_TRY(hipStreamSynchronize(s.stream));
    for (int i = 0; i < s.n; ++i) aos[(size_t)i * stride + c] = buf[(size_t)i];
  }
  return SPC2_OK;
}

// Verlet list from the positions s + kcdt * kprev (the state when kprev is
// null): bin them on the build grid, move the particle arrays into the build's
// cell order, then count, scan and fill the wavefront-interleaved list.
spc2_status verlet_build(Sim &s, double *const *kprev, double kcdt, bool k_xy_live) {
  if (kprev)
    launch_stage_bin(s, kprev[0], kprev[1], kprev[2], kcdt);
  else
    launch_stage_bin(s, nullptr, nullptr, nullptr, 0.0);
  launch_scan(s);
  launch_place(s);
  launch_rank_tiles(s);
  launch_rehome(s, k_xy_live);  // slot k becomes home k: tiles, cells and start[] are now by home
  HIP_TRY(hipMemcpyAsync(s.xb, s.t.x, sizeof(double) * (size_t)s.n, hipMemcpyDeviceToDevice, s.stream));
  HIP_TRY(hipMemcpyAsync(s.yb, s.t.y, sizeof(double) * (size_t)s.n, hipMemcpyDeviceToDevice, s.stream));
  // the LDS layout: every pass workgroup's halo, and the largest (the FP32
  // count pass also reads its candidates from the halos that fit its LDS)
  const bool try_lds = s.vlds_allowed && s.bh_max && verlet_lds_unbased(s.P.system);
  if (try_lds) {
    HIP_TRY(hipMemsetAsync(s.bh_max, 0, sizeof(int), s.stream));
    launch_block_halo(s);
  }
  launch_verlet_count(s, try_lds);
  launch_scan_array(s, s.wmax, s.nw, s.wrow);
  if (spc2_status st = check_launch("verlet count")) return st;
  // lane slots sorted by list length pay off when lists are long enough for the
  // saved wavefront iterations and list padding to outweigh the epilogue's
  // scattered accesses (measured on the vortex RK4 pass at 45 entries per
  // particle: 5% faster over 500 steps); the mean list length is estimated from
  // the mean density
  const double mean_list = (double)s.n / (s.P.Lx * s.P.Ly) * 3.14159265358979 * s.rlist * s.rlist;
  const bool want_sort = try_lds && (s.vsort_mode == 1 || (s.vsort_mode < 0 && mean_list > 32.0));
  int *wrow_l = want_sort ? s.wrow_s : s.wrow8;  // the LDS layout's offsets (groups of eight)
  if (try_lds) {
    if (want_sort) {
      launch_verlet_order(s);
      launch_scan_array(s, s.wmax_s, s.nw, s.wrow_s);
    } else {
      launch_scan_array(s, s.wmax8, s.nw, s.wrow8);
    }
  }
  int groups = 0, groups_l = 0, halo = 0;
  HIP_TRY(hipMemcpyAsync(&groups, s.wrow + s.nw, sizeof(int), hipMemcpyDeviceToHost, s.stream));
  if (try_lds) {
    HIP_TRY(hipMemcpyAsync(&groups_l, wrow_l + s.nw, sizeof(int), hipMemcpyDeviceToHost, s.stream));
    HIP_TRY(hipMemcpyAsync(&halo, s.bh_max, sizeof(int), hipMemcpyDeviceToHost, s.stream));
  }
  HIP_TRY(hipStreamSynchronize(s.stream));
  // pass 2 stages 20 B per halo particle in the FP32 form (32 B, 40 B in the
  // FP64 forms), plus one sentinel tile after the halo
  const int cap_limit = kLdsBytes / verlet_lds_tile_bytes(s.prec);
  s.vlds = try_lds && halo >= 0 && halo < cap_limit;  // halo is INT_MAX when a workgroup cannot be staged
  s.vlds_cap = s.vlds ? std::min(cap_limit, std::max(64, (halo + 1 + 63) / 64 * 64)) : 0;
  s.vsorted = s.vlds && want_sort;
  if (s.vlds) {  // the passes and the fill use the LDS layout's offsets
    std::swap(s.wrow, want_sort ? s.wrow_s : s.wrow8);
    groups = groups_l;
  }
  // groups = sum over wavefronts of their longest list in groups of four (LDS
  // layout: eight); more than 4096 entries per particle means a skin or density
  // the list is not meant for
  if (groups < 0 || (long long)groups * (s.vlds ? 8 : 4) > 4096LL * s.nw)
    return fail(SPC2_ERR_INVALID, "Verlet list of " + std::to_string(groups) + " groups for " +
                                      std::to_string(s.nw) + " wavefronts: skin or density too large for a neighbour list");
  const size_t need = ((size_t)groups + kVerletSpareGroups) * 256;
  if (need > s.vlist_cap || !s.vlist) {
    if (s.vlist) (void)hipFree(s.vlist);
    s.vlist = nullptr;
    s.vlist_cap = 0;
    const size_t cap = need + need / 8 + 64 * 256;
    HIP_TRY(dalloc(&s.vlist, cap));
    s.vlist_cap = cap;
  }
  s.vrho_fresh = false;
  if (s.vlds && s.prec == P32K && s.P.system == SPC2_CHIMERA && s.vrho_allowed) {
    const size_t bytes = need / 4;  // a byte per lane and group
    if (bytes > s.vrho_cap || !s.vrho) {
      if (s.vrho) (void)hipFree(s.vrho);
      s.vrho = nullptr;
      s.vrho_cap = 0;
      const size_t cap = bytes + bytes / 8 + 64 * 64;
      HIP_TRY(dalloc(&s.vrho, cap));
      s.vrho_cap = cap;
    }
  } else if (s.vrho) {
    (void)hipFree(s.vrho);
    s.vrho = nullptr;
    s.vrho_cap = 0;
  }
  if (s.vlds)
    launch_verlet_fill_lds(s);
  else
    launch_verlet_fill(s);
  HIP_TRY(hipMemsetAsync(s.vlist + (size_t)groups * 256, 0, sizeof(int) * 256 * kVerletSpareGroups, s.stream));
  if (spc2_status st = check_launch("verlet build")) return st;
  s.vbuilt = true;
  s.built = false;
  s.tile_fresh = false;
  s.steps_since_rehome = 0;
  s.drift = kprev ? std::numeric_limits<double>::infinity() : 0.0;
  s.vbuilds += 1;
  s.ventries = -1;
  return SPC2_OK;
}

// Largest displacement of the state since the Verlet build (synchronous).
spc2_status verlet_measure(Sim &s) {
  double d2 = 0.0;
  launch_drift(s, s.red + 2048);
  HIP_TRY(hipMemcpyAsync(&d2, s.red + 2048, sizeof(double), hipMemcpyDeviceToHost, s.stream));
  HIP_TRY(hipStreamSynchronize(s.stream));
  s.drift = std::sqrt(d2) * (1.0 + 1e-12);
  return SPC2_OK;
}

// Makes the list valid for evaluations at positions within `reach` of the
// state: stage positions are within drift + reach of the build positions, and
// every pair within rho at the stage was within rho + skin at the build while
// that stays <= skin / 2.
// Largest distance one SRK step can move a particle when the noise comes from
// the built-in generator (StochasticRungeKuttaStepperPtr.hpp:52-66 update):
//   v0 dt                                    the drift,
// + sqrt(2 D) v0 |dZ|, dZ = h1 (U1 + U2 / sqrt(3)), |U| <= 8.6 (Box-Muller's
//   radius at its smallest uniform, 2^-53, is 8.572),
// + dt^2 / 2 v0 (|a_phi| + D), a_phi the pass-1 phase derivative: |a_phi| <=
//   |sigma| (chimera: sigma times a mean of sines) or max(|mu+|, |mu-|) +
//   |kappa| (vortex arrays: a mean of W sin, |W| <= max(|mu+|, |mu-|), plus the
//   cohesion mean), with 0.1% for rounding.
// Supplied noise has no such bound: its steps are measured instead.
double srk_step_reach(const Sim &s, double dt) {
  const double v0 = std::fabs(s.P.v0), D = std::fabs(s.P.D);
  const double a_phi = 1.001 * (s.P.system == SPC2_CHIMERA
                                    ? std::fabs(s.P.sigma)
                                    : std::max(std::fabs(s.P.mu_plus), std::fabs(s.P.mu_minus)) + std::fabs(s.P.kappa));
  const double dZ = 0.5 * dt * std::sqrt(dt) * 8.6 * (1.0 + 1.0 / std::sqrt(3.0));
  return v0 * (dt + std::sqrt(2.0 * D) * dZ + 0.5 * dt * dt * (a_phi + D));
}

spc2_status verlet_ensure(Sim &s, double reach, bool k_xy_live) {
  const double half = 0.5 * s.skin * (1.0 - 1e-9);
  if (s.vbuilt && s.drift + reach <= half) return SPC2_OK;
  if (s.vbuilt && std::isinf(s.drift)) {
    if (spc2_status st = verlet_measure(s)) return st;
    if (s.drift + reach <= half) return SPC2_OK;
  }
  return verlet_build(s, nullptr, 0.0, k_xy_live);
}

// Stage + cell list + tiles for one RHS evaluation (Verlet: stage and tiles
// by home only; the list is kept valid by verlet_ensure).
void build(Sim &s, double *const *kprev, double kcdt) {
  s.tile_fresh = false;
  if (s.verlet) {
    if (kprev)
      launch_verlet_stage(s, kprev[0], kprev[1], kprev[2], kcdt);
    else
      launch_verlet_stage(s, nullptr, nullptr, nullptr, 0.0);
    return;
  }
  if (kprev)
    launch_stage_bin(s, kprev[0], kprev[1], kprev[2], kcdt);
  else
    launch_stage_bin(s, nullptr, nullptr, nullptr, 0.0);
  if (s.linked) {
    launch_scan(s);
    launch_place(s);
    launch_rank_tiles(s);
    s.built = true;
  }
}

// Moves the particle arrays to the cell order of the current state (used
// when a new state arrives).
spc2_status rehome_now(Sim &s) {
  if (s.verlet) return verlet_build(s, nullptr, 0.0, true);
  if (!s.linked) return SPC2_OK;
  build(s, nullptr, 0.0);
  launch_rehome(s);
  s.built = false;
  s.steps_since_rehome = 0;
  HIP_TRY(hipStreamSynchronize(s.stream));
  return SPC2_OK;
}

// After every rehome_every steps, moves the particle arrays to the cell order
// of the last build.
void maybe_rehome(Sim &s) {
  if (!s.linked || s.verlet) return;  // Verlet homes change only at list builds
  if (++s.steps_since_rehome < s.rehome_every || !s.built) return;
  launch_rehome(s);
  s.built = false;
  s.steps_since_rehome = 0;
}

// Verlet: the next pass launch also writes the next stage's tiles (mode 1:
// s + kcdt k of that pass-1 launch; 2: the state the launch writes, by pass 1
// of SE and RK4's last stage or by SRK's pass 2).
void next_stage(Sim &s, int mode, double kcdt) {
  s.vnext = mode;
  s.vnext_kcdt = kcdt;
}

// RHS passes on the handle's neighbour structure.
void rhs_pass1(Sim &s, int epi, double *ox, double *oy, double *op, const StepArgs &sa, double *const k1[3],
               double *const k2[3], double *const k3[3]) {
  if (s.verlet)
    launch_verlet_pass1(s, epi, ox, oy, op, sa, k1, k2, k3);
  else
    launch_pass1(s, epi, ox, oy, op, sa, k1, k2, k3);
}

void rhs_pass2(Sim &s, int epi, double *a0, double *a1, const StepArgs &sa) {
  if (s.verlet)
    launch_verlet_pass2(s, epi, a0, a1, sa);
  else
    launch_pass2(s, epi, a0, a1, sa);
}

StepArgs step_args(const Sim &s, double dt, uint64_t step) {
  StepArgs a{};
  a.dt = dt;
  a.sqdt = std::sqrt(dt);
  a.sq2D = std::sqrt(2.0 * s.P.D);
  a.h1 = 0.5 * dt * std::sqrt(dt);
  a.sqrt3 = std::sqrt(3.0);
  a.D = s.P.D;
  a.seed = s.P.seed;
  a.step = step;
  a.noise = nullptr;
  // natural frequencies enter the chimera drift only (the vortex-arrays
  // steppers have none, StochasticRungeKuttaStepperPtr.hpp:94)
  a.omega = s.P.system == SPC2_CHIMERA ? s.omega : nullptr;
  a.zero_in = 0;
  a.wrap = 1;
  return a;
}

spc2_status check_handle(spc2_sim_t h) {
  if (!h) return fail(SPC2_ERR_INVALID, "null handle");
  return SPC2_OK;
}

// Device staging of spc2_noise_noalloc: one 16 MiB array per GPU, held until
// the process ends (2^21 normals: the SRK draws of N = 1,048,576 particles).
// Host code never passes `scratch` itself to a kernel or copy (on the host the
// name is a stand-in, not a device address): it uses the address that
// hipGetSymbolAddress returns for the current device, kept per device in
// g_scratch.
constexpr size_t kScratchNormals = size_t(1) << 21;
static __device__ __align__(256) double scratch[kScratchNormals];

constexpr int kScratchDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch[kScratchDevices] = {};

// Looks up (once per device, with that device current) the address of
// `scratch` on `device`.
spc2_status scratch_lookup(int device) {
  if (device < 0 || device >= kScratchDevices) return fail(SPC2_ERR_INVALID, "device ordinal out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch[device]) return SPC2_OK;
  int cur = -1;
  HIP_TRY(hipGetDevice(&cur));
  if (cur != device) HIP_TRY(hipSetDevice(device));
  void *p = nullptr;
  hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (cur != device) (void)hipSetDevice(cur);
  if (e != hipSuccess || !p)
    return fail(SPC2_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
  g_scratch[device] = static_cast<double *>(p);
  return SPC2_OK;
}

}  // namespace

extern "C" {

void spc2_default_params(spc2_params *p) {
  if (!p) return;
  std::memset(p, 0, sizeof(*p));
  p->system = SPC2_CHIMERA;
  p->precision = SPC2_FP64;
  p->neighbour = SPC2_NEIGH_AUTO;
  p->n = 1000;
  p->Lx = p->Ly = 1.0;
  p->rho = 0.1;
  p->v0 = 0.01;
  p->sigma = 1.0;
  p->alpha = 1.45;
  p->D = 0.01;
  p->mu_plus = p->mu_minus = 100.0;
  p->xi_a = 0.01;
  p->xi_r = 0.005;
  p->kappa = 20.0;
  p->seed = 42;
}

const char *spc2_last_error(void) { return g_err.c_str(); }

spc2_status spc2_create(const spc2_params *p, spc2_sim_t *out) {
  g_err.clear();
  if (!p || !out) return fail(SPC2_ERR_INVALID, "null argument");
  *out = nullptr;
  if (p->n <= 0 || p->n >= (1LL << 31) - 1) return fail(SPC2_ERR_INVALID, "n must be in [1, 2^31 - 1)");
  if (!(p->Lx > 0.0) || !(p->Ly > 0.0) || !(p->rho > 0.0)) return fail(SPC2_ERR_INVALID, "Lx, Ly, rho must be > 0");
  if (p->system != SPC2_CHIMERA && p->system != SPC2_VORTEX) return fail(SPC2_ERR_INVALID, "unknown system");
  if (p->precision < SPC2_FP64 || p->precision > SPC2_FP64_FACTORED)
    return fail(SPC2_ERR_INVALID, "unknown precision");
  if (p->neighbour < SPC2_NEIGH_AUTO || p->neighbour > SPC2_NEIGH_VERLET)
    return fail(SPC2_ERR_INVALID, "unknown neighbour mode");
  if (!(p->verlet_skin >= 0.0)) return fail(SPC2_ERR_INVALID, "verlet_skin must be >= 0");
  if (!(p->D >= 0.0)) return fail(SPC2_ERR_INVALID, "D must be >= 0");
  spc2_sim *h = new (std::nothrow) spc2_sim();
  if (!h) return fail(SPC2_ERR_NOMEM, "host allocation failed");
  Sim &s = h->s;
  s.P = *p;
  s.prec = p->precision == SPC2_FP64 ? P64 : (p->precision == SPC2_FP32_KAHAN ? P32K : P64F);
  s.n = (int)p->n;
  if (p->neighbour == SPC2_NEIGH_AUTO)
    s.linked = p->rho <= 0.25 * p->Lx;  // ChimeraSystem.cpp:90, VortexArraysSystemPtr.cpp:113
  else
    s.linked = p->neighbour == SPC2_NEIGH_CELLS || p->neighbour == SPC2_NEIGH_VERLET;
  s.verlet = p->neighbour == SPC2_NEIGH_VERLET;
  s.nw = (s.n + 63) / 64;
  if (s.verlet) {
    if (p->n >= (1LL << 27)) {  // list entries pack the neighbour's home above 4 offset bits
      delete h;
      return fail(SPC2_ERR_INVALID, "Verlet lists support n < 2^27");
    }
    if (p->system == SPC2_VORTEX && p->rho <= 0.25 * p->Lx && p->rho <= 0.25 * p->Ly) {
      // the reference finds cohesion partners only within its 3 x 3 stencil of
      // cells of side L / int(L / rho) (VortexArraysSystemPtr.cpp:35-36, :449-489);
      // a list reaching further would add partners the reference never sees
      const double hx = p->Lx / (int)(p->Lx / p->rho), hy = p->Ly / (int)(p->Ly / p->rho);
      if (p->xi_r > std::min(hx, hy)) {
        delete h;
        return fail(SPC2_ERR_INVALID, "Verlet lists need xi_r <= the linked cells' side L / int(L / rho)");
      }
    }
    s.skin = p->verlet_skin > 0.0 ? p->verlet_skin : 0.2 * p->rho;
    s.rlist = (p->system == SPC2_VORTEX ? std::max(p->rho, p->xi_r) : p->rho) + s.skin;
    double fx = p->Lx / s.rlist, fy = p->Ly / s.rlist;
    if (fx > 2.0e9 || fy > 2.0e9 || fx * fy > 1.0e9 || (int)fx < 3 || (int)fy < 3) {
      delete h;
      return fail(SPC2_ERR_INVALID, "Verlet lists need 3 to 1e9 build cells (L / (rho + skin) >= 3)");
    }
    s.nx = (int)fx;  // build cells of side L / nx >= rho + skin
    s.ny = (int)fy;
    s.nc = s.nx * s.ny;
  } else if (s.linked) {
    double fx = p->Lx / p->rho, fy = p->Ly / p->rho;
    if (fx > 2.0e9 || fy > 2.0e9 || fx * fy > 1.0e9) {
      delete h;
      return fail(SPC2_ERR_INVALID, "too many cells for the linked-cell path");
    }
    s.nx = (int)(p->Lx / p->rho);  // num_subcells_, ChimeraSystem.cpp:32-33
    s.ny = (int)(p->Ly / p->rho);
    if (s.nx < 3 || s.ny < 3) {
      delete h;
      return fail(SPC2_ERR_INVALID, "linked cells need at least 3 cells per side (rho <= L/3)");
    }
    s.nc = s.nx * s.ny;
  }
  fill_kparams(s);
  configure_segments(s);
  if (const char *env = std::getenv("SPC2_REHOME_EVERY")) s.rehome_every = std::max(1, std::atoi(env));
  if (const char *env = std::getenv("SPC2_VERLET_LDS")) s.vlds_allowed = std::atoi(env) != 0;
  if (const char *env = std::getenv("SPC2_VERLET_SORT")) s.vsort_mode = std::atoi(env) != 0 ? 1 : 0;
  if (const char *env = std::getenv("SPC2_RK4_STORE_KXY")) s.rk4_store_kxy = std::atoi(env) != 0;
  if (const char *env = std::getenv("SPC2_VERLET_RHO_BITS")) s.vrho_allowed = std::atoi(env) != 0;
  hipError_t e = hipGetDevice(&s.device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    std::string m = std::string("HIP init: ") + hipGetErrorString(e);
    delete h;
    return fail(SPC2_ERR_HIP, m);
  }
  // set-up of the device's staging array for spc2_noise_noalloc (first handle
  // on a device), so that the entry point itself makes no lookup
  if (spc2_status st = scratch_lookup(s.device)) {
    (void)hipStreamDestroy(s.stream);
    delete h;
    return st;
  }
  spc2_status st = alloc_all(s);
  if (st == SPC2_OK) {
    launch_identity_homes(s);
    if (hipStreamSynchronize(s.stream) != hipSuccess) st = fail(SPC2_ERR_HIP, "initialising homes failed");
  }
  if (st != SPC2_OK) {
    std::string m = g_err;
    free_all(s);
    (void)hipStreamDestroy(s.stream);
    delete h;
    return fail(SPC2_ERR_NOMEM, m);
  }
  *out = h;
  return SPC2_OK;
}

spc2_status spc2_destroy(spc2_sim_t h) {
  if (!h) return SPC2_OK;
  Sim &s = h->s;
  (void)hipStreamSynchronize(s.stream);
  s.prof_collect();
  for (auto &pr : s.prof_free) {
    (void)hipEventDestroy(pr.first);
    (void)hipEventDestroy(pr.second);
  }
  free_all(s);
  (void)hipStreamDestroy(s.stream);
  delete h;
  return SPC2_OK;
}

spc2_status spc2_set_state(spc2_sim_t h, const double *state) {
  if (check_handle(h) || !state) return fail(SPC2_ERR_INVALID, "null argument");
  Sim &s = h->s;
  double *d[3] = {s.sx, s.sy, s.sp};
  if (spc2_status st = upload_aos(s, state, 3, d, 3)) return st;
  return rehome_now(s);
}

spc2_status spc2_get_state(spc2_sim_t h, double *state) {
  if (check_handle(h) || !state) return fail(SPC2_ERR_INVALID, "null argument");
  Sim &s = h->s;
  double *d[3] = {s.sx, s.sy, s.sp};
  return download_aos(s, state, 3, d, 3);
}

spc2_status spc2_init_random_state(spc2_sim_t h, double omega_std) {
  if (check_handle(h)) return SPC2_ERR_INVALID;
  if (!(omega_std >= 0.0)) return fail(SPC2_ERR_INVALID, "omega_std must be >= 0");
  Sim &s = h->s;
  launch_init_state(s, omega_std);
  if (spc2_status st = check_launch("init_state")) return st;
  HIP_TRY(hipStreamSynchronize(s.stream));
  return rehome_now(s);
}

spc2_status spc2_set_omega(spc2_sim_t h, const double *omega) {
  if (check_handle(h) || !omega) return fail(SPC2_ERR_INVALID, "null argument");
  Sim &s = h->s;
  HIP_TRY(hipMemcpyAsync(s.io[0], omega, sizeof(double) * (size_t)s.n, hipMemcpyHostToDevice, s.stream));
  launch_to_home(s, s.io[0], s.omega);
  HIP_TRY(hipStreamSynchronize(s.stream));
  return SPC2_OK;
}

spc2_status spc2_eval_rhs(spc2_sim_t h, const double *k_prev, double k_coef, double dt, double *k_next) {
  if (check_handle(h) || !k_prev || !k_next) return fail(SPC2_ERR_INVALID, "null argument");
  Sim &s = h->s;
  // The caller's k buffers go to scratch of their own: s.k holds the RK4
  // stepper's k_1..k_4, whose phi components carry over between its steps.
  for (auto &kk : s.kev)
    for (auto &p : kk)
      if (!p) HIP_TRY(dalloc(&p, (size_t)s.n));
  if (spc2_status st = upload_aos(s, k_prev, 3, s.kev[1], 3)) return st;
  if (s.verlet) {
    // reach of the stage: |k_coef dt| max_i |(k_x, k_y)_i|
    double kmax = 0.0;
    for (int64_t i = 0; i < s.n; ++i) kmax = std::max(kmax, std::hypot(k_prev[3 * i], k_prev[3 * i + 1]));
    const double reach = std::fabs(k_coef * dt) * kmax * (1.0 + 1e-12);
    const long long builds = s.vbuilds;
    spc2_status st = reach <= 0.5 * s.skin * (1.0 - 1e-9) ? verlet_ensure(s, reach, false)
                                                           : verlet_build(s, s.kev[1], k_coef * dt, false);
    if (st) return st;
    // a build re-homes the particles: k_prev again, in the new home order
    if (s.vbuilds != builds)
      if (spc2_status st2 = upload_aos(s, k_prev, 3, s.kev[1], 3)) return st2;
  }
  if (spc2_status st = upload_aos(s, k_next, 3, s.kev[0], 3)) return st;
  build(s, s.kev[1], k_coef * dt);
  StepArgs sa = step_args(s, dt, 0);
  rhs_pass1(s, EPI_K, s.kev[0][0], s.kev[0][1], s.kev[0][2], sa, nullptr, nullptr, nullptr);
  if (spc2_status st = check_launch("eval_rhs")) return st;
  return download_aos(s, k_next, 3, s.kev[0], 3);
}

spc2_status spc2_eval_rhs_ito(spc2_sim_t h, double *derivative, double *additional) {
  if (check_handle(h) || !derivative || !additional) return fail(SPC2_ERR_INVALID, "null argument");
  Sim &s = h->s;
  // the list first: a build re-homes the particles, so the in/out arrays are
  // uploaded after it, in the home order the passes use
  if (s.verlet)
    if (spc2_status st = verlet_ensure(s, 0.0, false)) return st;
  if (spc2_status st = upload_aos(s, derivative, 3, s.der, 3)) return st;
  if (spc2_status st = upload_aos(s, additional, 2, s.add, 2)) return st;
  build(s, nullptr, 0.0);
  StepArgs sa = step_args(s, 0.0, 0);
  rhs_pass1(s, EPI_ITO1, s.der[0], s.der[1], s.der[2], sa, nullptr, nullptr, nullptr);
  rhs_pass2(s, EPI_ITO2, s.add[0], s.add[1], sa);
  if (spc2_status st = check_launch("eval_rhs_ito")) return st;
  if (spc2_status st = download_aos(s, derivative, 3, s.der, 3)) return st;
  return download_aos(s, additional, 2, s.add, 2);
}

spc2_status spc2_step(spc2_sim_t h, int32_t method, double dt, int64_t step0, int64_t nsteps, const double *noise) {
  return spc2_step_ex(h, method, dt, step0, nsteps, noise, 0);
}

spc2_status spc2_step_ex(spc2_sim_t h, int32_t method, double dt, int64_t step0, int64_t nsteps, const double *noise,
                         int32_t flags) {
  if (check_handle(h)) return SPC2_ERR_INVALID;
  if (flags & ~SPC2_STEP_NO_WRAP) return fail(SPC2_ERR_INVALID, "unknown step flags");
  Sim &s = h->s;
  if (method < SPC2_SE || method > SPC2_RK4) return fail(SPC2_ERR_INVALID, "unknown method");
  if (!(dt > 0.0) || nsteps < 0 || step0 < 0) return fail(SPC2_ERR_INVALID, "dt must be > 0, steps >= 0");
  const int npp = method == SPC2_SRK ? 2 : 1;
  if (noise && method != SPC2_RK4) {
    size_t need = (size_t)nsteps * npp * (size_t)s.n;
    if (need > s.noise_cap) {
      if (s.noise) (void)hipFree(s.noise);
      s.noise = nullptr;
      s.noise_cap = 0;
      HIP_TRY(dalloc(&s.noise, need));
      s.noise_cap = need;
    }
    HIP_TRY(hipMemcpyAsync(s.noise, noise, sizeof(double) * need, hipMemcpyHostToDevice, s.stream));
  }
  for (int64_t k = 0; k < nsteps; ++k) {
    if (s.verlet) {
      // RK4 stages lie within v0 dt of the state (|k_x, k_y| = v0), SE and SRK
      // evaluate at the state.  A rebuild re-homes the particle arrays, so the
      // step's arguments are taken after it.
      // (the x / y components of k are dead here: a step writes them first)
      if (spc2_status st = verlet_ensure(s, method == SPC2_RK4 ? std::fabs(s.P.v0) * dt * (1.0 + 1e-12) : 0.0,
                                         false))
        return st;
    }
    StepArgs sa = step_args(s, dt, (uint64_t)(step0 + k));
    sa.wrap = (flags & SPC2_STEP_NO_WRAP) ? 0 : 1;
    if (noise && method != SPC2_RK4) sa.noise = s.noise + (size_t)k * npp * (size_t)s.n;
    // Verlet: each pass-1 launch of SE and RK4, and SRK's pass-2 launch, also
    // writes the tiles of the next stage (the next step's first stage after the
    // update), so a separate stage pass is needed only after a list build or
    // another call
    const bool fused = s.verlet;
    if (fused && !s.tile_fresh) build(s, nullptr, 0.0);
    if (method == SPC2_SE) {
      if (fused) {
        next_stage(s, 2, 0.0);
      } else {
        build(s, nullptr, 0.0);
      }
      rhs_pass1(s, EPI_SE, nullptr, nullptr, nullptr, sa, nullptr, nullptr, nullptr);
    } else if (method == SPC2_SRK) {
      if (!fused) build(s, nullptr, 0.0);
      sa.zero_in = 1;  // derivative zero-filled each step (StochasticRungeKuttaStepperPtr.hpp:43-46)
      // the derivative itself is read by nothing in the step (pass 2 takes a_phi
      // per slot and its own velocity), so pass 1 does not write it
      rhs_pass1(s, EPI_ITO1, nullptr, nullptr, nullptr, sa, nullptr, nullptr, nullptr);
      if (fused) next_stage(s, 2, 0.0);
      rhs_pass2(s, EPI_SRK, nullptr, nullptr, sa);
    } else {
      // RungeKutta4StepperPtr.cpp:25-43: k1 = f(s, k1, 0), k2 = f(s, k1, 1/2),
      // k3 = f(s, k2, 1/2), k4 = f(s, k3, 1), stage = s + (c * dt) * k
      double *const *k1 = s.k[0], *const *k2 = s.k[1], *const *k3 = s.k[2], *const *k4 = s.k[3];
      if (fused) {  // the first stage s + 0 k1 is the state itself
        // The epilogues form the next stage from k in registers, so k_1..k_3's
        // x / y components need not be stored: the combine rebuilds them from
        // the stage phases p, p + (dt / 2) k1_phi and p + (dt / 2) k2_phi, which
        // needs these stage coefficients to stay 0.5 dt (SPC2_RK4_STORE_KXY=1
        // stores and reads them instead).
        const bool kxy = s.rk4_store_kxy;
        double *const k1o[3] = {kxy ? k1[0] : nullptr, kxy ? k1[1] : nullptr, k1[2]};
        double *const k2o[3] = {kxy ? k2[0] : nullptr, kxy ? k2[1] : nullptr, k2[2]};
        double *const k3o[3] = {kxy ? k3[0] : nullptr, kxy ? k3[1] : nullptr, k3[2]};
        next_stage(s, 1, 0.5 * dt);
        rhs_pass1(s, EPI_K, k1o[0], k1o[1], k1o[2], sa, nullptr, nullptr, nullptr);
        next_stage(s, 1, 0.5 * dt);
        rhs_pass1(s, EPI_K, k2o[0], k2o[1], k2o[2], sa, nullptr, nullptr, nullptr);
        next_stage(s, 1, 1.0 * dt);
        rhs_pass1(s, EPI_K, k3o[0], k3o[1], k3o[2], sa, nullptr, nullptr, nullptr);
        next_stage(s, 2, 0.0);
        rhs_pass1(s, EPI_RK4_FINAL, k4[0], k4[1], k4[2], sa, k1o, k2o, k3o);
      } else {
        build(s, k1, 0.0 * dt);
        rhs_pass1(s, EPI_K, k1[0], k1[1], k1[2], sa, nullptr, nullptr, nullptr);
        build(s, k1, 0.5 * dt);
        rhs_pass1(s, EPI_K, k2[0], k2[1], k2[2], sa, nullptr, nullptr, nullptr);
        build(s, k2, 0.5 * dt);
        rhs_pass1(s, EPI_K, k3[0], k3[1], k3[2], sa, nullptr, nullptr, nullptr);
        build(s, k3, 1.0 * dt);
        rhs_pass1(s, EPI_RK4_FINAL, k4[0], k4[1], k4[2], sa, k1, k2, k3);
      }
    }
    if (fused) s.tile_fresh = true;
    if (s.verlet) {
      // SE and RK4 move a particle by at most v0 dt per step, SRK by at most
      // srk_step_reach with generated noise; with supplied noise SRK's
      // displacement is measured before the next use
      if (method == SPC2_SRK && sa.noise)
        s.drift = std::numeric_limits<double>::infinity();
      else
        s.drift += (method == SPC2_SRK ? srk_step_reach(s, dt) : std::fabs(s.P.v0) * dt) * (1.0 + 1e-12);
    }
    maybe_rehome(s);
    if (spc2_status st = check_launch("step")) return st;
  }
  HIP_TRY(hipStreamSynchronize(s.stream));
  return SPC2_OK;
}

spc2_status spc2_noise(spc2_sim_t h, int32_t method, int64_t step, double *out) {
  if (check_handle(h) || !out) return fail(SPC2_ERR_INVALID, "null argument");
  if (step < 0) return fail(SPC2_ERR_INVALID, "step must be >= 0");
  Sim &s = h->s;
  int npp = method == SPC2_SRK ? 2 : 1;
  double *buf = nullptr;
  HIP_TRY(dalloc(&buf, (size_t)npp * s.n));
  launch_noise(s, npp, (uint64_t)step, buf);
  spc2_status st = check_launch("noise");
  hipError_t e = st ? hipSuccess
                    : hipMemcpyAsync(out, buf, sizeof(double) * (size_t)npp * s.n, hipMemcpyDeviceToHost, s.stream);
  if (e == hipSuccess && !st) e = hipStreamSynchronize(s.stream);
  (void)hipFree(buf);
  if (st) return st;
  if (e != hipSuccess) return fail(SPC2_ERR_HIP, std::string("noise copy: ") + hipGetErrorString(e));
  return SPC2_OK;
}

// ---- exact checkpoints ---------------------------------------------------------
// Layout: CkptHeader, then by particle id the state (AoS x, y, phi), omega, the
// phase components of the RK4 stepper's k_1..k_4 (the only k components a step
// reads before writing them: a particle without partners keeps its phase
// derivative from the previous step) and, for Verlet handles, the build
// positions x, y of the current list.
namespace {
struct CkptHeader {
  char magic[8];
  uint32_t version;
  uint32_t has_list;  // the Verlet list's build positions and displacement bound are valid
  spc2_params params;
  double drift;
  int64_t n;
};
constexpr char kCkptMagic[8] = {'S', 'P', 'C', '2', 'C', 'K', 'P', 'T'};
constexpr uint32_t kCkptVersion = 1;

uint64_t ckpt_bytes(const Sim &s) {
  return sizeof(CkptHeader) + sizeof(double) * (uint64_t)s.n * (uint64_t)(3 + 1 + 4 + (s.verlet ? 2 : 0));
}

bool same_params(const spc2_params &a, const spc2_params &b) {
  return a.system == b.system && a.precision == b.precision && a.neighbour == b.neighbour && a.n == b.n &&
         a.Lx == b.Lx && a.Ly == b.Ly && a.rho == b.rho && a.v0 == b.v0 && a.sigma == b.sigma && a.alpha == b.alpha &&
         a.D == b.D && a.mu_plus == b.mu_plus && a.mu_minus == b.mu_minus && a.xi_a == b.xi_a && a.xi_r == b.xi_r &&
         a.kappa == b.kappa && a.seed == b.seed && a.verlet_skin == b.verlet_skin;
}
}  // namespace

spc2_status spc2_checkpoint_size(spc2_sim_t h, uint64_t *bytes) {
  if (check_handle(h) || !bytes) return fail(SPC2_ERR_INVALID, "null argument");
  *bytes = ckpt_bytes(h->s);
  return SPC2_OK;
}

spc2_status spc2_checkpoint_save(spc2_sim_t h, void *buf, uint64_t bytes) {
  if (check_handle(h) || !buf) return fail(SPC2_ERR_INVALID, "null argument");
  Sim &s = h->s;
  if (bytes < ckpt_bytes(s)) return fail(SPC2_ERR_INVALID, "checkpoint buffer smaller than spc2_checkpoint_size");
  CkptHeader hd;
  std::memset(&hd, 0, sizeof(hd));
  std::memcpy(hd.magic, kCkptMagic, sizeof(kCkptMagic));
  hd.version = kCkptVersion;
  hd.has_list = s.verlet && s.vbuilt ? 1u : 0u;
  hd.params = s.P;
  hd.drift = s.drift;
  hd.n = s.n;
  char *out = static_cast<char *>(buf);
  std::memcpy(out, &hd, sizeof(hd));
  out += sizeof(hd);
  const size_t n = (size_t)s.n;
  std::vector<double> tmp(3 * n);
  auto put = [&](size_t count) {
    std::memcpy(out, tmp.data(), sizeof(double) * count);
    out += sizeof(double) * count;
  };
  double *st[3] = {s.sx, s.sy, s.sp};
  if (spc2_status e = download_aos(s, tmp.data(), 3, st, 3)) return e;
  put(3 * n);
  if (spc2_status e = download_aos(s, tmp.data(), 1, &s.omega, 1)) return e;
  put(n);
  for (int j = 0; j < 4; ++j) {
    if (spc2_status e = download_aos(s, tmp.data(), 1, &s.k[j][2], 1)) return e;
    put(n);
  }
  if (s.verlet) {
    if (hd.has_list) {
      double *b[2] = {s.xb, s.yb};
      if (spc2_status e = download_aos(s, tmp.data(), 2, b, 2)) return e;
    } else {
      std::fill(tmp.begin(), tmp.begin() + 2 * n, 0.0);
    }
    put(2 * n);
  }
  return SPC2_OK;
}

spc2_status spc2_checkpoint_load(spc2_sim_t h, const void *buf, uint64_t bytes) {
  if (check_handle(h) || !buf) return fail(SPC2_ERR_INVALID, "null argument");
  Sim &s = h->s;
  CkptHeader hd;
  if (bytes < sizeof(hd)) return fail(SPC2_ERR_INVALID, "checkpoint too short");
  std::memcpy(&hd, buf, sizeof(hd));
  if (std::memcmp(hd.magic, kCkptMagic, sizeof(kCkptMagic)) != 0 || hd.version != kCkptVersion)
    return fail(SPC2_ERR_INVALID, "not an spc2 checkpoint (or another version)");
  if (hd.n != s.n || bytes < ckpt_bytes(s)) return fail(SPC2_ERR_INVALID, "checkpoint of another particle count");
  if (!same_params(hd.params, s.P))
    return fail(SPC2_ERR_INVALID, "checkpoint of a handle with other parameters (system, precision, neighbour "
                                  "mode, box, model parameters, seed or skin)");
  if (hd.has_list && !s.verlet) return fail(SPC2_ERR_INVALID, "checkpoint holds a Verlet list for a cell handle");
  const size_t n = (size_t)s.n;
  const char *in = static_cast<const char *>(buf) + sizeof(hd);
  auto take = [&](std::vector<double> &v, size_t count) {
    v.resize(count);
    std::memcpy(v.data(), in, sizeof(double) * count);
    in += sizeof(double) * count;
  };
  std::vector<double> state, omega, kphi[4], xyb;
  take(state, 3 * n);
  take(omega, n);
  for (auto &k : kphi) take(k, n);
  if (s.verlet) take(xyb, 2 * n);
  double *st[3] = {s.sx, s.sy, s.sp};
  if (hd.has_list) {
    // the list of the saving handle, rebuilt from its own build positions: the
    // same cells, homes, entries and summation order
    std::vector<double> pos(3 * n);
    for (size_t i = 0; i < n; ++i) {
      pos[3 * i] = xyb[2 * i];
      pos[3 * i + 1] = xyb[2 * i + 1];
      pos[3 * i + 2] = state[3 * i + 2];
    }
    if (spc2_status e = upload_aos(s, pos.data(), 3, st, 3)) return e;
    if (spc2_status e = verlet_build(s, nullptr, 0.0, true)) return e;
  } else {
    if (spc2_status e = upload_aos(s, state.data(), 3, st, 3)) return e;
    if (spc2_status e = rehome_now(s)) return e;
  }
  // everything by particle id into the homes the build (or re-homing) chose;
  // re-homing swaps the arrays with their spares, so the pointers are read again
  double *st2[3] = {s.sx, s.sy, s.sp};
  if (spc2_status e = upload_aos(s, state.data(), 3, st2, 3)) return e;
  if (spc2_status e = upload_aos(s, omega.data(), 1, &s.omega, 1)) return e;
  for (int j = 0; j < 4; ++j)
    if (spc2_status e = upload_aos(s, kphi[j].data(), 1, &s.k[j][2], 1)) return e;
  if (hd.has_list) s.drift = hd.drift;
  s.tile_fresh = false;
  s.built = false;
  HIP_TRY(hipStreamSynchronize(s.stream));
  return SPC2_OK;
}

spc2_status spc2_neighbour_counts(spc2_sim_t h, int32_t *n_rho, int32_t *n_xi_r) {
  if (check_handle(h) || !n_rho || !n_xi_r) return fail(SPC2_ERR_INVALID, "null argument");
  Sim &s = h->s;
  // pass 1 of the Ito form at the current state records each slot's counts
  // (its derivative output is not written)
  if (s.verlet)
    if (spc2_status st = verlet_ensure(s, 0.0, false)) return st;
  build(s, nullptr, 0.0);
  StepArgs sa = step_args(s, 0.0, 0);
  sa.zero_in = 1;
  rhs_pass1(s, EPI_ITO1, nullptr, nullptr, nullptr, sa, nullptr, nullptr, nullptr);
  // slots are homes on the Verlet and all-pairs paths, cell-sorted slots on the
  // linked-cell path
  const int *id_of_slot = (s.verlet || !s.linked) ? s.hid : s.slot_id;
  launch_slot_to_id(s, id_of_slot, s.n_slot, s.io[0]);
  launch_slot_to_id(s, id_of_slot, s.n2_slot, s.io[1]);
  if (spc2_status st = check_launch("neighbour counts")) return st;
  std::vector<double> a((size_t)s.n), b((size_t)s.n);
  HIP_TRY(hipMemcpyAsync(a.data(), s.io[0], sizeof(double) * (size_t)s.n, hipMemcpyDeviceToHost, s.stream));
  HIP_TRY(hipMemcpyAsync(b.data(), s.io[1], sizeof(double) * (size_t)s.n, hipMemcpyDeviceToHost, s.stream));
  HIP_TRY(hipStreamSynchronize(s.stream));
  for (int i = 0; i < s.n; ++i) {
    n_rho[i] = (int32_t)a[(size_t)i];
    n_xi_r[i] = (int32_t)b[(size_t)i];
  }
  return SPC2_OK;
}

spc2_status spc2_noise_noalloc(spc2_sim_t h, int32_t method, int64_t step, double *out) {
  if (check_handle(h) || !out) return fail(SPC2_ERR_INVALID, "null argument");
  if (step < 0) return fail(SPC2_ERR_INVALID, "step must be >= 0");
  Sim &s = h->s;
  const int npp = method == SPC2_SRK ? 2 : 1;
  const size_t need = (size_t)npp * (size_t)s.n;
  if (need > kScratchNormals)
    return fail(SPC2_ERR_INVALID, "spc2_noise_noalloc stages at most 2097152 normals (N <= 2097152 for SE, "
                                  "1048576 for SRK); this call needs " + std::to_string(need));
  double *buf = s.device >= 0 && s.device < kScratchDevices ? g_scratch[s.device] : nullptr;
  if (!buf) return fail(SPC2_ERR_INVALID, "no staging array for the handle's device");
  // the kernel writes every one of the `need` normals: nothing is cleared
  launch_noise(s, npp, (uint64_t)step, buf);
  if (spc2_status st = check_launch("noise")) return st;
  HIP_TRY(hipMemcpyAsync(out, buf, sizeof(double) * need, hipMemcpyDeviceToHost, s.stream));
  HIP_TRY(hipStreamSynchronize(s.stream));
  return SPC2_OK;
}

spc2_status spc2_order_parameter(spc2_sim_t h, double *R, double *psi) {
  if (check_handle(h) || !R || !psi) return fail(SPC2_ERR_INVALID, "null argument");
  Sim &s = h->s;
  double *out2 = s.red + 2 * 1024;
  launch_order_param(s, out2);
  if (spc2_status st = check_launch("order_parameter")) return st;
  double cs[2];
  HIP_TRY(hipMemcpyAsync(cs, out2, sizeof(cs), hipMemcpyDeviceToHost, s.stream));
  HIP_TRY(hipStreamSynchronize(s.stream));
  double cr = cs[0] / (double)s.n, ci = cs[1] / (double)s.n;
  *R = std::hypot(cr, ci);
  *psi = std::atan2(ci, cr);
  return SPC2_OK;
}

spc2_status spc2_get_cell_list(spc2_sim_t h, int32_t *cell_of, int32_t *cell_start, int32_t *sorted) {
  if (check_handle(h) || !cell_of || !cell_start || !sorted) return fail(SPC2_ERR_INVALID, "null argument");
  Sim &s = h->s;
  if (!s.linked) return fail(SPC2_ERR_INVALID, "no cell list on the all-pairs path");
  if (s.verlet) return fail(SPC2_ERR_INVALID, "no per-RHS cell list on the Verlet path");
  build(s, nullptr, 0.0);
  if (spc2_status st = check_launch("cell list")) return st;
  launch_from_home_int(s, s.cell_of, s.spare_i);
  HIP_TRY(hipMemcpyAsync(cell_of, s.spare_i, sizeof(int) * (size_t)s.n, hipMemcpyDeviceToHost, s.stream));
  HIP_TRY(hipMemcpyAsync(cell_start, s.start, sizeof(int) * ((size_t)s.nc + 1), hipMemcpyDeviceToHost, s.stream));
  HIP_TRY(hipMemcpyAsync(sorted, s.slot_id, sizeof(int) * (size_t)s.n, hipMemcpyDeviceToHost, s.stream));
  HIP_TRY(hipStreamSynchronize(s.stream));
  return SPC2_OK;
}

spc2_status spc2_verlet_stats(spc2_sim_t h, int64_t *builds, int64_t *entries) {
  if (check_handle(h) || !builds || !entries) return fail(SPC2_ERR_INVALID, "null argument");
  Sim &s = h->s;
  if (!s.verlet) return fail(SPC2_ERR_INVALID, "the handle does not use Verlet lists");
  if (s.ventries < 0 && s.vbuilt) {
    std::vector<int> c((size_t)s.n);
    HIP_TRY(hipMemcpyAsync(c.data(), s.vcnt, sizeof(int) * c.size(), hipMemcpyDeviceToHost, s.stream));
    HIP_TRY(hipStreamSynchronize(s.stream));
    long long t = 0;
    for (int v : c) t += v;
    s.ventries = t;
  }
  *builds = s.vbuilds;
  *entries = s.vbuilt ? s.ventries : 0;
  return SPC2_OK;
}

spc2_status spc2_verlet_layout(spc2_sim_t h, int32_t *lds_staged, int32_t *lds_capacity) {
  if (check_handle(h) || !lds_staged || !lds_capacity) return fail(SPC2_ERR_INVALID, "null argument");
  Sim &s = h->s;
  if (!s.verlet) return fail(SPC2_ERR_INVALID, "the handle does not use Verlet lists");
  *lds_staged = s.vbuilt && s.vlds ? 1 : 0;
  *lds_capacity = s.vbuilt && s.vlds ? s.vlds_cap : 0;
  return SPC2_OK;
}

spc2_status spc2_device_memory(uint64_t *free_bytes, uint64_t *total_bytes) {
  if (!free_bytes || !total_bytes) return fail(SPC2_ERR_INVALID, "null argument");
  size_t f = 0, t = 0;
  HIP_TRY(hipMemGetInfo(&f, &t));
  *free_bytes = f;
  *total_bytes = t;
  return SPC2_OK;
}

spc2_status spc2_num_cells(spc2_sim_t h, int32_t *nx, int32_t *ny) {
  if (check_handle(h) || !nx || !ny) return fail(SPC2_ERR_INVALID, "null argument");
  *nx = h->s.nx;
  *ny = h->s.ny;
  return SPC2_OK;
}

spc2_status spc2_set_counting(spc2_sim_t h, int32_t enable) {
  if (check_handle(h)) return SPC2_ERR_INVALID;
  Sim &s = h->s;
  s.counting = enable != 0;
  HIP_TRY(hipMemsetAsync(s.counts, 0, sizeof(unsigned long long) * 3 * kCountSlots, s.stream));
  HIP_TRY(hipStreamSynchronize(s.stream));
  return SPC2_OK;
}

spc2_status spc2_get_counts(spc2_sim_t h, int64_t *counts3) {
  if (check_handle(h) || !counts3) return fail(SPC2_ERR_INVALID, "null argument");
  Sim &s = h->s;
  std::vector<unsigned long long> c(3 * (size_t)kCountSlots);
  HIP_TRY(hipMemcpyAsync(c.data(), s.counts, sizeof(unsigned long long) * c.size(), hipMemcpyDeviceToHost, s.stream));
  HIP_TRY(hipStreamSynchronize(s.stream));
  for (int i = 0; i < 3; ++i) counts3[i] = 0;
  for (size_t q = 0; q < c.size(); ++q) counts3[q % 3] += (int64_t)c[q];
  return SPC2_OK;
}

spc2_status spc2_set_profiling(spc2_sim_t h, int32_t enable) {
  if (check_handle(h)) return SPC2_ERR_INVALID;
  Sim &s = h->s;
  s.prof_collect();
  s.profiling = enable != 0;
  s.prof_ms = 0.0;
  s.prof_launches = 0;
  return SPC2_OK;
}

spc2_status spc2_get_profile(spc2_sim_t h, double *pair_kernel_ms, int64_t *launches) {
  if (check_handle(h) || !pair_kernel_ms || !launches) return fail(SPC2_ERR_INVALID, "null argument");
  h->s.prof_collect();
  *pair_kernel_ms = h->s.prof_ms;
  *launches = h->s.prof_launches;
  return SPC2_OK;
}

spc2_status spc2_snapshot_async(spc2_sim_t h, float *dst) {
  if (check_handle(h) || !dst) return fail(SPC2_ERR_INVALID, "null argument");
  Sim &s = h->s;
  if (!s.snap) HIP_TRY(dalloc(&s.snap, 3 * (size_t)s.n));
  if (!s.snap_ev) HIP_TRY(hipEventCreateWithFlags(&s.snap_ev, hipEventDisableTiming));
  launch_snapshot(s, s.snap);
  if (spc2_status st = check_launch("snapshot")) return st;
  HIP_TRY(hipMemcpyAsync(dst, s.snap, sizeof(float) * 3 * (size_t)s.n, hipMemcpyDeviceToHost, s.stream));
  HIP_TRY(hipEventRecord(s.snap_ev, s.stream));
  s.snap_pending = true;
  return SPC2_OK;
}

spc2_status spc2_snapshot_wait(spc2_sim_t h) {
  if (check_handle(h)) return SPC2_ERR_INVALID;
  Sim &s = h->s;
  if (s.snap_pending) {
    HIP_TRY(hipEventSynchronize(s.snap_ev));
    s.snap_pending = false;
  }
  return SPC2_OK;
}

spc2_status spc2_host_alloc(uint64_t bytes, void **out) {
  if (!out) return fail(SPC2_ERR_INVALID, "null argument");
  *out = nullptr;
  HIP_TRY(hipHostMalloc(out, std::max<uint64_t>(bytes, 1), hipHostMallocDefault));
  return SPC2_OK;
}

spc2_status spc2_host_free(void *p) {
  if (p) HIP_TRY(hipHostFree(p));
  return SPC2_OK;
}

spc2_status spc2_synchronize(spc2_sim_t h) {
  if (check_handle(h)) return SPC2_ERR_INVALID;
  HIP_TRY(hipStreamSynchronize(h->s.stream));
  return SPC2_OK;
}

}  // extern "C"
