// Host-side internals shared by the spc2 HIP translation units: the device
// buffers of one simulation and the launchers each kernel file provides.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <utility>
#include <vector>

#include "../../include/spc2_hip.h"

namespace spc2 {

enum Prec : int { P64 = 0, P32K = 1, P64F = 2 };

// persistent per-particle double arrays moved by re-homing: x, y, phi, omega, k_1..k_4
constexpr int kPermuteArrays = 16;
// slot triples of the pair counters (candidates, pairs within rho, cohesion)
constexpr int kCountSlots = 1024;

// What a force kernel does with the per-particle result (its "epilogue").
enum Epi : int {
  EPI_K = 0,          // k-form: write k_next (phi only when a partner was found)
  EPI_ITO1 = 1,       // Ito pass 1: write derivative; keep a_phi per slot for pass 2
  EPI_SE = 2,         // pass 1 + Stochastic Euler update + wrap
  EPI_RK4_FINAL = 3,  // 4th RK4 stage + combine + wrap
  EPI_ITO2 = 4,       // Ito pass 2: write additional
  EPI_SRK = 5,        // pass 2 + stochastic Runge-Kutta update + wrap
};

// Scalars every force kernel needs (one copy per launch, by value).
struct KParams {
  int n, nx, ny, linked;
  double Lx, Ly, rLx, rLy, hx, hy;
  float hxf, hyf;
  double rho, rho2;
  float rho2f;
  double v0, sigma, alpha, D;
  // vortex
  double mu_plus, mu_minus, xi_a, xi_r, kappa;
  double outer2;  // max(rho, xi_r)^2: widest test radius
  float xa2f, xr2f, outer2f, xaf, rhof, mu_pf, mu_p_over_xa2f, c4f, kappaf, v0f, Df;
  // outer W(r) = c4 (r - xi_a)(rho - r) in Horner form r (w_a - c4 r) + w_b,
  // w_a = c4 (xi_a + rho), w_b = -c4 xi_a rho
  float w_af, w_bf;
  // FP32 cutoff band: |r^2 - cutoff^2| <= band * cutoff^2 is re-decided in FP64
  float band_rho2f, band_xr2f, band_xa2f;
  float rho2f_lo, rho2f_hi, xr2f_hi;  // rho2f -/+ band_rho2f, xr2f + band_xr2f
  // FP32 cohesion pairs closer than this (r^2) take FP64 geometry
  float close2f;
  // pair-kernel decomposition, segs workgroups per cell row: row chunks of
  // `chunk` slots (chunk > 0; the row's last workgroup takes the rest of the
  // row) or, with chunk 0, row segments of seg_w cells; LDS capacity in tile
  // elements per workgroup
  int seg_w, segs, lds_cap, chunk;
};

// Tile arrays in cell-sorted slot order, written by the rank kernel.
struct Tiles {
  double *x, *y, *p, *c, *s;  // FP64 forms: wrapped stage position, stage phase, cos/sin
  float4 *f4;                 // FP32 form: cell-local x, y and cos, sin of the phase
  int *cell;                  // cell of each slot
};

struct StepArgs {
  double dt, sqdt, sq2D, h1, sqrt3, D;  // h1 = 0.5 * dt * sqrt(dt)
  uint64_t seed, step;
  const double *noise;  // optional materialised noise for this step (by particle)
  const double *omega;  // natural frequencies (chimera) or null
  int zero_in;          // 1: the phi output starts from 0 (zero-filled derivative), else in/out
  int wrap;             // 1: apply the engine's periodic wrap after the update
};

struct Sim {
  spc2_params P;
  int prec = P64;
  int n = 0, nx = 0, ny = 0, nc = 0;
  bool linked = true;
  KParams kp{};
  hipStream_t stream = nullptr;
  int device = 0;

  // master state, FP64 SoA, original particle order
  double *sx = nullptr, *sy = nullptr, *sp = nullptr, *omega = nullptr;
  // k buffers (RK4 k_1..k_4, also the k-form in/out), SoA by particle
  double *k[4][3] = {};
  // Ito outputs
  double *der[3] = {}, *add[2] = {};
  // k_prev / k_next of the k-form evaluation entry point, allocated at its first call
  double *kev[2][3] = {};
  // stage positions (wrapped for binning when linked), by particle
  double *gx = nullptr, *gy = nullptr, *gp = nullptr;
  // Particle arrays (state, omega, k, derivative, stage) are stored in "home"
  // order h: hid[h] is the particle id at home h, hpos[id] its home.  Homes are
  // re-assigned every rehome_every steps to the current cell-sorted order so
  // that per-particle reads and writes stay coalesced.  hpos is rebuilt from
  // hid only when a transfer by id needs it (hpos_stale after a re-homing).
  int *hid = nullptr, *hpos = nullptr;
  bool hpos_stale = false;
  // cell list (per home: cell_of, local; per slot: tmp/tmpid while placing,
  // then slot_h / slot_id = home / id of the particle in each cell-sorted slot)
  int *cell_of = nullptr, *count = nullptr, *local = nullptr, *start = nullptr, *tmp = nullptr,
      *tmpid = nullptr, *slot_h = nullptr, *slot_id = nullptr;
  bool built = false;  // slot_h holds the cell order of the current homes
  int rehome_every = 16, steps_since_rehome = 0;
  // re-homing double buffers (state x, y, phi, omega, k_1..k_4) and id-order
  // transfer scratch for the ABI's host arrays
  double *spare[16] = {};
  int *spare_i = nullptr;
  double *io[3] = {};
  Tiles t{};
  // Verlet: the next stage's tiles and phases, written by a pass-1 epilogue
  // (t2.cell aliases t.cell); swapped with t / gp after that launch
  Tiles t2{};
  double *gp2 = nullptr;
  int vnext = 0;            // next launch's stage write: 0 none, 1 s + vnext_kcdt k, 2 the new state
  double vnext_kcdt = 0.0;
  bool tile_fresh = false;  // t / gp hold the tiles of the current state
  // pass-1 results kept per slot for pass 2
  double *aphi_slot = nullptr, *n_slot = nullptr, *n2_slot = nullptr;
  float *aphi_slot_f = nullptr;
  // materialised noise (optional)
  double *noise = nullptr;
  size_t noise_cap = 0;
  // counts: candidates, interactions, cohesion pairs
  unsigned long long *counts = nullptr;
  bool counting = false;
  // Verlet lists: the build grid is nx x ny (cells of side >= rlist); homes are
  // re-assigned to the build's cell order at each build
  bool verlet = false, vbuilt = false;
  double skin = 0.0, rlist = 0.0;
  double drift = 0.0;  // bound on the largest |state - build position| (inf: unknown)
  int nw = 0;          // wavefronts of 64 homes
  // wmax[w] / wrow[w]: wavefront w's group count and first group (see the list
  // layout in spc2_verlet.hip); vlist_cap in ints
  int *vcnt = nullptr, *wmax = nullptr, *wrow = nullptr, *vlist = nullptr;
  // candidate test bits of the count pass, read back by the fill (words by
  // word-major [w * n + i]; homes with more candidates are tested again)
  uint32_t *vmask = nullptr;
  int vmask_words = 0;
  // LDS-staged FP32 passes: per pass workgroup (kBlock homes) its halo as runs of
  // consecutive build cells (bh_run: first and last cell, bh_off: LDS offsets,
  // bh_nrun: run count or -1), bh_max: the largest halo; vlds when the current
  // list holds halo indices (eight 16-bit entries per group: wmax8 / wrow8 are
  // the wavefront group counts and offsets in groups of eight)
  bool vlds = false, vlds_allowed = true;
  int nblk = 0, vlds_cap = 0;
  int *bh_nrun = nullptr, *bh_run = nullptr, *bh_off = nullptr, *bh_max = nullptr;
  int *wmax8 = nullptr, *wrow8 = nullptr;
  // LDS layout lane slots (vslot[slot] = home) and their wavefront group counts
  // and offsets (groups of eight); while vlds, wrow holds the offsets of the
  // layout in use
  int *vslot = nullptr, *wmax_s = nullptr, *wrow_s = nullptr;
  bool vsorted = false;  // the current LDS-layout list uses the lane slots
  // chimera FP32 LDS layout: pass 1 of an Ito pair stores each entry's
  // within-rho decision (a byte per lane and group, the list's layout) for the
  // pass 2 that follows (vrho_fresh: the last pass-1 launch wrote them);
  // vrho_allowed: SPC2_VERLET_RHO_BITS
  uint8_t *vrho = nullptr;
  size_t vrho_cap = 0;
  bool vrho_fresh = false, vrho_allowed = true;
  int vsort_mode = -1;   // -1: by list length, 0: never, 1: always (SPC2_VERLET_SORT)
  // fused RK4 steps store k_1..k_3's x / y components instead of rebuilding
  // them in the combine (SPC2_RK4_STORE_KXY)
  bool rk4_store_kxy = false;
  size_t vlist_cap = 0;
  double *xb = nullptr, *yb = nullptr;  // build positions by home
  long long vbuilds = 0, ventries = 0;
  // float frame buffer for asynchronous snapshots and its completion event
  float *snap = nullptr;
  hipEvent_t snap_ev = nullptr;
  bool snap_pending = false;
  // scan workgroup totals, reduction scratch
  int *bsum = nullptr;
  double *red = nullptr;
  // profiling: HIP events around the pair kernels, read lazily
  bool profiling = false;
  double prof_ms = 0.0;
  long long prof_launches = 0;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> prof_pending, prof_free;
  void prof_take(hipEvent_t &b, hipEvent_t &e) {
    if (!prof_free.empty()) {
      b = prof_free.back().first;
      e = prof_free.back().second;
      prof_free.pop_back();
      return;
    }
    (void)hipEventCreate(&b);
    (void)hipEventCreate(&e);
  }
  // folds the finished event pairs into prof_ms (waits for the last one)
  void prof_collect() {
    if (!prof_pending.empty()) (void)hipEventSynchronize(prof_pending.back().second);
    for (auto &pr : prof_pending) {
      float ms = 0.0f;
      if (hipEventElapsedTime(&ms, pr.first, pr.second) == hipSuccess) {
        prof_ms += ms;
        prof_launches += 1;
      }
      prof_free.push_back(pr);
    }
    prof_pending.clear();
  }
};

// ---- launchers -----------------------------------------------------------------
// spc2_api.hip
spc2_status set_error(spc2_status st, const std::string &msg);
int precision_code(int32_t p);
// linked cells or all pairs, and the cell grid, as the reference decides them
// (ChimeraSystem.cpp:32-33, :90; VortexArraysSystemPtr.cpp:35-36, :113)
spc2_status cell_grid(const spc2_params &p, bool &linked, int &nx, int &ny);
void fill_kparams_from(KParams &k, const spc2_params &P, int n, int nx, int ny, bool linked);
// spc2_cells.hip
void launch_stage_bin(Sim &s, const double *kx, const double *ky, const double *kp, double kcdt);
int scan_blocks(int nc);
void launch_scan(Sim &s);
void launch_scan_array(Sim &s, int *count, int m, int *out);
void launch_scan_raw(hipStream_t st, int *bsum, int *count, int m, int *out);
void launch_place_raw(hipStream_t st, int n, const int *cell_of, const int *local, const int *start, const int *hid,
                      int *tmp, int *tmpid);
void launch_place(Sim &s);
void launch_rank_tiles(Sim &s);
// spc2_pairs.hip
constexpr double kSegTarget = 256.0;  // default particles per pair workgroup (row chunk or segment)
constexpr int kLdsBytes = 60 * 1024;  // dynamic LDS of one pair workgroup at most
int lds_elem_bytes(int prec);
void configure_segments_kp(KParams &k, int prec, double target);
void configure_segments(Sim &s);
void launch_pass1(Sim &s, int epi, double *ox, double *oy, double *op, const StepArgs &sa, double *const k1[3],
                  double *const k2[3], double *const k3[3]);
void launch_pass2(Sim &s, int epi, double *a0, double *a1, const StepArgs &sa);
// spc2_verlet.hip
// valid groups (64 lanes x 4 entries) kept after a Verlet list for the
// passes' read-ahead
constexpr int kVerletSpareGroups = 2;
// halo: the workgroups' halo runs are current (launch_block_halo ran since the re-homing)
void launch_verlet_count(Sim &s, bool halo);
void launch_verlet_fill(Sim &s);
constexpr int kVerletRunMax = 16;  // halo runs per pass workgroup (kRunMax of spc2_verlet.hip)
void launch_block_halo(Sim &s);
// the LDS layout's pass kernels of this system start their dynamic LDS at address 0
bool verlet_lds_unbased(int system);
int verlet_lds_tile_bytes(int prec);
void launch_verlet_fill_lds(Sim &s);
void launch_verlet_order(Sim &s);
void launch_verlet_stage(Sim &s, const double *kx, const double *ky, const double *kp, double kcdt);
void launch_verlet_pass1(Sim &s, int epi, double *ox, double *oy, double *op, const StepArgs &sa,
                         double *const k1[3], double *const k2[3], double *const k3[3]);
void launch_verlet_pass2(Sim &s, int epi, double *a0, double *a1, const StepArgs &sa);
void launch_drift(Sim &s, double *out);
// spc2_misc.hip
void launch_to_home(Sim &s, const double *src_id, double *dst_home);
void launch_from_home(Sim &s, const double *src_home, double *dst_id);
void launch_from_home_int(Sim &s, const int *src_home, int *dst_id);
void launch_slot_to_id(Sim &s, const int *id_of_slot, const double *src_slot, double *dst_id);
void launch_identity_homes(Sim &s);
void launch_snapshot(Sim &s, float *out);
// Re-homes the particle arrays to the last build's cell order; k_xy_live =
// false leaves the x / y components of k_1..k_4 behind (dead between steps).
void launch_rehome(Sim &s, bool k_xy_live = true);
void launch_order_param(Sim &s, double *out2);
void launch_init_state(Sim &s, double omega_std);
void launch_noise(Sim &s, int per_particle, uint64_t step, double *out);
void launch_fill(Sim &s, double *p, double v, size_t n);

}  // namespace spc2
