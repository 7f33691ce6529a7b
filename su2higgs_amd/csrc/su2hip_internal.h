// su2hip_internal.h -- device data structures shared by the kernels and the
// C ABI layer of libsu2hip.so.
//
// Device layout of one lattice (checkerboarded).  With Vh = V/2 sites per
// parity and s the index of a site inside its parity (s = natural_index >> 1
// for even L0, which is exactly the reference's "even sites first, order
// preserved" site order, src/layout.c:826-851):
//   links    U[d][p][s][c]   d<3, c<4    (24 Vh words; a 4-vector per link)
//   doublet  F[p][s][c]      c<4         (8 Vh; a 4-vector per site)
//   doublet2 F2[p][s][c]     c<4         (8 Vh; the second doublet, nhiggs = 2)
//   U(1)     A[d][p][s]                  (6 Vh)
//   triplet  G[c][p][s]      c<3         (6 Vh)
//   singlet  S[p][s]                     (2 Vh)
//   backup   B: the doublet's or the triplet's layout
// so the 64 lanes of a wavefront, which own 64 consecutive s of one parity,
// read one contiguous 1 KiB (FP32) segment per link or doublet load and one
// contiguous segment per component of the other fields.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/su2hip.h"

namespace su2 {

constexpr int BLK = 256;          // threads per workgroup in every site kernel
constexpr int NCOUNT = 12;        // counter rows: pass kinds SU2HIP_K_LINK_HEATBATH .. SINGLET_METROPOLIS,
                                  // then the second doublet's two kinds
constexpr int CNT_DOUBLET2_OR = 10;
constexpr int CNT_DOUBLET2_METRO = 11;
constexpr uint32_t MUCA_SITE = 0xFFFFFFFFu;
constexpr uint64_t HOTSTART_PASS = 0xFFFFFFFFFFFFFFFFull;
constexpr uint64_t WALL_PASS = 0xFFFFFFFFFFFFFFFEull;
constexpr uint32_t SCHED_SITE = 0xFFFFFFFEu;  // site of a sweep's pass-order draws

// Step of a sweep for the random pass order: bit 15 set for a scalar
// round, bits 8-14 the round, bits 0-7 the position (0-5 in a link round's
// parity/direction arrays, 0-1 in a scalar round's parities)
constexpr uint16_t STEP_SCALAR = 0x8000;
inline uint16_t step_info(bool scalar, int round, int pos) {
  return (uint16_t)((scalar ? STEP_SCALAR : 0) | ((round & 0x7F) << 8) | (pos & 0xFF));
}

// model bits; M_H2 (two doublets) always comes with M_NH
enum : int { M_NH = 1, M_U1 = 2, M_TR = 4, M_SG = 8, M_H2 = 16 };
// not field content: FP32 links stored in the packed 12-byte form (ldU)
constexpr int M_PK = 64;

inline int model_bits(const su2hip_params &p) {
  return (p.nhiggs ? M_NH : 0) | (p.nhiggs == 2 ? M_H2 : 0) | (p.u1 ? M_U1 : 0) |
         (p.triplet ? M_TR : 0) | (p.singlet ? M_SG : 0);
}

// fast division by a runtime constant (Granlund-Montgomery, round-up
// variant): q = (umulhi(m, n) + n) >> l for n < 2^31
struct FastDiv {
  uint32_t m, l, d;
};

inline FastDiv make_fastdiv(uint32_t d) {
  FastDiv f;
  f.d = d;
  uint32_t l = 0;
  while ((1ull << l) < d) l++;
  f.l = l;
  f.m = (uint32_t)((((1ull << l) - d) << 32) / d + 1);
  return f;
}

__host__ __device__ inline uint32_t fdiv(uint32_t n, const FastDiv &f) {
#ifdef __HIP_DEVICE_COMPILE__
  uint32_t t = __umulhi(f.m, n);
#else
  uint32_t t = (uint32_t)(((uint64_t)f.m * n) >> 32);
#endif
  return (uint32_t)(((uint64_t)t + n) >> f.l);
}

// Per-lattice descriptor, resident in device memory.
struct LatDesc {
  void *U, *A, *F, *G, *S, *B;  // field arrays (element type = precision)
  void *F2;                     // second doublet (nhiggs = 2), else null
  // one doublet, when the handle keeps the cache: the doublet's hopping
  // term of every link staple, H_mu(x) (k_hop), 4 values per link in the
  // link layout; else null
  void *H;
  int L0, L1, L2, h0;           // extents, h0 = L0/2
  int Vh, vol;                  // sites per parity, volume
  FastDiv div_h0, div_L1;
  int active;
  int pad0;
  int64_t budget;               // sweeps left (-1 unlimited, 0 = skip)
  uint64_t pass;                // Philox pass counter at the start of the sweep
  uint32_t key0, key1;          // Philox key
  int blk_base;                 // first logical block of this lattice
  int nblk;                     // logical blocks of this lattice
  double beta, betau1, r_u1, msq, lambda, msq_t, a2, b4;
  double b1_s, msq_s, b3_s, b4_s, a1_s, a2_s;
  // two-doublet potential (src/su2u1.c:537-554)
  double msq2, m12re, m12im, lam2, lam3, lam4;
  double lam5re, lam5im, lam6re, lam6im, lam7re, lam7im;
};

// Multicanonical state of one lattice on the device.
struct MucaDev {
  int enabled, orderparam, mode, cps;
  int bins, last_max, do_acceptance, muca_count;
  int update_interval, rejected;
  uint32_t check;
  int pad;
  double delta, min, max, wrk_min, wrk_max;
  double pv[2];
  long long total, accepted;
  double *pos, *W, *slope, *b, *hgram;
  int *hits;
  long long *gsum, *nsum;
};

// One logical chunk of BLK sites of one lattice.
struct BlockMap {
  int lat;
  int s0;
};

// ---- launch interface (implemented in su2hip_kernels.hip) ----
struct LaunchCtx {
  hipStream_t stream;
  LatDesc *desc;       // device, [max_slots]
  MucaDev *muca;       // device, [max_slots]
  const BlockMap *blk; // device, [nblk]
  int nblk;
  long long *cnt;      // device, [NCOUNT][2][nblk]
  double *partial;     // device, [nblk][SU2HIP_NSUMS]
  double *sums;        // device, [max_slots][SU2HIP_NSUMS]
  int max_slots;
  int model;
  int fp64;
  // when set, pass kernels are launched with hipExtLaunchKernel and these
  // events take the kernel's own start and end timestamps
  hipEvent_t ev_start = nullptr, ev_stop = nullptr;
  // threads per workgroup of pass kernels = sites per block-table chunk
  // (64, 128 or 256); blk/nblk/cnt then belong to the table of that chunk
  int block = BLK;
  // occupancy target of the pass kernel: 0 = the compiler's own register
  // allocation, 5 = registers capped for 5 waves per SIMD (link
  // overrelaxation only; see su2hip_set_waves)
  int waves = 0;
  // random pass order: per lattice and sweep step, parity | direction << 1
  // ([max_slots][npass]); null = the launch's parity and direction
  const uint8_t *order = nullptr;
  int npass = 0;
};

// One checkerboard pass (dir: the direction of a link kind, the doublet of
// a doublet kind).  muca_step < 0: plain pass; otherwise the pass is step
// `muca_step` of a multicanonical parity sweep.
hipError_t launch_pass(const LaunchCtx &c, int kind, int parity, int dir,
                       int pass_off, int muca_step, int reunit);
// the doublet's hopping term of every link staple (one doublet; a no-op
// otherwise), read by the link passes that follow until the doublet or the
// U(1) links change
hipError_t launch_hop(const LaunchCtx &c);
// multicanonical global accept/reject after a scalar pass step (hid: the
// doublet of a doublet kind)
hipError_t launch_muca_accept(const LaunchCtx &c, int kind, int hid, int parity,
                              int pass_off, int muca_step);
// random pass order of the coming sweep of every active lattice
// (src/update.c:292-300, :341-345) into order[max_slots][npass], from the
// lattice's Philox stream at (pass counter, SCHED_SITE)
hipError_t launch_order(const LaunchCtx &c, uint8_t *order, const uint16_t *info, int npass,
                        int link_rounds, int scalar_rounds);
// sweep epilogue: pass counter += passes, budget -= 1
hipError_t launch_end_sweep(const LaunchCtx &c, int passes);
hipError_t launch_measure(const LaunchCtx &c);
// one lattice's measure sums without the block table, partials of
// ceil(Vh / BLK) blocks in `partial` (su2hip_measure_noalloc)
hipError_t launch_measure_one(hipStream_t stream, int fp64, int model, const LatDesc &L,
                              double *partial, double *sums);
// both parities' order-parameter contributions of lattice `only` (< 0: all)
hipError_t launch_orderparam_init(const LaunchCtx &c, int only = -1);
hipError_t launch_cold(const LaunchCtx &c, int slot, double phi0, double sigma0,
                       double singlet0);
hipError_t launch_hot(const LaunchCtx &c, int slot);
hipError_t launch_to_device_layout(const LaunchCtx &c, int slot, int field,
                                   const double *host_layout_dev);
hipError_t launch_to_host_layout(const LaunchCtx &c, int slot, int field,
                                 double *host_layout_dev);
hipError_t launch_copy_probe(const double *src, double *dst, size_t n,
                             hipStream_t s);
// gradient flow on one lattice (slot) whose descriptor in c.desc points at
// the flow buffers: stage 0 = Euler, 1..3 = RK3 stages
// correlator plane sums of lattice `slot` along dir: buf holds 10 V
// doubles, planes 10 Ld
hipError_t launch_corr(const LaunchCtx &c, int slot, int dir, int Vh, int Ld, size_t P,
                       double *buf, double *planes);
// z-profile plane sums (4 values) of lattice `slot` along dir
hipError_t launch_zprof(const LaunchCtx &c, int slot, int dir, int Vh, int Ld, size_t P,
                        double *buf, double *planes);
// one blocking level: smeared fields of `fine` into `blocked` (bmask: the
// blocked directions), both descriptors by value
hipError_t launch_block(const LaunchCtx &c, const LatDesc &fine, const LatDesc &blocked, int bmask);
// two-phase wall start of lattice `slot` with the interface across zdir
hipError_t launch_wall(const LaunchCtx &c, int slot, int zdir);
hipError_t launch_flow_stage(const LaunchCtx &c, int slot, const void *Uin,
                             void *Uout, const void *Gin, void *Gout, void *Z,
                             void *ZG, double dt, double msq_t, int stage);
hipError_t launch_flow_force(const LaunchCtx &c, int slot, double *force);

}  // namespace su2
