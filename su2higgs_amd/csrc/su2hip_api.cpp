// su2hip_api.cpp -- the C ABI of libsu2hip.so (include/su2hip.h): lattice
// slots in HBM, the pass schedule of one compound sweep, hipGraph replay,
// measurement, multicanonical set-up, gradient flow and timing.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <string>
#include <vector>

#include <rocprofiler-sdk-roctx/roctx.h>

#include "su2hip_internal.h"

using namespace su2;

static thread_local std::string g_err;

static int fail(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define HIPTRY(x)                                                                    \
  do {                                                                               \
    hipError_t e_ = (x);                                                             \
    if (e_ != hipSuccess)                                                            \
      return fail(SU2HIP_EDEVICE, "%s failed: %s", #x, hipGetErrorString(e_));       \
  } while (0)

namespace {

// A gradient flow in flight: device descriptors, private measurement
// buffers and the rows it measures, on the stream it was enqueued on.
struct FlowJob {
  bool active = false;
  LatDesc *ddesc = nullptr;
  BlockMap *dblk = nullptr;
  double *d_partial = nullptr, *d_sums = nullptr, *d_rows = nullptr;
  std::vector<double> t;  // flow time of each row
  hipEvent_t done = nullptr;
};

struct Slot {
  bool used = false;
  su2hip_params p{};
  LatDesc hd{};
  void *mem = nullptr;
  size_t bytes = 0;
  su2hip_counters cnt{};
  void *bu_mem = nullptr;
  size_t bu_bytes = 0;
  void *muca_mem = nullptr;
  size_t muca_bytes = 0;
  MucaDev hm{};
  void *flow_mem = nullptr;
  size_t flow_bytes = 0;
  void *hop_mem = nullptr;  // hopping-term cache (LatDesc::H), when the handle keeps one
  size_t hop_bytes = 0;
  FlowJob fj;
};

struct Step {
  int kind, parity, dir, pass_off;
  bool muca;
  uint16_t info;  // place in the random pass order (step_info)
};

struct Ev {
  int kind;
  hipEvent_t a, b;
};

// Block table of pass kernels launched with a smaller workgroup (64 or 128
// sites per chunk), with its own per-block counters.
struct AltMap {
  int chunk = 0;
  BlockMap *d_blk = nullptr;
  long long *d_cnt = nullptr;
  int nblk = 0, cap = 0;
  std::vector<BlockMap> h;
};

constexpr int NCHUNK = 3;
constexpr int CHUNKS[NCHUNK] = {64, 128, 256};

}  // namespace

struct su2hip_batch {
  int max_slots = 0;
  uint64_t hbm_budget = 0, used = 0;
  int model = 0, fp64 = 1;
  su2hip_params tmpl{};
  std::vector<Slot> slots;
  hipStream_t stream = nullptr;
  LatDesc *d_desc = nullptr;
  MucaDev *d_muca = nullptr;
  BlockMap *d_blk = nullptr;
  int nblk = 0, nblk_cap = 0;
  std::vector<BlockMap> h_blk;
  long long *d_cnt = nullptr;
  double *d_partial = nullptr;
  size_t partial_cap = 0;
  double *d_sums = nullptr;
  double *d_stage = nullptr;
  size_t stage_cap = 0;
  bool map_dirty = true;
  std::vector<Step> sched;
  int passes = 0, muca_steps = 1;
  bool sched_dirty = true;
  bool use_graph = false;
  hipGraphExec_t gexec = nullptr;
  bool graph_valid = false;
  bool timing = false;
  std::vector<Ev> pending;
  std::vector<hipEvent_t> pool;
  double kms[SU2HIP_NKINDS] = {};
  long long kcount[SU2HIP_NKINDS] = {};
  AltMap alt[2];                  // chunks 64 and 128; 256 is the main table
  hipStream_t flow_stream = nullptr;  // asynchronous flows (su2hip_flow_start)
  int tune[SU2HIP_NKINDS] = {};   // chunk per pass kind, 0 = chosen by size
  int waves[SU2HIP_NKINDS] = {};  // occupancy target per pass kind (su2hip_set_waves)
  bool random_order = false;      // the reference's random_sweeps
  bool hop_on = false;            // hopping-term cache kept (su2hip_set_hop_cache)
  uint8_t *d_order = nullptr;     // [max_slots][passes] parity | dir << 1
  uint16_t *d_info = nullptr;     // [passes] step_info of each step
  size_t order_cap = 0, info_cap = 0;
};

// ---------------------------------------------------------------------------
// helpers

// A roctx range around a host entry point (rocprofv3 --marker-trace shows
// it; without a profiler the calls do nothing).  Sweeps, flows and
// measurements are enqueued asynchronously, so a range covers the host's
// enqueueing and any waits inside the call, not the kernels themselves.
struct TraceRange {
  explicit TraceRange(const char *what) { roctxRangePushA(what); }
  ~TraceRange() { roctxRangePop(); }
  TraceRange(const TraceRange &) = delete;
  TraceRange &operator=(const TraceRange &) = delete;
};

static size_t wsize(const su2hip_params &p) { return p.precision == SU2HIP_FP64 ? 8 : 4; }
static int vol_of(const su2hip_params &p) { return p.L[0] * p.L[1] * p.L[2]; }
// FP32 links of this lattice in the packed 12-byte form (link_format
// resolved by the handle; FP64 links always have four components)
static bool packed_links(const su2hip_params &p) {
  return p.precision == SU2HIP_FP32 && p.link_format == SU2HIP_LINKS_PACKED;
}
// words of the links of a lattice with Vh sites per parity: 3 directions x 2
// parities x the words of one stored link (3 packed, else 4)
static size_t link_words(size_t Vh, bool packed) { return 6 * Vh * (packed ? 3 : 4); }
static size_t link_words(const su2hip_params &p) {
  return link_words((size_t)vol_of(p) / 2, packed_links(p));
}

static size_t field_words(const su2hip_params &p) {
  size_t Vh = (size_t)vol_of(p) / 2;
  size_t w = link_words(p);
  if (p.u1) w += 6 * Vh;
  if (p.nhiggs) w += 8 * Vh * p.nhiggs;
  if (p.triplet) w += 6 * Vh;
  if (p.singlet) w += 2 * Vh;
  return w;
}

static size_t flow_words(const su2hip_params &p) {
  size_t Vh = (size_t)vol_of(p) / 2;
  size_t w = 2 * link_words(p) + 18 * Vh;
  if (p.triplet) w += 12 * Vh + 6 * Vh;
  return w;
}

static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

static bool supported_model(int m) {
  return m == 0 || m == M_NH || m == (M_NH | M_U1) || m == (M_NH | M_SG) ||
         m == (M_NH | M_U1 | M_SG) || m == M_TR || m == (M_NH | M_TR) || m == (M_NH | M_H2) ||
         m == (M_NH | M_H2 | M_U1);
}

static int validate(const su2hip_params *p) {
  if (!p) return fail(SU2HIP_EINVAL, "params is NULL");
  for (int d = 0; d < 3; d++)
    if (p->L[d] < 2 || (p->L[d] & 1))
      return fail(SU2HIP_EINVAL, "extent L%d = %d must be even and >= 2", d + 1, p->L[d]);
  long long V = (long long)p->L[0] * p->L[1] * p->L[2];
  if (V > (1ll << 28)) return fail(SU2HIP_ERANGE, "volume %lld too large", V);
  if (p->nhiggs < 0 || p->nhiggs > 2)
    return fail(SU2HIP_EUNSUPPORTED, "nhiggs = %d (0, 1 or 2 supported)", p->nhiggs);
  if (p->link_format < SU2HIP_LINKS_AUTO || p->link_format > SU2HIP_LINKS_PACKED)
    return fail(SU2HIP_EINVAL, "link_format %d", p->link_format);
  if (!supported_model(model_bits(*p)))
    return fail(SU2HIP_EUNSUPPORTED,
                "field content nhiggs=%d u1=%d triplet=%d singlet=%d is not supported",
                p->nhiggs, p->u1, p->triplet, p->singlet);
  if (p->precision != SU2HIP_FP64 && p->precision != SU2HIP_FP32)
    return fail(SU2HIP_EINVAL, "precision %d (64 or 32)", p->precision);
  if (p->algorithm_su2link != SU2HIP_HEATBATH && p->algorithm_su2link != SU2HIP_METROPOLIS)
    return fail(SU2HIP_EINVAL, "algorithm_su2link %d", p->algorithm_su2link);
  if (p->u1 && p->algorithm_u1link != SU2HIP_METROPOLIS)
    return fail(SU2HIP_EINVAL, "algorithm_u1link %d", p->algorithm_u1link);
  int sc[3] = {p->algorithm_su2doublet, p->algorithm_su2triplet, p->algorithm_singlet};
  int on[3] = {p->nhiggs, p->triplet, p->singlet};
  for (int k = 0; k < 3; k++)
    if (on[k] && sc[k] != SU2HIP_OVERRELAX && sc[k] != SU2HIP_METROPOLIS)
      return fail(SU2HIP_EINVAL, "scalar algorithm %d", sc[k]);
  if (p->update_links < 0 || p->link_overrelax < 0 || p->scalar_sweeps < 0 ||
      p->update_su2doublet < 0 || p->update_su2triplet < 0 || p->update_singlet < 0)
    return fail(SU2HIP_EINVAL, "negative update count");
  if (p->betasu2 == 0.0) return fail(SU2HIP_EINVAL, "betasu2 must be nonzero");
  return SU2HIP_OK;
}

// lattices of one batch share the pass schedule and the arithmetic
static bool compatible(const su2hip_params &a, const su2hip_params &b) {
  return a.nhiggs == b.nhiggs && a.u1 == b.u1 && a.triplet == b.triplet &&
         a.singlet == b.singlet && a.precision == b.precision &&
         a.algorithm_su2link == b.algorithm_su2link && a.algorithm_u1link == b.algorithm_u1link &&
         a.algorithm_su2doublet == b.algorithm_su2doublet &&
         a.algorithm_su2triplet == b.algorithm_su2triplet &&
         a.algorithm_singlet == b.algorithm_singlet && a.update_links == b.update_links &&
         a.link_overrelax == b.link_overrelax && a.scalar_sweeps == b.scalar_sweeps &&
         a.update_su2doublet == b.update_su2doublet &&
         a.update_su2triplet == b.update_su2triplet && a.update_singlet == b.update_singlet &&
         a.reunitarize == b.reunitarize;
}

static LaunchCtx ctx_of(su2hip_batch *b) {
  LaunchCtx c;
  c.stream = b->stream;
  c.desc = b->d_desc;
  c.muca = b->d_muca;
  c.blk = b->d_blk;
  c.nblk = b->nblk;
  c.cnt = b->d_cnt;
  c.partial = b->d_partial;
  c.sums = b->d_sums;
  c.max_slots = b->max_slots;
  c.model = b->model;
  c.fp64 = b->fp64;
  return c;
}

// workgroup width of a pass kind: the tuned one, else the widest that still
// gives at least 4 workgroups per CU (256 CUs), else 64
static int chunk_for(su2hip_batch *b, int kind) {
  if (kind >= 0 && kind < SU2HIP_NKINDS && b->tune[kind]) return b->tune[kind];
  if (b->nblk >= 1024) return 256;
  if (b->alt[1].nblk >= 1024) return 128;
  return 64;
}

static LaunchCtx ctx_for(su2hip_batch *b, int kind) {
  LaunchCtx c = ctx_of(b);
  if (kind >= 0 && kind < SU2HIP_NKINDS) c.waves = b->waves[kind];
  int ch = chunk_for(b, kind);
  for (int k = 0; k < 2; k++)
    if (ch == b->alt[k].chunk) {
      c.blk = b->alt[k].d_blk;
      c.nblk = b->alt[k].nblk;
      c.cnt = b->alt[k].d_cnt;
      c.block = ch;
    }
  return c;
}

static void fill_desc(Slot &s) {
  const su2hip_params &p = s.p;
  LatDesc &L = s.hd;
  size_t Vh = (size_t)vol_of(p) / 2, w = wsize(p);
  char *base = (char *)s.mem;
  size_t off = 0;
  auto take = [&](size_t words) {
    void *q = base + off;
    off += align256(words * w);
    return q;
  };
  L.U = take(link_words(p));
  L.A = p.u1 ? take(6 * Vh) : nullptr;
  L.F = p.nhiggs ? take(8 * Vh) : nullptr;
  L.F2 = p.nhiggs == 2 ? take(8 * Vh) : nullptr;
  L.G = p.triplet ? take(6 * Vh) : nullptr;
  L.S = p.singlet ? take(2 * Vh) : nullptr;
  L.H = s.hop_mem;
  L.B = s.bu_mem;
  L.L0 = p.L[0];
  L.L1 = p.L[1];
  L.L2 = p.L[2];
  L.h0 = p.L[0] / 2;
  L.Vh = (int)Vh;
  L.vol = vol_of(p);
  L.div_h0 = make_fastdiv((uint32_t)L.h0);
  L.div_L1 = make_fastdiv((uint32_t)L.L1);
  L.active = 1;
  L.key0 = (uint32_t)p.seed;
  L.key1 = (uint32_t)(p.seed >> 32);
  L.beta = p.betasu2;
  L.betau1 = p.betau1;
  L.r_u1 = p.r_u1;
  L.msq = p.msq_phi;
  L.lambda = p.lambda_phi;
  L.msq_t = p.msq_triplet;
  L.a2 = p.a2;
  L.b4 = p.b4;
  L.b1_s = p.b1_s;
  L.msq_s = p.msq_s;
  L.b3_s = p.b3_s;
  L.b4_s = p.b4_s;
  L.a1_s = p.a1_s;
  L.a2_s = p.a2_s;
  L.msq2 = p.msq_phi2;
  L.m12re = p.m12sq_re;
  L.m12im = p.m12sq_im;
  L.lam2 = p.lam2;
  L.lam3 = p.lam3;
  L.lam4 = p.lam4;
  L.lam5re = p.lam5_re;
  L.lam5im = p.lam5_im;
  L.lam6re = p.lam6_re;
  L.lam6im = p.lam6_im;
  L.lam7re = p.lam7_re;
  L.lam7im = p.lam7_im;
}

static size_t slot_bytes(const su2hip_params &p) {
  size_t Vh = (size_t)vol_of(p) / 2, w = wsize(p);
  size_t b = align256(link_words(p) * w);
  if (p.u1) b += align256(6 * Vh * w);
  if (p.nhiggs) b += align256(8 * Vh * w);
  if (p.nhiggs == 2) b += align256(8 * Vh * w);
  if (p.triplet) b += align256(6 * Vh * w);
  if (p.singlet) b += align256(2 * Vh * w);
  return b;
}

// the hopping-term cache of a lattice with one doublet: 4 values per link
static size_t hop_cache_bytes(const su2hip_params &p) {
  size_t Vh = (size_t)vol_of(p) / 2;
  return p.nhiggs == 1 ? align256(24 * Vh * wsize(p)) : 0;
}

// pull the device-owned descriptor fields (pass counter, budget)
static int pull_desc(su2hip_batch *b) {
  std::vector<LatDesc> tmp(b->max_slots);
  HIPTRY(hipMemcpyAsync(tmp.data(), b->d_desc, sizeof(LatDesc) * b->max_slots,
                        hipMemcpyDeviceToHost, b->stream));
  HIPTRY(hipStreamSynchronize(b->stream));
  for (int i = 0; i < b->max_slots; i++)
    if (b->slots[i].used) {
      b->slots[i].hd.pass = tmp[i].pass;
      b->slots[i].hd.budget = tmp[i].budget;
    }
  return SU2HIP_OK;
}

static int push_desc(su2hip_batch *b) {
  std::vector<LatDesc> tmp(b->max_slots);
  for (int i = 0; i < b->max_slots; i++) {
    if (b->slots[i].used) {
      tmp[i] = b->slots[i].hd;
    } else {
      memset(&tmp[i], 0, sizeof(LatDesc));
      tmp[i].active = 0;
    }
  }
  HIPTRY(hipMemcpyAsync(b->d_desc, tmp.data(), sizeof(LatDesc) * b->max_slots,
                        hipMemcpyHostToDevice, b->stream));
  HIPTRY(hipStreamSynchronize(b->stream));
  return SU2HIP_OK;
}

static void add_counts(su2hip_counters &c, int kind, long long acc, long long tot) {
  switch (kind) {
    case SU2HIP_K_LINK_HEATBATH:
    case SU2HIP_K_LINK_METROPOLIS:
      c.accepted_su2link += acc;
      c.total_su2link += tot;
      break;
    case SU2HIP_K_LINK_OVERRELAX:
      c.accepted_link_or += acc;
      c.total_link_or += tot;
      break;
    case SU2HIP_K_U1_METROPOLIS:
      c.accepted_u1link += acc;
      c.total_u1link += tot;
      break;
    case SU2HIP_K_DOUBLET_OVERRELAX:
      c.acc_overrelax_doublet += acc;
      c.total_overrelax_doublet += tot;
      break;
    case SU2HIP_K_DOUBLET_METROPOLIS:
      c.accepted_doublet += acc;
      c.total_doublet += tot;
      break;
    case SU2HIP_K_TRIPLET_OVERRELAX:
      c.acc_overrelax_triplet += acc;
      c.total_overrelax_triplet += tot;
      break;
    case SU2HIP_K_TRIPLET_METROPOLIS:
      c.accepted_triplet += acc;
      c.total_triplet += tot;
      break;
    case SU2HIP_K_SINGLET_OVERRELAX:
      c.acc_overrelax_singlet += acc;
      c.total_overrelax_singlet += tot;
      break;
    case SU2HIP_K_SINGLET_METROPOLIS:
      c.accepted_singlet += acc;
      c.total_singlet += tot;
      break;
    case CNT_DOUBLET2_OR:
      c.acc_overrelax_doublet2 += acc;
      c.total_overrelax_doublet2 += tot;
      break;
    case CNT_DOUBLET2_METRO:
      c.accepted_doublet2 += acc;
      c.total_doublet2 += tot;
      break;
  }
}

// fold per-block device counters of one block table into the slots' host
// counters and clear them
static int fold_table(su2hip_batch *b, long long *d_cnt, int nblk, const std::vector<BlockMap> &hb) {
  if (!d_cnt || nblk == 0) return SU2HIP_OK;
  size_t n = (size_t)NCOUNT * 2 * nblk;
  std::vector<long long> h(n);
  HIPTRY(hipMemcpyAsync(h.data(), d_cnt, n * sizeof(long long), hipMemcpyDeviceToHost, b->stream));
  HIPTRY(hipMemsetAsync(d_cnt, 0, n * sizeof(long long), b->stream));
  HIPTRY(hipStreamSynchronize(b->stream));
  for (int k = 0; k < NCOUNT; k++)
    for (int i = 0; i < nblk; i++) {
      long long a = h[(size_t)(k * 2) * nblk + i], t = h[(size_t)(k * 2 + 1) * nblk + i];
      if (a || t) add_counts(b->slots[hb[i].lat].cnt, k, a, t);
    }
  return SU2HIP_OK;
}

static int fold_counts(su2hip_batch *b) {
  int rc = fold_table(b, b->d_cnt, b->nblk, b->h_blk);
  for (int k = 0; k < 2 && !rc; k++) rc = fold_table(b, b->alt[k].d_cnt, b->alt[k].nblk, b->alt[k].h);
  return rc;
}

// logical chunks lattice by lattice, dealt to physical workgroups so that
// XCD j (workgroups j, j+8, ...) gets a contiguous run of chunks
static std::vector<BlockMap> deal_to_xcds(const std::vector<BlockMap> &logical) {
  int n = (int)logical.size();
  std::vector<BlockMap> phys(n);
  const int X = 8;
  int base = 0;
  for (int j = 0; j < X; j++) {
    int cnt = n / X + (j < n % X ? 1 : 0);
    for (int k = 0; k < cnt; k++) phys[j + X * k] = logical[base + k];
    base += cnt;
  }
  return phys;
}

static int build_alt(su2hip_batch *b, AltMap &m, int chunk) {
  m.chunk = chunk;
  std::vector<BlockMap> logical;
  for (int i = 0; i < b->max_slots; i++) {
    Slot &s = b->slots[i];
    if (!s.used) continue;
    for (int k = 0; k * chunk < s.hd.Vh; k++) logical.push_back(BlockMap{i, k * chunk});
  }
  int n = (int)logical.size();
  if (n > m.cap) {
    if (m.d_blk) HIPTRY(hipFree(m.d_blk));
    if (m.d_cnt) HIPTRY(hipFree(m.d_cnt));
    m.d_blk = nullptr;
    m.d_cnt = nullptr;
    int cap = std::max(n, 256);
    if (hipMalloc(&m.d_blk, sizeof(BlockMap) * cap) != hipSuccess ||
        hipMalloc(&m.d_cnt, sizeof(long long) * NCOUNT * 2 * cap) != hipSuccess)
      return fail(SU2HIP_ENOMEM, "block table allocation failed");
    m.cap = cap;
  }
  m.h = deal_to_xcds(logical);
  m.nblk = n;
  if (n) {
    HIPTRY(hipMemcpyAsync(m.d_blk, m.h.data(), sizeof(BlockMap) * n, hipMemcpyHostToDevice,
                          b->stream));
    HIPTRY(hipMemsetAsync(m.d_cnt, 0, sizeof(long long) * NCOUNT * 2 * n, b->stream));
  }
  return SU2HIP_OK;
}

// Block table: logical chunks lattice by lattice, then physical workgroup
// ids dealt so that XCD j (workgroups j, j+8, ...) gets a contiguous run of
// chunks and its L2 sees neighbouring rows.
static int rebuild_map(su2hip_batch *b) {
  if (!b->map_dirty) return SU2HIP_OK;
  int rc = fold_counts(b);
  if (rc) return rc;
  rc = pull_desc(b);
  if (rc) return rc;
  std::vector<BlockMap> logical;
  for (int i = 0; i < b->max_slots; i++) {
    Slot &s = b->slots[i];
    if (!s.used) continue;
    int nb = (s.hd.Vh + BLK - 1) / BLK;
    s.hd.blk_base = (int)logical.size();
    s.hd.nblk = nb;
    for (int k = 0; k < nb; k++) logical.push_back(BlockMap{i, k * BLK});
  }
  int n = (int)logical.size();
  std::vector<BlockMap> phys = deal_to_xcds(logical);
  if (n > b->nblk_cap) {
    if (b->d_blk) HIPTRY(hipFree(b->d_blk));
    if (b->d_cnt) HIPTRY(hipFree(b->d_cnt));
    b->d_blk = nullptr;
    b->d_cnt = nullptr;
    int cap = std::max(n, 64);
    if (hipMalloc(&b->d_blk, sizeof(BlockMap) * cap) != hipSuccess ||
        hipMalloc(&b->d_cnt, sizeof(long long) * NCOUNT * 2 * cap) != hipSuccess)
      return fail(SU2HIP_ENOMEM, "block table allocation failed");
    b->nblk_cap = cap;
  }
  size_t need = (size_t)std::max(n, 1) * SU2HIP_NSUMS * 2;
  if (need > b->partial_cap) {
    if (b->d_partial) HIPTRY(hipFree(b->d_partial));
    b->d_partial = nullptr;
    if (hipMalloc(&b->d_partial, need * sizeof(double)) != hipSuccess)
      return fail(SU2HIP_ENOMEM, "partial-sum buffer allocation failed");
    b->partial_cap = need;
  }
  b->nblk = n;
  b->h_blk = phys;
  if (n) {
    HIPTRY(hipMemcpyAsync(b->d_blk, phys.data(), sizeof(BlockMap) * n, hipMemcpyHostToDevice,
                          b->stream));
    HIPTRY(hipMemsetAsync(b->d_cnt, 0, sizeof(long long) * NCOUNT * 2 * n, b->stream));
  }
  for (int k = 0; k < 2 && !rc; k++) rc = build_alt(b, b->alt[k], CHUNKS[k]);
  if (rc) return rc;
  rc = push_desc(b);
  if (rc) return rc;
  b->map_dirty = false;
  b->graph_valid = false;
  b->sched_dirty = true;
  return SU2HIP_OK;
}

// field 1 the first doublet, 2 the triplet, 3 the second doublet
static bool muca_on_field(const MucaDev &m, int field) {
  if (!m.enabled || !m.do_acceptance) return false;
  if (field == 1) return m.orderparam == SU2HIP_PHISQ || m.orderparam == SU2HIP_PHI2MINUSSIGMA2;
  if (field == 2) return m.orderparam == SU2HIP_SIGMASQ || m.orderparam == SU2HIP_PHI2MINUSSIGMA2;
  if (field == 3) return m.orderparam == SU2HIP_PHI2SQ;
  return false;
}

static int muca_nsteps(int Vh, int cps) {
  int iv = Vh / cps;
  if (iv <= 0) iv = 1;
  int nfull = Vh / iv;
  return nfull + (((Vh % iv) != 0) || ((Vh % cps) != 0) ? 1 : 0);
}

static int scalar_round(int alg, bool last, int kor, int kme) {
  if (last) return alg == SU2HIP_OVERRELAX ? kme : -1;
  return alg == SU2HIP_OVERRELAX ? kor : kme;
}

// The pass sequence of update_lattice (src/update.c:270-430), with the link
// overrelaxation rounds after each heatbath/Metropolis link round.
static void build_sched(su2hip_batch *b) {
  if (!b->sched_dirty) return;
  const su2hip_params &p = b->tmpl;
  bool muca_f[4] = {false, false, false, false};
  int steps = 1;
  for (auto &s : b->slots) {
    if (!s.used) continue;
    for (int f = 1; f <= 3; f++)
      if (muca_on_field(s.hm, f)) {
        muca_f[f] = true;
        steps = std::max(steps, muca_nsteps(s.hd.Vh, std::max(1, s.hm.cps)));
      }
  }
  b->muca_steps = steps;
  std::vector<Step> &v = b->sched;
  v.clear();
  int off = 0;
  auto add = [&](int kind, int par, int dir, bool muca, uint16_t info) {
    v.push_back(Step{kind, par, dir, off++, muca, info});
  };
  // link rounds: position j = 2 dir + parity of the default order, which the
  // random order replaces by (par_a[j], dir_a[j]) of the round
  int lk = p.algorithm_su2link == SU2HIP_HEATBATH ? SU2HIP_K_LINK_HEATBATH : SU2HIP_K_LINK_METROPOLIS;
  for (int k = 0; k < p.update_links; k++) {
    for (int d = 0; d < 3; d++)
      for (int par = 0; par < 2; par++) add(lk, par, d, false, step_info(false, k, 2 * d + par));
    for (int r = 0; r < p.link_overrelax; r++)
      for (int d = 0; d < 3; d++)
        for (int par = 0; par < 2; par++)
          add(SU2HIP_K_LINK_OVERRELAX, par, d, false, step_info(false, k, 2 * d + par));
    if (p.u1)
      for (int d = 0; d < 3; d++)
        for (int par = 0; par < 2; par++)
          add(SU2HIP_K_U1_METROPOLIS, par, d, false, step_info(false, k, 2 * d + par));
  }
  int ns = p.scalar_sweeps + 1;
  for (int s = 0; s < ns; s++) {
    bool last = s >= ns - 1;
    if (p.nhiggs) {
      int kd = scalar_round(p.algorithm_su2doublet, last, SU2HIP_K_DOUBLET_OVERRELAX,
                            SU2HIP_K_DOUBLET_METROPOLIS);
      // first EVEN and ODD of doublet 1, then doublet 2 (src/update.c:361-373)
      if (kd >= 0)
        for (int k = 0; k < p.update_su2doublet; k++)
          for (int h = 0; h < p.nhiggs; h++)
            for (int par = 0; par < 2; par++)
              add(kd, par, h, muca_f[h ? 3 : 1], step_info(true, s, par));
    }
    if (p.triplet) {
      int kt = scalar_round(p.algorithm_su2triplet, last, SU2HIP_K_TRIPLET_OVERRELAX,
                            SU2HIP_K_TRIPLET_METROPOLIS);
      if (kt >= 0)
        for (int k = 0; k < p.update_su2triplet; k++)
          for (int par = 0; par < 2; par++) add(kt, par, 0, muca_f[2], step_info(true, s, par));
    }
    if (p.singlet) {
      int ks = scalar_round(p.algorithm_singlet, last, SU2HIP_K_SINGLET_OVERRELAX,
                            SU2HIP_K_SINGLET_METROPOLIS);
      if (ks >= 0)
        for (int k = 0; k < p.update_singlet; k++)
          for (int par = 0; par < 2; par++) add(ks, par, 0, false, step_info(true, s, par));
    }
  }
  b->passes = off;
  b->sched_dirty = false;
  b->graph_valid = false;
}

static hipEvent_t ev_get(su2hip_batch *b) {
  if (!b->pool.empty()) {
    hipEvent_t e = b->pool.back();
    b->pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

// run fn() bracketed by timing events when timing is on
template <typename Fn>
static hipError_t timed(su2hip_batch *b, int kind, Fn &&fn) {
  if (!b->timing) return fn();
  hipEvent_t a = ev_get(b), e = ev_get(b);
  if (!a || !e) return hipErrorOutOfMemory;
  hipError_t r = hipEventRecord(a, b->stream);
  if (r != hipSuccess) return r;
  r = fn();
  if (r != hipSuccess) return r;
  r = hipEventRecord(e, b->stream);
  b->pending.push_back(Ev{kind, a, e});
  return r;
}

// in_sweep: the step is part of a compound sweep, whose random pass order
// (when on) replaces the step's parity and link direction
static hipError_t enqueue_step(su2hip_batch *b, const Step &st, bool in_sweep = true) {
  LaunchCtx c = ctx_of(b);
  if (in_sweep && b->random_order) {
    c.order = b->d_order;
    c.npass = b->passes;
  }
  int reunit = b->tmpl.reunitarize && !b->fp64;
  if (!st.muca) {
    LaunchCtx ct = ctx_for(b, st.kind);
    ct.order = c.order;
    ct.npass = c.npass;
    return timed(b, st.kind, [&] {
      return launch_pass(ct, st.kind, st.parity, st.dir, st.pass_off, -1, reunit);
    });
  }
  c.waves = b->waves[st.kind];  // multicanonical steps keep the 256-site table
  for (int k = 0; k < b->muca_steps; k++) {
    hipError_t e = timed(b, st.kind, [&] {
      return launch_pass(c, st.kind, st.parity, st.dir, st.pass_off, k, reunit);
    });
    if (e != hipSuccess) return e;
    e = timed(b, SU2HIP_K_MUCA,
              [&] { return launch_muca_accept(c, st.kind, st.dir, st.parity, st.pass_off, k); });
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

// the SU(2) link kinds, whose staples read the hopping-term cache
static bool su2_link_kind(int kind) {
  return kind == SU2HIP_K_LINK_HEATBATH || kind == SU2HIP_K_LINK_METROPOLIS ||
         kind == SU2HIP_K_LINK_OVERRELAX;
}

static hipError_t enqueue_sweep(su2hip_batch *b) {
  if (b->random_order) {
    const su2hip_params &p = b->tmpl;
    LaunchCtx c = ctx_of(b);
    hipError_t e = timed(b, SU2HIP_K_AUX, [&] {
      return launch_order(c, b->d_order, b->d_info, b->passes, p.update_links, p.scalar_sweeps + 1);
    });
    if (e != hipSuccess) return e;
  }
  // the hopping-term cache is filled before each round of link passes: the
  // doublet and the U(1) links change only in the passes between rounds
  bool hop_fresh = false;
  LaunchCtx c = ctx_of(b);
  for (const Step &st : b->sched) {
    hipError_t e = hipSuccess;
    if (!su2_link_kind(st.kind))
      hop_fresh = false;
    else if (!hop_fresh && b->hop_on) {
      e = timed(b, SU2HIP_K_AUX, [&] { return launch_hop(c); });
      hop_fresh = true;
    }
    if (e == hipSuccess) e = enqueue_step(b, st);
    if (e != hipSuccess) return e;
  }
  return timed(b, SU2HIP_K_AUX, [&] { return launch_end_sweep(c, b->passes); });
}

// device tables of the random pass order for the current schedule
static int ensure_order(su2hip_batch *b) {
  if (!b->random_order || b->passes == 0) return SU2HIP_OK;
  size_t no = (size_t)b->max_slots * b->passes, ni = (size_t)b->passes;
  if (no > b->order_cap) {
    if (b->d_order) HIPTRY(hipFree(b->d_order));
    b->d_order = nullptr;
    b->order_cap = 0;
    if (hipMalloc(&b->d_order, no) != hipSuccess) return fail(SU2HIP_ENOMEM, "order table");
    b->order_cap = no;
  }
  if (ni > b->info_cap) {
    if (b->d_info) HIPTRY(hipFree(b->d_info));
    b->d_info = nullptr;
    b->info_cap = 0;
    if (hipMalloc(&b->d_info, ni * sizeof(uint16_t)) != hipSuccess)
      return fail(SU2HIP_ENOMEM, "order table");
    b->info_cap = ni;
  }
  std::vector<uint16_t> h(ni);
  for (size_t j = 0; j < ni; j++) h[j] = b->sched[j].info;
  HIPTRY(hipMemcpyAsync(b->d_info, h.data(), ni * sizeof(uint16_t), hipMemcpyHostToDevice, b->stream));
  HIPTRY(hipStreamSynchronize(b->stream));
  return SU2HIP_OK;
}

static int prepare(su2hip_batch *b) {
  int rc = rebuild_map(b);
  if (rc) return rc;
  bool rebuilt = b->sched_dirty;
  build_sched(b);
  if (rebuilt) return ensure_order(b);
  return SU2HIP_OK;
}

static int ensure_stage(su2hip_batch *b, size_t doubles) {
  if (doubles <= b->stage_cap) return SU2HIP_OK;
  if (b->d_stage) HIPTRY(hipFree(b->d_stage));
  b->d_stage = nullptr;
  if (hipMalloc(&b->d_stage, doubles * sizeof(double)) != hipSuccess)
    return fail(SU2HIP_ENOMEM, "staging buffer of %zu doubles", doubles);
  b->stage_cap = doubles;
  return SU2HIP_OK;
}

static int check_slot(su2hip_batch *b, int slot) {
  if (!b) return fail(SU2HIP_EINVAL, "NULL handle");
  if (slot < 0 || slot >= b->max_slots) return fail(SU2HIP_ERANGE, "slot %d out of range", slot);
  if (!b->slots[slot].used) return fail(SU2HIP_ERANGE, "slot %d is empty", slot);
  return SU2HIP_OK;
}

// The link storage of a handle (SU2HIP_LINKS_FULL or _PACKED) from its
// first lattice's link_format; AUTO measures the FP32 fields of the n
// lattices in the full form against SU2HIP_LINKS_PACKED_ABOVE (n = 0: the
// lattices come later, as in serving: packed).
static int resolve_links(int n, const su2hip_params *params) {
  const su2hip_params &p0 = params[0];
  if (p0.precision == SU2HIP_FP64) return SU2HIP_LINKS_FULL;
  if (p0.link_format != SU2HIP_LINKS_AUTO) return p0.link_format;
  if (n == 0) return SU2HIP_LINKS_PACKED;
  uint64_t bytes = 0;
  for (int i = 0; i < n; i++) {
    su2hip_params q = params[i];
    q.link_format = SU2HIP_LINKS_FULL;
    bytes += field_words(q) * wsize(q);
  }
  return bytes > SU2HIP_LINKS_PACKED_ABOVE ? SU2HIP_LINKS_PACKED : SU2HIP_LINKS_FULL;
}

// the slot's hopping-term cache, inside the handle's HBM budget
static int hop_alloc(su2hip_batch *b, int i) {
  Slot &s = b->slots[i];
  size_t bytes = hop_cache_bytes(s.p);
  if (s.hop_mem || bytes == 0) return SU2HIP_OK;
  if (b->hbm_budget && b->used + bytes > b->hbm_budget)
    return fail(SU2HIP_ENOMEM, "HBM budget: %llu + %zu bytes of hopping-term cache exceeds %llu",
                (unsigned long long)b->used, bytes, (unsigned long long)b->hbm_budget);
  if (hipMalloc(&s.hop_mem, bytes) != hipSuccess) {
    s.hop_mem = nullptr;
    return fail(SU2HIP_ENOMEM, "device allocation of %zu bytes failed", bytes);
  }
  s.hop_bytes = bytes;
  b->used += bytes;
  s.hd.H = s.hop_mem;
  return SU2HIP_OK;
}

static void hop_release(su2hip_batch *b, int i) {
  Slot &s = b->slots[i];
  if (s.hop_mem) (void)hipFree(s.hop_mem);
  b->used -= s.hop_bytes;
  s.hop_mem = nullptr;
  s.hop_bytes = 0;
  s.hd.H = nullptr;
}

// AUTO choice of the hopping-term cache for a handle created with n
// lattices: kept with one doublet, except where the fields fit the
// Infinity Cache and the fields and the cache together do not (the link
// passes save less than the other passes then lose to the extra footprint);
// a handle created empty keeps it
static bool resolve_hop(int n, const su2hip_params *params, int link_format) {
  if (params[0].nhiggs != 1) return false;
  if (n == 0) return true;
  uint64_t f = 0, h = 0;
  for (int i = 0; i < n; i++) {
    su2hip_params q = params[i];
    q.link_format = link_format;
    f += slot_bytes(q);
    h += hop_cache_bytes(q);
  }
  return !(f <= SU2HIP_INFINITY_CACHE_BYTES && f + h > SU2HIP_INFINITY_CACHE_BYTES);
}

static int alloc_slot(su2hip_batch *b, int i, const su2hip_params &p_in) {
  su2hip_params p = p_in;
  p.link_format = b->tmpl.link_format;  // the handle's resolved storage
  size_t bytes = slot_bytes(p);
  if (b->hbm_budget && b->used + bytes > b->hbm_budget)
    return fail(SU2HIP_ENOMEM, "HBM budget: %llu + %zu bytes exceeds %llu",
                (unsigned long long)b->used, bytes, (unsigned long long)b->hbm_budget);
  Slot &s = b->slots[i];
  s = Slot();
  s.p = p;
  if (hipMalloc(&s.mem, bytes) != hipSuccess) {
    s.mem = nullptr;
    return fail(SU2HIP_ENOMEM, "device allocation of %zu bytes failed", bytes);
  }
  s.bytes = bytes;
  s.used = true;
  b->used += bytes;
  fill_desc(s);
  s.hd.budget = -1;
  s.hd.pass = 0;
  s.hm = MucaDev();
  b->map_dirty = true;
  // a lattice whose cache does not fit the budget runs without one (its
  // link passes form the hopping term themselves)
  if (b->hop_on) (void)hop_alloc(b, i);
  return SU2HIP_OK;
}

static void flow_release(FlowJob &j) {
  if (j.ddesc) (void)hipFree(j.ddesc);
  if (j.dblk) (void)hipFree(j.dblk);
  if (j.d_partial) (void)hipFree(j.d_partial);
  if (j.d_sums) (void)hipFree(j.d_sums);
  if (j.d_rows) (void)hipFree(j.d_rows);
  if (j.done) (void)hipEventDestroy(j.done);
  j = FlowJob();
}

static void free_slot(su2hip_batch *b, int i) {
  Slot &s = b->slots[i];
  if (s.fj.active) {
    (void)hipEventSynchronize(s.fj.done);
    flow_release(s.fj);
  }
  if (s.mem) (void)hipFree(s.mem);
  if (s.bu_mem) (void)hipFree(s.bu_mem);
  if (s.muca_mem) (void)hipFree(s.muca_mem);
  if (s.flow_mem) (void)hipFree(s.flow_mem);
  if (s.hop_mem) (void)hipFree(s.hop_mem);
  b->used -= s.bytes + s.bu_bytes + s.flow_bytes + s.hop_bytes;
  s = Slot();
  b->map_dirty = true;
}

static int push_muca(su2hip_batch *b, int slot) {
  HIPTRY(hipMemcpyAsync(b->d_muca + slot, &b->slots[slot].hm, sizeof(MucaDev),
                        hipMemcpyHostToDevice, b->stream));
  HIPTRY(hipStreamSynchronize(b->stream));
  return SU2HIP_OK;
}

static int pull_muca(su2hip_batch *b, int slot) {
  HIPTRY(hipMemcpyAsync(&b->slots[slot].hm, b->d_muca + slot, sizeof(MucaDev),
                        hipMemcpyDeviceToHost, b->stream));
  HIPTRY(hipStreamSynchronize(b->stream));
  return SU2HIP_OK;
}

// both parity contributions of the order parameter from the fields
// (calc_orderparam) of slot `only`, or of every slot with only < 0
static int orderparam_refresh(su2hip_batch *b, int only = -1) {
  bool any = false;
  for (int i = 0; i < b->max_slots; i++)
    if (b->slots[i].used && b->slots[i].hm.enabled && (only < 0 || only == i)) any = true;
  if (!any) return SU2HIP_OK;
  int rc = prepare(b);
  if (rc) return rc;
  HIPTRY(launch_orderparam_init(ctx_of(b), only));
  HIPTRY(hipStreamSynchronize(b->stream));
  return SU2HIP_OK;
}

// ---------------------------------------------------------------------------
// C ABI

extern "C" {

const char *su2hip_error_string(int code) {
  switch (code) {
    case SU2HIP_OK: return "ok";
    case SU2HIP_EINVAL: return "invalid argument";
    case SU2HIP_ENOMEM: return "out of memory or HBM budget";
    case SU2HIP_EDEVICE: return "HIP runtime error";
    case SU2HIP_ERANGE: return "index or size out of range";
    case SU2HIP_EUNSUPPORTED: return "unsupported model";
    case SU2HIP_ESTATE: return "invalid state";
    case SU2HIP_EIO: return "I/O or parse error";
  }
  return "unknown error";
}

const char *su2hip_last_error(void) { return g_err.c_str(); }

// used by the host I/O layer (su2hip_io.c) to report its errors the same way
int su2hip_internal_fail(int code, const char *msg) {
  g_err = msg ? msg : "";
  return code;
}

// ---------------------------------------------------------------------------
// Scratch of su2hip_measure_noalloc: the block partial sums of one lattice,
// in a file-scope device array that every GPU holds from when the library
// is loaded until the process ends, so the entry point allocates nothing.
// In host code `noalloc_scratch` is only a stand-in for the symbol: its
// device address, which differs from GPU to GPU, comes from
// hipGetSymbolAddress once per device (noalloc_scratch_init, called by
// su2hip_create with the handle's device current) and only that address is
// passed to kernels.
static __device__ __align__(256) double noalloc_scratch[SU2HIP_NOALLOC_MAX_BLOCKS * SU2HIP_NSUMS];

static constexpr int kMaxDevices = 64;
static double *g_noalloc_ptr[kMaxDevices];
static std::mutex g_noalloc_mu;

static int noalloc_scratch_init() {
  int dev = 0;
  HIPTRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(SU2HIP_EDEVICE, "device %d beyond the table", dev);
  std::lock_guard<std::mutex> lock(g_noalloc_mu);
  if (g_noalloc_ptr[dev]) return SU2HIP_OK;
  void *ptr = nullptr;
  hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(noalloc_scratch));
  if (e != hipSuccess || !ptr)
    return fail(SU2HIP_EDEVICE, "scratch address on device %d: %s", dev, hipGetErrorString(e));
  g_noalloc_ptr[dev] = (double *)ptr;
  return SU2HIP_OK;
}

int su2hip_footprint(const su2hip_params *p, int with_flow, uint64_t *bytes) {
  int rc = validate(p);
  if (rc) return rc;
  if (!bytes) return fail(SU2HIP_EINVAL, "bytes is NULL");
  su2hip_params q = *p;
  q.link_format = resolve_links(1, &q);  // AUTO: as a handle of this one lattice would
  p = &q;
  size_t b = slot_bytes(*p);
  if (resolve_hop(1, p, p->link_format)) b += hop_cache_bytes(*p);
  size_t Vh = (size_t)vol_of(*p) / 2;
  if (p->nhiggs || p->triplet) b += align256((p->nhiggs ? 8 : 6) * Vh * wsize(*p));
  if (with_flow) b += align256(flow_words(*p) * wsize(*p)) + 4096;  // as su2hip_flow allocates
  *bytes = b;
  return SU2HIP_OK;
}

int su2hip_create(int n, const su2hip_params *params, int max_slots, uint64_t hbm_budget,
                  su2hip_batch_t **out) {
  if (!out) return fail(SU2HIP_EINVAL, "out is NULL");
  *out = nullptr;
  // n = 0 makes an empty handle whose lattices come later through su2hip_add;
  // params[0] still fixes the field content, precision and schedule
  if (n < 0 || !params) return fail(SU2HIP_EINVAL, "need the parameters of lattice 0");
  if (n == 0) {
    int rc = validate(params);
    if (rc) return rc;
    if (max_slots < 1) return fail(SU2HIP_EINVAL, "an empty handle needs max_slots >= 1");
  }
  if (max_slots < n) max_slots = n;
  for (int i = 0; i < n; i++) {
    int rc = validate(&params[i]);
    if (rc) return rc;
    if (!compatible(params[0], params[i]))
      return fail(SU2HIP_EINVAL, "lattice %d: field content, precision or schedule differs from lattice 0", i);
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    return fail(SU2HIP_EDEVICE, "no HIP device");
  int rs = noalloc_scratch_init();
  if (rs) return rs;
  su2hip_batch *b = new su2hip_batch();
  b->waves[SU2HIP_K_LINK_OVERRELAX] = 5;
  b->max_slots = max_slots;
  b->hbm_budget = hbm_budget;
  b->tmpl = params[0];
  b->tmpl.link_format = resolve_links(n, params);
  b->hop_on = resolve_hop(n, params, b->tmpl.link_format);
  b->model = model_bits(params[0]) | (packed_links(b->tmpl) ? M_PK : 0);
  b->fp64 = params[0].precision == SU2HIP_FP64;
  b->slots.resize(max_slots);
  int rc = SU2HIP_OK;
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
      hipMalloc(&b->d_desc, sizeof(LatDesc) * max_slots) != hipSuccess ||
      hipMalloc(&b->d_muca, sizeof(MucaDev) * max_slots) != hipSuccess ||
      hipMalloc(&b->d_sums, sizeof(double) * SU2HIP_NSUMS * max_slots) != hipSuccess) {
    su2hip_destroy(b);
    return fail(SU2HIP_ENOMEM, "handle allocation failed");
  }
  std::vector<MucaDev> zm(max_slots);
  (void)hipMemcpy(b->d_muca, zm.data(), sizeof(MucaDev) * max_slots, hipMemcpyHostToDevice);
  for (int i = 0; i < n && !rc; i++) rc = alloc_slot(b, i, params[i]);
  if (!rc) rc = push_desc(b);
  if (!rc) rc = prepare(b);
  for (int i = 0; i < n && !rc; i++) rc = su2hip_cold_start(b, i, 0.0, 0.0, 0.0);
  if (rc) {
    std::string keep = g_err;
    su2hip_destroy(b);
    g_err = keep;
    return rc;
  }
  *out = b;
  return SU2HIP_OK;
}

int su2hip_destroy(su2hip_batch_t *b) {
  if (!b) return SU2HIP_OK;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  for (int i = 0; i < b->max_slots; i++)
    if (b->slots[i].used) free_slot(b, i);
  if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
  for (auto &e : b->pending) {
    (void)hipEventDestroy(e.a);
    (void)hipEventDestroy(e.b);
  }
  for (auto e : b->pool) (void)hipEventDestroy(e);
  if (b->d_desc) (void)hipFree(b->d_desc);
  if (b->d_muca) (void)hipFree(b->d_muca);
  if (b->d_blk) (void)hipFree(b->d_blk);
  if (b->d_cnt) (void)hipFree(b->d_cnt);
  if (b->d_partial) (void)hipFree(b->d_partial);
  if (b->d_sums) (void)hipFree(b->d_sums);
  if (b->d_stage) (void)hipFree(b->d_stage);
  if (b->d_order) (void)hipFree(b->d_order);
  if (b->d_info) (void)hipFree(b->d_info);
  for (auto &m : b->alt) {
    if (m.d_blk) (void)hipFree(m.d_blk);
    if (m.d_cnt) (void)hipFree(m.d_cnt);
  }
  if (b->stream) (void)hipStreamDestroy(b->stream);
  if (b->flow_stream) (void)hipStreamDestroy(b->flow_stream);
  delete b;
  return SU2HIP_OK;
}

int su2hip_add(su2hip_batch_t *b, const su2hip_params *p, int *slot) {
  if (!b || !slot) return fail(SU2HIP_EINVAL, "NULL argument");
  int rc = validate(p);
  if (rc) return rc;
  if (!compatible(b->tmpl, *p))
    return fail(SU2HIP_EINVAL, "field content, precision or schedule differs from the batch");
  int i = 0;
  while (i < b->max_slots && b->slots[i].used) i++;
  if (i == b->max_slots) return fail(SU2HIP_ENOMEM, "all %d slots in use", b->max_slots);
  rc = rebuild_map(b);  // fold counters of the current table first
  if (rc) return rc;
  rc = alloc_slot(b, i, *p);
  if (rc) return rc;
  rc = prepare(b);
  if (rc) return rc;
  rc = su2hip_cold_start(b, i, 0.0, 0.0, 0.0);
  if (rc) return rc;
  *slot = i;
  return SU2HIP_OK;
}

int su2hip_remove(su2hip_batch_t *b, int slot) {
  int rc = check_slot(b, slot);
  if (rc) return rc;
  HIPTRY(hipStreamSynchronize(b->stream));
  rc = fold_counts(b);
  if (rc) return rc;
  rc = pull_desc(b);
  if (rc) return rc;
  free_slot(b, slot);
  MucaDev z{};
  b->slots[slot].hm = z;
  rc = push_muca(b, slot);
  if (rc) return rc;
  return prepare(b);
}

int su2hip_slots(su2hip_batch_t *b, int *max_slots, int *active) {
  if (!b) return fail(SU2HIP_EINVAL, "NULL handle");
  int n = 0;
  for (auto &s : b->slots) n += s.used;
  if (max_slots) *max_slots = b->max_slots;
  if (active) *active = n;
  return SU2HIP_OK;
}

int su2hip_slot_active(su2hip_batch_t *b, int slot, int *active) {
  if (!b || !active) return fail(SU2HIP_EINVAL, "NULL argument");
  if (slot < 0 || slot >= b->max_slots) return fail(SU2HIP_ERANGE, "slot %d out of range", slot);
  *active = b->slots[slot].used;
  return SU2HIP_OK;
}

int su2hip_set_fields(su2hip_batch_t *b, int slot, const double *su2link, const double *u1link,
                      const double *doublet, const double *triplet, const double *singlet) {
  TraceRange tr_("su2hip_set_fields");
  int rc = check_slot(b, slot);
  if (rc) return rc;
  if (b->slots[slot].fj.active)
    return fail(SU2HIP_ESTATE, "slot %d has a flow in flight", slot);
  rc = prepare(b);
  if (rc) return rc;
  const su2hip_params &p = b->slots[slot].p;
  size_t V = vol_of(p);
  rc = ensure_stage(b, V * 12);
  if (rc) return rc;
  // the second doublet follows the first in `doublet` (field 5 on the device)
  const double *src[6] = {su2link, u1link, doublet, triplet, singlet,
                          doublet && p.nhiggs == 2 ? doublet + V * 4 : nullptr};
  int have[6] = {1, p.u1, p.nhiggs, p.triplet, p.singlet, p.nhiggs == 2};
  int dofs[6] = {12, 3, 4, 3, 1, 4};
  LaunchCtx c = ctx_of(b);
  for (int f = 0; f < 6; f++) {
    if (!src[f]) continue;
    if (!have[f]) return fail(SU2HIP_EINVAL, "field %d is not part of this model", f);
    HIPTRY(hipMemcpyAsync(b->d_stage, src[f], V * dofs[f] * sizeof(double), hipMemcpyHostToDevice,
                          b->stream));
    HIPTRY(launch_to_device_layout(c, slot, f, b->d_stage));
    HIPTRY(hipStreamSynchronize(b->stream));
  }
  return orderparam_refresh(b, slot);
}

int su2hip_get_fields(su2hip_batch_t *b, int slot, double *su2link, double *u1link,
                      double *doublet, double *triplet, double *singlet) {
  TraceRange tr_("su2hip_get_fields");
  int rc = check_slot(b, slot);
  if (rc) return rc;
  rc = prepare(b);
  if (rc) return rc;
  const su2hip_params &p = b->slots[slot].p;
  size_t V = vol_of(p);
  rc = ensure_stage(b, V * 12);
  if (rc) return rc;
  double *dst[6] = {su2link, u1link, doublet, triplet, singlet,
                    doublet && p.nhiggs == 2 ? doublet + V * 4 : nullptr};
  int have[6] = {1, p.u1, p.nhiggs, p.triplet, p.singlet, p.nhiggs == 2};
  int dofs[6] = {12, 3, 4, 3, 1, 4};
  LaunchCtx c = ctx_of(b);
  for (int f = 0; f < 6; f++) {
    if (!dst[f]) continue;
    if (!have[f]) return fail(SU2HIP_EINVAL, "field %d is not part of this model", f);
    HIPTRY(launch_to_host_layout(c, slot, f, b->d_stage));
    HIPTRY(hipMemcpyAsync(dst[f], b->d_stage, V * dofs[f] * sizeof(double), hipMemcpyDeviceToHost,
                          b->stream));
    HIPTRY(hipStreamSynchronize(b->stream));
  }
  return SU2HIP_OK;
}

int su2hip_cold_start(su2hip_batch_t *b, int slot, double phi0, double sigma0, double singlet0) {
  int rc = check_slot(b, slot);
  if (rc) return rc;
  if (b->slots[slot].fj.active)
    return fail(SU2HIP_ESTATE, "slot %d has a flow in flight", slot);
  rc = prepare(b);
  if (rc) return rc;
  HIPTRY(launch_cold(ctx_of(b), slot, phi0, sigma0, singlet0));
  HIPTRY(hipStreamSynchronize(b->stream));
  return orderparam_refresh(b, slot);
}

int su2hip_hot_start(su2hip_batch_t *b, int slot) {
  int rc = check_slot(b, slot);
  if (rc) return rc;
  if (b->slots[slot].fj.active)
    return fail(SU2HIP_ESTATE, "slot %d has a flow in flight", slot);
  rc = prepare(b);
  if (rc) return rc;
  HIPTRY(launch_hot(ctx_of(b), slot));
  HIPTRY(hipStreamSynchronize(b->stream));
  return orderparam_refresh(b, slot);
}

// the reference's z direction: the longest, of equal ones the last
// (init_z_coord, src/z_coord.c:18-24; make_misc_tables, src/layout.c:771-777)
static int longest_dir(const su2hip_params &p) {
  int l = 2;
  for (int d = 2; d >= 0; d--)
    if (p.L[d] > p.L[l]) l = d;
  return l;
}

int su2hip_wall_start(su2hip_batch_t *b, int slot, int dir) {
  int rc = check_slot(b, slot);
  if (rc) return rc;
  if (b->slots[slot].fj.active)
    return fail(SU2HIP_ESTATE, "slot %d has a flow in flight", slot);
  if (dir > 2) return fail(SU2HIP_EINVAL, "direction %d", dir);
  if (dir < 0) dir = longest_dir(b->slots[slot].p);
  rc = prepare(b);
  if (rc) return rc;
  HIPTRY(launch_wall(ctx_of(b), slot, dir));
  HIPTRY(hipStreamSynchronize(b->stream));
  return orderparam_refresh(b, slot);
}

int su2hip_zprofile(su2hip_batch_t *b, int slot, int dir, double *out, int max_rows, int *nrows) {
  int rc = check_slot(b, slot);
  if (rc) return rc;
  const su2hip_params &p = b->slots[slot].p;
  if (dir > 2) return fail(SU2HIP_EINVAL, "direction %d", dir);
  if (dir < 0) dir = longest_dir(p);
  rc = prepare(b);
  if (rc) return rc;
  int Ld = p.L[dir];
  size_t V = (size_t)vol_of(p), P = V / Ld;
  rc = ensure_stage(b, 4 * V + 4 * (size_t)Ld);
  if (rc) return rc;
  double *buf = b->d_stage, *planes = b->d_stage + 4 * V;
  HIPTRY(launch_zprof(ctx_of(b), slot, dir, (int)(V / 2), Ld, P, buf, planes));
  std::vector<double> pl(4 * (size_t)Ld);
  HIPTRY(hipMemcpyAsync(pl.data(), planes, sizeof(double) * pl.size(), hipMemcpyDeviceToHost, b->stream));
  HIPTRY(hipStreamSynchronize(b->stream));
  // plane averages (src/z_coord.c:282-292)
  for (int z = 0; z < Ld && z < max_rows && out; z++)
    for (int k = 0; k < 4; k++) out[(size_t)z * 4 + k] = pl[(size_t)k * Ld + z] / (double)P;
  if (nrows) *nrows = Ld;
  return SU2HIP_OK;
}

int su2hip_update(su2hip_batch_t *b, int n_sweeps) {
  TraceRange tr_("su2hip_update");
  if (!b) return fail(SU2HIP_EINVAL, "NULL handle");
  if (n_sweeps < 0) return fail(SU2HIP_EINVAL, "n_sweeps < 0");
  int rc = prepare(b);
  if (rc) return rc;
  if (b->nblk == 0 || n_sweeps == 0) return SU2HIP_OK;
  if (b->use_graph && !b->timing) {
    if (!b->graph_valid) {
      if (b->gexec) HIPTRY(hipGraphExecDestroy(b->gexec));
      b->gexec = nullptr;
      hipGraph_t g = nullptr;
      HIPTRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
      hipError_t e = enqueue_sweep(b);
      hipError_t e2 = hipStreamEndCapture(b->stream, &g);
      if (e != hipSuccess || e2 != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return fail(SU2HIP_EDEVICE, "graph capture failed: %s",
                    hipGetErrorString(e != hipSuccess ? e : e2));
      }
      e = hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (e != hipSuccess) return fail(SU2HIP_EDEVICE, "graph instantiate: %s", hipGetErrorString(e));
      b->graph_valid = true;
    }
    for (int i = 0; i < n_sweeps; i++) HIPTRY(hipGraphLaunch(b->gexec, b->stream));
    return SU2HIP_OK;
  }
  for (int i = 0; i < n_sweeps; i++) {
    hipError_t e = enqueue_sweep(b);
    if (e != hipSuccess) return fail(SU2HIP_EDEVICE, "sweep launch: %s", hipGetErrorString(e));
  }
  return SU2HIP_OK;
}

int su2hip_set_budget(su2hip_batch_t *b, int slot, int64_t sweeps) {
  int rc = check_slot(b, slot);
  if (rc) return rc;
  rc = prepare(b);
  if (rc) return rc;
  rc = pull_desc(b);
  if (rc) return rc;
  if (b->slots[slot].fj.active && sweeps != 0)
    return fail(SU2HIP_ESTATE, "slot %d has a flow in flight; finish it before sweeping", slot);
  b->slots[slot].hd.budget = sweeps < 0 ? -1 : sweeps;
  return push_desc(b);
}

int su2hip_get_budgets(su2hip_batch_t *b, int64_t *budgets) {
  if (!b || !budgets) return fail(SU2HIP_EINVAL, "NULL argument");
  int rc = prepare(b);
  if (rc) return rc;
  rc = pull_desc(b);
  if (rc) return rc;
  for (int i = 0; i < b->max_slots; i++) budgets[i] = b->slots[i].used ? b->slots[i].hd.budget : -2;
  return SU2HIP_OK;
}

int su2hip_pass(su2hip_batch_t *b, int kind, int parity, int dir) {
  if (!b) return fail(SU2HIP_EINVAL, "NULL handle");
  if (kind < 0 || kind > SU2HIP_K_SINGLET_METROPOLIS) return fail(SU2HIP_EINVAL, "kind %d", kind);
  if (parity < 0 || parity > 1 || dir < 0 || dir > 2) return fail(SU2HIP_EINVAL, "parity/dir");
  bool dbl = kind == SU2HIP_K_DOUBLET_OVERRELAX || kind == SU2HIP_K_DOUBLET_METROPOLIS;
  if (dbl && b->tmpl.nhiggs < 2) dir = 0;  // one doublet: dir is ignored as for other scalars
  if (dbl && dir > 1) dir = 1;
  int rc = prepare(b);
  if (rc) return rc;
  int field = dbl ? (dir ? 3 : 1)
              : (kind == SU2HIP_K_TRIPLET_OVERRELAX || kind == SU2HIP_K_TRIPLET_METROPOLIS) ? 2
                                                                                             : 0;
  bool muca = false;
  for (auto &s : b->slots)
    if (s.used && field && muca_on_field(s.hm, field)) muca = true;
  hipError_t e = b->hop_on && su2_link_kind(kind) ? launch_hop(ctx_of(b)) : hipSuccess;
  if (e == hipSuccess) e = enqueue_step(b, Step{kind, parity, dir, 0, muca, 0}, false);
  if (e == hipErrorInvalidValue)
    return fail(SU2HIP_EINVAL, "kind %d is not part of this model", kind);
  if (e != hipSuccess) return fail(SU2HIP_EDEVICE, "pass launch: %s", hipGetErrorString(e));
  // one pass number consumed, budget untouched
  rc = pull_desc(b);
  if (rc) return rc;
  for (auto &s : b->slots)
    if (s.used && s.hd.budget != 0) s.hd.pass += 1;
  return push_desc(b);
}

int su2hip_use_graph(su2hip_batch_t *b, int enable) {
  if (!b) return fail(SU2HIP_EINVAL, "NULL handle");
  b->use_graph = enable != 0;
  return SU2HIP_OK;
}

int su2hip_set_random_order(su2hip_batch_t *b, int enable) {
  if (!b) return fail(SU2HIP_EINVAL, "NULL handle");
  if (b->random_order != (enable != 0)) {
    b->random_order = enable != 0;
    b->sched_dirty = true;
    b->graph_valid = false;
  }
  return prepare(b);
}

int su2hip_get_random_order(su2hip_batch_t *b, int *enabled) {
  if (!b || !enabled) return fail(SU2HIP_EINVAL, "NULL argument");
  *enabled = b->random_order ? 1 : 0;
  return SU2HIP_OK;
}

int su2hip_slot_params(su2hip_batch_t *b, int slot, su2hip_params *p) {
  int rc = check_slot(b, slot);
  if (rc) return rc;
  if (!p) return fail(SU2HIP_EINVAL, "NULL params");
  *p = b->slots[slot].p;
  return SU2HIP_OK;
}

int su2hip_muca_acceptance(su2hip_batch_t *b, int slot, int enable) {
  int rc = check_slot(b, slot);
  if (rc) return rc;
  rc = pull_muca(b, slot);
  if (rc) return rc;
  Slot &s = b->slots[slot];
  if (!s.hm.enabled) return fail(SU2HIP_ESTATE, "slot %d has no multicanonical weight", slot);
  s.hm.do_acceptance = enable != 0;
  rc = push_muca(b, slot);
  if (rc) return rc;
  b->sched_dirty = true;
  b->graph_valid = false;
  return prepare(b);
}

int su2hip_muca_orderparam(su2hip_batch_t *b, int slot, int recompute, double *pv, double *value) {
  int rc = check_slot(b, slot);
  if (rc) return rc;
  if (!b->slots[slot].hm.enabled)
    return fail(SU2HIP_ESTATE, "slot %d has no multicanonical weight", slot);
  if (recompute) {
    rc = orderparam_refresh(b, slot);
    if (rc) return rc;
  }
  rc = pull_muca(b, slot);
  if (rc) return rc;
  const MucaDev &h = b->slots[slot].hm;
  if (pv) {
    pv[0] = h.pv[0];
    pv[1] = h.pv[1];
  }
  if (value) *value = h.pv[0] + h.pv[1];
  return SU2HIP_OK;
}

int su2hip_muca_set_orderparam(su2hip_batch_t *b, int slot, const double *pv) {
  int rc = check_slot(b, slot);
  if (rc) return rc;
  if (!pv) return fail(SU2HIP_EINVAL, "NULL pv");
  rc = pull_muca(b, slot);
  if (rc) return rc;
  Slot &s = b->slots[slot];
  if (!s.hm.enabled) return fail(SU2HIP_ESTATE, "slot %d has no multicanonical weight", slot);
  s.hm.pv[0] = pv[0];
  s.hm.pv[1] = pv[1];
  return push_muca(b, slot);
}

int su2hip_sync(su2hip_batch_t *b) {
  if (!b) return fail(SU2HIP_EINVAL, "NULL handle");
  HIPTRY(hipStreamSynchronize(b->stream));
  return SU2HIP_OK;
}

int su2hip_get_counters(su2hip_batch_t *b, int slot, su2hip_counters *c) {
  int rc = check_slot(b, slot);
  if (rc) return rc;
  if (!c) return fail(SU2HIP_EINVAL, "NULL counters");
  rc = prepare(b);
  if (rc) return rc;
  rc = fold_counts(b);
  if (rc) return rc;
  rc = pull_muca(b, slot);
  if (rc) return rc;
  *c = b->slots[slot].cnt;
  c->total_muca = b->slots[slot].hm.total;
  c->accepted_muca = b->slots[slot].hm.accepted;
  return SU2HIP_OK;
}

int su2hip_reset_counters(su2hip_batch_t *b, int slot) {
  int rc = check_slot(b, slot);
  if (rc) return rc;
  rc = prepare(b);
  if (rc) return rc;
  rc = fold_counts(b);
  if (rc) return rc;
  b->slots[slot].cnt = su2hip_counters();
  rc = pull_muca(b, slot);
  if (rc) return rc;
  b->slots[slot].hm.total = 0;
  b->slots[slot].hm.accepted = 0;
  return push_muca(b, slot);
}

int su2hip_rng_get(su2hip_batch_t *b, int slot, uint64_t *seed, uint64_t *pass) {
  int rc = check_slot(b, slot);
  if (rc) return rc;
  rc = prepare(b);
  if (rc) return rc;
  rc = pull_desc(b);
  if (rc) return rc;
  const LatDesc &L = b->slots[slot].hd;
  if (seed) *seed = (uint64_t)L.key0 | ((uint64_t)L.key1 << 32);
  if (pass) *pass = L.pass;
  return SU2HIP_OK;
}

int su2hip_rng_set(su2hip_batch_t *b, int slot, uint64_t seed, uint64_t pass) {
  int rc = check_slot(b, slot);
  if (rc) return rc;
  rc = prepare(b);
  if (rc) return rc;
  rc = pull_desc(b);
  if (rc) return rc;
  LatDesc &L = b->slots[slot].hd;
  L.key0 = (uint32_t)seed;
  L.key1 = (uint32_t)(seed >> 32);
  L.pass = pass;
  b->slots[slot].p.seed = seed;
  return push_desc(b);
}

int su2hip_measure_all(su2hip_batch_t *b, double *sums) {
  TraceRange tr_("su2hip_measure");
  if (!b || !sums) return fail(SU2HIP_EINVAL, "NULL argument");
  int rc = prepare(b);
  if (rc) return rc;
  LaunchCtx c = ctx_of(b);
  hipError_t e = timed(b, SU2HIP_K_MEASURE, [&] { return launch_measure(c); });
  if (e != hipSuccess) return fail(SU2HIP_EDEVICE, "measure: %s", hipGetErrorString(e));
  HIPTRY(hipMemcpyAsync(sums, b->d_sums, sizeof(double) * SU2HIP_NSUMS * b->max_slots,
                        hipMemcpyDeviceToHost, b->stream));
  HIPTRY(hipStreamSynchronize(b->stream));
  return SU2HIP_OK;
}

int su2hip_measure(su2hip_batch_t *b, int slot, double *sums) {
  int rc = check_slot(b, slot);
  if (rc) return rc;
  if (!sums) return fail(SU2HIP_EINVAL, "NULL sums");
  std::vector<double> all((size_t)SU2HIP_NSUMS * b->max_slots);
  rc = su2hip_measure_all(b, all.data());
  if (rc) return rc;
  memcpy(sums, all.data() + (size_t)slot * SU2HIP_NSUMS, sizeof(double) * SU2HIP_NSUMS);
  return SU2HIP_OK;
}

int su2hip_measure_noalloc(su2hip_batch_t *b, int slot, double *sums_device) {
  if (!b) return fail(SU2HIP_EINVAL, "NULL handle");
  if (slot < 0 || slot >= b->max_slots) return fail(SU2HIP_ERANGE, "slot %d out of range", slot);
  if (!b->slots[slot].used) return fail(SU2HIP_ERANGE, "slot %d is empty", slot);
  if (!sums_device) return fail(SU2HIP_EINVAL, "NULL sums");
  const LatDesc &L = b->slots[slot].hd;
  int nblk = (L.Vh + BLK - 1) / BLK;
  if (nblk > SU2HIP_NOALLOC_MAX_BLOCKS)
    return fail(SU2HIP_ERANGE, "lattice of %d sites exceeds the scratch of su2hip_measure_noalloc "
                "(at most %d sites)", L.vol, 2 * BLK * SU2HIP_NOALLOC_MAX_BLOCKS);
  int dev = 0;
  HIPTRY(hipGetDevice(&dev));
  double *scratch = (dev >= 0 && dev < kMaxDevices) ? g_noalloc_ptr[dev] : nullptr;
  if (!scratch) return fail(SU2HIP_ESTATE, "no handle was created on device %d", dev);
  // every block writes all of its partials: nothing to clear
  hipError_t e = launch_measure_one(b->stream, b->fp64, b->model, L, scratch, sums_device);
  if (e != hipSuccess) return fail(SU2HIP_EDEVICE, "measure_noalloc: %s", hipGetErrorString(e));
  return SU2HIP_OK;
}

// column values in the order of measure()'s fprintf (src/measure.c:252-306)
int su2hip_measure_columns(const su2hip_params *p, const double *s, double weight,
                           double muca_param, double *cols, int *ncols) {
  if (!p || !s || !cols || !ncols) return fail(SU2HIP_EINVAL, "NULL argument");
  double vol = (double)((long)p->L[0] * p->L[1] * p->L[2]);
  int n = 0;
  cols[n++] = weight;
  cols[n++] = muca_param;
  cols[n++] = s[SU2HIP_SUM_ACTION];
  cols[n++] = s[SU2HIP_SUM_WILSON] / (vol * p->betasu2);
  if (p->u1) cols[n++] = s[SU2HIP_SUM_U1WILSON] / (vol * p->betau1);
  if (p->nhiggs) {
    cols[n++] = s[SU2HIP_SUM_COV_PHI] / (vol * 3);
    cols[n++] = s[SU2HIP_SUM_PHI2] / vol;
    cols[n++] = s[SU2HIP_SUM_PHI4] / vol;
  }
  if (p->nhiggs == 2) {
    cols[n++] = s[SU2HIP_SUM_COV_PHI_2] / (vol * 3);
    cols[n++] = s[SU2HIP_SUM_PHI2_2] / vol;
    cols[n++] = s[SU2HIP_SUM_PHI4_2] / vol;
    cols[n++] = s[SU2HIP_SUM_R12] / vol;
    cols[n++] = s[SU2HIP_SUM_I12] / vol;
  }
  if (p->triplet) {
    cols[n++] = s[SU2HIP_SUM_HOP_SIGMA] / (vol * 3);
    cols[n++] = s[SU2HIP_SUM_SIGMA2] / vol;
    cols[n++] = s[SU2HIP_SUM_SIGMA4] / vol;
    if (p->nhiggs) cols[n++] = s[SU2HIP_SUM_PHI2SIGMA2] / vol;
    cols[n++] = s[SU2HIP_SUM_MAGCHARGE];
    cols[n++] = s[SU2HIP_SUM_MAGCHARGE_ABS] / (2.0 * M_PI * sqrt(p->betasu2));
  }
  if (p->singlet) {
    cols[n++] = s[SU2HIP_SUM_S_COV] / (vol * 3);
    cols[n++] = s[SU2HIP_SUM_S] / vol;
    cols[n++] = s[SU2HIP_SUM_S2] / vol;
    cols[n++] = s[SU2HIP_SUM_S3] / vol;
    cols[n++] = s[SU2HIP_SUM_S4] / vol;
    if (p->nhiggs == 1) {
      cols[n++] = s[SU2HIP_SUM_SPHISQ] / vol;
      cols[n++] = s[SU2HIP_SUM_S2PHISQ] / vol;
    }
  }
  *ncols = n;
  return SU2HIP_OK;
}

int su2hip_muca_set(su2hip_batch_t *b, int slot, const su2hip_muca *m, const double *pos,
                    const double *W) {
  TraceRange tr_("su2hip_muca_set");
  int rc = check_slot(b, slot);
  if (rc) return rc;
  if (!m) return fail(SU2HIP_EINVAL, "NULL muca");
  Slot &s = b->slots[slot];
  if (m->enabled) {
    if (!pos || !W) return fail(SU2HIP_EINVAL, "pos/W required");
    if (m->bins < 2) return fail(SU2HIP_EINVAL, "bins = %d", m->bins);
    if (m->checks_per_sweep < 1) return fail(SU2HIP_EINVAL, "checks_per_sweep = %d", m->checks_per_sweep);
    if (m->mode < SU2HIP_READONLY || m->mode > SU2HIP_SLOW) return fail(SU2HIP_EINVAL, "mode %d", m->mode);
    int op = m->orderparam;
    bool ok = (op == SU2HIP_PHISQ && s.p.nhiggs) || (op == SU2HIP_SIGMASQ && s.p.triplet) ||
              (op == SU2HIP_PHI2MINUSSIGMA2 && s.p.nhiggs && s.p.triplet) ||
              (op == SU2HIP_PHI2SQ && s.p.nhiggs == 2);
    if (!ok) return fail(SU2HIP_EINVAL, "order parameter %d needs a field this model lacks", op);
  }
  HIPTRY(hipStreamSynchronize(b->stream));
  // Everything new is allocated and filled first; the slot's previous
  // weight (or its disabled state) is replaced only once that succeeded, so
  // a failure leaves the slot exactly as it was.
  MucaDev h{};
  void *newmem = nullptr, *newbu = nullptr;
  size_t bytes = 0, bub = 0;
  if (m->enabled) {
    // backup of the order-parameter field (dofs of the doublet or triplet)
    size_t Vh = (size_t)vol_of(s.p) / 2;
    bub = align256((s.p.nhiggs ? 8 : 6) * Vh * wsize(s.p));
    if (!s.bu_mem) {
      if (b->hbm_budget && b->used + bub > b->hbm_budget)
        return fail(SU2HIP_ENOMEM, "HBM budget exceeded by the multicanonical backup");
      if (hipMalloc(&newbu, bub) != hipSuccess) return fail(SU2HIP_ENOMEM, "backup allocation");
    }
    int n = m->bins + 2;
    bytes = sizeof(double) * n * 6 + sizeof(int) * n + sizeof(long long) * n * 2 + 1024;
    if (hipMalloc(&newmem, bytes) != hipSuccess) {
      if (newbu) (void)hipFree(newbu);
      return fail(SU2HIP_ENOMEM, "weight arrays");
    }
    char *q = (char *)newmem;
    auto take = [&](size_t sz) {
      void *r = q;
      q += (sz + 15) & ~(size_t)15;
      return r;
    };
    h.pos = (double *)take(sizeof(double) * n);
    h.W = (double *)take(sizeof(double) * n);
    h.slope = (double *)take(sizeof(double) * n);
    h.b = (double *)take(sizeof(double) * n);
    h.hgram = (double *)take(sizeof(double) * n * 2);
    h.hits = (int *)take(sizeof(int) * n);
    h.gsum = (long long *)take(sizeof(long long) * n);
    h.nsum = (long long *)take(sizeof(long long) * n);
    std::vector<double> slope(n), bb(n), hg(2 * n, 0.0);
    std::vector<int> hits(n, 0);
    std::vector<long long> gs(n, 5), ns(n, 1);
    // linearize_weight (src/multicanonical.c:337-346)
    slope[0] = 0;
    slope[n - 1] = 0;
    bb[0] = W[0];
    bb[n - 1] = W[n - 1];
    for (int i = 1; i < n - 1; i++) {
      slope[i] = (W[i + 1] - W[i]) / (pos[i + 1] - pos[i]);
      bb[i] = W[i] - slope[i] * pos[i];
    }
    if (m->mode == SU2HIP_SLOW)
      for (int i = 0; i < n; i++) hg[i] = exp(W[i]);
    const struct {
      void *dst;
      const void *src;
      size_t sz;
    } up[] = {{h.pos, pos, sizeof(double) * n},           {h.W, W, sizeof(double) * n},
              {h.slope, slope.data(), sizeof(double) * n}, {h.b, bb.data(), sizeof(double) * n},
              {h.hgram, hg.data(), sizeof(double) * 2 * n}, {h.hits, hits.data(), sizeof(int) * n},
              {h.gsum, gs.data(), sizeof(long long) * n},  {h.nsum, ns.data(), sizeof(long long) * n}};
    for (const auto &u : up) {
      hipError_t e = hipMemcpy(u.dst, u.src, u.sz, hipMemcpyHostToDevice);
      if (e != hipSuccess) {
        (void)hipFree(newmem);
        if (newbu) (void)hipFree(newbu);
        return fail(SU2HIP_EDEVICE, "weight upload: %s", hipGetErrorString(e));
      }
    }
    h.enabled = 1;
    h.orderparam = m->orderparam;
    h.mode = m->mode;
    h.cps = m->checks_per_sweep;
    h.bins = m->bins;
    h.last_max = m->last_max;
    h.do_acceptance = m->do_acceptance;
    h.muca_count = m->muca_count;
    h.update_interval = m->update_interval > 0 ? m->update_interval
                                               : (m->mode == SU2HIP_SLOW ? 5000 : 100);
    h.delta = m->delta;
    h.min = m->min;
    h.max = m->max;
    h.wrk_min = m->wrk_min;
    h.wrk_max = m->wrk_max;
  }
  // commit: nothing below allocates
  if (s.muca_mem) (void)hipFree(s.muca_mem);
  s.muca_mem = newmem;
  s.muca_bytes = bytes;
  if (newbu) {
    s.bu_mem = newbu;
    s.bu_bytes = bub;
    b->used += bub;
    s.hd.B = s.bu_mem;
  }
  s.hm = h;
  rc = push_muca(b, slot);
  if (rc) return rc;
  b->sched_dirty = true;
  b->graph_valid = false;
  rc = prepare(b);
  if (rc) return rc;
  rc = push_desc(b);
  if (rc) return rc;
  return orderparam_refresh(b, slot);
}

int su2hip_muca_get(su2hip_batch_t *b, int slot, su2hip_muca *m, double *pos, double *W,
                    int32_t *hits) {
  int rc = check_slot(b, slot);
  if (rc) return rc;
  if (!m) return fail(SU2HIP_EINVAL, "NULL muca");
  rc = pull_muca(b, slot);
  if (rc) return rc;
  const MucaDev &h = b->slots[slot].hm;
  memset(m, 0, sizeof(*m));
  m->enabled = h.enabled;
  m->orderparam = h.orderparam;
  m->mode = h.mode;
  m->checks_per_sweep = h.cps;
  m->bins = h.bins;
  m->last_max = h.last_max;
  m->do_acceptance = h.do_acceptance;
  m->muca_count = h.muca_count;
  m->update_interval = h.update_interval;
  m->delta = h.delta;
  m->min = h.min;
  m->max = h.max;
  m->wrk_min = h.wrk_min;
  m->wrk_max = h.wrk_max;
  m->param_value[0] = h.pv[0];
  m->param_value[1] = h.pv[1];
  if (h.enabled) {
    int n = h.bins + 2;
    if (pos) HIPTRY(hipMemcpy(pos, h.pos, sizeof(double) * n, hipMemcpyDeviceToHost));
    if (W) HIPTRY(hipMemcpy(W, h.W, sizeof(double) * n, hipMemcpyDeviceToHost));
    if (hits) HIPTRY(hipMemcpy(hits, h.hits, sizeof(int) * n, hipMemcpyDeviceToHost));
  }
  return SU2HIP_OK;
}

int su2hip_correlators(su2hip_batch_t *b, int slot, int dir, double *out, int max_dist,
                       int *ndist) {
  TraceRange tr_("su2hip_correlators");
  int rc = check_slot(b, slot);
  if (rc) return rc;
  const su2hip_params &p = b->slots[slot].p;
  if (dir < 0) dir = longest_dir(p);
  if (dir > 2) return fail(SU2HIP_EINVAL, "direction %d", dir);
  rc = prepare(b);
  if (rc) return rc;
  int Ld = p.L[dir];
  size_t V = (size_t)vol_of(p), P = V / Ld;
  rc = ensure_stage(b, 10 * V + 10 * (size_t)Ld);
  if (rc) return rc;
  double *buf = b->d_stage, *planes = b->d_stage + 10 * V;
  HIPTRY(launch_corr(ctx_of(b), slot, dir, (int)(V / 2), Ld, P, buf, planes));
  std::vector<double> pl(10 * (size_t)Ld);
  HIPTRY(hipMemcpyAsync(pl.data(), planes, sizeof(double) * pl.size(), hipMemcpyDeviceToHost, b->stream));
  HIPTRY(hipStreamSynchronize(b->stream));
  // higgs_correlator, triplet_correlator and photon_correl
  // (src/correlation.c:40-305) from the plane sums
  int maxd = Ld / 2;
  for (int d = 0; d < maxd && d < max_dist && out; d++) {
    double r[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int z = 0; z < Ld; z++) {
      int zd = (z + d) % Ld;
      auto a = [&](int k) { return pl[(size_t)k * Ld + z]; };
      auto c = [&](int k) { return pl[(size_t)k * Ld + zd]; };
      r[0] += a(0) * c(0);
      r[1] += a(1) * c(1);
      for (int k = 2; k < 10; k += 2) {
        // cmult(op1, conj(op2)) (src/generic/stddefs.h:93-98)
        double br = c(k), bi = -c(k + 1);
        r[k] += a(k) * br - a(k + 1) * bi;
        r[k + 1] += br * a(k + 1) + a(k) * bi;
      }
    }
    double *o = out + (size_t)d * 10;
    o[0] = 4.0 * r[0] / (double)V;
    for (int k = 1; k < 10; k++) o[k] = r[k] / (double)V;
  }
  if (ndist) *ndist = maxd;
  return SU2HIP_OK;
}

// blocked lattice of `base` in the directions of bmask, fields at U/G
static LatDesc blocked_desc(const LatDesc &base, int bmask, void *U, void *G) {
  LatDesc B = base;
  B.L0 = (bmask & 1) ? base.L0 / 2 : base.L0;
  B.L1 = (bmask & 2) ? base.L1 / 2 : base.L1;
  B.L2 = (bmask & 4) ? base.L2 / 2 : base.L2;
  B.h0 = B.L0 / 2;
  B.vol = B.L0 * B.L1 * B.L2;
  B.Vh = B.vol / 2;
  B.div_h0 = make_fastdiv((uint32_t)B.h0);
  B.div_L1 = make_fastdiv((uint32_t)B.L1);
  B.U = U;
  B.G = G;
  B.F = B.A = B.S = B.B = nullptr;
  B.blk_base = 0;
  B.nblk = (B.Vh + BLK - 1) / BLK;
  B.active = 1;
  return B;
}

int su2hip_blocked_correlators(su2hip_batch_t *b, int slot, int levels, double *out,
                               int *nlevels, int *ndist) {
  int rc = check_slot(b, slot);
  if (rc) return rc;
  const su2hip_params &p = b->slots[slot].p;
  int dir = longest_dir(p), maxd = p.L[dir] / 2;
  if (nlevels) *nlevels = 0;
  if (ndist) *ndist = maxd;
  // measure_blocked_correlators smears the triplet only and measures
  // nothing with a doublet (src/correlation.c:403-408)
  if (p.nhiggs || !p.triplet || levels <= 0) return SU2HIP_OK;
  int bmask = 7 & ~(1 << dir);
  // max_block_level (src/blocking.c:19-38)
  int len[3] = {p.L[0], p.L[1], p.L[2]}, maxl = 0;
  for (bool ok = true; ok;) {
    for (int d = 0; d < 3 && ok; d++)
      if ((bmask & (1 << d)) && (len[d] % 2 != 0 || len[d] <= 2)) ok = false;
    if (!ok) break;
    for (int d = 0; d < 3; d++)
      if (bmask & (1 << d)) len[d] /= 2;
    maxl++;
  }
  if (levels > maxl) levels = maxl;
  if (levels == 0) return SU2HIP_OK;
  rc = prepare(b);
  if (rc) return rc;
  rc = pull_desc(b);
  if (rc) return rc;
  LatDesc fine = b->slots[slot].hd;
  size_t w = wsize(p), Vh1 = (size_t)fine.Vh / 4;  // first level: a quarter of the sites
  size_t lw = link_words(Vh1, packed_links(p));
  size_t words = align256(lw * w) + align256(6 * Vh1 * w);
  char *mem = nullptr;
  LatDesc *ddesc = nullptr;
  if (hipMalloc(&mem, 2 * words) != hipSuccess) return fail(SU2HIP_ENOMEM, "blocked fields");
  if (hipMalloc(&ddesc, sizeof(LatDesc)) != hipSuccess) {
    (void)hipFree(mem);
    return fail(SU2HIP_ENOMEM, "blocked descriptor");
  }
  hipError_t e = hipSuccess;
  LatDesc base = fine;
  for (int k = 0; k < levels && e == hipSuccess; k++) {
    char *m = mem + (size_t)(k & 1) * words;
    LatDesc B = blocked_desc(base, bmask, m, m + align256(lw * w));
    e = launch_block(ctx_of(b), base, B, bmask);
    if (e == hipSuccess) e = hipMemcpyAsync(ddesc, &B, sizeof(LatDesc), hipMemcpyHostToDevice, b->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    if (e != hipSuccess) break;
    // correlators on the blocked lattice (su2hip_correlators' plane sums)
    size_t V = (size_t)B.vol, P = V / p.L[dir];
    if (ensure_stage(b, 10 * V + 10 * (size_t)p.L[dir])) {
      e = hipErrorOutOfMemory;
      break;
    }
    LaunchCtx c = ctx_of(b);
    c.desc = ddesc;
    double *buf = b->d_stage, *planes = b->d_stage + 10 * V;
    e = launch_corr(c, 0, dir, B.Vh, p.L[dir], P, buf, planes);
    std::vector<double> pl(10 * (size_t)p.L[dir]);
    if (e == hipSuccess)
      e = hipMemcpyAsync(pl.data(), planes, sizeof(double) * pl.size(), hipMemcpyDeviceToHost, b->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    if (e != hipSuccess) break;
    int Ld = p.L[dir];
    for (int d = 0; d < maxd && out; d++) {
      double r[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
      for (int z = 0; z < Ld; z++) {
        int zd = (z + d) % Ld;
        auto a = [&](int q) { return pl[(size_t)q * Ld + z]; };
        auto cc = [&](int q) { return pl[(size_t)q * Ld + zd]; };
        r[1] += a(1) * cc(1);
        for (int q = 2; q < 10; q += 2) {
          double br = cc(q), bi = -cc(q + 1);
          r[q] += a(q) * br - a(q + 1) * bi;
          r[q + 1] += br * a(q + 1) + a(q) * bi;
        }
      }
      double *o = out + ((size_t)k * maxd + d) * 10;
      o[0] = 0.0;
      for (int q = 1; q < 10; q++) o[q] = r[q] / (double)V;
    }
    base = B;
  }
  (void)hipStreamSynchronize(b->stream);
  (void)hipFree(ddesc);
  (void)hipFree(mem);
  if (e != hipSuccess) return fail(SU2HIP_EDEVICE, "blocked correlators: %s", hipGetErrorString(e));
  if (nlevels) *nlevels = levels;
  return SU2HIP_OK;
}

int su2hip_muca_set_sums(su2hip_batch_t *b, int slot, const int64_t *gsum, const int64_t *nsum) {
  int rc = check_slot(b, slot);
  if (rc) return rc;
  rc = pull_muca(b, slot);
  if (rc) return rc;
  const MucaDev &h = b->slots[slot].hm;
  if (!h.enabled) return fail(SU2HIP_ESTATE, "slot %d has no multicanonical weight", slot);
  int n = h.bins + 2;
  static_assert(sizeof(long long) == sizeof(int64_t), "gsum/nsum element size");
  if (gsum) HIPTRY(hipMemcpy(h.gsum, gsum, sizeof(long long) * n, hipMemcpyHostToDevice));
  if (nsum) HIPTRY(hipMemcpy(h.nsum, nsum, sizeof(long long) * n, hipMemcpyHostToDevice));
  return SU2HIP_OK;
}

int su2hip_muca_get_sums(su2hip_batch_t *b, int slot, int64_t *gsum, int64_t *nsum) {
  int rc = check_slot(b, slot);
  if (rc) return rc;
  HIPTRY(hipStreamSynchronize(b->stream));
  rc = pull_muca(b, slot);
  if (rc) return rc;
  const MucaDev &h = b->slots[slot].hm;
  if (!h.enabled) return fail(SU2HIP_ESTATE, "slot %d has no multicanonical weight", slot);
  int n = h.bins + 2;
  if (gsum) HIPTRY(hipMemcpy(gsum, h.gsum, sizeof(long long) * n, hipMemcpyDeviceToHost));
  if (nsum) HIPTRY(hipMemcpy(nsum, h.nsum, sizeof(long long) * n, hipMemcpyDeviceToHost));
  return SU2HIP_OK;
}

int su2hip_muca_weight_of(su2hip_batch_t *b, int slot, double v, double *weight) {
  int rc = check_slot(b, slot);
  if (rc) return rc;
  if (!weight) return fail(SU2HIP_EINVAL, "NULL weight");
  rc = pull_muca(b, slot);
  if (rc) return rc;
  const MucaDev &h = b->slots[slot].hm;
  if (!h.enabled) return fail(SU2HIP_ESTATE, "no weight installed");
  int n = h.bins + 2;
  std::vector<double> pos(n), bb(n), sl(n);
  HIPTRY(hipMemcpy(pos.data(), h.pos, sizeof(double) * n, hipMemcpyDeviceToHost));
  HIPTRY(hipMemcpy(bb.data(), h.b, sizeof(double) * n, hipMemcpyDeviceToHost));
  HIPTRY(hipMemcpy(sl.data(), h.slope, sizeof(double) * n, hipMemcpyDeviceToHost));
  // get_weight / whichbin (src/multicanonical.c:350-379, :671-704)
  if (v >= h.max) {
    *weight = bb[h.bins + 1];
    return SU2HIP_OK;
  }
  if (v < h.min) {
    *weight = bb[0];
    return SU2HIP_OK;
  }
  int lo = 1, hi = h.bins, bin = h.bins;
  while (lo <= hi) {
    int k = lo + (hi - lo) / 2;
    if ((v < pos[k + 1] && v > pos[k]) || fabs(v - pos[k]) < 1e-12) {
      bin = k;
      break;
    }
    if (v > pos[k])
      lo = k + 1;
    else
      hi = k - 1;
  }
  *weight = bin >= h.bins + 1 ? bb[h.bins + 1] : bb[bin] + sl[bin] * v;
  return SU2HIP_OK;
}

// remove_counterterms (src/gradflow.c:424-459): tree-level triplet mass
static double triplet_flow_msq(const su2hip_params &p) {
  double Sg = 3.17591153625, zeta = 0.08849, delta = 1.942130, rho = -0.313964;
  double k1 = 0.958382, k2 = 0.25 * Sg * Sg - 0.5 * delta - 0.25;
  double k3 = 0.751498, k4 = 1.204295;
  double gsq = 4.0 / p.betasu2, RG = gsq, b4 = p.b4;
  double ct = -(4.0 * gsq + 5.0 * b4) * Sg / (4.0 * M_PI);
  ct += -1.0 / (16.0 * M_PI * M_PI) *
        ((20.0 * b4 * gsq - 10.0 * b4 * b4) * (log(6.0 / (RG)) + zeta) +
         20.0 * b4 * gsq * (0.25 * Sg * Sg - delta) +
         2.0 * gsq * gsq *
             (1.25 * Sg * Sg + M_PI / 3.0 * Sg - 6.0 * delta - 6.0 * rho + 4.0 * k1 - k2 - k3 -
              3.0 * k4));
  return p.msq_triplet - ct;
}

// a one-lattice launch context for `slot` with the given descriptor array
static int single_ctx(su2hip_batch *b, int slot, LatDesc *ddesc, BlockMap *dblk, int nb,
                      LaunchCtx &c) {
  c = ctx_of(b);
  c.desc = ddesc;
  c.blk = dblk;
  c.nblk = nb;
  (void)slot;
  return SU2HIP_OK;
}

int su2hip_flow_force(su2hip_batch_t *b, int slot, double *force) {
  int rc = check_slot(b, slot);
  if (rc) return rc;
  if (!force) return fail(SU2HIP_EINVAL, "NULL force");
  rc = prepare(b);
  if (rc) return rc;
  const Slot &s = b->slots[slot];
  size_t V = vol_of(s.p);
  rc = ensure_stage(b, V * 12);
  if (rc) return rc;
  int nb = (s.hd.Vh + BLK - 1) / BLK;
  std::vector<BlockMap> bm(nb);
  for (int k = 0; k < nb; k++) bm[k] = BlockMap{slot, k * BLK};
  BlockMap *dblk = nullptr;
  if (hipMalloc(&dblk, sizeof(BlockMap) * nb) != hipSuccess) return fail(SU2HIP_ENOMEM, "block table");
  hipError_t e = hipMemcpy(dblk, bm.data(), sizeof(BlockMap) * nb, hipMemcpyHostToDevice);
  LaunchCtx c;
  single_ctx(b, slot, b->d_desc, dblk, nb, c);
  if (e == hipSuccess) e = launch_flow_force(c, slot, b->d_stage);
  if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
  if (e == hipSuccess)
    e = hipMemcpy(force, b->d_stage, sizeof(double) * V * 12, hipMemcpyDeviceToHost);
  (void)hipFree(dblk);
  if (e != hipSuccess) return fail(SU2HIP_EDEVICE, "flow force: %s", hipGetErrorString(e));
  return SU2HIP_OK;
}

// Enqueue a whole flow of `slot` on stream st: copy the links (+ triplet)
// into the flow buffers, run the integrator stages, and at every
// measurement point reduce the flowed fields into job-private buffers and
// copy the sums into the job's row array.  Nothing waits on the host.
static int flow_enqueue(su2hip_batch *b, int slot, int scheme, double dt, double t_max,
                        int meas_interval, hipStream_t st, bool use_timing, FlowJob &j) {
  Slot &s = b->slots[slot];
  const su2hip_params &p = s.p;
  size_t Vh = (size_t)vol_of(p) / 2, w = wsize(p);
  if (!s.flow_mem) {
    size_t fb = align256(flow_words(p) * w) + 4096;
    if (b->hbm_budget && b->used + fb > b->hbm_budget)
      return fail(SU2HIP_ENOMEM, "HBM budget exceeded by the flow buffers");
    if (hipMalloc(&s.flow_mem, fb) != hipSuccess) return fail(SU2HIP_ENOMEM, "flow buffers");
    s.flow_bytes = fb;
    b->used += fb;
  }
  char *q = (char *)s.flow_mem;
  auto take = [&](size_t words) {
    void *r = q;
    q += align256(words * w);
    return r;
  };
  void *Ub[2] = {take(link_words(p)), take(link_words(p))};
  void *Z = take(18 * Vh);
  void *Gb[2] = {nullptr, nullptr}, *ZG = nullptr;
  if (p.triplet) {
    Gb[0] = take(6 * Vh);
    Gb[1] = take(6 * Vh);
    ZG = take(6 * Vh);
  }
  double msq_t = p.triplet ? triplet_flow_msq(p) : p.msq_triplet;
  // the time of every row: grad_flow's loop (src/gradflow.c:91-141)
  j.t.clear();
  j.t.push_back(0.0);
  std::vector<std::pair<double, int>> steps;  // (dt, measure after)
  {
    double t = 0.0, h = dt;
    int iter = 1;
    while (t < t_max) {
      if (t + h > t_max) h = t_max - t;
      t += h;
      bool m = iter % meas_interval == 0;
      steps.push_back({h, m ? 1 : 0});
      if (m) j.t.push_back(t);
      iter++;
    }
    if ((iter - 1) % meas_interval != 0) {
      if (!steps.empty()) steps.back().second = 1;
      j.t.push_back(t);
    }
  }
  int nb = (int)((Vh + BLK - 1) / BLK);
  std::vector<LatDesc> hd(2 * (size_t)b->max_slots);
  for (int k = 0; k < 2; k++)
    for (int i = 0; i < b->max_slots; i++) {
      LatDesc &L = hd[(size_t)k * b->max_slots + i];
      if (i == slot) {
        L = s.hd;
        L.U = Ub[k];
        if (p.triplet) L.G = Gb[k];
        L.msq_t = msq_t;
        L.blk_base = 0;
        L.nblk = nb;
        L.active = 1;
        L.budget = -1;
      } else {
        memset(&L, 0, sizeof(L));
      }
    }
  std::vector<BlockMap> seq(nb);
  for (int k = 0; k < nb; k++) seq[k] = BlockMap{slot, k * BLK};
  // XCD j gets a contiguous run of chunks, as in the sweep's table
  std::vector<BlockMap> bm = deal_to_xcds(seq);
  size_t nrows = j.t.size();
  if (hipMalloc(&j.ddesc, sizeof(LatDesc) * hd.size()) != hipSuccess ||
      hipMalloc(&j.dblk, sizeof(BlockMap) * nb) != hipSuccess ||
      hipMalloc(&j.d_partial, sizeof(double) * SU2HIP_NSUMS * (size_t)nb) != hipSuccess ||
      hipMalloc(&j.d_sums, sizeof(double) * SU2HIP_NSUMS * (size_t)b->max_slots) != hipSuccess ||
      hipMalloc(&j.d_rows, sizeof(double) * SU2HIP_NSUMS * nrows) != hipSuccess ||
      hipEventCreateWithFlags(&j.done, hipEventDisableTiming) != hipSuccess) {
    flow_release(j);
    return fail(SU2HIP_ENOMEM, "flow descriptors and measurement buffers");
  }
  j.active = true;
  hipError_t e = hipMemcpyAsync(j.ddesc, hd.data(), sizeof(LatDesc) * hd.size(), hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipMemcpyAsync(j.dblk, bm.data(), sizeof(BlockMap) * nb, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipMemcpyAsync(Ub[0], s.hd.U, link_words(p) * w, hipMemcpyDeviceToDevice, st);
  if (e == hipSuccess && p.triplet) e = hipMemcpyAsync(Gb[0], s.hd.G, 6 * Vh * w, hipMemcpyDeviceToDevice, st);
  int cur = 0;
  size_t r = 0;
  auto ctx = [&]() {
    LaunchCtx c = ctx_of(b);
    c.stream = st;
    c.desc = j.ddesc + (size_t)cur * b->max_slots;
    c.blk = j.dblk;
    c.nblk = nb;
    c.partial = j.d_partial;
    c.sums = j.d_sums;
    return c;
  };
  auto row = [&]() -> hipError_t {
    LaunchCtx c = ctx();
    hipError_t x = use_timing ? timed(b, SU2HIP_K_MEASURE, [&] { return launch_measure(c); })
                              : launch_measure(c);
    if (x == hipSuccess)
      x = hipMemcpyAsync(j.d_rows + r * SU2HIP_NSUMS, j.d_sums + (size_t)slot * SU2HIP_NSUMS,
                         sizeof(double) * SU2HIP_NSUMS, hipMemcpyDeviceToDevice, st);
    r++;
    return x;
  };
  if (e == hipSuccess) e = row();
  int stages = scheme == SU2HIP_FLOW_EULER ? 1 : 3;
  for (size_t k = 0; k < steps.size() && e == hipSuccess; k++) {
    double h = steps[k].first;
    for (int sg = 0; sg < stages && e == hipSuccess; sg++) {
      LaunchCtx c = ctx();
      int stage = scheme == SU2HIP_FLOW_EULER ? 0 : sg + 1;
      auto go = [&] {
        return launch_flow_stage(c, slot, Ub[cur], Ub[1 - cur], Gb[cur], Gb[1 - cur], Z, ZG, h,
                                 msq_t, stage);
      };
      e = use_timing ? timed(b, SU2HIP_K_FLOW, go) : go();
      cur = 1 - cur;
    }
    if (e == hipSuccess && steps[k].second) e = row();
  }
  if (e == hipSuccess) e = hipEventRecord(j.done, st);
  if (e != hipSuccess) {
    (void)hipStreamSynchronize(st);
    flow_release(j);
    return fail(SU2HIP_EDEVICE, "flow: %s", hipGetErrorString(e));
  }
  return SU2HIP_OK;
}

// rows of a finished flow job, then release it
static int flow_collect(FlowJob &j, double *rows, int max_rows, int *nrows) {
  hipError_t e = hipEventSynchronize(j.done);
  size_t n = j.t.size();
  std::vector<double> h(n * SU2HIP_NSUMS);
  if (e == hipSuccess)
    e = hipMemcpy(h.data(), j.d_rows, sizeof(double) * h.size(), hipMemcpyDeviceToHost);
  if (e == hipSuccess && rows)
    for (size_t r = 0; r < n && (int)r < max_rows; r++) {
      double *o = rows + r * (1 + SU2HIP_NSUMS);
      o[0] = j.t[r];
      memcpy(o + 1, h.data() + r * SU2HIP_NSUMS, sizeof(double) * SU2HIP_NSUMS);
    }
  flow_release(j);
  if (e != hipSuccess) return fail(SU2HIP_EDEVICE, "flow: %s", hipGetErrorString(e));
  if (nrows) *nrows = (int)n;
  return SU2HIP_OK;
}

static int flow_args(su2hip_batch *b, int slot, int scheme, double dt, int meas_interval) {
  int rc = check_slot(b, slot);
  if (rc) return rc;
  if (scheme != SU2HIP_FLOW_EULER && scheme != SU2HIP_FLOW_RK3)
    return fail(SU2HIP_EINVAL, "flow scheme %d", scheme);
  if (!(dt > 0) || meas_interval < 1) return fail(SU2HIP_EINVAL, "dt/meas_interval");
  if (b->slots[slot].fj.active) return fail(SU2HIP_ESTATE, "slot %d already has a flow in flight", slot);
  return SU2HIP_OK;
}

int su2hip_flow(su2hip_batch_t *b, int slot, int scheme, double dt, double t_max,
                int meas_interval, double *rows, int max_rows, int *nrows) {
  TraceRange tr_("su2hip_flow");
  int rc = flow_args(b, slot, scheme, dt, meas_interval);
  if (rc) return rc;
  rc = prepare(b);
  if (rc) return rc;
  rc = pull_desc(b);
  if (rc) return rc;
  FlowJob &j = b->slots[slot].fj;
  rc = flow_enqueue(b, slot, scheme, dt, t_max, meas_interval, b->stream, true, j);
  if (rc) return rc;
  return flow_collect(j, rows, max_rows, nrows);
}

int su2hip_flow_start(su2hip_batch_t *b, int slot, int scheme, double dt, double t_max,
                      int meas_interval) {
  TraceRange tr_("su2hip_flow_start");
  int rc = flow_args(b, slot, scheme, dt, meas_interval);
  if (rc) return rc;
  rc = prepare(b);
  if (rc) return rc;
  rc = pull_desc(b);
  if (rc) return rc;
  if (b->slots[slot].hd.budget != 0)
    return fail(SU2HIP_ESTATE, "slot %d must have sweep budget 0 while its flow runs", slot);
  if (!b->flow_stream && hipStreamCreateWithFlags(&b->flow_stream, hipStreamNonBlocking) != hipSuccess)
    return fail(SU2HIP_EDEVICE, "flow stream");
  // the flow starts after the work already queued on the handle's stream
  hipEvent_t ev;
  HIPTRY(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
  hipError_t e = hipEventRecord(ev, b->stream);
  if (e == hipSuccess) e = hipStreamWaitEvent(b->flow_stream, ev, 0);
  (void)hipEventDestroy(ev);
  if (e != hipSuccess) return fail(SU2HIP_EDEVICE, "flow start: %s", hipGetErrorString(e));
  return flow_enqueue(b, slot, scheme, dt, t_max, meas_interval, b->flow_stream, false,
                      b->slots[slot].fj);
}

int su2hip_flow_poll(su2hip_batch_t *b, int slot, int *done) {
  int rc = check_slot(b, slot);
  if (rc) return rc;
  if (!done) return fail(SU2HIP_EINVAL, "NULL done");
  const FlowJob &j = b->slots[slot].fj;
  if (!j.active) return fail(SU2HIP_ESTATE, "slot %d has no flow in flight", slot);
  hipError_t e = hipEventQuery(j.done);
  if (e != hipSuccess && e != hipErrorNotReady) return fail(SU2HIP_EDEVICE, "flow: %s", hipGetErrorString(e));
  *done = e == hipSuccess;
  return SU2HIP_OK;
}

int su2hip_flow_finish(su2hip_batch_t *b, int slot, double *rows, int max_rows, int *nrows) {
  TraceRange tr_("su2hip_flow_finish");
  int rc = check_slot(b, slot);
  if (rc) return rc;
  FlowJob &j = b->slots[slot].fj;
  if (!j.active) return fail(SU2HIP_ESTATE, "slot %d has no flow in flight", slot);
  return flow_collect(j, rows, max_rows, nrows);
}

int su2hip_set_timing(su2hip_batch_t *b, int enable) {
  if (!b) return fail(SU2HIP_EINVAL, "NULL handle");
  b->timing = enable != 0;
  return SU2HIP_OK;
}

static int resolve_events(su2hip_batch *b) {
  if (b->pending.empty()) return SU2HIP_OK;
  HIPTRY(hipStreamSynchronize(b->stream));
  for (auto &ev : b->pending) {
    float ms = 0.f;
    HIPTRY(hipEventElapsedTime(&ms, ev.a, ev.b));
    b->kms[ev.kind] += ms;
    b->kcount[ev.kind]++;
    b->pool.push_back(ev.a);
    b->pool.push_back(ev.b);
  }
  b->pending.clear();
  return SU2HIP_OK;
}

int su2hip_kernel_times(su2hip_batch_t *b, double *ms, int64_t *launches) {
  if (!b) return fail(SU2HIP_EINVAL, "NULL handle");
  int rc = resolve_events(b);
  if (rc) return rc;
  for (int k = 0; k < SU2HIP_NKINDS; k++) {
    if (ms) ms[k] = b->kms[k];
    if (launches) launches[k] = b->kcount[k];
  }
  return SU2HIP_OK;
}

int su2hip_reset_timing(su2hip_batch_t *b) {
  if (!b) return fail(SU2HIP_EINVAL, "NULL handle");
  int rc = resolve_events(b);
  if (rc) return rc;
  for (int k = 0; k < SU2HIP_NKINDS; k++) {
    b->kms[k] = 0;
    b->kcount[k] = 0;
  }
  return SU2HIP_OK;
}

static int time_pass_impl(su2hip_batch *b, int kind, int repeats, double *ms_kernel,
                          double *ms_span) {
  if (!b || (!ms_kernel && !ms_span)) return fail(SU2HIP_EINVAL, "NULL argument");
  if (kind < 0 || kind >= NCOUNT || repeats < 1) return fail(SU2HIP_EINVAL, "kind/repeats");
  int rc = prepare(b);
  if (rc) return rc;
  // directions of link kinds, doublets of doublet kinds
  int ndir = (kind <= SU2HIP_K_U1_METROPOLIS) ? 3 : 1;
  if (kind == SU2HIP_K_DOUBLET_OVERRELAX || kind == SU2HIP_K_DOUBLET_METROPOLIS)
    ndir = std::max(1, (int)b->tmpl.nhiggs);
  // one start/stop pair per launch, stamped by the kernel dispatch itself
  std::vector<hipEvent_t> ev(2 * (size_t)repeats + 2, nullptr);
  hipError_t e = hipSuccess;
  for (size_t i = 0; i < ev.size() && e == hipSuccess; i++) e = hipEventCreate(&ev[i]);
  LaunchCtx c = ctx_for(b, kind);
  int reunit = b->tmpl.reunitarize && !b->fp64;
  hipEvent_t s0 = ev[2 * (size_t)repeats], s1 = ev[2 * (size_t)repeats + 1];
  // link passes read the hopping-term cache, filled once outside the timing
  if (e == hipSuccess && b->hop_on && su2_link_kind(kind)) e = launch_hop(ctx_of(b));
  if (e == hipSuccess) e = hipEventRecord(s0, b->stream);
  for (int i = 0; i < repeats && e == hipSuccess; i++) {
    c.ev_start = ev[2 * (size_t)i];
    c.ev_stop = ev[2 * (size_t)i + 1];
    e = launch_pass(c, kind, i & 1, (i >> 1) % ndir, i, -1, reunit);
  }
  if (e == hipSuccess) e = hipEventRecord(s1, b->stream);
  if (e == hipSuccess) e = hipEventSynchronize(s1);
  double sum = 0.0;
  float span = 0.f;
  for (int i = 0; i < repeats && e == hipSuccess; i++) {
    float t = 0.f;
    e = hipEventElapsedTime(&t, ev[2 * (size_t)i], ev[2 * (size_t)i + 1]);
    sum += t;
  }
  if (e == hipSuccess) e = hipEventElapsedTime(&span, s0, s1);
  for (hipEvent_t x : ev)
    if (x) (void)hipEventDestroy(x);
  c = ctx_of(b);
  if (e == hipSuccess) e = launch_end_sweep(c, repeats);
  if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
  if (e == hipErrorInvalidValue) return fail(SU2HIP_EINVAL, "kind %d is not part of this model", kind);
  if (e != hipSuccess) return fail(SU2HIP_EDEVICE, "time_pass: %s", hipGetErrorString(e));
  if (ms_kernel) *ms_kernel = sum / repeats;
  if (ms_span) *ms_span = (double)span / repeats;
  return SU2HIP_OK;
}

int su2hip_time_pass(su2hip_batch_t *b, int kind, int repeats, double *ms_kernel,
                     double *ms_span) {
  return time_pass_impl(b, kind, repeats, ms_kernel, ms_span);
}

int su2hip_set_chunk(su2hip_batch_t *b, int kind, int chunk) {
  if (!b) return fail(SU2HIP_EINVAL, "NULL handle");
  if (kind < 0 || kind >= NCOUNT) return fail(SU2HIP_EINVAL, "kind %d is not a pass kind", kind);
  if (chunk != 0 && chunk != 64 && chunk != 128 && chunk != 256)
    return fail(SU2HIP_EINVAL, "chunk %d (0, 64, 128 or 256)", chunk);
  b->tune[kind] = chunk;
  b->graph_valid = false;
  return SU2HIP_OK;
}

int su2hip_get_chunk(su2hip_batch_t *b, int kind, int *chunk) {
  if (!b || !chunk) return fail(SU2HIP_EINVAL, "NULL argument");
  if (kind < 0 || kind >= NCOUNT) return fail(SU2HIP_EINVAL, "kind %d is not a pass kind", kind);
  int rc = prepare(b);
  if (rc) return rc;
  *chunk = chunk_for(b, kind);
  return SU2HIP_OK;
}

// the occupancy targets a pass kind's kernel is built with
static bool waves_valid(int kind, int waves) {
  return waves == 0 || (waves == 5 && kind == SU2HIP_K_LINK_OVERRELAX);
}

int su2hip_set_waves(su2hip_batch_t *b, int kind, int waves) {
  if (!b) return fail(SU2HIP_EINVAL, "NULL handle");
  if (kind < 0 || kind >= NCOUNT) return fail(SU2HIP_EINVAL, "kind %d is not a pass kind", kind);
  if (!waves_valid(kind, waves))
    return fail(SU2HIP_EINVAL, "waves %d for kind %d (0, or 5 for link overrelaxation)", waves, kind);
  b->waves[kind] = waves;
  b->graph_valid = false;
  return SU2HIP_OK;
}

int su2hip_get_waves(su2hip_batch_t *b, int kind, int *waves) {
  if (!b || !waves) return fail(SU2HIP_EINVAL, "NULL argument");
  if (kind < 0 || kind >= NCOUNT) return fail(SU2HIP_EINVAL, "kind %d is not a pass kind", kind);
  *waves = b->waves[kind];
  return SU2HIP_OK;
}

int su2hip_set_hop_cache(su2hip_batch_t *b, int enable) {
  if (!b) return fail(SU2HIP_EINVAL, "NULL handle");
  bool on = enable != 0;
  if (on && b->tmpl.nhiggs != 1)
    return fail(SU2HIP_EINVAL, "the hopping-term cache needs a model with one doublet");
  HIPTRY(hipStreamSynchronize(b->stream));
  int rc = pull_desc(b);  // keep the device-owned pass counters and budgets
  if (rc) return rc;
  b->hop_on = on;
  for (int i = 0; i < b->max_slots && !rc; i++)
    if (b->slots[i].used) rc = on ? hop_alloc(b, i) : (hop_release(b, i), SU2HIP_OK);
  if (rc) {  // off everywhere, as documented
    std::string keep = g_err;
    for (int i = 0; i < b->max_slots; i++)
      if (b->slots[i].used) hop_release(b, i);
    b->hop_on = false;
    b->graph_valid = false;
    (void)push_desc(b);
    g_err = keep;
    return rc;
  }
  b->graph_valid = false;
  return push_desc(b);
}

// kept: on for the handle and present for each of its lattices
int su2hip_get_hop_cache(su2hip_batch_t *b, int *enabled) {
  if (!b || !enabled) return fail(SU2HIP_EINVAL, "NULL argument");
  bool on = b->hop_on;
  for (int i = 0; i < b->max_slots; i++)
    if (b->slots[i].used && !b->slots[i].hop_mem) on = false;
  *enabled = on ? 1 : 0;
  return SU2HIP_OK;
}

int su2hip_autotune(su2hip_batch_t *b, int repeats, double *table) {
  TraceRange tr_("su2hip_autotune");
  if (!b || repeats < 1) return fail(SU2HIP_EINVAL, "NULL handle or repeats < 1");
  int rc = prepare(b);
  if (rc) return rc;
  if (b->nblk == 0) return SU2HIP_OK;
  // the pass kinds of this batch's sweep
  bool in_sweep[NCOUNT] = {};
  for (const Step &st : b->sched) in_sweep[st.kind] = true;
  // snapshot: fields, pass counters, budgets and counters come back as they were
  HIPTRY(hipStreamSynchronize(b->stream));
  rc = fold_counts(b);
  if (rc) return rc;
  rc = pull_desc(b);
  if (rc) return rc;
  std::vector<void *> save(b->max_slots, nullptr);
  std::vector<LatDesc> hd(b->max_slots);
  std::vector<su2hip_counters> cnt(b->max_slots);
  auto release = [&] {
    for (void *q : save)
      if (q) (void)hipFree(q);
  };
  for (int i = 0; i < b->max_slots; i++) {
    Slot &sl = b->slots[i];
    if (!sl.used) continue;
    hd[i] = sl.hd;
    cnt[i] = sl.cnt;
    if (hipMalloc(&save[i], sl.bytes) != hipSuccess) {
      save[i] = nullptr;
      release();
      return fail(SU2HIP_ENOMEM, "autotune snapshot of %zu bytes", sl.bytes);
    }
    hipError_t e = hipMemcpyAsync(save[i], sl.mem, sl.bytes, hipMemcpyDeviceToDevice, b->stream);
    if (e != hipSuccess) {
      release();
      return fail(SU2HIP_EDEVICE, "autotune snapshot: %s", hipGetErrorString(e));
    }
  }
  int saved_tune[SU2HIP_NKINDS], saved_waves[SU2HIP_NKINDS];
  memcpy(saved_tune, b->tune, sizeof(saved_tune));
  memcpy(saved_waves, b->waves, sizeof(saved_waves));
  for (int k = 0; k < NCOUNT && !rc; k++) {
    if (!in_sweep[k]) continue;
    // candidates: every chunk, with each occupancy target the kind has; a
    // candidate's time is weighed by 1.02 unless its chunk is 256 and by
    // 1.01 unless its target is the kind's default, so a smaller chunk has
    // to win by 2% and the other target by 1% to be kept
    const int dflt = k == SU2HIP_K_LINK_OVERRELAX ? 5 : 0;
    double best = 0.0;
    int pick = 256, pick_w = dflt;
    bool first = true;
    for (int c = 0; c < NCHUNK && !rc; c++) {
      double row = 0.0;
      for (int w : {5, 0}) {
        if (!waves_valid(k, w) || rc) continue;
        b->tune[k] = CHUNKS[c];
        b->waves[k] = w;
        double ms = 0.0;
        rc = time_pass_impl(b, k, repeats, &ms, nullptr);
        if (rc) break;
        row = (row == 0.0 || ms < row) ? ms : row;
        double score = ms * (CHUNKS[c] == 256 ? 1.0 : 1.02) * (w == dflt ? 1.0 : 1.01);
        if (first || score < best) {
          best = score;
          pick = CHUNKS[c];
          pick_w = w;
          first = false;
        }
      }
      if (table && !rc) table[k * NCHUNK + c] = row;
    }
    b->tune[k] = rc ? saved_tune[k] : pick;
    b->waves[k] = rc ? saved_waves[k] : pick_w;
  }
  // restore the state
  int rc2 = fold_counts(b);
  for (int i = 0; i < b->max_slots && !rc2; i++) {
    Slot &sl = b->slots[i];
    if (!sl.used) continue;
    if (hipMemcpyAsync(sl.mem, save[i], sl.bytes, hipMemcpyDeviceToDevice, b->stream) != hipSuccess)
      rc2 = fail(SU2HIP_EDEVICE, "autotune restore");
    sl.hd.pass = hd[i].pass;
    sl.hd.budget = hd[i].budget;
    sl.cnt = cnt[i];
  }
  if (!rc2) rc2 = push_desc(b);
  if (!rc2 && hipStreamSynchronize(b->stream) != hipSuccess) rc2 = fail(SU2HIP_EDEVICE, "autotune sync");
  release();
  rc2 = rc2 ? rc2 : orderparam_refresh(b);
  b->graph_valid = false;
  return rc ? rc : rc2;
}

int su2hip_copy_probe(uint64_t bytes, int repeats, double *ms) {
  if (!ms || bytes < 16 || repeats < 1) return fail(SU2HIP_EINVAL, "copy probe arguments");
  size_t n = bytes / sizeof(double);
  double *a = nullptr, *c = nullptr;
  hipStream_t s = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  // every resource is released on every path: the first failure is kept in e
  hipError_t e = hipMalloc(&a, n * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&c, n * sizeof(double));
  if (e != hipSuccess) {
    if (a) (void)hipFree(a);
    return fail(SU2HIP_ENOMEM, "probe buffers of %zu bytes", n * sizeof(double));
  }
  const char *what = "stream";
  float t = 0.f;
  e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
  if (e == hipSuccess) { what = "events"; e = hipEventCreate(&e0); }
  if (e == hipSuccess) e = hipEventCreate(&e1);
  if (e == hipSuccess) { what = "fill"; e = hipMemsetAsync(a, 0, n * sizeof(double), s); }
  if (e == hipSuccess) { what = "copy kernel"; e = launch_copy_probe(a, c, n, s); }
  if (e == hipSuccess) e = hipEventRecord(e0, s);
  for (int r = 0; r < repeats && e == hipSuccess; r++) e = launch_copy_probe(a, c, n, s);
  if (e == hipSuccess) e = hipEventRecord(e1, s);
  if (e == hipSuccess) { what = "wait"; e = hipEventSynchronize(e1); }
  if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
  if (s) (void)hipStreamSynchronize(s);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (s) (void)hipStreamDestroy(s);
  (void)hipFree(a);
  (void)hipFree(c);
  if (e != hipSuccess) return fail(SU2HIP_EDEVICE, "copy probe %s: %s", what, hipGetErrorString(e));
  *ms = t / repeats;
  return SU2HIP_OK;
}

}  // extern "C"
