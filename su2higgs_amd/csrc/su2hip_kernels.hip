// su2hip_kernels.hip -- HIP kernels of the checkerboard update sweep for
// gfx950 (MI355X), and their launchers.
//
// One thread updates one site of one parity; a workgroup of BLK threads owns
// BLK consecutive sites (in the device site order) of one lattice, and the
// block table maps workgroups to (lattice, first site) for every lattice of
// the batch, in an XCD-aware order built by the host.  Every per-site
// routine names the reference function it restates (file:line under the
// reference's src/).  The SU(2) algebra is quaternion products; nothing here
// is a dense contraction, so MFMA is not used and the kernels are bound by
// HBM traffic and FP math.
#include <hip/hip_ext.h>
#include <type_traits>

#include "su2hip_device.h"

namespace su2 {

// ------------------------------------------------------------------------
// math wrappers (float / double overloads)

template <typename T> __device__ __forceinline__ T msqrt(T x) { if constexpr (sizeof(T) == 8) return ::sqrt(x); else return ::sqrtf(x); }
template <typename T> __device__ __forceinline__ T mlog(T x) { if constexpr (sizeof(T) == 8) return ::log(x); else return ::logf(x); }
template <typename T> __device__ __forceinline__ T mexp(T x) { if constexpr (sizeof(T) == 8) return ::exp(x); else return ::expf(x); }
template <typename T> __device__ __forceinline__ T mcos(T x) { if constexpr (sizeof(T) == 8) return ::cos(x); else return ::cosf(x); }
template <typename T> __device__ __forceinline__ T msin(T x) { if constexpr (sizeof(T) == 8) return ::sin(x); else return ::sinf(x); }

// Transcendentals of the sweep updates.  FP64 keeps the accurate library
// functions above (the reference's arithmetic, compared with it to 1e-12).
// FP32 uses the hardware instructions -- v_log_f32, v_exp_f32, v_sqrt_f32,
// v_rsq_f32, v_rcp_f32, and v_sin_f32/v_cos_f32 on an argument reduced to
// [-1/2, 1/2] revolutions -- whose errors of a few ulp sit far below what
// the FP32 mode is checked to (observables within statistical errors).
constexpr float LN2F = 0.693147180559945309f;
constexpr float LOG2EF = 1.442695040888963407f;
constexpr float INV2PIF = 0.159154943091895336f;
__device__ __forceinline__ float flog(float x) { return __builtin_amdgcn_logf(x) * LN2F; }
__device__ __forceinline__ float fexp(float x) { return __builtin_amdgcn_exp2f(x * LOG2EF); }
__device__ __forceinline__ float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ float frsqrt(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }
// cos and sin of 2 pi t
__device__ __forceinline__ float fcos2pi(float t) { return __builtin_amdgcn_cosf(t - rintf(t)); }
__device__ __forceinline__ float fsin2pi(float t) { return __builtin_amdgcn_sinf(t - rintf(t)); }
__device__ __forceinline__ float fcos(float x) { return fcos2pi(x * INV2PIF); }
__device__ __forceinline__ float fsin(float x) { return fsin2pi(x * INV2PIF); }
// exp for an accept test: accurate in FP64, the hardware exp in FP32
template <typename T> __device__ __forceinline__ T sexp(T x) { if constexpr (sizeof(T) == 8) return ::exp(x); else return fexp(x); }

constexpr double PI = 3.14159265358979323846;

template <int M, int KIND>
constexpr bool kind_valid() {
  if (KIND == SU2HIP_K_LINK_HEATBATH || KIND == SU2HIP_K_LINK_METROPOLIS ||
      KIND == SU2HIP_K_LINK_OVERRELAX)
    return true;
  if (KIND == SU2HIP_K_U1_METROPOLIS) return (M & M_U1) != 0;
  if (KIND == SU2HIP_K_DOUBLET_OVERRELAX || KIND == SU2HIP_K_DOUBLET_METROPOLIS)
    return (M & M_NH) != 0;
  if (KIND == SU2HIP_K_TRIPLET_OVERRELAX || KIND == SU2HIP_K_TRIPLET_METROPOLIS)
    return (M & M_TR) != 0;
  if (KIND == SU2HIP_K_SINGLET_OVERRELAX || KIND == SU2HIP_K_SINGLET_METROPOLIS)
    return (M & M_SG) != 0;
  return false;
}

// ------------------------------------------------------------------------
// local pieces of the action

// su2link_staple (src/staples.c:134-193) with su2staple_wilson (:84-115):
// -beta/2 sum_nu [U_nu(x+mu) U_mu(x+nu)^+ U_nu(x)^+ + U_nu(x+mu-nu)^+
// U_mu(x-nu)^+ U_nu(x-nu)] - 1/2 Phi(x+mu) exp(-i alpha sigma_3) Phi(x)^+
template <typename T, int M>
__device__ __forceinline__ Q<T> wilson_staple(const LatDesc &L, const T *U, const Site &x, int mu) {
  Q<T> W{(T)0, (T)0, (T)0, (T)0};
  Site xp = shift(L, x, mu, +1);
#pragma unroll
  for (int nu = 0; nu < 3; nu++) {
    if (nu == mu) continue;
    Site xn = shift(L, x, nu, +1), xm = shift(L, x, nu, -1), xpm = shift(L, xp, nu, -1);
    Q<T> up = qmul(qmul(U_at<M>(L, U, xp, nu), qconj(U_at<M>(L, U, xn, mu))), qconj(U_at<M>(L, U, x, nu)));
    W = qadd(W, up);
    Q<T> lo = qmul(qmul(qconj(U_at<M>(L, U, xpm, nu)), qconj(U_at<M>(L, U, xm, mu))), U_at<M>(L, U, xm, nu));
    W = qadd(W, lo);
  }
  return W;
}

// With one doublet the hopping term of the staple is read from the cache
// k_hop fills (the same arithmetic, done once per round of link passes),
// when the handle keeps one (LatDesc::H not null).
template <int M>
constexpr bool hop_cached() {
  return (M & M_NH) != 0 && (M & M_H2) == 0;
}

// -1/2 Phi(x+mu) exp(-i alpha sigma_3) Phi(x)^+ of link (x, mu)
template <typename T, int M>
__device__ __forceinline__ Q<T> hop_term(const LatDesc &L, const T *A, const T *F, const Site &x,
                                         int mu) {
  Q<T> b = F_at(L, F, shift(L, x, mu, +1));
  if constexpr ((M & M_U1) != 0) b = qmul_u1(b, u1twist<T>(A_at(L, A, x, mu)));
  Q<T> t = qmul(b, qconj(F_at(L, F, x)));
  return qscale(t, (T)-0.5);
}

template <typename T, int M>
__device__ __forceinline__ Q<T> link_staple(const LatDesc &L, const T *U, const T *A, const T *F,
                                            const Site &x, int mu) {
  Q<T> V = qscale(wilson_staple<T, M>(L, U, x, mu), (T)(-0.5 * L.beta));
  if constexpr (hop_cached<M>())
    if (L.H) return qadd(V, ld4((const T *)L.H + ((size_t)(mu * 2 + x.par) * L.Vh + x.s) * 4));
  if constexpr ((M & M_NH) != 0) V = qadd(V, hop_term<T, M>(L, A, F, x, mu));
  // the second doublet's hopping term (the doublet loop of
  // src/staples.c:150-190)
  if constexpr ((M & M_H2) != 0) V = qadd(V, hop_term<T, M>(L, A, (const T *)L.F2, x, mu));
  return V;
}

// staple_doublet (src/staples.c:243-302)
template <typename T, int M>
__device__ __forceinline__ void doublet_hops(const LatDesc &L, const T *U, const T *A, const T *F,
                                             const Site &x, int d, Q<T> &s) {
  Site xp = shift(L, x, d, +1), xm = shift(L, x, d, -1);
  Q<T> t = qmul(U_at<M>(L, U, x, d), F_at(L, F, xp));
  if constexpr ((M & M_U1) != 0) t = qmul_u1(t, u1twist<T>(A_at(L, A, x, d)));
  s = qadd(s, qscale(t, (T)-1));
  Q<T> t2 = qmul(qconj(U_at<M>(L, U, xm, d)), F_at(L, F, xm));
  if constexpr ((M & M_U1) != 0) t2 = qmul_u1(t2, qconj(u1twist<T>(A_at(L, A, xm, d))));
  s = qadd(s, qscale(t2, (T)-1));
}

template <typename T, int M>
__device__ __forceinline__ Q<T> doublet_staple(const LatDesc &L, const T *U, const T *A, const T *F,
                                               const Site &x) {
  Q<T> s{(T)0, (T)0, (T)0, (T)0};
  if constexpr ((M & M_U1) != 0 && sizeof(T) == 8) {
    // one direction at a time: six inlined FP64 sincos side by side would
    // hold 170 VGPRs (2 waves per SIMD)
#pragma unroll 1
    for (int d = 0; d < 3; d++) doublet_hops<T, M>(L, U, A, F, x, d, s);
    return s;
  }
#pragma unroll
  for (int d = 0; d < 3; d++) {
    Site xp = shift(L, x, d, +1), xm = shift(L, x, d, -1);
    Q<T> t = qmul(U_at<M>(L, U, x, d), F_at(L, F, xp));
    if constexpr ((M & M_U1) != 0) t = qmul_u1(t, u1twist<T>(A_at(L, A, x, d)));
    s = qadd(s, qscale(t, (T)-1));
    Q<T> t2 = qmul(qconj(U_at<M>(L, U, xm, d)), F_at(L, F, xm));
    if constexpr ((M & M_U1) != 0) t2 = qmul_u1(t2, qconj(u1twist<T>(A_at(L, A, xm, d))));
    s = qadd(s, qscale(t2, (T)-1));
  }
  return s;
}

// hopping staple of the triplet (src/overrelax.c:143-180)
template <typename T, int M>
__device__ __forceinline__ V3<T> triplet_staple(const LatDesc &L, const T *U, const T *G, const Site &x) {
  V3<T> s{(T)0, (T)0, (T)0};
#pragma unroll
  for (int d = 0; d < 3; d++) {
    Site xp = shift(L, x, d, +1), xm = shift(L, x, d, -1);
    V3<T> r = rotate(U_at<M>(L, U, x, d), G_at(L, G, xp));
    s.x -= r.x;
    s.y -= r.y;
    s.z -= r.z;
    r = rotate(qconj(U_at<M>(L, U, xm, d)), G_at(L, G, xm));
    s.x -= r.x;
    s.y -= r.y;
    s.z -= r.z;
  }
  return s;
}

// hopping_triplet_forward (src/su2u1.c:605-617) with link u
template <typename T>
__device__ __forceinline__ T hop_triplet(const Q<T> &u, const V3<T> &a, const V3<T> &b) {
  V3<T> r = rotate(u, b);
  return -(a.x * r.x + a.y * r.y + a.z * r.z);
}

// polysolve3 (src/overrelax.c:14-34): real root by Cardano's formula for a
// negative discriminant, else b/a as the reference returns
__device__ __forceinline__ double polysolve3(double a, double b, double c, double d) {
  double disc = 18.0 * a * b * c * d - 4.0 * b * b * b * d + b * b * c * c -
                4.0 * a * c * c * c - 27.0 * a * a * d * d;
  double d1 = 2.0 * b * b * b - 9.0 * a * b * c + 27.0 * a * a * d;
  if (disc < 0.0) {
    double A = 0.5 * (d1 + a * ::sqrt(-27.0 * disc));
    A = ::cbrt(A);
    return -1.0 * (b + A + (b * b - 3.0 * a * c) / A) / (3.0 * a);
  }
  return b / a;
}

// Cartesian (X,Y) overrelaxation of src/overrelax.c:54-128 given the staple
// norm F, X, |Y|^2 and the local B, C: returns the accept probability and
// the new X in nX.
__device__ __forceinline__ double xy_or(double F, double X, double Ysq, double B, double C,
                                        double &nX) {
  double al = C;
  double be = C * X;
  double cc = B + 2.0 * Ysq * C;
  double ga = (cc + C * X * X);
  double de = (F + cc * X + C * X * X * X);
  nX = polysolve3(al, be, ga, de);
  double dV = F + 2.0 * B * X + 4.0 * C * (X * X * X + Ysq * X);
  double dVn = F + 2.0 * B * nX + 4.0 * C * (nX * nX * nX + Ysq * nX);
  return ::fabs(dV / dVn);
}

// FP32 version of xy_or: the same move, with the cubic's real root found by
// Newton's method from the harmonic reflection -X - F/cc (the root for
// C = 0).  With cc > 0 the cubic is increasing, so it has one real root
// and Newton converges; otherwise the FP64 Cardano path decides, with the
// reference's fallback.
__device__ __forceinline__ float xy_or_f(float F, float X, float Ysq, float B, float C, float &nX) {
  float cc = B + 2.f * Ysq * C;
  float al = C, be = C * X, ga = cc + C * X * X, de = F + cc * X + C * X * X * X;
  float x;
  if (cc > 0.f) {
    x = -X - F * frcp(cc);
#pragma unroll 1
    for (int it = 0; it < 8; it++) {
      float g = ((al * x + be) * x + ga) * x + de;
      float dg = (3.f * al * x + 2.f * be) * x + ga;
      float dx = g * frcp(dg);
      x -= dx;
      if (fabsf(dx) <= 2.5e-7f * (1.f + fabsf(x))) break;
    }
  } else {
    x = (float)polysolve3(al, be, ga, de);
  }
  nX = x;
  float dV = F + 2.f * B * X + 4.f * C * (X * X * X + Ysq * X);
  float dVn = F + 2.f * B * x + 4.f * C * (x * x * x + Ysq * x);
  return fabsf(dV * frcp(dVn));
}

// accept rule of every Metropolis update (src/metropolis.c:31-40): a draw
// is taken only when the action grows
template <typename T>
__device__ __forceinline__ int metro_accept(T diff, Rng &r) {
  if (diff < (T)0) return 1;
  if (diff > (T)0 && sexp(-diff) > uni<T>(r)) return 1;
  return 0;
}

template <typename T>
__device__ __forceinline__ Q<T> normalise(const Q<T> &q) {
  T n;
  if constexpr (sizeof(T) == 8) n = 1.0 / ::sqrt(qdot(q, q));
  else n = frsqrt(qdot(q, q));
  return qscale(q, n);
}

template <typename T>
__device__ __forceinline__ Q<double> qd(const Q<T> &q) {
  return Q<double>{(double)q.a, (double)q.b, (double)q.c, (double)q.d};
}

// higgspotential of two doublets (src/su2u1.c:528-554), f1 = phi_1 and
// f2 = phi_2 at one site
__device__ __forceinline__ double pot2(const LatDesc &L, const Q<double> &h1, const Q<double> &h2) {
  double f11 = 0.5 * qdot(h1, h1), f22 = 0.5 * qdot(h2, h2);
  // get_phi12 (src/su2u1.c:513-520)
  double R = 0.5 * (h1.a * h2.a + h1.b * h2.b + h1.c * h2.c + h1.d * h2.d);
  double I = 0.5 * (h1.d * h2.a + h1.c * h2.b - h1.b * h2.c - h1.a * h2.d);
  return L.msq * f11 + L.lambda * f11 * f11 + L.msq2 * f22 + L.m12re * R - L.m12im * I +
         L.lam2 * f22 * f22 + L.lam3 * f11 * f22 + L.lam4 * (R * R + I * I) +
         L.lam5re * (R * R - I * I) - 2.0 * L.lam5im * R * I + f11 * (L.lam6re * R - L.lam6im * I) +
         f22 * (L.lam7re * R + L.lam7im * I);
}

// overrelax_higgs2's move (src/overrelax.c:295-417) for doublet h of two:
// f the updated doublet, H the other one, st its hopping staple; returns the
// unit direction e, X = f.e, the new X (nX) and the accept probability p.
// The potential along e with Y = f - X e fixed is the quartic
// b4 X^4 + b3 X^3 + b2 X^2 + b1 X, whose other root of V(X') = V(X) is the
// real root of a cubic (polysolve3).
__device__ __forceinline__ void higgs2_or(const LatDesc &L, int h, const Q<double> &f,
                                          const Q<double> &H, const Q<double> &st, Q<double> &e,
                                          double &X, double &nX, double &p) {
  double sg = h ? 1.0 : -1.0;
  Q<double> G{sg * H.d, sg * H.c, -sg * H.b, -sg * H.a};
  double lam12, msq, l67re, l67im, o67re, o67im;
  if (h == 0) {
    lam12 = L.lambda;
    msq = L.msq;
    l67re = L.lam7re;
    l67im = L.lam7im;
    o67re = L.lam6re;
    o67im = -L.lam6im;
  } else {
    lam12 = L.lam2;
    msq = L.msq2;
    l67re = L.lam6re;
    l67im = -L.lam6im;
    o67re = L.lam7re;
    o67im = L.lam7im;
  }
  double Hsq = 0.5 * qdot(H, H);
  double cH = 0.5 * (L.m12re + Hsq * l67re), cG = 0.5 * (-1.0 * L.m12im + Hsq * l67im);
  Q<double> s{-(st.a + cH * H.a + cG * G.a), -(st.b + cH * H.b + cG * G.b),
              -(st.c + cH * H.c + cG * G.c), -(st.d + cH * H.d + cG * G.d)};
  double F = ::sqrt(qdot(s, s));
  e = Q<double>{s.a / F, s.b / F, s.c / F, s.d / F};
  X = qdot(f, e);
  Q<double> Y{f.a - X * e.a, f.b - X * e.b, f.c - X * e.c, f.d - X * e.d};
  double Ysq = qdot(Y, Y);
  double Hf = qdot(H, e), Gf = qdot(G, e), HY = qdot(H, Y), GY = qdot(G, Y);
  double b4 = 0.25 * lam12;
  double b3 = 0.25 * (Hf * o67re + Gf * o67im);
  double b2 = 0.5 * msq + 1.0 * 3 + 0.5 * Ysq * lam12 + 0.25 * (HY * o67re + GY * o67im);
  double b1 = -1.0 * F + 0.25 * Ysq * (Hf * o67re + Gf * o67im);
  b2 += 0.5 * Hsq * L.lam3 + 0.25 * L.lam4 * (Hf * Hf + Gf * Gf) +
        0.25 * L.lam5re * (Hf * Hf - Gf * Gf) - 0.5 * L.lam5im * Gf * Hf;
  b1 += 0.5 * Hf * (HY * L.lam4 + HY * L.lam5re - GY * L.lam5im) +
        0.5 * Gf * (GY * L.lam4 - GY * L.lam5re - HY * L.lam5im);
  nX = polysolve3(b4, b3 + b4 * X, b2 + b3 * X + b4 * X * X, b1 + b2 * X + b3 * X * X + b4 * X * X * X);
  double dV = 4.0 * b4 * X * X * X + 3.0 * b3 * X * X + 2.0 * b2 * X + b1;
  double dVn = 4.0 * b4 * nX * nX * nX + 3.0 * b3 * nX * nX + 2.0 * b2 * nX + b1;
  p = ::fabs(dV / dVn);
}

// ------------------------------------------------------------------------
// A rejection loop shared by the lanes of a wavefront (the direction of the
// FP64 link heatbath, accepted with probability pi/6 per try).
//
// A lane's rejection loop runs until one of its tries accepts, and a
// wavefront runs until its slowest lane is done (about 7 tries).  Here, in
// each round, the 64 lanes make tries for the lanes that are still pending:
// pending lane of rank q (among the pending ones, P of them) gets the
// contiguous run of workers w with (w P) >> 6 == q, worker w making its try
// number tries_q + (w - lo(q)), lo(q) = ceil(64 q / P).  A pending lane
// takes the first accepted try of its run.  Draws are addressed by (pass,
// site, index), so a worker can make another site's try; every lane ends
// with the draws, the accepted try and the stream position of its own
// sequential loop, in about 3 rounds.  Callers use this only when all 64
// lanes are active.  (The Kennedy-Pendleton loop before it accepts most
// tries at once; sharing its rounds cost more than the waiting it saves.)

// draws n .. n+2: always two Philox blocks (words 0-2 of n>>1 and n>>1 + 1,
// or words 1-3 for odd n)
__device__ __forceinline__ void rng_u53_run3(const Rng &r, uint32_t site, uint32_t n, double u[3]) {
  uint32_t a0 = n >> 1, a1 = site, a2 = r.plo, a3 = r.phi;
  philox10(a0, a1, a2, a3, r.k0, r.k1);
  uint32_t b0 = (n >> 1) + 1, b1 = site, b2 = r.plo, b3 = r.phi;
  philox10(b0, b1, b2, b3, r.k0, r.k1);
  uint64_t w0 = ((uint64_t)a1 << 32) | a0, w1 = ((uint64_t)a3 << 32) | a2;
  uint64_t w2 = ((uint64_t)b1 << 32) | b0, w3 = ((uint64_t)b3 << 32) | b2;
  bool odd = n & 1;
  u[0] = (double)((odd ? w1 : w0) >> 11) * 0x1.0p-53;
  u[1] = (double)((odd ? w2 : w1) >> 11) * 0x1.0p-53;
  u[2] = (double)((odd ? w3 : w2) >> 11) * 0x1.0p-53;
}

// lane of the q-th (from 0) set bit of m; q < popcount(m)
__device__ __forceinline__ int kth_set_bit(uint64_t m, int q) {
  int pos = 0;
#pragma unroll
  for (int w = 32; w >= 1; w >>= 1) {
    int c = __popcll(m & ((1ull << w) - 1));
    if (q >= c) {
      q -= c;
      m >>= w;
      pos += w;
    }
  }
  return pos;
}

// this lane's place in a round: its target lane and try offset as a worker,
// its run [lo, hi) of workers as a pending target
struct CoopRound {
  int target, offset;
  uint64_t run;
};
__device__ __forceinline__ CoopRound coop_round(uint64_t pending, int lane) {
  const int P = __popcll(pending);
  CoopRound c;
  int q = (lane * P) >> 6;
  c.target = kth_set_bit(pending, q);
  c.offset = lane - (q * 64 + P - 1) / P;
  int mq = __popcll(pending & ((1ull << lane) - 1));
  int lo = (mq * 64 + P - 1) / P, hi = ((mq + 1) * 64 + P - 1) / P;
  uint64_t below_hi = hi >= 64 ? ~0ull : (1ull << hi) - 1;
  c.run = below_hi & ~((1ull << lo) - 1);
  return c;
}

// The direction's cube rejection of heatbath_su2link (src/heatbath.c:82-93)
// for the lanes with `want`, starting at draw r.k: the accepted try's
// r1, r2, r3, d and the number of tries.
__device__ void hb_dir_wave(const Rng &r, bool want, double &r1o, double &r2o, double &r3o,
                            double &dout, int &tries_out) {
  const int lane = __lane_id();
  bool pending = want;
  int tries = 0;
  for (;;) {
    uint64_t pm = __ballot(pending);
    if (pm == 0) break;
    const bool all = pm == ~0ull;  // every lane on its own try
    CoopRound c;
    int t;
    uint32_t tsite, tn;
    if (all) {
      c = CoopRound{lane, 0, 1ull << lane};
      t = tries;
      tsite = r.site;
      tn = r.k;
    } else {
      c = coop_round(pm, lane);
      t = __shfl(tries, c.target, 64) + c.offset;
      tsite = __shfl(r.site, c.target, 64);
      tn = __shfl(r.k, c.target, 64);
    }
    double u[3];
    rng_u53_run3(r, tsite, tn + 3u * (uint32_t)t, u);
    double r1 = 2.0 * u[0] - 1.0;
    double r2 = 2.0 * u[1] - 1.0;
    double r3 = 2.0 * u[2] - 1.0;
    double d = r1 * r1 + r2 * r2 + r3 * r3;
    bool acc = !(d > 1.0 || d == 0.0);
    uint64_t mine = __ballot(acc) & c.run;
    int win = mine ? __ffsll((unsigned long long)mine) - 1 : lane;
    double w1 = r1, w2 = r2, w3 = r3, wd = d;
    if (!all) {
      w1 = __shfl(r1, win, 64);
      w2 = __shfl(r2, win, 64);
      w3 = __shfl(r3, win, 64);
      wd = __shfl(d, win, 64);
    }
    if (pending) {
      int lo = __ffsll((unsigned long long)c.run) - 1;
      if (mine) {
        r1o = w1;
        r2o = w2;
        r3o = w3;
        dout = wd;
        tries += win - lo + 1;
        pending = false;
      } else {
        tries += __popcll(c.run);
      }
    }
  }
  tries_out = tries;
}

// ------------------------------------------------------------------------
// per-site updates; each returns 1 if accepted and stores the new value

struct Fields {
  void *U, *A, *F, *G, *S, *B;
};

template <typename T, int M, int KIND>
__device__ __forceinline__ int site_update(const LatDesc &L, const Site &x, int par, int s, int mu,
                                           Rng &r, int reunit) {
  T *U = (T *)L.U;
  T *A = (T *)L.A;
  T *F = (T *)L.F;
  T *G = (T *)L.G;
  T *S = (T *)L.S;
  const int Vh = L.Vh;

  if constexpr (KIND == SU2HIP_K_LINK_HEATBATH) {
    // heatbath_su2link (src/heatbath.c:15-124), Kennedy-Pendleton
    Q<T> old{};
    T oldact = 0;
    V3<T> a1{}, a2{};
    if constexpr ((M & M_TR) != 0) {
      old = ldU<M>(U, Vh, mu, par, s);
      a1 = ldG(G, Vh, par, s);
      a2 = G_at(L, G, shift(L, x, mu, +1));
      oldact = hop_triplet(old, a1, a2);
    }
    Q<T> V = link_staple<T, M>(L, U, A, F, x, mu);
    T r1, r2, r3, d, a0;
    int loop = 1;
    // a full wavefront shares the FP64 direction's rejection tries
    // (hb_dir_wave); the results are those of the per-lane loop
    const bool wave = sizeof(T) == 8 && __ballot(1) == ~0ull;
    if constexpr (sizeof(T) == 8) {
      T a = ::sqrt(qdot(V, V));
      V = qscale(V, (T)-1 / a);
      a *= (T)2;
      do {
        r1 = -::log((T)1 - uni<T>(r)) / a;
        r2 = -::log((T)1 - uni<T>(r)) / a;
        r3 = ::cospi((T)2 * uni<T>(r));
        r3 *= r3;
        d = r1 * r3 + r2;
        r3 = (T)1 - uni<T>(r);
        loop++;
      } while (r3 * r3 > (T)1 - (T)0.5 * d && loop <= 200);
    } else {
      // the same Kennedy-Pendleton loop on the hardware log/cos
      T ia = frsqrt(qdot(V, V));
      V = qscale(V, -ia);
      ia *= (T)0.5;  // 1 / (2 |V|)
      do {
        r1 = -flog((T)1 - uni<T>(r)) * ia;
        r2 = -flog((T)1 - uni<T>(r)) * ia;
        r3 = fcos2pi(uni<T>(r));
        r3 *= r3;
        d = r1 * r3 + r2;
        r3 = (T)1 - uni<T>(r);
        loop++;
      } while (r3 * r3 > (T)1 - (T)0.5 * d && loop <= 200);
    }
    a0 = (loop > 200) ? (T)1e-9 : (T)1 - d;
    T rad = (T)1 - a0 * a0;
    Q<T> nu;
    if constexpr (sizeof(T) == 8) {
      // isotropic direction by cube rejection, the reference's draw order
      // (src/heatbath.c:82-93)
      if (wave) {
        // every lane stays in the wavefront's rounds, as a worker at least
        int tries;
        hb_dir_wave(r, !(rad < (T)0), r1, r2, r3, d, tries);
        if (rad < (T)0) return 0;
        r.k += 3u * (uint32_t)tries;
        r.blk = 0xFFFFFFFFu;
      } else {
        if (rad < (T)0) return 0;
        d = (T)2;
        while (d > (T)1 || d == (T)0) {
          r1 = (T)2 * uni<T>(r) - (T)1;
          r2 = (T)2 * uni<T>(r) - (T)1;
          r3 = (T)2 * uni<T>(r) - (T)1;
          d = r1 * r1 + r2 * r2 + r3 * r3;
        }
      }
      rad = ::sqrt(rad);
      d = ::sqrt(d);
      nu = Q<T>{a0, r1 * (rad / d), r2 * (rad / d), r3 * (rad / d)};
    } else {
      if (rad < (T)0) return 0;
      // FP32: the same uniform direction on S^2 without rejection
      // (cos(theta) uniform in [-1,1], phi uniform), so a wavefront does
      // not wait for its slowest lane
      rad = fsqrt(rad);
      T cz = (T)2 * uni<T>(r) - (T)1;
      T ph = uni<T>(r);
      T sn = fsin2pi(ph), cs = fcos2pi(ph);
      T sr = fsqrt(fmaxf((T)1 - cz * cz, (T)0)) * rad;
      nu = Q<T>{a0, sr * cs, sr * sn, rad * cz};
    }
    nu = qmul(nu, qconj(V));
    if constexpr ((M & M_TR) != 0) {
      T diff = oldact - hop_triplet(nu, a1, a2);
      if (!(uni<T>(r) < sexp(diff))) return 0;
    }
    stU<M>(U, Vh, mu, par, s, nu);
    return 1;
  } else if constexpr (KIND == SU2HIP_K_LINK_METROPOLIS) {
    // metro_su2link (src/metropolis.c:15-41), random_su2link (src/init.c:43-58);
    // Delta S = 2 Re[(U'-U) V]_0 + Delta(triplet hop), from the staple
    Q<T> old = ldU<M>(U, Vh, mu, par, s);
    Q<T> V = link_staple<T, M>(L, U, A, F, x, mu);
    Q<T> R;
    R.a = (T)10 * (uni<T>(r) - (T)0.5);
    R.b = (T)1 * (uni<T>(r) - (T)0.5);
    R.c = (T)1 * (uni<T>(r) - (T)0.5);
    R.d = (T)1 * (uni<T>(r) - (T)0.5);
    if constexpr (sizeof(T) == 8) {
      T n = ::sqrt(qdot(R, R));
      R = Q<T>{R.a / n, R.b / n, R.c / n, R.d / n};
    } else {
      R = qscale(R, frsqrt(qdot(R, R)));
    }
    Q<T> nu = qmul(old, R);
    (void)reunit;  // FP32 links are renormalised by their packed storage
    Q<T> du{nu.a - old.a, nu.b - old.b, nu.c - old.c, nu.d - old.d};
    T diff = (T)2 * qmul(du, V).a;
    if constexpr ((M & M_TR) != 0) {
      V3<T> a1 = ldG(G, Vh, par, s), a2 = G_at(L, G, shift(L, x, mu, +1));
      diff += hop_triplet(nu, a1, a2) - hop_triplet(old, a1, a2);
    }
    if (!metro_accept(diff, r)) return 0;
    stU<M>(U, Vh, mu, par, s, nu);
    return 1;
  } else if constexpr (KIND == SU2HIP_K_LINK_OVERRELAX) {
    // SU(2) link overrelaxation, U' = W^+ U^+ W^+ with W = staple/|staple|
    // (no reference counterpart); with a triplet a Metropolis test on its
    // hopping term as in src/heatbath.c:107-120
    Q<T> old = ldU<M>(U, Vh, mu, par, s);
    Q<T> V = link_staple<T, M>(L, U, A, F, x, mu);
    Q<T> w;
    if constexpr (sizeof(T) == 8) {
      T a = ::sqrt(qdot(V, V));
      if (a == (T)0) return 1;
      w = qconj(Q<T>{V.a / a, V.b / a, V.c / a, V.d / a});
    } else {
      T n2 = qdot(V, V);
      if (n2 == (T)0) return 1;
      w = qconj(qscale(V, frsqrt(n2)));
    }
    Q<T> nu = qmul(qmul(w, qconj(old)), w);
    (void)reunit;  // FP32 links are renormalised by their packed storage
    if constexpr ((M & M_TR) != 0) {
      V3<T> a1 = ldG(G, Vh, par, s), a2 = G_at(L, G, shift(L, x, mu, +1));
      T diff = hop_triplet(old, a1, a2) - hop_triplet(nu, a1, a2);
      if (!(uni<T>(r) < sexp(diff))) return 0;
    }
    stU<M>(U, Vh, mu, par, s, nu);
    return 1;
  } else if constexpr (KIND == SU2HIP_K_U1_METROPOLIS) {
    // metro_u1link (src/metropolis.c:48-70) with localact_u1link
    // (src/su2u1.c:336-354) evaluated as a difference
    T al = ldA(A, Vh, mu, par, s);
    T na = al + (T)0.95 * (uni<T>(r) - (T)0.5);
    if constexpr (sizeof(T) == 4) {
      // FP32: keep the angle in [-pi, pi] so it keeps its precision; the
      // action is 2 pi periodic in it for an integer charge r_u1
      if (L.r_u1 == ::rint(L.r_u1)) na -= (T)(2.0 * PI) * rintf(na * (T)(0.5 / PI));
    }
    T rr = (T)L.r_u1;
    Site xp = shift(L, x, mu, +1);
    T dS = 0;
    if constexpr (sizeof(T) == 8) {
      // the plaquette differences cos A - cos B of localact_u1link as
      // -2 sin((A+B)/2) sin((A-B)/2), with m = (al+na)/2 and h = (na-al)/2
      // (5 sines where the differences take 8 cosines, and no cancellation);
      // the hopping term -(q0 cos a + q3 sin a), q = Phi(x)^+ U_mu(x)
      // Phi(x+mu), likewise
      T m = (T)0.5 * (al + na), h = (T)0.5 * (na - al);
      T sum = 0;
#pragma unroll
      for (int nu = 0; nu < 3; nu++) {
        if (nu == mu) continue;
        Site xn = shift(L, x, nu, +1), xm = shift(L, x, nu, -1), xpm = shift(L, xp, nu, -1);
        T rest_up = A_at(L, A, xp, nu) - A_at(L, A, xn, mu) - A_at(L, A, x, nu);
        T rest_dn = A_at(L, A, xm, mu) + A_at(L, A, xpm, nu) - A_at(L, A, xm, nu);
        sum += ::sin(rr * (m + rest_up)) - ::sin(rr * (rest_dn - m));
      }
      dS = (T)2 * (T)L.betau1 * ::sin(rr * h) * sum;
      if constexpr ((M & M_NH) != 0) {
        Q<T> q = qmul(qmul(qconj(ldF(F, Vh, par, s)), ldU<M>(U, Vh, mu, par, s)), F_at(L, F, xp));
        if constexpr ((M & M_H2) != 0) {
          const T *F2 = (const T *)L.F2;
          q = qadd(q, qmul(qmul(qconj(ldF(F2, Vh, par, s)), ldU<M>(U, Vh, mu, par, s)), F_at(L, F2, xp)));
        }
        T sm, cm;
        ::sincos(m, &sm, &cm);
        dS -= (T)2 * ::sin(h) * (q.d * cm - q.a * sm);
      }
    } else {
      // the same difference through cos A - cos B = -2 sin((A+B)/2) sin((A-B)/2)
      // with m = (al+na)/2, h = (na-al)/2: one sine of the small step h, the
      // rest of the accuracy of the hardware sine; exactly antisymmetric
      // under exchanging al and na
      T m = (T)0.5 * (al + na), h = (T)0.5 * (na - al);
      T sum = 0;
#pragma unroll
      for (int nu = 0; nu < 3; nu++) {
        if (nu == mu) continue;
        Site xn = shift(L, x, nu, +1), xm = shift(L, x, nu, -1), xpm = shift(L, xp, nu, -1);
        T rest_up = A_at(L, A, xp, nu) - A_at(L, A, xn, mu) - A_at(L, A, x, nu);
        T rest_dn = A_at(L, A, xm, mu) + A_at(L, A, xpm, nu) - A_at(L, A, xm, nu);
        sum += fsin(rr * (m + rest_up)) - fsin(rr * (rest_dn - m));
      }
      dS = (T)2 * (T)L.betau1 * fsin(rr * h) * sum;
      if constexpr ((M & M_NH) != 0) {
        Q<T> q = qmul(qmul(qconj(ldF(F, Vh, par, s)), ldU<M>(U, Vh, mu, par, s)), F_at(L, F, xp));
        if constexpr ((M & M_H2) != 0) {
          const T *F2 = (const T *)L.F2;
          q = qadd(q, qmul(qmul(qconj(ldF(F2, Vh, par, s)), ldU<M>(U, Vh, mu, par, s)), F_at(L, F2, xp)));
        }
        // -(q0 (cos na - cos al) + q3 (sin na - sin al))
        dS -= (T)2 * fsin(h) * (q.d * fcos(m) - q.a * fsin(m));
      }
    }
    if (!metro_accept(dS, r)) return 0;
    *gmem(A + (size_t)(mu * 2 + par) * Vh + s) = na;
    return 1;
  } else if constexpr ((KIND == SU2HIP_K_DOUBLET_OVERRELAX ||
                        KIND == SU2HIP_K_DOUBLET_METROPOLIS) && (M & M_H2) != 0) {
    // two doublets; mu is the one updated (the reference's higgs_id)
    T *Fh = mu ? (T *)L.F2 : F;
    const T *Fo = mu ? (const T *)F : (const T *)L.F2;
    Q<T> ft = ldF(Fh, Vh, par, s);
    Q<double> f = qd(ft), H = qd(ldF(Fo, Vh, par, s));
    Q<double> st = qd(doublet_staple<T, M>(L, U, A, Fh, x));
    Q<T> nf;
    if constexpr (KIND == SU2HIP_K_DOUBLET_OVERRELAX) {
      // overrelax_higgs2 (src/overrelax.c:295-417) in FP64 for both
      // precisions: Y kept, X moved to the other root of the quartic
      double nX, X, p;
      Q<double> e;
      higgs2_or(L, mu, f, H, st, e, X, nX, p);
      if (!(p >= (double)uni<T>(r))) return 0;
      double k = nX - X;
      nf = Q<T>{(T)(f.a + e.a * k), (T)(f.b + e.b * k), (T)(f.c + e.c * k), (T)(f.d + e.d * k)};
    } else {
      // metro_doublet (src/metropolis.c:77-106) with localact_doublet
      // (src/su2u1.c:495-510) as a difference: hopping, kinetic and the
      // two-doublet potential
      nf = Q<T>{ft.a + (T)1 * (uni<T>(r) - (T)0.5), ft.b + (T)1 * (uni<T>(r) - (T)0.5),
                ft.c + (T)1 * (uni<T>(r) - (T)0.5), ft.d + (T)1 * (uni<T>(r) - (T)0.5)};
      Q<double> n = qd(nf);
      double o2 = qdot(f, f), n2 = qdot(n, n);
      double pn = mu ? pot2(L, H, n) : pot2(L, n, H);
      double po = mu ? pot2(L, H, f) : pot2(L, f, H);
      double diff = (n.a - f.a) * st.a + (n.b - f.b) * st.b + (n.c - f.c) * st.c +
                    (n.d - f.d) * st.d + 3.0 * (n2 - o2) + (pn - po);
      if (!metro_accept((T)diff, r)) return 0;
    }
    stF(Fh, Vh, par, s, nf);
    return 1;
  } else if constexpr (KIND == SU2HIP_K_DOUBLET_OVERRELAX ||
                       KIND == SU2HIP_K_DOUBLET_METROPOLIS) {
    Q<T> f = ldF(F, Vh, par, s);
    Q<T> st = doublet_staple<T, M>(L, U, A, F, x);
    double B0 = 0.5 * L.msq + 1.0 * 3;
    if constexpr ((M & M_TR) != 0) {
      V3<T> g = ldG(G, Vh, par, s);
      B0 += 0.5 * L.a2 * (0.5 * ((double)g.x * g.x + (double)g.y * g.y + (double)g.z * g.z));
    }
    if constexpr ((M & M_SG) != 0) {
      double Sx = ldS(S, Vh, par, s);
      B0 += 0.25 * L.a1_s * Sx + 0.25 * L.a2_s * Sx * Sx;
    }
    double C = 0.25 * L.lambda;
    Q<T> nf;
    int acc;
    if constexpr (KIND == SU2HIP_K_DOUBLET_OVERRELAX && sizeof(T) == 4) {
      // overrelax_doublet in FP32
      T n2 = qdot(st, st);
      T iF = frsqrt(n2);
      T Fn = n2 * iF;
      Q<T> e = qscale(st, iF);
      T X = qdot(f, e);
      Q<T> y{f.a - X * e.a, f.b - X * e.b, f.c - X * e.c, f.d - X * e.d};
      T nX;
      T p = xy_or_f(Fn, X, qdot(y, y), (T)B0, (T)C, nX);
      if (!(p >= uni<T>(r))) return 0;
      T k = nX + X;
      nf = Q<T>{-f.a + k * e.a, -f.b + k * e.b, -f.c + k * e.c, -f.d + k * e.d};
    } else if constexpr (KIND == SU2HIP_K_DOUBLET_OVERRELAX) {
      // overrelax_doublet (src/overrelax.c:54-128)
      double s0 = st.a, s1 = st.b, s2 = st.c, s3 = st.d;
      double Fn = ::sqrt(s0 * s0 + s1 * s1 + s2 * s2 + s3 * s3);
      double X = ((double)f.a * s0 + (double)f.b * s1 + (double)f.c * s2 + (double)f.d * s3) / Fn;
      double y0 = f.a - X * s0 / Fn, y1 = f.b - X * s1 / Fn, y2 = f.c - X * s2 / Fn,
             y3 = f.d - X * s3 / Fn;
      double Ysq = y0 * y0 + y1 * y1 + y2 * y2 + y3 * y3;
      double nX;
      double p = xy_or(Fn, X, Ysq, B0, C, nX);
      acc = p >= rng_u53(r);
      if (!acc) return 0;
      double k = (nX + X);
      nf = Q<T>{(T)(-1.0 * f.a + k * s0 / Fn), (T)(-1.0 * f.b + k * s1 / Fn),
                (T)(-1.0 * f.c + k * s2 / Fn), (T)(-1.0 * f.d + k * s3 / Fn)};
    } else {
      // metro_doublet (src/metropolis.c:77-106); Delta S of
      // localact_doublet = (f'-f).s + B(|f'|^2-|f|^2) + C(|f'|^4-|f|^4)
      nf = Q<T>{f.a + (T)1 * (uni<T>(r) - (T)0.5), f.b + (T)1 * (uni<T>(r) - (T)0.5),
                f.c + (T)1 * (uni<T>(r) - (T)0.5), f.d + (T)1 * (uni<T>(r) - (T)0.5)};
      T o2 = qdot(f, f), n2 = qdot(nf, nf);
      T diff = (nf.a - f.a) * st.a + (nf.b - f.b) * st.b + (nf.c - f.c) * st.c +
               (nf.d - f.d) * st.d + (T)B0 * (n2 - o2) + (T)C * (n2 * n2 - o2 * o2);
      if (!metro_accept(diff, r)) return 0;
    }
    stF(F, Vh, par, s, nf);
    return 1;
  } else if constexpr (KIND == SU2HIP_K_TRIPLET_OVERRELAX ||
                       KIND == SU2HIP_K_TRIPLET_METROPOLIS) {
    V3<T> g = ldG(G, Vh, par, s);
    V3<T> st = triplet_staple<T, M>(L, U, G, x);
    double C = 0.25 * L.b4;
    V3<T> ng;
    if constexpr (KIND == SU2HIP_K_TRIPLET_OVERRELAX && sizeof(T) == 4) {
      // overrelax_triplet in FP32 (no portal term in B, src/overrelax.c:202)
      T n2 = st.x * st.x + st.y * st.y + st.z * st.z;
      T iF = frsqrt(n2);
      T Fn = n2 * iF;
      V3<T> e{st.x * iF, st.y * iF, st.z * iF};
      T X = g.x * e.x + g.y * e.y + g.z * e.z;
      V3<T> y{g.x - X * e.x, g.y - X * e.y, g.z - X * e.z};
      T nX;
      T p = xy_or_f(Fn, X, y.x * y.x + y.y * y.y + y.z * y.z, (T)(0.5 * L.msq_t + 3.0), (T)C, nX);
      if (!(p >= uni<T>(r))) return 0;
      T k = nX + X;
      ng = V3<T>{-g.x + k * e.x, -g.y + k * e.y, -g.z + k * e.z};
    } else if constexpr (KIND == SU2HIP_K_TRIPLET_OVERRELAX) {
      // overrelax_triplet (src/overrelax.c:137-236); the portal term is not
      // in B, as in the reference (src/overrelax.c:202)
      double s0 = st.x, s1 = st.y, s2 = st.z;
      double Fn = ::sqrt(s0 * s0 + s1 * s1 + s2 * s2);
      double X = ((double)g.x * s0 + (double)g.y * s1 + (double)g.z * s2) / Fn;
      double y0 = g.x - X * s0 / Fn, y1 = g.y - X * s1 / Fn, y2 = g.z - X * s2 / Fn;
      double Ysq = y0 * y0 + y1 * y1 + y2 * y2;
      double B0 = 0.5 * L.msq_t + 1.0 * 3;
      double nX;
      double p = xy_or(Fn, X, Ysq, B0, C, nX);
      if (!(p >= rng_u53(r))) return 0;
      double k = (nX + X);
      ng = V3<T>{(T)(-1.0 * g.x + k * s0 / Fn), (T)(-1.0 * g.y + k * s1 / Fn),
                 (T)(-1.0 * g.z + k * s2 / Fn)};
    } else {
      // metro_triplet (src/metropolis.c:113-143) with localact_triplet
      // (src/su2u1.c:657-679) as a difference
      double B0 = 0.5 * L.msq_t + 1.0 * 3;
      if constexpr ((M & M_NH) != 0) {
        Q<T> f = ldF(F, Vh, par, s);
        B0 += 0.5 * L.a2 * (0.5 * (double)qdot(f, f));
      }
      ng = V3<T>{g.x + (T)1 * (uni<T>(r) - (T)0.5), g.y + (T)1 * (uni<T>(r) - (T)0.5),
                 g.z + (T)1 * (uni<T>(r) - (T)0.5)};
      T o2 = g.x * g.x + g.y * g.y + g.z * g.z, n2 = ng.x * ng.x + ng.y * ng.y + ng.z * ng.z;
      T diff = (ng.x - g.x) * st.x + (ng.y - g.y) * st.y + (ng.z - g.z) * st.z +
               (T)B0 * (n2 - o2) + (T)C * (n2 * n2 - o2 * o2);
      if (!metro_accept(diff, r)) return 0;
    }
    stG(G, Vh, par, s, ng);
    return 1;
  } else if constexpr (KIND == SU2HIP_K_SINGLET_OVERRELAX ||
                       KIND == SU2HIP_K_SINGLET_METROPOLIS) {
    double Sx = ldS(S, Vh, par, s);
    double nb = 0.0;
#pragma unroll
    for (int d = 0; d < 3; d++)
      nb += (double)S_at(L, S, shift(L, x, d, +1)) + (double)S_at(L, S, shift(L, x, d, -1));
    double ph = 0.0;
    if constexpr ((M & M_NH) != 0) {
      Q<T> f = ldF(F, Vh, par, s);
      ph = 0.5 * (double)qdot(f, f);
    }
    double nS;
    if constexpr (KIND == SU2HIP_K_SINGLET_OVERRELAX) {
      // overrelax_singlet (src/overrelax.c:243-285)
      double c1 = L.b1_s - nb;
      double c2 = 3 + 0.5 * L.msq_s;
      double c3 = 1.0 / 3.0 * L.b3_s;
      double c4 = 0.25 * L.b4_s;
      if constexpr ((M & M_NH) != 0) {
        c1 += 0.5 * L.a1_s * ph;
        c2 += 0.5 * L.a2_s * ph;
      }
      double d0 = c1 + c2 * Sx + c3 * Sx * Sx + c4 * Sx * Sx * Sx;
      double d1 = c2 + c3 * Sx + c4 * Sx * Sx;
      double d2 = c3 + c4 * Sx;
      double d3 = c4;
      nS = polysolve3(d3, d2, d1, d0);
      double dV = c1 + 2.0 * c2 * Sx + 3.0 * c3 * Sx * Sx + 4.0 * c4 * Sx * Sx * Sx;
      double dVn = c1 + 2.0 * c2 * nS + 3.0 * c3 * nS * nS + 4.0 * c4 * nS * nS * nS;
      if (!(::fabs(dV / dVn) >= rng_u53(r))) return 0;
    } else {
      // metro_singlet (src/metropolis.c:148-167), localact_singlet
      // (src/su2u1.c:688-723) as a difference
      T u = uni<T>(r);
      nS = (double)(T)(Sx + (T)1 * (u - (T)0.5));
      auto pot = [&](double v) {
        double pv = L.b1_s * v + 0.5 * L.msq_s * v * v + 1.0 / 3.0 * L.b3_s * v * v * v +
                    0.25 * L.b4_s * v * v * v * v;
        if constexpr ((M & M_NH) != 0) pv += 0.5 * L.a1_s * v * ph + 0.5 * L.a2_s * v * v * ph;
        return pv;
      };
      double diff = 3.0 * (nS * nS - Sx * Sx) - (nS - Sx) * nb + (pot(nS) - pot(Sx));
      if (!metro_accept((T)diff, r)) return 0;
    }
    *gmem(S + (size_t)par * Vh + s) = (T)nS;
    return 1;
  }
  return 0;
}

// ------------------------------------------------------------------------
// multicanonical helpers (src/multicanonical.c, src/update.c:82-156)

// hid: the doublet of a doublet kind (the first carries PHISQ and
// PHI2MINUSSIGMA2, the second PHI2SQ, src/update.c:95-101)
__device__ __forceinline__ bool muca_field_on(const MucaDev &m, int kind, int hid) {
  if (!m.enabled || !m.do_acceptance) return false;
  bool dbl = kind == SU2HIP_K_DOUBLET_OVERRELAX || kind == SU2HIP_K_DOUBLET_METROPOLIS;
  bool trp = kind == SU2HIP_K_TRIPLET_OVERRELAX || kind == SU2HIP_K_TRIPLET_METROPOLIS;
  if (dbl && hid) return m.orderparam == SU2HIP_PHI2SQ;
  if (dbl) return m.orderparam == SU2HIP_PHISQ || m.orderparam == SU2HIP_PHI2MINUSSIGMA2;
  if (trp) return m.orderparam == SU2HIP_SIGMASQ || m.orderparam == SU2HIP_PHI2MINUSSIGMA2;
  return false;
}

// Sites [lo,hi) of step `step` of a parity sweep with checks_per_sweep cps
// (a check after every `iv` sites, plus a final check when Vh % cps != 0,
// src/update.c:105-153); returns false if the lattice has no such step.
__device__ __forceinline__ bool muca_step_range(int Vh, int cps, int step, int &lo, int &hi) {
  int iv = Vh / cps;
  if (iv <= 0) iv = 1;
  int nfull = Vh / iv;
  int extra = (Vh % iv != 0) || (Vh % cps != 0);
  if (step >= nfull + extra) return false;
  lo = step * iv;
  hi = step < nfull ? lo + iv : Vh;
  return true;
}

template <typename T, int M>
__device__ __forceinline__ double orderparam_at(const LatDesc &L, int op, int par, int s) {
  double v = 0.0;
  if constexpr ((M & M_NH) != 0) {
    if (op == SU2HIP_PHISQ || op == SU2HIP_PHI2MINUSSIGMA2) {
      Q<T> f = ldF((const T *)L.F, L.Vh, par, s);
      v += 0.5 * ((double)f.a * f.a + (double)f.b * f.b + (double)f.c * f.c + (double)f.d * f.d);
    }
  }
  if constexpr ((M & M_H2) != 0) {
    if (op == SU2HIP_PHI2SQ) {
      Q<T> f = ldF((const T *)L.F2, L.Vh, par, s);
      v += 0.5 * ((double)f.a * f.a + (double)f.b * f.b + (double)f.c * f.c + (double)f.d * f.d);
    }
  }
  if constexpr ((M & M_TR) != 0) {
    if (op == SU2HIP_SIGMASQ || op == SU2HIP_PHI2MINUSSIGMA2) {
      V3<T> g = ldG((const T *)L.G, L.Vh, par, s);
      double t = 0.5 * ((double)g.x * g.x + (double)g.y * g.y + (double)g.z * g.z);
      v = (op == SU2HIP_SIGMASQ) ? t : v - t;
    }
  }
  return v;
}

// ------------------------------------------------------------------------
// the pass kernel

template <typename T, int M, int KIND>
__device__ __forceinline__ void pass_body(const LatDesc *__restrict__ lats,
                                          const MucaDev *__restrict__ mucas,
                                          const BlockMap *__restrict__ blk, int par_in, int mu_in,
                                          int pass_off, int step, int reunit,
                                          long long *__restrict__ cnt, int nblk,
                                          double *__restrict__ partial,
                                          const uint8_t *__restrict__ order, int npass,
                                          double *lds) {
  const BlockMap bm = blk[blockIdx.x];
  const LatDesc &L = lats[bm.lat];
  if (!L.active || L.budget == 0) return;
  int par = par_in, mu = mu_in;
  if (order) {
    // random pass order: this lattice's parity (and link direction) of the step
    uint8_t o = order[(size_t)bm.lat * npass + pass_off];
    par = o & 1;
    if (KIND <= SU2HIP_K_U1_METROPOLIS) mu = o >> 1;
  }
  const int s = bm.s0 + (int)threadIdx.x;
  const bool valid = s < L.Vh;
  int lo = 0, hi = L.Vh;
  bool muca = false;
  int op = 0;
  if (step >= 0) {
    const MucaDev &m = mucas[bm.lat];
    muca = muca_field_on(m, KIND, (M & M_H2) ? mu : 0);
    if (muca) {
      if (!muca_step_range(L.Vh, m.cps, step, lo, hi)) return;
      op = m.orderparam;
    } else if (step > 0) {
      return;
    }
  }
  const bool upd = valid && s >= lo && s < hi;
  int acc = 0;
  if (upd) {
    if (muca) {
      // backup of every site of the step, accepted locally or not: a
      // rejected global check restores the whole step from it
      if constexpr (KIND == SU2HIP_K_DOUBLET_OVERRELAX || KIND == SU2HIP_K_DOUBLET_METROPOLIS)
        stF((T *)L.B, L.Vh, par, s, ldF((const T *)((M & M_H2) && mu ? L.F2 : L.F), L.Vh, par, s));
      if constexpr (KIND == SU2HIP_K_TRIPLET_OVERRELAX || KIND == SU2HIP_K_TRIPLET_METROPOLIS)
        stG((T *)L.B, L.Vh, par, s, ldG((const T *)L.G, L.Vh, par, s));
    }
    Site x = site_of(L, par, s);
    Rng r;
    rng_init(r, L.key0, L.key1, L.pass + (uint64_t)pass_off, (uint32_t)(par * L.Vh + s));
    if constexpr (KIND == SU2HIP_K_U1_METROPOLIS && sizeof(T) == 4) {
      // one inlined copy per link direction: the neighbour offsets and the
      // plaquettes' directions become constants, so a site's loads issue
      // together instead of one direction at a time (6.8 -> 6.4 us at 96^3).
      // The SU(2) link kinds and FP64 lose more to the extra registers
      // (a wave per SIMD fewer, or spills) than they gain.
      if (mu == 0)
        acc = site_update<T, M, KIND>(L, x, par, s, 0, r, reunit);
      else if (mu == 1)
        acc = site_update<T, M, KIND>(L, x, par, s, 1, r, reunit);
      else
        acc = site_update<T, M, KIND>(L, x, par, s, 2, r, reunit);
    } else {
      acc = site_update<T, M, KIND>(L, x, par, s, mu, r, reunit);
    }
  }
  int nacc = __syncthreads_count(acc);
  int ntot = __syncthreads_count(upd);
  if (threadIdx.x == 0) {
    int row = KIND;
    if constexpr ((M & M_H2) != 0 && KIND == SU2HIP_K_DOUBLET_OVERRELAX)
      row = mu ? CNT_DOUBLET2_OR : KIND;
    if constexpr ((M & M_H2) != 0 && KIND == SU2HIP_K_DOUBLET_METROPOLIS)
      row = mu ? CNT_DOUBLET2_METRO : KIND;
    cnt[(size_t)(row * 2 + 0) * nblk + blockIdx.x] += nacc;
    cnt[(size_t)(row * 2 + 1) * nblk + blockIdx.x] += ntot;
  }
  if (muca) {
    double v = valid ? orderparam_at<T, M>(L, op, par, s) : 0.0;
    double t = block_sum(v, lds);
    if (threadIdx.x == 0) partial[L.blk_base + bm.s0 / BLK] = t;
  }
}

#define K_PASS_ARGS                                                                          \
  const LatDesc *__restrict__ lats, const MucaDev *__restrict__ mucas,                       \
      const BlockMap *__restrict__ blk, int par_in, int mu_in, int pass_off, int step,       \
      int reunit, long long *__restrict__ cnt, int nblk, double *__restrict__ partial,       \
      const uint8_t *__restrict__ order, int npass

template <typename T, int M, int KIND>
__global__ __launch_bounds__(BLK) void k_pass(K_PASS_ARGS) {
  __shared__ double lds[BLK / 64];
  pass_body<T, M, KIND>(lats, mucas, blk, par_in, mu_in, pass_off, step, reunit, cnt, nblk,
                        partial, order, npass, lds);
}

// The same pass with an occupancy target of 5 waves per SIMD, built for the
// link overrelaxation only: whether it wins depends on the lattice (FP64
// 96^3 22.5 against 22.9 us, 192^3 232 against 222 us, in one GPU
// session), so the autotuner picks per handle; the other kinds either do
// not change or spill under it.
template <typename T, int M, int KIND>
__global__ __launch_bounds__(BLK) __attribute__((amdgpu_waves_per_eu(5))) void k_pass_w5(
    K_PASS_ARGS) {
  __shared__ double lds[BLK / 64];
  pass_body<T, M, KIND>(lats, mucas, blk, par_in, mu_in, pass_off, step, reunit, cnt, nblk,
                        partial, order, npass, lds);
}

// The hopping-term cache of one doublet: for every link, the doublet part
// of su2link_staple (src/staples.c:150-190) as the link passes would form
// it.  Neither the link passes nor the link staple change the doublet or
// the U(1) links, so a round of link passes reads 4 values per link here
// instead of two doublets and a U(1) link, and FP64 evaluates each U(1)
// sincos once per round instead of once per pass.
template <typename T, int M>
__global__ __launch_bounds__(BLK) void k_hop(const LatDesc *__restrict__ lats,
                                             const BlockMap *__restrict__ blk) {
  const BlockMap bm = blk[blockIdx.x];
  const LatDesc &L = lats[bm.lat];
  if (!L.active || L.budget == 0 || !L.H) return;
  const int s = bm.s0 + (int)threadIdx.x;
  if (s >= L.Vh) return;
  const T *A = (const T *)L.A, *F = (const T *)L.F;
  T *H = (T *)L.H;
#pragma unroll 1
  for (int par = 0; par < 2; par++) {
    const Site x = site_of(L, par, s);
#pragma unroll
    for (int mu = 0; mu < 3; mu++)
      st4(H + ((size_t)(mu * 2 + par) * L.Vh + s) * 4, hop_term<T, M>(L, A, F, x, mu));
  }
}

// get_weight (src/multicanonical.c:350-379) with whichbin (:671-704)
__device__ int muca_whichbin(const MucaDev &m, double val) {
  if (val < m.min) return 0;
  if (val >= m.max) return m.bins + 1;
  int lo = 1, hi = m.bins;
  while (lo <= hi) {
    int bin = lo + (hi - lo) / 2;
    double cur = m.pos[bin], nxt = m.pos[bin + 1];
    if ((val < nxt && val > cur) || (::fabs(val - cur) < 1e-12)) return bin;
    if (val > cur)
      lo = bin + 1;
    else
      hi = bin - 1;
  }
  return m.bins;
}

__device__ double muca_weight(const MucaDev &m, double val) {
  int n = m.bins;
  if (val >= m.max) return m.b[n + 1];
  if (val < m.min) return m.b[0];
  int bin = muca_whichbin(m, val);
  if (bin >= n + 1) return m.b[n + 1];
  return m.b[bin] + m.slope[bin] * val;
}

// The same lookup with the bin edges staged in LDS (pos = m.pos copied):
// the same comparisons in the same order, without a dependent global load
// per step of the search.
__device__ int muca_whichbin_staged(const MucaDev &m, const double *pos, double val) {
  if (val < m.min) return 0;
  if (val >= m.max) return m.bins + 1;
  int lo = 1, hi = m.bins;
  while (lo <= hi) {
    int bin = lo + (hi - lo) / 2;
    double cur = pos[bin], nxt = pos[bin + 1];
    if ((val < nxt && val > cur) || (::fabs(val - cur) < 1e-12)) return bin;
    if (val > cur)
      lo = bin + 1;
    else
      hi = bin - 1;
  }
  return m.bins;
}
__device__ double muca_weight_staged(const MucaDev &m, const double *pos, double val) {
  int n = m.bins;
  if (val >= m.max) return m.b[n + 1];
  if (val < m.min) return m.b[0];
  int bin = muca_whichbin_staged(m, pos, val);
  if (bin >= n + 1) return m.b[n + 1];
  return m.b[bin] + m.slope[bin] * val;
}

// bin edges that fit the accept kernel's LDS stage
constexpr int MUCA_STAGE = 2048;

// linearize_weight (src/multicanonical.c:337-346)
__device__ void muca_linearize(MucaDev &m) {
  int n = m.bins;
  m.slope[0] = 0;
  m.slope[n + 1] = 0;
  m.b[0] = m.W[0];
  m.b[n + 1] = m.W[n + 1];
  for (int i = 1; i < n + 1; i++) {
    m.slope[i] = (m.W[i + 1] - m.W[i]) / (m.pos[i + 1] - m.pos[i]);
    m.b[i] = m.W[i] - m.slope[i] * m.pos[i];
  }
}

// muca_accumulate_hits (src/multicanonical.c:451-498)
__device__ void muca_hits(MucaDev &m, double val) {
  int n = m.bins;
  int bin = muca_whichbin(m, val);
  if (bin == 0) {
    double dmin = m.pos[2] - m.pos[1];
    if (val < m.pos[1] - 0.5 * dmin) return;
    bin += 1;
  } else if (bin < n + 1) {
    if (val - m.pos[bin] > m.pos[bin + 1] - val) bin = bin + 1;
  } else {
    double dmax = m.pos[n + 1] - m.pos[n];
    if (val > m.pos[n + 1] + 0.5 * dmax) return;
  }
  if (m.mode == SU2HIP_FAST) {
    m.hits[bin] += 5;
    if (bin > 1) m.hits[bin - 1] += 3;
    if (bin > 2) m.hits[bin - 2] += 1;
    if (bin < n) m.hits[bin + 1] += 3;
    if (bin < n - 1) m.hits[bin + 2] += 1;
  } else if (m.mode == SU2HIP_SLOW) {
    m.hits[bin]++;
    m.hgram[bin] += ::exp(muca_weight(m, val));
  }
}

// update_weight (src/multicanonical.c:504-557), FAST recursion
__device__ void muca_update_fast(MucaDev &m) {
  int n = m.bins, first = -1;
  double delta = 0.0;
  for (int i = 0; i < n + 2; i++) {
    if (m.pos[i] >= m.wrk_min && m.wrk_max >= m.pos[i]) {
      if (first < 0) first = i;
      delta = (m.hits[i] - m.hits[first]) * m.delta / ((double)n);
      m.W[i] += delta;
    } else if (m.wrk_max < m.pos[i]) {
      m.W[i] += delta;
    }
  }
  m.W[0] = m.W[1];
  m.W[n + 1] = m.W[n];
  muca_linearize(m);
  int tunnel = 0;
  if (m.last_max && m.hits[1] > 0) {
    m.last_max = 0;
    tunnel = 1;
  } else if (!m.last_max && m.hits[n] > 0) {
    m.last_max = 1;
    tunnel = 1;
  }
  if (tunnel) m.delta /= 1.5;
  for (int i = 0; i < n + 2; i++) m.hits[i] = 0;
}

// update_weight_slow (src/multicanonical.c:561-666), hep-lat/9804019; the
// scratch copy of W lives in hgram's second half
__device__ void muca_update_slow(MucaDev &m) {
  int n = m.bins;
  double *wold = m.hgram + (n + 2);
  for (int i = 0; i < n + 2; i++) wold[i] = m.W[i];
  double dmin = m.pos[2] - m.pos[1];
  m.hgram[1] /= dmin;
  for (int i = 2; i < n + 1; i++) m.hgram[i] /= (m.pos[i + 1] - m.pos[i - 1]);
  int first = -1, last = -1;
  for (int i = 1; i < n + 1; i++) {
    if (m.pos[i] >= m.wrk_min && m.wrk_max >= m.pos[i]) {
      if (first < 0) {
        first = i;
        continue;
      }
      m.nsum[i] += m.hits[i];
      int g;
      if (m.hits[i] >= 8 && m.hits[i - 1] >= 8) {
        g = m.hits[i] + m.hits[i - 1];
      } else {
        m.W[i] = wold[i] - wold[i - 1] + m.W[i - 1];
        continue;
      }
      long long gnew = m.gsum[i] + g;
      double ln = m.hgram[i - 1] / m.hgram[i];
      ln = -1.0 * g * ::log(ln);
      m.W[i] = m.W[i - 1] + ((wold[i] - wold[i - 1]) * m.gsum[i] + ln) / ((double)gnew);
      m.gsum[i] = gnew;
    } else if (m.wrk_max < m.pos[i]) {
      if (last < 0) last = i - 1;
      m.W[i] += m.W[last] - wold[last];
    }
  }
  for (int i = 0; i < n + 2; i++) {
    m.hits[i] = 0;
    m.hgram[i] = 0.0;
  }
  m.W[0] = m.W[1];
  m.W[n + 1] = m.W[n];
  for (int i = 0; i < n + 2; i++) wold[i] = m.W[i];
  double corr = 0.0;
  dmin = (first == 1) ? m.pos[first + 1] - m.pos[first] : m.pos[first + 1] - m.pos[first - 1];
  for (int i = first + 1; i < n + 1; i++) {
    if (i <= last) {
      double d1 = m.pos[i + 1] - m.pos[i - 1];
      corr = 2.0 * ::log((dmin / d1) * m.nsum[i] / m.nsum[first]);
    }
    m.W[i] += corr;
  }
  m.W[0] = m.W[1];
  m.W[n + 1] = m.W[n];
  muca_linearize(m);
  for (int i = 0; i < n + 2; i++) m.W[i] = wold[i];
}

// muca_check (src/update.c:13-30) + calc_orderparam's reduction
// (src/multicanonical.c:709-751) + multicanonical_acceptance (:385-445);
// one workgroup per lattice, the sequential part in thread 0
__global__ __launch_bounds__(BLK) void k_muca_accept(const LatDesc *__restrict__ lats,
                                                     MucaDev *__restrict__ mucas,
                                                     const double *__restrict__ partial, int kind,
                                                     int hid, int par_in, int pass_off, int step,
                                                     const uint8_t *__restrict__ order, int npass) {
  __shared__ double lds[BLK / 64];
  __shared__ double pos[MUCA_STAGE];
  const LatDesc &L = lats[blockIdx.x];
  MucaDev &m = mucas[blockIdx.x];
  if (!L.active || L.budget == 0 || !muca_field_on(m, kind, hid)) return;
  const int par = order ? (order[(size_t)blockIdx.x * npass + pass_off] & 1) : par_in;
  int lo, hi;
  if (!muca_step_range(L.Vh, m.cps, step, lo, hi)) return;
  const bool staged = m.bins + 2 <= MUCA_STAGE;
  if (staged)
    for (int i = threadIdx.x; i < m.bins + 2; i += BLK) pos[i] = m.pos[i];
  // fixed-order sum of this lattice's block partials
  double v = 0.0;
  for (int i = threadIdx.x; i < L.nblk; i += BLK) v += partial[L.blk_base + i];
  double tot = block_sum(v, lds);  // its barriers also publish the staged edges
  if (threadIdx.x == 0) {
    double newp = tot / L.vol;
    double old = m.pv[0] + m.pv[1];
    m.pv[par] = newp;
    double nw = newp + m.pv[1 - par];
    double diff = staged ? muca_weight_staged(m, pos, nw) - muca_weight_staged(m, pos, old)
                         : muca_weight(m, nw) - muca_weight(m, old);
    Rng r;
    rng_init(r, L.key0, L.key1, L.pass + (uint64_t)pass_off, MUCA_SITE);
    r.k = (uint32_t)step;
    int accept = ::exp(-(diff)) > rng_u53(r);
    if (m.mode != SU2HIP_READONLY) {
      if (!accept) nw = old;
      muca_hits(m, nw);
      m.muca_count++;
      if (m.muca_count % m.update_interval == 0) {
        if (m.mode == SU2HIP_FAST)
          muca_update_fast(m);
        else if (m.mode == SU2HIP_SLOW)
          muca_update_slow(m);
        m.muca_count = 0;
      }
    }
    if (!accept) m.pv[par] = old - m.pv[1 - par];
    m.rejected = !accept;
    m.accepted += accept;
    m.total++;
  }
}

// undo a rejected step: copy the backup of the step's sites back
template <typename T>
__global__ __launch_bounds__(BLK) void k_muca_restore(const LatDesc *__restrict__ lats,
                                                      const MucaDev *__restrict__ mucas,
                                                      const BlockMap *__restrict__ blk, int kind,
                                                      int hid, int par_in, int pass_off, int step,
                                                      const uint8_t *__restrict__ order, int npass) {
  const BlockMap bm = blk[blockIdx.x];
  const LatDesc &L = lats[bm.lat];
  const MucaDev &m = mucas[bm.lat];
  if (!L.active || L.budget == 0 || !muca_field_on(m, kind, hid) || !m.rejected) return;
  const int par = order ? (order[(size_t)bm.lat * npass + pass_off] & 1) : par_in;
  int lo, hi;
  if (!muca_step_range(L.Vh, m.cps, step, lo, hi)) return;
  int s = bm.s0 + (int)threadIdx.x;
  if (s >= L.Vh || s < lo || s >= hi) return;
  bool dbl = kind == SU2HIP_K_DOUBLET_OVERRELAX || kind == SU2HIP_K_DOUBLET_METROPOLIS;
  if (dbl)
    stF((T *)(hid ? L.F2 : L.F), L.Vh, par, s, ldF((const T *)L.B, L.Vh, par, s));
  else
    stG((T *)L.G, L.Vh, par, s, ldG((const T *)L.B, L.Vh, par, s));
}

// order parameter of both parities (calc_orderparam at start-up,
// src/main.c:212-217): partial sums of parity p in partial[p*nblk + ...]
template <typename T, int M>
__global__ __launch_bounds__(BLK) void k_orderparam(const LatDesc *__restrict__ lats,
                                                    const MucaDev *__restrict__ mucas,
                                                    const BlockMap *__restrict__ blk, int nblk,
                                                    double *__restrict__ partial, int only) {
  __shared__ double lds[BLK / 64];
  const BlockMap bm = blk[blockIdx.x];
  const LatDesc &L = lats[bm.lat];
  const MucaDev &m = mucas[bm.lat];
  if (!L.active || !m.enabled || (only >= 0 && bm.lat != only)) return;
  int s = bm.s0 + (int)threadIdx.x;
  for (int par = 0; par < 2; par++) {
    double v = s < L.Vh ? orderparam_at<T, M>(L, m.orderparam, par, s) : 0.0;
    double t = block_sum(v, lds);
    if (threadIdx.x == 0) partial[(size_t)par * nblk + L.blk_base + bm.s0 / BLK] = t;
  }
}

__global__ __launch_bounds__(BLK) void k_orderparam_final(const LatDesc *__restrict__ lats,
                                                          MucaDev *__restrict__ mucas, int nblk,
                                                          const double *__restrict__ partial,
                                                          int only) {
  __shared__ double lds[BLK / 64];
  const LatDesc &L = lats[blockIdx.x];
  MucaDev &m = mucas[blockIdx.x];
  if (!L.active || !m.enabled || (only >= 0 && (int)blockIdx.x != only)) return;
  for (int par = 0; par < 2; par++) {
    double v = 0.0;
    for (int i = threadIdx.x; i < L.nblk; i += BLK) v += partial[(size_t)par * nblk + L.blk_base + i];
    double tot = block_sum(v, lds);
    if (threadIdx.x == 0) m.pv[par] = tot / L.vol;
  }
}

// shuffle (src/update.c:461-473) with iran() the next 64-bit draw
__device__ __forceinline__ void shuffle6(int *a, Rng &r) {
  for (int i = 5; i > 0; i--) {
    unsigned int j = (unsigned int)(rng_u64(r) % (unsigned long long)(i + 1));
    int t = a[i];
    a[i] = a[j];
    a[j] = t;
  }
}

// the random pass order of one sweep of every lattice (one thread per
// lattice): per link round the parities and then the directions shuffled,
// per scalar round a uniform for the first parity, in the reference's draw
// order (src/update.c:278-300, :337-345)
__global__ void k_order(const LatDesc *__restrict__ lats, int n, uint8_t *__restrict__ order,
                        const uint16_t *__restrict__ info, int npass, int link_rounds,
                        int scalar_rounds) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const LatDesc &L = lats[i];
  if (!L.active || L.budget == 0) return;
  Rng r;
  rng_init(r, L.key0, L.key1, L.pass, SCHED_SITE);
  uint8_t *o = order + (size_t)i * npass;
  for (int k = 0; k < link_rounds; k++) {
    int pa[6] = {0, 1, 0, 1, 0, 1}, da[6] = {0, 0, 1, 1, 2, 2};
    shuffle6(pa, r);
    shuffle6(da, r);
    for (int j = 0; j < npass; j++) {
      uint16_t in = info[j];
      if (!(in & STEP_SCALAR) && ((in >> 8) & 0x7F) == k) {
        int pos = in & 0xFF;
        o[j] = (uint8_t)(pa[pos] | (da[pos] << 1));
      }
    }
  }
  for (int s = 0; s < scalar_rounds; s++) {
    int p0 = rng_u53(r) < 0.5 ? 0 : 1;
    for (int j = 0; j < npass; j++) {
      uint16_t in = info[j];
      if ((in & STEP_SCALAR) && ((in >> 8) & 0x7F) == s)
        o[j] = (uint8_t)((in & 0xFF) == 0 ? p0 : 1 - p0);
    }
  }
}

// end of a compound sweep: advance the pass counter and the budget
__global__ void k_end_sweep(LatDesc *lats, int n, int passes) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  LatDesc &L = lats[i];
  if (!L.active || L.budget == 0) return;
  L.pass += (uint64_t)passes;
  if (L.budget > 0) L.budget--;
}

// ------------------------------------------------------------------------
// measure() site loop (src/measure.c:110-174) in FP64

struct Cm {
  double re[4], im[4];
};

__device__ __forceinline__ Cm cm_mul(const Cm &a, const Cm &b, bool dag) {
  Cm bb = b;
  if (dag) {
    bb.re[0] = b.re[0];
    bb.im[0] = -b.im[0];
    bb.re[1] = b.re[2];
    bb.im[1] = -b.im[2];
    bb.re[2] = b.re[1];
    bb.im[2] = -b.im[1];
    bb.re[3] = b.re[3];
    bb.im[3] = -b.im[3];
  }
  Cm t;
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++) {
      double sr = 0, si = 0;
#pragma unroll
      for (int k = 0; k < 2; k++) {
        double ar = a.re[2 * i + k], ai = a.im[2 * i + k];
        double br = bb.re[2 * k + j], bi = bb.im[2 * k + j];
        sr += ar * br - ai * bi;
        si += ar * bi + ai * br;
      }
      t.re[2 * i + j] = sr;
      t.im[2 * i + j] = si;
    }
  return t;
}

template <typename T>
__device__ __forceinline__ Cm halfproj(const V3<T> &g) {
  double a0 = g.x, a1 = g.y, a2 = g.z;
  double md = ::sqrt(a0 * a0 + a1 * a1 + a2 * a2);
  double n0 = a0 / md, n1 = a1 / md, n2 = a2 / md;
  Cm m;
  m.re[0] = 0.5 * (1 + n2);
  m.im[0] = 0.0;
  m.re[1] = 0.5 * n0;
  m.im[1] = -0.5 * n1;
  m.re[2] = 0.5 * n0;
  m.im[2] = 0.5 * n1;
  m.re[3] = 0.5 * (1 - n2);
  m.im[3] = 0.0;
  return m;
}

// project_u1 (src/magfield.c:115-184)
template <typename T, int M>
__device__ Cm project_u1(const LatDesc &L, const T *U, const T *G, const Site &x, int d) {
  Cm pl = halfproj(G_at(L, G, x)), pr = halfproj(G_at(L, G, shift(L, x, d, +1)));
  Q<T> q = U_at<M>(L, U, x, d);
  Cm u;
  u.re[0] = q.a;
  u.im[0] = q.d;
  u.re[1] = q.c;
  u.im[1] = q.b;
  u.re[2] = -q.c;
  u.im[2] = q.b;
  u.re[3] = q.a;
  u.im[3] = -q.d;
  return cm_mul(cm_mul(pl, u, false), pr, false);
}

// alpha_proj (src/magfield.c:189-216)
template <typename T, int M>
__device__ double alpha_proj(const LatDesc &L, const T *U, const T *G, const Site &x, int d1, int d2) {
  Cm u1 = project_u1<T, M>(L, U, G, x, d1);
  Cm u2 = project_u1<T, M>(L, U, G, shift(L, x, d1, +1), d2);
  Cm u3 = project_u1<T, M>(L, U, G, shift(L, x, d2, +1), d1);
  Cm u4 = project_u1<T, M>(L, U, G, x, d2);
  u1 = cm_mul(u1, u2, false);
  u1 = cm_mul(u1, u3, true);
  u1 = cm_mul(u1, u4, true);
  return ::atan2(u1.im[0] + u1.im[3], u1.re[0] + u1.re[3]) * ::sqrt(L.beta);
}

template <typename T, int M>
__device__ void measure_site(const LatDesc &L, int par, int s, double *acc) {
  const T *U = (const T *)L.U, *A = (const T *)L.A, *F = (const T *)L.F, *G = (const T *)L.G,
          *S = (const T *)L.S;
  Site x = site_of(L, par, s);
  // local_su2wilson (src/su2u1.c:143-157)
  double wl = 0.0;
#pragma unroll
  for (int d1 = 0; d1 < 3; d1++)
#pragma unroll
    for (int d2 = 0; d2 < d1; d2++) {
      Q<double> t = qmul(qmul(qmul(qd(U_at<M>(L, U, x, d2)), qd(U_at<M>(L, U, shift(L, x, d2, +1), d1))),
                              qconj(qd(U_at<M>(L, U, shift(L, x, d1, +1), d2)))),
                         qconj(qd(U_at<M>(L, U, x, d1))));
      wl += (1.0 - 0.5 * (2.0 * t.a));
    }
  wl *= L.beta;
  double action = wl;
  acc[SU2HIP_SUM_WILSON] += wl;
  if constexpr ((M & M_U1) != 0) {
    // local_u1wilson (src/su2u1.c:316-328)
    double u = 0.0;
#pragma unroll
    for (int d1 = 0; d1 < 3; d1++)
#pragma unroll
      for (int d2 = 0; d2 < d1; d2++) {
        double pl = (double)A_at(L, A, x, d2) + (double)A_at(L, A, shift(L, x, d2, +1), d1) -
                    (double)A_at(L, A, shift(L, x, d1, +1), d2) - (double)A_at(L, A, x, d1);
        u += (1.0 - ::cos(L.r_u1 * pl));
      }
    u *= L.betau1;
    action += u;
    acc[SU2HIP_SUM_U1WILSON] += u;
  }
  double mod = 0.0, tmod = 0.0;
  if constexpr ((M & M_NH) != 0) {
    Q<double> f = qd(F_at(L, F, x));
    mod = 0.5 * qdot(f, f);
    // covariant_doublet (src/su2u1.c:477-488)
    double cov = 0.0;
#pragma unroll
    for (int d = 0; d < 3; d++) {
      Q<double> t = qmul(qmul(qconj(f), qd(U_at<M>(L, U, x, d))), qd(F_at(L, F, shift(L, x, d, +1))));
      double hop;
      if constexpr ((M & M_U1) != 0) {
        double al = A_at(L, A, x, d);
        hop = -(t.a * ::cos(al) + t.d * ::sin(al));
      } else {
        hop = -t.a;
      }
      cov += 2.0 * mod + hop;
    }
    acc[SU2HIP_SUM_COV_PHI] += cov;
    acc[SU2HIP_SUM_PHI2] += mod;
    acc[SU2HIP_SUM_PHI4] += mod * mod;
    if constexpr ((M & M_H2) != 0) {
      // the second doublet's covariant derivative, moments and phi1^+ phi2;
      // pot2 holds the whole two-doublet potential
      const T *F2 = (const T *)L.F2;
      Q<double> f2 = qd(F_at(L, F2, x));
      double mod2 = 0.5 * qdot(f2, f2), cov2 = 0.0;
#pragma unroll
      for (int d = 0; d < 3; d++) {
        Q<double> t = qmul(qmul(qconj(f2), qd(U_at<M>(L, U, x, d))), qd(F_at(L, F2, shift(L, x, d, +1))));
        double hop;
        if constexpr ((M & M_U1) != 0) {
          double al = A_at(L, A, x, d);
          hop = -(t.a * ::cos(al) + t.d * ::sin(al));
        } else {
          hop = -t.a;
        }
        cov2 += 2.0 * mod2 + hop;
      }
      acc[SU2HIP_SUM_COV_PHI_2] += cov2;
      acc[SU2HIP_SUM_PHI2_2] += mod2;
      acc[SU2HIP_SUM_PHI4_2] += mod2 * mod2;
      acc[SU2HIP_SUM_R12] += 0.5 * (f.a * f2.a + f.b * f2.b + f.c * f2.c + f.d * f2.d);
      acc[SU2HIP_SUM_I12] += 0.5 * (f.d * f2.a + f.c * f2.b - f.b * f2.c - f.a * f2.d);
      action += cov + cov2 + pot2(L, f, f2);
    } else {
      action += cov + L.msq * mod + L.lambda * mod * mod;
    }
  }
  if constexpr ((M & M_TR) != 0) {
    V3<T> gt = G_at(L, G, x);
    V3<double> g{(double)gt.x, (double)gt.y, (double)gt.z};
    tmod = 0.5 * (g.x * g.x + g.y * g.y + g.z * g.z);
    double hop = 0.0;
#pragma unroll
    for (int d = 0; d < 3; d++) {
      V3<T> nb = G_at(L, G, shift(L, x, d, +1));
      V3<double> r = rotate(qd(U_at<M>(L, U, x, d)), V3<double>{(double)nb.x, (double)nb.y, (double)nb.z});
      hop += -(g.x * r.x + g.y * r.y + g.z * r.z);
    }
    acc[SU2HIP_SUM_SIGMA2] += tmod;
    acc[SU2HIP_SUM_SIGMA4] += tmod * tmod;
    acc[SU2HIP_SUM_HOP_SIGMA] += hop;
    // covariant_triplet + triplet potential (src/su2u1.c:558-565, :643-651)
    action += 3 * 2.0 * tmod + hop + L.msq_t * tmod + L.b4 * tmod * tmod;
    if constexpr ((M & M_NH) != 0) {
      action += L.a2 * mod * tmod;
      acc[SU2HIP_SUM_PHI2SIGMA2] += mod * tmod;
    }
    // the monopole charge of the cube at x is summed by k_magcharge
  }
  if constexpr ((M & M_SG) != 0) {
    double Sx = S_at(L, S, x);
    double kin = 3 * Sx * Sx, cov = 0.0;
#pragma unroll
    for (int d = 0; d < 3; d++) {
      double Sn = S_at(L, S, shift(L, x, d, +1));
      kin -= Sx * Sn;
      cov += Sx * Sx - Sx * Sn;
    }
    double pot = L.b1_s * Sx + 0.5 * L.msq_s * Sx * Sx + 1.0 / 3.0 * L.b3_s * Sx * Sx * Sx +
                 0.25 * L.b4_s * Sx * Sx * Sx * Sx;
    if constexpr ((M & M_NH) != 0) {
      pot += 0.5 * L.a1_s * Sx * mod + 0.5 * L.a2_s * Sx * Sx * mod;
      acc[SU2HIP_SUM_SPHISQ] += mod * Sx;
      acc[SU2HIP_SUM_S2PHISQ] += mod * Sx * Sx;
    }
    action += kin + pot;
    acc[SU2HIP_SUM_S_COV] += cov;
    acc[SU2HIP_SUM_S] += Sx;
    acc[SU2HIP_SUM_S2] += Sx * Sx;
    acc[SU2HIP_SUM_S3] += Sx * Sx * Sx;
    acc[SU2HIP_SUM_S4] += Sx * Sx * Sx * Sx;
  }
  acc[SU2HIP_SUM_ACTION] += action;
}

// One edge of a projected plaquette, given the projectors at its two ends:
// project_u1 (src/magfield.c:115-184) with halfproj evaluated by the caller.
template <typename T, int M>
__device__ __forceinline__ Cm edge_u1(const LatDesc &L, const T *U, const Cm &pl, const Site &x,
                                      int d, const Cm &pr) {
  Q<T> q = U_at<M>(L, U, x, d);
  Cm u;
  u.re[0] = q.a;
  u.im[0] = q.d;
  u.re[1] = q.c;
  u.im[1] = q.b;
  u.re[2] = -q.c;
  u.im[2] = q.b;
  u.re[3] = q.a;
  u.im[3] = -q.d;
  return cm_mul(cm_mul(pl, u, false), pr, false);
}

// alpha_proj (src/magfield.c:189-216) from its four projected edges
__device__ __forceinline__ double face_angle(const LatDesc &L, Cm u1, const Cm &u2, const Cm &u3,
                                             const Cm &u4) {
  u1 = cm_mul(u1, u2, false);
  u1 = cm_mul(u1, u3, true);
  u1 = cm_mul(u1, u4, true);
  return ::atan2(u1.im[0] + u1.im[3], u1.re[0] + u1.re[3]) * ::sqrt(L.beta);
}

// Monopole charge of the elementary cube at x, magcharge_cube
// (src/magfield.c:262-279): sum over dir of magfield(x + dir) - magfield(x),
// where magfield(y, 0) = +alpha(y; 1,2), magfield(y, 1) = -alpha(y; 0,2) and
// magfield(y, 2) = +alpha(y; 0,1) (src/magfield.c:221-256).  The cube's six
// faces share its eight corners; the projector of each corner is computed
// once instead of once per edge end, which leaves every value unchanged.
// Corner cabc sits at x + a e0 + b e1 + c e2.
template <typename T, int M>
__device__ __forceinline__ double magcharge_site(const LatDesc &L, int par, int s) {
  const T *U = (const T *)L.U, *G = (const T *)L.G;
  const Site c000 = site_of(L, par, s);
  const Site c100 = shift(L, c000, 0, +1), c010 = shift(L, c000, 1, +1), c001 = shift(L, c000, 2, +1);
  const Site c110 = shift(L, c100, 1, +1), c101 = shift(L, c100, 2, +1), c011 = shift(L, c010, 2, +1);
  const Site c111 = shift(L, c110, 2, +1);
  const Cm p000 = halfproj(G_at(L, G, c000)), p100 = halfproj(G_at(L, G, c100));
  const Cm p010 = halfproj(G_at(L, G, c010)), p001 = halfproj(G_at(L, G, c001));
  const Cm p110 = halfproj(G_at(L, G, c110)), p101 = halfproj(G_at(L, G, c101));
  const Cm p011 = halfproj(G_at(L, G, c011)), p111 = halfproj(G_at(L, G, c111));
  auto e = [&](const Cm &pl, const Site &y, int d, const Cm &pr) {
    return edge_u1<T, M>(L, U, pl, y, d, pr);
  };
  // alpha(y; d1,d2) = face(P_y U_d1 P_y+d1, P_y+d1 U_d2 P_y+d1+d2, P_y+d2 U_d1 P_y+d1+d2, P_y U_d2 P_y+d2)
  const double b0 = 0.0 + face_angle(L, e(p000, c000, 1, p010), e(p010, c010, 2, p011),
                                     e(p001, c001, 1, p011), e(p000, c000, 2, p001));
  const double b0n = 0.0 + face_angle(L, e(p100, c100, 1, p110), e(p110, c110, 2, p111),
                                      e(p101, c101, 1, p111), e(p100, c100, 2, p101));
  const double b1 = 0.0 - face_angle(L, e(p000, c000, 0, p100), e(p100, c100, 2, p101),
                                     e(p001, c001, 0, p101), e(p000, c000, 2, p001));
  const double b1n = 0.0 - face_angle(L, e(p010, c010, 0, p110), e(p110, c110, 2, p111),
                                      e(p011, c011, 0, p111), e(p010, c010, 2, p011));
  const double b2 = 0.0 + face_angle(L, e(p000, c000, 0, p100), e(p100, c100, 1, p110),
                                     e(p010, c010, 0, p110), e(p000, c000, 1, p010));
  const double b2n = 0.0 + face_angle(L, e(p001, c001, 0, p101), e(p101, c101, 1, p111),
                                      e(p011, c011, 0, p111), e(p001, c001, 1, p011));
  double ch = 0.0;
  ch += b0n - b0;
  ch += b1n - b1;
  ch += b2n - b2;
  return ch;
}

// Each thread measures site s of parity 0, then of parity 1, so every sum
// of a thread is (0 + site0) + site1 in that order.  The loop is kept
// rolled: interleaving the two sites doubles the live registers.
template <typename T, int M>
__global__ __launch_bounds__(BLK) void k_measure(const LatDesc *__restrict__ lats,
                                                 const BlockMap *__restrict__ blk,
                                                 double *__restrict__ partial) {
  __shared__ double lds[BLK / 64 * SU2HIP_NSUMS];
  const BlockMap bm = blk[blockIdx.x];
  const LatDesc &L = lats[bm.lat];
  if (!L.active) return;
  int s = bm.s0 + (int)threadIdx.x;
  double acc[SU2HIP_NSUMS];
#pragma unroll
  for (int j = 0; j < SU2HIP_NSUMS; j++) acc[j] = 0.0;
  if (s < L.Vh) {
#pragma unroll 1
    for (int par = 0; par < 2; par++) measure_site<T, M>(L, par, s, acc);
  }
  double t = block_sums<SU2HIP_NSUMS>(acc, lds);
  if (threadIdx.x < SU2HIP_NSUMS)
    partial[(size_t)(L.blk_base + bm.s0 / BLK) * SU2HIP_NSUMS + threadIdx.x] = t;
}

// The two monopole sums of the adjoint model, written into the rows
// k_measure has just filled (which hold 0 there): same threads, same sites,
// same order, so each is bitwise what one fused kernel would give.  Kept
// apart because the 24 projected links of a cube, in FP64, do not fit in
// registers beside the other sums.
template <typename T, int M>
__global__ __launch_bounds__(BLK) void k_magcharge(const LatDesc *__restrict__ lats,
                                                   const BlockMap *__restrict__ blk,
                                                   double *__restrict__ partial) {
  __shared__ double lds[BLK / 64 * 2];
  const BlockMap bm = blk[blockIdx.x];
  const LatDesc &L = lats[bm.lat];
  if (!L.active) return;
  int s = bm.s0 + (int)threadIdx.x;
  double acc[2] = {0.0, 0.0};
  if (s < L.Vh) {
#pragma unroll 1
    for (int par = 0; par < 2; par++) {
      double ch = magcharge_site<T, M>(L, par, s);
      acc[0] += ch;
      acc[1] += ::fabs(ch);
    }
  }
  double t = block_sums<2>(acc, lds);
  if (threadIdx.x < 2)
    partial[(size_t)(L.blk_base + bm.s0 / BLK) * SU2HIP_NSUMS + SU2HIP_SUM_MAGCHARGE + threadIdx.x] = t;
}

__global__ __launch_bounds__(BLK) void k_measure_final(const LatDesc *__restrict__ lats,
                                                       const double *__restrict__ partial,
                                                       double *__restrict__ sums) {
  __shared__ double lds[BLK / 64 * SU2HIP_NSUMS];
  const LatDesc &L = lats[blockIdx.x];
  double *out = sums + (size_t)blockIdx.x * SU2HIP_NSUMS;
  if (!L.active) {
    if (threadIdx.x < SU2HIP_NSUMS) out[threadIdx.x] = 0.0;
    return;
  }
  // each thread its rows in order, all sums of a row at once
  double v[SU2HIP_NSUMS];
#pragma unroll
  for (int j = 0; j < SU2HIP_NSUMS; j++) v[j] = 0.0;
  for (int i = threadIdx.x; i < L.nblk; i += BLK) {
    const double *row = partial + (size_t)(L.blk_base + i) * SU2HIP_NSUMS;
#pragma unroll
    for (int j = 0; j < SU2HIP_NSUMS; j++) v[j] += row[j];
  }
  double t = block_sums<SU2HIP_NSUMS>(v, lds);
  if (threadIdx.x < SU2HIP_NSUMS) out[threadIdx.x] = t;
}

// ------------------------------------------------------------------------
// correlators (the reference's CORRELATORS build, src/correlation.c):
// per-site operators stored by (value, plane along dir, position in the
// plane), then one fixed-order reduction per (value, plane)

// a.(b.(c.d)) in FP64
__device__ __forceinline__ Q<double> qmul4(const Q<double> &a, const Q<double> &b,
                                           const Q<double> &c, const Q<double> &d) {
  return qmul(a, qmul(b, qmul(c, d)));
}

// clover_su2 (src/su2u1.c:207-275): the four plaquettes of the (d1, d2)
// plane with a corner at x
template <typename T, int M>
__device__ Q<double> clover(const LatDesc &L, const T *U, const Site &x, int d1, int d2) {
  auto u = [&](const Site &c, int d) { return qd(U_at<M>(L, U, c, d)); };
  Site x1 = shift(L, x, d1, +1), x2 = shift(L, x, d2, +1);
  Site m1 = shift(L, x, d1, -1), m2 = shift(L, x, d2, -1);
  Site x2m1 = shift(L, x2, d1, -1), m1m2 = shift(L, m1, d2, -1), m2p1 = shift(L, m2, d1, +1);
  Q<double> c = qmul4(u(x, d1), u(x1, d2), qconj(u(x2, d1)), qconj(u(x, d2)));
  c = qadd(c, qmul4(u(x, d2), qconj(u(x2m1, d1)), qconj(u(m1, d2)), u(m1, d1)));
  c = qadd(c, qmul4(qconj(u(m1, d1)), qconj(u(m1m2, d2)), u(m1m2, d1), u(m2, d2)));
  c = qadd(c, qmul4(qconj(u(m2, d2)), u(m2, d1), u(m2p1, d2), qconj(u(x, d1))));
  return c;
}

template <typename T, int M>
__global__ __launch_bounds__(BLK) void k_corr_sites(const LatDesc *__restrict__ lats, int slot,
                                                    int dir, double *__restrict__ buf) {
  const LatDesc &L = lats[slot];
  int s = blockIdx.x * BLK + (int)threadIdx.x;
  if (s >= L.Vh) return;
  const T *U = (const T *)L.U, *F = (const T *)L.F, *G = (const T *)L.G;
  int Ld = dir == 0 ? L.L0 : (dir == 1 ? L.L1 : L.L2);
  size_t P = (size_t)L.vol / Ld, V = (size_t)L.vol;
  // transverse directions of projected_photon_operator (src/correlation.c:101-108)
  int d1 = dir == 0 ? 1 : 0, d2 = dir == 2 ? 1 : 2;
  for (int par = 0; par < 2; par++) {
    Site x = site_of(L, par, s);
    int cd = dir == 0 ? x.x : (dir == 1 ? x.y : x.z);
    size_t pos = dir == 0 ? (size_t)x.y + (size_t)L.L1 * x.z
                          : (dir == 1 ? (size_t)x.x + (size_t)L.L0 * x.z
                                      : (size_t)x.x + (size_t)L.L0 * x.y);
    double *o = buf + (size_t)cd * P + pos;
    double v[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    if constexpr ((M & M_NH) != 0) {
      Q<double> f = qd(F_at(L, F, x));
      v[0] = 0.5 * qdot(f, f);
    }
    if constexpr ((M & M_TR) != 0) {
      V3<T> gt = G_at(L, G, x);
      double a0 = gt.x, a1 = gt.y, a2 = gt.z;
      v[1] = 0.5 * (a0 * a0 + a1 * a1 + a2 * a2);
      Q<double> cl = clover<T, M>(L, U, x, d1, d2);
      double tr = a0 * (cl.b / 4.0) + a1 * (cl.c / 4.0) + a2 * (cl.d / 4.0);
      double al = 2.0 * alpha_proj<T, M>(L, U, G, x, d1, d2);
#pragma unroll
      for (int m = 1; m <= 2; m++) {
        // photon_correl's momentum: transverse momentum m along x only when
        // dir is not 0 (src/correlation.c:270-275)
        double px = dir != 0 ? (double)(m * x.x) / (double)L.L0 : 0.0;
        double cs = ::cos(2.0 * PI * px), sn = ::sin(2.0 * PI * px);
        v[2 * m] = tr * cs;
        v[2 * m + 1] = tr * sn;
        v[4 + 2 * m] = al * cs;
        v[4 + 2 * m + 1] = al * sn;
      }
    }
#pragma unroll
    for (int k = 0; k < 10; k++) o[(size_t)k * V] = v[k];
  }
}

// measure_along_z's site values (src/z_coord.c:206-263): Tr Phi^+Phi/2,
// Tr Sigma^2/2, S, S^2, stored like k_corr_sites' values
template <typename T, int M>
__global__ __launch_bounds__(BLK) void k_zprof_sites(const LatDesc *__restrict__ lats, int slot,
                                                     int dir, double *__restrict__ buf) {
  const LatDesc &L = lats[slot];
  int s = blockIdx.x * BLK + (int)threadIdx.x;
  if (s >= L.Vh) return;
  int Ld = dir == 0 ? L.L0 : (dir == 1 ? L.L1 : L.L2);
  size_t P = (size_t)L.vol / Ld, V = (size_t)L.vol;
  for (int par = 0; par < 2; par++) {
    Site x = site_of(L, par, s);
    int cd = dir == 0 ? x.x : (dir == 1 ? x.y : x.z);
    size_t pos = dir == 0 ? (size_t)x.y + (size_t)L.L1 * x.z
                          : (dir == 1 ? (size_t)x.x + (size_t)L.L0 * x.z
                                      : (size_t)x.x + (size_t)L.L0 * x.y);
    double *o = buf + (size_t)cd * P + pos;
    double v[4] = {0, 0, 0, 0};
    if constexpr ((M & M_NH) != 0) {
      Q<double> f = qd(F_at(L, (const T *)L.F, x));
      v[0] = 0.5 * qdot(f, f);
    }
    if constexpr ((M & M_TR) != 0) {
      V3<T> g = G_at(L, (const T *)L.G, x);
      v[1] = 0.5 * ((double)g.x * g.x + (double)g.y * g.y + (double)g.z * g.z);
    }
    if constexpr ((M & M_SG) != 0) {
      double sv = S_at(L, (const T *)L.S, x);
      v[2] = sv;
      v[3] = sv * sv;
    }
#pragma unroll
    for (int k = 0; k < 4; k++) o[(size_t)k * V] = v[k];
  }
}

// prepare_wall (src/z_coord.c:304-376): two phases with an interface at
// z = L[zdir]/2 in the reference's value ranges; the small deviations are
// uniform draws of the lattice's Philox stream (pass WALL_PASS) in the
// order the reference takes them at a site; U(1) angles are left alone
template <typename T, int M>
__global__ void k_wall(LatDesc L, int zdir) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= L.Vh) return;
  int Lz = zdir == 0 ? L.L0 : (zdir == 1 ? L.L1 : L.L2);
  for (int par = 0; par < 2; par++) {
    Site x = site_of(L, par, s);
    int zc = zdir == 0 ? x.x : (zdir == 1 ? x.y : x.z);
    bool lo = zc < 0.5 * Lz;
    Rng r;
    rng_init(r, L.key0, L.key1, WALL_PASS, (uint32_t)(par * L.Vh + s));
    if constexpr ((M & M_NH) != 0) {
      for (int h = 0; h < ((M & M_H2) ? 2 : 1); h++) {
        double a = lo ? 0.2 : 1.2, e = lo ? 0.01 : 0.05;
        double f0 = a + e * rng_u53(r), f1 = e * rng_u53(r), f2 = e * rng_u53(r), f3 = e * rng_u53(r);
        stF((T *)(h ? L.F2 : L.F), L.Vh, par, s, Q<T>{(T)f0, (T)f1, (T)f2, (T)f3});
      }
    }
    if constexpr ((M & M_TR) != 0) {
      double a = lo ? 1.5 : 0.2, e = lo ? 0.05 : 0.01;
      double g0 = a + e * rng_u53(r), g1 = e * rng_u53(r), g2 = e * rng_u53(r);
      stG((T *)L.G, L.Vh, par, s, V3<T>{(T)g0, (T)g1, (T)g2});
    }
    if constexpr ((M & M_SG) != 0) {
      double sv = (lo ? 0.8 : 0.1) + 0.01 * rng_u53(r);
      ((T *)L.S)[(size_t)par * L.Vh + s] = (T)sv;
    }
    for (int d = 0; d < 3; d++) {
      double u = 1.0 - (lo ? 0.03 : 0.4) * rng_u53(r);
      stU<M>((T *)L.U, L.Vh, d, par, s, Q<T>{(T)u, (T)::sqrt(1.0 - u * u), (T)0, (T)0});
    }
  }
}

// ------------------------------------------------------------------------
// blocking (the reference's BLOCKING build, src/blocking.c, su2u1.c:736-895)

// the site at coordinates (x, y, z)
__device__ __forceinline__ Site site_at(const LatDesc &L, int x, int y, int z) {
  Site c;
  c.x = x;
  c.y = y;
  c.z = z;
  c.par = (x + y + z) & 1;
  c.s = (x + L.L0 * (y + L.L1 * z)) >> 1;
  return c;
}

// su2staple_wilson_onedir with dagger = 1 (src/staples.c:197-234)
template <typename T, int M>
__device__ Q<double> staple_onedir_dag(const LatDesc &L, const T *U, const Site &x, int mu, int nu) {
  auto u = [&](const Site &c, int d) { return qd(U_at<M>(L, U, c, d)); };
  Site xp = shift(L, x, mu, +1), xn = shift(L, x, nu, +1), xm = shift(L, x, nu, -1);
  Site xpm = shift(L, xp, nu, -1);
  Q<double> up = qmul(u(xp, nu), qmul(qconj(u(xn, mu)), qconj(u(x, nu))));
  Q<double> lo = qmul(qconj(u(xpm, nu)), qmul(qconj(u(xm, mu)), u(xm, nu)));
  return Q<double>{up.a + lo.a, -up.b + -lo.b, -up.c + -lo.c, -up.d + -lo.d};
}

// smear_fields + make_blocked_fields (src/su2u1.c:752-895, blocking.c:566-596)
// for one level: each blocked site takes the smeared links (blocked
// directions) or the plain links, and the smeared triplet, of the fine site
// at twice its coordinates in the blocked directions
template <typename T, int M>
__global__ void k_block(LatDesc F, LatDesc B, int bmask) {
  int sb = blockIdx.x * blockDim.x + threadIdx.x;
  if (sb >= B.Vh) return;
  const T *U = (const T *)F.U, *G = (const T *)F.G;
  for (int par = 0; par < 2; par++) {
    Site cb = site_of(B, par, sb);
    Site x = site_at(F, (bmask & 1) ? 2 * cb.x : cb.x, (bmask & 2) ? 2 * cb.y : cb.y,
                     (bmask & 4) ? 2 * cb.z : cb.z);
    int paths = 1 + 2 * (__popc(bmask) - 1);
    for (int d = 0; d < 3; d++) {
      Q<double> r = qd(U_at<M>(F, U, x, d));
      if (bmask & (1 << d)) {
        // smear_link (src/su2u1.c:752-797)
        Site xd = shift(F, x, d, +1);
        Q<double> v2 = qd(U_at<M>(F, U, xd, d));
        for (int j = 0; j < 3; j++) {
          if (j == d || !(bmask & (1 << j))) continue;
          r = qadd(r, staple_onedir_dag<T, M>(F, U, x, d, j));
          v2 = qadd(v2, staple_onedir_dag<T, M>(F, U, xd, d, j));
        }
        r = Q<double>{r.a / paths, r.b / paths, r.c / paths, r.d / paths};
        v2 = Q<double>{v2.a / paths, v2.b / paths, v2.c / paths, v2.d / paths};
        r = qmul(r, v2);
        double det = ::sqrt(qdot(r, r));
        r = Q<double>{r.a / det, r.b / det, r.c / det, r.d / det};
      }
      stU<M>((T *)B.U, B.Vh, d, par, sb, Q<T>{(T)r.a, (T)r.b, (T)r.c, (T)r.d});
    }
    // smear_triplet (src/su2u1.c:810-857)
    V3<double> g0 = [&] { V3<T> t = G_at(F, G, x); return V3<double>{t.x, t.y, t.z}; }();
    V3<double> cov{0, 0, 0};
    int sites = 1;
    for (int d = 0; d < 3; d++) {
      if (!(bmask & (1 << d))) continue;
      Site xn = shift(F, x, d, +1), xm = shift(F, x, d, -1);
      V3<T> a = G_at(F, G, xn), b = G_at(F, G, xm);
      V3<double> r1 = rotate(qd(U_at<M>(F, U, x, d)), V3<double>{a.x, a.y, a.z});
      V3<double> r2 = rotate(qconj(qd(U_at<M>(F, U, xm, d))), V3<double>{b.x, b.y, b.z});
      cov.x += r1.x;
      cov.y += r1.y;
      cov.z += r1.z;
      cov.x += r2.x;
      cov.y += r2.y;
      cov.z += r2.z;
      sites += 2;
    }
    stG((T *)B.G, B.Vh, par, sb,
        V3<T>{(T)((g0.x + cov.x) / sites), (T)((g0.y + cov.y) / sites), (T)((g0.z + cov.z) / sites)});
  }
}

// planes[b] = sum of the P values of (value, plane) b, in a fixed order
__global__ __launch_bounds__(BLK) void k_corr_reduce(const double *__restrict__ buf, size_t P,
                                                     double *__restrict__ planes) {
  __shared__ double lds[BLK / 64];
  const double *src = buf + (size_t)blockIdx.x * P;
  double v = 0.0;
  for (size_t i = threadIdx.x; i < P; i += BLK) v += src[i];
  double t = block_sum(v, lds);
  if (threadIdx.x == 0) planes[blockIdx.x] = t;
}

// ------------------------------------------------------------------------
// start-up fields and layout conversion

// setfields (src/init.c:92-112)
template <typename T, int M>
__global__ void k_cold(LatDesc L, double phi0, double sigma0, double singlet0) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= L.Vh) return;
  for (int par = 0; par < 2; par++) {
    for (int d = 0; d < 3; d++) {
      stU<M>((T *)L.U, L.Vh, d, par, s, Q<T>{(T)1, (T)0, (T)0, (T)0});
      if constexpr ((M & M_U1) != 0) ((T *)L.A)[(size_t)(d * 2 + par) * L.Vh + s] = 0;
    }
    if constexpr ((M & M_NH) != 0) stF((T *)L.F, L.Vh, par, s, Q<T>{(T)phi0, (T)0, (T)0, (T)0});
    if constexpr ((M & M_H2) != 0) stF((T *)L.F2, L.Vh, par, s, Q<T>{(T)phi0, (T)0, (T)0, (T)0});
    if constexpr ((M & M_TR) != 0) stG((T *)L.G, L.Vh, par, s, V3<T>{(T)sigma0, (T)0, (T)0});
    if constexpr ((M & M_SG) != 0) ((T *)L.S)[(size_t)par * L.Vh + s] = (T)singlet0;
  }
}

__device__ __forceinline__ void gauss2(Rng &r, double &z0, double &z1) {
  double u1 = rng_u53(r), u2 = rng_u53(r);
  double rr = ::sqrt(-2.0 * ::log(1.0 - u1));
  double t = 2.0 * PI * u2;
  z0 = rr * ::cos(t);
  z1 = rr * ::sin(t);
}

// hot start, the draw order of the oracle's orc_hot_start
template <typename T, int M>
__global__ void k_hot(LatDesc L) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= L.Vh) return;
  for (int par = 0; par < 2; par++) {
    Rng r;
    rng_init(r, L.key0, L.key1, HOTSTART_PASS, (uint32_t)(par * L.Vh + s));
    for (int d = 0; d < 3; d++) {
      double z[4];
      gauss2(r, z[0], z[1]);
      gauss2(r, z[2], z[3]);
      double n = ::sqrt(z[0] * z[0] + z[1] * z[1] + z[2] * z[2] + z[3] * z[3]);
      stU<M>((T *)L.U, L.Vh, d, par, s, Q<T>{(T)(z[0] / n), (T)(z[1] / n), (T)(z[2] / n), (T)(z[3] / n)});
    }
    if constexpr ((M & M_U1) != 0)
      for (int d = 0; d < 3; d++)
        ((T *)L.A)[(size_t)(d * 2 + par) * L.Vh + s] = (T)(PI * (2.0 * rng_u53(r) - 1.0));
    if constexpr ((M & M_NH) != 0) {
      for (int h = 0; h < ((M & M_H2) ? 2 : 1); h++) {
        double z[4];
        gauss2(r, z[0], z[1]);
        gauss2(r, z[2], z[3]);
        stF((T *)(h ? L.F2 : L.F), L.Vh, par, s, Q<T>{(T)z[0], (T)z[1], (T)z[2], (T)z[3]});
      }
    }
    if constexpr ((M & M_TR) != 0) {
      double z[4];
      gauss2(r, z[0], z[1]);
      gauss2(r, z[2], z[3]);
      stG((T *)L.G, L.Vh, par, s, V3<T>{(T)z[0], (T)z[1], (T)z[2]});
    }
    if constexpr ((M & M_SG) != 0) {
      double z0, z1;
      gauss2(r, z0, z1);
      ((T *)L.S)[(size_t)par * L.Vh + s] = (T)z0;
    }
  }
}

// reference host layout (site i = par*Vh + s, components innermost)
// <-> device SoA; field 0 links, 1 U(1), 2 doublet, 3 triplet, 4 singlet,
// 5 second doublet
__host__ __device__ inline int field_dofs(int field) {
  if (field == 5) return 4;
  return field == 0 ? 12 : (field == 1 ? 3 : (field == 2 ? 4 : (field == 3 ? 3 : 1)));
}

__device__ __forceinline__ void *field_ptr(const LatDesc &L, int field) {
  switch (field) {
    case 0: return L.U;
    case 1: return L.A;
    case 2: return L.F;
    case 3: return L.G;
    case 4: return L.S;
    default: return L.F2;
  }
}

// index of component k of a non-link field (links go through ldU / stU,
// whose FP32 form is not one word per component)
__device__ __forceinline__ size_t soa_index(int field, int Vh, int par, int s, int k) {
  if (field == 2 || field == 5) return ((size_t)par * Vh + s) * 4 + k;
  return (size_t)(k * 2 + par) * Vh + s;
}

template <typename T, int M>
__global__ void k_to_device(LatDesc L, int field, const double *__restrict__ h) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= L.Vh) return;
  void *dst = field_ptr(L, field);
  int n = field_dofs(field);
  for (int par = 0; par < 2; par++) {
    size_t i = (size_t)par * L.Vh + s;
    if (field == 0) {
      for (int d = 0; d < 3; d++) {
        const double *q = h + i * 12 + d * 4;
        if constexpr (packed_links<M, T>()) {
          // the packed form holds unit links only: normalise in FP64 first
          double n = 1.0 / ::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
          stU<M>((T *)dst, L.Vh, d, par, s,
                 Q<T>{(T)(q[0] * n), (T)(q[1] * n), (T)(q[2] * n), (T)(q[3] * n)});
        } else {
          // four components: the values as given, rounded to T
          st4((T *)dst + ((size_t)(d * 2 + par) * L.Vh + s) * 4,
              Q<T>{(T)q[0], (T)q[1], (T)q[2], (T)q[3]});
        }
      }
      continue;
    }
    for (int k = 0; k < n; k++) ((T *)dst)[soa_index(field, L.Vh, par, s, k)] = (T)h[i * n + k];
  }
}

template <typename T, int M>
__global__ void k_to_host(LatDesc L, int field, double *__restrict__ h) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= L.Vh) return;
  const void *src = field_ptr(L, field);
  int n = field_dofs(field);
  for (int par = 0; par < 2; par++) {
    size_t i = (size_t)par * L.Vh + s;
    if (field == 0) {
      for (int d = 0; d < 3; d++) {
        Q<T> q = ldU<M>((const T *)src, L.Vh, d, par, s);
        double *o = h + i * 12 + d * 4;
        o[0] = (double)q.a;
        o[1] = (double)q.b;
        o[2] = (double)q.c;
        o[3] = (double)q.d;
      }
      continue;
    }
    for (int k = 0; k < n; k++) h[i * n + k] = (double)((const T *)src)[soa_index(field, L.Vh, par, s, k)];
  }
}

__global__ void k_copy(const double2 *__restrict__ a, double2 *__restrict__ b, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t st = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += st) b[i] = a[i];
}

// ------------------------------------------------------------------------
// gradient flow (src/gradflow.c), links double-buffered, one thread per
// site of both parities

// grad_force_link (src/gradflow.c:190-258): F^a = 2 (4/beta) (U s)_a with
// s = -beta/8 * Wilson staple (+ the triplet's -Sigma(x+d) U^+ Sigma(x))
template <typename T, int M>
__device__ __forceinline__ V3<T> link_force(const LatDesc &L, const T *U, const T *G, const Site &x,
                                           int d) {
  Q<T> st = qscale(wilson_staple<T, M>(L, U, x, d), (T)(-L.beta / 8.0));
  Q<T> u = U_at<M>(L, U, x, d);
  if constexpr ((M & M_TR) != 0) {
    V3<T> g2 = G_at(L, G, shift(L, x, d, +1)), g1 = G_at(L, G, x);
    Q<T> t = qmul(qmul(Q<T>{(T)0, g2.x, g2.y, g2.z}, qconj(u)), Q<T>{(T)0, g1.x, g1.y, g1.z});
    st = qadd(st, qscale(t, (T)0.25));
  }
  Q<T> m = qmul(u, st);
  T k = (T)(4.0 / L.beta);
  return V3<T>{(T)2 * m.b * k, (T)2 * m.c * k, (T)2 * m.d * k};
}

// grad_force_triplet (src/gradflow.c:267-330)
template <typename T, int M>
__device__ __forceinline__ V3<T> triplet_force(const LatDesc &L, const T *U, const T *F, const T *G,
                                              const Site &x, double msq_t) {
  V3<T> res = triplet_staple<T, M>(L, U, G, x);
  V3<T> g = G_at(L, G, x);
  T tsq = (T)0.5 * (g.x * g.x + g.y * g.y + g.z * g.z);
  T k = (T)(2.0 * 3 + msq_t) + (T)2 * (T)L.b4 * tsq;
  if constexpr ((M & M_NH) != 0) {
    Q<T> f = F_at(L, F, x);
    k += (T)L.a2 * (T)0.5 * qdot(f, f);
  }
  return V3<T>{-(res.x + k * g.x), -(res.y + k * g.y), -(res.z + k * g.z)};
}

// exp(i z.sigma) v.  The FP32 flow's links are renormalised by their
// packed storage at every stage: unrenormalised, the link norm would
// random-walk over the 1200 stages of an RK3 flow to t = 4 and the flowed
// 1 - P drift by 1e-2 .. 1e-1.  FP64 keeps the reference's arithmetic
// (src/gradflow.c:361-401).
template <typename T>
__device__ __forceinline__ Q<T> expmul(const V3<T> &z, const Q<T> &v) {
  T mod = z.x * z.x + z.y * z.y + z.z * z.z;
  if (mod <= (T)0) return v;
  mod = msqrt(mod);
  T sn = msin(mod) / mod;
  Q<T> u{mcos(mod), sn * z.x, sn * z.y, sn * z.z};
  return qmul(u, v);
}

// A flow stage runs one thread per link: six workgroups per block-table
// entry, one per (parity, direction).  Measured at 64^3 per RK3 stage:
// FP32 13.3-14.2 us per link against 21.9 us with one thread per site doing
// its six links, FP64 26.8 against 30.0 us.
constexpr int FLOW_GROUPS = 6;

// Work-groups go to the device's NXCD chiplets round-robin (block b to
// XCD b % NXCD), each chiplet with its own L2.  A grid of flow_grid(n)
// blocks covers block-table entries k < n with FLOW_GROUPS blocks each, all
// on XCD k % NXCD (the XCD the block table gives entry k), so that they
// share the entry's neighbourhood in one L2 instead of fetching it into six
// of them (2.7x the algorithmic bytes when consecutive blocks took the six
// groups of one entry).
constexpr unsigned NXCD = 8;

constexpr unsigned flow_grid(unsigned n) {
  return FLOW_GROUPS * NXCD * ((n + NXCD - 1) / NXCD);
}

// stage 0: Euler (flow_gauge, src/gradflow.c:361-401); stages 1..3: RK3.
// Group g of a block-table entry is parity g / 3, direction g % 3.  The
// links of a stage are independent (each reads Uin and its own Z, writes
// Uout and its own Z), so the split does not change a bit; the
// direction-0 links also step the triplet.
template <typename T, int M>
__global__ __launch_bounds__(BLK) void k_flow(const LatDesc *__restrict__ lats,
                                              const BlockMap *__restrict__ blk, int nblk,
                                              const T *__restrict__ Uin,
                                              T *__restrict__ Uout, const T *__restrict__ Gin,
                                              T *__restrict__ Gout, T *__restrict__ Z,
                                              T *__restrict__ ZG, double dt, double msq_t, int stage) {
  const unsigned xcd = blockIdx.x % NXCD, i = blockIdx.x / NXCD;
  const unsigned k = (i / FLOW_GROUPS) * NXCD + xcd;
  const int g = (int)(i % FLOW_GROUPS);
  if (k >= (unsigned)nblk) return;
  const BlockMap bm = blk[k];
  const LatDesc &L = lats[bm.lat];
  int s = bm.s0 + (int)threadIdx.x;
  if (s >= L.Vh) return;
  const double cs[4] = {1.0, 0.25, 8.0 / 9.0, 0.75};
  const double ds[4] = {0.0, 0.0, 17.0 / 9.0, 1.0};
  T c = (T)(cs[stage] * dt), dd = (T)ds[stage];
  const int p0 = g / 3, p1 = p0 + 1;
  const int d0 = g % 3, d1 = d0 + 1;
  for (int par = p0; par < p1; par++) {
    Site x = site_of(L, par, s);
    for (int d = d0; d < d1; d++) {
      V3<T> f = link_force<T, M>(L, Uin, Gin, x, d);
      V3<T> z{c * f.x, c * f.y, c * f.z};
      if (stage > 0) {
        size_t o = (size_t)((d * 3) * 2 + par) * L.Vh + s, st = (size_t)2 * L.Vh;
        if (stage > 1) {
          z.x -= dd * Z[o];
          z.y -= dd * Z[o + st];
          z.z -= dd * Z[o + 2 * st];
        }
        Z[o] = z.x;
        Z[o + st] = z.y;
        Z[o + 2 * st] = z.z;
      }
      stU<M>(Uout, L.Vh, d, par, s, expmul(z, ldU<M>(Uin, L.Vh, d, par, s)));
    }
    if constexpr ((M & M_TR) != 0) {
      if (d0 == 0) {
        V3<T> f = triplet_force<T, M>(L, Uin, (const T *)L.F, Gin, x, msq_t);
        V3<T> z{c * f.x, c * f.y, c * f.z};
        if (stage > 0) {
          size_t o = (size_t)par * L.Vh + s, st = (size_t)2 * L.Vh;
          if (stage > 1) {
            z.x -= dd * ZG[o];
            z.y -= dd * ZG[o + st];
            z.z -= dd * ZG[o + 2 * st];
          }
          ZG[o] = z.x;
          ZG[o + st] = z.y;
          ZG[o + 2 * st] = z.z;
        }
        V3<T> gg = ldG(Gin, L.Vh, par, s);
        stG(Gout, L.Vh, par, s, V3<T>{gg.x + z.x, gg.y + z.y, gg.z + z.z});
      }
    }
  }
}

template <typename T, int M>
__global__ __launch_bounds__(BLK) void k_flow_force(const LatDesc *__restrict__ lats,
                                                    const BlockMap *__restrict__ blk,
                                                    double *__restrict__ out) {
  const BlockMap bm = blk[blockIdx.x];
  const LatDesc &L = lats[bm.lat];
  int s = bm.s0 + (int)threadIdx.x;
  if (s >= L.Vh) return;
  for (int par = 0; par < 2; par++) {
    Site x = site_of(L, par, s);
    size_t i = (size_t)par * L.Vh + s;
    for (int d = 0; d < 3; d++) {
      V3<T> f = link_force<T, M>(L, (const T *)L.U, (const T *)L.G, x, d);
      double *o = out + (i * 3 + d) * 4;
      o[0] = 0.0;
      o[1] = f.x;
      o[2] = f.y;
      o[3] = f.z;
    }
  }
}

// ------------------------------------------------------------------------
// dispatch

template <typename Fn>
static hipError_t with_model(int model, Fn &&fn) {
  switch (model) {
    case 0: return fn(std::integral_constant<int, 0>{});
    case M_NH: return fn(std::integral_constant<int, M_NH>{});
    case M_NH | M_U1: return fn(std::integral_constant<int, M_NH | M_U1>{});
    case M_NH | M_SG: return fn(std::integral_constant<int, M_NH | M_SG>{});
    case M_NH | M_U1 | M_SG: return fn(std::integral_constant<int, M_NH | M_U1 | M_SG>{});
    case M_TR: return fn(std::integral_constant<int, M_TR>{});
    case M_NH | M_TR: return fn(std::integral_constant<int, M_NH | M_TR>{});
    case M_NH | M_H2: return fn(std::integral_constant<int, M_NH | M_H2>{});
    case M_NH | M_H2 | M_U1: return fn(std::integral_constant<int, M_NH | M_H2 | M_U1>{});
  }
  return hipErrorInvalidValue;
}

// The model bits of a launch in precision T: FP32 handles may add M_PK
// (packed links), FP64 never does (no FP64 kernel is built with it).
template <typename T, typename Fn>
static hipError_t with_model_of(int model, Fn &&fn) {
  if constexpr (sizeof(T) == 4) {
    if (model & M_PK)
      return with_model(model & ~M_PK, [&](auto m) -> hipError_t {
        return fn(std::integral_constant<int, decltype(m)::value | M_PK>{});
      });
  }
  if (model & M_PK) return hipErrorInvalidValue;
  return with_model(model, fn);
}

template <typename Fn>
static hipError_t with_prec(int fp64, Fn &&fn) {
  if (fp64) return fn(double(0));
  return fn(float(0));
}

template <typename T, int M, int KIND>
static hipError_t pass_one(const LaunchCtx &c, int parity, int dir, int pass_off, int step,
                           int reunit) {
  if constexpr (kind_valid<M, KIND>()) {
    // multicanonical steps use the 256-site table: their order-parameter
    // partials are indexed by 256-site chunk
    if (step >= 0 && c.block != BLK) return hipErrorInvalidValue;
    if constexpr (KIND == SU2HIP_K_LINK_OVERRELAX) {
      if (c.waves == 5) {
        if (c.ev_start || c.ev_stop) {
          hipExtLaunchKernelGGL((k_pass_w5<T, M, KIND>), dim3(c.nblk), dim3(c.block), 0, c.stream,
                                c.ev_start, c.ev_stop, 0, c.desc, c.muca, c.blk, parity, dir,
                                pass_off, step, reunit, c.cnt, c.nblk, c.partial, c.order, c.npass);
        } else {
          hipLaunchKernelGGL((k_pass_w5<T, M, KIND>), dim3(c.nblk), dim3(c.block), 0, c.stream,
                             c.desc, c.muca, c.blk, parity, dir, pass_off, step, reunit, c.cnt,
                             c.nblk, c.partial, c.order, c.npass);
        }
        return hipGetLastError();
      }
    }
    if (c.ev_start || c.ev_stop) {
      hipExtLaunchKernelGGL((k_pass<T, M, KIND>), dim3(c.nblk), dim3(c.block), 0, c.stream,
                            c.ev_start, c.ev_stop, 0, c.desc, c.muca, c.blk, parity, dir, pass_off,
                            step, reunit, c.cnt, c.nblk, c.partial, c.order, c.npass);
    } else {
      hipLaunchKernelGGL((k_pass<T, M, KIND>), dim3(c.nblk), dim3(c.block), 0, c.stream, c.desc,
                         c.muca, c.blk, parity, dir, pass_off, step, reunit, c.cnt, c.nblk,
                         c.partial, c.order, c.npass);
    }
    return hipGetLastError();
  }
  return hipErrorInvalidValue;
}

hipError_t launch_pass(const LaunchCtx &c, int kind, int parity, int dir, int pass_off,
                       int muca_step, int reunit) {
  if (c.nblk == 0) return hipSuccess;
  return with_prec(c.fp64, [&](auto t) {
    using T = decltype(t);
    return with_model_of<T>(c.model, [&](auto m) -> hipError_t {
      constexpr int M = decltype(m)::value;
      switch (kind) {
        case SU2HIP_K_LINK_HEATBATH: return pass_one<T, M, SU2HIP_K_LINK_HEATBATH>(c, parity, dir, pass_off, muca_step, reunit);
        case SU2HIP_K_LINK_METROPOLIS: return pass_one<T, M, SU2HIP_K_LINK_METROPOLIS>(c, parity, dir, pass_off, muca_step, reunit);
        case SU2HIP_K_LINK_OVERRELAX: return pass_one<T, M, SU2HIP_K_LINK_OVERRELAX>(c, parity, dir, pass_off, muca_step, reunit);
        case SU2HIP_K_U1_METROPOLIS: return pass_one<T, M, SU2HIP_K_U1_METROPOLIS>(c, parity, dir, pass_off, muca_step, reunit);
        case SU2HIP_K_DOUBLET_OVERRELAX: return pass_one<T, M, SU2HIP_K_DOUBLET_OVERRELAX>(c, parity, dir, pass_off, muca_step, reunit);
        case SU2HIP_K_DOUBLET_METROPOLIS: return pass_one<T, M, SU2HIP_K_DOUBLET_METROPOLIS>(c, parity, dir, pass_off, muca_step, reunit);
        case SU2HIP_K_TRIPLET_OVERRELAX: return pass_one<T, M, SU2HIP_K_TRIPLET_OVERRELAX>(c, parity, dir, pass_off, muca_step, reunit);
        case SU2HIP_K_TRIPLET_METROPOLIS: return pass_one<T, M, SU2HIP_K_TRIPLET_METROPOLIS>(c, parity, dir, pass_off, muca_step, reunit);
        case SU2HIP_K_SINGLET_OVERRELAX: return pass_one<T, M, SU2HIP_K_SINGLET_OVERRELAX>(c, parity, dir, pass_off, muca_step, reunit);
        case SU2HIP_K_SINGLET_METROPOLIS: return pass_one<T, M, SU2HIP_K_SINGLET_METROPOLIS>(c, parity, dir, pass_off, muca_step, reunit);
      }
      return hipErrorInvalidValue;
    });
  });
}

hipError_t launch_hop(const LaunchCtx &c) {
  if (c.nblk == 0) return hipSuccess;
  return with_prec(c.fp64, [&](auto t) {
    using T = decltype(t);
    return with_model_of<T>(c.model, [&](auto m) -> hipError_t {
      constexpr int M = decltype(m)::value;
      if constexpr (hop_cached<M>()) {
        hipLaunchKernelGGL((k_hop<T, M>), dim3(c.nblk), dim3(BLK), 0, c.stream, c.desc, c.blk);
        return hipGetLastError();
      }
      return hipSuccess;
    });
  });
}

hipError_t launch_muca_accept(const LaunchCtx &c, int kind, int hid, int parity, int pass_off,
                              int muca_step) {
  hipLaunchKernelGGL(k_muca_accept, dim3(c.max_slots), dim3(BLK), 0, c.stream, c.desc, c.muca,
                     c.partial, kind, hid, parity, pass_off, muca_step, c.order, c.npass);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || c.nblk == 0) return e;
  return with_prec(c.fp64, [&](auto t) {
    using T = decltype(t);
    hipLaunchKernelGGL(k_muca_restore<T>, dim3(c.nblk), dim3(BLK), 0, c.stream, c.desc, c.muca,
                       c.blk, kind, hid, parity, pass_off, muca_step, c.order, c.npass);
    return hipGetLastError();
  });
}

hipError_t launch_orderparam_init(const LaunchCtx &c, int only) {
  if (c.nblk == 0) return hipSuccess;
  hipError_t e = with_prec(c.fp64, [&](auto t) {
    using T = decltype(t);
    return with_model_of<T>(c.model, [&](auto m) -> hipError_t {
      constexpr int M = decltype(m)::value;
      hipLaunchKernelGGL((k_orderparam<T, M>), dim3(c.nblk), dim3(BLK), 0, c.stream, c.desc,
                         c.muca, c.blk, c.nblk, c.partial, only);
      return hipGetLastError();
    });
  });
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_orderparam_final, dim3(c.max_slots), dim3(BLK), 0, c.stream, c.desc,
                     c.muca, c.nblk, c.partial, only);
  return hipGetLastError();
}

hipError_t launch_order(const LaunchCtx &c, uint8_t *order, const uint16_t *info, int npass,
                        int link_rounds, int scalar_rounds) {
  int nb = (c.max_slots + 63) / 64;
  hipLaunchKernelGGL(k_order, dim3(nb), dim3(64), 0, c.stream, c.desc, c.max_slots, order, info,
                     npass, link_rounds, scalar_rounds);
  return hipGetLastError();
}

hipError_t launch_end_sweep(const LaunchCtx &c, int passes) {
  int nb = (c.max_slots + 63) / 64;
  hipLaunchKernelGGL(k_end_sweep, dim3(nb), dim3(64), 0, c.stream, c.desc, c.max_slots, passes);
  return hipGetLastError();
}

hipError_t launch_measure(const LaunchCtx &c) {
  if (c.nblk > 0) {
    hipError_t e = with_prec(c.fp64, [&](auto t) {
      using T = decltype(t);
      return with_model_of<T>(c.model, [&](auto m) -> hipError_t {
        constexpr int M = decltype(m)::value;
        hipLaunchKernelGGL((k_measure<T, M>), dim3(c.nblk), dim3(BLK), 0, c.stream, c.desc, c.blk,
                           c.partial);
        if constexpr ((M & M_TR) != 0)
          hipLaunchKernelGGL((k_magcharge<T, M>), dim3(c.nblk), dim3(BLK), 0, c.stream, c.desc,
                             c.blk, c.partial);
        return hipGetLastError();
      });
    });
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(k_measure_final, dim3(c.max_slots), dim3(BLK), 0, c.stream, c.desc,
                     c.partial, c.sums);
  return hipGetLastError();
}

// One lattice measured without the handle's block table: block k sums
// sites k*BLK .. k*BLK+BLK-1 of both parities into partial[k], the same
// blocks and order as the lattice's rows of the table, so the sums are
// bitwise those of k_measure + k_measure_final.
template <typename T, int M>
__global__ __launch_bounds__(BLK) void k_measure_one(LatDesc L, double *__restrict__ partial) {
  __shared__ double lds[BLK / 64 * SU2HIP_NSUMS];
  int s = (int)blockIdx.x * BLK + (int)threadIdx.x;
  double acc[SU2HIP_NSUMS];
#pragma unroll
  for (int j = 0; j < SU2HIP_NSUMS; j++) acc[j] = 0.0;
  if (s < L.Vh) {
#pragma unroll 1
    for (int par = 0; par < 2; par++) measure_site<T, M>(L, par, s, acc);
  }
  double t = block_sums<SU2HIP_NSUMS>(acc, lds);
  if (threadIdx.x < SU2HIP_NSUMS) partial[(size_t)blockIdx.x * SU2HIP_NSUMS + threadIdx.x] = t;
}

template <typename T, int M>
__global__ __launch_bounds__(BLK) void k_magcharge_one(LatDesc L, double *__restrict__ partial) {
  __shared__ double lds[BLK / 64 * 2];
  int s = (int)blockIdx.x * BLK + (int)threadIdx.x;
  double acc[2] = {0.0, 0.0};
  if (s < L.Vh) {
#pragma unroll 1
    for (int par = 0; par < 2; par++) {
      double ch = magcharge_site<T, M>(L, par, s);
      acc[0] += ch;
      acc[1] += ::fabs(ch);
    }
  }
  double t = block_sums<2>(acc, lds);
  if (threadIdx.x < 2)
    partial[(size_t)blockIdx.x * SU2HIP_NSUMS + SU2HIP_SUM_MAGCHARGE + threadIdx.x] = t;
}

__global__ __launch_bounds__(BLK) void k_measure_one_final(const double *__restrict__ partial,
                                                           int nblk, double *__restrict__ sums) {
  __shared__ double lds[BLK / 64 * SU2HIP_NSUMS];
  double v[SU2HIP_NSUMS];
#pragma unroll
  for (int j = 0; j < SU2HIP_NSUMS; j++) v[j] = 0.0;
  for (int i = threadIdx.x; i < nblk; i += BLK) {
    const double *row = partial + (size_t)i * SU2HIP_NSUMS;
#pragma unroll
    for (int j = 0; j < SU2HIP_NSUMS; j++) v[j] += row[j];
  }
  double t = block_sums<SU2HIP_NSUMS>(v, lds);
  if (threadIdx.x < SU2HIP_NSUMS) sums[threadIdx.x] = t;
}

hipError_t launch_measure_one(hipStream_t stream, int fp64, int model, const LatDesc &L,
                              double *partial, double *sums) {
  int nb = (L.Vh + BLK - 1) / BLK;
  hipError_t e = with_prec(fp64, [&](auto t) {
    using T = decltype(t);
    return with_model_of<T>(model, [&](auto m) -> hipError_t {
      constexpr int M = decltype(m)::value;
      hipLaunchKernelGGL((k_measure_one<T, M>), dim3(nb), dim3(BLK), 0, stream, L, partial);
      if constexpr ((M & M_TR) != 0)
        hipLaunchKernelGGL((k_magcharge_one<T, M>), dim3(nb), dim3(BLK), 0, stream, L, partial);
      return hipGetLastError();
    });
  });
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_measure_one_final, dim3(1), dim3(BLK), 0, stream, (const double *)partial, nb,
                     sums);
  return hipGetLastError();
}

hipError_t launch_corr(const LaunchCtx &c, int slot, int dir, int Vh, int Ld, size_t P,
                       double *buf, double *planes) {
  int nb = (Vh + BLK - 1) / BLK;
  hipError_t e = with_prec(c.fp64, [&](auto t) {
    using T = decltype(t);
    return with_model_of<T>(c.model, [&](auto m) -> hipError_t {
      constexpr int M = decltype(m)::value;
      hipLaunchKernelGGL((k_corr_sites<T, M>), dim3(nb), dim3(BLK), 0, c.stream, c.desc, slot, dir,
                         buf);
      return hipGetLastError();
    });
  });
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_corr_reduce, dim3(10 * Ld), dim3(BLK), 0, c.stream, (const double *)buf, P,
                     planes);
  return hipGetLastError();
}

hipError_t launch_block(const LaunchCtx &c, const LatDesc &fine, const LatDesc &blocked, int bmask) {
  int nb = (blocked.Vh + BLK - 1) / BLK;
  return with_prec(c.fp64, [&](auto t) {
    using T = decltype(t);
    return with_model_of<T>(c.model, [&](auto m) -> hipError_t {
      constexpr int M = decltype(m)::value;
      hipLaunchKernelGGL((k_block<T, M>), dim3(nb), dim3(BLK), 0, c.stream, fine, blocked, bmask);
      return hipGetLastError();
    });
  });
}

hipError_t launch_zprof(const LaunchCtx &c, int slot, int dir, int Vh, int Ld, size_t P,
                        double *buf, double *planes) {
  int nb = (Vh + BLK - 1) / BLK;
  hipError_t e = with_prec(c.fp64, [&](auto t) {
    using T = decltype(t);
    return with_model_of<T>(c.model, [&](auto m) -> hipError_t {
      constexpr int M = decltype(m)::value;
      hipLaunchKernelGGL((k_zprof_sites<T, M>), dim3(nb), dim3(BLK), 0, c.stream, c.desc, slot, dir,
                         buf);
      return hipGetLastError();
    });
  });
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_corr_reduce, dim3(4 * Ld), dim3(BLK), 0, c.stream, (const double *)buf, P,
                     planes);
  return hipGetLastError();
}

static LatDesc host_desc_of(const LaunchCtx &c, int slot, hipError_t &e) {
  LatDesc L;
  e = hipMemcpyAsync(&L, c.desc + slot, sizeof(LatDesc), hipMemcpyDeviceToHost, c.stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c.stream);
  return L;
}

hipError_t launch_cold(const LaunchCtx &c, int slot, double phi0, double sigma0, double singlet0) {
  hipError_t e;
  LatDesc L = host_desc_of(c, slot, e);
  if (e != hipSuccess) return e;
  int nb = (L.Vh + BLK - 1) / BLK;
  return with_prec(c.fp64, [&](auto t) {
    using T = decltype(t);
    return with_model_of<T>(c.model, [&](auto m) -> hipError_t {
      constexpr int M = decltype(m)::value;
      hipLaunchKernelGGL((k_cold<T, M>), dim3(nb), dim3(BLK), 0, c.stream, L, phi0, sigma0, singlet0);
      return hipGetLastError();
    });
  });
}

hipError_t launch_hot(const LaunchCtx &c, int slot) {
  hipError_t e;
  LatDesc L = host_desc_of(c, slot, e);
  if (e != hipSuccess) return e;
  int nb = (L.Vh + BLK - 1) / BLK;
  return with_prec(c.fp64, [&](auto t) {
    using T = decltype(t);
    return with_model_of<T>(c.model, [&](auto m) -> hipError_t {
      constexpr int M = decltype(m)::value;
      hipLaunchKernelGGL((k_hot<T, M>), dim3(nb), dim3(BLK), 0, c.stream, L);
      return hipGetLastError();
    });
  });
}

hipError_t launch_wall(const LaunchCtx &c, int slot, int zdir) {
  hipError_t e;
  LatDesc L = host_desc_of(c, slot, e);
  if (e != hipSuccess) return e;
  int nb = (L.Vh + BLK - 1) / BLK;
  return with_prec(c.fp64, [&](auto t) {
    using T = decltype(t);
    return with_model_of<T>(c.model, [&](auto m) -> hipError_t {
      constexpr int M = decltype(m)::value;
      hipLaunchKernelGGL((k_wall<T, M>), dim3(nb), dim3(BLK), 0, c.stream, L, zdir);
      return hipGetLastError();
    });
  });
}

hipError_t launch_to_device_layout(const LaunchCtx &c, int slot, int field, const double *h) {
  hipError_t e;
  LatDesc L = host_desc_of(c, slot, e);
  if (e != hipSuccess) return e;
  int nb = (L.Vh + BLK - 1) / BLK;
  return with_prec(c.fp64, [&](auto t) {
    using T = decltype(t);
    return with_model_of<T>(c.model, [&](auto m) -> hipError_t {
      constexpr int M = decltype(m)::value;
      hipLaunchKernelGGL((k_to_device<T, M>), dim3(nb), dim3(BLK), 0, c.stream, L, field, h);
      return hipGetLastError();
    });
  });
}

hipError_t launch_to_host_layout(const LaunchCtx &c, int slot, int field, double *h) {
  hipError_t e;
  LatDesc L = host_desc_of(c, slot, e);
  if (e != hipSuccess) return e;
  int nb = (L.Vh + BLK - 1) / BLK;
  return with_prec(c.fp64, [&](auto t) {
    using T = decltype(t);
    return with_model_of<T>(c.model, [&](auto m) -> hipError_t {
      constexpr int M = decltype(m)::value;
      hipLaunchKernelGGL((k_to_host<T, M>), dim3(nb), dim3(BLK), 0, c.stream, L, field, h);
      return hipGetLastError();
    });
  });
}

hipError_t launch_copy_probe(const double *src, double *dst, size_t n, hipStream_t s) {
  size_t n2 = n / 2;
  int nb = 256 * 16;
  hipLaunchKernelGGL(k_copy, dim3(nb), dim3(256), 0, s, (const double2 *)src, (double2 *)dst, n2);
  return hipGetLastError();
}

hipError_t launch_flow_stage(const LaunchCtx &c, int slot, const void *Uin, void *Uout,
                             const void *Gin, void *Gout, void *Z, void *ZG, double dt,
                             double msq_t, int stage) {
  (void)slot;
  if (c.nblk == 0) return hipSuccess;
  return with_prec(c.fp64, [&](auto t) {
    using T = decltype(t);
    return with_model_of<T>(c.model, [&](auto m) -> hipError_t {
      constexpr int M = decltype(m)::value;
      hipLaunchKernelGGL((k_flow<T, M>), dim3(flow_grid(c.nblk)), dim3(BLK), 0, c.stream,
                         c.desc, c.blk, c.nblk, (const T *)Uin, (T *)Uout, (const T *)Gin, (T *)Gout, (T *)Z, (T *)ZG, dt,
                         msq_t, stage);
      return hipGetLastError();
    });
  });
}

hipError_t launch_flow_force(const LaunchCtx &c, int slot, double *force) {
  (void)slot;
  if (c.nblk == 0) return hipSuccess;
  return with_prec(c.fp64, [&](auto t) {
    using T = decltype(t);
    return with_model_of<T>(c.model, [&](auto m) -> hipError_t {
      constexpr int M = decltype(m)::value;
      hipLaunchKernelGGL((k_flow_force<T, M>), dim3(c.nblk), dim3(BLK), 0, c.stream, c.desc, c.blk,
                         force);
      return hipGetLastError();
    });
  });
}

}  // namespace su2
