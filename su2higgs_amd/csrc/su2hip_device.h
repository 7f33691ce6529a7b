// su2hip_device.h -- device-side algebra, RNG and geometry for the sweep
// kernels (gfx950).  Included by su2hip_kernels.hip only.
#pragma once

#include "su2hip_internal.h"

namespace su2 {

// ------------------------------------------------------------------------
// Philox4x32-10 (Salmon et al., SC'11).  Counter = (block, site, pass_lo,
// pass_hi), key = lattice seed.  FP64 draws take 64 bits per uniform (two
// per block), FP32 draws 32 bits (four per block).

__device__ __forceinline__ void philox10(uint32_t &c0, uint32_t &c1, uint32_t &c2,
                                         uint32_t &c3, uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; r++) {
    if (r) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    // one 32x32 -> 64-bit multiply each (v_mad_u64_u32), high and low halves
    uint64_t p0 = (uint64_t)0xD2511F53u * c0, p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
    c0 = n0;
    c1 = lo1;
    c2 = n2;
    c3 = lo0;
  }
}

struct Rng {
  uint32_t k0, k1, site, plo, phi;
  uint32_t k, blk;
  uint32_t w0, w1, w2, w3;
};

__device__ __forceinline__ void rng_init(Rng &r, uint32_t k0, uint32_t k1,
                                         uint64_t pass, uint32_t site) {
  r.k0 = k0;
  r.k1 = k1;
  r.site = site;
  r.plo = (uint32_t)pass;
  r.phi = (uint32_t)(pass >> 32);
  r.k = 0;
  r.blk = 0xFFFFFFFFu;
}

__device__ __forceinline__ void rng_block(Rng &r, uint32_t b) {
  if (b != r.blk) {
    uint32_t c0 = b, c1 = r.site, c2 = r.plo, c3 = r.phi;
    philox10(c0, c1, c2, c3, r.k0, r.k1);
    r.w0 = c0;
    r.w1 = c1;
    r.w2 = c2;
    r.w3 = c3;
    r.blk = b;
  }
}

__device__ __forceinline__ double rng_u53(Rng &r) {
  uint32_t k = r.k++;
  rng_block(r, k >> 1);
  uint64_t x = (k & 1) ? (((uint64_t)r.w3 << 32) | r.w2) : (((uint64_t)r.w1 << 32) | r.w0);
  return (double)(x >> 11) * 0x1.0p-53;
}

// the 64-bit word of the next draw (the reference's iran(); rng_u53 is its
// top 53 bits)
__device__ __forceinline__ uint64_t rng_u64(Rng &r) {
  uint32_t k = r.k++;
  rng_block(r, k >> 1);
  return (k & 1) ? (((uint64_t)r.w3 << 32) | r.w2) : (((uint64_t)r.w1 << 32) | r.w0);
}

__device__ __forceinline__ float rng_u24(Rng &r) {
  uint32_t k = r.k++;
  rng_block(r, k >> 2);
  uint32_t j = k & 3;
  uint32_t w = j == 0 ? r.w0 : (j == 1 ? r.w1 : (j == 2 ? r.w2 : r.w3));
  return (float)(w >> 8) * 0x1.0p-24f;
}

template <typename T>
__device__ __forceinline__ T uni(Rng &r);
template <>
__device__ __forceinline__ double uni<double>(Rng &r) {
  return rng_u53(r);
}
template <>
__device__ __forceinline__ float uni<float>(Rng &r) {
  return rng_u24(r);
}

// ------------------------------------------------------------------------
// quaternion algebra, U = u0 + i u.sigma (product rule of the reference's
// su2rot, src/su2u1.c:60-70)

template <typename T>
struct Q {
  T a, b, c, d;
};

template <typename T>
__device__ __forceinline__ Q<T> qmul(const Q<T> &x, const Q<T> &y) {
  Q<T> r;
  r.a = x.a * y.a - x.b * y.b - x.c * y.c - x.d * y.d;
  r.b = x.a * y.b + x.b * y.a - (x.c * y.d - x.d * y.c);
  r.c = x.a * y.c + x.c * y.a - (x.d * y.b - x.b * y.d);
  r.d = x.a * y.d + x.d * y.a - (x.b * y.c - x.c * y.b);
  return r;
}
template <typename T>
__device__ __forceinline__ Q<T> qconj(const Q<T> &x) {
  return Q<T>{x.a, -x.b, -x.c, -x.d};
}
template <typename T>
__device__ __forceinline__ Q<T> qadd(const Q<T> &x, const Q<T> &y) {
  return Q<T>{x.a + y.a, x.b + y.b, x.c + y.c, x.d + y.d};
}
template <typename T>
__device__ __forceinline__ Q<T> qscale(const Q<T> &x, T s) {
  return Q<T>{x.a * s, x.b * s, x.c * s, x.d * s};
}
template <typename T>
__device__ __forceinline__ T qdot(const Q<T> &x, const Q<T> &y) {
  return x.a * y.a + x.b * y.b + x.c * y.c + x.d * y.d;
}
// exp(-i alpha sigma_3) as a quaternion (src/staples.c:161-177)
template <typename T>
__device__ __forceinline__ Q<T> u1twist(T alpha) {
  T s, c;
  if constexpr (sizeof(T) == 8) {
    sincos(alpha, &s, &c);
  } else {
    // FP32 sweeps: the hardware sine/cosine on [-1/2, 1/2] revolutions
    float t = alpha * 0.159154943091895336f;
    t -= rintf(t);
    s = __builtin_amdgcn_sinf(t);
    c = __builtin_amdgcn_cosf(t);
  }
  return Q<T>{c, (T)0, (T)0, -s};
}
// x times a quaternion (c, 0, 0, d): the U(1) twist (d = -sin alpha) or its
// conjugate (d = sin alpha), the product rule of qmul with the two zero
// components dropped (the same value, half the multiplies)
template <typename T>
__device__ __forceinline__ Q<T> qmul_u1(const Q<T> &x, const Q<T> &u) {
  const T c = u.a, d = u.d;
  Q<T> r;
  r.a = x.a * c - x.d * d;
  r.b = x.b * c - x.c * d;
  r.c = x.c * c + x.b * d;
  r.d = x.a * d + x.d * c;
  return r;
}

template <typename T>
struct V3 {
  T x, y, z;
};
// R(U) v = vector part of U (0,v) U^dagger
template <typename T>
__device__ __forceinline__ V3<T> rotate(const Q<T> &u, const V3<T> &v) {
  Q<T> q{(T)0, v.x, v.y, v.z};
  Q<T> t = qmul(qmul(u, q), qconj(u));
  return V3<T>{t.b, t.c, t.d};
}

// ------------------------------------------------------------------------
// geometry: a site carries its coordinates, parity and index in its parity;
// unit steps update the index incrementally (periodic wrap)

struct Site {
  int x, y, z;
  int par, s;
};

__device__ __forceinline__ Site site_of(const LatDesc &L, int par, int s) {
  uint32_t r = fdiv((uint32_t)s, L.div_h0);
  int xh = s - (int)r * L.h0;
  uint32_t z = fdiv(r, L.div_L1);
  int y = (int)r - (int)z * L.L1;
  Site c;
  c.y = y;
  c.z = (int)z;
  c.x = 2 * xh + ((par + y + (int)z) & 1);
  c.par = par;
  c.s = s;
  return c;
}

__device__ __forceinline__ Site shift(const LatDesc &L, const Site &c, int d, int sign) {
  Site r = c;
  r.par = c.par ^ 1;
  if (d == 0) {
    int nx = c.x + sign;
    nx = nx < 0 ? L.L0 - 1 : (nx >= L.L0 ? 0 : nx);
    r.x = nx;
    r.s = c.s - (c.x >> 1) + (nx >> 1);
  } else if (d == 1) {
    if (sign > 0) {
      bool w = c.y + 1 == L.L1;
      r.y = w ? 0 : c.y + 1;
      r.s = w ? c.s - L.h0 * (L.L1 - 1) : c.s + L.h0;
    } else {
      bool w = c.y == 0;
      r.y = w ? L.L1 - 1 : c.y - 1;
      r.s = w ? c.s + L.h0 * (L.L1 - 1) : c.s - L.h0;
    }
  } else {
    int plane = L.h0 * L.L1;
    if (sign > 0) {
      bool w = c.z + 1 == L.L2;
      r.z = w ? 0 : c.z + 1;
      r.s = w ? c.s - plane * (L.L2 - 1) : c.s + plane;
    } else {
      bool w = c.z == 0;
      r.z = w ? L.L2 - 1 : c.z - 1;
      r.s = w ? c.s + plane * (L.L2 - 1) : c.s - plane;
    }
  }
  return r;
}

// ------------------------------------------------------------------------
// field access (layout in su2hip_internal.h): links and doublet as
// 4-component site vectors (one 16-byte load per site in FP32), the others
// component-major

// Field arrays live in device global memory.  Their pointers reach the
// kernels through LatDesc (generic pointers), and a generic-to-global cast is
// folded away as a no-op, which leaves FLAT loads that count against both the
// vector-memory and the LDS/scalar wait counters.  Going through the integer
// value makes the loads and stores GLOBAL ones (SGPR base + VGPR offset).
template <typename T>
using gptr = __attribute__((address_space(1))) T *;
typedef float vf4 __attribute__((ext_vector_type(4)));
typedef double vd2 __attribute__((ext_vector_type(2)));
template <typename T>
__device__ __forceinline__ gptr<T> gmem(T *p) {
  return reinterpret_cast<gptr<T>>(reinterpret_cast<unsigned long>(p));
}

template <typename T>
__device__ __forceinline__ Q<T> ld4(const T *p);
template <>
__device__ __forceinline__ Q<float> ld4<float>(const float *p) {
  vf4 v = *gmem(reinterpret_cast<const vf4 *>(p));
  return Q<float>{v.x, v.y, v.z, v.w};
}
template <>
__device__ __forceinline__ Q<double> ld4<double>(const double *p) {
  vd2 a = *gmem(reinterpret_cast<const vd2 *>(p));
  vd2 b = *gmem(reinterpret_cast<const vd2 *>(p + 2));
  return Q<double>{a.x, a.y, b.x, b.y};
}
template <typename T>
__device__ __forceinline__ void st4(T *p, const Q<T> &q);
template <>
__device__ __forceinline__ void st4<float>(float *p, const Q<float> &q) {
  *gmem(reinterpret_cast<vf4 *>(p)) = vf4{q.a, q.b, q.c, q.d};
}
template <>
__device__ __forceinline__ void st4<double>(double *p, const Q<double> &q) {
  *gmem(reinterpret_cast<vd2 *>(p)) = vd2{q.a, q.b};
  *gmem(reinterpret_cast<vd2 *>(p + 2)) = vd2{q.c, q.d};
}

// FP32 links are stored in 12 bytes instead of 16, by stereographic
// projection from the pole opposite the real part's sign: a unit
// quaternion (a, b, c, d) is kept as v = 2 w, w = (b, c, d) / (1 + |a|),
// |w| <= 1, with the sign of a in the lowest mantissa bit of v.x (which
// moves it by at most one unit in the last place).  It comes back as
// a = +-(1 - s) / (1 + s), (b, c, d) = 2 w / (1 + s) = v / (1 + s) with
// s = |w|^2 = |v|^2 / 4 in [0, 1]: one reciprocal of a number in [1, 2]
// (the hardware one and a Newton step), no square root and no case by case
// placement, and unit in exact arithmetic for every v, so a loaded link is
// unit to FP32 rounding whatever was stored (the store needs no
// renormalisation of its own).  Keeping 2 w rather than w saves an
// instruction per load and rounds the same way (the factor is a power of 2).
// FP64 links keep all four components (the reference's arithmetic).
struct F3 {
  float x, y, z;
};

__device__ __forceinline__ Q<float> link_unpack(const F3 &v) {
  const uint32_t neg = __float_as_uint(v.x) & 1u;
  const float s4 = __builtin_fmaf(v.z, v.z, __builtin_fmaf(v.y, v.y, v.x * v.x));  // 4 s
  const float den = __builtin_fmaf(s4, 0.25f, 1.0f);
  float t = __builtin_amdgcn_rcpf(den);
  t = __builtin_fmaf(__builtin_fmaf(-den, t, 1.0f), t, t);
  Q<float> q;
  q.a = __uint_as_float(__float_as_uint(__builtin_fmaf(s4, -0.25f, 1.0f) * t) ^ (neg << 31));
  q.b = v.x * t;
  q.c = v.y * t;
  q.d = v.z * t;
  return q;
}

// |q|^2 - 1 of a link within rounding of unit, as an FMA chain started at
// a^2 - 1: the small result is never rounded on the float grid near 1,
// whose spacing changes there (a renormalisation through 1/sqrt(|q|^2),
// even a correctly rounded one, keeps a bias of about 1e-9 in the norm;
// the hardware estimate alone left links about 3e-8 short, which a long
// gradient flow turned into a plaquette 1.4e-7 low).
__device__ __forceinline__ float unit_excess(const Q<float> &q) {
  float dx = __builtin_fmaf(q.a, q.a, -1.0f);
  dx = __builtin_fmaf(q.b, q.b, dx);
  dx = __builtin_fmaf(q.c, q.c, dx);
  return __builtin_fmaf(q.d, q.d, dx);
}

// The projection of q / |q|: v = 2 (b, c, d) / (|q| + |a|), with
// |q| = 1 + dx/2 (an error of order dx^2 ~ 1e-14), so a link that a kernel
// computed a little off unit is stored as its normalised direction rather
// than tilted by the excess norm.
__device__ __forceinline__ F3 link_pack(const Q<float> &q) {
  const float den = __builtin_fmaf(0.5f, unit_excess(q), 1.0f + fabsf(q.a));
  float t = __builtin_amdgcn_rcpf(den);
  t = __builtin_fmaf(__builtin_fmaf(-den, t, 1.0f), t, t);
  t += t;
  F3 v{q.b * t, q.c * t, q.d * t};
  v.x = __uint_as_float((__float_as_uint(v.x) & ~1u) | (q.a < 0.0f ? 1u : 0u));
  return v;
}

// Link loads and stores.  M carries the model bits of the kernel; with
// M_PK an FP32 link is in the packed 12-byte form, otherwise (and always in
// FP64) it is the four components.  There is no default: every caller says
// which layout it reads.  Every FP32 store renormalises (the packed form by
// its construction); FP64 stores the reference's arithmetic as it is.
template <int M, typename T>
constexpr bool packed_links() {
  return sizeof(T) == 4 && (M & M_PK) != 0;
}

template <int M, typename T>
__device__ __forceinline__ Q<T> ldU(const T *__restrict__ U, int Vh, int d, int par, int s) {
  if constexpr (packed_links<M, T>()) {
    gptr<const float> p = gmem(U + ((size_t)(d * 2 + par) * Vh + s) * 3);
    return link_unpack(F3{p[0], p[1], p[2]});
  } else {
    return ld4(U + ((size_t)(d * 2 + par) * Vh + s) * 4);
  }
}
template <int M, typename T>
__device__ __forceinline__ void stU(T *__restrict__ U, int Vh, int d, int par, int s, const Q<T> &q) {
  if constexpr (packed_links<M, T>()) {
    F3 v = link_pack(q);
    gptr<float> p = gmem(U + ((size_t)(d * 2 + par) * Vh + s) * 3);
    p[0] = v.x;
    p[1] = v.y;
    p[2] = v.z;
  } else if constexpr (sizeof(T) == 4) {
    // the full FP32 form renormalises every link it stores: q (1 + e) with
    // e = 1/sqrt(1 + dx) - 1 to second order in the excess dx, one rounding
    // per component
    const float dx = unit_excess(q);
    const float e = dx * __builtin_fmaf(0.375f, dx, -0.5f);
    st4(U + ((size_t)(d * 2 + par) * Vh + s) * 4,
        Q<float>{__builtin_fmaf(q.a, e, q.a), __builtin_fmaf(q.b, e, q.b), __builtin_fmaf(q.c, e, q.c),
                 __builtin_fmaf(q.d, e, q.d)});
  } else {
    st4(U + ((size_t)(d * 2 + par) * Vh + s) * 4, q);
  }
}
template <typename T>
__device__ __forceinline__ Q<T> ldF(const T *__restrict__ F, int Vh, int par, int s) {
  return ld4(F + ((size_t)par * Vh + s) * 4);
}
template <typename T>
__device__ __forceinline__ void stF(T *__restrict__ F, int Vh, int par, int s, const Q<T> &q) {
  st4(F + ((size_t)par * Vh + s) * 4, q);
}
template <typename T>
__device__ __forceinline__ V3<T> ldG(const T *__restrict__ G, int Vh, int par, int s) {
  gptr<const T> b = gmem(G + (size_t)par * Vh + s);
  size_t st = (size_t)2 * Vh;
  return V3<T>{b[0], b[st], b[2 * st]};
}
template <typename T>
__device__ __forceinline__ void stG(T *__restrict__ G, int Vh, int par, int s, const V3<T> &q) {
  gptr<T> b = gmem(G + (size_t)par * Vh + s);
  size_t st = (size_t)2 * Vh;
  b[0] = q.x;
  b[st] = q.y;
  b[2 * st] = q.z;
}
template <typename T>
__device__ __forceinline__ T ldA(const T *__restrict__ A, int Vh, int d, int par, int s) {
  return *gmem(A + (size_t)(d * 2 + par) * Vh + s);
}
template <typename T>
__device__ __forceinline__ T ldS(const T *__restrict__ S, int Vh, int par, int s) {
  return *gmem(S + (size_t)par * Vh + s);
}

// load by site
template <int M, typename T>
__device__ __forceinline__ Q<T> U_at(const LatDesc &L, const T *U, const Site &c, int d) {
  return ldU<M>(U, L.Vh, d, c.par, c.s);
}
template <typename T>
__device__ __forceinline__ Q<T> F_at(const LatDesc &L, const T *F, const Site &c) {
  return ldF(F, L.Vh, c.par, c.s);
}
template <typename T>
__device__ __forceinline__ V3<T> G_at(const LatDesc &L, const T *G, const Site &c) {
  return ldG(G, L.Vh, c.par, c.s);
}
template <typename T>
__device__ __forceinline__ T A_at(const LatDesc &L, const T *A, const Site &c, int d) {
  return ldA(A, L.Vh, d, c.par, c.s);
}
template <typename T>
__device__ __forceinline__ T S_at(const LatDesc &L, const T *S, const Site &c) {
  return ldS(S, L.Vh, c.par, c.s);
}

// ------------------------------------------------------------------------
// block reductions (64-wide wavefronts, BLK threads)

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ int wave_sumi(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Sum over the block in a fixed order; result valid in thread 0.
__device__ __forceinline__ double block_sum(double v, double *lds) {
  v = wave_sum(v);
  int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) lds[w] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < BLK / 64; i++) r += lds[i];
  }
  __syncthreads();
  return r;
}

// N sums over the block at once, each in block_sum's fixed order (the
// wavefront butterfly, then the waves in order); sum j is returned in
// thread j < N.  One barrier pair for all of them; lds holds BLK/64 * N.
template <int N>
__device__ __forceinline__ double block_sums(const double *v, double *lds) {
  const int w = threadIdx.x >> 6;
#pragma unroll
  for (int j = 0; j < N; j++) {
    double t = wave_sum(v[j]);
    if ((threadIdx.x & 63) == 0) lds[w * N + j] = t;
  }
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x < N) {
#pragma unroll
    for (int i = 0; i < BLK / 64; i++) r += lds[i * N + threadIdx.x];
  }
  __syncthreads();
  return r;
}

}  // namespace su2
