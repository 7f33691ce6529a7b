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
    uint32_t hi0 = __umulhi(0xD2511F53u, c0), lo0 = 0xD2511F53u * c0;
    uint32_t hi1 = __umulhi(0xCD9E8D57u, c2), lo1 = 0xCD9E8D57u * c2;
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
    sincosf(alpha, &s, &c);
  }
  return Q<T>{c, (T)0, (T)0, -s};
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
// geometry

struct Site {
  int x, y, z;
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
  return c;
}

__device__ __forceinline__ int ext(const LatDesc &L, int d) {
  return d == 0 ? L.L0 : (d == 1 ? L.L1 : L.L2);
}

__device__ __forceinline__ Site shift(const LatDesc &L, Site c, int d, int sign) {
  int n = ext(L, d);
  if (d == 0) {
    c.x += sign;
    c.x = c.x < 0 ? n - 1 : (c.x >= n ? 0 : c.x);
  } else if (d == 1) {
    c.y += sign;
    c.y = c.y < 0 ? n - 1 : (c.y >= n ? 0 : c.y);
  } else {
    c.z += sign;
    c.z = c.z < 0 ? n - 1 : (c.z >= n ? 0 : c.z);
  }
  return c;
}

__device__ __forceinline__ int par_of(const Site &c) { return (c.x + c.y + c.z) & 1; }
__device__ __forceinline__ int s_of(const LatDesc &L, const Site &c) {
  return (c.x >> 1) + L.h0 * (c.y + L.L1 * c.z);
}

// ------------------------------------------------------------------------
// field access (layout in su2hip_internal.h)

template <typename T>
__device__ __forceinline__ Q<T> ldU(const T *__restrict__ U, int Vh, int d, int par, int s) {
  const T *b = U + (size_t)(d * 8 + par) * Vh + s;
  size_t st = (size_t)2 * Vh;
  return Q<T>{b[0], b[st], b[2 * st], b[3 * st]};
}
template <typename T>
__device__ __forceinline__ void stU(T *__restrict__ U, int Vh, int d, int par, int s, const Q<T> &q) {
  T *b = U + (size_t)(d * 8 + par) * Vh + s;
  size_t st = (size_t)2 * Vh;
  b[0] = q.a;
  b[st] = q.b;
  b[2 * st] = q.c;
  b[3 * st] = q.d;
}
template <typename T>
__device__ __forceinline__ Q<T> ldF(const T *__restrict__ F, int Vh, int par, int s) {
  const T *b = F + (size_t)par * Vh + s;
  size_t st = (size_t)2 * Vh;
  return Q<T>{b[0], b[st], b[2 * st], b[3 * st]};
}
template <typename T>
__device__ __forceinline__ void stF(T *__restrict__ F, int Vh, int par, int s, const Q<T> &q) {
  T *b = F + (size_t)par * Vh + s;
  size_t st = (size_t)2 * Vh;
  b[0] = q.a;
  b[st] = q.b;
  b[2 * st] = q.c;
  b[3 * st] = q.d;
}
template <typename T>
__device__ __forceinline__ V3<T> ldG(const T *__restrict__ G, int Vh, int par, int s) {
  const T *b = G + (size_t)par * Vh + s;
  size_t st = (size_t)2 * Vh;
  return V3<T>{b[0], b[st], b[2 * st]};
}
template <typename T>
__device__ __forceinline__ void stG(T *__restrict__ G, int Vh, int par, int s, const V3<T> &q) {
  T *b = G + (size_t)par * Vh + s;
  size_t st = (size_t)2 * Vh;
  b[0] = q.x;
  b[st] = q.y;
  b[2 * st] = q.z;
}
template <typename T>
__device__ __forceinline__ T ldA(const T *__restrict__ A, int Vh, int d, int par, int s) {
  return A[(size_t)(d * 2 + par) * Vh + s];
}
template <typename T>
__device__ __forceinline__ T ldS(const T *__restrict__ S, int Vh, int par, int s) {
  return S[(size_t)par * Vh + s];
}

// load by coordinates
template <typename T>
__device__ __forceinline__ Q<T> U_at(const LatDesc &L, const T *U, const Site &c, int d) {
  return ldU(U, L.Vh, d, par_of(c), s_of(L, c));
}
template <typename T>
__device__ __forceinline__ Q<T> F_at(const LatDesc &L, const T *F, const Site &c) {
  return ldF(F, L.Vh, par_of(c), s_of(L, c));
}
template <typename T>
__device__ __forceinline__ V3<T> G_at(const LatDesc &L, const T *G, const Site &c) {
  return ldG(G, L.Vh, par_of(c), s_of(L, c));
}
template <typename T>
__device__ __forceinline__ T A_at(const LatDesc &L, const T *A, const Site &c, int d) {
  return ldA(A, L.Vh, d, par_of(c), s_of(L, c));
}
template <typename T>
__device__ __forceinline__ T S_at(const LatDesc &L, const T *S, const Site &c) {
  return ldS(S, L.Vh, par_of(c), s_of(L, c));
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

}  // namespace su2
