// Tile-based fused BiCGStab for one or many independent tissues.
//
// A tile is a run of consecutive SELL-64 slices of one tissue (the tiles of
// a tissue never hold rows of another).  One workgroup works on one tile in
// every launch, so a single launch per stage serves a whole batch of meshes of
// uneven size.  Every tissue has its own scalars (rho, alpha, omega, |b|), its
// own convergence flag and iteration count; its dot products are the ordered
// sum of its tiles' partials.  Nothing a tissue computes depends on the other
// tissues of the batch, so a tissue solved in a batch gives bitwise the result
// of the same tissue solved alone with the same tile size.
//
// Per iteration: w = J^T rh once per solve makes (rh, v) = (w, ph) = (rh, J ph)
// available before the SpMV that computes v (since (J^T rh, ph) = (rh, J ph)).
//   tz_spmv_v:   convergence test, alpha = rho / (w, ph); v = J ph;
//                s = r - alpha v; sh = M^-1 s
//   tz_spmv_t:   t = J sh; partials (t, s), (t, t)
//   tz_update_x: omega; x += alpha ph + omega sh; r = s - omega t;
//                partials (rh, r), (r, r)
//   tz_update_p: beta; p = r + beta (p - omega v); ph = M^-1 p; partial (w, ph)
// Partial slots: 0 (rh, r), 1 (r, r), 2 (t, s), 3 (t, t), 4 (w, ph).
// The per-tissue scalar slots are written only by the workgroup of the
// tissue's first tile, and only slots that no workgroup of the same launch
// reads.
#include <cstring>
#include <type_traits>

#include <hip/hip_ext.h>

#include "tf_internal.h"

namespace tf {
namespace {

constexpr int kW = kBlock / 64;

__device__ inline void reduce_store(double* vals, int n, double* const* outs, int slot) {
  __shared__ double red[4][kW];
  for (int k = 0; k < n; ++k) {
    double x = vals[k];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    vals[k] = x;
  }
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  __syncthreads();
  if (lane == 0)
    for (int k = 0; k < n; ++k) red[k][wave] = vals[k];
  __syncthreads();
  if (threadIdx.x == 0)
    for (int k = 0; k < n; ++k) {
      double s = 0.0;
#pragma unroll
      for (int w = 0; w < kW; ++w) s += red[k][w];
      outs[k][slot] = s;
    }
}

struct TC {
  TileMap tm;
  int32_t n_nodes;
  const int32_t* slice_base;
  const int32_t* col;
  const int32_t* tslot;
  int64_t n_slots;
  int32_t* state;        // 2 per tissue: [K_* state, iterations]
  double* scal;          // 8 per tissue
  double* part;          // 5 slots x stride
  int stride;
  const int32_t* active; // per tissue: 0 = skip the solve (x = 0)
  double rtol;
  const double* rtol_scale;  // per tissue factor on rtol (null: 1)
};

__device__ inline double tc_rtol(const TC& c, int T) {
  return c.rtol_scale ? c.rtol * c.rtol_scale[T] : c.rtol;
}


__device__ inline double* TP(const TC& c, int k) { return c.part + (size_t)k * c.stride; }

struct TileCtx {
  int L, T, s0, s1, first;
};

__device__ inline bool tile_ctx(const TC& c, TileCtx& t) {
  t.L = xcd_item(blockIdx.x, gridDim.x);
  if (t.L >= c.tm.n_tiles) return false;
  t.T = c.tm.tile_tissue[t.L];
  t.s0 = c.tm.tile_slice[t.L];
  t.s1 = c.tm.tile_slice[t.L + 1];
  t.first = (t.L == c.tm.tissue_tile[t.T]);
  return true;
}

__device__ inline double tissue_sum(const TC& c, int k, int T) {
  return wg_ordered_sum(TP(c, k), c.tm.tissue_tile[T], c.tm.tissue_tile[T + 1]);
}

template <typename T>
__device__ inline void minv_apply(const T* m, const T* in, T* out, int i) {
  const T a = in[3 * i], b = in[3 * i + 1], d = in[3 * i + 2];
  const T* M = m + 9 * (size_t)i;
  out[3 * i] = M[0] * a + M[1] * b + M[2] * d;
  out[3 * i + 1] = M[3] * a + M[4] * b + M[5] * d;
  out[3 * i + 2] = M[6] * a + M[7] * b + M[8] * d;
}

// One SELL-64 row (lane) of a slice: a = sum over the row's blocks of
// J_block x_col.  U consecutive block steps are loaded before any is
// accumulated, so U column-index loads, 3U gathers and 9U value loads are in
// flight together; the accumulation order (block by block, in k order) is
// that of the plain loop, so the result does not depend on U.
#ifndef TF_SPMV_UNROLL32
#define TF_SPMV_UNROLL32 4
#endif
#ifndef TF_SPMV_UNROLL64
#define TF_SPMV_UNROLL64 2
#endif
template <typename T, typename Gather>
__device__ inline void row_spmv_g(int sl, int lane, const TC& c, const T* __restrict__ val,
                                  Gather gx, T& a0, T& a1, T& a2) {
  constexpr int U = sizeof(T) == 4 ? TF_SPMV_UNROLL32 : TF_SPMV_UNROLL64;
  const int64_t base = c.slice_base[sl];
  const int width = (int)((c.slice_base[sl + 1] - base) >> 6);
  a0 = 0; a1 = 0; a2 = 0;
  const T* v0 = val + base + lane;
  const int32_t* cp = c.col + base + lane;
  const int64_t ns = c.n_slots;
  int k = 0;
  for (; k + U <= width; k += U) {
    int cc[U];
    T xv[U][3], m[U][9];
#pragma unroll
    for (int u = 0; u < U; ++u) cc[u] = __builtin_nontemporal_load(cp + (int64_t)(k + u) * 64);
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const T* vv = v0 + (int64_t)(k + u) * 64;
#pragma unroll
      for (int q = 0; q < 9; ++q) m[u][q] = __builtin_nontemporal_load(vv + q * ns);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) gx(cc[u], xv[u][0], xv[u][1], xv[u][2]);
#pragma unroll
    for (int u = 0; u < U; ++u) {
      a0 += m[u][0] * xv[u][0] + m[u][1] * xv[u][1] + m[u][2] * xv[u][2];
      a1 += m[u][3] * xv[u][0] + m[u][4] * xv[u][1] + m[u][5] * xv[u][2];
      a2 += m[u][6] * xv[u][0] + m[u][7] * xv[u][1] + m[u][8] * xv[u][2];
    }
  }
  for (; k < width; ++k) {
    const int64_t o = (int64_t)k * 64;
    const int cc = __builtin_nontemporal_load(cp + o);
    const T* vv = v0 + o;
    const T m0 = __builtin_nontemporal_load(vv), m1 = __builtin_nontemporal_load(vv + ns),
            m2 = __builtin_nontemporal_load(vv + 2 * ns), m3 = __builtin_nontemporal_load(vv + 3 * ns),
            m4 = __builtin_nontemporal_load(vv + 4 * ns), m5 = __builtin_nontemporal_load(vv + 5 * ns),
            m6 = __builtin_nontemporal_load(vv + 6 * ns), m7 = __builtin_nontemporal_load(vv + 7 * ns),
            m8 = __builtin_nontemporal_load(vv + 8 * ns);
    T x0, x1, x2;
    gx(cc, x0, x1, x2);
    a0 += m0 * x0 + m1 * x1 + m2 * x2;
    a1 += m3 * x0 + m4 * x1 + m5 * x2;
    a2 += m6 * x0 + m7 * x1 + m8 * x2;
  }
}

// Plain gather of x.
template <typename T>
__device__ inline void row_spmv(int sl, int lane, const TC& c, const T* __restrict__ val,
                                const T* __restrict__ x, T& a0, T& a1, T& a2) {
  row_spmv_g<T>(sl, lane, c, val,
                [x](int cc, T& x0, T& x1, T& x2) {
                  x0 = x[3 * cc];
                  x1 = x[3 * cc + 1];
                  x2 = x[3 * cc + 2];
                },
                a0, a1, a2);
}

// r = rh = p = b, x = 0 (x = 0 also for tissues that are not solved).
template <typename T>
__global__ __launch_bounds__(kBlock) void tz_init_a(TC c, const double* b, T* r, T* rh, T* p,
                                                    T* x) {
  TileCtx t;
  if (!tile_ctx(c, t)) return;
  const int end = min(t.s1 * 64, c.n_nodes);
  for (int i = t.s0 * 64 + threadIdx.x; i < end; i += kBlock)
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const T bi = (T)b[3 * i + d];
      r[3 * i + d] = bi; rh[3 * i + d] = bi; p[3 * i + d] = bi; x[3 * i + d] = 0;
    }
}

// w = J^T rh through the transpose slot map.
template <typename T>
__global__ __launch_bounds__(kBlock) void tz_spmv_transposed(TC c, const T* val, const T* x,
                                                             T* y) {
  TileCtx t;
  if (!tile_ctx(c, t)) return;
  if (!c.active[t.T]) return;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int sl = t.s0 + wave; sl < t.s1; sl += kW) {
    const int64_t base = c.slice_base[sl];
    const int width = (int)((c.slice_base[sl + 1] - base) >> 6);
    T a0 = 0, a1 = 0, a2 = 0;
    for (int k = 0; k < width; ++k) {
      const int64_t o = base + (int64_t)k * 64 + lane;
      const int cc = c.col[o];
      const int64_t s = c.tslot[o];
      const T x0 = x[3 * cc], x1 = x[3 * cc + 1], x2 = x[3 * cc + 2];
      a0 += val[s] * x0 + val[3 * c.n_slots + s] * x1 + val[6 * c.n_slots + s] * x2;
      a1 += val[c.n_slots + s] * x0 + val[4 * c.n_slots + s] * x1 + val[7 * c.n_slots + s] * x2;
      a2 += val[2 * c.n_slots + s] * x0 + val[5 * c.n_slots + s] * x1 + val[8 * c.n_slots + s] * x2;
    }
    const int row = sl * 64 + lane;
    if (row < c.n_nodes) { y[3 * row] = a0; y[3 * row + 1] = a1; y[3 * row + 2] = a2; }
  }
}

// ph = M^-1 p; partials (rh, r) = (r, r), (r, r), (w, ph); tissue scalars reset.
template <typename T>
__global__ __launch_bounds__(kBlock) void tz_init_b(TC c, const T* r, const T* w, const T* p,
                                                    T* ph, const T* minv) {
  TileCtx t;
  if (!tile_ctx(c, t)) return;
  const int end = min(t.s1 * 64, c.n_nodes);
  double acc[3] = {0.0, 0.0, 0.0};
  for (int i = t.s0 * 64 + threadIdx.x; i < end; i += kBlock) {
    minv_apply(minv, p, ph, i);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      acc[0] += (double)r[3 * i + d] * (double)r[3 * i + d];
      acc[2] += (double)w[3 * i + d] * (double)ph[3 * i + d];
    }
  }
  acc[1] = acc[0];
  double* outs[3] = {TP(c, 0), TP(c, 1), TP(c, 4)};
  reduce_store(acc, 3, outs, t.L);
  if (t.first && threadIdx.x == 0) {
    c.scal[8 * t.T + TS_OMEGA] = 1.0;
    c.state[2 * t.T] = c.active[t.T] ? K_RUN : K_CONVERGED;
    c.state[2 * t.T + 1] = 0;
  }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void tz_spmv_v(TC c, const T* __restrict__ val,
                                                    const T* __restrict__ ph,
                                                    const T* __restrict__ r,
                                                    const T* __restrict__ minv,
                                                    T* __restrict__ v, T* __restrict__ s,
                                                    T* __restrict__ sh) {
  TileCtx t;
  if (!tile_ctx(c, t)) return;
  if (c.state[2 * t.T] != K_RUN) return;
  const double rho = tissue_sum(c, 0, t.T);
  const double rr = tissue_sum(c, 1, t.T);
  const double wp = tissue_sum(c, 4, t.T);
  const double rnorm = sqrt(rr);
  const bool first_it = (c.state[2 * t.T + 1] == 0);
  double* sc = c.scal + 8 * t.T;
  const double bnorm = first_it ? rnorm : sc[TS_BNORM];
  const double alpha = rho / wp;
  // a NaN or a residual grown far past |b| (divergence, typically of an FP32
  // solve on an ill-conditioned tissue) is a breakdown, never convergence
  int state = K_RUN;
  if (!isfinite(rnorm) || rnorm > kDivergence * bnorm) state = K_BREAKDOWN;
  else if (rnorm <= tc_rtol(c, t.T) * bnorm) state = K_CONVERGED;
  else if (!first_it && sc[TS_OMEGA] == 0.0) state = K_BREAKDOWN;
  else if (rho == 0.0 || wp == 0.0 || !isfinite(alpha)) state = K_BREAKDOWN;
  if (t.first && threadIdx.x == 0) {
    if (first_it) sc[TS_BNORM] = bnorm;
    sc[TS_RHO] = rho;
    sc[TS_ALPHA] = alpha;
    sc[TS_RNORM] = rnorm;
    if (state != K_RUN) c.state[2 * t.T] = state;
  }
  if (state != K_RUN) return;
  const T ta = (T)alpha;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int sl = t.s0 + wave; sl < t.s1; sl += kW) {
    T a0, a1, a2;
    row_spmv(sl, lane, c, val, ph, a0, a1, a2);
    const int row = sl * 64 + lane;
    if (row < c.n_nodes) {
      v[3 * row] = a0; v[3 * row + 1] = a1; v[3 * row + 2] = a2;
      s[3 * row] = r[3 * row] - ta * a0;
      s[3 * row + 1] = r[3 * row + 1] - ta * a1;
      s[3 * row + 2] = r[3 * row + 2] - ta * a2;
      minv_apply(minv, s, sh, row);
    }
  }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void tz_spmv_t(TC c, const T* __restrict__ val,
                                                    const T* __restrict__ sh,
                                                    const T* __restrict__ s, T* __restrict__ tv) {
  TileCtx t;
  if (!tile_ctx(c, t)) return;
  if (c.state[2 * t.T] != K_RUN) return;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  double acc[2] = {0.0, 0.0};
  for (int sl = t.s0 + wave; sl < t.s1; sl += kW) {
    T a0, a1, a2;
    row_spmv(sl, lane, c, val, sh, a0, a1, a2);
    const int row = sl * 64 + lane;
    if (row < c.n_nodes) {
      tv[3 * row] = a0; tv[3 * row + 1] = a1; tv[3 * row + 2] = a2;
      acc[0] += (double)a0 * (double)s[3 * row] + (double)a1 * (double)s[3 * row + 1] +
                (double)a2 * (double)s[3 * row + 2];
      acc[1] += (double)a0 * (double)a0 + (double)a1 * (double)a1 + (double)a2 * (double)a2;
    }
  }
  double* outs[2] = {TP(c, 2), TP(c, 3)};
  reduce_store(acc, 2, outs, t.L);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void tz_update_x(TC c, T* x, T* r, const T* rh, const T* ph,
                                                      const T* sh, const T* s, const T* tv) {
  TileCtx t;
  if (!tile_ctx(c, t)) return;
  if (c.state[2 * t.T] != K_RUN) return;
  const double ts = tissue_sum(c, 2, t.T);
  const double tt = tissue_sum(c, 3, t.T);
  const double omega = (tt > 0.0) ? ts / tt : 0.0;
  double* sc = c.scal + 8 * t.T;
  const T ta = (T)sc[TS_ALPHA], tw = (T)omega;
  const int end = min(t.s1 * 64, c.n_nodes);
  double acc[2] = {0.0, 0.0};
  for (int i = t.s0 * 64 + threadIdx.x; i < end; i += kBlock)
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const int k = 3 * i + d;
      x[k] = x[k] + ta * ph[k] + tw * sh[k];
      const T rk = s[k] - tw * tv[k];
      r[k] = rk;
      acc[0] += (double)rh[k] * (double)rk;
      acc[1] += (double)rk * (double)rk;
    }
  double* outs[2] = {TP(c, 0), TP(c, 1)};
  reduce_store(acc, 2, outs, t.L);
  if (t.first && threadIdx.x == 0) {
    sc[TS_OMEGA] = omega;
    c.state[2 * t.T + 1] = c.state[2 * t.T + 1] + 1;
  }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void tz_update_p(TC c, const T* r, T* p, T* ph, const T* v,
                                                      const T* w, const T* minv) {
  TileCtx t;
  if (!tile_ctx(c, t)) return;
  if (c.state[2 * t.T] != K_RUN) return;
  const double rho_new = tissue_sum(c, 0, t.T);
  const double* sc = c.scal + 8 * t.T;
  const double rho = sc[TS_RHO], alpha = sc[TS_ALPHA], omega = sc[TS_OMEGA];
  const double beta = (omega != 0.0 && rho != 0.0) ? (rho_new / rho) * (alpha / omega) : 0.0;
  const T tb = (T)beta, tw = (T)omega;
  const int end = min(t.s1 * 64, c.n_nodes);
  double acc[1] = {0.0};
  for (int i = t.s0 * 64 + threadIdx.x; i < end; i += kBlock) {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const int k = 3 * i + d;
      p[k] = r[k] + tb * (p[k] - tw * v[k]);
    }
    minv_apply(minv, p, ph, i);
#pragma unroll
    for (int d = 0; d < 3; ++d) acc[0] += (double)w[3 * i + d] * (double)ph[3 * i + d];
  }
  double* outs[1] = {TP(c, 4)};
  reduce_store(acc, 1, outs, t.L);
}

// ---------------------------------------------------------------------------
// Preconditioned CG for nearly symmetric Jacobians.  Every workgroup forms its
// tissue's scalars from the ordered sums of the tiles' partials (the same
// value in every workgroup), and the tissue's first workgroup stores them:
//   S1 (cg_s1): rho = (r, z), the convergence test on |r|, beta = rho / rho_old;
//   S2:         alpha = rho / (p, q).
// Without a coarse space, two launches per iteration:
//   tzc_spmv<FUSED>: S1; p = z + beta p_old formed on the fly for every
//                    gathered column and stored for the tile's own rows (into
//                    the other p buffer); q = J p; partial (p, q)
//   tzc_update:      S2; x += alpha p; r -= alpha q; z = M^-1 r; partials
//                    (r, z), (r, r)
// With the coarse space (z = M^-1 r + P A_c^-1 P^T r), five:
//   tzc_spmv:        q = J p; partial (p, q)
//   tzc_update:      S2; x += alpha p; r -= alpha q; z_D = M^-1 r; partials
//   coarse restriction; coarse product y = A_c^-1 P^T r with partials (w, y)
//   tzc_pupdate:     S1 (with the (w, y) partials); p = (z_D + P y) + beta p
// A non-positive (p, q) or rho (indefinite operator) is a breakdown: the caller
// falls back to BiCGStab.
enum { TC_RZ0 = 0, TC_RR0 = 1, TC_PQ = 2 };

template <typename T>
__device__ inline T fma_t(T a, T b, T c) {
  if constexpr (sizeof(T) == 4) return fmaf(a, b, c);
  else return fma(a, b, c);
}

__device__ inline CgS1 make_s1(const TC& c) {
  CgS1 s;
  s.rz = TP(c, TC_RZ0);
  s.rr = TP(c, TC_RR0);
  s.tissue_tile = c.tm.tissue_tile;
  s.scal = c.scal;
  s.state = c.state;
  s.rtol = c.rtol;
  s.rtol_scale = c.rtol_scale;
  return s;
}

// r = b, x = 0, p = 0, z = M^-1 r; partials (r, z), (r, r); rho_old = 1.  A
// tissue not solved (active 0) is converged with x = 0.
template <typename T>
__global__ __launch_bounds__(kBlock) void tzc_init(TC c, const double* b, T* r, T* z, T* x,
                                                   T* p, const T* minv) {
  TileCtx t;
  if (!tile_ctx(c, t)) return;
  const int end = min(t.s1 * 64, c.n_nodes);
  double acc[2] = {0.0, 0.0};
  for (int i = t.s0 * 64 + threadIdx.x; i < end; i += kBlock) {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      r[3 * i + d] = (T)b[3 * i + d];
      x[3 * i + d] = 0;
      p[3 * i + d] = 0;
    }
    minv_apply(minv, r, z, i);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      acc[0] += (double)r[3 * i + d] * (double)z[3 * i + d];
      acc[1] += (double)r[3 * i + d] * (double)r[3 * i + d];
    }
  }
  double* outs[2] = {TP(c, TC_RZ0), TP(c, TC_RR0)};
  reduce_store(acc, 2, outs, t.L);
  if (t.first && threadIdx.x == 0) {
    c.scal[8 * t.T + TS_RHO] = 1.0;
    c.state[2 * t.T] = c.active[t.T] ? K_RUN : K_CONVERGED;
    c.state[2 * t.T + 1] = 0;
  }
}

// q = J p, partial (p, q).  FUSED: S1, then p = z + beta p_old on the fly
// (p_old and z gathered, p stored into pnew for the tile's rows); else p is
// gathered as it is (pold) and pnew is not used.
template <typename T, bool FUSED>
__global__ __launch_bounds__(kBlock) void tzc_spmv(TC c, const T* __restrict__ val,
                                                   const T* __restrict__ z,
                                                   const T* __restrict__ pold,
                                                   T* __restrict__ pnew, T* __restrict__ q) {
  TileCtx t;
  if (!tile_ctx(c, t)) return;
  if (c.state[2 * t.T] != K_RUN) return;
  T tb = (T)0;
  if (FUSED) {
    double beta;
    if (cg_s1(make_s1(c), t.T, t.first, &beta) != K_RUN) return;
    tb = (T)beta;
  }
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  double acc[1] = {0.0};
  for (int sl = t.s0 + wave; sl < t.s1; sl += kW) {
    T a0, a1, a2;
    if (FUSED) {
      row_spmv_g<T>(sl, lane, c, val,
                    [z, pold, tb](int cc, T& x0, T& x1, T& x2) {
                      x0 = fma_t(tb, pold[3 * cc], z[3 * cc]);
                      x1 = fma_t(tb, pold[3 * cc + 1], z[3 * cc + 1]);
                      x2 = fma_t(tb, pold[3 * cc + 2], z[3 * cc + 2]);
                    },
                    a0, a1, a2);
    } else {
      row_spmv<T>(sl, lane, c, val, pold, a0, a1, a2);
    }
    const int row = sl * 64 + lane;
    if (row < c.n_nodes) {
      T p0, p1, p2;
      if (FUSED) {
        p0 = fma_t(tb, pold[3 * row], z[3 * row]);
        p1 = fma_t(tb, pold[3 * row + 1], z[3 * row + 1]);
        p2 = fma_t(tb, pold[3 * row + 2], z[3 * row + 2]);
        pnew[3 * row] = p0; pnew[3 * row + 1] = p1; pnew[3 * row + 2] = p2;
      } else {
        p0 = pold[3 * row]; p1 = pold[3 * row + 1]; p2 = pold[3 * row + 2];
      }
      q[3 * row] = a0; q[3 * row + 1] = a1; q[3 * row + 2] = a2;
      acc[0] += (double)p0 * (double)a0 + (double)p1 * (double)a1 + (double)p2 * (double)a2;
    }
  }
  double* outs[1] = {TP(c, TC_PQ)};
  reduce_store(acc, 1, outs, t.L);
}

// S2; x += alpha p; r -= alpha q; z = M^-1 r; partials (r, z), (r, r).  A
// thread's first kUpre rows (p, q, x, r and the block-Jacobi inverse) are
// loaded before (p, q) is reduced, so that their loads overlap the reduction.
constexpr int kUpre = 2;

template <typename T>
__device__ inline void update_row(T ta, const T (&pv)[3], const T (&qv)[3], T (&xv)[3],
                                  T (&rv)[3], const T (&m)[9], T* x, T* r, T* z, int i,
                                  double* acc) {
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    xv[d] = fma_t(ta, pv[d], xv[d]);
    rv[d] = fma_t(-ta, qv[d], rv[d]);
    x[3 * i + d] = xv[d];
    r[3 * i + d] = rv[d];
  }
  const T zv[3] = {m[0] * rv[0] + m[1] * rv[1] + m[2] * rv[2],
                   m[3] * rv[0] + m[4] * rv[1] + m[5] * rv[2],
                   m[6] * rv[0] + m[7] * rv[1] + m[8] * rv[2]};
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    z[3 * i + d] = zv[d];
    acc[0] += (double)rv[d] * (double)zv[d];
    acc[1] += (double)rv[d] * (double)rv[d];
  }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void tzc_update(TC c, T* x, T* r, T* z, const T* p,
                                                     const T* q, const T* minv) {
  TileCtx t;
  if (!tile_ctx(c, t)) return;
  if (c.state[2 * t.T] != K_RUN) return;
  const int beg = t.s0 * 64 + threadIdx.x;
  const int end = min(t.s1 * 64, c.n_nodes);
  T pv[kUpre][3], qv[kUpre][3], xv[kUpre][3], rv[kUpre][3], m[kUpre][9];
#pragma unroll
  for (int k = 0; k < kUpre; ++k) {
    const int i = beg + k * kBlock;
    const bool in = i < end;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      pv[k][d] = in ? p[3 * i + d] : (T)0;
      qv[k][d] = in ? q[3 * i + d] : (T)0;
      xv[k][d] = in ? x[3 * i + d] : (T)0;
      rv[k][d] = in ? r[3 * i + d] : (T)0;
    }
#pragma unroll
    for (int e = 0; e < 9; ++e) m[k][e] = in ? minv[9 * (size_t)i + e] : (T)0;
  }
  const double rho = c.scal[8 * t.T + TS_RHON];
  const double pq = tissue_sum(c, TC_PQ, t.T);
  const double alpha = rho / pq;
  if (!(pq > 0.0) || !isfinite(alpha)) {
    if (t.first && threadIdx.x == 0) c.state[2 * t.T] = K_BREAKDOWN;
    return;
  }
  const T ta = (T)alpha;
  double acc[2] = {0.0, 0.0};
#pragma unroll
  for (int k = 0; k < kUpre; ++k) {
    const int i = beg + k * kBlock;
    if (i < end) update_row<T>(ta, pv[k], qv[k], xv[k], rv[k], m[k], x, r, z, i, acc);
  }
  for (int i = beg + kUpre * kBlock; i < end; i += kBlock) {
    const T p1[3] = {p[3 * i], p[3 * i + 1], p[3 * i + 2]};
    const T q1[3] = {q[3 * i], q[3 * i + 1], q[3 * i + 2]};
    T x1[3] = {x[3 * i], x[3 * i + 1], x[3 * i + 2]};
    T r1[3] = {r[3 * i], r[3 * i + 1], r[3 * i + 2]};
    T m1[9];
#pragma unroll
    for (int e = 0; e < 9; ++e) m1[e] = minv[9 * (size_t)i + e];
    update_row<T>(ta, p1, q1, x1, r1, m1, x, r, z, i, acc);
  }
  double* outs[2] = {TP(c, TC_RZ0), TP(c, TC_RR0)};
  reduce_store(acc, 2, outs, t.L);
  if (t.first && threadIdx.x == 0) {
    c.scal[8 * t.T + TS_ALPHA] = alpha;
    c.scal[8 * t.T + TS_RHO] = rho;
    c.state[2 * t.T + 1] = c.state[2 * t.T + 1] + 1;
  }
}

// Coarse data of the CG update that also forms the restriction w = P^T r.
template <typename T>
struct CoarseU {
  const int32_t* agg_ptr;
  const int32_t* agg_nodes;
  const int32_t* agg_tissue;
  const int32_t* agg_cbase;
  const int32_t* seg_agg;  // first aggregate of each tissue
  const float* list_d;
  const uint8_t* list_mask;
  T* w;        // P^T r terms of the first half of every aggregate's list
  T* w2;       // and of the second half
  double* rz;  // per half-aggregate: partial (r, z_D)
  double* rr;  // per half-aggregate: partial (r, r)
  int32_t n_agg;
};

// One workgroup per half of an aggregate's node list: 630 aggregates at C3
// make 1260 workgroups, which spread over the CUs more evenly than 630.
constexpr int kUBlock = 256;

// Workgroup sum of n values per thread in a fixed order (wave trees, then the
// waves in order), valid in thread 0.
template <int N, int B>
__device__ inline void wg_sum(double (&v)[N]) {
  __shared__ double red[N][B / 64];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
  for (int k = 0; k < N; ++k) {
    double x = v[k];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    if (lane == 0) red[k][wave] = x;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < N; ++k) {
      double t = 0.0;
#pragma unroll
      for (int w = 0; w < B / 64; ++w) t += red[k][w];
      v[k] = t;
    }
  }
}

// The restriction terms of one node (r after the update) into acc[2..7]:
// translations r, rotations d x r, on the DOF that take part in the coarse space.
template <typename T>
__device__ inline void restrict_row(const T (&rv)[3], const float* d3, uint8_t mk, double* acc) {
  const double r0 = (mk & 1) ? 0.0 : (double)rv[0];
  const double r1 = (mk & 2) ? 0.0 : (double)rv[1];
  const double r2 = (mk & 4) ? 0.0 : (double)rv[2];
  const double d0 = d3[0], d1 = d3[1], d2 = d3[2];
  acc[2] += r0;
  acc[3] += r1;
  acc[4] += r2;
  acc[5] += d1 * r2 - d2 * r1;
  acc[6] += d2 * r0 - d0 * r2;
  acc[7] += d0 * r1 - d1 * r0;
}

// S2 and the update of tzc_update over the node list of one aggregate, which
// also forms the aggregate's w = P^T r from the updated r: the restriction
// needs no launch of its own.  Workgroup h takes half h & 1 of aggregate h >> 1
// and writes its terms of w to w (first half) or w2 (second half); the product
// adds the two.  The partials (r, z_D), (r, r) are per half-aggregate.  Every
// node of a tissue belongs to one aggregate; rows of no aggregate (the padding
// rows of a batch) hold x = r = z = 0 and are left alone.
// update_row with x, r at list position t and z at node i
template <typename T>
__device__ inline void update_row_l(T ta, const T (&pv)[3], const T (&qv)[3], T (&xv)[3],
                                    T (&rv)[3], const T (&m)[9], T* xl, T* rl, T* z, int t, int i,
                                    double* acc) {
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    xv[d] = fma_t(ta, pv[d], xv[d]);
    rv[d] = fma_t(-ta, qv[d], rv[d]);
    xl[3 * t + d] = xv[d];
    rl[3 * t + d] = rv[d];
  }
  const T zv[3] = {m[0] * rv[0] + m[1] * rv[1] + m[2] * rv[2],
                   m[3] * rv[0] + m[4] * rv[1] + m[5] * rv[2],
                   m[6] * rv[0] + m[7] * rv[1] + m[8] * rv[2]};
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    z[3 * i + d] = zv[d];
    acc[0] += (double)rv[d] * (double)zv[d];
    acc[1] += (double)rv[d] * (double)rv[d];
  }
}

// x, r and the block-Jacobi inverse are read and written in list order (xl,
// rl, ml: position t of agg_nodes), so that they stream; p and q are gathered
// and z is written at the node.
template <typename T>
__global__ __launch_bounds__(kUBlock) void tzc_update_agg(TC c, CoarseU<T> cu, T* xl, T* rl, T* z,
                                                         const T* p, const T* q,
                                                         const T* ml) {
  const int h = xcd_item(blockIdx.x, gridDim.x);
  if (h >= 2 * cu.n_agg) return;
  const int a = h >> 1, half = h & 1;
  const int T_ = cu.agg_tissue[a];
  if (c.state[2 * T_] != K_RUN) return;
  const bool first = a == cu.seg_agg[T_] && half == 0;
  const int a0 = cu.agg_ptr[a], a1 = cu.agg_ptr[a + 1], mid = a0 + (a1 - a0 + 1) / 2;
  const int beg = (half ? mid : a0) + threadIdx.x, end = half ? a1 : mid;
  int idx[kUpre];
  T pv[kUpre][3], qv[kUpre][3], xv[kUpre][3], rv[kUpre][3], m[kUpre][9];
#pragma unroll
  for (int k = 0; k < kUpre; ++k) {
    const int t = beg + k * kUBlock;
    idx[k] = t < end ? cu.agg_nodes[t] : -1;
  }
#pragma unroll
  for (int k = 0; k < kUpre; ++k) {
    const int i = idx[k], t = beg + k * kUBlock;
    const bool in = i >= 0;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      pv[k][d] = in ? p[3 * i + d] : (T)0;
      qv[k][d] = in ? q[3 * i + d] : (T)0;
      xv[k][d] = in ? xl[3 * t + d] : (T)0;
      rv[k][d] = in ? rl[3 * t + d] : (T)0;
    }
#pragma unroll
    for (int e = 0; e < 9; ++e) m[k][e] = in ? ml[9 * (size_t)t + e] : (T)0;
  }
  const double rho = c.scal[8 * T_ + TS_RHON];
  const double pq =
      wg_ordered_sum<kUBlock>(TP(c, TC_PQ), c.tm.tissue_tile[T_], c.tm.tissue_tile[T_ + 1]);
  const double alpha = rho / pq;
  if (!(pq > 0.0) || !isfinite(alpha)) {
    if (first && threadIdx.x == 0) c.state[2 * T_] = K_BREAKDOWN;
    return;
  }
  const T ta = (T)alpha;
  double acc[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int k = 0; k < kUpre; ++k) {
    const int i = idx[k];
    if (i < 0) continue;
    const int t = beg + k * kUBlock;
    update_row_l<T>(ta, pv[k], qv[k], xv[k], rv[k], m[k], xl, rl, z, t, i, acc);
    restrict_row<T>(rv[k], cu.list_d + 3 * (size_t)t, cu.list_mask[t], acc);
  }
  for (int t = beg + kUpre * kUBlock; t < end; t += kUBlock) {
    const int i = cu.agg_nodes[t];
    const T p1[3] = {p[3 * i], p[3 * i + 1], p[3 * i + 2]};
    const T q1[3] = {q[3 * i], q[3 * i + 1], q[3 * i + 2]};
    T x1[3] = {xl[3 * t], xl[3 * t + 1], xl[3 * t + 2]};
    T r1[3] = {rl[3 * t], rl[3 * t + 1], rl[3 * t + 2]};
    T m1[9];
#pragma unroll
    for (int e = 0; e < 9; ++e) m1[e] = ml[9 * (size_t)t + e];
    update_row_l<T>(ta, p1, q1, x1, r1, m1, xl, rl, z, t, i, acc);
    restrict_row<T>(r1, cu.list_d + 3 * (size_t)t, cu.list_mask[t], acc);
  }
  wg_sum<8, kUBlock>(acc);
  if (threadIdx.x == 0) {
    cu.rz[h] = acc[0];
    cu.rr[h] = acc[1];
    const int cb = cu.agg_cbase[a];
    T* wh = half ? cu.w2 : cu.w;
#pragma unroll
    for (int k = 0; k < 6; ++k) wh[cb + k] = (T)acc[2 + k];
    if (first) {
      c.scal[8 * T_ + TS_ALPHA] = alpha;
      c.scal[8 * T_ + TS_RHO] = rho;
      c.state[2 * T_ + 1] = c.state[2 * T_ + 1] + 1;
    }
  }
}

// Coarse data of the search-direction update (CG with a coarse space).
template <typename T>
struct CoarseP {
  const int32_t* node_cb;  // per node: first coarse DOF << 3 | excluded DOF bits (-1: none)
  const int32_t* seg_on;
  const float* node_d;
  const T* y;
};

// S1 (with the coarse partials (w, y)); p = (z_D + P y) + beta p, in place;
// the coarse term t + omega x d of the node's aggregate on the DOF that take
// part in the coarse space.  A thread's first kPre rows are loaded before the
// tissue's scalars are reduced, so that their loads overlap the reduction.
template <typename T>
__device__ inline void pupdate_row(const CoarseP<T>& cp, bool on, T tb, const T (&zi)[3],
                                   const T (&pi)[3], const T (&d)[3], int cb, T* p, int i) {
  T zz[3] = {zi[0], zi[1], zi[2]};
  if (on && cb >= 0) {
    const T* ya = cp.y + (cb >> 3);
    const T v[3] = {ya[0] + (ya[4] * d[2] - ya[5] * d[1]), ya[1] + (ya[5] * d[0] - ya[3] * d[2]),
                    ya[2] + (ya[3] * d[1] - ya[4] * d[0])};
#pragma unroll
    for (int q = 0; q < 3; ++q)
      if (!((cb >> q) & 1)) zz[q] += v[q];
  }
#pragma unroll
  for (int q = 0; q < 3; ++q) p[3 * i + q] = fma_t(tb, pi[q], zz[q]);
}

constexpr int kPre = 2;

template <typename T>
__global__ __launch_bounds__(kBlock) void tzc_pupdate(TC c, CgS1 s1, CoarseP<T> cp,
                                                      const T* __restrict__ z,
                                                      T* __restrict__ p) {
  TileCtx t;
  if (!tile_ctx(c, t)) return;
  if (c.state[2 * t.T] != K_RUN) return;
  const bool on = cp.seg_on[t.T] != 0;
  const int beg = t.s0 * 64 + threadIdx.x;
  const int end = min(t.s1 * 64, c.n_nodes);
  T zi[kPre][3], pi[kPre][3], d[kPre][3];
  int cb[kPre];
#pragma unroll
  for (int k = 0; k < kPre; ++k) {
    const int i = beg + k * kBlock;
    const bool in = i < end;
    cb[k] = in ? cp.node_cb[i] : -1;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      zi[k][q] = in ? z[3 * i + q] : (T)0;
      pi[k][q] = in ? p[3 * i + q] : (T)0;
      d[k][q] = in ? (T)cp.node_d[3 * i + q] : (T)0;
    }
  }
  double beta;
  if (cg_s1(s1, t.T, t.first, &beta) != K_RUN) return;
  const T tb = (T)beta;
#pragma unroll
  for (int k = 0; k < kPre; ++k) {
    const int i = beg + k * kBlock;
    if (i < end) pupdate_row<T>(cp, on, tb, zi[k], pi[k], d[k], cb[k], p, i);
  }
  for (int i = beg + kPre * kBlock; i < end; i += kBlock) {
    const T z1[3] = {z[3 * i], z[3 * i + 1], z[3 * i + 2]};
    const T p1[3] = {p[3 * i], p[3 * i + 1], p[3 * i + 2]};
    const T d1[3] = {(T)cp.node_d[3 * i], (T)cp.node_d[3 * i + 1], (T)cp.node_d[3 * i + 2]};
    pupdate_row<T>(cp, on, tb, z1, p1, d1, cp.node_cb[i], p, i);
  }
}

// |dx|^2 per tile (slot 0) and xk += dx.
__global__ __launch_bounds__(kBlock) void tz_norm_update(TC c, const double* dx, double* xk) {
  TileCtx t;
  if (!tile_ctx(c, t)) return;
  const int end = min(t.s1 * 64, c.n_nodes);
  double acc[1] = {0.0};
  for (int i = t.s0 * 64 + threadIdx.x; i < end; i += kBlock)
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const double v = dx[3 * i + d];
      acc[0] += v * v;
      xk[3 * i + d] = xk[3 * i + d] + v;
    }
  double* outs[1] = {TP(c, 0)};
  reduce_store(acc, 1, outs, t.L);
}

// |v|^2 per tile (slot k).
template <typename T>
__global__ __launch_bounds__(kBlock) void tz_sqnorm(TC c, const T* v, int slot) {
  TileCtx t;
  if (!tile_ctx(c, t)) return;
  const int end = min(t.s1 * 64, c.n_nodes);
  double acc[1] = {0.0};
  for (int i = t.s0 * 64 + threadIdx.x; i < end; i += kBlock)
#pragma unroll
    for (int d = 0; d < 3; ++d) acc[0] += (double)v[3 * i + d] * (double)v[3 * i + d];
  double* outs[1] = {TP(c, slot)};
  reduce_store(acc, 1, outs, t.L);
}

// out[T] = ordered sum of the tissue's tile partials of slot k; one workgroup
// per tissue.
__global__ __launch_bounds__(kBlock) void tz_tissue_sums(TC c, int slot, double* out) {
  const int T = blockIdx.x;
  if (T >= c.tm.n_tissues) return;
  const double s = tissue_sum(c, slot, T);
  if (threadIdx.x == 0) out[T] = s;
}

// bs = b / |b_T| per tissue (0 for tissues with |b_T| = 0 or not solved).
__global__ __launch_bounds__(kBlock) void tz_scale(TC c, const double* b, const double* norms,
                                                   double* bs) {
  TileCtx t;
  if (!tile_ctx(c, t)) return;
  const double nb = norms[t.T];
  const double f = (nb > 0.0 && c.active[t.T]) ? 1.0 / nb : 0.0;
  const int end = min(t.s1 * 64, c.n_nodes);
  for (int i = t.s0 * 64 + threadIdx.x; i < end; i += kBlock)
#pragma unroll
    for (int d = 0; d < 3; ++d) bs[3 * i + d] = b[3 * i + d] * f;
}

// x += |b_T| z per tissue.
__global__ __launch_bounds__(kBlock) void tz_axpy(TC c, const float* z, const double* norms,
                                                  double* x) {
  TileCtx t;
  if (!tile_ctx(c, t)) return;
  const double nb = c.active[t.T] ? norms[t.T] : 0.0;
  const int end = min(t.s1 * 64, c.n_nodes);
  for (int i = t.s0 * 64 + threadIdx.x; i < end; i += kBlock)
#pragma unroll
    for (int d = 0; d < 3; ++d) x[3 * i + d] = x[3 * i + d] + (double)z[3 * i + d] * nb;
}

// r = b - J x (FP32 values promoted, FP64 accumulation), per-tile |r|^2 in slot 1.
template <typename V>
__global__ __launch_bounds__(kBlock) void tz_residual(TC c, const V* __restrict__ val,
                                                      const double* __restrict__ x,
                                                      const double* __restrict__ b,
                                                      double* __restrict__ r) {
  TileCtx t;
  if (!tile_ctx(c, t)) return;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  double acc[1] = {0.0};
  for (int sl = t.s0 + wave; sl < t.s1; sl += kW) {
    const int64_t base = c.slice_base[sl];
    const int width = (int)((c.slice_base[sl + 1] - base) >> 6);
    double a0 = 0, a1 = 0, a2 = 0;
    for (int k = 0; k < width; ++k) {
      const int64_t o = base + (int64_t)k * 64 + lane;
      const int cc = c.col[o];
      const double x0 = x[3 * cc], x1 = x[3 * cc + 1], x2 = x[3 * cc + 2];
      const int64_t ns = c.n_slots;
      a0 += (double)val[o] * x0 + (double)val[ns + o] * x1 + (double)val[2 * ns + o] * x2;
      a1 += (double)val[3 * ns + o] * x0 + (double)val[4 * ns + o] * x1 + (double)val[5 * ns + o] * x2;
      a2 += (double)val[6 * ns + o] * x0 + (double)val[7 * ns + o] * x1 + (double)val[8 * ns + o] * x2;
    }
    const int row = sl * 64 + lane;
    if (row < c.n_nodes) {
      const double r0 = b[3 * row] - a0, r1 = b[3 * row + 1] - a1, r2 = b[3 * row + 2] - a2;
      r[3 * row] = r0; r[3 * row + 1] = r1; r[3 * row + 2] = r2;
      acc[0] += r0 * r0 + r1 * r1 + r2 * r2;
    }
  }
  double* outs[1] = {TP(c, 1)};
  reduce_store(acc, 1, outs, t.L);
}

// dst = src on the tissues being solved.
__global__ __launch_bounds__(kBlock) void tz_copy_active(TC c, const double* src, double* dst) {
  TileCtx t;
  if (!tile_ctx(c, t)) return;
  if (!c.active[t.T]) return;
  const int end = min(t.s1 * 64, c.n_nodes);
  for (int i = t.s0 * 64 + threadIdx.x; i < end; i += kBlock)
#pragma unroll
    for (int d = 0; d < 3; ++d) dst[3 * i + d] = src[3 * i + d];
}

// dst += src on the tissues being solved.
__global__ __launch_bounds__(kBlock) void tz_add_active(TC c, const double* src, double* dst) {
  TileCtx t;
  if (!tile_ctx(c, t)) return;
  if (!c.active[t.T]) return;
  const int end = min(t.s1 * 64, c.n_nodes);
  for (int i = t.s0 * 64 + threadIdx.x; i < end; i += kBlock)
#pragma unroll
    for (int d = 0; d < 3; ++d) dst[3 * i + d] += src[3 * i + d];
}

// Tissues that do not use their guess: x0 = 0, r0 = b.
__global__ __launch_bounds__(kBlock) void tz_guess_select(TC c, const int32_t* use,
                                                          const double* b, double* x0,
                                                          double* r0) {
  TileCtx t;
  if (!tile_ctx(c, t)) return;
  if (use[t.T]) return;
  const int end = min(t.s1 * 64, c.n_nodes);
  for (int i = t.s0 * 64 + threadIdx.x; i < end; i += kBlock)
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      x0[3 * i + d] = 0.0;
      r0[3 * i + d] = b[3 * i + d];
    }
}

// x0 from the displacements of a tissue's last steps: 0 (no history), h0
// (one step) or 2 h0 - h1 (two steps).
__global__ __launch_bounds__(kBlock) void tz_extrapolate(TC c, const int32_t* order,
                                                         const double* h0, const double* h1,
                                                         double* x0) {
  TileCtx t;
  if (!tile_ctx(c, t)) return;
  const int o = order[t.T];
  const int end = min(t.s1 * 64, c.n_nodes);
  for (int i = 3 * (t.s0 * 64) + threadIdx.x; i < 3 * end; i += kBlock)
    x0[i] = o >= 2 ? 2.0 * h0[i] - h1[i] : (o == 1 ? h0[i] : 0.0);
}

__global__ __launch_bounds__(kBlock) void difference_kernel(int64_t n, const double* a,
                                                            const double* b, double* d) {
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kBlock)
    d[i] = a[i] - b[i];
}

__global__ __launch_bounds__(kBlock) void tz_fill(TC c, double* v, double value) {
  TileCtx t;
  if (!tile_ctx(c, t)) return;
  const int end = min(t.s1 * 64, c.n_nodes);
  for (int i = t.s0 * 64 + threadIdx.x; i < end; i += kBlock)
#pragma unroll
    for (int d = 0; d < 3; ++d) v[3 * i + d] = value;
}

TC make_tc(const KrylovArgs& k, double rtol) {
  TC c;
  std::memset(&c, 0, sizeof(c));  // padding too: graphs are matched by memcmp
  c.tm = k.tm;
  c.n_nodes = k.n_nodes;
  c.slice_base = k.pat.slice_base;
  c.col = k.pat.col;
  c.tslot = k.pat.tslot;
  c.n_slots = k.pat.n_slots;
  c.state = k.ws->tstate;
  c.scal = k.ws->tscal;
  c.part = k.ws->tpart;
  c.stride = k.ws->tpart_stride;
  c.active = k.ws->tactive;
  c.rtol = rtol;
  c.rtol_scale = k.rtol_scale;
  return c;
}

template <typename T>
struct TV {
  T *r, *rh, *p, *v, *ph, *s, *sh, *t, *x, *w;
  const T* val;
  const T* minv;
};

// One chunk of `chunk` iterations (4 launches each), launched directly.
template <typename T>
void launch_chunk(const TC& c, const TV<T>& V, int G, hipStream_t st, int chunk, Prof* pf) {
  // profiled launches carry their own start/stop events (recorded by the
  // runtime at the kernel's start and end)
  auto timed = [&](int which, auto launch) {
    const bool on = pf && pf->on && pf->used < kProfPairs;
    if (on) {
      pf->kind[pf->used] = 2 * which + (sizeof(T) == 4 ? 1 : 0);
      launch(pf->ev[2 * pf->used], pf->ev[2 * pf->used + 1]);
      pf->used++;
    } else {
      launch(nullptr, nullptr);
    }
  };
  for (int it = 0; it < chunk; ++it) {
    if (pf && pf->on && pf->used + 2 > kProfPairs) {
      (void)hipStreamSynchronize(st);
      prof_harvest(pf);
    }
    timed(0, [&](hipEvent_t e0, hipEvent_t e1) {
      hipExtLaunchKernelGGL(tz_spmv_v<T>, dim3(G), dim3(kBlock), 0, st, e0, e1, 0, c, V.val,
                            (const T*)V.ph, (const T*)V.r, V.minv, V.v, V.s, V.sh);
    });
    timed(1, [&](hipEvent_t e0, hipEvent_t e1) {
      hipExtLaunchKernelGGL(tz_spmv_t<T>, dim3(G), dim3(kBlock), 0, st, e0, e1, 0, c, V.val,
                            (const T*)V.sh, (const T*)V.s, V.t);
    });
    hipLaunchKernelGGL(tz_update_x<T>, dim3(G), dim3(kBlock), 0, st, c, V.x, V.r, (const T*)V.rh,
                       (const T*)V.ph, (const T*)V.sh, (const T*)V.s, (const T*)V.t);
    hipLaunchKernelGGL(tz_update_p<T>, dim3(G), dim3(kBlock), 0, st, c, (const T*)V.r, V.p, V.ph,
                       (const T*)V.v, (const T*)V.w, V.minv);
  }
}

// Captured chunks: a full chunk of iterations is recorded once into a
// hipGraph per key (solver, precision, kernel arguments, chunk length, parity
// of the first iteration) and replayed with one hipGraphLaunch.  The kernels
// read every per-tissue scalar and flag from device memory, so a replay is
// the same work as the direct launches.  A chunk holds kernel launches only
// (no memset or copy nodes).
struct GraphKey {
  int kind, tsize, chunk, parity;
  TC c;
  const void* ptrs[12];
};
struct ChunkGraph {
  hipGraphExec_t exec = nullptr;
  GraphKey key;
  uint64_t used = 0;
};
constexpr int kGraphSlots = 6;
struct GraphCache {
  ChunkGraph g[kGraphSlots];
  uint64_t clock = 0;
};

// Launches `launch(prof)` directly when graphs are off or profiling is on;
// otherwise replays (capturing first if needed) the graph of `key`.
template <typename F>
hipError_t run_captured(const KrylovArgs& k, const GraphKey& key, F launch) {
  Workspace* ws = k.ws;
  hipStream_t st = k.L.stream;
  if (!ws->graphs_on || (k.prof && k.prof->on)) {
    launch(k.prof);
    return hipGetLastError();
  }
  GraphCache* gc = static_cast<GraphCache*>(ws->graphs);
  if (!gc) {
    gc = new GraphCache();
    ws->graphs = gc;
  }
  ChunkGraph* hit = nullptr;
  for (auto& g : gc->g)
    if (g.exec && std::memcmp(&g.key, &key, sizeof(GraphKey)) == 0) {
      hit = &g;
      break;
    }
  hipError_t e;
  if (!hit) {
    hit = &gc->g[0];
    for (auto& g : gc->g)
      if (g.used < hit->used) hit = &g;
    if (hit->exec) (void)hipGraphExecDestroy(hit->exec);
    hit->exec = nullptr;
    if ((e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal)) != hipSuccess) return e;
    launch(static_cast<Prof*>(nullptr));
    hipGraph_t graph = nullptr;
    hipError_t le = hipGetLastError();
    e = hipStreamEndCapture(st, &graph);
    if (le != hipSuccess) e = le;
    if (e == hipSuccess) e = hipGraphInstantiate(&hit->exec, graph, nullptr, nullptr, 0);
    if (graph) (void)hipGraphDestroy(graph);
    if (e != hipSuccess) {
      hit->exec = nullptr;
      return e;
    }
    hit->key = key;
  }
  hit->used = ++gc->clock;
  return hipGraphLaunch(hit->exec, st);
}

GraphKey make_key(int kind, int tsize, int chunk, int parity, const TC& c, const void* const* ptrs,
                  int n_ptrs) {
  GraphKey key;
  std::memset(&key, 0, sizeof(key));  // padding too: keys are compared with memcmp
  key.kind = kind;
  key.tsize = tsize;
  key.chunk = chunk;
  key.parity = parity;
  key.c = c;
  for (int i = 0; i < n_ptrs; ++i) key.ptrs[i] = ptrs[i];
  return key;
}

template <typename T>
hipError_t run_chunk(const KrylovArgs& k, const TC& c, const TV<T>& V, int chunk) {
  const void* ptrs[12] = {V.r, V.rh, V.p, V.v, V.ph, V.s, V.sh, V.t, V.x, V.w, V.val, V.minv};
  const GraphKey key = make_key(0, (int)sizeof(T), chunk, 0, c, ptrs, 12);
  const int G = k.tm.grid;
  hipStream_t st = k.L.stream;
  return run_captured(k, key, [&](Prof* pf) { launch_chunk<T>(c, V, G, st, chunk, pf); });
}

// CG vectors: r, z = M^-1 r, x, the two p buffers, q = J p.
template <typename T>
struct CV {
  T *r, *z, *x, *p0, *p1, *q;
  const T* val;
  const T* minv;
  const CoarseOp* co = nullptr;  // z = D^-1 r + P A_c^-1 P^T r (FP64: with the FP64 inverse)
  const uint8_t* flags = nullptr;
  int32_t n_nodes = 0;
};

CgS1 host_s1(const TC& c) {
  CgS1 s;
  s.rz = c.part + (size_t)TC_RZ0 * c.stride;
  s.rr = c.part + (size_t)TC_RR0 * c.stride;
  s.tissue_tile = c.tm.tissue_tile;
  s.scal = c.scal;
  s.state = c.state;
  s.rtol = c.rtol;
  s.rtol_scale = c.rtol_scale;
  return s;
}

// The coarse half of a CG iteration (with a coarse space): restriction,
// product, then S1 and p = (z_D + P y) + beta p.
template <typename T>
void coarse_half(const TC& c, const CV<T>& V, int G, hipStream_t st) {
  CgS1 s1 = host_s1(c);
  S1Pre pre;
  pre.rz = s1.rz;
  pre.rr = s1.rr;
  pre.part_off = s1.tissue_tile;
  pre.cpart_off = V.co->cpart_off;
  pre.out = V.co->s1pre;
  coarse_apply(*V.co, c.state, (const T*)V.r, pre, st);
  s1.cpart = V.co->cpart;
  s1.cpart_off = V.co->cpart_off;
  s1.pre = V.co->s1pre;
  const T* y;
  if constexpr (std::is_same<T, float>::value) y = V.co->y;
  else y = V.co->y64;
  const CoarseP<T> cp{V.co->node_cb, V.co->seg_on, V.co->node_d, y};
  hipLaunchKernelGGL(tzc_pupdate<T>, dim3(G), dim3(kBlock), 0, st, c, s1, cp, (const T*)V.z,
                     V.p0);
}

// The coarse half of a CG iteration whose update formed w: product, then S1
// (from the per-aggregate partials) and p = (z_D + P y) + beta p.
template <typename T>
void coarse_half_fused(const TC& c, const CV<T>& V, int G, hipStream_t st) {
  CgS1 s1 = host_s1(c);
  S1Pre pre;
  pre.rz = V.co->agg_rz;
  pre.rr = V.co->agg_rr;
  pre.part_off = V.co->seg_half_d;
  pre.cpart_off = V.co->cpart_off;
  pre.out = V.co->s1pre;
  coarse_product(*V.co, c.state, pre, (T*)nullptr, st);
  s1.cpart = V.co->cpart;
  s1.cpart_off = V.co->cpart_off;
  s1.pre = V.co->s1pre;
  const T* y;
  if constexpr (std::is_same<T, float>::value) y = V.co->y;
  else y = V.co->y64;
  const CoarseP<T> cp{V.co->node_cb, V.co->seg_on, V.co->node_d, y};
  hipLaunchKernelGGL(tzc_pupdate<T>, dim3(G), dim3(kBlock), 0, st, c, s1, cp, (const T*)V.z,
                     V.p0);
}

template <typename T>
CoarseU<T> coarse_u(const CoarseOp& co) {
  CoarseU<T> u;
  u.agg_ptr = co.agg_ptr;
  u.agg_nodes = co.agg_nodes;
  u.agg_tissue = co.agg_tissue;
  u.agg_cbase = co.agg_cbase;
  u.seg_agg = co.seg_agg_d;
  u.list_d = co.list_d;
  u.list_mask = co.list_mask;
  if constexpr (std::is_same<T, float>::value) {
    u.w = co.w;
    u.w2 = co.w2;
  } else {
    u.w = co.w64;
    u.w2 = nullptr;  // the FP64 CG keeps the tile update and the restriction launch
  }
  u.rz = co.agg_rz;
  u.rr = co.agg_rr;
  u.n_agg = co.n_agg;
  return u;
}

// The FP32 CG with the coarse space keeps x, r and the block-Jacobi inverse of
// the nodes of aggregates in list order: after tzc_init, r and the inverse are
// copied in and x starts at zero; at the end x goes back to node order (the
// rows of no aggregate keep tzc_init's zero).
__global__ __launch_bounds__(kBlock) void tzc_list_in(int32_t n_list,
                                                      const int32_t* __restrict__ nodes,
                                                      const float* __restrict__ r,
                                                      const float* __restrict__ minv,
                                                      float* __restrict__ xl, float* __restrict__ rl,
                                                      float* __restrict__ ml) {
  for (int t = blockIdx.x * kBlock + threadIdx.x; t < n_list; t += gridDim.x * kBlock) {
    const int i = nodes[t];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      rl[3 * t + d] = r[3 * i + d];
      xl[3 * t + d] = 0.f;
    }
#pragma unroll
    for (int e = 0; e < 9; ++e) ml[9 * (size_t)t + e] = minv[9 * (size_t)i + e];
  }
}

__global__ __launch_bounds__(kBlock) void tzc_list_out(int32_t n_list,
                                                       const int32_t* __restrict__ nodes,
                                                       const float* __restrict__ xl,
                                                       float* __restrict__ x) {
  for (int t = blockIdx.x * kBlock + threadIdx.x; t < n_list; t += gridDim.x * kBlock) {
    const int i = nodes[t];
#pragma unroll
    for (int d = 0; d < 3; ++d) x[3 * i + d] = xl[3 * t + d];
  }
}

inline int list_grid(int32_t n) { return std::max(1, std::min((n + kBlock - 1) / kBlock, 8192)); }

// r = b, x = 0, p = 0, z; with the coarse space also the first S1 and p.
template <typename T>
void cg_start(const TC& c, const CV<T>& V, const double* b, int G, hipStream_t st) {
  hipLaunchKernelGGL(tzc_init<T>, dim3(G), dim3(kBlock), 0, st, c, b, V.r, V.z, V.x, V.p0,
                     V.minv);
  if (V.co) {
    if constexpr (sizeof(T) == 4)
      hipLaunchKernelGGL(tzc_list_in, dim3(list_grid(V.co->agg_total)), dim3(kBlock), 0, st,
                         V.co->agg_total, V.co->agg_nodes, (const float*)V.r,
                         (const float*)V.minv, V.co->xl, V.co->rl, V.co->ml);
    coarse_half<T>(c, V, G, st);
  }
}

template <typename T>
void launch_cg_chunk(const TC& c, const CV<T>& V, int G, hipStream_t st, int chunk, int it0,
                     Prof* pf) {
  for (int it = 0; it < chunk; ++it) {
    if (pf && pf->on && pf->used + 1 > kProfPairs) {
      (void)hipStreamSynchronize(st);
      prof_harvest(pf);
    }
    const bool on = pf && pf->on && pf->used < kProfPairs;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    if (on) {
      pf->kind[pf->used] = sizeof(T) == 4 ? 1 : 0;
      ev0 = pf->ev[2 * pf->used];
      ev1 = pf->ev[2 * pf->used + 1];
      pf->used++;
    }
    if (V.co) {
      hipExtLaunchKernelGGL(tzc_spmv<T, false>, dim3(G), dim3(kBlock), 0, st, ev0, ev1, 0, c,
                            V.val, (const T*)V.z, (const T*)V.p0, (T*)nullptr, V.q);
      if constexpr (sizeof(T) == 4) {
        // FP32: the update walks the aggregate lists and forms the restriction
        hipLaunchKernelGGL(tzc_update_agg<T>, dim3(xcd_grid(2 * V.co->n_agg)), dim3(kUBlock), 0, st,
                           c, coarse_u<T>(*V.co), V.co->xl, V.co->rl, V.z, (const T*)V.p0,
                           (const T*)V.q, (const T*)V.co->ml);
        coarse_half_fused<T>(c, V, G, st);
      } else {
        // FP64: the list-order gathers of 8-byte rows cost the update more than the
        // restriction's own launch (72 against 44 + 19 us at C3)
        hipLaunchKernelGGL(tzc_update<T>, dim3(G), dim3(kBlock), 0, st, c, V.x, V.r, V.z,
                           (const T*)V.p0, (const T*)V.q, V.minv);
        coarse_half<T>(c, V, G, st);
      }
    } else {
      const int par = (it0 + it) & 1;
      const T* pold = par ? V.p1 : V.p0;
      T* pnew = par ? V.p0 : V.p1;
      hipExtLaunchKernelGGL(tzc_spmv<T, true>, dim3(G), dim3(kBlock), 0, st, ev0, ev1, 0, c,
                            V.val, (const T*)V.z, pold, pnew, V.q);
      hipLaunchKernelGGL(tzc_update<T>, dim3(G), dim3(kBlock), 0, st, c, V.x, V.r, V.z,
                         (const T*)pnew, (const T*)V.q, V.minv);
    }
  }
}

template <typename T>
hipError_t tiled_cg_impl(const KrylovArgs& k, const CV<T>& V, const double* b, double rtol,
                         int maxit, int sync_every, TiledResult* res) {
  Workspace* ws = k.ws;
  hipStream_t st = k.L.stream;
  const int G = k.tm.grid;
  const int nt = k.tm.n_tissues;
  const TC c = make_tc(k, rtol);
  cg_start<T>(c, V, b, G, st);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const void* ptrs[12] = {V.r, V.z, V.x, V.p0, V.p1, V.q, V.val, V.minv,
                          V.co ? (sizeof(T) == 4 ? (const void*)V.co->inv32 : (const void*)V.co->inv64)
                               : nullptr,
                          V.co ? V.co->node_d : nullptr,
                          V.co ? V.co->agg_nodes : nullptr,
                          V.co ? (const void*)(intptr_t)V.co->gen : nullptr};
  int done = 0;
  while (true) {
    const int chunk = std::min(sync_every, maxit - done);
    // the p buffers alternate only without the coarse space
    const GraphKey key = make_key(1, (int)sizeof(T), chunk, V.co ? 0 : (done & 1), c, ptrs, 12);
    const int it0 = done;
    e = run_captured(k, key, [&](Prof* pf) { launch_cg_chunk<T>(c, V, G, st, chunk, it0, pf); });
    if (e != hipSuccess) return e;
    done += chunk;
    if (done >= maxit && !V.co) {  // one more convergence test on the last residuals
      const int par = done & 1;
      hipLaunchKernelGGL((tzc_spmv<T, true>), dim3(G), dim3(kBlock), 0, st, c, V.val,
                         (const T*)V.z, (const T*)(par ? V.p1 : V.p0), par ? V.p0 : V.p1, V.q);
    }
    // with the coarse space every iteration ends with the next one's S1
    if ((e = hipMemcpyAsync(ws->host_tstate, ws->tstate, 2 * sizeof(int32_t) * nt,
                            hipMemcpyDeviceToHost, st)) != hipSuccess)
      return e;
    if ((e = hipStreamSynchronize(st)) != hipSuccess) return e;
    prof_harvest(k.prof);
    bool running = false;
    for (int q = 0; q < nt; ++q) running |= (ws->host_tstate[2 * q] == K_RUN);
    if (!running || done >= maxit) break;
  }
  if constexpr (sizeof(T) == 4) {
    if (V.co) {
      hipLaunchKernelGGL(tzc_list_out, dim3(list_grid(V.co->agg_total)), dim3(kBlock), 0, st,
                         V.co->agg_total, V.co->agg_nodes, (const float*)V.co->xl, (float*)V.x);
      if ((e = hipGetLastError()) != hipSuccess) return e;
    }
  }
  for (int q = 0; q < nt; ++q) {
    res->state[q] = ws->host_tstate[2 * q];
    res->iters[q] = ws->host_tstate[2 * q + 1];
  }
  return hipSuccess;
}

template <typename T>
hipError_t tiled_bicgstab_impl(const KrylovArgs& k, const TV<T>& V, const double* b, double rtol,
                               int maxit, int sync_every, TiledResult* res) {
  Workspace* ws = k.ws;
  hipStream_t st = k.L.stream;
  const int G = k.tm.grid;
  const int nt = k.tm.n_tissues;
  TC c = make_tc(k, rtol);
  hipLaunchKernelGGL(tz_init_a<T>, dim3(G), dim3(kBlock), 0, st, c, b, V.r, V.rh, V.p, V.x);
  hipLaunchKernelGGL(tz_spmv_transposed<T>, dim3(G), dim3(kBlock), 0, st, c, V.val,
                     (const T*)V.rh, V.w);
  hipLaunchKernelGGL(tz_init_b<T>, dim3(G), dim3(kBlock), 0, st, c, (const T*)V.r,
                     (const T*)V.w, (const T*)V.p, V.ph, V.minv);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  Prof* pf = k.prof;
  int done = 0;
  while (true) {
    const int chunk = std::min(sync_every, maxit - done);
    if ((e = run_chunk<T>(k, c, V, chunk)) != hipSuccess) return e;
    done += chunk;
    if (done >= maxit)  // one more convergence test on the last residuals
      hipLaunchKernelGGL(tz_spmv_v<T>, dim3(G), dim3(kBlock), 0, st, c, V.val, (const T*)V.ph,
                         (const T*)V.r, V.minv, V.v, V.s, V.sh);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if ((e = hipMemcpyAsync(ws->host_tstate, ws->tstate, 2 * sizeof(int32_t) * nt,
                            hipMemcpyDeviceToHost, st)) != hipSuccess)
      return e;
    if ((e = hipStreamSynchronize(st)) != hipSuccess) return e;
    prof_harvest(pf);
    bool running = false;
    for (int q = 0; q < nt; ++q) running |= (ws->host_tstate[2 * q] == K_RUN);
    if (!running || done >= maxit) break;
  }
  for (int q = 0; q < nt; ++q) {
    res->state[q] = ws->host_tstate[2 * q];
    res->iters[q] = ws->host_tstate[2 * q + 1];
  }
  return hipSuccess;
}

}  // namespace

hipError_t tiled_bicgstab(const KrylovArgs& k, const double* b, double* x, double rtol, int maxit,
                          int sync_every, TiledResult* res) {
  Workspace* ws = k.ws;
  TV<double> V{ws->r, ws->rh, ws->p, ws->v, ws->ph, ws->s, ws->sh, ws->t, x, ws->w, k.val, k.minv};
  return tiled_bicgstab_impl<double>(k, V, b, rtol, maxit, sync_every, res);
}

hipError_t tiled_cg(const KrylovArgs& k, const double* b, double* x, double rtol, int maxit,
                    int sync_every, TiledResult* res) {
  Workspace* ws = k.ws;
  CV<double> V{ws->r, ws->s, x, ws->p, ws->ph, ws->v, k.val, k.minv};
  if (k.coarse && k.coarse->active && k.coarse->inv64) {
    V.co = k.coarse;
    V.flags = k.dof_flags;
    V.n_nodes = k.n_nodes;
  }
  return tiled_cg_impl<double>(k, V, b, rtol, maxit, sync_every, res);
}

hipError_t tiled_mixed(const KrylovArgs& k, const double* b, double* x, double rtol,
                       double inner_rtol, int maxit, int sync_every, int method, TiledResult* res) {
  Workspace* ws = k.ws;
  hipStream_t st = k.L.stream;
  const int G = k.tm.grid;
  const int nt = k.tm.n_tissues;
  TC c = make_tc(k, inner_rtol);
  // x = 0 and |b_T|
  hipLaunchKernelGGL(tz_sqnorm<double>, dim3(G), dim3(kBlock), 0, st, c, b, 1);
  hipLaunchKernelGGL(tz_tissue_sums, dim3(nt), dim3(kBlock), 0, st, c, 1, ws->tnorm);
  hipError_t e = hipMemsetAsync(x, 0, sizeof(double) * 3 * (size_t)k.n_nodes, st);
  if (e != hipSuccess) return e;
  if ((e = hipMemcpyAsync(ws->host_tnorm, ws->tnorm, sizeof(double) * nt, hipMemcpyDeviceToHost,
                          st)) != hipSuccess)
    return e;
  if ((e = hipStreamSynchronize(st)) != hipSuccess) return e;
  std::vector<double> bnorm(nt), rnorm(nt);
  std::vector<int32_t> act(nt), act0(nt);
  if ((e = hipMemcpy(act0.data(), ws->tactive, sizeof(int32_t) * nt, hipMemcpyDeviceToHost)) !=
      hipSuccess)
    return e;
  for (int T = 0; T < nt; ++T) {
    bnorm[T] = sqrt(ws->host_tnorm[T]);
    rnorm[T] = bnorm[T];
    res->state[T] = K_CONVERGED;
    res->iters[T] = 0;
    act[T] = act0[T] && bnorm[T] > 0.0;
  }
  const double* rsrc = b;
  TV<float> V{ws->r32, ws->rh32, ws->p32, ws->v32, ws->ph32, ws->s32, ws->sh32, ws->t32, ws->x32,
              ws->w32, k.val32, k.minv32};
  TiledResult inner;
  for (int outer = 0; outer < 20; ++outer) {
    bool any = false;
    for (int T = 0; T < nt; ++T) {
      any |= act[T] != 0;
      ws->host_tnorm[T] = rnorm[T];
    }
    if (!any) break;
    if ((e = hipMemcpyAsync(ws->tactive, act.data(), sizeof(int32_t) * nt, hipMemcpyHostToDevice,
                            st)) != hipSuccess)
      return e;
    if ((e = hipMemcpyAsync(ws->tnorm, ws->host_tnorm, sizeof(double) * nt, hipMemcpyHostToDevice,
                            st)) != hipSuccess)
      return e;
    // scale each tissue's residual to O(1) before the inner solve rounds it to FP32
    hipLaunchKernelGGL(tz_scale, dim3(G), dim3(kBlock), 0, st, c, rsrc, ws->tnorm, ws->b);
    int left = maxit;
    for (int T = 0; T < nt; ++T) left = std::min(left, maxit - res->iters[T]);
    if (method == KRYLOV_CG) {
      CV<float> C{ws->r32, ws->s32, ws->x32, ws->p32, ws->ph32, ws->v32, k.val32, k.minv32};
      if (k.coarse && k.coarse->active) {
        C.co = k.coarse;
        C.flags = k.dof_flags;
        C.n_nodes = k.n_nodes;
      }
      e = tiled_cg_impl<float>(k, C, ws->b, inner_rtol, std::max(left, 1), sync_every, &inner);
    } else {
      e = tiled_bicgstab_impl<float>(k, V, ws->b, inner_rtol, std::max(left, 1), sync_every,
                                     &inner);
    }
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(tz_axpy, dim3(G), dim3(kBlock), 0, st, c, ws->x32, ws->tnorm, x);
    for (int T = 0; T < nt; ++T)
      if (act[T]) {
        res->iters[T] += inner.iters[T];
        res->state[T] = inner.state[T];
      }
    if (rtol >= inner_rtol) break;  // the NR loop (FP64 residual) is the refinement
    hipLaunchKernelGGL(tz_residual<float>, dim3(G), dim3(kBlock), 0, st, c, k.val32, x, b, ws->r);
    hipLaunchKernelGGL(tz_tissue_sums, dim3(nt), dim3(kBlock), 0, st, c, 1, ws->tnorm);
    if ((e = hipMemcpyAsync(ws->host_tnorm, ws->tnorm, sizeof(double) * nt,
                            hipMemcpyDeviceToHost, st)) != hipSuccess)
      return e;
    if ((e = hipStreamSynchronize(st)) != hipSuccess) return e;
    rsrc = ws->r;
    for (int T = 0; T < nt; ++T) {
      if (!act[T]) continue;
      rnorm[T] = sqrt(ws->host_tnorm[T]);
      const double rt = k.rtol_scale_h ? rtol * k.rtol_scale_h[T] : rtol;
      if (!(rnorm[T] > rt * bnorm[T])) { res->state[T] = K_CONVERGED; act[T] = 0; }
      else if (!std::isfinite(rnorm[T]) || inner.state[T] == K_BREAKDOWN) {
        res->state[T] = K_BREAKDOWN; act[T] = 0;
      } else if (res->iters[T] >= maxit) { res->state[T] = K_RUN; act[T] = 0; }
      else res->state[T] = K_RUN;
    }
  }
  // restore the caller's activity mask (act0 is pageable and local: wait for the copy)
  if ((e = hipMemcpyAsync(ws->tactive, act0.data(), sizeof(int32_t) * nt, hipMemcpyHostToDevice,
                          st)) != hipSuccess)
    return e;
  return hipStreamSynchronize(st);
}

hipError_t tiled_copy_active(const KrylovArgs& k, const double* src, double* dst) {
  TC c = make_tc(k, 0.0);
  hipLaunchKernelGGL(tz_copy_active, dim3(k.tm.grid), dim3(kBlock), 0, k.L.stream, c, src, dst);
  return hipGetLastError();
}

hipError_t tiled_add_active(const KrylovArgs& k, const double* src, double* dst) {
  TC c = make_tc(k, 0.0);
  hipLaunchKernelGGL(tz_add_active, dim3(k.tm.grid), dim3(kBlock), 0, k.L.stream, c, src, dst);
  return hipGetLastError();
}

hipError_t tiled_warm_start(const KrylovArgs& k, int mixed, const double* b, double* x0,
                            double* r0, double* scale_d, double* scale_h) {
  Workspace* ws = k.ws;
  hipStream_t st = k.L.stream;
  const int G = k.tm.grid;
  const int nt = k.tm.n_tissues;
  TC c = make_tc(k, 0.0);
  if (mixed)
    hipLaunchKernelGGL(tz_residual<float>, dim3(G), dim3(kBlock), 0, st, c, k.val32, x0, b, r0);
  else
    hipLaunchKernelGGL(tz_residual<double>, dim3(G), dim3(kBlock), 0, st, c, k.val, x0, b, r0);
  hipLaunchKernelGGL(tz_tissue_sums, dim3(nt), dim3(kBlock), 0, st, c, 1, ws->tnorm);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  if ((e = hipMemcpyAsync(ws->host_tnorm, ws->tnorm, sizeof(double) * nt, hipMemcpyDeviceToHost,
                          st)) != hipSuccess)
    return e;
  if ((e = hipMemcpyAsync(ws->host_gnorm, ws->gnorm, sizeof(double) * nt, hipMemcpyDeviceToHost,
                          st)) != hipSuccess)
    return e;
  if ((e = hipStreamSynchronize(st)) != hipSuccess) return e;
  std::vector<int32_t> use(nt, 0);
  bool all = true;
  for (int T = 0; T < nt; ++T) {
    const double bb = ws->host_gnorm[T], rr = ws->host_tnorm[T];
    // the guess must leave at most half of |b_T| (|r0|^2 <= |b|^2 / 4)
    use[T] = std::isfinite(rr) && std::isfinite(bb) && bb > 0.0 && rr > 0.0 && 4.0 * rr <= bb;
    scale_h[T] = use[T] ? std::sqrt(bb / rr) : 1.0;
    all &= use[T] != 0;
  }
  if ((e = hipMemcpyAsync(scale_d, scale_h, sizeof(double) * nt, hipMemcpyHostToDevice, st)) !=
      hipSuccess)
    return e;
  if (!all) {
    if ((e = hipMemcpyAsync(ws->guess_use, use.data(), sizeof(int32_t) * nt,
                            hipMemcpyHostToDevice, st)) != hipSuccess)
      return e;
    hipLaunchKernelGGL(tz_guess_select, dim3(G), dim3(kBlock), 0, st, c,
                       (const int32_t*)ws->guess_use, b, x0, r0);
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  // use is pageable and local: wait for its copy
  return hipStreamSynchronize(st);
}

hipError_t tiled_extrapolate(const KrylovArgs& k, const int32_t* order, const double* h0,
                             const double* h1, double* x0) {
  TC c = make_tc(k, 0.0);
  hipLaunchKernelGGL(tz_extrapolate, dim3(k.tm.grid), dim3(kBlock), 0, k.L.stream, c, order, h0,
                     h1, x0);
  return hipGetLastError();
}

hipError_t launch_difference(int64_t n, const double* a, const double* b, double* d,
                             hipStream_t st) {
  const int64_t g = std::min<int64_t>((n + kBlock - 1) / kBlock, 65536);
  hipLaunchKernelGGL(difference_kernel, dim3((unsigned)std::max<int64_t>(g, 1)), dim3(kBlock), 0,
                     st, n, a, b, d);
  return hipGetLastError();
}

hipError_t tiled_norm_update(const KrylovArgs& k, const double* dx, double* xk, double* norms) {
  Workspace* ws = k.ws;
  hipStream_t st = k.L.stream;
  TC c = make_tc(k, 0.0);
  const int nt = k.tm.n_tissues;
  hipLaunchKernelGGL(tz_norm_update, dim3(k.tm.grid), dim3(kBlock), 0, st, c, dx, xk);
  hipLaunchKernelGGL(tz_tissue_sums, dim3(nt), dim3(kBlock), 0, st, c, 0, ws->tnorm);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  if ((e = hipMemcpyAsync(ws->host_tnorm, ws->tnorm, sizeof(double) * nt, hipMemcpyDeviceToHost,
                          st)) != hipSuccess)
    return e;
  if ((e = hipStreamSynchronize(st)) != hipSuccess) return e;
  for (int T = 0; T < nt; ++T) norms[T] = sqrt(ws->host_tnorm[T]);
  return hipSuccess;
}

// Times `iters` iterations of the given method (FP64 or FP32 operator) on the
// current operator, for the tile-size autotuner.
hipError_t tiled_time_iterations(const KrylovArgs& k, int mixed, int method, int iters,
                                 float* ms) {
  Workspace* ws = k.ws;
  hipStream_t st = k.L.stream;
  const int G = k.tm.grid;
  TC c = make_tc(k, 0.0);  // rtol 0: never converges during the timing
  hipEvent_t e0, e1;
  hipError_t e;
  if ((e = hipEventCreate(&e0)) != hipSuccess) return e;
  if ((e = hipEventCreate(&e1)) != hipSuccess) return e;
  hipLaunchKernelGGL(tz_fill, dim3(G), dim3(kBlock), 0, st, c, ws->b, 1.0);
  auto run = [&](auto V) {
    using T = std::remove_pointer_t<decltype(V.r)>;
    hipLaunchKernelGGL(tz_init_a<T>, dim3(G), dim3(kBlock), 0, st, c, (const double*)ws->b, V.r,
                       V.rh, V.p, V.x);
    hipLaunchKernelGGL(tz_spmv_transposed<T>, dim3(G), dim3(kBlock), 0, st, c, V.val,
                       (const T*)V.rh, V.w);
    hipLaunchKernelGGL(tz_init_b<T>, dim3(G), dim3(kBlock), 0, st, c, (const T*)V.r,
                       (const T*)V.w, (const T*)V.p, V.ph, V.minv);
    (void)hipEventRecord(e0, st);
    for (int it = 0; it < iters; ++it) {
      hipLaunchKernelGGL(tz_spmv_v<T>, dim3(G), dim3(kBlock), 0, st, c, V.val, (const T*)V.ph,
                         (const T*)V.r, V.minv, V.v, V.s, V.sh);
      hipLaunchKernelGGL(tz_spmv_t<T>, dim3(G), dim3(kBlock), 0, st, c, V.val, (const T*)V.sh,
                         (const T*)V.s, V.t);
      hipLaunchKernelGGL(tz_update_x<T>, dim3(G), dim3(kBlock), 0, st, c, V.x, V.r,
                         (const T*)V.rh, (const T*)V.ph, (const T*)V.sh, (const T*)V.s,
                         (const T*)V.t);
      hipLaunchKernelGGL(tz_update_p<T>, dim3(G), dim3(kBlock), 0, st, c, (const T*)V.r, V.p,
                         V.ph, (const T*)V.v, (const T*)V.w, V.minv);
    }
    (void)hipEventRecord(e1, st);
  };
  auto run_cg = [&](auto V) {
    using T = std::remove_pointer_t<decltype(V.r)>;
    cg_start<T>(c, V, (const double*)ws->b, G, st);
    (void)hipEventRecord(e0, st);
    launch_cg_chunk<T>(c, V, G, st, iters, 0, nullptr);
    (void)hipEventRecord(e1, st);
  };
  if (method == KRYLOV_CG) {
    if (mixed) {
      CV<float> V{ws->r32, ws->s32, ws->x32, ws->p32, ws->ph32, ws->v32, k.val32, k.minv32};
      if (k.coarse && k.coarse->active) {
        V.co = k.coarse;
        V.flags = k.dof_flags;
        V.n_nodes = k.n_nodes;
      }
      run_cg(V);
    } else {
      CV<double> V{ws->r, ws->s, ws->x, ws->p, ws->ph, ws->v, k.val, k.minv};
      run_cg(V);
    }
  } else if (mixed) {
    TV<float> V{ws->r32, ws->rh32, ws->p32, ws->v32, ws->ph32, ws->s32, ws->sh32, ws->t32,
                ws->x32, ws->w32, k.val32, k.minv32};
    run(V);
  } else {
    TV<double> V{ws->r, ws->rh, ws->p, ws->v, ws->ph, ws->s, ws->sh, ws->t, ws->x, ws->w, k.val,
                 k.minv};
    run(V);
  }
  if ((e = hipStreamSynchronize(st)) != hipSuccess) return e;
  e = hipEventElapsedTime(ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return e;
}

// ---- restarted GMRES(m), FP64, right block-Jacobi preconditioning ----
// The last resort of the NR loop's solver chain: it needs no positive
// definiteness and cannot break down short of exact convergence, so it solves
// the indefinite Jacobians that node-node packing produces (a repulsive
// central force has a negative tangential stiffness) where BiCGStab breaks
// down.  Arnoldi with classical Gram-Schmidt applied twice, on the tiles and
// per-tissue ordered sums of the other solvers; Givens rotations and the small
// triangular solves of every tissue run on the host.  Partial slots: 0..m the
// projections (v_i, w), m + 1 the squared norm.
namespace {

__global__ __launch_bounds__(kBlock) void tg_residual(TC c, const int32_t* gact,
                                                      const double* __restrict__ val,
                                                      const double* __restrict__ x,
                                                      const double* __restrict__ b,
                                                      double* __restrict__ r, int x_zero,
                                                      int slot) {
  TileCtx t;
  if (!tile_ctx(c, t) || !gact[t.T]) return;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  double acc[1] = {0.0};
  for (int sl = t.s0 + wave; sl < t.s1; sl += kW) {
    double a0 = 0.0, a1 = 0.0, a2 = 0.0;
    if (!x_zero) row_spmv<double>(sl, lane, c, val, x, a0, a1, a2);
    const int row = sl * 64 + lane;
    if (row < c.n_nodes) {
      const double r0 = b[3 * row] - a0, r1 = b[3 * row + 1] - a1, r2 = b[3 * row + 2] - a2;
      r[3 * row] = r0;
      r[3 * row + 1] = r1;
      r[3 * row + 2] = r2;
      acc[0] += r0 * r0 + r1 * r1 + r2 * r2;
    }
  }
  double* outs[1] = {TP(c, slot)};
  reduce_store(acc, 1, outs, t.L);
}

// out[k * n_tissues + T] = ordered sum of slot k of tissue T, k < n_slots.
__global__ __launch_bounds__(kBlock) void tg_tissue_sums(TC c, int k0, int n_slots, double* out) {
  const int T = blockIdx.x;
  if (T >= c.tm.n_tissues) return;
  for (int k = k0; k < k0 + n_slots; ++k) {
    const double s = tissue_sum(c, k, T);
    if (threadIdx.x == 0) out[(size_t)k * c.tm.n_tissues + T] = s;
  }
}

// out = in * fac[T]
__global__ __launch_bounds__(kBlock) void tg_scale(TC c, const int32_t* gact, const double* in,
                                                   const double* fac, double* out) {
  TileCtx t;
  if (!tile_ctx(c, t) || !gact[t.T]) return;
  const double f = fac[t.T];
  const int end = min(t.s1 * 64, c.n_nodes);
  for (int i = t.s0 * 64 + threadIdx.x; i < end; i += kBlock)
#pragma unroll
    for (int d = 0; d < 3; ++d) out[3 * i + d] = in[3 * i + d] * f;
}

// z = M^-1 v
__global__ __launch_bounds__(kBlock) void tg_precond(TC c, const int32_t* gact,
                                                     const double* minv, const double* v,
                                                     double* z) {
  TileCtx t;
  if (!tile_ctx(c, t) || !gact[t.T]) return;
  const int end = min(t.s1 * 64, c.n_nodes);
  for (int i = t.s0 * 64 + threadIdx.x; i < end; i += kBlock) minv_apply(minv, v, z, i);
}

// w = J z
__global__ __launch_bounds__(kBlock) void tg_spmv(TC c, const int32_t* gact,
                                                  const double* __restrict__ val,
                                                  const double* __restrict__ z,
                                                  double* __restrict__ w) {
  TileCtx t;
  if (!tile_ctx(c, t) || !gact[t.T]) return;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int sl = t.s0 + wave; sl < t.s1; sl += kW) {
    double a0, a1, a2;
    row_spmv<double>(sl, lane, c, val, z, a0, a1, a2);
    const int row = sl * 64 + lane;
    if (row < c.n_nodes) {
      w[3 * row] = a0;
      w[3 * row + 1] = a1;
      w[3 * row + 2] = a2;
    }
  }
}

// partial (v_i, w) into slot i for i <= j
__global__ __launch_bounds__(kBlock) void tg_dots(TC c, const int32_t* gact, const double* V,
                                                  int64_t n3, int j, const double* w) {
  TileCtx t;
  if (!tile_ctx(c, t) || !gact[t.T]) return;
  const int end = min(t.s1 * 64, c.n_nodes);
  for (int i = 0; i <= j; ++i) {
    const double* v = V + (size_t)i * n3;
    double acc[1] = {0.0};
    for (int q = t.s0 * 64 + threadIdx.x; q < end; q += kBlock)
#pragma unroll
      for (int d = 0; d < 3; ++d) acc[0] += v[3 * q + d] * w[3 * q + d];
    double* outs[1] = {TP(c, i)};
    reduce_store(acc, 1, outs, t.L);
  }
}

// w -= sum_{i <= j} h[i, T] v_i; partial |w|^2 into slot norm_slot
__global__ __launch_bounds__(kBlock) void tg_orth(TC c, const int32_t* gact, const double* V,
                                                  int64_t n3, int j, const double* h, double* w,
                                                  int norm_slot) {
  TileCtx t;
  if (!tile_ctx(c, t) || !gact[t.T]) return;
  const int nt = c.tm.n_tissues;
  const int end = min(t.s1 * 64, c.n_nodes);
  double acc[1] = {0.0};
  for (int q = t.s0 * 64 + threadIdx.x; q < end; q += kBlock)
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      double a = w[3 * q + d];
      for (int i = 0; i <= j; ++i) a -= h[(size_t)i * nt + t.T] * V[(size_t)i * n3 + 3 * q + d];
      w[3 * q + d] = a;
      acc[0] += a * a;
    }
  double* outs[1] = {TP(c, norm_slot)};
  reduce_store(acc, 1, outs, t.L);
}

// u = sum_{i < k[T]} y[i, T] v_i, then x += M^-1 u (fused per node)
__global__ __launch_bounds__(kBlock) void tg_update(TC c, const int32_t* gupd, const double* V,
                                                    int64_t n3, const int32_t* kT,
                                                    const double* y, const double* minv,
                                                    double* x) {
  TileCtx t;
  if (!tile_ctx(c, t) || !gupd[t.T]) return;
  const int nt = c.tm.n_tissues;
  const int kk = kT[t.T];
  const int end = min(t.s1 * 64, c.n_nodes);
  for (int q = t.s0 * 64 + threadIdx.x; q < end; q += kBlock) {
    double u[3] = {0.0, 0.0, 0.0};
    for (int i = 0; i < kk; ++i) {
      const double yi = y[(size_t)i * nt + t.T];
#pragma unroll
      for (int d = 0; d < 3; ++d) u[d] += yi * V[(size_t)i * n3 + 3 * q + d];
    }
    const double* M = minv + 9 * (size_t)q;
#pragma unroll
    for (int d = 0; d < 3; ++d)
      x[3 * q + d] = x[3 * q + d] + (M[3 * d] * u[0] + M[3 * d + 1] * u[1] + M[3 * d + 2] * u[2]);
  }
}

}  // namespace

hipError_t tiled_gmres(const KrylovArgs& k, const double* b, double* x, double rtol, int maxit,
                       TiledResult* res) {
  Workspace* ws = k.ws;
  hipStream_t st = k.L.stream;
  const int G = k.tm.grid;
  const int nt = k.tm.n_tissues;
  const int m = kGmresRestart;
  const int64_t n3 = 3 * (int64_t)k.n_nodes;
  TC c = make_tc(k, rtol);
  c.part = ws->gpart;
  double* Vb = ws->gV;
  double *r = ws->r, *z = ws->s, *w = ws->v;
  hipError_t e;
  // host arrays uploaded asynchronously live until the final synchronisation
  std::vector<int32_t> act0(nt), gact(nt), kT(nt, 0), it(nt, 0), done(nt, 0), cyc(nt, 0);
  if ((e = hipMemcpy(act0.data(), ws->tactive, sizeof(int32_t) * nt, hipMemcpyDeviceToHost)) !=
      hipSuccess)
    return e;
  if ((e = hipMemsetAsync(x, 0, sizeof(double) * n3, st)) != hipSuccess) return e;
  std::vector<double> hs((size_t)(m + 2) * nt), hs2((size_t)(m + 2) * nt), bnorm(nt, 0.0);
  std::vector<double> fac(nt), y((size_t)m * nt);
  // per tissue: Hessenberg column-major (m+1) x m, rotations, rhs
  std::vector<double> H((size_t)nt * (m + 1) * m), cs((size_t)nt * m), sn((size_t)nt * m),
      g((size_t)nt * (m + 1));
  auto upload = [&](const void* src, void* dst, size_t bytes) {
    return hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st);
  };
  auto tissue_norms = [&](int slot, std::vector<double>& out) -> hipError_t {
    hipLaunchKernelGGL(tg_tissue_sums, dim3(nt), dim3(kBlock), 0, st, c, slot, 1, ws->gsum);
    hipError_t e2 = hipMemcpyAsync(out.data() + (size_t)slot * nt, ws->gsum + (size_t)slot * nt,
                                   sizeof(double) * nt, hipMemcpyDeviceToHost, st);
    if (e2 != hipSuccess) return e2;
    return hipStreamSynchronize(st);
  };
  for (int T = 0; T < nt; ++T) {
    gact[T] = act0[T];
    res->state[T] = act0[T] ? K_RUN : K_CONVERGED;
    res->iters[T] = 0;
  }
  for (int cycle = 0;; ++cycle) {
    // r = b - J x and |r| per tissue
    if ((e = upload(gact.data(), ws->gact, sizeof(int32_t) * nt)) != hipSuccess) return e;
    hipLaunchKernelGGL(tg_residual, dim3(G), dim3(kBlock), 0, st, c, ws->gact, k.val, x, b, r,
                       cycle == 0 ? 1 : 0, m + 1);
    if ((e = tissue_norms(m + 1, hs)) != hipSuccess) return e;
    bool any = false;
    for (int T = 0; T < nt; ++T) {
      if (!gact[T]) continue;
      const double beta = sqrt(hs[(size_t)(m + 1) * nt + T]);
      if (cycle == 0) bnorm[T] = beta;
      if (!std::isfinite(beta)) {
        res->state[T] = K_BREAKDOWN;
        gact[T] = 0;
      } else if (!(beta > rtol * bnorm[T])) {
        res->state[T] = K_CONVERGED;
        gact[T] = 0;
      } else if (it[T] >= maxit) {
        res->state[T] = K_RUN;
        gact[T] = 0;
      } else {
        fac[T] = 1.0 / beta;
        std::fill(g.begin() + (size_t)T * (m + 1), g.begin() + (size_t)(T + 1) * (m + 1), 0.0);
        g[(size_t)T * (m + 1)] = beta;
        kT[T] = 0;
        any = true;
      }
      done[T] = !gact[T];
    }
    if (!any) break;
    cyc = gact;  // tissues taking part in this cycle
    if ((e = upload(gact.data(), ws->gact, sizeof(int32_t) * nt)) != hipSuccess) return e;
    if ((e = upload(fac.data(), ws->gfac, sizeof(double) * nt)) != hipSuccess) return e;
    hipLaunchKernelGGL(tg_scale, dim3(G), dim3(kBlock), 0, st, c, ws->gact, r, ws->gfac, Vb);
    for (int j = 0; j < m; ++j) {
      double* vj = Vb + (size_t)j * n3;
      hipLaunchKernelGGL(tg_precond, dim3(G), dim3(kBlock), 0, st, c, ws->gact, k.minv, vj, z);
      hipLaunchKernelGGL(tg_spmv, dim3(G), dim3(kBlock), 0, st, c, ws->gact, k.val, z, w);
      // classical Gram-Schmidt, twice
      for (int pass = 0; pass < 2; ++pass) {
        hipLaunchKernelGGL(tg_dots, dim3(G), dim3(kBlock), 0, st, c, ws->gact, Vb, n3, j, w);
        double* hdev = ws->gsum + (size_t)pass * (m + 2) * nt;
        hipLaunchKernelGGL(tg_tissue_sums, dim3(nt), dim3(kBlock), 0, st, c, 0, j + 1, hdev);
        hipLaunchKernelGGL(tg_orth, dim3(G), dim3(kBlock), 0, st, c, ws->gact, Vb, n3, j, hdev,
                           w, m + 1);
      }
      hipLaunchKernelGGL(tg_tissue_sums, dim3(nt), dim3(kBlock), 0, st, c, m + 1, 1,
                         ws->gsum + (size_t)(m + 2) * nt);
      if ((e = hipGetLastError()) != hipSuccess) return e;
      if ((e = hipMemcpyAsync(hs.data(), ws->gsum, sizeof(double) * (m + 2) * nt,
                              hipMemcpyDeviceToHost, st)) != hipSuccess)
        return e;
      if ((e = hipMemcpyAsync(hs2.data(), ws->gsum + (size_t)(m + 2) * nt,
                              sizeof(double) * (m + 2) * nt, hipMemcpyDeviceToHost, st)) !=
          hipSuccess)
        return e;
      if ((e = hipStreamSynchronize(st)) != hipSuccess) return e;
      bool running = false;
      for (int T = 0; T < nt; ++T) {
        if (!gact[T]) continue;
        double* Hc = H.data() + ((size_t)T * m + j) * (m + 1);  // column j
        for (int i = 0; i <= j; ++i) Hc[i] = hs[(size_t)i * nt + T] + hs2[(size_t)i * nt + T];
        Hc[j + 1] = sqrt(hs2[(size_t)(m + 1) * nt + T]);
        const double hnext = Hc[j + 1];
        double* csT = cs.data() + (size_t)T * m;
        double* snT = sn.data() + (size_t)T * m;
        double* gT = g.data() + (size_t)T * (m + 1);
        for (int i = 0; i < j; ++i) {
          const double a = csT[i] * Hc[i] + snT[i] * Hc[i + 1];
          Hc[i + 1] = -snT[i] * Hc[i] + csT[i] * Hc[i + 1];
          Hc[i] = a;
        }
        const double den = std::hypot(Hc[j], Hc[j + 1]);
        csT[j] = den > 0 ? Hc[j] / den : 1.0;
        snT[j] = den > 0 ? Hc[j + 1] / den : 0.0;
        Hc[j] = den;
        Hc[j + 1] = 0.0;
        gT[j + 1] = -snT[j] * gT[j];
        gT[j] = csT[j] * gT[j];
        kT[T] = j + 1;
        it[T]++;
        const double est = std::fabs(gT[j + 1]);
        if (!std::isfinite(est) || !std::isfinite(den)) {
          res->state[T] = K_BREAKDOWN;
          gact[T] = 0;
          cyc[T] = 0;
          done[T] = 1;
        } else if (est <= rtol * bnorm[T] || !(hnext > 0.0) || it[T] >= maxit) {
          gact[T] = 0;  // ends this cycle; the true residual decides at the next
        } else {
          fac[T] = 1.0 / hnext;
          running = true;
        }
      }
      if (!running || j + 1 == m) break;
      if ((e = upload(gact.data(), ws->gact, sizeof(int32_t) * nt)) != hipSuccess) return e;
      if ((e = upload(fac.data(), ws->gfac, sizeof(double) * nt)) != hipSuccess) return e;
      hipLaunchKernelGGL(tg_scale, dim3(G), dim3(kBlock), 0, st, c, ws->gact, w, ws->gfac,
                         Vb + (size_t)(j + 1) * n3);
    }
    // y = H^-1 g per tissue, x += M^-1 V y
    for (int T = 0; T < nt; ++T) {
      if (!cyc[T]) continue;
      const int kk = kT[T];
      const double* HT = H.data() + (size_t)T * m * (m + 1);
      const double* gT = g.data() + (size_t)T * (m + 1);
      for (int i = kk - 1; i >= 0; --i) {
        double sacc = gT[i];
        for (int q = i + 1; q < kk; ++q) sacc -= HT[(size_t)q * (m + 1) + i] * y[(size_t)q * nt + T];
        y[(size_t)i * nt + T] = sacc / HT[(size_t)i * (m + 1) + i];
      }
    }
    if ((e = upload(cyc.data(), ws->gact, sizeof(int32_t) * nt)) != hipSuccess) return e;
    if ((e = upload(kT.data(), ws->gk, sizeof(int32_t) * nt)) != hipSuccess) return e;
    if ((e = upload(y.data(), ws->gy, sizeof(double) * m * nt)) != hipSuccess) return e;
    hipLaunchKernelGGL(tg_update, dim3(G), dim3(kBlock), 0, st, c, ws->gact, Vb, n3, ws->gk,
                       ws->gy, k.minv, x);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    for (int T = 0; T < nt; ++T) gact[T] = cyc[T];
  }
  for (int T = 0; T < nt; ++T) res->iters[T] = it[T];
  return hipStreamSynchronize(st);
}

// Per-tissue |g|_2 of the Newton residual into ws->host_gnorm (pinned),
// copied asynchronously: valid after the caller's next synchronisation.  A
// NaN or Inf here is reported as TF_ERR_NAN by the NR loop.
hipError_t tiled_residual_norms(const KrylovArgs& k, const double* g) {
  Workspace* ws = k.ws;
  hipStream_t st = k.L.stream;
  TC c = make_tc(k, 0.0);
  const int nt = k.tm.n_tissues;
  hipLaunchKernelGGL(tz_sqnorm<double>, dim3(k.tm.grid), dim3(kBlock), 0, st, c, g, 1);
  hipLaunchKernelGGL(tz_tissue_sums, dim3(nt), dim3(kBlock), 0, st, c, 1, ws->gnorm);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  return hipMemcpyAsync(ws->host_gnorm, ws->gnorm, sizeof(double) * nt, hipMemcpyDeviceToHost,
                        st);
}

// The linear solve of the NR step on a single-tissue handle: the replacement
// of the Pardiso call of NewtonRaphsonSolver::solveWithPardiso
// (NewtonRaphsonSolver.cpp:498-763).  A bound system arrives as its explicit
// operator N^T J N + I_slave, so every solve takes the same path.
hipError_t krylov_bicgstab(const KrylovArgs& k, const double* b, double* x, double rtol,
                           int maxit, int* iters, int* state, int sync_every) {
  static thread_local TiledResult res;
  hipError_t e = tiled_bicgstab(k, b, x, rtol, maxit, sync_every, &res);
  *iters = res.iters[0];
  *state = res.state[0];
  return e;
}

// Adds the durations of the profiled launches (HIP event pairs recorded by the
// runtime at each kernel's start and end) to the per-kind sums.
void prof_harvest(Prof* p) {
  if (!p || !p->on) return;
  for (int i = 0; i < p->used; ++i) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, p->ev[2 * i], p->ev[2 * i + 1]) == hipSuccess) {
      p->ms[p->kind[i]] += ms;
      p->count[p->kind[i]]++;
    }
  }
  p->used = 0;
}

void tiled_release(Workspace* ws) {
  GraphCache* gc = static_cast<GraphCache*>(ws->graphs);
  if (!gc) return;
  for (auto& g : gc->g)
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
  delete gc;
  ws->graphs = nullptr;
}

}  // namespace tf
