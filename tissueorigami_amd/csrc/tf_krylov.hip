// Krylov solver of the NR linear system J du = g: the replacement of the
// Pardiso call of NewtonRaphsonSolver::solveWithPardiso (NewtonRaphsonSolver.cpp:498-763).
//
// Right-preconditioned BiCGStab with 3x3 node-block Jacobi.  Everything stays
// on the device: the scalars (rho, alpha, omega) live in device memory and
// every kernel recomputes the dot products it needs from per-workgroup
// partial sums (summed in a fixed order, so results are run-to-run
// reproducible); the convergence test runs inside the first kernel of each
// iteration, which raises a device flag that makes the remaining launches of
// a batch return at once.  The host only looks at the flag every few
// iterations.
//
// SpMV: SELL-64 of 3x3 blocks, one wavefront per slice of 64 node rows, one
// lane per row; the 9 block values and the column index of a step are loaded
// with fully coalesced 64-lane loads from planar arrays.  Workgroups are
// remapped so that each XCD works on one contiguous eighth of the rows,
// keeping the x gathers of neighbouring rows in the same L2.
#include "tf_internal.h"

namespace tf {
namespace {

constexpr int kWaves = kBlock / 64;

__device__ inline int xcd_logical(int bid, int grid) {
  // workgroups are dispatched round-robin over the 8 XCDs; give XCD x the
  // contiguous logical range [x * grid/8, (x+1) * grid/8)
  const int per = grid >> 3;
  return (bid & 7) * per + (bid >> 3);
}

// Deterministic workgroup reduction of N values; thread 0 stores value k at
// out_k[blockIdx.x].
template <int N>
__device__ inline void block_reduce_store(double (&v)[N], double* o0, double* o1) {
  __shared__ double red[N][kWaves];
#pragma unroll
  for (int k = 0; k < N; ++k) {
    double x = v[k];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    v[k] = x;
  }
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  __syncthreads();  // a previous reduction of this kernel may still read red
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < N; ++k) red[k][wave] = v[k];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < N; ++k) {
      double s = 0.0;
#pragma unroll
      for (int w = 0; w < kWaves; ++w) s += red[k][w];
      if (k == 0) o0[blockIdx.x] = s;
      else o1[blockIdx.x] = s;
    }
  }
}

// Sum of the first n partials, identical in every thread of every workgroup.
__device__ inline double sum_partials(const double* part, int n) {
  __shared__ double red[kWaves];
  __shared__ double result;
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += kBlock) s += part[i];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  __syncthreads();
  if (lane == 0) red[wave] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) t += red[w];
    result = t;
  }
  __syncthreads();
  return result;
}

struct Ctl {
  int32_t* flag;     // [0] state, [1] iterations done
  double* scal;
  double* part;      // 5 * kParts
  int nparts;        // workgroups of the grid that produced the partials
  double rtol;
};

__device__ inline double* P(const Ctl& c, int k) { return c.part + (size_t)k * kParts; }

// ---------------------------------------------------------------- kernels

template <typename T>
__global__ __launch_bounds__(kBlock) void init_kernel(int n_nodes, int per, const double* b,
                                                      T* r, T* rh, T* p, T* v, T* x, Ctl c) {
  const int L = xcd_logical(blockIdx.x, gridDim.x);
  const int beg = L * per, end = min(n_nodes, beg + per);
  double acc[2] = {0.0, 0.0};
  for (int i = beg + threadIdx.x; i < end; i += kBlock) {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const T bi = (T)b[3 * i + d];
      r[3 * i + d] = bi; rh[3 * i + d] = bi;
      p[3 * i + d] = 0; v[3 * i + d] = 0; x[3 * i + d] = 0;
      acc[0] += (double)bi * (double)bi;
      acc[1] += (double)bi * (double)bi;
    }
  }
  block_reduce_store<2>(acc, P(c, 0), P(c, 1));
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    c.scal[S_RHO] = 1.0; c.scal[S_ALPHA] = 1.0; c.scal[S_OMEGA] = 1.0;
    c.flag[0] = K_RUN; c.flag[1] = 0;
  }
}

template <typename T>
__device__ inline void apply_minv(const T* m, const T* in, T* out, int i) {
  const T a = in[3 * i], b = in[3 * i + 1], d = in[3 * i + 2];
  const T* M = m + 9 * (size_t)i;
  out[3 * i] = M[0] * a + M[1] * b + M[2] * d;
  out[3 * i + 1] = M[3] * a + M[4] * b + M[5] * d;
  out[3 * i + 2] = M[6] * a + M[7] * b + M[8] * d;
}

// K1: convergence test, beta, p = r + beta (p - omega v), ph = M^-1 p.
template <typename T>
__global__ __launch_bounds__(kBlock) void bicg_p_kernel(int n_nodes, int per, const T* r, T* p,
                                                        const T* v, T* ph, const T* minv, Ctl c) {
  if (c.flag[0] != K_RUN) return;
  const double rho_new = sum_partials(P(c, 0), c.nparts);
  const double rr = sum_partials(P(c, 1), c.nparts);
  const double rnorm = sqrt(rr);
  const bool first = (c.flag[1] == 0);
  const double bnorm = first ? rnorm : c.scal[S_BNORM];
  int state = K_RUN;
  if (!(rnorm > c.rtol * bnorm) || rnorm == 0.0) state = K_CONVERGED;
  else if (rho_new == 0.0 || !isfinite(rho_new) || !isfinite(rnorm) || c.scal[S_OMEGA] == 0.0)
    state = K_BREAKDOWN;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    if (first) c.scal[S_BNORM] = bnorm;
    c.scal[S_RNORM] = rnorm;
    c.scal[S_RHO_NEW] = rho_new;
    if (state != K_RUN) c.flag[0] = state;
  }
  if (state != K_RUN) return;
  const double beta = (rho_new / c.scal[S_RHO]) * (c.scal[S_ALPHA] / c.scal[S_OMEGA]);
  const T tb = (T)beta, tw = (T)c.scal[S_OMEGA];
  const int L = xcd_logical(blockIdx.x, gridDim.x);
  const int beg = L * per, end = min(n_nodes, beg + per);
  for (int i = beg + threadIdx.x; i < end; i += kBlock) {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const int k = 3 * i + d;
      p[k] = r[k] + tb * (p[k] - tw * v[k]);
    }
    apply_minv(minv, p, ph, i);
  }
}

// K3: alpha = rho_new / (rh, v); s = r - alpha v; sh = M^-1 s.
template <typename T>
__global__ __launch_bounds__(kBlock) void bicg_s_kernel(int n_nodes, int per, const T* r,
                                                        const T* v, T* s, T* sh, const T* minv,
                                                        Ctl c) {
  if (c.flag[0] != K_RUN) return;
  const double rv = sum_partials(P(c, 2), c.nparts);
  const double alpha = c.scal[S_RHO_NEW] / rv;
  if (rv == 0.0 || !isfinite(alpha)) {
    if (blockIdx.x == 0 && threadIdx.x == 0) c.flag[0] = K_BREAKDOWN;
    return;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) c.scal[S_ALPHA] = alpha;
  const T ta = (T)alpha;
  const int L = xcd_logical(blockIdx.x, gridDim.x);
  const int beg = L * per, end = min(n_nodes, beg + per);
  for (int i = beg + threadIdx.x; i < end; i += kBlock) {
#pragma unroll
    for (int d = 0; d < 3; ++d) s[3 * i + d] = r[3 * i + d] - ta * v[3 * i + d];
    apply_minv(minv, s, sh, i);
  }
}

// K5: omega = (t, s) / (t, t); x += alpha ph + omega sh; r = s - omega t;
// partials of (rh, r) and (r, r) for the next iteration.
template <typename T>
__global__ __launch_bounds__(kBlock) void bicg_x_kernel(int n_nodes, int per, T* x, T* r,
                                                        const T* rh, const T* ph, const T* s,
                                                        const T* sh, const T* t, Ctl c) {
  if (c.flag[0] != K_RUN) return;
  const double ts = sum_partials(P(c, 3), c.nparts);
  const double tt = sum_partials(P(c, 4), c.nparts);
  const double omega = (tt > 0.0) ? ts / tt : 0.0;
  const double alpha = c.scal[S_ALPHA];
  const T ta = (T)alpha, tw = (T)omega;
  const int L = xcd_logical(blockIdx.x, gridDim.x);
  const int beg = L * per, end = min(n_nodes, beg + per);
  double acc[2] = {0.0, 0.0};
  for (int i = beg + threadIdx.x; i < end; i += kBlock) {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const int k = 3 * i + d;
      x[k] = x[k] + ta * ph[k] + tw * sh[k];
      const T rk = s[k] - tw * t[k];
      r[k] = rk;
      acc[0] += (double)rh[k] * (double)rk;
      acc[1] += (double)rk * (double)rk;
    }
  }
  block_reduce_store<2>(acc, P(c, 0), P(c, 1));
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    c.scal[S_OMEGA] = omega;
    c.scal[S_RHO] = c.scal[S_RHO_NEW];
    c.flag[1] = c.flag[1] + 1;
  }
}

// SpMV y = J x on the SELL-64 pattern; optional dots (d0, y) -> part k0 and
// (y, y) or (d1, y) -> part k1.  ND: number of dots (0, 1 or 2); SELF1: second
// dot is (y, y).
template <typename T, int ND, bool SELF1>
__global__ __launch_bounds__(kBlock) void spmv_kernel(int n_nodes, int n_slices, int per_slices,
                                                      const int32_t* __restrict__ slice_base,
                                                      const int32_t* __restrict__ col,
                                                      const T* __restrict__ val, int64_t n_slots,
                                                      const T* __restrict__ x, T* __restrict__ y,
                                                      const T* __restrict__ d0, Ctl c, int k0,
                                                      int k1, int check_flag) {
  if (check_flag && c.flag[0] != K_RUN) return;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int L = xcd_logical(blockIdx.x, gridDim.x);
  const int sbeg = L * per_slices, send = min(n_slices, sbeg + per_slices);
  double acc[ND > 0 ? ND : 1];
#pragma unroll
  for (int k = 0; k < (ND > 0 ? ND : 1); ++k) acc[k] = 0.0;
  for (int sl = sbeg + wave; sl < send; sl += kWaves) {
    const int64_t base = slice_base[sl];
    const int width = (int)((slice_base[sl + 1] - base) >> 6);
    T a0 = 0, a1 = 0, a2 = 0;
    const T* v0 = val + base + lane;
    const int32_t* cp = col + base + lane;
    for (int k = 0; k < width; ++k) {
      const int64_t o = (int64_t)k * 64;
      const int cc = cp[o];
      const T x0 = x[3 * cc], x1 = x[3 * cc + 1], x2 = x[3 * cc + 2];
      const T* vv = v0 + o;
      a0 += vv[0] * x0 + vv[n_slots] * x1 + vv[2 * n_slots] * x2;
      a1 += vv[3 * n_slots] * x0 + vv[4 * n_slots] * x1 + vv[5 * n_slots] * x2;
      a2 += vv[6 * n_slots] * x0 + vv[7 * n_slots] * x1 + vv[8 * n_slots] * x2;
    }
    const int row = sl * 64 + lane;
    if (row < n_nodes) {
      y[3 * row] = a0; y[3 * row + 1] = a1; y[3 * row + 2] = a2;
      if (ND >= 1) {
        acc[0] += (double)d0[3 * row] * (double)a0 + (double)d0[3 * row + 1] * (double)a1 +
                  (double)d0[3 * row + 2] * (double)a2;
      }
      if (ND >= 2) {
        acc[ND - 1] += (double)a0 * (double)a0 + (double)a1 * (double)a1 + (double)a2 * (double)a2;
      }
    }
  }
  if (ND > 0) block_reduce_store<(ND > 0 ? ND : 1)>(acc, P(c, k0), P(c, k1));
}

// Generic dots: (a0, b0) -> part k0 and optionally (a1, b1) -> part k1.
template <typename T>
__global__ __launch_bounds__(kBlock) void dot2_kernel(int n_nodes, int per, const T* a0,
                                                      const T* b0, const T* a1, const T* b1,
                                                      Ctl c, int k0, int k1, int check_flag) {
  if (check_flag && c.flag[0] != K_RUN) return;
  const int L = xcd_logical(blockIdx.x, gridDim.x);
  const int beg = L * per, end = min(n_nodes, beg + per);
  double acc[2] = {0.0, 0.0};
  for (int i = beg + threadIdx.x; i < end; i += kBlock)
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      acc[0] += (double)a0[3 * i + d] * (double)b0[3 * i + d];
      if (a1) acc[1] += (double)a1[3 * i + d] * (double)b1[3 * i + d];
    }
  if (a1) {
    block_reduce_store<2>(acc, P(c, k0), P(c, k1));
  } else {
    double one[1] = {acc[0]};
    block_reduce_store<1>(one, P(c, k0), nullptr);
  }
}

// Binding operator pieces (N^T J N + I_slave, NewtonRaphsonSolver.cpp:128-183).
template <typename T>
__global__ void copy_kernel(int n, const T* a, T* b, const int32_t* flag, int check_flag) {
  if (check_flag && flag[0] != K_RUN) return;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) b[i] = a[i];
}

template <typename T>
__global__ void bind_in_kernel(const int32_t* pairs, int np, const T* x, T* tmp,
                               const int32_t* flag, int check_flag) {
  if (check_flag && flag[0] != K_RUN) return;
  for (int p = blockIdx.x * kBlock + threadIdx.x; p < np; p += gridDim.x * kBlock)
    tmp[pairs[2 * p]] = x[pairs[2 * p + 1]];
}

template <typename T>
__global__ void bind_out_kernel(const int32_t* pairs, int np, const T* x, T* y,
                                const int32_t* flag, int check_flag) {
  if (check_flag && flag[0] != K_RUN) return;
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  for (int p = 0; p < np; ++p) y[pairs[2 * p + 1]] += y[pairs[2 * p]];
  for (int p = 0; p < np; ++p) y[pairs[2 * p]] = x[pairs[2 * p]];
}

// |dx|^2 partials and x += dx.
__global__ __launch_bounds__(kBlock) void norm_update_kernel(int n_nodes, int per,
                                                             const double* dx, double* xk, Ctl c) {
  const int L = xcd_logical(blockIdx.x, gridDim.x);
  const int beg = L * per, end = min(n_nodes, beg + per);
  double acc[1] = {0.0};
  for (int i = beg + threadIdx.x; i < end; i += kBlock)
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const double v = dx[3 * i + d];
      acc[0] += v * v;
      xk[3 * i + d] = xk[3 * i + d] + v;
    }
  block_reduce_store<1>(acc, P(c, 0), nullptr);
}

__global__ void finish_sum_kernel(const double* part, int n, double* out) {
  const double s = sum_partials(part, n);
  if (threadIdx.x == 0) *out = s;
}

// Mixed precision helpers.
__global__ void to_f32_kernel(int n, const double* a, float* b, double scale) {
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock)
    b[i] = (float)(a[i] * scale);
}

__global__ void scale_kernel(int n, const double* a, double* b, double scale) {
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock)
    b[i] = a[i] * scale;
}

__global__ void axpy_f32_kernel(int n, const float* z, double scale, double* x) {
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock)
    x[i] = x[i] + (double)z[i] * scale;
}

// r = b - J x with the FP32 block values promoted and FP64 accumulation;
// partial of (r, r) into part 1.
template <typename TV>
__global__ __launch_bounds__(kBlock) void residual_kernel(int n_nodes, int n_slices,
                                                          int per_slices,
                                                          const int32_t* __restrict__ slice_base,
                                                          const int32_t* __restrict__ col,
                                                          const TV* __restrict__ val,
                                                          int64_t n_slots,
                                                          const double* __restrict__ x,
                                                          const double* __restrict__ b,
                                                          double* __restrict__ r, Ctl c) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int L = xcd_logical(blockIdx.x, gridDim.x);
  const int sbeg = L * per_slices, send = min(n_slices, sbeg + per_slices);
  double acc[1] = {0.0};
  for (int sl = sbeg + wave; sl < send; sl += kWaves) {
    const int64_t base = slice_base[sl];
    const int width = (int)((slice_base[sl + 1] - base) >> 6);
    double a0 = 0, a1 = 0, a2 = 0;
    for (int k = 0; k < width; ++k) {
      const int64_t o = base + (int64_t)k * 64 + lane;
      const int cc = col[o];
      const double x0 = x[3 * cc], x1 = x[3 * cc + 1], x2 = x[3 * cc + 2];
      a0 += (double)val[o] * x0 + (double)val[n_slots + o] * x1 + (double)val[2 * n_slots + o] * x2;
      a1 += (double)val[3 * n_slots + o] * x0 + (double)val[4 * n_slots + o] * x1 +
            (double)val[5 * n_slots + o] * x2;
      a2 += (double)val[6 * n_slots + o] * x0 + (double)val[7 * n_slots + o] * x1 +
            (double)val[8 * n_slots + o] * x2;
    }
    const int row = sl * 64 + lane;
    if (row < n_nodes) {
      const double r0 = b[3 * row] - a0, r1 = b[3 * row + 1] - a1, r2 = b[3 * row + 2] - a2;
      r[3 * row] = r0; r[3 * row + 1] = r1; r[3 * row + 2] = r2;
      acc[0] += r0 * r0 + r1 * r1 + r2 * r2;
    }
  }
  block_reduce_store<1>(acc, P(c, 1), nullptr);
}


// ---------------------------------------------------------------------------
// Fused BiCGStab: four kernels per iteration with every vector update fused
// into an SpMV or into one of two streaming passes.
// With w = J^T rh (one transposed SpMV per solve), (rh, v) = (w, ph) is
// formed as a partial sum in the pass that produces ph, so alpha is known
// when the SpMV that computes v starts and s = r - alpha v, sh = M^-1 s fuse
// into it.  rho = (rh, r) is always a direct dot product: the recursive
// identity rho' = rho - alpha (rh, v) - omega (rh, t) loses the tight
// tolerances the NR step asks for.
//   fz_spmv_v:  convergence test, alpha = rho / (w, ph); v = J ph;
//               s = r - alpha v; sh = M^-1 s
//   fz_spmv_t:  t = J sh; partials (t, s), (t, t)
//   fz_update_x: omega; x += alpha ph + omega sh; r = s - omega t;
//               partials (rh, r), (r, r)
//   fz_update_p: beta = (rho'/rho)(alpha/omega); p = r + beta (p - omega v);
//               ph = M^-1 p; partial (w, ph)
// Partial slots: 0 (rh, r), 1 (r, r), 2 (t, s), 3 (t, t), 4 (w, ph).
// Scalars: fz_spmv_v writes RHO, ALPHA, BNORM; fz_update_x writes OMEGA;
// fz_update_p only reads.  No launch reads a slot its own workgroups write.

template <typename T>
__device__ inline void spmv_row(int sl, int lane, const int32_t* __restrict__ slice_base,
                                const int32_t* __restrict__ col, const T* __restrict__ val,
                                int64_t n_slots, const T* __restrict__ x, T& a0, T& a1, T& a2) {
  const int64_t base = slice_base[sl];
  const int width = (int)((slice_base[sl + 1] - base) >> 6);
  a0 = 0; a1 = 0; a2 = 0;
  const T* v0 = val + base + lane;
  const int32_t* cp = col + base + lane;
  for (int k = 0; k < width; ++k) {
    const int64_t o = (int64_t)k * 64;
    const int cc = __builtin_nontemporal_load(cp + o);
    const T x0 = x[3 * cc], x1 = x[3 * cc + 1], x2 = x[3 * cc + 2];
    const T* vv = v0 + o;
    a0 += __builtin_nontemporal_load(vv) * x0 + __builtin_nontemporal_load(vv + n_slots) * x1 +
          __builtin_nontemporal_load(vv + 2 * n_slots) * x2;
    a1 += __builtin_nontemporal_load(vv + 3 * n_slots) * x0 +
          __builtin_nontemporal_load(vv + 4 * n_slots) * x1 +
          __builtin_nontemporal_load(vv + 5 * n_slots) * x2;
    a2 += __builtin_nontemporal_load(vv + 6 * n_slots) * x0 +
          __builtin_nontemporal_load(vv + 7 * n_slots) * x1 +
          __builtin_nontemporal_load(vv + 8 * n_slots) * x2;
  }
}

struct SellView {
  int n_nodes, n_slices, per_slices;
  const int32_t* slice_base;
  const int32_t* col;
  const int32_t* tslot;
  int64_t n_slots;
};

// y = J^T x through the transpose slot map (block (R, C) of J^T is block
// (C, R) of J, transposed).  Used once per solve for w = J^T rh.
template <typename T>
__global__ __launch_bounds__(kBlock) void spmv_transposed_kernel(SellView m, const T* val,
                                                                 const T* x, T* y) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int L = xcd_logical(blockIdx.x, gridDim.x);
  const int sbeg = L * m.per_slices, send = min(m.n_slices, sbeg + m.per_slices);
  for (int sl = sbeg + wave; sl < send; sl += kWaves) {
    const int64_t base = m.slice_base[sl];
    const int width = (int)((m.slice_base[sl + 1] - base) >> 6);
    T a0 = 0, a1 = 0, a2 = 0;
    for (int k = 0; k < width; ++k) {
      const int64_t o = base + (int64_t)k * 64 + lane;
      const int cc = m.col[o];
      const int64_t t = m.tslot[o];
      const T x0 = x[3 * cc], x1 = x[3 * cc + 1], x2 = x[3 * cc + 2];
      // (B^T)[p][q] = B[q][p]
      a0 += val[t] * x0 + val[3 * m.n_slots + t] * x1 + val[6 * m.n_slots + t] * x2;
      a1 += val[m.n_slots + t] * x0 + val[4 * m.n_slots + t] * x1 + val[7 * m.n_slots + t] * x2;
      a2 += val[2 * m.n_slots + t] * x0 + val[5 * m.n_slots + t] * x1 + val[8 * m.n_slots + t] * x2;
    }
    const int row = sl * 64 + lane;
    if (row < m.n_nodes) { y[3 * row] = a0; y[3 * row + 1] = a1; y[3 * row + 2] = a2; }
  }
}

// r = rh = p = b, x = 0.
template <typename T>
__global__ __launch_bounds__(kBlock) void fz_init_a_kernel(int n, const double* b, T* r, T* rh,
                                                           T* p, T* x) {
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    const T bi = (T)b[i];
    r[i] = bi; rh[i] = bi; p[i] = bi; x[i] = 0;
  }
}

// ph = M^-1 p; partials (rh, r) = (r, r) into slots 0 and 1, (w, ph) into 4.
template <typename T>
__global__ __launch_bounds__(kBlock) void fz_init_b_kernel(int n_nodes, int per, const T* r,
                                                           const T* w, const T* p, T* ph,
                                                           const T* minv, Ctl c) {
  const int L = xcd_logical(blockIdx.x, gridDim.x);
  const int beg = L * per, end = min(n_nodes, beg + per);
  double acc[2] = {0.0, 0.0};
  for (int i = beg + threadIdx.x; i < end; i += kBlock) {
    apply_minv(minv, p, ph, i);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      acc[0] += (double)r[3 * i + d] * (double)r[3 * i + d];
      acc[1] += (double)w[3 * i + d] * (double)ph[3 * i + d];
    }
  }
  double rr[2] = {acc[0], acc[0]};
  block_reduce_store<2>(rr, P(c, 0), P(c, 1));
  double wp[1] = {acc[1]};
  block_reduce_store<1>(wp, P(c, 4), nullptr);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    c.scal[S_OMEGA] = 1.0;
    c.flag[0] = K_RUN; c.flag[1] = 0;
  }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void fz_spmv_v_kernel(SellView m, const T* __restrict__ val,
                                                           const T* __restrict__ ph,
                                                           const T* __restrict__ r,
                                                           const T* __restrict__ minv,
                                                           T* __restrict__ v, T* __restrict__ s,
                                                           T* __restrict__ sh, Ctl c) {
  if (c.flag[0] != K_RUN) return;
  const double rho = sum_partials(P(c, 0), c.nparts);
  const double rr = sum_partials(P(c, 1), c.nparts);
  const double wp = sum_partials(P(c, 4), c.nparts);
  const double rnorm = sqrt(rr);
  const bool first = (c.flag[1] == 0);
  const double bnorm = first ? rnorm : c.scal[S_BNORM];
  const double alpha = rho / wp;
  int state = K_RUN;
  if (!(rnorm > c.rtol * bnorm) || rnorm == 0.0) state = K_CONVERGED;
  else if (!first && c.scal[S_OMEGA] == 0.0) state = K_BREAKDOWN;
  else if (rho == 0.0 || wp == 0.0 || !isfinite(alpha) || !isfinite(rnorm)) state = K_BREAKDOWN;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    if (first) c.scal[S_BNORM] = bnorm;
    c.scal[S_RHO] = rho;
    c.scal[S_ALPHA] = alpha;
    c.scal[S_RNORM] = rnorm;
    if (state != K_RUN) c.flag[0] = state;
  }
  if (state != K_RUN) return;
  const T ta = (T)alpha;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int L = xcd_logical(blockIdx.x, gridDim.x);
  const int sbeg = L * m.per_slices, send = min(m.n_slices, sbeg + m.per_slices);
  for (int sl = sbeg + wave; sl < send; sl += kWaves) {
    T a0, a1, a2;
    spmv_row(sl, lane, m.slice_base, m.col, val, m.n_slots, ph, a0, a1, a2);
    const int row = sl * 64 + lane;
    if (row < m.n_nodes) {
      v[3 * row] = a0; v[3 * row + 1] = a1; v[3 * row + 2] = a2;
      s[3 * row] = r[3 * row] - ta * a0;
      s[3 * row + 1] = r[3 * row + 1] - ta * a1;
      s[3 * row + 2] = r[3 * row + 2] - ta * a2;
      apply_minv(minv, s, sh, row);
    }
  }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void fz_spmv_t_kernel(SellView m, const T* __restrict__ val,
                                                           const T* __restrict__ sh,
                                                           const T* __restrict__ s,
                                                           T* __restrict__ t, Ctl c) {
  if (c.flag[0] != K_RUN) return;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int L = xcd_logical(blockIdx.x, gridDim.x);
  const int sbeg = L * m.per_slices, send = min(m.n_slices, sbeg + m.per_slices);
  double acc[2] = {0.0, 0.0};
  for (int sl = sbeg + wave; sl < send; sl += kWaves) {
    T a0, a1, a2;
    spmv_row(sl, lane, m.slice_base, m.col, val, m.n_slots, sh, a0, a1, a2);
    const int row = sl * 64 + lane;
    if (row < m.n_nodes) {
      t[3 * row] = a0; t[3 * row + 1] = a1; t[3 * row + 2] = a2;
      acc[0] += (double)a0 * (double)s[3 * row] + (double)a1 * (double)s[3 * row + 1] +
                (double)a2 * (double)s[3 * row + 2];
      acc[1] += (double)a0 * (double)a0 + (double)a1 * (double)a1 + (double)a2 * (double)a2;
    }
  }
  block_reduce_store<2>(acc, P(c, 2), P(c, 3));
}

template <typename T>
__global__ __launch_bounds__(kBlock) void fz_update_x_kernel(int n_nodes, int per, T* x, T* r,
                                                             const T* rh, const T* ph,
                                                             const T* sh, const T* s, const T* t,
                                                             Ctl c) {
  if (c.flag[0] != K_RUN) return;
  const double ts = sum_partials(P(c, 2), c.nparts);
  const double tt = sum_partials(P(c, 3), c.nparts);
  const double omega = (tt > 0.0) ? ts / tt : 0.0;
  const T ta = (T)c.scal[S_ALPHA], tw = (T)omega;
  const int L = xcd_logical(blockIdx.x, gridDim.x);
  const int beg = L * per, end = min(n_nodes, beg + per);
  double acc[2] = {0.0, 0.0};
  for (int i = beg + threadIdx.x; i < end; i += kBlock) {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const int k = 3 * i + d;
      x[k] = x[k] + ta * ph[k] + tw * sh[k];
      const T rk = s[k] - tw * t[k];
      r[k] = rk;
      acc[0] += (double)rh[k] * (double)rk;
      acc[1] += (double)rk * (double)rk;
    }
  }
  block_reduce_store<2>(acc, P(c, 0), P(c, 1));
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    c.scal[S_OMEGA] = omega;
    c.flag[1] = c.flag[1] + 1;
  }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void fz_update_p_kernel(int n_nodes, int per, const T* r,
                                                             T* p, T* ph, const T* v, const T* w,
                                                             const T* minv, Ctl c) {
  if (c.flag[0] != K_RUN) return;
  const double rho_new = sum_partials(P(c, 0), c.nparts);
  const double rho = c.scal[S_RHO], alpha = c.scal[S_ALPHA], omega = c.scal[S_OMEGA];
  const double beta = (omega != 0.0 && rho != 0.0) ? (rho_new / rho) * (alpha / omega) : 0.0;
  const T tb = (T)beta, tw = (T)omega;
  const int L = xcd_logical(blockIdx.x, gridDim.x);
  const int beg = L * per, end = min(n_nodes, beg + per);
  double acc[1] = {0.0};
  for (int i = beg + threadIdx.x; i < end; i += kBlock) {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const int k = 3 * i + d;
      p[k] = r[k] + tb * (p[k] - tw * v[k]);
    }
    apply_minv(minv, p, ph, i);
#pragma unroll
    for (int d = 0; d < 3; ++d) acc[0] += (double)w[3 * i + d] * (double)ph[3 * i + d];
  }
  block_reduce_store<1>(acc, P(c, 4), nullptr);
}

int per_item(int64_t total, int grid) { return (int)((total + grid - 1) / grid); }

template <typename T>
struct Vecs {
  T *r, *rh, *p, *v, *ph, *s, *sh, *t, *x;
  const T* val;
  const T* minv;
};

template <typename T>
hipError_t spmv_op(const KrylovArgs& k, const Ctl& c, const T* val, const T* x, T* y, T* tmp,
                   const T* d0, int nd, bool self1, int k0, int k1) {
  const int G = k.L.grid;
  const int n = 3 * k.n_nodes;
  const int per_sl = per_item(k.pat.n_slices, G);
  const int per_n = per_item(k.n_nodes, G);
  hipStream_t st = k.L.stream;
  const T* xin = x;
  if (k.n_pairs > 0) {
    hipLaunchKernelGGL(copy_kernel<T>, dim3(G), dim3(kBlock), 0, st, n, x, tmp, c.flag, 1);
    hipLaunchKernelGGL(bind_in_kernel<T>, dim3((k.n_pairs + kBlock - 1) / kBlock), dim3(kBlock),
                       0, st, k.pairs, k.n_pairs, x, tmp, c.flag, 1);
    xin = tmp;
    hipLaunchKernelGGL((spmv_kernel<T, 0, false>), dim3(G), dim3(kBlock), 0, st, k.n_nodes,
                       k.pat.n_slices, per_sl, k.pat.slice_base, k.pat.col, val, k.pat.n_slots,
                       xin, y, (const T*)nullptr, c, 0, 0, 1);
    hipLaunchKernelGGL(bind_out_kernel<T>, dim3(1), dim3(64), 0, st, k.pairs, k.n_pairs, x, y,
                       c.flag, 1);
    if (nd > 0) {
      hipLaunchKernelGGL(dot2_kernel<T>, dim3(G), dim3(kBlock), 0, st, k.n_nodes, per_n, d0,
                         (const T*)y, nd > 1 ? (const T*)y : (const T*)nullptr,
                         nd > 1 ? (const T*)y : (const T*)nullptr, c, k0, k1, 1);
    }
    return hipGetLastError();
  }
  Prof* pf = k.prof;
  const bool timed = pf && pf->on && pf->used < kProfPairs;
  if (timed) {
    pf->kind[pf->used] = sizeof(T) == 4 ? 1 : 0;
    (void)hipEventRecord(pf->ev[2 * pf->used], st);
  }
  if (nd == 0)
    hipLaunchKernelGGL((spmv_kernel<T, 0, false>), dim3(G), dim3(kBlock), 0, st, k.n_nodes,
                       k.pat.n_slices, per_sl, k.pat.slice_base, k.pat.col, val, k.pat.n_slots,
                       xin, y, d0, c, k0, k1, 1);
  else if (nd == 1)
    hipLaunchKernelGGL((spmv_kernel<T, 1, false>), dim3(G), dim3(kBlock), 0, st, k.n_nodes,
                       k.pat.n_slices, per_sl, k.pat.slice_base, k.pat.col, val, k.pat.n_slots,
                       xin, y, d0, c, k0, k1, 1);
  else
    hipLaunchKernelGGL((spmv_kernel<T, 2, true>), dim3(G), dim3(kBlock), 0, st, k.n_nodes,
                       k.pat.n_slices, per_sl, k.pat.slice_base, k.pat.col, val, k.pat.n_slots,
                       xin, y, d0, c, k0, k1, 1);
  (void)self1;
  if (timed) {
    (void)hipEventRecord(pf->ev[2 * pf->used + 1], st);
    pf->used++;
  }
  return hipGetLastError();
}

template <typename T>
hipError_t bicgstab_impl(const KrylovArgs& k, const Vecs<T>& V, T* tmp, const double* b,
                         double rtol, int maxit, int* iters, int* state, int sync_every) {
  Workspace* ws = k.ws;
  const int G = k.L.grid;
  const int per_n = per_item(k.n_nodes, G);
  hipStream_t st = k.L.stream;
  Ctl c{ws->flag, ws->scal, ws->partials, G, rtol};
  hipLaunchKernelGGL(init_kernel<T>, dim3(G), dim3(kBlock), 0, st, k.n_nodes, per_n, b, V.r,
                     V.rh, V.p, V.v, V.x, c);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  int done = 0;
  *state = K_RUN;
  *iters = 0;
  while (done < maxit + 1) {
    const int chunk = std::min(sync_every, maxit + 1 - done);
    for (int it = 0; it < chunk; ++it) {
      if (k.prof && k.prof->on && k.prof->used + 2 > kProfPairs) {
        if ((e = hipStreamSynchronize(st)) != hipSuccess) return e;
        prof_harvest(k.prof);
      }
      hipLaunchKernelGGL(bicg_p_kernel<T>, dim3(G), dim3(kBlock), 0, st, k.n_nodes, per_n, V.r,
                         V.p, V.v, V.ph, V.minv, c);
      if ((e = spmv_op<T>(k, c, V.val, V.ph, V.v, tmp, V.rh, 1, false, 2, 0)) != hipSuccess)
        return e;
      hipLaunchKernelGGL(bicg_s_kernel<T>, dim3(G), dim3(kBlock), 0, st, k.n_nodes, per_n, V.r,
                         V.v, V.s, V.sh, V.minv, c);
      if ((e = spmv_op<T>(k, c, V.val, V.sh, V.t, tmp, V.s, 2, true, 3, 4)) != hipSuccess)
        return e;
      hipLaunchKernelGGL(bicg_x_kernel<T>, dim3(G), dim3(kBlock), 0, st, k.n_nodes, per_n, V.x,
                         V.r, V.rh, V.ph, V.s, V.sh, V.t, c);
    }
    done += chunk;
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if ((e = hipMemcpyAsync(ws->host_flag, ws->flag, 2 * sizeof(int32_t), hipMemcpyDeviceToHost,
                            st)) != hipSuccess)
      return e;
    if ((e = hipStreamSynchronize(st)) != hipSuccess) return e;
    prof_harvest(k.prof);
    *iters = ws->host_flag[1];
    if (ws->host_flag[0] != K_RUN) {
      *state = ws->host_flag[0];
      return hipSuccess;
    }
  }
  // iteration cap: run one more convergence test (K1 of the next iteration)
  hipLaunchKernelGGL(bicg_p_kernel<T>, dim3(G), dim3(kBlock), 0, st, k.n_nodes, per_n, V.r, V.p,
                     V.v, V.ph, V.minv, c);
  if ((e = hipMemcpyAsync(ws->host_flag, ws->flag, 2 * sizeof(int32_t), hipMemcpyDeviceToHost,
                          st)) != hipSuccess)
    return e;
  if ((e = hipStreamSynchronize(st)) != hipSuccess) return e;
  *iters = ws->host_flag[1];
  *state = ws->host_flag[0];
  return hipSuccess;
}


template <typename T>
hipError_t bicgstab_fused_impl(const KrylovArgs& k, const Vecs<T>& V, T* w, const double* b,
                               double rtol, int maxit, int* iters, int* state, int sync_every) {
  Workspace* ws = k.ws;
  const int G = k.L.grid;
  const int n = 3 * k.n_nodes;
  const int per_n = per_item(k.n_nodes, G);
  hipStream_t st = k.L.stream;
  Ctl c{ws->flag, ws->scal, ws->partials, G, rtol};
  SellView m{k.n_nodes, k.pat.n_slices, per_item(k.pat.n_slices, G), k.pat.slice_base,
             k.pat.col, k.pat.tslot, k.pat.n_slots};
  hipLaunchKernelGGL(fz_init_a_kernel<T>, dim3(G), dim3(kBlock), 0, st, n, b, V.r, V.rh, V.p, V.x);
  hipLaunchKernelGGL(spmv_transposed_kernel<T>, dim3(G), dim3(kBlock), 0, st, m, V.val,
                     (const T*)V.rh, w);
  hipLaunchKernelGGL(fz_init_b_kernel<T>, dim3(G), dim3(kBlock), 0, st, k.n_nodes, per_n,
                     (const T*)V.r, (const T*)w, (const T*)V.p, V.ph, V.minv, c);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  Prof* pf = k.prof;
  auto timed_spmv = [&](auto launch) {
    const bool timed = pf && pf->on && pf->used < kProfPairs;
    if (timed) {
      pf->kind[pf->used] = sizeof(T) == 4 ? 1 : 0;
      (void)hipEventRecord(pf->ev[2 * pf->used], st);
    }
    launch();
    if (timed) {
      (void)hipEventRecord(pf->ev[2 * pf->used + 1], st);
      pf->used++;
    }
  };
  int done = 0;
  *state = K_RUN;
  *iters = 0;
  while (true) {
    const int chunk = std::min(sync_every, maxit - done);
    for (int it = 0; it < chunk; ++it) {
      if (pf && pf->on && pf->used + 2 > kProfPairs) {
        if ((e = hipStreamSynchronize(st)) != hipSuccess) return e;
        prof_harvest(pf);
      }
      timed_spmv([&] {
        hipLaunchKernelGGL(fz_spmv_v_kernel<T>, dim3(G), dim3(kBlock), 0, st, m, V.val,
                           (const T*)V.ph, (const T*)V.r, V.minv, V.v, V.s, V.sh, c);
      });
      timed_spmv([&] {
        hipLaunchKernelGGL(fz_spmv_t_kernel<T>, dim3(G), dim3(kBlock), 0, st, m, V.val,
                           (const T*)V.sh, (const T*)V.s, V.t, c);
      });
      hipLaunchKernelGGL(fz_update_x_kernel<T>, dim3(G), dim3(kBlock), 0, st, k.n_nodes, per_n,
                         V.x, V.r, (const T*)V.rh, (const T*)V.ph, (const T*)V.sh, (const T*)V.s,
                         (const T*)V.t, c);
      hipLaunchKernelGGL(fz_update_p_kernel<T>, dim3(G), dim3(kBlock), 0, st, k.n_nodes, per_n,
                         (const T*)V.r, V.p, V.ph, (const T*)V.v, (const T*)w, V.minv, c);
    }
    done += chunk;
    if (done >= maxit) {
      // one more convergence test on the last residual
      hipLaunchKernelGGL(fz_spmv_v_kernel<T>, dim3(G), dim3(kBlock), 0, st, m, V.val,
                         (const T*)V.ph, (const T*)V.r, V.minv, V.v, V.s, V.sh, c);
    }
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if ((e = hipMemcpyAsync(ws->host_flag, ws->flag, 2 * sizeof(int32_t), hipMemcpyDeviceToHost,
                            st)) != hipSuccess)
      return e;
    if ((e = hipStreamSynchronize(st)) != hipSuccess) return e;
    prof_harvest(pf);
    *iters = ws->host_flag[1];
    if (ws->host_flag[0] != K_RUN || done >= maxit) {
      *state = ws->host_flag[0];
      return hipSuccess;
    }
  }
}

}  // namespace

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

hipError_t krylov_bicgstab(const KrylovArgs& k, const double* b, double* x, double rtol,
                           int maxit, int* iters, int* state, int sync_every) {
  Workspace* ws = k.ws;
  Vecs<double> V{ws->r, ws->rh, ws->p, ws->v, ws->ph, ws->s, ws->sh, ws->t, x, k.val, k.minv};
  if (k.n_pairs == 0 && k.fused)
    return bicgstab_fused_impl<double>(k, V, ws->w, b, rtol, maxit, iters, state, sync_every);
  return bicgstab_impl<double>(k, V, ws->tmp, b, rtol, maxit, iters, state, sync_every);
}

hipError_t apply_operator(const KrylovArgs& k, const double* x, double* y) {
  Ctl c{k.ws->flag, k.ws->scal, k.ws->partials, k.L.grid, 0.0};
  // the flag must read K_RUN for the launches to execute
  hipError_t e = hipMemsetAsync(k.ws->flag, 0, sizeof(int32_t), k.L.stream);
  if (e != hipSuccess) return e;
  return spmv_op<double>(k, c, k.val, x, y, k.ws->tmp, nullptr, 0, false, 0, 0);
}

hipError_t norm_and_update(int32_t n_nodes, const double* dx, double* xk, Workspace* ws, Launch L,
                           double* norm_out) {
  Ctl c{ws->flag, ws->scal, ws->partials, L.grid, 0.0};
  const int per_n = per_item(n_nodes, L.grid);
  hipLaunchKernelGGL(norm_update_kernel, dim3(L.grid), dim3(kBlock), 0, L.stream, n_nodes, per_n,
                     dx, xk, c);
  hipLaunchKernelGGL(finish_sum_kernel, dim3(1), dim3(kBlock), 0, L.stream, ws->partials, L.grid,
                     ws->scal + S_COUNT);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  e = hipMemcpyAsync(ws->host_pinned, ws->scal + S_COUNT, sizeof(double), hipMemcpyDeviceToHost,
                     L.stream);
  if (e != hipSuccess) return e;
  e = hipStreamSynchronize(L.stream);
  if (e != hipSuccess) return e;
  *norm_out = sqrt(ws->host_pinned[0]);
  return hipSuccess;
}

hipError_t krylov_mixed(const KrylovArgs& k, const double* b, double* x, double rtol,
                        double inner_rtol, int maxit, int* iters, int* state, int sync_every) {
  Workspace* ws = k.ws;
  const int G = k.L.grid;
  const int n = 3 * k.n_nodes;
  hipStream_t st = k.L.stream;
  Ctl c{ws->flag, ws->scal, ws->partials, G, 0.0};
  const int per_sl = per_item(k.pat.n_slices, G);
  hipError_t e = hipMemsetAsync(x, 0, sizeof(double) * n, st);
  if (e != hipSuccess) return e;
  // ||b||
  hipLaunchKernelGGL(dot2_kernel<double>, dim3(G), dim3(kBlock), 0, st, k.n_nodes,
                     per_item(k.n_nodes, G), b, b, (const double*)nullptr,
                     (const double*)nullptr, c, 1, 1, 0);
  hipLaunchKernelGGL(finish_sum_kernel, dim3(1), dim3(kBlock), 0, st, ws->partials + kParts, G,
                     ws->scal + S_COUNT);
  e = hipMemcpyAsync(ws->host_pinned, ws->scal + S_COUNT, sizeof(double), hipMemcpyDeviceToHost, st);
  if (e != hipSuccess) return e;
  if ((e = hipStreamSynchronize(st)) != hipSuccess) return e;
  const double bnorm = sqrt(ws->host_pinned[0]);
  *iters = 0;
  *state = K_CONVERGED;
  if (bnorm == 0.0) return hipSuccess;
  double rnorm = bnorm;
  const double* rsrc = b;
  Vecs<float> V{ws->r32, ws->rh32, ws->p32, ws->v32, ws->ph32, ws->s32, ws->sh32, ws->t32,
                ws->x32, k.val32, k.minv32};
  for (int outer = 0; outer < 20; ++outer) {
    // scale the residual to O(1) before the inner solve rounds it to FP32
    double* bs = ws->b;
    hipLaunchKernelGGL(scale_kernel, dim3(G), dim3(kBlock), 0, st, n, rsrc, bs, 1.0 / rnorm);
    int it = 0, stt = 0;
    if (k.n_pairs == 0 && k.fused)
      e = bicgstab_fused_impl<float>(k, V, ws->w32, bs, inner_rtol, maxit - *iters, &it, &stt,
                                     sync_every);
    else
      e = bicgstab_impl<float>(k, V, (float*)ws->tmp, bs, inner_rtol, maxit - *iters, &it, &stt,
                               sync_every);
    if (e != hipSuccess) return e;
    *iters += it;
    hipLaunchKernelGGL(axpy_f32_kernel, dim3(G), dim3(kBlock), 0, st, n, ws->x32, rnorm, x);
    // FP64 residual of the FP32-stored operator
    if (rtol >= inner_rtol) {
      // one FP32 solve reaches the requested tolerance; the NR loop (FP64
      // residual) is the refinement
      *state = (stt == K_CONVERGED) ? K_CONVERGED : stt;
      return hipSuccess;
    }
    if (k.n_pairs > 0) {
      // binding operator in FP64 through the promoted values is not
      // available in this path; the caller routes bound systems to FP64
      *state = K_BREAKDOWN;
      return hipSuccess;
    }
    hipLaunchKernelGGL(residual_kernel<float>, dim3(G), dim3(kBlock), 0, st, k.n_nodes,
                       k.pat.n_slices, per_sl, k.pat.slice_base, k.pat.col, k.val32,
                       k.pat.n_slots, x, b, ws->r, c);
    hipLaunchKernelGGL(finish_sum_kernel, dim3(1), dim3(kBlock), 0, st, ws->partials + kParts, G,
                       ws->scal + S_COUNT);
    e = hipMemcpyAsync(ws->host_pinned, ws->scal + S_COUNT, sizeof(double), hipMemcpyDeviceToHost,
                       st);
    if (e != hipSuccess) return e;
    if ((e = hipStreamSynchronize(st)) != hipSuccess) return e;
    rnorm = sqrt(ws->host_pinned[0]);
    rsrc = ws->r;
    if (!(rnorm > rtol * bnorm)) { *state = K_CONVERGED; return hipSuccess; }
    if (!std::isfinite(rnorm) || stt == K_BREAKDOWN) { *state = K_BREAKDOWN; return hipSuccess; }
    if (*iters >= maxit) { *state = K_RUN; return hipSuccess; }
  }
  *state = K_RUN;
  return hipSuccess;
}

}  // namespace tf
