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
  if (!isfinite(rnorm) || !isfinite(rho_new) || rnorm > kDivergence * bnorm) state = K_BREAKDOWN;
  else if (rnorm <= c.rtol * bnorm) state = K_CONVERGED;
  else if (rho_new == 0.0 || c.scal[S_OMEGA] == 0.0) state = K_BREAKDOWN;
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
    pf->kind[pf->used] = sizeof(T) == 4 ? 1 : 0;  // the v = J ph product
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
  if (k.n_pairs == 0) {
    // no binding: the tile-based solver (one tissue, or each tissue of a batch)
    static thread_local TiledResult res;
    hipError_t e = tiled_bicgstab(k, b, x, rtol, maxit, sync_every, &res);
    *iters = res.iters[0];
    *state = res.state[0];
    return e;
  }
  Vecs<double> V{ws->r, ws->rh, ws->p, ws->v, ws->ph, ws->s, ws->sh, ws->t, x, k.val, k.minv};
  return bicgstab_impl<double>(k, V, ws->tmp, b, rtol, maxit, iters, state, sync_every);
}

}  // namespace tf
