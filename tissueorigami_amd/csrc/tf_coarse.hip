// Two-level preconditioner for the FP32 CG solve of the mixed-precision step:
//   M^-1 r = D^-1 r + P A_c^-1 P^T r,   A_c = P^T A P.
//
// No counterpart in the reference, which factorises K with Pardiso.  Block
// Jacobi (D^-1) alone leaves the smooth, tissue-wide deformation modes to the
// Krylov iteration, so the iteration count grows with the mesh (about 200 CG
// iterations per NR iteration at 1.6 M prisms).  The coarse space P holds,
// for every aggregate of nodes, the six rigid-body modes of the aggregate
// (three translations, three rotations about its centroid); an aggregate is
// the set of nodes of one tissue whose reference (x, y) lies in one square
// cell and whose reference z lies in one band of the tissue's thickness.
// Fixed and slave DOF have zero rows in P, so the correction never moves them.
//
// Every tissue of a handle is a segment with its own aggregates, so A_c is
// block-diagonal: one dense block, inverse and Newton-Schulz state per
// segment, and a tissue's correction does not depend on the other tissues.
//
// Per setup: A_c is formed block by block from the FP32 operator (one
// workgroup per aggregate pair, a fixed summation order).  A reference
// inverse X_ref comes from rocSOLVER (getrf, getri) on the first setup; the
// operator changes little from one setup to the next, so every setup starts
// from X = X_ref and takes Newton-Schulz steps X <- X (2I - A_c X) (two
// rocBLAS GEMMs each, quadratic convergence) until max|I - A_c X| < 1e-5 (a
// preconditioner needs no more).  When the steps do not converge (residual
// above 0.25, or seven steps), X_ref is refactorised from the current A_c.
// The inverse is kept as a dense FP32 matrix.  Per CG iteration: w = P^T r
// (one workgroup per aggregate), y = A_c^-1 w (one wavefront per row) and
// the partial products (w, y) that complete (r, z); the CG's search-direction
// update adds P y (tzc_pupdate).  Every sum has a fixed order and rocBLAS runs
// with atomics off, so the solve stays bitwise reproducible.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include "tf_internal.h"

namespace tf {
namespace {

constexpr int kMinAggNodes = 16;   // smaller cells are merged into the nearest larger one
constexpr int kMaxAggregates = 8192;  // per segment
constexpr int kGemvRows = kBlock / 64;  // rows of A_c^-1 per product workgroup

template <typename T>
hipError_t upload(CoarseOp& c, T** p, const std::vector<T>& v, size_t count = 0) {
  const size_t n = std::max<size_t>(count ? count : v.size(), 1);
  void* q = nullptr;
  hipError_t e = hipMalloc(&q, n * sizeof(T));
  if (e != hipSuccess) return e;
  c.allocs.push_back(q);
  *p = (T*)q;
  if (!v.empty()) e = hipMemcpy(q, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice);
  return e;
}

// P row p of node i (masked), columns 0..5: translation e_p, rotation
// d(omega x d)_p / d omega_r.
__device__ inline void p_rows(const float* d, const uint8_t* flags, int i, double P[3][6]) {
  const double d0 = d[3 * i], d1 = d[3 * i + 1], d2 = d[3 * i + 2];
  const double R[3][3] = {{0.0, d2, -d1}, {-d2, 0.0, d0}, {d1, -d0, 0.0}};
#pragma unroll
  for (int p = 0; p < 3; ++p) {
    const bool off = flags && (flags[3 * i + p] & kCoarseExcluded);
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      P[p][a] = off ? 0.0 : (p == a ? 1.0 : 0.0);
      P[p][3 + a] = off ? 0.0 : R[p][a];
    }
  }
}

// Workgroup sum of n values per thread in a fixed order (wave trees, then
// waves in order); the result is valid in thread 0.
template <int N, int B = kBlock>
__device__ inline void block_sum(double (&v)[N], double (*red)[B / 64]) {
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
      double s = 0.0;
#pragma unroll
      for (int w = 0; w < B / 64; ++w) s += red[k][w];
      v[k] = s;
    }
  }
}

// (P_I^T A P_K) for every aggregate pair (I, K), from the operator slots
// listed for the pair.
template <typename V>
__global__ __launch_bounds__(kBlock) void coarse_pairs_kernel(
    int32_t n_pairs, const int32_t* __restrict__ cp_ptr, const int32_t* __restrict__ cp_slot,
    const int32_t* __restrict__ cp_row, const int32_t* __restrict__ col,
    const V* __restrict__ val, int64_t n_slots, const float* __restrict__ d,
    const uint8_t* __restrict__ flags, double* __restrict__ out) {
  __shared__ double red[36][kBlock / 64];
  const int pr = blockIdx.x;
  if (pr >= n_pairs) return;
  double acc[36];
#pragma unroll
  for (int q = 0; q < 36; ++q) acc[q] = 0.0;
  for (int t = cp_ptr[pr] + threadIdx.x; t < cp_ptr[pr + 1]; t += kBlock) {
    const int slot = cp_slot[t], i = cp_row[t], k = col[slot];
    double J[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) J[q] = (double)val[(size_t)q * n_slots + slot];
    double Pi[3][6], Pk[3][6];
    p_rows(d, flags, i, Pi);
    p_rows(d, flags, k, Pk);
    double T[3][6];  // J P_k
#pragma unroll
    for (int p = 0; p < 3; ++p)
#pragma unroll
      for (int b = 0; b < 6; ++b)
        T[p][b] = J[3 * p] * Pk[0][b] + J[3 * p + 1] * Pk[1][b] + J[3 * p + 2] * Pk[2][b];
#pragma unroll
    for (int a = 0; a < 6; ++a)
#pragma unroll
      for (int b = 0; b < 6; ++b)
        acc[6 * a + b] += Pi[0][a] * T[0][b] + Pi[1][a] * T[1][b] + Pi[2][a] * T[2][b];
  }
  block_sum<36>(acc, red);
  if (threadIdx.x == 0)
#pragma unroll
    for (int q = 0; q < 36; ++q) out[36 * (size_t)pr + q] = acc[q];
}

__global__ __launch_bounds__(kBlock) void coarse_fill_kernel(double* a, int64_t n, double v) {
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kBlock)
    a[i] = v;
}

// Pair blocks into the dense column-major A_c block of their segment (each
// pair owns its 6x6 block; both aggregates of a pair lie in one segment).
__global__ void coarse_scatter_kernel(int32_t n_pairs, const int32_t* cp_IK,
                                      const int32_t* agg_tissue, const int32_t* segd,
                                      const int64_t* segm, const double* blk, double* dense) {
  const int pr = blockIdx.x, q = threadIdx.x;
  if (pr >= n_pairs || q >= 36) return;
  const int I = cp_IK[2 * pr], K = cp_IK[2 * pr + 1];
  const int T = agg_tissue[I];
  const int a0 = segd[2 * T], nc = segd[2 * T + 1];
  const int a = q / 6, b = q % 6;
  dense[segm[T] + (int64_t)(6 * (K - a0) + b) * nc + 6 * (I - a0) + a] = blk[36 * (size_t)pr + q];
}

// A mode with no free DOF (all its rows masked) has a zero row and column:
// 1 on its diagonal keeps A_c invertible and leaves its y at 0.
__global__ __launch_bounds__(kBlock) void coarse_diag_kernel(int32_t n_c, double* dense) {
  const int r = blockIdx.x * kBlock + threadIdx.x;
  if (r < n_c && dense[(size_t)r * n_c + r] == 0.0) dense[(size_t)r * n_c + r] = 1.0;
}

// Row-major copy (row stride ld; FP32 or FP64) of the column-major FP64
// inverse, through 32 x 32 LDS tiles (grid: tiles of rows x tiles of columns).
template <typename O>
__global__ __launch_bounds__(kBlock) void coarse_convert_kernel(int32_t n_c, int32_t ld,
                                                                const double* X, O* inv32) {
  __shared__ O tile[32][33];
  const int r0 = blockIdx.x * 32, c0 = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  for (int k = ty; k < 32; k += 8) {
    const int r = r0 + tx, cc = c0 + k;  // read down a column of X
    tile[k][tx] = (r < n_c && cc < n_c) ? (O)X[(size_t)cc * n_c + r] : (O)0;
  }
  __syncthreads();
  for (int k = ty; k < 32; k += 8) {
    const int r = r0 + k, cc = c0 + tx;  // write along a row of inv32
    if (r < n_c && cc < n_c) inv32[(size_t)r * ld + cc] = tile[tx][k];
  }
}

// max |I - T| of every column (one workgroup per column; a NaN counts as a
// huge residual); the host takes the maximum over the columns.
__global__ __launch_bounds__(kBlock) void coarse_resid_kernel(int32_t n_c, const double* T,
                                                              double* colmax) {
  __shared__ double red[kBlock / 64];
  const int cc = blockIdx.x;
  double m = 0.0;
  for (int r = threadIdx.x; r < n_c; r += kBlock) {
    const double v = fabs((r == cc ? 1.0 : 0.0) - T[(size_t)cc * n_c + r]);
    m = fmax(m, v == v ? v : 1e300);
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = red[0];
#pragma unroll
    for (int k = 1; k < kBlock / 64; ++k) t = fmax(t, red[k]);
    colmax[cc] = t;
  }
}

// T <- 2I - T; grid: row blocks x columns.
__global__ __launch_bounds__(kBlock) void coarse_twoI_kernel(int32_t n_c, double* T) {
  const int r = blockIdx.x * kBlock + threadIdx.x, cc = blockIdx.y;
  if (r < n_c) {
    double& t = T[(size_t)cc * n_c + r];
    t = (r == cc ? 2.0 : 0.0) - t;
  }
}



// max |V - W| of every probe column (one workgroup per column), where W =
// A_c X_ref V: the probe estimate of how far X_ref is from the inverse of the
// current A_c.  V has entries +-1, so a row i contributes |sum_j R_ij v_j|,
// about the 2-norm of row i of R = I - A_c X_ref, never less than its largest
// entry on average.
__global__ __launch_bounds__(kBlock) void coarse_probe_kernel(int32_t n_c, const double* V,
                                                              const double* W, double* out) {
  __shared__ double red[kBlock / 64];
  const int j = blockIdx.x;
  double m = 0.0;
  for (int r = threadIdx.x; r < n_c; r += kBlock) {
    const double v = fabs(V[(size_t)j * n_c + r] - W[(size_t)j * n_c + r]);
    m = fmax(m, v == v ? v : 1e300);
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = red[0];
#pragma unroll
    for (int k = 1; k < kBlock / 64; ++k) t = fmax(t, red[k]);
    out[j] = t;
  }
}

// w = P^T r: one workgroup per aggregate over its node list.  The offsets and
// masks are stored in list order, only r is gathered; every thread has kRU
// list entries in flight (indices first, then offsets, masks and the gathered
// r) before it adds any of them, in list order.
constexpr int kRU = 4;
constexpr int kRBlock = 512;  // threads per aggregate: one pass over 2048 list entries
template <typename R>
__global__ __launch_bounds__(kRBlock) void coarse_restrict_kernel(
    int32_t n_agg, const int32_t* __restrict__ agg_ptr, const int32_t* __restrict__ agg_nodes,
    const int32_t* __restrict__ agg_tissue, const int32_t* __restrict__ agg_cbase,
    const int32_t* __restrict__ seg_on, const float* __restrict__ list_d,
    const uint8_t* __restrict__ list_mask, const R* __restrict__ r,
    const int32_t* __restrict__ state, R* __restrict__ w) {
  __shared__ double red[6][kRBlock / 64];
  const int a = blockIdx.x;
  if (a >= n_agg) return;
  const int T = agg_tissue[a];
  if (state[2 * T] != K_RUN || !seg_on[T]) return;
  double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  const int beg = agg_ptr[a], end = agg_ptr[a + 1];
  for (int t0 = beg + threadIdx.x; t0 < end; t0 += kRU * kRBlock) {
    int idx[kRU];
#pragma unroll
    for (int u = 0; u < kRU; ++u) {
      const int t = t0 + u * kRBlock;
      idx[u] = t < end ? agg_nodes[t] : -1;
    }
    double d[kRU][3], rv[kRU][3];
#pragma unroll
    for (int u = 0; u < kRU; ++u) {
      const int t = t0 + u * kRBlock;
      const uint8_t mk = idx[u] >= 0 ? list_mask[t] : (uint8_t)7;
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        d[u][k] = idx[u] >= 0 ? (double)list_d[3 * t + k] : 0.0;
        rv[u][k] = (mk >> k) & 1 ? 0.0 : (double)r[3 * idx[u] + k];
      }
    }
#pragma unroll
    for (int u = 0; u < kRU; ++u) {
      const double r0 = rv[u][0], r1 = rv[u][1], r2 = rv[u][2];
      const double d0 = d[u][0], d1 = d[u][1], d2 = d[u][2];
      acc[0] += r0;
      acc[1] += r1;
      acc[2] += r2;
      // rotation rows of P^T: d x r
      acc[3] += d1 * r2 - d2 * r1;
      acc[4] += d2 * r0 - d0 * r2;
      acc[5] += d0 * r1 - d1 * r0;
    }
  }
  block_sum<6, kRBlock>(acc, red);
  if (threadIdx.x == 0)
#pragma unroll
    for (int q = 0; q < 6; ++q) w[agg_cbase[a] + q] = (R)acc[q];
}

// Fixed / slave bits of every listed node (list order), from the DOF flags.
__global__ __launch_bounds__(kBlock) void coarse_mask_kernel(int32_t n_list,
                                                             const int32_t* __restrict__ agg_nodes,
                                                             const uint8_t* __restrict__ flags,
                                                             uint8_t* __restrict__ list_mask) {
  for (int t = blockIdx.x * kBlock + threadIdx.x; t < n_list; t += gridDim.x * kBlock) {
    const int i = agg_nodes[t];
    uint8_t m = 0;
    if (flags) {
#pragma unroll
      for (int p = 0; p < 3; ++p)
        if (flags[3 * i + p] & kCoarseExcluded) m |= (uint8_t)(1 << p);
    }
    list_mask[t] = m;
  }
}

// Per node: the first coarse DOF of its aggregate, shifted left by three,
// with bit p set when DOF p takes no part in the coarse space (fixed or
// slave); -1 for a node without an aggregate.  The search-direction update
// reads this one word instead of the aggregate, its base and the DOF flags.
__global__ __launch_bounds__(kBlock) void coarse_nodecb_kernel(int32_t n_nodes,
                                                               const int32_t* __restrict__ node_agg,
                                                               const int32_t* __restrict__ agg_cbase,
                                                               const uint8_t* __restrict__ flags,
                                                               int32_t* __restrict__ node_cb) {
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n_nodes; i += gridDim.x * kBlock) {
    const int a = node_agg[i];
    int32_t cb = -1;
    if (a >= 0) {
      int32_t m = 0;
      if (flags) {
#pragma unroll
        for (int p = 0; p < 3; ++p)
          if (flags[3 * i + p] & kCoarseExcluded) m |= 1 << p;
      }
      cb = (agg_cbase[a] << 3) | m;
    }
    node_cb[i] = cb;
  }
}

// The first two product workgroups of every segment also form its tissue's
// ordered sums of the (r, z_D) and (r, r) tile partials (the same sums the
// search-direction update would form), so that the update's workgroups read
// two numbers instead of every tile partial.
__device__ inline void s1_prereduce(const S1Pre& pre, int T, int block) {
  if (!pre.out) return;
  const int g0 = pre.cpart_off[T], g1 = pre.cpart_off[T + 1];
  // the segment's first workgroup forms the (r, z_D) sum, its second the
  // (r, r) sum (the first both when the segment has one product workgroup)
  const bool two = g1 - g0 > 1;
  const bool do_rz = block == g0, do_rr = two ? block == g0 + 1 : block == g0;
  if (!do_rz && !do_rr) return;
  const int t0 = pre.part_off[T], t1 = pre.part_off[T + 1];
  if (do_rz) {
    const double v = wg_ordered_sum(pre.rz, t0, t1);
    if (threadIdx.x == 0) pre.out[2 * T] = v;
  }
  if (do_rr) {
    const double v = wg_ordered_sum(pre.rr, t0, t1);
    if (threadIdx.x == 0) pre.out[2 * T + 1] = v;
  }
}

// y = A_c^-1 w, one wavefront per row (16-byte loads over the padded row
// stride ld); partial (w, y) per workgroup.  With w2, w is the sum w + w2 of
// the two half-aggregate terms, formed in that order.
typedef float v4f __attribute__((ext_vector_type(4)));
__global__ __launch_bounds__(kBlock) void coarse_gemv_kernel(
    const int32_t* __restrict__ gw, const int32_t* __restrict__ sgi,
    const int64_t* __restrict__ sgio, const int32_t* __restrict__ seg_on,
    const float* __restrict__ inv32, const float* __restrict__ w, const float* __restrict__ w2,
    const int32_t* __restrict__ state, float* __restrict__ y, double* __restrict__ cpart,
    S1Pre pre) {
  __shared__ double wy[kGemvRows];
  const int T = gw[2 * blockIdx.x], row0 = gw[2 * blockIdx.x + 1];
  if (state[2 * T] != K_RUN) return;
  s1_prereduce(pre, T, (int)blockIdx.x);
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int nc = sgi[3 * T], ld = sgi[3 * T + 1], off = sgi[3 * T + 2];
  const int row = row0 + wave;
  double dot = 0.0;
  if (seg_on[T] && row < nc) {
    const v4f* a = (const v4f*)(inv32 + sgio[T] + (size_t)row * ld);
    const float* ws = w + off;
    float s = 0.f;
    if (w2) {
      const float* ws2 = w2 + off;
      for (int j = lane; j < ld / 4; j += 64) {
        const v4f v = ((const v4f*)ws)[j] + ((const v4f*)ws2)[j];
        const v4f u = a[j];
        s += u.x * v.x + u.y * v.y + u.z * v.z + u.w * v.w;
      }
    } else {
      for (int j = lane; j < ld / 4; j += 64) {
        const v4f v = ((const v4f*)ws)[j];
        const v4f u = a[j];
        s += u.x * v.x + u.y * v.y + u.z * v.z + u.w * v.w;
      }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if (lane == 0) y[off + row] = s;
    const float wr = w2 ? ws[row] + w2[off + row] : ws[row];
    dot = (double)wr * (double)s;
  }
  if (lane == 0) wy[wave] = dot;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
#pragma unroll
    for (int k = 0; k < kGemvRows; ++k) t += wy[k];
    cpart[blockIdx.x] = t;
  }
}

// The FP32 product; w2 as above (null: w alone).
void launch_gemv32(const CoarseOp& c, const int32_t* state, const float* w2, const S1Pre& pre,
                   hipStream_t st) {
  hipLaunchKernelGGL(coarse_gemv_kernel, dim3(c.n_gemv), dim3(kBlock), 0, st, c.gw, c.sgi,
                     c.sgio, c.seg_on, c.inv32, c.w, w2, state, c.y, c.cpart, pre);
}

// The FP64 product for the FP64 CG: y = A_c^-1 w with the FP64 inverse.
typedef double v2d __attribute__((ext_vector_type(2)));
__global__ __launch_bounds__(kBlock) void coarse_gemv64_kernel(
    const int32_t* __restrict__ gw, const int32_t* __restrict__ sgi,
    const int64_t* __restrict__ sgio, const int32_t* __restrict__ seg_on,
    const double* __restrict__ inv64, const double* __restrict__ w,
    const int32_t* __restrict__ state, double* __restrict__ y, double* __restrict__ cpart,
    S1Pre pre) {
  __shared__ double wy[kGemvRows];
  const int T = gw[2 * blockIdx.x], row0 = gw[2 * blockIdx.x + 1];
  if (state[2 * T] != K_RUN) return;
  s1_prereduce(pre, T, (int)blockIdx.x);
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int nc = sgi[3 * T], ld = sgi[3 * T + 1], off = sgi[3 * T + 2];
  const int row = row0 + wave;
  double dot = 0.0;
  if (seg_on[T] && row < nc) {
    const v2d* a = (const v2d*)(inv64 + sgio[T] + (size_t)row * ld);
    const double* ws = w + off;
    double s = 0.0;
    for (int j = lane; j < ld / 2; j += 64) {
      const v2d v = ((const v2d*)ws)[j];
      const v2d u = a[j];
      s += u.x * v.x + u.y * v.y;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if (lane == 0) y[off + row] = s;
    dot = ws[row] * s;
  }
  if (lane == 0) wy[wave] = dot;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
#pragma unroll
    for (int k = 0; k < kGemvRows; ++k) t += wy[k];
    cpart[blockIdx.x] = t;
  }
}

int grid_nodes(int64_t n) {
  int64_t g = (n + kBlock - 1) / kBlock;
  return (int)std::max<int64_t>(1, std::min<int64_t>(g, 65535 * 4));
}

}  // namespace

void coarse_release(CoarseOp& c) {
  for (void* p : c.allocs) (void)hipFree(p);
  c.allocs.clear();
  if (c.blas) (void)rocblas_destroy_handle((rocblas_handle)c.blas);
  c = CoarseOp();
}

hipError_t coarse_build(CoarseOp& c, int32_t n_nodes, const double* X,
                        const std::vector<CoarseSeg>& segs, int z_bands,
                        const std::vector<int32_t>& slice_base, const std::vector<int32_t>& col,
                        const std::vector<int32_t>& row_len) {
  void* blas = c.blas;
  const int32_t gen = c.gen;
  c.blas = nullptr;
  coarse_release(c);
  c.blas = blas;
  c.gen = gen;
  if (n_nodes <= 0 || segs.empty() || z_bands < 1) return hipErrorInvalidValue;
  const int n_seg = (int)segs.size();
  std::vector<int32_t> agg(n_nodes, -1);
  std::vector<float> dn(3 * (size_t)n_nodes, 0.f);
  std::vector<int32_t> seg_agg(n_seg + 1, 0), na_all, agg_tissue;
  for (int T = 0; T < n_seg; ++T) {
    const int b0 = segs[T].begin, b1 = segs[T].end;
    const double cell = segs[T].cell;
    if (b1 <= b0 || !(cell > 0.0)) return hipErrorInvalidValue;
    const int nn = b1 - b0;
    // square cells over the segment's reference (x, y), each split into
    // z_bands equal bands of its reference z range
    double x0 = X[3 * b0], y0 = X[3 * b0 + 1], z0 = X[3 * b0 + 2], z1 = X[3 * b0 + 2];
    for (int i = b0; i < b1; ++i) {
      x0 = std::min(x0, X[3 * i]);
      y0 = std::min(y0, X[3 * i + 1]);
      z0 = std::min(z0, X[3 * i + 2]);
      z1 = std::max(z1, X[3 * i + 2]);
    }
    const double zb = z1 > z0 ? (z1 - z0) / z_bands : 1.0;
    std::vector<int64_t> key(nn);
    for (int q = 0; q < nn; ++q) {
      const double* P = X + 3 * (size_t)(b0 + q);
      const int64_t ix = (int64_t)std::floor((P[0] - x0) / cell);
      const int64_t iy = (int64_t)std::floor((P[1] - y0) / cell);
      const int64_t iz = std::min<int64_t>(z_bands - 1, (int64_t)std::floor((P[2] - z0) / zb));
      key[q] = (ix << 40) | (iy << 16) | iz;
    }
    std::vector<int64_t> keys(key);
    std::sort(keys.begin(), keys.end());
    keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
    const int n_cells = (int)keys.size();
    std::vector<int32_t> cell_of(nn), cnt(n_cells, 0);
    for (int q = 0; q < nn; ++q) {
      cell_of[q] = (int32_t)(std::lower_bound(keys.begin(), keys.end(), key[q]) - keys.begin());
      ++cnt[cell_of[q]];
    }
    std::vector<double> cen(3 * (size_t)n_cells, 0.0);
    for (int q = 0; q < nn; ++q)
      for (int k = 0; k < 3; ++k) cen[3 * cell_of[q] + k] += X[3 * (size_t)(b0 + q) + k];
    for (int g = 0; g < n_cells; ++g)
      for (int k = 0; k < 3; ++k) cen[3 * g + k] /= cnt[g];
    // small cells join the nearest cell of kMinAggNodes nodes or more (all one
    // aggregate when there is none)
    std::vector<int32_t> target(n_cells), agg_of_cell(n_cells, -1);
    std::vector<int32_t> big;
    for (int g = 0; g < n_cells; ++g)
      if (cnt[g] >= kMinAggNodes) big.push_back(g);
    for (int g = 0; g < n_cells; ++g) {
      if (cnt[g] >= kMinAggNodes) {
        target[g] = g;
      } else if (big.empty()) {
        target[g] = 0;
      } else {
        double best = 1e300;
        int bq = big[0];
        for (int bb : big) {
          const double dx = cen[3 * g] - cen[3 * bb], dy = cen[3 * g + 1] - cen[3 * bb + 1];
          const double dz = cen[3 * g + 2] - cen[3 * bb + 2];
          const double d2 = dx * dx + dy * dy + dz * dz;
          if (d2 < best) { best = d2; bq = bb; }
        }
        target[g] = bq;
      }
    }
    int n_loc = 0;
    for (int g = 0; g < n_cells; ++g)
      if (agg_of_cell[target[g]] < 0) agg_of_cell[target[g]] = n_loc++;
    const int a_base = seg_agg[T];
    if (n_loc > kMaxAggregates) return hipErrorInvalidValue;
    seg_agg[T + 1] = a_base + n_loc;
    for (int q = 0; q < nn; ++q) agg[b0 + q] = a_base + agg_of_cell[target[cell_of[q]]];
    // centroid and radius per aggregate; scaled offsets per node
    std::vector<double> ac(3 * (size_t)n_loc, 0.0), rad(n_loc, 0.0);
    std::vector<int32_t> na(n_loc, 0);
    for (int q = 0; q < nn; ++q) {
      const int a = agg[b0 + q] - a_base;
      ++na[a];
      for (int k = 0; k < 3; ++k) ac[3 * a + k] += X[3 * (size_t)(b0 + q) + k];
    }
    for (int a = 0; a < n_loc; ++a)
      for (int k = 0; k < 3; ++k) ac[3 * a + k] /= na[a];
    for (int q = 0; q < nn; ++q) {
      const int a = agg[b0 + q] - a_base;
      double sq = 0.0;
      for (int k = 0; k < 3; ++k) {
        const double t = X[3 * (size_t)(b0 + q) + k] - ac[3 * a + k];
        sq += t * t;
      }
      rad[a] = std::max(rad[a], std::sqrt(sq));
    }
    for (int q = 0; q < nn; ++q) {
      const int a = agg[b0 + q] - a_base;
      const double L = rad[a] > 0.0 ? rad[a] : 1.0;
      for (int k = 0; k < 3; ++k)
        dn[3 * (size_t)(b0 + q) + k] = (float)((X[3 * (size_t)(b0 + q) + k] - ac[3 * a + k]) / L);
    }
    for (int a = 0; a < n_loc; ++a) {
      na_all.push_back(na[a]);
      agg_tissue.push_back(T);
    }
  }
  const int n_agg = seg_agg[n_seg];
  std::vector<int32_t> aptr(n_agg + 1, 0);
  for (int a = 0; a < n_agg; ++a) aptr[a + 1] = aptr[a] + na_all[a];
  const int n_list = aptr[n_agg];
  std::vector<int32_t> anodes(n_list);
  {
    std::vector<int32_t> fill(aptr.begin(), aptr.end() - 1);
    for (int i = 0; i < n_nodes; ++i)
      if (agg[i] >= 0) anodes[fill[agg[i]]++] = i;
  }
  // segment layout: coarse DOF, vector offsets (multiples of 64), dense and
  // FP32 matrix offsets (multiples of 64 elements), product workgroups
  c.n_seg = n_seg;
  c.seg_agg = seg_agg;
  c.seg_nc.assign(n_seg, 0);
  c.seg_ld.assign(n_seg, 0);
  c.seg_coff.assign(n_seg, 0);
  c.seg_moff.assign(n_seg, 0);
  c.seg_ioff.assign(n_seg, 0);
  c.seg_goff.assign(n_seg + 1, 0);
  c.seg_poff.assign(n_seg, 0);
  c.seg_have_inv.assign(n_seg, 0);
  c.seg_active.assign(n_seg, 0);
  int32_t vlen = 0, poff = 0;
  int64_t moff = 0, ioff = 0;
  std::vector<int32_t> gw, segd(2 * (size_t)n_seg), cbase(n_agg), sgi(3 * (size_t)n_seg);
  std::vector<int64_t> segm(n_seg), sgio(n_seg);
  for (int T = 0; T < n_seg; ++T) {
    const int nc = 6 * (seg_agg[T + 1] - seg_agg[T]);
    const int ld = (nc + 63) / 64 * 64;
    c.seg_nc[T] = nc;
    c.seg_ld[T] = ld;
    c.seg_coff[T] = vlen;
    c.seg_moff[T] = moff;
    c.seg_ioff[T] = ioff;
    c.seg_poff[T] = poff;
    for (int a = seg_agg[T]; a < seg_agg[T + 1]; ++a) cbase[a] = vlen + 6 * (a - seg_agg[T]);
    segd[2 * T] = seg_agg[T];
    segd[2 * T + 1] = nc;
    segm[T] = moff;
    sgi[3 * T] = nc;
    sgi[3 * T + 1] = ld;
    sgi[3 * T + 2] = vlen;
    sgio[T] = ioff;
    for (int r0 = 0; r0 < nc; r0 += kGemvRows) {
      gw.push_back(T);
      gw.push_back(r0);
    }
    c.seg_goff[T + 1] = (int32_t)(gw.size() / 2);
    vlen += ld;
    moff += ((int64_t)nc * nc + 63) / 64 * 64;
    ioff += (int64_t)nc * ld;
    poff += nc;
  }
  // operator slots grouped by aggregate pair, ascending within each pair
  std::vector<int32_t> pIK, pcnt;
  std::vector<std::vector<std::pair<int32_t, int32_t>>> pair_of(n_agg);  // (K, pair id) per I
  auto pair_id = [&](int I, int K, bool create) -> int32_t {
    auto& v = pair_of[I];
    for (auto& e : v)
      if (e.first == K) return e.second;
    if (!create) return -1;
    const int32_t id = (int32_t)pcnt.size();
    v.push_back({K, id});
    pcnt.push_back(0);
    pIK.push_back(I);
    pIK.push_back(K);
    return id;
  };
  const int n_slices = (int)slice_base.size() - 1;
  auto for_slots = [&](auto f) {
    for (int sl = 0; sl < n_slices; ++sl) {
      const int w = (slice_base[sl + 1] - slice_base[sl]) / kSlice;
      for (int k = 0; k < w; ++k)
        for (int lane = 0; lane < kSlice; ++lane) {
          const int R = sl * kSlice + lane;
          if (R >= n_nodes || k >= row_len[R]) continue;
          const int slot = slice_base[sl] + k * kSlice + lane;
          const int C = col[slot];
          // A_c keeps one block per segment: couplings between tissues (none
          // in a batch of separate tissues) stay out of it
          if (agg[R] < 0 || agg[C] < 0 || agg_tissue[agg[R]] != agg_tissue[agg[C]]) continue;
          f(slot, R, C);
        }
    }
  };
  int64_t listed = 0;
  for_slots([&](int, int R, int C) {
    ++pcnt[pair_id(agg[R], agg[C], true)];
    ++listed;
  });
  const int n_pairs = (int)pcnt.size();
  std::vector<int32_t> pptr(n_pairs + 1, 0), pslot(listed), prow(listed);
  for (int q = 0; q < n_pairs; ++q) pptr[q + 1] = pptr[q] + pcnt[q];
  {
    std::vector<int32_t> fill(pptr.begin(), pptr.end() - 1);
    for_slots([&](int slot, int R, int C) {
      const int32_t id = pair_id(agg[R], agg[C], false);
      pslot[fill[id]] = slot;
      prow[fill[id]] = R;
      ++fill[id];
    });
  }
  // node offsets in list order
  std::vector<float> ldv(3 * (size_t)std::max(n_list, 1));
  for (int t = 0; t < n_list; ++t)
    for (int k = 0; k < 3; ++k) ldv[3 * (size_t)t + k] = dn[3 * (size_t)anodes[t] + k];
  c.n_agg = n_agg;
  c.n_nodes = n_nodes;
  c.n_c = 6 * n_agg;
  c.n_cpairs = n_pairs;
  c.n_gemv = (int32_t)(gw.size() / 2);
  c.vlen = vlen;
  c.agg_total = n_list;
  hipError_t e;
  if ((e = upload(c, &c.node_agg, agg)) != hipSuccess) return e;
  if ((e = upload(c, &c.node_cb, agg)) != hipSuccess) return e;
  if ((e = upload(c, &c.node_d, dn)) != hipSuccess) return e;
  if ((e = upload(c, &c.agg_ptr, aptr)) != hipSuccess) return e;
  if ((e = upload(c, &c.agg_nodes, anodes)) != hipSuccess) return e;
  if ((e = upload(c, &c.agg_tissue, agg_tissue)) != hipSuccess) return e;
  if ((e = upload(c, &c.agg_cbase, cbase)) != hipSuccess) return e;
  if ((e = upload(c, &c.list_d, ldv)) != hipSuccess) return e;
  if ((e = upload(c, &c.list_mask, std::vector<uint8_t>((size_t)std::max(n_list, 1), 0))) != hipSuccess) return e;
  if ((e = upload(c, &c.cp_ptr, pptr)) != hipSuccess) return e;
  if ((e = upload(c, &c.cp_slot, pslot)) != hipSuccess) return e;
  if ((e = upload(c, &c.cp_row, prow)) != hipSuccess) return e;
  if ((e = upload(c, &c.cp_IK, pIK)) != hipSuccess) return e;
  if ((e = upload(c, &c.gw, gw)) != hipSuccess) return e;
  if ((e = upload(c, &c.cpart_off, c.seg_goff)) != hipSuccess) return e;
  if ((e = upload(c, &c.seg_on, std::vector<int32_t>(n_seg, 0))) != hipSuccess) return e;
  if ((e = upload(c, &c.segd, segd)) != hipSuccess) return e;
  if ((e = upload(c, &c.segm, segm)) != hipSuccess) return e;
  if ((e = upload(c, &c.sgi, sgi)) != hipSuccess) return e;
  if ((e = upload(c, &c.sgio, sgio)) != hipSuccess) return e;
  if ((e = upload(c, &c.pairblk, std::vector<double>(), 36 * (size_t)n_pairs)) != hipSuccess) return e;
  if ((e = upload(c, &c.Ac, std::vector<double>(), (size_t)moff)) != hipSuccess) return e;
  if ((e = upload(c, &c.X, std::vector<double>(), (size_t)moff)) != hipSuccess) return e;
  if ((e = upload(c, &c.Xref, std::vector<double>(), (size_t)moff)) != hipSuccess) return e;
  if ((e = upload(c, &c.X2, std::vector<double>(), (size_t)moff)) != hipSuccess) return e;
  if ((e = upload(c, &c.T, std::vector<double>(), (size_t)moff)) != hipSuccess) return e;
  c.colmax_len = std::max(poff, 1);
  {
    // probe block V of every segment: kCoarseProbes columns of +-1 entries
    // from a fixed hash of (row, column), the same for every build
    std::vector<double> pv((size_t)kCoarseProbes * c.colmax_len, 0.0);
    for (int T = 0; T < n_seg; ++T)
      for (int j = 0; j < kCoarseProbes; ++j)
        for (int r = 0; r < c.seg_nc[T]; ++r) {
          const uint32_t hsh = (uint32_t)r * 2654435761u ^ (uint32_t)(j + 1) * 40503u;
          pv[(size_t)kCoarseProbes * c.seg_poff[T] + (size_t)j * c.seg_nc[T] + r] =
              ((hsh >> 15) & 1u) ? 1.0 : -1.0;
        }
    if ((e = upload(c, &c.probe_v, pv)) != hipSuccess) return e;
    if ((e = upload(c, &c.probe_u, std::vector<double>(), pv.size())) != hipSuccess) return e;
    if ((e = upload(c, &c.probe_w, std::vector<double>(), pv.size())) != hipSuccess) return e;
    if ((e = upload(c, &c.probe_max, std::vector<double>((size_t)kCoarseProbes * n_seg, 0.0))) !=
        hipSuccess)
      return e;
  }
  if ((e = upload(c, &c.colmax, std::vector<double>((size_t)c.colmax_len, 0.0))) != hipSuccess) return e;
  if ((e = upload(c, &c.ipiv, std::vector<int32_t>(), (size_t)std::max(poff, 1))) != hipSuccess) return e;
  if ((e = upload(c, &c.info, std::vector<int32_t>(), 2 * (size_t)n_seg)) != hipSuccess) return e;
  if ((e = upload(c, &c.inv32, std::vector<float>((size_t)std::max<int64_t>(ioff, 1), 0.f))) != hipSuccess) return e;
  if ((e = upload(c, &c.w, std::vector<float>((size_t)std::max(vlen, 1), 0.f))) != hipSuccess) return e;
  if ((e = upload(c, &c.y, std::vector<float>((size_t)std::max(vlen, 1), 0.f))) != hipSuccess) return e;
  if ((e = upload(c, &c.cpart, std::vector<double>((size_t)std::max(c.n_gemv, 1), 0.0))) != hipSuccess) return e;
  if ((e = upload(c, &c.s1pre, std::vector<double>(2 * (size_t)n_seg, 0.0))) != hipSuccess) return e;
  if ((e = upload(c, &c.seg_agg_d, seg_agg)) != hipSuccess) return e;
  {
    std::vector<int32_t> seg_half(seg_agg.size());
    for (size_t k = 0; k < seg_agg.size(); ++k) seg_half[k] = 2 * seg_agg[k];
    if ((e = upload(c, &c.seg_half_d, seg_half)) != hipSuccess) return e;
  }
  if ((e = upload(c, &c.w2, std::vector<float>((size_t)std::max(vlen, 1), 0.f))) != hipSuccess) return e;
  if ((e = upload(c, &c.xl, std::vector<float>(), 3 * (size_t)std::max(n_list, 1))) != hipSuccess) return e;
  if ((e = upload(c, &c.rl, std::vector<float>(), 3 * (size_t)std::max(n_list, 1))) != hipSuccess) return e;
  if ((e = upload(c, &c.ml, std::vector<float>(), 9 * (size_t)std::max(n_list, 1))) != hipSuccess) return e;
  if ((e = upload(c, &c.agg_rz, std::vector<double>(2 * (size_t)std::max(n_agg, 1), 0.0))) != hipSuccess) return e;
  if ((e = upload(c, &c.agg_rr, std::vector<double>(2 * (size_t)std::max(n_agg, 1), 0.0))) != hipSuccess) return e;
  c.built = true;
  c.active = false;
  return hipSuccess;
}

namespace {

// Pointers to one segment's dense matrices.
struct SegMats {
  int nc;
  double *Ac, *X, *Xref, *X2, *T;
  int32_t *ipiv, *info;
};

SegMats seg_mats(CoarseOp& c, int T) {
  const int64_t m = c.seg_moff[T];
  return SegMats{c.seg_nc[T], c.Ac + m, c.X + m, c.Xref + m, c.X2 + m, c.T + m,
                 c.ipiv + c.seg_poff[T], c.info + 2 * T};
}

// T = A_c X of one segment and the column maxima of |I - T| into its part of
// c.colmax (no synchronisation).
hipError_t residual_launch(CoarseOp& c, const SegMats& S, const double* X, rocblas_handle hb,
                           hipStream_t st, int32_t coff) {
  const int nc = S.nc;
  const double one = 1.0, zero = 0.0;
  if (rocblas_dgemm(hb, rocblas_operation_none, rocblas_operation_none, nc, nc, nc, &one, S.Ac,
                    nc, X, nc, &zero, S.T, nc) != rocblas_status_success)
    return hipErrorUnknown;
  hipLaunchKernelGGL(coarse_resid_kernel, dim3(nc), dim3(kBlock), 0, st, nc, S.T,
                     c.colmax + coff);
  return hipGetLastError();
}

// X_ref = A_c^-1 of every listed segment by LU (rocSOLVER), one
// synchronisation for all; ok[T] false when a segment's A_c is singular.
hipError_t refactor(CoarseOp& c, const std::vector<int>& segs, rocblas_handle hb, hipStream_t st,
                    std::vector<char>& ok) {
  hipError_t e;
  std::vector<char> launched(c.n_seg, 0);
  for (int T : segs) {
    const SegMats S = seg_mats(c, T);
    const int nc = S.nc;
    ok[T] = 0;
    if ((e = hipMemcpyAsync(S.Xref, S.Ac, sizeof(double) * (size_t)nc * nc,
                            hipMemcpyDeviceToDevice, st)) != hipSuccess)
      return e;
    launched[T] = rocsolver_dgetrf(hb, nc, nc, S.Xref, nc, S.ipiv, S.info) ==
                      rocblas_status_success &&
                  rocsolver_dgetri(hb, nc, S.Xref, nc, S.ipiv, S.info + 1) ==
                      rocblas_status_success;
  }
  std::vector<int32_t> info(2 * (size_t)c.n_seg, 1);
  if ((e = hipMemcpyAsync(info.data(), c.info, sizeof(int32_t) * info.size(),
                          hipMemcpyDeviceToHost, st)) != hipSuccess)
    return e;
  if ((e = hipStreamSynchronize(st)) != hipSuccess) return e;
  for (int T : segs) {
    ++c.refactors;
    ok[T] = launched[T] && info[2 * T] == 0 && info[2 * T + 1] == 0;
  }
  return hipSuccess;
}

// Largest residual max |I - A_c X| from which one Newton-Schulz step is taken
// without checking the result.
constexpr double kNsAccept = 1e-3;
// Largest probe estimate of |I - A_c X_ref| (largest row of the probe
// residual) at which X_ref is used as it is, with no Newton-Schulz step.
// X_ref is the exact inverse of an earlier A_c, so it is symmetric positive
// definite and the preconditioned CG stays valid; a coarse inverse within a
// few per cent of the current one changes the CG iteration count little.
constexpr double kProbeAccept = 2e-2;

// FP32 (and with want64 FP64) row-major copies of the segment's inverse Xa.
hipError_t accept_inverse(CoarseOp& c, int T, const double* Xa, hipStream_t st) {
  const int nc = c.seg_nc[T];
  hipLaunchKernelGGL(coarse_convert_kernel<float>, dim3((nc + 31) / 32, (nc + 31) / 32),
                     dim3(kBlock), 0, st, nc, c.seg_ld[T], Xa, c.inv32 + c.seg_ioff[T]);
  if (c.want64)
    hipLaunchKernelGGL(coarse_convert_kernel<double>, dim3((nc + 31) / 32, (nc + 31) / 32),
                       dim3(kBlock), 0, st, nc, c.seg_ld[T], Xa, c.inv64 + c.seg_ioff[T]);
  return hipGetLastError();
}

// Every listed segment that holds a reference inverse: W = A_c (X_ref V) for
// its probe block V and the largest |V - W| per probe column; one
// synchronisation for all.  probe[T] is the largest over the columns (1e300
// for a segment without a reference inverse).
hipError_t probe_segments(CoarseOp& c, const std::vector<int>& segs, rocblas_handle hb,
                          hipStream_t st, std::vector<double>& probe) {
  probe.assign(c.n_seg, 1e300);
  const double one = 1.0, zero = 0.0;
  bool any = false;
  for (int T : segs) {
    if (!c.seg_have_inv[T]) continue;
    const SegMats S = seg_mats(c, T);
    const int nc = S.nc;
    const size_t po = (size_t)kCoarseProbes * c.seg_poff[T];
    if (rocblas_dgemm(hb, rocblas_operation_none, rocblas_operation_none, nc, kCoarseProbes, nc,
                      &one, S.Xref, nc, c.probe_v + po, nc, &zero, c.probe_u + po, nc) !=
            rocblas_status_success ||
        rocblas_dgemm(hb, rocblas_operation_none, rocblas_operation_none, nc, kCoarseProbes, nc,
                      &one, S.Ac, nc, c.probe_u + po, nc, &zero, c.probe_w + po, nc) !=
            rocblas_status_success)
      return hipErrorUnknown;
    hipLaunchKernelGGL(coarse_probe_kernel, dim3(kCoarseProbes), dim3(kBlock), 0, st, nc,
                       c.probe_v + po, c.probe_w + po, c.probe_max + kCoarseProbes * T);
    any = true;
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || !any) return e;
  std::vector<double> pm((size_t)kCoarseProbes * c.n_seg);
  if ((e = hipMemcpyAsync(pm.data(), c.probe_max, sizeof(double) * pm.size(),
                          hipMemcpyDeviceToHost, st)) != hipSuccess)
    return e;
  if ((e = hipStreamSynchronize(st)) != hipSuccess) return e;
  for (int T : segs) {
    if (!c.seg_have_inv[T]) continue;
    double m = 0.0;
    for (int j = 0; j < kCoarseProbes; ++j) m = std::max(m, pm[(size_t)kCoarseProbes * T + j]);
    probe[T] = m;
  }
  return hipSuccess;
}

// The inverse of the A_c of every listed segment: Newton-Schulz steps from
// its X_ref, which is refactorised when they do not converge; ok[T] when the
// segment's inverse is usable.  The segments step together, with one
// synchronisation per round of residuals, and each segment's arithmetic is
// that of the segment alone.
hipError_t factor_segments(CoarseOp& c, const std::vector<int>& segs, rocblas_handle hb,
                           hipStream_t st, std::vector<char>& ok) {
  hipError_t e;
  for (int T : segs) {
    const int nc = c.seg_nc[T];
    ok[T] = 0;
    hipLaunchKernelGGL(coarse_diag_kernel, dim3((nc + kBlock - 1) / kBlock), dim3(kBlock), 0, st,
                       nc, c.Ac + c.seg_moff[T]);
  }
  if ((e = hipGetLastError()) != hipSuccess) return e;
  // a reference inverse still close to the inverse of the current A_c is
  // used as it is (no Newton-Schulz GEMM)
  {
    std::vector<double> probe;
    if ((e = probe_segments(c, segs, hb, st, probe)) != hipSuccess) return e;
    for (int T : segs)
      if (probe[T] < kProbeAccept) {
        if ((e = accept_inverse(c, T, c.Xref + c.seg_moff[T], st)) != hipSuccess) return e;
        ok[T] = 1;
        ++c.probe_accepts;
      }
  }
  // X from the reference inverse by Newton-Schulz steps against the current
  // A_c.  The result depends only on (A_c, X_ref), never on earlier solves,
  // so a repeated step repeats its bits; X_ref is refactorised (rocSOLVER)
  // only when the steps do not converge from it.
  std::vector<char> cur(c.n_seg, 0);   // 0: X holds the iterate, 1: X2
  std::vector<char> dead(c.n_seg, 0);  // singular A_c: the segment keeps block Jacobi
  for (int attempt = 0; attempt < 2; ++attempt) {
    std::vector<int> redo, run;
    for (int T : segs)
      if (!ok[T] && !dead[T] && (!c.seg_have_inv[T] || attempt == 1)) redo.push_back(T);
    if (!redo.empty()) {
      std::vector<char> fok(c.n_seg, 0);
      if ((e = refactor(c, redo, hb, st, fok)) != hipSuccess) return e;
      for (int T : redo) {
        c.seg_have_inv[T] = fok[T];
        dead[T] = !fok[T];
      }
    }
    for (int T : segs) {
      if (ok[T] || dead[T]) continue;
      const SegMats S = seg_mats(c, T);
      if ((e = hipMemcpyAsync(S.X, S.Xref, sizeof(double) * (size_t)S.nc * S.nc,
                              hipMemcpyDeviceToDevice, st)) != hipSuccess)
        return e;
      cur[T] = 0;
      run.push_back(T);
    }
    const double one = 1.0, zero = 0.0;
    for (int it = 0; !run.empty(); ++it) {
      for (int T : run) {
        const SegMats S = seg_mats(c, T);
        if ((e = residual_launch(c, S, cur[T] ? S.X2 : S.X, hb, st, c.seg_poff[T])) !=
            hipSuccess)
          return e;
      }
      std::vector<double> cm(c.colmax_len);
      if ((e = hipMemcpyAsync(cm.data(), c.colmax, sizeof(double) * cm.size(),
                              hipMemcpyDeviceToHost, st)) != hipSuccess)
        return e;
      if ((e = hipStreamSynchronize(st)) != hipSuccess) return e;
      std::vector<int> next;
      for (int T : run) {
        const SegMats S = seg_mats(c, T);
        const int nc = S.nc;
        double r = 0.0;
        for (int q = 0; q < nc; ++q) r = std::max(r, cm[c.seg_poff[T] + q]);
        double* X = cur[T] ? S.X2 : S.X;
        auto accept = [&](const double* Xa) {
          ok[T] = 1;
          return accept_inverse(c, T, Xa, st);
        };
        if (r < 1e-5) {
          if ((e = accept(X)) != hipSuccess) return e;
          continue;
        }
        if (!(r < 0.25) || it == 6) continue;  // too far from X_ref (or NaN)
        hipLaunchKernelGGL(coarse_twoI_kernel, dim3((nc + kBlock - 1) / kBlock, nc), dim3(kBlock),
                           0, st, nc, S.T);
        if (rocblas_dgemm(hb, rocblas_operation_none, rocblas_operation_none, nc, nc, nc, &one, X,
                          nc, S.T, nc, &zero, cur[T] ? S.X : S.X2, nc) != rocblas_status_success)
          return hipErrorUnknown;
        cur[T] ^= 1;
        // the step squares the residual, I - A_c X' = (I - A_c X)^2, so from
        // max |I - A_c X| = r every new entry is below nc r^2; from r below
        // kNsAccept the step is accepted without a residual product (a
        // preconditioner needs no more, and the step stays a function of
        // A_c and X_ref alone)
        if (r < kNsAccept) {
          if ((e = accept(cur[T] ? S.X2 : S.X)) != hipSuccess) return e;
          continue;
        }
        next.push_back(T);
      }
      run.swap(next);
    }
  }
  return hipSuccess;
}

}  // namespace

namespace {

template <typename V>
hipError_t coarse_factor_impl(CoarseOp& c, const int32_t* col, const V* val32, int64_t n_slots,
                              const uint8_t* dof_flags, const int32_t* run, hipStream_t st) {
  c.active = false;
  if (!c.built) return hipSuccess;
  if (!c.blas) {
    rocblas_handle hb = nullptr;
    if (rocblas_create_handle(&hb) != rocblas_status_success) return hipErrorUnknown;
    // deterministic GEMMs: no atomics in rocBLAS kernels
    if (rocblas_set_atomics_mode(hb, rocblas_atomics_not_allowed) != rocblas_status_success)
      return hipErrorUnknown;
    c.blas = hb;
  }
  rocblas_handle hb = (rocblas_handle)c.blas;
  if (rocblas_set_stream(hb, st) != rocblas_status_success) return hipErrorUnknown;
  // the segments' dense blocks lie in order: A_c ends with the last one
  const int last = c.n_seg - 1;
  const int64_t total = c.seg_moff[last] + (int64_t)c.seg_nc[last] * c.seg_nc[last];
  hipLaunchKernelGGL(coarse_pairs_kernel<V>, dim3(std::max(c.n_cpairs, 1)), dim3(kBlock), 0, st,
                     c.n_cpairs, c.cp_ptr, c.cp_slot, c.cp_row, col, val32, n_slots, c.node_d,
                     dof_flags, c.pairblk);
  hipLaunchKernelGGL(coarse_fill_kernel, dim3(grid_nodes(std::max<int64_t>(total, 1))),
                     dim3(kBlock), 0, st, c.Ac, total, 0.0);
  hipLaunchKernelGGL(coarse_scatter_kernel, dim3(std::max(c.n_cpairs, 1)), dim3(64), 0, st,
                     c.n_cpairs, c.cp_IK, c.agg_tissue, c.segd, c.segm, c.pairblk, c.Ac);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  // segments not iterating keep their coarse state
  std::vector<int> segs;
  for (int T = 0; T < c.n_seg; ++T)
    if (!run || run[T]) segs.push_back(T);
  std::vector<char> ok(c.n_seg, 0);
  if ((e = factor_segments(c, segs, hb, st, ok)) != hipSuccess) return e;
  for (int T : segs) c.seg_active[T] = ok[T];
  std::vector<int32_t> on(c.n_seg, 0);
  for (int T = 0; T < c.n_seg; ++T) {
    on[T] = c.seg_active[T] ? 1 : 0;
    c.active |= c.seg_active[T] != 0;
  }
  if ((e = hipMemcpyAsync(c.seg_on, on.data(), sizeof(int32_t) * c.n_seg, hipMemcpyHostToDevice,
                          st)) != hipSuccess)
    return e;
  hipLaunchKernelGGL(coarse_mask_kernel, dim3(grid_nodes(std::max(c.agg_total, 1))), dim3(kBlock),
                     0, st, c.agg_total, c.agg_nodes, dof_flags, c.list_mask);
  hipLaunchKernelGGL(coarse_nodecb_kernel, dim3(grid_nodes(std::max(c.n_nodes, 1))), dim3(kBlock),
                     0, st, c.n_nodes, c.node_agg, c.agg_cbase, dof_flags, c.node_cb);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  // the activity flags are pageable host memory: wait for their copy
  return hipStreamSynchronize(st);
}

}  // namespace

hipError_t coarse_factor(CoarseOp& c, const int32_t* col, const float* val32, int64_t n_slots,
                         const uint8_t* dof_flags, const int32_t* run, hipStream_t st) {
  c.want64 = 0;
  return coarse_factor_impl<float>(c, col, val32, n_slots, dof_flags, run, st);
}

hipError_t coarse_factor(CoarseOp& c, const int32_t* col, const double* val64, int64_t n_slots,
                         const uint8_t* dof_flags, const int32_t* run, hipStream_t st) {
  if (c.built && !c.inv64) {
    const int last = c.n_seg - 1;
    const int64_t n_inv = std::max<int64_t>(c.seg_ioff[last] + (int64_t)c.seg_nc[last] * c.seg_ld[last], 1);
    hipError_t e;
    if ((e = upload(c, &c.inv64, std::vector<double>((size_t)n_inv, 0.0))) != hipSuccess) return e;
    if ((e = upload(c, &c.w64, std::vector<double>((size_t)std::max(c.vlen, 1), 0.0))) != hipSuccess) return e;
    if ((e = upload(c, &c.y64, std::vector<double>((size_t)std::max(c.vlen, 1), 0.0))) != hipSuccess) return e;
  }
  c.want64 = 1;
  hipError_t e = coarse_factor_impl<double>(c, col, val64, n_slots, dof_flags, run, st);
  c.want64 = 0;
  return e;
}

void coarse_apply(const CoarseOp& c, const int32_t* state, const float* r, const S1Pre& pre,
                  hipStream_t st) {
  hipLaunchKernelGGL(coarse_restrict_kernel<float>, dim3(c.n_agg), dim3(kRBlock), 0, st, c.n_agg,
                     c.agg_ptr, c.agg_nodes, c.agg_tissue, c.agg_cbase, c.seg_on, c.list_d,
                     c.list_mask, r, state, c.w);
  launch_gemv32(c, state, nullptr, pre, st);
}

// After the FP32 update that formed w in two half-aggregate terms (w, w2).
void coarse_product(const CoarseOp& c, const int32_t* state, const S1Pre& pre, float* /*tag*/,
                    hipStream_t st) {
  launch_gemv32(c, state, c.w2, pre, st);
}

void coarse_product(const CoarseOp& c, const int32_t* state, const S1Pre& pre, double* /*tag*/,
                    hipStream_t st) {
  hipLaunchKernelGGL(coarse_gemv64_kernel, dim3(c.n_gemv), dim3(kBlock), 0, st, c.gw, c.sgi,
                     c.sgio, c.seg_on, c.inv64, c.w64, state, c.y64, c.cpart, pre);
}

void coarse_apply(const CoarseOp& c, const int32_t* state, const double* r, const S1Pre& pre,
                  hipStream_t st) {
  hipLaunchKernelGGL(coarse_restrict_kernel<double>, dim3(c.n_agg), dim3(kRBlock), 0, st, c.n_agg,
                     c.agg_ptr, c.agg_nodes, c.agg_tissue, c.agg_cbase, c.seg_on, c.list_d,
                     c.list_mask, r, state, c.w64);
  hipLaunchKernelGGL(coarse_gemv64_kernel, dim3(c.n_gemv), dim3(kBlock), 0, st, c.gw, c.sgi,
                     c.sgio, c.seg_on, c.inv64, c.w64, state, c.y64, c.cpart, pre);
}

}  // namespace tf
