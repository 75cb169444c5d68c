// Two-level preconditioner for the FP32 CG solve of the mixed-precision step:
//   M^-1 r = D^-1 r + P A_c^-1 P^T r,   A_c = P^T A P.
//
// No counterpart in the reference, which factorises K with Pardiso.  Block
// Jacobi (D^-1) alone leaves the smooth, tissue-wide deformation modes to the
// Krylov iteration, so the iteration count grows with the mesh (about 200 CG
// iterations per NR iteration at 1.6 M prisms).  The coarse space P holds,
// for every aggregate of nodes, the six rigid-body modes of the aggregate
// (three translations, three rotations about its centroid); an aggregate is
// the set of nodes whose reference (x, y) lies in one square cell, through
// all layers of the tissue.  Fixed and slave DOF have zero rows in P, so the
// correction never moves them.
//
// Per assembly: A_c is formed block by block from the FP32 operator (one
// workgroup per aggregate pair, a fixed summation order).  A reference
// inverse X_ref comes from rocSOLVER (getrf, getri) on the first assembly;
// the operator changes little from one assembly to the next, so every
// assembly starts from X = X_ref and takes Newton-Schulz steps
// X <- X (2I - A_c X) (two rocBLAS GEMMs each, quadratic convergence) until
// max|I - A_c X| < 1e-5 (a preconditioner needs no more).  When the steps do
// not converge (residual above 0.25, or seven steps), X_ref is refactorised
// from the current A_c.  The inverse is kept as a dense
// FP32 matrix.  Per CG iteration: w = P^T r (one workgroup per aggregate),
// y = A_c^-1 w (one wavefront per row), z += P y (one thread per node), and
// the partial products (w, y) that complete (r, z).  Every sum has a fixed
// order and rocBLAS runs with atomics off, so the solve stays bitwise
// reproducible.
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
constexpr int kMaxAggregates = 8192;
constexpr int kGemvRows = kBlock / 64;  // rows of A_c^-1 per product workgroup
constexpr int kSplit = 4;               // workgroups per aggregate in the restriction

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
    const bool off = flags && (flags[3 * i + p] & 3);
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      P[p][a] = off ? 0.0 : (p == a ? 1.0 : 0.0);
      P[p][3 + a] = off ? 0.0 : R[p][a];
    }
  }
}

// Workgroup sum of n values per thread in a fixed order (wave trees, then
// waves in order); the result is valid in thread 0.
template <int N>
__device__ inline void block_sum(double (&v)[N], double (*red)[kBlock / 64]) {
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
      for (int w = 0; w < kBlock / 64; ++w) s += red[k][w];
      v[k] = s;
    }
  }
}

// (P_I^T A P_K) for every aggregate pair (I, K), from the operator slots
// listed for the pair.
__global__ __launch_bounds__(kBlock) void coarse_pairs_kernel(
    int32_t n_pairs, const int32_t* __restrict__ cp_ptr, const int32_t* __restrict__ cp_slot,
    const int32_t* __restrict__ cp_row, const int32_t* __restrict__ col,
    const float* __restrict__ val, int64_t n_slots, const float* __restrict__ d,
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

// Pair blocks into the dense column-major A_c (each pair owns its 6x6 block).
__global__ void coarse_scatter_kernel(int32_t n_pairs, int32_t n_c, const int32_t* cp_IK,
                                      const double* blk, double* dense) {
  const int pr = blockIdx.x, q = threadIdx.x;
  if (pr >= n_pairs || q >= 36) return;
  const int I = cp_IK[2 * pr], K = cp_IK[2 * pr + 1];
  const int a = q / 6, b = q % 6;
  dense[(size_t)(6 * K + b) * n_c + 6 * I + a] = blk[36 * (size_t)pr + q];
}

// A mode with no free DOF (all its rows masked) has a zero row and column:
// 1 on its diagonal keeps A_c invertible and leaves its y at 0.
__global__ __launch_bounds__(kBlock) void coarse_diag_kernel(int32_t n_c, double* dense) {
  const int r = blockIdx.x * kBlock + threadIdx.x;
  if (r < n_c && dense[(size_t)r * n_c + r] == 0.0) dense[(size_t)r * n_c + r] = 1.0;
}

// FP32 row-major copy (row stride ld) of the column-major FP64 inverse,
// through 32 x 32 LDS tiles (grid: tiles of rows x tiles of columns).
__global__ __launch_bounds__(kBlock) void coarse_convert_kernel(int32_t n_c, int32_t ld,
                                                                const double* X, float* inv32) {
  __shared__ float tile[32][33];
  const int r0 = blockIdx.x * 32, c0 = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  for (int k = ty; k < 32; k += 8) {
    const int r = r0 + tx, cc = c0 + k;  // read down a column of X
    tile[k][tx] = (r < n_c && cc < n_c) ? (float)X[(size_t)cc * n_c + r] : 0.f;
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



// w = P^T r as kSplit partial vectors (w_s of stride ld, summed in order by
// the product kernel): kSplit workgroups per aggregate, each over a fixed
// part of the aggregate's node list; the offsets and masks are stored in list
// order, only r is gathered.
__global__ __launch_bounds__(kBlock) void coarse_restrict_kernel(
    int32_t n_agg, int32_t ld, const int32_t* __restrict__ agg_ptr,
    const int32_t* __restrict__ agg_nodes, const float* __restrict__ list_d,
    const uint8_t* __restrict__ list_mask, const float* __restrict__ r,
    const int32_t* __restrict__ state, float* __restrict__ w) {
  __shared__ double red[6][kBlock / 64];
  const int a = blockIdx.x / kSplit, part = blockIdx.x % kSplit;
  if (a >= n_agg || state[0] != K_RUN) return;
  double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  const int a0 = agg_ptr[a], len = agg_ptr[a + 1] - a0;
  const int beg = a0 + (int)((int64_t)len * part / kSplit);
  const int end = a0 + (int)((int64_t)len * (part + 1) / kSplit);
  for (int t = beg + threadIdx.x; t < end; t += kBlock) {
    const int i = agg_nodes[t];
    const uint8_t mk = list_mask[t];
    const double d0 = list_d[3 * t], d1 = list_d[3 * t + 1], d2 = list_d[3 * t + 2];
    const double r0 = (mk & 1) ? 0.0 : (double)r[3 * i];
    const double r1 = (mk & 2) ? 0.0 : (double)r[3 * i + 1];
    const double r2 = (mk & 4) ? 0.0 : (double)r[3 * i + 2];
    acc[0] += r0;
    acc[1] += r1;
    acc[2] += r2;
    // rotation rows of P^T: d x r
    acc[3] += d1 * r2 - d2 * r1;
    acc[4] += d2 * r0 - d0 * r2;
    acc[5] += d0 * r1 - d1 * r0;
  }
  block_sum<6>(acc, red);
  if (threadIdx.x == 0)
#pragma unroll
    for (int q = 0; q < 6; ++q) w[(size_t)part * ld + 6 * a + q] = (float)acc[q];
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
        if (flags[3 * i + p] & 3) m |= (uint8_t)(1 << p);
    }
    list_mask[t] = m;
  }
}

// y = A_c^-1 w, one wavefront per row (16-byte loads over the padded row
// stride ld); partial (w, y) per workgroup.
typedef float v4f __attribute__((ext_vector_type(4)));
__global__ __launch_bounds__(kBlock) void coarse_gemv_kernel(int32_t n_c, int32_t ld,
                                                             const float* __restrict__ inv32,
                                                             const float* __restrict__ w,
                                                             const int32_t* __restrict__ state,
                                                             float* __restrict__ y,
                                                             double* __restrict__ cpart) {
  __shared__ double wy[kGemvRows];
  if (state[0] != K_RUN) return;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int row = blockIdx.x * kGemvRows + wave;
  double dot = 0.0;
  if (row < n_c) {
    const v4f* a = (const v4f*)(inv32 + (size_t)row * ld);
    float s = 0.f;
    for (int j = lane; j < ld / 4; j += 64) {
      v4f v = ((const v4f*)w)[j];
#pragma unroll
      for (int p = 1; p < kSplit; ++p) v += ((const v4f*)(w + (size_t)p * ld))[j];
      const v4f u = a[j];
      s += u.x * v.x + u.y * v.y + u.z * v.z + u.w * v.w;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if (lane == 0) y[row] = s;
    float wr = w[row];
#pragma unroll
    for (int p = 1; p < kSplit; ++p) wr += w[(size_t)p * ld + row];
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

// z += P y.
__global__ __launch_bounds__(kBlock) void coarse_prolong_kernel(
    int32_t n_nodes, const int32_t* __restrict__ node_agg, const float* __restrict__ d,
    const uint8_t* __restrict__ flags, const float* __restrict__ y,
    const int32_t* __restrict__ state, float* __restrict__ z) {
  if (state[0] != K_RUN) return;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n_nodes; i += gridDim.x * kBlock) {
    const int a = node_agg[i];
    const float* ya = y + 6 * a;
    const float d0 = d[3 * i], d1 = d[3 * i + 1], d2 = d[3 * i + 2];
    // t + omega x d
    const float v[3] = {ya[0] + (ya[4] * d2 - ya[5] * d1), ya[1] + (ya[5] * d0 - ya[3] * d2),
                        ya[2] + (ya[3] * d1 - ya[4] * d0)};
#pragma unroll
    for (int p = 0; p < 3; ++p)
      if (!(flags && (flags[3 * i + p] & 3))) z[3 * i + p] += v[p];
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

hipError_t coarse_build(CoarseOp& c, int32_t n_nodes, const double* X, double cell,
                        int z_bands, const std::vector<int32_t>& slice_base,
                        const std::vector<int32_t>& col, const std::vector<int32_t>& row_len) {
  void* blas = c.blas;
  const int32_t gen = c.gen;
  c.blas = nullptr;
  coarse_release(c);
  c.blas = blas;
  c.gen = gen;
  if (n_nodes <= 0 || !(cell > 0.0) || z_bands < 1) return hipErrorInvalidValue;
  // square cells over the reference (x, y), each split into z_bands equal bands
  // of the reference z range
  double x0 = X[0], y0 = X[1], z0 = X[2], z1 = X[2];
  for (int i = 0; i < n_nodes; ++i) {
    x0 = std::min(x0, X[3 * i]);
    y0 = std::min(y0, X[3 * i + 1]);
    z0 = std::min(z0, X[3 * i + 2]);
    z1 = std::max(z1, X[3 * i + 2]);
  }
  const double zb = z1 > z0 ? (z1 - z0) / z_bands : 1.0;
  std::vector<int64_t> key(n_nodes);
  for (int i = 0; i < n_nodes; ++i) {
    const int64_t ix = (int64_t)std::floor((X[3 * i] - x0) / cell);
    const int64_t iy = (int64_t)std::floor((X[3 * i + 1] - y0) / cell);
    const int64_t iz = std::min<int64_t>(z_bands - 1, (int64_t)std::floor((X[3 * i + 2] - z0) / zb));
    key[i] = (ix << 40) | (iy << 16) | iz;
  }
  std::vector<int64_t> keys(key);
  std::sort(keys.begin(), keys.end());
  keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
  const int n_cells = (int)keys.size();
  std::vector<int32_t> cell_of(n_nodes), cnt(n_cells, 0);
  for (int i = 0; i < n_nodes; ++i) {
    cell_of[i] = (int32_t)(std::lower_bound(keys.begin(), keys.end(), key[i]) - keys.begin());
    ++cnt[cell_of[i]];
  }
  std::vector<double> cen(3 * (size_t)n_cells, 0.0);
  for (int i = 0; i < n_nodes; ++i)
    for (int k = 0; k < 3; ++k) cen[3 * cell_of[i] + k] += X[3 * i + k];
  for (int q = 0; q < n_cells; ++q)
    for (int k = 0; k < 3; ++k) cen[3 * q + k] /= cnt[q];
  // small cells join the nearest cell of kMinAggNodes nodes or more (all one
  // aggregate when there is none)
  std::vector<int32_t> target(n_cells), agg_of_cell(n_cells, -1);
  std::vector<int32_t> big;
  for (int q = 0; q < n_cells; ++q)
    if (cnt[q] >= kMinAggNodes) big.push_back(q);
  for (int q = 0; q < n_cells; ++q) {
    if (cnt[q] >= kMinAggNodes) {
      target[q] = q;
    } else if (big.empty()) {
      target[q] = 0;
    } else {
      double best = 1e300;
      int bq = big[0];
      for (int b : big) {
        const double dx = cen[3 * q] - cen[3 * b], dy = cen[3 * q + 1] - cen[3 * b + 1];
        const double dz = cen[3 * q + 2] - cen[3 * b + 2];
        const double d2 = dx * dx + dy * dy + dz * dz;
        if (d2 < best) { best = d2; bq = b; }
      }
      target[q] = bq;
    }
  }
  int n_agg = 0;
  for (int q = 0; q < n_cells; ++q)
    if (agg_of_cell[target[q]] < 0) agg_of_cell[target[q]] = n_agg++;
  if (n_agg > kMaxAggregates) return hipErrorInvalidValue;
  std::vector<int32_t> agg(n_nodes);
  for (int i = 0; i < n_nodes; ++i) agg[i] = agg_of_cell[target[cell_of[i]]];
  // centroid and radius per aggregate; scaled offsets per node
  std::vector<double> ac(3 * (size_t)n_agg, 0.0), rad(n_agg, 0.0);
  std::vector<int32_t> na(n_agg, 0);
  for (int i = 0; i < n_nodes; ++i) {
    ++na[agg[i]];
    for (int k = 0; k < 3; ++k) ac[3 * agg[i] + k] += X[3 * i + k];
  }
  for (int a = 0; a < n_agg; ++a)
    for (int k = 0; k < 3; ++k) ac[3 * a + k] /= na[a];
  for (int i = 0; i < n_nodes; ++i) {
    double s = 0.0;
    for (int k = 0; k < 3; ++k) {
      const double t = X[3 * i + k] - ac[3 * agg[i] + k];
      s += t * t;
    }
    rad[agg[i]] = std::max(rad[agg[i]], std::sqrt(s));
  }
  std::vector<float> dn(3 * (size_t)n_nodes);
  for (int i = 0; i < n_nodes; ++i) {
    const double L = rad[agg[i]] > 0.0 ? rad[agg[i]] : 1.0;
    for (int k = 0; k < 3; ++k) dn[3 * i + k] = (float)((X[3 * i + k] - ac[3 * agg[i] + k]) / L);
  }
  std::vector<int32_t> aptr(n_agg + 1, 0), anodes(n_nodes);
  for (int a = 0; a < n_agg; ++a) aptr[a + 1] = aptr[a] + na[a];
  {
    std::vector<int32_t> fill(aptr.begin(), aptr.end() - 1);
    for (int i = 0; i < n_nodes; ++i) anodes[fill[agg[i]]++] = i;
  }
  // operator slots grouped by aggregate pair, ascending within each pair
  std::vector<int32_t> pair_id((size_t)n_agg * n_agg, -1), pIK, pcnt;
  const int n_slices = (int)slice_base.size() - 1;
  auto for_slots = [&](auto f) {
    for (int sl = 0; sl < n_slices; ++sl) {
      const int w = (slice_base[sl + 1] - slice_base[sl]) / kSlice;
      for (int k = 0; k < w; ++k)
        for (int lane = 0; lane < kSlice; ++lane) {
          const int R = sl * kSlice + lane;
          if (R >= n_nodes || k >= row_len[R]) continue;
          const int slot = slice_base[sl] + k * kSlice + lane;
          f(slot, R, col[slot]);
        }
    }
  };
  int64_t listed = 0;
  for_slots([&](int slot, int R, int C) {
    int32_t& id = pair_id[(size_t)agg[R] * n_agg + agg[C]];
    if (id < 0) {
      id = (int32_t)pcnt.size();
      pcnt.push_back(0);
      pIK.push_back(agg[R]);
      pIK.push_back(agg[C]);
    }
    ++pcnt[id];
    ++listed;
  });
  const int n_pairs = (int)pcnt.size();
  std::vector<int32_t> pptr(n_pairs + 1, 0), pslot(listed), prow(listed);
  for (int p = 0; p < n_pairs; ++p) pptr[p + 1] = pptr[p] + pcnt[p];
  {
    std::vector<int32_t> fill(pptr.begin(), pptr.end() - 1);
    for_slots([&](int slot, int R, int C) {
      const int32_t id = pair_id[(size_t)agg[R] * n_agg + agg[C]];
      pslot[fill[id]] = slot;
      prow[fill[id]] = R;
      ++fill[id];
    });
  }
  // node offsets in list order
  std::vector<float> ldv(3 * (size_t)n_nodes);
  for (int t = 0; t < n_nodes; ++t)
    for (int k = 0; k < 3; ++k) ldv[3 * (size_t)t + k] = dn[3 * (size_t)anodes[t] + k];
  c.n_agg = n_agg;
  c.n_c = 6 * n_agg;
  c.n_cpairs = n_pairs;
  c.n_gemv = (c.n_c + kGemvRows - 1) / kGemvRows;
  c.ld32 = (c.n_c + 63) / 64 * 64;
  c.agg_total = n_nodes;
  hipError_t e;
  if ((e = upload(c, &c.node_agg, agg)) != hipSuccess) return e;
  if ((e = upload(c, &c.node_d, dn)) != hipSuccess) return e;
  if ((e = upload(c, &c.agg_ptr, aptr)) != hipSuccess) return e;
  if ((e = upload(c, &c.agg_nodes, anodes)) != hipSuccess) return e;
  if ((e = upload(c, &c.list_d, ldv)) != hipSuccess) return e;
  if ((e = upload(c, &c.list_mask, std::vector<uint8_t>((size_t)n_nodes, 0))) != hipSuccess) return e;
  if ((e = upload(c, &c.cp_ptr, pptr)) != hipSuccess) return e;
  if ((e = upload(c, &c.cp_slot, pslot)) != hipSuccess) return e;
  if ((e = upload(c, &c.cp_row, prow)) != hipSuccess) return e;
  if ((e = upload(c, &c.cp_IK, pIK)) != hipSuccess) return e;
  const size_t nc = (size_t)c.n_c, ld = (size_t)c.ld32;
  if ((e = upload(c, &c.pairblk, std::vector<double>(), 36 * (size_t)n_pairs)) != hipSuccess) return e;
  if ((e = upload(c, &c.Ac, std::vector<double>(), nc * nc)) != hipSuccess) return e;
  if ((e = upload(c, &c.X, std::vector<double>(), nc * nc)) != hipSuccess) return e;
  if ((e = upload(c, &c.Xref, std::vector<double>(), nc * nc)) != hipSuccess) return e;
  if ((e = upload(c, &c.X2, std::vector<double>(), nc * nc)) != hipSuccess) return e;
  if ((e = upload(c, &c.T, std::vector<double>(), nc * nc)) != hipSuccess) return e;
  if ((e = upload(c, &c.colmax, std::vector<double>(nc, 0.0))) != hipSuccess) return e;
  if ((e = upload(c, &c.ipiv, std::vector<int32_t>(), nc)) != hipSuccess) return e;
  if ((e = upload(c, &c.info, std::vector<int32_t>(), 2)) != hipSuccess) return e;
  if ((e = upload(c, &c.inv32, std::vector<float>(nc * ld, 0.f))) != hipSuccess) return e;
  if ((e = upload(c, &c.w, std::vector<float>(kSplit * ld, 0.f))) != hipSuccess) return e;
  if ((e = upload(c, &c.y, std::vector<float>(nc, 0.f))) != hipSuccess) return e;
  if ((e = upload(c, &c.cpart, std::vector<double>((size_t)c.n_gemv, 0.0))) != hipSuccess) return e;
  c.built = true;
  c.active = false;
  c.have_inv = false;
  return hipSuccess;
}

namespace {

// max |I - A_c X| on the host (one synchronisation).
hipError_t residual(CoarseOp& c, rocblas_handle hb, hipStream_t st, double* out) {
  const int nc = c.n_c;
  const int64_t n2 = (int64_t)nc * nc;
  const double one = 1.0, zero = 0.0;
  if (rocblas_dgemm(hb, rocblas_operation_none, rocblas_operation_none, nc, nc, nc, &one, c.Ac,
                    nc, c.X, nc, &zero, c.T, nc) != rocblas_status_success)
    return hipErrorUnknown;
  hipLaunchKernelGGL(coarse_resid_kernel, dim3(nc), dim3(kBlock), 0, st, nc, c.T, c.colmax);
  std::vector<double> cm(nc);
  hipError_t e;
  if ((e = hipMemcpyAsync(cm.data(), c.colmax, sizeof(double) * nc, hipMemcpyDeviceToHost,
                          st)) != hipSuccess)
    return e;
  if ((e = hipStreamSynchronize(st)) != hipSuccess) return e;
  double r = 0.0;
  for (double v : cm) r = std::max(r, v);
  *out = r;
  return hipSuccess;
}

// X_ref = A_c^-1 by LU (rocSOLVER); false when A_c is singular.
hipError_t refactor(CoarseOp& c, rocblas_handle hb, hipStream_t st, bool* ok) {
  const int nc = c.n_c;
  *ok = false;
  hipError_t e;
  if ((e = hipMemcpyAsync(c.Xref, c.Ac, sizeof(double) * (size_t)nc * nc,
                          hipMemcpyDeviceToDevice, st)) != hipSuccess)
    return e;
  if (rocsolver_dgetrf(hb, nc, nc, c.Xref, nc, c.ipiv, c.info) != rocblas_status_success ||
      rocsolver_dgetri(hb, nc, c.Xref, nc, c.ipiv, c.info + 1) != rocblas_status_success)
    return hipSuccess;
  int32_t info[2] = {1, 1};
  if ((e = hipMemcpyAsync(info, c.info, sizeof(info), hipMemcpyDeviceToHost, st)) != hipSuccess)
    return e;
  if ((e = hipStreamSynchronize(st)) != hipSuccess) return e;
  ++c.refactors;
  *ok = info[0] == 0 && info[1] == 0;
  return hipSuccess;
}

}  // namespace

hipError_t coarse_factor(CoarseOp& c, const int32_t* col, const float* val32, int64_t n_slots,
                         const uint8_t* dof_flags, hipStream_t st) {
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
  const int nc = c.n_c;
  const int64_t n2 = (int64_t)nc * nc;
  hipLaunchKernelGGL(coarse_pairs_kernel, dim3(c.n_cpairs), dim3(kBlock), 0, st, c.n_cpairs,
                     c.cp_ptr, c.cp_slot, c.cp_row, col, val32, n_slots, c.node_d, dof_flags,
                     c.pairblk);
  hipLaunchKernelGGL(coarse_fill_kernel, dim3(grid_nodes(n2)), dim3(kBlock), 0, st, c.Ac, n2,
                     0.0);
  hipLaunchKernelGGL(coarse_scatter_kernel, dim3(c.n_cpairs), dim3(64), 0, st, c.n_cpairs, nc,
                     c.cp_IK, c.pairblk, c.Ac);
  hipLaunchKernelGGL(coarse_diag_kernel, dim3((nc + kBlock - 1) / kBlock), dim3(kBlock), 0, st,
                     nc, c.Ac);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  // X from the reference inverse by Newton-Schulz steps against the current
  // A_c.  The result depends only on (A_c, X_ref), never on earlier solves,
  // so a repeated step repeats its bits; X_ref is refactorised (rocSOLVER)
  // only when the steps do not converge from it.
  const size_t bytes = sizeof(double) * (size_t)nc * nc;
  bool ok = false;
  for (int attempt = 0; attempt < 2 && !ok; ++attempt) {
    if (!c.have_inv || attempt == 1) {
      bool fok = false;
      if ((e = refactor(c, hb, st, &fok)) != hipSuccess) return e;
      c.have_inv = fok;
      if (!fok) return hipSuccess;  // singular: the solve keeps block Jacobi
    }
    if ((e = hipMemcpyAsync(c.X, c.Xref, bytes, hipMemcpyDeviceToDevice, st)) != hipSuccess)
      return e;
    const double one = 1.0, zero = 0.0;
    for (int it = 0;; ++it) {
      double r = 0.0;
      if ((e = residual(c, hb, st, &r)) != hipSuccess) return e;  // T = A_c X
      if (r < 1e-5) { ok = true; break; }
      if (!(r < 0.25) || it == 6) break;  // too far from X_ref (or NaN)
      hipLaunchKernelGGL(coarse_twoI_kernel, dim3((nc + kBlock - 1) / kBlock, nc), dim3(kBlock), 0,
                         st, nc, c.T);
      if (rocblas_dgemm(hb, rocblas_operation_none, rocblas_operation_none, nc, nc, nc, &one, c.X,
                        nc, c.T, nc, &zero, c.X2, nc) != rocblas_status_success)
        return hipErrorUnknown;
      std::swap(c.X, c.X2);
    }
  }
  if (!ok) return hipSuccess;
  hipLaunchKernelGGL(coarse_convert_kernel, dim3((nc + 31) / 32, (nc + 31) / 32), dim3(kBlock), 0,
                     st, nc, c.ld32, c.X, c.inv32);
  hipLaunchKernelGGL(coarse_mask_kernel, dim3(grid_nodes(c.agg_total)), dim3(kBlock), 0, st,
                     c.agg_total, c.agg_nodes, dof_flags, c.list_mask);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  c.active = true;
  return hipSuccess;
}

void coarse_apply(const CoarseOp& c, const int32_t* state, const float* r, float* z,
                  int32_t n_nodes, const uint8_t* dof_flags, hipStream_t st) {
  hipLaunchKernelGGL(coarse_restrict_kernel, dim3(c.n_agg * kSplit), dim3(kBlock), 0, st,
                     c.n_agg, c.ld32, c.agg_ptr, c.agg_nodes, c.list_d, c.list_mask, r, state,
                     c.w);
  hipLaunchKernelGGL(coarse_gemv_kernel, dim3(c.n_gemv), dim3(kBlock), 0, st, c.n_c, c.ld32,
                     c.inv32, c.w, state, c.y, c.cpart);
  hipLaunchKernelGGL(coarse_prolong_kernel, dim3(grid_nodes(n_nodes)), dim3(kBlock), 0, st,
                     n_nodes, c.node_agg, c.node_d, dof_flags, c.y, state, z);
}

}  // namespace tf
