// Internal declarations shared by the host driver (tf_nr.hip) and the device
// code (tf_element.hip, tf_assemble.hip, tf_tiled.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/tissue_nr.h"

namespace tf {

// Jacobian storage: SELL-64 of 3x3 node blocks.  Slice s covers block rows
// [64 s, 64 s + 64); every row of the slice is padded to the slice width w_s.
// Block slot of (row R, k-th stored column of R) is
//   slice_base[R / 64] + k * 64 + R % 64,
// values are planar: val[c * n_slots + slot], c = 3 p + q of the 3x3 block.
// Padding slots point at column R with zero values.
constexpr int kSlice = 64;

struct DevPattern {
  int32_t n_slices = 0;
  int64_t n_slots = 0;          // padded SELL size in blocks
  int32_t* slice_base = nullptr;  // n_slices + 1 (slot offsets)
  int32_t* col = nullptr;         // n_slots
  int32_t* diag_slot = nullptr;   // n_nodes
  int32_t* row_len = nullptr;     // n_nodes: stored blocks of each row
  int32_t* tslot = nullptr;       // n_slots: slot of the transposed block (C, R)
  int32_t* contrib_ptr = nullptr; // n_slots + 1
  int32_t* contrib_src = nullptr; // n_prisms * 36: e*36 + B*6 + A (K block A,B)
  // n_slices: the order in which the assembly visits slices (Morton order of
  // the slices' reference (x, y) centroids within each tissue), so that the
  // slices of one column of prisms, which gather from the same prisms' tangent
  // records, are assembled close together in time
  int32_t* slice_order = nullptr;
  int32_t* node_ptr = nullptr;    // n_nodes + 1
  int32_t* node_src = nullptr;    // n_prisms * 6: e*6 + a
  // contact slots of node-node packing (null when the handle does not pack):
  // slice s stores its rows' element blocks in columns [0, slice_ew[s]) and
  // slice_cap[s] contact columns after them; the e-th couple of row R (in
  // ascending partner order) sits in column slice_ew[s] + e, unused contact
  // slots point at column R with zero values
  int32_t* slice_ew = nullptr;    // n_slices
  int32_t* slice_cap = nullptr;   // n_slices
};

// The operator of a bound system, A_b = N^T J N + I_slave (slave/master
// binding, NewtonRaphsonSolver::calculateBoundKWithSlavesMasterDoF,
// NewtonRaphsonSolver.cpp:128-183), stored explicitly on its own SELL
// pattern.  Slot s of A_b copies J slot blockmap[s] when blockmap[s] >= 0;
// otherwise exception ex = -blockmap[s] - 1 lists, per entry q (0..8), the J
// values to sum (planar indices q' * n_slots_J + slot_J, in ascending DOF
// order) in ex_src[ex_ptr[9 ex + q] .. ex_ptr[9 ex + q + 1]), plus 1 where
// ex_one[9 ex + q] (a slave's diagonal).
struct BoundOp {
  bool ready = false;
  DevPattern pat;  // slice_base, col, diag_slot, row_len, tslot, n_slices, n_slots
  int32_t* blockmap = nullptr;
  int32_t* ex_ptr = nullptr;
  int32_t* ex_src = nullptr;
  uint8_t* ex_one = nullptr;
  int32_t n_ex = 0;
  double* val = nullptr;   // 9 * n_slots planar
  float* val32 = nullptr;
  double* minv = nullptr;  // block-Jacobi inverse of A_b
  float* minv32 = nullptr;
  // the pairs grouped by master: master bm_master[i] has the slaves
  // bm_slave[bm_ptr[i] .. bm_ptr[i+1]) in the order of the pair list
  int32_t* bm_ptr = nullptr;
  int32_t* bm_master = nullptr;
  int32_t* bm_slave = nullptr;
  int32_t n_masters = 0;
  // host copy of the pattern (for the coarse space's block lists)
  std::vector<int32_t> h_slice_base, h_col, h_row_len;
};

// Two-level preconditioner of the FP32 CG solve (tf_coarse.hip):
//   M^-1 r = D^-1 r + P A_c^-1 P^T r,   A_c = P^T A P,
// with D the block diagonal of the operator A, and P the rigid-body modes
// (three translations, three rotations about the centroid) of node
// aggregates: the nodes whose reference (x, y) falls in one square cell,
// through all layers.  Fixed and slave DOF have zero rows in P.  A_c is
// formed on the device from the FP32 operator at every assembly; its FP64
// inverse comes from Newton-Schulz steps X <- X (2I - A_c X) (rocBLAS GEMMs)
// started from a reference inverse (rocSOLVER getrf / getri, refactorised
// only when the steps do not converge); it is applied as a dense FP32
// product.
// Probe columns of the check of a reference coarse inverse against the
// current coarse operator.
constexpr int kCoarseProbes = 4;
struct CoarseOp {
  bool built = false;    // aggregates and block lists match the operator pattern
  bool for_bound = false;
  bool active = false;   // the factorisation of some segment's operator succeeded
  // One segment per tissue: its own aggregates, its own dense A_c block, inverse
  // and Newton-Schulz state.  A single-tissue handle has one segment.
  int32_t n_seg = 0;
  int32_t n_agg = 0, n_c = 0, n_cpairs = 0, n_gemv = 0, agg_total = 0;
  int32_t vlen = 0;             // padded coarse vector length (sum of the segments' ld)
  std::vector<int32_t> seg_agg;   // n_seg + 1: aggregates of each segment
  std::vector<int32_t> seg_nc;    // coarse DOF per segment (6 per aggregate)
  std::vector<int32_t> seg_ld;    // row stride of the segment's inv32 (n_c rounded up to 64)
  std::vector<int32_t> seg_coff;  // offset of the segment in the coarse vectors
  std::vector<int64_t> seg_moff;  // offset of the segment's dense FP64 matrices
  std::vector<int64_t> seg_ioff;  // offset of the segment's FP32 inverse
  std::vector<int32_t> seg_goff;  // n_seg + 1: product workgroups of each segment
  std::vector<int32_t> seg_poff;  // offset of the segment's pivots
  std::vector<char> seg_have_inv;  // Xref holds an inverse of the segment
  std::vector<char> seg_active;    // the segment's last factorisation succeeded
  int32_t n_nodes = 0;          // nodes covered by node_agg / node_cb
  int32_t* node_agg = nullptr;  // per node (-1: no aggregate)
  int32_t* node_cb = nullptr;   // per node: agg_cbase << 3 | coarse-excluded DOF bits (-1: none)
  float* node_d = nullptr;      // per node: (X - centroid) / radius, 3 floats
  int32_t* agg_ptr = nullptr;   // nodes of each aggregate, ascending
  int32_t* agg_nodes = nullptr;
  int32_t* agg_tissue = nullptr;  // per aggregate: its segment (tissue)
  int32_t* agg_cbase = nullptr;   // per aggregate: its first coarse DOF in the coarse vectors
  int32_t* cp_ptr = nullptr;    // operator slots of each aggregate pair, ascending
  int32_t* cp_slot = nullptr;
  int32_t* cp_row = nullptr;    // row (node) of each listed slot
  int32_t* cp_IK = nullptr;     // the pair's aggregates (I, K)
  int32_t* gw = nullptr;        // per product workgroup: segment, first row
  int32_t* cpart_off = nullptr; // n_seg + 1: product workgroups of each segment (device)
  int32_t* seg_on = nullptr;    // per segment: 1 when its inverse is in use
  int32_t* segd = nullptr;      // per segment: first aggregate, n_c (device, for the scatter)
  int64_t* segm = nullptr;      // per segment: dense offset (device)
  int32_t* sgi = nullptr;       // per segment: n_c, ld, vector offset (device, for the product)
  int64_t* sgio = nullptr;      // per segment: FP32 inverse offset (device)
  float* list_d = nullptr;      // node offsets in aggregate-list order (3 per entry)
  uint8_t* list_mask = nullptr; // fixed / slave bits of the listed node (bit p: DOF p)
  double* pairblk = nullptr;    // 36 per pair: (P_I^T A P_K), row-major 6x6
  double* Ac = nullptr;         // per segment n_c x n_c column-major
  double* Xref = nullptr;       // reference inverse (column-major, rocSOLVER)
  double* X = nullptr;          // inverse of the current A_c, and two work matrices
  double* X2 = nullptr;
  double* T = nullptr;
  double* colmax = nullptr;     // per column of every segment: max |I - A_c X|
  int32_t colmax_len = 0;
  int32_t refactors = 0;        // rocSOLVER factorisations so far
  int32_t probe_accepts = 0;    // setups that kept X_ref on its probe residual
  double* probe_v = nullptr;    // per segment kCoarseProbes columns of +-1 (column-major)
  double* probe_u = nullptr;    // X_ref V
  double* probe_w = nullptr;    // A_c X_ref V
  double* probe_max = nullptr;  // kCoarseProbes per segment: max |V - W| per column
  int32_t* ipiv = nullptr;
  int32_t* info = nullptr;      // 2 per segment: getrf, getri
  float* inv32 = nullptr;       // per segment A_c^-1, row-major FP32
  float* w = nullptr;           // P^T r (vlen entries, zero padding)
  float* w2 = nullptr;          // the FP32 update's second half-aggregate terms of P^T r
  // the FP32 CG's x, r (3 per list position) and block-Jacobi inverse (9), in
  // the order of agg_nodes, which the update streams
  float* xl = nullptr;
  float* rl = nullptr;
  float* ml = nullptr;
  float* y = nullptr;           // A_c^-1 P^T r
  // the same in FP64, for the FP64 CG (allocated by its first factorisation)
  double* inv64 = nullptr;
  double* w64 = nullptr;
  double* y64 = nullptr;
  int want64 = 0;               // the factorisation in progress also writes inv64
  double* cpart = nullptr;      // per workgroup of the product: partial (w, y)
  double* s1pre = nullptr;      // 2 per segment: ordered sums of the (r, z_D), (r, r) partials
  int32_t* seg_agg_d = nullptr; // n_seg + 1: first aggregate of each segment (device)
  int32_t* seg_half_d = nullptr; // n_seg + 1: twice seg_agg_d (first half-aggregate)
  double* agg_rz = nullptr;     // per half-aggregate: partial (r, z_D) of the FP32 update
  double* agg_rr = nullptr;     // per half-aggregate: partial (r, r)
  std::vector<void*> allocs;
  void* blas = nullptr;         // rocblas_handle
  int32_t gen = 0;              // incremented by every build (captured graphs key on it)
};
// One segment of a coarse build: the nodes [begin, end) of one tissue and the
// cell size of its aggregates.
struct CoarseSeg {
  int32_t begin, end;
  double cell;
};
// A_b from the assembled J (FP64, or FP32 with mixed) and its block-Jacobi
// inverse.
hipError_t launch_bound_gather(const BoundOp& b, int32_t n_nodes, int64_t n_slots_J,
                               const double* valJ, const float* valJ32, int mixed,
                               hipStream_t st);

// Tiles of SELL slices for the fused solver: tile t covers slices
// [tile_slice[t], tile_slice[t+1]) of tissue tile_tissue[t]; tissue T owns
// tiles [tissue_tile[T], tissue_tile[T+1]).  grid = n_tiles rounded up to a
// multiple of 8 (the XCD count).
struct TileMap {
  int32_t n_tiles = 0, grid = 0, n_tissues = 0, slices_per_tile = 0;
  const int32_t* tile_slice = nullptr;
  const int32_t* tile_tissue = nullptr;
  const int32_t* tissue_tile = nullptr;
};

constexpr int kMaxTissues = 4096;
struct TiledResult {
  int32_t state[kMaxTissues];
  int32_t iters[kMaxTissues];
};

struct Workspace {
  // per-prism outputs (block-major, reference K orientation):
  // elemK[(e*36 + B*6 + A)*9 + r*3 + c] = TriPointKe[3A+r][3B+c]: the six
  // blocks that row B of J gathers (J(R, C) = K(C, R)^T) are contiguous, and
  // those of the three nodes of one face fill one half of the prism's record
  double* elemKe = nullptr;
  double* elemKv = nullptr;
  float* elemKe32 = nullptr;
  float* elemKv32 = nullptr;
  double* elemFe = nullptr;  // n_prisms * 18, unmasked -sum_gp w g
  double* elemFv = nullptr;
  double* strain = nullptr;  // n_prisms * 6, last Gauss point of the last assembly
  // assembled operator
  double* val = nullptr;    // 9 * n_slots
  float* val32 = nullptr;   // 9 * n_slots (mixed precision)
  double* minv = nullptr;   // 9 * n_nodes block-Jacobi inverse
  float* minv32 = nullptr;
  // vectors (3 * n_nodes)
  double *un = nullptr, *uk = nullptr, *g = nullptr, *du = nullptr, *internal = nullptr;
  double *r = nullptr, *rh = nullptr, *p = nullptr, *v = nullptr, *ph = nullptr, *s = nullptr,
         *sh = nullptr, *t = nullptr, *x = nullptr, *b = nullptr, *tmp = nullptr, *w = nullptr;
  float *r32 = nullptr, *rh32 = nullptr, *p32 = nullptr, *v32 = nullptr, *ph32 = nullptr,
        *s32 = nullptr, *sh32 = nullptr, *t32 = nullptr, *x32 = nullptr, *w32 = nullptr;
  // per-tissue state of the tiled solver
  int32_t* tstate = nullptr;   // 2 per tissue
  double* tscal = nullptr;     // 8 per tissue
  double* tpart = nullptr;     // 5 x tpart_stride
  int tpart_stride = 0;
  int32_t* tactive = nullptr;  // per tissue
  double* tnorm = nullptr;     // per tissue
  int32_t* host_tstate = nullptr;
  double* host_tnorm = nullptr;
  double* gnorm = nullptr;       // per tissue: |g|_2 of the Newton residual
  double* host_gnorm = nullptr;
  // warm start of a step's first solve: the guess x0, its residual r0, the
  // per-tissue rtol factors (device and pinned host), the displacements of
  // the last two steps (hist0 the latest), allocated on first use
  double *x0 = nullptr, *r0 = nullptr, *rscale = nullptr, *host_rscale = nullptr;
  double *hist0 = nullptr, *hist1 = nullptr;
  int32_t* guess_use = nullptr;  // per tissue: the guess is used
  int32_t* hist_order = nullptr; // per tissue: steps of history in use (0, 1, 2)
  // restarted GMRES (FP64): basis of kGmresRestart + 1 vectors, per-tile
  // partials, per-tissue sums and host-computed coefficients
  double* gV = nullptr;
  double* gpart = nullptr;   // (kGmresRestart + 2) x tpart_stride
  double* gsum = nullptr;    // 2 (kGmresRestart + 2) x tissues
  double* gfac = nullptr;    // per tissue
  double* gy = nullptr;      // kGmresRestart x tissues
  int32_t* gact = nullptr;   // per tissue
  int32_t* gk = nullptr;     // per tissue
  int gpart_stride = 0;
  // captured iteration chunks of the tiled solver (hipGraph), see tf_tiled.hip
  void* graphs = nullptr;
  int graphs_on = 1;
};

constexpr int kBlock = 256;
// Per-DOF flags (dof_flags): bit 0 fixed, bit 1 slave of a binding pair, bit 2
// a DOF of a laser-ablated node.  The coarse space of the two-level CG has
// zero rows for all three, so its correction never moves them: an ablated node
// without live prisms has a residual that does not depend on its position, and
// only an exact solve of its identity rows keeps it in place.
constexpr uint8_t kDofFixed = 1, kDofSlave = 2, kDofAblated = 4;
constexpr uint8_t kCoarseExcluded = kDofFixed | kDofSlave | kDofAblated;
constexpr int kGmresRestart = 100;

enum KState { K_RUN = 0, K_CONVERGED = 1, K_BREAKDOWN = 2 };
// A Krylov residual above kDivergence |b| ends the solve as a breakdown.
constexpr double kDivergence = 1e8;

// Work item of workgroup bid in a grid that is a multiple of 8.  The eight
// XCDs take workgroups round-robin; this map gives each XCD a contiguous
// eighth of the items, so neighbouring items (slices and tiles that share
// prisms and gathered vector entries) meet in one XCD's L2.
__device__ inline int xcd_item(int bid, int grid) { return (bid & 7) * (grid >> 3) + (bid >> 3); }

// A grid of workgroups for n items, a multiple of 8 below the launch limit.
inline int xcd_grid(int64_t n) {
  const int64_t g = (n + 7) / 8 * 8;
  return (int)(g < 8 ? 8 : (g > 262136 ? 262136 : g));
}

// ---- per-tissue scalars of the tiled solvers ----
// Ordered sum of part[beg..end), identical in every thread of the workgroup.
// B is the workgroup size.
template <int B = kBlock>
__device__ inline double wg_ordered_sum(const double* part, int beg, int end) {
  constexpr int kWaves = B / 64;
  __shared__ double red[kWaves];
  __shared__ double result;
  double s = 0.0;
  for (int i = beg + (int)threadIdx.x; i < end; i += B) s += part[i];
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

// N ordered sums at once, each bit for bit the sum wg_ordered_sum gives, with
// one load round and one chain of barriers for all of them.
template <int N>
__device__ inline void wg_ordered_sums(const double* const (&part)[N], const int (&beg)[N],
                                       const int (&end)[N], double (&out)[N]) {
  constexpr int kWaves = kBlock / 64;
  __shared__ double red[N][kWaves];
  __shared__ double result[N];
  double s[N];
#pragma unroll
  for (int q = 0; q < N; ++q) {
    s[q] = 0.0;
    for (int i = beg[q] + (int)threadIdx.x; i < end[q]; i += kBlock) s[q] += part[q][i];
  }
#pragma unroll
  for (int q = 0; q < N; ++q)
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s[q] += __shfl_xor(s[q], off, 64);
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  __syncthreads();
  if (lane == 0)
#pragma unroll
    for (int q = 0; q < N; ++q) red[q][wave] = s[q];
  __syncthreads();
  if (threadIdx.x < N) {
    double t = 0.0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) t += red[threadIdx.x][w];
    result[threadIdx.x] = t;
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < N; ++q) out[q] = result[q];
}

// Slots of the per-tissue scalars (8 per tissue) of the tiled solvers.  A
// slot is written only by the first workgroup of its tissue, and only in a
// launch whose workgroups do not read it.  CG: TS_RHO holds rho of the last
// iteration (read by the start-of-iteration step), TS_RHON the current rho
// (written by that step, read by the update, which then moves it to TS_RHO).
enum { TS_RHO = 0, TS_ALPHA, TS_OMEGA, TS_BNORM, TS_RNORM, TS_BETA, TS_RHON };

// What the preconditioned CG's start-of-iteration step reads and writes for
// one tissue: the tiles' partials (r, z) and (r, r), optionally the coarse
// product's partials (w, y), the scalars and the state.
struct CgS1 {
  const double* rz = nullptr;
  const double* rr = nullptr;
  const int32_t* tissue_tile = nullptr;
  const double* cpart = nullptr;   // null: no coarse partials
  const int32_t* cpart_off = nullptr;
  // with the coarse partials: the ordered sums of rz and rr per tissue, formed
  // by the coarse product (two per tissue); null: sum the tile partials here
  const double* pre = nullptr;
  const double* rtol_scale = nullptr;  // per tissue factor on rtol (null: 1)
  double* scal = nullptr;
  int32_t* state = nullptr;
  double rtol = 0.0;
};

// What the coarse product needs to form the ordered sums of the (r, z_D) and
// (r, r) tile partials for the next CG scalar step (out: two per tissue).
struct S1Pre {
  const double* rz = nullptr;
  const double* rr = nullptr;
  // the partials of tissue T are rz/rr[part_off[T] .. part_off[T+1]): one per
  // tile, or one per aggregate after the update that forms the restriction
  const int32_t* part_off = nullptr;
  const int32_t* cpart_off = nullptr;  // first product workgroup of each segment
  double* out = nullptr;               // null: nothing to form
};

// rho = (r, z) [+ (w, y)], the convergence test on |r| (against |b| = |r| of
// the first step, before any update), beta = rho / rho_old.  Every thread of
// the workgroup calls it and gets beta and the state; with `write`, thread 0
// stores |b| (first step), |r|, beta, rho (TS_RHON) and a state that ends the
// solve.
__device__ inline int cg_s1(const CgS1& s, int T, bool write, double* beta_out) {
  const int t0 = s.tissue_tile[T], t1 = s.tissue_tile[T + 1];
  double rho, rr;
  if (s.cpart && s.pre) {
    rho = s.pre[2 * T] + wg_ordered_sum(s.cpart, s.cpart_off[T], s.cpart_off[T + 1]);
    rr = s.pre[2 * T + 1];
  } else if (s.cpart) {
    const double* const parts[3] = {s.rz, s.cpart, s.rr};
    const int beg[3] = {t0, s.cpart_off[T], t0}, end[3] = {t1, s.cpart_off[T + 1], t1};
    double sums[3];
    wg_ordered_sums<3>(parts, beg, end, sums);
    rho = sums[0] + sums[1];
    rr = sums[2];
  } else {
    const double* const parts[2] = {s.rz, s.rr};
    const int beg[2] = {t0, t0}, end[2] = {t1, t1};
    double sums[2];
    wg_ordered_sums<2>(parts, beg, end, sums);
    rho = sums[0];
    rr = sums[1];
  }
  double* sc = s.scal + 8 * T;
  const bool first = s.state[2 * T + 1] == 0;
  const double rnorm = sqrt(rr);
  const double bnorm = first ? rnorm : sc[TS_BNORM];
  const double beta = rho / sc[TS_RHO];
  int state = K_RUN;
  if (!isfinite(rnorm) || rnorm > kDivergence * bnorm) state = K_BREAKDOWN;
  else if (rnorm <= (s.rtol_scale ? s.rtol * s.rtol_scale[T] : s.rtol) * bnorm) state = K_CONVERGED;
  else if (!(rho > 0.0) || !isfinite(beta)) state = K_BREAKDOWN;
  if (write && threadIdx.x == 0) {
    if (first) sc[TS_BNORM] = bnorm;
    sc[TS_RNORM] = rnorm;
    sc[TS_BETA] = beta;
    sc[TS_RHON] = rho;
    if (state != K_RUN) s.state[2 * T] = state;
  }
  *beta_out = beta;
  return state;
}

struct Launch {
  hipStream_t stream;
};

// ---- per-prism kernel (tf_element.hip) ----
struct ElementArgs {
  int32_t n_prisms;
  const int32_t* conn;
  const double* refpos;
  const double* invFg;
  const double* lambda;
  const double* mu;
  const double* eta;
  const double* uk;
  const double* un;
  double dt;
  int any_visc;
  double* Ke;
  double* Kv;
  float* Ke32;
  float* Kv32;
  double* Fe;
  double* Fv;
  // optional: the Green-Lagrange strain of the last Gauss point, 6 per prism
  // in Voigt order (E00, E11, E22, 2E01, 2E21, 2E02); with ptissue set, only
  // prisms whose tissue is flagged in tactive are written
  double* strain = nullptr;
  const int32_t* ptissue = nullptr;
  const int32_t* tactive = nullptr;
};
// mode 0: FP64 tangents; 1: FP32 elastic and viscous tangents apart;
// 2: FP32 elastic + viscous summed per K block into Ke32, and the elastic +
// viscous forces summed into Fe (Fv not written)
hipError_t launch_element(const ElementArgs& a, int mode, hipStream_t st);

// ---- assembly (tf_assemble.hip) ----
struct AssembleArgs {
  int32_t n_nodes;
  DevPattern pat;
  const double* Ke;
  const double* Kv;
  const float* Ke32;
  const float* Kv32;
  const double* Fe;
  const double* Fv;
  int any_visc;
  int combined;  // mixed: Ke32 holds elastic + viscous blocks and Fe the summed forces (element mode 2)
  const double* ext_visc;
  const double* area;
  const uint8_t* dof_flags;  // bit0 fixed, bit1 slave
  // laser-ablated nodes (Simulation::AblatedNodes, nullptr: none): a zero
  // diagonal Jacobian entry becomes 1 (checkJacobianForAblatedNodes)
  const uint8_t* abl_node = nullptr;
  const double* uk;
  const double* un;
  double dt;
  double* val;
  float* val32;
  double* minv;
  float* minv32;
  double* g;
  double* internal;
  // node-node packing (pk_cnt == nullptr: none): couples per node, partners
  // (pk_cap per node, ascending), node masses, parameters (pk_par[0]: the
  // packing threshold), and the per-node packing force of the residual
  const int32_t* pk_cnt = nullptr;
  const int32_t* pk_inc = nullptr;
  int pk_cap = 0;
  const double* pk_mass = nullptr;
  const double* pk_par = nullptr;
  double* pk_force = nullptr;
};
hipError_t launch_assemble(const AssembleArgs& a, int mixed, hipStream_t st);
// Zero outputs (tangents in the element mode's buffers, and forces) of the
// laser-ablated prisms `list` (n of them) after the per-prism kernel: an
// ablated prism contributes no forces and zero tangents
// (NewtonRaphsonSolver.cpp:236, 344; ShapeBase.cpp:2850-2852, 2879-2882).
hipError_t launch_ablate_zero(const ElementArgs& a, int mode, const int32_t* list, int n,
                              hipStream_t st);

// ---- node-node packing (tf_packing.hip) ----
// One couple id0 < id1 of Simulation::calculatePackingForcesImplicit3D
// (Simulation.cpp:7338-7400) and calculatePackingJacobian3D
// (Simulation.cpp:7444-7490): the force F on id0 (-F on id1) and the
// symmetric D = dF0/dx0 = m [f/d (I - n n^T) + f' n n^T], with the constants
// of Simulation::detectPacingNodes (multiplier 5000, sigmoid saturation 5,
// Simulation.cpp:8018-8020), the sigmoid centred at 0.6 threshold and the
// distance clamped to 1e-3, in the reference's operation order.
__device__ inline void pack_couple(const double* p0, const double* p1, double m0, double m1,
                                   double thr, double F[3], double D[9]) {
#pragma clang fp contract(off)
  const double multiplier = 5000.0, sat = 5.0;
  const double shift = 0.6 * thr;
  const double dx = p0[0] - p1[0];
  const double dy = p0[1] - p1[1];
  const double dz = p0[2] - p1[2];
  double d = sqrt(dx * dx + dy * dy + dz * dz);
  const double avg_mass = 0.5 * (m0 + m1);
  double n[3] = {1.0, 0.0, 0.0};
  if (d > 1e-3) {
    n[0] = dx / d;
    n[1] = dy / d;
    n[2] = dz / d;
  } else {
    d = 1e-3;
  }
  const double shifted = d - shift;
  const double sig = 1 / (1 + exp(sat / thr * 2.0 * shifted));
  const double fmag = multiplier * avg_mass * sig;
  F[0] = fmag * n[0];
  F[1] = fmag * n[1];
  F[2] = fmag * n[2];
  const double dsig = (-sat / thr * 2.0) * sig * (1 - sig);
  const double scale = multiplier * avg_mass;
  const double sd = sig / d;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const double nn = n[i] * n[j];
      double v = (i == j ? 1.0 : 0.0) - nn;
      v = v * sd;
      v = v + nn * dsig;
      D[3 * i + j] = v * scale;
    }
}

// At most this many couples per node (detection reports more as an error).
constexpr int kPackLocal = 32;

struct PackDev {
  bool on = false;
  int32_t cap = 0;             // contact slots per row of the slices holding packable nodes
  int symmetric_y = 0;
  int32_t n_cand = 0;          // nodes that may pack
  int32_t hash_size = 0;       // power of two
  int32_t* place = nullptr;    // per node: 0 basal, 1 apical, -1 never packs
  double* vol = nullptr;       // per prism: reference volume
  double* mass = nullptr;      // per node (Node::mass)
  int32_t* cnt = nullptr;      // per node: couples of the last detection
  int32_t* inc = nullptr;      // per node: cap partners, ascending
  double* par = nullptr;       // [0] packing threshold, [1] largest detection threshold,
                               // [2..51] squared grid thresholds, [52..57] bounding box
  double* force = nullptr;     // per node: packing force of the last residual
  int32_t* head = nullptr;     // hash buckets
  int32_t* next = nullptr;     // per node: next node of its bucket
  int32_t* cell = nullptr;     // per prism: threshold-grid cell (-1: not columnar)
  double* eval = nullptr;      // per prism: its grid value, apical and basal contributions
  int32_t* ecnt = nullptr;     // per prism: bit 0 apical and bit 1 basal value counted
  double* part = nullptr;      // per chunk partial sums of the thresholds
  int32_t* flag = nullptr;     // [0] largest couple count of the last detection
  int32_t n_chunks = 0;
  bool have_mass = false, detected = false;
  int32_t n_pairs = 0;         // couples of the last detection
  int32_t max_count = 0;
};
// Reference volumes (Prism::calculateReferenceVolume, Prism.cpp:888-922).
hipError_t launch_ref_volume(int32_t n_prisms, const double* refpos, double* vol, hipStream_t st);
// Node::mass from F_g (Simulation::updateNodeMasses, Simulation.cpp:7587-7596).
hipError_t launch_node_mass(int32_t n_nodes, const int32_t* node_ptr, const int32_t* node_src,
                            const double* Fg, const double* vol, double* mass, hipStream_t st);
// Thresholds and couples at positions u (Simulation::detectPacingNodes).
hipError_t launch_packing_detect(PackDev& p, int32_t n_nodes, int32_t n_prisms,
                                 const int32_t* conn, const double* u, const int32_t* elem_place,
                                 const int32_t* elem_type, const uint8_t* spans,
                                 const DevPattern& pat, hipStream_t st);
// Column and transpose-slot entries of the contact slots for the last detection.
hipError_t launch_contact_slots(const PackDev& p, int32_t n_nodes, const DevPattern& pat,
                                hipStream_t st);
hipError_t launch_bind_forces(const BoundOp& b, double* g, hipStream_t st);
hipError_t launch_bind_increment(const int32_t* pairs, int n_pairs, double* du, hipStream_t st);
hipError_t launch_visc_surface(int32_t n_nodes, int32_t n_prisms, const int32_t* conn,
                               const double* u, const int32_t* node_ptr,
                               const int32_t* node_src, const int32_t* placement,
                               const int32_t* type, const uint8_t* spans, double* area,
                               hipStream_t st,
                               const uint8_t* ablated = nullptr);

hipError_t launch_grow(int32_t n_prisms, const double* rates, double dt, double* Fg,
                       double* invFg, hipStream_t st,
                       const uint8_t* ablated = nullptr);
hipError_t launch_inverse(int32_t n_prisms, const double* Fg, double* invFg, hipStream_t st);

// ---- growth producers (tf_growth.hip) ----
// Per-prism data of the growth functions (ShapeBase members): tissue type,
// placement, flags, growth weights of linker prisms (columnar, peripodial),
// relative z position in the tissue height (0 apical .. 1 basal).
constexpr uint8_t kGrowEcm = 1, kGrowEcmCirc = 2, kGrowActin = 4;
constexpr int kApplyColumnar = 1, kApplyPeripodial = 2, kApplyBasalEcm = 4, kApplyLateralEcm = 8;
struct GrowthPrisms {
  const int32_t* type = nullptr;
  const int32_t* place = nullptr;
  const uint8_t* flags = nullptr;
  const double* weights = nullptr;  // 2 per prism
  const double* rel_z = nullptr;
  int explicit_ecm = 0;
};
// One growth function: type 1 (uniform: rates and the rotation matrix of its
// angle) or 3 (grid of nx x ny points: its compatible corner data start at
// cell cell_off of the handle's grid arrays); active for t0 <= t < t1, for the
// tissue parts of `apply` and the relative z range [zmin, zmax].
struct GrowthFn {
  int type = 0, apply = 0, nx = 0, ny = 0;
  int64_t cell_off = 0;
  double t0 = 0, t1 = 0, zmin = 0, zmax = 1;
  double rate[3] = {0, 0, 0};
  double rot[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
};
void grid_compatible(int nx, int ny, const double* rates, const double* angles_deg, double* cg,
                     double* ca, uint8_t* ce);
// Relative (x, y) positions of the prisms in their tissue's bounding box at
// positions u (internal node numbering; tissue T owns nodes
// [node_rng[2T], node_rng[2T+1]) and the prisms with prism_tissue == T).
hipError_t launch_relative_positions(int32_t n_tissues, const int32_t* node_rng, int32_t n_prisms,
                                     const int32_t* conn, const int32_t* prism_tissue,
                                     const double* u, const uint8_t* excluded, int symmetric_y,
                                     const int32_t* apical_of, double* bb, double* raw,
                                     double* rel, hipStream_t st);
// Buffers of the column volume conservation (on = 1): per prism the rotated
// growth and shape-change increments, the growth rotation and whether it
// applies, the grown and the reference volume.
struct GrowthStage {
  int on = 0;
  double *G = nullptr, *SC = nullptr, *gsr = nullptr, *vol = nullptr;
  const double* vref = nullptr;
  uint8_t* rot = nullptr;
};
// The growth of one step at time t (Simulation::calculateGrowth and the
// growShapeByFg loop of Simulation::runOneStep); columns[n_columns][layers]
// (prism ids, -1 for none) when gs.on.
hipError_t launch_growth(int32_t n_prisms, const int32_t* conn, const double* refpos,
                         const double* u, const GrowthPrisms& gp, const GrowthFn* fns, int n_fn,
                         const double* cg, const double* ca, const uint8_t* ce, const double* relpos,
                         double t, double dt, int interp, const uint8_t* ablated, double* Fg,
                         double* invFg, const GrowthStage& gs, int32_t n_columns, int32_t layers,
                         const int32_t* columns, hipStream_t st);

// ---- Krylov (tf_tiled.hip) ----
// Optional event timing of the SpMV launches ([0] FP64 operator, [1] FP32).
constexpr int kProfPairs = 128;
// Event timing of the SpMV launches.  kind = 2 * which + (FP32 ? 1 : 0),
// which 0: the SpMV that forms v = J ph (with s, sh), 1: t = J sh (with dots).
struct Prof {
  int on = 0;
  hipEvent_t ev[2 * kProfPairs] = {};
  int kind[kProfPairs] = {};
  int used = 0;
  double ms[4] = {0.0, 0.0, 0.0, 0.0};
  int64_t count[4] = {0, 0, 0, 0};
};
void prof_harvest(Prof* p);

struct KrylovArgs {
  int32_t n_nodes;
  DevPattern pat;
  const double* val;
  const float* val32;
  const double* minv;
  const float* minv32;
  const int32_t* pairs;
  int32_t n_pairs;
  Workspace* ws;
  Launch L;
  Prof* prof;
  TileMap tm;
  const CoarseOp* coarse = nullptr;  // two-level FP32 CG preconditioner when active
  const uint8_t* dof_flags = nullptr;
  // per-tissue factor on rtol (device, and its host copy for host-side
  // tests): a solve from a warm start targets rtol |b_T| of the original
  // right-hand side while it is handed the smaller residual of the guess
  const double* rtol_scale = nullptr;
  const double* rtol_scale_h = nullptr;
};
// Solves J x = b (FP64).  Returns the iteration count in *iters, the Krylov
// state (K_CONVERGED/K_BREAKDOWN/K_RUN=cap reached) in *state.
hipError_t krylov_bicgstab(const KrylovArgs& k, const double* b, double* x, double rtol,
                           int maxit, int* iters, int* state, int sync_every);
// Tiled fused BiCGStab on all tissues of the handle (tissues whose tactive
// entry is 0 get x = 0); per-tissue state and iteration counts in *res.
hipError_t tiled_bicgstab(const KrylovArgs& k, const double* b, double* x, double rtol, int maxit,
                          int sync_every, TiledResult* res);
// Mixed precision on all tissues: per-tissue scaled FP32 inner solves to
// inner_rtol (at most maxit iterations in all), and FP64 residual refinement
// while rtol < inner_rtol.  A tissue whose FP32 solve breaks down, diverges
// or reaches maxit ends with that state.
// Krylov methods of the tiled solvers.
enum KrylovMethod { KRYLOV_BICGSTAB = 0, KRYLOV_CG = 1, KRYLOV_GMRES = 2 };
hipError_t tiled_mixed(const KrylovArgs& k, const double* b, double* x, double rtol,
                       double inner_rtol, int maxit, int sync_every, int method,
                       TiledResult* res);
// Preconditioned CG (FP64) on all active tissues; a non-positive curvature
// or rho ends a tissue's solve as a breakdown.
hipError_t tiled_cg(const KrylovArgs& k, const double* b, double* x, double rtol, int maxit,
                    int sync_every, TiledResult* res);
// Restarted GMRES(kGmresRestart) with right block-Jacobi preconditioning on
// the FP64 operator, all active tissues (workspace from the gV.. members).
hipError_t tiled_gmres(const KrylovArgs& k, const double* b, double* x, double rtol, int maxit,
                       TiledResult* res);
// x_k += dx and per-tissue |dx|_2 into norms (host).
hipError_t tiled_norm_update(const KrylovArgs& k, const double* dx, double* xk, double* norms);
// Per-tissue |g|_2 into ws->host_gnorm (asynchronous; read after a sync).
hipError_t tiled_residual_norms(const KrylovArgs& k, const double* g);
hipError_t tiled_time_iterations(const KrylovArgs& k, int mixed, int method, int iters,
                                 float* ms);
// Plain copy of n4 16-byte vectors (bandwidth reference); kCopyVariants
// shapes (vectors per thread, ordinary or non-temporal accesses).
constexpr int kCopyVariants = 6;
hipError_t launch_copy(const void* src, void* dst, int64_t n4, int variant, hipStream_t st);
// Releases the captured graphs of a workspace.
void tiled_release(Workspace* ws);
// dst = src on the tissues whose tactive entry is 1 (others untouched).
hipError_t tiled_copy_active(const KrylovArgs& k, const double* src, double* dst);
// Warm start of a solve of J x = b: r0 = b - J x0 with the FP32 operator
// (mixed) or the FP64 one, per tissue.  A tissue whose guess leaves at least
// half of |b_T| (or a non-finite residual) takes x0 = 0 and r0 = b instead.
// ws->gnorm must hold |b_T|^2 (tiled_residual_norms).  scale_h / scale_d get
// |b_T| / |r0_T| (1 where the guess is not used): the factor on rtol that
// keeps the solve's target at rtol |b_T|.  Synchronises the stream.
hipError_t tiled_warm_start(const KrylovArgs& k, int mixed, const double* b, double* x0,
                            double* r0, double* scale_d, double* scale_h);
// dst += src on the tissues whose tactive entry is 1.
hipError_t tiled_add_active(const KrylovArgs& k, const double* src, double* dst);
// Per tissue T: x0 = 0 (order[T] 0), h0 (1) or 2 h0 - h1 (2); order on the device.
hipError_t tiled_extrapolate(const KrylovArgs& k, const int32_t* order, const double* h0,
                             const double* h1, double* x0);
// d = a - b, n entries.
hipError_t launch_difference(int64_t n, const double* a, const double* b, double* d,
                             hipStream_t st);

// ---- adjoint (tf_adjoint.hip) ----
struct AdjointLaunch {
  int32_t n_prisms, n_nodes;
  const int32_t* conn;
  const double* refpos;
  const double* invFg;
  const double* dinvFg;
  const double* lambda;
  const double* mu;
  const double* eta;
  const double* ext_visc;
  const uint8_t* dof_flags;
  const double* uk;
  const double* un;
  double dt;
  const double* m;
  const int32_t* placement;
  const int32_t* type;
  const uint8_t* spans;
  const double* area;
  const int32_t* node_ptr;
  const int32_t* node_src;
  double* s_out;
  double* b_out;
  double* a_prev;
};
hipError_t launch_adjoint_element(const AdjointLaunch& l, hipStream_t st);
hipError_t launch_transpose_operator(int64_t n_slots, const int32_t* tslot, const double* val,
                                     double* valT, int32_t n_nodes, const double* minv,
                                     double* minvT, hipStream_t st);

// ---- two-level preconditioner (tf_coarse.hip) ----
// Aggregates and block lists for an operator pattern (host arrays of the
// pattern: slice_base, col, row_len) and the node reference coordinates X
// (3 per node), one segment (tissue) at a time: its nodes fall in square
// (x, y) cells of the segment's cell size (same units as X), each split into
// z_bands bands of the segment's reference thickness.  Aggregates never cross
// segments, so A_c is block-diagonal with one dense block per segment.
hipError_t coarse_build(CoarseOp& c, int32_t n_nodes, const double* X,
                        const std::vector<CoarseSeg>& segs, int z_bands,
                        const std::vector<int32_t>& slice_base, const std::vector<int32_t>& col,
                        const std::vector<int32_t>& row_len);
void coarse_release(CoarseOp& c);
// A_c of the FP32 operator (col / val32 of the pattern the lists were built
// for), its FP64 inverse and the FP32 copy, for every segment T with run[T]
// set (all segments when run is null; host array); c.active reports that
// some segment's inverse is in use.
hipError_t coarse_factor(CoarseOp& c, const int32_t* col, const float* val32, int64_t n_slots,
                         const uint8_t* dof_flags, const int32_t* run, hipStream_t st);
// The same from the FP64 operator, keeping also an FP64 copy of the inverse
// (for the FP64 CG).
hipError_t coarse_factor(CoarseOp& c, const int32_t* col, const double* val64, int64_t n_slots,
                         const uint8_t* dof_flags, const int32_t* run, hipStream_t st);
// y = A_c^-1 P^T r (FP32) in c.y, with the partial (w, y) per product
// workgroup in c.cpart; nothing for a tissue whose state (state[2 T]) is not
// running, and zero partials for a segment whose factorisation failed.  The
// correction P y is added where the search direction is formed.
void coarse_apply(const CoarseOp& c, const int32_t* state, const float* r, const S1Pre& pre,
                  hipStream_t st);
// FP64: y64 = A_c^-1 P^T r with the FP64 inverse, w64 and y64 in FP64.
void coarse_apply(const CoarseOp& c, const int32_t* state, const double* r, const S1Pre& pre,
                  hipStream_t st);
// y = A_c^-1 w and the (w, y) partials only, for a w that the CG update formed
// (the FP32 or the FP64 product, chosen by the tag's type).
void coarse_product(const CoarseOp& c, const int32_t* state, const S1Pre& pre, float* tag,
                    hipStream_t st);
void coarse_product(const CoarseOp& c, const int32_t* state, const S1Pre& pre, double* tag,
                    hipStream_t st);

void set_error(const std::string& s);

}  // namespace tf
