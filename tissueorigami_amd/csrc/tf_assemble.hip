// Deterministic slot-map assembly of the system Jacobian and residual.
//
// The reference scatters every element's 18x18 blocks into a dense K, all
// viscous blocks first and then all elastic ones, each pass in element order
// (NewtonRaphsonSolver::writeImplicitElementalKToJacobian,
// NewtonRaphsonSolver.cpp:350-360, ShapeBase.cpp:2935-2983), then adds the
// external-viscosity diagonal (NewtonRaphsonSolver.cpp:399-423) and clears
// fixed rows and columns (calcutateFixedK, NewtonRaphsonSolver.cpp:195-219).
// Here one thread owns one 3x3 block (slot) of J = K^T (the operator of the
// transposed Pardiso solve) and gathers its contributions through a
// precomputed list sorted by element id, so the sum of every entry is formed
// in exactly the reference's order and the result is bitwise reproducible.
// The same thread applies the external viscosity and the fixed-DOF masks and,
// for diagonal blocks, writes the block-Jacobi inverse.  After its slice's
// blocks, the workgroup forms the residual of the slice's 64 nodes (one
// component per wavefront), summing each node's element forces in element
// order.
#include "tf_internal.h"

#pragma clang fp contract(off)

namespace tf {
namespace {

__device__ inline void inv3_dev(const double A[9], double X[9]) {
  // partial-pivot LU, then L U X = P I
  double LU[9];
  int perm[3] = {0, 1, 2};
#pragma unroll
  for (int q = 0; q < 9; ++q) LU[q] = A[q];
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    int piv = j;
    double mx = fabs(LU[3 * j + j]);
#pragma unroll
    for (int i = j + 1; i < 3; ++i)
      if (fabs(LU[3 * i + j]) > mx) { mx = fabs(LU[3 * i + j]); piv = i; }
    if (piv != j) {
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        double t = LU[3 * j + k]; LU[3 * j + k] = LU[3 * piv + k]; LU[3 * piv + k] = t;
      }
      int t = perm[j]; perm[j] = perm[piv]; perm[piv] = t;
    }
    if (LU[3 * j + j] != 0.0) {
#pragma unroll
      for (int i = j + 1; i < 3; ++i) {
        LU[3 * i + j] /= LU[3 * j + j];
#pragma unroll
        for (int k = j + 1; k < 3; ++k) LU[3 * i + k] -= LU[3 * i + j] * LU[3 * j + k];
      }
    }
  }
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    double y[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      double s = (perm[i] == c) ? 1.0 : 0.0;
#pragma unroll
      for (int k = 0; k < i; ++k) s -= LU[3 * i + k] * y[k];
      y[i] = s;
    }
#pragma unroll
    for (int i = 2; i >= 0; --i) {
      double s = y[i];
#pragma unroll
      for (int k = i + 1; k < 3; ++k) s -= LU[3 * i + k] * X[3 * k + c];
      X[3 * i + c] = s / LU[3 * i + i];
    }
  }
}

template <typename KT>
__device__ inline void gather_blocks(const KT* __restrict__ Kel, const int32_t* __restrict__ src,
                                     int beg, int end, double J[9]) {
  for (int t = beg; t < end; ++t) {
    const KT* b = Kel + (size_t)src[t] * 9;
    double k[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) k[q] = (double)b[q];
    // J(R, C)[p][q] = K(C, R)[q][p]
#pragma unroll
    for (int p = 0; p < 3; ++p)
#pragma unroll
      for (int q = 0; q < 3; ++q) J[3 * p + q] = J[3 * p + q] + k[3 * q + p];
  }
}

// Node residual, component i of node R: gSum = ge + gvInternal + gvExternal +
// gExt (packing), fixed DOF zeroed (NewtonRaphsonSolver.cpp:362-397, 441-457,
// 209); element forces are summed in element order (ShapeBase.cpp:2362-2376).
// internal receives SystemForces (elastic + viscous per element, non-fixed
// DOF only).
__device__ inline void node_force(const AssembleArgs& a, int R, int i) {
  const double rdt = 1.0 / a.dt;
  const int beg = a.pat.node_ptr[R], end = a.pat.node_ptr[R + 1];
  const bool fx = a.dof_flags ? (a.dof_flags[3 * R + i] & 1) : false;
  double ge = 0.0, gv = 0.0, sf = 0.0;
#pragma unroll 4
  for (int t = beg; t < end; ++t) {
    const int s = a.pat.node_src[t];  // e*6 + a
    // combined: Fe holds elastic + viscous, and gv stays 0
    const double fe_raw = a.Fe[3 * (size_t)s + i];
    const double fv_raw = a.combined ? 0.0 : a.Fv[3 * (size_t)s + i];
    const double fe = fx ? 0.0 : fe_raw;
    const double fv = fx ? 0.0 : fv_raw;
    ge = ge + fe;
    gv = gv + fv;
    if (!fx) sf = sf + (fe_raw + fv_raw);
  }
  // PackingForces of R, couple by couple (Simulation.cpp:7396-7401)
  double fp = 0.0;
  if (a.pk_cnt) {
    const int nc = min(a.pk_cnt[R], a.pk_cap);
    for (int e = 0; e < nc; ++e) {
      const int o = a.pk_inc[(size_t)R * a.pk_cap + e];
      const int lo = min(R, o), hi = max(R, o);
      double F[3], D[9];
      pack_couple(a.uk + 3 * lo, a.uk + 3 * hi, a.pk_mass[lo], a.pk_mass[hi], a.pk_par[0], F, D);
      fp = (R == lo) ? fp + F[i] : fp - F[i];
    }
    if (a.pk_force) a.pk_force[3 * R + i] = fp;
  }
  const double sv = a.area[R] * a.ext_visc[3 * R + i];
  const double disp = (a.uk[3 * R + i] - a.un[3 * R + i]) * rdt;
  const double gve = -(sv * disp);
  double g = ge + gv + gve;
  // gExt = PackingForces + 0 (Simulation.cpp:8130-8151), added by
  // NewtonRaphsonSolver::addExernalForces
  g = g + (fp + 0.0);
  if (fx) g = 0.0;
  a.g[3 * R + i] = g;
  if (a.internal) a.internal[3 * R + i] = sf;
}

template <bool VISC, bool MIXED>
__global__ __launch_bounds__(kBlock) void assemble_kernel(AssembleArgs a) {
  const int64_t n_slots = a.pat.n_slots;
  // one thread per slot: a workgroup takes a slice, its four wavefronts take
  // every fourth column k of the slice, lane = row; consecutive lanes touch
  // consecutive slots (coalesced planar stores) and all slots of a slice are
  // gathered concurrently
  constexpr int kCols = kBlock / kSlice;
  const int lane = threadIdx.x % kSlice, kq = threadIdx.x / kSlice;
  for (int it = xcd_item(blockIdx.x, gridDim.x); it < a.pat.n_slices; it += gridDim.x) {
    const int sl = a.pat.slice_order ? a.pat.slice_order[it] : it;
    const int R = sl * kSlice + lane;
    const int64_t sbase = a.pat.slice_base[sl];
    const int width = (int)((a.pat.slice_base[sl + 1] - sbase) / kSlice);
    const int row_len = (R < a.n_nodes) ? a.pat.row_len[R] : 0;
    for (int kk = kq; kk < width; kk += kCols) {
    const int64_t slot = sbase + (int64_t)kk * kSlice + lane;
    if (kk >= row_len) {
      // padding slot (zero block), or the contact slot of a packing couple
      double J[9];
#pragma unroll
      for (int q = 0; q < 9; ++q) J[q] = 0.0;
      if (a.pk_cnt && R < a.n_nodes) {
        const int e = kk - a.pat.slice_ew[sl];
        if (e >= 0 && e < min(a.pk_cnt[R], a.pk_cap)) {
          const int C = a.pk_inc[(size_t)R * a.pk_cap + e];
          const int lo = min(R, C), hi = max(R, C);
          double F[3], D[9];
          pack_couple(a.uk + 3 * lo, a.uk + 3 * hi, a.pk_mass[lo], a.pk_mass[hi], a.pk_par[0],
                      F, D);
          // K(id0, id1) and K(id1, id0) receive dF0/dx0 on a zero entry
          // (Simulation.cpp:7557-7567); J(R, C) = K(C, R)^T
#pragma unroll
          for (int p = 0; p < 3; ++p)
#pragma unroll
            for (int q = 0; q < 3; ++q) J[3 * p + q] = D[3 * q + p] + 0.0;
          if (a.dof_flags) {
#pragma unroll
            for (int p = 0; p < 3; ++p)
#pragma unroll
              for (int q = 0; q < 3; ++q)
                if ((a.dof_flags[3 * R + p] | a.dof_flags[3 * C + q]) & 1) J[3 * p + q] = 0.0;
          }
        }
      }
      if (MIXED) {
#pragma unroll
        for (int q = 0; q < 9; ++q) a.val32[(size_t)q * n_slots + slot] = (float)J[q];
      }
      if (a.val) {
#pragma unroll
        for (int q = 0; q < 9; ++q) a.val[(size_t)q * n_slots + slot] = J[q];
      }
      continue;
    }
    const int C = a.pat.col[slot];
    double J[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) J[q] = 0.0;
    const int beg = a.pat.contrib_ptr[slot], end = a.pat.contrib_ptr[slot + 1];
    if (MIXED) {
      if (VISC) gather_blocks(a.Kv32, a.pat.contrib_src, beg, end, J);
      gather_blocks(a.Ke32, a.pat.contrib_src, beg, end, J);
    } else {
      if (VISC) gather_blocks(a.Kv, a.pat.contrib_src, beg, end, J);
      gather_blocks(a.Ke, a.pat.contrib_src, beg, end, J);
    }
    {
      if (R == C) {
        // external viscous drag (NewtonRaphsonSolver.cpp:414-420)
        const double spdt = a.area[R] / a.dt;
#pragma unroll
        for (int j = 0; j < 3; ++j) J[4 * j] = spdt * a.ext_visc[3 * R + j] + J[4 * j];
        if (a.pk_cnt) {
          // -dF/dx of every couple of R, in couple order (ascending partner:
          // the couples where R is id1 come first), Simulation.cpp:7557-7567
          const int nc = min(a.pk_cnt[R], a.pk_cap);
          for (int e = 0; e < nc; ++e) {
            const int o = a.pk_inc[(size_t)R * a.pk_cap + e];
            const int lo = min(R, o), hi = max(R, o);
            double F[3], D[9];
            pack_couple(a.uk + 3 * lo, a.uk + 3 * hi, a.pk_mass[lo], a.pk_mass[hi], a.pk_par[0],
                        F, D);
#pragma unroll
            for (int p = 0; p < 3; ++p)
#pragma unroll
              for (int q = 0; q < 3; ++q) J[3 * p + q] = -D[3 * q + p] + J[3 * p + q];
          }
        }
        // checkJacobianForAblatedNodes (NewtonRaphsonSolver.cpp:425-439), after the
        // drag and packing terms and before the fixed-DOF pass (Simulation.cpp:6691)
        if (a.abl_node && a.abl_node[R]) {
#pragma unroll
          for (int j = 0; j < 3; ++j)
            if (J[4 * j] == 0.0) J[4 * j] = 1.0;
        }
      }
      if (a.dof_flags) {
#pragma unroll
        for (int p = 0; p < 3; ++p)
#pragma unroll
          for (int q = 0; q < 3; ++q) {
            const bool fr = a.dof_flags[3 * R + p] & 1, fc = a.dof_flags[3 * C + q] & 1;
            if (fr || fc) J[3 * p + q] = (R == C && p == q) ? 1.0 : 0.0;
          }
      }
      if (R == C) {
        double M[9], X[9];
#pragma unroll
        for (int q = 0; q < 9; ++q) M[q] = J[q];
        if (a.dof_flags) {
#pragma unroll
          for (int p = 0; p < 3; ++p)
#pragma unroll
            for (int q = 0; q < 3; ++q)
              if ((a.dof_flags[3 * R + p] | a.dof_flags[3 * R + q]) & 2)
                M[3 * p + q] = (p == q) ? 1.0 : 0.0;
        }
        inv3_dev(M, X);
        if (MIXED) {
#pragma unroll
          for (int q = 0; q < 9; ++q) a.minv32[9 * R + q] = (float)X[q];
        }
#pragma unroll
        for (int q = 0; q < 9; ++q) a.minv[9 * R + q] = X[q];
      }
    }
    if (MIXED) {
#pragma unroll
      for (int q = 0; q < 9; ++q) a.val32[(size_t)q * n_slots + slot] = (float)J[q];
    }
    if (a.val) {
#pragma unroll
      for (int q = 0; q < 9; ++q) a.val[(size_t)q * n_slots + slot] = J[q];
    }
    }
    // the residual of the slice's nodes: three wavefronts, one component each
    if (kq < 3 && R < a.n_nodes) node_force(a, R, kq);
  }
}

// Slave/master binding of the residual, pair by pair in list order
// (NewtonRaphsonSolver.cpp:148-158).  One thread: the list is short and the
// reference order is sequential.
// One thread per master: its slaves' forces are added to it one by one in
// the order of the pair list, as the reference's serial loop adds them, and
// cleared.  A slave is never a master and belongs to one master, so no two
// threads touch the same entry.
__global__ void bind_forces_kernel(const int32_t* ptr, const int32_t* master,
                                   const int32_t* slave, int n_masters, double* g) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_masters) return;
  const int m = master[i];
  double masterForce = g[m];
  for (int k = ptr[i]; k < ptr[i + 1]; ++k) {
    const int s = slave[k];
    masterForce += g[s];
    g[s] = 0.0;
  }
  g[m] = masterForce;
}

// equateSlaveDisplacementsToMasters (NewtonRaphsonSolver.cpp:185-193), one
// thread per pair (masters are never slaves, so no entry is read after it is
// written).
__global__ void bind_increment_kernel(const int32_t* pairs, int n_pairs, double* du) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < n_pairs) du[pairs[2 * p]] = du[pairs[2 * p + 1]];
}

// Exposed-surface area of one triangle of a prism (Prism::calculateApicalArea /
// calculateBasalArea, Prism.cpp:1485-1556).
__device__ double tri_area(const double* u, int n0, int n1, int n2) {
  const double Threshold = 1E-5;
  double s1[3], s2[3], Side1 = 0, Side2 = 0, costet = 0, Area = 0;
  for (int i = 0; i < 3; ++i) {
    s1[i] = u[3 * n1 + i] - u[3 * n0 + i];
    s2[i] = u[3 * n2 + i] - u[3 * n0 + i];
    costet += s1[i] * s2[i];
    Side1 += s1[i] * s1[i];
    Side2 += s2[i] * s2[i];
  }
  if (Side1 > Threshold && Side2 > Threshold) {
    Side1 = sqrt(Side1);
    Side2 = sqrt(Side2);
    costet /= (Side1 * Side2);
    const double sinTetSq = 1 - costet * costet;
    double sintet = 0;
    if (sinTetSq > Threshold) sintet = sqrt(sinTetSq);
    Area = Side1 * Side2 * sintet / 2.0;
  }
  return Area;
}

// Node viscosity surface: for each node, the exposed faces of its prisms in
// prism order (Simulation::updateNodeViscositySurfaces, Simulation.cpp:7598-7621;
// ShapeBase::assignViscositySurfaceAreaToNodes, ShapeBase.cpp:3764-3775;
// Prism::assignExposedSurfaceAreaIndices, Prism.cpp:1650-1709).
__global__ __launch_bounds__(kBlock) void visc_surface_kernel(
    int32_t n_nodes, const int32_t* conn, const double* u, const int32_t* node_ptr,
    const int32_t* node_src, const int32_t* placement, const int32_t* type,
    const uint8_t* spans, const uint8_t* ablated, double* area) {
  for (int R = blockIdx.x * kBlock + threadIdx.x; R < n_nodes; R += gridDim.x * kBlock) {
    double A = 0.0;
    for (int t = node_ptr[R]; t < node_ptr[R + 1]; ++t) {
      const int s = node_src[t];
      const int e = s / 6, k = s % 6;
      const int tt = type[e], tp = placement[e];
      if (tt != 0 && tt != 1) continue;
      if (ablated && ablated[e]) continue;  // Simulation.cpp:7608-7621 skips ablated prisms
      const bool has_apical = (tp == 1) || (tp == 2 && spans[e]);
      const bool has_basal = (tp == 0) || (tp == 2 && spans[e]);
      const int* c = conn + 6 * e;
      const bool node_is_apical_face = (tt == 1) ? (k < 3) : (k >= 3);
      if (has_apical && node_is_apical_face) {
        const double ar = (tt == 1) ? tri_area(u, c[0], c[1], c[2]) : tri_area(u, c[3], c[4], c[5]);
        A += ar / 3;
      }
      if (has_basal && !node_is_apical_face) {
        const double ar = (tt == 1) ? tri_area(u, c[3], c[4], c[5]) : tri_area(u, c[0], c[1], c[2]);
        A += ar / 3;
      }
    }
    area[R] = A;
  }
}

// Pre-step growth on the device: the increment of ShapeBase::calculateFgFromRates
// (ShapeBase.cpp:471-505, no rotation, tissue weight 1),
// diag(exp(rx dt), exp(ry dt), exp(rz dt)), applied as ShapeBase::growShapeByFg
// does (ShapeBase.cpp:1452-1505) with identity plastic and shape-change
// increments: F_g <- inc F_g, then InvFg = inv(F_g) by the same LU as the
// element kernel.
__global__ __launch_bounds__(kBlock) void grow_kernel(int32_t n_prisms, const double* rates,
                                                      double dt, const uint8_t* ablated,
                                                      double* Fg, double* invFg) {
  for (int e = blockIdx.x * kBlock + threadIdx.x; e < n_prisms; e += gridDim.x * kBlock) {
    if (ablated && ablated[e]) continue;  // Simulation.cpp:6055: ablated prisms do not grow
    double inc[3];
    for (int i = 0; i < 3; ++i) inc[i] = exp(rates[3 * e + i] * dt);
    double F[9], X[9];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        // sum_k inc[i][k] Fg[k][j] with inc diagonal
        double s = 0.0;
        for (int k = 0; k < 3; ++k) s += ((i == k) ? inc[i] : 0.0) * Fg[9 * e + 3 * k + j];
        F[3 * i + j] = s;
      }
    for (int q = 0; q < 9; ++q) Fg[9 * e + q] = F[q];
    inv3_dev(F, X);
    for (int q = 0; q < 9; ++q) invFg[9 * e + q] = X[q];
  }
}

__global__ __launch_bounds__(kBlock) void inverse_kernel(int32_t n_prisms, const double* Fg,
                                                         double* invFg) {
  for (int e = blockIdx.x * kBlock + threadIdx.x; e < n_prisms; e += gridDim.x * kBlock) {
    double F[9], X[9];
    for (int q = 0; q < 9; ++q) F[q] = Fg[9 * e + q];
    inv3_dev(F, X);
    for (int q = 0; q < 9; ++q) invFg[9 * e + q] = X[q];
  }
}

int grid_for(int64_t n) {
  int64_t g = (n + kBlock - 1) / kBlock;
  if (g > 65535 * 4) g = 65535 * 4;
  if (g < 1) g = 1;
  return (int)g;
}

}  // namespace

hipError_t launch_assemble(const AssembleArgs& a, int mixed, hipStream_t st) {
  const int g1 = xcd_grid(a.pat.n_slices);
  if (a.any_visc && !(mixed && a.combined)) {
    if (mixed) hipLaunchKernelGGL((assemble_kernel<true, true>), dim3(g1), dim3(kBlock), 0, st, a);
    else hipLaunchKernelGGL((assemble_kernel<true, false>), dim3(g1), dim3(kBlock), 0, st, a);
  } else {
    if (mixed) hipLaunchKernelGGL((assemble_kernel<false, true>), dim3(g1), dim3(kBlock), 0, st, a);
    else hipLaunchKernelGGL((assemble_kernel<false, false>), dim3(g1), dim3(kBlock), 0, st, a);
  }
  return hipGetLastError();
}

namespace {
// A_b slot by slot (one thread per slot, as assemble_kernel), and the
// block-Jacobi inverse of its diagonal blocks.
template <typename T>
__global__ __launch_bounds__(kBlock) void bound_gather_kernel(BoundOp b, int32_t n_nodes,
                                                              int64_t nsJ, const T* __restrict__ valJ,
                                                              T* __restrict__ valB,
                                                              T* __restrict__ minvB) {
  constexpr int kCols = kBlock / kSlice;
  const int64_t nsB = b.pat.n_slots;
  const int lane = threadIdx.x % kSlice, kq = threadIdx.x / kSlice;
  for (int sl = xcd_item(blockIdx.x, gridDim.x); sl < b.pat.n_slices; sl += gridDim.x) {
    const int R = sl * kSlice + lane;
    const int64_t sbase = b.pat.slice_base[sl];
    const int width = (int)((b.pat.slice_base[sl + 1] - sbase) / kSlice);
    const int row_len = (R < n_nodes) ? b.pat.row_len[R] : 0;
    for (int kk = kq; kk < width; kk += kCols) {
      const int64_t slot = sbase + (int64_t)kk * kSlice + lane;
      double J[9];
      if (kk >= row_len) {
#pragma unroll
        for (int q = 0; q < 9; ++q) valB[(size_t)q * nsB + slot] = (T)0;
        continue;
      }
      const int bm = b.blockmap[slot];
      if (bm >= 0) {
#pragma unroll
        for (int q = 0; q < 9; ++q) J[q] = (double)valJ[(size_t)q * nsJ + bm];
      } else {
        const int ex = -bm - 1;
#pragma unroll
        for (int q = 0; q < 9; ++q) {
          double acc = 0.0;
          for (int t = b.ex_ptr[9 * ex + q]; t < b.ex_ptr[9 * ex + q + 1]; ++t)
            acc = acc + (double)valJ[b.ex_src[t]];
          if (b.ex_one[9 * ex + q]) acc = acc + 1.0;
          J[q] = acc;
        }
      }
#pragma unroll
      for (int q = 0; q < 9; ++q) valB[(size_t)q * nsB + slot] = (T)J[q];
      if (b.pat.col[slot] == R) {
        double M[9], X[9];
#pragma unroll
        for (int q = 0; q < 9; ++q) M[q] = (double)(T)J[q];
        inv3_dev(M, X);
#pragma unroll
        for (int q = 0; q < 9; ++q) minvB[9 * R + q] = (T)X[q];
      }
    }
  }
}
}  // namespace

hipError_t launch_bound_gather(const BoundOp& b, int32_t n_nodes, int64_t n_slots_J,
                               const double* valJ, const float* valJ32, int mixed,
                               hipStream_t st) {
  const int g1 = xcd_grid(b.pat.n_slices);
  if (mixed)
    hipLaunchKernelGGL(bound_gather_kernel<float>, dim3(g1), dim3(kBlock), 0, st, b, n_nodes,
                       n_slots_J, valJ32, b.val32, b.minv32);
  else
    hipLaunchKernelGGL(bound_gather_kernel<double>, dim3(g1), dim3(kBlock), 0, st, b, n_nodes,
                       n_slots_J, valJ, b.val, b.minv);
  return hipGetLastError();
}

hipError_t launch_bind_forces(const BoundOp& b, double* g, hipStream_t st) {
  if (b.n_masters == 0) return hipSuccess;
  hipLaunchKernelGGL(bind_forces_kernel, dim3(grid_for(b.n_masters)), dim3(kBlock), 0, st,
                     b.bm_ptr, b.bm_master, b.bm_slave, b.n_masters, g);
  return hipGetLastError();
}

hipError_t launch_bind_increment(const int32_t* pairs, int n_pairs, double* du, hipStream_t st) {
  if (n_pairs == 0) return hipSuccess;
  hipLaunchKernelGGL(bind_increment_kernel, dim3(grid_for(n_pairs)), dim3(kBlock), 0, st, pairs,
                     n_pairs, du);
  return hipGetLastError();
}

hipError_t launch_visc_surface(int32_t n_nodes, int32_t n_prisms, const int32_t* conn,
                               const double* u, const int32_t* node_ptr,
                               const int32_t* node_src, const int32_t* placement,
                               const int32_t* type, const uint8_t* spans, double* area,
                               hipStream_t st, const uint8_t* ablated) {
  (void)n_prisms;
  hipLaunchKernelGGL(visc_surface_kernel, dim3(grid_for(n_nodes)), dim3(kBlock), 0, st, n_nodes,
                     conn, u, node_ptr, node_src, placement, type, spans, ablated, area);
  return hipGetLastError();
}

hipError_t launch_grow(int32_t n_prisms, const double* rates, double dt, double* Fg,
                       double* invFg, hipStream_t st, const uint8_t* ablated) {
  hipLaunchKernelGGL(grow_kernel, dim3(grid_for(n_prisms)), dim3(kBlock), 0, st, n_prisms, rates,
                     dt, ablated, Fg, invFg);
  return hipGetLastError();
}

namespace {
// One workgroup per ablated prism: its 324 tangent entries (in the buffers the
// element mode writes), its 18 + 18 force entries and its strain are set to zero.
__global__ __launch_bounds__(kBlock) void ablate_zero_kernel(ElementArgs a, int mode,
                                                             const int32_t* list, int n) {
  const int i = blockIdx.x;
  if (i >= n) return;
  const size_t e = (size_t)list[i];
  for (int q = threadIdx.x; q < 324; q += kBlock) {
    if (mode == 0) {
      a.Ke[e * 324 + q] = 0.0;
      if (a.any_visc) a.Kv[e * 324 + q] = 0.0;
    } else {
      a.Ke32[e * 324 + q] = 0.f;
      if (a.any_visc && mode == 1) a.Kv32[e * 324 + q] = 0.f;
    }
  }
  for (int q = threadIdx.x; q < 18; q += kBlock) {
    a.Fe[e * 18 + q] = 0.0;
    a.Fv[e * 18 + q] = 0.0;
  }
  // an ablated prism is never evaluated by the reference, whose strain of it
  // stays at its initial zero
  if (a.strain && threadIdx.x < 6) a.strain[e * 6 + threadIdx.x] = 0.0;
}
}  // namespace

hipError_t launch_ablate_zero(const ElementArgs& a, int mode, const int32_t* list, int n,
                              hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(ablate_zero_kernel, dim3(n), dim3(kBlock), 0, st, a, mode, list, n);
  return hipGetLastError();
}

hipError_t launch_inverse(int32_t n_prisms, const double* Fg, double* invFg, hipStream_t st) {
  hipLaunchKernelGGL(inverse_kernel, dim3(grid_for(n_prisms)), dim3(kBlock), 0, st, n_prisms, Fg,
                     invFg);
  return hipGetLastError();
}

namespace {
// Plain 16-byte-vector copies: the bandwidth reference (STREAM copy).  A
// workgroup copies U x 256 consecutive vectors, each thread issuing its U
// loads before its U stores, with ordinary or non-temporal accesses.
typedef float v4f __attribute__((ext_vector_type(4)));
template <int U, bool NT>
__global__ __launch_bounds__(kBlock) void tf_copy_kernel(const v4f* __restrict__ src,
                                                         v4f* __restrict__ dst, int64_t n4) {
  const int64_t base = (int64_t)blockIdx.x * (kBlock * U) + threadIdx.x;
  if (base + (int64_t)(U - 1) * kBlock < n4) {
    v4f v[U];
#pragma unroll
    for (int u = 0; u < U; ++u)
      v[u] = NT ? __builtin_nontemporal_load(src + base + u * kBlock) : src[base + u * kBlock];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (NT) __builtin_nontemporal_store(v[u], dst + base + u * kBlock);
      else dst[base + u * kBlock] = v[u];
    }
  } else {
    for (int u = 0; u < U; ++u)
      if (base + u * kBlock < n4) dst[base + u * kBlock] = src[base + u * kBlock];
  }
}

template <int U, bool NT>
hipError_t launch_copy_t(const void* src, void* dst, int64_t n4, hipStream_t st) {
  const int64_t grid = (n4 + kBlock * U - 1) / (kBlock * U);
  if (grid > INT32_MAX) return hipErrorInvalidValue;
  hipLaunchKernelGGL((tf_copy_kernel<U, NT>), dim3((unsigned)grid), dim3(kBlock), 0, st,
                     (const v4f*)src, (v4f*)dst, n4);
  return hipGetLastError();
}
}  // namespace

hipError_t launch_copy(const void* src, void* dst, int64_t n4, int variant, hipStream_t st) {
  switch (variant) {
    case 0: return launch_copy_t<1, false>(src, dst, n4, st);
    case 1: return launch_copy_t<2, false>(src, dst, n4, st);
    case 2: return launch_copy_t<4, false>(src, dst, n4, st);
    case 3: return launch_copy_t<1, true>(src, dst, n4, st);
    case 4: return launch_copy_t<2, true>(src, dst, n4, st);
    default: return launch_copy_t<4, true>(src, dst, n4, st);
  }
}

}  // namespace tf
