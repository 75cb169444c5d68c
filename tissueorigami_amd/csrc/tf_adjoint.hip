// Adjoint of the implicit NR step with respect to per-prism growth parameters.
//
// One converged step solves g(u_n; u_{n-1}, theta) = 0 with J = -dg/du_n.  For
// a loss L(u_N), the implicit-function theorem gives, step by step backwards,
//   J^T mu_n = a_n,
//   dL/dtheta += mu_n^T dg/dtheta,   a_{n-1} = mu_n^T dg/du_{n-1},
// starting from a_N = dL/du_N.  Here theta is one scalar r_e per prism that
// enters the step only through the prism's inverse growth tensor, and the
// caller supplies dInvFg_e/dr_e for every step.  The pieces of dg are
// hand-derived from the forward definitions:
//   * elastic force of a prism, Prism::calculateCurrNodalForces
//     (Prism.cpp:766-886): g_a = -sum_gp w detF detdXde sigma(Fe) gradN_a with
//     sigma = Je^-1 [mu (Fe Fe^T - I) + lambda ln Je I], Fe = F InvFg, so for
//     a growth perturbation dFe = F dInvFg at fixed positions
//     dsigma = -tr(Fe^-1 dFe) sigma + Je^-1 [mu (dFe Fe^T + Fe dFe^T) + lambda tr(Fe^-1 dFe) I];
//   * internal viscous force (ShapeBase.cpp:2640-2843): depends on u_{n-1}
//     only through v = (u_n - u_{n-1})/dt, with d = sym(sum_c v_c (x) gradN_c);
//   * external drag (NewtonRaphsonSolver.cpp:362-397): g_i = -A eta_i v_i with
//     A the node's exposed area computed from u_{n-1}
//     (Prism::calculateApicalArea/BasalArea, Prism.cpp:1485-1556), whose
//     gradient with respect to a face vertex p0 is (p1 - p2) x n_hat / 2.
// Fixed DOFs carry g = 0 identically and are masked out of every product.
#include "tf_internal.h"

namespace tf {
namespace {

constexpr int kAdjThreads = 64;

__device__ inline double det3a(const double M[3][3]) {
  return M[0][0] * (M[1][1] * M[2][2] - M[1][2] * M[2][1]) -
         M[0][1] * (M[1][0] * M[2][2] - M[1][2] * M[2][0]) +
         M[0][2] * (M[1][0] * M[2][1] - M[1][1] * M[2][0]);
}

__device__ inline void inv3a(const double A[3][3], double X[3][3]) {
  const double d = det3a(A);
  const double r = 1.0 / d;
  X[0][0] = (A[1][1] * A[2][2] - A[1][2] * A[2][1]) * r;
  X[0][1] = (A[0][2] * A[2][1] - A[0][1] * A[2][2]) * r;
  X[0][2] = (A[0][1] * A[1][2] - A[0][2] * A[1][1]) * r;
  X[1][0] = (A[1][2] * A[2][0] - A[1][0] * A[2][2]) * r;
  X[1][1] = (A[0][0] * A[2][2] - A[0][2] * A[2][0]) * r;
  X[1][2] = (A[0][2] * A[1][0] - A[0][0] * A[1][2]) * r;
  X[2][0] = (A[1][0] * A[2][1] - A[1][1] * A[2][0]) * r;
  X[2][1] = (A[0][1] * A[2][0] - A[0][0] * A[2][1]) * r;
  X[2][2] = (A[0][0] * A[1][1] - A[0][1] * A[1][0]) * r;
}

__device__ inline void tri_area_grad(const double* u, int n0, int n1, int n2, double g[3][3]) {
  double p0[3], p1[3], p2[3], e1[3], e2[3], n[3];
  for (int i = 0; i < 3; ++i) {
    p0[i] = u[3 * n0 + i]; p1[i] = u[3 * n1 + i]; p2[i] = u[3 * n2 + i];
    e1[i] = p1[i] - p0[i]; e2[i] = p2[i] - p0[i];
  }
  n[0] = e1[1] * e2[2] - e1[2] * e2[1];
  n[1] = e1[2] * e2[0] - e1[0] * e2[2];
  n[2] = e1[0] * e2[1] - e1[1] * e2[0];
  const double nn = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
  for (int k = 0; k < 3; ++k)
    for (int i = 0; i < 3; ++i) g[k][i] = 0.0;
  // the forward area is 0 below its thresholds (degenerate triangle): no gradient
  const double s1 = e1[0] * e1[0] + e1[1] * e1[1] + e1[2] * e1[2];
  const double s2 = e2[0] * e2[0] + e2[1] * e2[1] + e2[2] * e2[2];
  if (!(s1 > 1E-5 && s2 > 1E-5) || nn == 0.0) return;
  const double c = (e1[0] * e2[0] + e1[1] * e2[1] + e1[2] * e2[2]) / sqrt(s1 * s2);
  if (!(1 - c * c > 1E-5)) return;
  const double h[3] = {n[0] / nn, n[1] / nn, n[2] / nn};
  const double* P[3] = {p0, p1, p2};
  for (int k = 0; k < 3; ++k) {
    const double* a = P[(k + 1) % 3];
    const double* b = P[(k + 2) % 3];
    const double d[3] = {a[0] - b[0], a[1] - b[1], a[2] - b[2]};
    g[k][0] = 0.5 * (d[1] * h[2] - d[2] * h[1]);
    g[k][1] = 0.5 * (d[2] * h[0] - d[0] * h[2]);
    g[k][2] = 0.5 * (d[0] * h[1] - d[1] * h[0]);
  }
}

struct AdjArgs {
  int32_t n_prisms;
  const int32_t* conn;
  const double* refpos;
  const double* invFg;
  const double* dinvFg;
  const double* lambda;
  const double* mu;
  const double* eta;
  const double* ext_visc;
  const uint8_t* dof_flags;
  const double* uk;   // u_n
  const double* un;   // u_{n-1}
  double dt;
  const double* m;    // adjoint mu_n (3 n_nodes)
  const int32_t* placement;
  const int32_t* type;
  const uint8_t* spans;
  double* s_out;      // per prism: mu^T dg/dr_e
  double* b_out;      // per prism x 18: contributions to a_{n-1}
  double sfd[6][3][6];
  double w;
};

// One thread per prism.
__global__ __launch_bounds__(kAdjThreads) void adjoint_element_kernel(AdjArgs a) {
  const int e = blockIdx.x * kAdjThreads + threadIdx.x;
  if (e >= a.n_prisms) return;
  int nd[6];
  double pos[6][3], ref[6][3], iFg[3][3], diFg[3][3], mu_e[6][3], vel[6][3];
  const double rdt = 1.0 / a.dt;
  for (int k = 0; k < 6; ++k) {
    nd[k] = a.conn[6 * e + k];
    for (int i = 0; i < 3; ++i) {
      pos[k][i] = a.uk[3 * nd[k] + i];
      ref[k][i] = a.refpos[18 * e + 3 * k + i];
      vel[k][i] = (a.uk[3 * nd[k] + i] - a.un[3 * nd[k] + i]) * rdt;
      const bool fx = a.dof_flags && (a.dof_flags[3 * nd[k] + i] & 1);
      mu_e[k][i] = fx ? 0.0 : a.m[3 * nd[k] + i];
    }
  }
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      iFg[i][j] = a.invFg[9 * e + 3 * i + j];
      diFg[i][j] = a.dinvFg[9 * e + 3 * i + j];
    }
  const double lambda = a.lambda[e], mu = a.mu[e], eta = a.eta[e];
  double s_acc = 0.0;
  double b[6][3];
  for (int k = 0; k < 6; ++k)
    for (int i = 0; i < 3; ++i) b[k][i] = 0.0;
  for (int g = 0; g < 6; ++g) {
    const double(*SFD)[6] = a.sfd[g];
    double dXde[3][3], Jac[3][3];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        double s = 0.0, t = 0.0;
        for (int k = 0; k < 6; ++k) {
          s += SFD[j][k] * ref[k][i];
          t += SFD[j][k] * pos[k][i];
        }
        dXde[i][j] = s;
        Jac[i][j] = t;
      }
    const double detdXde = det3a(dXde);
    double InvdXde[3][3], F[3][3], Fe[3][3], dFe[3][3], InvJ[3][3], Fe_inv[3][3];
    inv3a(dXde, InvdXde);
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        double s = 0.0;
        for (int k = 0; k < 3; ++k) s += Jac[i][k] * InvdXde[k][j];
        F[i][j] = s;
      }
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        double s = 0.0, t = 0.0;
        for (int k = 0; k < 3; ++k) {
          s += F[i][k] * iFg[k][j];
          t += F[i][k] * diFg[k][j];
        }
        Fe[i][j] = s;
        dFe[i][j] = t;
      }
    const double detF = det3a(F);
    const double Je = det3a(Fe);
    const double lnJ = log(Je);
    inv3a(Fe, Fe_inv);
    inv3a(Jac, InvJ);
    double dN[6][3];
    for (int k = 0; k < 6; ++k)
      for (int m = 0; m < 3; ++m) {
        double s = 0.0;
        for (int n = 0; n < 3; ++n) s += InvJ[n][m] * SFD[n][k];
        dN[k][m] = s;
      }
    const double vol = a.w * detF * detdXde;
    // sigma and its growth derivative
    double bb[3][3], sig[3][3], dsig[3][3];
    double tr = 0.0;
    for (int i = 0; i < 3; ++i)
      for (int k = 0; k < 3; ++k) tr += Fe_inv[i][k] * dFe[k][i];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        double s = 0.0, t = 0.0;
        for (int k = 0; k < 3; ++k) {
          s += Fe[i][k] * Fe[j][k];
          t += dFe[i][k] * Fe[j][k] + Fe[i][k] * dFe[j][k];
        }
        bb[i][j] = s;
        const double I = (i == j) ? 1.0 : 0.0;
        sig[i][j] = (mu * (s - I) + lambda * lnJ * I) / Je;
        dsig[i][j] = (mu * t + lambda * tr * I) / Je;
      }
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) dsig[i][j] -= tr * sig[i][j];
    // mu^T dg/dr: dg_a = -vol dsig gradN_a
    for (int k = 0; k < 6; ++k)
      for (int i = 0; i < 3; ++i) {
        double s = 0.0;
        for (int j = 0; j < 3; ++j) s += dsig[i][j] * dN[k][j];
        s_acc -= vol * s * mu_e[k][i];
      }
    // viscous: a_{n-1}[c][j] += (eta/dt) vol 1/2 sum_a [mu_a_j (gN_c . gN_a) + (mu_a . gN_c) gN_a_j]
    if (eta != 0.0) {
      const double f = 0.5 * eta * vol * rdt;
      for (int c = 0; c < 6; ++c)
        for (int ka = 0; ka < 6; ++ka) {
          const double dot = dN[c][0] * dN[ka][0] + dN[c][1] * dN[ka][1] + dN[c][2] * dN[ka][2];
          const double md = mu_e[ka][0] * dN[c][0] + mu_e[ka][1] * dN[c][1] + mu_e[ka][2] * dN[c][2];
          for (int j = 0; j < 3; ++j) b[c][j] += f * (mu_e[ka][j] * dot + md * dN[ka][j]);
        }
    }
  }
  // exposed faces: drag depends on u_{n-1} through the node areas
  const int tt = a.type[e], tp = a.placement[e];
  if (tt == 0 || tt == 1) {
    const bool has_apical = (tp == 1) || (tp == 2 && a.spans[e]);
    const bool has_basal = (tp == 0) || (tp == 2 && a.spans[e]);
    for (int face = 0; face < 2; ++face) {
      const bool apical = face == 0;
      if (apical ? !has_apical : !has_basal) continue;
      // apical face of a columnar prism is 3,4,5 (0,1,2 for peripodial)
      const bool top = apical ? (tt == 0) : (tt == 1);
      const int f0 = top ? 3 : 0;
      double q = 0.0;  // sum over the face's nodes of mu . dg/dA
      for (int k = f0; k < f0 + 3; ++k)
        for (int i = 0; i < 3; ++i) {
          const double vi = (a.uk[3 * nd[k] + i] - a.un[3 * nd[k] + i]) * rdt;
          q += mu_e[k][i] * (-a.ext_visc[3 * nd[k] + i] * vi);
        }
      double grad[3][3];
      tri_area_grad(a.un, nd[f0], nd[f0 + 1], nd[f0 + 2], grad);
      for (int k = 0; k < 3; ++k)
        for (int i = 0; i < 3; ++i) b[f0 + k][i] += (q / 3) * grad[k][i];
    }
  }
  a.s_out[e] = s_acc;
  for (int k = 0; k < 6; ++k)
    for (int i = 0; i < 3; ++i) a.b_out[18 * e + 3 * k + i] = b[k][i];
}

// a_{n-1} = sum of element contributions (element order) + the direct drag
// term A eta / dt mu (free DOFs only).
__global__ __launch_bounds__(kBlock) void adjoint_node_kernel(
    int32_t n_nodes, const int32_t* node_ptr, const int32_t* node_src, const double* b_el,
    const double* area, const double* ext_visc, const uint8_t* dof_flags, const double* m,
    double dt, double* a_prev) {
  for (int R = blockIdx.x * kBlock + threadIdx.x; R < n_nodes; R += gridDim.x * kBlock) {
    double acc[3] = {0.0, 0.0, 0.0};
    for (int t = node_ptr[R]; t < node_ptr[R + 1]; ++t) {
      const int s = node_src[t];
      for (int i = 0; i < 3; ++i) acc[i] += b_el[3 * (size_t)s + i];
    }
    for (int i = 0; i < 3; ++i) {
      const bool fx = dof_flags && (dof_flags[3 * R + i] & 1);
      const double mi = fx ? 0.0 : m[3 * R + i];
      a_prev[3 * R + i] = acc[i] + area[R] * ext_visc[3 * R + i] / dt * mi;
    }
  }
}

// J^T values: valT[c][slot] = val[c^T][tslot[slot]].
__global__ __launch_bounds__(kBlock) void transpose_values_kernel(int64_t n_slots,
                                                                  const int32_t* tslot,
                                                                  const double* val,
                                                                  double* valT) {
  for (int64_t s = (int64_t)blockIdx.x * kBlock + threadIdx.x; s < n_slots;
       s += (int64_t)gridDim.x * kBlock) {
    const int64_t t = tslot[s];
    for (int p = 0; p < 3; ++p)
      for (int q = 0; q < 3; ++q) valT[(size_t)(3 * p + q) * n_slots + s] = val[(size_t)(3 * q + p) * n_slots + t];
  }
}

__global__ __launch_bounds__(kBlock) void transpose_blocks_kernel(int32_t n, const double* a,
                                                                  double* b) {
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock)
    for (int p = 0; p < 3; ++p)
      for (int q = 0; q < 3; ++q) b[9 * (size_t)i + 3 * p + q] = a[9 * (size_t)i + 3 * q + p];
}

}  // namespace

hipError_t launch_adjoint_element(const AdjointLaunch& l, hipStream_t st) {
  AdjArgs a;
  a.n_prisms = l.n_prisms; a.conn = l.conn; a.refpos = l.refpos; a.invFg = l.invFg;
  a.dinvFg = l.dinvFg; a.lambda = l.lambda; a.mu = l.mu; a.eta = l.eta;
  a.ext_visc = l.ext_visc; a.dof_flags = l.dof_flags; a.uk = l.uk; a.un = l.un; a.dt = l.dt;
  a.m = l.m; a.placement = l.placement; a.type = l.type; a.spans = l.spans;
  a.s_out = l.s_out; a.b_out = l.b_out;
  const double sq = 1.0 / std::pow(3.0, 0.5);
  const double A = 1.0 / 6.0, B = 2.0 / 3.0;
  const double pts[6][3] = {{A, A, sq}, {B, A, sq}, {A, B, sq}, {A, A, -sq}, {B, A, -sq}, {A, B, -sq}};
  for (int g = 0; g < 6; ++g) {
    const double eta = pts[g][0], nu = pts[g][1], zeta = pts[g][2];
    const double alpha = (1 - zeta) / 2, beta = (1 + zeta) / 2, lam = 1 - eta - nu;
    const double r0[6] = {-alpha, alpha, 0, -beta, beta, 0};
    const double r1[6] = {-alpha, 0, alpha, -beta, 0, beta};
    const double r2[6] = {-lam / 2, -eta / 2, -nu / 2, lam / 2, eta / 2, nu / 2};
    for (int k = 0; k < 6; ++k) {
      a.sfd[g][0][k] = r0[k]; a.sfd[g][1][k] = r1[k]; a.sfd[g][2][k] = r2[k];
    }
  }
  a.w = 1.0 / 6.0;
  const int grid = (l.n_prisms + kAdjThreads - 1) / kAdjThreads;
  hipLaunchKernelGGL(adjoint_element_kernel, dim3(grid), dim3(kAdjThreads), 0, st, a);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  int g2 = (l.n_nodes + kBlock - 1) / kBlock;
  g2 = std::max(1, std::min(g2, 65535));
  hipLaunchKernelGGL(adjoint_node_kernel, dim3(g2), dim3(kBlock), 0, st, l.n_nodes, l.node_ptr,
                     l.node_src, l.b_out, l.area, l.ext_visc, l.dof_flags, l.m, l.dt, l.a_prev);
  return hipGetLastError();
}

hipError_t launch_transpose_operator(int64_t n_slots, const int32_t* tslot, const double* val,
                                     double* valT, int32_t n_nodes, const double* minv,
                                     double* minvT, hipStream_t st) {
  int g = (int)std::min<int64_t>((n_slots + kBlock - 1) / kBlock, 65535 * 4);
  hipLaunchKernelGGL(transpose_values_kernel, dim3(std::max(g, 1)), dim3(kBlock), 0, st, n_slots,
                     tslot, val, valT);
  int g2 = std::max(1, std::min((n_nodes + kBlock - 1) / kBlock, 65535));
  hipLaunchKernelGGL(transpose_blocks_kernel, dim3(g2), dim3(kBlock), 0, st, n_nodes, minv, minvT);
  return hipGetLastError();
}

}  // namespace tf
