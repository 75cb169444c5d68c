// CPU oracle for the implicit Newton-Raphson step on 6-node prism meshes.
//
// TEST INFRASTRUCTURE ONLY.  This file restates, in plain C++ with no GSL, the
// reference algorithm of TissueOrigami's NR step so that tests, the smoke check
// and the CPU-baseline leg of the benchmark have something to compare the HIP
// product path against.  Nothing in the product library links, loads or calls
// this code.
//
// Every function cites the reference file:line it follows (paths relative to the
// reference repository root, TissueFolding/SourceCode/).  Where the reference
// delegates to a third-party routine (GSL's LU inverse, GSL dnrm2, Pardiso), the
// routine is restated from its published algorithm and named in the comment.
//
// Pinning (what checks each function; see tests/test_oracle_*.py):
//   prism geometry         -> exact wedge volume by independent quadrature
//   element forces         -> Save_Force fixtures of run07007, objectivity,
//                             homogeneous-growth patch test
//   faithful elastic K     -> central finite differences of the forces, symmetry,
//                             rigid-translation null space
//   closed-form elastic K  -> faithful 6561-term loop (rel 1e-12) and FD
//   viscous K (transposed) -> central finite differences of the viscous forces
//   dense assembly / BCs   -> brute-force scatter in numpy, fixture replay
//   threshold CSR          -> scipy.sparse on the same dense matrix
//   dense transposed solve -> numpy.linalg.solve
//   NR step                -> run07007 per-step replay (positions, |du| log)
//   BSR + BiCGStab path    -> dense transposed solve on small meshes
//   Young's modulus series -> closed-form rates, known-answer grid lookups,
//                             multi-step products against the closed form
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// Gauss rule of the 6-node prism: Prism.cpp:99-131 (numberOfGaussPoints == 6).
// Columns are (eta, nu, zeta) in the reference's naming.
struct GaussRule {
  double p[6][3];
  double w[6];
  GaussRule() {
    const double s = 1.0 / std::pow(3.0, 0.5);
    const double a = 1.0 / 6.0, b = 2.0 / 3.0;
    const double pts[6][3] = {{a, a, s}, {b, a, s}, {a, b, s},
                              {a, a, -s}, {b, a, -s}, {a, b, -s}};
    for (int g = 0; g < 6; ++g) {
      for (int k = 0; k < 3; ++k) p[g][k] = pts[g][k];
      w[g] = 1.0 / 6.0;
    }
  }
};
const GaussRule kGauss;

// Shape-function derivative table of the prism, Prism.cpp:597-652.
// Row 0: d/d(eta), row 1: d/d(nu), row 2: d/d(zeta); columns are the 6 nodes.
// The reference calls setShapeFunctionDerivatives(M, eta, zeta, nu) with
// eta = gp[0], zeta = gp[2], nu = gp[1] (Prism.cpp:720-724).
void shape_der(double eta, double zeta, double nu, double D[3][6]) {
  const double alpha = (1 - zeta) / 2;
  const double beta = (1 + zeta) / 2;
  const double lambda = 1 - eta - nu;
  const double r0[6] = {-alpha, alpha, 0, -beta, beta, 0};
  const double r1[6] = {-alpha, 0, alpha, -beta, 0, beta};
  const double r2[6] = {-lambda / 2, -eta / 2, -nu / 2, lambda / 2, eta / 2, nu / 2};
  for (int a = 0; a < 6; ++a) {
    D[0][a] = r0[a];
    D[1][a] = r1[a];
    D[2][a] = r2[a];
  }
}

// Cofactor determinant, ShapeBase.cpp:2264-2271.
double det3(const double M[3][3]) {
  double det = 0.0;
  det = M[0][0] * (M[1][1] * M[2][2] - M[1][2] * M[2][1]);
  det -= M[0][1] * (M[1][0] * M[2][2] - M[1][2] * M[2][0]);
  det += M[0][2] * (M[1][0] * M[2][1] - M[1][1] * M[2][0]);
  return det;
}

// 3x3 inverse through LU with partial pivoting (ShapeBase.cpp:3546-3561 calls
// gsl_linalg_LU_decomp + gsl_linalg_LU_invert; GSL's published algorithm is
// Doolittle elimination with row pivoting on the largest |a_ij|, followed by
// the solution of L U X = P I).  Rounding may differ from a given GSL build in
// the last bit; parity is asserted at the 1e-12 level, never bitwise.
void inv3(const double A[3][3], double X[3][3]) {
  double LU[3][3];
  int perm[3] = {0, 1, 2};
  std::memcpy(LU, A, sizeof(LU));
  for (int j = 0; j < 2; ++j) {
    int piv = j;
    double mx = std::fabs(LU[j][j]);
    for (int i = j + 1; i < 3; ++i)
      if (std::fabs(LU[i][j]) > mx) { mx = std::fabs(LU[i][j]); piv = i; }
    if (piv != j) {
      for (int k = 0; k < 3; ++k) std::swap(LU[j][k], LU[piv][k]);
      std::swap(perm[j], perm[piv]);
    }
    if (LU[j][j] != 0.0) {
      for (int i = j + 1; i < 3; ++i) {
        LU[i][j] /= LU[j][j];
        for (int k = j + 1; k < 3; ++k) LU[i][k] -= LU[i][j] * LU[j][k];
      }
    }
  }
  for (int c = 0; c < 3; ++c) {
    double y[3];
    for (int i = 0; i < 3; ++i) {
      double s = (perm[i] == c) ? 1.0 : 0.0;
      for (int k = 0; k < i; ++k) s -= LU[i][k] * y[k];
      y[i] = s;
    }
    for (int i = 2; i >= 0; --i) {
      double s = y[i];
      for (int k = i + 1; k < 3; ++k) s -= LU[i][k] * X[k][c];
      X[i][c] = s / LU[i][i];
    }
  }
}

void matmul3(const double A[3][3], const double B[3][3], double C[3][3]) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double s = 0.0;
      for (int k = 0; k < 3; ++k) s += A[i][k] * B[k][j];
      C[i][j] = s;
    }
}

// Per-element reference geometry at every Gauss point:
// Prism::calculateElementShapeFunctionDerivatives, Prism.cpp:698-733.
struct PrismGeom {
  double SFD[6][3][6];   // shape function derivatives wrt (eta, nu, zeta)
  double InvdXde[6][3][3];
  double detdXde[6];
};

void prism_geometry(const double ref[6][3], PrismGeom& G) {
  for (int g = 0; g < 6; ++g) {
    const double eta = kGauss.p[g][0], nu = kGauss.p[g][1], zeta = kGauss.p[g][2];
    shape_der(eta, zeta, nu, G.SFD[g]);
    // dXde = (SFD * RefShape)^T  (Prism.cpp:727-728)
    double dXde[3][3];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        double s = 0.0;
        for (int a = 0; a < 6; ++a) s += G.SFD[g][j][a] * ref[a][i];
        dXde[i][j] = s;
      }
    G.detdXde[g] = det3(dXde);
    inv3(dXde, G.InvdXde[g]);
  }
}

// Quantities kept per Gauss point for the tangent integrals
// (ShapeBase members FeMatrices, detFs, detdXdes, D81, elasticStress,
// invJShapeFuncDerStack, Bmatrices, viscousStress; ShapeBase.h:53-69,129-143).
struct GPState {
  double Fe[3][3];
  double detF, detdXde, detFe, lnJ;
  double D81[3][3][3][3];
  double sigma[3][3];
  double dNdx[6][3];
  double B[6][18];
  double sigv[3][3];
};

// One Gauss point of Prism::calculateCurrNodalForces, Prism.cpp:766-886, with
// the ShapeBase helpers it calls (ShapeBase.cpp:2448-2845).
void gauss_point_forces(const double pos[6][3], const PrismGeom& G, int g,
                        const double invFg[3][3], double lambda, double mu,
                        double eta_visc, const double vel[6][3], GPState& S,
                        double currge[18], double currgv[18]) {
  const double (*SFD)[6] = G.SFD[g];
  // Jacobian = (ShapeFuncDer * CurrShape)^T  (Prism.cpp:782-783)
  double Jac[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double s = 0.0;
      for (int a = 0; a < 6; ++a) s += SFD[j][a] * pos[a][i];
      Jac[i][j] = s;
    }
  double F[3][3], Fe[3][3];
  matmul3(Jac, G.InvdXde[g], F);          // Prism.cpp:785
  matmul3(F, invFg, Fe);                  // Prism.cpp:788
  std::memcpy(S.Fe, Fe, sizeof(Fe));
  // C = Fe^T Fe  (ShapeBase.cpp:2448-2458)
  double C[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double s = 0.0;
      for (int k = 0; k < 3; ++k) s += Fe[k][i] * Fe[k][j];
      C[i][j] = s;
    }
  const double detFe = det3(Fe);          // Prism.cpp:802
  const double lnJ = std::log(detFe);     // Prism.cpp:804
  S.detFe = detFe;
  S.lnJ = lnJ;
  double InvC[3][3];
  inv3(C, InvC);                          // Prism.cpp:818-823
  // S = mu (I - C^-1) + lambda lnJ C^-1  (ShapeBase.cpp:2507-2525)
  double PK2[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      const double Iij = (i == j) ? 1.0 : 0.0;
      PK2[i][j] = InvC[i][j] * (lambda * lnJ) + (Iij - InvC[i][j]) * mu;
    }
  // D81 (ShapeBase.cpp:2527-2545)
  const double multiplier = 2 * (mu - lambda * lnJ);
  for (int I = 0; I < 3; ++I)
    for (int J = 0; J < 3; ++J)
      for (int K = 0; K < 3; ++K)
        for (int L = 0; L < 3; ++L) {
          const double Iijkl = 0.5 * (InvC[I][K] * InvC[J][L] + InvC[I][L] * InvC[J][K]);
          S.D81[I][J][K][L] = lambda * InvC[I][J] * InvC[K][L] + multiplier * Iijkl;
        }
  // sigma = Fe S Fe^T / detFe  (ShapeBase.cpp:2547-2568)
  double T1[3][3];
  matmul3(Fe, PK2, T1);
  const double inv_detFe = 1.0 / detFe;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double s = 0.0;
      for (int k = 0; k < 3; ++k) s += T1[i][k] * Fe[j][k];
      S.sigma[i][j] = s * inv_detFe;
    }
  const double compact[6] = {S.sigma[0][0], S.sigma[1][1], S.sigma[2][2],
                             S.sigma[0][1], S.sigma[2][1], S.sigma[0][2]};
  // Inverse Jacobian stack and B (ShapeBase.cpp:2605-2638): dN_a/dx_m =
  // sum_n dN_a/dxi_n InvJac(n, m); B uses the Voigt map of Prism.cpp:323-367.
  double InvJ[3][3];
  inv3(Jac, InvJ);
  for (int a = 0; a < 6; ++a)
    for (int m = 0; m < 3; ++m) {
      double s = 0.0;
      for (int n = 0; n < 3; ++n) s += InvJ[n][m] * SFD[n][a];
      S.dNdx[a][m] = s;
    }
  std::memset(S.B, 0, sizeof(S.B));
  for (int a = 0; a < 6; ++a) {
    const double nx = S.dNdx[a][0], ny = S.dNdx[a][1], nz = S.dNdx[a][2];
    S.B[0][3 * a + 0] = nx;
    S.B[1][3 * a + 1] = ny;
    S.B[2][3 * a + 2] = nz;
    S.B[3][3 * a + 0] = ny;
    S.B[3][3 * a + 1] = nx;
    S.B[4][3 * a + 1] = nz;
    S.B[4][3 * a + 2] = ny;
    S.B[5][3 * a + 0] = nz;
    S.B[5][3 * a + 2] = nx;
  }
  S.detF = det3(F);                       // Prism.cpp:850
  S.detdXde = G.detdXde[g];
  // currge = (B^T detF detdXde) compactStress  (Prism.cpp:858-860)
  double BT[18][6];
  for (int r = 0; r < 18; ++r)
    for (int v = 0; v < 6; ++v) {
      double x = S.B[v][r] * S.detF;
      x *= S.detdXde;
      BT[r][v] = x;
    }
  for (int r = 0; r < 18; ++r) {
    double s = 0.0;
    for (int v = 0; v < 6; ++v) s += BT[r][v] * compact[v];
    currge[r] = s;
  }
  // viscous stress and forces (Prism.cpp:862-871, ShapeBase.cpp:2640-2843)
  std::memset(S.sigv, 0, sizeof(S.sigv));
  for (int r = 0; r < 18; ++r) currgv[r] = 0.0;
  if (eta_visc != 0) {
    double l[6];
    for (int v = 0; v < 6; ++v) {
      double s = 0.0;
      for (int r = 0; r < 18; ++r) s += S.B[v][r] * vel[r / 3][r % 3];
      l[v] = s;
    }
    double d[3][3];
    d[0][0] = l[0]; d[1][1] = l[1]; d[2][2] = l[2];
    d[0][1] = 0.5 * l[3]; d[2][1] = 0.5 * l[4]; d[0][2] = 0.5 * l[5];
    d[1][0] = 0.5 * l[3]; d[1][2] = 0.5 * l[4]; d[2][0] = 0.5 * l[5];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) S.sigv[i][j] = d[i][j] * eta_visc;
    const double cv[6] = {S.sigv[0][0], S.sigv[1][1], S.sigv[2][2],
                          S.sigv[0][1], S.sigv[2][1], S.sigv[0][2]};
    for (int r = 0; r < 18; ++r) {
      double s = 0.0;
      for (int v = 0; v < 6; ++v) s += BT[r][v] * cv[v];
      currgv[r] = s;
    }
  }
}

// ShapeBase::calculateElasticKIntegral1, ShapeBase.cpp:3010-3061 (the 8-deep loop).
void elastic_k_integral1(const GPState& S, double K[18][18]) {
  for (int a = 0; a < 6; ++a)
    for (int b = 0; b < 6; ++b) {
      const double* DNa = S.dNdx[a];
      const double* DNb = S.dNdx[b];
      for (int i = 0; i < 3; ++i)
        for (int k = 0; k < 3; ++k) {
          double value = 0;
          for (int j = 0; j < 3; ++j)
            for (int l = 0; l < 3; ++l)
              for (int I = 0; I < 3; ++I)
                for (int J = 0; J < 3; ++J)
                  for (int Kk = 0; Kk < 3; ++Kk)
                    for (int L = 0; L < 3; ++L)
                      value += S.Fe[i][I] * S.Fe[j][J] * S.Fe[k][Kk] * S.Fe[l][L] *
                               S.D81[I][J][Kk][L] * DNb[l] * DNa[j];
          value *= S.detF * S.detdXde;
          value /= S.detFe;
          K[3 * a + i][3 * b + k] += value;
        }
    }
}

// ShapeBase::calculateElasticKIntegral2, ShapeBase.cpp:3064-3095.
void elastic_k_integral2(const GPState& S, double K[18][18]) {
  for (int a = 0; a < 6; ++a)
    for (int b = 0; b < 6; ++b) {
      double tmp1[3];
      for (int j = 0; j < 3; ++j) {
        double s = 0.0;
        for (int i = 0; i < 3; ++i) s += S.dNdx[a][i] * S.sigma[i][j];
        tmp1[j] = s;
      }
      double keab2 = 0.0;
      for (int j = 0; j < 3; ++j) keab2 += tmp1[j] * S.dNdx[b][j];
      const double value = keab2 * S.detF * S.detdXde;
      for (int i = 0; i < 3; ++i) K[3 * a + i][3 * b + i] += value;
    }
}

// Closed form of the elastic tangent at one Gauss point.  Because
// Fe C^-1 Fe^T = I, the push-forward of D81 collapses to
// Je^-1 [lambda d_ij d_kl + mu' (d_ik d_jl + d_il d_jk)], mu' = mu - lambda lnJ,
// so Integral1 + Integral2 of ShapeBase.cpp:3010-3095 equal
//   Ke_ab(i,k) = detF detdXde { Je^-1 [lambda dNa_i dNb_k
//                + mu' (d_ik dNa.dNb + dNa_k dNb_i)] + d_ik dNa.sigma.dNb }.
void elastic_k_closed(const GPState& S, double lambda, double mu, double K[18][18]) {
  const double vol = S.detF * S.detdXde;
  const double c = vol / S.detFe;
  const double cl = c * lambda;
  const double cm = c * (mu - lambda * S.lnJ);
  for (int a = 0; a < 6; ++a)
    for (int b = 0; b < 6; ++b) {
      const double* Na = S.dNdx[a];
      const double* Nb = S.dNdx[b];
      double dot = 0.0, geo = 0.0;
      for (int j = 0; j < 3; ++j) dot += Na[j] * Nb[j];
      for (int i = 0; i < 3; ++i) {
        double s = 0.0;
        for (int j = 0; j < 3; ++j) s += S.sigma[i][j] * Nb[j];
        geo += Na[i] * s;
      }
      const double diag = cm * dot + vol * geo;
      for (int i = 0; i < 3; ++i)
        for (int k = 0; k < 3; ++k) {
          double v = cl * Na[i] * Nb[k] + cm * Na[k] * Nb[i];
          if (i == k) v += diag;
          K[3 * a + i][3 * b + k] += v;
        }
    }
}

// ShapeBase::calculateImplicitKViscous with ViscousKIntegral1/2,
// ShapeBase.cpp:2868-2933, 3097-3209.  Stored transposed: row b*3+j, col a*3+i.
void viscous_k_gp(const GPState& S, double eta_visc, const double vel[6][3], double dt,
                  double K[18][18]) {
  double L[3][3] = {{0}};
  for (int c = 0; c < 6; ++c)
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) L[i][j] += vel[c][i] * S.dNdx[c][j];
  double P[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) P[i][j] = ((i == j) ? 1.0 : 0.0) * (1.0 / dt) - L[i][j];
  for (int a = 0; a < 6; ++a)
    for (int b = 0; b < 6; ++b) {
      double BaT[3][6], Bb[6][3];
      for (int i = 0; i < 6; ++i)
        for (int j = 0; j < 3; ++j) {
          BaT[j][i] = S.B[i][3 * a + j];
          Bb[i][j] = S.B[i][3 * b + j];
        }
      for (int i = 3; i < 6; ++i)
        for (int j = 0; j < 3; ++j) Bb[i][j] = 0.5 * Bb[i][j];
      double BaTBb[3][3];
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
          double s = 0.0;
          for (int k = 0; k < 6; ++k) s += BaT[i][k] * Bb[k][j];
          BaTBb[i][j] = s;
        }
      double K1[3][3];
      matmul3(BaTBb, P, K1);
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
          double v = K1[i][j] * S.detF;
          v *= S.detdXde;
          v *= eta_visc;
          K[3 * b + j][3 * a + i] += v;
        }
    }
  for (int a = 0; a < 6; ++a)
    for (int b = 0; b < 6; ++b) {
      double paren[3][3];
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
          paren[i][j] = S.dNdx[a][i] * S.dNdx[b][j] - S.dNdx[b][i] * S.dNdx[a][j];
      double K2[3][3];
      matmul3(S.sigv, paren, K2);
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
          double v = K2[i][j] * S.detF;
          v *= S.detdXde;
          K[3 * b + j][3 * a + i] += v;
        }
    }
}

struct ElementOut {
  double ge[18];     // ElementalElasticSystemForces (unmasked), = -sum_gp w g_gp
  double gv[18];     // ElementalInternalViscousSystemForces (unmasked)
  double Ke[18][18]; // TriPointKe
  double Kv[18][18]; // TriPointKv (reference storage orientation)
};

// ShapeBase::calculateForces3D (ShapeBase.cpp:2389-2445) +
// calculateImplicitKElastic (2846-2866) + calculateImplicitKViscous (2868-2933).
// tangent_mode 0: faithful 8-deep loop; 1: closed form; 2: forces only.
void element_eval(const double pos[6][3], const double ref[6][3], const double invFg[3][3],
                  double lambda, double mu, double eta_visc, const double vel[6][3], double dt,
                  int tangent_mode, ElementOut& out) {
  PrismGeom G;
  prism_geometry(ref, G);
  double tri_ge[18] = {0}, tri_gv[18] = {0};
  std::memset(out.Ke, 0, sizeof(out.Ke));
  std::memset(out.Kv, 0, sizeof(out.Kv));
  GPState S[6];
  for (int g = 0; g < 6; ++g) {
    double cge[18], cgv[18];
    gauss_point_forces(pos, G, g, invFg, lambda, mu, eta_visc, vel, S[g], cge, cgv);
    for (int r = 0; r < 18; ++r) {
      tri_ge[r] = tri_ge[r] + cge[r] * kGauss.w[g];
      tri_gv[r] = tri_gv[r] + cgv[r] * kGauss.w[g];
    }
  }
  for (int r = 0; r < 18; ++r) {
    out.ge[r] = 0.0 - tri_ge[r];
    out.gv[r] = 0.0 - tri_gv[r];
  }
  if (tangent_mode == 2) return;
  for (int g = 0; g < 6; ++g) {
    double curr[18][18];
    std::memset(curr, 0, sizeof(curr));
    if (tangent_mode == 0) {
      elastic_k_integral1(S[g], curr);
      elastic_k_integral2(S[g], curr);
    } else {
      elastic_k_closed(S[g], lambda, mu, curr);
    }
    for (int r = 0; r < 18; ++r)
      for (int c = 0; c < 18; ++c) out.Ke[r][c] += curr[r][c] * kGauss.w[g];
  }
  if (eta_visc != 0) {
    for (int g = 0; g < 6; ++g) {
      double curr[18][18];
      std::memset(curr, 0, sizeof(curr));
      viscous_k_gp(S[g], eta_visc, vel, dt, curr);
      for (int r = 0; r < 18; ++r)
        for (int c = 0; c < 18; ++c) out.Kv[r][c] += curr[r][c] * kGauss.w[g];
    }
  }
}

// Node-node packing state of one step (Simulation::pacingNodeCouples0/1,
// Node::mass, Simulation::packingThreshold), set by orc_set_packing.  While
// set, assemble_dense adds the packing forces and their Jacobian blocks.
struct PackState {
  std::vector<int32_t> pairs;  // id0, id1 per couple, in detection order
  std::vector<double> mass;    // per node
  double threshold = 0.0;
  bool on = false;
};
PackState g_pack;

// Laser ablation (ShapeBase::IsAblated per prism, Simulation::AblatedNodes),
// set by orc_set_ablation.  While set, an ablated prism contributes neither
// forces (NewtonRaphsonSolver.cpp:236, 344) nor tangents (its TriPointKe and
// TriPointKv are zero, ShapeBase.cpp:2850-2852, 2879-2882) nor viscosity
// surfaces (Simulation.cpp:7608-7621), and an ablated node whose Jacobian
// diagonal entry is exactly zero after the drag and packing terms gets 1 there
// (checkJacobianForAblatedNodes, NewtonRaphsonSolver.cpp:425-439, called at
// Simulation.cpp:6691 before the fixed-DOF pass).
struct AblationState {
  std::vector<uint8_t> el;     // per prism
  std::vector<int32_t> nodes;  // ablated node ids, in the reference's list order
  bool on = false;
};
AblationState g_abl;

bool ablated_el(int e) { return g_abl.on && !g_abl.el.empty() && g_abl.el[e]; }

// checkJacobianForAblatedNodes on a dense K
void ablated_node_diagonal(double* K, int n) {
  if (!g_abl.on) return;
  for (int nd : g_abl.nodes)
    for (int aa = 0; aa < 3; ++aa) {
      double& d = K[(size_t)(3 * nd + aa) * n + (3 * nd + aa)];
      if (d == 0) d = 1;
    }
}

// One couple of Simulation::calculatePackingForcesImplicit3D
// (Simulation.cpp:7338-7400) and calculatePackingJacobian3D
// (Simulation.cpp:7444-7490): force F on id0 (-F on id1) and the symmetric
// 3x3 derivative dF0/dx0 = m [f/d (I - n n^T) + f' n n^T], in the
// reference's operation order (multiplier 5000, sigmoid saturation 5,
// shift 0.6 threshold, zero-distance guard 1e-3; detectPacingNodes,
// Simulation.cpp:8018-8020).
void packing_couple(const double* p0, const double* p1, double m0, double m1, double thr,
                    double F[3], double D[9]) {
  const double zeroThreshold = 1E-3;
  const double multiplier = 5000.0, sigmoidSaturation = 5.0;
  const double distanceShiftScale = 0.6 * thr;
  const double dx = p0[0] - p1[0];
  const double dy = p0[1] - p1[1];
  const double dz = p0[2] - p1[2];
  double d = std::pow((dx * dx + dy * dy + dz * dz), 0.5);
  const double averageMass = 0.5 * (m0 + m1);
  double normal[3] = {1.0, 0.0, 0.0};
  if (d > zeroThreshold) {
    normal[0] = dx / d;
    normal[1] = dy / d;
    normal[2] = dz / d;
  } else {
    d = zeroThreshold;
  }
  const double shifted_d = d - distanceShiftScale;
  const double sigmoid = 1 / (1 + std::exp(sigmoidSaturation / thr * 2.0 * shifted_d));
  const double Fmagnitude = multiplier * averageMass * sigmoid;
  for (int i = 0; i < 3; ++i) F[i] = Fmagnitude * normal[i];
  const double dSdXi = (-sigmoidSaturation / thr * 2.0) * sigmoid * (1 - sigmoid);
  const double scale = multiplier * averageMass;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      const double nn = normal[i] * normal[j];       // gsl_blas_dgemm(n, n^T)
      double v = (i == j ? 1.0 : 0.0) - nn;          // I - n n^T
      v = v * (sigmoid / d);                         // gsl_matrix_scale
      v = v + nn * dSdXi;                            // + n n^T dS/dd
      D[3 * i + j] = v * scale;                      // * multiplier * mass
    }
}

struct Problem {
  int n_nodes, n_el;
  const int32_t* conn;
  const double* refpos;
  const double* invFg;
  const double* lambda;
  const double* mu;
  const double* eta;
  const double* ext_visc;
  const double* visc_surf;
  const uint8_t* fixed;
  const int32_t* pairs;
  int n_pairs;
};

void gather_element(const Problem& P, int e, const double* uk, const double* un, double dt,
                    double pos[6][3], double ref[6][3], double iFg[3][3], double vel[6][3]) {
  const double rdt = 1.0 / dt;
  for (int a = 0; a < 6; ++a) {
    const int nd = P.conn[6 * e + a];
    for (int i = 0; i < 3; ++i) {
      pos[a][i] = uk[3 * nd + i];
      ref[a][i] = P.refpos[18 * e + 3 * a + i];
      // displacementPerDt = (uk - un) * (1/dt)  (NewtonRaphsonSolver.cpp:119-126)
      vel[a][i] = (uk[3 * nd + i] - un[3 * nd + i]) * rdt;
    }
  }
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) iFg[i][j] = P.invFg[9 * e + 3 * i + j];
}

// Dense assembly of one NR iteration, following Simulation::updateStepNR
// (Simulation.cpp:6643-6709): element loop (NewtonRaphsonSolver.cpp:221-241),
// force scatter (342-348 -> ShapeBase.cpp:2362-2387), Kv then Ke scatter
// (350-360 -> ShapeBase.cpp:2935-2983), external viscous forces and diagonal
// (362-423), sum (441-457), fixed DOF (195-219) and slave/master binding (128-183).
// K is row-major n x n in the reference orientation.
void assemble_dense(const Problem& P, const double* uk, const double* un, double dt,
                    int tangent_mode, double* K, double* gsum, double* ge_out) {
  const int n = 3 * P.n_nodes;
  std::vector<ElementOut> outs(P.n_el);
#pragma omp parallel for schedule(static)
  for (int e = 0; e < P.n_el; ++e) {
    double pos[6][3], ref[6][3], iFg[3][3], vel[6][3];
    if (ablated_el(e)) {
      outs[e] = ElementOut();
      continue;
    }
    gather_element(P, e, uk, un, dt, pos, ref, iFg, vel);
    element_eval(pos, ref, iFg, P.lambda[e], P.mu[e], P.eta[e], vel, dt, tangent_mode, outs[e]);
  }
  std::vector<double> ge(n, 0.0), gvi(n, 0.0), gve(n, 0.0);
  std::memset(K, 0, sizeof(double) * (size_t)n * n);
  for (int e = 0; e < P.n_el; ++e) {
    if (ablated_el(e)) continue;  // writeInternalForcesTogeAndgv is skipped
    for (int a = 0; a < 6; ++a) {
      const int nd = P.conn[6 * e + a];
      for (int i = 0; i < 3; ++i) {
        // calculateForces3D writes only non-fixed DOFs (ShapeBase.cpp:2407-2438)
        const bool fx = P.fixed && P.fixed[3 * nd + i];
        const double fe = fx ? 0.0 : outs[e].ge[3 * a + i];
        const double fv = fx ? 0.0 : outs[e].gv[3 * a + i];
        ge[3 * nd + i] = ge[3 * nd + i] + fe;
        gvi[3 * nd + i] = gvi[3 * nd + i] + fv;
      }
    }
  }
  for (int pass = 0; pass < 2; ++pass) {
    for (int e = 0; e < P.n_el; ++e) {
      if (pass == 0 && P.eta[e] == 0) continue;  // writeKviscousToMainKatrix guard
      const double(*Kel)[18] = pass == 0 ? outs[e].Kv : outs[e].Ke;
      for (int a = 0; a < 6; ++a)
        for (int b = 0; b < 6; ++b) {
          const int r0 = 3 * P.conn[6 * e + a], c0 = 3 * P.conn[6 * e + b];
          for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
              double& dst = K[(size_t)(r0 + i) * n + (c0 + j)];
              dst = dst + Kel[3 * a + i][3 * b + j];
            }
        }
    }
  }
  const double rdt = 1.0 / dt;
  for (int i = 0; i < P.n_nodes; ++i)
    for (int j = 0; j < 3; ++j) {
      const double sv = P.visc_surf[i] * P.ext_visc[3 * i + j];
      const double disp = (uk[3 * i + j] - un[3 * i + j]) * rdt;
      gve[3 * i + j] = -(sv * disp);
    }
  for (int i = 0; i < P.n_nodes; ++i) {
    const double spdt = P.visc_surf[i] / dt;
    for (int j = 0; j < 3; ++j) {
      double& d = K[(size_t)(3 * i + j) * n + (3 * i + j)];
      d = spdt * P.ext_visc[3 * i + j] + d;
    }
  }
  // packing between node couples (Simulation.cpp:6680-6681, 6698): forces
  // accumulated couple by couple into PackingForces, Jacobian blocks added
  // after the external-viscosity diagonal and before the fixed-DOF pass
  std::vector<double> pf(n, 0.0);
  if (g_pack.on) {
    const double thr = g_pack.threshold;
    for (size_t c = 0; c + 1 < g_pack.pairs.size(); c += 2) {
      const int id0 = g_pack.pairs[c], id1 = g_pack.pairs[c + 1];
      double F[3], D[9];
      packing_couple(uk + 3 * id0, uk + 3 * id1, g_pack.mass[id0], g_pack.mass[id1], thr, F, D);
      for (int i = 0; i < 3; ++i) {
        pf[3 * id0 + i] += F[i];
        pf[3 * id1 + i] -= F[i];
      }
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
          const double v = D[3 * i + j];
          double& a = K[(size_t)(3 * id0 + i) * n + (3 * id0 + j)];
          a = -v + a;
          double& b = K[(size_t)(3 * id0 + i) * n + (3 * id1 + j)];
          b = v + b;
          double& c2 = K[(size_t)(3 * id1 + i) * n + (3 * id1 + j)];
          c2 = -v + c2;
          double& d2 = K[(size_t)(3 * id1 + i) * n + (3 * id0 + j)];
          d2 = v + d2;
        }
    }
  }
  ablated_node_diagonal(K, n);
  for (int i = 0; i < n; ++i) {
    gsum[i] = ge[i] + gvi[i] + gve[i];
    // addPackingForces puts PackingForces + 0 into gExt (Simulation.cpp:8130-8151),
    // addExernalForces adds gExt (NewtonRaphsonSolver.cpp:447-457)
    gsum[i] = gsum[i] + (pf[i] + 0.0);
  }
  if (ge_out)
    for (int i = 0; i < n; ++i) ge_out[i] = ge[i] + gvi[i];
  if (P.fixed) {
    for (int idx = 0; idx < n; ++idx) {
      if (!P.fixed[idx]) continue;
      gsum[idx] = 0.0;
      for (int k = 0; k < n; ++k) {
        const double value = (idx == k) ? 1.0 : 0.0;
        K[(size_t)idx * n + k] = value;
        K[(size_t)k * n + idx] = value;
      }
    }
  }
  for (int p = 0; p < P.n_pairs; ++p) {
    const int s = P.pairs[2 * p], m = P.pairs[2 * p + 1];
    gsum[m] = gsum[m] + gsum[s];
    gsum[s] = 0;
    for (int k = 0; k < n; ++k) {
      K[(size_t)m * n + k] += K[(size_t)s * n + k];
      K[(size_t)s * n + k] = 0;
    }
    for (int k = 0; k < n; ++k) {
      K[(size_t)k * n + m] += K[(size_t)k * n + s];
      K[(size_t)k * n + s] = 0;
    }
    K[(size_t)s * n + s] = 1;
  }
}

// constructiaForPardiso, NewtonRaphsonSolver.cpp:766-787.
int csr_threshold(const double* K, int n, int32_t* ia, int32_t* ja, double* a, int cap) {
  const double neg = -1E-13, posv = 1E-13;
  int counter = 0;
  for (int i = 0; i < n; ++i) {
    bool wrote = false;
    for (int j = 0; j < n; ++j) {
      const double v = K[(size_t)i * n + j];
      if (v > posv || v < neg) {
        if (ja && counter < cap) { ja[counter] = j; a[counter] = v; }
        if (!wrote) { if (ia) ia[i] = counter; wrote = true; }
        counter++;
      }
    }
    if (!wrote && ia) ia[i] = counter;  // the reference leaves such rows unset
  }
  if (ia) ia[n] = counter;
  return counter;
}

// Dense LU with partial pivoting solving K^T x = b: the Pardiso stand-in
// (solveWithPardiso with iparm[11] = 1, NewtonRaphsonSolver.cpp:716-723).
// The matrix is the thresholded CSR of constructiaForPardiso expanded back
// to dense, so dropped entries are exactly zero as in the Pardiso input.
int dense_solve_transposed(const double* K, int n, const double* b, double* x) {
  std::vector<double> A((size_t)n * n, 0.0);
  const double neg = -1E-13, posv = 1E-13;
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      const double v = K[(size_t)i * n + j];
      if (v > posv || v < neg) A[(size_t)j * n + i] = v;  // transpose
    }
  std::vector<int> perm(n);
  for (int i = 0; i < n; ++i) perm[i] = i;
  for (int j = 0; j < n; ++j) {
    int piv = j;
    double mx = std::fabs(A[(size_t)j * n + j]);
    for (int i = j + 1; i < n; ++i) {
      const double v = std::fabs(A[(size_t)i * n + j]);
      if (v > mx) { mx = v; piv = i; }
    }
    if (mx == 0.0) return 1;
    if (piv != j) {
      for (int k = 0; k < n; ++k) std::swap(A[(size_t)j * n + k], A[(size_t)piv * n + k]);
      std::swap(perm[j], perm[piv]);
    }
    const double d = A[(size_t)j * n + j];
    for (int i = j + 1; i < n; ++i) {
      double& lij = A[(size_t)i * n + j];
      if (lij == 0.0) continue;
      lij /= d;
      const double l = lij;
      double* ri = &A[(size_t)i * n];
      const double* rj = &A[(size_t)j * n];
      for (int k = j + 1; k < n; ++k) ri[k] -= l * rj[k];
    }
  }
  std::vector<double> y(n);
  for (int i = 0; i < n; ++i) {
    double s = b[perm[i]];
    const double* ri = &A[(size_t)i * n];
    for (int k = 0; k < i; ++k) s -= ri[k] * y[k];
    y[i] = s;
  }
  for (int i = n - 1; i >= 0; --i) {
    double s = y[i];
    const double* ri = &A[(size_t)i * n];
    for (int k = i + 1; k < n; ++k) s -= ri[k] * x[k];
    x[i] = s / ri[i];
  }
  return 0;
}

// gsl_blas_dnrm2 (used at NewtonRaphsonSolver.cpp:815): BLAS dnrm2's scaled
// sum of squares, restated from the reference BLAS algorithm.
double dnrm2(const double* x, int n) {
  double scale = 0.0, ssq = 1.0;
  for (int i = 0; i < n; ++i) {
    if (x[i] != 0.0) {
      const double ax = std::fabs(x[i]);
      if (scale < ax) {
        ssq = 1.0 + ssq * (scale / ax) * (scale / ax);
        scale = ax;
      } else {
        ssq += (ax / scale) * (ax / scale);
      }
    }
  }
  return scale * std::sqrt(ssq);
}

// ---------------------------------------------------------------------------
// Sparse variant (block-CSR of 3x3 node blocks + block-Jacobi BiCGStab) used
// where the reference's dense K cannot be allocated.  Assembly order per entry
// is the reference's: all Kv contributions in element order, then all Ke.
struct Bsr {
  int nb;                       // block rows (nodes)
  std::vector<int> rowptr, col; // block pattern, ascending columns
  std::vector<double> val;      // 9 doubles per block, row-major, J = K^T orientation
};

void build_pattern(const Problem& P, Bsr& M) {
  const int N = P.n_nodes;
  std::vector<std::vector<int>> nb(N);
  for (int e = 0; e < P.n_el; ++e)
    for (int a = 0; a < 6; ++a)
      for (int b = 0; b < 6; ++b) nb[P.conn[6 * e + a]].push_back(P.conn[6 * e + b]);
  M.nb = N;
  M.rowptr.assign(N + 1, 0);
  for (int i = 0; i < N; ++i) {
    nb[i].push_back(i);
    std::sort(nb[i].begin(), nb[i].end());
    nb[i].erase(std::unique(nb[i].begin(), nb[i].end()), nb[i].end());
    M.rowptr[i + 1] = M.rowptr[i] + (int)nb[i].size();
  }
  M.col.resize(M.rowptr[N]);
  for (int i = 0; i < N; ++i)
    std::copy(nb[i].begin(), nb[i].end(), M.col.begin() + M.rowptr[i]);
  M.val.assign((size_t)9 * M.rowptr[N], 0.0);
}

int find_block(const Bsr& M, int r, int c) {
  const int* b = M.col.data() + M.rowptr[r];
  const int* e = M.col.data() + M.rowptr[r + 1];
  const int* p = std::lower_bound(b, e, c);
  return (p != e && *p == c) ? (int)(p - M.col.data()) : -1;
}

// Assembles J = K^T (the operator of the transposed solve) into the BSR.
// Fixed DOFs get the zero row/column and unit diagonal of calcutateFixedK.
void assemble_bsr(const Problem& P, const double* uk, const double* un, double dt, Bsr& M,
                  double* gsum) {
  const int n = 3 * P.n_nodes;
  std::vector<ElementOut> outs(P.n_el);
#pragma omp parallel for schedule(static)
  for (int e = 0; e < P.n_el; ++e) {
    double pos[6][3], ref[6][3], iFg[3][3], vel[6][3];
    if (ablated_el(e)) {
      outs[e] = ElementOut();
      continue;
    }
    gather_element(P, e, uk, un, dt, pos, ref, iFg, vel);
    element_eval(pos, ref, iFg, P.lambda[e], P.mu[e], P.eta[e], vel, dt, 1, outs[e]);
  }
  std::fill(M.val.begin(), M.val.end(), 0.0);
  std::vector<double> ge(n, 0.0), gvi(n, 0.0);
  for (int e = 0; e < P.n_el; ++e)
    for (int a = 0; a < 6; ++a) {
      if (ablated_el(e)) break;
      const int nd = P.conn[6 * e + a];
      for (int i = 0; i < 3; ++i) {
        const bool fx = P.fixed && P.fixed[3 * nd + i];
        ge[3 * nd + i] += fx ? 0.0 : outs[e].ge[3 * a + i];
        gvi[3 * nd + i] += fx ? 0.0 : outs[e].gv[3 * a + i];
      }
    }
  for (int pass = 0; pass < 2; ++pass)
    for (int e = 0; e < P.n_el; ++e) {
      if (pass == 0 && P.eta[e] == 0) continue;
      const double(*Kel)[18] = pass == 0 ? outs[e].Kv : outs[e].Ke;
      for (int a = 0; a < 6; ++a)
        for (int b = 0; b < 6; ++b) {
          // K block (node_a, node_b) -> J block (node_b, node_a) transposed
          const int blk = find_block(M, P.conn[6 * e + b], P.conn[6 * e + a]);
          double* dst = &M.val[(size_t)9 * blk];
          for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) dst[3 * j + i] += Kel[3 * a + i][3 * b + j];
        }
    }
  const double rdt = 1.0 / dt;
  for (int i = 0; i < P.n_nodes; ++i) {
    const int d = find_block(M, i, i);
    const double spdt = P.visc_surf[i] / dt;
    for (int j = 0; j < 3; ++j) {
      double& v = M.val[(size_t)9 * d + 4 * j];
      v = spdt * P.ext_visc[3 * i + j] + v;
      const double sv = P.visc_surf[i] * P.ext_visc[3 * i + j];
      const double disp = (uk[3 * i + j] - un[3 * i + j]) * rdt;
      gsum[3 * i + j] = ge[3 * i + j] + gvi[3 * i + j] + (-(sv * disp));
    }
  }
  if (g_abl.on)
    for (int nd : g_abl.nodes) {
      const int d = find_block(M, nd, nd);
      for (int aa = 0; aa < 3; ++aa) {
        double& v = M.val[(size_t)9 * d + 4 * aa];
        if (v == 0) v = 1;
      }
    }
  if (P.fixed) {
    for (int r = 0; r < P.n_nodes; ++r)
      for (int k = M.rowptr[r]; k < M.rowptr[r + 1]; ++k) {
        const int c = M.col[k];
        for (int i = 0; i < 3; ++i)
          for (int j = 0; j < 3; ++j) {
            const bool fr = P.fixed[3 * r + i], fc = P.fixed[3 * c + j];
            if (fr || fc) M.val[(size_t)9 * k + 3 * i + j] = (r == c && i == j) ? 1.0 : 0.0;
          }
      }
    for (int i = 0; i < n; ++i)
      if (P.fixed[i]) gsum[i] = 0.0;
  }
}

void bsr_spmv(const Bsr& M, const double* x, double* y) {
#pragma omp parallel for schedule(static)
  for (int r = 0; r < M.nb; ++r) {
    double a0 = 0, a1 = 0, a2 = 0;
    for (int k = M.rowptr[r]; k < M.rowptr[r + 1]; ++k) {
      const double* v = &M.val[(size_t)9 * k];
      const double* xc = x + 3 * M.col[k];
      a0 += v[0] * xc[0] + v[1] * xc[1] + v[2] * xc[2];
      a1 += v[3] * xc[0] + v[4] * xc[1] + v[5] * xc[2];
      a2 += v[6] * xc[0] + v[7] * xc[1] + v[8] * xc[2];
    }
    y[3 * r] = a0; y[3 * r + 1] = a1; y[3 * r + 2] = a2;
  }
}

double dot(const std::vector<double>& a, const std::vector<double>& b) {
  double s = 0.0;
  const int n = (int)a.size();
#pragma omp parallel for reduction(+ : s) schedule(static)
  for (int i = 0; i < n; ++i) s += a[i] * b[i];
  return s;
}

// Matrix-free binding operator N^T J N + I_slave applied around the BSR SpMV
// (the same operator calculateBoundKWithSlavesMasterDoF forms explicitly,
// NewtonRaphsonSolver.cpp:128-183).
struct Op {
  const Bsr& M;
  const int32_t* pairs;
  int n_pairs;
  mutable std::vector<double> tmp_in, tmp_out;
  void apply(const std::vector<double>& x, std::vector<double>& y) const {
    if (n_pairs == 0) { bsr_spmv(M, x.data(), y.data()); return; }
    tmp_in = x;
    for (int p = 0; p < n_pairs; ++p) tmp_in[pairs[2 * p]] = x[pairs[2 * p + 1]];
    tmp_out.resize(x.size());
    bsr_spmv(M, tmp_in.data(), tmp_out.data());
    y = tmp_out;
    for (int p = 0; p < n_pairs; ++p) {
      y[pairs[2 * p + 1]] += tmp_out[pairs[2 * p]];
    }
    for (int p = 0; p < n_pairs; ++p) y[pairs[2 * p]] = x[pairs[2 * p]];
  }
};

// Right-preconditioned BiCGStab with 3x3 node-block Jacobi.
int bicgstab(const Op& A, const std::vector<double>& Minv, const std::vector<double>& b,
             std::vector<double>& x, double rtol, int maxit, int* iters) {
  const int n = (int)b.size();
  const int N = n / 3;
  auto prec = [&](const std::vector<double>& in, std::vector<double>& out) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < N; ++i) {
      const double* m = &Minv[9 * (size_t)i];
      const double* v = &in[3 * (size_t)i];
      out[3 * i] = m[0] * v[0] + m[1] * v[1] + m[2] * v[2];
      out[3 * i + 1] = m[3] * v[0] + m[4] * v[1] + m[5] * v[2];
      out[3 * i + 2] = m[6] * v[0] + m[7] * v[1] + m[8] * v[2];
    }
  };
  std::vector<double> r(n), rh(n), p(n, 0.0), v(n, 0.0), ph(n), s(n), sh(n), t(n);
  std::fill(x.begin(), x.end(), 0.0);
  r = b;
  rh = r;
  const double bnorm = std::sqrt(dot(b, b));
  *iters = 0;
  if (bnorm == 0.0) return 0;
  double rho = 1, alpha = 1, omega = 1;
  for (int it = 0; it < maxit; ++it) {
    const double rho_new = dot(rh, r);
    if (rho_new == 0.0) return 3;
    const double beta = (rho_new / rho) * (alpha / omega);
    for (int i = 0; i < n; ++i) p[i] = r[i] + beta * (p[i] - omega * v[i]);
    prec(p, ph);
    A.apply(ph, v);
    alpha = rho_new / dot(rh, v);
    for (int i = 0; i < n; ++i) s[i] = r[i] - alpha * v[i];
    prec(s, sh);
    A.apply(sh, t);
    const double tt = dot(t, t);
    omega = tt > 0 ? dot(t, s) / tt : 0.0;
    for (int i = 0; i < n; ++i) {
      x[i] += alpha * ph[i] + omega * sh[i];
      r[i] = s[i] - omega * t[i];
    }
    rho = rho_new;
    *iters = it + 1;
    if (std::sqrt(dot(r, r)) <= rtol * bnorm) return 0;
    if (omega == 0.0) return 3;
  }
  return 4;
}

void block_jacobi(const Bsr& M, const int32_t* pairs, int n_pairs, std::vector<double>& Minv) {
  Minv.assign((size_t)9 * M.nb, 0.0);
  std::vector<char> slave(3 * (size_t)M.nb, 0);
  for (int p = 0; p < n_pairs; ++p) slave[pairs[2 * p]] = 1;
  for (int i = 0; i < M.nb; ++i) {
    const int d = find_block(M, i, i);
    double A[3][3], X[3][3];
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) {
        A[r][c] = M.val[(size_t)9 * d + 3 * r + c];
        if (slave[3 * i + r] || slave[3 * i + c]) A[r][c] = (r == c) ? 1.0 : 0.0;
      }
    inv3(A, X);
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) Minv[(size_t)9 * i + 3 * r + c] = X[r][c];
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI of the oracle (consumed by tests/ and bench.py's cpu_baseline through
// ctypes).  Element arrays: conn n_el x 6 (int32), refpos n_el x 18, invFg
// n_el x 9 row-major, positions 3 x n_nodes interleaved (DOF = 3*node + dim).
extern "C" {

int orc_num_threads() {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

// Prism::setElasticProperties / updateElasticProperties (Prism.cpp:512-513, 526-527).
void orc_lame(int n, const double* E, const double* nu, double* lambda, double* mu) {
  for (int i = 0; i < n; ++i) {
    const double v = nu[i];
    lambda[i] = E[i] * v / (1 + v) / (1 - 2.0 * v);
    mu[i] = E[i] / 2.0 / (1 + v);
  }
}

// Prism geometry at the 6 Gauss points for one element: InvdXde (6x9), detdXde (6).
void orc_prism_geometry(const double* refpos18, double* invdXde54, double* detdXde6) {
  double ref[6][3];
  for (int a = 0; a < 6; ++a)
    for (int i = 0; i < 3; ++i) ref[a][i] = refpos18[3 * a + i];
  PrismGeom G;
  prism_geometry(ref, G);
  for (int g = 0; g < 6; ++g) {
    detdXde6[g] = G.detdXde[g];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) invdXde54[9 * g + 3 * i + j] = G.InvdXde[g][i][j];
  }
}

// ShapeBase::Strain after ShapeBase::calculateForces3D: the loop over the six
// Gauss points (ShapeBase.cpp:2407-2416) leaves the strain of the last one,
// E = (C - I) / 2 with C = Fe^T Fe (Prism.cpp:782-799, 826-837;
// ShapeBase.cpp:2448-2472), in the Voigt order E00, E11, E22, 2E01, 2E21, 2E02.
// pos and refpos: 18 per prism; invFg: 9 per prism; strain: 6 per prism;
// gp: the Gauss point (5 for the reference's output, others for tests that
// show the saved strain tells the points apart).
void orc_element_strain(int n_el, const double* pos, const double* refpos, const double* invFg,
                        int gp, double* strain) {
#pragma omp parallel for schedule(static)
  for (int e = 0; e < n_el; ++e) {
    double p[6][3], r[6][3], f[3][3];
    for (int a = 0; a < 6; ++a)
      for (int i = 0; i < 3; ++i) {
        p[a][i] = pos[18 * e + 3 * a + i];
        r[a][i] = refpos[18 * e + 3 * a + i];
      }
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) f[i][j] = invFg[9 * e + 3 * i + j];
    PrismGeom G;
    prism_geometry(r, G);
    const int g = gp;
    double Jac[3][3], F[3][3], Fe[3][3], C[3][3], E[3][3];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        double s = 0.0;
        for (int a = 0; a < 6; ++a) s += G.SFD[g][j][a] * p[a][i];
        Jac[i][j] = s;
      }
    matmul3(Jac, G.InvdXde[g], F);
    matmul3(F, f, Fe);
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        double s = 0.0;
        for (int k = 0; k < 3; ++k) s += Fe[k][i] * Fe[k][j];
        C[i][j] = s;
      }
    // calculateEForNodalForcesKirshoff: copy C, subtract I, scale by 0.5
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) E[i][j] = (C[i][j] - (i == j ? 1.0 : 0.0)) * 0.5;
    double* out = strain + 6 * (size_t)e;
    out[0] = E[0][0];
    out[1] = E[1][1];
    out[2] = E[2][2];
    out[3] = 2.0 * E[0][1];
    out[4] = 2.0 * E[2][1];
    out[5] = 2.0 * E[0][2];
  }
}

void orc_inv3(const double* A9, double* X9) {
  double A[3][3], X[3][3];
  std::memcpy(A, A9, sizeof(A));
  inv3(A, X);
  std::memcpy(X9, X, sizeof(X));
}

// Element forces and tangents for n_el elements whose node positions are given
// element-wise (pos n_el x 18, vel n_el x 18).  Outputs per element:
// ge/gv (18), Ke/Kv (324, reference storage orientation).
void orc_elements(int n_el, const double* pos, const double* refpos, const double* invFg,
                  const double* lambda, const double* mu, const double* eta, const double* vel,
                  double dt, int tangent_mode, double* ge, double* gv, double* Ke, double* Kv) {
#pragma omp parallel for schedule(static)
  for (int e = 0; e < n_el; ++e) {
    double p[6][3], r[6][3], f[3][3], v[6][3];
    for (int a = 0; a < 6; ++a)
      for (int i = 0; i < 3; ++i) {
        p[a][i] = pos[18 * e + 3 * a + i];
        r[a][i] = refpos[18 * e + 3 * a + i];
        v[a][i] = vel ? vel[18 * e + 3 * a + i] : 0.0;
      }
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) f[i][j] = invFg[9 * e + 3 * i + j];
    ElementOut out;
    element_eval(p, r, f, lambda[e], mu[e], eta[e], v, dt, tangent_mode, out);
    for (int k = 0; k < 18; ++k) {
      if (ge) ge[18 * e + k] = out.ge[k];
      if (gv) gv[18 * e + k] = out.gv[k];
    }
    if (Ke) std::memcpy(Ke + (size_t)324 * e, out.Ke, sizeof(out.Ke));
    if (Kv) std::memcpy(Kv + (size_t)324 * e, out.Kv, sizeof(out.Kv));
  }
}

// Dense assembly (reference K orientation) and gSum for one NR iteration.
// ge_int (optional): internal force sum ge + gvInternal before BCs, i.e. the
// SystemForces of Save_Force for non-fixed DOFs.
void orc_assemble_dense(int n_nodes, int n_el, const int32_t* conn, const double* refpos,
                        const double* invFg, const double* lambda, const double* mu,
                        const double* eta, const double* ext_visc, const double* visc_surf,
                        const uint8_t* fixed, const int32_t* pairs, int n_pairs,
                        const double* uk, const double* un, double dt, int tangent_mode,
                        double* K, double* gsum, double* ge_int) {
  Problem P{n_nodes, n_el, conn, refpos, invFg, lambda, mu, eta, ext_visc, visc_surf,
            fixed, pairs, n_pairs};
  assemble_dense(P, uk, un, dt, tangent_mode, K, gsum, ge_int);
}

int orc_csr_threshold(const double* K, int n, int32_t* ia, int32_t* ja, double* a, int cap) {
  return csr_threshold(K, n, ia, ja, a, cap);
}

int orc_dense_solve_transposed(const double* K, int n, const double* b, double* x) {
  return dense_solve_transposed(K, n, b, x);
}

double orc_dnrm2(const double* x, int n) { return dnrm2(x, n); }

// Lateral/top/bottom exposed surface areas assigned to nodes
// (Simulation::updateNodeViscositySurfaces, Simulation.cpp:7598-7621;
// Prism::calculateApicalArea/BasalArea, Prism.cpp:1485-1556;
// ShapeBase::assignViscositySurfaceAreaToNodes, ShapeBase.cpp:3764-3775;
// Prism::assignExposedSurfaceAreaIndices, Prism.cpp:1650-1709).
// elem_placement: tissuePlacement per element; elem_type: tissueType;
// spans: spansWholeTissue flag.
void orc_visc_surfaces(int n_nodes, int n_el, const int32_t* conn, const double* u,
                       const int32_t* elem_placement, const int32_t* elem_type,
                       const uint8_t* spans, double* A) {
  for (int i = 0; i < n_nodes; ++i) A[i] = 0.0;
  auto tri_area = [&](int e, int i0, int i1, int i2) {
    const double Threshold = 1E-5;
    double s1[3], s2[3], Side1 = 0, Side2 = 0, costet = 0, Area = 0;
    const int n0 = conn[6 * e + i0], n1 = conn[6 * e + i1], n2 = conn[6 * e + i2];
    for (int i = 0; i < 3; ++i) {
      s1[i] = u[3 * n1 + i] - u[3 * n0 + i];
      s2[i] = u[3 * n2 + i] - u[3 * n0 + i];
      costet += s1[i] * s2[i];
      Side1 += s1[i] * s1[i];
      Side2 += s2[i] * s2[i];
    }
    if (Side1 > Threshold && Side2 > Threshold) {
      Side1 = std::pow(Side1, 0.5);
      Side2 = std::pow(Side2, 0.5);
      costet /= (Side1 * Side2);
      const double sinTetSq = 1 - costet * costet;
      double sintet = 0;
      if (sinTetSq > Threshold) sintet = std::pow(sinTetSq, 0.5);
      Area = Side1 * Side2 * sintet / 2.0;
    }
    return Area;
  };
  for (int e = 0; e < n_el; ++e) {
    const int tt = elem_type[e], tp = elem_placement[e];
    if (tt != 0 && tt != 1) continue;
    if (ablated_el(e)) continue;
    int apical[3] = {3, 4, 5}, basal[3] = {0, 1, 2};
    if (tt == 1) { apical[0] = 0; apical[1] = 1; apical[2] = 2; basal[0] = 3; basal[1] = 4; basal[2] = 5; }
    const bool has_apical = (tp == 1) || (tp == 2 && spans[e]);
    const bool has_basal = (tp == 0) || (tp == 2 && spans[e]);
    if (has_apical) {
      // calculateApicalArea uses nodes 3,4,5 (0,1,2 for peripodial)
      const double area = tt == 1 ? tri_area(e, 0, 1, 2) : tri_area(e, 3, 4, 5);
      for (int k = 0; k < 3; ++k) A[conn[6 * e + apical[k]]] += area / 3;
    }
    if (has_basal) {
      const double area = tt == 1 ? tri_area(e, 3, 4, 5) : tri_area(e, 0, 1, 2);
      for (int k = 0; k < 3; ++k) A[conn[6 * e + basal[k]]] += area / 3;
    }
  }
}

// Full NR step with the dense Jacobian and the dense transposed LU as the
// Pardiso stand-in: Simulation::updateStepNR, Simulation.cpp:6605-6746.
// Returns the number of iterations; du_norms receives |du|_2 per iteration
// (capacity maxit + 1).  *converged is 0 when the iteration cap was hit (the
// reference prints "did not converge" and accepts the iterate).
int orc_nr_step_dense(int n_nodes, int n_el, const int32_t* conn, const double* refpos,
                      const double* invFg, const double* lambda, const double* mu,
                      const double* eta, const double* ext_visc, const double* visc_surf,
                      const uint8_t* fixed, const int32_t* pairs, int n_pairs,
                      const double* u_n, double dt, double tol, int maxit, int tangent_mode,
                      double* u_out, double* du_norms, int* converged) {
  Problem P{n_nodes, n_el, conn, refpos, invFg, lambda, mu, eta, ext_visc, visc_surf,
            fixed, pairs, n_pairs};
  const int n = 3 * n_nodes;
  std::vector<double> un(u_n, u_n + n), uk(un), K((size_t)n * n), g(n), du(n);
  int it = 0;
  bool conv = false, hit_cap = false;
  while (!conv) {
    assemble_dense(P, uk.data(), un.data(), dt, tangent_mode, K.data(), g.data(), nullptr);
    if (dense_solve_transposed(K.data(), n, g.data(), du.data()) != 0) {
      *converged = 0;
      return -1;
    }
    for (int p = 0; p < n_pairs; ++p) du[pairs[2 * p]] = du[pairs[2 * p + 1]];
    const double d = dnrm2(du.data(), n);
    du_norms[it] = d;
    conv = !(d > tol);
    for (int i = 0; i < n; ++i) uk[i] = uk[i] + du[i];
    it++;
    if (!conv && it > maxit) { conv = true; hit_cap = true; }
  }
  std::memcpy(u_out, uk.data(), sizeof(double) * n);
  *converged = hit_cap ? 0 : 1;
  return it;
}

// Same step on the block-sparse Jacobian with block-Jacobi BiCGStab (relative
// residual krylov_rtol).  krylov_iters (capacity maxit + 1) receives the
// BiCGStab iteration count of every NR iteration.
int orc_nr_step_sparse(int n_nodes, int n_el, const int32_t* conn, const double* refpos,
                       const double* invFg, const double* lambda, const double* mu,
                       const double* eta, const double* ext_visc, const double* visc_surf,
                       const uint8_t* fixed, const int32_t* pairs, int n_pairs,
                       const double* u_n, double dt, double tol, int maxit, double krylov_rtol,
                       int krylov_maxit, double* u_out, double* du_norms, int* krylov_iters,
                       int* converged) {
  Problem P{n_nodes, n_el, conn, refpos, invFg, lambda, mu, eta, ext_visc, visc_surf,
            fixed, pairs, n_pairs};
  const int n = 3 * n_nodes;
  Bsr M;
  build_pattern(P, M);
  std::vector<double> un(u_n, u_n + n), uk(un), g(n), du(n), Minv;
  Op A{M, pairs, n_pairs, {}, {}};
  int it = 0;
  bool conv = false, hit_cap = false;
  while (!conv) {
    assemble_bsr(P, uk.data(), un.data(), dt, M, g.data());
    for (int p = 0; p < n_pairs; ++p) {
      g[pairs[2 * p + 1]] = g[pairs[2 * p + 1]] + g[pairs[2 * p]];
      g[pairs[2 * p]] = 0;
    }
    block_jacobi(M, pairs, n_pairs, Minv);
    int kit = 0;
    const int st = bicgstab(A, Minv, g, du, krylov_rtol, krylov_maxit, &kit);
    krylov_iters[it] = kit;
    if (st == 3) { *converged = 0; return -1; }
    for (int p = 0; p < n_pairs; ++p) du[pairs[2 * p]] = du[pairs[2 * p + 1]];
    const double d = dnrm2(du.data(), n);
    du_norms[it] = d;
    conv = !(d > tol);
    for (int i = 0; i < n; ++i) uk[i] = uk[i] + du[i];
    it++;
    if (!conv && it > maxit) { conv = true; hit_cap = true; }
  }
  std::memcpy(u_out, uk.data(), sizeof(double) * n);
  *converged = hit_cap ? 0 : 1;
  return it;
}

// The sparse Jacobian J = K^T at (uk, un) and the residual gSum: block rows
// rowptr (n_nodes + 1), block columns col, 9 values per block (row-major).
// Called with col == nullptr it returns the number of blocks only.
long orc_assemble_bsr(int n_nodes, int n_el, const int32_t* conn, const double* refpos,
                      const double* invFg, const double* lambda, const double* mu,
                      const double* eta, const double* ext_visc, const double* visc_surf,
                      const uint8_t* fixed, const double* uk, const double* un, double dt,
                      int32_t* rowptr, int32_t* col, double* val, double* g) {
  Problem P{n_nodes, n_el, conn, refpos, invFg, lambda, mu, eta, ext_visc, visc_surf,
            fixed, nullptr, 0};
  Bsr M;
  build_pattern(P, M);
  if (!col) return (long)M.col.size();
  std::vector<double> gs(3 * (size_t)n_nodes);
  assemble_bsr(P, uk, un, dt, M, gs.data());
  std::copy(M.rowptr.begin(), M.rowptr.end(), rowptr);
  std::copy(M.col.begin(), M.col.end(), col);
  std::copy(M.val.begin(), M.val.end(), val);
  if (g) std::copy(gs.begin(), gs.end(), g);
  return (long)M.col.size();
}

// One NR iteration of the sparse path with a fixed Krylov iteration budget;
// the CPU-baseline timing unit.  Returns BiCGStab iterations run.
int orc_nr_iteration_sparse(int n_nodes, int n_el, const int32_t* conn, const double* refpos,
                            const double* invFg, const double* lambda, const double* mu,
                            const double* eta, const double* ext_visc, const double* visc_surf,
                            const uint8_t* fixed, const double* uk, const double* un, double dt,
                            double krylov_rtol, int krylov_maxit, double* du_out,
                            double* t_assembly_s, double* t_solve_s) {
  Problem P{n_nodes, n_el, conn, refpos, invFg, lambda, mu, eta, ext_visc, visc_surf,
            fixed, nullptr, 0};
  const int n = 3 * n_nodes;
  Bsr M;
  build_pattern(P, M);
  std::vector<double> g(n), du(n), Minv;
#ifdef _OPENMP
  const double t0 = omp_get_wtime();
#endif
  assemble_bsr(P, uk, un, dt, M, g.data());
  block_jacobi(M, nullptr, 0, Minv);
#ifdef _OPENMP
  const double t1 = omp_get_wtime();
#endif
  Op A{M, nullptr, 0, {}, {}};
  int kit = 0;
  bicgstab(A, Minv, g, du, krylov_rtol, krylov_maxit, &kit);
#ifdef _OPENMP
  const double t2 = omp_get_wtime();
  if (t_assembly_s) *t_assembly_s = t1 - t0;
  if (t_solve_s) *t_solve_s = t2 - t1;
#else
  if (t_assembly_s) *t_assembly_s = 0;
  if (t_solve_s) *t_solve_s = 0;
#endif
  if (du_out) std::memcpy(du_out, du.data(), sizeof(double) * n);
  return kit;
}


// ---------------------------------------------------------------------------
// Node-node packing: the pre-step detection of Simulation::runOneStep
// (Simulation.cpp:6065-6076) and the per-iteration couple terms.

// Sets (pairs != NULL) or clears the packing couples that assemble_dense and
// the dense NR step apply.  mass: Node::mass per node; threshold:
// Simulation::packingThreshold.
void orc_set_packing(const int32_t* pairs, int n_pairs, const double* mass, int n_nodes,
                     double threshold) {
  g_pack = PackState();
  if (!pairs) return;
  g_pack.pairs.assign(pairs, pairs + 2 * n_pairs);
  g_pack.mass.assign(mass, mass + n_nodes);
  g_pack.threshold = threshold;
  g_pack.on = true;
}

void orc_set_ablation(const uint8_t* ablated_el, int n_el, const int32_t* nodes, int n_nodes) {
  g_abl = AblationState();
  if (!ablated_el && !nodes) return;
  if (ablated_el) g_abl.el.assign(ablated_el, ablated_el + n_el);
  if (nodes) g_abl.nodes.assign(nodes, nodes + n_nodes);
  g_abl.on = true;
}

void orc_packing_couple(const double* p0, const double* p1, double m0, double m1, double thr,
                        double* F, double* D) {
  packing_couple(p0, p1, m0, m1, thr, F, D);
}

// Prism::calculateReferenceVolume (Prism.cpp:888-922) through
// ShapeBase::calculateVolumeForInputShapeStructure (ShapeBase.cpp:3364-3398):
// eight triangles (two bases, each rectangle split in two), each with the
// element centre as apex of a tetrahedron.
void orc_reference_volumes(int n_el, const double* refpos, double* vol) {
  static const int faces[8][3] = {{0, 1, 2}, {3, 4, 5}, {0, 1, 3}, {1, 3, 4},
                                  {1, 2, 4}, {2, 4, 5}, {0, 2, 3}, {2, 3, 5}};
  for (int e = 0; e < n_el; ++e) {
    const double* P = refpos + 18 * e;
    double centre[3] = {0.0, 0.0, 0.0};
    for (int i = 0; i < 6; ++i)
      for (int j = 0; j < 3; ++j) centre[j] += P[3 * i + j];
    for (int j = 0; j < 3; ++j) centre[j] /= 6;
    double total = 0;
    for (int f = 0; f < 8; ++f) {
      const double* a = P + 3 * faces[f][0];
      const double* b = P + 3 * faces[f][1];
      const double* c = P + 3 * faces[f][2];
      double v1[3], v2[3], vm[3], base[3];
      for (int j = 0; j < 3; ++j) {
        v1[j] = b[j] - a[j];
        v2[j] = c[j] - a[j];
        vm[j] = centre[j] - a[j];
      }
      base[0] = v1[1] * v2[2] - v1[2] * v2[1];
      base[1] = v1[2] * v2[0] - v1[0] * v2[2];
      base[2] = v1[0] * v2[1] - v1[1] * v2[0];
      const double nb = std::pow(base[0] * base[0] + base[1] * base[1] + base[2] * base[2], 0.5);
      const double area = nb / 2;
      double height = (vm[0] * base[0] + vm[1] * base[1] + vm[2] * base[2]) / nb;
      if (height < 0) height *= (-1.0);
      total += 1.0 / 3.0 * (height * area);
    }
    vol[e] = total;
  }
}

// Simulation::updateNodeMasses (Simulation.cpp:7587-7596) with
// ShapeBase::assignVolumesToNodes (ShapeBase.cpp:3749-3753): every element adds
// VolumePerNode = det(F_g) V_ref / 6 (ShapeBase.cpp:1497-1499) to its nodes,
// in element order.
void orc_node_masses(int n_nodes, int n_el, const int32_t* conn, const double* vol,
                     const double* Fg, double* mass) {
  for (int i = 0; i < n_nodes; ++i) mass[i] = 0;
  for (int e = 0; e < n_el; ++e) {
    if (ablated_el(e)) continue;  // Simulation::updateNodeMasses skips ablated prisms
    double M[3][3];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) M[i][j] = Fg[9 * e + 3 * i + j];
    const double grown = det3(M) * vol[e];
    const double per_node = grown / 6;
    for (int a = 0; a < 6; ++a) mass[conn[6 * e + a]] += per_node;
  }
}

// Simulation::calculateBoundingBox (Simulation.cpp:8388-8425) over the nodes
// (no explicit ECM), with its `else if`: a node that lowers the minimum of a
// coordinate is not tested against the maximum.  bb = {min x, y, z, max x, y, z}.
void orc_bounding_box(int n_nodes, const double* u, int symmetric_y, double* bb) {
  double lo[3] = {100000.0, 100000.0, 100000.0}, hi[3] = {-100000.0, -100000.0, -100000.0};
  for (int k = 0; k < n_nodes; ++k)
    for (int i = 0; i < 3; ++i) {
      const double v = u[3 * k + i];
      if (v < lo[i]) lo[i] = v;
      else if (v > hi[i]) hi[i] = v;
    }
  if (symmetric_y) lo[1] = (-1.0) * hi[1];
  for (int i = 0; i < 3; ++i) { bb[i] = lo[i]; bb[3 + i] = hi[i]; }
}

// Thresholds of Simulation::detectPacingNodes (Simulation.cpp:8002-8020) from
// Simulation::getAverageSideLength (Simulation.cpp:3197-3260), for a tissue of
// columnar elements (tissueType 0, no peripodial membrane):
//  - every non-ablated columnar element adds its apical side-length average
//    (Prism::getApicalSideLengthAverage, Prism.cpp:1329-1360) and, if basal or
//    spanning the whole tissue, its basal one (1362-1393) to the column mean;
//  - the element's last value goes to the 10 x 5 grid cell of its relative
//    position in the bounding box (ShapeBase::calculateRelativePosInBoundingBox,
//    ShapeBase.cpp:2091-2095, centre by getCentre, 247-257: cell
//    (floor(rx), floor(ry / 2)), clamped);
//  - packingThreshold = 0.4 * column mean; grid = 0.4 * cell means.
// grid50 receives the squared detection thresholds (packingDetectionThresholdGridSq).
void orc_packing_thresholds(int n_nodes, int n_el, const int32_t* conn, const double* u,
                            const int32_t* elem_placement, const int32_t* elem_type,
                            const uint8_t* spans, int symmetric_y, double* grid50,
                            double* packing_threshold, double* bb) {
  orc_bounding_box(n_nodes, u, symmetric_y, bb);
  const double size[2] = {bb[3] - bb[0], bb[4] - bb[1]};
  double sum[10][5] = {}, cnt[10][5] = {};
  double dsumCol = 0.0;
  int colCounter = 0;
  auto side_avg = [&](int e, const int pr[3][2]) {
    double dsum = 0.0;
    int counter = 0;
    for (int i = 0; i < 3; ++i) {
      const int a = conn[6 * e + pr[i][0]], b = conn[6 * e + pr[i][1]];
      const double dx = u[3 * a] - u[3 * b];
      const double dy = u[3 * a + 1] - u[3 * b + 1];
      const double dz = u[3 * a + 2] - u[3 * b + 2];
      const double dsum2 = (dx * dx + dy * dy + dz * dz);
      if (dsum2 > 0) {
        dsum += std::pow((dx * dx + dy * dy + dz * dz), 0.5);
        counter++;
      }
    }
    if (counter > 0) dsum /= counter;
    return dsum;
  };
  static const int apical[3][2] = {{3, 4}, {3, 5}, {4, 5}};
  static const int basal[3][2] = {{0, 1}, {0, 2}, {1, 2}};
  for (int e = 0; e < n_el; ++e) {
    if (elem_type[e] != 0) continue;
    double currValue = side_avg(e, apical);
    if (currValue > 0) { dsumCol += currValue; colCounter++; }
    if (elem_placement[e] == 0 || spans[e]) {
      currValue = side_avg(e, basal);
      if (currValue > 0) { dsumCol += currValue; colCounter++; }
    }
    double c[3] = {0.0, 0.0, 0.0};
    for (int a = 0; a < 6; ++a)
      for (int j = 0; j < 3; ++j) c[j] += u[3 * conn[6 * e + a] + j];
    c[0] /= 6; c[1] /= 6; c[2] /= 6;
    const double rx = (c[0] - bb[0]) / size[0];
    const double ry = (c[1] - bb[1]) / size[1];
    int relX = (int)std::floor(rx);
    int relY = (int)std::floor(ry / 2.0);
    if (relX < 0) relX = 0;
    if (relX > 9) relX = 9;
    if (relY < 0) relY = 0;
    if (relY > 4) relY = 4;
    sum[relX][relY] += currValue;
    cnt[relX][relY]++;
  }
  const double colAverageSideLength = dsumCol / (double)colCounter;
  for (int i = 0; i < 10; ++i)
    for (int j = 0; j < 5; ++j) {
      double g = sum[i][j];
      if (cnt[i][j] > 0) g /= cnt[i][j];
      g = 0.4 * g;
      grid50[5 * i + j] = g * g;
    }
  *packing_threshold = 0.4 * colAverageSideLength;
}

// The couple search of Simulation::detectPacingNodes (Simulation.cpp:8035-8078):
// nodes that can pack (Node::checkIfNodeHasPacking, Node.cpp:128-146: mass
// non-zero, no lateral owner (node_excluded), apical or basal) are tested
// against every later node that can pack, of the same placement, that is not
// an immediate neighbour (shares no element, ShapeBase.cpp:3260-3280); a couple
// is kept when its squared distance is below the master's (lower id's) grid
// threshold, indexed by the master's position in the bounding box.  The
// curved-region release (thereIsEmergentEllipseMarking) and collapsed nodes
// are not modelled.  Returns the number of couples (pairs filled up to cap).
int orc_detect_packing(int n_nodes, int n_el, const int32_t* conn, const double* u,
                       const int32_t* node_placement, const uint8_t* node_excluded,
                       const double* mass, const double* grid50, const double* bb, int32_t* pairs,
                       int cap) {
  std::vector<std::vector<int32_t>> neig(n_nodes);
  for (int e = 0; e < n_el; ++e)
    for (int a = 0; a < 6; ++a)
      for (int b = 0; b < 6; ++b)
        if (a != b) neig[conn[6 * e + a]].push_back(conn[6 * e + b]);
  for (auto& v : neig) {
    std::sort(v.begin(), v.end());
    v.erase(std::unique(v.begin(), v.end()), v.end());
  }
  auto can_pack = [&](int k) {
    if (mass[k] == 0) return false;
    if (node_excluded && node_excluded[k]) return false;
    return node_placement[k] == 0 || node_placement[k] == 1;
  };
  const double size[2] = {bb[3] - bb[0], bb[4] - bb[1]};
  int count = 0;
  for (int k = 0; k < n_nodes; ++k) {
    if (!can_pack(k)) continue;
    const double* pos = u + 3 * k;
    int gx = (int)std::floor((pos[0] - bb[0]) / size[0]);
    int gy = (int)std::floor((pos[1] - bb[1]) / size[1]);
    gx = std::min(std::max(gx, 0), 9);  // in range for positions inside the box
    gy = std::min(std::max(gy, 0), 4);
    const double t2 = grid50[5 * gx + gy];
    for (int o = k + 1; o < n_nodes; ++o) {
      if (!can_pack(o)) continue;
      if (node_placement[k] != node_placement[o]) continue;
      if (std::binary_search(neig[k].begin(), neig[k].end(), o)) continue;
      const double* ps = u + 3 * o;
      const double dx = pos[0] - ps[0];
      const double dy = pos[1] - ps[1];
      const double dz = pos[2] - ps[2];
      const double d2 = dx * dx + dy * dy + dz * dz;
      if (d2 < t2) {
        if (count < cap) { pairs[2 * count] = k; pairs[2 * count + 1] = o; }
        count++;
      }
    }
  }
  return count;
}

// ---- Young's modulus time series (YoungsModulusModifier.cpp) ----

// YoungsModulusModifier::CalculateYoungsModulusChangeRate
// (YoungsModulusModifier.cpp:15-46): rate[t][i][j] from the grids
// grids[t][i][j] (t < n_times, i < ny, j < nx).
void orc_stiffness_rates(int n_times, const double* times, int nx, int ny, const double* grids,
                         int linear, double* rate) {
  for (int t = 0; t + 1 < n_times; ++t) {
    const double deltaT = times[t + 1] - times[t];
    for (int i = 0; i < ny; ++i)
      for (int j = 0; j < nx; ++j) {
        const double g0 = grids[((size_t)t * ny + i) * nx + j];
        const double g1 = grids[((size_t)(t + 1) * ny + i) * nx + j];
        rate[((size_t)t * ny + i) * nx + j] = linear ? ((g1 / g0) - 1) / deltaT : log(g1 / g0) / deltaT;
      }
  }
}

// YoungsModulusModifier::calculateCurrentRateGridIndex
// (YoungsModulusModifier.cpp:48-72): the interval holding t, -1 when none.
int orc_stiffness_interval(int n_times, const double* times, double t) {
  if (t > times[n_times - 1]) return -1;
  for (int i = 0; i < n_times - 1; ++i)
    if (t >= times[i] && t < times[i + 1]) return i;
  return -1;
}

// ShapeBase::convertRelativePosToGridIndex (ShapeBase.cpp:834-860).
static void orc_grid_index(const double* rel, int nx, int ny, int* ix, int* iy, double* fx,
                           double* fy) {
  double r0 = rel[0], r1 = rel[1];
  r0 *= (float)(nx - 1);
  r1 *= (float)(ny - 1);
  *ix = (int)floor(r0);
  *fx = r0 - *ix;
  *iy = (int)floor(r1);
  *fy = r1 - *iy;
  if (*ix >= nx - 1) {
    *ix = nx - 2;
    *fx = 1.0;
  } else if (*ix < 0) {
    *ix = 0;
    *fx = 0.0;
  }
  if (*iy >= ny - 1) {
    *iy = ny - 2;
    *fy = 1.0;
  } else if (*iy < 0) {
    *iy = 0;
    *fy = 0.0;
  }
}

// YoungsModulusModifier::GetCurrentMultiplierChangeRate
// (YoungsModulusModifier.cpp:106-156) for one position.
double orc_stiffness_rate_at(const double* rate_t, int nx, int ny, int method, const double* rel) {
  int ix, iy;
  double fx, fy;
  orc_grid_index(rel, nx, ny, &ix, &iy, &fx, &fy);
  double nb[4];
  nb[0] = rate_t[(size_t)iy * nx + ix];
  nb[1] = rate_t[(size_t)iy * nx + ix + 1];
  nb[2] = rate_t[(size_t)(iy + 1) * nx + ix];
  nb[3] = rate_t[(size_t)(iy + 1) * nx + ix + 1];
  double out = 0.0;
  if (method == 0) {
    if (fx > 0.5) out = fy > 0.5 ? nb[3] : nb[1];
    else out = fy > 0.5 ? nb[2] : nb[0];
  } else if (method == 1) {
    out = nb[0] * (1.0 - fx) * (1.0 - fy) + nb[1] * fx * (1.0 - fy) +
          nb[2] * (1.0 - fx) * fy + nb[3] * fx * fy;
  } else if (method == 2) {
    out = nb[0];
  }
  return out;
}

// YoungsModulusModifier::updateTimeSeriesStiffnessMultiplier
// (YoungsModulusModifier.cpp:158-177) over every prism, then
// Prism::updateElasticProperties (Prism.cpp:526-527) with E the prism's E
// times its stiffness multiplier.  stm, lambda and mu are updated in place for
// the applicable prisms; returns the interval used (-1: nothing applied).
int orc_stiffness_apply(int n_el, double t, double dt, int n_times, const double* times, int nx,
                        int ny, const double* rate, int linear, int method,
                        const uint8_t* applicable, const double* rel_pos, const double* E,
                        const double* nu, double* stm, double* lambda, double* mu) {
  const int k = orc_stiffness_interval(n_times, times, t);
  if (k < 0) return -1;
  const double* rt = rate + (size_t)k * nx * ny;
  for (int e = 0; e < n_el; ++e) {
    if (!applicable[e]) continue;
    const double r = orc_stiffness_rate_at(rt, nx, ny, method, rel_pos + 2 * e);
    stm[e] = linear ? stm[e] + r * dt : stm[e] * exp(r * dt);
    const double v = nu[e];
    lambda[e] = stm[e] * E[e] * v / (1 + v) / (1 - 2.0 * v);
    mu[e] = stm[e] * E[e] / 2.0 / (1 + v);
  }
  return k;
}

}  // extern "C"
