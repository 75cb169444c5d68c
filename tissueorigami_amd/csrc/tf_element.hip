// Per-prism kernel: shape functions at the 6 Gauss points, Neo-Hookean elastic
// and viscous nodal forces, 18x18 elastic and viscous tangents.
//
// Follows, per Gauss point, Prism::calculateCurrNodalForces (Prism.cpp:766-886)
// with the ShapeBase helpers (ShapeBase.cpp:2448-2845); the elastic tangent
// uses the exact closed form of ShapeBase::calculateElasticKIntegral1 +
// calculateElasticKIntegral2 (ShapeBase.cpp:3010-3095): because
// Fe C^-1 Fe^T = I the push-forward of the Lagrangian tensor D81 reduces to
// Je^-1 [lambda d_ij d_kl + mu' (d_ik d_jl + d_il d_jk)], mu' = mu - lambda ln Je.
// The viscous tangent follows ShapeBase::calculateImplicitKViscous
// (ShapeBase.cpp:2868-2933, 3097-3209) and is stored transposed as the
// reference stores it.
//
// Work decomposition: one workgroup handles kEPB prisms.
//   phase 1: one thread per (prism, Gauss point) computes the kinematics and
//            stresses and leaves grad N, sigma and the scalar factors in LDS;
//   phase 2: one thread per (prism, 3x3 block) sums its block over the Gauss
//            points and writes 9 contiguous values (coalesced across lanes);
//            for the summed FP32 operator of the mixed-precision step, one
//            thread per (prism, node a) forms the six blocks (a, 0..5);
//   phase 3: one thread per (prism, node) sums the nodal forces.
// Every floating-point operation is written in the order of the CPU oracle's
// restatement and compiled without contraction, so the FP64 outputs agree
// with the oracle to the last bit up to the device log().
#include <type_traits>

#include "tf_internal.h"

#pragma clang fp contract(off)

namespace tf {
namespace {

#ifndef TF_ELEM_PER_BLOCK
#define TF_ELEM_PER_BLOCK 32
#endif
constexpr int kEPB = TF_ELEM_PER_BLOCK;  // prisms per workgroup
constexpr int kGP = kEPB * 6;             // (prism, Gauss point) threads of phases 1 and 3
constexpr int kThreads = (kGP + 63) / 64 * 64;  // whole wavefronts

struct ElemConst {
  double sfd[6][3][6];  // shape function derivatives per Gauss point
  double w;             // Gauss weight (1/6 for all six points)
};

struct GPData {
  double dN[6][3];
  double sig[9];
  double detF, detdXde, cl, cm;
};

struct GPVisc {
  double sv[9];
  double P[9];
  double eta;
};

__device__ inline double det3(const double M[3][3]) {
  double det = 0.0;
  det = M[0][0] * (M[1][1] * M[2][2] - M[1][2] * M[2][1]);
  det -= M[0][1] * (M[1][0] * M[2][2] - M[1][2] * M[2][0]);
  det += M[0][2] * (M[1][0] * M[2][1] - M[1][1] * M[2][0]);
  return det;
}

// LU with partial pivoting, then L U X = P I (the GSL algorithm the reference
// calls at ShapeBase.cpp:3546-3561).
__device__ inline void inv3(const double A[3][3], double X[3][3]) {
  double LU[3][3];
  int perm[3] = {0, 1, 2};
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) LU[i][j] = A[i][j];
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    int piv = j;
    double mx = fabs(LU[j][j]);
#pragma unroll
    for (int i = j + 1; i < 3; ++i)
      if (fabs(LU[i][j]) > mx) { mx = fabs(LU[i][j]); piv = i; }
    if (piv != j) {
#pragma unroll
      for (int k = 0; k < 3; ++k) { double t = LU[j][k]; LU[j][k] = LU[piv][k]; LU[piv][k] = t; }
      int t = perm[j]; perm[j] = perm[piv]; perm[piv] = t;
    }
    if (LU[j][j] != 0.0) {
#pragma unroll
      for (int i = j + 1; i < 3; ++i) {
        LU[i][j] /= LU[j][j];
#pragma unroll
        for (int k = j + 1; k < 3; ++k) LU[i][k] -= LU[i][j] * LU[j][k];
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
      for (int k = 0; k < i; ++k) s -= LU[i][k] * y[k];
      y[i] = s;
    }
#pragma unroll
    for (int i = 2; i >= 0; --i) {
      double s = y[i];
#pragma unroll
      for (int k = i + 1; k < 3; ++k) s -= LU[i][k] * X[k][c];
      X[i][c] = s / LU[i][i];
    }
  }
}

__device__ inline void matmul3(const double A[3][3], const double B[3][3], double C[3][3]) {
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      double s = 0.0;
#pragma unroll
      for (int k = 0; k < 3; ++k) s += A[i][k] * B[k][j];
      C[i][j] = s;
    }
}

// Mixed precision, phase 2: the FP32 tangent block (a, b) of one prism from
// its FP64 Gauss-point data, in closed forms that need about a fifth of the
// operations of the reference-order evaluation, with contraction allowed:
//   elastic  K_ik = sum_gp w [cl Na_i Nb_k + cm Na_k Nb_i
//                             + d_ik (cm Na.Nb + vol Na.sigma.Nb)]
//   viscous  B_a^T B~_b = (Na.Nb I + Nb (x) Na) / 2 (the Voigt B of node a
//            times that of node b with halved shear rows), so
//            (B_a^T B~_b P)_ij = (Na.Nb P_ij + Nb_i (Na^T P)_j) / 2 and
//            (S (Na Nb^T - Nb Na^T))_ij = (S Na)_i Nb_j - (S Nb)_i Na_j;
//            Kv = sum_gp w detF detdXde [eta B_a^T B~_b P + S (...)].
// Kv is returned in the same (untransposed) order as phase 2b's K.
__device__ inline void mixed_blocks(const GPData* sd, const GPVisc* svd, int el, int ia, int ib,
                                    bool visc, float w, float Ke[9], float Kv[9]) {
#pragma clang fp contract(fast)
#pragma unroll
  for (int q = 0; q < 9; ++q) { Ke[q] = 0.f; Kv[q] = 0.f; }
  const bool has_eta = visc && svd[el * 6].eta != 0;
#pragma unroll 1
  for (int g = 0; g < 6; ++g) {
    const GPData& d = sd[el * 6 + g];
    float Na[3], Nb[3], sg[9];
#pragma unroll
    for (int j = 0; j < 3; ++j) { Na[j] = (float)d.dN[ia][j]; Nb[j] = (float)d.dN[ib][j]; }
#pragma unroll
    for (int q = 0; q < 9; ++q) sg[q] = (float)d.sig[q];
    const float cl = (float)d.cl, cm = (float)d.cm;
    const float vol = (float)(d.detF * d.detdXde);
    const float dot = Na[0] * Nb[0] + Na[1] * Nb[1] + Na[2] * Nb[2];
    float sNb[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) sNb[i] = sg[3 * i] * Nb[0] + sg[3 * i + 1] * Nb[1] + sg[3 * i + 2] * Nb[2];
    const float geo = Na[0] * sNb[0] + Na[1] * sNb[1] + Na[2] * sNb[2];
    const float diag = cm * dot + vol * geo;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        float v = cl * Na[i] * Nb[k] + cm * Na[k] * Nb[i];
        if (i == k) v += diag;
        Ke[3 * i + k] += v * w;
      }
    if (has_eta) {
      const GPVisc& dv = svd[el * 6 + g];
      float P[9], S[9];
#pragma unroll
      for (int q = 0; q < 9; ++q) { P[q] = (float)dv.P[q]; S[q] = (float)dv.sv[q]; }
      float NaP[3], SNa[3], SNb[3];
#pragma unroll
      for (int j = 0; j < 3; ++j) NaP[j] = Na[0] * P[j] + Na[1] * P[3 + j] + Na[2] * P[6 + j];
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        SNa[i] = S[3 * i] * Na[0] + S[3 * i + 1] * Na[1] + S[3 * i + 2] * Na[2];
        SNb[i] = S[3 * i] * Nb[0] + S[3 * i + 1] * Nb[1] + S[3 * i + 2] * Nb[2];
      }
      const float eta = (float)dv.eta;
      const float f = vol * w;
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
          const float k1 = 0.5f * (dot * P[3 * i + j] + Nb[i] * NaP[j]);
          const float k2 = SNa[i] * Nb[j] - SNb[i] * Na[j];
          Kv[3 * i + j] += (eta * k1 + k2) * f;
        }
    }
  }
}

// Green-Lagrange strain E = (C - I) / 2 of Prism::calculateCurrNodalForces
// (Prism.cpp:826-837, ShapeBase.cpp:2459-2472), from C = Fe^T Fe: the
// reference keeps the value of the last Gauss point it evaluates, in the Voigt
// order E00, E11, E22, 2 E01, 2 E21, 2 E02 (2 (C01 / 2) = C01 exactly).
__device__ inline void store_strain(const ElementArgs& a, int e, const double C[3][3]) {
  if (!a.strain || (a.ptissue && !a.tactive[a.ptissue[e]])) return;
  double* s = a.strain + 6 * (size_t)e;
  s[0] = (C[0][0] - 1.0) * 0.5;
  s[1] = (C[1][1] - 1.0) * 0.5;
  s[2] = (C[2][2] - 1.0) * 0.5;
  s[3] = C[0][1];
  s[4] = C[2][1];
  s[5] = C[0][2];
}

// MODE 0: FP64 tangents (oracle operation order); 1: FP32 elastic and
// viscous tangents stored apart (Ke32, Kv32, for element-level parity).
// The mixed-precision NR step uses element_kernel_mix.
template <bool VISC, int MODE>
__global__ __launch_bounds__(kThreads) void element_kernel(ElemConst c, ElementArgs a) {
  constexpr bool MIXED = MODE != 0;
  __shared__ GPData sd[kEPB * 6];
  __shared__ GPVisc svd[VISC ? kEPB * 6 : 1];
  const int tid = threadIdx.x;
  const int e0 = blockIdx.x * kEPB;

  // ---------------- phase 1: one thread per (prism, Gauss point) ----------
  if (tid < kGP) {
    const int el = tid / 6;
    const int g = tid % 6;
    const int e = e0 + el;
    if (e < a.n_prisms) {
      int nd[6];
#pragma unroll
      for (int k = 0; k < 6; ++k) nd[k] = a.conn[6 * e + k];
      double pos[6][3], ref[6][3], iFg[3][3];
#pragma unroll
      for (int k = 0; k < 6; ++k)
#pragma unroll
        for (int i = 0; i < 3; ++i) {
          pos[k][i] = a.uk[3 * nd[k] + i];
          ref[k][i] = a.refpos[18 * e + 3 * k + i];
        }
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) iFg[i][j] = a.invFg[9 * e + 3 * i + j];
      const double lambda = a.lambda[e];
      const double mu = a.mu[e];
      const double(*SFD)[6] = c.sfd[g];
      // reference geometry (Prism.cpp:715-732)
      double dXde[3][3], InvdXde[3][3];
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
          double s = 0.0;
#pragma unroll
          for (int k = 0; k < 6; ++k) s += SFD[j][k] * ref[k][i];
          dXde[i][j] = s;
        }
      const double detdXde = det3(dXde);
      inv3(dXde, InvdXde);
      // current Jacobian, F, Fe (Prism.cpp:782-788)
      double Jac[3][3], F[3][3], Fe[3][3];
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
          double s = 0.0;
#pragma unroll
          for (int k = 0; k < 6; ++k) s += SFD[j][k] * pos[k][i];
          Jac[i][j] = s;
        }
      matmul3(Jac, InvdXde, F);
      matmul3(F, iFg, Fe);
      double C[3][3];
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
          double s = 0.0;
#pragma unroll
          for (int k = 0; k < 3; ++k) s += Fe[k][i] * Fe[k][j];
          C[i][j] = s;
        }
      if (g == 5) store_strain(a, e, C);
      const double detFe = det3(Fe);
      const double lnJ = log(detFe);
      double InvC[3][3];
      inv3(C, InvC);
      double PK2[3][3];
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
          const double Iij = (i == j) ? 1.0 : 0.0;
          PK2[i][j] = InvC[i][j] * (lambda * lnJ) + (Iij - InvC[i][j]) * mu;
        }
      double T1[3][3];
      matmul3(Fe, PK2, T1);
      const double inv_detFe = 1.0 / detFe;
      GPData& out = sd[tid];
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
          double s = 0.0;
#pragma unroll
          for (int k = 0; k < 3; ++k) s += T1[i][k] * Fe[j][k];
          out.sig[3 * i + j] = s * inv_detFe;
        }
      double InvJ[3][3];
      inv3(Jac, InvJ);
#pragma unroll
      for (int k = 0; k < 6; ++k)
#pragma unroll
        for (int m = 0; m < 3; ++m) {
          double s = 0.0;
#pragma unroll
          for (int n = 0; n < 3; ++n) s += InvJ[n][m] * SFD[n][k];
          out.dN[k][m] = s;
        }
      const double detF = det3(F);
      out.detF = detF;
      out.detdXde = detdXde;
      const double vol = detF * detdXde;
      const double cc = vol / detFe;
      out.cl = cc * lambda;
      out.cm = cc * (mu - lambda * lnJ);
      if (VISC) {
        const double eta = a.eta[e];
        GPVisc& ov = svd[tid];
        ov.eta = eta;
        if (eta != 0) {
          const double rdt = 1.0 / a.dt;
          double vel[6][3];
#pragma unroll
          for (int k = 0; k < 6; ++k)
#pragma unroll
            for (int i = 0; i < 3; ++i)
              vel[k][i] = (a.uk[3 * nd[k] + i] - a.un[3 * nd[k] + i]) * rdt;
          // l = B v (ShapeBase.cpp:2640-2716), Voigt xx, yy, zz, xy, yz, xz
          double l0 = 0.0, l1 = 0.0, l2 = 0.0, l3 = 0.0, l4 = 0.0, l5 = 0.0;
#pragma unroll
          for (int k = 0; k < 6; ++k) {
            const double nx = out.dN[k][0], ny = out.dN[k][1], nz = out.dN[k][2];
            l0 += nx * vel[k][0];
            l1 += ny * vel[k][1];
            l2 += nz * vel[k][2];
            l3 += ny * vel[k][0];
            l3 += nx * vel[k][1];
            l4 += nz * vel[k][1];
            l4 += ny * vel[k][2];
            l5 += nz * vel[k][0];
            l5 += nx * vel[k][2];
          }
          double d[9];
          d[0] = l0; d[4] = l1; d[8] = l2;
          d[1] = 0.5 * l3; d[7] = 0.5 * l4; d[2] = 0.5 * l5;
          d[3] = 0.5 * l3; d[5] = 0.5 * l4; d[6] = 0.5 * l5;
#pragma unroll
          for (int q = 0; q < 9; ++q) ov.sv[q] = d[q] * eta;
          // velocity gradient and (I/dt - L) (ShapeBase.cpp:2908-2916, 3097-3121)
          double L[9];
#pragma unroll
          for (int q = 0; q < 9; ++q) L[q] = 0.0;
#pragma unroll
          for (int k = 0; k < 6; ++k)
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
              for (int j = 0; j < 3; ++j) L[3 * i + j] += vel[k][i] * out.dN[k][j];
#pragma unroll
          for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j)
              ov.P[3 * i + j] = ((i == j) ? 1.0 : 0.0) * (1.0 / a.dt) - L[3 * i + j];
        } else {
#pragma unroll
          for (int q = 0; q < 9; ++q) { ov.sv[q] = 0.0; ov.P[q] = 0.0; }
        }
      }
    }
  }
  __syncthreads();

  const double w = c.w;
  if (MIXED) {
    // ---------------- phase 2 (mixed): FP32 elastic and viscous blocks -----
    for (int task = tid; task < kEPB * 36; task += kThreads) {
      const int el = task / 36;
      const int ab = task % 36;
      const int e = e0 + el;
      if (e >= a.n_prisms) break;
      const int ia = ab / 6, ib = ab % 6;
      const size_t base = ((size_t)e * 36 + ib * 6 + ia) * 9;
      float Ke[9], Kv[9];
      mixed_blocks(sd, svd, el, ia, ib, VISC, (float)w, Ke, Kv);
#pragma unroll
      for (int q = 0; q < 9; ++q) a.Ke32[base + q] = Ke[q];
      if (VISC) {
        // reference storage: TriPointKv[3b+j][3a+i] -> block (b, a), entry (j, i)
        const size_t bv = ((size_t)e * 36 + ia * 6 + ib) * 9;
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
          for (int j = 0; j < 3; ++j) a.Kv32[bv + 3 * j + i] = Kv[3 * i + j];
      }
    }
  } else {
  // ---------------- phase 2a: elastic 3x3 blocks ---------------------------
  using T = double;
  const T tw = (T)w;
  for (int task = tid; task < kEPB * 36; task += kThreads) {
    const int el = task / 36;
    const int ab = task % 36;
    const int e = e0 + el;
    if (e >= a.n_prisms) break;
    const int ia = ab / 6, ib = ab % 6;
    T K[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) K[q] = (T)0;
#pragma unroll
    for (int g = 0; g < 6; ++g) {
      const GPData& d = sd[el * 6 + g];
      T Na[3], Nb[3], sg[9];
#pragma unroll
      for (int j = 0; j < 3; ++j) { Na[j] = (T)d.dN[ia][j]; Nb[j] = (T)d.dN[ib][j]; }
#pragma unroll
      for (int q = 0; q < 9; ++q) sg[q] = (T)d.sig[q];
      const T cl = (T)d.cl, cm = (T)d.cm;
      const T vol = (T)d.detF * (T)d.detdXde;
      T dot = (T)0, geo = (T)0;
#pragma unroll
      for (int j = 0; j < 3; ++j) dot += Na[j] * Nb[j];
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        T s = (T)0;
#pragma unroll
        for (int j = 0; j < 3; ++j) s += sg[3 * i + j] * Nb[j];
        geo += Na[i] * s;
      }
      const T diag = cm * dot + vol * geo;
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          T v = cl * Na[i] * Nb[k] + cm * Na[k] * Nb[i];
          if (i == k) v += diag;
          K[3 * i + k] += v * tw;
        }
    }
    const size_t base = ((size_t)e * 36 + ib * 6 + ia) * 9;
#pragma unroll
    for (int q = 0; q < 9; ++q) a.Ke[base + q] = K[q];
  }

  // ---------------- phase 2b: viscous 3x3 blocks (transposed storage) -----
  if (VISC) {
    for (int task = tid; task < kEPB * 36; task += kThreads) {
      const int el = task / 36;
      const int ab = task % 36;
      const int e = e0 + el;
      if (e >= a.n_prisms) break;
      const int ia = ab / 6, ib = ab % 6;
      T K[9];  // K[3i+j] -> TriPointKv[3b+j][3a+i]
#pragma unroll
      for (int q = 0; q < 9; ++q) K[q] = (T)0;
      if (svd[el * 6].eta != 0) {
#pragma unroll 1
        for (int g = 0; g < 6; ++g) {
          const GPData& d = sd[el * 6 + g];
          const GPVisc& dv = svd[el * 6 + g];
          T Na[3], Nb[3];
#pragma unroll
          for (int j = 0; j < 3; ++j) { Na[j] = (T)d.dN[ia][j]; Nb[j] = (T)d.dN[ib][j]; }
          // B columns of nodes a and b (Voigt rows xx, yy, zz, xy, yz, xz)
          T Ba[3][6], Bb[3][6];
#pragma unroll
          for (int j = 0; j < 3; ++j)
#pragma unroll
            for (int r = 0; r < 6; ++r) { Ba[j][r] = (T)0; Bb[j][r] = (T)0; }
          Ba[0][0] = Na[0]; Ba[0][3] = Na[1]; Ba[0][5] = Na[2];
          Ba[1][1] = Na[1]; Ba[1][3] = Na[0]; Ba[1][4] = Na[2];
          Ba[2][2] = Na[2]; Ba[2][4] = Na[1]; Ba[2][5] = Na[0];
          Bb[0][0] = Nb[0]; Bb[0][3] = Nb[1]; Bb[0][5] = Nb[2];
          Bb[1][1] = Nb[1]; Bb[1][3] = Nb[0]; Bb[1][4] = Nb[2];
          Bb[2][2] = Nb[2]; Bb[2][4] = Nb[1]; Bb[2][5] = Nb[0];
#pragma unroll
          for (int j = 0; j < 3; ++j)
#pragma unroll
            for (int r = 3; r < 6; ++r) Bb[j][r] = (T)0.5 * Bb[j][r];
          T BB[3][3];
#pragma unroll
          for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) {
              T s = (T)0;
#pragma unroll
              for (int r = 0; r < 6; ++r) s += Ba[i][r] * Bb[j][r];
              BB[i][j] = s;
            }
          T P[3][3], K1[3][3], S[3][3], paren[3][3], K2[3][3];
#pragma unroll
          for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) {
              P[i][j] = (T)dv.P[3 * i + j];
              S[i][j] = (T)dv.sv[3 * i + j];
              paren[i][j] = Na[i] * Nb[j] - Nb[i] * Na[j];
            }
#pragma unroll
          for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) {
              T s1 = (T)0, s2 = (T)0;
#pragma unroll
              for (int q = 0; q < 3; ++q) {
                s1 += BB[i][q] * P[q][j];
                s2 += S[i][q] * paren[q][j];
              }
              K1[i][j] = s1;
              K2[i][j] = s2;
            }
          const T detF = (T)d.detF, detdXde = (T)d.detdXde, eta = (T)dv.eta;
#pragma unroll
          for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) {
              T v1 = K1[i][j] * detF;
              v1 *= detdXde;
              v1 *= eta;
              T v2 = K2[i][j] * detF;
              v2 *= detdXde;
              T curr = (T)0;
              curr += v1;
              curr += v2;
              K[3 * i + j] += curr * tw;
            }
        }
      }
      // reference storage: TriPointKv[3b+j][3a+i] -> block (b, a), entry (j, i)
      const size_t base = ((size_t)e * 36 + ia * 6 + ib) * 9;
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) a.Kv[base + 3 * j + i] = K[3 * i + j];
    }
  }

  }  // FP64 phase 2

  // ---------------- phase 3: nodal forces ----------------------------------
  if (tid < kGP) {
    const int el = tid / 6;
    const int k = tid % 6;
    const int e = e0 + el;
    if (e < a.n_prisms) {
      double te[3] = {0.0, 0.0, 0.0}, tv[3] = {0.0, 0.0, 0.0};
#pragma unroll
      for (int g = 0; g < 6; ++g) {
        const GPData& d = sd[el * 6 + g];
        const double* s = d.sig;
        // compact stress [xx, yy, zz, xy, zy, xz] (ShapeBase.cpp:2560-2566)
        const double c0 = s[0], c1 = s[4], c2 = s[8], c3 = s[1], c4 = s[7], c5 = s[2];
        double bx = d.dN[k][0] * d.detF; bx *= d.detdXde;
        double by = d.dN[k][1] * d.detF; by *= d.detdXde;
        double bz = d.dN[k][2] * d.detF; bz *= d.detdXde;
        double f0 = 0.0; f0 += bx * c0; f0 += by * c3; f0 += bz * c5;
        double f1 = 0.0; f1 += by * c1; f1 += bx * c3; f1 += bz * c4;
        double f2 = 0.0; f2 += bz * c2; f2 += by * c4; f2 += bx * c5;
        te[0] = te[0] + f0 * w;
        te[1] = te[1] + f1 * w;
        te[2] = te[2] + f2 * w;
        if (VISC) {
          const double* v = svd[el * 6 + g].sv;
          const double v0 = v[0], v1 = v[4], v2 = v[8], v3 = v[1], v4 = v[7], v5 = v[2];
          double h0 = 0.0; h0 += bx * v0; h0 += by * v3; h0 += bz * v5;
          double h1 = 0.0; h1 += by * v1; h1 += bx * v3; h1 += bz * v4;
          double h2 = 0.0; h2 += bz * v2; h2 += by * v4; h2 += bx * v5;
          tv[0] = tv[0] + h0 * w;
          tv[1] = tv[1] + h1 * w;
          tv[2] = tv[2] + h2 * w;
        }
      }
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        a.Fe[18 * e + 3 * k + i] = 0.0 - te[i];
        a.Fv[18 * e + 3 * k + i] = 0.0 - tv[i];
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Mixed-precision per-prism kernel (mode 2): the FP32 K blocks of the summed
// elastic + viscous tangent and the FP64 nodal forces (elastic + viscous
// summed), which the mixed NR step needs.  FP64 results agree with the
// reference-order evaluation to rounding, not bit for bit:
//   - sigma = (lambda ln J I + mu (b - I)) / J with b = Fe Fe^T (equal to
//     Fe PK2 Fe^T / J since Fe C^-1 Fe^T = I), which needs no inverse of C;
//   - the inverses of dX/dxi and of the current Jacobian are adj(M) / det(M).
// Phase 1 (one thread per (prism, Gauss point)) leaves in LDS both the FP64
// data of the forces and the FP32 data of the tangent, already scaled by the
// Gauss weight and the volume; phase 2 (one thread per (prism, node a))
// forms the six K blocks (a, 0..5) and node a's force.
#ifndef TF_MIX_EPB
#define TF_MIX_EPB 8
#endif
constexpr int kMEPB = TF_MIX_EPB;
constexpr int kMGP = kMEPB * 6;
constexpr int kMThreads = (kMGP + 63) / 64 * 64;

struct GPMix {
  double fN[6][3];  // grad N * det F * det dX/dxi
  double s6[6];     // sigma: xx, yy, zz, xy, yz, xz
  double v6[6];     // eta d, same order (viscous)
  float N[6][3];    // grad N
  float sg[9];      // sigma
  float clw, cmw, volw;
  float Sp[9], Pp[9];  // w vol eta d and (w vol eta / 2) (I/dt - L)
  // padding to 110 dwords: the Gauss points that the threads of a wavefront
  // read together then fall in different LDS banks
  float pad[2];
};
static_assert(sizeof(GPMix) == 440, "GPMix layout");

// adj(M) / det(M), and det(M)
__device__ inline double inv3_adj(const double M[3][3], double X[3][3]) {
  const double c00 = M[1][1] * M[2][2] - M[1][2] * M[2][1];
  const double c01 = M[1][2] * M[2][0] - M[1][0] * M[2][2];
  const double c02 = M[1][0] * M[2][1] - M[1][1] * M[2][0];
  const double det = M[0][0] * c00 + M[0][1] * c01 + M[0][2] * c02;
  const double r = 1.0 / det;
  X[0][0] = c00 * r;
  X[1][0] = c01 * r;
  X[2][0] = c02 * r;
  X[0][1] = (M[0][2] * M[2][1] - M[0][1] * M[2][2]) * r;
  X[1][1] = (M[0][0] * M[2][2] - M[0][2] * M[2][0]) * r;
  X[2][1] = (M[0][1] * M[2][0] - M[0][0] * M[2][1]) * r;
  X[0][2] = (M[0][1] * M[1][2] - M[0][2] * M[1][1]) * r;
  X[1][2] = (M[0][2] * M[1][0] - M[0][0] * M[1][2]) * r;
  X[2][2] = (M[0][0] * M[1][1] - M[0][1] * M[1][0]) * r;
  return det;
}

template <bool VISC>
__global__ __launch_bounds__(kMThreads) void element_kernel_mix(ElemConst c, ElementArgs a) {
#pragma clang fp contract(fast)
  __shared__ GPMix sd[kMGP];
  const int tid = threadIdx.x;
  const int e0 = blockIdx.x * kMEPB;
  const double w = c.w;
  const float wf = (float)w;

  // ---------------- phase 1: one thread per (prism, Gauss point) ----------
  if (tid < kMGP) {
    const int el = tid / 6, g = tid % 6;
    const int e = e0 + el;
    if (e < a.n_prisms) {
      int nd[6];
#pragma unroll
      for (int k = 0; k < 6; ++k) nd[k] = a.conn[6 * e + k];
      double pos[6][3], ref[6][3], iFg[3][3];
#pragma unroll
      for (int k = 0; k < 6; ++k)
#pragma unroll
        for (int i = 0; i < 3; ++i) {
          pos[k][i] = a.uk[3 * nd[k] + i];
          ref[k][i] = a.refpos[18 * e + 3 * k + i];
        }
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) iFg[i][j] = a.invFg[9 * e + 3 * i + j];
      const double lambda = a.lambda[e], mu = a.mu[e];
      const double(*SFD)[6] = c.sfd[g];
      double dXde[3][3], InvdXde[3][3], Jac[3][3], InvJ[3][3];
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
          double s = 0.0, t = 0.0;
#pragma unroll
          for (int k = 0; k < 6; ++k) {
            s += SFD[j][k] * ref[k][i];
            t += SFD[j][k] * pos[k][i];
          }
          dXde[i][j] = s;
          Jac[i][j] = t;
        }
      const double detdXde = inv3_adj(dXde, InvdXde);
      double F[3][3], Fe[3][3];
      matmul3(Jac, InvdXde, F);
      matmul3(F, iFg, Fe);
      if (g == 5 && a.strain) {
        double C[3][3];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
          for (int j = i; j < 3; ++j) {
            C[i][j] = Fe[0][i] * Fe[0][j] + Fe[1][i] * Fe[1][j] + Fe[2][i] * Fe[2][j];
            C[j][i] = C[i][j];
          }
        store_strain(a, e, C);
      }
      const double detFe = det3(Fe);
      const double lnJ = log(detFe);
      const double rJ = 1.0 / detFe;
      double sig[9];
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = i; j < 3; ++j) {
          const double b = Fe[i][0] * Fe[j][0] + Fe[i][1] * Fe[j][1] + Fe[i][2] * Fe[j][2];
          const double v = (i == j ? lambda * lnJ + mu * (b - 1.0) : mu * b) * rJ;
          sig[3 * i + j] = v;
          sig[3 * j + i] = v;
        }
      inv3_adj(Jac, InvJ);
      double dN[6][3];
#pragma unroll
      for (int k = 0; k < 6; ++k)
#pragma unroll
        for (int m = 0; m < 3; ++m)
          dN[k][m] = InvJ[0][m] * SFD[0][k] + InvJ[1][m] * SFD[1][k] + InvJ[2][m] * SFD[2][k];
      const double detF = det3(F);
      const double vol = detF * detdXde;
      const double cc = vol * rJ;
      GPMix& o = sd[tid];
#pragma unroll
      for (int k = 0; k < 6; ++k)
#pragma unroll
        for (int m = 0; m < 3; ++m) {
          o.fN[k][m] = dN[k][m] * vol;
          o.N[k][m] = (float)dN[k][m];
        }
      o.s6[0] = sig[0]; o.s6[1] = sig[4]; o.s6[2] = sig[8];
      o.s6[3] = sig[1]; o.s6[4] = sig[5]; o.s6[5] = sig[2];
#pragma unroll
      for (int q = 0; q < 9; ++q) o.sg[q] = (float)sig[q];
      o.clw = (float)(cc * lambda * w);
      o.cmw = (float)(cc * (mu - lambda * lnJ) * w);
      o.volw = (float)(vol * w);
      if (VISC) {
        const double eta = a.eta[e];
        double sv[9], P[9];
        if (eta != 0) {
          const double rdt = 1.0 / a.dt;
          double vel[6][3];
#pragma unroll
          for (int k = 0; k < 6; ++k)
#pragma unroll
            for (int i = 0; i < 3; ++i)
              vel[k][i] = (a.uk[3 * nd[k] + i] - a.un[3 * nd[k] + i]) * rdt;
          // velocity gradient L (ShapeBase.cpp:2908-2916, 3097-3121), its
          // symmetric part d and (I/dt - L)
          double L[9];
#pragma unroll
          for (int q = 0; q < 9; ++q) L[q] = 0.0;
#pragma unroll
          for (int k = 0; k < 6; ++k)
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
              for (int j = 0; j < 3; ++j) L[3 * i + j] += vel[k][i] * dN[k][j];
#pragma unroll
          for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) {
              sv[3 * i + j] = 0.5 * (L[3 * i + j] + L[3 * j + i]) * eta;
              P[3 * i + j] = (i == j ? rdt : 0.0) - L[3 * i + j];
            }
        } else {
#pragma unroll
          for (int q = 0; q < 9; ++q) { sv[q] = 0.0; P[q] = 0.0; }
        }
        o.v6[0] = sv[0]; o.v6[1] = sv[4]; o.v6[2] = sv[8];
        o.v6[3] = sv[1]; o.v6[4] = sv[5]; o.v6[5] = sv[2];
        const double vw = vol * w, pe = 0.5 * vol * w * eta;
#pragma unroll
        for (int q = 0; q < 9; ++q) {
          o.Sp[q] = (float)(vw * sv[q]);
          o.Pp[q] = (float)(pe * P[q]);
        }
      }
    }
  }
  __syncthreads();

  // ---------------- phase 2: one thread per (prism, node a) ----------------
  if (tid >= kMGP) return;
  const int el = tid / 6, ia = tid % 6;
  const int e = e0 + el;
  if (e >= a.n_prisms) return;
  const bool has_eta = VISC && a.eta[e] != 0;
  float K[6][9];
#pragma unroll
  for (int b = 0; b < 6; ++b)
#pragma unroll
    for (int q = 0; q < 9; ++q) K[b][q] = 0.f;
  double fe[3] = {0.0, 0.0, 0.0};
#pragma unroll 1
  for (int g = 0; g < 6; ++g) {
    const GPMix& d = sd[el * 6 + g];
    // the force of node a: (sigma + eta d) grad N_a vol w
    {
      const double bx = d.fN[ia][0], by = d.fN[ia][1], bz = d.fN[ia][2];
      double s0 = d.s6[0], s1 = d.s6[1], s2 = d.s6[2], s3 = d.s6[3], s4 = d.s6[4], s5 = d.s6[5];
      if (VISC) {
        s0 += d.v6[0]; s1 += d.v6[1]; s2 += d.v6[2];
        s3 += d.v6[3]; s4 += d.v6[4]; s5 += d.v6[5];
      }
      fe[0] += (bx * s0 + by * s3 + bz * s5) * w;
      fe[1] += (by * s1 + bx * s3 + bz * s4) * w;
      fe[2] += (bz * s2 + by * s4 + bx * s5) * w;
    }
    // the Gauss point's FP32 data in registers, loaded once
    float N[6][3], sg[9], Sp[9], Pp[9];
#pragma unroll
    for (int b = 0; b < 6; ++b)
#pragma unroll
      for (int j = 0; j < 3; ++j) N[b][j] = d.N[b][j];
#pragma unroll
    for (int q = 0; q < 9; ++q) sg[q] = d.sg[q];
    if (has_eta) {
#pragma unroll
      for (int q = 0; q < 9; ++q) { Sp[q] = d.Sp[q]; Pp[q] = d.Pp[q]; }
    }
    float Na[3];
#pragma unroll
    for (int j = 0; j < 3; ++j) Na[j] = N[ia][j];
    const float clw = d.clw, cmw = d.cmw, volw = d.volw;
    float sTNa[3], V[3];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      sTNa[j] = sg[j] * Na[0] + sg[3 + j] * Na[1] + sg[6 + j] * Na[2];
      V[j] = cmw * Na[j];
    }
    if (has_eta) {
#pragma unroll
      for (int j = 0; j < 3; ++j)
        V[j] -= Sp[3 * j] * Na[0] + Sp[3 * j + 1] * Na[1] + Sp[3 * j + 2] * Na[2];
    }
#pragma unroll
    for (int b = 0; b < 6; ++b) {
      float Nb[3];
#pragma unroll
      for (int j = 0; j < 3; ++j) Nb[j] = N[b][j];
      const float dot = Na[0] * Nb[0] + Na[1] * Nb[1] + Na[2] * Nb[2];
      const float diag = volw * (sTNa[0] * Nb[0] + sTNa[1] * Nb[1] + sTNa[2] * Nb[2]) + cmw * dot;
      float U[3];
#pragma unroll
      for (int j = 0; j < 3; ++j) U[j] = clw * Nb[j];
      if (has_eta) {
        float W[3];
#pragma unroll
        for (int j = 0; j < 3; ++j) {
          U[j] += Sp[3 * j] * Nb[0] + Sp[3 * j + 1] * Nb[1] + Sp[3 * j + 2] * Nb[2];
          W[j] = Pp[j] * Nb[0] + Pp[3 + j] * Nb[1] + Pp[6 + j] * Nb[2];
        }
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
          for (int cc = 0; cc < 3; ++cc)
            K[b][3 * r + cc] += Na[r] * U[cc] + Nb[r] * V[cc] + Na[cc] * W[r] + dot * Pp[3 * cc + r];
      } else {
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
          for (int cc = 0; cc < 3; ++cc) K[b][3 * r + cc] += Na[r] * U[cc] + Nb[r] * V[cc];
      }
#pragma unroll
      for (int r = 0; r < 3; ++r) K[b][4 * r] += diag;
    }
  }
#pragma unroll
  for (int b = 0; b < 6; ++b) {
    float* out = a.Ke32 + ((size_t)e * 36 + b * 6 + ia) * 9;
#pragma unroll
    for (int q = 0; q < 9; ++q) out[q] = K[b][q];
  }
#pragma unroll
  for (int i = 0; i < 3; ++i) a.Fe[18 * e + 3 * ia + i] = 0.0 - fe[i];
}

ElemConst make_const() {
  ElemConst c;
  // Gauss rule (Prism.cpp:99-131) and shape-function derivative table
  // (Prism.cpp:597-652), evaluated on the host with the reference's expressions.
  const double s = 1.0 / std::pow(3.0, 0.5);
  const double A = 1.0 / 6.0, B = 2.0 / 3.0;
  const double pts[6][3] = {{A, A, s}, {B, A, s}, {A, B, s}, {A, A, -s}, {B, A, -s}, {A, B, -s}};
  for (int g = 0; g < 6; ++g) {
    const double eta = pts[g][0], nu = pts[g][1], zeta = pts[g][2];
    const double alpha = (1 - zeta) / 2;
    const double beta = (1 + zeta) / 2;
    const double lam = 1 - eta - nu;
    const double r0[6] = {-alpha, alpha, 0, -beta, beta, 0};
    const double r1[6] = {-alpha, 0, alpha, -beta, 0, beta};
    const double r2[6] = {-lam / 2, -eta / 2, -nu / 2, lam / 2, eta / 2, nu / 2};
    for (int k = 0; k < 6; ++k) {
      c.sfd[g][0][k] = r0[k];
      c.sfd[g][1][k] = r1[k];
      c.sfd[g][2][k] = r2[k];
    }
  }
  c.w = 1.0 / 6.0;
  return c;
}

}  // namespace

hipError_t launch_element(const ElementArgs& a, int mode, hipStream_t st) {
  static const ElemConst c = make_const();
  if (a.n_prisms == 0) return hipSuccess;
  const int grid = (a.n_prisms + kEPB - 1) / kEPB;
  auto go = [&](auto kernel) {
    hipLaunchKernelGGL(kernel, dim3(grid), dim3(kThreads), 0, st, c, a);
  };
  if (mode == 2) {
    // the mixed NR step: summed FP32 blocks and summed FP64 forces (in Fe)
    const int gm = (a.n_prisms + kMEPB - 1) / kMEPB;
    if (a.any_visc)
      hipLaunchKernelGGL(element_kernel_mix<true>, dim3(gm), dim3(kMThreads), 0, st, c, a);
    else
      hipLaunchKernelGGL(element_kernel_mix<false>, dim3(gm), dim3(kMThreads), 0, st, c, a);
    return hipGetLastError();
  }
  if (a.any_visc) {
    if (mode == 1) go(element_kernel<true, 1>);
    else go(element_kernel<true, 0>);
  } else {
    if (mode == 1) go(element_kernel<false, 1>);
    else go(element_kernel<false, 0>);
  }
  return hipGetLastError();
}

}  // namespace tf
