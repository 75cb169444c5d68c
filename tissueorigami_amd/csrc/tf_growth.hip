// Pre-step growth producers on the device: the relative positions of the
// prisms in the tissue's bounding box, and the growth deformation gradient
// update of Simulation::calculateGrowth with uniform (type 1) and grid-based
// (type 3) growth functions, the growth re-orientation by each prism's rigid
// rotation about z, and ShapeBase::growShapeByFg.
//
// The relative positions (once per step, from the step's start positions):
//   Simulation::calculateBoundingBox (Simulation.cpp:8388-8425),
//   ShapeBase::calculateRelativePosInBoundingBox (ShapeBase.cpp:2091-2095) of
//   the prism centre (ShapeBase::getCentre, ShapeBase.cpp:247-257), then the
//   apical prism's position for every prism of its column
//   (Simulation::updateRelativePositionsToApicalPositioning, Simulation.cpp:6147-6167).
// The growth of one prism (one thread):
//   its rigid rotation from the polar factor of the Gauss-averaged deformation
//   gradient (ShapeBase::CalculateGrowthRotationByF, ShapeBase.cpp:1913-1933,
//   calculate3DRotMatFromF :1999-2040, disassembleRotationMatrixForZ :1950-1996);
//   the growth increment of every active growth function
//   (Simulation::calculateGrowthUniform, Simulation.cpp:8765-8802;
//   calculateGrowthGridBased, :9038-9085; ShapeBase::calculateFgFromRates,
//   ShapeBase.cpp:471-505; calculateFgFromGridCorners, :507-658;
//   isGrowthRateApplicable, :297-335; updateGrowthIncrement, :764-790);
//   then F_g <- (SC (P G)) F_g with G rotated when the prism rotated, the
//   plastic (P) and shape-change (SC) increments identity, and InvFg by LU
//   (ShapeBase::growShapeByFg, ShapeBase.cpp:1452-1505).
// 3x3 products keep the reference's order (row by column, k ascending).
#include <cmath>

#include "tf_internal.h"

#pragma clang fp contract(off)

namespace tf {
namespace {

constexpr int kGB = 256;
// zeta of the Gauss points, 1 / pow(3, 0.5) as Prism.cpp:99-131 forms it, on the host
const double kGaussZ = 1.0 / std::pow(3.0, 0.5);

int blocks_for(int64_t n) {
  int64_t g = (n + kGB - 1) / kGB;
  if (g > 65535 * 4) g = 65535 * 4;
  return (int)std::max<int64_t>(g, 1);
}

__device__ inline void mm3(const double A[9], const double B[9], double C[9]) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double s = 0.0;
      for (int k = 0; k < 3; ++k) s += A[3 * i + k] * B[3 * k + j];
      C[3 * i + j] = s;
    }
}
// C = A^T B
__device__ inline void mtm3(const double A[9], const double B[9], double C[9]) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double s = 0.0;
      for (int k = 0; k < 3; ++k) s += A[3 * k + i] * B[3 * k + j];
      C[3 * i + j] = s;
    }
}
// C = A B^T
__device__ inline void mmt3(const double A[9], const double B[9], double C[9]) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double s = 0.0;
      for (int k = 0; k < 3; ++k) s += A[3 * i + k] * B[3 * j + k];
      C[3 * i + j] = s;
    }
}

__device__ inline double det3d(const double M[9]) {
  double det = 0.0;
  det = M[0] * (M[4] * M[8] - M[5] * M[7]);
  det -= M[1] * (M[3] * M[8] - M[5] * M[6]);
  det += M[2] * (M[3] * M[7] - M[4] * M[6]);
  return det;
}

// 3x3 inverse by LU with partial pivoting (the element kernel's, GSL's
// gsl_linalg_LU_decomp / LU_invert algorithm).
__device__ inline void inv3g(const double A[9], double X[9]) {
  double LU[9];
  int perm[3] = {0, 1, 2};
  for (int q = 0; q < 9; ++q) LU[q] = A[q];
  for (int j = 0; j < 2; ++j) {
    int piv = j;
    double mx = fabs(LU[3 * j + j]);
    for (int i = j + 1; i < 3; ++i)
      if (fabs(LU[3 * i + j]) > mx) { mx = fabs(LU[3 * i + j]); piv = i; }
    if (piv != j) {
      for (int k = 0; k < 3; ++k) {
        const double t = LU[3 * j + k]; LU[3 * j + k] = LU[3 * piv + k]; LU[3 * piv + k] = t;
      }
      const int t = perm[j]; perm[j] = perm[piv]; perm[piv] = t;
    }
    if (LU[3 * j + j] != 0.0)
      for (int i = j + 1; i < 3; ++i) {
        LU[3 * i + j] /= LU[3 * j + j];
        for (int k = j + 1; k < 3; ++k) LU[3 * i + k] -= LU[3 * i + j] * LU[3 * j + k];
      }
  }
  for (int c = 0; c < 3; ++c) {
    double y[3];
    for (int i = 0; i < 3; ++i) {
      double s = (perm[i] == c) ? 1.0 : 0.0;
      for (int k = 0; k < i; ++k) s -= LU[3 * i + k] * y[k];
      y[i] = s;
    }
    for (int i = 2; i >= 0; --i) {
      double s = y[i];
      for (int k = i + 1; k < 3; ++k) s -= LU[3 * i + k] * X[3 * k + c];
      X[3 * i + c] = s / LU[3 * i + i];
    }
  }
}

// Shape-function derivatives of the wedge at (eta, nu, zeta), rows d/d(eta),
// d/d(nu), d/d(zeta) (Prism.cpp:597-652).
__device__ inline void shape_der_g(double eta, double zeta, double nu, double D[3][6]) {
  const double alpha = (1 - zeta) / 2, beta = (1 + zeta) / 2, lam = 1 - eta - nu;
  const double r0[6] = {-alpha, alpha, 0, -beta, beta, 0};
  const double r1[6] = {-alpha, 0, alpha, -beta, 0, beta};
  const double r2[6] = {-lam / 2, -eta / 2, -nu / 2, lam / 2, eta / 2, nu / 2};
  for (int a = 0; a < 6; ++a) { D[0][a] = r0[a]; D[1][a] = r1[a]; D[2][a] = r2[a]; }
}

// TriPointF = sum over the six Gauss points of w (SFD pos)^T InvdXde
// (ShapeBase::calculateTriPointFForRatation, ShapeBase.cpp:1935-1948, with
// Prism::calculateCurrTriPointFForRotation, Prism.cpp:735-761).
__device__ void tri_point_F(const double ref[6][3], const double pos[6][3], double s3, double F[9]) {
  const double a = 1.0 / 6.0, b = 2.0 / 3.0;
  const double pts[6][3] = {{a, a, s3}, {b, a, s3}, {a, b, s3}, {a, a, -s3}, {b, a, -s3}, {a, b, -s3}};
  for (int q = 0; q < 9; ++q) F[q] = 0.0;
  for (int g = 0; g < 6; ++g) {
    double D[3][6];
    shape_der_g(pts[g][0], pts[g][2], pts[g][1], D);
    double dXde[9], inv[9], J[9], cur[9];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        double s = 0.0, t = 0.0;
        for (int k = 0; k < 6; ++k) {
          s += D[j][k] * ref[k][i];
          t += D[j][k] * pos[k][i];
        }
        dXde[3 * i + j] = s;
        J[3 * i + j] = t;
      }
    inv3g(dXde, inv);
    mm3(J, inv, cur);
    for (int q = 0; q < 9; ++q) F[q] = F[q] + cur[q] * (1.0 / 6.0);
  }
}

// Polar factor R = U V^T of F = U S V^T (the rotation the reference takes
// from gsl_linalg_SV_decomp; unique for det F > 0), by the Newton iteration
// X <- (X + X^-T) / 2 from X = F.  Returns false for det F <= 0.
__device__ bool polar_rotation(const double F[9], double R[9]) {
  if (!(det3d(F) > 0.0)) return false;
  double X[9];
  for (int q = 0; q < 9; ++q) X[q] = F[q];
  for (int it = 0; it < 60; ++it) {
    double Xi[9], Y[9];
    inv3g(X, Xi);
    double d = 0.0;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        Y[3 * i + j] = 0.5 * (X[3 * i + j] + Xi[3 * j + i]);
        d = fmax(d, fabs(Y[3 * i + j] - X[3 * i + j]));
      }
    for (int q = 0; q < 9; ++q) X[q] = Y[q];
    if (d <= 1e-15) break;
  }
  for (int q = 0; q < 9; ++q) R[q] = X[q];
  return true;
}

// The growth increment of one growth function for one source tissue,
// written into inc as the reference writes its increment matrix: the
// diagonal is set, then the matrix is rotated in place (uniform growth;
// grid growth when its interpolated angle is non-zero).  inc holds the
// caller's matrix on entry (identity, or an earlier branch's result).
__device__ void growth_increment(const GrowthFn& f, int source, const GrowthPrisms& gp, int e,
                                 const double* relpos, const double* cg, const double* ca,
                                 const uint8_t* ce, double dt, int interp, double inc[9]) {
  // isGrowthRateApplicable (ShapeBase.cpp:297-335)
  const double rz = gp.rel_z[e];
  const int tt = gp.type[e];
  double w = 0.0;
  bool ok = false;
  if (!(rz < f.zmin || rz > f.zmax)) {
    if (source == 0) {
      if (tt == 0) { w = 1.0; ok = true; }
      else if (tt == 2) { w = gp.weights[2 * e]; ok = true; }
    } else {
      if (tt == 1) { w = 1.0; ok = true; }
      else if (tt == 2) { w = gp.weights[2 * e + 1]; ok = true; }
    }
  }
  if (!ok) {
    for (int q = 0; q < 9; ++q) inc[q] = (q % 4 == 0) ? 1.0 : 0.0;
    return;
  }
  const bool actin = gp.flags[e] & kGrowActin;
  if (f.type == 1) {
    double x = f.rate[0], y = f.rate[1], z = f.rate[2];
    if (actin) { x = 0; y = 0; z = 0; }
    inc[0] = exp(x * w * dt);
    inc[4] = exp(y * w * dt);
    inc[8] = exp(z * w * dt);
    double T[9];
    mm3(f.rot, inc, T);
    mmt3(T, f.rot, inc);
    return;
  }
  // grid: the cell of the prism's relative position (convertRelativePosToGridIndex)
  double r0 = relpos[2 * e], r1 = relpos[2 * e + 1];
  r0 *= (float)(f.nx - 1);
  r1 *= (float)(f.ny - 1);
  int ix = (int)floor(r0), iy = (int)floor(r1);
  double fx = r0 - ix, fy = r1 - iy;
  if (ix >= f.nx - 1) { ix = f.nx - 2; fx = 1.0; }
  else if (ix < 0) { ix = 0; fx = 0.0; }
  if (iy >= f.ny - 1) { iy = f.ny - 2; fy = 1.0; }
  else if (iy < 0) { iy = 0; fy = 0.0; }
  const int64_t cell = f.cell_off + (int64_t)ix * (f.ny - 1) + iy;
  const double* gc = cg + 12 * cell;   // [corner][axis]
  const double* an = ca + 4 * cell;
  const uint8_t* el = ce + 4 * cell;
  double growth[3] = {0.0, 0.0, 0.0};
  double angle = 0.0;
  if (interp == 0) {
    const int c = fx > 0.5 ? (fy > 0.5 ? 3 : 1) : (fy > 0.5 ? 2 : 0);
    for (int i = 0; i < 3; ++i) growth[i] = gc[3 * c + i];
  } else if (interp == 1) {
    double fe = 0.0;
    if (el[0]) fe += (1.0 - fx) * (1.0 - fy);
    if (el[1]) fe += fx * (1.0 - fy);
    if (el[2]) fe += (1.0 - fx) * fy;
    if (el[3]) fe += fx * fy;
    angle = an[0] * (1.0 - fx) * (1.0 - fy) + an[1] * fx * (1.0 - fy) + an[2] * (1.0 - fx) * fy +
            an[3] * fx * fy;
    if (fe > 0) {
      if (fe >= 0.9999999) angle = 0.0;
      else angle /= (1.0 - fe);
    }
    for (int k = 0; k < 3; ++k) {
      growth[k] = gc[k] * (1.0 - fx) * (1.0 - fy) + gc[3 + k] * fx * (1.0 - fy) +
                  gc[6 + k] * (1.0 - fx) * fy + gc[9 + k] * fx * fy;
      growth[k] *= w;
    }
  }
  if (actin) growth[2] = 0;
  for (int k = 0; k < 3; ++k) inc[4 * k] = exp(growth[k] * dt);
  if (angle != 0.0) {
    const double c = cos(angle), s = sin(angle);
    const double Rm[9] = {c, -1.0 * s, 0.0, s, c, 0.0, 0.0, 0.0, 1.0};
    double T[9];
    mm3(Rm, inc, T);
    mmt3(T, Rm, inc);
  }
}

// Rotation of a 3x3 increment by the growth rotation: M <- R^T M R
// (ShapeBase::growShapeByFg, ShapeBase.cpp:1462-1468).
__device__ inline void rotate_back(const double gsr[9], double M[9]) {
  double T[9];
  mtm3(gsr, M, T);
  mm3(T, gsr, M);
}

// F_g <- (SC (P G)) F_g with P the identity, and InvFg (ShapeBase.cpp:1475-1486).
__device__ inline void grow_by(const double G[9], const double SC[9], double* Fg, double* invFg) {
  const double P[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
  double t1[9], t2[9], t3[9], Fo[9], X[9];
  for (int q = 0; q < 9; ++q) Fo[q] = Fg[q];
  mm3(P, G, t1);
  mm3(SC, t1, t2);
  mm3(t2, Fo, t3);
  inv3g(t3, X);
  for (int q = 0; q < 9; ++q) {
    Fg[q] = t3[q];
    invFg[q] = X[q];
  }
}

// stage 0: the growth of every prism, applied at once.
// stage 1 (column volume conservation): the increments are rotated as
// ShapeBase::tempgrowShapeByFg does (ShapeBase.cpp:1396-1450, in place, so
// growShapeByFg rotates them a second time) and kept in Gs / SCs with the
// rotation, and vol gets the grown volume det(SC (P G) F_g) V_ref; F_g is not
// changed.  An ablated prism keeps det(F_g) V_ref, the volume its last
// growShapeByFg left.
__global__ __launch_bounds__(kGB) void growth_kernel(int32_t n_prisms, const int32_t* conn,
                                                     const double* refpos, const double* u,
                                                     GrowthPrisms gp, const GrowthFn* fns, int n_fn,
                                                     const double* cg, const double* ca,
                                                     const uint8_t* ce, const double* relpos,
                                                     double t, double dt, int interp,
                                                     const uint8_t* ablated, double gauss_z,
                                                     double* Fg, double* invFg, GrowthStage gs) {
  for (int e = blockIdx.x * kGB + threadIdx.x; e < n_prisms; e += gridDim.x * kGB) {
    if (ablated && ablated[e]) {  // Simulation.cpp:6055, 4416
      if (gs.on) gs.vol[e] = det3d(Fg + 9 * (size_t)e) * gs.vref[e];
      continue;
    }
    // rigid rotation about z at the step's start positions
    double ref[6][3], pos[6][3];
    for (int a = 0; a < 6; ++a) {
      const int nd = conn[6 * (size_t)e + a];
      for (int i = 0; i < 3; ++i) {
        ref[a][i] = refpos[18 * (size_t)e + 3 * a + i];
        pos[a][i] = u[3 * (size_t)nd + i];
      }
    }
    double F[9], R[9];
    tri_point_F(ref, pos, gauss_z, F);
    bool rotated = false;
    double gsr[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    if (polar_rotation(F, R)) {
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
          if (i != j && (R[3 * i + j] > 0.017 || R[3 * i + j] < (-1.0 * 0.017))) rotated = true;
      if (rotated) {
        const double th = atan2(R[1], R[0]);
        if (th > 0.017 || th < -0.017) {
          const double c = cos(th), s = sin(th);
          // the rotation matrix, transposed (GrowthStrainsRotMat)
          const double rm[9] = {c, -1.0 * s, 0.0, s, c, 0.0, 0.0, 0.0, 1.0};
          for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) gsr[3 * i + j] = rm[3 * j + i];
        } else {
          rotated = false;
        }
      }
    }
    // the growth increment of the active functions (cleanUpGrowthRates: identity)
    double G[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    const int tt = gp.type[e], pl = gp.place[e];
    const uint8_t fl = gp.flags[e];
    const bool ecm = fl & kGrowEcm, ecm_circ = fl & kGrowEcmCirc;
    for (int k = 0; k < n_fn; ++k) {
      const GrowthFn& f = fns[k];
      if (!(t >= f.t0 && t < f.t1)) continue;
      double col[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1}, per[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
      if ((f.apply & kApplyBasalEcm) && ecm && pl == 0)
        growth_increment(f, 0, gp, e, relpos, cg, ca, ce, dt, interp, col);
      if ((f.apply & kApplyLateralEcm) && ecm_circ && !(pl == 0))
        growth_increment(f, 0, gp, e, relpos, cg, ca, ce, dt, interp, col);
      if (!gp.explicit_ecm || !ecm) {
        if (f.apply & kApplyColumnar) growth_increment(f, 0, gp, e, relpos, cg, ca, ce, dt, interp, col);
        if (f.apply & kApplyPeripodial) growth_increment(f, 1, gp, e, relpos, cg, ca, ce, dt, interp, per);
      }
      double T[9];
      if (tt == 0) {
        mm3(col, G, T);
        for (int q = 0; q < 9; ++q) G[q] = T[q];
      } else if (tt == 1) {
        mm3(per, G, T);
        for (int q = 0; q < 9; ++q) G[q] = T[q];
      } else if (tt == 2) {
        mm3(col, G, T);
        mm3(per, T, G);
      }
    }
    double SC[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    if (rotated) {
      rotate_back(gsr, G);
      rotate_back(gsr, SC);
    }
    double* Fe = Fg + 9 * (size_t)e;
    if (!gs.on) {
      grow_by(G, SC, Fe, invFg + 9 * (size_t)e);
      continue;
    }
    const double P[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    double t1[9], t2[9], t3[9];
    mm3(P, G, t1);
    mm3(SC, t1, t2);
    mm3(t2, Fe, t3);
    gs.vol[e] = det3d(t3) * gs.vref[e];
    for (int q = 0; q < 9; ++q) {
      gs.G[9 * (size_t)e + q] = G[q];
      gs.SC[9 * (size_t)e + q] = SC[q];
      gs.gsr[9 * (size_t)e + q] = gsr[q];
    }
    gs.rot[e] = rotated ? 1 : 0;
  }
}

// Simulation::conserveColumnVolume (Simulation.cpp:6518-6600): per column (the
// columnar prisms of an apical prism's stack, ECM and actin prisms left out),
// the factor (sum of reference volumes) / (sum of grown volumes), in stack
// order, and its cube root on the diagonal of an increment applied before each
// member's growth increment (ShapeBase::updateGrowthIncrement, ShapeBase.cpp:764-790,
// as for a columnar prism).
__global__ __launch_bounds__(kGB) void column_volume_kernel(int32_t n_columns, int32_t layers,
                                                            const int32_t* columns,
                                                            const int32_t* type, GrowthStage gs) {
  for (int c = blockIdx.x * kGB + threadIdx.x; c < n_columns; c += gridDim.x * kGB) {
    const int32_t* m = columns + (size_t)c * layers;
    double grown = 0, ref = 0;
    for (int i = 0; i < layers; ++i) {
      if (m[i] < 0) continue;
      grown += gs.vol[m[i]];
      ref += gs.vref[m[i]];
    }
    const double f = ref / grown;
    const double g = pow(f, 1.0 / 3.0);
    const double D[9] = {g, 0, 0, 0, g, 0, 0, 0, g};
    const double I[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    for (int i = 0; i < layers; ++i) {
      const int e = m[i];
      if (e < 0) continue;
      double Ge[9], T[9];
      for (int q = 0; q < 9; ++q) Ge[q] = gs.G[9 * (size_t)e + q];
      const int tt = type[e];
      if (tt == 0 || tt == 2) {
        mm3(D, Ge, T);
        if (tt == 2) mm3(I, T, Ge);
        else for (int q = 0; q < 9; ++q) Ge[q] = T[q];
      } else if (tt == 1) {
        mm3(I, Ge, T);
        for (int q = 0; q < 9; ++q) Ge[q] = T[q];
      }
      for (int q = 0; q < 9; ++q) gs.G[9 * (size_t)e + q] = Ge[q];
    }
  }
}

// ShapeBase::growShapeByFg after the conservation: the kept increments rotated
// (again) when the prism rotated, then F_g and InvFg.
__global__ __launch_bounds__(kGB) void grow_staged_kernel(int32_t n_prisms, const uint8_t* ablated,
                                                          GrowthStage gs, double* Fg,
                                                          double* invFg) {
  for (int e = blockIdx.x * kGB + threadIdx.x; e < n_prisms; e += gridDim.x * kGB) {
    if (ablated && ablated[e]) continue;
    double G[9], SC[9], gsr[9];
    for (int q = 0; q < 9; ++q) {
      G[q] = gs.G[9 * (size_t)e + q];
      SC[q] = gs.SC[9 * (size_t)e + q];
      gsr[q] = gs.gsr[9 * (size_t)e + q];
    }
    if (gs.rot[e]) {
      rotate_back(gsr, G);
      rotate_back(gsr, SC);
    }
    grow_by(G, SC, Fg + 9 * (size_t)e, invFg + 9 * (size_t)e);
  }
}

// Bounding box of the kept nodes of one tissue (one workgroup per tissue,
// nodes [node_rng[2 T], node_rng[2 T + 1])) with the reference's `else if` (a node
// that lowers a coordinate's minimum is not tested against the maximum):
// plain minima and maxima over the nodes after the first kept one, and the
// sequential rule when that first node holds the strict maximum.
// bb[6 T ..] = {min x, y, z, max x, y, z}.
__global__ __launch_bounds__(1024) void relpos_bbox_kernel(const int32_t* node_rng, const double* u,
                                                           const uint8_t* excluded,
                                                           int symmetric_y, double* bb) {
  __shared__ double lo_s[1024][3], hi_s[1024][3];
  __shared__ int first_s[1024];
  const int T = blockIdx.x;
  const int n0 = node_rng[2 * T], n1 = node_rng[2 * T + 1];
  int first = n1;
  for (int k = n0 + (int)threadIdx.x; k < n1; k += 1024)
    if (!excluded || !excluded[k]) { first = k; break; }
  first_s[threadIdx.x] = first;
  __syncthreads();
  for (int w = 512; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) first_s[threadIdx.x] = min(first_s[threadIdx.x], first_s[threadIdx.x + w]);
    __syncthreads();
  }
  const int k0 = first_s[0];
  double lo[3] = {100000.0, 100000.0, 100000.0}, hi[3] = {-1e300, -1e300, -1e300};
  for (int k = n0 + (int)threadIdx.x; k < n1; k += 1024) {
    if (excluded && excluded[k]) continue;
    for (int i = 0; i < 3; ++i) {
      const double v = u[3 * (size_t)k + i];
      lo[i] = fmin(lo[i], v);
      if (k > k0) hi[i] = fmax(hi[i], v);
    }
  }
  for (int i = 0; i < 3; ++i) {
    lo_s[threadIdx.x][i] = lo[i];
    hi_s[threadIdx.x][i] = hi[i];
  }
  __syncthreads();
  for (int w = 512; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int i = 0; i < 3; ++i) {
        lo_s[threadIdx.x][i] = fmin(lo_s[threadIdx.x][i], lo_s[threadIdx.x + w][i]);
        hi_s[threadIdx.x][i] = fmax(hi_s[threadIdx.x][i], hi_s[threadIdx.x + w][i]);
      }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    double* out = bb + 6 * (size_t)T;
    for (int i = 0; i < 3; ++i) {
      double l = lo_s[0][i], h = hi_s[0][i];
      const double v0 = k0 < n1 ? u[3 * (size_t)k0 + i] : 100000.0;
      if (!(h >= v0) || !(v0 < 100000.0)) {
        l = 100000.0;
        h = -100000.0;
        for (int k = n0; k < n1; ++k) {
          if (excluded && excluded[k]) continue;
          const double v = u[3 * (size_t)k + i];
          if (v < l) l = v;
          else if (v > h) h = v;
        }
      } else {
        h = fmax(h, -100000.0);
      }
      out[i] = l;
      out[3 + i] = h;
    }
    if (symmetric_y) out[1] = (-1.0) * out[4];
  }
}

// Centre of every prism relative to its tissue's box (x, y).
__global__ __launch_bounds__(kGB) void relpos_kernel(int32_t n_prisms, const int32_t* conn,
                                                     const double* u, const int32_t* prism_tissue,
                                                     const double* bbs, double* rel) {
  for (int e = blockIdx.x * kGB + threadIdx.x; e < n_prisms; e += gridDim.x * kGB) {
    const double* bb = bbs + 6 * (size_t)(prism_tissue ? prism_tissue[e] : 0);
    const double bx = bb[0], by = bb[1], sx = bb[3] - bb[0], sy = bb[4] - bb[1];
    double c[3] = {0.0, 0.0, 0.0};
    for (int a = 0; a < 6; ++a) {
      const int nd = conn[6 * (size_t)e + a];
      for (int j = 0; j < 3; ++j) c[j] += u[3 * (size_t)nd + j];
    }
    c[0] /= 6; c[1] /= 6;
    rel[2 * (size_t)e] = (c[0] - bx) / sx;
    rel[2 * (size_t)e + 1] = (c[1] - by) / sy;
  }
}

// Every prism takes the position of its column's apical prism.
__global__ __launch_bounds__(kGB) void relpos_apical_kernel(int32_t n_prisms,
                                                            const int32_t* apical_of,
                                                            const double* raw, double* rel) {
  for (int e = blockIdx.x * kGB + threadIdx.x; e < n_prisms; e += gridDim.x * kGB) {
    const int a = apical_of ? apical_of[e] : -1;
    const int s = a >= 0 ? a : e;
    rel[2 * (size_t)e] = raw[2 * (size_t)s];
    rel[2 * (size_t)e + 1] = raw[2 * (size_t)s + 1];
  }
}

}  // namespace

// GridBasedGrowthFunction's constructor and preCalculateAnglesForCompatibleAveraging
// (GrowthFunctionTypes.h:290-546): per grid cell (ix, iy) < (nx - 1, ny - 1),
// the growth rates at its four corners (ix, iy), (ix+1, iy), (ix, iy+1),
// (ix+1, iy+1), their orientation angles brought within 45 degrees of the first
// corner whose aspect ratio counts (x and y rates swapped where the angle
// moves by 90 degrees), and which corners' angles are eliminated (aspect
// ratio below 1.1).  rates[(i * ny + j) * 3 + k] in 1/s, angles_deg[i * ny + j].
void grid_compatible(int nx, int ny, const double* rates, const double* angles_deg, double* cg,
                     double* ca, uint8_t* ce) {
  const double kPi = M_PI;
  std::vector<double> ang((size_t)nx * ny);
  std::vector<char> ar_over((size_t)nx * ny, 1);
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j < ny; ++j) {
      const size_t ij = (size_t)i * ny + j;
      ang[ij] = angles_deg[ij] * kPi / 180.0;
      double AR = rates[3 * ij] / rates[3 * ij + 1];
      if (AR < 0) AR *= -1.0;
      if (AR < 1.0) AR = 1 / AR;
      if (AR < 1.1) ar_over[ij] = 0;
    }
  for (int ix = 0; ix < nx - 1; ++ix)
    for (int iy = 0; iy < ny - 1; ++iy) {
      const size_t corner[4] = {(size_t)ix * ny + iy, (size_t)(ix + 1) * ny + iy,
                                (size_t)ix * ny + iy + 1, (size_t)(ix + 1) * ny + iy + 1};
      double g[4][3], a[4] = {0.0, 0.0, 0.0, 0.0};
      bool elim[4] = {false, false, false, false};
      for (int c = 0; c < 4; ++c) {
        for (int k = 0; k < 3; ++k) g[c][k] = rates[3 * corner[c] + k];
        if (ar_over[corner[c]]) a[c] = ang[corner[c]];
        else elim[c] = true;
      }
      for (int c = 0; c < 4; ++c) {
        while (a[c] < 0) a[c] += kPi;
        while (a[c] > kPi) a[c] -= kPi;
      }
      bool selected = false;
      int sel = 3;
      double tet = 0;
      for (int c = 0; c < 4; ++c) {
        if (!elim[c] && !selected) {
          tet = a[c];
          selected = true;
        }
        if (selected) {
          sel = c;
          break;
        }
      }
      bool flip[4] = {false, false, false, false};
      for (int c = sel; c < 4; ++c) {
        const double d = a[c] - tet;
        if (d > 0) {
          if (d >= kPi / 4.0 && d < 3.0 * kPi / 4.0) { a[c] = a[c] - kPi / 2.0; flip[c] = true; }
          else if (d >= 3.0 * kPi / 4.0 && d < 5.0 * kPi / 4.0) a[c] = a[c] - kPi;
          else if (d >= 5.0 * kPi / 4.0 && d < 7.0 * kPi / 4.0) { a[c] = a[c] - 6.0 * kPi / 4.0; flip[c] = true; }
          else if (d >= 7.0 * kPi / 4.0) a[c] = a[c] - 2.0 * kPi;
        } else if (d < 0) {
          if (d <= -kPi / 4.0 && d > -3.0 * kPi / 4.0) { a[c] = a[c] + kPi / 2.0; flip[c] = true; }
          else if (d <= -3.0 * kPi / 4.0 && d > -5.0 * kPi / 4.0) a[c] = a[c] + kPi;
          else if (d <= -5.0 * kPi / 4.0 && d > -7.0 * kPi / 4.0) { a[c] = a[c] + 6.0 * kPi / 4.0; flip[c] = true; }
          else if (d <= -7.0 * kPi / 4.0) a[c] = a[c] + 2.0 * kPi;
        }
      }
      for (int c = 0; c < 4; ++c)
        if (flip[c]) std::swap(g[c][0], g[c][1]);
      const size_t cell = (size_t)ix * (ny - 1) + iy;
      for (int c = 0; c < 4; ++c) {
        for (int k = 0; k < 3; ++k) cg[12 * cell + 3 * c + k] = g[c][k];
        ca[4 * cell + c] = a[c];
        ce[4 * cell + c] = elim[c] ? 1 : 0;
      }
    }
}

hipError_t launch_relative_positions(int32_t n_tissues, const int32_t* node_rng, int32_t n_prisms,
                                     const int32_t* conn, const int32_t* prism_tissue,
                                     const double* u, const uint8_t* excluded, int symmetric_y,
                                     const int32_t* apical_of, double* bb, double* raw,
                                     double* rel, hipStream_t st) {
  hipLaunchKernelGGL(relpos_bbox_kernel, dim3(n_tissues), dim3(1024), 0, st, node_rng, u,
                     excluded, symmetric_y, bb);
  hipLaunchKernelGGL(relpos_kernel, dim3(blocks_for(n_prisms)), dim3(kGB), 0, st, n_prisms, conn,
                     u, prism_tissue, (const double*)bb, raw);
  hipLaunchKernelGGL(relpos_apical_kernel, dim3(blocks_for(n_prisms)), dim3(kGB), 0, st, n_prisms,
                     apical_of, (const double*)raw, rel);
  return hipGetLastError();
}

hipError_t launch_growth(int32_t n_prisms, const int32_t* conn, const double* refpos,
                         const double* u, const GrowthPrisms& gp, const GrowthFn* fns, int n_fn,
                         const double* cg, const double* ca, const uint8_t* ce, const double* relpos,
                         double t, double dt, int interp, const uint8_t* ablated, double* Fg,
                         double* invFg, const GrowthStage& gs, int32_t n_columns, int32_t layers,
                         const int32_t* columns, hipStream_t st) {
  hipLaunchKernelGGL(growth_kernel, dim3(blocks_for(n_prisms)), dim3(kGB), 0, st, n_prisms, conn,
                     refpos, u, gp, fns, n_fn, cg, ca, ce, relpos, t, dt, interp, ablated,
                     kGaussZ, Fg, invFg, gs);
  if (gs.on) {
    hipLaunchKernelGGL(column_volume_kernel, dim3(blocks_for(n_columns)), dim3(kGB), 0, st,
                       n_columns, layers, columns, gp.type, gs);
    hipLaunchKernelGGL(grow_staged_kernel, dim3(blocks_for(n_prisms)), dim3(kGB), 0, st, n_prisms,
                       ablated, gs, Fg, invFg);
  }
  return hipGetLastError();
}

}  // namespace tf
