// CPU oracle of the pre-step growth producers (TEST INFRASTRUCTURE ONLY, built
// into liboracle.so beside oracle.cpp; the product library never uses it).
//
// Restates, in plain C++ with contraction off:
//   - the relative positions of the prisms in the tissue's bounding box
//     (Simulation::calculateBoundingBox, Simulation.cpp:8388-8428, with its
//     sequential `else if`; ShapeBase::getCentre, ShapeBase.cpp:247-257;
//     ShapeBase::calculateRelativePosInBoundingBox, ShapeBase.cpp:2091-2095;
//     Simulation::updateRelativePositionsToApicalPositioning, Simulation.cpp:6147-6167);
//   - the compatible corner data of a growth grid
//     (GridBasedGrowthFunction, GrowthFunctionTypes.h:290-546);
//   - the growth rotation about z (ShapeBase::CalculateGrowthRotationByF,
//     ShapeBase.cpp:1913-2040), whose polar factor the reference takes from
//     GSL's SVD (R = U V^T); here from the eigen-decomposition of F^T F by
//     cyclic Jacobi rotations (R = F V diag(1/sqrt(lambda)) V^T), a method
//     independent of the product's Newton iteration;
//   - one step of Simulation::calculateGrowth (Simulation.cpp:8640-8660) with
//     uniform (:8762-8802) and grid-based (:9038-9085) functions
//     (ShapeBase::calculateFgFromRates :471-505, calculateFgFromGridCorners
//     :507-658, isGrowthRateApplicable :297-335, updateGrowthIncrement
//     :764-790), the column volume conservation (Simulation.cpp:6040-6046,
//     6518-6600; ShapeBase::tempgrowShapeByFg, ShapeBase.cpp:1396-1450) and
//     ShapeBase::growShapeByFg (ShapeBase.cpp:1452-1505).
//
// Pinning (tests/test_oracle_growth.py): polar factor against numpy's SVD;
// grid corner data against known answers (aspect-ratio elimination, the
// 45-degree window, x/y swaps); the growth step against closed forms (an
// unrotated prism grows by diag(exp(r w dt)); a rigidly turned tissue grows in
// its turned frame; a constant grid equals uniform growth; the conserved
// column keeps its reference volume) and against the Save_Growth fixtures of
// run07007 (grid growth, column conservation) and run07008 (uniform growth,
// rotated prisms).
#include <cmath>
#include <cstdint>
#include <cstring>
#include <utility>
#include <vector>

namespace {

typedef double M3[3][3];

void mul(const M3 A, const M3 B, M3 C) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double s = 0.0;
      for (int k = 0; k < 3; ++k) s += A[i][k] * B[k][j];
      C[i][j] = s;
    }
}
// A^T B
void mul_tn(const M3 A, const M3 B, M3 C) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double s = 0.0;
      for (int k = 0; k < 3; ++k) s += A[k][i] * B[k][j];
      C[i][j] = s;
    }
}
// A B^T
void mul_nt(const M3 A, const M3 B, M3 C) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double s = 0.0;
      for (int k = 0; k < 3; ++k) s += A[i][k] * B[j][k];
      C[i][j] = s;
    }
}
void eye(M3 A) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) A[i][j] = i == j ? 1.0 : 0.0;
}
void copy(const M3 A, M3 B) { std::memcpy(B, A, sizeof(M3)); }

double det(const M3 M) {
  double d = 0.0;
  d = M[0][0] * (M[1][1] * M[2][2] - M[1][2] * M[2][1]);
  d -= M[0][1] * (M[1][0] * M[2][2] - M[1][2] * M[2][0]);
  d += M[0][2] * (M[1][0] * M[2][1] - M[1][1] * M[2][0]);
  return d;
}

// LU with partial pivoting, then L U X = P I (GSL's LU_decomp / LU_invert).
void inverse(const M3 A, M3 X) {
  double a[3][3];
  int p[3] = {0, 1, 2};
  std::memcpy(a, A, sizeof(a));
  for (int c = 0; c < 2; ++c) {
    int r = c;
    for (int i = c + 1; i < 3; ++i)
      if (std::fabs(a[i][c]) > std::fabs(a[r][c])) r = i;
    if (r != c) {
      for (int k = 0; k < 3; ++k) std::swap(a[c][k], a[r][k]);
      std::swap(p[c], p[r]);
    }
    if (a[c][c] == 0.0) continue;
    for (int i = c + 1; i < 3; ++i) {
      a[i][c] /= a[c][c];
      for (int k = c + 1; k < 3; ++k) a[i][k] -= a[i][c] * a[c][k];
    }
  }
  for (int col = 0; col < 3; ++col) {
    double y[3];
    for (int i = 0; i < 3; ++i) {
      double s = p[i] == col ? 1.0 : 0.0;
      for (int k = 0; k < i; ++k) s -= a[i][k] * y[k];
      y[i] = s;
    }
    for (int i = 2; i >= 0; --i) {
      double s = y[i];
      for (int k = i + 1; k < 3; ++k) s -= a[i][k] * X[k][col];
      X[i][col] = s / a[i][i];
    }
  }
}

// Eigen-decomposition of a symmetric 3x3 matrix by cyclic Jacobi rotations:
// S = V diag(l) V^T.
void sym_eigen(const M3 S, double l[3], M3 V) {
  M3 a;
  copy(S, a);
  eye(V);
  for (int sweep = 0; sweep < 100; ++sweep) {
    const double off = a[0][1] * a[0][1] + a[0][2] * a[0][2] + a[1][2] * a[1][2];
    const double scale = a[0][0] * a[0][0] + a[1][1] * a[1][1] + a[2][2] * a[2][2];
    if (off <= 1e-34 * scale) break;
    for (int p = 0; p < 2; ++p)
      for (int q = p + 1; q < 3; ++q) {
        if (a[p][q] == 0.0) continue;
        const double th = (a[q][q] - a[p][p]) / (2.0 * a[p][q]);
        const double t = (th >= 0 ? 1.0 : -1.0) / (std::fabs(th) + std::sqrt(th * th + 1.0));
        const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < 3; ++k) {  // a <- J^T a J
          const double akp = a[k][p], akq = a[k][q];
          a[k][p] = c * akp - s * akq;
          a[k][q] = s * akp + c * akq;
        }
        for (int k = 0; k < 3; ++k) {
          const double apk = a[p][k], aqk = a[q][k];
          a[p][k] = c * apk - s * aqk;
          a[q][k] = s * apk + c * aqk;
        }
        for (int k = 0; k < 3; ++k) {
          const double vkp = V[k][p], vkq = V[k][q];
          V[k][p] = c * vkp - s * vkq;
          V[k][q] = s * vkp + c * vkq;
        }
      }
  }
  for (int i = 0; i < 3; ++i) l[i] = a[i][i];
}

// Rotation factor of F = R U (R = U_svd V_svd^T for det F > 0).
bool polar(const M3 F, M3 R) {
  if (!(det(F) > 0.0)) return false;
  M3 C, V;
  mul_tn(F, F, C);
  double l[3];
  sym_eigen(C, l, V);
  M3 W;  // V diag(1/sqrt(l)) V^T
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double s = 0.0;
      for (int k = 0; k < 3; ++k) s += V[i][k] * V[j][k] / std::sqrt(l[k]);
      W[i][j] = s;
    }
  mul(F, W, R);
  return true;
}

// Gauss-averaged deformation gradient of the current positions against the
// reference shape, without F_g (ShapeBase::calculateTriPointFForRatation,
// ShapeBase.cpp:1935-1948; Prism::calculateCurrTriPointFForRotation,
// Prism.cpp:735-761; shape-function derivatives Prism.cpp:597-652).
void tri_point_f(const double ref[6][3], const double pos[6][3], M3 F) {
  const double z = 1.0 / std::pow(3.0, 0.5);
  const double a = 1.0 / 6.0, b = 2.0 / 3.0;
  const double pts[6][3] = {{a, a, z}, {b, a, z}, {a, b, z}, {a, a, -z}, {b, a, -z}, {a, b, -z}};
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) F[i][j] = 0.0;
  for (int g = 0; g < 6; ++g) {
    const double eta = pts[g][0], nu = pts[g][1], zeta = pts[g][2];
    const double al = (1 - zeta) / 2, be = (1 + zeta) / 2, la = 1 - eta - nu;
    const double D[3][6] = {{-al, al, 0, -be, be, 0},
                            {-al, 0, al, -be, 0, be},
                            {-la / 2, -eta / 2, -nu / 2, la / 2, eta / 2, nu / 2}};
    M3 dX, dx, inv, cur;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        double s = 0.0, t = 0.0;
        for (int k = 0; k < 6; ++k) {
          s += D[j][k] * ref[k][i];
          t += D[j][k] * pos[k][i];
        }
        dX[i][j] = s;
        dx[i][j] = t;
      }
    inverse(dX, inv);
    mul(dx, inv, cur);
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) F[i][j] = F[i][j] + cur[i][j] * (1.0 / 6.0);
  }
}

// GrowthStrainsRotMat of a prism (ShapeBase::CalculateGrowthRotationByF,
// ShapeBase.cpp:1913-1933; disassembleRotationMatrixForZ :1950-1996): the
// transposed z rotation by theta = atan2(R01, R00) when an off-diagonal
// entry of R and theta exceed 0.017 in magnitude.
bool growth_rotation(const double ref[6][3], const double pos[6][3], M3 gsr) {
  M3 F, R;
  tri_point_f(ref, pos, F);
  if (!polar(F, R)) return false;
  bool any = false;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      if (i != j && std::fabs(R[i][j]) > 0.017) any = true;
  if (!any) return false;
  const double th = std::atan2(R[0][1], R[0][0]);
  if (!(std::fabs(th) > 0.017)) return false;
  const double c = std::cos(th), s = std::sin(th);
  gsr[0][0] = c;  gsr[0][1] = s;  gsr[0][2] = 0.0;
  gsr[1][0] = -1.0 * s; gsr[1][1] = c; gsr[1][2] = 0.0;
  gsr[2][0] = 0.0; gsr[2][1] = 0.0; gsr[2][2] = 1.0;
  return true;
}

// Compatible corner data of one grid (GrowthFunctionTypes.h:310-546), for cell
// (ix, iy) at cell index ix * (ny - 1) + iy.
struct Grid {
  int nx = 0, ny = 0;
  std::vector<double> g;   // [cell][corner][axis]
  std::vector<double> a;   // [cell][corner]
  std::vector<uint8_t> e;  // [cell][corner]
};

void compatible(int nx, int ny, const double* rates, const double* angles_deg, Grid& G) {
  G.nx = nx;
  G.ny = ny;
  const int cells = (nx - 1) * (ny - 1);
  G.g.assign(12 * (size_t)cells, 0.0);
  G.a.assign(4 * (size_t)cells, 0.0);
  G.e.assign(4 * (size_t)cells, 0);
  auto rate = [&](int i, int j, int k) { return rates[((size_t)i * ny + j) * 3 + k]; };
  auto keeps_angle = [&](int i, int j) {  // aspect ratio of x / y rates at least 1.1
    double ar = rate(i, j, 0) / rate(i, j, 1);
    if (ar < 0) ar *= -1.0;
    if (ar < 1.0) ar = 1 / ar;
    return !(ar < 1.1);
  };
  const double pi = M_PI;
  for (int ix = 0; ix + 1 < nx; ++ix)
    for (int iy = 0; iy + 1 < ny; ++iy) {
      const int ci[4] = {ix, ix + 1, ix, ix + 1}, cj[4] = {iy, iy, iy + 1, iy + 1};
      double gr[4][3], an[4];
      bool el[4];
      for (int c = 0; c < 4; ++c) {
        for (int k = 0; k < 3; ++k) gr[c][k] = rate(ci[c], cj[c], k);
        el[c] = !keeps_angle(ci[c], cj[c]);
        an[c] = el[c] ? 0.0 : angles_deg[(size_t)ci[c] * ny + cj[c]] * pi / 180.0;
        while (an[c] < 0) an[c] += pi;
        while (an[c] > pi) an[c] -= pi;
      }
      // the first corner whose angle counts is the base (corner 3 when none)
      int base = 3;
      for (int c = 0; c < 4; ++c)
        if (!el[c]) { base = c; break; }
      const double tet = base < 3 || !el[3] ? an[base] : 0.0;
      for (int c = base; c < 4; ++c) {
        const double d = an[c] - tet;
        bool swap_xy = false;
        if (d > 0) {
          if (d >= pi / 4.0 && d < 3.0 * pi / 4.0) { an[c] -= pi / 2.0; swap_xy = true; }
          else if (d >= 3.0 * pi / 4.0 && d < 5.0 * pi / 4.0) an[c] -= pi;
          else if (d >= 5.0 * pi / 4.0 && d < 7.0 * pi / 4.0) { an[c] -= 6.0 * pi / 4.0; swap_xy = true; }
          else if (d >= 7.0 * pi / 4.0) an[c] -= 2.0 * pi;
        } else if (d < 0) {
          if (d <= -pi / 4.0 && d > -3.0 * pi / 4.0) { an[c] += pi / 2.0; swap_xy = true; }
          else if (d <= -3.0 * pi / 4.0 && d > -5.0 * pi / 4.0) an[c] += pi;
          else if (d <= -5.0 * pi / 4.0 && d > -7.0 * pi / 4.0) { an[c] += 6.0 * pi / 4.0; swap_xy = true; }
          else if (d <= -7.0 * pi / 4.0) an[c] += 2.0 * pi;
        }
        if (swap_xy) std::swap(gr[c][0], gr[c][1]);
      }
      const size_t cell = (size_t)ix * (ny - 1) + iy;
      for (int c = 0; c < 4; ++c) {
        for (int k = 0; k < 3; ++k) G.g[12 * cell + 3 * c + k] = gr[c][k];
        G.a[4 * cell + c] = an[c];
        G.e[4 * cell + c] = el[c];
      }
    }
}

struct Fn {
  int type, apply;
  double t0, t1, zmin, zmax;
  double rate[3];
  M3 rot;
  int grid = -1;
};

struct PrismInfo {
  const int32_t* type;
  const int32_t* place;
  const uint8_t* flags;
  const double* weights;
  const double* rel_z;
  int explicit_ecm;
};

// isGrowthRateApplicable (ShapeBase.cpp:297-335): the weight, or < 0 when the
// function does not apply to this prism for this source tissue.
double applicable_weight(const PrismInfo& P, int e, int source, double zmin, double zmax) {
  if (P.rel_z[e] < zmin || P.rel_z[e] > zmax) return -1.0;
  const int tt = P.type[e];
  if (tt == source) return 1.0;
  if (tt == 2) return P.weights ? P.weights[2 * e + source] : 0.0;
  return -1.0;
}

// One function's increment for one source tissue (calculateFgFromRates /
// calculateFgFromGridCorners), written into inc.
void increment(const Fn& f, const std::vector<Grid>& grids, const PrismInfo& P, int e,
               int source, const double* rel, double dt, int interp, M3 inc) {
  const double w = applicable_weight(P, e, source, f.zmin, f.zmax);
  if (w < 0) {
    eye(inc);
    return;
  }
  const bool actin = P.flags[e] & 4;
  if (f.type == 1) {
    const double x = actin ? 0.0 : f.rate[0], y = actin ? 0.0 : f.rate[1],
                 z = actin ? 0.0 : f.rate[2];
    inc[0][0] = std::exp(x * w * dt);
    inc[1][1] = std::exp(y * w * dt);
    inc[2][2] = std::exp(z * w * dt);
    M3 T;
    mul(f.rot, inc, T);
    mul_nt(T, f.rot, inc);
    return;
  }
  const Grid& G = grids[f.grid];
  // ShapeBase::convertRelativePosToGridIndex (ShapeBase.cpp:834-860)
  double px = rel[2 * e], py = rel[2 * e + 1];
  px *= (float)(G.nx - 1);
  py *= (float)(G.ny - 1);
  int ix = (int)std::floor(px), iy = (int)std::floor(py);
  double fx = px - ix, fy = py - iy;
  if (ix >= G.nx - 1) { ix = G.nx - 2; fx = 1.0; }
  else if (ix < 0) { ix = 0; fx = 0.0; }
  if (iy >= G.ny - 1) { iy = G.ny - 2; fy = 1.0; }
  else if (iy < 0) { iy = 0; fy = 0.0; }
  const size_t cell = (size_t)ix * (G.ny - 1) + iy;
  const double* g = &G.g[12 * cell];
  const double* an = &G.a[4 * cell];
  const uint8_t* el = &G.e[4 * cell];
  // bilinear weights of the corners, each product rounded in the reference's
  // left-to-right order: v * (1 - fx) * (1 - fy) + v * fx * (1 - fy) + ...
  auto blend = [&](const double v[4]) {
    return v[0] * (1.0 - fx) * (1.0 - fy) + v[1] * fx * (1.0 - fy) + v[2] * (1.0 - fx) * fy +
           v[3] * fx * fy;
  };
  double gr[3] = {0.0, 0.0, 0.0}, angle = 0.0;
  if (interp == 0) {
    const int c = (fx > 0.5 ? 1 : 0) + (fy > 0.5 ? 2 : 0);
    for (int k = 0; k < 3; ++k) gr[k] = g[3 * c + k];
  } else if (interp == 1) {
    const double wc[4] = {(1.0 - fx) * (1.0 - fy), fx * (1.0 - fy), (1.0 - fx) * fy, fx * fy};
    double gone = 0.0;
    for (int c = 0; c < 4; ++c)
      if (el[c]) gone += wc[c];
    angle = blend(an);
    if (gone > 0) angle = gone >= 0.9999999 ? 0.0 : angle / (1.0 - gone);
    for (int k = 0; k < 3; ++k) {
      const double v[4] = {g[k], g[3 + k], g[6 + k], g[9 + k]};
      gr[k] = blend(v) * w;
    }
  }
  if (actin) gr[2] = 0;
  for (int k = 0; k < 3; ++k) inc[k][k] = std::exp(gr[k] * dt);
  if (angle != 0.0) {
    const double c = std::cos(angle), s = std::sin(angle);
    const M3 Rm = {{c, -1.0 * s, 0.0}, {s, c, 0.0}, {0.0, 0.0, 1.0}};
    M3 T;
    mul(Rm, inc, T);
    mul_nt(T, Rm, inc);
  }
}

}  // namespace

extern "C" {

void orc_grid_compatible(int nx, int ny, const double* rates, const double* angles_deg,
                         double* cg, double* ca, uint8_t* ce) {
  Grid G;
  compatible(nx, ny, rates, angles_deg, G);
  std::memcpy(cg, G.g.data(), sizeof(double) * G.g.size());
  std::memcpy(ca, G.a.data(), sizeof(double) * G.a.size());
  std::memcpy(ce, G.e.data(), G.e.size());
}

int orc_polar_rotation(const double* F9, double* R9) {
  M3 F, R;
  std::memcpy(F, F9, sizeof(M3));
  if (!polar(F, R)) return 0;
  std::memcpy(R9, R, sizeof(M3));
  return 1;
}

// 1 and GrowthStrainsRotMat (row-major) when the prism's growth is rotated.
int orc_growth_rotation(const double* ref18, const double* pos18, double* gsr9) {
  double ref[6][3], pos[6][3];
  std::memcpy(ref, ref18, sizeof(ref));
  std::memcpy(pos, pos18, sizeof(pos));
  M3 g;
  if (!growth_rotation(ref, pos, g)) return 0;
  std::memcpy(gsr9, g, sizeof(M3));
  return 1;
}

// Relative (x, y) positions of one tissue's prisms at positions u.
void orc_relative_positions(int n_nodes, const double* u, const uint8_t* excluded,
                            int symmetric_y, int n_prisms, const int32_t* conn,
                            const int32_t* apical_of, double* rel) {
  double lo[3] = {100000.0, 100000.0, 100000.0}, hi[3] = {-100000.0, -100000.0, -100000.0};
  for (int k = 0; k < n_nodes; ++k) {
    if (excluded && excluded[k]) continue;
    for (int i = 0; i < 3; ++i) {
      const double v = u[3 * k + i];
      if (v < lo[i]) lo[i] = v;
      else if (v > hi[i]) hi[i] = v;
    }
  }
  if (symmetric_y) lo[1] = (-1.0) * hi[1];
  const double sx = hi[0] - lo[0], sy = hi[1] - lo[1];
  std::vector<double> raw(2 * (size_t)n_prisms);
  for (int e = 0; e < n_prisms; ++e) {
    double c[3] = {0.0, 0.0, 0.0};
    for (int a = 0; a < 6; ++a)
      for (int j = 0; j < 3; ++j) c[j] += u[3 * conn[6 * e + a] + j];
    for (int j = 0; j < 3; ++j) c[j] /= 6;
    raw[2 * e] = (c[0] - lo[0]) / sx;
    raw[2 * e + 1] = (c[1] - lo[1]) / sy;
  }
  for (int e = 0; e < n_prisms; ++e) {
    const int s = apical_of && apical_of[e] >= 0 ? apical_of[e] : e;
    rel[2 * e] = raw[2 * s];
    rel[2 * e + 1] = raw[2 * s + 1];
  }
}

// One step's growth of every prism at time t from positions u, updating Fg and
// invFg (n_prisms x 9).  Functions: fn_type (1 uniform, 3 grid), fn_apply
// (1 columnar, 2 peripodial, 4 basal ECM, 8 lateral ECM), fn_t (t0, t1; kept
// as float), fn_z (zmin, zmax), fn_rate (x, y, z in 1/s), fn_angle (radians),
// fn_grid (nx, ny, offset of the grid's first point in grid_rates / grid_angles).
void orc_growth_step(int n_prisms, const int32_t* conn, const double* refpos, const double* u,
                     const int32_t* type, const int32_t* place, const uint8_t* flags,
                     const double* weights, const double* rel_z, int explicit_ecm, int n_fn,
                     const int32_t* fn_type, const int32_t* fn_apply, const double* fn_t,
                     const double* fn_z, const double* fn_rate, const double* fn_angle,
                     const int32_t* fn_grid, const double* grid_rates,
                     const double* grid_angles_deg, const double* rel, double t, double dt,
                     int interp, const uint8_t* ablated, int n_columns, int layers,
                     const int32_t* columns, const double* ref_volume, double* Fg,
                     double* invFg) {
  PrismInfo P{type, place, flags, weights, rel_z, explicit_ecm};
  std::vector<Fn> fns(n_fn);
  std::vector<Grid> grids;
  for (int k = 0; k < n_fn; ++k) {
    Fn& f = fns[k];
    f.type = fn_type[k];
    f.apply = fn_apply[k];
    f.t0 = (float)fn_t[2 * k];
    f.t1 = (float)fn_t[2 * k + 1];
    f.zmin = fn_z[2 * k];
    f.zmax = fn_z[2 * k + 1];
    for (int i = 0; i < 3; ++i) f.rate[i] = fn_rate[3 * k + i];
    const double c = std::cos(fn_angle[k]), s = std::sin(fn_angle[k]);
    const M3 R = {{c, -1.0 * s, 0.0}, {s, c, 0.0}, {0.0, 0.0, 1.0}};
    copy(R, f.rot);
    if (f.type == 3) {
      grids.emplace_back();
      const int nx = fn_grid[3 * k], ny = fn_grid[3 * k + 1], off = fn_grid[3 * k + 2];
      compatible(nx, ny, grid_rates + 3 * (size_t)off, grid_angles_deg + off, grids.back());
      f.grid = (int)grids.size() - 1;
    }
  }
  const bool conserve = n_columns > 0;
  std::vector<M3> G(n_prisms), SC(n_prisms), GSR(n_prisms);
  std::vector<char> turned(n_prisms, 0);
  std::vector<double> grown(n_prisms, 0.0);
  for (int e = 0; e < n_prisms; ++e) {
    M3 Fe;
    std::memcpy(Fe, Fg + 9 * (size_t)e, sizeof(M3));
    if (ablated && ablated[e]) {
      grown[e] = det(Fe) * (ref_volume ? ref_volume[e] : 0.0);
      continue;
    }
    double ref[6][3], pos[6][3];
    for (int a = 0; a < 6; ++a)
      for (int i = 0; i < 3; ++i) {
        ref[a][i] = refpos[18 * (size_t)e + 3 * a + i];
        pos[a][i] = u[3 * (size_t)conn[6 * e + a] + i];
      }
    turned[e] = growth_rotation(ref, pos, GSR[e]);
    eye(G[e]);
    const int tt = type[e], pl = place[e];
    const bool ecm = flags[e] & 1, ecm_circ = flags[e] & 2;
    for (const Fn& f : fns) {
      if (!(t >= f.t0 && t < f.t1)) continue;
      M3 col, per, T;
      eye(col);
      eye(per);
      if ((f.apply & 4) && ecm && pl == 0) increment(f, grids, P, e, 0, rel, dt, interp, col);
      if ((f.apply & 8) && ecm_circ && pl != 0) increment(f, grids, P, e, 0, rel, dt, interp, col);
      if (!explicit_ecm || !ecm) {
        if (f.apply & 1) increment(f, grids, P, e, 0, rel, dt, interp, col);
        if (f.apply & 2) increment(f, grids, P, e, 1, rel, dt, interp, per);
      }
      if (tt == 0 || tt == 2) {
        mul(col, G[e], T);
        copy(T, G[e]);
      }
      if (tt == 1 || tt == 2) {
        mul(per, G[e], T);
        copy(T, G[e]);
      }
    }
    eye(SC[e]);
    if (turned[e]) {  // R^T M R
      M3 T;
      mul_tn(GSR[e], G[e], T);
      mul(T, GSR[e], G[e]);
      mul_tn(GSR[e], SC[e], T);
      mul(T, GSR[e], SC[e]);
    }
    if (conserve) {
      M3 P1, T1, T2, T3;
      eye(P1);
      mul(P1, G[e], T1);
      mul(SC[e], T1, T2);
      mul(T2, Fe, T3);
      grown[e] = det(T3) * ref_volume[e];
    }
  }
  if (conserve) {
    for (int c = 0; c < n_columns; ++c) {
      double sg = 0, sr = 0;
      for (int i = 0; i < layers; ++i) {
        const int e = columns[(size_t)c * layers + i];
        if (e < 0) continue;
        sg += grown[e];
        sr += ref_volume[e];
      }
      const double g = std::pow(sr / sg, 1.0 / 3.0);
      const M3 D = {{g, 0.0, 0.0}, {0.0, g, 0.0}, {0.0, 0.0, g}};
      for (int i = 0; i < layers; ++i) {
        const int e = columns[(size_t)c * layers + i];
        if (e < 0) continue;
        M3 T;
        mul(D, G[e], T);
        copy(T, G[e]);
      }
    }
    for (int e = 0; e < n_prisms; ++e)
      if (turned[e] && !(ablated && ablated[e])) {  // growShapeByFg rotates them again
        M3 T;
        mul_tn(GSR[e], G[e], T);
        mul(T, GSR[e], G[e]);
        mul_tn(GSR[e], SC[e], T);
        mul(T, GSR[e], SC[e]);
      }
  }
  for (int e = 0; e < n_prisms; ++e) {
    if (ablated && ablated[e]) continue;
    M3 P1, T1, T2, T3, Fe, X;
    std::memcpy(Fe, Fg + 9 * (size_t)e, sizeof(M3));
    eye(P1);
    mul(P1, G[e], T1);
    mul(SC[e], T1, T2);
    mul(T2, Fe, T3);
    inverse(T3, X);
    std::memcpy(Fg + 9 * (size_t)e, T3, sizeof(M3));
    std::memcpy(invFg + 9 * (size_t)e, X, sizeof(M3));
  }
}

}  // extern "C"
