// tissue_nr.hpp -- C++ drop-in surface over the C ABI of tissue_nr.h.
//
// Keeps the names of the reference's NR path so that a caller written against
// TissueOrigami's Node / ShapeBase / Prism / NewtonRaphsonSolver
// (TissueFolding/SourceCode/Node.h:15-87, ShapeBase.h:301-324, Prism.h:15-67,
// NewtonRaphsonSolver.h:22-104) can drive the device step with the same loop
// shape as Simulation::updateStepNR (Simulation.cpp:6605-6746):
//
//   solver.setMatricesToZeroAtTheBeginningOfIteration(false);
//   solver.constructUnMatrix(Nodes);
//   solver.initialteUkMatrix();
//   while (!converged) {
//     solver.calculateDisplacementMatrix(dt);
//     solver.calculateForcesAndJacobianMatrixNR(Nodes, Elements, dt);   // + scatter, BCs
//     solver.solveForDeltaU();
//     converged = solver.checkConvergenceViaDeltaU();
//     solver.updateUkInIteration();
//     ...
//   }
//
// or in one call: solver.updateStepNR(Nodes, Elements, dt).  The Jacobian never
// leaves the device; un, uk, deltaU and gSum are host vectors as the
// reference's gsl matrices were.  Header-only; link with -ltissue_nr.
#ifndef TISSUE_NR_HPP
#define TISSUE_NR_HPP

#include <array>
#include <cmath>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "tissue_nr.h"

namespace tissue_nr {

inline void check(tf_status s) {
  if (s != TF_OK) throw std::runtime_error(std::string("tissue_nr: ") + tf_last_error());
}

// Node state read by the NR step (Node.h:21-58).
struct Node {
  int Id = 0;
  std::array<double, 3> Position{};
  std::array<bool, 3> FixedPos{};
  std::array<double, 3> externalViscosity{};
  double viscositySurface = 0.0;
  int tissuePlacement = 2;
  int tissueType = 0;
  double mass = 0.0;                   // Node::mass (Simulation::updateNodeMasses)
  bool hasLateralElementOwner = false;
  // Node::allOwnersECMMimicing: left out of the bounding box (Simulation.cpp:8397)
  bool allOwnersECMMimicing = false;
};

// A growth function (GrowthFunctionBase.h:13-60 with UniformGrowthFunction,
// type 1, GrowthFunctionTypes.h:14-97, and GridBasedGrowthFunction, type 3,
// GrowthFunctionTypes.h:290-546), holding what the reference's objects hold
// after ModelInputObject has read them: times in s, rates in 1/s, the uniform
// angle in radians, the grid angles in degrees.
struct GrowthFunction {
  int Type = 1;
  float initTime = 0.0f, endTime = 0.0f;
  bool applyToColumnarLayer = false, applyToPeripodialMembrane = false,
       applyToBasalECM = false, applyToLateralECM = false;
  double zMin = 0.0, zMax = 1.0;
  double GrowthRate[3] = {0.0, 0.0, 0.0};
  double angle = 0.0;
  int nGridX = 0, nGridY = 0;
  std::vector<std::vector<std::array<double, 3>>> GrowthMatrix;  // [i][j][x, y, z]
  std::vector<std::vector<double>> AngleMatrix;                  // [i][j]
};

// Growth options of Simulation (Simulation.h:349-360) the growth step reads.
struct GrowthOptions {
  int gridGrowthsInterpolationType = 1;
  bool GridGrowthsPinnedOnInitialMesh = false;
  std::vector<int> growthPinUpdateTime;  // s
  bool thereIsExplicitECM = false;
  bool conservingColumnVolumes = false;
  bool symmetricY = false;
  int TissueHeightDiscretisationLayers = 1;
};

// Element state read by the NR step (ShapeBase.h:53-69, 129-143).
class ShapeBase {
 public:
  virtual ~ShapeBase() = default;
  std::array<int, 6> NodeIds{};
  double ReferencePositions[6][3] = {};
  double E = 10.0, v = 0.3, internalViscosity = 0.0;
  // ShapeBase::stiffnessMultiplier (ShapeBase.h:146): the Lame parameters are
  // formed from it times E (Prism::updateElasticProperties, Prism.cpp:526-527)
  double stiffnessMultiplier = 1.0;
  double InvFg[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  bool IsAblated = false;  // ShapeBase::IsAblated: no forces, zero tangents
  int tissuePlacement = 2, tissueType = 0;  // ShapeBase.cpp:905-1037
  bool spansWholeTissue = false;
  // Young's modulus time series (ShapeBase.h:147): the multiplier the series
  // set, the ECM flags its applicability test reads, and the relative (x, y)
  // position in the tissue's bounding box it looks the rate up at
  double StiffnessTimeSeriesMultiplier = 1.0;
  bool isECMMimicing = false, isECMMimimcingAtCircumference = false, atBasalBorderOfECM = false;
  double relativePosInBoundingBox[2] = {0.0, 0.0};
  // growth (ShapeBase.h:120-160): the growth tensor, the linker prisms'
  // weights, the relative z position, actin, the stack of the prism's column
  // (basal first) and the pinned relative position
  double Fg[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  double columnarGrowthWeight = 0.0, peripodialGrowthWeight = 0.0;
  double initialRelativePositionInZ = 0.0;
  bool isActinMimicing = false;
  std::vector<int> elementsIdsOnSameColumn;
  double initialRelativePosInBoundingBox[2] = {0.0, 0.0};
  // ShapeBase::Strain (ShapeBase.h:159): Green-Lagrange strain of the last
  // Gauss point of the last force evaluation, E00, E11, E22, 2E01, 2E21, 2E02
  // (Prism.cpp:826-837); set by the solver after each assembly and step
  double Strain[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  int getNodeNumber() const { return 6; }
  const int* getNodeIds() const { return NodeIds.data(); }
  // ShapeBase::updatePositions (ShapeBase.cpp:3279-3285): element positions
  // follow the nodes after every NR update.
  void updatePositions(const std::vector<std::unique_ptr<Node>>& Nodes) {
    for (int i = 0; i < 6; ++i)
      for (int j = 0; j < 3; ++j) Positions[i][j] = Nodes[NodeIds[i]]->Position[j];
  }
  double Positions[6][3] = {};
};

// 6-node wedge (Prism.h:15-67).
class Prism : public ShapeBase {
 public:
  Prism(const int* ids, const std::vector<std::unique_ptr<Node>>& Nodes) {
    for (int i = 0; i < 6; ++i) NodeIds[i] = ids[i];
    for (int i = 0; i < 6; ++i)
      for (int j = 0; j < 3; ++j) ReferencePositions[i][j] = Nodes[ids[i]]->Position[j];
    updatePositions(Nodes);
  }
};

class NewtonRaphsonSolver {
 public:
  // NewtonRaphsonSolver(int nDim, int nNodes) (NewtonRaphsonSolver.cpp:16-41);
  // the mesh is bound at the first call that sees the elements.
  NewtonRaphsonSolver(int nDim, int nNodes) : nDim(nDim), nNodes(nNodes) {
    if (nDim != 3) throw std::invalid_argument("only 3D prism meshes");
    un.assign(3 * (size_t)nNodes, 0.0);
    uk = un;
    deltaU = un;
    gSum = un;
    displacementPerDt = un;
  }
  ~NewtonRaphsonSolver() { tf_nr_destroy(h); }
  NewtonRaphsonSolver(const NewtonRaphsonSolver&) = delete;
  NewtonRaphsonSolver& operator=(const NewtonRaphsonSolver&) = delete;

  double threshold = 1E-8;                    // NewtonRaphsonSolver.cpp:22
  std::vector<double> un, uk, deltaU, gSum, displacementPerDt;
  std::vector<std::vector<int>> slaveMasterList;
  // Simulation::AblatedNodes, the list checkJacobianForAblatedNodes applies
  std::vector<int> AblatedNodes;
  bool boundNodesWithSlaveMasterDefinition = false;
  int precision = TF_PREC_FP64;
  double krylovTolerance = 1e-10;
  int lastKrylovIterations = 0;

  void setMatricesToZeroAtTheBeginningOfIteration(bool) {
    std::fill(gSum.begin(), gSum.end(), 0.0);
    std::fill(deltaU.begin(), deltaU.end(), 0.0);
  }
  void constructUnMatrix(const std::vector<std::unique_ptr<Node>>& Nodes) {
    for (int i = 0; i < nNodes; ++i)
      for (int j = 0; j < 3; ++j) un[3 * i + j] = Nodes[i]->Position[j];
  }
  void initialteUkMatrix() { uk = un; }
  void calculateDisplacementMatrix(double dt) {
    for (size_t i = 0; i < uk.size(); ++i) displacementPerDt[i] = (uk[i] - un[i]) * (1.0 / dt);
  }
  // Element loop, scatter, external viscosity, fixed DOF and binding: the
  // device counterpart of calculateForcesAndJacobianMatrixNR +
  // writeForcesTogeAndgvInternal + writeImplicitElementalKToJacobian +
  // calculateExternalViscousForcesForNR + addImplicitKViscousExternalToJacobian +
  // calculateSumOfInternalForces + calcutateFixedK + calculateBoundKWithSlavesMasterDoF.
  void calculateForcesAndJacobianMatrixNR(const std::vector<std::unique_ptr<Node>>& Nodes,
                                          const std::vector<std::unique_ptr<ShapeBase>>& Elements,
                                          double dt) {
    bind(Nodes, Elements);
    check(tf_nr_assemble(h, uk.data(), un.data(), dt, gSum.data(), nullptr));
    lastDt = dt;
    readStrains(Elements);
  }
  // The elements' ShapeBase::Strain from the device's last assembly.
  void readStrains(const std::vector<std::unique_ptr<ShapeBase>>& Elements) {
    std::vector<double> s(6 * Elements.size());
    check(tf_nr_element_strain(h, s.data()));
    for (size_t e = 0; e < Elements.size(); ++e)
      for (int q = 0; q < 6; ++q) Elements[e]->Strain[q] = s[6 * e + q];
  }
  // checkJacobianForAblatedNodes (NewtonRaphsonSolver.cpp:425-439): a zero
  // Jacobian diagonal entry of a listed node becomes 1.  The device assembly
  // applies it to the list bound with the other inputs; called after
  // calculateForcesAndJacobianMatrixNR with a list that differs from that one,
  // it binds the list and assembles again at the same state.
  void checkJacobianForAblatedNodes(const std::vector<int>& nodes) {
    AblatedNodes = nodes;
    if (!h || boundAblNodes == nodes) return;
    bindAblation(ablFlags);
    if (lastDt > 0) check(tf_nr_assemble(h, uk.data(), un.data(), lastDt, gSum.data(), nullptr));
  }
  // solveForDeltaU (NewtonRaphsonSolver.cpp:461-488) with the device Krylov
  // solver in place of Pardiso; slaves follow their masters.
  void solveForDeltaU() {
    int32_t it = 0;
    check(tf_nr_solve(h, gSum.data(), deltaU.data(), krylovTolerance, 50000, &it));
    lastKrylovIterations = it;
    for (auto& p : slaveMasterList) deltaU[p[0]] = deltaU[p[1]];
  }
  bool checkConvergenceViaDeltaU() {
    double s = 0.0;
    for (double d : deltaU) s += d * d;
    return !(std::sqrt(s) > threshold);
  }
  void updateUkInIteration() {
    for (size_t i = 0; i < uk.size(); ++i) uk[i] = uk[i] + deltaU[i];
  }
  void updateNodePositionsNR(const std::vector<std::unique_ptr<Node>>& Nodes) const {
    for (int i = 0; i < nNodes; ++i)
      for (int j = 0; j < 3; ++j) Nodes[i]->Position[j] = uk[3 * i + j];
  }

  // The whole of Simulation::updateStepNR in one device call.  Returns the
  // step statistics; node and element positions are updated in place.
  tf_step_stats updateStepNR(const std::vector<std::unique_ptr<Node>>& Nodes,
                             const std::vector<std::unique_ptr<ShapeBase>>& Elements, double dt,
                             int maxIteration = 20) {
    bind(Nodes, Elements);
    constructUnMatrix(Nodes);
    tf_step_stats st;
    const tf_status s = tf_nr_step(h, un.data(), dt, threshold, maxIteration, precision,
                                   krylovTolerance, uk.data(), &st);
    if (s != TF_OK && s != TF_ERR_NOT_CONVERGED) check(s);
    updateNodePositionsNR(Nodes);
    for (auto& e : Elements) e->updatePositions(Nodes);
    readStrains(Elements);
    return st;
  }

  // Uploads the per-step inputs (material with each element's stiffness
  // multiplier, growth, viscosity surfaces, fixing, binding, ablation) from the
  // node and element objects.  Each input goes to the device only when it
  // differs from what was bound last, so the NR iterations of a step, which
  // call this at every assembly, upload nothing after the first.
  void bind(const std::vector<std::unique_ptr<Node>>& Nodes,
            const std::vector<std::unique_ptr<ShapeBase>>& Elements) {
    const int ne = (int)Elements.size();
    if (!h) {
      std::vector<int32_t> conn(6 * (size_t)ne);
      std::vector<double> ref(18 * (size_t)ne);
      for (int e = 0; e < ne; ++e)
        for (int a = 0; a < 6; ++a) {
          conn[6 * e + a] = Elements[e]->NodeIds[a];
          for (int j = 0; j < 3; ++j) ref[18 * e + 3 * a + j] = Elements[e]->ReferencePositions[a][j];
        }
      check(tf_nr_create(nNodes, ne, conn.data(), ref.data(), &h));
    }
    std::vector<double> E(ne), nu(ne), eta(ne), iFg(9 * (size_t)ne), ev(3 * (size_t)nNodes),
        A(nNodes);
    std::vector<uint8_t> fixed(3 * (size_t)nNodes), abl(ne);
    for (int e = 0; e < ne; ++e) {
      E[e] = Elements[e]->stiffnessMultiplier * Elements[e]->E;
      nu[e] = Elements[e]->v;
      eta[e] = Elements[e]->internalViscosity;
      abl[e] = Elements[e]->IsAblated ? 1 : 0;
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) iFg[9 * e + 3 * i + j] = Elements[e]->InvFg[i][j];
    }
    for (int i = 0; i < nNodes; ++i) {
      A[i] = Nodes[i]->viscositySurface;
      for (int j = 0; j < 3; ++j) {
        ev[3 * i + j] = Nodes[i]->externalViscosity[j];
        fixed[3 * i + j] = Nodes[i]->FixedPos[j] ? 1 : 0;
      }
    }
    std::vector<int32_t> pairs;
    for (auto& p : slaveMasterList) {
      pairs.push_back(p[0]);
      pairs.push_back(p[1]);
    }
    if (!bound || E != bE || nu != bNu || eta != bEta || ev != bEv || fixed != bFixed ||
        pairs != bPairs) {
      check(tf_nr_set_material(h, E.data(), nu.data(), eta.data(), ev.data(), fixed.data(),
                               pairs.empty() ? nullptr : pairs.data(),
                               (int32_t)(pairs.size() / 2)));
      bE = E; bNu = nu; bEta = eta; bEv = ev; bFixed = fixed; bPairs = pairs;
    }
    if (!bound || iFg != bFg) {
      check(tf_nr_set_growth(h, iFg.data()));
      bFg = iFg;
    }
    if (!bound || A != bA) {
      check(tf_nr_set_visc_surface(h, A.data()));
      bA = A;
    }
    if (abl != ablFlags || AblatedNodes != boundAblNodes) bindAblation(abl);
    bound = true;
  }

  // Node-node packing: Simulation::detectPacingNodes (Simulation.cpp:7997-8078)
  // at the nodes' current positions, with their masses and placements and the
  // elements' placement and type.  The couples found hold for the following
  // updateStepNR calls, whose NR iterations then add the couples' forces and
  // Jacobian blocks (Simulation.cpp:6680-6698).  Returns the number of couples.
  int detectPacingNodes(const std::vector<std::unique_ptr<Node>>& Nodes,
                        const std::vector<std::unique_ptr<ShapeBase>>& Elements,
                        bool symmetricY = false) {
    bind(Nodes, Elements);
    const int ne = (int)Elements.size();
    std::vector<int32_t> place(ne), type(ne);
    std::vector<uint8_t> spans(ne);
    for (int e = 0; e < ne; ++e) {
      place[e] = Elements[e]->tissuePlacement;
      type[e] = Elements[e]->tissueType;
      spans[e] = Elements[e]->spansWholeTissue ? 1 : 0;
    }
    check(tf_nr_set_surface_flags(h, place.data(), type.data(), spans.data()));
    if (!packingOn) {
      std::vector<int32_t> npl(nNodes);
      std::vector<uint8_t> ex(nNodes);
      for (int i = 0; i < nNodes; ++i) {
        npl[i] = Nodes[i]->tissuePlacement;
        ex[i] = Nodes[i]->hasLateralElementOwner ? 1 : 0;
      }
      check(tf_nr_set_packing(h, npl.data(), ex.data(), symmetricY ? 1 : 0, 2));
      packingOn = true;
    }
    std::vector<double> m(nNodes), u(3 * (size_t)nNodes);
    for (int i = 0; i < nNodes; ++i) {
      m[i] = Nodes[i]->mass;
      for (int j = 0; j < 3; ++j) u[3 * i + j] = Nodes[i]->Position[j];
    }
    check(tf_nr_set_node_mass(h, m.data()));
    int32_t n = 0;
    check(tf_nr_detect_packing(h, u.data(), &n));
    return n;
  }

  // The YoungsModulusModifier constructor (YoungsModulusModifier.cpp:5-13):
  // TimeSeriesGrid[t][iy][ix] applied at TimeToApply[t], the linear or
  // exponential change, the rate method (0 nearest, 1 bilinear, 2 lower
  // corner) and WhereToApply (TimeSeriesPhysicalProperties.cpp:112-120).
  // The prisms it applies to follow YoungsModulusModifier::IsItApplicable
  // (YoungsModulusModifier.cpp:74-100).
  void addYoungsModulusModifier(const std::vector<std::vector<std::vector<double>>>& TimeSeriesGrid,
                                const std::vector<double>& TimeToApply, bool linear, int method,
                                const std::array<bool, 7>& WhereToApply,
                                const std::vector<std::unique_ptr<Node>>& Nodes,
                                const std::vector<std::unique_ptr<ShapeBase>>& Elements) {
    bind(Nodes, Elements);
    const int nt = (int)TimeSeriesGrid.size();
    const int ny = nt ? (int)TimeSeriesGrid[0].size() : 0;
    const int nx = ny ? (int)TimeSeriesGrid[0][0].size() : 0;
    std::vector<double> g;
    for (const auto& grid : TimeSeriesGrid)
      for (const auto& row : grid) g.insert(g.end(), row.begin(), row.end());
    if (g.size() != (size_t)nt * nx * ny) throw std::runtime_error("ragged time-series grid");
    const int ne = (int)Elements.size();
    std::vector<uint8_t> app(ne);
    std::vector<double> rel(2 * (size_t)ne);
    const bool cl = WhereToApply[0], pm = WhereToApply[1], ap = WhereToApply[2],
               mid = WhereToApply[3], bas = WhereToApply[4], cecm = WhereToApply[6];
    for (int e = 0; e < ne; ++e) {
      const ShapeBase& el = *Elements[e];
      const int pl = el.tissuePlacement;
      bool on = false;
      if ((cl && el.tissueType == 0) || (pm && el.tissueType == 1)) {
        if (el.isECMMimicing)
          on = (cecm && pl == 0) || (cecm && el.isECMMimimcingAtCircumference && pl != 0);
        else
          on = (ap && (pl == 1 || pl == 4)) || (mid && pl == 2) ||
               (bas && (pl == 0 || el.atBasalBorderOfECM));
      }
      app[e] = on ? 1 : 0;
      rel[2 * e] = el.relativePosInBoundingBox[0];
      rel[2 * e + 1] = el.relativePosInBoundingBox[1];
    }
    check(tf_nr_add_stiffness_series(h, nt, TimeToApply.data(), nx, ny, g.data(), linear ? 1 : 0,
                                     method, app.data(), rel.data()));
  }

  // Simulation::ApplyAllYoungsModulusModifiers (Simulation.cpp:878-884) at
  // currSimTimeSec: the multipliers come back into the elements.
  void ApplyAllYoungsModulusModifiers(double currSimTimeSec, double dt,
                                      const std::vector<std::unique_ptr<ShapeBase>>& Elements) {
    if (!h) throw std::runtime_error("ApplyAllYoungsModulusModifiers: no modifier added");
    check(tf_nr_apply_stiffness_series(h, currSimTimeSec, dt));
    std::vector<double> m(Elements.size());
    check(tf_nr_get_stiffness_multiplier(h, m.data()));
    for (size_t e = 0; e < Elements.size(); ++e) Elements[e]->StiffnessTimeSeriesMultiplier = m[e];
  }

  // The growth functions and options of a Simulation, with its elements'
  // growth data: from now on calculateRelativePositions and calculateGrowth
  // run them on the device, and the stiffness series take their relative
  // positions from there.  The elements' Fg is the starting growth.
  void setGrowthFunctions(const std::vector<GrowthFunction>& functions, const GrowthOptions& opt,
                          const std::vector<std::unique_ptr<Node>>& Nodes,
                          const std::vector<std::unique_ptr<ShapeBase>>& Elements) {
    bind(Nodes, Elements);
    const int ne = (int)Elements.size();
    const int L = opt.TissueHeightDiscretisationLayers;
    std::vector<int32_t> type(ne), place(ne), apical(ne, -1);
    std::vector<uint8_t> flags(ne), excl(nNodes);
    std::vector<double> w(2 * (size_t)ne), rz(ne), Fg(9 * (size_t)ne);
    std::vector<int32_t> cols;
    for (int e = 0; e < ne; ++e) {
      const ShapeBase& el = *Elements[e];
      type[e] = el.tissueType;
      place[e] = el.tissuePlacement;
      flags[e] = (el.isECMMimicing ? 1 : 0) | (el.isECMMimimcingAtCircumference ? 2 : 0) |
                 (el.isActinMimicing ? 4 : 0);
      w[2 * e] = el.columnarGrowthWeight;
      w[2 * e + 1] = el.peripodialGrowthWeight;
      rz[e] = el.initialRelativePositionInZ;
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) Fg[9 * e + 3 * i + j] = el.Fg[i][j];
    }
    for (int e = 0; e < ne; ++e) {
      const ShapeBase& el = *Elements[e];
      if (el.tissueType != 0 || el.tissuePlacement != 1 || (int)el.elementsIdsOnSameColumn.size() < L)
        continue;
      // Simulation::updateRelativePositionsToApicalPositioning (Simulation.cpp:6147-6167)
      for (int i = 0; i < L - 2; ++i) apical[el.elementsIdsOnSameColumn[i]] = e;
      // Simulation::conserveColumnVolume (Simulation.cpp:6518-6540)
      for (int i = 0; i < L; ++i) {
        const int m = el.elementsIdsOnSameColumn[i];
        const ShapeBase& c = *Elements[m];
        cols.push_back(c.tissueType == 0 && !c.isECMMimicing && !c.isActinMimicing ? m : -1);
      }
    }
    for (int i = 0; i < nNodes; ++i) excl[i] = Nodes[i]->allOwnersECMMimicing ? 1 : 0;
    check(tf_nr_set_growth_prisms(h, type.data(), place.data(), flags.data(), w.data(), rz.data(),
                                  apical.data(), excl.data(), opt.symmetricY ? 1 : 0,
                                  opt.thereIsExplicitECM ? 1 : 0));
    for (const GrowthFunction& f : functions) {
      const int apply = (f.applyToColumnarLayer ? 1 : 0) | (f.applyToPeripodialMembrane ? 2 : 0) |
                        (f.applyToBasalECM ? 4 : 0) | (f.applyToLateralECM ? 8 : 0);
      if (f.Type == 1) {
        check(tf_nr_add_uniform_growth(h, f.initTime, f.endTime, apply, f.zMin, f.zMax,
                                       f.GrowthRate, f.angle));
      } else if (f.Type == 3) {
        std::vector<double> r, a;
        for (int i = 0; i < f.nGridX; ++i)
          for (int j = 0; j < f.nGridY; ++j) {
            for (int k = 0; k < 3; ++k) r.push_back(f.GrowthMatrix.at(i).at(j)[k]);
            a.push_back(f.AngleMatrix.at(i).at(j));
          }
        check(tf_nr_add_grid_growth(h, f.initTime, f.endTime, apply, f.zMin, f.zMax, f.nGridX,
                                    f.nGridY, r.data(), a.data()));
      } else {
        throw std::invalid_argument("growth function type " + std::to_string(f.Type) +
                                    " is not supported");
      }
    }
    std::vector<double> pins(opt.growthPinUpdateTime.begin(), opt.growthPinUpdateTime.end());
    check(tf_nr_set_growth_options(h, opt.gridGrowthsInterpolationType,
                                   opt.GridGrowthsPinnedOnInitialMesh ? 1 : 0, (int32_t)pins.size(),
                                   pins.empty() ? nullptr : pins.data()));
    if (opt.conservingColumnVolumes && !cols.empty())
      check(tf_nr_set_column_volume_conservation(h, (int32_t)(cols.size() / L), L, cols.data()));
    check(tf_nr_growth_tensor(h, nullptr, Fg.data()));
  }

  // Simulation::calculateBoundingBox with ShapeBase::calculateRelativePosInBoundingBox
  // and updateRelativePositionsToApicalPositioning (Simulation.cpp:5976-5983) at
  // the nodes' positions, on the device; the positions come back into the
  // elements (the first call also sets the pinned positions).
  void calculateRelativePositions(const std::vector<std::unique_ptr<Node>>& Nodes,
                                  const std::vector<std::unique_ptr<ShapeBase>>& Elements) {
    constructUnMatrix(Nodes);
    check(tf_nr_update_relative_positions(h, un.data()));
    const size_t ne = Elements.size();
    std::vector<double> cur(2 * ne), pin(2 * ne);
    check(tf_nr_relative_positions(h, cur.data(), pin.data()));
    for (size_t e = 0; e < ne; ++e)
      for (int i = 0; i < 2; ++i) {
        Elements[e]->relativePosInBoundingBox[i] = cur[2 * e + i];
        Elements[e]->initialRelativePosInBoundingBox[i] = pin[2 * e + i];
      }
  }

  // The growth of one step at currSimTimeSec (the pinning updates,
  // Simulation::updateGrowthRotationMatrices, calculateGrowth, the column
  // volume conservation and growShapeByFg of Simulation::runOneStep,
  // Simulation.cpp:6012-6063) from the nodes' positions.  Fg and InvFg come
  // back into the elements; the device keeps its copy.
  void calculateGrowth(const std::vector<std::unique_ptr<Node>>& Nodes,
                       const std::vector<std::unique_ptr<ShapeBase>>& Elements,
                       double currSimTimeSec, double dt) {
    constructUnMatrix(Nodes);
    check(tf_nr_grow(h, un.data(), currSimTimeSec, dt));
    const size_t ne = Elements.size();
    std::vector<double> Fg(9 * ne);
    check(tf_nr_growth_tensor(h, Fg.data(), nullptr));
    for (size_t e = 0; e < ne; ++e) {
      ShapeBase& el = *Elements[e];
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) el.Fg[i][j] = Fg[9 * e + 3 * i + j];
      invert3(el.Fg, el.InvFg);
      for (int q = 0; q < 9; ++q) bFg[9 * e + q] = el.InvFg[q / 3][q % 3];
    }
  }

  tf_handle* handle() const { return h; }

 private:
  // 3x3 inverse by LU with partial pivoting, the device's (and GSL's
  // LU_decomp / LU_invert) order of operations, so that the InvFg the elements
  // get back equals the device's and bind() has nothing to upload.
  static void invert3(const double A[3][3], double X[3][3]) {
    double a[3][3];
    int p[3] = {0, 1, 2};
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) a[i][j] = A[i][j];
    for (int c = 0; c < 2; ++c) {
      int r = c;
      double mx = std::fabs(a[c][c]);
      for (int i = c + 1; i < 3; ++i)
        if (std::fabs(a[i][c]) > mx) { mx = std::fabs(a[i][c]); r = i; }
      if (r != c) {
        for (int k = 0; k < 3; ++k) { const double t = a[c][k]; a[c][k] = a[r][k]; a[r][k] = t; }
        const int t = p[c]; p[c] = p[r]; p[r] = t;
      }
      if (a[c][c] != 0.0)
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

  // the ablation flags and node list of the next assemblies
  void bindAblation(const std::vector<uint8_t>& abl) {
    std::vector<int32_t> nodes(AblatedNodes.begin(), AblatedNodes.end());
    bool any = false;
    for (uint8_t a : abl) any |= a != 0;
    bool before = !boundAblNodes.empty();
    for (uint8_t a : ablFlags) before |= a != 0;
    if (any || !nodes.empty() || before)
      check(tf_nr_set_ablation(h, any ? abl.data() : nullptr, nodes.empty() ? nullptr : nodes.data(),
                             (int32_t)nodes.size()));
    ablFlags = abl;
    boundAblNodes = AblatedNodes;
  }

  int nDim, nNodes;
  tf_handle* h = nullptr;
  bool packingOn = false;
  double lastDt = 0.0;
  // what bind() last uploaded
  bool bound = false;
  std::vector<double> bE, bNu, bEta, bEv, bFg, bA;
  std::vector<uint8_t> bFixed, ablFlags;
  std::vector<int32_t> bPairs;
  std::vector<int> boundAblNodes;
};

}  // namespace tissue_nr

#endif  // TISSUE_NR_HPP
