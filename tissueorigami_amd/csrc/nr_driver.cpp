// Command-line driver of the device NR step written against the C++ drop-in
// surface (include/tissue_nr.hpp): it builds Node / Prism objects and runs the
// iteration loop of Simulation::updateStepNR (Simulation.cpp:6605-6746) with
// the reference's method names, so a TissueOrigami-style caller can be
// exercised end to end.
//
//   nr_driver <input.txt> <output.txt> [loop|step]
//
// Input (whitespace separated): n_nodes n_prisms dt tol maxit
//   n_nodes lines:  x y z ext_visc_x ext_visc_y ext_visc_z viscosity_surface fx fy fz
//   n_prisms lines: 6 node ids, 18 reference coordinates, E nu eta, 9 InvFg entries
//   optionally: n_pairs, then n_pairs lines "slave_dof master_dof"
//     (NewtonRaphsonSolver::slaveMasterList, as Simulation::bindCircumferenceXY
//     fills it), then any of these sections:
//     "stiffness" and n_prisms values of ShapeBase::stiffnessMultiplier;
//     "ablated_prisms" k and k prism ids (ShapeBase::IsAblated);
//     "ablated_nodes" k and k node ids (Simulation::AblatedNodes, applied by
//     checkJacobianForAblatedNodes after the assembly, Simulation.cpp:6691);
//     growth (GrowthFunction / GrowthOptions of tissue_nr.hpp):
//     "growth_prisms" then per prism: tissueType tissuePlacement flags
//       (1 ECM, 2 ECM at circumference, 4 actin) columnarGrowthWeight
//       peripodialGrowthWeight initialRelativePositionInZ, the column length c
//       and c prism ids (elementsIdsOnSameColumn);
//     "ecm_nodes" k and k node ids (Node::allOwnersECMMimicing);
//     "growth_uniform" t0 t1 apply zmin zmax rx ry rz angle;
//     "growth_grid" t0 t1 apply zmin zmax nx ny, then per grid point (i, j),
//       i outer: rx ry rz angle_deg;
//     "growth_options" interpolation pinned conserve symmetricY explicitECM
//       layers n_pins pin_times...;
//     "steps" K t0: K steps of Simulation::runOneStep's growth path
//       (calculateRelativePositions, calculateGrowth at t, updateStepNR).
// Output: one line per node with the converged x y z (%.17g), then a line
//   "iterations <k>" and one "du <norm>" line per NR iteration (of the last
//   step), with growth one "Fg" line of 9 values per prism, and one "strain"
//   line of 6 values per prism (ShapeBase::Strain after the last step).
#include <cstdio>
#include <fstream>
#include <iostream>
#include <memory>
#include <string>
#include <vector>

#include "../../include/tissue_nr.hpp"

using tissue_nr::Node;
using tissue_nr::Prism;
using tissue_nr::ShapeBase;

int main(int argc, char** argv) {
  if (argc < 3) {
    std::cerr << "usage: nr_driver input output [loop|step]\n";
    return 2;
  }
  const std::string mode = argc > 3 ? argv[3] : "loop";
  std::ifstream in(argv[1]);
  int nn = 0, ne = 0, maxit = 20;
  double dt = 0, tol = 1e-8;
  in >> nn >> ne >> dt >> tol >> maxit;
  if (!in || nn <= 0 || ne <= 0) {
    std::cerr << "bad input header\n";
    return 2;
  }
  std::vector<std::unique_ptr<Node>> Nodes;
  for (int i = 0; i < nn; ++i) {
    auto n = std::make_unique<Node>();
    n->Id = i;
    int f[3];
    in >> n->Position[0] >> n->Position[1] >> n->Position[2] >> n->externalViscosity[0] >>
        n->externalViscosity[1] >> n->externalViscosity[2] >> n->viscositySurface >> f[0] >> f[1] >>
        f[2];
    for (int j = 0; j < 3; ++j) n->FixedPos[j] = f[j] != 0;
    Nodes.push_back(std::move(n));
  }
  std::vector<std::unique_ptr<ShapeBase>> Elements;
  for (int e = 0; e < ne; ++e) {
    int ids[6];
    for (int& id : ids) in >> id;
    auto p = std::make_unique<Prism>(ids, Nodes);
    for (int a = 0; a < 6; ++a)
      for (int j = 0; j < 3; ++j) in >> p->ReferencePositions[a][j];
    in >> p->E >> p->v >> p->internalViscosity;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) in >> p->InvFg[i][j];
    Elements.push_back(std::move(p));
  }
  if (!in) {
    std::cerr << "bad input body\n";
    return 2;
  }
  std::vector<std::vector<int>> slaveMasterList;
  int n_pairs = 0;
  if (in >> n_pairs) {
    for (int p = 0; p < n_pairs; ++p) {
      int sl = -1, ms = -1;
      in >> sl >> ms;
      slaveMasterList.push_back({sl, ms});
    }
    if (!in) {
      std::cerr << "bad pair list\n";
      return 2;
    }
  }
  std::vector<int> ablatedNodes;
  std::vector<tissue_nr::GrowthFunction> growth;
  tissue_nr::GrowthOptions gopt;
  int n_steps = 0;
  double t0 = 0.0;
  in.clear();
  std::string key;
  while (in >> key) {
    if (key == "growth_prisms") {
      for (auto& el : Elements) {
        int fl = 0, c = 0;
        in >> el->tissueType >> el->tissuePlacement >> fl >> el->columnarGrowthWeight >>
            el->peripodialGrowthWeight >> el->initialRelativePositionInZ >> c;
        el->isECMMimicing = fl & 1;
        el->isECMMimimcingAtCircumference = fl & 2;
        el->isActinMimicing = fl & 4;
        el->elementsIdsOnSameColumn.resize(c);
        for (int& id : el->elementsIdsOnSameColumn) in >> id;
      }
    } else if (key == "ecm_nodes") {
      int k = 0;
      in >> k;
      for (int i = 0; i < k; ++i) {
        int id = -1;
        in >> id;
        if (id >= 0 && id < nn) Nodes[id]->allOwnersECMMimicing = true;
      }
    } else if (key == "growth_uniform" || key == "growth_grid") {
      tissue_nr::GrowthFunction f;
      int apply = 0;
      in >> f.initTime >> f.endTime >> apply >> f.zMin >> f.zMax;
      f.applyToColumnarLayer = apply & 1;
      f.applyToPeripodialMembrane = apply & 2;
      f.applyToBasalECM = apply & 4;
      f.applyToLateralECM = apply & 8;
      if (key == "growth_uniform") {
        f.Type = 1;
        in >> f.GrowthRate[0] >> f.GrowthRate[1] >> f.GrowthRate[2] >> f.angle;
      } else {
        f.Type = 3;
        in >> f.nGridX >> f.nGridY;
        if (f.nGridX < 2 || f.nGridY < 2 || f.nGridX > 4096 || f.nGridY > 4096) {
          std::cerr << "bad grid size\n";
          return 2;
        }
        f.GrowthMatrix.assign(f.nGridX, std::vector<std::array<double, 3>>(f.nGridY));
        f.AngleMatrix.assign(f.nGridX, std::vector<double>(f.nGridY));
        for (int i = 0; i < f.nGridX; ++i)
          for (int j = 0; j < f.nGridY; ++j)
            in >> f.GrowthMatrix[i][j][0] >> f.GrowthMatrix[i][j][1] >> f.GrowthMatrix[i][j][2] >>
                f.AngleMatrix[i][j];
      }
      growth.push_back(f);
    } else if (key == "growth_options") {
      int pinned = 0, conserve = 0, sym = 0, ecm = 0, np = 0;
      in >> gopt.gridGrowthsInterpolationType >> pinned >> conserve >> sym >> ecm >>
          gopt.TissueHeightDiscretisationLayers >> np;
      gopt.GridGrowthsPinnedOnInitialMesh = pinned;
      gopt.conservingColumnVolumes = conserve;
      gopt.symmetricY = sym;
      gopt.thereIsExplicitECM = ecm;
      gopt.growthPinUpdateTime.resize(np > 0 ? np : 0);
      for (int& t : gopt.growthPinUpdateTime) in >> t;
    } else if (key == "steps") {
      in >> n_steps >> t0;
    } else if (key == "stiffness") {
      for (auto& el : Elements) in >> el->stiffnessMultiplier;
    } else if (key == "ablated_prisms" || key == "ablated_nodes") {
      int k = 0;
      in >> k;
      for (int i = 0; i < k; ++i) {
        int id = -1;
        in >> id;
        if (key == "ablated_nodes") ablatedNodes.push_back(id);
        else if (id >= 0 && id < ne) Elements[id]->IsAblated = true;
      }
    } else {
      std::cerr << "unknown section " << key << "\n";
      return 2;
    }
    if (!in) {
      std::cerr << "bad section " << key << "\n";
      return 2;
    }
  }
  std::vector<double> norms;
  try {
    tissue_nr::NewtonRaphsonSolver NRSolver(3, nn);
    NRSolver.threshold = tol;
    NRSolver.krylovTolerance = 1e-13;
    NRSolver.slaveMasterList = slaveMasterList;
    NRSolver.boundNodesWithSlaveMasterDefinition = !slaveMasterList.empty();
    if (n_steps > 0) {
      // Simulation::runOneStep's order: relative positions, growth at the
      // step's start time, then the NR step
      NRSolver.AblatedNodes = ablatedNodes;
      if (!growth.empty()) NRSolver.setGrowthFunctions(growth, gopt, Nodes, Elements);
      double t = t0;
      for (int k = 0; k < n_steps; ++k) {
        if (!growth.empty()) {
          NRSolver.calculateRelativePositions(Nodes, Elements);
          NRSolver.calculateGrowth(Nodes, Elements, t, dt);
        }
        tf_step_stats st = NRSolver.updateStepNR(Nodes, Elements, dt, maxit);
        norms.clear();
        for (int i = 0; i < st.nr_iterations; ++i) norms.push_back(st.du_norms[i]);
        t += dt;
      }
    } else if (mode == "step") {
      NRSolver.AblatedNodes = ablatedNodes;
      tf_step_stats st = NRSolver.updateStepNR(Nodes, Elements, dt, maxit);
      for (int k = 0; k < st.nr_iterations; ++k) norms.push_back(st.du_norms[k]);
    } else {
      // the loop of Simulation::updateStepNR with the reference's call names
      int iteratorK = 0;
      bool converged = false;
      NRSolver.setMatricesToZeroAtTheBeginningOfIteration(false);
      NRSolver.constructUnMatrix(Nodes);
      NRSolver.initialteUkMatrix();
      while (!converged) {
        NRSolver.calculateDisplacementMatrix(dt);
        NRSolver.calculateForcesAndJacobianMatrixNR(Nodes, Elements, dt);
        NRSolver.checkJacobianForAblatedNodes(ablatedNodes);
        NRSolver.solveForDeltaU();
        converged = NRSolver.checkConvergenceViaDeltaU();
        double s = 0.0;
        for (double d : NRSolver.deltaU) s += d * d;
        norms.push_back(std::sqrt(s));
        NRSolver.updateUkInIteration();
        NRSolver.updateNodePositionsNR(Nodes);
        for (auto& el : Elements) el->updatePositions(Nodes);
        iteratorK++;
        if (!converged && iteratorK > maxit) {
          std::cerr << "Error: did not converge!!!" << std::endl;
          converged = true;
        }
      }
    }
  } catch (const std::exception& ex) {
    std::cerr << ex.what() << "\n";
    return 1;
  }
  FILE* out = std::fopen(argv[2], "w");
  if (!out) return 1;
  for (auto& n : Nodes)
    std::fprintf(out, "%.17g %.17g %.17g\n", n->Position[0], n->Position[1], n->Position[2]);
  std::fprintf(out, "iterations %zu\n", norms.size());
  for (double d : norms) std::fprintf(out, "du %.17g\n", d);
  if (!growth.empty())
    for (auto& el : Elements) {
      std::fprintf(out, "Fg");
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) std::fprintf(out, " %.17g", el->Fg[i][j]);
      std::fprintf(out, "\n");
    }
  for (auto& el : Elements) {
    std::fprintf(out, "strain");
    for (int q = 0; q < 6; ++q) std::fprintf(out, " %.17g", el->Strain[q]);
    std::fprintf(out, "\n");
  }
  std::fclose(out);
  return 0;
}
