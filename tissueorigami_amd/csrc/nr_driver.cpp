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
//     checkJacobianForAblatedNodes after the assembly, Simulation.cpp:6691)
// Output: one line per node with the converged x y z (%.17g), then a line
//   "iterations <k>" and one "du <norm>" line per NR iteration.
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
  in.clear();
  std::string key;
  while (in >> key) {
    if (key == "stiffness") {
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
    if (mode == "step") {
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
  std::fclose(out);
  return 0;
}
