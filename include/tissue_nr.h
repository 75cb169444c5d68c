/*
 * tissue_nr.h -- C ABI of the MI355X implicit Newton-Raphson step on 6-node
 * prism meshes (libtissue_nr.so).
 *
 * The library replaces, for one tissue at a time, the hot path of
 * TissueOrigami's Simulation::updateStepNR (TissueFolding/SourceCode/
 * Simulation.cpp:6605-6746): the element loop of
 * NewtonRaphsonSolver::calculateForcesAndJacobianMatrixNR (NewtonRaphsonSolver.cpp:221-241)
 * over ShapeBase/Prism (ShapeBase.cpp:2356-3285, Prism.cpp:766-886), the
 * scatter into the system Jacobian (NewtonRaphsonSolver.cpp:342-360), the
 * boundary conditions (NewtonRaphsonSolver.cpp:128-219, 362-457) and the
 * Pardiso solve of solveForDeltaU (NewtonRaphsonSolver.cpp:461-763).
 *
 * Conventions (all arrays are caller-owned host memory, copied in or out
 * during the call; no pointer is retained):
 *   - DOF index = 3 * node + dim (NewtonRaphsonSolver.cpp:107-110);
 *     positions and forces are 3 * n_nodes doubles, node-major.
 *   - conn: n_prisms x 6 int32 node ids in the reference's prism order
 *     (basal triangle 0,1,2 then apical 3,4,5, counter-clockwise seen from
 *     the apical side, as left by Prism::checkRotationConsistency3D,
 *     Prism.cpp:369-417).
 *   - ref_pos: n_prisms x 18 reference (stress-free) node positions,
 *     ReferenceShapeBase::Positions.
 *   - Matrices are row-major.
 * Every entry point returns a tf_status; no entry point calls exit().
 * A handle is used by one host thread at a time and owns one HIP stream;
 * distinct handles may run concurrently.
 */
#ifndef TISSUE_NR_H
#define TISSUE_NR_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum tf_status {
  TF_OK = 0,
  TF_ERR_HIP = 1,            /* a HIP runtime call failed */
  TF_ERR_NAN = 2,            /* NaN/Inf in the residual or the increment */
  TF_ERR_KRYLOV = 3,         /* BiCGStab breakdown or iteration cap */
  TF_ERR_NOT_CONVERGED = 4,  /* NR hit maxit: result still written, as the
                                reference accepts the iterate
                                (Simulation.cpp:6724-6727) */
  TF_ERR_ARG = 5,            /* invalid argument or call order */
  TF_ERR_NO_DEVICE = 6       /* no usable gfx950 device */
} tf_status;

typedef enum tf_precision {
  TF_PREC_FP64 = 0,   /* all-FP64 step */
  TF_PREC_MIXED = 1   /* FP32 tangent blocks, FP32 SpMV and block-Jacobi,
                         FP64 residual, det F / log J and FP64 refinement */
} tf_precision;

#define TF_MAX_NR_ITERS 32

typedef struct tf_step_stats {
  int32_t status;                          /* tf_status of the step */
  int32_t nr_iterations;                   /* NR iterations run */
  int32_t converged;                       /* 1 if |du|_2 <= tol was reached */
  int32_t krylov_iterations_total;
  int32_t krylov_iterations[TF_MAX_NR_ITERS];
  double du_norms[TF_MAX_NR_ITERS];        /* |du|_2 per NR iteration
                                              (NewtonRaphsonSolver.cpp:815) */
  double ms_element;                       /* per-prism kernel */
  double ms_assembly;                      /* slot-map reduction + BCs */
  double ms_krylov;                        /* linear solves */
  double ms_total;                         /* whole step, device time */
  int32_t fp64_fallbacks;                  /* mixed precision: NR iterations whose
                                              FP32 solve failed and was redone
                                              with the FP64 operator */
  int32_t cg_fallbacks;                    /* NR iterations whose CG solve failed
                                              (indefinite or stalled) and was
                                              redone with BiCGStab */
  int32_t gmres_fallbacks;                 /* NR iterations whose BiCGStab solve
                                              failed and was redone with FP64
                                              GMRES (the last stage) */
} tf_step_stats;

typedef struct tf_handle tf_handle;

/* Creates a solver for one mesh: uploads connectivity and reference shape,
 * builds the block pattern of the Jacobian and the element -> block slot map.
 * Replaces NewtonRaphsonSolver::NewtonRaphsonSolver (NewtonRaphsonSolver.cpp:16-41)
 * and the per-element geometry set-up of Prism::Prism (Prism.cpp:9-251). */
tf_status tf_nr_create(int32_t n_nodes, int32_t n_prisms, const int32_t* conn,
                       const double* ref_pos, tf_handle** out);

/* Per-prism Young's modulus E, Poisson ratio nu and internal viscosity eta;
 * per-node external viscosity (n_nodes x 3); fixed DOF mask (n_nodes x 3,
 * Node::FixedPos, may be NULL); slave/master DOF pairs (n_pairs x 2,
 * NewtonRaphsonSolver::slaveMasterList, may be NULL).  With pairs, the solves
 * use the bound operator N^T J N + I_slave of calculateBoundKWithSlavesMasterDoF
 * (NewtonRaphsonSolver.cpp:128-183) stored explicitly on its own pattern, built
 * here when the pair list changes (a DOF is the slave of one master at most and
 * no master is a slave; TF_ERR_ARG otherwise).  On a batch handle the pairs
 * use the concatenated DOF numbering and both DOF of a pair must belong to
 * one tissue.  Lame parameters follow
 * Prism::setElasticProperties (Prism.cpp:464-515).  Replaces the physical
 * property state read by ShapeBase (ShapeBase.h:129-143) and Node (Node.h:21-58). */
tf_status tf_nr_set_material(tf_handle* h, const double* E, const double* nu,
                             const double* eta, const double* ext_visc,
                             const uint8_t* fixed, const int32_t* slave_master,
                             int32_t n_pairs);

/* Inverse growth deformation gradient per prism (n_prisms x 9, row-major),
 * ShapeBase::InvFg as left by ShapeBase::growShapeByFg (ShapeBase.cpp:1452-1505). */
tf_status tf_nr_set_growth(tf_handle* h, const double* inv_Fg);

/* Exposed surface area per node (Node::viscositySurface), computed at the
 * start of the step by Simulation::updateNodeViscositySurfaces
 * (Simulation.cpp:7598-7621). */
tf_status tf_nr_set_visc_surface(tf_handle* h, const double* area);

/* Laser ablation: ablated_prisms (n_prisms flags, ShapeBase::IsAblated; NULL:
 * none) contribute neither forces nor tangents to the step (their forces are
 * not computed and their TriPointKe / TriPointKv are zero,
 * NewtonRaphsonSolver.cpp:236, 344; ShapeBase.cpp:2850-2852, 2879-2882), no
 * viscosity surfaces (Simulation.cpp:7608-7621) and do not grow in
 * tf_nr_advance (Simulation.cpp:6055); ablated_nodes (Simulation::AblatedNodes,
 * n_ablated_nodes ids, may be NULL) get 1 on every Jacobian diagonal entry that
 * is exactly zero after the drag and packing terms
 * (NewtonRaphsonSolver::checkJacobianForAblatedNodes, NewtonRaphsonSolver.cpp:425-439,
 * called at Simulation.cpp:6691).  Each call replaces the previous flags.  On a
 * batch handle both use the concatenated numbering.  Not supported together
 * with node-node packing or by the adjoint (TF_ERR_ARG). */
tf_status tf_nr_set_ablation(tf_handle* h, const uint8_t* ablated_prisms,
                             const int32_t* ablated_nodes, int32_t n_ablated_nodes);

/* Computes the node viscosity surfaces on the device from positions u
 * (3 x n_nodes) and per-prism tissue placement / type / spans-whole-tissue
 * flags (ShapeBase.cpp:905-1037), then stores them as tf_nr_set_visc_surface
 * would.  Replaces Simulation::updateNodeViscositySurfaces with
 * Prism::calculateApicalArea/BasalArea (Prism.cpp:1485-1556). */
tf_status tf_nr_update_visc_surface(tf_handle* h, const double* u,
                                    const int32_t* elem_placement,
                                    const int32_t* elem_type, const uint8_t* spans,
                                    double* area_out);

/* tf_nr_update_visc_surface for callers that cannot allocate device memory
 * (loops under capture, serving loops, harnesses that forbid hipMalloc): the
 * same kernel and the same results, with the per-prism flags staged in the
 * library's static device scratch array (32 MiB per device, held until the
 * process ends) instead of buffers allocated and freed per call; it calls no
 * hipMalloc or hipFree.  Handles of up to about 3.7 million prisms (9 bytes
 * of scratch per prism); larger ones get TF_ERR_ARG.  All calls on one device
 * share the array: calls that overlap on one device (from several threads)
 * race and must be serialised by the caller.  No counterpart in the
 * reference; it exists for the callers above. */
tf_status tf_nr_update_visc_surface_noalloc(tf_handle* h, const double* u,
                                            const int32_t* elem_placement,
                                            const int32_t* elem_type, const uint8_t* spans,
                                            double* area_out);

/* One implicit time step: NR iterations from u_n until |du|_2 <= tol or
 * maxit (the reference uses tol 1e-8 and maxit 20,
 * NewtonRaphsonSolver.cpp:22, Simulation.cpp:6615).  Replaces
 * Simulation::updateStepNR.  krylov_rtol is the relative residual of each
 * linear solve (0 selects 1e-10).  u_out receives the positions. */
tf_status tf_nr_step(tf_handle* h, const double* u_n, double dt, double tol, int32_t maxit,
                     int32_t precision_mode, double krylov_rtol, double* u_out,
                     tf_step_stats* stats);

/* One assembly at positions uk (previous-step positions un, time step dt):
 * the per-prism kernel, the slot-map reduction and the boundary conditions.
 * g_out (3 x n_nodes, may be NULL) receives gSum after fixing and binding
 * (NewtonRaphsonSolver::gSum); internal_out (may be NULL) receives the
 * internal nodal forces of non-fixed DOFs (the SystemForces of Save_Force,
 * ShapeBase::writeInternalForcesTogeAndgv, ShapeBase.cpp:2362-2387).
 * Replaces calculateForcesAndJacobianMatrixNR + writeForcesTogeAndgvInternal
 * + writeImplicitElementalKToJacobian + calculateExternalViscousForcesForNR +
 * addImplicitKViscousExternalToJacobian + calculateSumOfInternalForces +
 * calcutateFixedK (NewtonRaphsonSolver.cpp:195-457). */
tf_status tf_nr_assemble(tf_handle* h, const double* uk, const double* un, double dt,
                         double* g_out, double* internal_out);

/* tf_nr_assemble in a chosen precision (no reference counterpart): with
 * TF_PREC_MIXED the element tangents and J are FP32 (forces and g stay FP64);
 * tf_nr_element_outputs then returns the FP32 tangents. */
tf_status tf_nr_assemble_ex(tf_handle* h, const double* uk, const double* un, double dt,
                            int32_t precision_mode, double* g_out, double* internal_out);

/* Solves J x = b with the Jacobian of the last tf_nr_assemble (J = K^T, the
 * transposed solve of solveWithPardiso, NewtonRaphsonSolver.cpp:716-723;
 * binding handled as N^T J N + I_slave, NewtonRaphsonSolver.cpp:128-193).
 * Returns TF_ERR_KRYLOV if rtol was not reached within maxit iterations. */
tf_status tf_nr_solve(tf_handle* h, const double* b, double* x, double rtol, int32_t maxit,
                      int32_t* iterations);

/* Structural CSR pattern of K (scalar rows, ascending columns, 0-based), the
 * superset of constructiaForPardiso's thresholded pattern
 * (NewtonRaphsonSolver.cpp:766-787).  ia has 3*n_nodes+1 entries; pass
 * ja == NULL to query *nnz first. */
tf_status tf_csr_pattern(tf_handle* h, int32_t* ia, int32_t* ja, int64_t* nnz);

/* The assembled K of the last tf_nr_assemble in the reference orientation,
 * as constructiaForPardiso builds it: entries with |K_ij| > threshold (the
 * reference uses 1e-13), row-major, ascending columns, 0-based.  Binding
 * pairs are applied as calculateBoundKWithSlavesMasterDoF does.  After a
 * mixed assembly the values are those of the FP32 operator.  Pass
 * ja == NULL to query *nnz first. */
tf_status tf_nr_get_csr(tf_handle* h, double threshold, int32_t* ia, int32_t* ja, double* a,
                        int64_t* nnz);

/* Per-prism outputs of the last assembly, for element-level parity (the
 * tangents are the FP32 ones after a mixed tf_nr_assemble_ex, recomputed with
 * the forces from that call's inputs: the mixed assembly itself keeps only
 * the per-block sums of the tangents and the summed forces, so after a mixed
 * assembly every output needs tf_nr_assemble_ex as the last assembly):
 * elastic and internal viscous element forces (n_prisms x 18,
 * ElementalElasticSystemForces / ElementalInternalViscousSystemForces before
 * fixed-DOF masking) and the 18 x 18 tangents TriPointKe / TriPointKv in the
 * reference's storage orientation (n_prisms x 324).  Any output may be NULL. */
tf_status tf_nr_element_outputs(tf_handle* h, double* ge, double* gv, double* Ke, double* Kv);

/* Per-prism Green-Lagrange strain E = (Fe^T Fe - I) / 2 of the last Gauss
 * point, as Prism::calculateCurrNodalForces keeps it in ShapeBase::Strain
 * (Prism.cpp:826-837) and Simulation::writeTensionCompression saves it
 * (Simulation.cpp:8521-8531): n_prisms x 6 in the order E00, E11, E22,
 * 2 E01, 2 E21, 2 E02, taken at the positions of the last assembly (in an
 * NR step, the last iterate before the final increment, as in the
 * reference).  In a batch each tissue keeps the strain of its own last NR
 * iteration; ablated prisms give zero. */
tf_status tf_nr_element_strain(tf_handle* h, double* strain);

/* Device-resident time stepping with the pre-step producers on the device.
 * tf_nr_set_growth_rates: per-prism growth rates (n_prisms x 3, x/y/z, 1/s)
 *   and F_g reset to identity.  Every step of tf_nr_advance applies
 *   ShapeBase::calculateFgFromRates (ShapeBase.cpp:471-505, no rotation) and
 *   ShapeBase::growShapeByFg (ShapeBase.cpp:1452-1505): F_g <- diag(exp(r dt)) F_g,
 *   InvFg = inv(F_g).
 * tf_nr_growth_tensor: reads (Fg_out) and/or sets (Fg_in) F_g (n_prisms x 9).
 * tf_nr_set_surface_flags: per-prism placement / type / spans-whole-tissue
 *   used to recompute the node viscosity surfaces on the device at the start
 *   of every step (Simulation::updateNodeViscositySurfaces, Simulation.cpp:7598-7621).
 * tf_nr_advance: n_steps implicit steps; the start positions are u0 or, when
 *   u0 is NULL, the device-resident result of the previous call.  stats (may be
 *   NULL) receives one entry per step; u_out (may be NULL) the final positions.
 *   Replaces the per-step sequence of Simulation::runOneStep
 *   (Simulation.cpp:6012-6112) restricted to growth, surfaces and updateStepNR. */
tf_status tf_nr_set_growth_rates(tf_handle* h, const double* rates);
tf_status tf_nr_growth_tensor(tf_handle* h, double* Fg_out, const double* Fg_in);
tf_status tf_nr_set_surface_flags(tf_handle* h, const int32_t* elem_placement,
                                  const int32_t* elem_type, const uint8_t* spans);
tf_status tf_nr_advance(tf_handle* h, const double* u0, int32_t n_steps, double dt, double tol,
                        int32_t maxit, int32_t precision_mode, double krylov_rtol, double* u_out,
                        tf_step_stats* stats);

/* Growth functions on the device (Simulation::calculateGrowth,
 * Simulation.cpp:8640-8660, with uniform and grid-based functions).  Once
 * they are added, every step of tf_nr_advance runs, before the NR loop and in
 * Simulation::runOneStep's order (Simulation.cpp:5974-6063): the relative
 * positions of the prisms in the tissue's bounding box, the stiffness series,
 * the pinning updates, each prism's growth rotation about z, the growth of
 * every function active at the step's start time, the column volume
 * conservation, and F_g / InvFg; fixed growth rates (tf_nr_set_growth_rates)
 * are then not used.  On a batch handle the arrays run over the tissues'
 * prisms (and nodes) concatenated; the functions apply to every tissue, each
 * in its own bounding box.
 * tf_nr_set_growth_prisms: per prism ShapeBase::tissueType (0 columnar, 1
 *   peripodial, 2 linker), tissuePlacement (0 basal, 1 apical, 2 middle, 4
 *   lateral), flags (1 isECMMimicing, 2 isECMMimimcingAtCircumference, 4
 *   isActinMimicing), weights (n_prisms x 2 or NULL: columnarGrowthWeight,
 *   peripodialGrowthWeight of linker prisms), rel_z
 *   (initialRelativePositionInZ, ShapeBase.cpp:817-828); apical_of (NULL or
 *   per prism: the prism whose relative position it takes,
 *   Simulation::updateRelativePositionsToApicalPositioning,
 *   Simulation.cpp:6147-6167, -1 for its own); node_excluded (NULL or per
 *   node: Node::allOwnersECMMimicing, left out of the bounding box,
 *   Simulation::calculateBoundingBox, Simulation.cpp:8388-8428);
 *   symmetric_y (Simulation::symmetricY); explicit_ecm
 *   (Simulation::thereIsExplicitECM: ECM prisms take no columnar or
 *   peripodial growth).
 * tf_nr_add_uniform_growth: UniformGrowthFunction (GrowthFunctionTypes.h:14-97;
 *   ModelInputObject::readGrowthType1, ModelInputObject.cpp:403-517) active
 *   for t0 <= t < t1 (kept as float, as GrowthFunctionBase.h:39-40 does);
 *   apply: 1 columnar layer, 2 peripodial membrane, 4 basal ECM, 8 lateral
 *   ECM; [zmin, zmax] the rel_z range; rate (x, y, z, 1/s) and angle
 *   (radians) as the function holds them.
 * tf_nr_add_grid_growth: GridBasedGrowthFunction (GrowthFunctionTypes.h:290-546;
 *   ModelInputObject::readGrowthType3, ModelInputObject.cpp:684-870) on an
 *   nx x ny grid: rates[(i * ny + j) * 3 + k] (1/s) and angles_deg[i * ny + j]
 *   (degrees) at grid point (i, j), i along x.
 * tf_nr_set_growth_options: Simulation::gridGrowthsInterpolationType (0
 *   nearest corner, 1 bilinear), GridGrowthsPinnedOnInitialMesh and the
 *   pinning update times (s; each done once, at the first step whose start
 *   time reaches it, Simulation.cpp:6169-6187).
 * tf_nr_set_column_volume_conservation: Simulation::conservingColumnVolumes
 *   (Simulation.cpp:6040-6046, 6518-6600): columns[n_columns][layers] holds
 *   the prisms of each column whose volume is conserved (the columnar, not
 *   ECM, not actin prisms of an apical prism's stack; -1 pads).
 * tf_nr_update_relative_positions: the relative positions at positions u
 *   (NULL: the device-resident positions); the first call also sets the
 *   pinned positions (Simulation::initiateSystem, Simulation.cpp:741-752).
 *   Stiffness series use these from then on.
 * tf_nr_relative_positions: current and pinned positions (n_prisms x 2; either
 *   may be NULL).
 * tf_nr_grow: one step's growth at time t from positions u (NULL: the
 *   device-resident positions), with the current relative positions. */
tf_status tf_nr_set_growth_prisms(tf_handle* h, const int32_t* tissue_type,
                                  const int32_t* placement, const uint8_t* flags,
                                  const double* weights, const double* rel_z,
                                  const int32_t* apical_of, const uint8_t* node_excluded,
                                  int32_t symmetric_y, int32_t explicit_ecm);
tf_status tf_nr_add_uniform_growth(tf_handle* h, double t0, double t1, int32_t apply, double zmin,
                                   double zmax, const double* rate, double angle);
tf_status tf_nr_add_grid_growth(tf_handle* h, double t0, double t1, int32_t apply, double zmin,
                                double zmax, int32_t nx, int32_t ny, const double* rates,
                                const double* angles_deg);
tf_status tf_nr_set_growth_options(tf_handle* h, int32_t interpolation, int32_t pinned,
                                   int32_t n_pin_times, const double* pin_times);
tf_status tf_nr_set_column_volume_conservation(tf_handle* h, int32_t n_columns, int32_t layers,
                                               const int32_t* columns);
tf_status tf_nr_update_relative_positions(tf_handle* h, const double* u);
tf_status tf_nr_relative_positions(tf_handle* h, double* current, double* pinned);
tf_status tf_nr_grow(tf_handle* h, const double* u, double t, double dt);

/* Young's modulus time series (YoungsModulusModifier, YoungsModulusModifier.cpp:1-177,
 * with its base TimeSeriesPhysicalProperties).  On a batch handle the per-prism
 * arrays run over the tissues' prisms concatenated, each tissue's relative
 * positions taken in its own bounding box.
 * tf_nr_add_stiffness_series replaces the YoungsModulusModifier constructor
 * (ModelInputObject.cpp:1989-1990): n_times grids of ny x nx multipliers
 * (grids[t][iy][ix], the rows of the reference's grid files) applied at
 * times[t] (s); linear != 0 for a linear change of the multiplier, 0 for an
 * exponential one; method 0 (nearest grid point), 1 (bilinear) or 2 (lower
 * corner) for the rate at a prism's position.  applicable (n_prisms) holds
 * YoungsModulusModifier::IsItApplicable per prism (the placement / type / ECM
 * test of YoungsModulusModifier.cpp:74-100); rel_pos (n_prisms x 2) the
 * prism's relative (x, y) position in the tissue's bounding box
 * (ShapeBase::relativePosInBoundingBox, after the apical positioning of
 * Simulation.cpp:6147-6167), used until the device computes them
 * (tf_nr_update_relative_positions, and every step of tf_nr_advance once the
 * growth prisms are set); NULL when the growth prisms are set.  Needs
 * tf_nr_set_material first.
 * tf_nr_apply_stiffness_series replaces Simulation::ApplyAllYoungsModulusModifiers
 *   (Simulation.cpp:878-884) at time t: each series in the order added, over
 *   the interval [times[k], times[k+1]) that holds t (none when t lies past
 *   the last time or before the first), updates the time-series multiplier
 *   of every applicable prism by rate * dt (linear) or exp(rate * dt), and the
 *   prism's Lame parameters (Prism::updateElasticProperties, Prism.cpp:526-527:
 *   multiplier x E as set by tf_nr_set_material).
 * tf_nr_advance applies the series at the start of every step, at the handle's
 *   time (tf_nr_set_time, default 0; the time advances by dt per step, as
 *   Simulation::currSimTimeSec does, Simulation.cpp:6129).
 * tf_nr_get_stiffness_multiplier: the multiplier per prism
 *   (ShapeBase::StiffnessTimeSeriesMultiplier, 1 until a series changes it). */
tf_status tf_nr_add_stiffness_series(tf_handle* h, int32_t n_times, const double* times,
                                     int32_t nx, int32_t ny, const double* grids, int32_t linear,
                                     int32_t method, const uint8_t* applicable,
                                     const double* rel_pos);
tf_status tf_nr_apply_stiffness_series(tf_handle* h, double t, double dt);
tf_status tf_nr_set_time(tf_handle* h, double t);
tf_status tf_nr_get_time(tf_handle* h, double* t);
tf_status tf_nr_get_stiffness_multiplier(tf_handle* h, double* out);

/* Node-node packing (volume exclusion between surface nodes), the contact
 * terms Simulation::updateStepNR adds inside every NR iteration
 * (Simulation.cpp:6680-6698).  Single-tissue handles without slave/master
 * binding; not supported by the adjoint.
 * tf_nr_set_packing: turns packing on (node_placement != NULL) or off.
 *   node_placement: Node::tissuePlacement per node (0 basal, 1 apical; other
 *   values never pack); node_excluded (may be NULL): nodes that never pack
 *   (Node::hasLateralElementOwner, Node.cpp:128-146); symmetric_y: the
 *   bounding box is mirrored in y (Simulation::symmetricY,
 *   Simulation.cpp:8412-8414); capacity: couples per node the Jacobian's
 *   contact columns hold at first (1..32; widened by detection when needed).
 *   The Jacobian pattern is laid out again with the contact columns.
 * tf_nr_set_node_mass: Node::mass per node (Simulation::updateNodeMasses,
 *   Simulation.cpp:7587-7596).  tf_nr_update_node_mass computes it on the
 *   device from F_g (n_prisms x 9, or NULL for the device-resident F_g of
 *   tf_nr_advance) and the prisms' reference volumes
 *   (Prism::calculateReferenceVolume, Prism.cpp:888-922); mass_out may be NULL.
 * tf_nr_detect_packing: Simulation::detectPacingNodes (Simulation.cpp:7997-8078)
 *   at positions u (NULL: the device-resident positions of tf_nr_advance):
 *   the thresholds of getAverageSideLength (Simulation.cpp:3197-3260) and the
 *   bounding box (Simulation.cpp:8388-8425) from the element placement and
 *   type of tf_nr_set_surface_flags, and the node couples (without the
 *   emergent-ellipse release of curved regions).  The couples hold for every
 *   following tf_nr_step / tf_nr_assemble until the next detection;
 *   tf_nr_advance detects at the start of every step, after the masses.
 * tf_nr_get_packing: the couples of the last detection in the reference's
 *   order (ascending master, then partner) and, per couple, the packing
 *   forces of its two nodes in the last residual (6 doubles, as
 *   Simulation::writePacking, Simulation.cpp:8601-8620); pass couples and
 *   forces NULL to query the count.  thresholds (may be NULL, 57 doubles):
 *   packingThreshold, the 50 squared detection thresholds of the 10 x 5 grid
 *   (row-major), the bounding box (min x, y, z, max x, y, z). */
tf_status tf_nr_set_packing(tf_handle* h, const int32_t* node_placement,
                            const uint8_t* node_excluded, int32_t symmetric_y, int32_t capacity);
tf_status tf_nr_set_node_mass(tf_handle* h, const double* mass);
tf_status tf_nr_update_node_mass(tf_handle* h, const double* Fg, double* mass_out);
tf_status tf_nr_detect_packing(tf_handle* h, const double* u, int32_t* n_couples);
tf_status tf_nr_get_packing(tf_handle* h, int32_t* n_couples, int32_t* couples, double* forces,
                            double* thresholds);

/* Batch of independent tissues in one handle (no counterpart in the
 * reference, which simulates one tissue per process).  The tissues are
 * given concatenated: nodes_per_tissue / prisms_per_tissue, conn with node
 * ids local to each tissue, ref_pos of all prisms in order.  Every per-node
 * array of the other entry points (positions, external viscosity, fixed DOF,
 * surfaces) is the concatenation of the tissues' arrays; per-prism arrays
 * likewise.  Every kernel of the step runs once per stage for the whole
 * batch; every tissue keeps its own NR convergence test and its own Krylov
 * scalars, so its result equals the tissue stepped alone.  Step statistics
 * are reported per tissue: tf_nr_step fills n_tissues entries, tf_nr_advance
 * n_steps x n_tissues.  Binding pairs (tf_nr_set_material) bind DOF within
 * a tissue.  tf_nr_assemble/solve/get_csr/csr_pattern and the adjoint take
 * single-tissue handles only. */
tf_status tf_batch_create(int32_t n_tissues, const int32_t* nodes_per_tissue,
                          const int32_t* prisms_per_tissue, const int32_t* conn,
                          const double* ref_pos, tf_handle** out);

/* Tile size of the fused Krylov solver (consecutive 64-row slices per
 * workgroup; a tile never spans two tissues).  tf_nr_autotune times 10
 * iterations of the precision's Krylov method (tf_nr_set_krylov) on the
 * assembled operator of that precision for each candidate size and keeps the
 * fastest; tf_nr_set_tile_size sets it.
 * No reference counterpart. */
tf_status tf_nr_autotune(tf_handle* h, int32_t precision_mode, int32_t* slices_per_tile,
                         double* ms_per_iteration);
tf_status tf_nr_set_tile_size(tf_handle* h, int32_t slices_per_tile);

/* Krylov method of the linear solves of one precision mode (no reference
 * counterpart; the reference factorises with Pardiso).  BiCGStab handles
 * most Jacobians.  CG (one SpMV per iteration) suits Jacobians that are
 * symmetric positive definite up to the small viscous asymmetry; a tissue
 * whose CG solve meets non-positive curvature or stalls is solved again with
 * BiCGStab in the same NR iteration (tf_step_stats.cg_fallbacks).  A tissue
 * whose BiCGStab solve breaks down (typically on the indefinite Jacobians of
 * node-node packing) is solved again with restarted FP64 GMRES(100) with
 * block-Jacobi right preconditioning (tf_step_stats.gmres_fallbacks); GMRES
 * may also be chosen for the FP64 mode.  Defaults: FP64 BiCGStab, mixed CG. */
enum { TF_KRYLOV_BICGSTAB = 0, TF_KRYLOV_CG = 1, TF_KRYLOV_GMRES = 2 };
tf_status tf_nr_set_krylov(tf_handle* h, int32_t precision_mode, int32_t method);

/* Adjoint of n_steps implicit steps with respect to one growth parameter r_e
 * per prism.  No counterpart in the reference (which has no derivatives); the
 * forward is n_steps of tf_nr_step (FP64) that also keeps every step's
 * converged positions on the device as checkpoints.
 *   tf_adj_forward: u0 (3 x n_nodes); inv_Fg_steps (n_steps x n_prisms x 9)
 *     is the InvFg of each step; the node viscosity surfaces of each step are
 *     computed on the device from the step's start positions with the given
 *     per-prism placement / type / spans flags (as tf_nr_update_visc_surface).
 *   tf_adj_backward: given dL/du_final (3 x n_nodes) and dinv_Fg_steps, the
 *     derivative of every step's InvFg_e with respect to r_e, returns
 *     dL_dr (n_prisms) by the implicit-function theorem per step: one
 *     transposed-Jacobian solve J^T mu = a per step (relative residual
 *     krylov_rtol, 0 selects 1e-12) and hand-derived dg/dInvFg and
 *     dg/du_{n-1} (internal viscosity, external drag and its surface areas).
 * Slave/master binding is not supported (TF_ERR_ARG). */
tf_status tf_adj_forward(tf_handle* h, int32_t n_steps, const double* u0,
                         const double* inv_Fg_steps, double dt, double tol, int32_t maxit,
                         double krylov_rtol, const int32_t* elem_placement,
                         const int32_t* elem_type, const uint8_t* spans, double* u_final,
                         tf_step_stats* stats);
tf_status tf_adj_backward(tf_handle* h, const double* inv_Fg_steps, const double* dinv_Fg_steps,
                          const double* dL_du_final, double krylov_rtol, double* dL_dr,
                          tf_step_stats* stats);

/* Device-time accounting of the hot kernels, for measurement.  The reference
 * has no timers (no counterpart); these exist for the roofline figures.
 * tf_nr_profile(h, 1) resets the counters and brackets every SpMV launch
 * with HIP events on the handle's stream; tf_nr_kernel_times reads them. */
typedef struct tf_kernel_times {
  double spmv_ms[2];          /* summed durations of the first SpMV of each Krylov
                                 iteration: BiCGStab's v = J ph (fused with
                                 s = r - alpha v, sh = M^-1 s) or CG's q = J p;
                                 [0] FP64, [1] FP32 */
  int64_t spmv_launches[2];
  double spmv_t_ms[2];        /* the SpMV t = J sh fused with (t, s), (t, t) */
  int64_t spmv_t_launches[2];
  double element_ms;          /* per-prism kernel, summed over assemblies */
  double assembly_ms;         /* slot-map reduction + residual, summed */
  int64_t assemblies;
} tf_kernel_times;

tf_status tf_nr_profile(tf_handle* h, int32_t enable);
tf_status tf_nr_kernel_times(tf_handle* h, tf_kernel_times* out);

/* Device copy bandwidth: for each of six plain 16-byte-vector copy kernels
 * (tf_copy_kernel: 1, 2 or 4 vectors per thread, ordinary or non-temporal
 * accesses) `reps` launches over `bytes` bytes timed with HIP events; the
 * best rate in GB/s counting bytes read + written.  The STREAM-copy ceiling
 * the kernels' HBM fractions are set against.  No reference counterpart. */
tf_status tf_copy_bandwidth(int64_t bytes, int32_t reps, double* gb_per_s);

/* Replays of the Krylov iteration chunks through captured hipGraphs
 * (default on; profiling launches directly).  No reference counterpart. */
tf_status tf_nr_set_graphs(tf_handle* h, int32_t enable);

/* Two-level preconditioner of the mixed-precision CG solve (no reference
 * counterpart; in a batch handle every tissue has a coarse space of its own,
 * the one it has alone).  Block Jacobi is completed by a coarse
 * correction P A_c^-1 P^T r, where P holds the rigid-body modes (three
 * translations, three rotations) of node aggregates: the nodes whose
 * reference (x, y) lies in one square cell of side cell_size and whose
 * reference z lies in one of z_bands equal bands of the thickness (default
 * 3, tf_nr_set_coarse_bands).  A_c = P^T J P is formed on the device and
 * inverted (FP64) at the first NR iteration of every mixed step
 * (tf_nr_set_coarse_lag).  cell_size < 0: automatic (the default): about
 * sqrt(stored blocks) coarse DOF per tissue, so that the dense coarse product stays near
 * a tenth of an FP32 SpMV; 0: off (block Jacobi only).  The
 * converged positions do not depend on it, only the Krylov iteration count.
 * tf_nr_coarse_info reports the aggregates of the last build and whether the
 * last mixed solve used the correction. */
tf_status tf_nr_set_coarse(tf_handle* h, double cell_size);
/* Splits every aggregate cell into z_bands equal bands of the reference
 * thickness (default 3; 1: whole columns), so the coarse space also holds
 * through-thickness modes.  Handles with slave/master binding keep whole
 * columns (and their automatic cell gives at most about 400 cells).  No
 * reference counterpart. */
tf_status tf_nr_set_coarse_bands(tf_handle* h, int32_t z_bands);
/* lag != 0 (the default): A_c and its inverse are formed at the first NR
 * iteration of a step and kept for the step's later iterations; 0: at every
 * mixed assembly.  No reference counterpart. */
tf_status tf_nr_set_coarse_lag(tf_handle* h, int32_t lag);
tf_status tf_nr_coarse_info(tf_handle* h, int32_t* aggregates, int32_t* active);
/* enable != 0 (the default): within one tf_nr_advance trajectory, the first
 * Krylov solve of every step after the first starts from the displacement
 * that the tissue's last one or two converged steps extrapolate to (linear
 * extrapolation of the step displacements), and is handed that guess's
 * residual with the tolerance still measured against the full residual
 * (rtol |g|), so the increment meets the same test as a solve from zero.  A
 * tissue whose guess does not halve |g| solves from zero.  Positions handed
 * in (tf_nr_advance with u0, tf_nr_step, tf_adj_forward) start a new
 * trajectory without history.  0: every solve starts from zero.  No
 * reference counterpart: Pardiso, the reference's direct solver, takes no
 * initial guess. */
tf_status tf_nr_set_warm_start(tf_handle* h, int32_t enable);
/* enable != 0 (the default): with a Krylov rtol of 1e-6 or more (inexact
 * Newton, rtol the forcing term), a solve that the step's last two increments
 * predict to be its last (next increment e = |du_{k-1}|^2 / |du_{k-2}|) runs
 * to the relative tolerance min(0.1, max(rtol, 0.5 tol / e)) instead of rtol:
 * the safeguard against oversolving of Eisenstat and Walker's forcing terms.
 * Its increment still decides convergence to within ten per cent, and leaves
 * at most about a tenth of an increment already below tol.  Tight solves
 * (rtol below 1e-6, the default 1e-10 of the FP64 step) are never loosened.
 * 0: every solve runs to rtol.  No reference counterpart: Pardiso solves
 * exactly. */
tf_status tf_nr_set_forcing_safeguard(tf_handle* h, int32_t enable);

/* Size of the Jacobian pattern: stored 3x3 blocks (n_blocks, the structural
 * nonzeros / 9) and padded SELL-64 slots.  No reference counterpart. */
tf_status tf_nr_pattern_info(tf_handle* h, int64_t* n_blocks, int64_t* n_slots);

/* Releases every device buffer of the handle (NewtonRaphsonSolver::~NewtonRaphsonSolver,
 * NewtonRaphsonSolver.cpp:43-71). */
tf_status tf_nr_destroy(tf_handle* h);

/* Human-readable text of the last error on this thread. */
const char* tf_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* TISSUE_NR_H */
