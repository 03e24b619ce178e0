/*
 * akx.h -- C ABI of the MI355X explicit-dynamics engine (libakx.so).
 *
 * The engine replaces, on one GPU, the data-parallel hot path of Akantu's
 * SolidMechanicsModel explicit central-difference step with lumped mass:
 * FEEngine::gradientOnIntegrationPoints, Material::computeStress,
 * FEEngine::computeBtD + integrate, DOFManager::assembleElementalArrayLocalArray
 * and the NewmarkBeta(0, 1/2) predictor / lumped solve / corrector.
 *
 * Conventions (kept from the reference, src/common/aka_array.hh:139-224,
 * aka_types.hh:879,1068):
 *   - nodal arrays are (nb_nodes, dim) row-major float64 (blocked DOFs: uint8);
 *   - connectivity is (nb_elements, nodes_per_element) uint32, reference local
 *     node order;
 *   - per-quadrature-point d x d tensors are column-major, points element-major
 *     (index = element * nb_qp + q).
 * Host pointers are owned by the caller; the context owns every device
 * buffer.  All calls return AKX_OK (0) or a negative status; akx_last_error()
 * holds the message (the reference throws debug::Exception, aka_error.hh:100).
 * One host thread per context.  Calls are asynchronous on the context's stream
 * except the ones that return host data, which synchronise.
 */
#ifndef AKX_H
#define AKX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct akx_ctx akx_ctx;

enum akx_status {
  AKX_OK = 0,
  AKX_ERR_ARGUMENT = -1,
  AKX_ERR_STATE = -2,
  AKX_ERR_HIP = -3,
  AKX_ERR_JACOBIAN = -4,
  AKX_ERR_UNSUPPORTED = -5
};

/* element types (src/common/aka_element_classes_info.hh.in) */
enum akx_element_type {
  AKX_TRIANGLE_3 = 0,
  AKX_QUADRANGLE_4 = 1,
  AKX_TETRAHEDRON_4 = 2,
  AKX_HEXAHEDRON_8 = 3
};

/* constitutive laws */
enum akx_material_kind {
  AKX_MATERIAL_ELASTIC = 0,           /* MaterialElastic<dim>                 */
  AKX_MATERIAL_NEOHOOKEAN = 1,        /* MaterialNeohookean<dim>              */
  AKX_MATERIAL_NEOHOOKEAN_DAMAGE = 2  /* neo-Hookean + Marigo isotropic damage */
};

/* arithmetic of the element kernels */
enum akx_precision {
  AKX_FP64 = 0,  /* reference arithmetic, element-order sums (bit-exact path) */
  AKX_MIXED = 1  /* FP32 element kernels, FP64 nodal state + compensated sums  */
};

/* fields for upload / download */
enum akx_field {
  AKX_DISPLACEMENT = 0,          /* (N,d) f64 */
  AKX_VELOCITY = 1,              /* (N,d) f64 */
  AKX_ACCELERATION = 2,          /* (N,d) f64 */
  AKX_EXTERNAL_FORCE = 3,        /* (N,d) f64 */
  AKX_INTERNAL_FORCE = 4,        /* (N,d) f64, of the last step / assembly (writable) */
  AKX_MASS = 5,                  /* (N,d) f64, lumped (download only) */
  AKX_BLOCKED_DOFS = 6,          /* (N,d) u8 */
  AKX_PREVIOUS_DISPLACEMENT = 7, /* (N,d) f64, u before the last step's predictor */
  AKX_DISPLACEMENT_INCREMENT = 8,/* (N,d) f64, u - previous after the last step */
  AKX_GRAD_U = 9,                /* (E*Q, d*d) f64, on the current displacement */
  AKX_STRESS = 10,               /* (E*Q, d*d) f64, Cauchy (small strain) or PK2 */
  AKX_DAMAGE = 11,               /* (E*Q) f64 */
  AKX_SHAPES_DERIVATIVES = 12,   /* (E*Q, d*n) f64, B = dN/dX, column-major d x n */
  AKX_JACOBIANS = 13,            /* (E*Q) f64, det(J) * w */
  AKX_CURRENT_POSITION = 14      /* (N,d) f64, X + u (download only) */
};

enum akx_energy_kind {
  AKX_ENERGY_KINETIC = 0,   /* SolidMechanicsModel::getKineticEnergy      */
  AKX_ENERGY_POTENTIAL = 1, /* Material::getPotentialEnergy summed        */
  AKX_ENERGY_EXTERNAL_WORK = 2 /* SolidMechanicsModel::getExternalWork    */
};

/* akx_ctx_create flags */
enum akx_ctx_flags {
  /* Keep the caller's node and element numbering on the device.  By default
   * both are renumbered along a Morton curve for memory locality; host arrays
   * crossing the ABI are always in the caller's numbering. */
  AKX_CTX_KEEP_NUMBERING = 1u
};

/* Context (device, stream, buffers).  Replaces the construction of
 * SolidMechanicsModel (src/model/solid_mechanics/solid_mechanics_model.cc:71-103). */
int akx_ctx_create(int device, unsigned flags, akx_ctx** out);
int akx_ctx_destroy(akx_ctx* ctx);
const char* akx_last_error(const akx_ctx* ctx);
/* Use a caller-owned hipStream_t for all work of this context (NULL: the
 * context's own stream). */
int akx_set_stream(akx_ctx* ctx, void* hip_stream);

/* Mesh upload.  Replaces Mesh::getNodes / getConnectivity as seen by
 * FEEngineTemplate (src/mesh/mesh.hh:100-611); builds the node -> element
 * slot CSR of MeshUtils::buildNode2Elements (src/mesh_utils/mesh_utils.cc:51-95)
 * in assembly order. */
int akx_mesh_set(akx_ctx* ctx, int element_type, int dim, int64_t nb_nodes,
                 const double* nodes, int64_t nb_elements, const uint32_t* conn);

/* Materials.  Replaces MaterialFactory instantiation + Material::initMaterial
 * (src/model/solid_mechanics/material.hh:59-60,180; material_elastic.cc:83-107;
 * material_neohookean.cc:83-89).  Returns the material index in *mat_id. */
int akx_material_add(akx_ctx* ctx, int kind, double E, double nu, double rho,
                     int plane_stress, int* mat_id);
/* Damage parameters of a NEOHOOKEAN_DAMAGE material: Sd and the per-point
 * thresholds Yd (E*Q values in global element order; only the material's own
 * points are read).  Replaces MaterialMarigo's Sd / Yd random internal field
 * (material_marigo.cc:39-57, random_internal_field_tmpl.hh:97-102). */
int akx_material_damage(akx_ctx* ctx, int mat_id, double Sd, const double* Yd);
/* Element -> material map (SolidMechanicsModel::assignMaterialToElements,
 * solid_mechanics_model_material.cc:120-165).  NULL: every element -> 0. */
int akx_assign_materials(akx_ctx* ctx, const int32_t* element_material);

/* Precomputation: B = dN/dX and det(J) w at the reference configuration, lumped
 * mass.  Replaces FEEngine::initShapeFunctions (fe_engine_template_tmpl.hh:154-167,
 * shape_lagrange_inline_impl.hh:204-252, integrator_gauss_inline_impl.hh:208-252,
 * 396-423) and SolidMechanicsModel::assembleMassLumped
 * (solid_mechanics_model_mass.cc:58-102).  Fails with AKX_ERR_JACOBIAN on a
 * negative Jacobian (integrator_gauss_inline_impl.hh:182-205). */
int akx_init(akx_ctx* ctx);

/* Host <-> device copies of the public arrays (SolidMechanicsModel::getX(),
 * solid_mechanics_model.hh:363-397).  bytes must equal the field size. */
int akx_field_upload(akx_ctx* ctx, int field, const void* host, size_t bytes);
int akx_field_download(akx_ctx* ctx, int field, void* host, size_t bytes);
/* Asynchronous output (the reference's dumpers read the same arrays
 * synchronously: Model::dump, src/model/model.cc, through
 * src/io/dumper/dumper_iohelper.cc).  akx_field_download_async snapshots
 * `field` as it stands after the work already enqueued on the context's
 * stream -- in the caller's numbering and layout, the same bytes
 * akx_field_download writes -- and copies it to `host` on a second stream
 * while later steps run.  `host` must be page-locked (akx_host_alloc) and
 * left untouched until akx_download_wait(ticket) has returned or
 * akx_download_query reports it done.  Not for the static tables
 * (AKX_SHAPES_DERIVATIVES, AKX_JACOBIANS).  akx_download_wait and
 * akx_download_query may be called from another host thread than the one
 * that steps the context. */
int akx_host_alloc(size_t bytes, void** host);
int akx_host_free(void* host);
int akx_field_download_async(akx_ctx* ctx, int field, void* host, size_t bytes, uint64_t* ticket);
int akx_download_query(akx_ctx* ctx, uint64_t ticket, int* done);
int akx_download_wait(akx_ctx* ctx, uint64_t ticket);
/* Device address of a nodal field (for zero-copy interop on the same device).
 * Device arrays are in the internal node numbering (see akx_node_numbering). */
int akx_field_device_pointer(akx_ctx* ctx, int field, void** device_ptr);
/* internal_to_caller[k] = caller id of internal node k (nb_nodes entries). */
int akx_node_numbering(akx_ctx* ctx, int64_t* internal_to_caller);

/* SolidMechanicsModel::setTimeStep (solid_mechanics_model.cc:438). */
int akx_set_time_step(akx_ctx* ctx, double dt);

/* nsteps x SolidMechanicsModel::solveStep() with the explicit lumped-mass
 * solver: NonLinearSolverLumped::solve (non_linear_solver_lumped.cc:58-82) ->
 * predictor (newmark-beta.cc:60-89) -> assembleResidual
 * (solid_mechanics_model.cc:279-293, 375-414; integration_scheme_2nd_order.cc:79-102)
 * -> solveLumped (non_linear_solver_lumped.cc:85-98) -> corrector
 * (newmark-beta.cc:176-210).  Per-step Dirichlet increments registered with
 * akx_set_dirichlet_increment are applied before each predictor. */
int akx_solve_step(akx_ctx* ctx, int nsteps, int precision);

/* akx_solve_step as one HIP graph launch: the 2 nsteps + 1 kernel launches of
 * the steps are captured once and replayed; the capture is redone when the
 * step count, the precision, the time step(s), the Dirichlet increment, the
 * presence of external forces or the launch shape changes.  Same results,
 * bit for bit, as akx_solve_step; no per-kernel timing.  Replaces the
 * driver loop `for (s...) model.solveStep();` (examples/explicit/
 * explicit_dynamic.cc:88-99) for loops without host work between steps. */
int akx_solve_step_graph(akx_ctx* ctx, int nsteps, int precision);

/* One time step per node (nb_nodes values, caller numbering; NULL restores
 * the single akx_set_time_step value): independent bodies meshed into one
 * context (no element joins two bodies) then advance together, each with the
 * time step its own SolidMechanicsModel::setTimeStep
 * (solid_mechanics_model.cc:438) would have received.  The energies that
 * scale with the time step (external work) use the single value. */
int akx_set_node_time_steps(akx_ctx* ctx, const double* dt_per_node);

/* SolidMechanicsModel::getStableTimeStep (solid_mechanics_model.cc:482-533)
 * restricted to the elements of each material: dt[m] for every material m
 * (one value per akx_material_add), the stable step of body m when each body
 * has its own material. */
int akx_stable_time_step_per_material(akx_ctx* ctx, double* dt);

/* SolidMechanicsModel::assembleInternalForces (solid_mechanics_model.cc:375-414)
 * on the current displacement; the result is the AKX_INTERNAL_FORCE field. */
int akx_assemble_internal_forces(akx_ctx* ctx, int precision);

/* SolidMechanicsModel::getStableTimeStep (solid_mechanics_model.cc:482-533). */
int akx_stable_time_step(akx_ctx* ctx, double* dt);

/* SolidMechanicsModel::getEnergy (solid_mechanics_model.cc:665-683). */
int akx_energy(akx_ctx* ctx, int kind, double* value);

/* Blocks `axis` of the given nodes and adds `increment` to their displacement
 * before every subsequent step: the per-step
 * applyBC(BC::Dirichlet::IncrementValue(increment, axis), group) of a driver
 * (boundary_condition_functor.hh:100-125, boundary_condition_tmpl.hh:62-88). */
int akx_set_dirichlet_increment(akx_ctx* ctx, int64_t nb, const int64_t* nodes,
                                int axis, double increment);

/* Wait for all work queued on the context's stream. */
int akx_synchronize(akx_ctx* ctx);

/* Per-kernel timing with HIP events recorded on the context's stream around
 * every element and node kernel of akx_solve_step (off by default).
 * akx_last_step_timing returns the averages over the last akx_solve_step, ms. */
int akx_enable_timing(akx_ctx* ctx, int on);
int akx_last_step_timing(akx_ctx* ctx, double* element_ms, double* node_ms);

/* Launch shape of the element and node kernels: threads per block (64..256,
 * 0 keeps the current value), XCD-aware block mapping (1/0, -1 keeps),
 * whether B and det(J) w are rebuilt from the coordinates (1) or read from
 * the tables (0) -- the FP64 rebuild repeats the precompute's arithmetic and
 * is bit-identical to the tables --, the layout of the element force slots:
 * element-ordered (0), node-ordered (1), node-ordered written by LDS element
 * patches (2, mixed path) or one partial sum per (patch, node), summed in
 * FP64 and stored in FP64 (3, mixed path) or FP32 (4, mixed path) in patch
 * order, or in FP32 at the node's place in a node-major list (5, mixed
 * path), or no slots at all (6: one fused launch per step over node blocks
 * that recompute the elements around their nodes, on either path; stateless
 * laws of one kind on elements of up to 4 nodes -- elsewhere the FP64 path
 * runs 6 as node-ordered slots and the mixed path as 5), and whether the node
 * kernel copies its block's contiguous slot range to LDS first (node_stage
 * 1/0) (-1 keeps).  On the FP64 path none of these changes a result bit; on
 * the mixed path layouts 0/1, 2, 3 and 4/5 differ in rounding only
 * (patch-relative FP32 node values, partial sums), and so do the tables and
 * the rebuilt geometry (for hexahedron_8 rebuilt in the Walsh basis of the
 * trilinear shape functions, without B).  The mixed fused step (6) always
 * rebuilds B; on triangle_3 with mixed_recompute 1 from one 32-byte record
 * per element evaluation streamed in the blocks' order (coordinate
 * differences, node ids, LDS slots, material; meshes of up to 2^23 nodes
 * and 2^17 materials), with 0 from coordinates gathered through element ids:
 * the same result bit for bit. */
int akx_set_launch(akx_ctx* ctx, int element_block, int node_block, int xcd_remap,
                   int mixed_recompute, int slot_layout, int node_stage);

/* Scalar facts about the context, for diagnostics and byte models:
 *   "nb_nodes", "nb_elements", "csr_entries", "patch_blocks",
 *   "patch_nodes_total" (sum over LDS patches of their distinct nodes),
 *   "patch_max_nodes", "patch_available" (1/0), "patch_packed" (1/0: 16-bit
 *   patch-local node | rank words), "fused_available" (1/0: slot layout 6),
 *   "fused_elements" (element evaluations of one fused step),
 *   "fused_records" (1/0: per-evaluation records of the mixed fused step),
 *   "element_block", "node_block", "xcd_remap", "mixed_recompute", "slot_layout". */
int akx_query(akx_ctx* ctx, const char* key, double* value);

/* Autotuner: times every launch shape (element and node block sizes, XCD
 * mapping, mixed-path table/recompute variant, slot layout) over `nsteps`
 * steps of the given precision on the current mesh and keeps the fastest.
 * The model state (u, v, a, damage, ...) is restored afterwards.  chosen
 * (optional, 6 ints) receives element_block, node_block, xcd_remap,
 * mixed_recompute, slot_layout, node_stage; best_ms (optional) the fastest
 * time per step. */
int akx_autotune(akx_ctx* ctx, int precision, int nsteps, int* chosen, double* best_ms);

/* ---- constitutive surrogate -------------------------------------------------
 * A 7 -> 128 x 4 -> 7 network with the smoothstep swish activation
 * a(z) = z S(z), S = t^2 (3 - 2 t), t = clamp(z / 6 + 1/2, 0, 1) (trained in
 * BF16, evaluated in FP16 or BF16, FP32 accumulation, gfx950 MFMA) that stands in for
 * the neo-Hookean part of a
 * neohookean_damage material (material_neohookean_inline_impl.hh:46-65,
 * 145-158); the Marigo damage update (material_marigo_inline_impl.hh:67-76)
 * stays closed-form because it needs each point's own threshold Yd.
 *   inputs  (7): Green-Lagrange strain E (xx, yy, zz, yz, xz, xy) / eps_max,
 *                damage
 *   outputs (7): (S_NH - S_lin) / sS (same order), (W_NH - W_lin) / sW,
 *                S_lin = lambda tr(E) I + 2 mu E, W_lin = lambda/2 tr(E)^2 + mu E:E,
 *                sS = (lambda + 2 mu) eps_max^2, sW = (lambda + 2 mu) eps_max^3.
 * No counterpart in the reference (it has no learned material). */

/* Bind the surrogate to material mat_id (kind neohookean_damage, 3D) with the
 * strain envelope |E_k| <= eps_max; initialises the network. */
int akx_surrogate_setup(akx_ctx* ctx, int mat_id, double eps_max);
/* Training set on the GPU: `paths` random monotone / cyclic strain paths of
 * `steps` steps (Gaussian increments of standard deviation inc_std, reflected
 * at eps_max), damage along each path from the plain law with a Weibull
 * threshold Yd = yd_scale * Weibull(shape yd_shape); paths * steps samples. */
int akx_surrogate_generate(akx_ctx* ctx, int64_t paths, int steps, double inc_std,
                           double yd_scale, double yd_shape, uint64_t seed);
/* Minibatch Adam on the GPU's own forward and backward kernels (batch a
 * multiple of 256; learning rate warm-up then cosine decay to lr / 100);
 * final_loss (optional): mean squared error of the last batch. */
int akx_surrogate_train(akx_ctx* ctx, int iterations, int batch, double lr, uint64_t seed,
                        double* final_loss);
/* FP32 parameters (W0, b0, ..., W4, b4; row-major W[out][in]); count must be
 * the network's parameter count (akx_query "surrogate_parameters"). */
int akx_surrogate_params(akx_ctx* ctx, float* params, int64_t count, int upload);
/* Gradient of the mean squared error over the batch (x, t: n x 7 scaled
 * inputs / targets, n a multiple of 256) with respect to every parameter, from
 * the training kernels, without an update; loss (optional). */
int akx_surrogate_gradients(akx_ctx* ctx, int64_t n, const float* x, const float* t,
                            float* grad, double* loss);
/* Network outputs for n scaled inputs (n x 7 -> n x 7), inference kernel. */
int akx_surrogate_infer(akx_ctx* ctx, int64_t n, const float* x, float* y);
/* Samples [first, first + count) of the training set (scaled, n x 7 each). */
int akx_surrogate_dataset(akx_ctx* ctx, int64_t first, int64_t count, float* x, float* t);
/* Mixed-precision steps evaluate neohookean_damage points through the
 * network (1) or the plain law (0).  Points outside the envelope fall back to
 * the plain law and are counted (akx_surrogate_fallbacks). */
int akx_surrogate_enable(akx_ctx* ctx, int on);
int akx_surrogate_fallbacks(akx_ctx* ctx, uint64_t* count, int reset);
/* 16-bit format of the inference kernel's weights and activations (FP32
 * accumulation in both): AKX_SURROGATE_FP16 (the default: the activation on
 * packed FP16) or AKX_SURROGATE_BF16 (the training format: the activation in
 * FP32, rounded to BF16).  The same master weights serve both; akx_query
 * "surrogate_format" returns the current one. */
enum akx_surrogate_format { AKX_SURROGATE_FP16 = 0, AKX_SURROGATE_BF16 = 1 };
int akx_surrogate_format(akx_ctx* ctx, int format);

/* ---------------------------------------------------------------------------
 * FEEngine operators on a caller's field (FEEngine, src/fe_engine/fe_engine.hh;
 * FEEngineTemplate<IntegratorGauss, ShapeLagrange> for the element type of the
 * mesh), FP64 in the reference's operation order (bit-identical results).
 * `elements` lists caller element ids (the reference's filter_elements),
 * nullptr / 0 = every element in order; per-element and per-point arrays
 * follow that list (row = listed element * Q + q).  Nodal arrays are
 * (N, nb_dof) in the caller's node numbering.  Synchronous.
 * ------------------------------------------------------------------------- */
/* FEEngine::gradientOnIntegrationPoints (fe_engine.hh:149-153;
 * shape_lagrange_inline_impl.hh:353-379, shape_functions_inline_impl.hh:352-395):
 * u (N, nb_dof) -> nablauq (n_el*Q, nb_dof*d), per point nb_dof x d
 * column-major, [j*nb_dof + i] = du_i/dx_j.  1 <= nb_dof <= 9. */
int akx_fe_gradient_on_integration_points(akx_ctx* ctx, const double* u, int nb_dof,
                                          const uint32_t* elements, int64_t nb_elements,
                                          double* nablauq);
/* FEEngine::interpolateOnIntegrationPoints (fe_engine.hh:157-161;
 * shape_functions_inline_impl.hh:300-335): u (N, nb_dof) -> uq (n_el*Q, nb_dof). */
int akx_fe_interpolate_on_integration_points(akx_ctx* ctx, const double* u, int nb_dof,
                                             const uint32_t* elements, int64_t nb_elements,
                                             double* uq);
/* FEEngine::computeBtD (fe_engine.hh:170-175; shape_lagrange_inline_impl.hh:382-423):
 * Ds (n_el*Q, rows*d), per point rows x d column-major -> BtDs (n_el*Q, rows*n),
 * per point D * B, rows x n column-major ([a*rows + i]).  1 <= rows <= 9. */
int akx_fe_compute_btd(akx_ctx* ctx, const double* Ds, int rows, const uint32_t* elements,
                       int64_t nb_elements, double* BtDs);
/* FEEngine::computeNtb (fe_engine.hh:185-189; shape_lagrange_inline_impl.hh:495-533):
 * bs (n_el*Q, nb_dof) -> Ntbs (n_el*Q, nb_dof*n), [a*nb_dof + i] = b_i N_a. */
int akx_fe_compute_ntb(akx_ctx* ctx, const double* bs, int nb_dof, const uint32_t* elements,
                       int64_t nb_elements, double* Ntbs);
/* FEEngine::integrate (fe_engine.hh:95-100; integrator_gauss_inline_impl.hh:426-483):
 * f (n_el*Q, nb_dof) -> intf (n_el, nb_dof) = sum_q f det(J) w.  nb_dof <= 192. */
int akx_fe_integrate(akx_ctx* ctx, const double* f, int nb_dof, const uint32_t* elements,
                     int64_t nb_elements, double* intf);
/* FEEngine::integrate of a scalar field (fe_engine.hh:103-106;
 * integrator_gauss_inline_impl.hh:525-541): the per-element integrals reduced
 * with Math::reduce (aka_math.cc:225-250: sorted, then summed pairwise). */
int akx_fe_integrate_scalar(akx_ctx* ctx, const double* f, const uint32_t* elements,
                            int64_t nb_elements, double* value);
/* FEEngine::integrateOnIntegrationPoints (fe_engine.hh:110-114;
 * integrator_gauss_inline_impl.hh:545-597): out = f det(J) w per point. */
int akx_fe_integrate_on_integration_points(akx_ctx* ctx, const double* f, int nb_dof,
                                           const uint32_t* elements, int64_t nb_elements,
                                           double* out);
/* FEEngine::assembleFieldLumped, row sum (fe_engine.hh:240-244;
 * fe_engine_template_tmpl_field.hh:126-199, DOFManager element-order assembly,
 * dof_manager.cc:83-137): field (E*Q, nb_dof) for every element ->
 * lumped (N, nb_dof) += sum over elements of integral N_a field.  nb_dof <= 24. */
int akx_fe_assemble_field_lumped(akx_ctx* ctx, const double* field, int nb_dof, double* lumped);

/* ---------------------------------------------------------------------------
 * Material operators on the elements of one material (Material,
 * src/model/solid_mechanics/material.hh), FP64, reference order.  A
 * material's points are its elements in ascending caller id (its element
 * filter), Q points each.  The material's gradient, stress and damage
 * internals live on the device (akx_field_download of AKX_DAMAGE reads the
 * damage).
 * ------------------------------------------------------------------------- */
/* Material::getElementFilter: number of elements (and, when ids is not null,
 * their caller ids, ascending). */
int akx_material_elements(akx_ctx* ctx, int mat_id, int64_t* count, uint32_t* ids);
/* Material::computeStress (material.hh:110-113): stress of the material's
 * points from the given gradients (n_mat*Q, d*d) -> stress (same shape; may be
 * null).  Updates the damage history of those points (MaterialMarigo) and
 * keeps gradient and stress as the material's internals. */
int akx_material_compute_stress(akx_ctx* ctx, int mat_id, const double* grad_u, double* stress);
/* Material::computeAllStresses (material.hh:195; material.cc:254-279): gradient
 * of the current displacement on the material's points, then computeStress.
 * grad_u / stress (n_mat*Q, d*d, optional) receive the result. */
int akx_material_compute_all_stresses(akx_ctx* ctx, int mat_id, double* grad_u, double* stress);
/* Material::assembleInternalForces (material.hh:186; material.cc:178-246,
 * 637-722): -integral B^T sigma (or the finite-strain residual) from the
 * material's internals, added to the context's internal force
 * (AKX_INTERNAL_FORCE) node by node in element order. */
int akx_material_assemble_internal_forces(akx_ctx* ctx, int mat_id);
/* Material::getPotentialEnergy (material.cc:743-757) of one material, on the
 * current displacement. */
int akx_material_potential_energy(akx_ctx* ctx, int mat_id, double* value);

/* ---------------------------------------------------------------------------
 * Allocation-free variant.  No counterpart in the reference: for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).
 * ------------------------------------------------------------------------- */
/* akx_energy for AKX_ENERGY_KINETIC (velocity (N, dim), mass one value per
 * node) or AKX_ENERGY_EXTERNAL_WORK (velocity, external_force, internal_force
 * (N, dim), blocked (N, dim) u8, time_step), on caller device arrays; the
 * energy is written to the device address `energy`, on `hip_stream` (null:
 * the default stream), asynchronously.  Allocates nothing: its reduction
 * scratch is a fixed device array of 4096 + 1 doubles per GPU, held until the
 * process ends; any nb_nodes fits it.  Calls that overlap on one GPU race on
 * that array.  Bit-identical to akx_energy on the same arrays (same kernels,
 * same grid); AKX_ENERGY_POTENTIAL is not offered (AKX_ERR_UNSUPPORTED). */
int akx_energy_noalloc(int kind, int dim, int64_t nb_nodes, const double* velocity,
                       const double* mass, const double* external_force,
                       const double* internal_force, const uint8_t* blocked, double time_step,
                       double* energy, void* hip_stream);

#ifdef __cplusplus
}
#endif

#endif /* AKX_H */
