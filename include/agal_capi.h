/*
 * agal_capi.h -- C ABI of the MI355X-native octree-AMR lattice-Boltzmann step.
 *
 * This is the drop-in boundary.  The reference wires its LBM solver into the mesh through the
 * abstract class `Solver<ufloat_t,ufloat_g_t,AP>` (inc/solver.h:62-71) and keeps all device state
 * in `Mesh<...>` (inc/mesh.h:137-765).  Here one opaque context owns the device state of one mesh
 * and one solver on one GPU; every entry point below replaces one method of those classes and is
 * annotated with the reference declaration it stands for.  Plain C types only: pointers, sizes,
 * ints and doubles.  All functions return an `int` status (AGAL_OK == 0) instead of calling
 * exit() as the reference's gpuErrchk does (inc/cppspec.h:146-155).
 *
 * Ordering: every call is stream-ordered on the context's single HIP stream (the reference uses
 * one stream per device, inc/mesh.h:482, src/mesh/mesh_init.cu:236).  Calls on one context are
 * not thread-safe.  Download functions synchronise the stream before returning.
 *
 * Populations are exchanged in physical order: entry q of a cell is the pre-collision
 * population of direction q.  (The reference stores direction p in slot pb(p) between steps,
 * src/solver_lbm/solver_lbm_kernels.cu:251; that is an internal convention, not part of the ABI.)
 */
#ifndef AGAL_CAPI_H
#define AGAL_CAPI_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes --------------------------------------------------------------------------- */
#define AGAL_OK                 0
#define AGAL_ERR_ARG            1   /* invalid argument / unsupported combination               */
#define AGAL_ERR_HIP            2   /* a HIP runtime call failed                                */
#define AGAL_ERR_ALLOC          3   /* device or host allocation failed                         */
#define AGAL_ERR_CAPACITY       4   /* root grid does not fit the block pool (mesh_init.cu:171) */
#define AGAL_ERR_POOL_EXHAUSTED 5   /* refinement cancelled: not enough free block IDs
                                       (src/mesh/mesh_amr.cu:2190-2206).  The mesh is left
                                       consistent (marks reverted); the call still completes. */
#define AGAL_ERR_STATE          6   /* call out of order (e.g. step before set_ic)              */
#define AGAL_ERR_OVERFLOW       7   /* an internal bounded list overflowed                      */
#define AGAL_ERR_NOT_BUILT      8   /* library built without the requested variant            */

/* ---- enumerations (values follow the reference where it has them) ---------------------------- */
#define AGAL_D2Q9   9
#define AGAL_D3Q19 19
#define AGAL_D3Q27 27

#define AGAL_FP32      0    /* populations stored and computed in FP32                         */
#define AGAL_FP64      1    /* populations stored and computed in FP64                         */
/* 16-bit pools (no counterpart in the reference): each population is stored as its deviation
 * from the lattice weight, f - w_q, rounded to nearest-even, and every kernel computes in FP32.
 * Population and macroscopic-field transfers use FP32 (decoded values) for these. */
#define AGAL_FP16      2    /* f - w_q stored as IEEE binary16                                   */
#define AGAL_BF16      3    /* f - w_q stored as bfloat16                                        */

#define AGAL_BGK       0    /* CollisionOperator::BGK (inc/solver_lbm.h:19-24)                  */
/* CollisionOperator::MRT.  The reference's MRT exists only as dead legacy code for D3Q19
 * (src/solver_old/solver_lbm_collide_mrt_d3q19.cu:14-1460); here f* = f - M^T D^-1 S M (f - feq)
 * with integer orthogonal bases: d'Humieres et al. (2002) for D3Q19 (the legacy rows), tensor
 * products of {1, c, 3c^2-2} with the second-order block split into trace and deviator for D2Q9
 * and D3Q27.  Stress moments relax with omega = dx_L/tau_L, the others with s1..s6. */
#define AGAL_MRT       2

#define AGAL_INTERP_LINEAR 0 /* InterpOrder::Linear (inc/solver_lbm.h:25-30)                    */
#define AGAL_INTERP_CUBIC  2 /* InterpOrder::Cubic: Cu_Interpolate_Cubic (solver_lbm_kernels_comm.cu:
                                406-638); the advance then averages in BLOCK mode (solver_lbm_init.cu:94-97) */

/* Boundary-condition presets of the reference's device hook Cu_ImposeBC (inc/custom.h:66-123). */
#define AGAL_BC_NONE    0   /* hook returns f unchanged on every face: plain bounce-back        */
#define AGAL_BC_LDC2D   1   /* moving lid on face id -4 (custom.h:80-84)                         */
#define AGAL_BC_LDC3D   2   /* moving lid on face id -6 (custom.h:100-104)                       */
#define AGAL_BC_FPSC2D  3   /* inflow on -1,-3,-4, anti-bounce-back outflow on -2 (custom.h:87-97)*/
#define AGAL_BC_FPSC3D  4   /* inflow on -1,-3..-6, anti-bounce-back outflow on -2 (:107-117)     */

/* Interpolation / averaging modes (inc/mesh.h:58-62). */
#define AGAL_INTERP_INTERFACE 0
#define AGAL_INTERP_ADDED     1
#define AGAL_AVERAGE_INTERFACE 0
#define AGAL_AVERAGE_BLOCK     1
#define AGAL_AVERAGE_GRID      2

/* Refinement states (inc/mesh.h:34-46) and special neighbour IDs (inc/cppspec.h:124-125). */
#define AGAL_REF_UNREFINED          0
#define AGAL_REF_REFINED            1
#define AGAL_REF_REFINED_WCHILD     2
#define AGAL_REF_REFINED_PERM       3
#define AGAL_REF_MARK_REFINE        4
#define AGAL_REF_MARK_COARSEN       5
#define AGAL_REF_NEW                6
#define AGAL_REF_REMOVE             7
#define AGAL_REF_INACTIVE           8
#define AGAL_REF_UNREFINED_VIO      9
#define AGAL_REF_REFINED_WCHILD_PERM 10
#define AGAL_N_SKIPID  (-999)
#define AGAL_N_SPECID  (-998)

/* Cell masks (inc/mesh.h:82-98). */
#define AGAL_CELLMASK_INTERIOR  0
#define AGAL_CELLMASK_INTERFACE 1
#define AGAL_CELLMASK_GHOST     2
#define AGAL_CELLMASK_SOLID    (-1)
#define AGAL_CELLMASK_BOUNDARY (-6)

/* ---- parameters: the input.txt keys the step needs (inc/parser.h:62-144) ----------------------
 * Field names keep the reference's key names.  Values not listed in input.txt by the reference
 * (velocity set, precision, presets) are compile-time packs there (src/main.cu:48-57) and runtime
 * fields here.  agal_params_default() fills the reference's parser defaults. */
typedef struct agal_params {
    int    velocity_set;          /* AGAL_D2Q9 | AGAL_D3Q19 | AGAL_D3Q27 (VSET, src/Makefile:160)   */
    int    precision;             /* AGAL_FP32 | AGAL_FP64 (REAL_s, src/main.cu:60) | FP16 | BF16  */
    int    collision;             /* AGAL_BGK | AGAL_MRT (S_COLLISION)                            */
    int    interp_order;          /* AGAL_INTERP_LINEAR | AGAL_INTERP_CUBIC (S_INTERP)            */
    int    Nx;                    /* root-grid cells along x (Nx)                                 */
    double L_c, L_fy, L_fz;       /* domain length and aspect ratios                              */
    int    PERIODIC_X, PERIODIC_Y, PERIODIC_Z;
    int    MAX_LEVELS_WALL, MAX_LEVELS_INTERIOR, N_LEVEL_START;
    double v0;                    /* kinematic viscosity in the reference's units (tau = 3 v0 + dx/2) */
    int    S_CRITERION;           /* 0: |vorticity|, 1: Q-like invariant (solver_lbm_criterion.cu:
                                     544-547); 2: |grad u| Frobenius norm (no reference counterpart) */
    double N_REFINE_START, N_REFINE_INC, N_REFINE_MAX;
    int    P_REFINE;              /* coarse steps between solution-based rebuilds                 */
    int    bc_preset;             /* AGAL_BC_* : the Cu_ImposeBC hook                             */
    double bc_velocity;           /* wall / inflow speed used by the hook (0.05 in custom.h)      */
    double ic_rho, ic_u, ic_v, ic_w; /* Cu_ComputeIC hook: uniform state (custom.h:48-60)         */
    double s1, s2, s3, s4, s5, s6;   /* MRT relaxation rates of the non-hydrodynamic moments
                                        (input.txt keys s1..s6, solver_lbm_init.cu:55-61)       */
    double G_NEAR_WALL_DISTANCE;  /* width of the refined band around a geometry on level 0
                                     (parser.h:121; halved per level, mesh_voxelizer.cu:39-41) */
    /* Face-cell links of a geometry (parser.h:73-89): G_BIN_* size the face bins
     * (geometry_bins.h:118-123), G_BIN_MAKE_LINKS 1 computes the cut-link fractions
     * (mesh_advance.cu:106-111); S_BC_TYPE 1 applies the linear interpolated bounce-back on them
     * (solver_lbm_kernels.cu:796-875); S_FORCE_TYPE 0 with S_FORCE_ORDER 1 / 2 is the
     * momentum-exchange force on the linked cells (solver_lbm_kernels_forces.cu:322-540). */
    int    G_BIN_DENSITY, G_BIN_LEVELS, G_BIN_SPEC, G_BIN_MAKE_LINKS;
    int    S_BC_TYPE, S_FORCE_TYPE, S_FORCE_ORDER;
    /* Control volume of the force type 1 (parser.h:132-137; snapped to root-block edges,
     * solver_lbm_init.cu:36-41), as fractions of the domain: x in (Xm, XM), y in (Ym, YM), z in (Zm, ZM). */
    double S_FORCEVOLUME_Xm, S_FORCEVOLUME_XM, S_FORCEVOLUME_Ym, S_FORCEVOLUME_YM, S_FORCEVOLUME_Zm, S_FORCEVOLUME_ZM;
} agal_params;

typedef struct agal_ctx agal_ctx;

/* Sizes of the block pool and its per-level ID sets (inc/mesh.h:430-462). */
typedef struct agal_sizes {
    int  n_dim, n_q, n_q_max, cells_per_block, n_children, max_levels;
    long long n_maxcblocks, n_maxcells;
    int  n_coarsecblocks;
    int  n_ids[8];                /* active blocks per level                                       */
    int  id_max[9];               /* 1 + largest active ID per level; [max_levels] = overall       */
    int  n_gaps;                  /* free block IDs on the gap stack                               */
} agal_sizes;

/* Fill `p` with the reference parser's defaults (inc/parser.h:62-144). */
void agal_params_default(agal_params *p);

/* Mesh<...>::M_Init (src/mesh/mesh_init.cu:10-567) + Solver_LBM::S_Init (solver_lbm_init.cu:4-102).
 * `block_capacity` replaces the reference's pool sizing from free memory (mesh_init.cu:155-169):
 * the pool holds exactly that many cell-blocks.  Builds the root grid, boundary IDs and the gap
 * stack; allocates device memory.  Returns AGAL_ERR_CAPACITY if the root grid does not fit. */
int  agal_create(const agal_params *p, long long block_capacity, int device, agal_ctx **out);
/* Mesh<...>::M_Dest (src/mesh/mesh_dest.cu:11-98). */
int  agal_destroy(agal_ctx *ctx);

int  agal_get_sizes(agal_ctx *ctx, agal_sizes *out);
int  agal_sync(agal_ctx *ctx);
const char *agal_status_string(int status);
/* Last error text recorded by the context (HIP error name, argument that failed). */
const char *agal_last_error(agal_ctx *ctx);

/* ---- the solver interface (inc/solver.h:62-71) ------------------------------------------------ */
/* Solver::S_SetIC (solver_lbm_kernels.cu:79-98). */
int  agal_set_ic(agal_ctx *ctx, int L);
/* Solver::S_Advance (solver_lbm_advance.cu:30-174): one coarse step of level L0 with recursive
 * sub-cycling of the finer levels. */
int  agal_advance(agal_ctx *ctx, int L0);
/* One fused collide + boundary + stream sub-step of level L: S_Collide, S_ImposeBC and S_Stream
 * (solver_lbm_kernels.cu:464-510, 890-920, 635-660) in one HBM pass. */
int  agal_collide_stream(agal_ctx *ctx, int L);
/* Solver::S_Interpolate (solver_lbm_kernels_comm.cu:141-155), var = AGAL_INTERP_*. */
int  agal_interpolate(agal_ctx *ctx, int L, int var);
/* Solver::S_Average (solver_lbm_kernels_comm.cu:351-370), var = AGAL_AVERAGE_*. */
int  agal_average(agal_ctx *ctx, int L, int var);
/* Solver::S_RefreshVariables (solver_lbm_kernels.cu:161-175): rho,u,v,w of level L into the
 * macroscopic arrays. */
int  agal_refresh(agal_ctx *ctx, int L);
/* Solver::S_ComputeRefCriteria (solver_lbm_criterion.cu:594-613): marks blocks of every level. */
int  agal_ref_criteria(agal_ctx *ctx);

/* ---- the mesh interface ---------------------------------------------------------------------- */
/* Mesh::M_RefineAndCoarsenBlocks (src/mesh/mesh_amr.cu:2161-2744).  out_counts (may be NULL)
 * receives {marked for refinement, marked for coarsening after quality control, blocks added,
 * blocks removed}.  Returns AGAL_ERR_POOL_EXHAUSTED when refinement was cancelled. */
int  agal_refine_coarsen(agal_ctx *ctx, int out_counts[4]);
/* Mesh::M_Advance_RefineWithSolution (src/mesh/mesh_advance.cu:184-243). */
int  agal_refine_with_solution(agal_ctx *ctx, int out_counts[4]);
/* Mesh::M_ComputeRefCriteria(.., V_MESH_REF_UNIFORM) (src/mesh/mesh_criterion.cu:102-111). */
int  agal_mark_uniform(agal_ctx *ctx, int L);
/* Sets cblock_ID_ref of the listed blocks to `mark` (the reference's refine-region hook
 * Cu_RefineRegion, inc/custom.h:15-42, marks blocks the same way). */
int  agal_mark_blocks(agal_ctx *ctx, const int *ids, int n, int mark);
/* Mesh::M_FreezeRefinedCells (src/mesh/mesh_amr.cu:2049-2074), var 0 freeze / 1 unfreeze. */
int  agal_freeze_refined(agal_ctx *ctx, int var);
/* Mesh::M_Advance_Step (src/mesh/mesh_advance.cu:155-182) followed, when (i+1)%P_REFINE==0, by
 * M_Advance_RefineWithSolution -- the body of M_AdvanceLoop (mesh_advance.cu:421-463) for
 * coarse steps [i0, i0+n).  Rebuild time is reported separately in out_rebuild_ms (may be NULL). */
int  agal_advance_loop(agal_ctx *ctx, int i0, int n, double *out_step_ms, double *out_rebuild_ms);
/* Mesh::M_Advance_RefineNearWall (src/mesh/mesh_advance.cu:53-121): uniform refinement up to
 * N_LEVEL_START, then per wall level the geometry's voxelization and block marks (when a geometry
 * is set), the rebuild, and the initial condition of the level; then boundary cells and the freeze
 * of the refined blocks.  Returns AGAL_ERR_POOL_EXHAUSTED if a rebuild was cancelled. */
int  agal_refine_near_wall(agal_ctx *ctx);

/* ---- learned refinement predictor (no counterpart in the reference) --------------------------
 * A 12-64-64-1 network (ReLU, sigmoid output) over 12 features per block, FP16 weights with FP32
 * accumulation, trained online with Adam on the plain criterion's labels (the criterion of
 * agal_ref_criteria stays the fallback and the referee).  Features of block b, from the stored
 * macroscopic fields: plog2(x + 2^-30)/16 of the block's max |u|, max |rho-1|, max |grad u|,
 * max |vorticity| (plog2 = piecewise-linear log2, exact at powers of two), level/(levels-1), a
 * signed distance (to the geometry's nearest triangle centroid, else to the nearest domain face,
 * in root blocks, clamped to +-8, /8), the neighbour maxima of the four statistics, the fraction
 * of neighbours with children and the neighbour minimum of the distance.  Parameters, flat FP32:
 * W1[12][64], b1[64], W2[64][64], b2[64], W3[64], b3 (AGAL_MLP_NPARAM values). */
#define AGAL_MLP_NFEAT  12
#define AGAL_MLP_NPARAM 5057

typedef struct agal_learn_opts {
    int    train_steps;      /* coarse steps with online training (2000)                          */
    int    infer_every;      /* coarse steps between feature snapshots / inferences (25)          */
    int    batch;            /* rows per Adam step (4096)                                         */
    int    adam_steps;       /* Adam steps per rebuild while training (8)                         */
    double lr;               /* Adam learning rate (1e-3)                                         */
    double threshold;        /* a block is predicted to need children when p > threshold (0.5)    */
    double pos_weight;       /* loss weight of positive labels (4)                                */
    unsigned long long seed; /* weight initialisation and batch sampling                          */
    int    use_predictions;  /* 0: plain criterion only (training and audit counts still run);
                                1: the network's predictions refine ahead at every inference step;
                                controls without the network -- 2: the plain criterion's refinement
                                marks (no coarsening) at the same inference steps; 3: the plain
                                criterion with its refinement marks grown by one block (same level)
                                at the rebuilds                                                     */
} agal_learn_opts;

typedef struct agal_learn_stats {
    long long flagged, missed;             /* after training: blocks the criterion wants refined at
                                              rebuilds, and those of them still unrefined          */
    long long flagged_train, missed_train; /* the same counts during training (plain method)       */
    long long predicted_refinements;       /* blocks refined because the network predicted them    */
    int    adam_steps_done, rebuilds, early_rebuilds;
    double last_loss;                      /* mean loss of the last training batch                 */
    double step_ms, rebuild_ms, predictor_ms;
} agal_learn_stats;

/* Allocates the predictor (training buffers for up to batch_capacity rows) and initialises the
 * weights He-uniformly from a splitmix64 stream seeded with `seed`; Adam state reset. */
int  agal_predictor_init(agal_ctx *ctx, unsigned long long seed, int batch_capacity);
/* Sets the FP32 master parameters (AGAL_MLP_NPARAM values); Adam state reset. */
int  agal_predictor_set(agal_ctx *ctx, const float *params);
/* which: 0 master parameters, 1 their FP16 copy (as FP32), 2 last gradient, 3 / 4 Adam m / v. */
int  agal_predictor_get(agal_ctx *ctx, int which, float *out);
/* Logits and probabilities of n feature rows (n <= batch_capacity); either output may be NULL. */
int  agal_predictor_forward(agal_ctx *ctx, const float *x, int n, float *z_out, float *p_out);
/* One Adam step (beta 0.9 / 0.999, eps 1e-8) on the weighted binary cross-entropy of n rows,
 * dL/dz = w (sigmoid(z) - y) / n with w = pos_weight for y = 1; the mean loss in *loss. */
int  agal_predictor_train_step(agal_ctx *ctx, const float *x, const float *y, int n, double lr,
                               double pos_weight, double *loss);
/* Feature rows of block IDs [0, n_blocks) from the stored macroscopic fields (zeros for IDs that
 * are inactive or beyond the largest active ID). */
int  agal_predictor_features(agal_ctx *ctx, float *out, long long n_blocks);
/* The plain criterion's label per block ID: 1 wants children, 0 not, -1 inactive.  The stored
 * refinement marks are left unchanged. */
int  agal_criterion_labels(agal_ctx *ctx, int *out, long long n_blocks);
void agal_learn_opts_default(agal_learn_opts *opts);
/* M_AdvanceLoop (src/mesh/mesh_advance.cu:421-463) with the predictor: snapshots and training
 * during the first train_steps coarse steps, then predicted blocks refined every infer_every steps
 * and kept from coarsening at rebuilds, the plain criterion refining what it flags; audit counts
 * and timings in *stats. */
int  agal_advance_loop_learned(agal_ctx *ctx, int i0, int n, const agal_learn_opts *opts,
                               agal_learn_stats *stats);

/* ---- geometry (3D velocity sets) -------------------------------------------------------------
 * A closed triangle surface, 9 floats per triangle (x1 y1 z1 x2 y2 z2 x3 y3 z3), in domain
 * coordinates.  Replaces Geometry::G_Init_Arrays_CoordsList_CPU + G_MakeBinsGPU
 * (src/geometry/geometry_init.cu:79-212, geometry_bins_gpu.cu:586-1040): the library copies the
 * triangles and bins them on a (y, z) grid for the voxelizer's rays. */
int  agal_set_geometry(agal_ctx *ctx, const float *tri, long long n_tri);
/* Mesh::M_Geometry_Voxelize_S1 (src/mesh/mesh_voxelizer.cu:33-136) on level L: a cell is SOLID when
 * the +x ray from its centre crosses the surface an odd number of times; blocks holding one get
 * the SOLID block mask; then the boundary / exterior / near-wall-band refinement marks with
 * int(R/(4 sqrt2 dx_L))+1 propagation passes, R = G_NEAR_WALL_DISTANCE / 2^L. */
int  agal_voxelize(agal_ctx *ctx, int L);
/* Mesh::M_Geometry_Voxelize_S2 (mesh_voxelizer.cu:139-174) on level L: BOUNDARY cell masks (non-
 * solid cells with a solid neighbour) in leaf SOLIDB / SOLIDA blocks. */
int  agal_check_masks(agal_ctx *ctx, int L);
/* Solver_LBM::S_ReportForces (solver_lbm_compute.cu:165-218).  Force type 0: the momentum-exchange
 * sums of Cu_ComputeForcesMEA (solver_lbm_kernels_forces.cu:322-540), which every coarse step forms
 * around the finest level's last sub-step (after the collision and after the stream) on the cut
 * links of boundary cells -- order 1: f c_p; order 2 (S_FORCE_ORDER 2): ((1/2 + q) f + (1/2 - q)
 * f_m) c_p -- per block times dx^3, summed over blocks in FP64 and times 1/dx of the finest level;
 * needs G_BIN_MAKE_LINKS 1.  Force type 1: the control-volume balance of Cu_ComputeForcesCV
 * (:13-246) over the box S_FORCEVOLUME_*, formed at the start of each coarse step (momentum of the
 * leaf cells inside; root-level populations leaving through the box faces after a BGK collision
 * with the root rate) and at its end (after a full fine-to-coarse averaging and the interface
 * interpolations: the momentum again, and the root-level populations that entered), times 1/dx of
 * the root level.  out = (x, y, z) of sum(first part) - sum(second part); the block sums are then
 * reset.  Other force types: AGAL_ERR_STATE.  Drag / lift coefficients are left to the caller. */
int  agal_forces(agal_ctx *ctx, double out[3]);
/* Face-cell links after agal_refine_near_wall with G_BIN_MAKE_LINKS 1 (Mesh::M_Geometry_Voxelize_S3
 * and M_LinkLengthComputation, mesh_voxelizer.cu:177-296): onb_solid[k] = rank of block k among the
 * blocks holding boundary cells (cblock_ID_onb_solid, -1 otherwise) for k < n_blocks; X_b = the
 * first n_values cut-link fractions, [rank*64 + cell + q*64*n_solidb] over the 27 directions
 * (cells_f_X_b, -1 = no cut); info = {n_solidb (ranks rounded up to a multiple of 128), corrections
 * counted by the validation, links present}.  NULL pointers skip a part. */
int  agal_download_links(agal_ctx *ctx, int *onb_solid, long long n_blocks, float *X_b, long long n_values,
                         long long info[3]);
/* The face bins of the link search (Geometry::Bins::G_MakeBinsGPU_MD, geometry_bins_gpu.cu:1043-
 * 1336): bin_start (n_bins + 1 offsets) and faces (per bin, ordered by overlap rank then face);
 * info = {n_bins, entries, bins per unit length, faces kept by the ray-indicator filter}. */
int  agal_download_link_bins(agal_ctx *ctx, int *bin_start, long long n_start, int *faces, long long n_faces,
                             long long info[4]);
/* Host helpers, no context.  Icosphere: the icosahedron subdivided `subdivisions` times (20*4^s
 * triangles), each vertex at radius r (1 + noise u), u uniform in [-1, 1) from a splitmix64
 * stream seeded with `seed`, in vertex order.  ASCII STL: the facets of the file (the format of
 * the reference's geometry inputs).  Both return the triangle count (out may be NULL to query
 * it) or a negative value on error (-1 bad argument / unreadable file, -2 capacity too small,
 * -3 malformed file). */
long long agal_icosphere(double cx, double cy, double cz, double radius, int subdivisions, double noise,
                         unsigned long long seed, float *out, long long capacity);
long long agal_load_stl(const char *path, float *out, long long capacity);

/* ---- data exchange (M_RetrieveFromGPU / M_LoadToGPU, src/mesh/mesh_output.cu:4-82) ---------
 * Arrays are copied for block IDs [0, n_blocks).  Element layouts follow the reference's SoA:
 *   populations: Q arrays of n_blocks*cells_per_block values (dtype = precision), physical order
 *   macros:      4 arrays (rho,u,v,w) of n_blocks*cells_per_block
 *   cell masks:  n_blocks*cells_per_block ints
 *   nbr, nbr_child: n_q_max arrays of n_blocks ints (velocity-ordered, self at p=0)
 *   per-block ints (ref, level, onb, block mask): n_blocks
 *   coords: n_dim arrays of n_blocks floats
 * `dst_stride` / `src_stride` is the distance in elements between consecutive arrays. */
int  agal_download_populations(agal_ctx *ctx, void *dst, long long n_blocks, long long dst_stride);
int  agal_upload_populations(agal_ctx *ctx, const void *src, long long n_blocks, long long src_stride);
/* No counterpart in the reference.  agal_download_populations for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc): the same output, staged through a fixed device array of AGAL_NOALLOC_SCRATCH_BYTES
 * that the library holds on each GPU until the process ends, instead of a buffer allocated per
 * call.  Calls no hipMalloc / hipFree.  Returns AGAL_ERR_ARG (message in agal_last_error) when
 * Q * n_blocks * cells_per_block transfer values exceed the array.  All calls on one GPU share
 * its array: calls that overlap in time on the same GPU race. */
#define AGAL_NOALLOC_SCRATCH_BYTES (64ull << 20)
int  agal_download_populations_noalloc(agal_ctx *ctx, void *dst, long long n_blocks, long long dst_stride);
int  agal_download_macros(agal_ctx *ctx, void *dst, long long n_blocks, long long dst_stride);
int  agal_download_cell_masks(agal_ctx *ctx, int *dst, long long n_blocks);
/* M_LoadToGPU (src/mesh/mesh_output.cu:4-42) for one array: overwrite cell masks / a block table
 * (e.g. a restart or a voxelised geometry); per-level lists are rebuilt afterwards. */
int  agal_upload_cell_masks(agal_ctx *ctx, const int *src, long long n_blocks);
int  agal_upload_block_ints(agal_ctx *ctx, int which, const int *src, long long n_blocks, long long src_stride);
int  agal_download_block_ints(agal_ctx *ctx, int which, int *dst, long long n_blocks, long long dst_stride);
int  agal_download_coords(agal_ctx *ctx, float *dst, long long n_blocks, long long dst_stride);
int  agal_download_id_set(agal_ctx *ctx, int L, int *dst, int n);
int  agal_download_gap_set(agal_ctx *ctx, int *dst, int n);

/* M_LoadToGPU (src/mesh/mesh_output.cu:4-42) for the whole pool at once, as a restart load
 * (mesh_restart.cu:110-220) fills it: every block table for IDs [0, n_blocks), the cell masks,
 * the per-level ID sets (concatenated, each ascending), the gap stack and, when `populations` is
 * not NULL, the populations (physical order, the transfer type of agal_download_populations).
 * Blocks from n_blocks on become inactive.  The per-level lists are rebuilt from ref / level and
 * must reproduce the given ID sets, else AGAL_ERR_ARG and the context keeps the imported tables
 * with the rebuilt lists. */
typedef struct agal_pool {
    long long  n_blocks;
    const int *nbr, *nbr_child;            /* n_q_max rows of n_blocks (velocity order)       */
    const int *ref, *level, *onb, *blk_mask;
    const float *coords;                   /* n_dim rows of n_blocks                           */
    const int *cell_mask;                  /* n_blocks * cells_per_block                       */
    const int *n_ids;                      /* max_levels counts                                */
    const int *id_sets;                    /* sum(n_ids) IDs, level after level                */
    int        n_gaps;
    const int *gap_set;
    const void *populations;               /* n_q rows of n_blocks * cells_per_block, or NULL  */
} agal_pool;
int  agal_import_pool(agal_ctx *ctx, const agal_pool *pool);
/* `which` values for agal_download_block_ints. */
#define AGAL_BLK_NBR        0
#define AGAL_BLK_NBR_CHILD  1
#define AGAL_BLK_REF        2
#define AGAL_BLK_LEVEL      3
#define AGAL_BLK_ONB        4
#define AGAL_BLK_MASK       5

/* ---- options (no counterpart in the reference, which fixes these at compile time) ------------
 * AGAL_OPT_GRAPHS: 1 (default) records each coarse step of agal_advance / agal_advance_loop as a
 * HIP graph and replays it until the mesh changes; 0 launches every kernel directly. */
#define AGAL_OPT_GRAPHS 1
/* AGAL_OPT_MRT_MFMA: 1 runs the MRT moment transform of D3Q19 / D3Q27 FP32-compute steps (FP32, FP16
 * and BF16 pools) on matrix cores: per block, (moments x directions) . (directions x 64 cells)
 * products in v_mfma_f32_16x16x32_bf16 with the non-equilibrium parts and the relaxed moments
 * split into two bf16 terms (16 significant bits of a small correction; FP32 sums in the hardware's
 * order).  0 (default) keeps the vector path, whose sums follow a fixed order shared with the CPU
 * restatement. */
#define AGAL_OPT_MRT_MFMA 2
/* AGAL_OPT_EXC_CAP: entries per stream-guard exception list (default and maximum:
 * min(cells of the pool, 2^22) per level).  A sub-step whose list overflows drops links; the next
 * agal_sync / agal_advance_loop reports AGAL_ERR_OVERFLOW and the context refuses further steps
 * (AGAL_ERR_STATE). */
#define AGAL_OPT_EXC_CAP 3
/* AGAL_OPT_GRAPH_STEPS: coarse steps of agal_advance_loop recorded into one graph at most (1..64,
 * default 4).  Longer graphs amortize the gap between graph launches; each distinct length is a
 * separate recording per mesh epoch. */
#define AGAL_OPT_GRAPH_STEPS 4
/* AGAL_OPT_LINK_MASKS: 1 makes the boundary cells and face-cell links of the next
 * agal_refine_near_wall agree with the cell masks (no counterpart in the reference): every non-solid
 * cell with a SOLID neighbour becomes a BOUNDARY cell in every leaf block (the reference examines
 * SOLIDA / SOLIDB blocks only, not REGULAR ones nor SOLID ones that still hold fluid cells); after the reference's validation a BOUNDARY
 * cell's fraction towards direction p is kept only when the partner cell is SOLID, and a SOLID
 * partner without a cut face gets 1/2.  The links are then exactly the pairs that plain
 * bounce-back reflects, and the momentum exchange of a fluid at rest sums to zero.  0 (default)
 * keeps the reference's rules, under which fluid cells beside a solid cell of a neighbouring block
 * are not linked, a face cut inside a link to a fluid cell stays a link, and the validation reads
 * the partner's mask at its rank. */
#define AGAL_OPT_LINK_MASKS 5
int  agal_set_option(agal_ctx *ctx, int option, long long value);

/* ---- measurement ----------------------------------------------------------------------------- */
/* Average device time (ms) per launch of the fused level kernel since the last reset, measured
 * with HIP events on the context stream; `launches` receives the launch count. */
int  agal_timing_reset(agal_ctx *ctx, int enable);
int  agal_timing_get(agal_ctx *ctx, double *fused_ms_total, long long *fused_launches,
                     double *fused_cells_total);

#ifdef __cplusplus
}
#endif
#endif /* AGAL_CAPI_H */
