/*
 * topgpu.h -- C ABI of the MI355X (gfx950) TOP_GPU finite-element hot path.
 *
 * Two layers:
 *
 *  1. The six MEX entry points of the reference, one C function each, with the
 *     mexFunction calling convention (nlhs/plhs, nrhs/prhs), the same argument order,
 *     in-place output contract, class/shape checks and error identifiers.  A MATLAB
 *     MEX shim only has to wrap each mxGPUArray as a tg_array and forward the call.
 *
 *  2. A device-resident model/solver API used by the host mirror of TOP_GPU.m
 *     (top_gpu_on_matlab_amd/top_gpu.py): voxel discretisation + multigrid hierarchy
 *     (bit-exact index maps), boundary conditions, Galerkin assembly, matrix-free K*u,
 *     V-cycle, PCG, compliance/sensitivities, PDE filter and the OC update.
 *
 * Conventions (all entry points):
 *   - Arrays are column-major; index arrays are int32 and 1-based, 0 meaning "none".
 *   - Device pointers are HIP device pointers; `stream` is a hipStream_t (NULL = default).
 *   - Every function returns TG_OK (0) or a TG_ERR_* code; tg_last_error_id() then holds an
 *     identifier in the reference's "<Entry>:<Kind>" form and tg_last_error_message() the text.
 *     Error state is thread-local.
 *   - Results are deterministic: no floating-point atomics; every sum has a fixed order.
 */
#ifndef TOPGPU_H
#define TOPGPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------------------------------
 * Status, errors, version
 * ------------------------------------------------------------------------------------- */
enum {
    TG_OK = 0,
    TG_ERR_ARGS = 1,   /* wrong number of inputs/outputs, null pointer, bad option */
    TG_ERR_TYPE = 2,   /* class mismatch (single/double/int32/uint32) */
    TG_ERR_SHAPE = 3,  /* dimension mismatch */
    TG_ERR_HIP = 4,    /* HIP runtime / rocBLAS failure */
    TG_ERR_STATE = 5,  /* call out of order (e.g. solve before assemble) */
    TG_ERR_NUMERIC = 6 /* coarsest matrix not positive definite, non-finite values */
};

const char *tg_last_error_id(void);
const char *tg_last_error_message(void);
const char *tg_version(void);

/* ---------------------------------------------------------------------------------------
 * Array descriptor: stand-in for an mxGPUArray (device pointer, class, column-major dims).
 * ------------------------------------------------------------------------------------- */
enum { TG_CLASS_NONE = 0, TG_SINGLE = 1, TG_DOUBLE = 2, TG_INT32 = 3, TG_UINT32 = 4 };

typedef struct tg_array {
    void *data;       /* device pointer (may be NULL when every dim is 0) */
    int32_t cls;      /* TG_SINGLE / TG_DOUBLE / TG_INT32 / TG_UINT32 */
    int32_t ndim;     /* 1..4; trailing singleton dims may be omitted */
    int64_t dims[4];
} tg_array;

/* ---------------------------------------------------------------------------------------
 * Layer 1: the six MEX entry points.  `stream` orders the work; entry points whose
 * reference version ended with cudaDeviceSynchronize synchronise `stream` before returning.
 * ------------------------------------------------------------------------------------- */

/* Gathering_inplace(gUMat[Ne x 24] T, gU[Nn x 3] T, gE[Ne x 8] int32|uint32):
 * gUMat(e, 3n+c) = gU(gE(e,n), c).  nrhs == 3, nlhs == 0.  T in {single, double}.
 * Replaces Gathering_inplace.cu:75-171 (kernel :22-54).  Node ids outside [1, Nn] write 0
 * (the reference reads out of bounds there). */
int Gathering_inplace(int nlhs, tg_array *plhs, int nrhs, const tg_array *prhs, void *stream);

/* Scattering_inplace(Y[Nn x 3] T, eNodMat[Ne x 8] int32, uMat[Ne x 24] T):
 * Y(eNodMat(e,j), c) += uMat(e, c+3j); node ids outside [1, Nn] are skipped.  nrhs == 3.
 * Deterministic node-owned summation (elements in ascending order) instead of atomics.
 * uMat must have Y's class (the reference reinterprets it silently).  Synchronises.
 * Replaces Scattering_inplace.cu:81-146 (kernel :45-78). */
int Scattering_inplace(int nlhs, tg_array *plhs, int nrhs, const tg_array *prhs, void *stream);

/* scatter_accum3_inplace(keys[N] int32, Tmp[N x 3] T, xFiner[M x 3] T):
 * xFiner(keys(i), c) += Tmp(i, c); keys outside [1, M] skipped.  nrhs == 3, nlhs == 0.
 * Deterministic (rows added in ascending i per key).  keys must hold exactly N entries.
 * Replaces scatter_accum3_inplace.cu:59-145 (kernel :32-56). */
int scatter_accum3_inplace(int nlhs, tg_array *plhs, int nrhs, const tg_array *prhs, void *stream);

/* Scattering_inplace_noalloc: Scattering_inplace (same arguments, checks, error identifiers and
 * results) that allocates nothing.  No counterpart in the reference: it is for callers that cannot
 * allocate device memory (code under HIP graph capture, serving loops, harnesses that forbid hipMalloc).
 * The incidence is sorted on every call in a fixed 128 MiB device array of the library, so inputs are
 * limited to what fits there (16 x 8 Ne bytes of keys and ids plus the sort's storage: about 1.0M
 * elements); larger ones fail with Scattering_inplace:shape before any device work.  All calls on one
 * GPU share that array: calls that overlap on one GPU race and must be serialised by the caller.
 * Synchronises. */
int Scattering_inplace_noalloc(int nlhs, tg_array *plhs, int nrhs, const tg_array *prhs, void *stream);

/* assembleKs_level2_inplace(d_Ks[24x24xNe], d_upMap[Ne x nCh] int32, d_eleModulus, d_K_0[576],
 *                           d_Pch[24x24xnCh], d_mapUnique int32, d_uniqueFree[576 x nU] | empty,
 *                           d_uniqueFixed[576 x nU] | empty):
 * Ks_e = sum_s Pch_s' * Kch_s * Pch_s.  nrhs == 8.  plhs[0], when nlhs > 0, receives d_Ks.
 * Synchronises.  Replaces assembleKs_level2_inplace.cu:165-284 (kernel :58-160). */
int assembleKs_level2_inplace(int nlhs, tg_array *plhs, int nrhs, const tg_array *prhs, void *stream);

/* assembleKs_higherLevel_inplace(d_Ks, d_upMap, d_KsPrev[24x24xNf], d_Psub[24x24xnSub]):
 * Ks_e = sum_s Psub_s' * KsPrev(:,:,upMap(e,s)) * Psub_s.  nrhs == 4.  Synchronises.
 * Replaces assembleKs_higherLevel_inplace.cu:134-217 (kernel :59-129). */
int assembleKs_higherLevel_inplace(int nlhs, tg_array *plhs, int nrhs, const tg_array *prhs, void *stream);

/* assembleKs_level2_superEle_inplace(d_Ks, d_upMap, d_Ks0[576 x Nf], d_Psub, d_mapUnique int32,
 *                                    d_eleModulus[64 x Nf], d_uniqueFree[576x64xnU | 576*64 x nU] | empty,
 *                                    d_uniqueFixed[576 x nU] | empty).  nrhs == 8.  Synchronises.
 * Replaces assembleKs_level2_superEle_inplace.cu:178-325 (kernel :70-173). */
int assembleKs_level2_superEle_inplace(int nlhs, tg_array *plhs, int nrhs, const tg_array *prhs, void *stream);

/* ---------------------------------------------------------------------------------------
 * Layer 2: model (host) -- CreateVoxelFEAmodel / FEA_VoxelBasedDiscretization /
 * Solving_BuildingMeshHierarchy / FEA_ApplyBoundaryCondition / Solving_SetupKeWithFixedDOFs.
 * ------------------------------------------------------------------------------------- */
typedef struct tg_params {
    double youngs_modulus;          /* modulus_ (1.0),            TOP_GPU.m:388 */
    double poisson_ratio;           /* poissonRatio_ (0.3),       TOP_GPU.m:389 */
    double cell_size;               /* cellSize_ (1/nelx or 1),   TOP_GPU.m:392 */
    double weight_factor_jacobi;    /* weightFactorJacobi_ (0.6), TOP_GPU.m:397 */
    int32_t coarsest_resolution_control; /* 20000,                TOP_GPU.m:42 */
    int32_t non_dyadic;             /* nonDyadic_ (1),            TOP_GPU.m:399 */
    int32_t vcycle_type;            /* 0 'Standard', 1 'Adapted', TOP_GPU.m:398 */
    int32_t super_element;          /* Super_element_ (0/1), 4x4x4 sub-voxels */
    int32_t precision;              /* TG_PREC_FP64 or TG_PREC_MIXED */
} tg_params;

enum { TG_PREC_FP64 = 0, TG_PREC_MIXED = 1 };

void tg_params_default(tg_params *p);

typedef struct tg_model tg_model;

/* Discretise a voxel mask (uint8 [nely x nelx x nelz], column-major, nonzero = solid) and build
 * the multigrid hierarchy.  TOP_GPU.m:2019-2119 and 1276-1434. */
int tg_model_create(const uint8_t *voxels, int32_t nely, int32_t nelx, int32_t nelz,
                    const tg_params *params, tg_model **out);
void tg_model_destroy(tg_model *m);

/* Built-in cuboid load cases: 1 = cantilever, 2 = MBB (TOP_GPU.m:505-534). */
int tg_model_builtin_case(tg_model *m, int32_t opt_case);
/* External conditions in solid-node numbering: fixing [nFix x 4] (node, fx, fy, fz flags),
 * loads [nLoad x 4] (node, Fx, Fy, Fz), both column-major double. */
int tg_model_set_conditions(tg_model *m, const double *fixing, int64_t n_fix,
                            const double *loads, int64_t n_load);
/* One more load case, [nLoad x 4] as above, after tg_model_set_conditions (a .TopVoxel file's further
 * load blocks, loadingCond_{2..}, TOP_GPU.m:453-461): F_ gets a column per case, each case's loaded
 * nodes leave the fixations as the first case's do (TOP_GPU.m:1511-1524), and the optimisation solves
 * every case and weights them equally (objWeightingList_, TOP_GPU.m:155, 186, 461, 599). */
int tg_model_add_loads(tg_model *m, const double *loads, int64_t n_load);
/* FEA_ApplyBoundaryCondition + Solving_SetupKeWithFixedDOFs (TOP_GPU.m:1505-1727). */
int tg_model_finalize(tg_model *m);

typedef struct tg_level_info {
    int32_t resX, resY, resZ, spanWidth;
    int64_t numElements, numNodes, numDOFs, intermediateNumNodes;
} tg_level_info;

int32_t tg_model_num_levels(const tg_model *m);
int tg_model_level_info(const tg_model *m, int32_t level, tg_level_info *info); /* level is 1-based */

/* Index maps of meshHierarchy_(level), exactly as the reference stores them (int32). */
enum {
    TG_MAP_ENODMAT = 1,            /* [Ne x 8] */
    TG_MAP_ELEMAPBACK = 2,         /* [Ne] */
    TG_MAP_ELEMAPFORWARD = 3,      /* [resX*resY*resZ] */
    TG_MAP_NODMAPBACK = 4,         /* [Nn] (finest-grid node numbers on levels >= 2) */
    TG_MAP_NODMAPFORWARD = 5,      /* [(resX+1)(resY+1)(resZ+1)] */
    TG_MAP_NODESONBOUNDARY = 6,    /* [nB] */
    TG_MAP_ELEMENTSONBOUNDARY = 7, /* level 1 */
    TG_MAP_ELEMENTUPWARDMAP = 8,   /* [Ne x span^3], levels >= 2 */
    TG_MAP_TRANSFERMAT = 9,        /* [(span+1)^3 x Ne], levels >= 2 */
    TG_MAP_SOLIDNODEMAP = 10,      /* solidNodeMapCoarser2Finer [Nn(level-1)], levels >= 2 */
    TG_MAP_MAPUNIQUEKES = 11,      /* mapUniqueKes_ [Ne(1)] (level 1) */
    TG_MAP_ELEWITHFIXEDDOFS = 12   /* d_eleWithFixedDOFs_ [nU] (level 1) */
};
int64_t tg_model_map_size(const tg_model *m, int32_t level, int32_t which);
int tg_model_get_map(const tg_model *m, int32_t level, int32_t which, int32_t *out);

/* Float tables: */
enum {
    TG_TAB_TRANSFERCOEFF = 1,      /* transferMatCoeffi [intermediateNumNodes], levels >= 2 */
    TG_TAB_KE = 2,                 /* Ke [24 x 24] */
    TG_TAB_LOADS = 3,              /* F_ [3Nn x nLoadCases] (interleaved DOFs) */
    TG_TAB_UNIQUEKESFREE = 4,      /* [576 x nU] or [576 x 64 x nU] (super-element) */
    TG_TAB_UNIQUEKESFIXED = 5,     /* [576 x nU] */
    TG_TAB_IMAT = 6,               /* I_mat [576 x 64] (super-element) */
    TG_TAB_FIXEDDOFS = 7           /* fixed-DOF flags [3Nn] as 0/1 */
};
int64_t tg_model_table_size(const tg_model *m, int32_t level, int32_t which);
int tg_model_get_table(const tg_model *m, int32_t level, int32_t which, double *out);

/* ---------------------------------------------------------------------------------------
 * Layer 2: device solver.  Vectors at this boundary are DOF-interleaved [3*Nn] FP64 in the
 * reference's solid-node numbering; internal layout is the level's full structured grid.
 * ------------------------------------------------------------------------------------- */
typedef struct tg_solver tg_solver;

int tg_solver_create(const tg_model *m, void *stream, tg_solver **out);
void tg_solver_destroy(tg_solver *s);

/* meshHierarchy_(1).d_eleModulus: device FP64 [Ne] (standard) or [64 x Ne] (super-element). */
int tg_solver_set_modulus(tg_solver *s, const double *d_E);
/* Solving_AssembleFEAstencil (TOP_GPU.m:1033-1267): Galerkin Ks on every coarse level,
 * Jacobi diagonals, coarsest assembly + factorisation, all on the device. */
int tg_solver_assemble(tg_solver *s);
/* Solving_KbyU_MatrixFree (TOP_GPU.m:669-765) at `level` (1-based). */
int tg_solver_kbyu(tg_solver *s, int32_t level, const double *d_u, double *d_y);
/* Solving_Vcycle (TOP_GPU.m:766-816). */
int tg_solver_vcycle(tg_solver *s, const double *d_r, double *d_z);
/* Solving_RestrictResidual / Solving_InterpolationDeviation (TOP_GPU.m:817-878);
 * coarse_level is the 1-based index of the coarser level (2..numLevels). */
int tg_solver_restrict(tg_solver *s, int32_t coarse_level, const double *d_rf, double *d_rc);
int tg_solver_interpolate(tg_solver *s, int32_t coarse_level, const double *d_xc, double *d_xf);
/* Coarsest direct solve x = K_c^{-1} r. */
int tg_solver_coarse_solve(tg_solver *s, const double *d_r, double *d_x);
/* Solving_PCG(@Solving_KbyU_MatrixFree, @Solving_Vcycle, b, tol, maxIT) (TOP_GPU.m:602-667).
 * d_y holds the warm start on entry when warm != 0 (zeros otherwise) and the solution on exit. */
int tg_solver_pcg(tg_solver *s, const double *d_b, double *d_y, double tol, int32_t max_it,
                  int32_t warm, int32_t *its, double *resnorm);
/* meshHierarchy_(level).diagK [3Nn] and d_Ks [24 x 24 x Ne] (levels >= 2), copied to device. */
int tg_solver_get_diag(tg_solver *s, int32_t level, double *d_out);
int tg_solver_get_ks(tg_solver *s, int32_t level, double *d_out);
/* Wall time of the last tg_solver_pcg call (ms, HIP events on the solver stream). */
int tg_solver_last_timing(const tg_solver *s, double *ms_total);
/* Average duration (ms) of one matrix-free K*u kernel launch at `level`, timed with HIP events
 * around `reps` back-to-back launches; precision 0 = FP64 operator, 1 = FP32 V-cycle operator. */
int tg_solver_time_kbyu(tg_solver *s, int32_t level, int32_t precision, int32_t reps, double *ms_per_launch);
/* Launch tuning of the finest-level standard-voxel K*u (no reference counterpart: the reference's
 * K*u is a MATLAB pagemtimes, TOP_GPU.m:1729-1741).  Keys:
 *   TG_TUNE_KFINE_KERNEL  0 = element-centric LDS-tiled z-march (default), 1 = node-centric gather
 *   TG_TUNE_KFINE_ZCHUNK  node planes per workgroup of the tiled kernel (1..4096, default 12)
 *   TG_TUNE_TIME_KFINE    1 = bracket every tiled finest K*u launch with HIP events, 2 = only the
 *                         plain products (the PCG operator), 0 = stop; 1 and 2 clear the record
 *   TG_TUNE_KELEM_PARTS   threads per element of the FP64 per-element-matrix K*u (coarse levels and
 *                         the super-element finest level): 1 (default) = the packed upper triangles of
 *                         the symmetric element matrices, 300 of 576 entries; 4, 6, 8, 12 or 24 = the
 *                         full matrices, split over that many threads; the FP32 V-cycle's product is
 *                         always the packed one
 *   TG_TUNE_COARSE_GRAPH  1 = replay the coarsest solve's step-by-step launch sequence as a captured
 *                         HIP graph, 0 = launch it kernel by kernel (measured 1% faster per PCG iteration)
 *   TG_TUNE_COARSE_PERSISTENT  1 = run the coarsest solve as one persistent launch whose workgroups pass
 *                         the recurrence vectors to each other directly (default; planes of at most 1024
 *                         DOFs, no graph requested), 0 = step by step
 *   TG_TUNE_COARSE_POLLS  polls of an input (a device-coherent load each, a memory round trip apart)
 *                         before a stage of the persistent coarsest solve gives up (default 131072,
 *                         ~0.1 s); one give-up ends every other wait of the launch and makes the solve
 *                         fail with TG_ERR_HIP.  Small values force that path in tests.
 *   TG_TUNE_COARSE_STREAMING  1 = the persistent coarsest solve reads its factors with non-temporal
 *                         (streaming) loads (default; 174 against 178-185 us at 128 x 256 x 128),
 *                         0 = ordinary loads
 *   TG_TUNE_KFINE_MFMA    element products of the mixed V-cycle's finest K*u: 0 = Walsh form on the
 *                         VALU (default, the fastest measured), 1 = dense 24 x 24 on MFMA f32 32x32x2,
 *                         2 = dense on MFMA f32 32x32x16 bf16 (BF16 operands, FP32 accumulation),
 *                         3 = as 2 with each operand split into two BF16 parts (three products) */
enum { TG_TUNE_KFINE_KERNEL = 1, TG_TUNE_KFINE_ZCHUNK = 2, TG_TUNE_TIME_KFINE = 3, TG_TUNE_KELEM_PARTS = 4,
       TG_TUNE_COARSE_GRAPH = 5, TG_TUNE_KFINE_MFMA = 6, TG_TUNE_COARSE_PERSISTENT = 7,
       TG_TUNE_COARSE_STREAMING = 8, TG_TUNE_COARSE_POLLS = 9 };
int tg_solver_set_tuning(tg_solver *s, int32_t key, int32_t value);
/* Damped-Jacobi weight of the V-cycle smoother (the reference's global weightFactorJacobi_,
 * TOP_GPU.m:397, 784, 809-811; default 0.6).  0 < omega < 2.  The smoother only reduces every error
 * mode when omega * lambda_max(D^-1 K) < 2: on ragged non-cuboid domains lambda_max grows past
 * 2 / 0.6 and PCG with the default weight stalls. */
int tg_solver_set_jacobi_weight(tg_solver *s, double omega);
/* Launch count and summed duration (ms) of the timed finest K*u launches since timing was enabled,
 * for precision 0 (FP64) / 1 (FP32) and epilogue 0 (K*u), 1 (r - K*u), 2 (damped Jacobi),
 * 3 (r - K*x with the pre-smoothing x = omega r ./ diag fused in). */
int tg_solver_kfine_timing(tg_solver *s, int32_t precision, int32_t epilogue, int64_t *launches, double *ms_total);

/* ---------------------------------------------------------------------------------------
 * Layer 2: objective, sensitivities, filters and the OC update (device arrays in the
 * reference's element order; `stream` orders the work).
 * ------------------------------------------------------------------------------------- */

/* TopOpti_ComputeUnitCompliance (TOP_GPU.m:580-600): ce_e = u_e' Ke u_e with the unmodified Ke;
 * super-element mode: TopOpti_ComputeUnitCompliance_Super (TOP_GPU.m:2622-2645), u_e' K_mat_e u_e.
 * d_U: [3Nn] interleaved; d_ce: [Ne].  Synchronises. */
int tg_solver_unit_compliance(tg_solver *s, const double *d_U, double *d_ce);
/* pre_compute_der (TOP_GPU.m:2646-2671): d_dc(q, e) = u_e' I_q u_e, [64 x Ne] (super-element mode). */
int tg_solver_subvoxel_energy(tg_solver *s, const double *d_U, double *d_dc);

/* SIMP modulus E = Emin + x.^p (E0 - Emin)  (TOP_GPU.m:168, 172). */
int tg_simp_modulus(const double *d_x, int64_t n, double Emin, double E0, double penal, double *d_E, void *stream);
/* dc = -(p (E0 - Emin) x.^(p-1)) .* (ce / cSolid), ce indexed by i / per  (TOP_GPU.m:169-198, 2641-2643). */
int tg_sensitivity(const double *d_x, const double *d_ce, int64_t n, int32_t per, double Emin, double E0,
                   double penal, double cSolid, double *d_dc, void *stream);

enum { TG_EW_MUL = 0, TG_EW_DIV_FLOOR = 1, TG_EW_SET_WHERE = 2, TG_EW_AXPY = 3 };
/* out = a .* b (TG_EW_MUL) or a ./ max(scalar, b) (TG_EW_DIV_FLOOR)  (ft = 1 filtering, TOP_GPU.m:214);
 * out = (b ~= 0) ? scalar : a (TG_EW_SET_WHERE: xPhys(passiveElements) = 1, TOP_GPU.m:270, 273);
 * out = a + scalar .* b (TG_EW_AXPY, out may be a: the load cases' weighted unit compliances,
 * ceList * objWeightingList_, TOP_GPU.m:599, 2644, 2671). */
int tg_elementwise(int32_t op, const double *d_a, const double *d_b, double scalar, int64_t n, double *d_out,
                   void *stream);

enum { TG_RED_SUM = 0, TG_RED_DOT = 1, TG_RED_MAXABSDIFF = 2, TG_RED_SHARPNESS = 3 };
/* Deterministic reductions: sum(a), a.b, max|a-b|, sum(a.*(1-a)); result to host *out.  Synchronises. */
int tg_reduce(int32_t op, const double *d_a, const double *d_b, int64_t n, double *out, void *stream);

/* Optimality-criteria update with bisection on the Lagrange multiplier (TOP_GPU.m:233-253):
 * l1 = 0, l2 = 1e9, move; xnew = max(0,max(x-move,min(1,min(x+move,x.*sqrt(-dc./dv/lmid))))),
 * xnew(passive) = 1, gt = fval + mean(xnew - x).  d_passive: NULL (no passive elements) or [n] with a
 * nonzero entry where xnew is held at 1 in every bisection step (TOP_GPU.m:239, 247; in super-element
 * mode the caller marks sub-voxel row 4 of each passive element, as xnew(Num_, passiveElements) does).
 * Writes d_xnew, *change = max|xnew - x|, *bisections = steps run (a multiple of 16; trailing steps are
 * no-ops).  Synchronises. */
int tg_oc_update(const double *d_x, const double *d_dc, const double *d_dv, const double *d_passive, int64_t n,
                 double fval, double move, double *d_xnew, double *change, int32_t *bisections, void *stream);

/* PDE (Helmholtz) density filter on the finest grid (TOP_GPU.m:880-1031). */
typedef struct tg_filter tg_filter;
int tg_filter_create(const tg_model *m, double rMin, void *stream, tg_filter **out);
void tg_filter_destroy(tg_filter *f);
/* tar = A K_F^{-1} A' src: element->node spread, Jacobi-PCG to `tol`, node->element average.
 * The reference calls it with tol = tol_/10 and maxIT = maxIT_ (TOP_GPU.m:1001).  Synchronises. */
int tg_filter_apply(tg_filter *f, const double *d_src, double *d_tar, double tol, int32_t max_it, int32_t *its);

/* Compliance and its design gradient, standard voxel mode (the sensitivity of TOP_GPU.m:158-174,
 * 189-209 chained through the PDE filter, :214-218, without the 1/cSolid normalisation):
 *   xPhys = filter(x) (f may be NULL: xPhys = x),  E = Emin + xPhys^p (E0 - Emin),
 *   K(E) U = F (load case 1) by the MG-PCG (warm start from d_U when warm != 0; U returned in d_U,
 *   compressed DOF order),  c = F'U = sum_e E_e u_e'K0 u_e,
 *   dc/dxPhys_e = -p (E0 - Emin) xPhys_e^(p-1) u_e'K0 u_e,  d_dcdx = filter^T(dc/dxPhys) = filter(...).
 * Compliance is self-adjoint (adjoint state -U), so the gradient reuses the forward solve.
 * d_dcdx may be NULL (compliance only).  Synchronises. */
int tg_compliance_grad(tg_solver *s, tg_filter *f, const double *d_x, double Emin, double E0, double penal,
                       double tol, int32_t max_it, double filter_tol, int32_t warm, double *d_U, double *c_out,
                       double *d_dcdx, int32_t *its_out);

/* Super-element density filter: imfilter with the cone max(0, r - |d|) on the num-times refined box,
 * zero boundary (TOP_GPU.m:112-121, 220-227, 262-265); arrays are [num^3 x Ne] element-major. */
typedef struct tg_convfilter tg_convfilter;
int tg_convfilter_create(int32_t nelx, int32_t nely, int32_t nelz, double rMin_subvoxels, int32_t num, void *stream,
                         tg_convfilter **out);
void tg_convfilter_destroy(tg_convfilter *f);
/* xPhys = imfilter(x, h) ./ Hs */
int tg_convfilter_forward(tg_convfilter *f, const double *d_x, double *d_out);
/* dc = imfilter(dc ./ Hs, h) */
int tg_convfilter_adjoint(tg_convfilter *f, const double *d_dc, double *d_out);

#ifdef __cplusplus
}
#endif
#endif /* TOPGPU_H */
