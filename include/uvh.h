/*
 * uvh.h -- C ABI of the MI355X uvh assembly library (libuvh.so).
 *
 * Replaces the residual/Jacobian assembly behind Úa's
 *     [UserVar,RunInfo,R,K] = uvhAssembly(UserVar,RunInfo,CtrlVar,MUA,F0,F1,l1,BCs1)
 * (UaMain/uvhAssembly.m:1, body UaMain/uvhMatrixAssembly.m:6-585). A MATLAB MEX
 * gateway (uasource_amd/mex/uvhAssemblyGPU_mex.c) or any other host binds these
 * entry points; plain pointers and sizes only.
 *
 * Conventions
 *  - Connectivity is 0-based int32, element-major (conn[e*nod + a]); the MEX gateway
 *    converts MATLAB's 1-based Nele x nod column-major array once per mesh.
 *  - Nodal arrays are FP64, length nnodes.
 *  - R is 3*nnodes, ordered [x(1..Nn); y(1..Nn); h(1..Nn)] (uvhMatrixAssembly.m:185-186,425-439),
 *    R = Tint - Fext (uvhMatrixAssembly.m:455).
 *  - K is returned as values on a fixed CSR pattern of the 3Nn x 3Nn matrix with blocks
 *    [Kxu Kxv Kxh; Kyu Kyv Kyh; Khu Khv Khh] (uvhMatrixAssembly.m:25-27,478-502). Row
 *    (blk, node) holds, in order, the columns adj(node), adj(node)+Nn, adj(node)+2Nn, where
 *    adj(node) is the sorted list of nodes sharing an element with node (itself included).
 *    The pattern is structurally symmetric, so the same index arrays describe CSC.
 *    Entries that sum to exactly zero are kept (MATLAB's sparse would drop them).
 *  - Status codes mirror the reference's errors: UVH_E_NAN_R is
 *    error("uvhMatrixAssembly:NaNinR") (uvhMatrixAssembly.m:581-584), UVH_E_NAN_K is
 *    error(' NaN in K ') (uvhMatrixAssembly.m:575-579).
 */
#ifndef UVH_H
#define UVH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct uvh_ctx uvh_ctx;

enum uvh_status {
    UVH_OK = 0,
    UVH_E_ARG = 1,          /* bad argument (size, null pointer, unsupported element type) */
    UVH_E_HIP = 2,          /* HIP runtime error */
    UVH_E_NAN_R = 3,        /* NaN in R:  uvhMatrixAssembly:NaNinR */
    UVH_E_NAN_K = 4,        /* NaN on the diagonal of K */
    UVH_E_UNSUPPORTED = 5,  /* option outside this library: caller falls back to uvhMatrixAssembly */
    UVH_E_STATE = 6,        /* call out of order (e.g. assemble before set_mesh) */
    UVH_E_NOMEM = 7,        /* device allocation failed */
    UVH_E_NOCONV = 8        /* linear solve: zero pivot, breakdown or tolerance not reached
                               (solveKApe's gmres flag != 0) */
};

enum uvh_flags {
    UVH_RONLY = 1u,         /* CtrlVar.uvhMatrixAssembly.Ronly: R only, K untouched */
    UVH_ZEROFIELDS = 2u,    /* CtrlVar.uvhMatrixAssembly.ZeroFields (uvhMatrixAssembly.m:48-101) */
    UVH_FP32 = 4u,          /* precision form: FP32 element arithmetic, FP64 accumulation */
    UVH_NO_NAN_CHECK = 8u   /* skip the NaN checks of uvhMatrixAssembly.m:575-584 */
};

enum uvh_supg_tau { UVH_TAU_S = 0, UVH_TAU_T = 1, UVH_TAU_1 = 2, UVH_TAU_2 = 3, UVH_TAU_0 = 4 };

/* CtrlVar.SlidingLaw (UaMain/BasalDrag.m:182-257): "Weertman"/"W"; "Budd"/"W-N0" (needs q);
 * "Coulomb"/"C" (needs muk); "Tsai"/"minCW-N0" (needs muk: the smaller in magnitude of the
 * Weertman and Coulomb drag at each point); "Cornford"/"rpCW-N0" (rpCWN0.m, needs muk);
 * "Umbi"/"rCW-N0" (rCWN0.m, needs muk); "Joughin"/"rCW-v0" (rCWV0.m, needs V0). The laws
 * with muk use the effective pressure N0 (BasalDrag.m:500-521): rho g max(h - hf, 0) +
 * CtrlVar.Nzero. Other codes return UVH_E_UNSUPPORTED. */
enum uvh_sliding_law {
    UVH_SLIDING_WEERTMAN = 0,
    UVH_SLIDING_BUDD = 1,
    UVH_SLIDING_COULOMB = 2,
    UVH_SLIDING_TSAI = 3,
    UVH_SLIDING_CORNFORD = 4,
    UVH_SLIDING_UMBI = 5,
    UVH_SLIDING_JOUGHIN = 6
};

/* Scalars the assembly reads from CtrlVar and F (defaults: UaMain/Ua2D_DefaultParameters.m). */
typedef struct {
    double dt;                  /* CtrlVar.dt */
    double theta;               /* CtrlVar.theta */
    double supg_beta0;          /* CtrlVar.SUPG.beta0 */
    double supg_tau_multiplier; /* CtrlVar.uvh.SUPG.tauMultiplier */
    double speed_zero;          /* CtrlVar.SpeedZero */
    double eps_zero;            /* CtrlVar.EpsZero */
    double eta_zero;            /* CtrlVar.etaZero */
    double czero;               /* CtrlVar.Czero */
    double kH;                  /* CtrlVar.kH */
    double Hh0;                 /* CtrlVar.Hh0 */
    double g;                   /* F.g */
    double rhow;                /* F.rhow */
    double alpha;               /* F.alpha */
    double thick_min;           /* CtrlVar.ThickMin */
    int32_t supg_tau;           /* CtrlVar.uvh.SUPG.tau as enum uvh_supg_tau */
    int32_t reset_thickness;    /* ResetThicknessToMinThickness && ResetThicknessInNonLinLoop */
    int32_t enforce_b_above_B;  /* CtrlVar.Enforce_bAboveB */
    int32_t sliding_law;        /* CtrlVar.SlidingLaw as enum uvh_sliding_law */
    double nzero;               /* CtrlVar.Nzero (1e-20), lower bound of the effective pressure */
} uvh_ctrl;

/* Nodal fields of F1 (current) and F0 (previous time step); each nnodes long.
 * dadh may be NULL (zero: MassBalanceGeometryFeedback < 2). q (F1.q), muk (F1.muk) and V0
 * (F1.V0) are read only by the sliding laws that use them (NULL otherwise). */
typedef struct {
    const double *h, *ub, *vb, *S, *B, *rho, *C, *m, *n, *AGlen, *as, *ab;
    const double *h0, *ub0, *vb0, *as0, *ab0;
    const double *dadh;
    const double *q, *muk, *V0;
} uvh_fields;

/* Context on HIP device `device` (one per process and device; not thread-safe). */
int uvh_create(uvh_ctx **ctx, int device);
void uvh_destroy(uvh_ctx *ctx);
const char *uvh_last_error(const uvh_ctx *ctx);

/* Mesh (MUA) upload and sparsity pattern build, once per mesh.
 * Replaces the per-mesh part of MUA (connectivity, coordinates, points, weights:
 * UaMain/CreateMUA.m:31-61) that uvhMatrixAssembly reads; derivatives and detJ
 * (CalcMuaMeshDerivatives.m) are recomputed on the device from the coordinates.
 * nod in {3, 6, 10}; 1 <= nip <= 32; pts is nip x 2 (xi, eta), wts is nip. */
int uvh_set_mesh(uvh_ctx *ctx, int32_t nele, int32_t nod, int32_t nnodes, const int32_t *conn,
                 const double *x, const double *y, int32_t nip, const double *pts, const double *wts);

/* Pattern of K: neq = 3*nnodes rows, nnz values. */
int uvh_pattern_info(const uvh_ctx *ctx, int64_t *neq, int64_t *nnz);
/* rowptr: neq+1 entries, colind: nnz entries (host arrays). */
int uvh_get_pattern(const uvh_ctx *ctx, int64_t *rowptr, int32_t *colind);

/* CSC order of K, for hosts that store sparse matrices by column (MATLAB's sparse, the
 * reference's K: UaMain/uvhMatrixAssembly.m:509 via sparseUA.m:17). The pattern is structurally
 * symmetric, so the CSC column pointers and row indices are the CSR rowptr and colind of
 * uvh_get_pattern; the values are permuted: csc_val[k] = Kval[map[k]], map has nnz entries. */
int uvh_get_csc_map(const uvh_ctx *ctx, int64_t *map);

/* Host-buffer assembly: the uvhAssembly drop-in (uvhMatrixAssembly.m:6-585).
 * Fields are host arrays; R (3Nn) is required; Kval (nnz) may be NULL or is ignored with
 * UVH_RONLY; Tint and Fext (3Nn) may be NULL. Blocks until the results are on the host. */
int uvh_assemble(uvh_ctx *ctx, const uvh_ctrl *ctrl, const uvh_fields *fields, uint32_t flags,
                 double *R, double *Kval, double *Tint, double *Fext);

/* Device-buffer assembly, enqueued on `stream` (a hipStream_t; NULL is the HIP null stream).
 * All pointers are device pointers; nothing is copied to or from the host and the call
 * returns once the work is enqueued. NaN checks are deferred: uvh_sync_status() waits
 * for the context's last assembly (not for the device) and returns UVH_E_NAN_R / UVH_E_NAN_K
 * if it met a NaN.
 * One context runs one assembly at a time: its calls share the context's work buffers and
 * status word. A call on another stream than the context's previous call is ordered after
 * that call (the library makes the new stream wait for it), except under graph capture,
 * where the caller orders the replays. The status word holds the last call's NaN flags only.
 * Same seam as uvh_assemble (uvhMatrixAssembly.m:6-585), with device buffers. */
int uvh_assemble_device(uvh_ctx *ctx, const uvh_ctrl *ctrl, const uvh_fields *dfields, uint32_t flags,
                        double *dR, double *dKval, double *dTint, double *dFext, void *stream);
int uvh_sync_status(uvh_ctx *ctx);

/* Enqueues on `stream` a copy of the context's status word (bit 0: NaN in R, bit 1: NaN on
 * the diagonal of K, as uvhMatrixAssembly.m:575-584 tests) into `dst` (pinned host or device
 * memory), for callers that read each assembly's errors without waiting for the device (the
 * serving driver). Enqueue it on the stream of the assembly it reports, right after it. No
 * counterpart in the reference: it splits uvhMatrixAssembly's error check from the wait. */
int uvh_status_async(uvh_ctx *ctx, int32_t *dst, void *stream);

/* uvh_assemble_device without any allocation: the same arguments, kernels and results (bit for
 * bit), with the work buffers (status word, nodal lower surface, element residual and Jacobian
 * rows) in a 256 MiB device array of the library instead of the context's own buffers. No
 * counterpart in the reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc). A mesh whose
 * work buffers exceed 256 MiB (R+K: about 2.9 KB per 6-node element, so up to ~90k of them) is
 * refused with UVH_E_ARG. Every call on one GPU shares the array: calls that overlap on one GPU
 * race, whatever their contexts. NaN checks are deferred to uvh_sync_status as for
 * uvh_assemble_device. Always the two-pass kernels (the UVH_ASSEMBLY, UVH_CHUNKS and UVH_SCATTER
 * options do not apply), and no per-kernel timing. */
int uvh_assemble_device_noalloc(uvh_ctx *ctx, const uvh_ctrl *ctrl, const uvh_fields *dfields, uint32_t flags,
                                double *dR, double *dKval, double *dTint, double *dFext, void *stream);

/* Per-kernel timing with HIP events on the launch stream (for measurement).
 * uvh_profile(ctx, 1) starts recording; uvh_kernel_times fills up to `cap` entries of
 * (name, total ms, launches) accumulated since the last reset, and returns the count. */
int uvh_profile(uvh_ctx *ctx, int enable);
int uvh_kernel_times(uvh_ctx *ctx, const char **names, double *total_ms, int64_t *launches, int cap);

/* ---------------------------------------------------------------- uv (diagnostic) assembly
 * The velocity-only SSTREAM system on the same context and mesh:
 *   [Ruv,Kuv,Tint,Fext]=uvMatrixAssemblySSTREAM(CtrlVar,MUA,F,BCs)
 * (UaMain/uvMatrixAssemblySSTREAM.m:6, reached through uvMatrixAssembly.m), 2025 form.
 * R is 2Nn ([x; y], R = Tint - Fext); K is on the pattern of uvh_uv_get_pattern: 2Nn rows,
 * row (b, n) = [adj(n); adj(n) + Nn], 4*nnzN values, blocks [Kxu Kxv; Kyu Kyv]. The reference
 * symmetrises K as (K + K')/2 (:545-549); the uv Jacobian is a Hessian, and the library's K
 * is symmetric to rounding (within 1e-15 of the symmetrised matrix) without that step.
 * s is the upper surface (b = s - h at the points, :219); C and AGlen are clipped at Cmin and
 * AGlenmin (:185, :211); nrviscosity scales the viscosity-derivative terms (Dvisk, :72-76:
 * CtrlVar.NRviscosity, or 0 with CtrlVar.Picard). Flags: UVH_RONLY, UVH_FP32,
 * UVH_NO_NAN_CHECK; UVH_ZEROFIELDS is refused (UVH_E_UNSUPPORTED): the reference's uv
 * ZeroFields sets u, v to zero except at the fixed nodes (:24-28), which the caller passes
 * as ub, vb directly. */
typedef struct {
    const double *h, *s, *ub, *vb, *S, *B, *rho, *C, *m, *n, *AGlen;
    const double *q, *muk, *V0;   /* as in uvh_fields */
} uvh_uv_fields;

typedef struct {
    double cmin;         /* CtrlVar.Cmin */
    double aglenmin;     /* CtrlVar.AGlenmin */
    double nrviscosity;  /* CtrlVar.NRviscosity, 0 when CtrlVar.Picard */
} uvh_uv_ctrl;

/* Pattern of the uv K: neq = 2*nnodes rows, nnz = 4*nnzN values; rowptr neq+1, colind nnz. */
int uvh_uv_pattern_info(const uvh_ctx *ctx, int64_t *neq, int64_t *nnz);
int uvh_uv_get_pattern(const uvh_ctx *ctx, int64_t *rowptr, int32_t *colind);
/* Host buffers (R 2Nn required; Kval, Tint, Fext may be NULL); blocks until done. */
int uvh_uv_assemble(uvh_ctx *ctx, const uvh_ctrl *ctrl, const uvh_uv_ctrl *uvc, const uvh_uv_fields *fields,
                    uint32_t flags, double *R, double *Kval, double *Tint, double *Fext);
/* Device buffers, enqueued on `stream`; NaN checks deferred to uvh_sync_status; ordered with the
 * context's other calls as uvh_assemble_device is. */
int uvh_uv_assemble_device(uvh_ctx *ctx, const uvh_ctrl *ctrl, const uvh_uv_ctrl *uvc,
                           const uvh_uv_fields *dfields, uint32_t flags, double *dR, double *dKval, double *dTint,
                           double *dFext, void *stream);

/* ---------------------------------------------------------------- Newton-step linear solve
 * K d = rhs on the device with Dirichlet-fixed degrees of freedom eliminated: the branch
 * CtrlVar.AsymmSolver = 'EliminateBCsSolveSystemIterativly' of solveKApe
 * (UaMain/solveKApe.m:117-151: constrained rows and columns removed, a no-fill incomplete LU,
 * a preconditioned Krylov solve to a relative residual tolerance, CtrlVar.LinSolveTol =
 * 1e-8 by default). K is taken in 3 x 3 node blocks (block ILU(0) of the node graph) and the
 * Krylov method is BiCGStab. A fixed degree of freedom gets an identity row and column and a
 * zero right-hand side, so the other unknowns are those of the eliminated system and the
 * fixed ones are zero (the Newton-step form of x(iConstrainedDOF) = g).
 *
 * uvh_lin_create: for the mesh of `ctx` (set_mesh first; a later set_mesh invalidates the
 *   solver); `fixed` is a host mask of 3*Nn bytes in [ub; vb; h] order (nonzero = fixed), or
 *   NULL. Analyses the pattern once. Every node must belong to an element (a node in none has
 *   an empty row, K is singular there): UVH_E_ARG otherwise.
 * uvh_lin_factor: block ILU(0) of the device K values `dKval` (the context's CSR order, as
 *   uvh_assemble_device writes them), computed in FP64; the triangular solves read the factors
 *   rounded to FP32 (sums and vectors in FP64), or the FP64 factors when the environment
 *   variable UVH_LIN_FACTOR is "fp64" at uvh_lin_create. UVH_E_NOCONV on a zero pivot.
 * uvh_lin_solve: dx = K^-1 drhs (device vectors of 3*Nn in [ub; vb; h] order) to
 *   ||rhs - K dx|| <= rtol ||rhs|| within maxit iterations; *iterations and *relres (the true
 *   relative residual of dx; either may be NULL) are set on every return; iterations counts
 *   BiCGStab iterations plus GMRES steps. UVH_E_NOCONV if the tolerance is not reached; dx then
 *   holds the best iterate seen. BiCGStab keeps its scalars on the device and blocks the host
 *   once per iteration, with the next iteration already queued; GMRES blocks once per step. */
typedef struct uvh_lin uvh_lin;

/* Krylov methods (uvh_lin_set_krylov): BiCGStab; GMRES(restart), the reference branch's
 * method (gmres with restarts, solveKApe.m:143-144); AUTO (the default): BiCGStab, and when it
 * breaks down or misses the tolerance, GMRES(restart) from its best iterate. */
enum { UVH_LIN_BICGSTAB = 0, UVH_LIN_GMRES = 1, UVH_LIN_AUTO = 2 };

int uvh_lin_create(uvh_ctx *ctx, const uint8_t *fixed, uvh_lin **lin);
/* method: UVH_LIN_*; restart: GMRES cycle length, 2..64 (default 30). */
int uvh_lin_set_krylov(uvh_lin *lin, int32_t method, int32_t restart);
void uvh_lin_destroy(uvh_lin *lin);
const char *uvh_lin_last_error(const uvh_lin *lin);
int uvh_lin_factor(uvh_lin *lin, const double *dKval, void *stream);
int uvh_lin_solve(uvh_lin *lin, const double *drhs, double *dx, double rtol, int32_t maxit, int32_t *iterations,
                  double *relres, void *stream);

/* Batched solves. The reference solves one system per call (UaMain/solveKApe.m:117-151); an
 * ensemble of members each takes that branch in turn. Here the BiCGStab iterations of n
 * members' systems advance together, one launch per step of the iteration for all of them
 * (flat work lists of the members' vector chunks and block rows, and per colour index the
 * members' rows of that colour), each member being a uvh_lin of its own mesh, factored with
 * uvh_lin_factor before the solve. The convergence test, the near-breakdown restarts and the
 * iteration limit are per member and on the device; the host waits once per iteration, with
 * the next iteration already queued, until every member has stopped. For callers that solve
 * many members' Newton steps together (the batched Newton form).
 *
 * uvhb_lin_create: members[0..n-1], all on one device, each of at most 17,301,504 unknowns
 *   (UVH_E_ARG otherwise); destroy the batch before any member.
 * uvhb_lin_solve: dx[m] = K_m^-1 drhs[m] (device vectors of 3*Nn_m in [ub; vb; h] order) for
 *   every member with active[m] != 0 (active NULL: all), to ||rhs - K dx|| <= rtol ||rhs||
 *   within maxit iterations; iterations[m], relres[m] (true relative residual) and converged[m]
 *   are set for every member, 0 for inactive ones (each array may be NULL). UVH_E_NOCONV if an
 *   active member missed the tolerance; its dx then holds its best iterate.
 * The same solve in steps, for callers that capture the iterations in a HIP graph:
 * uvhb_lin_begin: the right-hand sides and their norms (one host synchronisation) and every
 *   member's start; uvhb_lin_iterate: `iterations` iterations of every member, launches only (no
 *   synchronisation, no allocation: capturable), then the members' stop words copied to pinned
 *   host memory; iterations after a member has stopped do nothing for it; uvhb_lin_stopped: 1
 *   when every stop word read by the last uvhb_lin_iterate (the caller synchronises first)
 *   says stopped, else 0; uvhb_lin_end: best iterates, true residuals and the solutions, as
 *   uvhb_lin_solve returns them. */
typedef struct uvhb_lin uvhb_lin;

int uvhb_lin_create(uvh_lin *const *members, int32_t n, uvhb_lin **batch);
void uvhb_lin_destroy(uvhb_lin *batch);
const char *uvhb_lin_last_error(const uvhb_lin *batch);
int uvhb_lin_solve(uvhb_lin *batch, const double *const *drhs, double *const *dx, const uint8_t *active, double rtol,
                   int32_t maxit, int32_t *iterations, double *relres, int32_t *converged, void *stream);
int uvhb_lin_begin(uvhb_lin *batch, const double *const *drhs, const uint8_t *active, double rtol, int32_t maxit,
                   void *stream);
int uvhb_lin_iterate(uvhb_lin *batch, int32_t iterations, void *stream);
int uvhb_lin_stopped(const uvhb_lin *batch);
int uvhb_lin_end(uvhb_lin *batch, double *const *dx, int32_t *iterations, double *relres, int32_t *converged,
                 void *stream);

/* ---------------------------------------------------------------- batched assembly
 * Many independent meshes (ensemble members with perturbed C and AGlen, mixed 3-, 6- and
 * 10-node elements) assembled together: per element type, one element launch and one row
 * launch over a flattened (member, block) work list (one workgroup per item; with the
 * environment variable UVH_BATCH=persistent read at uvhb_create, persistent workgroups that
 * take items from an atomic counter instead). Replaces N separate uvhAssembly
 * calls (UaMain/uvhAssembly.m:1), one per member; Úa itself spreads such work over MATLAB
 * workers (uvhMatrixAssemblySSTREAM_SPMD.m), which here share one GPU instead.
 * Each member's results are those of uvh_assemble on that member alone to rounding (the
 * kernels are compiled separately, so FMA contraction may differ) and bit for bit the same
 * from run to run. */
typedef struct uvh_batch uvh_batch;

int uvhb_create(uvh_batch **batch, int device);
void uvhb_destroy(uvh_batch *batch);
const char *uvhb_last_error(const uvh_batch *batch);
/* Adds a mesh (arguments as uvh_set_mesh); *member receives its index. */
int uvhb_add_mesh(uvh_batch *batch, int32_t nele, int32_t nod, int32_t nnodes, const int32_t *conn,
                  const double *x, const double *y, int32_t nip, const double *pts, const double *wts,
                  int32_t *member);
int uvhb_size(const uvh_batch *batch);
/* The member's context (owned by the batch): pattern, CSC map and single-mesh assembly. */
uvh_ctx *uvhb_member(uvh_batch *batch, int32_t member);
/* Element tile (elements per 256-thread workgroup) of every member with `nod` nodes, for
 * the FP64 (fp32 = 0) or FP32 element arithmetic; the autotuner's output. */
int uvhb_set_tile(uvh_batch *batch, int32_t nod, int32_t fp32, int32_t epb);
int uvhb_get_tile(const uvh_batch *batch, int32_t nod, int32_t fp32);
/* Per member m of the batch: rr[m] = sum_i (w_i R_i)^2 and d2[m] = -sum_i w_i R_i d_i over the
 * member's 3 Nn entries, with weights w (0 at fixed dofs, 1 elsewhere; w NULL or w[m] NULL: all 1):
 * the numerator of rForce and the D2 of CalcCostFunctionNRuvh.m:45-85 for each member's Newton
 * step d. R, d, w are arrays of the members' device pointers; rr and d2 device arrays of nmem
 * doubles. Sums in a fixed order with no atomics and nothing cleared, so the call can be captured
 * in a HIP graph (the first call sets up the batch's tables and must not be captured). */
int uvhb_cost_device(uvh_batch *b, const double *const *dR, const double *const *dd, const double *const *dw,
                     double *rr, double *d2, void *stream);
/* Device-buffer assembly of every member: dfields, dR and dK are arrays of uvh_size()
 * entries (dK may be NULL with UVH_RONLY). Enqueued on `stream`; uvhb_sync_status waits and
 * fills member_status (may be NULL) with UVH_OK / UVH_E_NAN_R / UVH_E_NAN_K per member. */
int uvhb_assemble_device(uvh_batch *batch, const uvh_ctrl *ctrl, const uvh_fields *dfields, uint32_t flags,
                         double *const *dR, double *const *dK, void *stream);
int uvhb_sync_status(uvh_batch *batch, int32_t *member_status);

#ifdef __cplusplus
}
#endif

#endif /* UVH_H */
