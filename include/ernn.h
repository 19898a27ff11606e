/* C ABI of libernn_hip.so: the ERNN forward recurrence and its reverse-time
 * adjoint on AMD Instinct MI355X (gfx950).
 *
 * The reference (guoziye97/Elastodynamics-encoded-RNN) has no FFI layer: its
 * boundary is the nn.Module contract of
 *   CFRPWaveSimulatorRicker.forward           Parameter_inversion/Inversion_adam.py:240-307
 *   SpatiallyVaryingAnisotropicSimulator.forward  Damage_inversion/Inversion_adam.py:331-393
 * and the autograd reverse pass that loss.backward() runs through it
 *   Parameter_inversion/Inversion_adam.py:447, Damage_inversion/Inversion_adam.py:511.
 * ernn_forward replaces the time loop of those forward methods (one call per
 * forward instead of ~60 tensor ops per step); ernn_adjoint replaces the
 * autograd replay (one call per backward).  The Python package binds these
 * entry points with ctypes behind a torch.autograd.Function.
 *
 * Conventions
 *  - A field is nx x ny, row-major with y (index j) contiguous, exactly the
 *    reference's (1, nx, ny) tensors (PI:186-191, PI:255).
 *  - Stiffness order is the reference's C11, C22, C12, C16, C26, C66
 *    (PI:86-98), i.e. C11, C33, C13, C15, C35, C55 in x-z notation.
 *  - All pointers in ernn_call_t are device pointers (HBM).  The library
 *    allocates no device memory per call: the caller provides the workspace
 *    (size from ernn_plan_info) and the outputs.  Device memory is allocated
 *    only by ernn_plan_create (tables) and freed by ernn_plan_destroy.
 *  - All work is enqueued on the caller's stream; no call synchronizes the
 *    host except ernn_plan_create/destroy and ernn_time_steps.
 *  - Every function returns 0 on success or a negative ERNN_E* code; the
 *    message of the last failure on the calling thread is ernn_last_error().
 *  - A plan must not be used by two streams at the same time.
 */
#ifndef ERNN_H_
#define ERNN_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ERNN_ABI_VERSION 7

#define ERNN_OK 0
#define ERNN_E_INVALID (-1)   /* bad argument or descriptor */
#define ERNN_E_HIP (-2)       /* HIP runtime error */
#define ERNN_E_CAPACITY (-3)  /* size beyond what the plan supports */

/* Parameter modes. */
#define ERNN_PARAM_CONST6 0  /* six constants per set: PI:166-171 */
#define ERNN_PARAM_DFIELD 1  /* C_p(cell) = D(cell) * C0_p: one damage field + six constants */
#define ERNN_PARAM_FIELD6 2  /* six per-cell fields per set: DI:245-258 */

/* One instance = one shot (source time function + point forces) on one grid
 * with one parameter set.  Instances of a set share the set's parameters and
 * sum their gradients into it (shots of one model); sets of the field modes
 * require equal grid sizes for all their instances. */
typedef struct {
  int32_t nx, ny;       /* grid (PI:186) */
  int32_t nt;           /* time steps (PI:268) */
  int32_t set;          /* parameter set index, 0 <= set < n_sets */
  double h, dt, rho;    /* spacing, step, density (PI:139-147) */
  int32_t src_begin, src_count;  /* slice of the source array */
  int32_t rec_begin, rec_count;  /* slice of the receiver array (receiver mode) */
} ernn_instance_t;

/* A point force: weight w on cell (i, j), component comp (0 = x, 1 = y).
 * At step n it adds wavelet[n] * w to the operator before the division by
 * rho, like force_amplitude * gaussian_dist (PI:270-277, PI:103-104). */
typedef struct {
  int32_t i, j;
  float w;
  int32_t comp;
} ernn_source_t;

typedef struct {
  int32_t i, j;
} ernn_receiver_t;

typedef struct {
  int32_t n_inst;
  int32_t n_sets;
  int32_t param_mode;   /* ERNN_PARAM_* */
  int32_t rec_stride;   /* P: frame k records u^{kP+1} (PI:294-301); 1 = save_all_steps */
  int32_t snapshot;     /* 1: frames are whole fields (PI:304-305); 0: receiver traces */
  int32_t ckpt_every;   /* K >= 2: checkpoint interval of the adjoint recompute */
  int32_t fwd_v, fwd_r; /* forward tile: V in {1,2,4} columns per lane, R in {1,2,4,8,16}
                           rows up to 8, V * R <= 16 (six fields: V * R <= 8); 0, 0 = default (1, 4) */
  int32_t adj_v, adj_r; /* adjoint tile: as above, damage field V * R <= 8 and R <= 4, six
                           fields V * R <= 4 and R <= 2; 0, 0 = default (1, 4), six fields (1, 2) */
  int32_t use_graph;    /* 1: replay each pass as a captured hipGraph */
  int32_t fwd_bpc;      /* forward grid cap in blocks per CU (0 = one tile per wavefront);
                           below the tile count the wavefronts loop over tiles */
  int32_t adj_bpc;      /* adjoint grid cap */
  int32_t fuse;         /* 1: run each adjoint step together with one recompute step of the
                           previous checkpoint segment (one launch; K more slots per instance) */
  int32_t blk_v, blk_rw, blk_ks; /* time-blocked forward: blk_ks steps per launch on blocks of 4
                           wavefronts x blk_rw rows x 64 blk_v columns; blk_ks <= 1 = one step per
                           launch.  V in {1,2}, RW in {4,8}, KS in {2,4,8}, 4 RW > 2 KS, or
                           (1, 8, 5); field modes exclude (V, RW) = (2, 8).  Used by the
                           forward pass and the segment recompute. */
  int32_t ablk_v, ablk_rw, ablk_ks; /* time-blocked adjoint (CONST6, DFIELD): ablk_ks reverse
                           steps per launch.  Unfused backward: V = 1 with RW in {4,8}, or
                           (1, 6, 4), or (CONST6) (1, 8, 5) or V = 2 with RW = 4 and KS in
                           {2,4}.  fuse = 1:
                           each launch also runs ablk_ks recompute steps of the previous
                           segment with the same shape, (1, 4, 2), (1, 4, 4) or (CONST6) (1, 8, 4).
                           ablk_ks <= 1 = off */
  int32_t ablk_ff_rw;   /* fuse = 1 with a time-blocked adjoint: rows per wavefront of the
                           fused kernel's recompute half, 0 = ablk_rw, or 8 beside ablk_rw = 4
                           with ablk_ks in {2,4} (shorter grids: every block resident at once);
                           otherwise 0 */
  int32_t fixed_c0;     /* DFIELD only, 0 or 1.  1: the six background constants are known
                           (stiffness-degradation imaging, C_p = D C0_p with C0 fixed) and the
                           adjoint returns dJ/dD only.  dJ/dD(cell) = sum over steps of
                           lam^{m+1} . (dt^2/rho) A(C0) u^m needs u^m only through A(C0) u^m,
                           which the forward computes anyway: the checkpoint bank then holds
                           (dt^2/rho) A(C0) u^m instead of u^m (same size), and the adjoint
                           reads it pointwise instead of taking a stencil of u.  Requires
                           fuse = 0; grad_params is not written. */
  int32_t reverse_u;    /* ERNN_PARAM_DFIELD (joint, or with fixed_c0 = 1), 0 or 1.  1: no
                           bank and no recompute; the adjoint runs the forward recurrence
                           backward in time from the checkpoints beside lam (the leapfrog
                           step is time-symmetric: u^{m-1} = u^m - v^m, v^m = v^{m+1} - d^m),
                           forming A(C0) u^m on the way (and, for the joint gradient, the
                           constant-gradient sums).  The checkpoints hold (u, c, v) at every
                           K-th level and each instance's last level.  Requires fuse = 0 and
                           a time-blocked adjoint (ablk_v, ablk_rw, ablk_ks) = (1, RW, KS),
                           RW in {4, 6, 8}, KS in {2, 4}, or RW in {8, 10}, KS in {5, 6}. */
  int32_t ortho;        /* with fixed_c0 = 1, 0 or 1.  1: the caller promises C16 = C26 = 0 in
                           every set's constants (a background orthotropic along the grid axes);
                           the time-blocked forward and the reversed adjoint drop those terms of
                           the operator sums (the same bits as with them, fewer instructions).
                           Constants that break the promise give wrong results. */
  const ernn_instance_t* inst;  /* host array [n_inst] */
  const ernn_source_t* src;     /* host array [n_src] */
  int64_t n_src;
  const ernn_receiver_t* rec;   /* host array [n_rec] */
  int64_t n_rec;
} ernn_desc_t;

typedef struct {
  int64_t workspace_bytes;  /* caller-provided workspace, 256-byte aligned */
  int64_t trace_floats;     /* total floats of the trace array */
  int64_t grad_field_doubles; /* total doubles of grad_field (field modes) */
  int64_t field_floats;     /* total floats of the external field input (field modes) */
  int32_t n_frames_max;
  int32_t ckpt_every;
  int32_t n_ckpt;           /* checkpoints per instance */
  int32_t fwd_v, fwd_r, adj_v, adj_r;
  int32_t fwd_tiles, adj_tiles;
  int32_t use_graph;
  int32_t fwd_bpc, adj_bpc;
  int32_t fwd_blocks, adj_blocks;
  int32_t n_cu;             /* compute units of the plan's device */
  int32_t fuse;
  int32_t blk_v, blk_rw, blk_ks; /* time-blocked forward shape (blk_ks = 0: off) */
  int32_t blk_blocks;       /* blocks per time-blocked launch */
  int32_t ablk_v, ablk_rw, ablk_ks; /* time-blocked adjoint shape (ablk_ks = 0: off) */
  int32_t ablk_blocks;
  int32_t ablk_ff_rw;       /* rows per wavefront of the fused recompute half (0: no fused
                               time-blocked backward) */
  int32_t fixed_c0;         /* the descriptor's fixed_c0 (DFIELD: dJ/dD only) */
  int32_t reverse_u;        /* the descriptor's reverse_u */
  int32_t ortho;            /* the descriptor's ortho */
} ernn_plan_info_t;

/* Per-call device pointers.
 *  params      [n_sets][6] fp32 constants (CONST6, DFIELD; C0 for DFIELD)
 *  field       per set, [n_coef][nx*ny] fp32, sets back to back (DFIELD: D,
 *              n_coef = 1; FIELD6: the six C fields, n_coef = 6)
 *  wavelets    per instance, [nt] fp32 source time function (PI:210-217),
 *              instances back to back
 *  traces      forward output.  Instance b's block starts at
 *              ernn_trace_offsets()[b] and is [2][n_frames][X] with X =
 *              rec_count (receiver mode, ordered as the instance's receiver
 *              slice) or nx*ny (snapshot mode); component 0 = ux, 1 = uy.
 *  grad_traces adjoint input dJ/d traces, same layout as traces
 *  grad_params [n_sets][6] fp64 output (CONST6, DFIELD without fixed_c0)
 *  grad_field  fp64 output, per set [n_coef][nx*ny] (DFIELD: dJ/dD; FIELD6:
 *              dJ/dC_p fields), sets back to back */
typedef struct {
  void* workspace;
  const float* params;
  const float* field;
  const float* wavelets;
  float* traces;
  const float* grad_traces;
  double* grad_params;
  double* grad_field;
} ernn_call_t;

/* Version of this ABI (ERNN_ABI_VERSION). */
int ernn_abi_version(void);

/* Message of the last failure on this thread ("" if none). */
const char* ernn_last_error(void);

/* Validate a descriptor, build tile and event tables on the current device.
 * Replaces the set-up of CFRPWaveSimulatorRicker.__init__ (PI:139-208) and
 * SpatiallyVaryingAnisotropicSimulator.__init__ (DI:208-296). */
int ernn_plan_create(const ernn_desc_t* desc, void** plan_out);

int ernn_plan_destroy(void* plan);

int ernn_plan_info(const void* plan, ernn_plan_info_t* info);

/* trace_off[b] for b in [0, n_inst] (n_inst + 1 entries). */
int ernn_trace_offsets(const void* plan, int64_t* trace_off);

/* Zero a new workspace (ghost cells must be zero; the kernels never write
 * them).  Uses a fill kernel, never hipMemset, so it may be graph-captured. */
int ernn_workspace_init(void* plan, void* workspace, void* stream);

/* Forward recurrence of every instance (PI:240-307, DI:331-393): writes the
 * traces and the checkpoints the adjoint needs into the workspace. */
int ernn_forward(void* plan, const ernn_call_t* call, void* stream);

/* Gradient of a scalar loss J given dJ/d traces (replaces loss.backward(),
 * PI:447 / DI:511).  Needs the workspace of the matching ernn_forward.
 * Recomputes the forward from checkpoints segment by segment and runs the
 * reverse-time adjoint recurrence.  Accumulation: the constant gradients
 * are summed in FP32 over one launch's cells and steps of a wavefront, then
 * in FP64 (per-wavefront partials, a fixed-order FP64 reduction per set that
 * depends only on the set's own instances); the per-cell gradients are
 * summed in FP32 over one launch (at most 8 steps) by the time-blocked
 * adjoint, or over one checkpoint segment by the one-step adjoint, and in
 * FP64 across launches or segments.  The forward pass also
 * leaves every state of the last checkpoint segment in the workspace; the
 * first adjoint on that workspace after it starts from them instead of
 * recomputing the segment (later adjoints, or a workspace touched by
 * ernn_time_steps or ernn_workspace_init since, recompute it). */
int ernn_adjoint(void* plan, const ernn_call_t* call, void* stream);

/* Change the tile (V, R) and grid cap (blocks per CU, 0 = none) of the
 * forward (which = 0) or adjoint (which = 1) kernel.  Used by the autotuner;
 * synchronizes the device and drops captured graphs. */
int ernn_plan_set_tiling(void* plan, int which, int v, int r, int bpc);

/* Switch the forward (which = 0) or the adjoint (which = 1) to ks
 * steps per launch on blocks of the given shape (ks <= 1: one step per
 * launch); same limits as ernn_desc_t.blk_* / ablk_*.  which = 2 (fuse plans
 * with a time-blocked adjoint) sets the rows per wavefront of the fused
 * kernel's recompute half: v and ks must be the adjoint's, rw the adjoint's
 * or 8 beside an adjoint RW of 4 with ks in {2,4}; setting the adjoint shape
 * resets it to the adjoint's.  No counterpart in the reference: its time loop
 * (PI:268-305) and autograd replay run one step at a time.  Synchronizes the
 * device and drops captured graphs. */
int ernn_plan_set_blocking(void* plan, int which, int v, int rw, int ks);

/* Time `steps` forward (which = 0), adjoint (which = 1) or fused adjoint +
 * recompute (which = 2, fuse plans) steps with the current tiling, or `steps`
 * time-blocked forward (which = 3) or adjoint (which = 4; fuse plans: fused
 * adjoint + recompute) launches, using hipEvents on `stream`; synchronizes.
 * The workspace must be initialized.  ms_out receives the elapsed
 * milliseconds. */
int ernn_time_steps(void* plan, int which, int steps, const ernn_call_t* call, void* stream,
                    float* ms_out);

/* ------------------------------------------------------------------------
 * FEM observed-data generator: the reference's explicit P1-triangle solver
 *   solve_2d_anisotropic_wave_fem_ricker_fixed
 *     Parameter_inversion/FEM_generate_data.py:196-590 (constant C),
 *     Damage_inversion/FEM_generate_data.py:770-1224 (damaged C)
 * on its checkerboard mesh (PF:117-189): lumped mass (PF:327-339), element
 * forces of PF:417-457 (per-element C at the node nearest the centroid,
 * DF:1003-1026), central differences with the boundary held at zero
 * (PF:481-497), frames scaled by rec_scale (0.7, PF:525-526).  ernn_fem_run
 * replaces the whole time loop (PF:481-513, ~100 tensor ops per step) by one
 * launch per step; nodes gather their forces, no atomics.
 * Fields are nx x ny row-major with y contiguous, the reference's
 * (nx, ny, t) output order (PF:517-523); frame k holds rec_scale * u^{k P}.
 * ------------------------------------------------------------------------ */
#define ERNN_FEM_CONST 0   /* one stiffness matrix (PF:402-406) */
#define ERNN_FEM_DAMAGE 1  /* C = D(node nearest the element centroid) * C0 (DF:183-212, DF:1017-1022) */

typedef struct {
  int32_t nx, ny;        /* nodes per side, boundary included (PF:117-147) */
  int32_t nt;            /* time samples; steps n = 1 .. nt-2 produce u^{n+1} (PF:481) */
  int32_t mode;          /* ERNN_FEM_* */
  double dx, dy;         /* Lx / (nx - 1), Ly / (ny - 1) (PF:140-141) */
  double dt, rho;        /* time step after the CFL check (PF:280-296), density */
  int32_t rec_stride;    /* P >= 1: frame k = u^{k P}, k < ceil(nt / P) (PF:926-930) */
  float rec_scale;       /* frame factor (0.7 in PF:525-526) */
  int32_t tile_v, tile_r; /* tile: a wavefront sweeps R node rows x 63 node columns, one column per
                            lane: (V, R) = (1, 1), (1, 2), (1, 4) or (1, 8); 0, 0 = (1, 1) */
  int32_t use_graph;     /* 1: replay the run as a captured hipGraph */
  int32_t blk_rw, blk_ks; /* time-blocked: blk_ks steps per launch on blocks of 4 wavefronts (8 for
                            RW = 2) x blk_rw node rows x 64 columns, (RW, KS) in (2, 2) (2, 4) (4, 2)
                            (4, 4) (8, 2) (8, 4) (8, 8);
                            blk_ks <= 1: one step per launch with the tile above.  Same results bit
                            for bit. */
  const ernn_source_t* loads; /* host array: nodal loads w at (i, j) on component comp, the
                            Gaussian of PF:298-301 at the nodes (PF:322); f_ext = w(t_n) w / rho * M
                            (PF:482-484).  Loads on boundary nodes have no effect and are dropped */
  int64_t n_loads;
} ernn_fem_desc_t;

typedef struct {
  int64_t workspace_bytes; /* caller-provided workspace, 256-byte aligned */
  int64_t frame_floats;    /* floats of each frame array, n_frames * nx * ny */
  int32_t n_frames;
  int32_t tile_v, tile_r;
  int32_t n_tiles, blocks;
  int32_t n_loads;         /* loads kept (interior nodes, one per node and component) */
  int32_t use_graph;
  int32_t blk_rw, blk_ks;  /* time-blocked shape (blk_ks = 0: off) */
  int32_t blk_blocks;      /* blocks per time-blocked launch */
} ernn_fem_info_t;

/* Per-run device pointers. */
typedef struct {
  void* workspace;
  const float* params;   /* [6] C11, C22, C12, C16, C26, C66 (DAMAGE: the intact C0) */
  const float* field;    /* DAMAGE: per-node factor D, nx * ny (1 intact, 0.5 damaged in DF) */
  const float* wavelet;  /* [nt] time factors w(t_n) (PF:459-463) */
  float* frames_x;       /* [n_frames][nx][ny] output, ux */
  float* frames_y;       /* [n_frames][nx][ny] output, uy */
} ernn_fem_call_t;

/* Validate a descriptor and build the tile and load tables on the current
 * device (replaces the set-up of PF:300-463). */
int ernn_fem_plan_create(const ernn_fem_desc_t* desc, void** plan_out);

int ernn_fem_plan_destroy(void* plan);

int ernn_fem_plan_info(const void* plan, ernn_fem_info_t* info);

/* The whole run (PF:479-526): zero state and frames, then steps n = 1 ..
 * nt-2, writing frame k = rec_scale * u^{kP} (frame 0 = 0).  Enqueued on
 * `stream`; no host synchronization. */
int ernn_fem_run(void* plan, const ernn_fem_call_t* call, void* stream);

/* Change the tile (V, R); synchronizes the device and drops the captured
 * graph.  Used by the autotuner. */
int ernn_fem_set_tiling(void* plan, int v, int r);

/* Switch to ks steps per launch on blocks of the given shape (ks <= 1: one
 * step per launch); same limits as ernn_fem_desc_t.blk_*.  Synchronizes the
 * device and drops the captured graph.  Used by the autotuner. */
int ernn_fem_set_blocking(void* plan, int rw, int ks);

/* Time steps n = 1 .. min(steps, nt - 2) after the run's set-up, with
 * hipEvents on `stream`; synchronizes.  ms_out receives the milliseconds. */
int ernn_fem_time_steps(void* plan, int steps, const ernn_fem_call_t* call, void* stream, float* ms_out);

/* ------------------------------------------------------------------------
 * Total-variation regularisation + Adam update of per-cell parameter fields
 * (Damage_inversion/Inversion_adam.py): theta_p are the trainable fields,
 * C_p = clamp(exp(theta_p), lo_p, hi_p) (DI:304-313; log_params = 0:
 * clamp(theta_p), DI:315-320), the loss adds lambda * sum_p TV(C_p) with
 * TV = sum |C[i+1,j] - C[i,j]| + sum |C[i,j+1] - C[i,j]| (DI:179-195,
 * DI:453-460), and torch.optim.Adam steps the thetas (DI:478-483, DI:514).
 * ernn_tv_adam_step replaces the TV forward, its autograd backward and the
 * optimizer step (~30 tensor ops per field) by two passes over the fields.
 * ------------------------------------------------------------------------ */
#define ERNN_TV_MAX_FIELDS 8

typedef struct {
  int32_t nx, ny;        /* field shape, row-major, y contiguous */
  int32_t n_fields;      /* 1 .. ERNN_TV_MAX_FIELDS */
  int32_t log_params;    /* 1: C = clamp(exp(theta)); 0: C = clamp(theta) */
  int32_t step;          /* Adam step count t >= 1, this update included */
  int32_t pad;
  double lambda_tv;      /* TV weight (DI:404) */
  double lr, beta1, beta2, eps, weight_decay;   /* torch.optim.Adam arguments (DI:478-483) */
  float lo[ERNN_TV_MAX_FIELDS], hi[ERNN_TV_MAX_FIELDS]; /* clamp bounds per field (DI:265-270) */
} ernn_tv_adam_desc_t;

/* Device pointers, one per field (fields need not be contiguous). */
typedef struct {
  float* theta[ERNN_TV_MAX_FIELDS]; /* updated in place */
  float* grad[ERNN_TV_MAX_FIELDS];  /* in: dJ_data/dtheta; out: the total gradient
                                       dJ_data/dtheta + lambda dTV/dtheta (what loss.backward()
                                       leaves in theta.grad for the whole loss) */
  float* m[ERNN_TV_MAX_FIELDS];     /* Adam first moments (zero before step 1), updated in place */
  float* v[ERNN_TV_MAX_FIELDS];     /* Adam second moments, updated in place */
  double* tv;                       /* optional [n_fields] output: TV(C_p) of the fields before the
                                       update (the loss term without lambda); NULL = skip */
  void* workspace;                  /* ernn_tv_adam_workspace_bytes, 8-byte aligned */
} ernn_tv_adam_call_t;

/* Workspace bytes of ernn_tv_adam_step for this descriptor (validates it). */
int ernn_tv_adam_workspace_bytes(const ernn_tv_adam_desc_t* desc, int64_t* bytes);

/* One regularised Adam step: grad += lambda dTV/dtheta, then Adam on theta,
 * m, v (torch's foreach arithmetic order).  Enqueued on `stream`. */
int ernn_tv_adam_step(const ernn_tv_adam_desc_t* desc, const ernn_tv_adam_call_t* call, void* stream);

/* ernn_tv_adam_step without a workspace: call->workspace is ignored and the
 * FP64 partial sums go to a static device array of 65536 doubles (512 KiB)
 * that each GPU holds until the process ends, so the call allocates nothing.
 * No counterpart in the reference: it is for callers that cannot allocate
 * (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc).  Limit: n_fields x ceil(nx * ny / 256) <= 65536 (about 2.8 M
 * cells per field for six fields), else ERNN_E_CAPACITY.  Calls on one GPU
 * share the array: calls that overlap (other streams or threads) race. */
int ernn_tv_adam_step_static(const ernn_tv_adam_desc_t* desc, const ernn_tv_adam_call_t* call, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* ERNN_H_ */
