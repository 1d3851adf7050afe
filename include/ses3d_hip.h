/*
 * ses3d_hip.h -- C-ABI of the MI355X SES3D time-stepper (libses3d_hip.so).
 *
 * The reference has no argument-passing interface for its hot path: the
 * Fortran driver SOURCE/ses3d_main.f90 calls argument-less subroutines that
 * share `module variables` (SOURCE/ses3d_modules.f90:33-233).  Each entry
 * point below replaces one of those calls or the module state it reads or
 * writes; the citation on each names what it replaces.
 *
 * Conventions
 *  - Every function returns 0 on success and a negative code on failure;
 *    ses3d_hip_last_error() returns the message of the last failure on the
 *    calling thread.
 *  - Host arrays are caller-owned, read or written only during the call.
 *  - Element-local arrays use the reference layout: Fortran column-major
 *    (0:nx,0:ny,0:nz,0:4,0:4,0:4) with the element index along theta fastest.
 *    Global arrays use (0:4(nx+1),0:4(ny+1),0:4(nz+1)).
 *  - Angles are radians and lengths metres, i.e. the values the reference
 *    holds after SOURCE/ses3d_input.f90 converted them (x*pi/180 in float32
 *    with pi=3.1415926535898).
 *  - One context holds one mesh and model (one section) and n_events events
 *    that advance together; a context is used from one host thread.
 *  - Results are bitwise reproducible run to run (no floating-point atomics).
 */
#ifndef SES3D_HIP_H
#define SES3D_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SES3D_HIP_OK 0
#define SES3D_HIP_EINVAL (-1)
#define SES3D_HIP_EDEVICE (-2)
#define SES3D_HIP_ESTATE (-3)
#define SES3D_HIP_ENOMEM (-4)

typedef struct ses3d_hip_context ses3d_hip_context;

/* Run parameters.  Replaces the scalars ses3d_input reads from the boxfile,
 * setup, event and relax files (SOURCE/ses3d_input.f90:49-399) and the
 * compile-time `pml` (SOURCE/ses3d_modules.f90:21). */
typedef struct {
    int32_t nx, ny, nz;        /* elements-1 along theta, phi, r (boxfile nx_max-nx_min) */
    int32_t lpd;               /* Lagrange degree; must be 4 */
    int32_t pml;               /* absorbing layer width in elements (reference: 2) */
    int32_t is_diss;           /* 1: three-mechanism anelastic stress */
    int32_t adjoint_flag;      /* 0 forward, 1 forward storing the forward field, 2 adjoint */
    int32_t samp_ad;           /* forward-field sampling interval */
    int32_t nt;                /* time steps */
    float dt;                  /* time increment [s] */
    float xmin, xmax;          /* colatitude bounds of the box [rad] */
    float ymin, ymax;          /* longitude bounds [rad] */
    float zmin, zmax;          /* radius bounds [m] */
    float zmax_global;         /* radius of the surface from the setup file [m] */
    float tau_p[3], D_p[3];    /* relaxation times and weights (INPUT/relax) */
} ses3d_hip_config;

/* Create a context on the current HIP device for n_events events sharing
 * one mesh and model.  Replaces the static module storage of
 * SOURCE/ses3d_modules.f90:33-233 (all sizes are run-time values here). */
int ses3d_hip_create(const ses3d_hip_config *cfg, int32_t n_events, ses3d_hip_context **out);

/* Release all device memory of a context. */
void ses3d_hip_destroy(ses3d_hip_context *ctx);

/* Message of the last failure on this thread (empty string if none). */
const char *ses3d_hip_last_error(void);

/* Model fields, float32, reference layout; Q may be NULL when is_diss==0.
 * Replaces the model reads of SOURCE/ses3d_input.f90:472-499. */
int ses3d_hip_set_model(ses3d_hip_context *ctx, const float *rhoinv, const float *mu,
                        const float *lambda, const float *A, const float *B, const float *C,
                        const float *Q);

/* ses3d_hip_set_model from device memory: the same fields, float32, reference
 * layout, at device addresses readable from the context's device (e.g. the
 * buffers of torch tensors), read on the context's stream without a host
 * round trip.  Their contents must be complete when the call is made (the
 * call orders nothing against the caller's streams); it returns once they
 * have been read. */
int ses3d_hip_set_model_device(ses3d_hip_context *ctx, const float *rhoinv, const float *mu,
                               const float *lambda, const float *A, const float *B, const float *C,
                               const float *Q);

/* Source of event ev: source_type 1/2/3 (point force theta/phi/r) or 10
 * (moment tensor); xxs, yys in rad, zzs depth in m; mom = (Mtt,Mpp,Mrr,
 * Mtp,Mtr,Mpr); so = nt samples of the source time function.  Replaces the
 * source state of SOURCE/ses3d_input.f90:201-230,405-437. */
int ses3d_hip_set_source(ses3d_hip_context *ctx, int32_t ev, int32_t source_type, float xxs,
                         float yys, float zzs, const float *mom, const float *so);

/* Receivers of event ev: rec = (nr,3) rows (colat rad, lon rad, depth m).
 * Replaces recloc_global (SOURCE/ses3d_input.f90:531-554). */
int ses3d_hip_set_receivers(ses3d_hip_context *ctx, int32_t ev, int32_t nr, const float *rec);

/* Adjoint point sources of event ev: loc = (n,3) rows (rad, rad, m);
 * stf = (n, nt, 3) rows as stored in ADJOINT/<ev>/ad_src_<k> (already time
 * reversed).  Replaces SOURCE/ses3d_input.f90:584-617 and
 * SOURCE/ses3d_init.f90:445-500. */
int ses3d_hip_set_adjoint_sources(ses3d_hip_context *ctx, int32_t ev, int32_t n, const float *loc,
                                  const float *stf);

/* Initialise: material coefficients, GLL tables, geometry, mass matrix,
 * receiver/source location, absorbing profiles, zero all wave fields.
 * Replaces SOURCE/ses3d_init.f90:7-729. */
int ses3d_hip_init(ses3d_hip_context *ctx);

/* Advance all events by n_steps iterations it_begin .. it_begin+n_steps-1 of
 * the reference time loop (SOURCE/ses3d_main.f90:95-178): the time step
 * (SOURCE/ses3d_evolution.f90:7-373 with the sources of
 * SOURCE/ses3d_miscellaneous.f90:11-150), seismogram recording
 * (SOURCE/ses3d_output.f90:456-497) in forward modes, forward-field storing
 * in HBM when adjoint_flag==1 (replaces ses3d_store, SOURCE/ses3d_output.f90:5-164)
 * and kernel accumulation when adjoint_flag==2 (SOURCE/ses3d_grad.f90:7-89).
 * Stream-ordered; call ses3d_hip_synchronize before reading results. */
int ses3d_hip_run(ses3d_hip_context *ctx, int32_t it_begin, int32_t n_steps);

/* Wait for all work of the context. */
int ses3d_hip_synchronize(ses3d_hip_context *ctx);

/* Number of receivers located in the box for event ev (`nr` after
 * SOURCE/ses3d_init.f90:292-350). */
int ses3d_hip_get_nr(ses3d_hip_context *ctx, int32_t ev, int32_t *nr);

/* Index (0-based, into the rows given to set_receivers) of each located
 * receiver, in the reference's order (station_name_local). */
int ses3d_hip_get_receiver_index(ses3d_hip_context *ctx, int32_t ev, int32_t *idx);

/* Seismograms of event ev: out = (3, nr, nt) float32 velocity, components
 * theta, phi, r (seismogram_x/y/z, SOURCE/ses3d_modules.f90:196-198). */
int ses3d_hip_get_seismograms(ses3d_hip_context *ctx, int32_t ev, float *out);

/* Named wave field of event ev in reference layout, float32:
 *   "vx","vy","vz"        element-local copy of the current global velocity,
 *                         i.e. what the next step gathers.  The reference's
 *                         snapshot of step it (SOURCE/ses3d_output.f90:304-306)
 *                         holds the velocity gathered at the start of step it,
 *                         so it equals this field read after step it-1.
 *   "vx_global",...       global velocity (vx_global)
 *   "dxux","dyuy","dzuz","dxuy","dyux","dxuz","dzux","dyuz","dzuy"
 *   "sxx","syy","szz","sxy","syx","sxz","szx","syz","szy"   (the _pml stresses)
 *   "Mxx","Myy","Mzz","Mxy","Mxz","Myz"  (3 mechanisms; out holds 3 fields)
 *   "MM_global"           mass matrix
 */
int ses3d_hip_get_field(ses3d_hip_context *ctx, int32_t ev, const char *name, float *out);

/* Frechet kernels of event ev, reference layout: grad_rho, grad_cp,
 * grad_csh, grad_csv (SOURCE/ses3d_grad.f90:77-83) and, when is_diss==1,
 * grad_Q_mu (SOURCE/ses3d_grad_Q.f90:111-128,151-153); any pointer may be
 * NULL.  Accumulated in float64 by default, or in float32 (the type of the
 * reference's grad_* arrays, widened exactly on output) after
 * ses3d_hip_set_kernel_precision.  Replaces the kernel files of
 * SOURCE/ses3d_output.f90:318-343. */
int ses3d_hip_get_kernels(ses3d_hip_context *ctx, int32_t ev, double *grad_rho, double *grad_cp,
                          double *grad_csh, double *grad_csv, double *grad_Q_mu);

/* ses3d_hip_get_kernels into device memory: the same float64 values in the
 * same reference layout, written on the context's stream to device buffers
 * of NPTS doubles (any may be NULL); returns once they are written. */
int ses3d_hip_get_kernels_device(ses3d_hip_context *ctx, int32_t ev, double *grad_rho, double *grad_cp,
                                 double *grad_csh, double *grad_csv, double *grad_Q_mu);

/* Forward-field history (adjoint_flag==1 writes it, adjoint_flag==2 reads
 * it).  A stored step holds 9 float32 fields in reference layout, in the
 * order vx vy vz exx eyy ezz exy exz eyz -- the nine records ses3d_store
 * writes per step (SOURCE/ses3d_main.f90:142-166).  get copies stored step
 * `step` (1-based forward step) out; put uploads one for an adjoint run.
 * saving_vector(it)=it if step it is stored, else 0
 * (SOURCE/ses3d_output.f90:379-392). */
int ses3d_hip_get_history(ses3d_hip_context *ctx, int32_t ev, int32_t step, float *out9);
int ses3d_hip_put_history(ses3d_hip_context *ctx, int32_t ev, int32_t step, const float *in9);
int ses3d_hip_get_saving_vector(ses3d_hip_context *ctx, int32_t *sv);

/* Polynomial degree `deg` (0..4) of the stored forward fields: the `deg`
 * argument of ses3d_store / ses3d_restore (SOURCE/ses3d_output.f90:5-267),
 * which the reference's call sites fix at lpd=4 (SOURCE/ses3d_main.f90:148-158,
 * SOURCE/ses3d_grad.f90:33-45).  With deg < 4 each element's field is stored
 * as its values at the (deg+1)^3 degree-deg knots (ses3d_output.f90:36-130)
 * and interpolated back to the GLL nodes where the kernels read it
 * (:242-265), with the reference's float32 arithmetic and summation order:
 * (deg+1)^3/125 of the memory and traffic of the stored fields, lossy.  A
 * stored step then holds, per field, the (0:nx,0:ny,0:nz,0:deg,0:deg,0:deg)
 * array ses3d_store writes; get/put_history use that layout.  Set it before
 * init, equal in the forward and the adjoint context; changing it releases
 * the stored steps.  history_size gives the floats of one stored step. */
int ses3d_hip_set_history_degree(ses3d_hip_context *ctx, int32_t deg);
int ses3d_hip_history_size(ses3d_hip_context *ctx, int64_t *n);

/* Type of the kernel accumulators of an adjoint_flag==2 context: bits = 64
 * (default, float64) or 32 (float32, as the reference's grad_* arrays,
 * SOURCE/ses3d_modules.f90:221-222, summed in the order of
 * SOURCE/ses3d_grad.f90:77-83 and SOURCE/ses3d_grad_Q.f90:111-128, so the
 * kernels equal the reference's bit for bit; 40 fewer bytes per point per
 * sampled adjoint step).  Takes effect at the next init, which zeroes the
 * kernels. */
int ses3d_hip_set_kernel_precision(ses3d_hip_context *ctx, int32_t bits);

/* Gradient without disk history: for an adjoint_flag==2 context, take the
 * forward wave field from HBM checkpoints.  ses3d_hip_forward_checkpointed
 * runs the forward problem of context fwd (adjoint_flag==1) for all nt
 * steps keeping a restart state every k_ckpt steps in HBM instead of every
 * sampled step (and the sampled steps of the last segment);
 * ses3d_hip_adjoint_checkpointed then runs the adjoint of context adj,
 * replaying each earlier forward segment from its checkpoint into an HBM
 * buffer of sampled steps just before the adjoint consumes it.  The
 * kernels equal those of the stored-history path bit for bit.  Replaces the
 * disk round trip of ses3d_store/ses3d_restore (SOURCE/ses3d_output.f90:5-267,
 * SOURCE/ses3d_grad.f90:33-45). */
int ses3d_hip_forward_checkpointed(ses3d_hip_context *fwd, int32_t k_ckpt);
int ses3d_hip_adjoint_checkpointed(ses3d_hip_context *adj, ses3d_hip_context *fwd);

/* Memory for replay: the next forward_checkpointed keeps the sampled steps of
 * its last `segments` segments (default 1) in HBM, so the adjoint replays
 * only the earlier ones: (segments-1)*k_ckpt fewer forward steps for a
 * sampled-step buffer `segments` times larger (9 floats per GLL point per
 * sampled step).  Kernels are identical for every value.  No counterpart in
 * the reference, which keeps every sampled step on disk
 * (SOURCE/ses3d_output.f90:5-267). */
int ses3d_hip_set_checkpoint_keep(ses3d_hip_context *fwd, int32_t segments);

/* Set-up half of ses3d_hip_forward_checkpointed: allocate (and clear) the
 * restart states for interval k_ckpt and the sampled-step buffer for the
 * current kept-segment count, so that the next forward_checkpointed with the
 * same k_ckpt allocates nothing.  Call after init; a checkpointed forward
 * run before it is discarded (adjoint_checkpointed then refuses until the
 * next forward_checkpointed).  Optional: forward_checkpointed allocates what
 * is missing itself.  Replaces the reference's up-front sizing of its module
 * arrays (SOURCE/ses3d_modules.f90:33-233), which allocates nothing per run. */
int ses3d_hip_prepare_checkpoints(ses3d_hip_context *fwd, int32_t k_ckpt);

/* Launch structure: the step number lives in HBM and each step is a fixed
 * sequence of kernels, so `steps` consecutive steps are captured once into a
 * hipGraph and replayed (default 16; 0 launches every kernel directly).  No
 * counterpart in the reference (it has no launches); results are identical
 * either way. */
int ses3d_hip_set_graph_steps(ses3d_hip_context *ctx, int32_t steps);

/* Implementation state after init (no counterpart in the reference):
 * out[0] reciprocal quotients by tau_p in use, out[1] by 3 in use,
 * out[2..5] float32 numerators failing the exhaustive check for tau_p[0..2]
 * and 3, out[6] anisotropic terms in use, out[7] steps per captured graph. */
int ses3d_hip_get_info(ses3d_hip_context *ctx, int64_t *out, int32_t n);

/* Device memory held by the context [bytes]. */
int ses3d_hip_memory_bytes(ses3d_hip_context *ctx, int64_t *bytes);

/* Timing of the hot kernels over the last ses3d_hip_run call, measured with
 * HIP events on the context's stream: total ms, and the summed ms of the
 * element kernel and the node kernel (zero unless profiling is enabled). */
int ses3d_hip_set_profiling(ses3d_hip_context *ctx, int32_t on);
int ses3d_hip_get_timing(ses3d_hip_context *ctx, double *total_ms, double *element_ms,
                         double *node_ms, int32_t *n_element_launches);

/* Launch schedule (no counterpart in the reference, which has no launches):
 * sched[0] elements per element-kernel workgroup (1, 2, 4), sched[1] lanes per
 * node-kernel workgroup (128, 256, 512), sched[2] / sched[3] XCD-aware
 * workgroup order of the element / node kernel (0, 1).  Every schedule gives
 * bit-identical results.  ses3d_hip_autotune advances steps it_begin ..
 * it_begin+n_steps-1 exactly as ses3d_hip_run would (n_steps >= 6), timing 6
 * trial schedules over the first 6*(n_steps/6) of them, and keeps the fastest
 * element and node schedules for later runs; get_schedule returns the current
 * schedule and, when trial_ms is not NULL, the 6 trials' fastest element and
 * node kernel times (12 floats, -1 before any autotune). */
int ses3d_hip_set_schedule(ses3d_hip_context *ctx, const int32_t *sched);
int ses3d_hip_get_schedule(ses3d_hip_context *ctx, int32_t *sched, float *trial_ms);

/* Cache policy of the element kernel's state streams (no counterpart in the
 * reference): mode -1 (default) streams the displacement gradient, split
 * stresses, memory variables, coefficients and sampled fields with
 * non-temporal loads and stores when they exceed the 256 MiB last-level cache
 * and keeps them cached otherwise; 0 always cached; 1 always non-temporal.
 * Both variants give bit-identical results.  get_streaming returns the mode
 * and, in *in_use, whether the non-temporal variant runs (either pointer may
 * be NULL). */
int ses3d_hip_set_streaming(ses3d_hip_context *ctx, int32_t mode);
int ses3d_hip_get_streaming(ses3d_hip_context *ctx, int32_t *mode, int32_t *in_use);
int ses3d_hip_autotune(ses3d_hip_context *ctx, int32_t it_begin, int32_t n_steps);

/* Groups: several initialised contexts (sections of any size, each with its
 * own events, receivers and sources; all elastic or all anelastic; one
 * adjoint_flag) advanced together by one element-kernel launch and one
 * node-kernel launch per step, captured into hipGraphs like a single context.
 * No counterpart in the reference, where each section is a separate MPI job
 * (SOURCE/ses3d_main.f90:95-178 per event).  group_run(g, it_begin, n) gives
 * every member exactly the results of ses3d_hip_run(member, it_begin, n);
 * members keep their own outputs (seismograms, fields, kernels) and may still
 * be run alone.  The group has its own stream and schedule; work issued on a
 * member's stream after a group call is ordered after it.  Members must
 * outlive the group. */
typedef struct ses3d_hip_group ses3d_hip_group;
int ses3d_hip_group_create(ses3d_hip_context *const *ctxs, int32_t n, ses3d_hip_group **out);
void ses3d_hip_group_destroy(ses3d_hip_group *grp);
int ses3d_hip_group_run(ses3d_hip_group *grp, int32_t it_begin, int32_t n_steps);
int ses3d_hip_group_autotune(ses3d_hip_group *grp, int32_t it_begin, int32_t n_steps);
int ses3d_hip_group_synchronize(ses3d_hip_group *grp);
int ses3d_hip_group_set_schedule(ses3d_hip_group *grp, const int32_t *sched);
int ses3d_hip_group_get_schedule(ses3d_hip_group *grp, int32_t *sched, float *trial_ms);
int ses3d_hip_group_set_streaming(ses3d_hip_group *grp, int32_t mode);   /* by the members' summed state */
int ses3d_hip_group_get_streaming(ses3d_hip_group *grp, int32_t *mode, int32_t *in_use);
int ses3d_hip_group_set_graph_steps(ses3d_hip_group *grp, int32_t steps);
int ses3d_hip_group_set_profiling(ses3d_hip_group *grp, int32_t on);
int ses3d_hip_group_get_timing(ses3d_hip_group *grp, double *total_ms, double *element_ms,
                               double *node_ms, int32_t *n_element_launches);

/* Projection of a GLL field onto a block grid, as TOOLS/SOURCE/project_kernel.f90
 * (:575-715) projects sensitivity kernels for one subvolume: block edges
 * bx_deg[nbx] (colatitude, degrees), by_deg[nby] (longitude, degrees),
 * bz_km[nbz] (radius, km), increasing; each block value is the sum over the
 * elements it overlaps of field * int L_l * int L_m * int L_n (GLL Lagrange
 * polynomials integrated between the block's limits in each element's local
 * coordinates) divided by the first block's width product in degrees^2 km.
 * field: float32 in the reference layout (0:nx,0:ny,0:nz,0:4,0:4,0:4); out:
 * (nbx-1)(nby-1)(nbz-1) floats, colatitude slowest and radius fastest (the
 * tool's output order). */
int ses3d_hip_project(ses3d_hip_context *ctx, const float *field, int32_t nbx, const float *bx_deg,
                      int32_t nby, const float *by_deg, int32_t nbz, const float *bz_km, float *out);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under graph capture, tight serving loops, harnesses that forbid
 * hipMalloc): ses3d_hip_project with the same arguments and results, but the
 * field, block tables and output are staged in a 64 MiB device array the
 * library holds on each GPU until the process ends (looked up once per device
 * when a context is created) instead of a buffer allocated per call, so it
 * calls no hipMalloc or hipFree.  Projections needing more than 64 MiB (the
 * field's 4 bytes per GLL point, about 16 million points, plus tables and
 * output) fail with SES3D_HIP_EINVAL.  Every call on one GPU shares that
 * array: calls that overlap on one GPU (other host threads, other contexts
 * on the same device) race and must be serialised by the caller. */
int ses3d_hip_project_noalloc(ses3d_hip_context *ctx, const float *field, int32_t nbx, const float *bx_deg,
                              int32_t nby, const float *by_deg, int32_t nbz, const float *bz_km, float *out);

#ifdef __cplusplus
}
#endif
#endif /* SES3D_HIP_H */
