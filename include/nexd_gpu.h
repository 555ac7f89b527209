/*
 * nexd_gpu.h -- C-ABI of the MI355X time-stepping library libnexd_gpu.so.
 *
 * The library replaces the body of the reference's time loop
 * (src/timeloopMod.f90:357-589) for a single partition that holds the whole
 * mesh on one GPU.  The caller keeps the reference's set-up
 * (src/timeloopMod.f90:135-356) and hands the library the arrays that set-up
 * holds, in the reference's own layout: Fortran column-major, 1-based index
 * values, REAL(8).  Every pointer argument is a plain host pointer; the
 * library copies (and re-lays out) what it needs to device memory and owns
 * that memory.  No call keeps a host pointer after it returns.
 *
 * Conventions
 *   - every function returns 0 on success and a negative NEXD_E* code on
 *     failure; nexd_last_error() then describes the failure;
 *   - one context per host thread; calls on one context are not
 *     thread-safe; independent contexts are;
 *   - nexd_run enqueues work on the context's HIP stream and waits for it;
 *     nexd_get_* return results that are complete.
 *
 * Shapes below are the Fortran shapes; "nglob" = nelem*Np.
 */
#ifndef NEXD_GPU_H
#define NEXD_GPU_H

#ifdef __cplusplus
extern "C" {
#endif

#define NEXD_OK 0
#define NEXD_E_ARG (-1)        /* invalid argument or inconsistent sizes */
#define NEXD_E_HIP (-2)        /* HIP runtime failure */
#define NEXD_E_STATE (-3)      /* call out of order (e.g. run before upload) */
#define NEXD_E_UNSUPPORTED (-4)/* feature outside what the library implements */

/* timeint values: the reference's 1 (Euler), 2 (TVD RK2), 3 (TVD RK3)
 * (src/timeloopMod.f90:230-248), and 4, an extension without a reference
 * counterpart: the five-stage fourth-order low-storage (2N-storage) RK of
 * Carpenter and Kennedy (NASA TM-109112, 1994) */
#define NEXD_EULER 1
#define NEXD_RK2 2
#define NEXD_RK3 3
#define NEXD_LSRK4 4

/* compat bits: reproduce reference behaviour bit for bit */
#define NEXD_COMPAT_RK3_FLOAT_WEIGHTS 1  /* 1/3, 2/3 rounded to float (src/timeloopMod.f90:129-133) */
#define NEXD_COMPAT_PML_LAST_ELEMENT 2   /* PML ADE uses the flux of the last element
                                            (src/rungekuttaMod.f90:57-59,80-82) */
#define NEXD_COMPAT_ALL 3

typedef struct nexd_config {
    int nelem;             /* elements of the (single) partition */
    int order;             /* polynomial order N, 1..5 (src/constants.h:26) */
    int nsrc;              /* point sources in this partition */
    int nrec;              /* receivers in this partition */
    int nt;                /* time steps (par%nt) */
    int timeint;           /* NEXD_EULER / NEXD_RK2 / NEXD_RK3 (par%timeint) / NEXD_LSRK4 */
    int attenuation;       /* par%attenuation: 3 standard linear solids */
    int pml;               /* par%set_pml */
    int compat;            /* NEXD_COMPAT_* bits */
    int track_displacement;/* keep ux,uy,uz for every node (movie/log); receivers
                              are exact either way */
    int energy;            /* compute the per-step energy of PML runs
                              (src/timeloopMod.f90:508-516) */
    int device;            /* HIP device ordinal */
    double dt;             /* deltat = dtfactor*cfl (src/timeloopMod.f90:199) */
} nexd_config;

typedef struct nexd_ctx nexd_ctx;

/* Creates a context and allocates device state for the configuration.
 * Replaces the allocation block src/timeloopMod.f90:162-170.  The kernels
 * address one field plane (nelem*Np doubles) and the per-element face
 * arrays (nelem*max(4*NpF, 36) doubles) with 32-bit byte offsets: a mesh
 * whose arrays exceed 4 GiB gets NEXD_E_UNSUPPORTED. */
int nexd_create(const nexd_config *cfg, nexd_ctx **out);

/* Reference-element operators: Dr, Ds, Dt, mass (Np,Np); lift (Np,4*NpF);
 * fmask (NpF,4).  mesh%Dr/Ds/Dt/mass/lift/fmask of typeMod (src/typeMod.f90:74-78). */
int nexd_upload_operators(nexd_ctx *ctx, const double *dr, const double *ds, const double *dt,
                          const double *lift, const double *mass, const int *fmask);

/* Mesh: ibool (Np,nelem) must be (ie-1)*Np+j as the mesher writes it
 * (src/meshMod.f90:796-811); rx..tz, jacobian (nglob); fscale (4*NpF,nelem);
 * neighbor, face (4,nelem); ibt, ibn (NpF,4,nelem); nnormal, stangential,
 * ttangential (3,4,nelem).  Partition interfaces (neighbor == -1) are not
 * supported: the whole mesh is one partition. */
int nexd_upload_mesh(nexd_ctx *ctx, const int *ibool,
                     const double *rx, const double *ry, const double *rz,
                     const double *sx, const double *sy, const double *sz,
                     const double *tx, const double *ty, const double *tz,
                     const double *jacobian, const double *fscale,
                     const int *neighbor, const int *face, const int *ibt, const int *ibn,
                     const double *nnormal, const double *stangential, const double *ttangential);

/* Materials (nelem each): mesh%vpu, vsu, rho, lambdau, muu (flux matrices and
 * elasticfluxvar, src/timeloopMod.f90:257-273) and mesh%mu, lambda (energy). */
int nexd_upload_material(nexd_ctx *ctx, const double *vpu, const double *vsu, const double *rho,
                         const double *lambdau, const double *muu, const double *mu,
                         const double *lambda);

/* Attenuation: mesh%wl (nMB) and anelasticvar (3*nMB*nelem)
 * (src/timeloopMod.f90:274-282). */
int nexd_upload_anelastic(nexd_ctx *ctx, const double *wl, const double *anelasticvar);

/* PML: pmlcheck (nelem, 0/1 as C int) and the per-node profiles ddx..kz (nglob)
 * of PMLPreCalculation (src/pmlMod.f90:90-165). */
int nexd_upload_pml(nexd_ctx *ctx, const int *pmlcheck,
                    const double *ddx, const double *ddy, const double *ddz,
                    const double *alphax, const double *alphay, const double *alphaz,
                    const double *kx, const double *ky, const double *kz);

/* Sources: src%srcelem (nsrc, 1-based), src%srctype (nsrc), srcArray
 * (Np,3,nsrc,nt) and srcArrayM (Np,6,nsrc,nt) as built at
 * src/timeloopMod.f90:327-356.  Either array may be NULL if no source of
 * that type exists. */
int nexd_upload_sources(nexd_ctx *ctx, const int *srcelem, const int *srctype,
                        const double *src_array, const double *src_array_m);

/* Receivers: rec%recelem (nrec, 1-based), recInt (Np,nrec)
 * (src/timeloopMod.f90:292-300). */
int nexd_upload_receivers(nexd_ctx *ctx, const int *recelem, const double *rec_int);

/* Resets the wavefield to the reference's initial state (every value the
 * default-REAL literal 1e-24, i.e. 1.0000000195414813e-24 in REAL(8);
 * PML auxiliaries 0; src/timeloopMod.f90:184-195,252-254) and the step
 * counter to 1. */
int nexd_reset(nexd_ctx *ctx);

/* Advances steps it_begin..it_end (1-based, inclusive, it_begin must be the
 * next step).  Records receiver rows it_begin..it_end (state before each
 * step, src/timeloopMod.f90:358-373).  If energy is non-NULL and the
 * context computes energy, energy[k] receives all_energy(it_begin+k). */
int nexd_run(nexd_ctx *ctx, int it_begin, int it_end, double *energy);

/* No counterpart in the reference; for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  nexd_run without any allocation: the step kernels run
 * directly on the context stream (no graph is instantiated) and keep their
 * scratch -- the energy block partials and the last element's fluxes --
 * in a static per-GPU array of 2^20 doubles (8 MiB) instead of the
 * context's buffers.  Needs a context that was reset (nexd_reset or
 * nexd_set_state; NEXD_E_STATE otherwise); contexts whose scratch exceeds
 * the array (about 4 million elements) get NEXD_E_UNSUPPORTED.  All calls on
 * one GPU share the array: calls that overlap in time on one GPU race.
 * Same arguments and results as nexd_run. */
int nexd_run_noalloc(nexd_ctx *ctx, int it_begin, int it_end, double *energy);

/* PML switch used by the caller's STA/LTA trigger (src/timeloopMod.f90:532-539). */
int nexd_set_pml_active(nexd_ctx *ctx, int active);

/* Seismograms plotux..plotaz, each (nrec,nt), rows recorded so far. */
int nexd_get_seismograms(nexd_ctx *ctx, double *ux, double *uy, double *uz,
                         double *vx, double *vy, double *vz,
                         double *ax, double *ay, double *az);

/* Current wavefield q (nglob,9); ux,uy,uz (nglob) if the context tracks
 * displacement (else NEXD_E_STATE for non-NULL ux). */
int nexd_get_state(nexd_ctx *ctx, double *q, double *ux, double *uy, double *uz);

/* Start from a given state instead of the reference's initial one
 * (src/timeloopMod.f90:184-195,252-254; the reference itself has no
 * restart): q (nglob,9), theta (nglob,6,nMB) with
 * attenuation, fprime/gprime/hprime (nglob,9) with PML, and the displacement
 * ux,uy,uz (nglob) after `it` completed steps (0 <= it < nt).  NULL leaves a
 * field as it is (nexd_reset's initial state on a fresh context).  The next
 * nexd_run must begin at it+1.  Receiver row it+1 is sampled from the given
 * state; its acceleration (the previous step's last-stage RHS) is not part
 * of the state and stays at the initial value. */
int nexd_set_state(nexd_ctx *ctx, int it, const double *q, const double *theta, const double *fprime,
                   const double *gprime, const double *hprime, const double *ux, const double *uy,
                   const double *uz);

/* Device time of the last nexd_run in milliseconds (HIP events). */
int nexd_last_run_ms(nexd_ctx *ctx, double *ms);

/* Measurement only (no reference counterpart; the reference times whole
 * steps with system_clock, src/timeloopMod.f90:544-566): launches the stage
 * kernel of one element list for every RK stage in turn, nrep rounds, on the
 * context stream between HIP events, and returns the mean duration of one
 * launch (averaged over the stages, as in the time loop) and the number of
 * elements it advances.  list 0: elements outside the PML (all elements
 * without PML), list 1: PML elements.  The stages overwrite the state: call
 * nexd_reset before running again.  Needs a multi-stage scheme. */
int nexd_time_stage_kernel(nexd_ctx *ctx, int list, int nrep, double *avg_ms, int *nelem_launch);

void nexd_destroy(nexd_ctx *ctx);

/* Description of the last failure on ctx, or of the last nexd_create
 * failure when ctx is NULL. */
const char *nexd_last_error(const nexd_ctx *ctx);

/* ==========================================================================
 * Batched shots (no reference counterpart: the reference runs one source
 * configuration per job).  Many shots on one mesh advance together in every
 * stage launch; geometry, operators and receivers are shared, materials,
 * PML profiles and sources are per shot.  Arithmetic is FP32 (the fast path)
 * or FP64.  Sources are given factored: a spatial operand per source and
 * field times the source time function of the step (the reference's
 * srcArray(:,c,i,it) = spatial(:,c,i)*stf(it,i), src/timeloopMod.f90:327-356).
 * ========================================================================== */
#define NEXD_FP64 0
#define NEXD_FP32 1

typedef struct nexd_batch_config {
    int nelem, order, nshot;
    int nsrc;             /* sources per shot */
    int nrec;             /* receivers, shared by all shots */
    int nt, timeint, attenuation, pml, compat;
    int precision;        /* NEXD_FP32 or NEXD_FP64 */
    int shared_material;  /* 1: one material/PML set for every shot */
    int compensated;      /* 1: compensated RK stage sum (state kept as hi+lo
                             pairs of the working precision) */
    int energy;           /* with pml: each shot's energy after every step
                             (src/timeloopMod.f90:508-516), nexd_batch_get_energy */
    int device;
    double dt;
} nexd_batch_config;

typedef struct nexd_batch nexd_batch;

/* The same size limit as nexd_create, per shot and in the batch's
 * precision (a field plane of nelem*Np reals below 4 GiB). */
int nexd_batch_create(const nexd_batch_config *cfg, nexd_batch **out);
/* same arguments and layouts as nexd_upload_operators / nexd_upload_mesh */
int nexd_batch_upload_operators(nexd_batch *b, const double *dr, const double *ds, const double *dt,
                                const double *lift, const double *mass, const int *fmask);
int nexd_batch_upload_mesh(nexd_batch *b, const int *ibool,
                           const double *rx, const double *ry, const double *rz,
                           const double *sx, const double *sy, const double *sz,
                           const double *tx, const double *ty, const double *tz,
                           const double *jacobian, const double *fscale,
                           const int *neighbor, const int *face, const int *ibt, const int *ibn,
                           const double *nnormal, const double *stangential, const double *ttangential);
/* per shot (shot 0..nshot-1; with shared_material only shot 0); the same
 * arrays as nexd_upload_material (mu, lambda enter only the energy) */
int nexd_batch_upload_material(nexd_batch *b, int shot, const double *vpu, const double *vsu,
                               const double *rho, const double *lambdau, const double *muu,
                               const double *mu, const double *lambda);
int nexd_batch_upload_anelastic(nexd_batch *b, int shot, const double *wl, const double *anelasticvar);
/* pmlcheck is a mesh property and must be the same for every shot */
int nexd_batch_upload_pml(nexd_batch *b, int shot, const int *pmlcheck,
                          const double *ddx, const double *ddy, const double *ddz,
                          const double *alphax, const double *alphay, const double *alphaz,
                          const double *kx, const double *ky, const double *kz);
/* srcelem (nsrc, 1-based), srctype (nsrc), spatial (Np,6,nsrc): columns 1..3
 * of a force source, 1..6 of a moment-tensor source; stf (nt,nsrc) */
int nexd_batch_upload_sources(nexd_batch *b, int shot, const int *srcelem, const int *srctype,
                              const double *spatial, const double *stf);
int nexd_batch_upload_receivers(nexd_batch *b, const int *recelem, const double *rec_int);
int nexd_batch_reset(nexd_batch *b);
int nexd_batch_run(nexd_batch *b, int it_begin, int it_end);
/* seismograms of one shot, each (nrec,nt) in double */
int nexd_batch_get_seismograms(nexd_batch *b, int shot, double *ux, double *uy, double *uz,
                               double *vx, double *vy, double *vz, double *ax, double *ay, double *az);
/* wavefield q (nglob,9) of one shot after the steps run so far, in double */
int nexd_batch_get_state(nexd_batch *b, int shot, double *q);
/* energy of one shot after every step recorded so far (nt values; config
 * energy and pml), as nexd_run returns it for one context */
int nexd_batch_get_energy(nexd_batch *b, int shot, double *energy);
/* one shot's PML switch for its STA/LTA trigger (src/timeloopMod.f90:532-539),
 * as nexd_set_pml_active for one context; the other shots keep theirs */
int nexd_batch_set_pml_active(nexd_batch *b, int shot, int active);
/* Block-shape autotuner (no reference counterpart).  FP32 batches with a
 * TVD scheme have two shapes of the stage launches: 0 sizes a block for
 * four blocks per compute unit, 1 for two (twice the LDS, so more elements
 * per block, and twice the registers per thread).  nexd_batch_autotune times
 * each on the batch's own element lists (nrep rounds of every RK stage per
 * list) and keeps the fastest; *tile receives it and ms[0..1] (may be NULL)
 * the mean stage time of each shape.  Results are cached per kind of batch
 * for the process.  The timing overwrites the state: the batch is reset
 * afterwards.  Other batches have one shape (tile 0). */
int nexd_batch_autotune(nexd_batch *b, int nrep, int *tile, double *ms);
int nexd_batch_set_tile(nexd_batch *b, int tile);
/* A step per element (nelem values, the host's double), for instances of
 * different time step packed into one mesh: each element advances by its
 * own step (no reference counterpart; the reference has one step per run). */
int nexd_batch_upload_element_dt(nexd_batch *b, const double *dt_elem);
int nexd_batch_last_run_ms(nexd_batch *b, double *ms);
/* measurement: mean duration of one stage launch of element list 0/1,
 * averaged over the RK stages (as nexd_time_stage_kernel); overwrites the
 * state */
int nexd_batch_time_stage_kernel(nexd_batch *b, int list, int nrep, double *avg_ms, int *nelem_launch);
void nexd_batch_destroy(nexd_batch *b);
const char *nexd_batch_last_error(const nexd_batch *b);

/* ==========================================================================
 * Groups of batches (no reference counterpart): independent instances of
 * different meshes, sizes, orders and precisions advanced together.  Each
 * member's step graph becomes a child node of one parent HIP graph without
 * dependencies between members, launched once per step, so small instances
 * overlap on the device instead of running one after another.
 * Members must share nt and the device and stand at the same step; they
 * keep their own state, seismograms and getters.  Do not upload to a member
 * or run it on its own while it is in a group; destroy the group first.
 * ========================================================================== */
typedef struct nexd_group nexd_group;

int nexd_group_create(nexd_batch *const *members, int n, nexd_group **out);
/* advance every member from step it_begin to it_end (1-based, inclusive) */
int nexd_group_run(nexd_group *g, int it_begin, int it_end);
int nexd_group_last_run_ms(nexd_group *g, double *ms);
void nexd_group_destroy(nexd_group *g);
const char *nexd_group_last_error(const nexd_group *g);

/* ==========================================================================
 * Multilayer perceptron (no reference counterpart): the learned
 * source-to-seismogram surrogate, trained and evaluated by the library's own
 * kernels.  widths[0..nlayers]: input, hidden and output widths; hidden
 * layers apply GELU (erf form), the output layer is linear.  Weights are
 * FP32 master copies with BF16 operands for the matrix-core products (FP32
 * accumulation); Adam updates.  Arrays are host, C row-major: W [out][in],
 * b [out], x [B][in], y and targets [B][out].  Batches B are multiples of 64
 * up to max_batch; training needs an output width that is a multiple of 64.
 * ========================================================================== */
typedef struct nexd_mlp nexd_mlp;

int nexd_mlp_create(int nlayers, const int *widths, int max_batch, int device, nexd_mlp **out);
/* sets one layer's weights and biases (and restarts the optimiser) */
int nexd_mlp_set_params(nexd_mlp *m, int layer, const float *W, const float *b);
/* which 0: weights and biases; 1: their gradients from the last
 * nexd_mlp_gradients or training step.  W or b may be NULL. */
int nexd_mlp_get_params(nexd_mlp *m, int layer, int which, float *W, float *b);
int nexd_mlp_forward(nexd_mlp *m, const float *x, int B, float *y);
/* the training set, kept on the device: x [n][in], targets [n][out] */
int nexd_mlp_upload_dataset(nexd_mlp *m, const float *x, const float *t, int n);
/* mean-squared-error loss and all gradients of the dataset rows idx[0..B),
 * without an update */
int nexd_mlp_gradients(nexd_mlp *m, const int *idx, int B, float *loss);
/* nsteps Adam steps over consecutive batches idx[s*B .. s*B+B); losses[s]
 * (may be NULL) is the loss of step s before its update */
int nexd_mlp_train(nexd_mlp *m, const int *idx, int B, int nsteps, float lr, float beta1, float beta2, float eps,
                   float *losses);
/* device time of the last nexd_mlp_train in milliseconds */
int nexd_mlp_last_run_ms(nexd_mlp *m, double *ms);
void nexd_mlp_destroy(nexd_mlp *m);
const char *nexd_mlp_last_error(const nexd_mlp *m);

#ifdef __cplusplus
}
#endif
#endif
