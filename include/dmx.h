/*
 * dmx.h -- C ABI of the MI355X-native DUMSES-Hybrid MHD step (libdmx.so).
 *
 * One context owns one device-resident ghosted domain (nghost = 3) and runs the
 * reference's second-order Godunov MHD step on it: MUSCL-Hancock slopes and
 * trace, 1D (HLLD/HLL/LLF/upwind) face fluxes, 2D (HLLD/LLF/upwind/HLLA) edge
 * EMFs with constrained transport, CFL timestep, ghost fills (periodic,
 * zero-gradient, shearing-periodic x) and FARGO orbital advection.
 *
 * The entry points replace the reference's external Fortran subroutines
 * (no explicit interface, all state in module globals), which a Fortran
 * iso_c_binding shim forwards here:
 *
 *   dmx_compute_dt  <- compute_dt(dt)           src/subroutines/physics.f90:18
 *   dmx_godunov     <- godunov                  src/subroutines/scheme.f90:104
 *                      (includes update + constrained_transport,
 *                       src/subroutines/scheme.f90:1764 and :1940)
 *   dmx_fargo       <- fargo_update             src/subroutines/fargo.f90:17
 *   dmx_boundary    <- boundary                 src/subroutines/boundary.f90:18
 *   dmx_upload      <- the host->device copy of uin at the data region
 *                      (src/dumses.f90:59, after init / restart_run)
 *   dmx_download    <- "!$acc update host(uin)" before output/history/special
 *                      (src/dumses.f90:62,122,133,138,144)
 *   dmx_history     <- the FP64 volume sums of history
 *                      (src/subroutines/history.f90:62-118)
 *
 * Conventions (the reference has none: it stops or ignores errors):
 *   - every function returns 0 on success and a negative DMX_E* code on error;
 *   - dmx_last_error() returns a message for the last error of the thread;
 *   - calls are synchronous at the ABI (internally asynchronous on one HIP
 *     stream); a context is not thread-safe, as the reference's module state.
 *
 * Arrays crossing the ABI use the reference layout of uin:
 *   uin(iu1:iu2, ju1:ju2, ku1:ku2, 1:11), Fortran column-major (i fastest,
 *   variable slowest), iu1 = -2, iu2 = nx+3 (src/subroutines/init.f90:34-42,
 *   :357), variables rho, rho*u, rho*v, rho*w, E, Bx_L, By_L, Bz_L, Bx_R, By_R,
 *   Bz_R (src/modules/commons.f90:72-80).
 */
#ifndef DMX_H
#define DMX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Riemann solver ids (src/modules/commons.f90:104-111) */
#define DMX_LLF     2
#define DMX_HLL     3
#define DMX_HLLD    4
#define DMX_UPWIND  5
#define DMX_HLLA    8

/* boundary types (the bdtype{x,y,z} strings, src/modules/commons.f90:128-130) */
#define DMX_BC_PERIODIC       0
#define DMX_BC_ZERO_GRADIENT  1
#define DMX_BC_SHEARINGBOX    2   /* x only */

/* arithmetic of the step */
#define DMX_FP64  0   /* everything in float64 (the reference's arithmetic) */
#define DMX_FP32  1   /* bulk float32; float64 E, face B and CT, pressure
                         recovery, the Riemann solvers' energy chain, dt
                         and history sums */

/* error codes */
#define DMX_OK          0
#define DMX_E_PARAM    -1   /* invalid parameter (unknown solver, size, bc) */
#define DMX_E_NOMEM    -2   /* device allocation failed */
#define DMX_E_HIP      -3   /* HIP runtime error */
#define DMX_E_NONFINITE -4  /* dt not finite or not positive */
#define DMX_E_STATE    -5   /* call out of order (e.g. step before upload) */

/* Parameters of the reference namelists scheme_params/model_params
 * (src/subroutines/init.f90:153-173, defaults :264-299). */
typedef struct dmx_params {
    int32_t nx, ny, nz;          /* interior cells, each >= 3                 */
    int32_t iso;                 /* 1: isothermal (reference ISO=1 build)     */
    int32_t slope_type;          /* 0 none, 1 minmod, 2 MC, 3 multidim,
                                    4 learned (dmx_set_limiter; no counterpart
                                    in the reference)                         */
    int32_t riemann;             /* DMX_LLF | DMX_HLL | DMX_HLLD | DMX_UPWIND  */
    int32_t riemann2d;           /* DMX_LLF | DMX_HLLD | DMX_UPWIND | DMX_HLLA */
    int32_t bdtype[3];           /* DMX_BC_* per direction                    */
    int32_t fargo;               /* 1: FARGO orbital advection (needs omega0) */
    int32_t precision;           /* DMX_FP64 | DMX_FP32                       */
    double  xmin, xmax, ymin, ymax, zmin, zmax;
    double  gamma, ciso, omega0, courant;
    double  nu, eta;             /* viscosity, resistivity (>= 0; 0 = off)    */
} dmx_params;

typedef struct dmx_ctx dmx_ctx;

/* Creates a context on the current HIP device and allocates the domain.
 * Replaces allocate_workspace (src/subroutines/init.f90:348) for the step. */
int  dmx_create(const dmx_params* p, dmx_ctx** ctx);
void dmx_destroy(dmx_ctx* ctx);

/* Host <-> device copies of the full ghosted uin (reference layout above). */
int  dmx_upload(dmx_ctx* ctx, const double* uin);
int  dmx_download(dmx_ctx* ctx, double* uin);

/* CFL timestep: dt = courant * min 1/(cfx/dx + cfy/dy + cfz/dz) over the
 * interior (src/subroutines/physics.f90:18-132). */
int  dmx_compute_dt(dmx_ctx* ctx, double* dt);

/* CFL timestep with no device allocation.  No counterpart in the reference:
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, serving loops or harnesses that forbid hipMalloc).  The same
 * arithmetic as dmx_compute_dt, on a float64 state the caller already holds
 * in device memory: uin is a DEVICE pointer to the reference layout above
 * (the cell field is the mean of the left and right face fields, variables
 * 6..8 and 9..11); *dt, a DEVICE pointer, receives the timestep.  The two
 * kernels are enqueued on `stream` (a hipStream_t; NULL: the null stream)
 * and the call returns without waiting.  p->precision must be DMX_FP64.
 * The block minima go to a static 16 KiB device array (2048 doubles) that
 * every call on a device shares: any domain size fits (the first kernel
 * strides over the cells with at most 2048 blocks), and calls that overlap
 * on one device race. */
int  dmx_compute_dt_noalloc(const dmx_params* p, const double* uin, double* dt, void* stream);

/* One Godunov step of length dt at time `time` (time enters the shearing-box
 * flux/EMF remap at time + dt/2, src/subroutines/scheme.f90:176-182). */
int  dmx_godunov(dmx_ctx* ctx, double time, double dt);

/* FARGO orbital advection over dt (src/subroutines/fargo.f90:17-124). */
int  dmx_fargo(dmx_ctx* ctx, double dt);

/* Ghost fill at the end of a step (shearing remap at time + dt,
 * src/subroutines/boundary.f90:18-92). */
int  dmx_boundary(dmx_ctx* ctx, double time, double dt);

/* Dissipation over dt (src/subroutines/physics.f90:272-664): Ohmic
 * resistivity (edge EMFs -eta*J*dt and constrained transport) if eta > 0,
 * Navier-Stokes viscosity (momentum and energy fluxes) if nu > 0, then the
 * ghost fill.  The reference calls it after boundary when nu > 0 or eta > 0
 * (src/dumses.f90:103-105); dmx_advance does the same. */
int  dmx_dissipation(dmx_ctx* ctx, double time, double dt);

/* Runs nsteps of the reference time loop (src/dumses.f90:75-120) on the
 * device without host synchronisation per step: compute_dt, godunov,
 * [fargo], boundary, [dissipation], time += dt.  *time is read and updated; if dts is not
 * NULL it receives the nsteps timesteps. */
int  dmx_advance(dmx_ctx* ctx, int32_t nsteps, double* time, double* dts);

/* FP64 volume sums of src/subroutines/history.f90:62-118, in the column order
 * of history.txt after time and dt: mass, maxwell, reynolds, maxwell+reynolds,
 * magp, meanBx, meanBy, meanBz, divB. */
int  dmx_history(dmx_ctx* ctx, double out[9]);

/* The same sums with the shearing-box stresses that problem plugins such as
 * the MRI setup substitute for the history routine
 * (src/problem/mri/history.f90:54-90): maxwell = -<Bx By>,
 * reynolds = <rho dvx dvy> (velocity fluctuations about the yz means). */
int  dmx_history_shear(dmx_ctx* ctx, double out[9]);

/* Writes the context's state as the reference's binary dump
 * (src/subroutines/output.f90:224-235, io_type='binary'): five Fortran
 * unformatted sequential records -- int32 nx,ny,nz,1,1,1; f64 time,dt,dx,dy,dz;
 * f64 xmin..zmax; f64 gamma,ciso; f64 uin -- each framed by 4-byte
 * little-endian length markers.  `path` is the full file name (the reference
 * uses output_NNNNNN/data.000000.bin, src/modules/file.f90:21-51); its
 * directory must exist.  The uin record must stay below 2 GiB (about 270^3). */
int  dmx_write_dump(dmx_ctx* ctx, const char* path, double time, double dt);

/* Restart: reads a binary dump written by the reference or dmx_write_dump
 * (src/subroutines/restart.f90:212-223) into the context and returns its
 * time and dt.  Sizes, single-rank slicing, domain bounds, gamma and ciso must
 * match the context (the reference allocates before restart and would read
 * a mismatched file out of bounds; here it is DMX_E_PARAM). */
int  dmx_read_dump(dmx_ctx* ctx, const char* path, double* time, double* dt);

/* Synchronise and time: milliseconds of the last dmx_advance on the device. */
int  dmx_last_advance_ms(dmx_ctx* ctx, double* ms);

/* Instrumented dmx_advance: runs nsteps (<= 100000) with HIP events around
 * every kernel group of every step on the context's stream, synchronises
 * once at the end, and returns in ms[id] the mean device time per step of
 * each group id < nkernels (names from dmx_kernel_name; groups that did not
 * run report 0); dmx_last_advance_ms then gives the whole run.  No
 * counterpart in the reference, whose timers come from a source
 * preprocessor (utils/pyutils/preproc/preprocess.py). */
int  dmx_profile(dmx_ctx* ctx, int32_t nsteps, double* time, double* ms, int32_t nkernels);
const char* dmx_kernel_name(int32_t id);

/* Launch-shape autotuner: on the uploaded state, times the candidate shapes
 * of the trace -- mode 0: a 16x4x4 cell tile per block; mode 1: a 32x8
 * column tile marching 8, 16 or 32 z-planes -- and of the face/edge Riemann
 * stage -- mode 0: one fused kernel marching 1, 2, 4, 8, 16 or 32 z-planes
 * per block; mode 1: separate face and edge kernels; mode 2: the fused kernel
 * carrying the previous plane's states in LDS (8, 16 or 32 planes) -- keeps
 * the fastest of each, then times mode 3, one kernel for the trace and the
 * face/edge stage together (the trace summary stays in LDS; 4, 8, 16 or 32
 * planes), against the best separate pair and keeps it if faster.  Returns
 * the face/edge stage's time per launch (mode 3: trace included) in
 * *best_ms and {face/edge mode, planes, trace mode, planes} in shape[4]
 * (either may be NULL).  Every shape gives the same results in float64.
 * The state is unchanged.  No counterpart in the reference's step: it takes
 * the role of utils/pyutils/gpuconfig/gpuconfig.py, which sweeps OpenACC
 * vector sizes by re-running the code. */
int  dmx_tune(dmx_ctx* ctx, double* best_ms, int32_t* shape);
/* Sets the launch shapes without timing them: shape[4] as dmx_tune returns
 * it (face/edge mode 0..3 with 4, 8, 16 or 32 planes, mode 2 not with 4,
 * modes 0 and 1 also with 1 or 2;
 * trace mode 0, or 1 with 1..64 planes; mode 3 makes the trace shape
 * unused); for runs that must repeat a shape tuned before.  No counterpart
 * in the reference. */
int  dmx_set_shape(dmx_ctx* ctx, const int32_t* shape);

/* ---- batched ensemble -------------------------------------------------
 * Several independent contexts of one precision advanced together: every
 * kernel of the step (dt reduction, trace, face/edge stage, update, ghost
 * fill) is ONE launch over the tiles of all boxes, which fills the GPU with
 * many small boxes of uneven sizes where per-context launches would leave
 * it idle.  Each box keeps its own dt and time, and its results are bit for
 * bit those of dmx_advance on its own context.  No counterpart in the
 * reference (one domain per run).  Requirements: every context uploaded,
 * all of one precision and one Riemann solver pair, no shearing box, FARGO,
 * dissipation or learned limiter.  The contexts must outlive the batch and
 * are not used concurrently with it; their state stays in them (download
 * each with dmx_download).  The boxes are dealt, by decreasing size, to
 * nstreams groups (0: 2, at most 8), each launched on a stream of its own,
 * so that kernels of different kinds of the groups overlap on the GPU.  A
 * group whose boxes are all periodic (and free of what dmx_advance cannot
 * fuse the dt with) fuses the next dt into its update, as dmx_advance does. */
typedef struct dmx_batch dmx_batch;
int  dmx_batch_create(dmx_ctx* const* ctxs, int32_t n, int32_t nstreams, dmx_batch** batch);
void dmx_batch_destroy(dmx_batch* batch);
/* nsteps steps of every box; times[n] holds each box's start time on entry
 * and its end time on return */
int  dmx_batch_advance(dmx_batch* batch, int32_t nsteps, double* times);
/* launch-shape autotuner of the batch: the trace (the 16x4x4 tile, or the
 * z-march over 8/16/32 planes), then the face/edge stage (fused or carrying
 * kernel, 4/8/16/32 planes per block) after that trace against the trace
 * and face/edge stage in one kernel (mode 3, 8/16/32 planes); state
 * unchanged; best_ms is the chosen stage kernel's time, shape[4] as
 * dmx_tune.  A batch starts with context 0's shapes (a z-march or mode 3
 * over another plane count takes the nearest of 8, 16, 32; face/edge mode
 * 1 takes mode 0). */
int  dmx_batch_tune(dmx_batch* batch, double* best_ms, int32_t* shape);
/* milliseconds of the last dmx_batch_advance on the device */
int  dmx_batch_last_ms(dmx_batch* batch, double* ms);

/* ---- learned slope limiter (slope_type 4) ------------------------------
 * No counterpart in the reference.  A 40-64-64-8 GELU network maps the
 * scale-free features of the 5-cell stencil of the 8 primitive variables
 * along one direction to corrections of the 8 van Leer slopes of the centre
 * cell (slope = van Leer + output x stencil scale); it runs on bf16 MFMA
 * (float accumulation) before the trace.  A cell monotone in all 8 variables along the direction
 * takes the MC slopes instead.  A slope above the TVD bound (twice the
 * smaller one-sided difference) is clamped to it; one of the wrong sign or
 * at an extremum, or a cell whose density or pressure would turn
 * non-positive at a face, falls back to van Leer.  Parameters are 7304 floats: W1[40][64], b1,
 * W2[64][64], b2, W3[64][8], b3 (row = input).  dumses-hybrid_amd/dmx/learned.py
 * is the float64 reference of features, network and fallback. */
int  dmx_set_limiter(dmx_ctx* ctx, const float* theta, int32_t n);
/* Learned slopes of the current state: out receives 24*nx*ny*nz floats
 * [direction][variable][k][j][i] (interior); fallback (may be NULL)
 * receives 2 values: the fraction of cell-directions that fell back to van
 * Leer for all 8 variables (positivity) and the fraction of single-variable
 * slopes that left the TVD region (clamped, or replaced by van Leer). */
int  dmx_learned_slopes(dmx_ctx* ctx, float* out, double* fallback);

/* Training of the network on the GPU (hand-written float32 forward and
 * backward kernels, mean-squared error, Adam). */
typedef struct dmx_mlp dmx_mlp;
int  dmx_mlp_create(int32_t max_batch, dmx_mlp** mlp);
void dmx_mlp_destroy(dmx_mlp* mlp);
int  dmx_mlp_set_params(dmx_mlp* mlp, const float* theta, int32_t n);   /* resets Adam */
int  dmx_mlp_get_params(dmx_mlp* mlp, float* theta, int32_t n);
/* Y[n][8] for X[n][40] (host arrays, n <= max_batch) */
int  dmx_mlp_forward(dmx_mlp* mlp, const float* X, int32_t n, float* Y);
/* loss = 0.5 * sum((Y - T)^2) / n and its gradient (7304 floats) */
int  dmx_mlp_loss_grad(dmx_mlp* mlp, const float* X, const float* T, int32_t n, float* loss, float* grad);
/* copies a data set of n rows (X[n][40], T[n][8]) to the device */
int  dmx_mlp_load_data(dmx_mlp* mlp, const float* X, const float* T, int64_t n);
/* `steps` Adam steps on minibatches of `batch` rows drawn from the loaded
 * data; losses (may be NULL) receives the minibatch loss of every step */
int  dmx_mlp_train(dmx_mlp* mlp, int32_t steps, int32_t batch, float lr, uint64_t seed, float* losses);
/* D = A B (16x32 by 32x16, exact in bf16) through the bf16 MFMA lane layout
 * the limiter uses; out[16*16] row-major, for a host-side check */
int  dmx_mfma_selftest(float* out);

const char* dmx_last_error(void);
const char* dmx_version(void);

#ifdef __cplusplus
}
#endif
#endif /* DMX_H */
