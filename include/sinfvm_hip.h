/*
 * sinfvm_hip.h — C ABI of the MI355X (gfx950) shallow-water backend.
 *
 * This is the boundary a host language binds to (Julia `ccall`, Python
 * `ctypes`).  Every entry point takes plain pointers and sizes.  The library
 * owns all device memory behind an opaque handle; host pointers are borrowed
 * for the duration of a call only.
 *
 * Array layout (all host buffers): column-major, x fastest, one plane per
 * conserved variable — byte-identical to the reference's `Volume._data`
 * (`zeros(realtype, nxT, nyT, nvar)`, src/backends/buffer.jl:29-31), so a
 * Julia caller passes `pointer(vol._data)` without a transpose.  Variable 1
 * holds the SURFACE ELEVATION w = h + B (the reference names it `:h`,
 * src/equations/swe_2D.jl:67), variables 2 and 3 hold hu and hv.
 * Bottom topography is the corner array of shape (nxT+1, nyT+1)
 * (src/bottom_topography.jl:78-82), also x fastest.
 *
 * Threading: calls on one handle must not overlap.  Independent handles may
 * be driven from different host threads (each handle owns its own stream).
 */
#ifndef SINFVM_HIP_H
#define SINFVM_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes (mirror the reference's exception types) ------------ */
enum sfv_status {
    SFV_OK = 0,
    SFV_ERR_DOMAIN = 1,       /* DomainError: array sizes (bottom_topography.jl:79-81,
                                 source_term_infiltration.jl:61-65, source_term_rain.jl:36-38) */
    SFV_ERR_ARGUMENT = 2,     /* ArgumentError: non-zero B without SourceTermBottom (equation.jl:74-84),
                                 or an invalid enum / parameter */
    SFV_ERR_ASSERTION = 3,    /* AssertionError: gc < 2 with linear reconstruction (linearreconstruction.jl:59) */
    SFV_ERR_OUT_OF_MEMORY = 4,
    SFV_ERR_NONFINITE_DT = 5, /* dt came out NaN or +-Inf (all-dry domain, simulator.jl:190-192) */
    SFV_ERR_HIP = 6,          /* a HIP runtime call failed */
    SFV_ERR_STATE = 7,        /* call made in the wrong order (e.g. step before set_state) */
    SFV_ERR_BUDGET = 8        /* serving: request does not fit the HBM byte budget */
};

/* ---- configuration enums (mirror the reference's Julia types) ---------- */
enum sfv_bc { SFV_BC_WALL = 0, SFV_BC_PERIODIC = 1, SFV_BC_NEUMANN = 2 };          /* grid.jl:21-28 */
enum sfv_stepper { SFV_FORWARD_EULER = 1, SFV_RUNGE_KUTTA2 = 2 };                   /* timestepper/ */
enum sfv_reconstruction { SFV_NO_RECONSTRUCTION = 0, SFV_LINEAR_RECONSTRUCTION = 1 };
enum sfv_friction {                                                                 /* friction.jl:30-43 */
    SFV_FRICTION_NONE = 0, SFV_FRICTION_BH2021 = 1, SFV_FRICTION_BSA2012 = 2, SFV_FRICTION_FCG2016 = 3
};
enum sfv_rain {                                                                     /* source_term_rain.jl */
    SFV_RAIN_NONE = 0, SFV_RAIN_CONSTANT = 1, SFV_RAIN_TIME_TABLE = 2,
    SFV_RAIN_FIELD = 3    /* FunctionalRain: per-cell rates the caller sets before each step (sfv_set_rain_field) */
};
enum sfv_infiltration {                                                             /* source_term_infiltration.jl */
    SFV_INFILTRATION_NONE = 0, SFV_INFILTRATION_CONSTANT = 1, SFV_INFILTRATION_HORTON = 2
};
enum sfv_precision { SFV_F64 = 0, SFV_F32_STORAGE = 1 };
enum sfv_layout { SFV_LAYOUT_FULL = 0, SFV_LAYOUT_INTERIOR = 1 };

/*
 * Everything the reference spreads over constructor keyword arguments:
 *   CartesianGrid(nx, ny; gc, boundary, extent)          grid.jl:82-90
 *   ShallowWaterEquations(B; ρ, g, depth_cutoff, desingularizing_kappa)  swe_2D.jl:28
 *   LinearReconstruction(theta) / NoReconstruction()     linearreconstruction.jl:22-25
 *   RungeKutta2() / ForwardEulerStepper(), Simulator(...; cfl)  simulator.jl:72
 *   sources = [SourceTermBottom(), rain, infiltration]   system.jl:37
 *   ImplicitFriction(; Cz, friction_function)             friction.jl:23-27
 * dx, dy are grid.Δ as the reference computes them: (x1 - x0) / nx.
 */
typedef struct sfv_config {
    int32_t nx, ny, gc;            /* nx <= 2^27 - 17 (the kernels address a row with 31-bit
                                      byte offsets), else SFV_ERR_DOMAIN */
    double dx, dy;
    int32_t bc;                    /* enum sfv_bc */
    double g, rho, depth_cutoff, desingularizing_kappa;
    int32_t reconstruction;        /* enum sfv_reconstruction */
    double theta;
    int32_t stepper;               /* enum sfv_stepper */
    double cfl;
    int32_t bottom_source;         /* 1 if SourceTermBottom is in the source list */
    int32_t friction;              /* enum sfv_friction (ImplicitFriction), 0 = no implicit source */
    double Cz;
    int32_t rain;                  /* enum sfv_rain */
    double rain_rate;              /* ConstantRain rate, in the reference's unit (divided by 3600 on use) */
    int32_t rain_table_len;        /* TimeDependentRain: rates (reference unit) and start times */
    const double* rain_table_rates;
    const double* rain_table_times;
    int32_t infiltration;          /* enum sfv_infiltration */
    double infiltration_rate;      /* ConstantInfiltration rate (used as is) */
    double horton_f0, horton_fc, horton_k;
    double infiltration_factor_scalar;   /* HortonInfiltration factor when no field is given */
    const double* infiltration_factor;   /* (nxT, nyT) field or NULL */
    int32_t rain_before_infiltration;    /* 1: rain precedes infiltration in the source list */
    int32_t precision;             /* enum sfv_precision */
    int32_t embedded_1d;           /* 1: this (nx, 3) grid carries a CartesianGrid{1} run in every row
                                      (bottom replicated along y), and B_cell is the 1D lookup
                                      0.5*(B[i] + B[i+1]) (bottom_topography.jl:92) so the middle row
                                      equals the 1D scheme bitwise; 0 otherwise */
} sfv_config;

/* Fill *cfg with the reference's defaults (gc=2, WallBC, g=9.81, ρ=1,
 * depth_cutoff = desingularizing_kappa = 1e-5, θ=1.2, RK2, cfl=0.25, no sources). */
void sfv_config_defaults(sfv_config* cfg, int32_t nx, int32_t ny, double dx, double dy);

typedef struct sfv_handle sfv_handle;

/* Lifecycle.  Replaces ConservedSystem(...) + Simulator(...) allocation
 * (system.jl:37-53, simulator.jl:72-91).  Validates the configuration the way
 * the reference constructors do; the topography is checked in
 * sfv_set_topography and when stepping starts. */
int sfv_create(const sfv_config* cfg, sfv_handle** out);
void sfv_destroy(sfv_handle* h);
/* Text of the last error on this handle ("" if none).  Valid until the next call. */
const char* sfv_last_error(const sfv_handle* h);
/* Text of the last error of sfv_create or sfv_advance_noalloc (thread-local). */
const char* sfv_create_error(void);

/* BottomTopography2D(B, backend, grid) (bottom_topography.jl:51-57): corner
 * array (nbx, nby) must equal (nxT+1, nyT+1) else SFV_ERR_DOMAIN.  A non-zero
 * B without bottom_source gives SFV_ERR_ARGUMENT (equation.jl:74-84). */
int sfv_set_topography(sfv_handle* h, const double* B, int64_t nbx, int64_t nby);

/* set_current_state! (simulator.jl:119-135).  SFV_LAYOUT_INTERIOR: U is
 * (nx, ny, 3); the interior is written and the boundary condition applied,
 * other ghost cells keep their values.  SFV_LAYOUT_FULL: U is (nxT, nyT, 3)
 * and is copied verbatim (a checkpoint restore).  Element type is double for
 * SFV_F64 and float for SFV_F32_STORAGE. */
int sfv_set_state(sfv_handle* h, const void* U, int32_t layout);
/* collect(current_state(sim)) / current_interior_state (simulator.jl:98-113). */
int sfv_get_state(sfv_handle* h, void* U, int32_t layout);

/* FunctionalRain (source_term_rain.jl:57-68): the rates rain_function(t, x, y)
 * at the cell centres of ALL cells, ghosts included (cell_center,
 * grid.jl:117-121), as a (nxT, nyT) array, x fastest, added to w as they are
 * (no /3600).  The reference evaluates the function at the step's start time
 * t for both stages, so the caller sets the field for t before each sfv_step.
 * Only for a handle created with rain = SFV_RAIN_FIELD; such a handle steps
 * with sfv_step only (sfv_run / sfv_advance / batches return
 * SFV_ERR_ARGUMENT, since the field must change every step). */
int sfv_set_rain_field(sfv_handle* h, const double* field);

/* perform_step!(sim, max_dt) (simulator.jl:145-177): one full time step with
 * sources evaluated at time t.  Writes the step's dt (current_timestep) to
 * *dt_out.  Does NOT advance t — the caller does, as simulate_to_time does
 * (simulator.jl:205). */
int sfv_step(sfv_handle* h, double max_dt, double t, double* dt_out);

/* simulate_to_time(sim, t_end; match_endtime, maximum_timestep) without a
 * callback (simulator.jl:179-212), run device-resident: time and dt live in
 * device memory and the host does not synchronise per step.  *t_io is the
 * start time on entry and the end time on exit; at most max_steps steps are
 * taken (max_steps <= 0: no limit).  maximum_timestep <= 0 means none. */
int sfv_run(sfv_handle* h, double t_end, int32_t match_endtime, double maximum_timestep,
            int64_t max_steps, double* t_io, int64_t* steps_out);

/* nsteps calls of perform_step!(sim, max_dt) with t += dt after each
 * (the fixed-step-count workload), device-resident.  *t_io in/out. */
int sfv_advance(sfv_handle* h, int64_t nsteps, double max_dt, double* t_io);

/* No counterpart in the reference: for callers that cannot allocate device
 * memory (tight serving loops, harnesses that forbid hipMalloc).  sfv_advance
 * on buffers the caller owns, with no handle and no allocation: nsteps calls
 * of perform_step!(sim, max_dt) with t += dt after each, device-resident, on
 * the full state U_dev ((nxT, nyT, 3) in the configuration's precision, in
 * device memory, 16-byte aligned; replaced by the state after nsteps steps,
 * ghost cells included, as sfv_get_state with SFV_LAYOUT_FULL would return
 * it) over the corner topography B_dev ((nxT+1, nyT+1) double in device
 * memory).  *t_io is the start time on entry and the end time on exit;
 * *dt_out (may be NULL) receives the last step's dt.  The scratch the steps
 * need (the stage right-hand side, the second state of the ping-pong pair and
 * the step-control block) is one static device array of
 * SFV_NOALLOC_SCRATCH_BYTES that each GPU the library runs on holds until the
 * process ends: grids of at most 1,398,085 cells ghosts included (f64;
 * 2,796,170 with f32 storage) are accepted, larger ones return
 * SFV_ERR_DOMAIN.  The configuration must suit the fused RK2 kernels:
 * RungeKutta2, WallBC or PeriodicBC with nx, ny >= gc >= 2, rain none or
 * constant, infiltration none or constant (no device tables); other ones
 * return SFV_ERR_ARGUMENT.  Work goes on `stream` (a hipStream_t; NULL: the
 * null stream) and the call returns when it has finished.  Results are
 * bitwise those of sfv_advance.  Every call on a GPU shares that GPU's
 * scratch: calls that overlap on one GPU race.  The message of a failed call
 * is sfv_create_error()'s. */
#define SFV_NOALLOC_SCRATCH_BYTES (64ull * 1024 * 1024)
int sfv_advance_noalloc(const sfv_config* cfg, const double* B_dev, void* U_dev, int64_t nsteps, double max_dt,
                        double* t_io, double* dt_out, void* stream);

/* dt of the most recent step (current_timestep(sim), simulator.jl:141). */
int sfv_current_timestep(sfv_handle* h, double* dt_out);

/* Block until all work queued on the handle's stream has finished. */
int sfv_synchronize(sfv_handle* h);

/* Bytes of device memory the handle holds (for budgeting). */
int64_t sfv_device_bytes(const sfv_handle* h);
/* Bytes of device memory a handle with this configuration would hold. */
int64_t sfv_device_bytes_for(const sfv_config* cfg);

/* Kernel-timing hook for measurement: run n RK2 steps' kernels from the
 * current state (stage 1, stage 2 with the dt stage 1's wave speeds give,
 * boundary condition) into a scratch output, with HIP events on the handle's
 * stream between consecutive kernels, after one warm-up step, and report the
 * named kernel's mean device time in milliseconds ("stage1", "stage2", "bc",
 * or "step" for the three).  "copy" times a plain device copy of one state (3
 * planes), the known-traffic kernel for calibrating HBM byte counters.  The
 * state is left unchanged.  No counterpart in the reference. */
int sfv_time_kernel(sfv_handle* h, const char* which, int32_t n, double* ms_out);

/* Number of visible HIP devices (0 when none).  Needs no handle. */
int sfv_device_count(void);

/* Select the kernels of a step for this handle: 2 = fused RK2 stage kernels,
 * one wavefront per 60-column strip marching in y (the default whenever the
 * boundary-condition sources of the ghost cells are interior cells; unless a
 * variant is set here, the first step's tuning switches to 1 when that is
 * faster for the grid); 1 = fused RK2 stage kernels tiled through LDS; 3 =
 * every right-hand side from
 * the march stage-1 kernel, the updates and boundary conditions from the
 * per-pass kernels (any stepper and boundary condition with gc >= 2: the
 * default for NeumannBC and ForwardEuler); 0 = one kernel per reference pass
 * (the device-side reference).  All give bitwise identical results.  No
 * counterpart in the reference. */
int sfv_set_fused(sfv_handle* h, int32_t fused);
/* The kernel variant a step of this handle runs (0 to 3 as above; -1 for a
 * NULL handle).  No counterpart in the reference. */
int32_t sfv_get_fused(const sfv_handle* h);

/* Rows per wave segment of the march kernels (sfv_set_fused(h, 2)).  0 (the
 * default) lets the library choose on the first step: candidate heights are
 * timed on the handle's own state and the fastest is kept for grids of that
 * size class (quarter octaves of nx and ny) on that device; the environment
 * variable SFV_AUTOTUNE=0 turns the measurement off (the height that makes
 * the launch's workgroups the fewest rounds of resident workgroups is used).
 * rows > 0 forces a height.  Results are bitwise the same for every height.
 * No counterpart in the reference. */
int sfv_set_march_rows(sfv_handle* h, int32_t rows);
/* The height in use (0 while not yet chosen). */
int32_t sfv_get_march_rows(const sfv_handle* h);

/* ------------------------------------------------------------------ adjoint
 * Reverse-mode gradient of J = sum over interior cells of (h_T - h_target)^2,
 * h = w - B_cell, where h_T is the state after nsteps calls of
 * perform_step!(sim, Inf) (simulator.jl:145-177) from the handle's current
 * state and *t_io.  Gradients (each may be NULL): dJ_dB over the corners
 * (nxT+1, nyT+1), dJ_dCz (ImplicitFriction's Cz; Cz = n^2 gives
 * dJ/dn = 2n dJ/dCz), dJ_dU0 over the interior initial state (nx, ny, 3).
 * The derivative is the one forward-mode differentiation of the reference
 * (ForwardDiff through simulate_to_time) gives: every branch the run took is
 * differentiated as taken, and dt is differentiated through its maximum
 * wave speed.  Checkpoints every checkpoint_every steps bound the memory.
 * Float64, RungeKutta2, LinearReconstruction, WallBC or PeriodicBC with
 * nx, ny >= gc, no FunctionalRain.  On return the handle holds the final
 * state, *t_io its time and *J_out the objective.  No counterpart in the
 * reference, whose optimisation example uses ForwardDiff. */
int sfv_adjoint(sfv_handle* h, int64_t nsteps, int32_t checkpoint_every, double* t_io, const double* h_target,
                double* J_out, double* dJ_dB, double* dJ_dCz, double* dJ_dU0);

/* ---------------------------------------------------------------- ensembles
 * No counterpart in the reference (a reference user runs one Simulator at a
 * time).  A batch groups existing handles — grids of any size, each with its
 * own configuration, topography, state, time and time step — and advances
 * them together: one persistent launch per RK stage covers every member and
 * the step sequence is a captured HIP graph.  Each member's results are
 * bitwise those of stepping it alone.  Members must use RungeKutta2 with the
 * fused kernels (ghost-cell sources inside the grid: WallBC or PeriodicBC
 * with nx, ny >= gc), one device and one precision.  The batch borrows the
 * handles: destroy it before them, and do not step a member through its own
 * handle while a batch run is in progress.  Read results with
 * sfv_get_state on each member. */
typedef struct sfv_batch sfv_batch;

int sfv_batch_create(sfv_handle* const* handles, int32_t n, sfv_batch** out);
void sfv_batch_destroy(sfv_batch* b);
/* Message of the last failed call on b (b = NULL: of sfv_batch_create). */
const char* sfv_batch_last_error(const sfv_batch* b);
int32_t sfv_batch_size(const sfv_batch* b);

/* simulate_to_time (simulator.jl:179-212) for every member at once, each
 * from its own t_io[i]: with t_end (n values) member i runs until t_end[i]
 * (match_endtime and maximum_timestep as in sfv_run), otherwise exactly
 * max_steps steps each (max_dt = maximum_timestep if > 0).  max_steps > 0
 * also caps the steps of a t_end run.  On return t_io[i] is member i's time
 * and steps_out[i] (may be NULL) its step count.  Status SFV_ERR_NONFINITE_DT
 * names the first member whose dt was NaN or Inf. */
int sfv_batch_run(sfv_batch* b, const double* t_end, int32_t match_endtime, double maximum_timestep,
                  int64_t max_steps, double* t_io, int64_t* steps_out);
/* Segment height of the march kernels that the last sfv_batch_run used for
 * members without a height set by sfv_set_march_rows (0 before any run).
 * A run times the height rule's choice against its neighbours on its first
 * replays and keeps the fastest (SFV_AUTOTUNE=0: the rule's choice); the
 * choice is kept for later runs with the same member sizes.  No height
 * changes any result. */
int sfv_batch_get_march_rows(const sfv_batch* b, int32_t* rows);

/* ------------------------------------------------------------------ serving
 * No counterpart in the reference.  A server runs flood scenarios submitted
 * at any time (continuous batching): a worker thread admits queued requests
 * in arrival order while their device bytes (sfv_device_bytes_for) fit the
 * HBM budget, advances everything admitted with the batched kernels (each
 * request with its own dt, until its own t_end, as simulate_to_time with
 * match_endtime does), and retires finished requests: the final interior
 * state is copied to host memory and the device memory is freed.  A result
 * is bitwise that of running the request alone.  Requests use RungeKutta2,
 * the server's precision, and ghost-cell sources inside the grid (WallBC or
 * PeriodicBC with nx, ny >= gc).  All calls are thread-safe. */
typedef struct sfv_server sfv_server;

enum sfv_request_status { SFV_REQ_QUEUED = 0, SFV_REQ_RUNNING = 1, SFV_REQ_DONE = 2, SFV_REQ_FAILED = 3 };

typedef struct sfv_request_info {
    int64_t id;
    int32_t status;        /* enum sfv_request_status */
    int32_t error;         /* status code of a failed (or non-finite-dt) request */
    double t;              /* simulated time reached */
    int64_t steps;
    int64_t cell_updates;  /* steps * nx * ny */
    int64_t device_bytes;  /* what the request holds on the device while it runs */
    double submit_s, admit_s, done_s;   /* seconds since the server started; -1 = not yet */
} sfv_request_info;

typedef struct sfv_server_stats {
    int64_t submitted, admitted, completed, failed, queued, running;
    int64_t bytes_in_use, bytes_peak, budget_bytes;
    int64_t cell_updates;  /* of completed requests */
    int64_t replays;       /* graph replays (8 steps of every running request each) */
    double busy_s;         /* time with a replay in flight */
    double uptime_s;
    double host_admit_s;   /* worker time creating admitted requests' handles and uploading their inputs */
    double host_retire_s;  /* worker time downloading finished requests' results */
} sfv_server_stats;

/* Start a server on the current device with an HBM byte budget. */
int sfv_server_create(int64_t budget_bytes, int32_t precision, sfv_server** out);
/* Stop the worker (queued and running requests fail) and free everything. */
void sfv_server_destroy(sfv_server* s);
const char* sfv_server_last_error(const sfv_server* s);
/* Queue a request: configuration (tables it points to are copied), corner
 * topography (nxT+1, nyT+1) and the initial interior state (nx, ny, 3) in the
 * server's precision; it runs from t0 to t_end.  Inputs are copied.  A
 * request larger than the whole budget fails with SFV_ERR_BUDGET. */
int sfv_server_submit(sfv_server* s, const sfv_config* cfg, const double* B, int64_t nbx, int64_t nby,
                      const void* U_interior, double t0, double t_end, int32_t match_endtime,
                      double maximum_timestep, int64_t* id_out);
/* Status of a request. */
int sfv_server_poll(sfv_server* s, int64_t id, sfv_request_info* info);
/* Block until the request is done or failed (timeout_s < 0: no limit);
 * SFV_ERR_STATE on timeout. */
int sfv_server_wait(sfv_server* s, int64_t id, double timeout_s, sfv_request_info* info);
/* Copy a finished request's final interior state (nx, ny, 3) to U_interior_out
 * (may be NULL) and forget the request.  Returns the request's own status
 * code for a failed request. */
int sfv_server_fetch(sfv_server* s, int64_t id, void* U_interior_out, sfv_request_info* info);
int sfv_server_get_stats(sfv_server* s, sfv_server_stats* out);

#ifdef __cplusplus
}
#endif

#endif /* SINFVM_HIP_H */
