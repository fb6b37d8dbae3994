/*
 * gpf.h -- C ABI of the MI355X spectral-step library (libgopf_amd.so).
 *
 * The entry points are what a cgo binding of gopf's spectral time step
 * needs (see INTEGRATION.md for the Go side).  Each group replaces one
 * reference interface, cited as path:line in the gopf repository:
 *
 *   gpf_model_*   pf.Model (pf/model.go:116-260): fields, scalars, equation
 *                 strings, registered terms; gpf_model_init is Model.Init and
 *                 parses the equations exactly as pf/rhsBuilder.go:26-242.
 *   gpf_solver_*  pf.Solver + pf.TimeStepper (pf/solver.go:15-120,
 *                 pf/euler.go:16-59, pf/rk4.go:29-135): fields live in HBM,
 *                 steps run on one MI355X.
 *   gpf_fft_*     pf.FourierTransform implemented by pfutil.FFTWWrapper
 *                 (pf/solver.go:22-26, pfutil/fftWrap.go:16-95).
 *   gpf_vandeven_* pf.Vandeven (pf/vandeven.go:13-40).
 *
 * Conventions
 *   - Every function returns 0 on success and a non-zero GPF_E* code on
 *     failure; gpf_last_error() returns the message of the last failure on
 *     the calling thread.  (The reference panics; the Go wrapper turns a
 *     non-zero code into a panic with that message.)
 *   - Complex host arrays are interleaved complex128 (double re, double im),
 *     the memory layout of Go's []complex128.  Host buffers are only
 *     borrowed for the duration of a call.
 *   - Grids follow FFTW's row-major layout, plan n = [D0, D1(, D2)], last
 *     index fastest (pfutil/fftWrap.go:19).
 *   - Handles are not thread-safe; one handle owns one HIP stream.
 */
#ifndef GPF_H
#define GPF_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GPF_OK 0
#define GPF_EINVAL 1     /* invalid argument / unsupported configuration */
#define GPF_EPARSE 2     /* equation parsing failed (the reference panics) */
#define GPF_EDEVICE 3    /* HIP runtime error / no device */
#define GPF_ESTATE 4     /* call out of order (e.g. solver before Model.Init) */

#define GPF_FP64 0       /* plain FP64: complex128 transforms, FP64 pointwise work */
#define GPF_FP32 1       /* FP32 bulk; FP64 k = 0 mode, denominators and reductions */

#define GPF_EULER 0      /* pf.Euler */
#define GPF_RK4 1        /* pf.RK4 */
#define GPF_IMPLICIT_EULER 2  /* pf.ImplicitEuler (FP64 only) */
#define GPF_SDD 3        /* pf.SDD, shrinking-dimer dynamics (FP64 only) */

typedef struct gpf_model gpf_model;
typedef struct gpf_solver gpf_solver;
typedef struct gpf_fft gpf_fft;
typedef struct gpf_batch gpf_batch;

/* Message of the last failed call on this thread ("" if none). */
const char* gpf_last_error(void);

/* ---------------- pf.Model ---------------- */

/* pf.NewModel (pf/model.go:130-138) */
int gpf_model_create(gpf_model** out);
void gpf_model_destroy(gpf_model* m);
/* pf.NewField + Model.AddField (pf/model.go:44-57, 141-144); data may be NULL (zeros) */
int gpf_model_add_field(gpf_model* m, const char* name, const double* data, long n);
/* pf.NewScalar + Model.AddScalar (pf/model.go:92-98, 147-149) */
int gpf_model_add_scalar(gpf_model* m, const char* name, double re, double im);
/* Model.AddEquation (pf/model.go:157-162): strips spaces, creates derived fields */
int gpf_model_add_equation(gpf_model* m, const char* eq);
/* pf.NewSquareGradient + Model.RegisterExplicitTerm (pf/squareGradientTerm.go:24-37,
 * pf/model.go:362-364); implicit != 0 registers it with RegisterImplicitTerm */
int gpf_model_register_square_gradient(gpf_model* m, const char* term, const char* field, double factor,
                                       int implicit);
/* pf.NewHomogeneousModolus + RegisterExplicitTerm (pf/homoLinElast.go:137-150):
 * C is C_ijkl (81 doubles, index i*27+j*9+k*3+l), misfit is 3x3 row-major */
int gpf_model_register_homogeneous_modulus(gpf_model* m, const char* term, const char* field,
                                           const double* C, const double* misfit);
/* pf.NewVolumeConservingLP + RegisterExplicitTerm (pf/volumeConserving.go:8-61): the term
 * Multiplier * indicator, whose Lagrange multiplier is updated after every step
 * (OnStepFinished) from the change of sum_i field[i], i < num_nodes; indicator names a
 * derived field (a registered function) or a field */
int gpf_model_register_volume_conserving(gpf_model* m, const char* term, const char* field, const char* indicator,
                                         double dt, long num_nodes);
/* pf.SpectralViscosity + RegisterImplicitTerm (pf/spectralViscosity.go:10-58): the
 * implicit multiplier -eps Q(|f|) |f|^power with Tadmor's interpolant Q between
 * threshold / 3 and threshold */
int gpf_model_register_spectral_viscosity(gpf_model* m, const char* term, double eps, double threshold, int power);
/* TensorialHessian + RegisterImplicitTerm (pf/tensorialHessian.go:19-66): the implicit
 * multiplier sum_ij K_ij (i 2 pi f_i)(i 2 pi f_j) of d^2/dx_i dx_j; K is row-major with
 * nk = 4 (2D) or 9 (3D) entries; field is kept for the reference's struct and unused */
int gpf_model_register_tensorial_hessian(gpf_model* m, const char* term, const char* field, const double* K,
                                         int nk);
/* Model.RegisterFunction(name, WhiteNoise{strength}.Generate) (pf/noise.go:10-22): a derived
 * field of Gaussian noise with variance 2 strength, redrawn at every SyncDerivedFields.  The
 * reference draws from the global math/rand source; the device draws from Philox4x32-10
 * keyed by seed on (node, draw count, derived-field index), so runs are reproducible. */
int gpf_model_register_white_noise(gpf_model* m, const char* name, double strength, unsigned long long seed);
/* NewConservativeNoise(strength, dim) + RegisterExplicitTerm(term, &noise,
 * noise.RequiredDerivedFields(N)) (pf/noise.go:24-100): dim derived Gaussian current fields
 * and the term sum_c i 2 sin(pi f_c) FFT(current_c) (0 on Nyquist rows), which conserves the
 * field's integral */
int gpf_model_register_conservative_noise(gpf_model* m, const char* term, double strength, int dim,
                                          unsigned long long seed);
/* Model.RegisterFunction (pf/model.go:400-412) for a polynomial of the fields' real parts:
 * value = sum_t coefs[t] * prod_f re(field_f)^powers[t * nfields + f]; the device cannot run
 * Go closures, so the function is given as data. */
int gpf_model_register_polynomial(gpf_model* m, const char* name, int nterms, const double* coefs,
                                  const double* powers, int nfields);
/* Model.Init (pf/model.go:244-260): Build every equation (pf/rhsBuilder.go:26-54) */
int gpf_model_init(gpf_model* m);

/* introspection of the parsed model (Model.AllFieldNames, Model.RHS) */
typedef struct {
    int kind;          /* 0 ConcreteTerm, 1 SquaredGradient, 2 HomogeneousModulusLinElast,
                          3 VolumeConservingLP, 4 SpectralViscosity, 5 TensorialHessian,
                          6 ConservativeNoise */
    double coef_re;    /* concrete: sign * prod Pow(scalar, p); user terms: +-1 */
    double coef_im;
    int brick;         /* brick index (fields, then derived fields) or -1 */
    int has_lap;       /* the term string contains LAP */
    int lap;           /* LaplacianN power (concrete) / summed LAP prefixes (user terms) */
    int user;          /* index among the registered terms of its kind, or -1 */
} gpf_term;

int gpf_model_num_bricks(const gpf_model* m, int* nfields, int* nderived);
int gpf_model_brick_name(const gpf_model* m, int idx, char* buf, int buflen);
int gpf_model_num_terms(const gpf_model* m, int eq, int* nterms, int* ndenum);
int gpf_model_get_term(const gpf_model* m, int eq, int denum, int idx, gpf_term* out);

/* ---------------- pf.Solver / pf.TimeStepper ---------------- */

typedef struct {
    int precision;     /* GPF_FP64 or GPF_FP32 */
    int stepper;       /* GPF_EULER or GPF_RK4 (Solver.SetStepper, pf/solver.go:88-103), or
                          GPF_IMPLICIT_EULER (solver.Stepper = &pf.ImplicitEuler{...},
                          pf/implicitEuler.go:20-29) */
    int filter_order;  /* > 0: Stepper.SetFilter(NewVandeven(order)); 0: no filter */
    int fused;         /* 1: allow the fused single-field path; 0: general path only */
    int device;        /* HIP device ordinal */
    int autotune;      /* 1: on the first step, time every pass of the fused step for each
                          admissible launch shape (lines per block, lines per thread, transform
                          schedule, 3D update axis) and keep the fastest */
} gpf_options;

void gpf_options_default(gpf_options* o);

/* pf.NewSolver (pf/solver.go:40-63) with the stepper chosen by o; the model must
 * have been initialised.  Uploads the fields; they stay resident in HBM. */
int gpf_solver_create(gpf_model* m, const int* dims, int ndim, double dt, const gpf_options* o,
                      gpf_solver** out);
void gpf_solver_destroy(gpf_solver* s);
/* Solver.Propagate (pf/solver.go:70-85): nsteps x (Stepper.Step + OnStepFinished) */
int gpf_solver_propagate(gpf_solver* s, int nsteps);
/* Euler.Propagate / RK4.Propagate (pf/euler.go:55-59, pf/rk4.go:130-135) */
int gpf_solver_stepper_propagate(gpf_solver* s, int nsteps);
/* No reference counterpart (the reference plans nothing ahead of Step): does ahead of time
 * what the first Step would, without advancing the fields or the time -- the autotuner
 * (when gpf_options.autotune is set), the energy monitor's buffers for nsteps steps and
 * the captured step graph that Propagate replays -- so that later steps only launch.
 * *ms (may be NULL) receives the host milliseconds it took.  A no-op on the general path. */
int gpf_solver_prepare(gpf_solver* s, int nsteps, double* ms);
/* TimeStepper.GetTime (pf/euler.go:50-52, pf/rk4.go:139-141) */
int gpf_solver_get_time(gpf_solver* s, double* t);
/* copy field idx (Model.Fields order) to / from host complex128 (imag set to 0 / ignored) */
int gpf_solver_get_field(gpf_solver* s, int idx, double* out, long n);
int gpf_solver_set_field(gpf_solver* s, int idx, const double* in, long n);
/* free energy F = sum_x sum_q poly[q] c^q + kappa/2 sum_x |grad c|^2 and mass sum_x c of field
 * idx, in FP64 (compensated).  No reference counterpart: a monitor (cf. pf/monitor.go). */
int gpf_solver_energy(gpf_solver* s, int idx, int npoly, const double* poly, double kappa,
                      double* energy, double* mass);
/* FP64 reductions of a real field (Neumaier-compensated sums; min and max over the finite
 * values, +inf / -inf when there is none; the count of NaN and +-inf values).  No reference
 * counterpart: it replaces the host scan for NaN that a gopf driver runs over
 * m.Fields[i].Data to stop an unstable run (examples/solverStability/main.go:58-67). */
typedef struct {
    double sum;        /* sum_x c(x): the field's mass */
    double sum_sq;     /* sum_x c(x)^2 */
    double min;
    double max;
    long nonfinite;    /* values that are NaN or +-inf (left out of the other four) */
} gpf_reductions;

/* the reductions of field idx (Model.Fields order) as gpf_solver_get_field would return it */
int gpf_solver_reduce(gpf_solver* s, int idx, gpf_reductions* out);
/* the reductions of n real values of the given precision (GPF_FP32: float, GPF_FP64:
 * double) in device memory of `device`, computed on `stream` (a hipStream_t; NULL is the
 * default stream), which is synchronised before return.  Allocates its partial sums. */
int gpf_reduce(const void* data, long n, int precision, int device, void* stream, gpf_reductions* out);
/* No reference counterpart.  For callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops): gpf_reduce with its partial sums in a static
 * device array of the library (160 KiB per device, held until the process ends) and the
 * result left in device memory: result[0..4] = sum, sum_sq, min, max, nonfinite (as a
 * double), written on `stream` without synchronising it.  It calls no allocating routine.
 * n may be at most 2^28 (268435456) values; larger n fails with GPF_EINVAL.  Every call on
 * one device shares the array: calls that overlap on one device race. */
int gpf_reduce_noalloc(const void* data, long n, int precision, int device, void* stream, double* result);

/* record the energy of field 0 before every step into a device series (fused path) */
int gpf_solver_monitor_energy(gpf_solver* s, int npoly, const double* poly, double kappa);
int gpf_solver_energy_series(gpf_solver* s, double* energy, double* mass, long cap, long* n);
int gpf_solver_synchronize(gpf_solver* s);
/* 1 if the fused single-field path runs, 0 for the general path */
int gpf_solver_is_fused(const gpf_solver* s);
/* per-pass device timing: enable, then read the summed milliseconds and launch counts
 * of each pass recorded with HIP events on the solver's stream */
int gpf_solver_enable_timing(gpf_solver* s, int on);
int gpf_solver_pass_times(gpf_solver* s, int maxpass, double* ms, long* launches, char* names, int namelen);
/* launch configuration of the fused step: lines per block of the rows, forward, update and
 * inverse passes (0 = default heuristic), in 3D the axis of the update pass (2 = z, 1 = y),
 * the transform schedule of the row pass (0 plain, 1 low-register), the lines per thread
 * of the forward and inverse pencil passes and their transform schedules; returns the
 * number of values written */
int gpf_solver_tuning(gpf_solver* s, int* tws, int maxn);
/* applies a launch layout of 10 values as gpf_solver_tuning returns them (from an earlier
 * run of the same grid and precision) instead of running the autotuner; GPF_EINVAL when
 * the layout is not admissible for this grid and precision */
int gpf_solver_set_tuning(gpf_solver* s, const int* tws, int n);
/* The persistent step of square 2D fused grids of side 32, 64, 128 or 256 (no counterpart in
 * the reference: a launch layout of Solver.Propagate's steps, pf/solver.go:106-120): a run
 * of steps is one cooperative launch whose workgroups meet at device-wide barriers between
 * the passes.  mode -1 (default): taken for sides up to 128 when the device places the
 * launch's workgroups on one XCD; 0: never (per-pass launches and replayed graphs); 1:
 * wherever admissible; 2: as 1 with agent-scope fences at every barrier (the layout used
 * when the workgroups are not on one XCD).  threads per block: 0 (the default, 64) or 64,
 * 128, 256, 512; GPF_EINVAL when not admissible for the grid.  A barrier wait that exceeds
 * its limit ends the launch and the next synchronising call fails with GPF_EDEVICE. */
int gpf_solver_set_persistent(gpf_solver* s, int mode, int threads);
/* the number of steps the persistent step has run */
int gpf_solver_persistent_steps(gpf_solver* s, long* n);
/* the HIP stream the solver launches on (hipStream_t) */
void* gpf_solver_stream(gpf_solver* s);
/* TimeStepper.SetFilter (pf/euler.go:62-64, pf/rk4.go:138-140) with the filter as the
 * 1000-entry table of a pf.Vandeven (pf/vandeven.go:8-40; n must be 1000), evaluated as
 * Vandeven.Eval at |f| 2/pi; table NULL removes the filter */
int gpf_solver_set_filter(gpf_solver* s, const double* table, long n);

/* ---------------- pf.SDD (pf/sdd.go) ---------------- */

/* SDD.Init (pf/sdd.go:382-402): orientation = final - init, normalised, and the initial dimer
 * length = |final - init|; init and final hold every field's real values back to back
 * (n = nfields * nodes) */
int gpf_solver_sdd_init(gpf_solver* s, const double* init, const double* final, long n);
/* SDD.SetInitialOrientation (pf/sdd.go:404-420) */
int gpf_solver_sdd_set_orientation(gpf_solver* s, const double* orient, long n);
/* SDD's parameters: Alpha, TimeConstants.Orientation, TimeConstants.DimerLength,
 * InitDimerLength (a negative value keeps the current one) and MinDimerLength
 * (pf/sdd.go:84-130; NewSDD's defaults 0.5, 1, 1, -, 0) */
int gpf_solver_sdd_set_params(gpf_solver* s, double alpha, double tau_orientation, double tau_length,
                              double init_length, double min_length);
/* SDDMonitor (pf/sdd.go:25-52): out[0] MaxForce, [1] ForcePowerSpectrum, [2] MaxTorque,
 * [3] FieldNorm, [4] FieldNormChange, [5] the torque's projection on the orientation
 * (the reference's sanity check, sdd.go:268-275) */
int gpf_solver_sdd_monitor(gpf_solver* s, double* out);
/* the current orientation vector (SaveOrientation's data, pf/sdd.go:422-435) */
int gpf_solver_sdd_orientation(gpf_solver* s, double* out, long n);
/* SDD.DimerLength(t) (pf/sdd.go:356-363) */
int gpf_solver_sdd_dimer_length(gpf_solver* s, double t, double* length);

/* ---------------- field output (pf/fileIO.go, pf/xdmf.go) ---------------- */

#define GPF_IO_FLOAT64 0  /* Float64IO: real part, big-endian float64 (pf/fileIO.go:45-62, 87-95) */
#define GPF_IO_UINT8 1    /* Uint8IO: real part as uint8 between min and max (pf/fileIO.go:15-42, util.go:108-117) */

/* Float64IO.SaveFields / Uint8IO.SaveFields (pf/fileIO.go:28-42, 57-62) as a solver
 * callback: writes <prefix>_<field>_<epoch>.bin for every field.  Asynchronous: the device
 * packs the file bytes on the solver's stream, a copy on a second stream moves them to
 * pinned host memory and a host thread writes the files while the solver keeps stepping
 * (two snapshots in flight at most).  gpf_solver_flush_io waits for every queued file and
 * reports the first write error (the reference prints it, pf/fileIO.go:89-92). */
int gpf_solver_save_fields(gpf_solver* s, const char* prefix, int epoch, int format);
int gpf_solver_flush_io(gpf_solver* s);
/* CsvIO.SaveFields (pf/fileIO.go:103-115): <prefix>_<epoch>.csv through SaveCsv */
int gpf_solver_save_csv(gpf_solver* s, const char* prefix, int epoch);
/* SaveFloat64 / LoadFloat64 (pf/fileIO.go:65-95); gpf_load_float64 stores the value count
 * in *n and copies at most cap values (out may be NULL to query the count) */
int gpf_save_float64(const char* path, const double* data, long n);
int gpf_load_float64(const char* path, double* out, long cap, long* n);
/* SaveCsv (pf/fileIO.go:147-200); values holds nfields arrays of n doubles back to back */
int gpf_save_csv(const char* path, const char* const* names, int nfields, const double* values, long n,
                 const int* dims, int ndim);
/* WriteXDMF (pf/xdmf.go:163-173): CreateXDMF(fields, prefix, num, dims) encoded with a
 * four-space indent.  gpf_create_xdmf returns the text (indent NULL = four spaces, "" = no
 * indentation as a bare xml.Encoder), its length in *len, NUL-terminated in buf up to cap. */
int gpf_write_xdmf(const char* path, const char* const* fields, int nfields, const char* prefix, int num,
                   const int* dims, int ndim);
int gpf_create_xdmf(const char* const* fields, int nfields, const char* prefix, int num, const int* dims, int ndim,
                    const char* indent, char* buf, long cap, long* len);

/* The nonlinear solver of an ImplicitEuler stepper:ImplicitEuler.NonlinSolver, a
 * nonlin.NewtonKrylov (pf/implicitEuler.go:26-28, 221-229; defaults Maxiter 50,
 * StepSize 1e-3, Tol 1e-7, Stencil 6).  stencil is the number of points of the central
 * difference that approximates the Jacobian-vector product (2, 4, 6 or 8; 1 selects a
 * forward difference); restart is the GMRES restart length (1..64, default 30).  A step
 * has converged when max_i |F_i| < tol. */
int gpf_solver_set_newton_krylov(gpf_solver* s, int maxiter, double step_size, double tol, int stencil,
                                 int restart);
/* Statistics of the last ImplicitEuler step: Newton iterations, GMRES iterations, whether
 * it converged (the reference logs "Iterative solver did not converge" otherwise,
 * pf/implicitEuler.go:201-204) and the last max-norm residual.  Any pointer may be NULL. */
int gpf_solver_newton_stats(gpf_solver* s, int* newton_iters, int* krylov_iters, int* converged,
                            double* residual);

/* Reverse-mode gradient (no reference counterpart; it replaces nothing in gopf and
 * differentiates the Euler.Step of pf/euler.go:16-47 on the fused single-field class).
 * From the solver's current field c_0 it runs nsteps Euler steps, checkpointing the
 * state every checkpoint_every steps, and returns J = sum_x (c_T(x) - target(x))^2,
 * dJ/dc_0 (grad_field, n doubles, may be NULL) and dJ/ds for each named scalar s
 * (at most 8 names).  Earlier states are recomputed from the checkpoints.  On return
 * the solver holds c_T and its time has advanced by nsteps * dt. */
int gpf_solver_adjoint(gpf_solver* s, int nsteps, int checkpoint_every, const double* target, long n,
                       double* loss, double* grad_field, int nscalars, const char* const* scalar_names,
                       double* grad_scalars);

/* Device copy bandwidth probe: average milliseconds of one read+write copy of `bytes`
 * bytes with the library's own copy kernel (no reference counterpart: the measured
 * copy bandwidth the roofline is also reported against, and the known-size kernel
 * the HBM traffic counters are calibrated with). */
int gpf_device_copy_bandwidth(long bytes, int reps, int device, double* ms_per_copy);
/* Achievable-bandwidth probes over `bytes` bytes (at least 1 MiB; rounded down to whole
 * 16 KiB blocks): kind 0 copies with nontemporal 16-byte loads and stores (bytes read +
 * bytes written per pass), kind 1 only reads, kind 2 only writes.  Average milliseconds of
 * one pass over reps passes.  No reference counterpart: the measured rates the roofline
 * fractions are also reported against. */
int gpf_device_bandwidth(long bytes, int reps, int kind, int device, double* ms_per_pass);

/* ---------------- pf.FourierTransform (pfutil.FFTWWrapper) ---------------- */

/* pfutil.NewFFTW (pfutil/fftWrap.go:16-23).  The plan is host data; device buffers are
 * taken on the first transform. */
int gpf_fft_create(const int* dims, int ndim, int precision, int device, gpf_fft** out);
void gpf_fft_destroy(gpf_fft* f);
/* FFTWWrapper.FFT / IFFT (pfutil/fftWrap.go:26-39): in place on host complex128, unnormalised.
 * n is the number of complex values in data and must equal the plan's size (GPF_EINVAL
 * otherwise; FFTWWrapper copies min(len) values, pfutil/fftWrap.go:27,35). */
int gpf_fft_forward(gpf_fft* f, double* data, long n);
int gpf_fft_inverse(gpf_fft* f, double* data, long n);
/* FFTWWrapper.Freq (pfutil/fftWrap.go:57-74): out[ndim] */
int gpf_fft_freq(const int* dims, int ndim, long i, double* out);
/* FFTWWrapper.ConjugateNode (pfutil/fftWrap.go:78-95) */
long gpf_fft_conjugate_node(const int* dims, int ndim, long i);

/* ---------------- batched parameter sweeps ---------------- */

/* Many independent single-field simulations of uneven 2D sizes that share one model
 * (equation strings and scalar names) and differ in scalar values and initial fields:
 * the sweep the reference runs as one pf.Solver per sim (pf/solver.go:40-63), with each
 * sim stored under its own simID (pf/database.go:67-80).  shapes holds [D0, D1] per class,
 * counts the sims per class; sims are numbered class by class.  Fused single-field
 * Euler models only (the Cahn-Hilliard class). */
int gpf_batch_create(gpf_model* m, int nclasses, const int* shapes, const int* counts, double dt,
                     const gpf_options* o, gpf_batch** out);
void gpf_batch_destroy(gpf_batch* b);
int gpf_batch_num_sims(const gpf_batch* b);
int gpf_batch_set_scalar(gpf_batch* b, int sim, const char* name, double value);
int gpf_batch_set_field(gpf_batch* b, int sim, const double* data, long n);
int gpf_batch_get_field(gpf_batch* b, int sim, double* out, long n);
/* nsteps Euler steps of every sim; blocks of 10 steps run as one replayed HIP graph.  With
 * gpf_options.autotune, the first call times the launch layouts on scratch copies: lines per
 * block and transform schedule of each class, and the packing -- one row launch and one
 * update launch per class, or one grouped row launch and one grouped update launch that
 * advance every class (classes of power-of-two sides 64..1024) -- and keeps the fastest. */
int gpf_batch_propagate(gpf_batch* b, int nsteps);
int gpf_batch_synchronize(gpf_batch* b);
/* the launch layout: per class (rows lines per block, update lines per block, rows transform
 * schedule), then the threads per block of the grouped launches, 0 when each class launches
 * on its own; returns the number of values (3 per class + 1) */
int gpf_batch_tuning(gpf_batch* b, int* tws, int maxn);
/* applies a layout in gpf_batch_tuning's format instead of the tuner (n = 3 per class + 1) */
int gpf_batch_set_layout(gpf_batch* b, const int* v, int n);

/* ---------------- pf.Vandeven ---------------- */

/* NewVandeven (pf/vandeven.go:13-27): 1000-entry table */
int gpf_vandeven_table(int order, double* table);
/* Vandeven.Eval (pf/vandeven.go:30-40) */
double gpf_vandeven_eval(const double* table, double x);

#ifdef __cplusplus
}
#endif

#endif /* GPF_H */
