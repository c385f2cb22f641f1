/*
 * pfc_hip.h -- C ABI of the MI355X (gfx950) spectral PFC library.
 *
 * This is the drop-in boundary: a C++ host (the pfc::fft::IFFT adapter and
 * the device-resident PFC model in openpfc_amd/host/) or any FFI (ctypes,
 * cgo, JNI) binds these symbols.  No C++ or HIP types cross it: handles are
 * opaque pointers, device buffers are plain `void *`, streams are passed as
 * `void *` (a hipStream_t value; NULL is the default stream).
 *
 * Conventions
 *   - Every function returns 0 on success or a PFC_HIP_E* code; the message
 *     of the last failure on the calling thread is pfc_hip_last_error().
 *     No C++ exception crosses this boundary.
 *   - Transforms are enqueued on `stream` and return without waiting; the
 *     caller synchronises the stream before reading results on the host.
 *   - Layout is the reference's one-rank layout (r2c_direction = 0):
 *     real fields are nx*ny*nz values with x fastest; spectra are
 *     (nx/2+1)*ny*nz complex values (re, im interleaved), kx fastest.
 *   - Precision PFC_HIP_F64 uses double / double-complex buffers,
 *     PFC_HIP_F32 float / float-complex buffers.
 *
 * Supported extents: every axis length in {1,2,4,...,1024} or {12,24,48,96,192};
 * nx even (nx/2 also in that set).  Other lengths are rejected with
 * PFC_HIP_EUNSUPPORTED.  The transposed step layouts need power-of-two
 * extents of at least 128; other extents step in the natural layout.
 */
#ifndef PFC_HIP_H
#define PFC_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PFC_HIP_ABI_VERSION 1

enum {
  PFC_HIP_OK = 0,
  PFC_HIP_EINVAL = 1,       /* bad argument (null pointer, bad size, ...) */
  PFC_HIP_EUNSUPPORTED = 2, /* extent or model not supported */
  PFC_HIP_EDEVICE = 3,      /* HIP runtime error (launch, copy, ...) */
  PFC_HIP_ENOMEM = 4,       /* device allocation failed */
  PFC_HIP_ESTATE = 5,       /* call out of order (e.g. step before operators) */
  PFC_HIP_EIO = 6           /* file could not be opened, read or written */
};

enum { PFC_HIP_F64 = 0, PFC_HIP_F32 = 1 };

/*
 * Plan flag (pfc_hip_plan_create): with PFC_HIP_F32, the passes of the PFC
 * step keep FP32 fields and spectra in HBM and LDS but compute every
 * butterfly (twiddle product and in-register DFT) in FP64 from FP64
 * twiddles, rounding to FP32 once per radix pass.  No reference
 * counterpart: it lowers the FP32-mixed step's arithmetic error (the
 * mean-split mode, pfc_hip_pfc_set_mean).  Needs the transposed step
 * layout (power-of-two extents, every axis >= 128, nx >= 256); such plans run
 * the PFC model only, and their plain transforms, free energy and adjoint use
 * FP32 arithmetic.
 */
#define PFC_HIP_FLAG_F64_BUTTERFLIES 1

enum { PFC_HIP_MODEL_PFC = 0, PFC_HIP_MODEL_TUNGSTEN = 1 };

/* Number of parameters of PFC_HIP_MODEL_TUNGSTEN (see pfc_hip_pfc_set_operators). */
#define PFC_HIP_TUNGSTEN_NPARAMS 17

typedef struct pfc_hip_plan_s *pfc_hip_plan_t;

/* ABI version of the loaded library (PFC_HIP_ABI_VERSION). */
int pfc_hip_abi_version(void);

/* Message of the last failed call on this thread ("" if none). */
const char *pfc_hip_last_error(void);

/*
 * Create a transform plan for an nx*ny*nz grid on the current device.
 * Replaces the construction of heffte::fft3d_r2c<rocfft> in
 *   src/openpfc/runtime/hip/fft_hip.cpp:19-37 (create_hip) and
 *   src/openpfc/runtime/cpu/fft.cpp:114-121,131-191 (create / create_with_backend)
 * at one rank: inbox [0,nx-1]x[0,ny-1]x[0,nz-1], outbox [0,nx/2]x[0,ny-1]x[0,nz-1].
 * flags: 0, or PFC_HIP_FLAG_F64_BUTTERFLIES with PFC_HIP_F32.
 */
int pfc_hip_plan_create(int nx, int ny, int nz, int precision, int flags,
                        pfc_hip_plan_t *plan);

/*
 * A plan over `batch` independent fields of one nx*ny*nz grid, stored one
 * after another (field b starts at b*nx*ny*nz values), stepped together by
 * pfc_hip_pfc_step / _step_timed / _autotune.  In the natural layout one
 * launch per pass covers every field; in the transposed layout (the
 * default where the extents have it) the fields are stepped one after
 * another, every pass of a field before the next field, so that a field's
 * spectra stay in the last-level cache between its passes; _step_timed then
 * reports the passes of field 0.  Each field has its own eps and mean
 * (pfc_hip_pfc_set_members).  Entry points that act on one field (r2c, c2r,
 * initial and boundary conditions, operator tables, the adjoint, the
 * Tungsten model, set_mean) return PFC_HIP_EUNSUPPORTED on it; sizes are
 * per field.  No reference counterpart: the reference steps one Model per
 * Simulator (docs/api/examples/10_complete_pfc_model.cpp:121-144).
 */
int pfc_hip_plan_create_batched(int nx, int ny, int nz, int batch, int precision, int flags,
                                pfc_hip_plan_t *plan);

/* Release a plan and every device buffer it owns. NULL is accepted. */
int pfc_hip_plan_destroy(pfc_hip_plan_t plan);

/*
 * Element counts of the real box, the complex box and the transform
 * workspace (complex elements).  Replaces IFFT::size_inbox/size_outbox/
 * size_workspace (include/openpfc/kernel/fft/fft_interface.hpp:66-71).
 */
int pfc_hip_sizes(pfc_hip_plan_t plan, size_t *inbox, size_t *outbox,
                  size_t *workspace);

/* Device bytes currently owned by the plan (tables, workspace, step state).
 * Replaces IFFT::get_allocated_memory_bytes (fft_interface.hpp:73). */
int pfc_hip_allocated_bytes(pfc_hip_plan_t plan, size_t *bytes);

/*
 * Forward R2C, unnormalised: out[k] = sum_x in[x] exp(-2 pi i k.x/n).
 * in: inbox real values; out: outbox complex values (device pointers).
 * Replaces FFT_Impl<rocfft>::forward(DataBuffer, DataBuffer)
 * (include/openpfc/kernel/fft/detail/fft_heffte_backend.hpp:146-170).
 */
int pfc_hip_r2c(pfc_hip_plan_t plan, const void *in, void *out, void *stream);

/*
 * Backward C2R scaled by 1/(nx*ny*nz) (heffte::scale::full).  `in` is not
 * modified; the plan's workspace holds the intermediate spectrum.
 * Replaces FFT_Impl<rocfft>::backward(DataBuffer, DataBuffer)
 * (fft_heffte_backend.hpp:200-227).
 */
int pfc_hip_c2r(pfc_hip_plan_t plan, const void *in, void *out, void *stream);

/*
 * Build the ETD1 operators of a spectral model on the device and allocate
 * the step state.  Replaces the operator precomputation of the model's
 * initialize(dt):
 *   model_kind PFC_HIP_MODEL_PFC, params = {dx, dy, dz, eps} (nparams 4):
 *     opL = exp(L dt), opN = (opL-1)/L * (-k^2) (dt*(-k^2) when |L| <= 1e-14),
 *     L = -k^2 (eps + (1-k^2)^2)
 *     (docs/api/examples/10_complete_pfc_model.cpp:166-213).
 *   model_kind PFC_HIP_MODEL_TUNGSTEN, params (nparams 17) =
 *     {dx, dy, dz, alpha2, lambda2, alpha_farTol, alpha_highOrd, Bx, T, T0,
 *      stabP, p2_bar, q2_bar, p3_bar, p4_bar, q3_bar, q4_bar}
 *     with alpha2 = 2*alpha^2 and lambda2 = 2*lambda^2 (the OperatorParams
 *     of apps/tungsten/include/tungsten/common/tungsten_spectral.hpp:55-73
 *     followed by the nonlinear coefficients): the mean-field filter
 *     fMF = exp(kLap/lambda2), exp(L dt) and n_weight = kLap*phi1(L) with
 *     L = kLap*opCk (tungsten_spectral.hpp:83-104,
 *     tungsten_etd_workspace.hpp:99-130).  Replaces
 *     Tungsten::prepare_operators (cpu/tungsten_model.hpp:223-279).
 * The step evaluates these per mode on the fly; nothing is tabulated.
 */
int pfc_hip_pfc_set_operators(pfc_hip_plan_t plan, int model_kind,
                              const double *params, int nparams, double dt);

/*
 * Advance psi (device, inbox real values, updated in place) by nsteps ETD1
 * steps: psi_hat <- opL*F(psi) + opN*F(psi^3); psi <- F^-1(psi_hat).
 * Replaces PFCModel::step (10_complete_pfc_model.cpp:121-144) called
 * nsteps times.  Requires pfc_hip_pfc_set_operators.  With nsteps >= 2 in
 * the transposed layout and role 16 set to 1, the inverse x pass of each
 * step and the forward x pass of the next run as one fused pass that keeps
 * psi on chip (role 15), so psi is written once at the end: 2R fewer bytes
 * per step boundary.  The fused pass may pair different transform plans
 * than the two passes it replaces, so such a call agrees with repeated
 * single steps to rounding, not bit for bit.
 */
int pfc_hip_pfc_step(pfc_hip_plan_t plan, void *psi, int nsteps, void *stream);

/*
 * Advance psi (device, inbox real values, in place) by nsteps Tungsten
 * mean-field ETD1 steps (requires the PFC_HIP_MODEL_TUNGSTEN operators):
 *   psi_MF = F^-1(fMF * F(psi)),
 *   N = p3*psi^2 + p4*psi^3 + q3*psi_MF^2 + q4*psi_MF^3 - stabP*psi,
 *   psi <- F^-1(exp(L dt)*F(psi) + n_weight*F(N)).
 * psiMF (device, inbox real values, may be NULL) receives psi_MF of the last
 * step's input field, as TungstenHIP::get_psiMF holds after step(); it must
 * not alias psi.  pfc_hip_pfc_step on a Tungsten plan is this call with
 * psiMF = NULL.  Replaces Tungsten::step / TungstenHIP::step
 * (apps/tungsten/include/tungsten/cpu/tungsten_model.hpp:306-366;
 * hip/tungsten_model.hpp:263-303).
 */
int pfc_hip_tungsten_step(pfc_hip_plan_t plan, void *psi, int nsteps, void *psiMF,
                          void *stream);

/*
 * Adjoint of one plain PFC ETD1 step taken from psi_n (device, inbox reals):
 * lambda (device, inbox reals, the adjoint field after the step) becomes
 * A lambda + 3 psi_n^2 (B lambda), A = F^-1 opL F and B = F^-1 opN F (both
 * symmetric: opL and opN are real and even in k).  With dr != NULL,
 * <lambda, d step(psi_n) / d eps> is added to *dr first (the stream is
 * synchronised), the derivative taken through opL and opN.  Chained from the
 * last step back, this gives the gradients of a loss on psi_T with respect
 * to psi_0 and eps.  Requires the PFC_HIP_MODEL_PFC operators; lambda must
 * not alias psi_n.  In the mean-split mode (pfc_hip_pfc_set_mean) psi_n holds
 * the deviation d and the adjoint is taken at psi = mean + d: the same step
 * map, so lambda stays the gradient with respect to the full field.  No
 * reference counterpart.
 */
int pfc_hip_pfc_adjoint_step(pfc_hip_plan_t plan, const void *psi_n, void *lambda,
                             double *dr, void *stream);

/*
 * Free energy of psi under the plain PFC model of the plan's operators,
 *   F = dV sum_x [ 1/2 psi (eps + (1 + lap)^2) psi + 1/4 psi^4 ],
 * the functional whose conserved gradient flow is the step's PDE
 * (L = -k^2 (eps + (1-k^2)^2), N = psi^3), evaluated spectrally and summed in
 * double with a fixed order.  In the mean-split mode psi holds the deviation
 * and the mean's share is formed in double from the plan's exact mean.  The
 * stream is synchronised.  No reference counterpart (the reference computes
 * no free energy).
 */
int pfc_hip_pfc_free_energy(pfc_hip_plan_t plan, const void *psi, double *F, void *stream);

/*
 * Mean-split step (the FP32-mixed mode): with enable != 0 the field passed
 * to pfc_hip_pfc_step holds the deviation d = psi - mean, and `mean` (kept in
 * double by the plan) is the conserved average of psi.  The step then forms
 * the nonlinearity as d*(3 mean^2 + (3 mean + d)*d) = psi^3 - mean^3 and keeps
 * the k = 0 mode of d at zero, so <psi> = mean is conserved exactly and the
 * precision of the field is spent on the deviation.  Mathematically the same
 * step as PFCModel::step (opN(0) = 0 there).  enable = 0 restores the plain
 * step.  No reference counterpart.
 */
int pfc_hip_pfc_set_mean(pfc_hip_plan_t plan, int enable, double mean);

/*
 * Per-field eps and mean of a plan's fields (after pfc_hip_pfc_set_operators
 * with the PFC model, whose spacing and dt all fields share).  eps: `batch`
 * values, or NULL to keep the operators' eps for every field.  mean_split
 * != 0 selects the mean-split step of pfc_hip_pfc_set_mean for every field,
 * with mean[b] the conserved average of field b (then not NULL).  On a plan
 * with batch > 1 eps and the means live in device memory read by the
 * natural-layout kernels and are launch arguments of the transposed ones;
 * on a plan with batch 1 they are the plan's scalars.  An ensemble
 * (pfc_hip_ensemble_step) captures its graph again when a group plan changed
 * anything its launches take as an argument (scalars, the mean-split mode,
 * variants), so later replays always use the values set here.  No reference
 * counterpart.
 */
int pfc_hip_pfc_set_members(pfc_hip_plan_t plan, const double *eps, const double *mean,
                            int mean_split);

/*
 * Measurement variant of pfc_hip_pfc_step: same work, with HIP events
 * around every kernel of the step.  On return (the stream is synchronised)
 * kernel_ms[i] holds the mean duration in milliseconds of pass i over the
 * nsteps steps.  PFC (nkernels >= 5): 0 forward-x (psi, psi^3), 1 forward-y
 * (both spectra), 2 z (forward, combine, inverse), 3 inverse-y, 4 inverse-x.
 * Tungsten (nkernels >= 9): 0 forward-x (psi), 1 forward-y, 2 z (forward,
 * mean-field filter, inverse), 3 inverse-y, 4 x (inverse, N, forward),
 * 5 forward-y, 6 z (forward, combine, inverse), 7 inverse-y, 8 inverse-x.
 * PFC with nkernels >= 6 and nsteps >= 2 where pfc_hip_pfc_step fuses step
 * boundaries: the launches of that fused run, kernel_ms[0] the forward x of
 * the first step, 1..3 the mean y, z and inverse y passes, 4 the inverse x of
 * the last step and 5 the mean fused boundary pass (otherwise kernel_ms[5]
 * = 0).  No reference counterpart (profiling aid).
 */
int pfc_hip_pfc_step_timed(pfc_hip_plan_t plan, void *psi, int nsteps,
                           void *stream, double *kernel_ms, int nkernels);

/*
 * Initial conditions applied on the device to a real field of the plan's
 * inbox (x = origin + i*spacing, x fastest).  All enqueue on `stream`.
 *
 * pfc_hip_ic_constant: psi = value everywhere.
 *   Replaces Constant::apply
 *   (include/openpfc/kernel/simulation/initial_conditions/constant.hpp:82-89).
 * pfc_hip_ic_single_seed: one-mode bcc seed rho + sum_q 2*amp*cos(q.x), q the
 *   six <110>/sqrt(2) vectors, inside x^2+y^2+z^2 < radius^2.  Outside it
 *   writes 0.0 when keep_outside == 0 (the reference behaviour, where radius
 *   is 64.0) or leaves psi unchanged when keep_outside == 1.
 *   Replaces SingleSeed::apply (single_seed.hpp:51-83).
 * pfc_hip_ic_add_noise: psi += lo + (hi-lo)*u_i with
 *   u_i = (splitmix64(splitmix64(seed) + i) >> 11) * 2^-53, i the linear
 *   index.  No reference counterpart (the reference has no noise IC); the
 *   stream is counter-based so any host can reproduce it exactly.
 */
int pfc_hip_ic_constant(pfc_hip_plan_t plan, void *psi, double value,
                        void *stream);
int pfc_hip_ic_single_seed(pfc_hip_plan_t plan, void *psi, const double *origin3,
                           const double *spacing3, double rho, double amp,
                           double radius, int keep_outside, void *stream);
int pfc_hip_ic_add_noise(pfc_hip_plan_t plan, void *psi, unsigned long long seed,
                         double lo, double hi, void *stream);

/*
 * Fixed boundary condition on the device.  With x = x0 + i*dx and
 * xpos = nx*dx - xwidth, every point with |x - xpos| < xwidth becomes
 * rho_low*S + rho_high*(1 - S), S = 1/(1 + exp(-alpha*(x - xpos))); other
 * points are unchanged.  Replaces FixedBC::apply
 * (include/openpfc/kernel/simulation/boundary_conditions/fixed_bc.hpp; the
 * reference uses xwidth = 20, alpha = 1).
 */
int pfc_hip_bc_fixed(pfc_hip_plan_t plan, void *psi, double x0, double dx, double rho_low,
                     double rho_high, double xwidth, double alpha, void *stream);

/*
 * The field of a list of rotated seeds on the device.  seeds (host memory)
 * holds nseeds rows of 24 doubles: location (3), the six rotated wave
 * vectors q_i (18), rho, radius, amplitude.  At X = origin + i*spacing the
 * first seed with |X - location|^2 < radius^2 gives
 * rho + sum_i 2*amplitude*cos(q_i . X) (X absolute); outside every seed the
 * field becomes 0.0, or is left unchanged with keep_outside.  Replaces the
 * field::apply of SeedGrid and RandomSeeds
 * (include/openpfc/kernel/simulation/initial_conditions/seed_grid.hpp:108-118,
 * random_seeds.hpp:80-89) over Seed::is_inside / Seed::get_value
 * (initial_conditions/seed.hpp:124-141).  The host builds the list (the
 * seeds' rotations and their mt19937_64 draws).
 */
int pfc_hip_ic_seeds(pfc_hip_plan_t plan, void *psi, const double *origin3,
                     const double *spacing3, const double *seeds, int nseeds, int keep_outside,
                     void *stream);

/*
 * Column profile of MovingBC: xline[i] (host memory, nx doubles) = the
 * maximum of the field over every (y, z) at x index i.  Synchronises the
 * stream.  Replaces the xline loop and MPI_Reduce(MPI_MAX) of
 * MovingBC::apply at one rank
 * (include/openpfc/kernel/simulation/boundary_conditions/moving_bc.hpp:120-140).
 */
int pfc_hip_column_max(pfc_hip_plan_t plan, const void *psi, double *xline, void *stream);

/*
 * No counterpart in the reference: pfc_hip_column_max for callers that
 * cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  It allocates nothing:
 * the partial maxima live in a static device array of the library (4 MiB,
 * held by each GPU the library runs on until the process ends), and the
 * result goes to xline_dev, a device buffer of nx doubles the caller owns.
 * Asynchronous on `stream` (no synchronisation, no host copy).  The scratch
 * holds min(ny*nz, 256) partial rows of nx columns, so nx <= 2048; a larger
 * nx gives PFC_HIP_EINVAL.  Every call on one GPU shares that scratch:
 * calls that overlap on one GPU (different streams, no ordering between
 * them) race and give wrong maxima.
 */
int pfc_hip_column_max_noalloc(pfc_hip_plan_t plan, const void *psi, double *xline_dev,
                               void *stream);

/*
 * MovingBC's band on the device: with l = nx*dx and p = fmod(xpos, l), every
 * point whose x = x0 + i*dx lies within xwidth of p, or of p - l / p + l
 * where the band wraps round the periodic box, becomes
 * rho_low*S + rho_high*(1 - S), S = 1/(1 + exp(-alpha*d)), d the signed
 * distance; other points are unchanged.  Replaces MovingBC::fill_bc
 * (moving_bc.hpp:187-219); the host keeps the boundary position.
 */
int pfc_hip_bc_moving(pfc_hip_plan_t plan, void *psi, double x0, double dx, double xpos,
                      double rho_low, double rho_high, double xwidth, double alpha,
                      void *stream);

/*
 * Operator tables of the fused step, written in the outbox layout (W = nx/2+1
 * pitch) to opL and opN (device, outbox real values each), scaled by
 * 1/(nx*ny*nz).  The step itself evaluates them per mode on the fly with the
 * same device code; this entry point exists to inspect and test them.
 * Counterpart of m_operator_L / m_operator_N
 * (docs/api/examples/10_complete_pfc_model.cpp:166-213).
 */
int pfc_hip_pfc_operators(pfc_hip_plan_t plan, void *opL, void *opN, void *stream);

/*
 * Tungsten counterpart: filterMF, exp(L dt) and n_weight in the outbox
 * layout, each scaled by 1/(nx*ny*nz) as the step applies them.
 * Counterpart of filterMF and the TungstenEtdWorkspace weights
 * (apps/tungsten/include/tungsten/cpu/tungsten_model.hpp:223-279).
 */
int pfc_hip_tungsten_operators(pfc_hip_plan_t plan, void *filterMF, void *expL,
                               void *nW, void *stream);

/*
 * Launch variants.  Every pass of the transforms and of the step exists in
 * several compile-time shapes (radix plan, values per thread, column tile
 * width, rows or lines per block, LDS exchange mode), and the step has two
 * spectral layouts.  Roles:
 *   0 x forward (plain R2C)  1 x inverse (C2R and natural step)  2 y columns
 *   3 z columns (plain transforms)  4 natural step forward x
 *   5 natural step z pass  6 step layout (0 transposed T1/T2, 1 natural)
 *   7..11 transposed step: forward x, forward y, z pass, inverse y, inverse x
 *   12 natural Tungsten step: fused x pass (inverse, N, forward)
 *   13 transposed Tungsten step: forward x of psi
 *   14 transposed Tungsten step: fused x pass
 *   15 transposed PFC multi-step runs: fused step boundary (inverse x of a
 *      step with the forward x of the next)
 *   16 multi-step runs: 0 as single steps (default), 1 with fused step
 *      boundaries (transposed layout; the autotuner picks 1 when the fused
 *      pass takes under 95% of the time of the two it replaces)
 * The Tungsten step also uses roles 0, 1, 2, 5 (natural) and 8..11
 * (transposed).  Index 0 is the default of each role.  No reference counterpart: the
 * reference delegates shapes to rocFFT; its GPU autotuner only times 1-D
 * block sizes (include/openpfc/runtime/common/gpu_autotune.hpp:295-333).
 */
enum {
  PFC_HIP_ROLE_X_R2C = 0,
  PFC_HIP_ROLE_X_C2R = 1,
  PFC_HIP_ROLE_COL_Y = 2,
  PFC_HIP_ROLE_COL_Z = 3,
  PFC_HIP_ROLE_PFC_X = 4,
  PFC_HIP_ROLE_PFC_Z = 5,
  PFC_HIP_ROLE_STEP_LAYOUT = 6,
  PFC_HIP_ROLE_T_X_FWD = 7,
  PFC_HIP_ROLE_T_Y_FWD = 8,
  PFC_HIP_ROLE_T_Z = 9,
  PFC_HIP_ROLE_T_Y_INV = 10,
  PFC_HIP_ROLE_T_X_INV = 11,
  PFC_HIP_ROLE_X_TUNG = 12,
  PFC_HIP_ROLE_T_X_R2C = 13,
  PFC_HIP_ROLE_T_X_TUNG = 14,
  PFC_HIP_ROLE_T_X_FUSE = 15,
  PFC_HIP_ROLE_STEP_FUSE = 16,
  PFC_HIP_ROLE_COUNT = 17
};
int pfc_hip_variant_count(pfc_hip_plan_t plan, int role, int *count);
/* Name of a variant ("r16x16x4/E16/L8/split"), NULL if out of range. */
const char *pfc_hip_variant_name(pfc_hip_plan_t plan, int role, int index);
int pfc_hip_variant_get(pfc_hip_plan_t plan, int role, int *index);
/* Selecting a step layout releases the other layout's step buffers. */
int pfc_hip_variant_set(pfc_hip_plan_t plan, int role, int index);

/*
 * Time every variant of every step role on full steps of psi (`reps` timed
 * steps after one warm-up each; the three fastest of a role are timed twice
 * more, round-robin), keep the fastest per role, then keep the faster step
 * layout.  psi is restored before return (the stream is synchronised).
 * times_ms (may be NULL) receives the fastest time per step of the passes
 * each variant runs, for every variant tried, roles in the order
 * 4, 2, 5, 1, 7, 8, 9, 10, 11, 15 (PFC) or 0, 2, 5, 12, 1, 13, 8, 9, 10, 14, 11
 * (Tungsten); roles of an unsupported layout contribute no entries (role 15,
 * the fused step boundary, is timed on its own launches).  Then
 * follow the fastest step times of the transposed and of the natural layout (0 if
 * not run); up to max_times entries are written.
 */
int pfc_hip_pfc_autotune(pfc_hip_plan_t plan, void *psi, int reps, void *stream,
                         double *times_ms, int max_times);

/*
 * Raw binary field files (the reference's BinaryWriter / BinaryReader:
 * include/openpfc/frontend/io/binary_writer.hpp:76-205,
 * include/openpfc/kernel/simulation/binary_reader.hpp:79-200).  A file holds
 * the global array of global3[0]*global3[1]*global3[2] values, x fastest
 * (the MPI_ORDER_FORTRAN subarray layout), each a native little-endian f64
 * (kind PFC_HIP_IO_COMPLEX: f64 re, f64 im), no header.  `field` holds the
 * brick local3 at offset3 of that array, x fastest, in `precision` (FP32
 * values are widened to f64 on write and rounded on read).  The domain
 * must satisfy 0 <= offset and offset + local <= global with positive extents
 * (domain_geometry.hpp:36-60), else PFC_HIP_EINVAL.  truncate != 0 empties
 * the file first (the reference's single-writer MPI_File_set_size(0));
 * bricks of one global array written by several calls use truncate = 0
 * after the first.  A file shorter than the domain needs gives PFC_HIP_EIO
 * ("Short read ...").
 *
 * pfc_hip_field_write / _read: `field` is device memory; the data are staged
 * through pinned chunks, copies on `stream` overlapping the file IO; the
 * call returns when the file (or the device buffer) is complete.
 * pfc_hip_field_write_host / _read_host: `field` is host memory; no device
 * is needed.
 */
enum { PFC_HIP_IO_REAL = 0, PFC_HIP_IO_COMPLEX = 1 };
int pfc_hip_field_write(const char *path, const void *field, int precision, int kind,
                        const int *global3, const int *local3, const int *offset3,
                        int truncate, void *stream);
int pfc_hip_field_read(const char *path, void *field, int precision, int kind,
                       const int *global3, const int *local3, const int *offset3,
                       void *stream);
int pfc_hip_field_write_host(const char *path, const void *field, int precision, int kind,
                             const int *global3, const int *local3, const int *offset3,
                             int truncate);
int pfc_hip_field_read_host(const char *path, void *field, int precision, int kind,
                            const int *global3, const int *local3, const int *offset3);

/*
 * Ensembles: independent PFC runs of uneven sizes advanced together.  Run r
 * has extents ext[3r..3r+2] and params[6r..6r+5] = {dx, dy, dz, eps, dt,
 * mean}.  Runs with equal extents, spacing and dt form a group stepped by
 * one batched plan; the fields of all runs live in one caller-owned device
 * buffer, each group's fields consecutive (groups in order of their first
 * run, runs in order within a group).  No reference counterpart: the
 * reference advances one Model per Simulator
 * (include/openpfc/frontend/simulator.hpp:549-553).
 *
 * pfc_hip_ensemble_layout: offsets[r] (may be NULL) receives the index of
 * the first value of run r, *total (may be NULL) the values of the buffer.
 * Pure host computation.
 */
typedef struct pfc_hip_ensemble_s *pfc_hip_ensemble_t;
int pfc_hip_ensemble_layout(int nruns, const int *ext, const double *params, size_t *offsets,
                            size_t *total);
/*
 * Create the groups' plans over `fields` (the layout above, precision
 * PFC_HIP_F64 or PFC_HIP_F32).  mean_split != 0: every field holds the
 * deviation from its run's mean (the mean-split step of
 * pfc_hip_pfc_set_mean); 0: the plain step (means ignored).
 */
int pfc_hip_ensemble_create(int nruns, const int *ext, const double *params, int precision,
                            int mean_split, void *fields, pfc_hip_ensemble_t *ens);
int pfc_hip_ensemble_destroy(pfc_hip_ensemble_t ens);
/* Number of groups; plan, batch and first value of group g (any may be NULL). */
int pfc_hip_ensemble_groups(pfc_hip_ensemble_t ens, int *ngroups);
int pfc_hip_ensemble_group(pfc_hip_ensemble_t ens, int group, pfc_hip_plan_t *plan, int *batch,
                           size_t *offset);
/* pfc_hip_pfc_autotune on every group (the fields are restored). */
int pfc_hip_ensemble_autotune(pfc_hip_ensemble_t ens, int reps, void *stream);
/*
 * Advance every run by nsteps.  use_graph != 0: one step of all groups (on
 * parallel branches) is captured once into a hipGraph -- again after an
 * autotune, or when a group plan changed a launch argument since -- and each
 * step is one hipGraphLaunch on `stream`; 0: the same launches are enqueued
 * directly.  Per-run eps, means and the mean-split mode may be changed
 * between steps through pfc_hip_pfc_set_members on a group's plan.
 */
int pfc_hip_ensemble_step(pfc_hip_ensemble_t ens, int nsteps, int use_graph, void *stream);

/* ---- measurement aid ------------------------------------------------------
 * No reference counterpart.  A plain device-to-device streaming copy (one
 * 16-byte value per thread, one read and one write per byte): the bandwidth
 * the step's passes can be set against on the same GPU.  `bytes` must be a
 * multiple of 16 and both pointers 16-byte aligned (PFC_HIP_EINVAL
 * otherwise); src and dst must not overlap.  Asynchronous on `stream`.
 */
int pfc_hip_stream_copy(const void *src, void *dst, size_t bytes, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* PFC_HIP_H */
