/*
 * myql_capi.h -- C-ABI of the MI355X Monte Carlo engine (libmyql_hip.so).
 *
 * Plain C: POD structs, raw pointers and sizes, integer status codes. No
 * C++ or torch types cross this boundary. The C++20 drop-in headers under
 * include/myql/ (MonteCarloPricer, MultiLevelMonteCarloPricer, ...) marshal
 * their template parameters into these structs and rethrow status codes as
 * the exceptions the reference library throws.
 *
 * Each entry point replaces one reference interface:
 *   myql_mc_calculate        MonteCarloPricer<...>::calculate
 *                            (reference: include/myql/pricers/montecarlo/
 *                             MonteCarloPricer.hpp:95-481)
 *   myql_mc_calculate_batch  several MonteCarloPricer::calculate calls (one
 *                            per job) fused into one persistent launch; the
 *                            reference runs them one after another
 *                            (MonteCarloPricer.hpp:95, called per maturity in
 *                             benchmarks/test_smile_suite.cpp:72-73)
 *   myql_mlmc_calculate      MultiLevelMonteCarloPricer<...>::calculate
 *                            (reference: include/myql/pricers/mlmc/
 *                             MultiLevelMonteCarloPricer.hpp:46-241)
 *   myql_create/destroy      (no counterpart: the reference has no device
 *                            state; a context owns a HIP stream, cached
 *                            stream states and NCI tables)
 *   myql_fourier_calculate_batch  FourierPricer<...>::calculate, one call per
 *                            job (reference: include/myql/pricers/fourier/
 *                             FourierPricer.hpp:34-407)
 *   myql_tuning_report       launch shapes picked by MYQL_FLAG_AUTOTUNE (no
 *                            counterpart: the reference has no launch shape)
 *   myql_cv_train,           learned martingale control variate (no
 *   myql_mc_calculate_batch_cv  counterpart: README.md:332 lists control
 *                            variates as future work)
 *
 * Ownership: the caller owns every pointer it passes; the library never
 * frees or retains caller memory past the call. Calls are synchronous on the
 * context's HIP stream. One context per host thread; different contexts may
 * be used concurrently.
 */
#ifndef MYQL_CAPI_H
#define MYQL_CAPI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
enum {
  MYQL_OK = 0,
  MYQL_E_INVALID_ARG = 1, /* std::invalid_argument in the C++ layer */
  MYQL_E_NV_FELLER = 2,   /* std::runtime_error("SchemeNV violation: ...") */
  MYQL_E_HIP = 3,         /* HIP runtime failure (launch, copy, device) */
  MYQL_E_OOM = 4,         /* device allocation failed */
  MYQL_E_UNSUPPORTED = 5  /* combination not implemented */
};

/* ---- enums mirroring the reference's template parameters --------------- */
enum { MYQL_GREEK_NONE = 0, MYQL_GREEK_ESSENTIAL = 1, MYQL_GREEK_FULL = 2 };
/* VolSchemes.hpp: NullVolScheme, SchemeNV, SchemeNCI, SchemeExact */
enum {
  MYQL_SCHEME_NULL = 0,
  MYQL_SCHEME_NV = 1,
  MYQL_SCHEME_NCI = 2,
  MYQL_SCHEME_EXACT = 3
};
/* JumpPolicies.hpp: NoJumps, MertonJump, KouJump */
enum { MYQL_JUMP_NONE = 0, MYQL_JUMP_MERTON = 1, MYQL_JUMP_KOU = 2 };
/* instruments/options: European, Asian (Geo/Arith tracker), Lookback, Barrier */
enum {
  MYQL_INSTR_EUROPEAN = 0,
  MYQL_INSTR_GEO_ASIAN = 1,
  MYQL_INSTR_ARITH_ASIAN = 2,
  MYQL_INSTR_LOOKBACK = 3,
  MYQL_INSTR_BARRIER = 4
};
/* Payoffs.hpp: PayoffVanilla, PayoffCashOrNothing, PayoffAssetOrNothing */
enum { MYQL_PAYOFF_VANILLA = 0, MYQL_PAYOFF_CASH = 1, MYQL_PAYOFF_ASSET = 2 };
enum { MYQL_CALL = 0, MYQL_PUT = 1 };
enum { MYQL_BARRIER_UP = 0, MYQL_BARRIER_DOWN = 1 };
enum { MYQL_BARRIER_IN = 0, MYQL_BARRIER_OUT = 1 };
enum {
  MYQL_REBATE_NONE = 0,
  MYQL_REBATE_AT_HIT = 1,
  MYQL_REBATE_AT_MATURITY = 2
};
/* Arithmetic type of the path state. Sums are always FP64. */
enum { MYQL_FP64 = 0, MYQL_FP32 = 1 };
/* Stream layout: how the S per-lane xoshiro256++ streams are derived.
 *  JUMP   stream s = splitmix64-seeded state of `seed`, jumped s * 2^128.
 *  SEEDED stream s = state seeded with seed + s * 999983, the reference's
 *         per-OpenMP-thread seeding (MonteCarloPricer.hpp:191); with S equal
 *         to OMP_NUM_THREADS the draws equal the reference's.
 * In both layouts stream s prices the contiguous block of paths that an
 * OpenMP static schedule over S threads would give thread s.
 *  SOBOL  randomized quasi-Monte Carlo (no counterpart in the reference,
 *         whose README.md:332 lists quasi-Monte Carlo as future work): the
 *         d-th uniform a path draws is coordinate d of a Sobol' point
 *         (direction integers from primitive polynomials in order of
 *         degree, Bratley & Fox), XORed with a random 32-bit digital shift
 *         per coordinate, u = (x + 1/2) 2^-32. The call is priced as
 *         qmc_replicates independent shifts (default 16) of num_paths /
 *         qmc_replicates points each (a power of two keeps each replicate a
 *         complete net); every value is the mean over the replicates and its
 *         standard error is their standard deviation / sqrt(replicates).
 *         Paths must draw a bounded number of uniforms: SchemeNV or
 *         SchemeNCI factors (or none), no jumps or Merton jumps, and at most
 *         1,024 coordinates per path -- per step 1 for the stock, 1 per NV
 *         and 2 per NCI factor and 2 for a path-dependent instrument's
 *         Merton jump, plus 2 for a European's terminal jump -- else
 *         MYQL_E_UNSUPPORTED. */
enum { MYQL_STREAMS_JUMP = 0, MYQL_STREAMS_SEEDED = 1, MYQL_STREAMS_SOBOL = 2 };
/* Config flags.
 *  REFERENCE_NCI_ROWS  use the reference's SchemeNCI table semantics
 *      (VolSchemes.hpp:245-255, 342-360): n_max from v_max = max(v0, 3 theta),
 *      every u < 0.999 interpolated, the chi-squared law inverted exactly for
 *      u >= 0.999 and above the last row. Without the flag: rows up to
 *      n_ext = ceil(l + 6 sqrt(l)), l the Poisson intensity at
 *      v = max(v0, 25 theta) (at most 1023 rows), so the exact inversion --
 *      thousands of cycles, run by the whole wavefront -- is essentially never
 *      needed for the row count; u >= 0.999 from a tail table on
 *      w = -log(1 - u); and rows N >= 2 inverted exactly in the first grid
 *      interval u < 0.999/9999, where interpolating from the quantile-0 node
 *      at u = 0 lowers the row's mean by up to 3e-5 -- a per-step bias of
 *      the variance that grows as the step shrinks.
 *  AUTOTUNE  the first pricing call of a context that meets a path kernel
 *      (per path precision, factor count, lockstep width, path dependence)
 *      times its launch shapes -- 256, 128 or 64 streams per work unit
 *      (block size); for the lockstep European kernels also the strike loop
 *      over staged samples unrolled by 4 or 1 and, at 64-stream units, a
 *      variant compiled for four waves per SIMD; 4 or 1 persistent blocks per
 *      resident slot -- on trial launches of that call's own jobs with an
 *      eighth of each stream's samples, keeps the fastest for the context and
 *      uses it from then on. Streams, draws and paths are those of the
 *      default launch; only the FP64 summation order of the per-unit partial
 *      sums can differ (results agree to ~1e-15 relative). Without the flag
 *      every kernel runs its default shape (256, 4, 4). Paths per lane stay
 *      num_paths / num_streams: changing them would change the streams. */
enum { MYQL_FLAG_REFERENCE_NCI_ROWS = 1, MYQL_FLAG_AUTOTUNE = 2 };

/* ---- model (ModelParams.hpp:7-47, ASVJmodel.hpp:14-87) ----------------- */
typedef struct {
  double kappa, theta, sigma, rho, v0; /* HestonParams */
} myql_heston_params;

typedef struct {
  int32_t n_factors;   /* 0: Zero-, 1: Single-, 2: DoubleFactorModel */
  int32_t jump_kind;   /* MYQL_JUMP_* */
  double vol;          /* ZeroFactorModel::vol */
  myql_heston_params heston[2];
  /* MertonParams {lambda, mu, delta, -}; KouParams {lambda, p_up, eta1, eta2} */
  double jump[4];
} myql_model;

/* ---- instrument (European.hpp, Asian.hpp, Lookback.hpp, Barrier.hpp) ---- */
typedef struct {
  int32_t kind;           /* MYQL_INSTR_* */
  int32_t payoff;         /* MYQL_PAYOFF_* */
  int32_t option_type;    /* MYQL_CALL / MYQL_PUT */
  int32_t fixed_strike;   /* Asian/Lookback: 1 fixed strike, 0 floating */
  int32_t barrier_dir;    /* MYQL_BARRIER_UP / DOWN */
  int32_t barrier_action; /* MYQL_BARRIER_IN / OUT */
  int32_t rebate_timing;  /* MYQL_REBATE_* */
  int32_t is_strip;       /* 0: StrikeContainer=double, 1: std::vector */
  double maturity;
  double barrier, rebate, rebate_rate; /* BarrierOption B, rebate, r */
  double payoff_eps;      /* digital smoothing bandwidth; < 0 = automatic */
  double past_avg_log, past_avg_S, t_elapsed; /* in-flight Asian */
  double hist_min_S, hist_max_S;              /* in-flight Lookback (0 = none) */
  const double *strikes;  /* n_strikes strikes (host memory) */
  int64_t n_strikes;
} myql_instrument;

/* ---- MonteCarloConfig (MonteCarloPricer.hpp:16-26) + engine options ---- */
typedef struct {
  uint64_t num_paths;
  uint64_t time_steps;
  uint64_t seed;
  double fd_bump, vol_bump, t_bump, r_bump;
  int32_t precision;     /* MYQL_FP64 / MYQL_FP32 */
  int32_t stream_layout; /* MYQL_STREAMS_* */
  uint64_t num_streams;  /* S; 0 = library default (65536, at most num_paths) */
  uint32_t flags;        /* MYQL_FLAG_* */
  uint32_t qmc_replicates; /* MYQL_STREAMS_SOBOL: independent shifts; 0 = 16 */
} myql_mc_config;

/* ---- MonteCarloResult<Mode, T> (PricingTypes.hpp:46-97) ----------------
 * Every pointer is a caller-owned array of n_strikes doubles, or NULL to
 * skip that field. Fields the Greek mode does not produce are left alone. */
typedef struct {
  double *price, *price_std_err;
  double *delta, *delta_std_err;
  double *gamma, *gamma_std_err;
  double *vega[2], *vega_std_err[2];
  double *theta, *theta_std_err;
  double *rho, *rho_std_err;
  double *vanna[2], *vanna_std_err[2];
  double *charm, *charm_std_err;
  uint64_t total_steps;    /* simulated path-steps, all Greek variants */
  double device_ms;        /* GPU time of the pricing kernels (HIP events) */
} myql_mc_result;

/* One MonteCarloPricer::calculate call inside a batch. */
typedef struct {
  const myql_model *model;
  int32_t schemes[2];
  int32_t greek_mode;
  const myql_instrument *instrument;
  const myql_mc_config *config;
  double S0, r, q;
  myql_mc_result *result;
} myql_mc_job;

/* ---- MLMCConfig (MultiLevelMonteCarloPricer.hpp:16-30) -----------------
 * estimator (no counterpart in the reference, whose README.md:332-335 lists
 * Richardson-Romberg and ML2R as future work):
 *  MYQL_MLMC_GILES  the reference's multilevel estimator: sum of the level
 *      means, N_l from the level variances, levels added until the last
 *      level's mean / (2^alpha - 1) is below epsilon / sqrt 2
 *      (MultiLevelMonteCarloPricer.hpp:46-241).
 *  MYQL_MLMC_ML2R   multilevel Richardson-Romberg (Lemaire & Pages, 2017)
 *      on the same coupled levels (refiners n_l = 2^l): the level means are
 *      weighted by W_l = sum_{j >= l} w_j, where w solves the Vandermonde
 *      system sum_j w_j = 1, sum_j w_j n_j^(-alpha k) = 0 (k = 1..L) with
 *      alpha = the config's alpha (the weak order), so the bias terms
 *      h^alpha .. h^(alpha L) of the finest level cancel; N_l minimises the
 *      cost at variance epsilon^2 / 2 of the weighted sum; levels are added
 *      until two successive depths' estimates differ by less than
 *      epsilon / sqrt 2. L_min = L_max = 1 is the Richardson-Romberg
 *      extrapolation 2^alpha/(2^alpha - 1) P_fine - P_coarse/(2^alpha - 1). */
enum { MYQL_MLMC_GILES = 0, MYQL_MLMC_ML2R = 1 };

typedef struct {
  double epsilon;
  uint64_t L_min, L_max, N0, base_steps;
  double alpha, beta;
  int32_t adaptive_rates;
  uint64_t seed;
  double fd_bump;
  int32_t precision;     /* path arithmetic; level sums are FP64 */
  int32_t stream_layout;
  uint64_t num_streams;
  uint32_t flags;        /* MYQL_FLAG_* */
  int32_t estimator;     /* MYQL_MLMC_GILES (0) or MYQL_MLMC_ML2R */
} myql_mlmc_config;

/* MlmcResult<Mode, double> (PricingTypes.hpp:102-140). The arrays are
 * caller-owned with L_max + 1 entries (level_weights may be NULL); n_levels
 * reports how many are used. Under GreekMode::Essential and Full, delta
 * and gamma are the multilevel estimates of the spot-bumped payoffs'
 * central differences -- from the up and down payoff buffers the
 * reference's evaluate_level computes and drops
 * (MultiLevelMonteCarloPricer.hpp:321-324) -- on the price's level sample
 * sizes, with their standard errors; under None they are 0, and the
 * reference fills none of them in any mode (no counterpart). Full mode's
 * other Greeks are not estimated by either (the drop-in returns them as 0,
 * as the reference does). level_weights: the weight of each level's mean in
 * the estimate (1 for MYQL_MLMC_GILES). */
typedef struct {
  double price, price_std_err;
  uint64_t total_steps;
  int32_t n_levels;
  uint64_t *paths_per_level;
  double *mean_diffs, *var_diffs;
  double alpha_estimated, beta_estimated;
  double device_ms;
  double delta, delta_std_err, gamma, gamma_std_err;
  double *level_weights;
} myql_mlmc_result;

typedef struct myql_ctx myql_ctx;

/* Context on HIP device `device`; NULL on failure. */
myql_ctx *myql_create(int device);
void myql_destroy(myql_ctx *ctx);
/* Message of the last failed call on this context ("" if none). */
const char *myql_last_error(const myql_ctx *ctx);
/* Launch shapes the autotuner (MYQL_FLAG_AUTOTUNE) chose on this context, one
 * line per kernel with its trial times, written NUL-terminated into buf
 * (truncated to len bytes). No counterpart in the reference, whose OpenMP
 * loop has no launch shape. */
int myql_tuning_report(const myql_ctx *ctx, char *buf, int64_t len);

int myql_mc_calculate(myql_ctx *ctx, const myql_model *model,
                      const int32_t schemes[2], const myql_instrument *instr,
                      int32_t greek_mode, const myql_mc_config *cfg,
                      double S0, double r, double q, myql_mc_result *res);

int myql_mc_calculate_batch(myql_ctx *ctx, const myql_mc_job *jobs,
                            int64_t n_jobs);

/* myql_mc_calculate without device allocation. No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). Same arguments, kernels and results as myql_mc_calculate,
 * but the call's device memory (stream states, job blocks, work units,
 * queue counters, per-unit accumulators, partial sums) is carved from a
 * file-scope device array of 64 MiB that each GPU holds until the process
 * ends; it calls no hipMalloc or hipFree. MYQL_E_INVALID_ARG when the call
 * needs more than that (about 65,536 streams x 64 strikes in Full mode;
 * fewer streams or strikes fit), MYQL_E_UNSUPPORTED for SchemeNCI (its
 * chi-squared table is allocated). MYQL_FLAG_AUTOTUNE is ignored (the
 * context's tuned or default launch shape is used). All such calls on one
 * GPU share the array: calls that overlap on one GPU race. */
int myql_mc_calculate_noalloc(myql_ctx *ctx, const myql_model *model,
                              const int32_t schemes[2], const myql_instrument *instr,
                              int32_t greek_mode, const myql_mc_config *cfg,
                              double S0, double r, double q, myql_mc_result *res);

int myql_mlmc_calculate(myql_ctx *ctx, const myql_model *model,
                        const int32_t schemes[2], const myql_instrument *instr,
                        int32_t greek_mode, const myql_mlmc_config *cfg,
                        double S0, double r, double q, myql_mlmc_result *res);

/* ---- Gil-Pelaez Fourier pricer ------------------------------------------
 * FourierPricer<Model, EuropeanOption<Payoff>, Mode>::calculate(S0, r, q,
 * instr) (reference: include/myql/pricers/fourier/FourierPricer.hpp:34-407,
 * engines/fourier/FourierEngine.hpp, kernels/GilPelaezKernel.hpp). European
 * vanilla, cash-or-nothing and asset-or-nothing calls and puts, scalar or
 * strike strip, any ASVJ model; price, and per Greek mode delta, gamma,
 * vega[2], theta (dP/dT), rho, vanna[2], charm exactly as the reference
 * assembles them from its integrals. */
/* FourierEngine::Config (FourierEngine.hpp:8-11): start_bound_guess is the
 * first point of the integration-bound search (start / sqrt(T) for T < 1,
 * then x1.5 until both price integrands are below 1e-12). tolerance is the
 * reference's adaptive-Simpson target; the engine integrates by composite
 * Gauss-Legendre (64 geometric panels x 16 nodes), whose error is well below
 * it, so the field is accepted and not used. */
typedef struct {
  double tolerance;
  double start_bound_guess;
} myql_fourier_config;

typedef struct {
  const myql_model *model;
  const myql_instrument *instrument; /* kind MYQL_INSTR_EUROPEAN */
  int32_t greek_mode;
  const myql_fourier_config *config; /* NULL: {1e-8, 20} */
  double S0, r, q;
  myql_mc_result *result; /* price and Greeks per strike; the *_std_err
                             arrays and total_steps are left alone */
} myql_fourier_job;

/* All jobs' strikes in one launch (one wavefront per option). */
int myql_fourier_calculate_batch(myql_ctx *ctx, const myql_fourier_job *jobs, int64_t n_jobs);

/* ---- learned martingale control variate ---------------------------------
 * No counterpart in the reference (its README.md:332 lists control variates
 * as future work). A network u(t, log S, v1, v2, T) -> R^64 gives, for
 * every strike K_k of the strip at once (a 64-output head over the strikes,
 * shared by the call's maturities), u_k(t, X) ~ E[e^{-r(T-t)} P_k | X_t]:
 * an MLP 5 -> 64 -> 64 -> 64 -> 64 with SiLU activations (12,864
 * parameters), trained in FP32 on engine paths and evaluated on FP16 MFMA
 * (FP32 accumulate) in forward mode. Per path and strike the estimator is
 *   P_k - c_A M_A - c_V M_V - c_J M_J - c_D D
 * with zero-mean terms
 *   M_A = sum_j e^{r(T-t_j)} du_k/dx (t_j, X_j) A_j          (stock noise)
 *   M_V = sum_j e^{r(T-t_j)} sum_f (du_k/dx C2_f + du_k/dv_f) V_f,j
 *                                      (variance noise and its stock part)
 *   M_J = e^{r(T-t_last)} du_k/dx (e^J - 1)      (the European's one jump)
 *   D   = P_k - E[P_k | S_T-]  (Merton jumps, vanilla and sharp digitals:
 *                               the conditional expectation in closed form)
 * where A_j, V_f,j sum the stepper's own increments with exact zero
 * conditional mean over `points` control-point blocks (x = log S, C2_f the
 * log-step's loading on factor f's new variance). The coefficients c are
 * fitted per (maturity, strike) by least squares on the training paths, so
 * they are independent of the pricing paths and the estimator is unbiased.
 * Europeans, GreekMode None or Essential (smoothed digitals; delta and gamma
 * are the plain estimates of the same paths); all jobs of one call share
 * the model, schemes, payoff, option type, strike strip (<= 64 strikes),
 * Greek mode and (S0, r, q); at most 16 jobs; default NCI semantics. */
enum { MYQL_CV_NPARAM = 12864, MYQL_CV_MAX_POINTS = 16, MYQL_CV_MAX_JOBS = 16, MYQL_CV_NREG = 4 };

typedef struct {
  uint64_t train_paths;  /* paths per job in the training run */
  uint64_t train_seed;   /* seed of the training paths and minibatches */
  uint32_t steps;        /* Adam steps */
  uint32_t batch;        /* rows per Adam step (multiple of 32) */
  double lr;             /* Adam learning rate */
  uint32_t points;       /* control points per path (1..16) */
  uint32_t init_seed;    /* weight initialisation */
} myql_cv_config;

typedef struct {
  /* W1[64][5], b1[64], W2[64][64], b2[64], W3[64][64], b3[64], W4[64][64], b4[64] */
  float params[MYQL_CV_NPARAM];
  float norm[6];       /* feature scaling fixed by the training call */
  int32_t n_jobs, n_strikes;
  double maturity[MYQL_CV_MAX_JOBS];
  /* regressor coefficients per trained job (maturity): M_A, M_V, M_J, D */
  double coef[MYQL_CV_MAX_JOBS][MYQL_CV_NREG][64];
  double final_loss;   /* mean squared error of the normalised targets, last steps */
  double train_ms;     /* GPU time of recording, training and fitting */
} myql_cv_net;

/* Per job of myql_mc_calculate_batch_cv; every pointer is an n_strikes
 * array or NULL. */
typedef struct {
  double *plain_price, *plain_std_err; /* plain estimate on the same paths */
  double *var_ratio;                   /* plain variance / CV variance */
  int32_t *used_cv;                    /* 1: result holds the CV estimate; 0:
                                          CV variance not lower, plain kept */
} myql_cv_result;

/* Train a network for the option sets of `jobs` (their model, strikes and
 * maturities; their num_paths and seed are not used). */
int myql_cv_train(myql_ctx *ctx, const myql_mc_job *jobs, int64_t n_jobs, const myql_cv_config *cfg,
                  myql_cv_net *net);

/* Price each job (its own paths, seed, precision) with the control variate
 * of `net`: result->price / price_std_err hold the CV estimate per strike,
 * or the plain one where the CV does not lower the variance. The jobs' paths
 * are recorded in one grid and evaluated in another (per precision, factor
 * count and Greek mode), in chunks of at most 32 GiB of recordings;
 * result->device_ms is the whole call's device time. The context keeps the
 * recording buffers for later calls until myql_destroy. */
int myql_mc_calculate_batch_cv(myql_ctx *ctx, const myql_mc_job *jobs, int64_t n_jobs,
                               const myql_cv_config *cfg, const myql_cv_net *net,
                               myql_cv_result *cv_results);

/* ---- diagnostics used by the parity tests ------------------------------
 * First n_per_stream raw xoshiro256++ outputs of streams [0, n_streams) for
 * (seed, layout), written stream-major into out (host memory). The draws
 * come from the same device code the pricing kernels use. */
int myql_debug_stream_draws(myql_ctx *ctx, uint64_t seed, int32_t layout,
                            uint64_t n_streams, uint64_t n_per_stream,
                            uint64_t *out);
/* Device Acklam inverse normal (FP64 and FP32 paths) on n host inputs. */
int myql_debug_normcdfinv(myql_ctx *ctx, const double *p, int64_t n,
                          double *out64, float *out32);
/* Device Giles inverse Poisson CDF on n (lambda, U) pairs (FP64). */
int myql_debug_poissinv(myql_ctx *ctx, const double *lambda, const double *u,
                        int64_t n, double *out);
/* Debug: the Poisson counts the path kernels use for n draws of n_slots
 * (1 or 5) slots sharing one uniform u[i], intensities lambda[i * n_slots +
 * k] (slot 0 the base path, the others bumped): an FP32 first pass decides
 * Giles' exact quantile where u clears its error bounds, the bumped slots
 * follow from slot 0's margins, and the FP64 method (myql_debug_poissinv)
 * finishes the rest. out[i * n_slots + k] equals Giles' count for
 * (lambda[i * n_slots + k], u[i]). No counterpart in the reference. */
int myql_debug_poisson_counts(myql_ctx *ctx, const double *lambda, int32_t n_slots, const double *u, int64_t n,
                              int32_t *out);
/* Debug: the randomized quasi-Monte Carlo coordinates the path kernels
 * draw (MYQL_STREAMS_SOBOL) for replicate `replicate` of seed `seed`:
 * out[i * dims + d] is coordinate d (as the 32-bit integer x of
 * u = (x + 1/2) 2^-32) of point i < n_points when the points are spread
 * over n_streams lanes (a multiple of 64) as a pricing call spreads them.
 * dims <= 1,024, n_points <= 2^24. No counterpart in the reference. */
int myql_debug_qmc_points(myql_ctx *ctx, uint64_t seed, uint64_t replicate, int32_t dims, uint64_t n_points,
                          uint64_t n_streams, uint32_t *out);
/* Device chi-squared quantile (inverse regularised incomplete gamma). */
int myql_debug_chi2_quantile(myql_ctx *ctx, const double *dof, const double *p,
                             int64_t n, double *out);
/* Device trackers and payoffs on given log-price paths: logS holds n_paths
 * rows of n_steps + 1 values (log S at t = 0, dt, ..., T; dt = the
 * instrument's maturity / n_steps); per path and strike the base, spot-up
 * and spot-down payoffs of Instrument::calculate_to_buffer (rows of
 * n_strikes in base / up / dn) after the tracker's init, per-step update
 * and finalize (PathTrackers.hpp and the option headers; FP64). There is
 * no model, so no automatic smoothing bandwidth: a cash/asset digital under
 * MYQL_GREEK_ESSENTIAL or MYQL_GREEK_FULL needs payoff_eps > 0
 * (MYQL_E_INVALID_ARG otherwise). */
int myql_debug_path_payoffs(myql_ctx *ctx, const myql_instrument *instrument,
                            int32_t greek_mode, double S0, double fd_bump,
                            const double *logS, int64_t n_paths,
                            int64_t n_steps, double *base, double *up,
                            double *dn);
/* Fix the launch shape of one path kernel on this context, as if the
 * autotuner had chosen it: precision MYQL_FP32/FP64, factor count, lockstep
 * width (1 for replays; 5 / 7 for 0-1 / 2-factor lockstep Europeans; 7 / 9
 * path-dependent), path dependence, block (256, 128 or 64), unroll (4; or 1
 * for lockstep Europeans), grid factor (>= 1), four_wave (1: the variant
 * compiled for four waves per SIMD; lockstep Europeans at block 64 only).
 * MYQL_E_UNSUPPORTED for a shape that is not compiled. */
int myql_debug_set_launch_shape(myql_ctx *ctx, int32_t precision, int32_t n_factors,
                                int32_t lockstep_width, int32_t path_dependent,
                                int32_t block, int32_t unroll, int32_t grid_factor,
                                int32_t four_wave);
/* Control-variate network on n caller feature rows (n x 5, scaled
 * features): the FP16 MFMA forward-mode tangents of the 64 outputs along
 * f1, f2, f3 used by the pricing kernel (out: n x 64 x 3), and the FP32
 * training gradient of the mean over rows and outputs of (net - target)^2
 * (targets n x 64; n a multiple of 32; grad: MYQL_CV_NPARAM; loss: the
 * mean). */
int myql_debug_cv_input_grad(myql_ctx *ctx, const float *params, const float *feats,
                             int64_t n, float *out);
int myql_debug_cv_param_grad(myql_ctx *ctx, const float *params, const float *feats,
                             const float *targets, int64_t n, float *grad, double *loss);
/* The control variate's terminal-jump expectation E[P_k | S_T- = S]
 * (Merton or Kou jumps over the instrument's maturity, strikes k of the
 * instrument; vanilla and, in Greek mode NONE, sharp digital payoffs), at
 * n given spots: tab[i * n_strikes + k] as the estimator reads it from its
 * table on the log-spot grid [log S0 - 3, log S0 + 2], dir the same summed
 * directly, ddir the derivative in log S of its n >= 1 part.
 * *table_on = 0 (outputs untouched) where the estimator goes without the
 * term (smoothed digitals, jump counts the table does not cover). No
 * counterpart in the reference. */
int myql_debug_cv_jump_expectation(myql_ctx *ctx, const myql_model *model, const myql_instrument *instrument,
                                   int32_t greek_mode, double S0, const double *spots, int64_t n, double *tab,
                                   double *dir, double *ddir, int32_t *table_on);

#ifdef __cplusplus
}
#endif

#endif /* MYQL_CAPI_H */
