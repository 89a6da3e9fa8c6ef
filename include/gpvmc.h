/* gpvmc.h - C ABI of libgpvmc_hip.so, the MI355X (gfx950) Metropolis-VMC hot path.
 *
 * Plain C types only; every call returns GPVMC_OK (0) or a negative error code, with a
 * message from gpvmc_last_error().  No exceptions cross the ABI.  The context owns all
 * device memory; host buffers passed in are caller-owned and copied.  One host thread
 * per context; several contexts may coexist.
 *
 * A context holds a batch of independent Markov chains ("walkers") of one lattice size
 * and channel.  Each walker is the state of one reference process: LatticeState
 * occupations, the two SlaterDeterminant inverse matrices, the log-magnitude/phase of
 * the amplitude (the BigComplex m_amp role), a Philox counter and FP64 estimator sums.
 *
 * Reference interfaces each entry point replaces (file:line in EPFL-LQM/gpvmc):
 *   gpvmc_set_orbitals    WaveFunction::init_matrices / MatEl (WaveFunction.cpp:27-37,
 *                         WaveFunction.h:61-66), restricted to occupied orbitals
 *   gpvmc_init_walkers    LatticeState::FockInit + SlaterDeterminant::Init
 *                         (LatticeState.cpp:84-113, SlaterDeterminant.cpp:60-92)
 *   gpvmc_init_random     the random start with redraws of vmc.cpp:221-242
 *                         (LatticeState::RanInit, LatticeState.cpp:37-82)
 *   gpvmc_walk            MetroMC::Walk (MetroMC.cpp:45-88) driving MetroMC::Step,
 *                         LatticeStepper::trystep/step (LatticeStepper.cpp:36-147),
 *                         SlaterDeterminant::VirtUpdate/Update (SlaterDeterminant.cpp:96-326)
 *                         and Quantity::measure of ProjHeis (ProjHeis.cpp:40-166),
 *                         StagMagnZ (StagMagnZ.cpp:12-43) and OverlapTrack (OverlapTrack.cpp:26-29)
 *   gpvmc_measure         Quantity::measure of the same estimators at the current state
 *   gpvmc_refresh         SlaterDeterminant::Init on the current configuration (FP64 LU
 *                         recomputation of the inverse; the reference never refreshes)
 *   gpvmc_read_ratio_table SlaterDeterminant::VirtUpdate single-swap ratios
 *   gpvmc_read_*          Quantity::save (ScalarQuantity.cpp:10-14, MatrixQuantity.cpp:12-17),
 *                         LatticeState::Getfs/Getpt (State.h:93-97), SlaterDeterminant::Amp
 */
#ifndef GPVMC_H
#define GPVMC_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GPVMC_OK 0
#define GPVMC_ERR_ARG -1
#define GPVMC_ERR_HIP -2
#define GPVMC_ERR_STATE -3
#define GPVMC_ERR_UNSUPPORTED -4

/* arithmetic of the inverse matrices and ratios */
#define GPVMC_FP64 0   /* plain FP64 (the parity mode) */
#define GPVMC_MIXED 1  /* FP32 inverse and ratios; FP64 LU refresh, log-det, sums */

/* estimator mask for gpvmc_walk */
#define GPVMC_MEAS_PROJHEIS 1u    /* ProjHeis local energy per site (ground state) */
#define GPVMC_MEAS_STAGMAGNZ 2u   /* StagMagnZ staggered magnetisation */
#define GPVMC_MEAS_OVERLAP 4u     /* OverlapTrack weight |amp|^2 (as log) */
#define GPVMC_MEAS_STATSPIN 8u    /* StatSpinStruct 5 x Ns^2 correlators */

typedef struct gpvmc_ctx gpvmc_ctx;

typedef struct {
    int32_t L;             /* lattice side; Ns = L*L sites, N = Ns/2 particles per flavour */
    int32_t n_walkers;     /* independent chains */
    int32_t n_groups;      /* parameter groups (distinct orbital tables), >= 1 */
    int32_t precision;     /* GPVMC_FP64 or GPVMC_MIXED */
    int32_t device;        /* HIP device ordinal */
    int32_t lanes_per_row; /* 0 = library default; else 1,2,4,8,16 (tuning).  Channel
                              contexts: 0 = register-resident kernel where the walker
                              fits, -1 = streamed kernel */
    uint64_t seed;         /* Philox4x32-10 key */
} gpvmc_config;

const char *gpvmc_last_error(void);
int gpvmc_version(void);

int gpvmc_create(const gpvmc_config *cfg, gpvmc_ctx **out);
int gpvmc_destroy(gpvmc_ctx *ctx);

/* Occupied-orbital table of one parameter group: occ[fl][site][p] = Phi_fl(orbital of
 * wavefunction particle p, site), complex interleaved, 2*Ns*N complex values. */
int gpvmc_set_orbitals(gpvmc_ctx *ctx, int32_t group, const double *occ);
/* walker -> parameter group map (n_walkers entries); default all 0 */
int gpvmc_set_groups(gpvmc_ctx *ctx, const int32_t *group_of_walker);

/* Initial occupations spins[w][site] in {0 = flavour 0 (up), 1 = flavour 1 (down)},
 * N of each.  Labels are assigned in increasing site order (State::InitFock), the inverse
 * is formed by FP64 LU with partial pivoting.  ok[w] = 1 when the amplitude is non-zero
 * (SlaterDeterminant::m_amp_ok), else 0.  `ok` may be NULL.  Resets the accumulators and
 * the Philox counters. */
int gpvmc_init_walkers(gpvmc_ctx *ctx, const int32_t *spins, int32_t *ok);

/* The random start of vmc.cpp:221-242 on the device: every walker draws its configuration
 * by LatticeState::RanInit (LatticeState.cpp:37-82: flavour 0 takes site floor(u Ns) of each
 * uniform until its particles sit on distinct sites, flavour 1 the rest), is initialised as
 * by gpvmc_init_walkers, and the walkers whose amplitude is zero draw again, up to
 * max_attempts draws in all.  Uniform i of walker w's draw t is Philox4x32-10 with key seed
 * and counter (i, t, w, 2), u = (x0 + 1/2) 2^-32 (the host's gpvmc::RanInit makes the same
 * draws).  ok[w] = 1 for a walker with a non-zero amplitude; attempts[w] = the draws it
 * used.  Either may be NULL.  Channel contexts draw their channel's particle numbers. */
int gpvmc_init_random(gpvmc_ctx *ctx, uint64_t seed, int32_t max_attempts, int32_t *ok, int32_t *attempts);

/* Sweep kernel variant of a ground-state context.  GPVMC_SWEEP_ONE_MOVE updates the ratio
 * tables after each accepted move.  GPVMC_SWEEP_WINDOWS runs delayed-update windows where the
 * kernel has them (mixed precision, N = 128, i.e. L = 16): up to 16 steps and 4 accepted
 * moves are decided on the window's sub-blocks of the tables, and the tables take one rank-m
 * update per window.  GPVMC_SWEEP_DEFAULT (0) is the library's choice for the shape (the
 * faster of the two as measured).  All take MetroMC::Step's decisions (MetroMC.cpp:90-103) on
 * the same uniforms; their tables differ by rounding.  Setting GPVMC_SWEEP_WINDOWS on a shape
 * without windows fails with GPVMC_ERR_UNSUPPORTED.  gpvmc_sweep_variant returns the variant
 * that runs (ONE_MOVE or WINDOWS), -1 for a null context.  No counterpart in the reference,
 * whose chain updates the inverse after every accepted move (SlaterDeterminant::Update,
 * SlaterDeterminant.cpp:270-326). */
/* The lanes_per_row values of the ground-state sweep shapes compiled for (precision, L),
 * the default first: writes at most `max` of them and returns how many there are (a
 * negative error code on bad arguments).  For launch-shape tuning; no counterpart in the
 * reference. */
int gpvmc_sweep_shapes(int32_t precision, int32_t L, int32_t *lanes_per_row, int32_t max);

#define GPVMC_SWEEP_DEFAULT 0
#define GPVMC_SWEEP_ONE_MOVE 1
#define GPVMC_SWEEP_WINDOWS 2
int gpvmc_set_sweep_variant(gpvmc_ctx *ctx, int32_t variant);
int gpvmc_sweep_variant(const gpvmc_ctx *ctx);

/* Recompute every walker's inverse and log-amplitude from its configuration in FP64
 * (drift refresh).  drift[w] (may be NULL) receives max|Ai_old - Ai_new| / max|Ai_new|. */
int gpvmc_refresh(gpvmc_ctx *ctx, double *drift);

/* Runs n_steps Metropolis steps on every walker (MetroMC::Walk(len=n_steps, meas)).
 * After step s (0-based within this call) the estimators in `mask` are measured when
 * meas > 0 and (meas == 1 or s % meas == 0).  Uniforms: Philox4x32-10 (key = seed,
 * counter = (step, walker)) when `uniforms` is NULL, else the replay stream
 * uniforms[w][s][4] (host array, values in [0,1)).  Optional traces (host arrays, NULL to
 * skip): noop/accept[w][s] (int8), ratio[w][s] (double), and per measurement
 * eloc[w][m] (complex), mz[w][m] (complex), logw[w][m] (double, ln |amp|^2).
 * Ground state: meas <= 2^30; a replayed or traced walk takes n_steps <= 2^30, a longer
 * untraced walk runs as launches of whole measurement intervals. */
int gpvmc_walk(gpvmc_ctx *ctx, int64_t n_steps, int64_t meas, uint32_t mask,
               const double *uniforms, int8_t *noop, int8_t *accept, double *ratio,
               double *eloc, double *mz, double *logw);

/* Measures the estimators in `mask` once at every walker's current configuration,
 * without stepping and without touching the accumulators (Quantity::measure outside a
 * walk).  Outputs (host, NULL to skip): eloc[w] complex, mz[w] complex, logw[w]. */
int gpvmc_measure(gpvmc_ctx *ctx, uint32_t mask, double *eloc, double *mz, double *logw);

/* Per-walker FP64 sums since the last gpvmc_reset_accumulators / init, 12 doubles per
 * walker: [0,1] sum E_loc (complex), [2] sum |E_loc|^2, [3] n_meas, [4,5] sum m_z,
 * [6] sum |m_z|^2, [7] steps, [8] accepted moves, [9] rejected proposals, [10] no-op
 * proposals, [11] sum ln|amp|^2. */
#define GPVMC_ACC_STRIDE 12
int gpvmc_read_accumulators(gpvmc_ctx *ctx, double *acc);
int gpvmc_reset_accumulators(gpvmc_ctx *ctx);
/* StatSpinStruct sums (GPVMC_MEAS_STATSPIN) since the last reset, as the reference's
 * MatrixQuantity m_vals (StatSpinStruct.cpp:12-14,77-81) before Quantity::save divides
 * them by the count: five Ns x Ns complex matrices szz, spm, smp, spp, smm per walker,
 * acc[k][5][Ns][Ns] complex for walkers w0 .. w0+nw-1, and n_meas[k] measurements.
 * spp and smm are zero for two flavours.  Either output may be NULL. */
int gpvmc_read_statspin(gpvmc_ctx *ctx, int32_t w0, int32_t nw, double *acc, double *n_meas);
/* The same sums added over the walkers of each parameter group: acc[g][5][Ns][Ns]
 * complex, n_meas[g] (the reduction over chains that the reference leaves to its
 * post-processing, which averages the per-rank files). */
int gpvmc_read_statspin_groups(gpvmc_ctx *ctx, double *acc, double *n_meas);

/* ---- channels (exciton states) -------------------------------------------------
 * A channel context walks on w = sum_k |a_k|^2 over the N_k states of a channel
 * (StagFluxTransExciton / StagFluxLongExciton; LatticeStepper::trystep,
 * LatticeStepper.cpp:81-95), keeping the inverse of its dominant ("current") state and
 * hopping the wavefunction to the largest |a_k|^2 after every accepted move
 * (LatticeStepper::step, :102-113; SlaterDeterminant::Update, Woodbury).  The graph is
 * WaveFunction::m_exc (WaveFunction.cpp:39-67) flattened: hop_n[c][k][fl] orbital hops
 * (at most 2) from state c to state k, hop_orb[c][k][fl][h] = (from, to) orbitals,
 * hopsign[c][k] = State::HopSign of that path on state c (State.cpp:86-115), and
 * state_orb[k] = the n_up then n_dn occupied orbitals of state k, each increasing.
 * ProjHeis is accumulated as the reference's Nexc > 1 matrices (ProjHeis.cpp:146-153):
 * H[k1][k2] = sum conj(a_k1) h_k2 / w and O[k1][k2] = sum conj(a_k1) a_k2 / w. */
typedef struct {
    int32_t n_up, n_dn;         /* particles per flavour, n_up + n_dn = L*L */
    int32_t nk;                 /* channel states */
    const int32_t *state_orb;   /* [nk][n_up + n_dn] */
    const int32_t *hop_n;       /* [nk][nk][2] */
    const int32_t *hop_orb;     /* [nk][nk][2][2][2] */
    const int32_t *hopsign;     /* [nk][nk] */
} gpvmc_channel;

/* StagFluxTransExciton / StagFluxLongExciton construction + SlaterDeterminant on them
 * (vmc.cpp:140-156).  The context's part arrays are [w][2][max(n_up, n_dn)]. */
int gpvmc_create_channel(const gpvmc_config *cfg, const gpvmc_channel *ch, gpvmc_ctx **out);
/* Several channels of equal n_up, n_dn and nk in one context (for instance one per
 * momentum q, each of which the reference runs as its own vmc process with --qx --qy,
 * vmc.cpp:140-156); walker w walks on channel w mod n_channels. */
int gpvmc_create_channels(const gpvmc_config *cfg, int32_t n_channels, const gpvmc_channel *chs,
                          gpvmc_ctx **out);
/* Full orbital table of one group: table[fl][orbital][site] complex, 2*Ns*Ns values
 * (WaveFunction::init_matrices / MatEl, WaveFunction.cpp:27-37). */
int gpvmc_set_orbital_table(gpvmc_ctx *ctx, int32_t group, const double *table);
/* ProjHeis sums ho[k][2 nk][nk] complex (rows 0..nk-1: H, nk..2nk-1: O) of walkers
 * w0 .. w0+nw-1, their current state and weight sum_k |a_k / a_current|^2; any output may
 * be NULL.  The measurement count is acc[3] of gpvmc_read_accumulators. */
int gpvmc_read_channel(gpvmc_ctx *ctx, int32_t w0, int32_t nw, double *ho, int32_t *cstate,
                       double *weight);

/* 1 if the context runs the register-resident channel kernel (each walker's inverse and
 * new-orbital rows held in VGPRs, chosen when lanes_per_row is 0 and the walker fits),
 * 0 for the streamed one (lanes_per_row -1, or a walker too large for registers). */
int gpvmc_channel_kernel(gpvmc_ctx *ctx, int32_t *register_resident);
/* Amplitude of every channel state at each walker's current configuration,
 * amp[w][k] = (ln|a_k|, arg a_k) with a_k the product of both flavours' determinants times
 * HopSign, as SlaterDeterminant::VirtUpdate returns it for k (SlaterDeterminant.cpp:96-268)
 * times the current amplitude (SlaterDeterminant::Amp). */
int gpvmc_read_amplitudes(gpvmc_ctx *ctx, double *amp);

/* ---- batches: contexts of any sizes and channels swept together --------------------
 * vmc.cpp runs one chain per process and has no counterpart; a batch is how walkers of
 * uneven lattice sizes and channels share one GPU.  Each member contributes one sweep
 * (n_steps steps with a measurement every `meas` steps, as gpvmc_walk without traces).
 * The first gpvmc_batch_sweep captures one sweep of every member into a hipGraph, each
 * member's kernels on a stream of its own forked from the batch's stream, so the members
 * run concurrently; later calls replay it.  Adding a member drops the graph.  Members
 * must not be walked through other calls while a sweep runs; refresh (gpvmc_refresh)
 * between sweeps as usual.  gpvmc_batch_sweep returns after the sweeps complete and
 * reports their device time.  A ground-state member takes n_steps, meas <= 2^30. */
typedef struct gpvmc_batch gpvmc_batch;
int gpvmc_batch_create(int32_t device, gpvmc_batch **out);
int gpvmc_batch_add(gpvmc_batch *b, gpvmc_ctx *ctx, int64_t n_steps, int64_t meas, uint32_t mask);
int gpvmc_batch_sweep(gpvmc_batch *b, int32_t n_sweeps, double *device_ms);
int gpvmc_batch_destroy(gpvmc_batch *b);

/* Current state: spins[w][site], labels part[w][2][N] (site of each particle),
 * logabs[w] = ln|amp|, phase[w] = arg(amp); any pointer may be NULL. */
int gpvmc_read_state(gpvmc_ctx *ctx, int32_t *spins, int32_t *part, double *logabs,
                     double *phase);
/* FP64 inverse matrices of walker w formed by the last gpvmc_init_walkers or
 * gpvmc_refresh, complex [2][N][N], Ai[fl][label][slot] = (A_fl^-1)(label, slot)
 * (the reference's m_mati, SlaterDeterminant.h:66, transposed to row-major). */
int gpvmc_read_inverse(gpvmc_ctx *ctx, int32_t walker, double *ai);
/* Current ratio tables of walker w as FP64 complex [2][N][N]:
 * G[fl][i][j] = det(A_fl with particle i moved to the site of particle j of the other
 * flavour) / det(A_fl) = sum_f Ai_fl[i][f] Phi_fl(orb_f, site_other(j)): the ratios
 * SlaterDeterminant::VirtUpdate computes for every single swap (SlaterDeterminant.cpp:96-268). */
int gpvmc_read_ratio_table(gpvmc_ctx *ctx, int32_t walker, double *g);

/* ---- dense linear algebra on caller device memory --------------------------------
 * Batched linalg::DetInv (linalg.cpp:9-44): for b < batch, the inverse of the complex
 * n x n matrix a[b] (column-major, element (i, j) at a[b][i + j n], complex interleaved)
 * into ai[b] in the same layout, logabs[b] = ln|det|, phase[b] = arg det and singular[b] =
 * 1 when |det|^2 < 1e-9 (the reference's test) or a pivot vanishes.  Gauss-Jordan
 * elimination with partial pivoting (largest |re| + |im| of the column, the zgetrf choice)
 * in FP64.  All array arguments are device pointers; `stream` is a hipStream_t (NULL: the
 * null stream) and the call returns once the work is queued on it.  1 <= n <= 256.  For
 * n > 128 the kernel works in a device workspace of gpvmc_det_inv_workspace_bytes(n) per
 * matrix, which the library allocates on first need and keeps (grown as needed) for later
 * calls on the device; calls that overlap on one device share it and race. */
int gpvmc_det_inv(int32_t device, int32_t n, int32_t batch, const double *a, double *ai,
                  double *logabs, double *phase, int32_t *singular, void *stream);
/* *bytes = workspace bytes per matrix of gpvmc_det_inv at size n (0 when none is needed) */
int gpvmc_det_inv_workspace_bytes(int32_t n, int64_t *bytes);
/* gpvmc_det_inv without any device allocation.  No counterpart in the reference: for
 * callers that cannot allocate (code under HIP graph capture, serving loops, harnesses
 * that forbid hipMalloc).  The workspace of n > 128 is a fixed 64 MiB device array per
 * device, held until the process ends: at most floor(64 MiB / workspace bytes) matrices per
 * call (81 at n = 200, 64 at n = 256); larger batches fail with GPVMC_ERR_ARG before any
 * work is queued.  Every call on a device shares that array, so calls that overlap on one
 * device (other streams, other threads) race; the kernel fills its workspace before
 * reading it, so nothing carries over between calls. */
int gpvmc_det_inv_noalloc(int32_t device, int32_t n, int32_t batch, const double *a, double *ai,
                          double *logabs, double *phase, int32_t *singular, void *stream);

/* Cumulative device time of the sweep / refresh kernels (HIP events on the context's
 * stream) and the number of launches since creation or the last reset. */
int gpvmc_kernel_time(gpvmc_ctx *ctx, double *sweep_ms, int64_t *sweep_launches,
                      double *refresh_ms, int64_t *refresh_launches);
int gpvmc_reset_kernel_time(gpvmc_ctx *ctx);

#ifdef __cplusplus
}
#endif
#endif
