/*
 * C-ABI of the MI355X CT-HYB sampling core.
 *
 * This is the drop-in boundary for the Z-space sampling path of the reference
 * solver module MCTQMC (w2dynamics src/ctqmc_fortran/CTQMC.F90).  Each entry point
 * names the reference interface it replaces.  Conventions:
 *   - plain pointers and sizes only; arrays are column-major (Fortran order) as
 *     the reference receives them through f2py;
 *   - every call returns an int status (W2D_OK = 0), nothing is thrown across
 *     the boundary; the reference stops the process instead (MatrixUpdate.F90:69-98,
 *     CTQMC.F90:4062, Trace.F90:668);
 *   - one handle owns a batch of independent Markov chains on one GPU; chain c
 *     is seeded with seed0 + c, so chain 0 reproduces the reference's stream for
 *     init_rnd(seed0) and the batch plays the role of MPI ranks NSeed + rank
 *     (cthyb:144);
 *   - a handle is not thread-safe; one host thread drives it.
 */
#ifndef W2DYN_CTQMC_H
#define W2DYN_CTQMC_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define W2D_OK 0
#define W2D_ERR_ARG 1          /* invalid argument or unsupported parameter value */
#define W2D_ERR_HIP 2          /* HIP runtime error (allocation, launch) */
#define W2D_ERR_CAPACITY 3     /* expansion order or window length above the capacity */
#define W2D_ERR_SINGULAR 4     /* "DetRat too small in removal" (CTQMC.F90:4062) */
#define W2D_ERR_ZERO_TRACE 5   /* "Zero trace initial state" (Trace.F90:668) */
#define W2D_ERR_STATE 6        /* call out of order (e.g. results before measure) */
#define W2D_ERR_PARTIAL 7      /* some chains stopped on an error, the others completed:
                                  their results are valid (w2d_ctqmc_chain_status) */

typedef struct w2d_ctqmc w2d_ctqmc;

/* Per-move record used by the parity tests; the same fields the reference driver
 * reads after each Step* call (CTQMC.F90:3353-4835). */
typedef struct {
    int32_t kind;      /* 1 add, 2 remove, 3 global, 4 tau shift */
    int32_t outcome;   /* 0 rejected, 1 accepted, 2 rejected before the weight (QN / no pair) */
    int32_t noper;     /* operators in the configuration after the move */
    int32_t outer_sst; /* outer superstate after the move */
    int32_t window;    /* sliding-window position after the move */
    int32_t pad;
    double trace_log, trace_sign, det_log, det_sign, cfgsign;
} w2d_move_record;

/* Replaces init_paras (CTQMC.F90:9412) + init_solver (CTQMC.F90:7894) + init_rnd
 * (CTQMC.F90:9404) for a batch of chains.
 *   params     : the "key = value #\n" string (Parameters.F90:281-320)
 *   u_matrix   : (2N,2N,2N,2N), flavour index 2*(b-1)+s   (CTQMC.F90:7899)
 *   ftau_full  : (N,2,N,2,nftau), only the flavour-diagonal part is used
 *   muimp_full : (N,2,N,2)
 *   capacity   : initial pairs per flavour per chain (0 = default 256); doubled
 *                on demand up to 4096 (w2d_ctqmc_chain_status); a negative
 *                value fixes the capacity at -capacity (chains that outgrow
 *                it stop with W2D_ERR_CAPACITY, the others carry on)
 * Screening (Uw) is not supported; offdiag, segment, statesampling, worm and
 * four-operator moves must be off; so must the measurements this library does
 * not produce (MeasGiw, MeasG2iw, MeasGSigmaiw, FourPnt, MeasSusz, MeasSsquare,
 * MeasExpResDensityMatrix, MeasGtauDetRat, MeasurementTiming >= 0): W2D_ERR_ARG.
 * AccumGtau 1 (mean) and 2 (binning) are supported. */
int w2d_ctqmc_create(const char *params, int nbands, int nftau, const double *u_matrix,
                     const double *ftau_full, const double *muimp_full, int nchains,
                     uint32_t seed0, int capacity, int device, w2d_ctqmc **out);

/* Replaces set_taudiffmax (CTQMC.F90:8180): a new window for every chain.
 * After the chains have moved (the calibration of cthyb_solver.py:899-944)
 * their configurations, random streams and inverses are kept. */
int w2d_ctqmc_set_taudiffmax(w2d_ctqmc *h, double taudiff_max);

/* Replaces init_counters (CTQMC.F90:847-899): move counters, the accepted-pair
 * histogram AccPair and the index of the calibration list. */
int w2d_ctqmc_init_counters(w2d_ctqmc *h);

/* Replaces ctqmc_calibrate(list, phase1pairnum, phase2taugrid, phase2pairnum,
 * nprogress) (CTQMC.F90:8088-8178) for every chain: Metropolis steps as in
 * warm-up until, checked every min(1000, nprogress) steps, phase 1 (list != 0)
 * has filled the chain's AccPairTau list with phase1pairnum accepted-pair
 * distances, or phase 2 (list = 0) has phase2pairnum accepted pairs of every
 * flavour in AccPair over phase2taugrid bins of [0, taudiff_max].  Each chain
 * stops on its own condition; the call returns when all have. */
int w2d_ctqmc_calibrate(w2d_ctqmc *h, int list, int phase1pairnum, int phase2taugrid,
                        int phase2pairnum, int nprogress);

/* Accepted-pair statistics of one chain (AccPairTau, AccPair; CTQMC.F90:66-70):
 * accpairtau[*napt] the distances recorded by calibration phase 1, accpair
 * [2N][*naccpair] the accepted insertions and removals per flavour and
 * distance bin since the last init_counters (counts).  Pointers may be NULL. */
int w2d_ctqmc_get_pair_stats(w2d_ctqmc *h, int chain, double *accpairtau, int *napt, double *accpair,
                             int *naccpair);

/* MeasOuterHistograms (CTQMC.F90:3285-3339) and AccPair summed over groups of
 * chains as w2d_ctqmc_get_group_sums does, raw (before the division by the
 * measurement count and visits of CTQMC.F90:8757-8764):
 *   outer_state [ngroups][nstates]  OuterStateHisto (visits of each outer state)
 *   sign_sst    [ngroups][nsst]     SignHistoSuperstates (sum of cfgsign)
 *   sign_state  [ngroups][nstates]  SignHistoStates
 *   contrib_sst [ngroups][nsst]     TraceContribSuperstates
 *   contrib_state [ngroups][nstates] TraceContribStates
 *   accpair     [ngroups][2N][naccpair capacity]  (w2d_ctqmc_dims' ntau or 1024,
 *               whichever is larger; the first naccpair bins are used)
 * Any pointer may be NULL. */
int w2d_ctqmc_get_group_histograms(w2d_ctqmc *h, int ngroups, double *outer_state, double *sign_sst,
                                   double *sign_state, double *contrib_sst, double *contrib_state,
                                   double *accpair);

/* Step-resolved measurements of one chain (Gtau_mean_step, Gtau_mid_step,
 * sign_step; CTQMC.F90:2394-2402), raw, [*nmstep][2N] and [*nmstep], where
 * *nmstep = Nmeas when any of them is enabled, else 0.  Pointers may be NULL. */
int w2d_ctqmc_get_steps(w2d_ctqmc *h, int chain, int *nmstep, double *gtau_mean_step,
                        double *gtau_mid_step, double *sign_step);

/* Replaces ctqmc_warmup (CTQMC.F90:8188): nwarmups Metropolis steps per chain.
 * Returns W2D_ERR_PARTIAL if some (not all) chains stopped on an error. */
int w2d_ctqmc_warmup(w2d_ctqmc *h, int64_t nwarmups);

/* Replaces ctqmc_measure(1,1) (CTQMC.F90:8355): nmeas x (NCorr steps + Z-space
 * measurements) per chain; accumulates into the handle's per-chain buffers,
 * which are cleared first (init_counters, CTQMC.F90:847). */
int w2d_ctqmc_measure(w2d_ctqmc *h, int64_t nmeas);

/* Sizes of the result arrays. */
int w2d_ctqmc_dims(const w2d_ctqmc *h, int *nbands, int *nchains, int *ntau, int *nlegmax,
                   int *nsst, int *nstates, int *maxhisto);

/* Raw per-chain accumulators after measure, before the normalisation of
 * CTQMC.F90:8663-8771 (which the host applies, so that per-chain samples give
 * the across-chain error of DistributedSample, statistics.py:180-209).
 * Any pointer may be NULL.  Shapes (C order, chain slowest):
 *   gtau   [nchains][2N][ntau]    sum over measurements of the binned M^{-1}
 *   gleg   [nchains][2N][nlegmax]
 *   dm     [nchains][nstates][nstates]  beta/2 density matrix, eigenbasis,
 *          block-diagonal, column-major within a chain
 *   histo  [nchains][2N][maxhisto+1]
 *   outer  [nchains][nsst]
 *   scalars[nchains][14]: sum cfgsign, measurements, acc/try add, rem, glob, shift,
 *          attempted moves, qn-rejected moves, algorithmic flops, algorithmic bytes
 * flavour index f = b + N*s (orbital fastest). */
int w2d_ctqmc_get_accumulators(w2d_ctqmc *h, double *gtau, double *gleg, double *dm,
                               double *histo, double *outer, double *scalars);

/* The same accumulators summed over `ngroups` contiguous groups of chains
 * (group g = chains [g*nchains/ngroups, (g+1)*nchains/ngroups)), leaving out
 * chains stopped on an error; the sum runs on the GPU.  Shapes as above with
 * the group in place of the chain, except
 *   dm_blocks [ngroups][sum_s n_s^2]  the density matrix's diagonal blocks in
 *             superstate order, each n_s x n_s column-major
 *   nused     [ngroups]  chains summed into each group
 * Pooled sums give the sign-ratio estimate sum(G sign) / sum(sign) of the
 * normalisation CTQMC.F90:8663-8771 and, over groups, its jackknife error. */
int w2d_ctqmc_get_group_sums(w2d_ctqmc *h, int ngroups, double *gtau, double *gleg, double *dm_blocks,
                             double *histo, double *outer, double *scalars, int32_t *nused);

/* w2d_ctqmc_get_group_sums without any allocation: no counterpart in the
 * reference; for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops).  Same arguments and results.  Its work
 * area (ngroups x (accumulator doubles per chain + 15) x 8 bytes) lives in the
 * library's static 16 MiB device scratch; larger requests return
 * W2D_ERR_CAPACITY with a message on stderr.  The scratch is shared by every
 * call on one GPU: overlapping calls on one GPU race. */
int w2d_ctqmc_get_group_sums_noalloc(w2d_ctqmc *h, int ngroups, double *gtau, double *gleg,
                                     double *dm_blocks, double *histo, double *outer,
                                     double *scalars, int32_t *nused);

/* AccumGtau = 2 (autocorracc, Accumulators.F90:283-362): binning analysis of
 * one chain's raw G(tau) bins, blocks[l][2N][ntau] = variance of the means of
 * bins of 2^l measurements (gtau_blocks, CTQMC.F90:8733-8739); *nlevels =
 * ilog2(Nmeas) (blocks may be NULL to query it).  W2D_ERR_STATE unless the
 * handle was created with AccumGtau = 2. */
int w2d_ctqmc_get_gtau_blocks(w2d_ctqmc *h, int chain, int *nlevels, double *blocks);

/* MeasGiw (measure_giw_nfft, CTQMC.F90:1015-1059, with the MeasGiwFlushLookup
 * path's sign, CTQMC.F90:2403-2417): raw sums over the measurements of
 *   cfgsign * sum_{a,c} M[c,a] exp(i w_n (tau_a - tau_c)),  w_n = (2n+1) pi / beta,
 * for n = 0..Niw-1 (G(i w_{-n-1}) is the complex conjugate), evaluated as the
 * direct sum (the NFFT fallback of Nfft_z.F90:64-90) on the GPU after every
 * measurement.  The normalised Giw of CTQMC.F90:8816 is sum / (count x mean
 * sign x beta) with the scalars of the same chains.  ngroups = 0: per chain,
 * out[nchains][2N][Niw][2] (re, im); ngroups > 0: sums over the contiguous
 * groups of chains of w2d_ctqmc_get_group_sums (chains stopped on an error left
 * out), out[ngroups][2N][Niw][2].  *niw = Niw (0 when the handle was created
 * without MeasGiw); out = NULL queries it.  W2D_ERR_STATE with MeasGiw off or
 * before a measure.  With MeasGiw on, measure launches end on every
 * measurement (one launch per NCorr moves). */
int w2d_ctqmc_get_giw(w2d_ctqmc *h, int ngroups, int *niw, double *out);

/* Eigenbasis of the local Hamiltonian built by create: per-superstate sizes and
 * offsets, eigenvalues (EigenstatesEnergies), occupation-basis state of each
 * index (States), and the block eigenvectors (column-major, concatenated);
 * needed to transform the density matrix to the occupation basis
 * (Transform_DensityMatrix_EB_to_OCCB, CTQMC.F90:3163). */
int w2d_ctqmc_get_model(const w2d_ctqmc *h, int *sizes, int *offsets, double *evals,
                        int *states, double *evecs, double *egs);

/* Parity entry point: run n steps of chain `chain` alone (other chains idle) and
 * record one w2d_move_record per step. */
int w2d_ctqmc_moves_log(w2d_ctqmc *h, int chain, int64_t n, w2d_move_record *rec);

/* Replaces get_mc_config_scalars / get_mc_config (CTQMC.F90:9358-9376) for one
 * chain: returns the configuration (0-based orbital/spin, ca 1 = creator,
 * 2 = annihilator), weights[5] = trace log/sign, det log/sign, cfgsign. */
int w2d_ctqmc_get_config(w2d_ctqmc *h, int chain, int cap, int *n, double *tau, int *orb,
                         int *spin, int *ca, int *outer_sst, double *weights);

/* Replaces set_mc_config (CTQMC.F90:9378) for one chain: full trace, full
 * inverse hybridisation matrices and sign of the given configuration. */
int w2d_ctqmc_set_config(w2d_ctqmc *h, int chain, int n, const double *tau, const int *orb,
                         const int *spin, const int *ca, int outer_sst, double *weights);

/* Inverse hybridisation block of flavour f of one chain (column-major k x k,
 * rows creators, columns annihilators as Trace.F90:2279-2333); *k receives the
 * order, out may be NULL to query it. */
int w2d_ctqmc_get_minv(w2d_ctqmc *h, int chain, int f, int *k, double *out);

/* Kernel time of the last warmup/measure call in milliseconds (HIP events on
 * the handle's stream around its launches) and the number of launches. */
int w2d_ctqmc_last_kernel_ms(const w2d_ctqmc *h, double *ms, int64_t *launches);

/* Per-chain outcome of the last warmup / measure: status[c] = W2D_OK or the
 * error that stopped chain c (W2D_ERR_SINGULAR for the reference's "DetRat too
 * small in removal" stop, CTQMC.F90:4062, W2D_ERR_CAPACITY for a chain that
 * would outgrow 4096 pairs per flavour); a stopped chain's accumulators are
 * left out by the hosts and the other chains carry on.  kfmax[c][2N] = the
 * largest order each flavour of chain c has reached; *capacity = current
 * pairs-per-flavour capacity.  A chain that runs out of room is not an error:
 * it stops before the move's first random draw, the capacity of every chain is
 * doubled and it continues (the reference allocates on demand,
 * MatrixUpdate.F90:410-482).  Any pointer may be NULL.  No reference
 * counterpart (the reference stops the process, one rank per chain). */
int w2d_ctqmc_chain_status(const w2d_ctqmc *h, int32_t *status, int32_t *kfmax, int *capacity);

/* Bytes of the eigenbasis operator blocks (Psis_EB, States.F90:40-42) the
 * trace reads; shared by all chains, so the roofline charges them once per
 * launch. */
int w2d_ctqmc_op_table_bytes(const w2d_ctqmc *h, int64_t *bytes);

/* Operators in the configuration (2 x expansion order) and outer superstate of
 * every chain (get_mc_config_scalars, CTQMC.F90:9358, for all chains at once);
 * outer_sst may be NULL. */
int w2d_ctqmc_get_orders(const w2d_ctqmc *h, int32_t *nop, int32_t *outer_sst);

/* Abort word of the handle: flag != 0 makes running and later sweeps stop at
 * the next measurement boundary (every 128 moves in warm-up), keeping the data
 * measured so far, as the reference does when a trapped signal fires
 * (any_signal_fired, CTQMC.F90:8212, 8436, 8648-8658); flag = 0 clears it.
 * Async-signal-safe (one store to pinned host memory). */
int w2d_ctqmc_set_abort(w2d_ctqmc *h, int flag);

/* `aborted` of the last warmup / measure (CTQMC.F90:8648-8658): 0 = ran to the
 * end, 1 = stopped early with data, 2 = stopped before any measurement. */
int w2d_ctqmc_aborted(const w2d_ctqmc *h, int *aborted);

/* register_trap for SIGTERM, SIGINT and SIGUSR1 (CTQMC.F90:8076-8078): the
 * handler raises the abort word of every live handle. */
int w2d_ctqmc_trap_signals(void);

/* ---- launch machinery (no counterpart in the reference: one process there
 * runs one chain) ----------------------------------------------------------
 * graphs != 0: the sweep launches of a warmup / measure pass are captured into
 * one HIP graph and launched together (default on); order_chains != 0: before
 * every sweep launch the chains are ordered by expansion order, longest first,
 * so that a launch's tail is short (default off: runs of the trace engine for
 * blocks above 16 states with it on ended in GPU memory faults);
 * moves_per_launch: moves per
 * chain in one sweep launch (rounded to whole measurement intervals; 0 = the
 * default of about 4000).  None of them changes any chain's results. */
int w2d_ctqmc_set_launch(w2d_ctqmc *h, int graphs, int order_chains, int64_t moves_per_launch);

/* Trace engine = padded block size: 1 (scalar blocks), 4 (register walks), 16
 * (per-lane products) or 128 (FP64 MFMA tiles), at least nm_min, the smallest
 * the model's quantum-number blocks allow (the default).  Switching re-lays out
 * the chains (state kept); nm <= 0 selects nm_min.  Engine 128 is admissible
 * only where nm_min is 128 (blocks above 16 states): W2D_ERR_ARG otherwise. */
int w2d_ctqmc_set_engine(w2d_ctqmc *h, int nm);
int w2d_ctqmc_get_engine(const w2d_ctqmc *h, int *nm, int *nm_min);

/* Autotuner: times every admissible engine over `moves` warm-up moves per
 * chain (after moves/4 untimed ones that rebuild its caches) and keeps the
 * fastest; rates[4] (may be NULL) receives moves/s for engines 1, 4, 16, 128
 * (0 where not admissible).  The moves are warm-up moves of the chains. */
int w2d_ctqmc_autotune(w2d_ctqmc *h, int64_t moves, int *nm, double *rates);

/* ---- parameter derivatives (no counterpart in the reference) ---------------
 * For callers that differentiate observables with respect to model parameters
 * theta (score-function estimators, d<O>/dtheta = <O s> - <O><s> + <dO/dtheta>
 * with s = d ln|w| / dtheta of the configuration weight
 * w = det(Delta) Tr[...], CTQMC.F90's weight of MatrixUpdate.F90:41-99 and
 * Trace.F90:5734-6102).
 *
 * Derivatives of the create inputs for ntheta (1..4) parameters, each array
 * laid out like its create counterpart and repeated ntheta times:
 *   du     [ntheta][(2N)^4]           d u_matrix / d theta (H must stay linear)
 *   dmuimp [ntheta][N][2][N][2]       d muimp_full / d theta
 *   dftau  [ntheta][N][2][N][2][Nftau] d ftau_full / d theta (diagonal read)
 * dH/dtheta must conserve the quantum numbers (else W2D_ERR_ARG).  Available
 * with the trace engines for blocks up to 16 states (W2D_ERR_ARG with the MFMA
 * engine, which set_engine / autotune then leave out).  score != 0
 * adds the score sums below to every later measurement; chains keep their
 * state (their storage is re-laid out). */
int w2d_ctqmc_set_derivatives(w2d_ctqmc *h, int ntheta, const double *du, const double *dmuimp,
                              const double *dftau, int score);

/* d ln|det Delta| / d theta_t -> out[2 t] and d ln|Tr| / d theta_t -> out[2 t + 1]
 * of chain `chain`'s current configuration (the trace part without the Egs
 * shift of the weights that get_config reports). */
int w2d_ctqmc_dlogw(w2d_ctqmc *h, int chain, double *out);

/* Score sums of the last measure over contiguous groups of chains (like
 * w2d_ctqmc_get_group_sums), with s_t of each measured configuration:
 *   ssum   [ngroups][ntheta]               sum of sign x s_t
 *   gtau   [ngroups][ntheta][2N][ntau]     sum of s_t x the G(tau) bins of a measurement
 *   dm     [ngroups][ntheta][dmsize]       sum of s_t x the density-matrix sample (blocks)
 *   dmx    [ngroups][ntheta][dmsize]       sum of d(sample)/d theta_t at fixed configuration
 * (NULL skips an array).  W2D_ERR_STATE without score sums. */
int w2d_ctqmc_get_score_sums(w2d_ctqmc *h, int ngroups, double *ssum, double *gtau, double *dm,
                             double *dmx);

/* Replaces dest_ctqmc (CTQMC.F90:901). */
int w2d_ctqmc_destroy(w2d_ctqmc *h);

/* Library/ABI version string. */
const char *w2d_ctqmc_version(void);

#ifdef __cplusplus
}
#endif
#endif
