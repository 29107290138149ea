// Device-side data layout of the batched CT-HYB engine.
#pragma once

#include <cstdint>
#include <hip/hip_runtime.h>

namespace w2d {

constexpr int WAVE = 64;       // one wavefront drives one Markov chain
constexpr int MAXF = 14;       // flavours (2*nbands, nbands <= 7)
constexpr int MAXSST = 256;    // superstates (quantum-number blocks) a model may have

// read-only model and parameters, passed by value to the kernels
struct DevModel {
    int nb, nf, nsst, nmax, nstates, nftau;
    double beta, egs;
    const int *sizes, *offsets, *connect, *first_state, *state_sst;
    const long *opoff;
    const int *opoff32;                 // the same offsets as int (HBM operator-block table for blocks above 16 states)
    // blocks above 16 states: every operator block zero-padded to multiples of
    // 16 rows and columns (column-major, leading dimension 16 ceil(n_t / 16)),
    // offsets indexed like opoff; the MFMA walks read whole 16 x 4 fragments
    // without bounds checks
    const double *ops16;
    const int *opoff16;
    const double *ops, *evals, *ftau;   // ftau[f][nftau]
    const long *dmoff;                  // block-diagonal density-matrix offsets
    const double *legc;                 // Bonnet coefficients (2l+1)/(l+1), l/(l+1) per l
    const int4 *trans;                  // per (superstate, operator): target, its size, block offset, its energies offset
    long dmsize;
    long nops;                          // size of ops
    double p_shift, p_glob, p_outer, p_4op;
    long long nslide;
    int ntau, nleg, maxhisto, has_szt, meas_dm;
    int accum, nlev;                    // AccumGtau (1 mean, 2 binning) and its levels
    int nmstep;                         // length of the step-resolved arrays (Nmeas), 0 = off
    int step_mean, step_mid, step_sign; // Gtau_mean_step, Gtau_mid_step, sign_step
    // parameter derivatives: dH/dtheta blocks in the eigenbasis [ntheta][dmsize]
    // (at dmoff, column-major) and dF/dtheta [ntheta][nf][nftau]; score != 0
    // adds the score-function sums to every measurement
    int ntheta, score;
    const double *dblk, *dftau;
    // user symmetry moves (NSymMove, SymMove01..): [2 nsym][nf] 1-based
    // flavour permutations, each one's inverse in row nsym + i
    int nsym;
    const int *symmoves;
};

// per-chain scalar state (kept contiguously for all chains, after the chain blocks)
struct ChainHdr {
    int nop, outer, wpos, nwin;
    int err, cache_pos, kst_sst, mti;
    int wia, wib, wvalid, pc_valid; // cached index range of the window's operators; pair counts kept
    int rk_n, lb_n;                // valid entries of the window's prefix kets / suffix bras
    double taudiff_max, tauwin_min, tauwin_max, kst_log;
    double trace_log, trace_sign, det_log, det_sign, cfgsign;
    double meas_sign;
    long long cnt_meas;
    long long try_add, try_rem, try_glob, try_shift;
    long long acc_add, acc_rem, acc_glob, acc_shift;
    long long qn_rej, moves;
    double alg_flops, alg_bytes;   // algorithmic work of this chain's moves and measurements
    long long pos;                 // moves done in the current warmup / measure call
    int halt;                      // HALT_* : why the chain stopped before its target
    int naccpair;                  // bins of the accepted-pair histogram (NAccPair)
    double accpair_max;            // its range (AccPairMax)
    int apt_index, apt_n;          // next entry / size of the accepted-pair distance list (AccPairTau)
    int kf[MAXF];
    int kfmax[MAXF];               // largest order each flavour has reached
};

// reasons a chain stops before the launch's target position (chain state intact)
enum { HALT_NONE = 0, HALT_CAP = 1, HALT_ABORT = 2, HALT_DONE = 3 };

// byte offsets of the per-chain arrays inside one chain block
struct ChainLayout {
    long stride;          // bytes per chain
    long hdrs;            // offset of the header array (ChainHdr[nchains]) after the chain blocks
    long mt;              // uint32[624]
    long tau, code;       // double[cap], int[cap]: configuration sorted by tau
    long fc, fa;          // double[nf][kcap]: creator / annihilator taus per flavour
    long fcs, fas;        // int[nf][kcap]: physical row / column slot of each rank
    long minv;            // double[nf][kcap*kcap]: inverse hybridisation, column-major
    long scratch;         // double[kmat + 4*cap]: matrix scratch, then operator scratch
    long kst;             // double[nmax*nmax]: ket at the window start
    long bnd;             // double[nbound][nmax*nmax]: bras at window boundaries
    long bndlog, bndsst;  // double[nbound], int[nbound]
    long part;            // double[2*nmax]: part traces (log, sign)
    long kwork;           // double[2][nmax*nmax]: work matrices (large blocks)
    long ltau, lcode;     // trial operator list (double[lcap], int[lcap])
    long gtau, gleg, dm, histo, outer; // accumulators
    long gbuf, gacc;      // AccumGtau = 2: one measurement's G(tau); binning levels [3][nlev][nf*ntau]
    // MeasOuterHistograms (CTQMC.F90:3285-3339): visits of each outer state,
    // signs of outer superstates / states, trace contributions of both
    long ohst, shss, shst, tcss, tcst;
    long accpair;         // accepted pairs per flavour and distance bin [nf][naccpair_cap]
    long gmstep, gdstep, sstep;   // step-resolved G(tau) mean / mid, sign [nmstep][nf], [nmstep]
    long aptau;           // accepted-pair distances [apt_cap] (calibration)
    // score-function sums [ntheta]: sign x s, s x G-sample [nf][ntau], s x
    // rho-sample and d(rho-sample)/dtheta [dmsize]; s = d ln|w| / d theta
    long scsum, scg, scdm, scdmx;
    long dwork;           // derivative work: (4 + 3 ntheta) matrices nmax x nmax and nmax factors
    int naccpair_cap, apt_cap;
    int cap, kcap, nbound, lcap;
    long kmat;            // max(kcap*kcap, nmax*nmax)
    // dynamic LDS carve-up (byte offsets; l_ops / l_k = -1: kept in global memory)
    long l_hdr, l_mt, l_sz, l_off, l_cn, l_oo, l_ev, l_ops, l_wtau, l_wcode, l_ttau, l_tcode;
    long l_dvec, l_q, l_r, l_col, l_row, l_pl, l_k, l_legc, l_tr, l_pc, lds_bytes;
    int lcap_lds;         // operators each LDS list holds (window list, trial list)
    // prefix kets / suffix bras of the window's operators (blocks up to 4 states)
    int sc;               // window operators they cover (0: not kept)
    long l_rk, l_lb, l_rkx, l_lbx, l_rks, l_lbs, l_qa, l_qlc, l_qrec, l_qsst;
    long l_strip;         // 16-column strip of the state for the MFMA walks (blocks above 16 states)
};

}  // namespace w2d
