// Device side of the batched CT-HYB (matrix/"M-V" Z-space) Monte Carlo engine
// for MI355X (gfx950).
//
// One 64-lane wavefront owns one Markov chain and runs NCorr-long sweeps of
// Metropolis steps inside one persistent launch.  Control logic (random draws,
// proposals, window bookkeeping) runs uniformly on all lanes against a chain
// header held in LDS; the heavy parts are lane-parallel:
//   * local trace (Trace.F90:5354-6102): block matrix products
//       K <- Psi_EB(b,s,ca) * diag(exp(-(E-E0) dtau)) * K
//     over the operators of the sliding window, starting from a cached ket at the
//     window's lower edge and contracted with a cached bra at its upper edge; K
//     holds one column per state of the outer superstate, lanes own (row, col)
//     entries, one log scale per matrix, one barrier per operator;
//   * determinant ratio S - R^T M Q and the rank-1 grow/shrink of M = Delta^-1
//     (MatrixUpdate.F90:392-482): lanes own rows of the k x k column-major block;
//   * G(tau)/Legendre accumulation (CTQMC.F90:2252-2440): lanes own (a, c) pairs.
// The random-number stream, proposal and acceptance formulas and the window
// schedule follow the reference exactly, so chain c seeded with seed0 + c makes
// the reference's accept/reject decisions for init_rnd(seed0 + c).
#pragma once

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <cstdint>

#include "../../include/w2dyn_ctqmc.h"
#include "engine.hpp"

namespace w2d {

#define DEV __device__ __forceinline__

// ---------------------------------------------------------------------------
// wave helpers (one wavefront per block)
// ---------------------------------------------------------------------------
DEV int lane_id() { return (int)threadIdx.x; }

// Cross-lane reductions with DPP row operations (VALU, no LDS traffic): quad
// permutes and row mirrors reduce each 16-lane row, readlane combines the four
// rows; the result is uniform.
template <int CTRL>
DEV int dpp_i(int v)
{
    return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xf, 0xf, false);
}

template <int CTRL>
DEV double dpp_d(double v)
{
    const int lo = __double2loint(v), hi = __double2hiint(v);
    return __hiloint2double(dpp_i<CTRL>(hi), dpp_i<CTRL>(lo));
}

DEV double readlane_d(double v, int l)
{
    return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), l),
                            __builtin_amdgcn_readlane(__double2loint(v), l));
}

constexpr int DPP_QP_1032 = 0xB1, DPP_QP_2301 = 0x4E, DPP_ROW_HALF_MIRROR = 0x141, DPP_ROW_MIRROR = 0x140;

DEV double wave_max(double v)
{
    v = fmax(v, dpp_d<DPP_QP_1032>(v));
    v = fmax(v, dpp_d<DPP_QP_2301>(v));
    v = fmax(v, dpp_d<DPP_ROW_HALF_MIRROR>(v));
    v = fmax(v, dpp_d<DPP_ROW_MIRROR>(v));
    return fmax(fmax(readlane_d(v, 0), readlane_d(v, 16)), fmax(readlane_d(v, 32), readlane_d(v, 48)));
}

DEV double wave_sum(double v)
{
    v += dpp_d<DPP_QP_1032>(v);
    v += dpp_d<DPP_QP_2301>(v);
    v += dpp_d<DPP_ROW_HALF_MIRROR>(v);
    v += dpp_d<DPP_ROW_MIRROR>(v);
    return (readlane_d(v, 0) + readlane_d(v, 16)) + (readlane_d(v, 32) + readlane_d(v, 48));
}

DEV int wave_isum(int v)
{
    v += dpp_i<DPP_QP_1032>(v);
    v += dpp_i<DPP_QP_2301>(v);
    v += dpp_i<DPP_ROW_HALF_MIRROR>(v);
    v += dpp_i<DPP_ROW_MIRROR>(v);
    return (__builtin_amdgcn_readlane(v, 0) + __builtin_amdgcn_readlane(v, 16)) +
           (__builtin_amdgcn_readlane(v, 32) + __builtin_amdgcn_readlane(v, 48));
}

DEV int wave_imax(int v)
{
    v = max(v, dpp_i<DPP_QP_1032>(v));
    v = max(v, dpp_i<DPP_QP_2301>(v));
    v = max(v, dpp_i<DPP_ROW_HALF_MIRROR>(v));
    v = max(v, dpp_i<DPP_ROW_MIRROR>(v));
    return max(max(__builtin_amdgcn_readlane(v, 0), __builtin_amdgcn_readlane(v, 16)),
               max(__builtin_amdgcn_readlane(v, 32), __builtin_amdgcn_readlane(v, 48)));
}

DEV int wave_imin(int v) { return -wave_imax(-v); }

DEV int wave_excl_scan(int v, int *total)
{
    int x = v;
    for (int o = 1; o < WAVE; o <<= 1) {
        int y = __shfl_up(x, o, WAVE);
        if (lane_id() >= o) x += y;
    }
    *total = __shfl(x, WAVE - 1, WAVE);
    return x - v;
}

// hardware FP64 add to device (global) memory, result unused
DEV void global_add(double *p, double v)
{
    typedef __attribute__((address_space(1))) double gdouble;
    (void)__builtin_amdgcn_global_atomic_fadd_f64((gdouble *)p, v);
}

DEV int popc64(unsigned long long m) { return __popcll(m); }

DEV unsigned long long lanemask_le() { return (lane_id() == 63) ? ~0ull : ((2ull << lane_id()) - 1ull); }

DEV void wsync() { __syncthreads(); }

// ---------------------------------------------------------------------------
// random numbers: mt19937 + libstdc++ generate_canonical<double,53> and
// uniform_int_distribution (src/mtrng/random.cc:30-79).  The 624-word state is
// shared in LDS; every lane draws the same numbers, the twist is lane-parallel.
// ---------------------------------------------------------------------------
struct Rng {
    uint32_t *mt;   // LDS
    int *idx;       // LDS (uniform)
};

DEV uint32_t mt_twist_word(uint32_t a, uint32_t b, uint32_t c)
{
    uint32_t y = (a & 0x80000000u) | (b & 0x7fffffffu);
    uint32_t v = c ^ (y >> 1);
    if (y & 1u) v ^= 0x9908b0dfu;
    return v;
}

// new mt[i] depends on mt[i+1] (old) and mt[i+397]: old for i < 227, new for
// i >= 227 (from index i-227).  Phases [0,227) [227,454) [454,623) {623}.
DEV void mt_regen(uint32_t *mt)
{
    const int l = lane_id();
    uint32_t nv[4];
    int c = 0;
    for (int i = l; i < 227; i += WAVE) nv[c++] = mt_twist_word(mt[i], mt[i + 1], mt[i + 397]);
    wsync();
    c = 0;
    for (int i = l; i < 227; i += WAVE) mt[i] = nv[c++];
    wsync();
    c = 0;
    for (int i = 227 + l; i < 454; i += WAVE) nv[c++] = mt_twist_word(mt[i], mt[i + 1], mt[i - 227]);
    wsync();
    c = 0;
    for (int i = 227 + l; i < 454; i += WAVE) mt[i] = nv[c++];
    wsync();
    c = 0;
    for (int i = 454 + l; i < 623; i += WAVE) nv[c++] = mt_twist_word(mt[i], mt[i + 1], mt[i - 227]);
    wsync();
    c = 0;
    for (int i = 454 + l; i < 623; i += WAVE) mt[i] = nv[c++];
    wsync();
    uint32_t last = mt_twist_word(mt[623], mt[0], mt[396]);
    wsync();
    if (l == 0) mt[623] = last;
    wsync();
}

DEV uint32_t mt_temper(uint32_t y)
{
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
}

DEV uint32_t mt_next(Rng &r)
{
    int i = *r.idx;
    if (i >= 624) {
        mt_regen(r.mt);
        i = 0;
    }
    const uint32_t y = r.mt[i];
    *r.idx = i + 1;   // all lanes store the same value
    return mt_temper(y);
}

// generate_canonical<double,53>: (g1 + g2 2^32) / 2^64, clamped below 1.  The
// two words are read together when both are left in the current block.
DEV double grnd(Rng &r)
{
    const int i = *r.idx;
    uint32_t g1, g2;
    if (i + 2 <= 624) {
        g1 = mt_temper(r.mt[i]);
        g2 = mt_temper(r.mt[i + 1]);
        *r.idx = i + 2;
    } else {
        g1 = mt_next(r);
        g2 = mt_next(r);
    }
    double sum = (double)g1;
    const double tmp = 4294967296.0;
    sum = __dadd_rn(sum, __dmul_rn((double)g2, tmp));
    double ret = sum / (tmp * tmp);
    if (ret >= 1.0) ret = 0x1.fffffffffffffp-1;
    return ret;
}

DEV int randint(Rng &r, int a, int b)
{
    uint32_t range = (uint32_t)b - (uint32_t)a + 1u;
    uint64_t prod = (uint64_t)mt_next(r) * (uint64_t)range;
    uint32_t low = (uint32_t)prod;
    if (low < range) {
        uint32_t thr = (uint32_t)(0u - range) % range;
        while (low < thr) {
            prod = (uint64_t)mt_next(r) * (uint64_t)range;
            low = (uint32_t)prod;
        }
    }
    return a + (int)(prod >> 32);
}

// ---------------------------------------------------------------------------
// operator codes: orb | spin << 4 | (ca-1) << 5 ; flavour f = orb + nb*spin
// ---------------------------------------------------------------------------
DEV int op_orb(int c) { return c & 15; }
DEV int op_spin(int c) { return (c >> 4) & 1; }
DEV int op_ca(int c) { return ((c >> 5) & 1) + 1; }
DEV int mk_code(int orb, int spin, int ca) { return orb | (spin << 4) | ((ca - 1) << 5); }

// ---------------------------------------------------------------------------
// chain context: LDS views of the header, model tables and work buffers
// ---------------------------------------------------------------------------
struct Ctx {
    const DevModel &M;
    const ChainLayout &L;
    char *base;          // chain block in global memory
    ChainHdr *ghdr;      // the chain's header in the contiguous header array (HBM)
    ChainHdr &h;         // LDS copy
    Rng rng;
    // model tables (LDS)
    const int *sz, *off, *cn, *oo;
    const int4 *tr;      // per (superstate, operator): target, target size, block offset, target energies offset
    const double *ev, *opsp;   // opsp: operator blocks (LDS or global)
    // chain arrays (global)
    double *tau;
    int *code;
    double *fc, *fa, *minv, *scratch, *kst, *bnd, *bndlog, *part;
    int *fcs, *fas;      // physical row (creator) / column (annihilator) slot of each rank
    int *bndsst;
    // work buffers
    double *wtau;        // window list (LDS): operators [wia, wib) while h.wvalid
    int *wcode;
    int *pc;             // removable-pair count of each window operator (LDS, while h.pc_valid)
    double *ttau;        // trial list / staging buffer (LDS), same capacity
    int *tcode;
    double *K0, *K1;     // trace work matrices (LDS for small blocks)
    double *strip;       // LDS strip of the state for the MFMA walks (blocks above 16 states)
    double *dvec, *Q, *R, *colb, *rowb, *pl;
    // window prefix kets rk[j] (after the first j window operators) and suffix
    // bras lb[m] (after the last m), with their logs and superstates (LDS)
    double *rk, *lb, *rkx, *lbx;
    int *rks, *lbs;
    // quad-walk chunk buffers (LDS): entries, path records, targets, log factors
    double *qa, *qlc;
    int4 *qrec;
    int *qsst;
    const double *legc;  // Legendre recursion coefficients (LDS)
    int nm;              // padded block size used for matrix strides
    // algorithmic work of this launch, added to the header by save_ctx
    double wflops = 0.0, wbytes = 0.0;
    DEV Ctx(const DevModel &m, const ChainLayout &l, char *b, ChainHdr &hd) : M(m), L(l), base(b), h(hd) {}
    DEV double *mi(int f) { return minv + (long)f * L.kcap * L.kcap; }
    DEV int fl(int c) const { return op_orb(c) + M.nb * op_spin(c); }
    DEV int conn(int s, int c, bool adjoint) const
    {
        int ca = op_ca(c);
        if (adjoint) ca = 3 - ca;
        return cn[((s * M.nb + op_orb(c)) * 2 + op_spin(c)) * 2 + (ca - 1)];
    }
    DEV int4 trans(int s, int c, bool adjoint) const
    {
        int ca = op_ca(c);
        if (adjoint) ca = 3 - ca;
        return tr[((s * M.nb + op_orb(c)) * 2 + op_spin(c)) * 2 + (ca - 1)];
    }
    DEV const double *opblock(int s, int c, bool adjoint) const
    {
        int ca = op_ca(c);
        if (adjoint) ca = 3 - ca;
        return opsp + oo[((s * M.nb + op_orb(c)) * 2 + op_spin(c)) * 2 + (ca - 1)];
    }
};

// get_HybrF (Trace.F90:1014-1033)
DEV double hybr(const DevModel &M, double tau, int f)
{
    double sig = 1.0;
    if (tau < 0.0) {
        tau = tau + M.beta;
        sig = -1.0;
    }
    double n = tau / M.beta * (double)(M.nftau - 1);
    int i = (int)n;
    const double *F = M.ftau + (long)f * M.nftau;
    return sig * (F[i] + (n - (double)i) * (F[i + 1] - F[i]));
}

// ---------------------------------------------------------------------------
// window (Trace.F90:1608-1751)
// ---------------------------------------------------------------------------
DEV void set_window(Ctx &c)
{
    if (c.h.nwin > 1) {
        c.h.tauwin_min = ((double)c.h.wpos / (double)(c.h.nwin + 1)) * c.M.beta;
        c.h.tauwin_max = ((double)(c.h.wpos + 2) / (double)(c.h.nwin + 1)) * c.M.beta;
    }
    c.h.wvalid = 0;
}

DEV void shift_window(Ctx &c)
{
    if (c.h.nwin > 1) {
        c.h.wpos++;
        if (c.h.wpos == c.h.nwin) c.h.wpos = 0;
        set_window(c);
    } else {
        c.h.wpos = 0;
    }
}

DEV double bound_tau(int m, int nwin, double beta)
{
    if (nwin <= 1) return m == 0 ? 0.0 : beta;
    return ((double)m / (double)(nwin + 1)) * beta;
}

DEV int lower_bound_tau(const double *tau, int n, double x)
{
    int lo = 0, hi = n;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (tau[mid] < x) lo = mid + 1; else hi = mid;
    }
    return lo;
}

DEV int upper_bound_tau(const double *tau, int n, double x)
{
    int lo = 0, hi = n;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (tau[mid] <= x) lo = mid + 1; else hi = mid;
    }
    return lo;
}

// indices [wia, wib) of the operators inside [tauwin_min, tauwin_max] and
// their copy in the LDS window list; kept across moves until the window
// shifts or a global / tau-shift move rewrites the configuration.  Returns
// false if the window does not fit the LDS list.
DEV bool ensure_window(Ctx &c)
{
    if (c.h.wvalid) return true;
    const int ia = lower_bound_tau(c.tau, c.h.nop, c.h.tauwin_min);
    const int ib = upper_bound_tau(c.tau, c.h.nop, c.h.tauwin_max);
    if (ib - ia + 2 > c.L.lcap_lds) return false;
    for (int j = lane_id(); j < ib - ia; j += WAVE) {
        c.wtau[j] = c.tau[ia + j];
        c.wcode[j] = c.code[ia + j];
    }
    c.h.wia = ia;
    c.h.wib = ib;
    c.h.wvalid = 1;
    c.h.pc_valid = 0;
    c.h.rk_n = 0;
    c.h.lb_n = 0;
    wsync();
    return true;
}

// the trial list becomes the window list (after an accepted add / remove)
DEV void swap_lists(Ctx &c)
{
    double *t = c.wtau;
    c.wtau = c.ttau;
    c.ttau = t;
    int *k = c.wcode;
    c.wcode = c.tcode;
    c.tcode = k;
}

// ---------------------------------------------------------------------------
// local trace
// ---------------------------------------------------------------------------
// algorithmic work of one operator application on `no` outer states: a
// (nt x ns) block times an (ns x no) state matrix plus the diagonal propagator;
// bytes = states in and out (the operator blocks are shared by all chains and
// charged once per launch by the host).  Accumulated in registers during a
// walk and added to the chain header once.
struct Work {
    long long ntns = 0, nss = 0, nts = 0;
    int nout = 0;
    DEV void op(int nt, int ns, int no)
    {
        ntns += nt * ns;
        nss += ns;
        nts += nt;
        nout = no;
    }
    DEV void commit(Ctx &c) const
    {
        c.wflops += (double)(nout * (2 * ntns + nss));
        c.wbytes += (double)(8LL * nout * (nss + nts));
    }
};

// exp(-(E_k - E0) dt) of superstate s into LDS dvec; returns the log factor
// -(E0 - Egs) dt.  Ends with a barrier.  (Blocks above 16 states.)
DEV double prop_factors(Ctx &c, int s, double dt)
{
    const int ns = c.sz[s];
    const double *E = c.ev + c.off[s];
    const double e0 = E[0];
    for (int k = lane_id(); k < ns; k += WAVE) c.dvec[k] = (ns > 1) ? exp(-(E[k] - e0) * dt) : 1.0;
    wsync();
    return -(e0 - c.M.egs) * dt;
}

// Lane k's propagator factor exp(-(E_k - E0) dt) of superstate s (1 beyond the
// block); the common log factor -(E0 - Egs) dt is added to lg.  No LDS traffic:
// the factors are broadcast by readlane inside apply_small.
DEV double prop_lane(const Ctx &c, int s, double dt, double &lg)
{
    const int ns = c.sz[s];
    const double *E = c.ev + c.off[s];
    const double e0 = E[0];
    lg += -(e0 - c.M.egs) * dt;
    const int k = lane_id();
    const double ek = k < ns ? E[k] : e0;
    return exp(-(ek - e0) * dt);   // exactly 1 beyond the block and for 1-state blocks
}

// prop_lane for a block given by its energies E and size ns
DEV double prop_e(const Ctx &c, const double *E, int ns, double dt, double &lg)
{
    const double e0 = E[0];
    lg += -(e0 - c.M.egs) * dt;
    const int k = lane_id();
    const double ek = k < ns ? E[k] : e0;
    return exp(-(ek - e0) * dt);
}

// Thresholds outside which a state matrix is renormalised during a walk; in
// between, the entries are left unscaled (rescaling by powers of two is exact,
// so skipping it changes no rounding).
constexpr double STATE_TINY = 0x1p-64, STATE_HUGE = 0x1p64;

// Output element u of this lane in a (rows x no) row-major state: row rr[u],
// column ii[u]; computed once per walk.
template <int NM>
struct LaneSlots {
    static constexpr int PER = (NM * NM + WAVE - 1) / WAVE;
    int rr[PER], ii[PER];
    DEV explicit LaneSlots(int no)
    {
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            const int idx = lane_id() + u * WAVE;
            rr[u] = idx / no;
            ii[u] = idx - rr[u] * no;
        }
    }
};

// One operator on blocks of at most 16 states: Kout (nt x no) = P * diag(d) *
// Kin (ns x no), d broadcast from the lanes of prop_lane.  One barrier.
template <int NM>
DEV void apply_small(const double *P, const double *Kin, double *Kout, int nt, int ns, int no, double d,
                     const LaneSlots<NM> &ls, double &lg)
{
    constexpr int PER = LaneSlots<NM>::PER;
    // block and state entries first (independent of the propagator factors),
    // for blocks up to 4 states held in registers
    constexpr bool PRE = NM <= 4;
    double pv[PRE ? PER : 1][PRE ? NM : 1], kv[PRE ? PER : 1][PRE ? NM : 1];
    if constexpr (PRE) {
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            const int r = ls.rr[u] < nt ? ls.rr[u] : 0, i = ls.ii[u];
#pragma unroll
            for (int k = 0; k < NM; ++k) {
                const int kk = k < ns ? k : 0;
                pv[u][k] = P[r + kk * nt];
                kv[u][k] = Kin[kk * no + i];
            }
        }
    }
    double dk[NM];
#pragma unroll
    for (int k = 0; k < NM; ++k) dk[k] = readlane_d(d, k);
    // columns k >= ns contribute zero: their factor is zeroed and their
    // addresses clamped, so the unrolled products run without branches
#pragma unroll
    for (int k = 0; k < NM; ++k) dk[k] = k < ns ? dk[k] : 0.0;
    double v[PER];
    double m = 0.0;
#pragma unroll
    for (int u = 0; u < PER; ++u) {
        double acc = 0.0;
        const int r = ls.rr[u] < nt ? ls.rr[u] : 0, i = ls.ii[u];
#pragma unroll
        for (int k = 0; k < NM; ++k) {
            const int kk = k < ns ? k : 0;
            if constexpr (PRE)
                acc = fma(pv[u][k] * dk[k], kv[u][k], acc);
            else
                acc = fma(P[r + kk * nt] * dk[k], Kin[kk * no + i], acc);
        }
        acc = ls.rr[u] < nt ? acc : 0.0;
        v[u] = acc;
        m = fmax(m, fabs(acc));
    }
    if (__ballot(m >= STATE_TINY) == 0ull || __ballot(m > STATE_HUGE) != 0ull) {
        const double mm = wave_max(m);
        if (mm > 0.0) {
            const int e = ilogb(mm);
            lg += (double)e * 0.69314718055994530942;
#pragma unroll
            for (int u = 0; u < PER; ++u) v[u] = ldexp(v[u], -e);
        }
    }
#pragma unroll
    for (int u = 0; u < PER; ++u)
        if (ls.rr[u] < nt) Kout[lane_id() + u * WAVE] = v[u];
    wsync();
}

// Final propagation of a walk (rows scaled by lane r's factor d) and
// normalisation to a maximum modulus of one.
template <int NM>
DEV void finish_small(double *K, int ns, int no, double d, const LaneSlots<NM> &ls, double &lg)
{
    constexpr int PER = LaneSlots<NM>::PER;
    double v[PER];
    double m = 0.0;
#pragma unroll
    for (int u = 0; u < PER; ++u) {
        const int r = ls.rr[u];
        const double dr = __shfl(d, r < WAVE ? r : 0, WAVE);
        v[u] = r < ns ? K[lane_id() + u * WAVE] * dr : 0.0;
        m = fmax(m, fabs(v[u]));
    }
    m = wave_max(m);
    if (m > 0.0) {
        const double inv = 1.0 / m;
        lg += log(m);
#pragma unroll
        for (int u = 0; u < PER; ++u) v[u] *= inv;
    }
#pragma unroll
    for (int u = 0; u < PER; ++u)
        if (ls.rr[u] < ns) K[lane_id() + u * WAVE] = v[u];
    wsync();
}

// ---------------------------------------------------------------------------
// Walks on blocks of at most 4 states, in registers.  The (ns x no) state is
// zero-padded to 4 x 4 and lane l holds its entry (r, i) = ((l / 4) % 4, l % 4);
// the four 16-lane rows of the wave hold the same matrix.  An operator scales
// row r by the propagator factor of state r (each lane its own), then forms
// K'(r, i) = sum_k P(r, k) Kd(k, i): rotations of a 16-lane row by 4, 8 and 12
// lanes (DPP) bring Kd(k, i) of the other three rows to lane (r, i).  Operator
// entries and energies are each lane's own LDS reads, issued one operator
// ahead; no state leaves registers and no barrier is taken per operator.
// ---------------------------------------------------------------------------
constexpr int DPP_ROW_ROR4 = 0x124, DPP_ROW_ROR8 = 0x128, DPP_ROW_ROR12 = 0x12C;

struct Quad {
    int r, i;    // entry held by this lane
    int k[4];    // row of the entry that the rotation by 4m lanes brings here (k[0] = r)
    DEV Quad()
    {
        const int l = lane_id() & 15;
        r = l >> 2;
        i = l & 3;
        k[0] = r;
        k[1] = dpp_i<DPP_ROW_ROR4>(r);
        k[2] = dpp_i<DPP_ROW_ROR8>(r);
        k[3] = dpp_i<DPP_ROW_ROR12>(r);
    }
    // entry of a row-major (ns x no) state, 0 outside
    DEV double load(const double *K, int ns, int no) const
    {
        const bool in = r < ns && i < no;
        const double x = K[in ? r * no + i : 0];
        return in ? x : 0.0;
    }
    DEV void store(double *K, int ns, int no, double v) const
    {
        if (lane_id() < 16 && r < ns && i < no) K[r * no + i] = v;
    }
    // P(r, k[m]) of an (nt x ns) column-major block; zeros for a record
    // without target
    DEV void entries(const double *ops, int4 tr, int ns, double (&p)[4]) const
    {
        const double *P = ops + (tr.x >= 0 ? tr.z : 0);
        const int nt = tr.x >= 0 ? tr.y : 0;
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            const bool in = r < nt && k[m] < ns;
            const double x = P[in ? r + k[m] * nt : 0];
            p[m] = in ? x : 0.0;
        }
    }
    // exp(-(E_r - E_0) dt) of this lane's row in a block of energies E (1
    // beyond the block); the common log factor -(E_0 - Egs) dt is returned in lc
    DEV double prop(const double *E, int ns, double dt, double egs, double &lc) const
    {
        const double e0 = E[0];
        const double er = E[r < ns ? r : 0];
        lc = -(e0 - egs) * dt;
        return exp(-(er - e0) * dt);
    }
};

// one operator on a quad state with this lane's entries p[m] = A(r, k[m]) of
// A = P diag(d); power-of-two renormalisation when the largest entry leaves
// [2^-64, 2^64]
DEV double quad_apply(double v, const double (&p)[4], double &lg)
{
    const double x1 = dpp_d<DPP_ROW_ROR4>(v), x2 = dpp_d<DPP_ROW_ROR8>(v), x3 = dpp_d<DPP_ROW_ROR12>(v);
    double w = fma(p[1], x1, p[0] * v) + fma(p[3], x3, p[2] * x2);
    const double m = fabs(w);
    if (__ballot(m >= STATE_TINY) == 0ull || __ballot(m > STATE_HUGE) != 0ull) {
        const double mm = wave_max(m);
        if (mm > 0.0) {
            const int e = ilogb(mm);
            lg += (double)e * 0.69314718055994530942;
            w = ldexp(w, -e);
        }
    }
    return w;
}

// final propagation (lane factor d); renormalised by a power of two only when
// the largest entry leaves [2^-64, 2^64]
DEV double quad_finish(double v, double d, double &lg)
{
    v *= d;
    const double m = fabs(v);
    if (__ballot(m >= STATE_TINY) == 0ull || __ballot(m > STATE_HUGE) != 0ull) {
        const double mm = wave_max(m);
        if (mm > 0.0) {
            const int e = ilogb(mm);
            lg += (double)e * 0.69314718055994530942;
            v = ldexp(v, -e);
        }
    }
    return v;
}

// Trace of a quad ket v (superstate s, log scale lg) against the row-major
// bra B at the same time: column sums by row rotations, part traces of the
// outer states to c.pl, the total in (tl, ts) (the plain sum of the columns:
// every column carries the same scale exp(lg)).
DEV void quad_contract(Ctx &c, const Quad &q, double v, const double *B, int s, int no, double lg, double &tl,
                       double &ts)
{
    const double x = v * q.load(B, c.sz[s], no);
    const double col = (x + dpp_d<DPP_ROW_ROR8>(x)) + (dpp_d<DPP_ROW_ROR4>(x) + dpp_d<DPP_ROW_ROR12>(x));
    if (lane_id() < no) {
        c.pl[2 * lane_id()] = lg + log(fabs(col));
        c.pl[2 * lane_id() + 1] = copysign(1.0, col);
    }
    const double S = (readlane_d(col, 0) + readlane_d(col, 1)) + (readlane_d(col, 2) + readlane_d(col, 3));
    if (S == 0.0) {
        tl = -DBL_MAX;
        ts = 0.0;
    } else {
        tl = lg + log(fabs(S));
        ts = copysign(1.0, S);
    }
    wsync();
}

constexpr int QCH = 16;   // operators per chunk of a quad walk

// Walk over the n operators of a list: ascending from entry 0 (ket, times
// tstart < tau_0 < ... < tend) or, with BRA, descending from entry n-1 with the
// adjoint blocks (tstart > tau_{n-1} > ... > tend).  In chunks of QCH
// operators: (1) the superstate path, one transition record per operator
// (uniform; stops at an operator without target block); (2) the entries
// A(r, k) = P(r, k) exp(-(E_k - E_0) dt) of each operator's block with its
// propagator, four operators per pass over the wave, one exp per lane, into
// LDS; (3) the products in registers.  after(k, v, s, ns, lg) sees the state
// after the k-th operator walked, at that operator's time.  Returns false at a
// quantum-number violation; otherwise dlast holds the lane factor of the
// propagation to tend (its log factor already in lg).
template <bool BRA, typename F>
DEV bool quad_walk(Ctx &c, const Quad &q, double &v, int &s, double &lg, int no, double tstart, const double *otau,
                   const int *ocode, int n, double tend, double &dlast, Work &w, F &&after)
{
    const double egs = c.M.egs;
    const int l = lane_id();
    // pass (2): lane l computes entry (r, m) = (q.r, q.i), i.e. A(r, k[m])
    const int km = q.i == 0 ? q.k[0] : q.i == 1 ? q.k[1] : q.i == 2 ? q.k[2] : q.k[3];
    int ns = c.sz[s], eo = c.off[s];
    const int nb4 = 4 * c.M.nb;
    double tlast = tstart;
    for (int c0 = 0; c0 < n; c0 += QCH) {
        const int nc = min(QCH, n - c0);
        // (1) path: record {block offset, target size, source size, source
        // energies offset}; lane kk decodes the column of operator kk first
        int colv = 0;
        if (l < nc) {
            const int cd = ocode[BRA ? n - 1 - (c0 + l) : c0 + l];
            const int ca = BRA ? 2 - op_ca(cd) : op_ca(cd) - 1;
            colv = (op_orb(cd) * 2 + op_spin(cd)) * 2 + ca;
        }
        for (int kk = 0; kk < nc; ++kk) {
            const int4 rec = c.tr[s * nb4 + __builtin_amdgcn_readlane(colv, kk)];
            if (rec.x < 0) return false;
            c.qrec[kk] = make_int4(rec.z, rec.y, ns, eo);
            c.qsst[kk] = rec.x;
            w.op(rec.y, ns, no);
            s = rec.x;
            ns = rec.y;
            eo = rec.w;
        }
        wsync();
        // (2) entries of the chunk's operators
        for (int kb = 0; kb < nc; kb += 4) {
            const int kk = kb + (l >> 4);
            const bool act = kk < nc;
            const int kq = act ? kk : kb;
            const int4 rec = c.qrec[kq];
            const int j = c0 + kq;
            const int idx = BRA ? n - 1 - j : j;
            const double tj = otau[idx];
            const double tp = j > 0 ? otau[BRA ? idx + 1 : idx - 1] : tstart;
            const double dt = BRA ? tp - tj : tj - tp;
            const double e0 = c.ev[rec.w];
            const bool kin = km < rec.z;
            const double ek = c.ev[rec.w + (kin ? km : 0)];
            const bool in = kin && q.r < rec.y;
            const double pv = c.opsp[in ? rec.x + q.r + km * rec.y : 0];
            const double a = in ? pv * exp(-(ek - e0) * dt) : 0.0;
            if (act) {
                c.qa[kk * 16 + (l & 15)] = a;
                c.qlc[kk] = -(e0 - egs) * dt;
            }
        }
        wsync();
        // (3) products; the next operator's entries, log factor and target
        // are read one operator ahead
        const double *pa = c.qa + q.r * 4;
        double p[4] = {pa[0], pa[1], pa[2], pa[3]};
        double lcv = c.qlc[0];
        int sv = c.qsst[0], ntv = c.qrec[0].y;
        for (int kk = 0; kk < nc; ++kk) {
            const int kn = kk + 1 < nc ? kk + 1 : kk;
            const double pn[4] = {pa[kn * 16], pa[kn * 16 + 1], pa[kn * 16 + 2], pa[kn * 16 + 3]};
            const double lcn = c.qlc[kn];
            const int svn = c.qsst[kn], ntn = c.qrec[kn].y;
            lg += lcv;
            v = quad_apply(v, p, lg);
            after(c0 + kk, v, sv, ntv, lg);
#pragma unroll
            for (int m = 0; m < 4; ++m) p[m] = pn[m];
            lcv = lcn;
            sv = svn;
            ntv = ntn;
        }
        tlast = otau[BRA ? n - c0 - nc : c0 + nc - 1];
        wsync();
    }
    double lc;
    dlast = q.prop(c.ev + eo, ns, BRA ? tlast - tend : tend - tlast, egs, lc);
    lg += lc;
    return true;
}

struct NoAfter {
    DEV void operator()(int, double, int, int, double) const {}
};

typedef double dbl4 __attribute__((ext_vector_type(4)));

// One operator on blocks above 16 states with the FP64 matrix cores:
// Kout (nt x no, row-major) = P diag(dvec) (P: nt x ns, column-major) Kin
// (ns x no, row-major), in 16 x 16 output tiles with k-steps of 4.  For
// v_mfma_f64_16x16x4f64 lane l holds A(l%16, l/16), B(l/16, l%16) and the
// outputs D(l/16 + 4e, l%16), e = 0..3.  Renormalised (by a power of two)
// only when the largest entry leaves [2^-64, 2^64].  Blocks up to 128 states.
DEV void apply_mfma(Ctx &c, const double *P, const double *Kin, double *Kout, int nt, int ns, int no,
                    double &lg)
{
    // a 16-column strip of the output at a time, all its row tiles in
    // registers: per k-step one B fragment serves every row tile
    constexpr int RT = 8;                    // row tiles held (blocks up to 128 states)
    const int l = lane_id();
    const int lr = l & 15, lk = l >> 4;
    const int ntr = (nt + 15) >> 4;
    double m = 0.0;
    for (int i0 = 0; i0 < no; i0 += 16) {
        const int i = i0 + lr;
        dbl4 acc[RT];
#pragma unroll
        for (int t = 0; t < RT; ++t) acc[t] = dbl4{0.0, 0.0, 0.0, 0.0};
#pragma unroll 2
        for (int k0 = 0; k0 < ns; k0 += 4) {
            const int k = k0 + lk;
            const bool kin = k < ns;
            const double b = (kin && i < no) ? Kin[(long)k * no + i] : 0.0;
            const double dk = kin ? c.dvec[k] : 0.0;
#pragma unroll
            for (int t = 0; t < RT; ++t) {
                if (t < ntr) {
                    const int r = t * 16 + lr;
                    const double a = (r < nt && kin) ? P[r + (long)k * nt] * dk : 0.0;
                    acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[t], 0, 0, 0);
                }
            }
        }
#pragma unroll
        for (int t = 0; t < RT; ++t) {
            if (t < ntr) {
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int ro = t * 16 + lk + 4 * e;
                    if (ro < nt && i < no) {
                        Kout[(long)ro * no + i] = acc[t][e];
                        m = fmax(m, fabs(acc[t][e]));
                    }
                }
            }
        }
    }
    wsync();
    if (__ballot(m >= STATE_TINY) == 0ull || __ballot(m > STATE_HUGE) != 0ull) {
        const double mm = wave_max(m);
        if (mm > 0.0) {
            const int e = ilogb(mm);
            lg += (double)e * 0.69314718055994530942;
            for (int idx = l; idx < nt * no; idx += WAVE) Kout[idx] = ldexp(Kout[idx], -e);
        }
        wsync();
    }
}

// What an operator walk over blocks above 16 states needs, by value (the walk
// is out of line; a reference to the chain context would live in scratch).
struct MfmaEnv {
    const int *sz, *off;   // block sizes and first eigenvalue index per superstate
    const int4 *tr;        // transition records (target, its size, block offset, its energies offset)
    const double *ev, *ops;
    const double *ops16;   // zero-padded operator blocks (multiples of 16 x 16)
    const int *off16;      // their offsets, indexed like tr
    double *dvec;          // LDS: propagator factors of the current block
    double *strip;         // LDS: one 16-column strip of the state (128 x 16)
    double egs;
    int nb;
};

struct MfmaOut {
    double lg;             // log scale added by the walk
    double ntns, nss, nts; // algorithmic work: sum of nt*ns, ns, nt over the operators
    int s;                 // superstate after the walk (-1: quantum-number violation)
};

DEV int mfma_tidx(const MfmaEnv &e, int s, int c, bool adjoint)
{
    int ca = op_ca(c);
    if (adjoint) ca = 3 - ca;
    return ((s * e.nb + op_orb(c)) * 2 + op_spin(c)) * 2 + (ca - 1);
}

DEV int4 mfma_trans(const MfmaEnv &e, int s, int c, bool adjoint) { return e.tr[mfma_tidx(e, s, c, adjoint)]; }

// exp(-(E_k - E_0) dt) of superstate s into LDS, zero up to the next multiple
// of 16 (all lanes read it after the barrier)
DEV void mfma_prop(const MfmaEnv &e, int s, double dt)
{
    const int ns = e.sz[s];
    const double *E = e.ev + e.off[s];
    const double e0 = E[0];
    for (int k = lane_id(); k < (ns + 15) / 16 * 16; k += WAVE)
        e.dvec[k] = k < ns ? (ns > 1 ? exp(-(E[k] - e0) * dt) : 1.0) : 0.0;
    wsync();
}

typedef __attribute__((address_space(3))) double lds_d;
typedef __attribute__((address_space(1))) double glb_d;

// out[tp] (tp < NTP) += P diag(dvec) S over nkt groups of 16 k-steps: P the
// zero-padded operator block (leading dimension ldp, global), S the LDS strip
// (rows beyond the block zero).  For v_mfma_f64_16x16x4f64 lane l holds
// A(l%16, l/16) = P(16 tp + l%16, k), B(l/16, l%16) = S(k, l%16), k = k0 + l/16.
// The A fragments of group kt + 1 are loaded before the MFMAs of group kt, so
// that their latency overlaps matrix work.
template <int NTP>
DEV void mfma_kloop(dbl4 (&out)[8], const glb_d *P, int ldp, int nkt, const lds_d *sb, const lds_d *dvec, int lr,
                    int lk)
{
    double a0[4][NTP], a1[4][NTP];
    auto load_a = [&](int kt, double (&a)[4][NTP]) {
#pragma unroll
        for (int ee = 0; ee < 4; ++ee) {
            const glb_d *Pk = P + (long)(16 * kt + 4 * ee + lk) * ldp + lr;
#pragma unroll
            for (int tp = 0; tp < NTP; ++tp) a[ee][tp] = Pk[16 * tp];
        }
    };
    auto step = [&](int kt, const double (&a)[4][NTP]) {
#pragma unroll
        for (int ee = 0; ee < 4; ++ee) {
            const int k = 16 * kt + 4 * ee + lk;
            const double b = sb[k * 16 + lr] * dvec[k];
#pragma unroll
            for (int tp = 0; tp < NTP; ++tp)
                out[tp] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[ee][tp], b, out[tp], 0, 0, 0);
        }
    };
    load_a(0, a0);
    int kt = 0;
    for (; kt + 1 < nkt; kt += 2) {
        load_a(kt + 1, a1);
        step(kt, a0);
        if (kt + 2 < nkt) load_a(kt + 2, a0);
        step(kt + 1, a1);
    }
    if (kt < nkt) step(kt, a0);
}

// Walk of a state (ns x no, row-major, entry (r, i) at K[r * no + i]) over
// operators otau/ocode (ascending tau; BRA: taken in descending order with the
// adjoint blocks) from tstart to tend, for blocks above 16 states, on the FP64
// matrix cores (v_mfma_f64_16x16x4f64).  The state is processed in strips of
// 16 columns (outer states); a strip stays in LDS (rows x 16) through every
// operator of the walk, so the state never goes back to memory between
// operators.  Per operator and k-step of 4 rows, lane l reads the B entry
// (row k0 + l/16, column l%16) of the strip times the source block's
// propagator exp(-(E_k - E_0) dt), and one entry of the operator block (nt x
// ns column-major) for each 16-row output tile; the output tiles accumulate
// in registers and are written back to the strip.  Each strip is
// renormalised by powers of two (exactly) when its largest entry leaves
// [2^-64, 2^64]; at the end the strips are brought to a common power of two
// (exactly) and written to Kout.  The common factors exp(-(E_0 - Egs) dt) go
// into the returned log scale.
template <bool BRA>
__device__ __attribute__((noinline)) MfmaOut walk_mfma(const MfmaEnv e, const double *Kin, double *Kout, int s0,
                                                      int no, double tstart, const double *otau, const int *ocode,
                                                      int n, double tend)
{
    MfmaOut o{0.0, 0.0, 0.0, 0.0, -1};
    const int l = lane_id(), lr = l & 15, lk = l >> 4;
    // address spaces made explicit (LDS and global instructions instead of
    // flat ones): the strip and the propagator factors live in LDS, the
    // operator blocks and the state matrices in global memory (blocks above
    // 16 states)
    lds_d *sb = (lds_d *)e.strip;   // [128 rows][16 columns]
    lds_d *dvec = (lds_d *)e.dvec;
    const glb_d *Kg = (const glb_d *)Kin;
    glb_d *Kog = (glb_d *)Kout;
    // superstate path, quantum numbers and the common log factor (uniform)
    int s = s0;
    double t = tstart;
    for (int jj = 0; jj < n; ++jj) {
        const int j = BRA ? n - 1 - jj : jj;
        const double tj = otau[j];
        const int4 tr = mfma_trans(e, s, ocode[j], BRA);
        if (tr.x < 0) return o;
        o.lg += -(e.ev[e.off[s]] - e.egs) * (BRA ? t - tj : tj - t);
        const double ns = (double)e.sz[s], nt = (double)tr.y;
        o.ntns += nt * ns;
        o.nss += ns;
        o.nts += nt;
        s = tr.x;
        t = tj;
    }
    o.lg += -(e.ev[e.off[s]] - e.egs) * (BRA ? t - tend : tend - t);
    const int sfin = s, nsfin = e.sz[sfin], ns0 = e.sz[s0];
    int emax = -2147483647, es_of[8];
    for (int q = 0; q < 8; ++q) es_of[q] = 0;
    for (int q = 0; q * 16 < no; ++q) {
        // rows beyond the block stay zero up to the 16-row padding
        for (int idx = l; idx < (ns0 + 15) / 16 * 256; idx += WAVE) {
            const int r = idx >> 4, i = q * 16 + (idx & 15);
            sb[idx] = (r < ns0 && i < no) ? Kg[(long)r * no + i] : 0.0;
        }
        wsync();
        int es = 0, ns = ns0;
        s = s0;
        t = tstart;
        for (int jj = 0; jj < n; ++jj) {
            const int j = BRA ? n - 1 - jj : jj;
            const double tj = otau[j];
            const int ti = mfma_tidx(e, s, ocode[j], BRA);
            const int4 tr = e.tr[ti];
            // uniform bounds (scalar loop control, no exec-mask branches)
            const int nt = __builtin_amdgcn_readfirstlane(tr.y);
            const int nkt = __builtin_amdgcn_readfirstlane((ns + 15) >> 4);
            const int ntp = __builtin_amdgcn_readfirstlane((nt + 15) >> 4);
            const int ldp = 16 * ntp;
            mfma_prop(e, s, BRA ? t - tj : tj - t);
            const glb_d *P = (const glb_d *)(e.ops16 + __builtin_amdgcn_readfirstlane(e.off16[ti]));
            dbl4 out[8];
#pragma unroll
            for (int tp = 0; tp < 8; ++tp) out[tp] = dbl4{0.0, 0.0, 0.0, 0.0};
            switch (ntp) {   // row tiles as a compile-time count: straight-line loads and MFMAs
            case 1: mfma_kloop<1>(out, P, ldp, nkt, sb, dvec, lr, lk); break;
            case 2: mfma_kloop<2>(out, P, ldp, nkt, sb, dvec, lr, lk); break;
            case 3: mfma_kloop<3>(out, P, ldp, nkt, sb, dvec, lr, lk); break;
            case 4: mfma_kloop<4>(out, P, ldp, nkt, sb, dvec, lr, lk); break;
            case 5: mfma_kloop<5>(out, P, ldp, nkt, sb, dvec, lr, lk); break;
            case 6: mfma_kloop<6>(out, P, ldp, nkt, sb, dvec, lr, lk); break;
            case 7: mfma_kloop<7>(out, P, ldp, nkt, sb, dvec, lr, lk); break;
            default: mfma_kloop<8>(out, P, ldp, nkt, sb, dvec, lr, lk); break;
            }
            // rows beyond the target block come out exactly zero (padded A)
            double m = 0.0;
#pragma unroll
            for (int tp = 0; tp < 8; ++tp)
                if (tp < ntp)
#pragma unroll
                    for (int ee = 0; ee < 4; ++ee) m = fmax(m, fabs(out[tp][ee]));
            int ex = 0;
            if (__ballot(m >= STATE_TINY) == 0ull || __ballot(m > STATE_HUGE) != 0ull) {
                const double mm = wave_max(m);
                if (mm > 0.0) ex = ilogb(mm);
            }
            es += ex;
            wsync();   // every lane has read the strip and dvec
#pragma unroll
            for (int tp = 0; tp < 8; ++tp)
                if (tp < ntp)
#pragma unroll
                    for (int ee = 0; ee < 4; ++ee) {
                        const int r = 16 * tp + lk + 4 * ee;
                        sb[r * 16 + lr] = ex ? ldexp(out[tp][ee], -ex) : out[tp][ee];
                    }
            wsync();
            s = tr.x;
            ns = nt;
            t = tj;
        }
        // propagation from the last operator to tend, then the strip to Kout
        mfma_prop(e, s, BRA ? t - tend : tend - t);
        for (int idx = l; idx < ns * 16; idx += WAVE) {
            const int r = idx >> 4, i = q * 16 + (idx & 15);
            if (i < no) Kog[(long)r * no + i] = sb[idx] * dvec[r];
        }
        wsync();
        es_of[q] = es;
        emax = max(emax, es);
    }
    // one power of two for the whole state (exact)
    for (int q = 0; q * 16 < no; ++q)
        if (es_of[q] != emax)
            for (int idx = l; idx < nsfin * 16; idx += WAVE) {
                const int r = idx >> 4, i = q * 16 + (idx & 15);
                if (i < no) Kog[(long)r * no + i] = ldexp(Kog[(long)r * no + i], es_of[q] - emax);
            }
    wsync();
    o.lg += (double)emax * 0.69314718055994530942;
    o.s = sfin;
    return o;
}

// scale rows by dvec and renormalise (end-of-sweep propagation)
DEV void scale_rows(Ctx &c, double *K, int ns, int no, double &lg)
{
    double m = 0.0;
    for (int idx = lane_id(); idx < ns * no; idx += WAVE) {
        double x = K[idx] * c.dvec[idx / no];
        K[idx] = x;
        m = fmax(m, fabs(x));
    }
    m = wave_max(m);
    wsync();
    if (m > 0.0) {
        double inv = 1.0 / m;
        lg += log(m);
        for (int idx = lane_id(); idx < ns * no; idx += WAVE) K[idx] *= inv;
    }
    wsync();
}

DEV MfmaEnv mfma_env(const Ctx &c)
{
    MfmaEnv e;
    e.sz = c.sz, e.off = c.off, e.tr = c.tr, e.ev = c.ev, e.ops = c.opsp, e.dvec = c.dvec, e.strip = c.strip;
    e.ops16 = c.M.ops16, e.off16 = c.M.opoff16;
    e.egs = c.M.egs, e.nb = c.M.nb;
    return e;
}

// algorithmic work of an MFMA walk (as Work::commit counts it)
DEV void mfma_work(Ctx &c, const MfmaOut &r, int no)
{
    c.wflops += (double)no * (2.0 * r.ntns + r.nss);
    c.wbytes += 8.0 * no * (r.nss + r.nts);
}

// Ket sweep over a list (ascending tau) from t0 to t1.  Returns false if an
// operator has no target block (quantum-number violation).
template <int NM>
DEV bool ket_walk(Ctx &c, double *&K, double *&Kw, int &s, double &lg, int no, double t0,
                  const double *otau, const int *ocode, int n, double t1)
{
    double t = t0;
    Work w;
    if constexpr (NM <= 4) {
        const Quad q;
        double v = q.load(K, c.sz[s], no);
        double d;
        const bool ok = quad_walk<false>(c, q, v, s, lg, no, t0, otau, ocode, n, t1, d, w, NoAfter());
        w.commit(c);
        if (!ok) return false;
        v = quad_finish(v, d, lg);
        q.store(Kw, c.sz[s], no, v);
        double *x = K;
        K = Kw;
        Kw = x;
        wsync();
        return true;
    } else if constexpr (NM <= 16) {
        // One LDS round trip per operator: the transition record of the next
        // operator, its target block's energies, the current block entries and
        // state are read together; operator codes and times two ahead.
        const LaneSlots<NM> ls(no);
        double tj = n > 0 ? otau[0] : t1;
        int cj = n > 0 ? ocode[0] : 0;
        double tn = n > 1 ? otau[1] : t1;
        int cn = n > 1 ? ocode[1] : 0;
        int ns = c.sz[s];
        double d = prop_e(c, c.ev + c.off[s], ns, tj - t, lg);
        int4 tr = n > 0 ? c.trans(s, cj, false) : make_int4(0, 0, 0, 0);
        for (int j = 0; j < n; ++j) {
            if (tr.x < 0) return w.commit(c), false;
            const double tnn = j + 2 < n ? otau[j + 2] : t1;
            const int cnn = j + 2 < n ? ocode[j + 2] : 0;
            const int4 trn = j + 1 < n ? c.trans(tr.x, cn, false) : make_int4(0, 0, 0, 0);
            const double dn = prop_e(c, c.ev + tr.w, tr.y, tn - tj, lg);
            w.op(tr.y, ns, no);
            apply_small<NM>(c.opsp + tr.z, K, Kw, tr.y, ns, no, d, ls, lg);
            double *x = K;
            K = Kw;
            Kw = x;
            s = tr.x;
            ns = tr.y;
            tr = trn;
            d = dn;
            tj = tn;
            tn = tnn;
            cj = cn;
            cn = cnn;
        }
        finish_small<NM>(K, ns, no, d, ls, lg);
    } else {
        const MfmaOut r = walk_mfma<false>(mfma_env(c), K, Kw, s, no, t, otau, ocode, n, t1);
        mfma_work(c, r, no);
        if (r.s < 0) return false;
        lg += r.lg;
        s = r.s;
        double *x = K;
        K = Kw;
        Kw = x;
        return true;
    }
    w.commit(c);
    return true;
}

// Bra sweep downward from t1 over a list (taken in descending order) to t0.
template <int NM>
DEV void bra_walk(Ctx &c, double *&B, double *&Bw, int &s, double &lg, int no, double t1,
                  const double *otau, const int *ocode, int n, double t0)
{
    double t = t1;
    Work w;
    if constexpr (NM <= 4) {
        const Quad q;
        double v = q.load(B, c.sz[s], no);
        double d;
        quad_walk<true>(c, q, v, s, lg, no, t1, otau, ocode, n, t0, d, w, NoAfter());
        w.commit(c);
        v = quad_finish(v, d, lg);
        q.store(Bw, c.sz[s], no, v);
        double *x = B;
        B = Bw;
        Bw = x;
        wsync();
        return;
    } else if constexpr (NM <= 16) {
        const LaneSlots<NM> ls(no);
        double tj = n > 0 ? otau[n - 1] : t0;
        int cj = n > 0 ? ocode[n - 1] : 0;
        double tn = n > 1 ? otau[n - 2] : t0;
        int cn = n > 1 ? ocode[n - 2] : 0;
        int ns = c.sz[s];
        double d = prop_e(c, c.ev + c.off[s], ns, t - tj, lg);
        int4 tr = n > 0 ? c.trans(s, cj, true) : make_int4(0, 0, 0, 0);
        for (int j = n - 1; j >= 0; --j) {
            const double tnn = j >= 2 ? otau[j - 2] : t0;
            const int cnn = j >= 2 ? ocode[j - 2] : 0;
            const int4 trn = j >= 1 ? c.trans(tr.x, cn, true) : make_int4(0, 0, 0, 0);
            const double dn = prop_e(c, c.ev + tr.w, tr.y, tj - tn, lg);
            w.op(tr.y, ns, no);
            apply_small<NM>(c.opsp + tr.z, B, Bw, tr.y, ns, no, d, ls, lg);
            double *x = B;
            B = Bw;
            Bw = x;
            s = tr.x;
            ns = tr.y;
            tr = trn;
            d = dn;
            tj = tn;
            tn = tnn;
            cj = cn;
            cn = cnn;
        }
        finish_small<NM>(B, ns, no, d, ls, lg);
    } else {
        const MfmaOut r = walk_mfma<true>(mfma_env(c), B, Bw, s, no, t, otau, ocode, n, t0);
        mfma_work(c, r, no);
        if (r.s < 0) return;
        lg += r.lg;
        s = r.s;
        double *x = B;
        B = Bw;
        Bw = x;
        return;
    }
    w.commit(c);
}

// stage global operators [j0, j1) into the LDS list (chunks of lcap)
DEV int stage_chunk(Ctx &c, int j0, int j1)
{
    const int n = min(j1 - j0, c.L.lcap_lds);
    for (int j = lane_id(); j < n; j += WAVE) {
        c.ttau[j] = c.tau[j0 + j];
        c.tcode[j] = c.code[j0 + j];
    }
    wsync();
    return n;
}

// ket over global operators [j0, j1) in LDS-staged chunks
template <int NM>
DEV bool ket_walk_global(Ctx &c, double *&K, double *&Kw, int &s, double &lg, int no, double t0, int j0,
                         int j1, double t1)
{
    // one walk call site: an empty range is one chunk of no operators
    double t = t0;
    int b = j0;
    do {
        const int n = b < j1 ? stage_chunk(c, b, j1) : 0;
        const double tend = (b + n < j1) ? c.ttau[n - 1] : t1;
        // walk the chunk; the final propagation of a non-last chunk is zero
        if (!ket_walk<NM>(c, K, Kw, s, lg, no, t, c.ttau, c.tcode, n, tend)) return false;
        t = tend;
        b += n;
        wsync();
    } while (b < j1);
    return true;
}

template <int NM>
DEV void bra_walk_global(Ctx &c, double *&B, double *&Bw, int &s, double &lg, int no, double t1, int j0,
                         int j1, double t0)
{
    double t = t1;
    int e = j1;
    do {
        const int b = max(j0, e - c.L.lcap_lds);
        const int n = e > j0 ? stage_chunk(c, b, e) : 0;
        const double tend = (b > j0) ? c.ttau[0] : t0;
        bra_walk<NM>(c, B, Bw, s, lg, no, t, c.ttau, c.tcode, n, tend);
        t = tend;
        e = b;
        wsync();
    } while (e > j0);
}

DEV void set_identity(double *K, int no)
{
    for (int idx = lane_id(); idx < no * no; idx += WAVE) K[idx] = (idx / no == idx % no) ? 1.0 : 0.0;
    wsync();
}

DEV void copy_mat(double *dst, const double *src, int n)
{
    for (int i = lane_id(); i < n; i += WAVE) dst[i] = src[i];
    wsync();
}

// signed log-sum (addLogTr, Trace.F90:286-298)
DEV void log_add(double &al, double &as, double bl, double bs)
{
    if (al >= bl) {
        al = al + log(1.0 + as * bs * exp(bl - al));
    } else {
        double nl = bl + log(1.0 + as * bs * exp(al - bl));
        al = nl;
        as = bs;
    }
}

// Contract ket K and bra B (same superstate s, both at the same tau) column by
// column; part traces to c.pl, total in (tl, ts).
DEV void contract(Ctx &c, const double *K, const double *B, int s, int no, double lg, double &tl, double &ts)
{
    // every column carries the same scale exp(lg), so the total is the plain
    // sum of the columns (the log-space sum of addLogTr, Trace.F90:286-298, of
    // the same numbers)
    const int ns = c.sz[s];
    double part = 0.0;
    for (int i = lane_id(); i < no; i += WAVE) {
        double acc = 0.0;
        for (int r = 0; r < ns; ++r) acc = fma(K[r * no + i], B[r * no + i], acc);
        c.pl[2 * i] = lg + log(fabs(acc));
        c.pl[2 * i + 1] = copysign(1.0, acc);
        part += acc;
    }
    const double S = wave_sum(part);
    if (S == 0.0) {
        tl = -DBL_MAX;
        ts = 0.0;
    } else {
        tl = lg + log(fabs(S));
        ts = copysign(1.0, S);
    }
    wsync();
}

// Full trace of the current configuration (global moves, tau shifts,
// set_config, initial state): identity ket on the outer superstate at tau = 0
// through every operator, contracted with the identity bra at beta.
// A configuration that breaks the quantum numbers (an operator without a
// target block) has a zero trace: returns false with tl = -DBL_MAX, ts = 0.
template <int NM>
DEV bool full_trace(Ctx &c, int outer, double &tl, double &ts)
{
    const int no = c.sz[outer];
    double *K = c.K0, *Kw = c.K1;
    set_identity(K, no);
    int s = outer;
    double lg = 0.0;
    if (!ket_walk_global<NM>(c, K, Kw, s, lg, no, 0.0, 0, c.h.nop, c.M.beta) || s != outer) {
        tl = -DBL_MAX;
        ts = 0.0;
        wsync();
        return false;
    }
    set_identity(Kw, no);
    contract(c, K, Kw, s, no, lg, tl, ts);
    return true;
}

// ---------------------------------------------------------------------------
// boundary cache: ket at the window's lower edge, bras at every upper edge
// ---------------------------------------------------------------------------
template <int NM>
DEV void rebuild_cache(Ctx &c)
{
    const int outer = c.h.outer, no = c.sz[outer];
    const int p = c.h.nwin > 1 ? c.h.wpos : 0;
    const int nw = c.h.nwin;
    const int top = (nw > 1) ? nw + 1 : 2;
    const int stride = c.nm * c.nm;
    double *B = c.K0, *Bw = c.K1;
    set_identity(B, no);
    int s = outer;
    double lg = 0.0;
    double tb = c.M.beta;
    int hiidx = c.h.nop;
    copy_mat(c.bnd + (long)top * stride, B, no * no);
    if (lane_id() == 0) {
        c.bndlog[top] = lg;
        c.bndsst[top] = s;
    }
    for (int m = top - 1; m >= p + 2; --m) {
        const double tm = bound_tau(m, nw, c.M.beta);
        const int loidx = upper_bound_tau(c.tau, c.h.nop, tm);
        bra_walk_global<NM>(c, B, Bw, s, lg, no, tb, loidx, hiidx, tm);
        copy_mat(c.bnd + (long)m * stride, B, c.sz[s] * no);
        if (lane_id() == 0) {
            c.bndlog[m] = lg;
            c.bndsst[m] = s;
        }
        tb = tm;
        hiidx = loidx;
    }
    double *K = c.K0, *Kw = c.K1;
    set_identity(K, no);
    s = outer;
    lg = 0.0;
    const double tp = bound_tau(p, nw, c.M.beta);
    const int n0 = lower_bound_tau(c.tau, c.h.nop, tp);
    ket_walk_global<NM>(c, K, Kw, s, lg, no, 0.0, 0, n0, tp);
    copy_mat(c.kst, K, c.sz[s] * no);
    c.h.kst_log = lg;
    c.h.kst_sst = s;
    c.h.cache_pos = p;
    wsync();
}

// Bring the caches to the current window position: advance the ket across
// passed half windows, or rebuild after a wrap / invalidation.
template <int NM>
DEV void ensure_cache(Ctx &c)
{
    const int p = c.h.nwin > 1 ? c.h.wpos : 0;
    if (c.h.cache_pos == p) return;
    c.h.rk_n = 0;
    c.h.lb_n = 0;
    if (c.h.cache_pos < 0 || c.h.cache_pos > p) {
        rebuild_cache<NM>(c);
        return;
    }
    const int no = c.sz[c.h.outer];
    int s = c.h.kst_sst;
    double lg = c.h.kst_log;
    double *K = c.K0, *Kw = c.K1;
    copy_mat(K, c.kst, c.sz[s] * no);
    const double ta = bound_tau(c.h.cache_pos, c.h.nwin, c.M.beta);
    const double tb = bound_tau(p, c.h.nwin, c.M.beta);
    const int ia = lower_bound_tau(c.tau, c.h.nop, ta);
    const int ib = lower_bound_tau(c.tau, c.h.nop, tb);
    ket_walk_global<NM>(c, K, Kw, s, lg, no, ta, ia, ib, tb);
    copy_mat(c.kst, K, c.sz[s] * no);
    c.h.kst_log = lg;
    c.h.kst_sst = s;
    c.h.cache_pos = p;
    wsync();
}

// Trace of the LDS trial list between the cached ket and bra; returns false
// on a quantum-number violation (check_sst_sequence, Trace.F90:1243-1302, on a
// superset of the changed span).
template <int NM>
DEV bool window_trace(Ctx &c, int nl, double &tl, double &ts)
{
    const int no = c.sz[c.h.outer];
    const int p = c.h.nwin > 1 ? c.h.wpos : 0;
    const int top = p + 2;
    const double ta = bound_tau(p, c.h.nwin, c.M.beta), tb = bound_tau(top, c.h.nwin, c.M.beta);
    int s = c.h.kst_sst;
    double lg = c.h.kst_log;
    if constexpr (NM <= 4) {
        const Quad q;
        double v = q.load(c.kst, c.sz[s], no);
        double d;
        Work w;
        const bool ok = quad_walk<false>(c, q, v, s, lg, no, ta, c.ttau, c.tcode, nl, tb, d, w, NoAfter());
        w.commit(c);
        if (!ok || s != c.bndsst[top]) return false;
        quad_contract(c, q, v * d, c.bnd + (long)top * c.nm * c.nm, s, no, lg + c.bndlog[top], tl, ts);
        return true;
    }
    double *K = c.K0, *Kw = c.K1;
    copy_mat(K, c.kst, c.sz[s] * no);
    if (!ket_walk<NM>(c, K, Kw, s, lg, no, ta, c.ttau, c.tcode, nl, tb)) return false;
    if (s != c.bndsst[top]) return false;
    contract(c, K, c.bnd + (long)top * c.nm * c.nm, s, no, lg + c.bndlog[top], tl, ts);
    return true;
}

// ---------------------------------------------------------------------------
// Window segment caches (blocks of up to 4 states).  rk[j] is the ket after the
// window's first j operators, at the time of operator j-1 (rk[0] = the cached
// ket at the window's lower edge); lb[m] is the bra after the window's last m
// operators, at the time of operator nw-m (lb[0] = the cached bra at the upper
// edge).  A proposal then walks only the operators between its changes: an
// insertion at window positions q1 <= q2 walks the two new operators and the
// q2-q1 between them, a removal of w1 < w2 the w2-w1-1 between them.  The
// caches are extended on demand and cut back to their unchanged parts after
// an accepted move; both restart when the window list is reloaded or the
// boundary caches change.
// ---------------------------------------------------------------------------
template <int NM>
DEV void rk_extend(Ctx &c, int to)
{
    constexpr int S2 = NM * NM;
    const int no = c.sz[c.h.outer];
    const double ta = bound_tau(c.h.nwin > 1 ? c.h.wpos : 0, c.h.nwin, c.M.beta);
    int j = c.h.rk_n - 1;
    if (j < 0) {
        copy_mat(c.rk, c.kst, c.sz[c.h.kst_sst] * no);
        c.rkx[0] = c.h.kst_log;
        c.rks[0] = c.h.kst_sst;
        j = 0;
    }
    if (j < to) {
        Work w;
        int s = c.rks[j];
        double lg = c.rkx[j];
        const Quad q;
        double v = q.load(c.rk + (long)j * S2, c.sz[s], no);
        double d;
        const int j0 = j;
        quad_walk<false>(c, q, v, s, lg, no, j0 > 0 ? c.wtau[j0 - 1] : ta, c.wtau + j0, c.wcode + j0, to - j0,
                         c.wtau[to - 1], d, w, [&](int k, double vk, int sk, int nsk, double lgk) {
                             q.store(c.rk + (long)(j0 + k + 1) * S2, nsk, no, vk);
                             c.rkx[j0 + k + 1] = lgk;
                             c.rks[j0 + k + 1] = sk;
                         });
        w.commit(c);
    }
    c.h.rk_n = max(c.h.rk_n, to + 1);
    wsync();
}

template <int NM>
DEV void lb_extend(Ctx &c, int nw, int to)
{
    constexpr int S2 = NM * NM;
    const int no = c.sz[c.h.outer];
    const int top = (c.h.nwin > 1 ? c.h.wpos : 0) + 2;
    const double tb = bound_tau(top, c.h.nwin, c.M.beta);
    int m = c.h.lb_n - 1;
    if (m < 0) {
        copy_mat(c.lb, c.bnd + (long)top * c.nm * c.nm, c.sz[c.bndsst[top]] * no);
        c.lbx[0] = c.bndlog[top];
        c.lbs[0] = c.bndsst[top];
        m = 0;
    }
    if (m < to) {
        Work w;
        int s = c.lbs[m];
        double lg = c.lbx[m];
        const Quad q;
        double v = q.load(c.lb + (long)m * S2, c.sz[s], no);
        double d;
        const int m0 = m;
        // the window operators nw-to .. nw-1-m0, walked downward
        quad_walk<true>(c, q, v, s, lg, no, m0 > 0 ? c.wtau[nw - m0] : tb, c.wtau + (nw - to),
                        c.wcode + (nw - to), to - m0, c.wtau[nw - to], d, w,
                        [&](int k, double vk, int sk, int nsk, double lgk) {
                            q.store(c.lb + (long)(m0 + k + 1) * S2, nsk, no, vk);
                            c.lbx[m0 + k + 1] = lgk;
                            c.lbs[m0 + k + 1] = sk;
                        });
        w.commit(c);
    }
    c.h.lb_n = max(c.h.lb_n, to + 1);
    wsync();
}

// Trace of the window with the trial list's entries [a, a + n) replacing the
// window's operators between rk[j0] and lb[m0]; t0 / t1 are the times of
// those two states.  Returns false on a quantum-number violation.
template <int NM>
DEV bool segment_trace(Ctx &c, int nw, int j0, int m0, int a, int n, double t0, double t1, double &tl, double &ts)
{
    constexpr int S2 = NM * NM;
    const int no = c.sz[c.h.outer];
    if (c.h.rk_n <= j0) rk_extend<NM>(c, j0);
    if (c.h.lb_n <= m0) lb_extend<NM>(c, nw, m0);
    int s = c.rks[j0];
    double lg = c.rkx[j0];
    if constexpr (NM <= 4) {
        const Quad q;
        double v = q.load(c.rk + (long)j0 * S2, c.sz[s], no);
        double d;
        Work w;
        const bool ok = quad_walk<false>(c, q, v, s, lg, no, t0, c.ttau + a, c.tcode + a, n, t1, d, w, NoAfter());
        w.commit(c);
        if (!ok || s != c.lbs[m0]) return false;
        quad_contract(c, q, v * d, c.lb + (long)m0 * S2, s, no, lg + c.lbx[m0], tl, ts);
        return true;
    }
    double *K = c.K0, *Kw = c.K1;
    copy_mat(K, c.rk + (long)j0 * S2, c.sz[s] * no);
    if (!ket_walk<NM>(c, K, Kw, s, lg, no, t0, c.ttau + a, c.tcode + a, n, t1)) return false;
    if (s != c.lbs[m0]) return false;
    contract(c, K, c.lb + (long)m0 * S2, s, no, lg + c.lbx[m0], tl, ts);
    return true;
}

// ---------------------------------------------------------------------------
// removable pairs (ensure_valid_NPairs_diag, Trace.F90:1769-1810) in a list
// ---------------------------------------------------------------------------
DEV int pairs_from(const Ctx &c, const double *tau, const int *code, int n, int j)
{
    const double te = tau[j];
    if (te < c.h.tauwin_min || te > c.h.tauwin_max) return 0;
    const int ce = code[j];
    const int fe = c.fl(ce), ae = op_ca(ce);
    // partners: later operators up to min(window end, te + taudiff_max); the
    // list is read four entries at a time so that the LDS loads overlap
    const double lim = fmin(c.h.tauwin_max, te + c.h.taudiff_max);
    int cnt = 0;
    for (int k = j + 1; k < n; k += 4) {
        double tk[4];
        int ck[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            tk[u] = k + u < n ? tau[k + u] : DBL_MAX;
            ck[u] = k + u < n ? code[k + u] : 0;
        }
        bool stop = false;
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            stop = stop || tk[u] > lim;
            if (!stop && c.fl(ck[u]) == fe && op_ca(ck[u]) != ae) ++cnt;
        }
        if (stop) break;
    }
    return cnt;
}

DEV int npairs_list(const Ctx &c, const double *tau, const int *code, int n)
{
    int cnt = 0;
    for (int j = lane_id(); j < n; j += WAVE) cnt += pairs_from(c, tau, code, n, j);
    return wave_isum(cnt);
}

// ---------------------------------------------------------------------------
// configuration sign: get_Sign with get_BlockSign and get_permaltca
// (Trace.F90:3479-3615).  Parities of
//   BlockSign: sum over operators j of the counts of later operators of the
//     flavours (b', s') with b' < o_j, plus (o_j, s'=0) when s_j = 1;
//   permaltca: sum over creators of the count of later same-flavour operators,
//     plus 1 for an odd number of pairs;
// computed per flavour g with ballots over chunks of 64 operators.
// ---------------------------------------------------------------------------
DEV double config_sign(Ctx &c)
{
    const int n = c.h.nop, nb = c.M.nb, nf = c.M.nf;
    // stage the operator codes in LDS (the trial buffer is free here)
    const int *codes = c.code;
    if (n <= 2 * c.L.lcap_lds) {
        int *st = reinterpret_cast<int *>(c.ttau);
        for (int j = lane_id(); j < n; j += WAVE) st[j] = c.code[j];
        wsync();
        codes = st;
    }
    int parity = 0, alt = 0;
    for (int g = 0; g < nf; ++g) {
        const int bg = g % nb, sg = g / nb;
        const int mg = 2 * c.h.kf[g];
        int before = 0;   // flavour-g operators before the current chunk
        for (int b0 = 0; b0 < n; b0 += WAVE) {
            const int j = b0 + lane_id();
            int cj = 0;
            bool isg = false;
            if (j < n) {
                cj = codes[j];
                isg = c.fl(cj) == g;
            }
            const unsigned long long mask = __ballot(isg);
            if (j < n) {
                const int upto = before + popc64(mask & lanemask_le());   // flavour g up to j
                const int later = mg - upto;
                const int oj = op_orb(cj), sj = op_spin(cj);
                const bool in_set = (bg < oj) || (sj == 1 && bg == oj && sg == 0);
                if (in_set) parity ^= later & 1;
                if (isg && op_ca(cj) == 1) alt ^= later & 1;
            }
            before += popc64(mask);
        }
    }
    wsync();
    parity = wave_isum(parity) & 1;
    alt = wave_isum(alt) & 1;
    if ((n / 2) % 2 == 1) alt ^= 1;
    const double block = parity ? -1.0 : 1.0, altsign = alt ? -1.0 : 1.0;
    return c.h.trace_sign * c.h.det_sign * block * altsign;
}

// ---------------------------------------------------------------------------
// hybridisation matrix algebra (MatrixUpdate.F90)
// ---------------------------------------------------------------------------
// S - R^T M Q for an order-k block (get_DetRatAdd)
DEV double detrat_add(Ctx &c, const double *Mf, int k, double S)
{
    // lanes over the k^2 elements: every load of M is issued at once
    const int ld = c.L.kcap;
    double part = 0.0;
    for (int idx = lane_id(); idx < k * k; idx += WAVE) {
        const int i = idx % k, j = idx / k;
        part = fma(c.R[i] * Mf[i + (long)j * ld], c.Q[j], part);
    }
    return S - wave_sum(part);
}

// Bordering update to order k+1, get_MatAdd (MatrixUpdate.F90:410-442), on
// physical storage: the new creator takes physical row k, the new annihilator
// physical column k, so no entry moves.  c.Q holds Q by physical column, c.R
// holds R by physical row.
//   col(i) = -(M Q)_i / dr,  row(j) = -(R M)_j / dr,
//   M(i,j) += dr col(i) row(j),  M(i,k) = col(i),  M(k,j) = row(j),  M(k,k) = 1/dr
DEV void mat_add_phys(Ctx &c, double *Mf, int k, double dr)
{
    const int ld = c.L.kcap;
    const int l = lane_id();
    double *col = c.colb, *row = c.rowb;
    for (int i = l; i < k; i += WAVE) {
        double v = 0.0;
#pragma unroll 8
        for (int j = 0; j < k; ++j) v = fma(Mf[i + (long)j * ld], c.Q[j], v);
        col[i] = -v / dr;
    }
    for (int j = l; j < k; j += WAVE) {
        double v = 0.0;
        const double *mc = Mf + (long)j * ld;
#pragma unroll 8
        for (int i = 0; i < k; ++i) v = fma(c.R[i], mc[i], v);
        row[j] = -v / dr;
    }
    wsync();
    for (int idx = l; idx < k * k; idx += WAVE) {
        const int i = idx % k, j = idx / k;
        Mf[i + (long)j * ld] = fma(dr * col[i], row[j], Mf[i + (long)j * ld]);
    }
    for (int i = l; i < k; i += WAVE) {
        Mf[i + (long)k * ld] = col[i];
        Mf[k + (long)i * ld] = row[i];
    }
    if (l == 0) Mf[k + (long)k * ld] = 1.0 / dr;
    wsync();
}

// Removal of the creator in physical row sc and the annihilator in physical
// column sa, get_MatRem (MatrixUpdate.F90:455-482):
//   M(i,j) -= M(i,sa) M(sc,j) / M(sc,sa);
// then the last physical row / column move into the freed slots.
DEV void mat_rem_phys(Ctx &c, double *Mf, int k, int sc, int sa)
{
    const int ld = c.L.kcap;
    const int l = lane_id();
    double *colp = c.colb, *rowp = c.rowb;
    for (int i = l; i < k; i += WAVE) {
        colp[i] = Mf[i + (long)sa * ld];
        rowp[i] = Mf[sc + (long)i * ld];
    }
    wsync();
    const double inv = 1.0 / rowp[sa];
    for (int idx = l; idx < k * k; idx += WAVE) {
        const int i = idx % k, j = idx / k;
        Mf[i + (long)j * ld] = fma(-colp[i] * inv, rowp[j], Mf[i + (long)j * ld]);
    }
    wsync();
    const int last = k - 1;
    if (sc != last) {
        for (int j = l; j < k; j += WAVE) Mf[sc + (long)j * ld] = Mf[last + (long)j * ld];
        wsync();
    }
    if (sa != last) {
        for (int i = l; i < k; i += WAVE) Mf[i + (long)sa * ld] = Mf[i + (long)last * ld];
        wsync();
    }
}

// LU with partial pivoting of A (k x k, ld) in place; log|det| and sign as
// get_LogDetFull (MatrixUpdate.F90:302-339).  Returns false if singular.
DEV bool lu_inplace(double *A, int k, int ld, double &dl, double &ds, int *piv)
{
    const int l = lane_id();
    dl = 0.0;
    ds = 1.0;
    for (int col = 0; col < k; ++col) {
        double best = -1.0;
        int bi = col;
        for (int r = col + l; r < k; r += WAVE) {
            double v = fabs(A[r + (long)col * ld]);
            if (v > best) best = v, bi = r;
        }
        // largest modulus, lowest row among ties
        const double wb = wave_max(best);
        bi = wave_imin(best == wb ? bi : 0x7fffffff);
        best = wb;
        if (best == 0.0) return false;
        if (l == 0) piv[col] = bi;
        if (bi != col) {
            for (int j = l; j < k; j += WAVE) {
                double t = A[col + (long)j * ld];
                A[col + (long)j * ld] = A[bi + (long)j * ld];
                A[bi + (long)j * ld] = t;
            }
        }
        wsync();
        double d = A[col + (long)col * ld];
        ds *= (bi != col ? -1.0 : 1.0) * (d >= 0 ? 1.0 : -1.0);
        dl += log(fabs(d));
        for (int r = col + 1 + l; r < k; r += WAVE) A[r + (long)col * ld] /= d;
        wsync();
        // trailing update, lanes over its elements
        const int m = k - col - 1;
        for (int idx = l; idx < m * m; idx += WAVE) {
            const int r = col + 1 + idx % m, j = col + 1 + idx / m;
            A[r + (long)j * ld] -= A[r + (long)col * ld] * A[col + (long)j * ld];
        }
        wsync();
    }
    return true;
}

// Inverse from an in-place LU (row swaps piv) into out (k x k, leading
// dimension ldo): out = U^-1 L^-1 P, by substitution on all columns at once,
// lanes over the elements of each step.
DEV void lu_inverse(const double *LU, int k, int ld, const int *piv, double *out, int ldo)
{
    const int l = lane_id();
    for (int idx = l; idx < k * k; idx += WAVE) out[idx % k + (long)(idx / k) * ldo] = 0.0;
    wsync();
    // row r of the identity ends in row pos after the swaps
    for (int r = l; r < k; r += WAVE) {
        int pos = r;
        for (int i = 0; i < k; ++i) {
            const int p = piv[i];
            pos = pos == i ? p : (pos == p ? i : pos);
        }
        out[pos + (long)r * ldo] = 1.0;
    }
    wsync();
    for (int i = 0; i + 1 < k; ++i) {
        const int m = k - i - 1;
        for (int idx = l; idx < m * k; idx += WAVE) {
            const int r = i + 1 + idx % m, cc = idx / m;
            out[r + (long)cc * ldo] -= LU[r + (long)i * ld] * out[i + (long)cc * ldo];
        }
        wsync();
    }
    for (int i = k - 1; i >= 0; --i) {
        for (int cc = l; cc < k; cc += WAVE) out[i + (long)cc * ldo] /= LU[i + (long)i * ld];
        wsync();
        for (int idx = l; idx < i * k; idx += WAVE) {
            const int r = idx % i, cc = idx / i;
            out[r + (long)cc * ldo] -= LU[r + (long)i * ld] * out[i + (long)cc * ldo];
        }
        wsync();
    }
}

// Delta_f(row = annihilator rank, col = creator rank) = F(tau_a - tau_c)
// (get_InvFullHybr, Trace.F90:2279-2333) from the per-flavour lists
DEV void build_delta(Ctx &c, int f, int k, double *D, int ld, const double *fcl, const double *fal)
{
    const double *ct = fcl + (long)f * c.L.kcap, *at = fal + (long)f * c.L.kcap;
    for (int idx = lane_id(); idx < k * k; idx += WAVE) {
        int r = idx % k, col = idx / k;
        D[r + (long)col * ld] = hybr(c.M, at[r] - ct[col], f);
    }
    wsync();
}

// ---------------------------------------------------------------------------
// array edits (lane-parallel shifts; chunks ordered so nothing is overwritten
// before it is read)
// ---------------------------------------------------------------------------
// Move arr[from, to) by `by` places (either sign).  Rounds of 8 elements per
// lane: all loads of a round complete before its stores, and rounds run from
// the end that is written into first, so nothing is overwritten unread.
template <typename T>
DEV void block_move(T *arr, int from, int to, int by)
{
    constexpr int U = 8, R = U * WAVE;
    const int l = lane_id();
    const int cnt = to - from;
    if (cnt <= 0 || by == 0) return;
    const int nr = (cnt + R - 1) / R;
    for (int q = 0; q < nr; ++q) {
        const int r = by > 0 ? nr - 1 - q : q;
        const int b = from + r * R;
        T v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int i = b + u * WAVE + l;
            if (i < to) v[u] = arr[i];
        }
        wsync();
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int i = b + u * WAVE + l;
            if (i < to) arr[i + by] = v[u];
        }
        wsync();
    }
}

// parallel count of entries < x in a sorted list of n
DEV int count_below_par(const double *arr, int n, double x)
{
    int cnt = 0;
    for (int j = lane_id(); j < n; j += WAVE) cnt += arr[j] < x;
    return wave_isum(cnt);
}

// ---------------------------------------------------------------------------
// moves
// ---------------------------------------------------------------------------
enum { OUT_REJ = 0, OUT_ACC = 1, OUT_QN = 2 };
enum { ERR_CAP = 3, ERR_SING = 4, ERR_STATE_DEV = 6 };

DEV void store_part(Ctx &c, int no)
{
    for (int i = lane_id(); i < 2 * no; i += WAVE) c.part[i] = c.pl[i];
    wsync();
}

// write the LDS trial list into the global operator list at wia (its length
// nl replaces the window's nw operators: the tail [wib, nop) moves first); the
// trial list then serves as the window list
DEV void commit_window(Ctx &c, int nl)
{
    const int nw = c.h.wib - c.h.wia;
    block_move(c.tau, c.h.wib, c.h.nop, nl - nw);
    block_move(c.code, c.h.wib, c.h.nop, nl - nw);
    for (int j = lane_id(); j < nl; j += WAVE) {
        c.tau[c.h.wia + j] = c.ttau[j];
        c.code[c.h.wia + j] = c.tcode[j];
    }
    wsync();
    c.h.nop += nl - nw;
    c.h.wib = c.h.wia + nl;
    swap_lists(c);
}

// An accepted insertion or removal of a pair of flavour f spanning dt
// (CTQMC.F90:3579-3587, 4119-4127): the histogram AccPair over NAccPair bins
// of [0, AccPairMax] and, while the calibration list is open, the distance.
DEV void record_pair(Ctx &c, int f, double dt)
{
    if (lane_id() == 0) {
        double *ap = reinterpret_cast<double *>(c.base + c.L.accpair) + (long)f * c.L.naccpair_cap;
        const int bin = min(c.h.naccpair, (int)floor(dt / c.h.accpair_max * (double)c.h.naccpair) + 1);
        ap[bin - 1] += 1.0;
        if (c.h.apt_index <= c.h.apt_n) reinterpret_cast<double *>(c.base + c.L.aptau)[c.h.apt_index - 1] = dt;
    }
    if (c.h.apt_index <= c.h.apt_n) c.h.apt_index++;
}

// StepAdd (CTQMC.F90:3353-3612), change_outer = false
template <int NM>
DEV int step_add(Ctx &c)
{
    const DevModel &M = c.M;
    c.h.try_add++;
    const double width = c.h.tauwin_max - c.h.tauwin_min;
    const double tref = c.h.tauwin_min + width * grnd(c.rng);
    double tdf = width;
    const int orb = randint(c.rng, 1, M.nb) - 1;
    const int spin = randint(c.rng, 1, 2) - 1;
    const double tmin = fmax(tref - c.h.taudiff_max, c.h.tauwin_min);
    const double w2 = fmin(tref + c.h.taudiff_max, c.h.tauwin_max) - tmin;
    const double t2 = tmin + w2 * grnd(c.rng);
    tdf *= w2;
    const int f = orb + M.nb * spin;
    if (t2 == tref) return c.h.qn_rej++, OUT_QN;
    const int k = c.h.kf[f];
    // room_for_move checked these before the move's first draw; a failure here
    // is a logic error, not a full chain
    if (!ensure_window(c) || c.h.wib - c.h.wia + 2 > c.L.lcap_lds || k + 1 > c.L.kcap ||
        c.h.nop + 2 > c.L.cap) {
        c.h.err = ERR_STATE_DEV;
        return OUT_REJ;
    }
    const int nw = c.h.wib - c.h.wia;
    ensure_cache<NM>(c);
    // equal times (process_OperAdd, Trace.F90:4287-4380) and insertion points
    const double ta = tref, tc = t2;
    const double t1 = fmin(ta, tc), t2o = fmax(ta, tc);
    const int code_a = mk_code(orb, spin, 2), code_c = mk_code(orb, spin, 1);
    const int code1 = ta < tc ? code_a : code_c, code2 = ta < tc ? code_c : code_a;
    int eq = 0, q1 = 0, q2 = 0;
    for (int j = lane_id(); j < nw; j += WAVE) {
        const double x = c.wtau[j];
        eq |= (x == t1) | (x == t2o);
        q1 += x < t1;
        q2 += x < t2o;
    }
    if (wave_isum(eq)) return c.h.qn_rej++, OUT_QN;
    q1 = wave_isum(q1);
    q2 = wave_isum(q2);
    // trial list: the window with t1 before index q1 and t2o before index q2
    for (int j = lane_id(); j < nw + 2; j += WAVE) {
        double x;
        int cd;
        if (j < q1) x = c.wtau[j], cd = c.wcode[j];
        else if (j == q1) x = t1, cd = code1;
        else if (j <= q2) x = c.wtau[j - 1], cd = c.wcode[j - 1];
        else if (j == q2 + 1) x = t2o, cd = code2;
        else x = c.wtau[j - 2], cd = c.wcode[j - 2];
        c.ttau[j] = x;
        c.tcode[j] = cd;
    }
    wsync();
    const int nl = nw + 2;
    double tl, ts;
    const bool seg = NM <= 4 && nw <= c.L.sc;
    if (seg) {
        const int p = c.h.nwin > 1 ? c.h.wpos : 0;
        const double t0 = q1 > 0 ? c.wtau[q1 - 1] : bound_tau(p, c.h.nwin, c.M.beta);
        const double t1e = q2 < nw ? c.wtau[q2] : bound_tau(p + 2, c.h.nwin, c.M.beta);
        if (!segment_trace<NM>(c, nw, q1, nw - q2, q1, q2 - q1 + 2, t0, t1e, tl, ts))
            return c.h.qn_rej++, OUT_QN;
    } else if (!window_trace<NM>(c, nl, tl, ts)) {
        return c.h.qn_rej++, OUT_QN;
    }
    // det ratio, get_LogDetRatPairAdd (Trace.F90:2775-2832): Q (annihilators) and
    // R (creators) placed by physical slot; FPos = 1 + earlier operators of the type
    const long fo = (long)f * c.L.kcap;
    const double *ctl = c.fc + fo, *atl = c.fa + fo;
    const int *csl = c.fcs + fo, *asl = c.fas + fo;
    int fpc = 0, fpa = 0;
    for (int j = lane_id(); j < k; j += WAVE) {
        const double xa = atl[j], xc = ctl[j];
        c.Q[asl[j]] = hybr(M, xa - tc, f);
        c.R[csl[j]] = hybr(M, ta - xc, f);
        fpa += xa < ta;
        fpc += xc < tc;
    }
    const int fp_c = wave_isum(fpc) + 1, fp_a = wave_isum(fpa) + 1;
    wsync();
    const double S = detrat_add(c, c.mi(f), k, hybr(M, ta - tc, f));
    c.wflops += 2.0 * k * k;
    c.wbytes += 8.0 * k * k;
    const double sig = (abs(fp_c - fp_a) % 2) ? -1.0 : 1.0;
    const double dr_log = log(fabs(S)), dr_sign = copysign(1.0, S) * sig;
    const double rnd = grnd(c.rng);
    // NPairs of the new configuration: pairs with both ends in the window all
    // lie in the trial list
    const int np = npairs_list(c, c.ttau, c.tcode, nl);
    const double wr = exp((tl - c.h.trace_log) + dr_log);
    const double prob = fabs(((1.0 * tdf) * (double)(2 * M.nb)) / (double)np * wr);
    if (!(rnd < prob)) return OUT_REJ;
    // accept: M, operator list, flavour lists, weights, sign
    mat_add_phys(c, c.mi(f), k, S);
    c.wflops += 6.0 * k * k;
    c.wbytes += 8.0 * k * k + 8.0 * (k + 1) * (k + 1);
    commit_window(c, nl);
    c.h.pc_valid = 0;
    // segment caches: the first q1 and the last nw - q2 window operators are unchanged
    c.h.rk_n = seg ? min(c.h.rk_n, q1 + 1) : 0;
    c.h.lb_n = seg ? min(c.h.lb_n, nw - q2 + 1) : 0;
    double *fcf = c.fc + fo, *faf = c.fa + fo;
    int *fcsf = c.fcs + fo, *fasf = c.fas + fo;
    block_move(fcf, fp_c - 1, k, 1);
    block_move(fcsf, fp_c - 1, k, 1);
    block_move(faf, fp_a - 1, k, 1);
    block_move(fasf, fp_a - 1, k, 1);
    if (lane_id() == 0) {
        fcf[fp_c - 1] = tc;
        fcsf[fp_c - 1] = k;
        faf[fp_a - 1] = ta;
        fasf[fp_a - 1] = k;
    }
    wsync();
    c.h.kf[f] = k + 1;
    c.h.kfmax[f] = max(c.h.kfmax[f], k + 1);
    record_pair(c, f, t2o - t1);
    c.h.det_log += dr_log;
    c.h.det_sign *= dr_sign;
    c.h.trace_log = tl;
    c.h.trace_sign = ts;
    store_part(c, c.sz[c.h.outer]);
    c.h.cfgsign = -config_sign(c);
    c.h.acc_add++;
    return OUT_ACC;
}

// rank whose slot equals `slot` in a slot list of length k (-1 if none)
DEV int find_slot(const int *sl, int k, int slot)
{
    int r = -1;
    for (int j = lane_id(); j < k; j += WAVE)
        if (sl[j] == slot) r = j;
    return wave_imax(r);
}

// StepRem (CTQMC.F90:3892-4150) with gen_OperRemove (Trace.F90:4851-4977)
template <int NM>
DEV int step_rem(Ctx &c)
{
    const DevModel &M = c.M;
    c.h.try_rem++;
    if (!ensure_window(c)) {
        c.h.err = ERR_STATE_DEV;
        return OUT_REJ;
    }
    const int nw = c.h.wib - c.h.wia;
    ensure_cache<NM>(c);
    // pair counts per window operator (the window holds every removable pair)
    // (kept in LDS until the window list changes)
    if (!c.h.pc_valid) {
        for (int j = lane_id(); j < nw; j += WAVE) c.pc[j] = pairs_from(c, c.wtau, c.wcode, nw, j);
        c.h.pc_valid = 1;
        wsync();
    }
    int np = 0;
    for (int j0 = 0; j0 < nw; j0 += WAVE) {
        const int j = j0 + lane_id();
        np += wave_isum(j < nw ? c.pc[j] : 0);
    }
    if (np <= 0) return c.h.qn_rej++, OUT_QN;
    int pick = randint(c.rng, 1, np);
    int w1 = -1, w2 = -1;
    {
        int base = 0;
        for (int j0 = 0; j0 < nw && w1 < 0; j0 += WAVE) {
            const int j = j0 + lane_id();
            const int cnt = j < nw ? c.pc[j] : 0;
            int tot;
            const int ex = wave_excl_scan(cnt, &tot);
            int found = (pick > base + ex && pick <= base + ex + cnt) ? j : -1;
            found = wave_imax(found);
            if (found >= 0) {
                int want = pick - base - __shfl(ex, found - j0, WAVE);
                const int ce = c.wcode[found];
                const int fe = c.fl(ce), ae = op_ca(ce);
                int kk = found + 1;
                for (; kk < nw; ++kk) {
                    const int ck = c.wcode[kk];
                    if (c.fl(ck) == fe && op_ca(ck) != ae)
                        if (--want <= 0) break;
                }
                w1 = found;
                w2 = kk;
            }
            base += tot;
        }
    }
    if (w1 < 0 || w2 >= nw) {
        c.h.err = ERR_STATE_DEV;
        return OUT_REJ;
    }
    const int c1 = c.wcode[w1];
    const double t1 = c.wtau[w1], t2 = c.wtau[w2];
    const int f = c.fl(c1);
    const int k = c.h.kf[f];
    const double tc = op_ca(c1) == 1 ? t1 : t2, ta = op_ca(c1) == 1 ? t2 : t1;
    const double hprf = (double)np /
                        ((c.h.tauwin_max - c.h.tauwin_min) *
                         (fmin(ta + c.h.taudiff_max, c.h.tauwin_max) - fmax(ta - c.h.taudiff_max, c.h.tauwin_min)));
    // trial list: the window without entries w1 < w2
    const int nl = nw - 2;
    for (int j = lane_id(); j < nl; j += WAVE) {
        const int src = j < w1 ? j : (j < w2 - 1 ? j + 1 : j + 2);
        c.ttau[j] = c.wtau[src];
        c.tcode[j] = c.wcode[src];
    }
    wsync();
    double tl, ts;
    const bool seg = NM <= 4 && nw <= c.L.sc;
    if (seg) {
        const int p = c.h.nwin > 1 ? c.h.wpos : 0;
        const double t0 = w1 > 0 ? c.wtau[w1 - 1] : bound_tau(p, c.h.nwin, c.M.beta);
        const double t1e = w2 + 1 < nw ? c.wtau[w2 + 1] : bound_tau(p + 2, c.h.nwin, c.M.beta);
        if (!segment_trace<NM>(c, nw, w1, nw - 1 - w2, w1, w2 - w1 - 1, t0, t1e, tl, ts))
            return c.h.qn_rej++, OUT_QN;
    } else if (!window_trace<NM>(c, nl, tl, ts)) {
        return c.h.qn_rej++, OUT_QN;
    }
    const long fo = (long)f * c.L.kcap;
    const int fp_c = count_below_par(c.fc + fo, k, tc) + 1, fp_a = count_below_par(c.fa + fo, k, ta) + 1;
    const int sc = c.fcs[fo + fp_c - 1], sa = c.fas[fo + fp_a - 1];
    // get_LogDetRatPairRem (Trace.F90:2937-2963): M(FPos(1), FPos(2))
    const int ld = c.L.kcap;
    const double S = c.mi(f)[sc + (long)sa * ld];
    const double sig = (abs(fp_c - fp_a) % 2) ? -1.0 : 1.0;
    const double dr_log = log(fabs(S)), dr_sign = copysign(1.0, S) * sig;
    const double wr = exp((tl - c.h.trace_log) + dr_log);
    const double prob = fabs(((1.0 * hprf) / (double)(2 * M.nb)) * wr);
    const double rnd = grnd(c.rng);
    if (!(rnd < prob)) return OUT_REJ;
    if (fabs(dr_sign * exp(dr_log)) < 1e-12) {
        c.h.err = ERR_SING;
        return OUT_REJ;
    }
    mat_rem_phys(c, c.mi(f), k, sc, sa);
    c.wflops += 3.0 * k * k;
    c.wbytes += 8.0 * k * k + 8.0 * (k - 1) * (k - 1);
    // the operators that owned the last physical row / column take the freed slots
    int *fcsf = c.fcs + fo, *fasf = c.fas + fo;
    const int rl = find_slot(fcsf, k, k - 1), al = find_slot(fasf, k, k - 1);
    if (lane_id() == 0) {
        fcsf[rl] = sc;
        fasf[al] = sa;
    }
    wsync();
    commit_window(c, nl);
    c.h.pc_valid = 0;
    // segment caches: the first w1 and the last nw - 1 - w2 window operators are unchanged
    c.h.rk_n = seg ? min(c.h.rk_n, w1 + 1) : 0;
    c.h.lb_n = seg ? min(c.h.lb_n, nw - w2) : 0;
    block_move(c.fc + fo, fp_c, k, -1);
    block_move(fcsf, fp_c, k, -1);
    block_move(c.fa + fo, fp_a, k, -1);
    block_move(fasf, fp_a, k, -1);
    c.h.kf[f] = k - 1;
    record_pair(c, f, t2 - t1);
    c.h.det_log += dr_log;
    c.h.det_sign *= dr_sign;
    c.h.trace_log = tl;
    c.h.trace_sign = ts;
    store_part(c, c.sz[c.h.outer]);
    c.h.cfgsign = -config_sign(c);
    c.h.acc_rem++;
    return OUT_ACC;
}

// relabel flavours / swap c and c+ (globalUpdate, Trace.F90:2126-2188)
DEV void global_update(Ctx &c, const int *gu)
{
    const DevModel &M = c.M;
    for (int j = lane_id(); j < c.h.nop; j += WAVE) {
        const int cj = c.code[j];
        const int idx = op_orb(cj) + M.nb * op_spin(cj) + 1;
        const int g = gu[idx];
        int ca = op_ca(cj);
        if (gu[0] != 0) ca = 3 - ca;
        c.code[j] = mk_code((g - 1) % M.nb, (g - 1) / M.nb, ca);
    }
    wsync();
}

// per-flavour tau lists and orders from the operator list: for every flavour,
// ballot-compact the creators / annihilators of each 64-operator chunk.  With
// slot lists given, slots are reset to the identity (logical = physical order).
DEV void rebuild_flavour_lists(Ctx &c, double *dfc, double *dfa, int *dkf, int *dfcs, int *dfas)
{
    const DevModel &M = c.M;
    const long kc = c.L.kcap;
    for (int g = 0; g < M.nf; ++g) {
        int nc = 0, na = 0;
        for (int b0 = 0; b0 < c.h.nop; b0 += WAVE) {
            const int j = b0 + lane_id();
            int cj = 0;
            double tj = 0.0;
            if (j < c.h.nop) {
                cj = c.code[j];
                tj = c.tau[j];
            }
            const bool isc = j < c.h.nop && c.fl(cj) == g && op_ca(cj) == 1;
            const bool isa = j < c.h.nop && c.fl(cj) == g && op_ca(cj) == 2;
            const unsigned long long mc = __ballot(isc), ma = __ballot(isa);
            const unsigned long long below = (lane_id() == 0) ? 0ull : (~0ull >> (64 - lane_id()));
            if (isc) dfc[g * kc + nc + popc64(mc & below)] = tj;
            if (isa) dfa[g * kc + na + popc64(ma & below)] = tj;
            nc += popc64(mc);
            na += popc64(ma);
        }
        if (dfcs)
            for (int r = lane_id(); r < nc; r += WAVE) {
                dfcs[g * kc + r] = r;
                dfas[g * kc + r] = r;
            }
        if (lane_id() == 0) dkf[g] = nc;
        wsync();
    }
    wsync();
}

DEV bool qn_cycle_global(Ctx &c, int s0)
{
    int s = s0;
    for (int j = 0; j < c.h.nop && s >= 0; ++j) s = c.conn(s, c.code[j], false);
    return s == s0;
}

// lane-parallel: which superstates give a valid cyclic sequence (lane = start
// superstate, up to 64 per pass); bit s of the returned masks
// okmask: one bit per superstate (at most MAXSST)
DEV int count_valid_outer(Ctx &c, unsigned long long *okmask)
{
    int total = 0;
    for (int s0 = 0; s0 < c.M.nsst; s0 += WAVE) {
        const int s = s0 + lane_id();
        bool ok = false;
        if (s < c.M.nsst) {
            int cur = s;
            for (int j = 0; j < c.h.nop && cur >= 0; ++j) cur = c.conn(cur, c.code[j], false);
            ok = (cur == s);
        }
        const unsigned long long m = __ballot(ok);
        okmask[s0 / WAVE] = m;
        total += popc64(m);
    }
    return total;
}

// StepGlob (CTQMC.F90:4433-4554)
template <int NM>
DEV int step_glob(Ctx &c)
{
    const DevModel &M = c.M;
    c.h.try_glob++;
    if (c.h.nop == 0) return OUT_REJ;
    const int nf = M.nf, nb = M.nb;
    int gu[MAXF + 1], igu[MAXF + 1];
    const double r = grnd(c.rng);
    bool spinflip = false;
    if (r < 0.3) {
        spinflip = true;
        gu[0] = 0;
        for (int i = 1; i <= nb; ++i) gu[i] = i + nb;
        for (int i = nb + 1; i <= 2 * nb; ++i) gu[i] = i - nb;
    } else if (r < 0.6) {
        gu[0] = 1;
        for (int i = 1; i <= nf; ++i) gu[i] = i;
    } else if (NM <= 16 && r < 0.8 && M.nsym > 0) {
        // gen_SymUpdate (Trace.F90:1166-1182): a user permutation or its inverse
        // (engines up to 16 states; the MFMA engine refuses NSymMove)
        if constexpr (NM <= 16) {
            const int isym = randint(c.rng, 1, 2 * M.nsym);
            gu[0] = 0;
            for (int i = 1; i <= nf; ++i) gu[i] = M.symmoves[(isym - 1) * nf + (i - 1)];
        }
    } else {
        for (int i = 1; i <= nf; ++i) gu[i] = i;
        for (int j = 1; j <= nf; ++j) {   // rand_perm (Trace.F90:1065-1080)
            const int i = randint(c.rng, j, nf);
            const int t = gu[j];
            gu[j] = gu[i];
            gu[i] = t;
        }
        gu[0] = randint(c.rng, 0, 1);
    }
    for (int i = 1; i <= nf; ++i) igu[gu[i]] = i;
    igu[0] = gu[0];
    const int old_outer = c.h.outer;
    double sst_factor;
    unsigned long long okm[MAXSST / WAVE];   // valid outer superstates, one bit each
    if (spinflip && M.has_szt) {
        // perform_spinflip_check_qn (Trace.F90:1306-1339)
        const int st = M.first_state[c.h.outer];
        const int dn = st & ((1 << nb) - 1), up = (st >> nb) & ((1 << nb) - 1);
        const int nst = up | (dn << nb);
        c.h.outer = M.state_sst[nst];
        global_update(c, gu);
        sst_factor = qn_cycle_global(c, c.h.outer) ? 1.0 : 0.0;
    } else {
        // perform_global_check_qn (Trace.F90:1345-1383)
        const int pold = count_valid_outer(c, okm);
        global_update(c, gu);
        const int pnew = count_valid_outer(c, okm);
        if (pnew <= 0) {
            sst_factor = 0.0;
        } else {
            sst_factor = (double)pnew / (double)pold;
            int kk = randint(c.rng, 1, pnew);
            for (int s = 0; s < M.nsst; ++s)
                if (((okm[s / WAVE] >> (s % WAVE)) & 1ull) && --kk <= 0) {
                    c.h.outer = s;
                    break;
                }
        }
    }
    if (sst_factor == 0.0) {
        c.h.outer = old_outer;
        global_update(c, igu);
        return c.h.qn_rej++, OUT_QN;
    }
    // flavour lists of the relabelled configuration in the operator scratch
    const long kc = c.L.kcap;
    double *tfc = c.scratch + c.L.kmat, *tfa = tfc + nf * kc;
    int *tkf = reinterpret_cast<int *>(tfa + nf * kc);
    rebuild_flavour_lists(c, tfc, tfa, tkf, nullptr, nullptr);
    // det(new Delta) * det(old M) (get_InvFullHybr, get_FullLogDet); the old
    // inverse is gathered into logical order first
    int *piv = reinterpret_cast<int *>(c.R);
    double dnl = 0.0, dns = 1.0, dol = 0.0, dos = 1.0;
    const int ld = c.L.kcap;
    // LDS work area for the k x k blocks (blocks up to 4 states: the window
    // segment caches, restarted after this move); HBM scratch otherwise
    double *wa = nullptr;
    long wcap = 0;
    if constexpr (NM <= 4) {
        wa = c.rk;
        wcap = 2L * (c.L.sc + 1) * NM * NM;
        c.h.rk_n = 0;
        c.h.lb_n = 0;
    }
    bool sing = false;
    for (int f = 0; f < nf; ++f) {
        const int k = tkf[f];
        if (k > 0) {
            const bool inl = (long)k * k <= wcap;
            double *A = inl ? wa : c.scratch;
            const int lda = inl ? k : ld;
            build_delta(c, f, k, A, lda, tfc, tfa);
            double l, s;
            if (!lu_inplace(A, k, lda, l, s, piv)) sing = true;
            dnl += l;
            dns *= s;
        }
        const double k3 = (double)k * k * k;
        c.wflops += (2.0 / 3.0) * k3 * 2.0;
        c.wbytes += 8.0 * (double)k * k;
    }
    // det(old M) = 1 / det(old Delta), which the chain carries
    dol = -c.h.det_log;
    dos = c.h.det_sign;
    double tl, ts;
    full_trace<NM>(c, c.h.outer, tl, ts);
    const double wr = exp((tl - c.h.trace_log) + (dnl + dol));
    const double rr = grnd(c.rng);   // drawn even for a singular Delta (prob = 0)
    if (!sing && rr < fabs(wr * sst_factor)) {
        rebuild_flavour_lists(c, c.fc, c.fa, c.h.kf, c.fcs, c.fas);
        for (int f = 0; f < nf; ++f) {
            const int k = c.h.kf[f];
            if (k == 0) continue;
            double l, s;
            if (2L * k * k <= wcap) {
                double *A = wa, *X = wa + (long)k * k;
                build_delta(c, f, k, A, k, c.fc, c.fa);
                lu_inplace(A, k, k, l, s, piv);
                lu_inverse(A, k, k, piv, X, k);
                double *Mf = c.mi(f);
                for (int idx = lane_id(); idx < k * k; idx += WAVE)
                    Mf[idx % k + (long)(idx / k) * ld] = X[idx];
                wsync();
            } else {
                build_delta(c, f, k, c.scratch, ld, c.fc, c.fa);
                lu_inplace(c.scratch, k, ld, l, s, piv);
                lu_inverse(c.scratch, k, ld, piv, c.mi(f), ld);
            }
        }
        c.h.det_log = dnl;
        c.h.det_sign = dns;
        c.h.trace_log = tl;
        c.h.trace_sign = ts;
        store_part(c, c.sz[c.h.outer]);
        c.h.cfgsign = -config_sign(c);
        c.h.cache_pos = -1;
        c.h.wvalid = 0;   // relabelled codes: reload the window list
        c.h.acc_glob++;
        return OUT_ACC;
    }
    c.h.outer = old_outer;
    global_update(c, igu);
    return OUT_REJ;
}

// empty-trace weights of every superstate into c.Q (sum returned)
template <int NM>
DEV double empty_weights(Ctx &c)
{
    double sum = 0.0;
    for (int q = 0; q < c.M.nsst; ++q) {
        double tl, ts;
        full_trace<NM>(c, q, tl, ts);
        const double v = ts * exp(tl);
        if (lane_id() == 0) c.Q[q] = v;
        sum += v;
    }
    wsync();
    return sum;
}

// choose the outer superstate proportional to the empty traces (init_Trace,
// Trace.F90:640-670; StepChangeOuter, CTQMC.F90:4569-4637)
template <int NM>
DEV void choose_outer(Ctx &c)
{
    const double sum = empty_weights<NM>(c);
    const double r = grnd(c.rng);
    double acc = 0.0;
    int pick = c.M.nsst - 1;
    for (int q = 0; q < c.M.nsst; ++q) {
        const double v = c.Q[q] / sum;
        acc += v;
        if (acc >= r && v != 0.0) {
            pick = q;
            break;
        }
    }
    c.h.outer = pick;
    double tl, ts;
    full_trace<NM>(c, pick, tl, ts);
    c.h.trace_log = tl;
    c.h.trace_sign = ts;
    store_part(c, c.sz[pick]);
    c.h.cache_pos = -1;
}

// StepShiftTau_proper (CTQMC.F90:4641-4835) + update_InvFullHybrShiftTau_diag
// (Trace.F90:2355-2457)
template <int NM>
DEV int step_shift(Ctx &c)
{
    const DevModel &M = c.M;
    if (c.h.nop == 0) {
        choose_outer<NM>(c);
        c.h.cfgsign = -config_sign(c);
        return OUT_REJ;
    }
    const double dtau = grnd(c.rng) * M.beta;
    if (dtau == 0.0 || dtau == M.beta) return OUT_REJ;
    c.h.try_shift++;
    const int n = c.h.nop, nf = M.nf;
    int first = lower_bound_tau(c.tau, n, dtau);
    if (first == n) first = 0;
    int s = c.h.outer;
    for (int i = 0; i < first; ++i) s = c.conn(s, c.code[i], false);
    // rotate through the operator scratch area behind the matrix scratch
    double *opscr = c.scratch + c.L.kmat;
    double *old_tau = opscr;
    int *old_code = reinterpret_cast<int *>(opscr + n);
    for (int i = lane_id(); i < n; i += WAVE) {
        old_tau[i] = c.tau[i];
        old_code[i] = c.code[i];
    }
    wsync();
    for (int i = lane_id(); i < n; i += WAVE) {
        const int src = (first + i) % n;
        const double t = old_tau[src] - dtau;
        c.tau[i] = t - floor(t / M.beta) * M.beta;
        c.code[i] = old_code[src];
    }
    wsync();
    const int old_outer = c.h.outer, old_pos = c.h.wpos;
    const double old_min = c.h.tauwin_min, old_max = c.h.tauwin_max;
    c.h.outer = s;
    if (c.h.nwin > 1) {
        const int nw = c.h.nwin;
        const int a = (int)floor((dtau / M.beta) * (double)(nw + 1));
        const int am = ((a % nw) + nw) % nw;
        const int p = old_pos - am;
        c.h.wpos = ((p % nw) + nw) % nw;
    } else {
        c.h.wpos = 0;
    }
    set_window(c);
    double tl, ts;
    full_trace<NM>(c, c.h.outer, tl, ts);
    const double ratio = ts * c.h.trace_sign * exp(tl - c.h.trace_log);
    if (grnd(c.rng) < fabs(1.0 * ratio * 1.0)) {
        c.h.trace_log = tl;
        c.h.trace_sign = ts;
        store_part(c, c.sz[c.h.outer]);
        double detrat = 1.0;
        const int ld = c.L.kcap;
        const long kc = c.L.kcap;
        for (int f = 0; f < nf; ++f) {
            const int k = c.h.kf[f];
            double *fcf = c.fc + f * kc, *faf = c.fa + f * kc;
            int *fcsf = c.fcs + f * kc, *fasf = c.fas + f * kc;
            const int w1 = count_below_par(fcf, k, dtau);
            const int w2 = count_below_par(faf, k, dtau);
            double *Mf = c.mi(f);
            if (!((w1 == 0 && w2 == 0) || (w1 == k && w2 == k))) {
                // negate the physical rows of wrapped creators and columns of
                // wrapped annihilators (entries wrapped on both sides keep sign)
                double *rneg = c.colb, *cneg = c.rowb;
                for (int r = lane_id(); r < k; r += WAVE) {
                    rneg[fcsf[r]] = r < w1 ? 1.0 : 0.0;
                    cneg[fasf[r]] = r < w2 ? 1.0 : 0.0;
                }
                wsync();
                for (int idx = lane_id(); idx < k * k; idx += WAVE) {
                    const int i = idx % k, j = idx / k;
                    if (rneg[i] != cneg[j]) Mf[i + (long)j * ld] = -Mf[i + (long)j * ld];
                }
                wsync();
            }
            // rotate ranks: the wrapped operators move to the end, taus shifted
            // like the operator list
            int *rs = reinterpret_cast<int *>(c.R);
            for (int r = lane_id(); r < k; r += WAVE) {
                const int src = (r + w1) % k;
                const double t = fcf[src] - dtau;
                c.Q[r] = t - floor(t / M.beta) * M.beta;
                rs[r] = fcsf[src];
            }
            wsync();
            for (int r = lane_id(); r < k; r += WAVE) {
                fcf[r] = c.Q[r];
                fcsf[r] = rs[r];
            }
            wsync();
            for (int r = lane_id(); r < k; r += WAVE) {
                const int src = (r + w2) % k;
                const double t = faf[src] - dtau;
                c.Q[r] = t - floor(t / M.beta) * M.beta;
                rs[r] = fasf[src];
            }
            wsync();
            for (int r = lane_id(); r < k; r += WAVE) {
                faf[r] = c.Q[r];
                fasf[r] = rs[r];
            }
            wsync();
            if (k % 2 != 0 && (w1 + w2) % 2 != 0) detrat = -detrat;
        }
        c.h.det_sign *= detrat;
        c.h.cfgsign = -config_sign(c);
        c.h.cache_pos = -1;
        c.h.acc_shift++;
        return OUT_ACC;
    }
    for (int i = lane_id(); i < n; i += WAVE) {
        c.tau[i] = old_tau[i];
        c.code[i] = old_code[i];
    }
    wsync();
    c.h.outer = old_outer;
    c.h.wpos = old_pos;
    c.h.tauwin_min = old_min;
    c.h.tauwin_max = old_max;
    c.h.wvalid = 0;
    return OUT_REJ;
}

// one Metropolis step with the dispatch of ctqmc_warmup / ctqmc_measure
// (CTQMC.F90:8213-8254, 8432-8533) and the window-shift rule
// Whether the next move fits the chain's memory whatever it proposes: every
// flavour can take one more pair and the window plus a pair fits the LDS
// lists.  Checked before the move's first random draw, so a chain that has no
// room stops with its state and random stream intact and continues exactly
// where it stopped once the host has enlarged its memory.  Loads the window
// list (as the next insertion or removal would).
DEV bool room_for_move(Ctx &c)
{
    int km = 0;
    for (int f = 0; f < c.M.nf; ++f) km = max(km, c.h.kf[f]);
    if (km + 1 > c.L.kcap || c.h.nop + 2 > c.L.cap) return false;
    if (!ensure_window(c)) return false;
    return c.h.wib - c.h.wia + 2 <= c.L.lcap_lds;
}

template <int NM>
DEV void one_move(Ctx &c, int &kind, int &out)
{
    const DevModel &M = c.M;
    const double r = grnd(c.rng);
    if (r < (1.0 - M.p_shift - M.p_glob) / 2.0) {
        kind = 1;
        (void)grnd(c.rng);               // Percentage4OperatorMove draw
        (void)(grnd(c.rng) < M.p_outer); // PercentageOuterMove draw
        out = step_add<NM>(c);
    } else if (r <= 1.0 - M.p_shift - M.p_glob) {
        kind = 2;
        (void)grnd(c.rng);
        (void)(grnd(c.rng) < M.p_outer);
        out = step_rem<NM>(c);
    } else if (r <= 1.0 - M.p_shift) {
        kind = 3;
        out = step_glob<NM>(c);
    } else {
        kind = 4;
        out = step_shift<NM>(c);
    }
    c.h.moves++;
    if ((c.h.try_add + c.h.try_rem) % M.nslide == 0) shift_window(c);
}

// ---------------------------------------------------------------------------
// measurements
// ---------------------------------------------------------------------------
// Reduce-scatter of 16 values over the 16 lanes of a DPP row: afterwards lane
// i of the row holds the sum over the row of v[i & 15].  Pairing steps: row
// mirror (i <-> 15-i), half-row mirror (i <-> 7-i), quad swaps (i <-> i^2,
// i <-> i^1); at each step a lane keeps the half its position selects and
// receives the partner's copy of that half.
DEV double row_reduce_scatter16(double (&v)[16])
{
    const int i = lane_id() & 15;
    {
        const bool hi = i >= 8;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const double keep = hi ? v[j + 8] : v[j], give = hi ? v[j] : v[j + 8];
            v[j] = keep + dpp_d<DPP_ROW_MIRROR>(give);
        }
    }
    {
        const bool hi = (i & 7) >= 4;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const double keep = hi ? v[j + 4] : v[j], give = hi ? v[j] : v[j + 4];
            v[j] = keep + dpp_d<DPP_ROW_HALF_MIRROR>(give);
        }
    }
    {
        const bool hi = (i & 3) >= 2;
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const double keep = hi ? v[j + 2] : v[j], give = hi ? v[j] : v[j + 2];
            v[j] = keep + dpp_d<DPP_QP_2301>(give);
        }
    }
    const bool hi = (i & 1) != 0;
    const double keep = hi ? v[1] : v[0], give = hi ? v[0] : v[1];
    return keep + dpp_d<DPP_QP_1032>(give);
}

// MeasGtauRem (CTQMC.F90:2252-2440): lanes over (annihilator, creator) pairs,
// up to four pairs per lane.  G(tau) bins by atomic adds; the Legendre sums run
// the Bonnet recursion per pair and fold 16 orders at a time with
// row_reduce_scatter16 into per-lane accumulators (up to 128 orders per pass
// over the pairs), combined across rows once per flavour.  Kept out of line: inlined into the sweep kernel, the
// ROCm 7.2 backend emits an invalid V_CMP (src_shared_base operand) for it.
struct GtauArgs {
    int nf, ntau, nleg, kcap;
    double beta, cfgsign;
    const double *legc;
    const int *kf;
    const double *fc, *fa, *minv;
    const int *fcs, *fas;
    double *gtau, *gleg;
    double *mstep, *dstep;   // this measurement's row of Gtau_mean_step / Gtau_mid_step (or null)
};

// One pass over the pairs of flavour f for the Legendre orders [l0, l0 + 128):
// every lane takes PPL pairs (p = p0 + lane + 64u), runs their Bonnet
// recursions side by side and folds the PPL products of each order before the
// reduce-scatter, so one reduce-scatter serves 64 PPL pairs.
template <int PPL>
DEV void gtau_pass(const GtauArgs &a, int f, int l0, double (&acc)[8])
{
    constexpr int LPASS = 128, NCH = LPASS / 16;
    const int ld = a.kcap;
    const int k = a.kf[f];
    const double *ctl = a.fc + (long)f * a.kcap, *atl = a.fa + (long)f * a.kcap;
    const int *csl = a.fcs + (long)f * a.kcap, *asl = a.fas + (long)f * a.kcap;
    const double *Mf = a.minv + (long)f * a.kcap * a.kcap;
    const int npair = k * k;
    const int lend = min(a.nleg, l0 + LPASS);
    for (int p0 = 0; p0 < npair; p0 += WAVE * PPL) {
        double Mv[PPL], x[PPL], pl[PPL], plm1[PPL];
#pragma unroll
        for (int u = 0; u < PPL; ++u) {
            const int p = p0 + lane_id() + WAVE * u;
            Mv[u] = 0.0;
            x[u] = 0.0;
            if (p < npair) {
                const int ai = p / k, ci = p - ai * k;    // annihilator rank, creator rank
                const double ta = atl[ai], tc = ctl[ci];
                double m = Mf[csl[ci] + (long)asl[ai] * ld];
                int idt;
                if (ta - tc > 0.0) {
                    idt = (int)((ta - tc) / a.beta * (double)a.ntau) + 1;
                    x[u] = 2.0 / a.beta * (ta - tc) - 1.0;
                } else {
                    idt = (int)((a.beta + ta - tc) / a.beta * (double)a.ntau) + 1;
                    x[u] = 2.0 / a.beta * (a.beta + ta - tc) - 1.0;
                    m = -m;
                }
                Mv[u] = m * a.cfgsign;
                if (idt > a.ntau) idt = a.ntau;
                if (l0 == 0) {
                    global_add(&a.gtau[(long)f * a.ntau + idt - 1], Mv[u]);
                    // step-resolved sums (CTQMC.F90:2394-2401): M * cfgsign of
                    // every pair, and of pairs with |ta - tc| / beta in [0.4, 0.6]
                    if (a.mstep) global_add(&a.mstep[f], Mv[u] * a.cfgsign);
                    if (a.dstep) {
                        const double r = fabs(ta - tc) / a.beta;
                        if (r >= 0.4 && r <= 0.6) global_add(&a.dstep[f], Mv[u] * a.cfgsign);
                    }
                }
            }
            pl[u] = 0.0;
            plm1[u] = 0.0;
        }
        // Bonnet recursion from order 0 up to the start of this pass
        for (int l = 0; l < l0; ++l) {
#pragma unroll
            for (int u = 0; u < PPL; ++u) {
                const double plp1 = l == 0 ? 1.0 : (l == 1 ? x[u] : a.legc[2 * (l - 1)] * x[u] * pl[u] - a.legc[2 * (l - 1) + 1] * plm1[u]);
                plm1[u] = pl[u];
                pl[u] = plp1;
            }
        }
#pragma unroll
        for (int q = 0; q < NCH; ++q) {
            const int lb = l0 + 16 * q;
            if (lb >= lend) break;
            double v[16];
#pragma unroll
            for (int j = 0; j < 16; ++j) {
                const int l = lb + j;
                double sum = 0.0;
                if (l < lend) {
                    const double c1 = l >= 2 ? a.legc[2 * (l - 1)] : 0.0, c2 = l >= 2 ? a.legc[2 * (l - 1) + 1] : 0.0;
#pragma unroll
                    for (int u = 0; u < PPL; ++u) {
                        const double plp1 = l == 0 ? 1.0 : (l == 1 ? x[u] : c1 * x[u] * pl[u] - c2 * plm1[u]);
                        plm1[u] = pl[u];
                        pl[u] = plp1;
                        sum = fma(Mv[u], plp1, sum);
                    }
                }
                v[j] = sum;
            }
            acc[q] += row_reduce_scatter16(v);
        }
    }
}

// all flavours; WIDE: flavours of more than 64 pairs take 2 or 4 pairs per lane
template <bool WIDE>
DEV double2 gtau_all(const GtauArgs &a)
{
    constexpr int LPASS = 128, NCH = LPASS / 16;
    double wflops = 0.0, wbytes = 0.0;
    for (int f = 0; f < a.nf; ++f) {
        const int k = a.kf[f];
        wflops += (double)k * k * (5.0 * a.nleg + 4.0);
        wbytes += 8.0 * k * k;
        const int npair = k * k;
        for (int l0 = 0; l0 < a.nleg; l0 += LPASS) {
            double acc[NCH];
#pragma unroll
            for (int q = 0; q < NCH; ++q) acc[q] = 0.0;
            if (WIDE && npair > 2 * WAVE) gtau_pass<4>(a, f, l0, acc);
            else if (WIDE && npair > WAVE) gtau_pass<2>(a, f, l0, acc);
            else gtau_pass<1>(a, f, l0, acc);
            // rows -> lanes 0..15, added to the chain's Legendre sums
#pragma unroll
            for (int q = 0; q < NCH; ++q) {
                const int lb = l0 + 16 * q;
                if (lb >= a.nleg) break;
                double r = acc[q];
                r += __shfl_xor(r, 16, WAVE);
                r += __shfl_xor(r, 32, WAVE);
                const int l = lb + lane_id();
                if (lane_id() < 16 && l < a.nleg) a.gleg[(long)f * a.nleg + l] += r;
            }
        }
    }
    wsync();
    return make_double2(wflops, wbytes);
}

// Two out-of-line entry points: the caller's registers saved around the call
// depend on the callee's footprint, so kernels whose chains stay at a few
// pairs per flavour (one-state blocks) call the narrow one.
__device__ __attribute__((noinline)) double2 meas_gtau(const GtauArgs a) { return gtau_all<true>(a); }
__device__ __attribute__((noinline)) double2 meas_gtau_narrow(const GtauArgs a) { return gtau_all<false>(a); }

// gtau_accum%add() of the binning accumulator (AccumGtau = 2,
// autocorracc_add, Accumulators.F90:298-340): this measurement's G(tau) bins
// (gbuf, filled by atomics) go into the sum and into level 0 by the Welford
// update; when the low bits of the count n are ones the finished bin of each
// level (mean of its halves) rolls over to the next.  Lanes over components;
// the roll-over depth depends on n only, so it is uniform.
DEV void autocorr_add(const DevModel &M, double *gbuf, double *gacc, double *gtau, long long n0)
{
    const long nc = (long)M.nf * M.ntau;
    double *spill = gacc, *mean = gacc + nc * M.nlev, *sq = gacc + 2 * nc * M.nlev;
    wsync();
    for (long i = lane_id(); i < nc; i += WAVE) {
        // the bins were written by atomics at L2: read past the vector L1
        double v = __hip_atomic_load(gbuf + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        gbuf[i] = 0.0;
        gtau[i] += v;
        long long n = n0;
        spill[i] += v;
        double d1 = v - mean[i];
        mean[i] += d1 / (double)(n + 1);
        sq[i] += d1 * (v - mean[i]);
        for (int lvl = 1; (n & 1) != 0 && lvl < M.nlev; ++lvl) {
            n >>= 1;
            const long lo = (long)(lvl - 1) * nc + i, hi = (long)lvl * nc + i;
            const double u = spill[lo] / 2.0;
            spill[lo] = 0.0;
            spill[hi] += u;
            const double e1 = u - mean[hi];
            mean[hi] += e1 / (double)(n + 1);
            sq[hi] += e1 * (u - mean[hi]);
        }
    }
    wsync();
}

// MeasDensityMatrix_beta_half (CTQMC.F90:3088-3151) with update_b2_states_eb
// (Trace.F90:6213-6469): ket from 0 to beta/2, bra from beta to beta/2
template <int NM>
DEV void meas_dm(Ctx &c, double *dm)
{
    const DevModel &M = c.M;
    const int outer = c.h.outer, no = c.sz[outer];
    if (c.h.nop != 0) {
        const double bh = M.beta / 2.0;
        const int nk = upper_bound_tau(c.tau, c.h.nop, bh);   // ops with tau <= beta/2
        double *K = c.K0, *Kw = c.K1;
        set_identity(K, no);
        int s = outer;
        double kl = 0.0;
        ket_walk_global<NM>(c, K, Kw, s, kl, no, 0.0, 0, nk, bh);
        double *Ks = c.scratch;
        copy_mat(Ks, K, c.sz[s] * no);
        double *B = c.K0, *Bw = c.K1;
        set_identity(B, no);
        int sb = outer;
        double bl = 0.0;
        bra_walk_global<NM>(c, B, Bw, sb, bl, no, M.beta, nk, c.h.nop, bh);
        const int n = c.sz[s];
        double *blk = dm + M.dmoff[s];
        const double fac = c.h.trace_sign * (-c.h.cfgsign) * exp(kl + bl - c.h.trace_log);
        for (int idx = lane_id(); idx < n * n; idx += WAVE) {
            const int s1 = idx % n, s2 = idx / n;
            double acc = 0.0;
            for (int i = 0; i < no; ++i) acc = fma(B[s1 * no + i], Ks[s2 * no + i], acc);
            blk[s1 + (long)s2 * n] += fac * acc;
        }
        wsync();
    } else {
        double *blk = dm + M.dmoff[outer];
        for (int i = lane_id(); i < no; i += WAVE) {
            const double v = c.part[2 * i + 1] * c.h.trace_sign * exp(c.part[2 * i] - c.h.trace_log);
            blk[i + (long)i * no] -= c.h.cfgsign * v;
        }
        wsync();
    }
}

// ---------------------------------------------------------------------------
// parameter derivatives (no counterpart in the reference)
// ---------------------------------------------------------------------------
// d ln|w| / d theta of a configuration, w = det(Delta) Tr[...]:
//  * d ln|det Delta| = sum_f tr(Delta_f^{-1} dDelta_f): each (creator,
//    annihilator) pair of a flavour times dF/dtheta(tau_a - tau_c), looked up
//    like get_HybrF (Trace.F90:1014-1033) on the derivative table;
//  * d ln|Tr| by forward tangents: each state matrix K carries dK_t beside it.
//    Over dt in block s (levels E_k relative to the block's lowest)
//      K <- P K,   dK_t <- P dK_t - (D_t o Phi) K,
//    P = diag(exp(-dt E_k)), D_t = V^T (dH/dtheta_t) V in the eigenbasis and
//    Phi_kl = (exp(-dt E_k) - exp(-dt E_l)) / (E_l - E_k) (dt exp(-dt E_k) for
//    equal levels); an operator block maps K and every dK_t alike, and one
//    power of two rescales them together.  The ket from 0 and the bra from
//    beta meet at beta/2 (as the density-matrix measurement), where the trace,
//    its tangents, the density-matrix sample and that sample's explicit
//    derivative are contracted.
constexpr int MAXTHETA = 4;

struct TanEnv {
    const DevModel *M;
    const int *sz, *off;
    const int4 *tr;
    const double *ev, *ops, *tau;
    const int *code;
    double *w;            // derivative work area of the chain
    int nm, nth;
};

DEV TanEnv tan_env(const Ctx &c)
{
    TanEnv e;
    e.M = &c.M, e.sz = c.sz, e.off = c.off, e.tr = c.tr, e.ev = c.ev, e.ops = c.opsp, e.tau = c.tau;
    e.code = c.code, e.w = reinterpret_cast<double *>(c.base + c.L.dwork), e.nm = c.nm, e.nth = c.M.ntheta;
    return e;
}

// divided difference of exp(-dt E) at two levels relative to the block's lowest
DEV double phi_div(double ek, double el, double dt)
{
    const double a = fmin(ek, el), y = fabs(ek - el) * dt;
    const double g = y > 0.0 ? -expm1(-y) / y : 1.0;
    return dt * exp(-a * dt) * g;
}

DEV int4 tan_trans(const TanEnv &e, int s, int code, bool adjoint)
{
    int ca = op_ca(code);
    if (adjoint) ca = 3 - ca;
    return e.tr[((s * e.M->nb + op_orb(code)) * 2 + op_spin(code)) * 2 + (ca - 1)];
}

// segment of length dt in block s: dKw_t = P dK_t - (D_t o Phi) K, then K <- P K
DEV void tan_prop(const TanEnv &e, int s, double dt, double *K, const double *dK, double *dKw, int no)
{
    const DevModel &M = *e.M;
    const int ns = e.sz[s];
    const long nm2 = (long)e.nm * e.nm;
    const double *E = e.ev + e.off[s];
    double *Phi = e.w, *pv = e.w + (4L + 3L * e.nth) * nm2;
    for (int idx = lane_id(); idx < ns * ns; idx += WAVE)
        Phi[idx] = phi_div(E[idx % ns] - E[0], E[idx / ns] - E[0], dt);
    for (int k = lane_id(); k < ns; k += WAVE) pv[k] = exp(-(E[k] - E[0]) * dt);
    wsync();
    for (int t = 0; t < e.nth; ++t) {
        const double *D = M.dblk + (long)t * M.dmsize + M.dmoff[s];
        const double *dk = dK + t * nm2;
        double *dw = dKw + t * nm2;
        for (int idx = lane_id(); idx < ns * no; idx += WAVE) {
            const int k = idx / no, i = idx - k * no;
            double acc = pv[k] * dk[idx];
            for (int l = 0; l < ns; ++l) acc = fma(-D[k + l * ns] * Phi[k + l * ns], K[l * no + i], acc);
            dw[idx] = acc;
        }
    }
    wsync();
    for (int idx = lane_id(); idx < ns * no; idx += WAVE) K[idx] *= pv[idx / no];
    wsync();
}

// operator block P (nt x ns, column-major): Kw = P K, dKw_t = P dK_t
DEV void tan_apply(const TanEnv &e, const double *P, int nt, int ns, int no, const double *K, double *Kw,
                   const double *dK, double *dKw)
{
    const long nm2 = (long)e.nm * e.nm;
    for (int t = -1; t < e.nth; ++t) {
        const double *src = t < 0 ? K : dK + t * nm2;
        double *dst = t < 0 ? Kw : dKw + t * nm2;
        for (int idx = lane_id(); idx < nt * no; idx += WAVE) {
            const int r = idx / no, i = idx - r * no;
            double acc = 0.0;
            for (int c = 0; c < ns; ++c) acc = fma(P[r + c * nt], src[c * no + i], acc);
            dst[idx] = acc;
        }
    }
    wsync();
}

// one power of two for K and its tangents when their largest entry leaves [2^-64, 2^64]
DEV void tan_renorm(const TanEnv &e, double *K, double *dK, int n, double &lg)
{
    const long nm2 = (long)e.nm * e.nm;
    double m = 0.0;
    for (int idx = lane_id(); idx < n; idx += WAVE) {
        m = fmax(m, fabs(K[idx]));
        for (int t = 0; t < e.nth; ++t) m = fmax(m, fabs(dK[t * nm2 + idx]));
    }
    const double mm = wave_max(m);
    if (mm > 0.0 && (mm < STATE_TINY || mm > STATE_HUGE)) {
        const int ex = ilogb(mm);
        for (int idx = lane_id(); idx < n; idx += WAVE) {
            K[idx] = ldexp(K[idx], -ex);
            for (int t = 0; t < e.nth; ++t) dK[t * nm2 + idx] = ldexp(dK[t * nm2 + idx], -ex);
        }
        lg += (double)ex * 0.69314718055994530942;
    }
    wsync();
}

// Walk over operators [j0, j1) of the configuration (BRA: descending, adjoint
// blocks) from tstart to tend with tangents; false on a quantum-number violation
template <bool BRA>
DEV bool tan_walk(const TanEnv &e, double *&K, double *&Kw, double *&dK, double *&dKw, int &s, double &lg, int no,
                  double tstart, int j0, int j1, double tend)
{
    const DevModel &M = *e.M;
    double t = tstart;
    for (int jj = 0; jj < j1 - j0; ++jj) {
        const int j = BRA ? j1 - 1 - jj : j0 + jj;
        const double tj = e.tau[j];
        const int4 tr = tan_trans(e, s, e.code[j], BRA);
        if (tr.x < 0) return false;
        const double dt = BRA ? t - tj : tj - t;
        tan_prop(e, s, dt, K, dK, dKw, no);
        lg += -(e.ev[e.off[s]] - M.egs) * dt;
        tan_apply(e, e.ops + tr.z, tr.y, e.sz[s], no, K, Kw, dKw, dK);
        double *x = K;
        K = Kw, Kw = x;
        tan_renorm(e, K, dK, tr.y * no, lg);
        s = tr.x;
        t = tj;
    }
    const double dt = BRA ? t - tend : tend - t;
    tan_prop(e, s, dt, K, dK, dKw, no);
    lg += -(e.ev[e.off[s]] - M.egs) * dt;
    double *x = dK;
    dK = dKw, dKw = x;
    tan_renorm(e, K, dK, e.sz[s] * no, lg);
    return true;
}

struct HalfTangents {
    const double *K, *dK, *B, *dB;   // ket and bra at beta/2 (rows of block smid x outer states) and tangents
    double kl, bl;                    // their log scales
    int smid, no;
    bool ok;
};

// ket from 0 and bra from beta to beta/2 with tangents
DEV HalfTangents half_tangents(const TanEnv &e, int nop, int outer)
{
    const DevModel &M = *e.M;
    const long nm2 = (long)e.nm * e.nm;
    double *A0 = e.w + nm2, *A1 = A0 + nm2, *A2 = A1 + nm2;
    double *T0 = A2 + nm2, *T1 = T0 + e.nth * nm2, *T2 = T1 + e.nth * nm2;
    HalfTangents r;
    const int no = e.sz[outer];
    r.no = no;
    const double bh = M.beta / 2.0;
    const int nk = upper_bound_tau(e.tau, nop, bh);
    double *K = A0, *Kw = A1, *dK = T0, *dKw = T1;
    for (int idx = lane_id(); idx < no * no; idx += WAVE) {
        K[idx] = (idx / no == idx % no) ? 1.0 : 0.0;
        for (int t = 0; t < e.nth; ++t) dK[t * nm2 + idx] = 0.0;
    }
    wsync();
    int s = outer;
    r.kl = 0.0;
    r.ok = tan_walk<false>(e, K, Kw, dK, dKw, s, r.kl, no, 0.0, 0, nk, bh);
    r.smid = s;
    const int nmid = e.sz[s];
    for (int idx = lane_id(); idx < nmid * no; idx += WAVE) {
        A2[idx] = K[idx];
        for (int t = 0; t < e.nth; ++t) T2[t * nm2 + idx] = dK[t * nm2 + idx];
    }
    wsync();
    double *B = A0, *Bw = A1, *dB = T0, *dBw = T1;
    for (int idx = lane_id(); idx < no * no; idx += WAVE) {
        B[idx] = (idx / no == idx % no) ? 1.0 : 0.0;
        for (int t = 0; t < e.nth; ++t) dB[t * nm2 + idx] = 0.0;
    }
    wsync();
    int sb = outer;
    r.bl = 0.0;
    r.ok = r.ok && tan_walk<true>(e, B, Bw, dB, dBw, sb, r.bl, no, M.beta, nk, nop, bh) && sb == s;
    r.K = A2, r.dK = T2, r.B = B, r.dB = dB;
    return r;
}

// d ln|Tr| / d theta from the beta/2 contraction
DEV void trace_tangent(const TanEnv &e, const HalfTangents &h, double *dlt)
{
    const long nm2 = (long)e.nm * e.nm;
    const int n = e.sz[h.smid] * h.no;
    double tr = 0.0;
    for (int idx = lane_id(); idx < n; idx += WAVE) tr += h.B[idx] * h.K[idx];
    tr = wave_sum(tr);
    for (int t = 0; t < e.nth; ++t) {
        double d = 0.0;
        for (int idx = lane_id(); idx < n; idx += WAVE)
            d += h.dB[t * nm2 + idx] * h.K[idx] + h.B[idx] * h.dK[t * nm2 + idx];
        dlt[t] = wave_sum(d) / tr;
    }
}

// What a derivative pass needs, by value (the pass is out of line so that it
// does not add to the sweep's registers; a reference to the chain context
// would live in scratch).
struct ScoreArgs {
    TanEnv e;
    const double *fc, *fa, *minv;
    const int *fcs, *fas;
    int kf[MAXF];
    int kcap, nop, outer;
    double cfgsign, trace_sign, trace_log;
    double *scsum, *scg, *scdm, *scdmx;   // score sums of the chain
    double *gbuf, *gtau;                  // this measurement's G(tau) bins, the G(tau) sums
};

DEV ScoreArgs score_args(const Ctx &c, double *gbuf, double *gtau)
{
    ScoreArgs a;
    a.e = tan_env(c);
    a.fc = c.fc, a.fa = c.fa, a.minv = c.minv, a.fcs = c.fcs, a.fas = c.fas;
    for (int f = 0; f < MAXF; ++f) a.kf[f] = c.h.kf[f];
    a.kcap = c.L.kcap, a.nop = c.h.nop, a.outer = c.h.outer;
    a.cfgsign = c.h.cfgsign, a.trace_sign = c.h.trace_sign, a.trace_log = c.h.trace_log;
    a.scsum = reinterpret_cast<double *>(c.base + c.L.scsum);
    a.scg = reinterpret_cast<double *>(c.base + c.L.scg);
    a.scdm = reinterpret_cast<double *>(c.base + c.L.scdm);
    a.scdmx = reinterpret_cast<double *>(c.base + c.L.scdmx);
    a.gbuf = gbuf, a.gtau = gtau;
    return a;
}

// d ln|det| / d theta = sum_f tr(Delta_f^{-1} dDelta_f)
DEV void det_tangent(const ScoreArgs &a, double *dld)
{
    const DevModel &M = *a.e.M;
    const int ld = a.kcap;
    double acc[MAXTHETA] = {0.0, 0.0, 0.0, 0.0};
    for (int f = 0; f < M.nf; ++f) {
        const int k = a.kf[f];
        const double *ctl = a.fc + (long)f * ld, *atl = a.fa + (long)f * ld;
        const int *csl = a.fcs + (long)f * ld, *asl = a.fas + (long)f * ld;
        const double *Mf = a.minv + (long)f * ld * ld;
        for (int p = lane_id(); p < k * k; p += WAVE) {
            const int ai = p / k, ci = p - ai * k;
            const double m = Mf[csl[ci] + (long)asl[ai] * ld];
            double tau = atl[ai] - ctl[ci], sig = 1.0;
            if (tau < 0.0) {
                tau += M.beta;
                sig = -1.0;
            }
            const double x = tau / M.beta * (double)(M.nftau - 1);
            const int i = (int)x;
            for (int t = 0; t < M.ntheta; ++t) {
                const double *F = M.dftau + ((long)t * M.nf + f) * M.nftau;
                acc[t] += m * sig * (F[i] + (x - (double)i) * (F[i + 1] - F[i]));
            }
        }
    }
    for (int t = 0; t < M.ntheta; ++t) dld[t] = wave_sum(acc[t]);
}

// One derivative pass over the configuration.  out != null: d ln|det| /
// d theta_t -> out[2 t], d ln|Tr| / d theta_t -> out[2 t + 1].  accumulate:
// the score-function sums of a measurement (s_t = d ln|w| / d theta_t):
// Sum sign s_t, Sum s_t x this measurement's G(tau) bins (gbuf, which is then
// added to the G(tau) sums unless the binning accumulator does that), Sum
// s_t x density-matrix sample and Sum d(sample)/dtheta_t at fixed
// configuration (the beta/2 sample depends on H explicitly).
__device__ __attribute__((noinline)) void score_pass(const ScoreArgs a, double *out, bool accumulate)
{
    const DevModel &M = *a.e.M;
    const TanEnv &e = a.e;
    const long nm2 = (long)e.nm * e.nm;
    double dld[MAXTHETA], dlt[MAXTHETA], sc[MAXTHETA];
    det_tangent(a, dld);
    const HalfTangents h = half_tangents(e, a.nop, a.outer);
    trace_tangent(e, h, dlt);
    if (out && lane_id() == 0)
        for (int t = 0; t < M.ntheta; ++t) {
            out[2 * t] = dld[t];
            out[2 * t + 1] = h.ok ? dlt[t] : __builtin_nan("");
        }
    if (!accumulate) return;
    for (int t = 0; t < M.ntheta; ++t) sc[t] = dld[t] + dlt[t];
    if (lane_id() == 0)
        for (int t = 0; t < M.ntheta; ++t) a.scsum[t] += a.cfgsign * sc[t];
    if (M.meas_dm) {
        const int n = e.sz[h.smid], no = h.no;
        const long off = M.dmoff[h.smid];
        const double fac = a.trace_sign * (-a.cfgsign) * exp(h.kl + h.bl - a.trace_log);
        for (int idx = lane_id(); idx < n * n; idx += WAVE) {
            const int s1 = idx % n, s2 = idx / n;
            double smp = 0.0;
            for (int i = 0; i < no; ++i) smp = fma(h.B[s1 * no + i], h.K[s2 * no + i], smp);
            smp *= fac;
            for (int t = 0; t < M.ntheta; ++t) {
                double x = 0.0;
                for (int i = 0; i < no; ++i)
                    x += h.dB[t * nm2 + s1 * no + i] * h.K[s2 * no + i] + h.B[s1 * no + i] * h.dK[t * nm2 + s2 * no + i];
                a.scdm[t * M.dmsize + off + idx] += sc[t] * smp;
                a.scdmx[t * M.dmsize + off + idx] += fac * x - smp * dlt[t];
            }
        }
    }
    // this measurement's G(tau) bins: written by atomics at L2, read past the vector L1
    wsync();
    const long nc = (long)M.nf * M.ntau;
    for (long i = lane_id(); i < nc; i += WAVE) {
        const double v = __hip_atomic_load(a.gbuf + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        for (int t = 0; t < M.ntheta; ++t) a.scg[t * nc + i] += sc[t] * v;
        if (M.accum != 2) {
            a.gtau[i] += v;
            a.gbuf[i] = 0.0;
        }
    }
    wsync();
}

// MeasSign, MeasOuterHistograms, MeasExpOrder (CTQMC.F90:3264-3350): visits
// of the outer superstate and of each of its states, superstate sign, state
// signs (trace sign x part-trace sign x configuration sign), trace
// contributions exp(log part trace) per state and summed per superstate, and
// the order histogram
DEV void meas_misc(Ctx &c, double *histo, double *outer_hist)
{
    const DevModel &M = c.M;
    const int outer = c.h.outer, no = c.sz[outer], off = c.off[outer];
    double *ohst = reinterpret_cast<double *>(c.base + c.L.ohst);
    double *shss = reinterpret_cast<double *>(c.base + c.L.shss);
    double *shst = reinterpret_cast<double *>(c.base + c.L.shst);
    double *tcss = reinterpret_cast<double *>(c.base + c.L.tcss);
    double *tcst = reinterpret_cast<double *>(c.base + c.L.tcst);
    c.h.meas_sign += c.h.cfgsign;
    double tc = 0.0;
    for (int i = lane_id(); i < no; i += WAVE) {
        const double e = exp(c.part[2 * i]);
        ohst[off + i] += 1.0;
        shst[off + i] += c.h.trace_sign * c.part[2 * i + 1] * c.h.cfgsign;
        tcst[off + i] += e;
        tc += e;
    }
    tc = wave_sum(tc);
    if (lane_id() == 0) {
        outer_hist[outer] += 1.0;
        shss[outer] += c.h.cfgsign;
        tcss[outer] += tc;
        for (int f = 0; f < M.nf; ++f) {
            int k = c.h.kf[f];
            if (k > M.maxhisto) k = M.maxhisto;
            histo[(long)f * (M.maxhisto + 1) + k] += 1.0;
        }
    }
    wsync();
}


// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
// Map the chain block and carve the dynamic LDS; copies header, RNG state and
// the model tables into LDS.
template <int NM>
DEV Ctx make_ctx(const DevModel &M, const ChainLayout &L, char *base, int chain, char *lds)
{
    ChainHdr &hd = *reinterpret_cast<ChainHdr *>(lds + L.l_hdr);
    Ctx c(M, L, base, hd);
    c.nm = NM;
    c.tau = reinterpret_cast<double *>(base + L.tau);
    c.code = reinterpret_cast<int *>(base + L.code);
    c.fc = reinterpret_cast<double *>(base + L.fc);
    c.fa = reinterpret_cast<double *>(base + L.fa);
    c.fcs = reinterpret_cast<int *>(base + L.fcs);
    c.fas = reinterpret_cast<int *>(base + L.fas);
    c.minv = reinterpret_cast<double *>(base + L.minv);
    c.scratch = reinterpret_cast<double *>(base + L.scratch);
    c.kst = reinterpret_cast<double *>(base + L.kst);
    c.bnd = reinterpret_cast<double *>(base + L.bnd);
    c.bndlog = reinterpret_cast<double *>(base + L.bndlog);
    c.bndsst = reinterpret_cast<int *>(base + L.bndsst);
    c.part = reinterpret_cast<double *>(base + L.part);
    uint32_t *mt = reinterpret_cast<uint32_t *>(lds + L.l_mt);
    int *sz = reinterpret_cast<int *>(lds + L.l_sz);
    int *off = reinterpret_cast<int *>(lds + L.l_off);
    int *cn = reinterpret_cast<int *>(lds + L.l_cn);
    int *oo = reinterpret_cast<int *>(lds + L.l_oo);
    double *ev = reinterpret_cast<double *>(lds + L.l_ev);
    c.wtau = reinterpret_cast<double *>(lds + L.l_wtau);
    c.wcode = reinterpret_cast<int *>(lds + L.l_wcode);
    c.pc = reinterpret_cast<int *>(lds + L.l_pc);
    c.ttau = reinterpret_cast<double *>(lds + L.l_ttau);
    c.tcode = reinterpret_cast<int *>(lds + L.l_tcode);
    c.dvec = reinterpret_cast<double *>(lds + L.l_dvec);
    c.Q = reinterpret_cast<double *>(lds + L.l_q);
    c.R = reinterpret_cast<double *>(lds + L.l_r);
    c.colb = reinterpret_cast<double *>(lds + L.l_col);
    c.rowb = reinterpret_cast<double *>(lds + L.l_row);
    c.pl = reinterpret_cast<double *>(lds + L.l_pl);
    if constexpr (NM <= 4) {
        c.rk = reinterpret_cast<double *>(lds + L.l_rk);
        c.lb = reinterpret_cast<double *>(lds + L.l_lb);
        c.rkx = reinterpret_cast<double *>(lds + L.l_rkx);
        c.lbx = reinterpret_cast<double *>(lds + L.l_lbx);
        c.rks = reinterpret_cast<int *>(lds + L.l_rks);
        c.lbs = reinterpret_cast<int *>(lds + L.l_lbs);
        c.qa = reinterpret_cast<double *>(lds + L.l_qa);
        c.qlc = reinterpret_cast<double *>(lds + L.l_qlc);
        c.qrec = reinterpret_cast<int4 *>(lds + L.l_qrec);
        c.qsst = reinterpret_cast<int *>(lds + L.l_qsst);
    }
    {
        double *lc = reinterpret_cast<double *>(lds + L.l_legc);
        for (int i = lane_id(); i < 2 * M.nleg; i += WAVE) lc[i] = M.legc[i];
        c.legc = lc;
    }
    // work matrices: LDS for blocks up to 16 states (decided at compile time so
    // that their accesses are LDS instructions, not generic ones), HBM above
    if constexpr (NM <= 16) {
        c.K0 = reinterpret_cast<double *>(lds + L.l_k);
    } else {
        c.K0 = reinterpret_cast<double *>(base + L.kwork);
    }
    c.K1 = c.K0 + (long)NM * NM;
    c.strip = L.l_strip >= 0 ? reinterpret_cast<double *>(lds + L.l_strip) : nullptr;
    // header, RNG state and model tables into LDS
    c.ghdr = reinterpret_cast<ChainHdr *>(base - (long)chain * L.stride + L.hdrs) + chain;
    const int *gh = reinterpret_cast<const int *>(c.ghdr);
    int *lh = reinterpret_cast<int *>(&hd);
    for (int i = lane_id(); i < (int)(sizeof(ChainHdr) / 4); i += WAVE) lh[i] = gh[i];
    const uint32_t *gmt = reinterpret_cast<const uint32_t *>(base + L.mt);
    for (int i = lane_id(); i < 624; i += WAVE) mt[i] = gmt[i];
    for (int i = lane_id(); i < M.nsst; i += WAVE) {
        sz[i] = M.sizes[i];
        off[i] = M.offsets[i];
    }
    const int ncn = M.nsst * M.nb * 4;
    if constexpr (NM <= 16) {
        for (int i = lane_id(); i < ncn; i += WAVE) {
            cn[i] = M.connect[i];
            oo[i] = (int)M.opoff[i];
        }
    }
    if constexpr (NM <= 16) {
        for (int i = lane_id(); i < M.nstates; i += WAVE) ev[i] = M.evals[i];
        int4 *trt = reinterpret_cast<int4 *>(lds + L.l_tr);
        for (int i = lane_id(); i < ncn; i += WAVE) trt[i] = M.trans[i];
        c.tr = trt;
    } else {
        c.tr = M.trans;
    }
    // operator blocks: always LDS for blocks up to 4 states (known at compile
    // time, so the block reads are LDS instructions); LDS or HBM above
    if constexpr (NM <= 4) {
        double *lops = reinterpret_cast<double *>(lds + L.l_ops);
        for (long i = lane_id(); i < M.nops; i += WAVE) lops[i] = M.ops[i];
        c.opsp = lops;
    } else if (L.l_ops >= 0) {
        double *lops = reinterpret_cast<double *>(lds + L.l_ops);
        for (long i = lane_id(); i < M.nops; i += WAVE) lops[i] = M.ops[i];
        c.opsp = lops;
    } else {
        c.opsp = M.ops;
    }
    c.sz = sz, c.off = off;
    if constexpr (NM <= 16) c.cn = cn, c.oo = oo;
    else c.cn = M.connect, c.oo = M.opoff32;
    if constexpr (NM <= 16) c.ev = ev;
    else c.ev = M.evals;
    c.rng.mt = mt;
    c.rng.idx = &hd.mti;
    hd.wvalid = 0;   // the LDS window list does not survive a launch
    wsync();
    return c;
}

DEV void save_ctx(Ctx &c)
{
    c.h.alg_flops += c.wflops;
    c.h.alg_bytes += c.wbytes;
    uint32_t *gmt = reinterpret_cast<uint32_t *>(c.base + c.L.mt);
    for (int i = lane_id(); i < 624; i += WAVE) gmt[i] = c.rng.mt[i];
    int *gh = reinterpret_cast<int *>(c.ghdr);
    const int *lh = reinterpret_cast<const int *>(&c.h);
    wsync();
    for (int i = lane_id(); i < (int)(sizeof(ChainHdr) / 4); i += WAVE) gh[i] = lh[i];
    wsync();
}

// Sweep kernel: each block advances chain blockIdx.x until it has done
// `target` moves in the current warmup / measure call (header field pos); if
// ncorr > 0 it measures after every ncorr-th move of the call.  A chain stops
// early, state intact, when the next move might not fit its memory
// (HALT_CAP) or when the host has raised the abort word (HALT_ABORT, polled
// at measurement boundaries and every 128 warm-up moves; the reference checks
// its signal flag before every move, CTQMC.F90:8212, 8436).  rec != nullptr
// records every step of chain `only` at rec[pos - rec_base] (parity mode;
// other chains return at once).
// SCORE: the measurements add the score-function sums (a separate
// instantiation, so that the derivative pass costs the plain sweep nothing)
template <int NM, bool SCORE>
__global__ void __launch_bounds__(WAVE) sweep_kernel(DevModel M, ChainLayout L, char *chains, long long target,
                                                     long long ncorr, int only, w2d_move_record *rec,
                                                     long long rec_base, const int *abort_word, int3 calib,
                                                     const int *order)
{
    // order: the chain each workgroup runs (longest first, so that the
    // launch's tail is short), or null for workgroup = chain.  The order is
    // rewritten by order_kernel before every launch: it is read with a
    // device-coherent vector load, never through the scalar cache (which is
    // not coherent with the vector stores of an earlier kernel)
    const int chain = order ? __builtin_amdgcn_readfirstlane(__hip_atomic_load(
                                  order + blockIdx.x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
                            : (int)blockIdx.x;
    if (only >= 0 && chain != only) return;
    extern __shared__ __align__(16) char lds[];
    char *base = chains + (long)chain * L.stride;
    Ctx c = make_ctx<NM>(M, L, base, chain, lds);
    double *gtau = reinterpret_cast<double *>(base + L.gtau);
    double *gleg = reinterpret_cast<double *>(base + L.gleg);
    double *dm = reinterpret_cast<double *>(base + L.dm);
    double *histo = reinterpret_cast<double *>(base + L.histo);
    double *ohist = reinterpret_cast<double *>(base + L.outer);
    double *gbuf = reinterpret_cast<double *>(base + L.gbuf);
    double *gacc = reinterpret_cast<double *>(base + L.gacc);
    while (c.h.pos < target && c.h.err == 0 && c.h.halt == HALT_NONE) {
        if (!room_for_move(c)) {
            c.h.halt = HALT_CAP;
            break;
        }
        int kind, out;
        one_move<NM>(c, kind, out);
        const long long m = c.h.pos++;
        // ctqmc_calibrate (CTQMC.F90:8128-8175): every calib.y moves, phase 1
        // (calib.x = 1) ends once the distance list is full, phase 2
        // (calib.x = 2) once every flavour has calib.z accepted pairs
        if (calib.x && (m + 1) % calib.y == 0) {
            bool done = calib.x == 1 && c.h.apt_index > c.h.apt_n;
            if (calib.x == 2) {
                const double *ap = reinterpret_cast<const double *>(base + L.accpair);
                int short_f = 0;
                for (int f = 0; f < M.nf; ++f) {
                    double n = 0.0;
                    for (int b = lane_id(); b < c.h.naccpair; b += WAVE) n += ap[(long)f * L.naccpair_cap + b];
                    short_f += wave_sum(n) < (double)calib.z;
                }
                done = short_f == 0;
            }
            if (done) c.h.halt = HALT_DONE;
        }
        if (rec && lane_id() == 0) {
            w2d_move_record &r = rec[m - rec_base];
            r.kind = kind;
            r.outcome = out;
            r.noper = c.h.nop;
            r.outer_sst = c.h.outer;
            r.window = c.h.wpos;
            r.trace_log = c.h.trace_log;
            r.trace_sign = c.h.trace_sign;
            r.det_log = c.h.det_log;
            r.det_sign = c.h.det_sign;
            r.cfgsign = c.h.cfgsign;
        }
        if (ncorr <= 0) {
            if ((m + 1) % 128 == 0 && __hip_atomic_load(abort_word, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM))
                c.h.halt = HALT_ABORT;
        } else if ((m + 1) % ncorr == 0) {
            c.h.cnt_meas++;
            meas_misc(c, histo, ohist);
            if (M.meas_dm) meas_dm<NM>(c, dm);
            GtauArgs ga;
            ga.nf = M.nf, ga.ntau = M.ntau, ga.nleg = M.nleg, ga.kcap = L.kcap;
            ga.beta = M.beta, ga.cfgsign = c.h.cfgsign, ga.legc = c.legc;
            ga.kf = c.h.kf, ga.fc = c.fc, ga.fa = c.fa, ga.minv = c.minv, ga.fcs = c.fcs, ga.fas = c.fas;
            constexpr bool score = SCORE;
            ga.gtau = (M.accum == 2 || score) ? gbuf : gtau, ga.gleg = gleg;
            const long long mi = c.h.cnt_meas - 1;   // g_inmeas - 1
            const bool st = mi < M.nmstep;
            ga.mstep = (st && M.step_mean) ? reinterpret_cast<double *>(base + L.gmstep) + mi * M.nf : nullptr;
            ga.dstep = (st && M.step_mid) ? reinterpret_cast<double *>(base + L.gdstep) + mi * M.nf : nullptr;
            // sign_step is set when the configuration has a pair (CTQMC.F90:2402)
            if (st && M.step_sign && c.h.nop > 0 && lane_id() == 0)
                reinterpret_cast<double *>(base + L.sstep)[mi] = c.h.cfgsign;
            const double2 wk = NM == 1 ? meas_gtau_narrow(ga) : meas_gtau(ga);
            c.wflops += wk.x;
            c.wbytes += wk.y;
            if constexpr (SCORE) score_pass(score_args(c, gbuf, gtau), nullptr, true);
            if (M.accum == 2) autocorr_add(M, gbuf, gacc, gtau, c.h.cnt_meas - 1);
            if (__hip_atomic_load(abort_word, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM)) c.h.halt = HALT_ABORT;
        }
    }
    save_ctx(c);
}

// d ln|det| / d theta and d ln|Tr| / d theta of one chain's configuration:
// out[2 t], out[2 t + 1]
template <int NM>
__global__ void __launch_bounds__(WAVE) dlogw_kernel(DevModel M, ChainLayout L, char *chains, int chain,
                                                       double *out)
{
    extern __shared__ __align__(16) char lds[];
    char *base = chains + (long)chain * L.stride;
    Ctx c = make_ctx<NM>(M, L, base, chain, lds);
    score_pass(score_args(c, nullptr, nullptr), out, false);
}

// Copy every chain from layout A to layout B (larger capacity): header, random
// state, configuration, flavour lists, inverse blocks re-strided to B's
// leading dimension, trace caches and accumulators.  Work buffers are not
// copied (they hold nothing between moves).  Physical row / column slots stay
// valid since they are below A's capacity.
DEV void copy_words(char *dst, const char *src, long bytes)
{
    int *d = reinterpret_cast<int *>(dst);
    const int *s = reinterpret_cast<const int *>(src);
    for (long i = lane_id(); i < bytes / 4; i += WAVE) d[i] = s[i];
}

__global__ void __launch_bounds__(WAVE) migrate_kernel(ChainLayout A, ChainLayout B, const char *src, char *dst,
                                                       int nf, int nm, long accbytes)
{
    const int chain = blockIdx.x;
    const char *s = src + (long)chain * A.stride;
    char *d = dst + (long)chain * B.stride;
    copy_words(dst + B.hdrs + (long)chain * sizeof(ChainHdr), src + A.hdrs + (long)chain * sizeof(ChainHdr),
               sizeof(ChainHdr));
    copy_words(d + B.mt, s + A.mt, 624 * 4);
    copy_words(d + B.tau, s + A.tau, (long)A.cap * 8);
    copy_words(d + B.code, s + A.code, (long)A.cap * 4);
    for (int f = 0; f < nf; ++f) {
        copy_words(d + B.fc + (long)f * B.kcap * 8, s + A.fc + (long)f * A.kcap * 8, (long)A.kcap * 8);
        copy_words(d + B.fa + (long)f * B.kcap * 8, s + A.fa + (long)f * A.kcap * 8, (long)A.kcap * 8);
        copy_words(d + B.fcs + (long)f * B.kcap * 4, s + A.fcs + (long)f * A.kcap * 4, (long)A.kcap * 4);
        copy_words(d + B.fas + (long)f * B.kcap * 4, s + A.fas + (long)f * A.kcap * 4, (long)A.kcap * 4);
        const double *ms = reinterpret_cast<const double *>(s + A.minv) + (long)f * A.kcap * A.kcap;
        double *md = reinterpret_cast<double *>(d + B.minv) + (long)f * B.kcap * B.kcap;
        for (long e = lane_id(); e < (long)A.kcap * A.kcap; e += WAVE) {
            const long i = e % A.kcap, j = e / A.kcap;
            md[i + j * B.kcap] = ms[e];
        }
    }
    // boundary caches: as many as both layouts hold (a new window count
    // invalidates them anyway: the host sets cache_pos = -1)
    const long nbd = min(A.nbound, B.nbound);
    copy_words(d + B.kst, s + A.kst, (long)nm * nm * 8);
    copy_words(d + B.bnd, s + A.bnd, nbd * nm * nm * 8);
    copy_words(d + B.bndlog, s + A.bndlog, nbd * 8);
    copy_words(d + B.bndsst, s + A.bndsst, nbd * 4);
    copy_words(d + B.part, s + A.part, 2L * nm * 8 + 16);
    copy_words(d + B.gtau, s + A.gtau, accbytes);
}

// Accumulators (the contiguous tail [gtau, stride) of each chain block, as
// doubles) summed over contiguous groups of chains: group g holds chains
// [g*nc/ng, (g+1)*nc/ng) that have not stopped on an error (header err == 0).
// Grid: x over elements, y over groups; consecutive threads read consecutive
// doubles of one chain.
// Launch order of the chains by expected cost: a counting sort of the chains
// by expansion order (nop >> shift, clamped to 1023), largest first, in one
// workgroup (LDS counters; order within a bucket is arbitrary: chains are
// independent, so their results do not depend on it).  A kernel rather than a
// host sort, so that it can sit in a captured graph between sweep launches.
constexpr int ORDER_BUCKETS = 1024;
__global__ void __launch_bounds__(1024) order_kernel(const ChainHdr *hdrs, int nc, int shift, int *order)
{
    __shared__ int cnt[ORDER_BUCKETS];
    const int t = threadIdx.x;
    cnt[t] = 0;
    __syncthreads();
    auto key = [&](int c) { return ORDER_BUCKETS - 1 - min(hdrs[c].nop >> shift, ORDER_BUCKETS - 1); };
    for (int c = t; c < nc; c += 1024) atomicAdd(&cnt[key(c)], 1);
    __syncthreads();
    if (t == 0) {   // exclusive scan (1024 entries)
        int run = 0;
        for (int b = 0; b < ORDER_BUCKETS; ++b) {
            const int x = cnt[b];
            cnt[b] = run;
            run += x;
        }
    }
    __syncthreads();
    for (int c = t; c < nc; c += 1024) order[atomicAdd(&cnt[key(c)], 1)] = c;
}

// MeasGiw (measure_giw_nfft, CTQMC.F90:1015-1059; the sum is the direct one of
// the NFFT fallback, Nfft_z.F90:64-90).  After a measurement, for every
// flavour f and n = 0..niw-1:
//   giw[chain][f][n] += cfgsign * sum_{a,c} M[c,a] exp(i w_n (tau_a - tau_c)),
// w_n = (2n+1) pi / beta, M the inverse hybridisation block (rows creators,
// columns annihilators, by physical slot), with the sign convention of the
// lookup path (CTQMC.F90:2403-2417 and :8816); negative frequencies are the
// complex conjugates.  The kernel follows every sweep launch of whole
// measurement intervals (the chain's state at the end of such a launch is the
// state it measured) and acts on the chains whose measurement count moved
// since its last visit: seen[chain][f] = {count at the last visit, G(i w)
// measurements made}; the host checks that the second equals the count.
// One workgroup per (chain, flavour).  A thread takes GIW_R consecutive
// frequencies; creators go in tiles of GIW_TC whose phases stay in registers,
// annihilators in tiles of GIW_TA with the M tile in LDS; the phase of the
// next frequency is one complex product with exp(2 pi i tau / beta).
constexpr int GIW_THREADS = 256, GIW_R = 4, GIW_TC = 8, GIW_TA = 64;

__global__ void __launch_bounds__(GIW_THREADS) giw_kernel(ChainLayout L, const char *chains, int nf, int niw,
                                                         double beta, double *giw, long long *seen)
{
    __shared__ double s_m[GIW_TA][GIW_TC];
    __shared__ double s_ct[GIW_TC], s_at[GIW_TA];
    __shared__ double2 s_cs[GIW_TC], s_as[GIW_TA];
    __shared__ int s_bad;
    const int chain = blockIdx.x, f = blockIdx.y, tid = threadIdx.x, nt = blockDim.x;
    // header fields written by the sweep kernels: device-coherent loads
    const ChainHdr *hp = reinterpret_cast<const ChainHdr *>(chains + L.hdrs) + chain;
    const long long cnt = __hip_atomic_load(&hp->cnt_meas, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const int err = __hip_atomic_load(&hp->err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    long long *sv = seen + 2 * ((long)chain * nf + f);
    const long long last = sv[0];
    if (err != 0 || cnt == last) return;   // uniform over the workgroup
    const int k = __hip_atomic_load(&hp->kf[f], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const double sg = __longlong_as_double(__hip_atomic_load(reinterpret_cast<const long long *>(&hp->cfgsign),
                                                             __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    if (tid == 0) s_bad = (k < 0 || k > L.kcap) ? 1 : 0;
    __syncthreads();
    const int kk = s_bad ? 0 : k;
    const char *base = chains + (long)chain * L.stride;
    const double *tcl = reinterpret_cast<const double *>(base + L.fc) + (long)f * L.kcap;
    const double *tal = reinterpret_cast<const double *>(base + L.fa) + (long)f * L.kcap;
    const int *csl = reinterpret_cast<const int *>(base + L.fcs) + (long)f * L.kcap;
    const int *asl = reinterpret_cast<const int *>(base + L.fas) + (long)f * L.kcap;
    const double *Mf = reinterpret_cast<const double *>(base + L.minv) + (long)f * L.kcap * L.kcap;
    double *out = giw + ((long)chain * nf + f) * (long)niw * 2;
    const double ib = 1.0 / beta;
    for (int n0 = 0; n0 < niw; n0 += nt * GIW_R) {
        const int nb = n0 + tid * GIW_R;            // first frequency of this thread
        const double xb = 2.0 * nb + 1.0;           // w_nb tau / pi = xb * tau / beta
        double gr[GIW_R], gi[GIW_R];
#pragma unroll
        for (int r = 0; r < GIW_R; ++r) gr[r] = gi[r] = 0.0;
        for (int c0 = 0; c0 < kk; c0 += GIW_TC) {
            const int nc = min(GIW_TC, kk - c0);
            __syncthreads();
            if (tid < GIW_TC) {
                const double t = tid < nc ? tcl[c0 + tid] * ib : 0.0;
                double s, c;
                sincospi(2.0 * t, &s, &c);
                s_ct[tid] = t;
                s_cs[tid] = make_double2(c, s);
            }
            __syncthreads();
            // exp(-i w_n tau_c) for the thread's frequencies; zero past the tile
            double cr[GIW_R][GIW_TC], ci[GIW_R][GIW_TC];
#pragma unroll
            for (int j = 0; j < GIW_TC; ++j) {
                double s, c;
                sincospi(xb * s_ct[j], &s, &c);
                const double2 st = s_cs[j];
                const double on = j < nc ? 1.0 : 0.0;
                double pr = c * on, pi = -s * on;
#pragma unroll
                for (int r = 0; r < GIW_R; ++r) {
                    cr[r][j] = pr;
                    ci[r][j] = pi;
                    const double t = pr * st.x + pi * st.y;
                    pi = pi * st.x - pr * st.y;
                    pr = t;
                }
            }
            for (int a0 = 0; a0 < kk; a0 += GIW_TA) {
                const int na = min(GIW_TA, kk - a0);
                __syncthreads();
                for (int e = tid; e < GIW_TA * GIW_TC; e += nt) {
                    const int ai = e / GIW_TC, cj = e - ai * GIW_TC;
                    double v = 0.0;
                    if (ai < na && cj < nc) {
                        const int rs = csl[c0 + cj], cs = asl[a0 + ai];
                        if (rs >= 0 && rs < L.kcap && cs >= 0 && cs < L.kcap) v = Mf[rs + (long)cs * L.kcap];
                        else s_bad = 1;
                    }
                    s_m[ai][cj] = v;
                }
                for (int ai = tid; ai < GIW_TA; ai += nt) {
                    const double t = ai < na ? tal[a0 + ai] * ib : 0.0;
                    double s, c;
                    sincospi(2.0 * t, &s, &c);
                    s_at[ai] = t;
                    s_as[ai] = make_double2(c, s);
                }
                __syncthreads();
                for (int ai = 0; ai < na; ++ai) {
                    double ur[GIW_R], ui[GIW_R];
#pragma unroll
                    for (int r = 0; r < GIW_R; ++r) ur[r] = ui[r] = 0.0;
#pragma unroll
                    for (int j = 0; j < GIW_TC; ++j) {
                        const double m = s_m[ai][j];
#pragma unroll
                        for (int r = 0; r < GIW_R; ++r) {
                            ur[r] = fma(m, cr[r][j], ur[r]);
                            ui[r] = fma(m, ci[r][j], ui[r]);
                        }
                    }
                    double s, c;
                    sincospi(xb * s_at[ai], &s, &c);
                    const double2 st = s_as[ai];
#pragma unroll
                    for (int r = 0; r < GIW_R; ++r) {
                        gr[r] += c * ur[r] - s * ui[r];
                        gi[r] += c * ui[r] + s * ur[r];
                        const double t = c * st.x - s * st.y;
                        s = c * st.y + s * st.x;
                        c = t;
                    }
                }
            }
        }
#pragma unroll
        for (int r = 0; r < GIW_R; ++r) {
            const int n = nb + r;
            if (n < niw) {
                out[2 * n] += sg * gr[r];
                out[2 * n + 1] += sg * gi[r];
            }
        }
    }
    __syncthreads();
    if (tid == 0) {
        sv[0] = cnt;
        if (!s_bad && cnt == last + 1) sv[1] += 1;
    }
}

__global__ void __launch_bounds__(256) group_sum_kernel(const char *chains, long stride, long acc_off, long nel,
                                                        int nc, int ng, const ChainHdr *hdrs, double *out)
{
    const long e = (long)blockIdx.x * 256 + threadIdx.x;
    const int g = blockIdx.y;
    if (e >= nel) return;
    const int c0 = (int)((long)g * nc / ng), c1 = (int)((long)(g + 1) * nc / ng);
    double acc = 0.0;
    for (int c = c0; c < c1; ++c)
        // the error flag is written by the sweep kernels: a device-coherent
        // load (a uniform address would otherwise go through the scalar cache)
        if (__hip_atomic_load(&hdrs[c].err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0)
            acc += reinterpret_cast<const double *>(chains + (long)c * stride + acc_off)[e];
    out[(long)g * nel + e] = acc;
}

// The header scalars of the same groups: sum of signs, measurements,
// accepted / tried moves of each kind, attempted and QN-rejected moves,
// algorithmic flops and bytes (scal[g][14]), and the chains summed (used[g]).
// One thread per group.
__global__ void __launch_bounds__(64) group_scalars_kernel(const ChainHdr *hdrs, int nc, int ng, double *scal,
                                                           int *used)
{
    const int g = blockIdx.x * 64 + threadIdx.x;
    if (g >= ng) return;
    const int c0 = (int)((long)g * nc / ng), c1 = (int)((long)(g + 1) * nc / ng);
    double S[14];
    for (int i = 0; i < 14; ++i) S[i] = 0.0;
    int n = 0;
    for (int c = c0; c < c1; ++c) {
        const ChainHdr &x = hdrs[c];
        if (x.err != 0) continue;
        S[0] += x.meas_sign, S[1] += (double)x.cnt_meas;
        S[2] += (double)x.acc_add, S[3] += (double)x.try_add;
        S[4] += (double)x.acc_rem, S[5] += (double)x.try_rem;
        S[6] += (double)x.acc_glob, S[7] += (double)x.try_glob;
        S[8] += (double)x.acc_shift, S[9] += (double)x.try_shift;
        S[10] += (double)x.moves, S[11] += (double)x.qn_rej;
        S[12] += x.alg_flops, S[13] += x.alg_bytes;
        ++n;
    }
    for (int i = 0; i < 14; ++i) scal[(long)g * 14 + i] = S[i];
    used[g] = n;
}

// Initial state of every chain: empty configuration, initial outer superstate
// drawn proportional to the empty traces (init_Trace, Trace.F90:640-670).
template <int NM>
__global__ void __launch_bounds__(WAVE) init_kernel(DevModel M, ChainLayout L, char *chains)
{
    extern __shared__ __align__(16) char lds[];
    char *base = chains + (long)blockIdx.x * L.stride;
    Ctx c = make_ctx<NM>(M, L, base, (int)blockIdx.x, lds);
    choose_outer<NM>(c);
    if (!(c.h.trace_log > -DBL_MAX)) c.h.err = 5;
    save_ctx(c);
}

// Adopt the configuration already written into tau/code/nop/outer: full trace,
// flavour lists, full inverses, sign (set_mc_config_from_arrays,
// Trace.F90:6872-7000).
template <int NM>
__global__ void __launch_bounds__(WAVE) setconfig_kernel(DevModel M, ChainLayout L, char *chains, int only)
{
    if ((int)blockIdx.x != only) return;
    extern __shared__ __align__(16) char lds[];
    char *base = chains + (long)blockIdx.x * L.stride;
    Ctx c = make_ctx<NM>(M, L, base, (int)blockIdx.x, lds);
    c.h.wpos = 0;
    set_window(c);
    double tl, ts;
    if (!full_trace<NM>(c, c.h.outer, tl, ts)) c.h.err = 5;   // W2D_ERR_ZERO_TRACE
    c.h.trace_log = tl;
    c.h.trace_sign = ts;
    store_part(c, c.sz[c.h.outer]);
    rebuild_flavour_lists(c, c.fc, c.fa, c.h.kf, c.fcs, c.fas);
    int *piv = reinterpret_cast<int *>(c.R);
    double dl = 0.0, dsg = 1.0;
    for (int f = 0; f < M.nf; ++f) {
        const int k = c.h.kf[f];
        if (k == 0) continue;
        build_delta(c, f, k, c.scratch, L.kcap, c.fc, c.fa);
        double l, sg;
        if (!lu_inplace(c.scratch, k, L.kcap, l, sg, piv)) {
            c.h.err = ERR_SING;
            break;
        }
        lu_inverse(c.scratch, k, L.kcap, piv, c.mi(f), L.kcap);
        dl += l;
        dsg *= sg;
    }
    c.h.det_log = dl;
    c.h.det_sign = dsg;
    c.h.cfgsign = -config_sign(c);
    c.h.cache_pos = -1;
    save_ctx(c);
}

}  // namespace w2d
