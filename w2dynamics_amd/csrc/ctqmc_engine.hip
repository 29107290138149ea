// Batched CT-HYB (matrix/"M-V" Z-space) Monte Carlo engine for MI355X (gfx950).
//
// One 64-lane wavefront owns one Markov chain and runs NCorr-long sweeps of
// Metropolis steps inside a single persistent launch.  Control logic (random
// draws, proposals, window bookkeeping) runs uniformly on all lanes; the heavy
// parts are lane-parallel:
//   * local trace (Trace.F90:5354-6102): block matrix products
//       K <- Psi_EB(b,s,ca) * diag(exp(-(E-E0) dtau)) * K
//     over the operators of the sliding window, starting from a cached ket at the
//     window's lower edge and contracted with a cached bra at its upper edge;
//     K holds one column per state of the outer superstate, lanes own (row, col)
//     entries, one log scale per matrix;
//   * determinant ratio S - R^T M Q and the rank-1 grow/shrink of M = Delta^-1
//     (MatrixUpdate.F90:392-482): lanes own rows of the k x k column-major block;
//   * G(tau)/Legendre accumulation (CTQMC.F90:2252-2440): lanes own (a, c) pairs.
// The random-number stream, proposal and acceptance formulas and the window
// schedule follow the reference exactly, so chain c seeded with seed0 + c makes
// the reference's accept/reject decisions for init_rnd(seed0 + c).
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <cstdint>

#include "engine.hpp"
#include "../../include/w2dyn_ctqmc.h"

namespace w2d {

#define DEV __device__ __forceinline__

// ---------------------------------------------------------------------------
// wave helpers (one wavefront per block)
// ---------------------------------------------------------------------------
DEV int lane_id() { return (int)threadIdx.x; }

DEV double wave_max(double v)
{
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, WAVE));
    return v;
}

DEV double wave_sum(double v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, WAVE);
    return v;
}

DEV int wave_isum(int v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, WAVE);
    return v;
}

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

DEV void wsync() { __syncthreads(); }

// ---------------------------------------------------------------------------
// random numbers: mt19937 + libstdc++ generate_canonical<double,53> and
// uniform_int_distribution (src/mtrng/random.cc:30-79).  The 624-word state is
// shared in LDS; every lane draws the same numbers, the twist is lane-parallel.
// ---------------------------------------------------------------------------
struct Rng {
    uint32_t *mt;   // LDS
    int idx;        // uniform
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

DEV uint32_t mt_next(Rng &r)
{
    if (r.idx >= 624) {
        mt_regen(r.mt);
        r.idx = 0;
    }
    uint32_t y = r.mt[r.idx++];
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
}

DEV double grnd(Rng &r)
{
    double sum = (double)mt_next(r);
    const double tmp = 4294967296.0;
    sum = __dadd_rn(sum, __dmul_rn((double)mt_next(r), tmp));
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
DEV int op_fl(const DevModel &M, int c) { return op_orb(c) + M.nb * op_spin(c); }

DEV int conn(const DevModel &M, int s, int code, bool adjoint)
{
    int ca = op_ca(code);
    if (adjoint) ca = 3 - ca;
    return M.connect[((s * M.nb + op_orb(code)) * 2 + op_spin(code)) * 2 + (ca - 1)];
}

DEV const double *opblock(const DevModel &M, int s, int code, bool adjoint)
{
    int ca = op_ca(code);
    if (adjoint) ca = 3 - ca;
    long o = M.opoff[((s * M.nb + op_orb(code)) * 2 + op_spin(code)) * 2 + (ca - 1)];
    return M.ops + o;
}

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
// chain context
// ---------------------------------------------------------------------------
template <int NM>
struct Ctx {
    const DevModel &M;
    ChainHdr h;          // uniform copy in registers
    Rng rng;
    char *base;
    const ChainLayout &L;
    double *tau;
    int *code;
    double *fc, *fa, *minv, *scratch, *kst, *bnd, *bndlog, *part;
    int *bndsst;
    double *ltau;
    int *lcode;
    double *K0, *K1;     // work matrices
    double *dvec;        // LDS exp factors
    double *Q, *R;       // LDS det vectors
    double *pl;          // LDS part traces scratch
    DEV Ctx(const DevModel &m, const ChainLayout &l, char *b) : M(m), base(b), L(l) {}
    DEV double *mi(int f) { return minv + (long)f * L.kcap * L.kcap; }
};

// ---------------------------------------------------------------------------
// window (Trace.F90:1608-1751)
// ---------------------------------------------------------------------------
template <int NM>
DEV void set_window(Ctx<NM> &c)
{
    if (c.h.nwin > 1) {
        c.h.tauwin_min = ((double)c.h.wpos / (double)(c.h.nwin + 1)) * c.M.beta;
        c.h.tauwin_max = ((double)(c.h.wpos + 2) / (double)(c.h.nwin + 1)) * c.M.beta;
    }
}

template <int NM>
DEV void shift_window(Ctx<NM> &c)
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

// ---------------------------------------------------------------------------
// local trace: matrix propagation with one log scale per matrix
// ---------------------------------------------------------------------------
// Apply exp(-(E - E0) dt) per row of K (ns x no, row-major K[r*no + i]).
template <int NM>
DEV void prop_rows(Ctx<NM> &c, double *K, int s, int no, double dt, double &lg)
{
    const DevModel &M = c.M;
    int ns = M.sizes[s];
    const double *E = M.evals + M.offsets[s];
    lg += -(E[0] - M.egs) * dt;
    if (ns == 1) return;
    for (int r = lane_id(); r < ns; r += WAVE) c.dvec[r] = exp(-(E[r] - E[0]) * dt);
    wsync();
    for (int idx = lane_id(); idx < ns * no; idx += WAVE) K[idx] *= c.dvec[idx / no];
    wsync();
}

template <int NM>
DEV void normalise(Ctx<NM> &c, double *K, int n, double &lg)
{
    double m = 0.0;
    for (int idx = lane_id(); idx < n; idx += WAVE) m = fmax(m, fabs(K[idx]));
    m = wave_max(m);
    if (m > 0.0) {
        double inv = 1.0 / m;
        for (int idx = lane_id(); idx < n; idx += WAVE) K[idx] *= inv;
        lg += log(m);
    }
    wsync();
}

// Kout (nt x no) = P (nt x ns, column-major) * Kin (ns x no)
template <int NM>
DEV void apply_block(const double *P, const double *Kin, double *Kout, int nt, int ns, int no)
{
    for (int idx = lane_id(); idx < nt * no; idx += WAVE) {
        int r = idx / no, i = idx - r * no;
        double acc = 0.0;
        for (int k = 0; k < ns; ++k) acc = fma(P[r + (long)k * nt], Kin[k * no + i], acc);
        Kout[idx] = acc;
    }
    wsync();
}

// algorithmic work of one operator application on `no` outer states: a
// (nt x ns) block times an (ns x no) state matrix plus the diagonal propagator;
// bytes = operator block + states in and out
template <int NM>
DEV void trace_work(Ctx<NM> &c, int nt, int ns, int no)
{
    c.h.alg_flops += (double)no * (2.0 * nt * ns + ns);
    c.h.alg_bytes += 8.0 * nt * ns + 8.0 * no * (ns + nt);
}

// Propagate a ket matrix upward in tau through `n` listed operators starting at
// tau t0, ending at tau t1.  *K (in/out) is swapped between K0/K1.
template <int NM>
DEV void ket_sweep(Ctx<NM> &c, double *&K, double *&Kw, int &s, double &lg, int no, double t0,
                   const double *otau, const int *ocode, int n, double t1)
{
    const DevModel &M = c.M;
    double t = t0;
    for (int j = 0; j < n; ++j) {
        double tj = otau[j];
        int cj = ocode[j];
        prop_rows(c, K, s, no, tj - t, lg);
        int tsst = conn(M, s, cj, false);
        trace_work(c, M.sizes[tsst], M.sizes[s], no);
        apply_block<NM>(opblock(M, s, cj, false), K, Kw, M.sizes[tsst], M.sizes[s], no);
        double *x = K;
        K = Kw;
        Kw = x;
        s = tsst;
        normalise(c, K, M.sizes[s] * no, lg);
        t = tj;
    }
    prop_rows(c, K, s, no, t1 - t, lg);
    normalise(c, K, M.sizes[s] * no, lg);
}

// Propagate a bra matrix downward from t1 through the listed operators (taken
// in descending order) to t0.
template <int NM>
DEV void bra_sweep(Ctx<NM> &c, double *&B, double *&Bw, int &s, double &lg, int no, double t1,
                   const double *otau, const int *ocode, int n, double t0)
{
    const DevModel &M = c.M;
    double t = t1;
    for (int j = n - 1; j >= 0; --j) {
        double tj = otau[j];
        int cj = ocode[j];
        prop_rows(c, B, s, no, t - tj, lg);
        int tsst = conn(M, s, cj, true);
        trace_work(c, M.sizes[tsst], M.sizes[s], no);
        apply_block<NM>(opblock(M, s, cj, true), B, Bw, M.sizes[tsst], M.sizes[s], no);
        double *x = B;
        B = Bw;
        Bw = x;
        s = tsst;
        normalise(c, B, M.sizes[s] * no, lg);
        t = tj;
    }
    prop_rows(c, B, s, no, t - t0, lg);
    normalise(c, B, M.sizes[s] * no, lg);
}

template <int NM>
DEV void set_identity(Ctx<NM> &c, double *K, int no)
{
    for (int idx = lane_id(); idx < no * no; idx += WAVE) K[idx] = (idx / no == idx % no) ? 1.0 : 0.0;
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
// column; writes part traces to c.pl and returns the total in (tl, ts).
template <int NM>
DEV void contract(Ctx<NM> &c, const double *K, const double *B, int s, int no, double lg,
                  double &tl, double &ts)
{
    int ns = c.M.sizes[s];
    for (int i = lane_id(); i < no; i += WAVE) {
        double acc = 0.0;
        for (int r = 0; r < ns; ++r) acc = fma(K[r * no + i], B[r * no + i], acc);
        c.pl[2 * i] = lg + log(fabs(acc));
        c.pl[2 * i + 1] = copysign(1.0, acc);
    }
    wsync();
    tl = -DBL_MAX;
    ts = 0.0;
    for (int i = 0; i < no; ++i) log_add(tl, ts, c.pl[2 * i], c.pl[2 * i + 1]);
}

// Full trace of the current configuration from tau = 0 to beta (global moves,
// tau shifts, set_config): K = identity on the outer superstate, through all
// operators, contracted with the identity bra at beta.
template <int NM>
DEV void full_trace(Ctx<NM> &c, int outer, double &tl, double &ts)
{
    const DevModel &M = c.M;
    int no = M.sizes[outer];
    double *K = c.K0, *Kw = c.K1;
    set_identity(c, K, no);
    int s = outer;
    double lg = 0.0;
    ket_sweep(c, K, Kw, s, lg, no, 0.0, c.tau, c.code, c.h.nop, M.beta);
    // bra at beta = identity on the outer superstate
    set_identity(c, Kw, no);
    contract(c, K, Kw, s, no, lg, tl, ts);
}

// ---------------------------------------------------------------------------
// boundary cache: ket at the window's lower edge, bras at every upper edge
// ---------------------------------------------------------------------------
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

template <int NM>
DEV void copy_mat(double *dst, const double *src, int n)
{
    for (int i = lane_id(); i < n; i += WAVE) dst[i] = src[i];
    wsync();
}

// Rebuild the caches for window position p: ket over ops with tau < b_p, bras
// over ops with tau > b_m for m = p+2 .. nwin+1.
template <int NM>
DEV void rebuild_cache(Ctx<NM> &c)
{
    const DevModel &M = c.M;
    int outer = c.h.outer, no = M.sizes[outer];
    int p = c.h.nwin > 1 ? c.h.wpos : 0;
    int nw = c.h.nwin;
    int top = (nw > 1) ? nw + 1 : 2;
    // bras from beta downward
    double *B = c.K0, *Bw = c.K1;
    set_identity(c, B, no);
    int s = outer;
    double lg = 0.0;
    double tb = M.beta;
    int hiidx = c.h.nop;
    copy_mat<NM>(c.bnd + (long)top * NM * NM, B, no * no);
    if (lane_id() == 0) {
        c.bndlog[top] = lg;
        c.bndsst[top] = s;
    }
    for (int m = top - 1; m >= p + 2; --m) {
        double tm = bound_tau(m, nw, M.beta);
        int loidx = upper_bound_tau(c.tau, c.h.nop, tm);
        bra_sweep(c, B, Bw, s, lg, no, tb, c.tau + loidx, c.code + loidx, hiidx - loidx, tm);
        copy_mat<NM>(c.bnd + (long)m * NM * NM, B, M.sizes[s] * no);
        if (lane_id() == 0) {
            c.bndlog[m] = lg;
            c.bndsst[m] = s;
        }
        tb = tm;
        hiidx = loidx;
    }
    // ket at the lower edge
    double *K = c.K0, *Kw = c.K1;
    set_identity(c, K, no);
    s = outer;
    lg = 0.0;
    double tp = bound_tau(p, nw, M.beta);
    int n0 = lower_bound_tau(c.tau, c.h.nop, tp);
    ket_sweep(c, K, Kw, s, lg, no, 0.0, c.tau, c.code, n0, tp);
    copy_mat<NM>(c.kst, K, M.sizes[s] * no);
    c.h.kst_log = lg;
    c.h.kst_sst = s;
    c.h.cache_pos = p;
    wsync();
}

// Bring the caches to the current window position: advance the ket across
// the passed half windows, or rebuild after a wrap / invalidation.
template <int NM>
DEV void ensure_cache(Ctx<NM> &c)
{
    const DevModel &M = c.M;
    int p = c.h.nwin > 1 ? c.h.wpos : 0;
    if (c.h.cache_pos == p) return;
    if (c.h.cache_pos < 0 || c.h.cache_pos > p) {
        rebuild_cache(c);
        return;
    }
    int no = M.sizes[c.h.outer];
    int s = c.h.kst_sst;
    double lg = c.h.kst_log;
    double *K = c.K0, *Kw = c.K1;
    copy_mat<NM>(K, c.kst, M.sizes[s] * no);
    double ta = bound_tau(c.h.cache_pos, c.h.nwin, M.beta);
    double tb = bound_tau(p, c.h.nwin, M.beta);
    int ia = lower_bound_tau(c.tau, c.h.nop, ta);
    int ib = lower_bound_tau(c.tau, c.h.nop, tb);
    ket_sweep(c, K, Kw, s, lg, no, ta, c.tau + ia, c.code + ia, ib - ia, tb);
    copy_mat<NM>(c.kst, K, M.sizes[s] * no);
    c.h.kst_log = lg;
    c.h.kst_sst = s;
    c.h.cache_pos = p;
    wsync();
}

// Trace of a trial window list (the window's operators after the proposed
// change) between the cached ket and bra.
template <int NM>
DEV void window_trace(Ctx<NM> &c, int nl, double &tl, double &ts)
{
    const DevModel &M = c.M;
    int no = M.sizes[c.h.outer];
    int p = c.h.nwin > 1 ? c.h.wpos : 0;
    int top = p + 2;
    double ta = bound_tau(p, c.h.nwin, M.beta), tb = bound_tau(top, c.h.nwin, M.beta);
    double *K = c.K0, *Kw = c.K1;
    int s = c.h.kst_sst;
    copy_mat<NM>(K, c.kst, M.sizes[s] * no);
    double lg = c.h.kst_log;
    ket_sweep(c, K, Kw, s, lg, no, ta, c.ltau, c.lcode, nl, tb);
    contract(c, K, c.bnd + (long)top * NM * NM, s, no, lg + c.bndlog[top], tl, ts);
}

// QN walk over the trial list from the cached lower-edge superstate; valid iff
// every step exists and it ends in the cached upper-edge superstate
// (check_sst_sequence, Trace.F90:1243-1302, on a superset of the changed span).
template <int NM>
DEV bool window_qn(Ctx<NM> &c, int nl)
{
    int p = c.h.nwin > 1 ? c.h.wpos : 0;
    int s = c.h.kst_sst;
    for (int j = 0; j < nl && s >= 0; ++j) s = conn(c.M, s, c.lcode[j], false);
    return s >= 0 && s == c.bndsst[p + 2];
}

// indices [ia, ib) of the operators inside the window [tauwin_min, tauwin_max]
template <int NM>
DEV void window_range(Ctx<NM> &c, int &ia, int &ib)
{
    ia = lower_bound_tau(c.tau, c.h.nop, c.h.tauwin_min);
    ib = upper_bound_tau(c.tau, c.h.nop, c.h.tauwin_max);
}

// ---------------------------------------------------------------------------
// removable pairs (ensure_valid_NPairs_diag, Trace.F90:1769-1810) over a list
// ---------------------------------------------------------------------------
template <int NM>
DEV int count_pairs_from(Ctx<NM> &c, const double *tau, const int *code, int n, int j, double wmin,
                         double wmax)
{
    const DevModel &M = c.M;
    double te = tau[j];
    if (te < wmin || te > wmax) return 0;
    int fe = op_fl(M, code[j]), ce = op_ca(code[j]);
    int cnt = 0;
    for (int k = j + 1; k < n; ++k) {
        double tk = tau[k];
        if (tk > wmax || tk > te + c.h.taudiff_max) break;
        int ck = code[k];
        if (op_fl(M, ck) == fe && op_ca(ck) != ce) ++cnt;
    }
    return cnt;
}

template <int NM>
DEV int npairs_list(Ctx<NM> &c, const double *tau, const int *code, int n)
{
    int cnt = 0;
    for (int j = lane_id(); j < n; j += WAVE)
        cnt += count_pairs_from(c, tau, code, n, j, c.h.tauwin_min, c.h.tauwin_max);
    return wave_isum(cnt);
}

// ---------------------------------------------------------------------------
// configuration sign: get_Sign with get_BlockSign and get_permaltca
// (Trace.F90:3479-3615), lane-parallel over chunks of the operator list
// ---------------------------------------------------------------------------
template <int NM>
DEV double config_sign(Ctx<NM> &c)
{
    const DevModel &M = c.M;
    const int n = c.h.nop, nb = M.nb, nf = M.nf;
    const int chunk = (n + WAVE - 1) / WAVE;
    const int j0 = lane_id() * chunk, j1 = min(n, j0 + chunk);
    int cnt[MAXF];
    for (int f = 0; f < nf; ++f) cnt[f] = 0;
    for (int j = j0; j < j1; ++j) cnt[op_fl(M, c.code[j])]++;
    // suffix counts: operators of each flavour after this lane's chunk
    int after[MAXF];
    for (int f = 0; f < nf; ++f) {
        int tot;
        int ex = wave_excl_scan(cnt[f], &tot);
        after[f] = tot - ex - cnt[f];
    }
    int perm = 0, alt = 0;
    for (int j = j1 - 1; j >= j0; --j) {
        int cj = c.code[j];
        int o = op_orb(cj), sp = op_spin(cj), f = o + nb * sp;
        // TNFlav after the decrement equals the count of flavour-f operators
        // after j; BlockSign adds flavour counts of lower blocks
        if (sp == 1) {
            for (int b = 0; b <= o; ++b) perm += after[b];
        } else {
            for (int b = 0; b < o; ++b) perm += after[b + nb];
        }
        for (int b = 0; b < o; ++b) perm += after[b + nb * sp];
        if (op_ca(cj) == 1) alt += after[f] & 1;
        after[f]++;
    }
    perm = wave_isum(perm & 1) & 1;
    alt = wave_isum(alt & 1) & 1;
    if ((n / 2) % 2 == 1) alt ^= 1;
    double block = perm ? -1.0 : 1.0, altsign = alt ? -1.0 : 1.0;
    return c.h.trace_sign * c.h.det_sign * block * altsign;
}

// ---------------------------------------------------------------------------
// hybridisation matrix algebra (MatrixUpdate.F90)
// ---------------------------------------------------------------------------
// S - R^T M Q for an order-k block (get_DetRatAdd)
template <int NM>
DEV double detrat_add(Ctx<NM> &c, const double *Mf, int k, double S)
{
    const int ld = c.L.kcap;
    double part = 0.0;
    for (int i = lane_id(); i < k; i += WAVE) {
        double v = 0.0;
        for (int j = 0; j < k; ++j) v = fma(Mf[i + (long)j * ld], c.Q[j], v);
        part = fma(c.R[i], v, part);
    }
    return S - wave_sum(part);
}

// In-place bordering update to order k+1 with the new row p1 / column p2
// (1-based), get_MatAdd (MatrixUpdate.F90:410-442).  Uses LDS scratch
// col[k+1] = new column p2, row[k+1] = new row p1 (c.Q / c.R are overwritten).
template <int NM>
DEV void mat_add(Ctx<NM> &c, double *Mf, int k, double dr, int p1, int p2, double *col, double *row)
{
    const int ld = c.L.kcap;
    const int l = lane_id();
    // col: -(M Q)/dr inserted at p1 -> 1/dr ; row: -(R M)/dr
    for (int i = l; i < k; i += WAVE) {
        double v = 0.0;
        for (int j = 0; j < k; ++j) v = fma(Mf[i + (long)j * ld], c.Q[j], v);
        col[i < p1 - 1 ? i : i + 1] = -v / dr;
    }
    for (int j = l; j < k; j += WAVE) {
        double v = 0.0;
        for (int i = 0; i < k; ++i) v = fma(c.R[i], Mf[i + (long)j * ld], v);
        row[j < p2 - 1 ? j : j + 1] = -v / dr;
    }
    wsync();
    if (l == 0) {
        col[p1 - 1] = 1.0 / dr;
        row[p2 - 1] = 1.0 / dr;
    }
    wsync();
    // new(i', j') = old(i, j) + dr * col(i') * row(j'); columns and row chunks in
    // descending order so that no old entry is overwritten before it is read
    for (int jp = k; jp >= 0; --jp) {
        int j = jp < p2 - 1 ? jp : jp - 1;
        int nchunk = (k + 1 + WAVE - 1) / WAVE;
        for (int ch = nchunk - 1; ch >= 0; --ch) {
            int ip = ch * WAVE + l;
            double val = 0.0;
            bool act = ip <= k;
            if (act) {
                if (jp == p2 - 1) {
                    val = col[ip];
                } else if (ip == p1 - 1) {
                    val = row[jp];
                } else {
                    int i = ip < p1 - 1 ? ip : ip - 1;
                    val = Mf[i + (long)j * ld] + dr * col[ip] * row[jp];
                }
            }
            wsync();
            if (act) Mf[ip + (long)jp * ld] = val;
            wsync();
        }
    }
}

// In-place removal of row p1 / column p2 (1-based), get_MatRem
// (MatrixUpdate.F90:455-482).
template <int NM>
DEV void mat_rem(Ctx<NM> &c, double *Mf, int k, int p1, int p2, double *colp, double *rowp)
{
    const int ld = c.L.kcap;
    const int l = lane_id();
    for (int i = l; i < k; i += WAVE) {
        colp[i] = Mf[i + (long)(p2 - 1) * ld];
        rowp[i] = Mf[(p1 - 1) + (long)i * ld];
    }
    wsync();
    const double piv = rowp[p2 - 1];
    const int km = k - 1;
    for (int j = 0; j < km; ++j) {
        int jo = j < p2 - 1 ? j : j + 1;
        int nchunk = (km + WAVE - 1) / WAVE;
        for (int ch = 0; ch < nchunk; ++ch) {
            int i = ch * WAVE + l;
            double val = 0.0;
            bool act = i < km;
            if (act) {
                int io = i < p1 - 1 ? i : i + 1;
                val = Mf[io + (long)jo * ld] - colp[io] * rowp[jo] / piv;
            }
            wsync();
            if (act) Mf[i + (long)j * ld] = val;
            wsync();
        }
    }
}

// LU with partial pivoting of A (k x k, ld) in place; log|det| and sign as
// get_LogDetFull (MatrixUpdate.F90:302-339).  Returns false if singular.
template <int NM>
DEV bool lu_inplace(Ctx<NM> &c, double *A, int k, int ld, double &dl, double &ds, int *piv)
{
    const int l = lane_id();
    dl = 0.0;
    ds = 1.0;
    for (int col = 0; col < k; ++col) {
        // pivot search
        double best = -1.0;
        int bi = col;
        for (int r = col + l; r < k; r += WAVE) {
            double v = fabs(A[r + (long)col * ld]);
            if (v > best) best = v, bi = r;
        }
        for (int o = 32; o > 0; o >>= 1) {
            double ob = __shfl_xor(best, o, WAVE);
            int oi = __shfl_xor(bi, o, WAVE);
            if (ob > best || (ob == best && oi < bi)) best = ob, bi = oi;
        }
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
        for (int j = col + 1; j < k; ++j) {
            double u = A[col + (long)j * ld];
            for (int r = col + 1 + l; r < k; r += WAVE) A[r + (long)j * ld] -= A[r + (long)col * ld] * u;
        }
        wsync();
    }
    return true;
}

// Inverse from an in-place LU (A overwritten by its inverse); uses `work`
// (k x k, ld) as target.
template <int NM>
DEV void lu_inverse(Ctx<NM> &c, const double *LU, int k, int ld, const int *piv, double *out)
{
    const int l = lane_id();
    for (int col = l; col < k; col += WAVE) {
        double *x = out + (long)col * ld;
        for (int i = 0; i < k; ++i) x[i] = (i == col) ? 1.0 : 0.0;
        for (int i = 0; i < k; ++i) {
            int p = piv[i];
            double t = x[i];
            x[i] = x[p];
            x[p] = t;
        }
        for (int i = 0; i < k; ++i) {
            double xi = x[i];
            for (int r = i + 1; r < k; ++r) x[r] -= LU[r + (long)i * ld] * xi;
        }
        for (int i = k - 1; i >= 0; --i) {
            double xi = x[i] / LU[i + (long)i * ld];
            x[i] = xi;
            for (int r = 0; r < i; ++r) x[r] -= LU[r + (long)i * ld] * xi;
        }
    }
    wsync();
}

// Delta_f(row = annihilator rank, col = creator rank) = F(tau_a - tau_c)
// (get_InvFullHybr, Trace.F90:2279-2333) from the per-flavour lists
template <int NM>
DEV void build_delta(Ctx<NM> &c, int f, int k, double *D)
{
    const int ld = c.L.kcap;
    const double *ct = c.fc + (long)f * c.L.kcap, *at = c.fa + (long)f * c.L.kcap;
    for (int idx = lane_id(); idx < k * k; idx += WAVE) {
        int r = idx % k, col = idx / k;
        D[r + (long)col * ld] = hybr(c.M, at[r] - ct[col], f);
    }
    wsync();
}

// ---------------------------------------------------------------------------
// per-flavour tau lists
// ---------------------------------------------------------------------------
template <int NM>
DEV void list_insert(double *arr, int n, int pos, double v)
{
    // shift arr[pos..n) up by one, descending chunks
    const int l = lane_id();
    int cnt = n - pos;
    for (int ch = (cnt + WAVE - 1) / WAVE - 1; ch >= 0; --ch) {
        int i = pos + ch * WAVE + l;
        double x = 0.0;
        if (i < n) x = arr[i];
        wsync();
        if (i < n) arr[i + 1] = x;
        wsync();
    }
    if (l == 0) arr[pos] = v;
    wsync();
}

template <int NM>
DEV void list_remove(double *arr, int n, int pos)
{
    const int l = lane_id();
    int cnt = n - pos - 1;
    for (int ch = 0; ch < (cnt + WAVE - 1) / WAVE; ++ch) {
        int i = pos + 1 + ch * WAVE + l;
        double x = 0.0;
        if (i < n) x = arr[i];
        wsync();
        if (i < n) arr[i - 1] = x;
        wsync();
    }
}

template <typename T>
DEV void ops_shift(T *arr, int n, int from, int by)
{
    // move arr[from..n) by `by` (positive: up, descending; negative: down, ascending)
    const int l = lane_id();
    int cnt = n - from;
    int nch = (cnt + WAVE - 1) / WAVE;
    if (by > 0) {
        for (int ch = nch - 1; ch >= 0; --ch) {
            int i = from + ch * WAVE + l;
            T x{};
            if (i < n) x = arr[i];
            wsync();
            if (i < n) arr[i + by] = x;
            wsync();
        }
    } else {
        for (int ch = 0; ch < nch; ++ch) {
            int i = from + ch * WAVE + l;
            T x{};
            if (i < n) x = arr[i];
            wsync();
            if (i < n) arr[i + by] = x;
            wsync();
        }
    }
}

DEV int count_below(const double *arr, int n, double x)
{
    int lo = 0, hi = n;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (arr[mid] < x) lo = mid + 1; else hi = mid;
    }
    return lo;
}

// ---------------------------------------------------------------------------
// moves
// ---------------------------------------------------------------------------
enum { OUT_REJ = 0, OUT_ACC = 1, OUT_QN = 2 };
enum { ERR_CAP = 3, ERR_SING = 4, ERR_STATE_DEV = 6 };

template <int NM>
DEV void store_part(Ctx<NM> &c, int no)
{
    for (int i = lane_id(); i < 2 * no; i += WAVE) c.part[i] = c.pl[i];
    wsync();
}

// StepAdd (CTQMC.F90:3353-3612), change_outer = false
template <int NM>
DEV int step_add(Ctx<NM> &c)
{
    const DevModel &M = c.M;
    c.h.try_add++;
    double width = c.h.tauwin_max - c.h.tauwin_min;
    double tref = c.h.tauwin_min + width * grnd(c.rng);
    double tdf = width;
    int orb = randint(c.rng, 1, M.nb) - 1;
    int spin = randint(c.rng, 1, 2) - 1;
    double tmin = fmax(tref - c.h.taudiff_max, c.h.tauwin_min);
    double w2 = fmin(tref + c.h.taudiff_max, c.h.tauwin_max) - tmin;
    double t2 = tmin + w2 * grnd(c.rng);
    tdf *= w2;
    const int f = orb + M.nb * spin;
    // equal times (process_OperAdd, Trace.F90:4287-4300, 4330-4380)
    bool eq = (t2 == tref);
    {
        int hit = 0;
        for (int j = lane_id(); j < c.h.nop; j += WAVE) hit |= (c.tau[j] == tref) | (c.tau[j] == t2);
        eq = eq || (wave_isum(hit) != 0);
    }
    if (eq) return c.h.qn_rej++, OUT_QN;
    const int k = c.h.kf[f];
    if (k + 1 > c.L.kcap || c.h.nop + 2 > c.L.cap) {
        c.h.err = ERR_CAP;
        return OUT_REJ;
    }
    // time order: annihilator at tref, creator at t2
    double ta = tref, tc = t2;
    double t1 = fmin(ta, tc), t2o = fmax(ta, tc);
    int code_a = mk_code(orb, spin, 2), code_c = mk_code(orb, spin, 1);
    int code1 = ta < tc ? code_a : code_c, code2 = ta < tc ? code_c : code_a;
    ensure_cache(c);
    // trial window list: window operators with the two new ones merged in
    int ia, ib;
    window_range(c, ia, ib);
    int nw = ib - ia;
    if (nw + 2 > c.L.lcap) {
        c.h.err = ERR_CAP;
        return OUT_REJ;
    }
    int q1 = lower_bound_tau(c.tau + ia, nw, t1);     // positions in window
    int q2 = lower_bound_tau(c.tau + ia, nw, t2o);
    for (int j = lane_id(); j < nw; j += WAVE) {
        int d = j + (j >= q1 ? 1 : 0) + (j >= q2 ? 1 : 0);
        c.ltau[d] = c.tau[ia + j];
        c.lcode[d] = c.code[ia + j];
    }
    if (lane_id() == 0) {
        c.ltau[q1] = t1;
        c.lcode[q1] = code1;
        c.ltau[q2 + 1] = t2o;
        c.lcode[q2 + 1] = code2;
    }
    wsync();
    const int nl = nw + 2;
    if (!window_qn(c, nl)) return c.h.qn_rej++, OUT_QN;
    // FPos: 1-based ranks of the new creator / annihilator in their lists
    const double *ctl = c.fc + (long)f * c.L.kcap, *atl = c.fa + (long)f * c.L.kcap;
    const int fp_c = count_below(ctl, k, tc) + 1, fp_a = count_below(atl, k, ta) + 1;
    // Q (annihilators), R (creators): get_LogDetRatPairAdd (Trace.F90:2775-2832)
    for (int j = lane_id(); j < k; j += WAVE) {
        c.Q[j] = hybr(M, atl[j] - tc, f);
        c.R[j] = hybr(M, ta - ctl[j], f);
    }
    wsync();
    double S = detrat_add(c, c.mi(f), k, hybr(M, ta - tc, f));
    c.h.alg_flops += 2.0 * k * k;
    c.h.alg_bytes += 8.0 * k * k;
    double sig = (abs(fp_c - fp_a) % 2) ? -1.0 : 1.0;
    double dr_log = log(fabs(S)), dr_sign = copysign(1.0, S) * sig;
    double tl, ts;
    window_trace(c, nl, tl, ts);
    double rnd = grnd(c.rng);
    // NPairs of the new configuration: pairs with both ends in the window all
    // lie in the trial list
    int np = npairs_list(c, c.ltau, c.lcode, nl);
    double wr = exp((tl - c.h.trace_log) + dr_log);
    double prob = fabs(((1.0 * tdf) * (double)(2 * M.nb)) / (double)np * wr);
    if (!(rnd < prob)) return OUT_REJ;
    // accept: M, operator list, flavour lists, weights, sign
    double *colb = c.dvec, *rowb = c.dvec + c.L.kcap + 1;
    mat_add(c, c.mi(f), k, S, fp_c, fp_a, colb, rowb);
    c.h.alg_flops += 6.0 * k * k;
    c.h.alg_bytes += 8.0 * k * k + 8.0 * (k + 1) * (k + 1);
    int g1 = ia + q1, g2 = ia + q2 + 1;
    ops_shift(c.tau, c.h.nop, ia + q2, 2);
    ops_shift(c.code, c.h.nop, ia + q2, 2);
    ops_shift(c.tau, ia + q2, ia + q1, 1);
    ops_shift(c.code, ia + q2, ia + q1, 1);
    if (lane_id() == 0) {
        c.tau[g1] = t1;
        c.code[g1] = code1;
        c.tau[g2] = t2o;
        c.code[g2] = code2;
    }
    wsync();
    list_insert<NM>(c.fc + (long)f * c.L.kcap, k, fp_c - 1, tc);
    list_insert<NM>(c.fa + (long)f * c.L.kcap, k, fp_a - 1, ta);
    c.h.nop += 2;
    c.h.kf[f] = k + 1;
    c.h.det_log += dr_log;
    c.h.det_sign *= dr_sign;
    c.h.trace_log = tl;
    c.h.trace_sign = ts;
    store_part(c, M.sizes[c.h.outer]);
    c.h.cfgsign = -config_sign(c);
    c.h.acc_add++;
    return OUT_ACC;
}

// StepRem (CTQMC.F90:3892-4150) with gen_OperRemove (Trace.F90:4851-4977)
template <int NM>
DEV int step_rem(Ctx<NM> &c)
{
    const DevModel &M = c.M;
    c.h.try_rem++;
    ensure_cache(c);
    int ia, ib;
    window_range(c, ia, ib);
    int nw = ib - ia;
    if (nw > c.L.lcap) {
        c.h.err = ERR_CAP;
        return OUT_REJ;
    }
    // pair counts per window operator, prefix-summed over the wave
    int np = 0;
    int i1 = -1, i2 = -1;
    {
        int base = 0;
        for (int j0 = 0; j0 < nw; j0 += WAVE) {
            int j = j0 + lane_id();
            int cnt = j < nw ? count_pairs_from(c, c.tau, c.code, c.h.nop, ia + j, c.h.tauwin_min,
                                                c.h.tauwin_max) : 0;
            int tot;
            wave_excl_scan(cnt, &tot);
            base += tot;
        }
        np = base;
    }
    if (np <= 0) return c.h.qn_rej++, OUT_QN;
    int pick = randint(c.rng, 1, np);
    if (nw < 2) {
        c.h.err = ERR_STATE_DEV;
        return OUT_REJ;
    }
    {
        int base = 0;
        for (int j0 = 0; j0 < nw; j0 += WAVE) {
            int j = j0 + lane_id();
            int cnt = j < nw ? count_pairs_from(c, c.tau, c.code, c.h.nop, ia + j, c.h.tauwin_min,
                                                c.h.tauwin_max) : 0;
            int tot;
            int ex = wave_excl_scan(cnt, &tot);
            int found = (pick > base + ex && pick <= base + ex + cnt) ? j : -1;
            for (int o = 32; o > 0; o >>= 1) found = max(found, __shfl_xor(found, o, WAVE));
            if (found >= 0) {
                int e = ia + found;
                int want = pick - base - __shfl(ex, found - j0, WAVE);
                int fe = op_fl(M, c.code[e]), ce = op_ca(c.code[e]);
                int kk = e + 1;
                for (; kk < c.h.nop; ++kk) {
                    int ck = c.code[kk];
                    if (op_fl(M, ck) == fe && op_ca(ck) != ce)
                        if (--want <= 0) break;
                }
                i1 = e;
                i2 = kk;
                break;
            }
            base += tot;
        }
    }
    if (i1 < 0 || i2 >= c.h.nop) {
        c.h.err = ERR_STATE_DEV;
        return OUT_REJ;
    }
    const int c1 = c.code[i1], c2 = c.code[i2];
    const double t1 = c.tau[i1], t2 = c.tau[i2];
    const int f = op_fl(M, c1);
    const int k = c.h.kf[f];
    const double tc = op_ca(c1) == 1 ? t1 : t2, ta = op_ca(c1) == 1 ? t2 : t1;
    const double *ctl = c.fc + (long)f * c.L.kcap, *atl = c.fa + (long)f * c.L.kcap;
    const int fp_c = count_below(ctl, k, tc) + 1, fp_a = count_below(atl, k, ta) + 1;
    double hprf = (double)np /
                  ((c.h.tauwin_max - c.h.tauwin_min) *
                   (fmin(ta + c.h.taudiff_max, c.h.tauwin_max) - fmax(ta - c.h.taudiff_max, c.h.tauwin_min)));
    // trial list without the pair
    for (int j = lane_id(); j < nw; j += WAVE) {
        int g = ia + j;
        if (g == i1 || g == i2) continue;
        int d = j - (g > i1 ? 1 : 0) - (g > i2 ? 1 : 0);
        c.ltau[d] = c.tau[g];
        c.lcode[d] = c.code[g];
    }
    wsync();
    const int nl = nw - 2;
    if (!window_qn(c, nl)) return c.h.qn_rej++, OUT_QN;
    // get_LogDetRatPairRem (Trace.F90:2937-2963): M(FPos(1), FPos(2))
    const int ld = c.L.kcap;
    double S = c.mi(f)[(fp_c - 1) + (long)(fp_a - 1) * ld];
    double sig = (abs(fp_c - fp_a) % 2) ? -1.0 : 1.0;
    double dr_log = log(fabs(S)), dr_sign = copysign(1.0, S) * sig;
    double tl, ts;
    window_trace(c, nl, tl, ts);
    double wr = exp((tl - c.h.trace_log) + dr_log);
    double prob = fabs(((1.0 * hprf) / (double)(2 * M.nb)) * wr);
    double rnd = grnd(c.rng);
    if (!(rnd < prob)) return OUT_REJ;
    if (fabs(dr_sign * exp(dr_log)) < 1e-12) {
        c.h.err = ERR_SING;
        return OUT_REJ;
    }
    mat_rem(c, c.mi(f), k, fp_c, fp_a, c.dvec, c.dvec + c.L.kcap);
    c.h.alg_flops += 3.0 * k * k;
    c.h.alg_bytes += 8.0 * k * k + 8.0 * (k - 1) * (k - 1);
    ops_shift(c.tau, i2, i1 + 1, -1);
    ops_shift(c.code, i2, i1 + 1, -1);
    ops_shift(c.tau, c.h.nop, i2 + 1, -2);
    ops_shift(c.code, c.h.nop, i2 + 1, -2);
    list_remove<NM>(c.fc + (long)f * c.L.kcap, k, fp_c - 1);
    list_remove<NM>(c.fa + (long)f * c.L.kcap, k, fp_a - 1);
    c.h.nop -= 2;
    c.h.kf[f] = k - 1;
    c.h.det_log += dr_log;
    c.h.det_sign *= dr_sign;
    c.h.trace_log = tl;
    c.h.trace_sign = ts;
    store_part(c, M.sizes[c.h.outer]);
    c.h.cfgsign = -config_sign(c);
    c.h.acc_rem++;
    return OUT_ACC;
}

// relabel flavours / swap c and c+ (globalUpdate, Trace.F90:2126-2188) and
// rebuild the per-flavour tau lists
template <int NM>
DEV void global_update(Ctx<NM> &c, const int *gu)
{
    const DevModel &M = c.M;
    for (int j = lane_id(); j < c.h.nop; j += WAVE) {
        int cj = c.code[j];
        int idx = op_orb(cj) + M.nb * op_spin(cj) + 1;
        int g = gu[idx];
        int ca = op_ca(cj);
        if (gu[0] != 0) ca = 3 - ca;
        c.code[j] = mk_code((g - 1) % M.nb, (g - 1) / M.nb, ca);
    }
    wsync();
}

template <int NM>
DEV void rebuild_flavour_lists(Ctx<NM> &c)
{
    const DevModel &M = c.M;
    // one lane per flavour (nf <= 14)
    int f = lane_id();
    if (f < M.nf) {
        int nc = 0, na = 0;
        for (int j = 0; j < c.h.nop; ++j) {
            int cj = c.code[j];
            if (op_fl(M, cj) != f) continue;
            if (op_ca(cj) == 1) c.fc[(long)f * c.L.kcap + nc++] = c.tau[j];
            else c.fa[(long)f * c.L.kcap + na++] = c.tau[j];
        }
    }
    wsync();
    for (int g = 0; g < M.nf; ++g) {
        int n = 0;
        for (int j = lane_id(); j < c.h.nop; j += WAVE) n += (op_fl(M, c.code[j]) == g && op_ca(c.code[j]) == 1);
        c.h.kf[g] = wave_isum(n);
    }
}

template <int NM>
DEV bool qn_full(Ctx<NM> &c, int s)
{
    for (int j = 0; j < c.h.nop && s >= 0; ++j) s = conn(c.M, s, c.code[j], false);
    return s >= 0;
}

template <int NM>
DEV bool qn_cycle(Ctx<NM> &c, int s0)
{
    int s = s0;
    for (int j = 0; j < c.h.nop && s >= 0; ++j) s = conn(c.M, s, c.code[j], false);
    return s == s0;
}

// StepGlob (CTQMC.F90:4433-4554)
template <int NM>
DEV int step_glob(Ctx<NM> &c)
{
    const DevModel &M = c.M;
    c.h.try_glob++;
    if (c.h.nop == 0) return OUT_REJ;
    const int nf = M.nf, nb = M.nb;
    int gu[MAXF + 1], igu[MAXF + 1];
    double r = grnd(c.rng);
    bool spinflip = false;
    if (r < 0.3) {
        spinflip = true;
        gu[0] = 0;
        for (int i = 1; i <= nb; ++i) gu[i] = i + nb;
        for (int i = nb + 1; i <= 2 * nb; ++i) gu[i] = i - nb;
    } else if (r < 0.6) {
        gu[0] = 1;
        for (int i = 1; i <= nf; ++i) gu[i] = i;
    } else {
        for (int i = 1; i <= nf; ++i) gu[i] = i;
        for (int j = 1; j <= nf; ++j) {   // rand_perm (Trace.F90:1065-1080)
            int i = randint(c.rng, j, nf);
            int t = gu[j];
            gu[j] = gu[i];
            gu[i] = t;
        }
        gu[0] = randint(c.rng, 0, 1);
    }
    for (int i = 1; i <= nf; ++i) igu[gu[i]] = i;
    igu[0] = gu[0];
    const int old_outer = c.h.outer;
    double sst_factor;
    if (spinflip && M.has_szt) {
        // perform_spinflip_check_qn (Trace.F90:1306-1339)
        int st = M.first_state[c.h.outer];
        int dn = st & ((1 << nb) - 1), up = (st >> nb) & ((1 << nb) - 1);
        int nst = up | (dn << nb);
        c.h.outer = M.state_sst[nst];
        global_update(c, gu);
        sst_factor = qn_cycle(c, c.h.outer) ? 1.0 : 0.0;
    } else {
        // perform_global_check_qn (Trace.F90:1345-1383)
        int pold = 0;
        for (int s = 0; s < M.nsst; ++s) pold += qn_cycle(c, s);
        global_update(c, gu);
        int pnew = 0;
        for (int s = 0; s < M.nsst; ++s) pnew += qn_cycle(c, s);
        if (pnew <= 0) {
            sst_factor = 0.0;
        } else {
            sst_factor = (double)pnew / (double)pold;
            int kk = randint(c.rng, 1, pnew);
            for (int s = 0; s < M.nsst; ++s)
                if (qn_cycle(c, s) && --kk <= 0) {
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
    // flavour lists of the relabelled configuration
    int kold[MAXF];
    for (int f = 0; f < nf; ++f) kold[f] = c.h.kf[f];
    rebuild_flavour_lists(c);
    // determinant ratio: det(new Delta) * det(old M) (get_InvFullHybr,
    // get_FullLogDet, Trace.F90:2279-2333, 2665-2685)
    int *piv = reinterpret_cast<int *>(c.R);
    double dnl = 0.0, dns = 1.0, dol = 0.0, dos = 1.0;
    const int ld = c.L.kcap;
    bool sing = false;
    for (int f = 0; f < nf; ++f) {
        int k = c.h.kf[f];
        if (k > 0) {
            build_delta(c, f, k, c.scratch);
            double l, s;
            if (!lu_inplace(c, c.scratch, k, ld, l, s, piv)) sing = true;
            dnl += l;
            dns *= s;
        }
        int ko = kold[f];
        if (ko == 1) {
            double v = c.mi(f)[0];
            dol += log(fabs(v));
            dos *= copysign(1.0, v);
        } else if (ko > 1) {
            for (int idx = lane_id(); idx < ko * ko; idx += WAVE) {
                int i = idx % ko, j = idx / ko;
                c.scratch[i + (long)j * ld] = c.mi(f)[i + (long)j * ld];
            }
            wsync();
            double l, s;
            lu_inplace(c, c.scratch, ko, ld, l, s, piv);
            dol += l;
            dos *= s;
        }
    }
    for (int f = 0; f < nf; ++f) {
        double k3 = (double)c.h.kf[f] * c.h.kf[f] * c.h.kf[f], ko3 = (double)kold[f] * kold[f] * kold[f];
        c.h.alg_flops += (2.0 / 3.0) * (k3 + ko3) * 2.0;
        c.h.alg_bytes += 8.0 * ((double)c.h.kf[f] * c.h.kf[f] + (double)kold[f] * kold[f]);
    }
    double tl, ts;
    full_trace(c, c.h.outer, tl, ts);
    double wr = exp((tl - c.h.trace_log) + (dnl + dol));
    double rr = grnd(c.rng);   // drawn even for a singular Delta (prob = 0)
    bool acc = !sing && rr < fabs(wr * sst_factor);
    if (acc) {
        for (int f = 0; f < nf; ++f) {
            int k = c.h.kf[f];
            if (k == 0) continue;
            build_delta(c, f, k, c.scratch);
            double l, s;
            lu_inplace(c, c.scratch, k, ld, l, s, piv);
            lu_inverse(c, c.scratch, k, ld, piv, c.mi(f));
        }
        c.h.det_log = dnl;
        c.h.det_sign = dns;
        c.h.trace_log = tl;
        c.h.trace_sign = ts;
        store_part(c, M.sizes[c.h.outer]);
        c.h.cfgsign = -config_sign(c);
        c.h.cache_pos = -1;
        c.h.acc_glob++;
        return OUT_ACC;
    }
    c.h.outer = old_outer;
    global_update(c, igu);
    rebuild_flavour_lists(c);
    return OUT_REJ;
}

// StepChangeOuter (CTQMC.F90:4569-4637) for the empty configuration
template <int NM>
DEV int step_change_outer(Ctx<NM> &c)
{
    const DevModel &M = c.M;
    // empty trace of superstate s: sum_i exp(-beta (E_i - Egs))
    double sum = 0.0;
    for (int s = 0; s < M.nsst; ++s) {
        double tl, ts;
        full_trace(c, s, tl, ts);
        sum += ts * exp(tl);
    }
    double r = grnd(c.rng), acc = 0.0;
    int pick = M.nsst - 1;
    for (int s = 0; s < M.nsst; ++s) {
        double tl, ts;
        full_trace(c, s, tl, ts);
        double v = ts * exp(tl) / sum;
        acc += v;
        if (acc >= r && v != 0.0) {
            pick = s;
            break;
        }
    }
    c.h.outer = pick;
    double tl, ts;
    full_trace(c, pick, tl, ts);
    c.h.trace_log = tl;
    c.h.trace_sign = ts;
    store_part(c, M.sizes[pick]);
    c.h.cfgsign = -config_sign(c);
    c.h.cache_pos = -1;
    return OUT_REJ;
}

// StepShiftTau_proper (CTQMC.F90:4641-4835) + update_InvFullHybrShiftTau_diag
// (Trace.F90:2355-2457)
template <int NM>
DEV int step_shift(Ctx<NM> &c)
{
    const DevModel &M = c.M;
    if (c.h.nop == 0) return step_change_outer(c);
    double dtau = grnd(c.rng) * M.beta;
    if (dtau == 0.0 || dtau == M.beta) return OUT_REJ;
    c.h.try_shift++;
    const int n = c.h.nop, nf = M.nf;
    int first = lower_bound_tau(c.tau, n, dtau);
    if (first == n) first = 0;
    int s = c.h.outer;
    for (int i = 0; i < first; ++i) s = conn(M, s, c.code[i], false);
    // rotate through the operator scratch area behind the matrix scratch
    double *opscr = c.scratch + c.L.kmat;
    double *tmp_tau = opscr;
    int *tmp_code = reinterpret_cast<int *>(opscr + n);
    for (int i = lane_id(); i < n; i += WAVE) {
        int src = (first + i) % n;
        double t = c.tau[src] - dtau;
        tmp_tau[i] = t - floor(t / M.beta) * M.beta;
        tmp_code[i] = c.code[src];
    }
    wsync();
    // keep the old configuration in the scratch area behind the new one
    double *old_tau = opscr + 2 * (long)n;
    int *old_code = reinterpret_cast<int *>(opscr + 3 * (long)n);
    for (int i = lane_id(); i < n; i += WAVE) {
        old_tau[i] = c.tau[i];
        old_code[i] = c.code[i];
        c.tau[i] = tmp_tau[i];
        c.code[i] = tmp_code[i];
    }
    wsync();
    const int old_outer = c.h.outer, old_pos = c.h.wpos;
    const double old_min = c.h.tauwin_min, old_max = c.h.tauwin_max;
    c.h.outer = s;
    if (c.h.nwin > 1) {
        int nw = c.h.nwin;
        int a = (int)floor((dtau / M.beta) * (double)(nw + 1));
        int am = ((a % nw) + nw) % nw;
        int p = old_pos - am;
        c.h.wpos = ((p % nw) + nw) % nw;
    } else {
        c.h.wpos = 0;
    }
    set_window(c);
    double tl, ts;
    full_trace(c, c.h.outer, tl, ts);
    double ratio = ts * c.h.trace_sign * exp(tl - c.h.trace_log);
    if (grnd(c.rng) < fabs(1.0 * ratio * 1.0)) {
        c.h.trace_log = tl;
        c.h.trace_sign = ts;
        store_part(c, M.sizes[c.h.outer]);
        // wrapped operators per flavour and type
        int wrapped[MAXF][2];
        for (int f = 0; f < nf; ++f) wrapped[f][0] = wrapped[f][1] = 0;
        for (int f = 0; f < nf; ++f) {
            const int k = c.h.kf[f];
            int w1 = count_below(c.fc + (long)f * c.L.kcap, k, dtau);
            int w2 = count_below(c.fa + (long)f * c.L.kcap, k, dtau);
            wrapped[f][0] = w1;
            wrapped[f][1] = w2;
        }
        double detrat = 1.0;
        const int ld = c.L.kcap;
        for (int f = 0; f < nf; ++f) {
            const int k = c.h.kf[f], w1 = wrapped[f][0], w2 = wrapped[f][1];
            double *Mf = c.mi(f);
            if (!((w1 == 0 && w2 == 0) || (w1 == k && w2 == k))) {
                for (int idx = lane_id(); idx < k * k; idx += WAVE) {
                    int i = idx % k, j = idx / k;
                    // new(i, j) from old with cyclic shift and negation of wrapped parts
                    int io = i < k - w1 ? i + w1 : i - (k - w1);
                    int jo = j < k - w2 ? j + w2 : j - (k - w2);
                    double sg = ((i >= k - w1) != (j >= k - w2)) ? -1.0 : 1.0;
                    c.scratch[idx] = sg * Mf[io + (long)jo * ld];
                }
                wsync();
                for (int idx = lane_id(); idx < k * k; idx += WAVE) {
                    int i = idx % k, j = idx / k;
                    Mf[i + (long)j * ld] = c.scratch[idx];
                }
                wsync();
            }
            if (k % 2 != 0 && (w1 + w2) % 2 != 0) detrat = -detrat;
        }
        c.h.det_sign *= detrat;
        rebuild_flavour_lists(c);
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
    return OUT_REJ;
}

// one Metropolis step with the dispatch of ctqmc_warmup / ctqmc_measure
// (CTQMC.F90:8213-8254, 8432-8533) and the window-shift rule
template <int NM>
DEV void one_move(Ctx<NM> &c, int &kind, int &out)
{
    double r = grnd(c.rng);
    const DevModel &M = c.M;
    if (r < (1.0 - M.p_shift - M.p_glob) / 2.0) {
        kind = 1;
        (void)grnd(c.rng);               // Percentage4OperatorMove draw
        (void)(grnd(c.rng) < M.p_outer); // PercentageOuterMove draw
        out = step_add(c);
    } else if (r <= 1.0 - M.p_shift - M.p_glob) {
        kind = 2;
        (void)grnd(c.rng);
        (void)(grnd(c.rng) < M.p_outer);
        out = step_rem(c);
    } else if (r <= 1.0 - M.p_shift) {
        kind = 3;
        out = step_glob(c);
    } else {
        kind = 4;
        out = step_shift(c);
    }
    c.h.moves++;
    if ((c.h.try_add + c.h.try_rem) % M.nslide == 0) shift_window(c);
}

// ---------------------------------------------------------------------------
// measurements
// ---------------------------------------------------------------------------
// MeasGtauRem (CTQMC.F90:2252-2440): lanes over (annihilator, creator) pairs
template <int NM>
DEV void meas_gtau(Ctx<NM> &c, double *gtau, double *gleg)
{
    const DevModel &M = c.M;
    const int ld = c.L.kcap;
    for (int f = 0; f < M.nf; ++f) {
        const int k = c.h.kf[f];
        const double *ctl = c.fc + (long)f * c.L.kcap, *atl = c.fa + (long)f * c.L.kcap;
        const double *Mf = c.mi(f);
        c.h.alg_flops += (double)k * k * (5.0 * M.nleg + 4.0);
        c.h.alg_bytes += 8.0 * k * k;
        for (int p0 = 0; p0 < k * k; p0 += WAVE) {
            int p = p0 + lane_id();
            bool act = p < k * k;
            double Mv = 0.0, x = 0.0;
            int idt = 0;
            if (act) {
                int ai = p / k, ci = p % k;    // annihilator rank, creator rank
                double ta = atl[ai], tc = ctl[ci];
                Mv = Mf[ci + (long)ai * ld];
                if (ta - tc > 0.0) {
                    idt = (int)((ta - tc) / M.beta * (double)M.ntau) + 1;
                    x = 2.0 / M.beta * (ta - tc) - 1.0;
                } else {
                    idt = (int)((M.beta + ta - tc) / M.beta * (double)M.ntau) + 1;
                    x = 2.0 / M.beta * (M.beta + ta - tc) - 1.0;
                    Mv = -Mv;
                }
                Mv = Mv * c.h.cfgsign;
                if (idt > M.ntau) idt = M.ntau;
                atomicAdd(&gtau[(long)f * M.ntau + idt - 1], Mv);
            }
            // Legendre: Bonnet recursion per lane, wave sums per order
            double pl = 0.0, plm1 = 0.0, plp1 = 0.0;
            for (int l = 0; l < M.nleg; ++l) {
                if (l == 0) plp1 = 1.0;
                else if (l == 1) plp1 = x;
                else {
                    int ll = l - 1;
                    plp1 = (2.0 * ll + 1.0) / (double)(ll + 1) * x * pl - (double)ll / (double)(ll + 1) * plm1;
                }
                double v = wave_sum(act ? Mv * plp1 : 0.0);
                if (lane_id() == 0) gleg[(long)f * M.nleg + l] += v;
                plm1 = pl;
                pl = plp1;
            }
        }
    }
    wsync();
}

// MeasDensityMatrix_beta_half (CTQMC.F90:3088-3151) with update_b2_states_eb
// (Trace.F90:6213-6469): ket from 0 to beta/2, bra from beta to beta/2
template <int NM>
DEV void meas_dm(Ctx<NM> &c, double *dm)
{
    const DevModel &M = c.M;
    const int outer = c.h.outer, no = M.sizes[outer];
    if (c.h.nop != 0) {
        double bh = M.beta / 2.0;
        int nk = upper_bound_tau(c.tau, c.h.nop, bh);   // ops with tau <= beta/2
        double *K = c.K0, *Kw = c.K1;
        set_identity(c, K, no);
        int s = outer;
        double kl = 0.0;
        ket_sweep(c, K, Kw, s, kl, no, 0.0, c.tau, c.code, nk, bh);
        // keep K in scratch while the bra is built
        double *Ks = c.scratch;
        copy_mat<NM>(Ks, K, M.sizes[s] * no);
        double *B = c.K0, *Bw = c.K1;
        set_identity(c, B, no);
        int sb = outer;
        double bl = 0.0;
        bra_sweep(c, B, Bw, sb, bl, no, M.beta, c.tau + nk, c.code + nk, c.h.nop - nk, bh);
        const int n = M.sizes[s];
        double *blk = dm + M.dmoff[s];
        const double fac = c.h.trace_sign * (-c.h.cfgsign) * exp(kl + bl - c.h.trace_log);
        // rho(s1, s2) += sum_i bra(s1, i) ket(s2, i) * fac
        for (int idx = lane_id(); idx < n * n; idx += WAVE) {
            int s1 = idx % n, s2 = idx / n;
            double acc = 0.0;
            for (int i = 0; i < no; ++i) acc = fma(B[s1 * no + i], Ks[s2 * no + i], acc);
            blk[s1 + (long)s2 * n] += fac * acc;
        }
        wsync();
    } else {
        double *blk = dm + M.dmoff[outer];
        for (int i = lane_id(); i < no; i += WAVE) {
            double v = c.part[2 * i + 1] * c.h.trace_sign * exp(c.part[2 * i] - c.h.trace_log);
            blk[i + (long)i * no] -= c.h.cfgsign * v;
        }
        wsync();
    }
}

// MeasSign, MeasOuterHistograms (visits), MeasExpOrder (CTQMC.F90:3264-3350)
template <int NM>
DEV void meas_misc(Ctx<NM> &c, double *histo, double *outer_hist)
{
    const DevModel &M = c.M;
    c.h.meas_sign += c.h.cfgsign;
    if (lane_id() == 0) {
        outer_hist[c.h.outer] += 1.0;
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
constexpr int KC_LDS = 512;   // largest per-flavour order held in LDS vectors

template <int NM>
struct Lds {
    uint32_t mt[624];
    double dvec[2 * KC_LDS + 4];  // exp factors / det row+col
    double Q[KC_LDS], R[KC_LDS];
    double pl[2 * (NM > 128 ? 128 : NM) + 2];
    double K[NM <= 16 ? 2 * NM * NM : 1];
};

template <int NM>
DEV void setup_ctx(Ctx<NM> &c, Lds<NM> &s)
{
    const ChainLayout &L = c.L;
    char *b = c.base;
    c.h = *reinterpret_cast<ChainHdr *>(b);
    c.tau = reinterpret_cast<double *>(b + L.tau);
    c.code = reinterpret_cast<int *>(b + L.code);
    c.fc = reinterpret_cast<double *>(b + L.fc);
    c.fa = reinterpret_cast<double *>(b + L.fa);
    c.minv = reinterpret_cast<double *>(b + L.minv);
    c.scratch = reinterpret_cast<double *>(b + L.scratch);
    c.kst = reinterpret_cast<double *>(b + L.kst);
    c.bnd = reinterpret_cast<double *>(b + L.bnd);
    c.bndlog = reinterpret_cast<double *>(b + L.bndlog);
    c.bndsst = reinterpret_cast<int *>(b + L.bndsst);
    c.part = reinterpret_cast<double *>(b + L.part);
    c.ltau = reinterpret_cast<double *>(b + L.ltau);
    c.lcode = reinterpret_cast<int *>(b + L.lcode);
    if (NM <= 16) {
        c.K0 = s.K;
        c.K1 = s.K + NM * NM;
    } else {
        c.K0 = reinterpret_cast<double *>(b + L.kwork);
        c.K1 = c.K0 + (long)NM * NM;
    }
    c.dvec = s.dvec;
    c.Q = s.Q;
    c.R = s.R;
    c.pl = s.pl;
    const uint32_t *gmt = reinterpret_cast<const uint32_t *>(b + L.mt);
    for (int i = lane_id(); i < 624; i += WAVE) s.mt[i] = gmt[i];
    c.rng.mt = s.mt;
    c.rng.idx = c.h.mti;
    wsync();
}

template <int NM>
DEV void save_ctx(Ctx<NM> &c, Lds<NM> &s)
{
    char *b = c.base;
    uint32_t *gmt = reinterpret_cast<uint32_t *>(b + c.L.mt);
    for (int i = lane_id(); i < 624; i += WAVE) gmt[i] = s.mt[i];
    c.h.mti = c.rng.idx;
    wsync();
    if (lane_id() == 0) *reinterpret_cast<ChainHdr *>(b) = c.h;
    wsync();
}

// Sweep kernel: each block advances chain blockIdx.x by `nmoves` steps; if
// ncorr > 0 it measures after every ncorr steps.  rec != nullptr records
// every step of chain `only` (parity mode; other chains return at once).
template <int NM>
__global__ void __launch_bounds__(WAVE) sweep_kernel(DevModel M, ChainLayout L, char *chains, long long nmoves,
                                                     long long ncorr, int only, w2d_move_record *rec)
{
    const int chain = blockIdx.x;
    if (only >= 0 && chain != only) return;
    __shared__ Lds<NM> s;
    Ctx<NM> c(M, L, chains + (long)chain * L.stride);
    setup_ctx(c, s);
    double *gtau = reinterpret_cast<double *>(c.base + L.gtau);
    double *gleg = reinterpret_cast<double *>(c.base + L.gleg);
    double *dm = reinterpret_cast<double *>(c.base + L.dm);
    double *histo = reinterpret_cast<double *>(c.base + L.histo);
    double *ohist = reinterpret_cast<double *>(c.base + L.outer);
    for (long long m = 0; m < nmoves && c.h.err == 0; ++m) {
        int kind, out;
        one_move(c, kind, out);
        if (rec && lane_id() == 0) {
            w2d_move_record &r = rec[m];
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
        if (ncorr > 0 && (m + 1) % ncorr == 0) {
            c.h.cnt_meas++;
            meas_misc(c, histo, ohist);
            if (M.meas_dm) meas_dm(c, dm);
            meas_gtau(c, gtau, gleg);
        }
    }
    save_ctx(c, s);
}

// Initial state of every chain: empty configuration, initial outer superstate
// drawn proportional to the empty traces (init_Trace, Trace.F90:640-670).
template <int NM>
__global__ void __launch_bounds__(WAVE) init_kernel(DevModel M, ChainLayout L, char *chains)
{
    const int chain = blockIdx.x;
    __shared__ Lds<NM> s;
    Ctx<NM> c(M, L, chains + (long)chain * L.stride);
    setup_ctx(c, s);
    double sum = 0.0;
    for (int q = 0; q < M.nsst; ++q) {
        double tl, ts;
        full_trace(c, q, tl, ts);
        sum += ts * exp(tl);
    }
    double r = grnd(c.rng), acc = 0.0;
    int pick = M.nsst - 1;
    for (int q = 0; q < M.nsst; ++q) {
        double tl, ts;
        full_trace(c, q, tl, ts);
        double v = ts * exp(tl) / sum;
        acc += v;
        if (acc >= r && v != 0.0) {
            pick = q;
            break;
        }
    }
    c.h.outer = pick;
    double tl, ts;
    full_trace(c, pick, tl, ts);
    c.h.trace_log = tl;
    c.h.trace_sign = ts;
    store_part(c, M.sizes[pick]);
    if (!(tl > -DBL_MAX)) c.h.err = 5;
    c.h.cache_pos = -1;
    save_ctx(c, s);
}

// Adopt the configuration already written into tau/code/nop/outer: full trace,
// flavour lists, full inverses, sign (set_mc_config_from_arrays,
// Trace.F90:6872-7000).
template <int NM>
__global__ void __launch_bounds__(WAVE) setconfig_kernel(DevModel M, ChainLayout L, char *chains, int only)
{
    const int chain = blockIdx.x;
    if (chain != only) return;
    __shared__ Lds<NM> s;
    Ctx<NM> c(M, L, chains + (long)chain * L.stride);
    setup_ctx(c, s);
    c.h.wpos = 0;
    set_window(c);
    double tl, ts;
    full_trace(c, c.h.outer, tl, ts);
    c.h.trace_log = tl;
    c.h.trace_sign = ts;
    store_part(c, M.sizes[c.h.outer]);
    rebuild_flavour_lists(c);
    int *piv = reinterpret_cast<int *>(c.R);
    double dl = 0.0, dsg = 1.0;
    for (int f = 0; f < M.nf; ++f) {
        int k = c.h.kf[f];
        if (k == 0) continue;
        build_delta(c, f, k, c.scratch);
        double l, sg;
        if (!lu_inplace(c, c.scratch, k, L.kcap, l, sg, piv)) {
            c.h.err = ERR_SING;
            break;
        }
        lu_inverse(c, c.scratch, k, L.kcap, piv, c.mi(f));
        dl += l;
        dsg *= sg;
    }
    c.h.det_log = dl;
    c.h.det_sign = dsg;
    c.h.cfgsign = -config_sign(c);
    c.h.cache_pos = -1;
    save_ctx(c, s);
}

}  // namespace w2d

// ===========================================================================
// host side: the C-ABI (include/w2dyn_ctqmc.h)
// ===========================================================================
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "model.hpp"

struct w2d_ctqmc {
    std::string params;
    w2d::LocalModel model;
    w2d::DevModel dm{};
    w2d::ChainLayout L{};
    int nchains = 0, nm = 0, device = 0, kcap = 0, ntau = 0, nleg = 0, maxhisto = 0;
    int nwin_alloc = 0;
    uint32_t seed0 = 0;
    double beta = 0.0, taudiff_param = -1.0;
    long long ncorr = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double last_ms = 0.0;
    long long last_launches = 0;
    std::vector<void *> dev_allocs;
    char *chains = nullptr;
    bool measured = false;
};

namespace {

using w2d::ChainHdr;

#define HIPCHK(x)                                                                     \
    do {                                                                              \
        hipError_t e_ = (x);                                                          \
        if (e_ != hipSuccess) {                                                       \
            std::fprintf(stderr, "w2d_ctqmc: HIP error %s at %s:%d\n", hipGetErrorString(e_), \
                         __FILE__, __LINE__);                                         \
            return W2D_ERR_HIP;                                                       \
        }                                                                             \
    } while (0)

template <typename T>
int upload(w2d_ctqmc *h, const std::vector<T> &v, const T **out)
{
    void *p = nullptr;
    size_t n = std::max<size_t>(v.size(), 1) * sizeof(T);
    HIPCHK(hipMalloc(&p, n));
    h->dev_allocs.push_back(p);
    if (!v.empty()) HIPCHK(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    *out = static_cast<const T *>(p);
    return W2D_OK;
}

long align256(long x) { return (x + 255) / 256 * 256; }

void seed_mt(uint32_t *mt, uint32_t s)
{
    mt[0] = s;
    for (int i = 1; i < 624; ++i) mt[i] = 1812433253u * (mt[i - 1] ^ (mt[i - 1] >> 30)) + (uint32_t)i;
}

// init_window (Trace.F90:1608-1646) on a host copy of the header
void init_window(ChainHdr &hd, double tdm, double beta)
{
    hd.wpos = 0;
    if (tdm > 0.0) {
        hd.taudiff_max = tdm;
        int nw = (int)std::floor(beta / tdm) - 1;
        hd.nwin = nw > 1 ? nw : 1;
        hd.tauwin_min = ((double)hd.wpos / (double)(hd.nwin + 1)) * beta;
        hd.tauwin_max = ((double)(hd.wpos + 2) / (double)(hd.nwin + 1)) * beta;
    } else {
        hd.taudiff_max = beta;
        hd.nwin = 1;
        hd.tauwin_min = 0.0;
        hd.tauwin_max = beta;
    }
    hd.cache_pos = -1;
}

int nbound_for(double tdm, double beta)
{
    if (tdm <= 0.0) return 3;
    int nw = (int)std::floor(beta / tdm) - 1;
    return std::max(nw, 1) + 2;
}

int layout_and_alloc(w2d_ctqmc *h, int nbound)
{
    using namespace w2d;
    ChainLayout &L = h->L;
    const int nf = 2 * h->model.nbands, nm = h->nm, kc = h->kcap;
    L.kcap = kc;
    L.cap = 2 * nf * kc;
    L.lcap = L.cap + 2;
    L.nbound = nbound;
    L.kmat = std::max<long>((long)kc * kc, (long)nm * nm);
    long o = align256(sizeof(ChainHdr));
    auto take = [&](long bytes) {
        long at = o;
        o = align256(o + bytes);
        return at;
    };
    L.mt = take(624 * 4);
    L.tau = take((long)L.cap * 8);
    L.code = take((long)L.cap * 4);
    L.fc = take((long)nf * kc * 8);
    L.fa = take((long)nf * kc * 8);
    L.minv = take((long)nf * kc * kc * 8);
    L.scratch = take((L.kmat + 4L * L.cap) * 8);
    L.kst = take((long)nm * nm * 8);
    L.bnd = take((long)nbound * nm * nm * 8);
    L.bndlog = take((long)nbound * 8);
    L.bndsst = take((long)nbound * 4);
    L.part = take(2L * nm * 8 + 16);
    L.kwork = take(nm > 16 ? 2L * nm * nm * 8 : 8);
    L.ltau = take((long)L.lcap * 8);
    L.lcode = take((long)L.lcap * 4);
    L.gtau = take((long)nf * h->ntau * 8);
    L.gleg = take((long)nf * h->nleg * 8);
    L.dm = take(h->dm.dmsize * 8);
    L.histo = take((long)nf * (h->maxhisto + 1) * 8);
    L.outer = take((long)h->model.nsst * 8);
    L.stride = o;
    if (h->chains) {
        HIPCHK(hipFree(h->chains));
        h->chains = nullptr;
    }
    HIPCHK(hipMalloc(&h->chains, (size_t)L.stride * h->nchains));
    HIPCHK(hipMemset(h->chains, 0, (size_t)L.stride * h->nchains));
    h->nwin_alloc = nbound;
    return W2D_OK;
}

template <int NM>
void launch_init(w2d_ctqmc *h)
{
    hipLaunchKernelGGL(w2d::init_kernel<NM>, dim3(h->nchains), dim3(w2d::WAVE), 0, h->stream, h->dm, h->L,
                       h->chains);
}

template <int NM>
void launch_sweep(w2d_ctqmc *h, long long nmoves, long long ncorr, int only, w2d_move_record *rec)
{
    hipLaunchKernelGGL(w2d::sweep_kernel<NM>, dim3(h->nchains), dim3(w2d::WAVE), 0, h->stream, h->dm, h->L,
                       h->chains, nmoves, ncorr, only, rec);
}

template <int NM>
void launch_setconfig(w2d_ctqmc *h, int chain)
{
    hipLaunchKernelGGL(w2d::setconfig_kernel<NM>, dim3(h->nchains), dim3(w2d::WAVE), 0, h->stream, h->dm,
                       h->L, h->chains, chain);
}

#define DISPATCH(fn, ...)                    \
    switch (h->nm) {                         \
    case 1: fn<1>(__VA_ARGS__); break;       \
    case 4: fn<4>(__VA_ARGS__); break;       \
    case 16: fn<16>(__VA_ARGS__); break;     \
    default: fn<128>(__VA_ARGS__); break;    \
    }

int read_headers(w2d_ctqmc *h, std::vector<ChainHdr> &hd)
{
    hd.resize(h->nchains);
    HIPCHK(hipStreamSynchronize(h->stream));
    for (int c = 0; c < h->nchains; ++c)
        HIPCHK(hipMemcpy(&hd[c], h->chains + (long)c * h->L.stride, sizeof(ChainHdr), hipMemcpyDeviceToHost));
    return W2D_OK;
}

int write_headers(w2d_ctqmc *h, const std::vector<ChainHdr> &hd)
{
    for (int c = 0; c < h->nchains; ++c)
        HIPCHK(hipMemcpy(h->chains + (long)c * h->L.stride, &hd[c], sizeof(ChainHdr), hipMemcpyHostToDevice));
    return W2D_OK;
}

int first_error(const std::vector<ChainHdr> &hd)
{
    for (const auto &x : hd)
        if (x.err) return x.err;
    return W2D_OK;
}

// (re)initialise every chain: seeded MT state, empty configuration, window from
// the TaudiffMax parameter, then the initial outer-superstate draw on the GPU
int init_chains(w2d_ctqmc *h)
{
    std::vector<uint32_t> mt(624);
    for (int c = 0; c < h->nchains; ++c) {
        ChainHdr hd;
        std::memset(&hd, 0, sizeof(hd));
        hd.det_sign = 1.0;
        hd.trace_sign = 1.0;
        hd.cfgsign = -1.0;
        hd.mti = 624;
        init_window(hd, h->taudiff_param, h->beta);
        char *base = h->chains + (long)c * h->L.stride;
        HIPCHK(hipMemcpy(base, &hd, sizeof(hd), hipMemcpyHostToDevice));
        seed_mt(mt.data(), h->seed0 + (uint32_t)c);
        HIPCHK(hipMemcpy(base + h->L.mt, mt.data(), 624 * 4, hipMemcpyHostToDevice));
    }
    DISPATCH(launch_init, h);
    HIPCHK(hipGetLastError());
    std::vector<ChainHdr> hd;
    int st = read_headers(h, hd);
    if (st) return st;
    return first_error(hd) ? W2D_ERR_ZERO_TRACE : W2D_OK;
}

void reset_counters(ChainHdr &x)
{
    x.try_add = x.try_rem = x.try_glob = x.try_shift = 0;
    x.acc_add = x.acc_rem = x.acc_glob = x.acc_shift = 0;
    x.qn_rej = x.moves = 0;
    x.alg_flops = x.alg_bytes = 0.0;
}

int map_err(int e)
{
    switch (e) {
    case 0: return W2D_OK;
    case 3: return W2D_ERR_CAPACITY;
    case 4: return W2D_ERR_SINGULAR;
    case 5: return W2D_ERR_ZERO_TRACE;
    default: return W2D_ERR_STATE;
    }
}

// run nmoves steps per chain in launches of whole measurement intervals
int run_moves(w2d_ctqmc *h, long long nmoves, long long ncorr)
{
    long long per = ncorr > 0 ? ncorr * std::max<long long>(1, 4000 / ncorr) : 4000;
    HIPCHK(hipEventRecord(h->ev0, h->stream));
    h->last_launches = 0;
    for (long long done = 0; done < nmoves; done += per, ++h->last_launches) {
        long long n = std::min(per, nmoves - done);
        DISPATCH(launch_sweep, h, n, ncorr, -1, nullptr);
        HIPCHK(hipGetLastError());
    }
    HIPCHK(hipEventRecord(h->ev1, h->stream));
    HIPCHK(hipEventSynchronize(h->ev1));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, h->ev0, h->ev1));
    h->last_ms = ms;
    std::vector<ChainHdr> hd;
    int st = read_headers(h, hd);
    if (st) return st;
    return map_err(first_error(hd));
}

}  // namespace

extern "C" {

const char *w2d_ctqmc_version(void) { return "w2dynamics_amd ctqmc 0.1 (gfx950)"; }

int w2d_ctqmc_create(const char *params, int nbands, int nftau, const double *u_matrix,
                     const double *ftau_full, const double *muimp_full, int nchains, uint32_t seed0,
                     int capacity, int device, w2d_ctqmc **out)
{
    if (!params || !u_matrix || !ftau_full || !muimp_full || !out || nbands < 1 || nbands > 7 || nftau < 2 ||
        nchains < 1)
        return W2D_ERR_ARG;
    *out = nullptr;
    w2d::ParamList P(params);
    // unsupported branches of the reference (kept out of this path)
    if (P.integer("segment") != 0 || P.integer("offdiag") != 0 || P.integer("statesampling") != 0 ||
        P.integer("Uw") == 1 || P.real("PercentageWormInsert") != 0.0 || P.real("Percentage4OperatorMove") != 0.0 ||
        P.real("PercentageOuterMove") != 0.0 || P.str("EnableOuterTruncation") == "YES" ||
        P.integer("NSymMove") > 0 || P.integer("Eigenbasis") == 0)
        return W2D_ERR_ARG;
    if (P.integer("Nd") != nbands) return W2D_ERR_ARG;
    auto h = std::make_unique<w2d_ctqmc>();
    h->params = params;
    h->nchains = nchains;
    h->seed0 = seed0;
    h->device = device;
    h->kcap = capacity > 0 ? capacity : 256;
    if (h->kcap >= w2d::KC_LDS) return W2D_ERR_ARG;
    h->beta = P.real("beta");
    h->ntau = (int)P.integer("Ntau");
    h->nleg = (int)P.integer("NLegMax");
    h->maxhisto = (int)P.integer("MaxHisto");
    h->ncorr = P.integer("NCorr");
    h->taudiff_param = P.real("TaudiffMax");
    if (h->beta <= 0 || h->ntau < 1 || h->nleg < 1 || h->maxhisto < 0 || h->ncorr < 1) return W2D_ERR_ARG;
    std::string err = w2d::build_local_model(nbands, u_matrix, muimp_full, P.str("QuantumNumbers"), h->model);
    if (!err.empty()) {
        std::fprintf(stderr, "w2d_ctqmc: %s\n", err.c_str());
        return W2D_ERR_ARG;
    }
    const w2d::LocalModel &m = h->model;
    if (m.nmax <= 1) h->nm = 1;
    else if (m.nmax <= 4) h->nm = 4;
    else if (m.nmax <= 16) h->nm = 16;
    else if (m.nmax <= 128) h->nm = 128;
    else return W2D_ERR_ARG;
    HIPCHK(hipSetDevice(device));
    HIPCHK(hipStreamCreate(&h->stream));
    HIPCHK(hipEventCreate(&h->ev0));
    HIPCHK(hipEventCreate(&h->ev1));
    // device model
    w2d::DevModel &d = h->dm;
    const int nf = 2 * nbands;
    d.nb = nbands;
    d.nf = nf;
    d.nsst = m.nsst;
    d.nmax = m.nmax;
    d.nstates = m.nstates;
    d.nftau = nftau;
    d.beta = h->beta;
    d.egs = m.egs;
    std::vector<double> fdiag((size_t)nf * nftau);
    for (int b = 0; b < nbands; ++b)
        for (int s = 0; s < 2; ++s) {
            long base = b + 2L * nbands * 0 + 0;
            (void)base;
            for (int t = 0; t < nftau; ++t) {
                // ftau_full(b, s, b, s, t) in Fortran order (N,2,N,2,nftau)
                long idx = b + (long)nbands * (s + 2L * (b + (long)nbands * (s + 2L * t)));
                fdiag[(size_t)(b + nbands * s) * nftau + t] = ftau_full[idx];
            }
        }
    std::vector<long> dmoff(m.nsst);
    long dms = 0;
    for (int s = 0; s < m.nsst; ++s) {
        dmoff[s] = dms;
        dms += (long)m.sizes[s] * m.sizes[s];
    }
    d.dmsize = dms;
    std::vector<int> first_state(m.nsst);
    for (int s = 0; s < m.nsst; ++s) first_state[s] = m.states[m.offsets[s]];
    int st = 0;
    if ((st = upload(h.get(), m.sizes, &d.sizes)) || (st = upload(h.get(), m.offsets, &d.offsets)) ||
        (st = upload(h.get(), m.connect, &d.connect)) || (st = upload(h.get(), first_state, &d.first_state)) ||
        (st = upload(h.get(), m.state_sst, &d.state_sst)) || (st = upload(h.get(), m.opoff, &d.opoff)) ||
        (st = upload(h.get(), m.ops, &d.ops)) || (st = upload(h.get(), m.evals, &d.evals)) ||
        (st = upload(h.get(), fdiag, &d.ftau)) || (st = upload(h.get(), dmoff, &d.dmoff)))
        return st;
    d.p_shift = P.real("PercentageTauShiftMove");
    d.p_glob = P.real("PercentageGlobalMove");
    d.p_outer = 0.0;
    d.p_4op = 0.0;
    d.ntau = h->ntau;
    d.nleg = h->nleg;
    d.maxhisto = h->maxhisto;
    d.has_szt = P.str("QuantumNumbers").find("Szt") != std::string::npos;
    d.meas_dm = P.integer("MeasDensityMatrix") != 0;
    // NSlide = max(2*NBands, floor(NCorr/num_windows)) with the window of the
    // parameter file (CTQMC.F90:8069)
    int nw0 = nbound_for(h->taudiff_param, h->beta) - 2;
    d.nslide = std::max<long long>(2LL * nbands, (long long)std::floor((double)h->ncorr / (double)nw0));
    if ((st = layout_and_alloc(h.get(), nbound_for(h->taudiff_param, h->beta)))) return st;
    if ((st = init_chains(h.get()))) return st;
    *out = h.release();
    return W2D_OK;
}

int w2d_ctqmc_set_taudiffmax(w2d_ctqmc *h, double tdm)
{
    if (!h) return W2D_ERR_ARG;
    if (!(tdm > 0.0)) return W2D_OK;   // set_taudiffmax with a non-positive value keeps one window
    int nb = nbound_for(tdm, h->beta);
    if (nb > h->nwin_alloc) {
        // re-create the chains' memory; the chains are still in their initial
        // state, which is re-derived from the seeds
        int st = layout_and_alloc(h, nb);
        if (st) return st;
        if ((st = init_chains(h))) return st;
    }
    std::vector<ChainHdr> hd;
    int st = read_headers(h, hd);
    if (st) return st;
    for (auto &x : hd) init_window(x, tdm, h->beta);
    return write_headers(h, hd);
}

int w2d_ctqmc_warmup(w2d_ctqmc *h, int64_t nwarmups)
{
    if (!h || nwarmups < 0) return W2D_ERR_ARG;
    std::vector<ChainHdr> hd;
    int st = read_headers(h, hd);
    if (st) return st;
    for (auto &x : hd) reset_counters(x);
    if ((st = write_headers(h, hd))) return st;
    return run_moves(h, nwarmups, 0);
}

int w2d_ctqmc_measure(w2d_ctqmc *h, int64_t nmeas)
{
    if (!h || nmeas < 0) return W2D_ERR_ARG;
    std::vector<ChainHdr> hd;
    int st = read_headers(h, hd);
    if (st) return st;
    for (auto &x : hd) {
        reset_counters(x);
        x.meas_sign = 0.0;
        x.cnt_meas = 0;
    }
    if ((st = write_headers(h, hd))) return st;
    // clear the accumulators of every chain
    const w2d::ChainLayout &L = h->L;
    for (int c = 0; c < h->nchains; ++c)
        HIPCHK(hipMemsetAsync(h->chains + (long)c * L.stride + L.gtau, 0, L.stride - L.gtau, h->stream));
    st = run_moves(h, nmeas * h->ncorr, h->ncorr);
    h->measured = (st == W2D_OK);
    return st;
}

int w2d_ctqmc_dims(const w2d_ctqmc *h, int *nbands, int *nchains, int *ntau, int *nlegmax, int *nsst,
                   int *nstates, int *maxhisto)
{
    if (!h) return W2D_ERR_ARG;
    if (nbands) *nbands = h->model.nbands;
    if (nchains) *nchains = h->nchains;
    if (ntau) *ntau = h->ntau;
    if (nlegmax) *nlegmax = h->nleg;
    if (nsst) *nsst = h->model.nsst;
    if (nstates) *nstates = h->model.nstates;
    if (maxhisto) *maxhisto = h->maxhisto;
    return W2D_OK;
}

int w2d_ctqmc_get_accumulators(w2d_ctqmc *h, double *gtau, double *gleg, double *dm, double *histo,
                               double *outer, double *scalars)
{
    if (!h) return W2D_ERR_ARG;
    if (!h->measured) return W2D_ERR_STATE;
    const w2d::ChainLayout &L = h->L;
    const w2d::LocalModel &m = h->model;
    const int nf = 2 * m.nbands;
    HIPCHK(hipStreamSynchronize(h->stream));
    std::vector<double> blk(h->dm.dmsize);
    std::vector<ChainHdr> hd;
    int st = read_headers(h, hd);
    if (st) return st;
    for (int c = 0; c < h->nchains; ++c) {
        char *base = h->chains + (long)c * L.stride;
        if (gtau)
            HIPCHK(hipMemcpy(gtau + (long)c * nf * h->ntau, base + L.gtau, (size_t)nf * h->ntau * 8,
                             hipMemcpyDeviceToHost));
        if (gleg)
            HIPCHK(hipMemcpy(gleg + (long)c * nf * h->nleg, base + L.gleg, (size_t)nf * h->nleg * 8,
                             hipMemcpyDeviceToHost));
        if (histo)
            HIPCHK(hipMemcpy(histo + (long)c * nf * (h->maxhisto + 1), base + L.histo,
                             (size_t)nf * (h->maxhisto + 1) * 8, hipMemcpyDeviceToHost));
        if (outer)
            HIPCHK(hipMemcpy(outer + (long)c * m.nsst, base + L.outer, (size_t)m.nsst * 8, hipMemcpyDeviceToHost));
        if (dm) {
            HIPCHK(hipMemcpy(blk.data(), base + L.dm, blk.size() * 8, hipMemcpyDeviceToHost));
            double *D = dm + (long)c * m.nstates * m.nstates;
            std::fill(D, D + (long)m.nstates * m.nstates, 0.0);
            long o = 0;
            for (int s = 0; s < m.nsst; ++s) {
                int n = m.sizes[s], off = m.offsets[s];
                for (int j = 0; j < n; ++j)
                    for (int i = 0; i < n; ++i) D[(off + i) + (long)m.nstates * (off + j)] = blk[o + i + (long)j * n];
                o += (long)n * n;
            }
        }
        if (scalars) {
            double *S = scalars + 14L * c;
            const ChainHdr &x = hd[c];
            S[0] = x.meas_sign;
            S[1] = (double)x.cnt_meas;
            S[2] = (double)x.acc_add, S[3] = (double)x.try_add;
            S[4] = (double)x.acc_rem, S[5] = (double)x.try_rem;
            S[6] = (double)x.acc_glob, S[7] = (double)x.try_glob;
            S[8] = (double)x.acc_shift, S[9] = (double)x.try_shift;
            S[10] = (double)x.moves, S[11] = (double)x.qn_rej;
            S[12] = x.alg_flops, S[13] = x.alg_bytes;
        }
    }
    return W2D_OK;
}

int w2d_ctqmc_get_model(const w2d_ctqmc *h, int *sizes, int *offsets, double *evals, int *states,
                        double *evecs, double *egs)
{
    if (!h) return W2D_ERR_ARG;
    const w2d::LocalModel &m = h->model;
    if (sizes) std::copy(m.sizes.begin(), m.sizes.end(), sizes);
    if (offsets) std::copy(m.offsets.begin(), m.offsets.end(), offsets);
    if (evals) std::copy(m.evals.begin(), m.evals.end(), evals);
    if (states) std::copy(m.states.begin(), m.states.end(), states);
    if (evecs) std::copy(m.evecs.begin(), m.evecs.end(), evecs);
    if (egs) *egs = m.egs;
    return W2D_OK;
}

int w2d_ctqmc_moves_log(w2d_ctqmc *h, int chain, int64_t n, w2d_move_record *rec)
{
    if (!h || chain < 0 || chain >= h->nchains || n < 0 || (!rec && n > 0)) return W2D_ERR_ARG;
    const long long per = 4000;
    w2d_move_record *d = nullptr;
    HIPCHK(hipMalloc(&d, sizeof(w2d_move_record) * per));
    int st = W2D_OK;
    for (long long done = 0; done < n && st == W2D_OK; done += per) {
        long long k = std::min<long long>(per, n - done);
        DISPATCH(launch_sweep, h, k, 0, chain, d);
        if (hipGetLastError() != hipSuccess || hipStreamSynchronize(h->stream) != hipSuccess ||
            hipMemcpy(rec + done, d, sizeof(w2d_move_record) * k, hipMemcpyDeviceToHost) != hipSuccess)
            st = W2D_ERR_HIP;
    }
    (void)hipFree(d);
    if (st) return st;
    std::vector<ChainHdr> hd;
    st = read_headers(h, hd);
    if (st) return st;
    return map_err(hd[chain].err);
}

int w2d_ctqmc_get_config(w2d_ctqmc *h, int chain, int cap, int *n, double *tau, int *orb, int *spin,
                         int *ca, int *outer_sst, double *weights)
{
    if (!h || chain < 0 || chain >= h->nchains || !n) return W2D_ERR_ARG;
    HIPCHK(hipStreamSynchronize(h->stream));
    char *base = h->chains + (long)chain * h->L.stride;
    ChainHdr x;
    HIPCHK(hipMemcpy(&x, base, sizeof(x), hipMemcpyDeviceToHost));
    *n = x.nop;
    if (outer_sst) *outer_sst = x.outer;
    if (weights) {
        weights[0] = x.trace_log, weights[1] = x.trace_sign;
        weights[2] = x.det_log, weights[3] = x.det_sign, weights[4] = x.cfgsign;
    }
    int m = std::min(cap, x.nop);
    if (m > 0 && tau && orb && spin && ca) {
        std::vector<int> code(m);
        HIPCHK(hipMemcpy(tau, base + h->L.tau, (size_t)m * 8, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(code.data(), base + h->L.code, (size_t)m * 4, hipMemcpyDeviceToHost));
        for (int i = 0; i < m; ++i) {
            orb[i] = code[i] & 15;
            spin[i] = (code[i] >> 4) & 1;
            ca[i] = ((code[i] >> 5) & 1) + 1;
        }
    }
    return W2D_OK;
}

int w2d_ctqmc_set_config(w2d_ctqmc *h, int chain, int n, const double *tau, const int *orb, const int *spin,
                         const int *ca, int outer_sst, double *weights)
{
    if (!h || chain < 0 || chain >= h->nchains || n < 0 || n > h->L.cap || outer_sst < 0 ||
        outer_sst >= h->model.nsst)
        return W2D_ERR_ARG;
    std::vector<int> cnt(2 * h->model.nbands * 2, 0);
    std::vector<int> code(n);
    for (int i = 0; i < n; ++i) {
        if (orb[i] < 0 || orb[i] >= h->model.nbands || spin[i] < 0 || spin[i] > 1 || ca[i] < 1 || ca[i] > 2)
            return W2D_ERR_ARG;
        if (i > 0 && !(tau[i] > tau[i - 1])) return W2D_ERR_ARG;
        if (tau[i] < 0.0 || tau[i] > h->beta) return W2D_ERR_ARG;
        code[i] = orb[i] | (spin[i] << 4) | ((ca[i] - 1) << 5);
        int f = orb[i] + h->model.nbands * spin[i];
        cnt[2 * f + ca[i] - 1]++;
    }
    for (int f = 0; f < 2 * h->model.nbands; ++f)
        if (cnt[2 * f] != cnt[2 * f + 1] || cnt[2 * f] >= h->kcap) return W2D_ERR_ARG;
    char *base = h->chains + (long)chain * h->L.stride;
    HIPCHK(hipStreamSynchronize(h->stream));
    ChainHdr x;
    HIPCHK(hipMemcpy(&x, base, sizeof(x), hipMemcpyDeviceToHost));
    x.nop = n;
    x.outer = outer_sst;
    x.err = 0;
    HIPCHK(hipMemcpy(base, &x, sizeof(x), hipMemcpyHostToDevice));
    if (n > 0) {
        HIPCHK(hipMemcpy(base + h->L.tau, tau, (size_t)n * 8, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(base + h->L.code, code.data(), (size_t)n * 4, hipMemcpyHostToDevice));
    }
    DISPATCH(launch_setconfig, h, chain);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(h->stream));
    HIPCHK(hipMemcpy(&x, base, sizeof(x), hipMemcpyDeviceToHost));
    if (weights) {
        weights[0] = x.trace_log, weights[1] = x.trace_sign;
        weights[2] = x.det_log, weights[3] = x.det_sign, weights[4] = x.cfgsign;
    }
    return map_err(x.err);
}

int w2d_ctqmc_get_minv(w2d_ctqmc *h, int chain, int f, int *k, double *out)
{
    if (!h || chain < 0 || chain >= h->nchains || f < 0 || f >= 2 * h->model.nbands || !k) return W2D_ERR_ARG;
    HIPCHK(hipStreamSynchronize(h->stream));
    char *base = h->chains + (long)chain * h->L.stride;
    ChainHdr x;
    HIPCHK(hipMemcpy(&x, base, sizeof(x), hipMemcpyDeviceToHost));
    *k = x.kf[f];
    if (!out) return W2D_OK;
    const int kk = x.kf[f], ld = h->L.kcap;
    const char *mf = base + h->L.minv + (long)f * ld * ld * 8;
    for (int j = 0; j < kk; ++j)
        HIPCHK(hipMemcpy(out + (long)j * kk, mf + (long)j * ld * 8, (size_t)kk * 8, hipMemcpyDeviceToHost));
    return W2D_OK;
}

int w2d_ctqmc_last_kernel_ms(const w2d_ctqmc *h, double *ms, int64_t *launches)
{
    if (!h || !ms) return W2D_ERR_ARG;
    *ms = h->last_ms;
    if (launches) *launches = h->last_launches;
    return W2D_OK;
}

int w2d_ctqmc_destroy(w2d_ctqmc *h)
{
    if (!h) return W2D_OK;
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    for (void *p : h->dev_allocs) (void)hipFree(p);
    if (h->chains) (void)hipFree(h->chains);
    if (h->ev0) (void)hipEventDestroy(h->ev0);
    if (h->ev1) (void)hipEventDestroy(h->ev1);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return W2D_OK;
}

}  // extern "C"
