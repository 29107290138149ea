/*
 * CPU restatement of the reference CT-HYB Z-space sampling path (w2dynamics,
 * src/ctqmc_fortran, matrix solver, diagonal hybridisation, no state sampling,
 * no outer truncation, no worms, no 4-operator moves).
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the oracle the HIP product is checked
 * against; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * load it.  It is deliberately simple: plain arrays instead of the reference's
 * linked lists, and the local trace recomputed in full for every proposal (the
 * reference's bra/ket caching, Trace.F90:5354-5549, only changes which partial
 * products are reused, not the value).
 *
 * Every function names the reference code it follows.  The random-number draw
 * order reproduces the reference exactly, so that a chain seeded like
 * init_rnd(seed) makes the same accept/reject decisions.
 */
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* Mersenne twister mt19937 and the libstdc++ distributions used by          */
/* src/mtrng/random.cc:30-79 (std::mt19937, uniform_real_distribution<double>,*/
/* uniform_int_distribution<int>)                                             */
/* ------------------------------------------------------------------------ */
typedef struct {
    uint32_t mt[624];
    int idx;
} orng_t;

/* std::mersenne_twister_engine::seed (C++ [rand.eng.mers]) */
static void orng_seed(orng_t *r, uint32_t s)
{
    r->mt[0] = s;
    for (int i = 1; i < 624; ++i)
        r->mt[i] = 1812433253u * (r->mt[i - 1] ^ (r->mt[i - 1] >> 30)) + (uint32_t)i;
    r->idx = 624;
}

/* Matsumoto & Nishimura (1998) tempered output */
static uint32_t orng_next(orng_t *r)
{
    if (r->idx >= 624) {
        for (int i = 0; i < 624; ++i) {
            uint32_t y = (r->mt[i] & 0x80000000u) | (r->mt[(i + 1) % 624] & 0x7fffffffu);
            uint32_t v = r->mt[(i + 397) % 624] ^ (y >> 1);
            if (y & 1u) v ^= 0x9908b0dfu;
            r->mt[i] = v;
        }
        r->idx = 0;
    }
    uint32_t y = r->mt[r->idx++];
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
}

/* generate_canonical<double,53> with a 32-bit engine: two draws,
 * (g1 + g2*2^32) / 2^64, clamped below 1 (libstdc++ bits/random.tcc). */
static double orng_grnd(orng_t *r)
{
    double sum = (double)orng_next(r);
    double tmp = 4294967296.0;
    sum += (double)orng_next(r) * tmp;
    double ret = sum / (tmp * tmp);
    if (ret >= 1.0) ret = nextafter(1.0, 0.0);
    return ret;
}

/* uniform_int_distribution<int>(a,b): Lemire's nearly divisionless method on a
 * 64-bit product (libstdc++ bits/uniform_int_dist.h, _S_nd). */
static int orng_randint(orng_t *r, int a, int b)
{
    uint32_t range = (uint32_t)b - (uint32_t)a + 1u;
    uint64_t prod = (uint64_t)orng_next(r) * (uint64_t)range;
    uint32_t low = (uint32_t)prod;
    if (low < range) {
        uint32_t thr = (uint32_t)(-range) % range;
        while (low < thr) {
            prod = (uint64_t)orng_next(r) * (uint64_t)range;
            low = (uint32_t)prod;
        }
    }
    return a + (int)(prod >> 32);
}

/* ------------------------------------------------------------------------ */
/* signed logarithms: TLogTr / TLogDet (Trace.F90:80-321, MatrixUpdate.F90:5-36)*/
/* ------------------------------------------------------------------------ */
typedef struct {
    double log, sign;
} olog_t;

static olog_t olog_add(olog_t a, olog_t b)   /* addLogTr, Trace.F90:286-298 */
{
    olog_t r;
    if (a.log >= b.log) {
        r.log = a.log + log(1.0 + a.sign * b.sign * exp(b.log - a.log));
        r.sign = a.sign;
    } else {
        r.log = b.log + log(1.0 + a.sign * b.sign * exp(a.log - b.log));
        r.sign = b.sign;
    }
    return r;
}

/* ------------------------------------------------------------------------ */
/* model and chain                                                          */
/* ------------------------------------------------------------------------ */
typedef struct {
    double tau;
    int orb, spin, ca;  /* orb 0..N-1, spin 0..1, ca 1 = creator, 2 = annihilator */
} oop_t;

typedef struct {
    /* model (built by oracle/model.py) */
    int nb, nsst, nmax, nstates;
    const int *sizes, *offsets, *connect; /* connect[((s*nb+b)*2+sp)*2+ca0]   */
    const long *opoff;                    /* same index, offset into ops or -1 */
    const double *ops, *evals, *evec;
    const long *evoff;
    const int *first_state, *state_sst;
    double egs;
    /* hybridisation F(tau): [flavour f = b + nb*sp][nftau] */
    int nftau;
    const double *ftau;
    double beta;
    /* parameters */
    double p_shift, p_glob, p_outer, p_4op;
    long nslide;
    int ntau, nleg, maxhisto, meas_dm;
    int has_szt;
    /* window (Trace.F90:1608-1751) */
    double taudiff_max, tauwin_min, tauwin_max;
    int num_windows, window_position;
    /* configuration */
    oop_t *op;
    int nop, cap;
    int outer_sst;
    double *minv[14];  /* column-major k x k; rows creators, cols annihilators */
    int k[14];
    olog_t trace, det;
    double cfgsign;
    olog_t *parttrace; /* per outer state of the last accepted evaluation */
    olog_t *tpart;
    orng_t rng;
    /* counters (CTQMC.F90:168-179) */
    long try_add, try_rem, try_glob, try_shift;
    long acc_add, acc_rem, acc_glob, acc_shift, accqn_add, accqn_rem, accqn_glob;
    /* accumulators */
    double *gtau_buf, *gleg, *dm, *histo, *outer_hist, *outer_state_hist;
    double *sign_hist_sst, *sign_hist_states, *tc_sst, *tc_states;
    /* MeasGiw: raw G(i w_n) sums [nf][niw][2] (re, im) for n = 0..niw-1 (niw = 0: off) */
    int niw;
    double *giw;
    double meas_sign;
    long cnt_meas;
    /* accepted-pair statistics (CTQMC.F90:66-70): AccPair [nf][ACC_CAP],
     * NAccPair bins of [0, AccPairMax]; AccPairTau list of apt_n entries */
    double *accpair, accpair_max, *aptau;
    int naccpair, apt_index, apt_n;
    /* step-resolved G(tau) mean / mid [nmstep][nf] and sign [nmstep] */
    double *mstep, *dstep, *sstep;
    int nmstep, step_mean, step_mid, step_sign;
    /* AccumGtau = 2: binning accumulator of G(tau) (autocorracc) */
    int accum, nlev;
    long acc_n;
    double *gmeas, *acc_spill, *acc_mean, *acc_sq;   /* [ncomp], [nlev][ncomp] x 3 */
    /* parameter derivatives (no reference counterpart): dH/dtheta blocks in
     * the eigenbasis [ntheta][sum n_s^2] at dboff[s], dF/dtheta
     * [ntheta][nf][nftau]; score sums (Sum sign s, Sum G-sample s, Sum
     * rho-sample s, Sum d(rho-sample)/dtheta) */
    int nsym, *symmoves;   /* user symmetry moves [2 nsym][nf], 1-based flavours (inverses second) */
    int ntheta, score;
    long dbsize, *dboff;
    double *dblk, *dftau;
    double *sc_sum, *sc_g, *sc_dm, *sc_dmx;
    /* scratch */
    double *ket, *tmp, *dket;
    double *kets_b2, *bras_b2, *klog_b2, *blog_b2;
    int sst_b2;
} ochain_t;

#define NF(c) (2 * (c)->nb)
#define FL(o) ((o).orb + ch->nb * (o).spin)

static int conn(const ochain_t *ch, int s, int b, int sp, int ca)
{
    return ch->connect[((s * ch->nb + b) * 2 + sp) * 2 + (ca - 1)];
}

static const double *opblk(const ochain_t *ch, int s, int b, int sp, int ca)
{
    long o = ch->opoff[((s * ch->nb + b) * 2 + sp) * 2 + (ca - 1)];
    return o < 0 ? NULL : ch->ops + o;
}

/* get_HybrF, Trace.F90:1014-1033 */
static double hybr(const ochain_t *ch, double tau, int f)
{
    double sig = 1.0;
    if (tau < 0.0) {
        tau = tau + ch->beta;
        sig = -1.0;
    }
    double n = tau / ch->beta * (double)(ch->nftau - 1);
    int i = (int)n;
    const double *F = ch->ftau + (long)f * ch->nftau;
    return sig * (F[i] + (n - (double)i) * (F[i + 1] - F[i]));
}

/* ---------------------------- window ------------------------------------ */
/* set_window, Trace.F90:1691-1751 (edge values only; list pointers are
 * implicit in the sorted array) */
static void set_window(ochain_t *ch)
{
    if (ch->num_windows > 1) {
        ch->tauwin_min = ((double)ch->window_position / (double)(ch->num_windows + 1)) * ch->beta;
        ch->tauwin_max = ((double)(ch->window_position + 2) / (double)(ch->num_windows + 1)) * ch->beta;
    }
}

/* init_window, Trace.F90:1608-1646 */
static void init_window(ochain_t *ch, double tdm)
{
    ch->window_position = 0;
    if (tdm > 0.0) {
        ch->taudiff_max = tdm;
        int nw = (int)floor(ch->beta / tdm) - 1;
        ch->num_windows = nw > 1 ? nw : 1;
        ch->tauwin_min = ((double)ch->window_position / (double)(ch->num_windows + 1)) * ch->beta;
        ch->tauwin_max = ((double)(ch->window_position + 2) / (double)(ch->num_windows + 1)) * ch->beta;
    } else {
        ch->taudiff_max = ch->beta;
        ch->num_windows = 1;
        ch->tauwin_min = 0.0;
        ch->tauwin_max = ch->beta;
    }
    set_window(ch);
}

/* shift_window, Trace.F90:1650-1680 */
static void shift_window(ochain_t *ch)
{
    if (ch->num_windows > 1) {
        ch->window_position++;
        if (ch->window_position == ch->num_windows) ch->window_position = 0;
        set_window(ch);
    } else {
        ch->window_position = 0;
    }
}

/* ------------------------- configuration helpers ------------------------ */
static int count_flavour(const ochain_t *ch, int f)
{
    int n = 0;
    for (int i = 0; i < ch->nop; ++i) n += (FL(ch->op[i]) == f);
    return n;
}

/* Number of removable pairs: ensure_valid_NPairs_diag, Trace.F90:1769-1810 */
static int npairs(const ochain_t *ch)
{
    int np = 0;
    for (int i = 0; i < ch->nop; ++i) {
        const oop_t *e = &ch->op[i];
        if (e->tau < ch->tauwin_min) continue;
        if (e->tau > ch->tauwin_max) break;
        for (int j = i + 1; j < ch->nop; ++j) {
            const oop_t *m = &ch->op[j];
            if (m->tau > ch->tauwin_max) break;
            if (m->tau > e->tau + ch->taudiff_max) break;
            if (FL(*m) == FL(*e) && m->ca != e->ca) ++np;
        }
    }
    return np;
}

/* 1-based rank of op i among same flavour and same c/c+ type (FPos) */
static int fpos_of(const ochain_t *ch, int i)
{
    int r = 1;
    for (int j = 0; j < i; ++j)
        if (FL(ch->op[j]) == FL(ch->op[i]) && ch->op[j].ca == ch->op[i].ca) ++r;
    return r;
}

/* Superstate sequence from outer_sst through all operators; 1 if valid and
 * cyclic.  Equivalent to check_sst_sequence (Trace.F90:1243-1302) for the
 * changed span since the unchanged parts were valid before. */
static int qn_valid(const ochain_t *ch, int outer)
{
    int s = outer;
    for (int i = 0; i < ch->nop; ++i) {
        s = conn(ch, s, ch->op[i].orb, ch->op[i].spin, ch->op[i].ca);
        if (s < 0) return 0;
    }
    return s == outer;
}

static void insert_op(ochain_t *ch, int pos, oop_t o)
{
    memmove(ch->op + pos + 1, ch->op + pos, sizeof(oop_t) * (size_t)(ch->nop - pos));
    ch->op[pos] = o;
    ch->nop++;
}

static void remove_op(ochain_t *ch, int pos)
{
    memmove(ch->op + pos, ch->op + pos + 1, sizeof(oop_t) * (size_t)(ch->nop - pos - 1));
    ch->nop--;
}

static int insert_pos(const ochain_t *ch, double tau)
{
    int p = 0;
    while (p < ch->nop && ch->op[p].tau < tau) ++p;
    return p;
}

/* ------------------------------ trace ----------------------------------- */
/* split_log, Trace.F90:5341-5348 */
static void split_log(double *v, int n, double *lf)
{
    double mv = 0.0;
    for (int i = 0; i < n; ++i) mv = fmax(mv, fabs(v[i]));
    if (mv > 0.0)
        for (int i = 0; i < n; ++i) v[i] /= mv;
    *lf += log(mv);
}

/* Local trace over the outer superstate with a full bra/ket-free recompute:
 * for every outer state the ket is propagated from tau=0 through all
 * operators to beta.  Values follow get_Trace_EB_fullmatrix and
 * get_Trace_EB_onlyscalars (Trace.F90:5556-6102): per-state time evolution
 * relative to the block's lowest level, per-vector log rescaling, part traces
 * summed with addLogTr. */
static olog_t local_trace(ochain_t *ch, olog_t *part)
{
    olog_t tr = {-DBL_MAX, 0.0}; /* TLogTr(log = -huge, sign = 0) */
    int outer = ch->outer_sst;
    int no = ch->sizes[outer];
    double *v = ch->ket, *w = ch->tmp;
    for (int i0 = 0; i0 < no; ++i0) {
        int s = outer, ns = no;
        double lg = 0.0, tr0 = 0.0;
        for (int i = 0; i < ns; ++i) v[i] = 0.0;
        v[i0] = 1.0;
        for (int j = 0; j < ch->nop; ++j) {
            const oop_t *o = &ch->op[j];
            int t = conn(ch, s, o->orb, o->spin, o->ca);
            if (t < 0) {   /* no Connect target: the outer superstate's trace is zero */
                s = -1;
                break;
            }
            int nt = ch->sizes[t];
            const double *E = ch->evals + ch->offsets[s];
            double dt = o->tau - tr0;
            lg += -(E[0] - ch->egs) * dt;
            const double *P = opblk(ch, s, o->orb, o->spin, o->ca);
            if (ch->nmax > 1 && ns > 1) {
                for (int i = 0; i < ns; ++i) v[i] *= exp(-(E[i] - E[0]) * dt);
                split_log(v, ns, &lg);
                for (int r = 0; r < nt; ++r) {
                    double acc = 0.0;
                    for (int c = 0; c < ns; ++c) acc += P[r + (long)c * nt] * v[c];
                    w[r] = acc;
                }
                split_log(w, nt, &lg);
            } else {
                for (int r = 0; r < nt; ++r) w[r] = v[0] * P[r];
            }
            memcpy(v, w, sizeof(double) * (size_t)nt);
            s = t;
            ns = nt;
            tr0 = o->tau;
        }
        if (s != outer) {
            /* a configuration that breaks the quantum numbers (set from
             * outside; the moves never propose one) has zero weight */
            if (part)
                for (int q = 0; q < no; ++q) part[q] = (olog_t){-DBL_MAX, 0.0};
            return (olog_t){-DBL_MAX, 0.0};
        }
        /* propagate to beta and project on <i0| */
        const double *E = ch->evals + ch->offsets[s];
        double dt = ch->beta - tr0;
        lg += -(E[0] - ch->egs) * dt;
        double val;
        if (ch->nmax > 1) {
            for (int i = 0; i < ns; ++i) v[i] *= exp(-(E[i] - E[0]) * dt);
            split_log(v, ns, &lg);
            val = v[i0];
        } else {
            val = v[0];
        }
        olog_t p; /* log(abs(braket)), sign(1.0, braket) */
        p.log = lg + log(fabs(val));
        p.sign = copysign(1.0, val);
        if (part) part[i0] = p;
        tr = olog_add(tr, p);
    }
    return tr;
}

/* ket at beta/2 (ops with tau <= beta/2) and bra at beta/2 (ops with tau >
 * beta/2), update_b2_states_eb (Trace.F90:6213-6469) */
static void beta_half_states(ochain_t *ch)
{
    int outer = ch->outer_sst, no = ch->sizes[outer];
    double bh = ch->beta / 2.0;
    int nmax = ch->nmax;
    for (int i0 = 0; i0 < no; ++i0) {
        /* ket */
        int s = outer, ns = no;
        double lg = 0.0, t0 = 0.0;
        double *v = ch->kets_b2 + (long)i0 * nmax;
        double *w = ch->tmp;
        for (int i = 0; i < ns; ++i) v[i] = 0.0;
        v[i0] = 1.0;
        for (int j = 0; j < ch->nop && ch->op[j].tau <= bh; ++j) {
            const oop_t *o = &ch->op[j];
            int t = conn(ch, s, o->orb, o->spin, o->ca), nt = ch->sizes[t];
            const double *E = ch->evals + ch->offsets[s];
            double dt = o->tau - t0;
            for (int i = 0; i < ns; ++i) v[i] *= exp(-(E[i] - E[0]) * dt);
            lg += -(E[0] - ch->egs) * dt;
            split_log(v, ns, &lg);
            const double *P = opblk(ch, s, o->orb, o->spin, o->ca);
            for (int r = 0; r < nt; ++r) {
                double acc = 0.0;
                for (int c = 0; c < ns; ++c) acc += P[r + (long)c * nt] * v[c];
                w[r] = acc;
            }
            split_log(w, nt, &lg);
            memcpy(v, w, sizeof(double) * (size_t)nt);
            s = t, ns = nt, t0 = o->tau;
        }
        {
            const double *E = ch->evals + ch->offsets[s];
            for (int i = 0; i < ns; ++i) v[i] *= exp(-(E[i] - E[0]) * (bh - t0));
            split_log(v, ns, &lg);
            ch->klog_b2[i0] = lg - (E[0] - ch->egs) * (bh - t0);
        }
        ch->sst_b2 = s;
        /* bra */
        s = outer, ns = no, lg = 0.0, t0 = ch->beta;
        v = ch->bras_b2 + (long)i0 * nmax;
        for (int i = 0; i < ns; ++i) v[i] = 0.0;
        v[i0] = 1.0;
        for (int j = ch->nop - 1; j >= 0 && ch->op[j].tau > bh; --j) {
            const oop_t *o = &ch->op[j];
            /* bra side applies the adjoint: c <-> c+ block (Trace.F90:5929) */
            int t = conn(ch, s, o->orb, o->spin, 3 - o->ca), nt = ch->sizes[t];
            const double *E = ch->evals + ch->offsets[s];
            double dt = t0 - o->tau;
            for (int i = 0; i < ns; ++i) v[i] *= exp(-(E[i] - E[0]) * dt);
            lg += -(E[0] - ch->egs) * dt;
            split_log(v, ns, &lg);
            const double *P = opblk(ch, s, o->orb, o->spin, 3 - o->ca);
            for (int r = 0; r < nt; ++r) {
                double acc = 0.0;
                for (int c = 0; c < ns; ++c) acc += P[r + (long)c * nt] * v[c];
                w[r] = acc;
            }
            split_log(w, nt, &lg);
            memcpy(v, w, sizeof(double) * (size_t)nt);
            s = t, ns = nt, t0 = o->tau;
        }
        {
            const double *E = ch->evals + ch->offsets[s];
            for (int i = 0; i < ns; ++i) v[i] *= exp((E[i] - E[0]) * (bh - t0));
            split_log(v, ns, &lg);
            ch->blog_b2[i0] = lg + (E[0] - ch->egs) * (bh - t0);
        }
    }
}

/* ---------------------------- determinants ------------------------------ */
/* LU with partial pivoting; log|det| and sign as get_MatLogDetFull /
 * get_LogDetFull (MatrixUpdate.F90:41-99, 302-339).  If inv != NULL the
 * inverse is written there.  Returns 0 on a singular matrix. */
static int lu_logdet(int k, const double *A, olog_t *det, double *inv)
{
    det->log = 0.0;
    det->sign = 1.0;
    if (k == 0) return 1;
    double *a = malloc(sizeof(double) * (size_t)k * k);
    int *piv = malloc(sizeof(int) * (size_t)k);
    memcpy(a, A, sizeof(double) * (size_t)k * k);
    for (int c = 0; c < k; ++c) {
        int p = c;
        double m = fabs(a[c + (long)c * k]);
        for (int r = c + 1; r < k; ++r)
            if (fabs(a[r + (long)c * k]) > m) m = fabs(a[r + (long)c * k]), p = r;
        piv[c] = p;
        if (m == 0.0) {
            free(a), free(piv);
            det->log = -HUGE_VAL, det->sign = 0.0;
            return 0;
        }
        if (p != c)
            for (int j = 0; j < k; ++j) {
                double t = a[c + (long)j * k];
                a[c + (long)j * k] = a[p + (long)j * k];
                a[p + (long)j * k] = t;
            }
        double d = a[c + (long)c * k];
        for (int r = c + 1; r < k; ++r) {
            double l = a[r + (long)c * k] / d;
            a[r + (long)c * k] = l;
            for (int j = c + 1; j < k; ++j) a[r + (long)j * k] -= l * a[c + (long)j * k];
        }
    }
    for (int i = 0; i < k; ++i) {
        double d = a[i + (long)i * k];
        det->sign *= (piv[i] != i ? -1.0 : 1.0) * (d >= 0 ? 1.0 : -1.0);
        det->log += log(fabs(d));
    }
    if (inv) {
        /* solve A X = I column by column */
        for (int col = 0; col < k; ++col) {
            double *x = inv + (long)col * k;
            for (int i = 0; i < k; ++i) x[i] = (i == col);
            for (int i = 0; i < k; ++i) {
                double t = x[i];
                x[i] = x[piv[i]];
                x[piv[i]] = t;
            }
            for (int i = 0; i < k; ++i)
                for (int r = i + 1; r < k; ++r) x[r] -= a[r + (long)i * k] * x[i];
            for (int i = k - 1; i >= 0; --i) {
                x[i] /= a[i + (long)i * k];
                for (int r = 0; r < i; ++r) x[r] -= a[r + (long)i * k] * x[i];
            }
        }
    }
    free(a), free(piv);
    return 1;
}

/* Hybridisation matrix of flavour f, Delta(row = annihilator rank,
 * col = creator rank) = F(tau_a - tau_c): get_InvFullHybr, Trace.F90:2279-2333 */
static int build_delta(const ochain_t *ch, int f, double *D)
{
    int k = 0;
    for (int i = 0; i < ch->nop; ++i)
        if (FL(ch->op[i]) == f && ch->op[i].ca == 1) ++k;
    int col = 0;
    for (int i = 0; i < ch->nop; ++i) {
        const oop_t *c = &ch->op[i];
        if (FL(*c) != f || c->ca != 1) continue;
        int row = 0;
        for (int j = 0; j < ch->nop; ++j) {
            const oop_t *a = &ch->op[j];
            if (FL(*a) != f || a->ca != 2) continue;
            D[row + (long)col * k] = hybr(ch, a->tau - c->tau, f);
            ++row;
        }
        ++col;
    }
    return k;
}

/* full inverse + determinant of all flavours (get_InvFullHybr) */
static olog_t full_inverse(ochain_t *ch, double **newinv, int *newk)
{
    olog_t det = {0.0, 1.0};
    for (int f = 0; f < NF(ch); ++f) {
        int k = count_flavour(ch, f) / 2;
        newk[f] = k;
        newinv[f] = malloc(sizeof(double) * (size_t)(k > 0 ? k * k : 1));
        if (k == 0) continue;
        double *D = malloc(sizeof(double) * (size_t)k * k);
        build_delta(ch, f, D);
        olog_t d;
        lu_logdet(k, D, &d, newinv[f]);
        det.log += d.log;
        det.sign *= d.sign;
        free(D);
    }
    return det;
}

/* get_DetRatAdd (MatrixUpdate.F90:392-407): S - sum_ij R_i M_ij Q_j */
static double detrat_add(const double *M, int k, const double *Q, const double *R, double S)
{
    double r = S;
    for (int i = 0; i < k; ++i)
        for (int j = 0; j < k; ++j) r -= R[i] * M[i + (long)j * k] * Q[j];
    return r;
}

/* get_MatAdd (MatrixUpdate.F90:410-442); p1 row (creator), p2 col, 1-based */
static double *mat_add(const double *N, int k, const double *Q, const double *R, double dr,
                       int p1, int p2)
{
    int k1 = k + 1;
    double *A = malloc(sizeof(double) * (size_t)k1 * k1);
#define AE(i, j) A[((i) - 1) + (long)((j) - 1) * k1]
#define NE(i, j) N[((i) - 1) + (long)((j) - 1) * k]
    AE(p1, p2) = 1.0 / dr;
    for (int i = 1; i <= k; ++i) {
        double s = 0.0;
        for (int j = 1; j <= k; ++j) s += NE(i, j) * Q[j - 1];
        AE(i < p1 ? i : i + 1, p2) = -s / dr;
    }
    for (int j = 1; j <= k; ++j) {
        double s = 0.0;
        for (int i = 1; i <= k; ++i) s += R[i - 1] * NE(i, j);
        AE(p1, j < p2 ? j : j + 1) = -s / dr;
    }
    for (int i = 1; i <= k; ++i) {
        int in = i < p1 ? i : i + 1;
        for (int j = 1; j <= k; ++j) {
            int jn = j < p2 ? j : j + 1;
            AE(in, jn) = NE(i, j) + dr * AE(in, p2) * AE(p1, jn);
        }
    }
#undef AE
#undef NE
    return A;
}

/* get_MatRem (MatrixUpdate.F90:455-482) */
static double *mat_rem(const double *N, int k, int p1, int p2)
{
    int km = k - 1;
    double *A = malloc(sizeof(double) * (size_t)(km > 0 ? km * km : 1));
#define NE(i, j) N[((i) - 1) + (long)((j) - 1) * k]
    for (int i = 1; i <= km; ++i) {
        int io = i < p1 ? i : i + 1;
        for (int j = 1; j <= km; ++j) {
            int jo = j < p2 ? j : j + 1;
            A[(i - 1) + (long)(j - 1) * km] = NE(io, jo) - NE(io, p2) * NE(p1, jo) / NE(p1, p2);
        }
    }
#undef NE
    return A;
}

/* ------------------------------ sign ------------------------------------ */
/* get_Sign (Trace.F90:3479-3496) with get_BlockSign (3594-3615) and
 * get_permaltca (3549-3587) */
static double config_sign(const ochain_t *ch)
{
    int nb = ch->nb;
    int tn[14];
    for (int f = 0; f < NF(ch); ++f) tn[f] = count_flavour(ch, f);
    long perm = 0;
    for (int i = 0; i < ch->nop; ++i) {
        const oop_t *o = &ch->op[i];
        tn[FL(*o)]--;
        if (o->spin == 1) {
            for (int b = 0; b <= o->orb; ++b) perm += tn[b + 0 * nb];
        } else {
            for (int b = 0; b < o->orb; ++b) perm += tn[b + 1 * nb];
        }
        for (int b = 0; b < o->orb; ++b) perm += tn[b + nb * o->spin];
    }
    double block = (perm % 2) ? -1.0 : 1.0;
    for (int f = 0; f < NF(ch); ++f) tn[f] = count_flavour(ch, f);
    long pa = 0;
    for (int i = 0; i < ch->nop; ++i) {
        const oop_t *o = &ch->op[i];
        tn[FL(*o)]--;
        pa += tn[FL(*o)] % (3 - o->ca);
    }
    if ((ch->nop / 2) % 2 == 1) pa++;
    double alt = (pa % 2) ? -1.0 : 1.0;
    return ch->trace.sign * ch->det.sign * block * alt;
}

/* ------------------------------ moves ----------------------------------- */
enum { MV_ADD = 1, MV_REM = 2, MV_GLOB = 3, MV_SHIFT = 4 };

static void accept_trace(ochain_t *ch, olog_t tr)
{
    ch->trace = tr;
    int no = ch->sizes[ch->outer_sst];
    memcpy(ch->parttrace, ch->tpart, sizeof(olog_t) * (size_t)no);
}

#define ACC_CAP 8192
#define APT_CAP 65536

/* accepted pair of flavour f spanning dt (CTQMC.F90:3579-3587, 4119-4127) */
static void record_pair(ochain_t *ch, int f, double dt)
{
    int bin = (int)floor(dt / ch->accpair_max * (double)ch->naccpair) + 1;
    if (bin > ch->naccpair) bin = ch->naccpair;
    ch->accpair[(long)f * ACC_CAP + bin - 1] += 1.0;
    if (ch->apt_index <= ch->apt_n) {
        ch->aptau[ch->apt_index - 1] = dt;
        ch->apt_index++;
    }
}

/* StepAdd (CTQMC.F90:3353-3612) with pair_OperAdd (Trace.F90:3657-3698) and
 * process_OperAdd (Trace.F90:4275-4462); change_outer is always false here */
static int step_add(ochain_t *ch)
{
    ch->try_add++;
    oop_t o1, o2;
    double width = ch->tauwin_max - ch->tauwin_min;
    double tref = ch->tauwin_min + width * orng_grnd(&ch->rng);
    double tdf = width;
    o1.tau = tref;
    o1.ca = 2;
    o1.orb = orng_randint(&ch->rng, 1, ch->nb) - 1;
    o1.spin = orng_randint(&ch->rng, 1, 2) - 1;
    double tmin = fmax(tref - ch->taudiff_max, ch->tauwin_min);
    double w2 = fmin(tref + ch->taudiff_max, ch->tauwin_max) - tmin;
    o2.tau = tmin + w2 * orng_grnd(&ch->rng);
    tdf *= w2;
    o2.ca = 1;
    o2.orb = o1.orb;
    o2.spin = o1.spin;
    /* equal times */
    if (o1.tau == o2.tau) return ch->accqn_add++, 2;
    for (int i = 0; i < ch->nop; ++i)
        if (ch->op[i].tau == o1.tau || ch->op[i].tau == o2.tau) return ch->accqn_add++, 2;
    if (o1.tau > o2.tau) {
        oop_t t = o1;
        o1 = o2;
        o2 = t;
    }
    int f = FL(o1);
    int k = count_flavour(ch, f) / 2;
    /* FPos(CA): rank in the new configuration */
    int fpos[3];
    int oi1 = insert_pos(ch, o1.tau);
    insert_op(ch, oi1, o1);
    int oi2 = insert_pos(ch, o2.tau);
    insert_op(ch, oi2, o2);
    fpos[o1.ca] = fpos_of(ch, oi1);
    fpos[o2.ca] = fpos_of(ch, oi2);
    if (!qn_valid(ch, ch->outer_sst)) {
        remove_op(ch, oi2);
        remove_op(ch, oi1);
        ch->accqn_add++;
        return 2;
    }
    /* det ratio, get_LogDetRatPairAdd (Trace.F90:2775-2832) */
    const oop_t *cr = o1.ca == 1 ? &o1 : &o2;
    const oop_t *an = o1.ca == 1 ? &o2 : &o1;
    double *Q = malloc(sizeof(double) * (size_t)(k + 1));
    double *R = malloc(sizeof(double) * (size_t)(k + 1));
    int iq = 0, ir = 0;
    for (int i = 0; i < ch->nop; ++i) {
        const oop_t *e = &ch->op[i];
        if (FL(*e) != f || i == oi1 || i == oi2) continue;
        if (e->ca == 2)
            Q[iq++] = hybr(ch, e->tau - cr->tau, f);
        else
            R[ir++] = hybr(ch, an->tau - e->tau, f);
    }
    double S = hybr(ch, an->tau - cr->tau, f);
    S = detrat_add(ch->minv[f], k, Q, R, S);
    double sig = (abs(fpos[1] - fpos[2]) % 2) ? -1.0 : 1.0;
    olog_t dr = {log(fabs(S)), (S >= 0 ? 1.0 : -1.0) * sig};
    olog_t trn = local_trace(ch, ch->tpart);
    double rnd = orng_grnd(&ch->rng);
    int np = npairs(ch);
    double wr = exp((trn.log - ch->trace.log) + dr.log);
    double prob = fabs(((1.0 * tdf) * (double)(2 * ch->nb)) / (double)np * wr);
    if (rnd < prob) {
        double *nm = mat_add(ch->minv[f], k, Q, R, S, fpos[1], fpos[2]);
        free(ch->minv[f]);
        ch->minv[f] = nm;
        ch->k[f] = k + 1;
        ch->det.log += dr.log;
        ch->det.sign *= dr.sign;
        accept_trace(ch, trn);
        ch->cfgsign = -config_sign(ch);
        ch->acc_add++;
        record_pair(ch, f, o2.tau - o1.tau);
        free(Q), free(R);
        return 1;
    }
    remove_op(ch, oi2);
    remove_op(ch, oi1);
    free(Q), free(R);
    return 0;
}

/* StepRem (CTQMC.F90:3892-4150) with gen_OperRemove (Trace.F90:4851-4977) */
static int step_rem(ochain_t *ch)
{
    ch->try_rem++;
    int np = npairs(ch);
    if (np <= 0) return ch->accqn_rem++, 2;
    int pick = orng_randint(&ch->rng, 1, np);
    int i1 = -1, i2 = -1;
    for (int i = 0; i < ch->nop && i1 < 0; ++i) {
        const oop_t *e = &ch->op[i];
        if (e->tau < ch->tauwin_min) continue;
        for (int j = i + 1; j < ch->nop; ++j) {
            const oop_t *m = &ch->op[j];
            if (m->tau > ch->tauwin_max || m->tau > e->tau + ch->taudiff_max) break;
            if (FL(*m) == FL(*e) && m->ca != e->ca) {
                if (--pick <= 0) {
                    i1 = i, i2 = j;
                    break;
                }
            }
        }
    }
    oop_t o1 = ch->op[i1], o2 = ch->op[i2];
    int f = FL(o1);
    int k = count_flavour(ch, f) / 2;
    int fpos[3];
    fpos[o1.ca] = fpos_of(ch, i1);
    fpos[o2.ca] = fpos_of(ch, i2);
    double ta = o1.ca == 2 ? o1.tau : o2.tau;
    double hprf = (double)np /
                  ((ch->tauwin_max - ch->tauwin_min) *
                   (fmin(ta + ch->taudiff_max, ch->tauwin_max) - fmax(ta - ch->taudiff_max, ch->tauwin_min)));
    remove_op(ch, i2);
    remove_op(ch, i1);
    if (!qn_valid(ch, ch->outer_sst)) {
        insert_op(ch, i1, o1);
        insert_op(ch, i2, o2);
        ch->accqn_rem++;
        return 2;
    }
    /* get_LogDetRatPairRem (Trace.F90:2937-2963): M(FPos(1), FPos(2)) */
    double S = ch->minv[f][(fpos[1] - 1) + (long)(fpos[2] - 1) * k];
    double sig = (abs(fpos[1] - fpos[2]) % 2) ? -1.0 : 1.0;
    olog_t dr = {log(fabs(S)), (S >= 0 ? 1.0 : -1.0) * sig};
    olog_t trn = local_trace(ch, ch->tpart);
    double wr = exp((trn.log - ch->trace.log) + dr.log);
    double prob = fabs(((1.0 * hprf) / (double)(2 * ch->nb)) * wr);
    double rnd = orng_grnd(&ch->rng);
    if (rnd < prob) {
        if (fabs(dr.sign * exp(dr.log)) < 1e-12) {
            fprintf(stderr, "oracle: DetRat too small in removal\n");
            abort();
        }
        double *nm = mat_rem(ch->minv[f], k, fpos[1], fpos[2]);
        free(ch->minv[f]);
        ch->minv[f] = nm;
        ch->k[f] = k - 1;
        ch->det.log += dr.log;
        ch->det.sign *= dr.sign;
        accept_trace(ch, trn);
        ch->cfgsign = -config_sign(ch);
        ch->acc_rem++;
        record_pair(ch, f, o2.tau - o1.tau);
        return 1;
    }
    insert_op(ch, i1, o1);
    insert_op(ch, i2, o2);
    return 0;
}

/* globalUpdate (Trace.F90:2126-2188): relabel flavours, optionally swap c/c+ */
static void global_update(ochain_t *ch, const int *gu)
{
    int nb = ch->nb;
    for (int i = 0; i < ch->nop; ++i) {
        oop_t *o = &ch->op[i];
        int idx = o->orb + nb * o->spin + 1; /* 1-based flavour index */
        int g = gu[idx];
        o->orb = (g - 1) % nb;
        o->spin = (g - 1) / nb;
        if (gu[0] != 0) o->ca = 3 - o->ca;
    }
}

static void inverse_gu(const int *gu, int *inv, int nf)
{
    inv[0] = gu[0];
    for (int i = 1; i <= nf; ++i) inv[gu[i]] = i;
}

/* number of superstates for which the whole operator string is valid
 * (check_qn_all_ssts, Trace.F90:1386-1433); ok[s] marks them */
static int count_valid_outer(const ochain_t *ch, int *ok)
{
    int n = 0;
    for (int s = 0; s < ch->nsst; ++s) {
        ok[s] = qn_valid(ch, s);
        n += ok[s];
    }
    return n;
}

/* StepGlob (CTQMC.F90:4433-4554) */
static int step_glob(ochain_t *ch)
{
    ch->try_glob++;
    if (ch->nop == 0) return 0;
    int nf = NF(ch), nb = ch->nb;
    int gu[15], igu[15];
    double r = orng_grnd(&ch->rng);
    int spinflip = 0;
    if (r < 0.3) {
        spinflip = 1; /* gen_SpinFlipUpdate, Trace.F90:1143-1159 */
        gu[0] = 0;
        for (int i = 1; i <= nb; ++i) gu[i] = i + nb;
        for (int i = nb + 1; i <= 2 * nb; ++i) gu[i] = i - nb;
    } else if (r < 0.6) {
        gu[0] = 1; /* gen_CAFlip, Trace.F90:1178-1190 */
        for (int i = 1; i <= nf; ++i) gu[i] = i;
    } else if (r < 0.8 && ch->nsym > 0) {
        /* gen_SymUpdate (Trace.F90:1166-1182): one of the user's flavour
         * permutations or its inverse, no c/c+ flip */
        int isym = orng_randint(&ch->rng, 1, 2 * ch->nsym);
        gu[0] = 0;
        for (int i = 1; i <= nf; ++i) gu[i] = ch->symmoves[(isym - 1) * nf + (i - 1)];
    } else {
        /* gen_GlobalUpdate_new + rand_perm (Trace.F90:1065-1103) */
        int a[15];
        for (int i = 1; i <= nf; ++i) a[i] = i;
        for (int j = 1; j <= nf; ++j) {
            int i = orng_randint(&ch->rng, j, nf);
            int t = a[j];
            a[j] = a[i];
            a[i] = t;
        }
        for (int i = 1; i <= nf; ++i) gu[i] = a[i];
        gu[0] = orng_randint(&ch->rng, 0, 1);
    }
    int old_outer = ch->outer_sst;
    double sst_factor;
    int *ok = malloc(sizeof(int) * (size_t)ch->nsst);
    if (spinflip && ch->has_szt) {
        /* perform_spinflip_check_qn (Trace.F90:1306-1339) */
        int st = ch->first_state[ch->outer_sst];
        int dn = st & ((1 << nb) - 1), up = (st >> nb) & ((1 << nb) - 1);
        int ns = up | (dn << nb);
        ch->outer_sst = ch->state_sst[ns];
        global_update(ch, gu);
        sst_factor = qn_valid(ch, ch->outer_sst) ? 1.0 : 0.0;
    } else {
        /* perform_global_check_qn (Trace.F90:1345-1383) */
        int pold = count_valid_outer(ch, ok);
        global_update(ch, gu);
        int pnew = count_valid_outer(ch, ok);
        if (pnew <= 0) {
            sst_factor = 0.0;
        } else {
            sst_factor = (double)pnew / (double)pold;
            int kk = orng_randint(&ch->rng, 1, pnew);
            for (int s = 0; s < ch->nsst; ++s)
                if (ok[s] && --kk <= 0) {
                    ch->outer_sst = s;
                    break;
                }
        }
    }
    free(ok);
    if (sst_factor == 0.0) {
        ch->outer_sst = old_outer;
        inverse_gu(gu, igu, nf);
        global_update(ch, igu);
        ch->accqn_glob++;
        return 2;
    }
    double *ninv[14];
    int nk[14];
    olog_t dnew = full_inverse(ch, ninv, nk);
    olog_t dold = {0.0, 1.0}; /* get_FullLogDet of the old inverse */
    for (int f = 0; f < nf; ++f) {
        if (ch->k[f] == 0) continue;
        olog_t d;
        if (ch->k[f] == 1) {
            d.log = log(fabs(ch->minv[f][0]));
            d.sign = ch->minv[f][0] >= 0 ? 1.0 : -1.0;
        } else {
            lu_logdet(ch->k[f], ch->minv[f], &d, NULL);
        }
        dold.log += d.log;
        dold.sign *= d.sign;
    }
    olog_t dr = {dnew.log + dold.log, dnew.sign * dold.sign};
    olog_t trn = local_trace(ch, ch->tpart);
    double wr = exp((trn.log - ch->trace.log) + dr.log);
    if (orng_grnd(&ch->rng) < fabs(wr * sst_factor)) {
        ch->det = dnew;
        accept_trace(ch, trn);
        for (int f = 0; f < nf; ++f) {
            free(ch->minv[f]);
            ch->minv[f] = ninv[f];
            ch->k[f] = nk[f];
        }
        ch->cfgsign = -config_sign(ch);
        ch->acc_glob++;
        return 1;
    }
    for (int f = 0; f < nf; ++f) free(ninv[f]);
    ch->outer_sst = old_outer;
    inverse_gu(gu, igu, nf);
    global_update(ch, igu);
    return 0;
}

/* StepChangeOuter (CTQMC.F90:4569-4637), empty configuration */
static int step_change_outer(ochain_t *ch)
{
    double *vals = malloc(sizeof(double) * (size_t)ch->nsst);
    double sum = 0.0;
    for (int s = 0; s < ch->nsst; ++s) {
        ch->outer_sst = s;
        olog_t t = local_trace(ch, NULL);
        vals[s] = t.sign * exp(t.log);
        sum += vals[s];
    }
    for (int s = 0; s < ch->nsst; ++s) vals[s] /= sum;
    double r = orng_grnd(&ch->rng), acc = 0.0;
    for (int s = 0; s < ch->nsst; ++s) {
        ch->outer_sst = s;
        acc += vals[s];
        if (acc >= r && vals[s] != 0.0) break;
    }
    free(vals);
    accept_trace(ch, local_trace(ch, ch->tpart));
    ch->cfgsign = -config_sign(ch);
    return 0;
}

/* StepShiftTau_proper (CTQMC.F90:4641-4835) with
 * update_InvFullHybrShiftTau_diag (Trace.F90:2355-2457) */
static int step_shift(ochain_t *ch)
{
    if (ch->nop == 0) return step_change_outer(ch);
    double dtau = orng_grnd(&ch->rng) * ch->beta;
    if (dtau == 0.0 || dtau == ch->beta) return 0;
    ch->try_shift++;
    int nf = NF(ch), n = ch->nop;
    int wrapped[14][3];
    memset(wrapped, 0, sizeof(wrapped));
    int first = n; /* index of first op with tau >= dtau */
    for (int i = n - 1; i >= 0; --i)
        if (ch->op[i].tau >= dtau) first = i;
    for (int i = 0; i < n; ++i)
        if (ch->op[i].tau < dtau) wrapped[FL(ch->op[i])][ch->op[i].ca]++;
    if (first == n) first = 0;
    /* new outer superstate: the one preceding the new first operator */
    int s = ch->outer_sst;
    for (int i = 0; i < first; ++i) s = conn(ch, s, ch->op[i].orb, ch->op[i].spin, ch->op[i].ca);
    oop_t *old = malloc(sizeof(oop_t) * (size_t)n);
    memcpy(old, ch->op, sizeof(oop_t) * (size_t)n);
    int old_outer = ch->outer_sst, old_pos = ch->window_position;
    double old_min = ch->tauwin_min, old_max = ch->tauwin_max;
    for (int i = 0; i < n; ++i) {
        oop_t o = old[(first + i) % n];
        double t = o.tau - dtau;
        o.tau = t - floor(t / ch->beta) * ch->beta; /* Fortran modulo */
        ch->op[i] = o;
    }
    ch->outer_sst = s;
    if (ch->num_windows > 1) {
        int nw = ch->num_windows;
        int a = (int)floor((dtau / ch->beta) * (double)(nw + 1));
        int am = ((a % nw) + nw) % nw;
        int p = old_pos - am;
        ch->window_position = ((p % nw) + nw) % nw;
    } else {
        ch->window_position = ((old_pos % 1) + 1) % 1;
    }
    set_window(ch);
    olog_t trn = local_trace(ch, ch->tpart);
    double ratio = trn.sign * ch->trace.sign * exp(trn.log - ch->trace.log);
    if (orng_grnd(&ch->rng) < fabs(1.0 * ratio * 1.0)) {
        accept_trace(ch, trn);
        double detrat = 1.0;
        for (int f = 0; f < nf; ++f) {
            int k = ch->k[f], w1 = wrapped[f][1], w2 = wrapped[f][2];
            if (!((w1 == 0 && w2 == 0) || (w1 == k && w2 == k))) {
                double *M = ch->minv[f];
                double *A = malloc(sizeof(double) * (size_t)k * k);
#define ME(i, j) M[((i) - 1) + (long)((j) - 1) * k]
#define AE(i, j) A[((i) - 1) + (long)((j) - 1) * k]
                for (int j = 1; j <= k - w2; ++j) {
                    for (int i = 1; i <= k - w1; ++i) AE(i, j) = ME(i + w1, j + w2);
                    for (int i = 1; i <= w1; ++i) AE(i + k - w1, j) = -ME(i, j + w2);
                }
                for (int j = 1; j <= w2; ++j) {
                    for (int i = 1; i <= k - w1; ++i) AE(i, j + k - w2) = -ME(i + w1, j);
                    for (int i = 1; i <= w1; ++i) AE(i + k - w1, j + k - w2) = ME(i, j);
                }
#undef ME
#undef AE
                free(M);
                ch->minv[f] = A;
            }
            if (k % 2 != 0 && (w1 + w2) % 2 != 0) detrat = -detrat;
        }
        ch->det.sign *= detrat;
        ch->cfgsign = -config_sign(ch);
        ch->acc_shift++;
        free(old);
        return 1;
    }
    memcpy(ch->op, old, sizeof(oop_t) * (size_t)n);
    ch->outer_sst = old_outer;
    ch->window_position = old_pos;
    ch->tauwin_min = old_min;
    ch->tauwin_max = old_max;
    free(old);
    return 0;
}

/* ------------------------- binning accumulator ------------------------------ */
/* autocorracc_add (Accumulators.F90:298-340): level 0 takes the new value by
 * the Welford update of mean and squared deviations; whenever the low bits of
 * the count are ones, the finished bin of a level (the mean of its two halves)
 * rolls over into the next level.  Arrays [nlev][ncomp]; `n` values so far. */
static void oacc_add(long ncomp, int nlev, long n, const double *val, double *spill, double *mean,
                     double *sq)
{
    for (long i = 0; i < ncomp; ++i) {
        double v = val[i];
        spill[i] += v;
        double d1 = v - mean[i];
        mean[i] += d1 / (double)(n + 1);
        sq[i] += d1 * (v - mean[i]);
    }
    int lvl = 0;
    while ((n & 1) != 0) {
        if (lvl == nlev - 1) break;
        n /= 2;
        ++lvl;
        double *sp0 = spill + (long)(lvl - 1) * ncomp, *sp = spill + (long)lvl * ncomp;
        double *mn = mean + (long)lvl * ncomp, *s2 = sq + (long)lvl * ncomp;
        for (long i = 0; i < ncomp; ++i) {
            double v = sp0[i] / 2.0;
            sp0[i] = 0.0;
            sp[i] += v;
            double d1 = v - mn[i];
            mn[i] += d1 / (double)(n + 1);
            s2[i] += d1 * (v - mn[i]);
        }
    }
}

/* autocorracc_blocks (Accumulators.F90:342-362): variance of the bin means of
 * level l, sqdiff / (n_l - 1) with n_l = n >> l */
static void oacc_blocks(long ncomp, int nlev, long n, const double *sq, double *out)
{
    for (int l = 0; l < nlev; ++l) {
        for (long i = 0; i < ncomp; ++i)
            out[(long)l * ncomp + i] = sq[(long)l * ncomp + i] / (n > 1 ? (double)(n - 1) : 0.0);
        n /= 2;
    }
}

/* the accumulator on a given sequence (data [n][ncomp]): mean [ncomp] and
 * blocks [nlev][ncomp]; pinned by the closed form of the blocking variances */
void oracle_autocorr(long ncomp, int nlev, long n, const double *data, double *mean_out, double *blocks_out)
{
    double *sp = calloc((size_t)nlev * ncomp, sizeof(double));
    double *mn = calloc((size_t)nlev * ncomp, sizeof(double));
    double *s2 = calloc((size_t)nlev * ncomp, sizeof(double));
    for (long j = 0; j < n; ++j) oacc_add(ncomp, nlev, j, data + j * ncomp, sp, mn, s2);
    memcpy(mean_out, mn, sizeof(double) * (size_t)ncomp);
    oacc_blocks(ncomp, nlev, n, s2, blocks_out);
    free(sp), free(mn), free(s2);
}

/* ilog2 (Accumulators.F90:89-101): floor(log2(v)), -1 for v < 1 */
static int ilog2_l(long v)
{
    int r = -1;
    while (v > 0) ++r, v /= 2;
    return r;
}

/* ------------------------- measurements --------------------------------- */
/* MeasGtauRem (CTQMC.F90:2252-2440) */
static void meas_gtau(ochain_t *ch)
{
    int nf = NF(ch);
    int cpos[14], apos[14];
    memset(apos, 0, sizeof(apos));
    for (int ia = 0; ia < ch->nop; ++ia) {
        const oop_t *a = &ch->op[ia];
        if (a->ca != 2) continue;
        int f = FL(*a);
        apos[f]++;
        cpos[f] = 0;
        int k = ch->k[f];
        for (int ic = 0; ic < ch->nop; ++ic) {
            const oop_t *c = &ch->op[ic];
            if (FL(*c) != f || c->ca != 1) continue;
            cpos[f]++;
            double M = ch->minv[f][(cpos[f] - 1) + (long)(apos[f] - 1) * k];
            /* MeasGiw with NLookup_nfft = 0: measure_giw_nfft (CTQMC.F90:1015-1059)
             * sums M(c,a) exp(i w_n (tau_a - tau_c)) over every pair of the
             * flavour, w_n = (2n+1) pi / beta, with the direct sum of the NFFT
             * fallback (Nfft_z.F90:64-90); the sign convention is that of the
             * lookup path (CTQMC.F90:2403-2417 with :8816): cfgsign M, so that
             * G(i w) is the Fourier transform of the G(tau) and G_l estimators */
            if (ch->niw > 0) {
                const double mw = M * ch->cfgsign, d = a->tau - c->tau;
                double *gw = ch->giw + (long)f * ch->niw * 2;
                for (int n = 0; n < ch->niw; ++n) {
                    const double ph = 3.14159265358979323846 * (2.0 * n + 1.0) * d / ch->beta;
                    gw[2 * n] += mw * cos(ph);
                    gw[2 * n + 1] += mw * sin(ph);
                }
            }
            int idt;
            double x;
            if (a->tau - c->tau > 0.0) {
                idt = (int)((a->tau - c->tau) / ch->beta * (double)ch->ntau) + 1;
                x = 2.0 / ch->beta * (a->tau - c->tau) - 1.0;
            } else {
                idt = (int)((ch->beta + a->tau - c->tau) / ch->beta * (double)ch->ntau) + 1;
                x = 2.0 / ch->beta * (ch->beta + a->tau - c->tau) - 1.0;
                M = -M;
            }
            M = M * ch->cfgsign;
            double pl = 0.0, plm1 = 0.0, plp1 = 0.0;
            for (int l = 0; l < ch->nleg; ++l) {
                if (l == 0)
                    plp1 = 1.0;
                else if (l == 1)
                    plp1 = x;
                else {
                    int ll = l - 1;
                    plp1 = (2.0 * ll + 1.0) / (double)(ll + 1) * x * pl - (double)ll / (double)(ll + 1) * plm1;
                }
                ch->gleg[(long)f * ch->nleg + l] += M * plp1;
                plm1 = pl;
                pl = plp1;
            }
            if (idt > ch->ntau) idt = ch->ntau;
            (ch->accum == 2 ? ch->gmeas : ch->gtau_buf)[(long)f * ch->ntau + idt - 1] += M;
            /* step-resolved sums (CTQMC.F90:2394-2402), g_inmeas = cnt_meas */
            long mi = ch->cnt_meas - 1;
            if (mi < ch->nmstep) {
                if (ch->step_mean) ch->mstep[mi * nf + f] += M * ch->cfgsign;
                double r = fabs(a->tau - c->tau) / ch->beta;
                if (ch->step_mid && r >= 0.4 && r <= 0.6) ch->dstep[mi * nf + f] += M * ch->cfgsign;
                if (ch->step_sign) ch->sstep[mi] = ch->cfgsign;
            }
        }
    }
    /* gtau_accum%add() (CTQMC.F90:2439) */
    if (ch->accum == 2) {
        long nc = (long)nf * ch->ntau;
        for (long i = 0; i < nc; ++i) ch->gtau_buf[i] += ch->gmeas[i];
        oacc_add(nc, ch->nlev, ch->acc_n, ch->gmeas, ch->acc_spill, ch->acc_mean, ch->acc_sq);
        ch->acc_n++;
        memset(ch->gmeas, 0, sizeof(double) * (size_t)nc);
    }
}

/* MeasDensityMatrix_beta_half (CTQMC.F90:3088-3151) */
static void meas_dm(ochain_t *ch)
{
    int outer = ch->outer_sst, no = ch->sizes[outer];
    long ns = ch->nstates;
    if (ch->nop != 0) {
        beta_half_states(ch);
        int sst = ch->sst_b2, n = ch->sizes[sst], off = ch->offsets[sst];
        for (int i0 = 0; i0 < no; ++i0) {
            const double *bra = ch->bras_b2 + (long)i0 * ch->nmax;
            const double *ket = ch->kets_b2 + (long)i0 * ch->nmax;
            /* 1/trval(Trace / TLogTr(ket+bra, -cfgsign)) */
            double fac = ch->trace.sign * (-ch->cfgsign) * exp(ch->klog_b2[i0] + ch->blog_b2[i0] - ch->trace.log);
            for (int s1 = 0; s1 < n; ++s1) {
                double f1 = bra[s1] * fac;
                for (int s2 = 0; s2 < n; ++s2) ch->dm[(off + s1) + ns * (off + s2)] += f1 * ket[s2];
            }
        }
    } else {
        int off = ch->offsets[outer];
        for (int i0 = 0; i0 < no; ++i0) {
            olog_t p = ch->parttrace[i0];
            double v = p.sign * ch->trace.sign * exp(p.log - ch->trace.log);
            ch->dm[(off + i0) + ns * (off + i0)] -= ch->cfgsign * v;
        }
    }
}

/* MeasSign, MeasOuterHistograms, MeasExpOrder (CTQMC.F90:3264-3350) */
static void meas_misc(ochain_t *ch)
{
    ch->meas_sign += ch->cfgsign;
    int outer = ch->outer_sst, no = ch->sizes[outer], off = ch->offsets[outer];
    ch->outer_hist[outer] += 1.0;
    ch->sign_hist_sst[outer] += ch->cfgsign;
    double tc = 0.0;
    for (int i = 0; i < no; ++i) {
        ch->outer_state_hist[off + i] += 1.0;
        ch->sign_hist_states[off + i] += ch->trace.sign * ch->parttrace[i].sign * ch->cfgsign;
        tc += exp(ch->parttrace[i].log);
        ch->tc_states[off + i] += exp(ch->parttrace[i].log);
    }
    ch->tc_sst[outer] += tc;
    for (int f = 0; f < NF(ch); ++f) {
        int k = count_flavour(ch, f) / 2;
        if (k > ch->maxhisto) k = ch->maxhisto;
        ch->histo[(long)f * (ch->maxhisto + 1) + k] += 1.0;
    }
}

/* ------------------------ parameter derivatives -------------------------- */
/* d ln|w| / d theta of a configuration (no reference counterpart; pinned by
 * finite differences of local_trace and full_inverse at theta +- delta).
 * w = det(Delta) Tr[e^{-(beta - tau_n) H} O_n ... O_1 e^{-tau_1 H}]:
 *   d ln|det Delta| = sum_f tr(Delta_f^{-1} dDelta_f), dDelta from dF/dtheta;
 *   d ln|Tr| from tangent vectors dv carried beside every state vector v:
 *   over a segment dt in block s (eigenbasis, levels E_k),
 *   d e^{-dt H} = -(D o Phi) with D = V^T dH V and the divided difference
 *   Phi_kl = (e^{-dt E_k} - e^{-dt E_l}) / (E_l - E_k) (dt e^{-dt E_k} if equal);
 *   an operator block O maps v and dv alike. */
static double phi_div(double ek, double el, double dt)
{
    /* levels relative to the block's lowest; evaluated from the smaller one */
    double a = ek < el ? ek : el, y = fabs(ek - el) * dt;
    double g = y > 0.0 ? -expm1(-y) / y : 1.0;
    return dt * exp(-a * dt) * g;
}

/* v <- P v, dv_t <- P dv_t - (D_t o Phi) v over dt in block s (relative levels) */
static void prop_tangent(const ochain_t *ch, int s, double dt, double *v, double *dv, int ns)
{
    const double *E = ch->evals + ch->offsets[s];
    double *tmp = ch->tmp;
    for (int t = 0; t < ch->ntheta; ++t) {
        const double *D = ch->dblk + (long)t * ch->dbsize + ch->dboff[s];
        double *d = dv + (long)t * ch->nmax;
        for (int k = 0; k < ns; ++k) {
            double acc = exp(-(E[k] - E[0]) * dt) * d[k];
            for (int l = 0; l < ns; ++l)
                acc -= D[k + (long)l * ns] * phi_div(E[k] - E[0], E[l] - E[0], dt) * v[l];
            tmp[k] = acc;
        }
        memcpy(d, tmp, sizeof(double) * (size_t)ns);
    }
    for (int k = 0; k < ns; ++k) v[k] *= exp(-(E[k] - E[0]) * dt);
}

static void apply_tangent(const ochain_t *ch, const double *P, int nt, int ns, double *v, double *dv)
{
    double *w = ch->tmp;
    for (int r = 0; r < nt; ++r) {
        double acc = 0.0;
        for (int c = 0; c < ns; ++c) acc += P[r + (long)c * nt] * v[c];
        w[r] = acc;
    }
    memcpy(v, w, sizeof(double) * (size_t)nt);
    for (int t = 0; t < ch->ntheta; ++t) {
        double *d = dv + (long)t * ch->nmax;
        for (int r = 0; r < nt; ++r) {
            double acc = 0.0;
            for (int c = 0; c < ns; ++c) acc += P[r + (long)c * nt] * d[c];
            w[r] = acc;
        }
        memcpy(d, w, sizeof(double) * (size_t)nt);
    }
}

/* rescale v and its tangents by one factor (the largest entry of any of them) */
static void split_log_tangent(const ochain_t *ch, double *v, double *dv, int n, double *lf)
{
    double mv = 0.0;
    for (int i = 0; i < n; ++i) mv = fmax(mv, fabs(v[i]));
    for (int t = 0; t < ch->ntheta; ++t)
        for (int i = 0; i < n; ++i) mv = fmax(mv, fabs(dv[(long)t * ch->nmax + i]));
    if (mv > 0.0) {
        for (int i = 0; i < n; ++i) v[i] /= mv;
        for (int t = 0; t < ch->ntheta; ++t)
            for (int i = 0; i < n; ++i) dv[(long)t * ch->nmax + i] /= mv;
        *lf += log(mv);
    }
}

/* walk of the unit vector i0 of the outer superstate from tau = 0 through every
 * operator to beta with tangents: value, tangents and log scale of <i0| ... |i0> */
static void walk_tangent_full(const ochain_t *ch, int i0, double *val, double *dval, double *lgo)
{
    int s = ch->outer_sst, ns = ch->sizes[s];
    double *v = ch->ket, *dv = ch->dket;
    double lg = 0.0, t0 = 0.0;
    memset(v, 0, sizeof(double) * (size_t)ch->nmax);
    memset(dv, 0, sizeof(double) * (size_t)ch->nmax * ch->ntheta);
    v[i0] = 1.0;
    for (int j = 0; j < ch->nop; ++j) {
        const oop_t *o = &ch->op[j];
        int t = conn(ch, s, o->orb, o->spin, o->ca), nt = ch->sizes[t];
        const double *E = ch->evals + ch->offsets[s];
        prop_tangent(ch, s, o->tau - t0, v, dv, ns);
        lg += -(E[0] - ch->egs) * (o->tau - t0);
        apply_tangent(ch, opblk(ch, s, o->orb, o->spin, o->ca), nt, ns, v, dv);
        split_log_tangent(ch, v, dv, nt, &lg);
        s = t, ns = nt, t0 = o->tau;
    }
    const double *E = ch->evals + ch->offsets[s];
    prop_tangent(ch, s, ch->beta - t0, v, dv, ns);
    lg += -(E[0] - ch->egs) * (ch->beta - t0);
    *val = v[i0];
    for (int t = 0; t < ch->ntheta; ++t) dval[t] = dv[(long)t * ch->nmax + i0];
    *lgo = lg;
}

/* d ln|Tr| / d theta: sum over the outer states of the walks above */
static void trace_tangent(const ochain_t *ch, double *dlt)
{
    int no = ch->sizes[ch->outer_sst];
    double *val = malloc(sizeof(double) * (size_t)no), *lg = malloc(sizeof(double) * (size_t)no);
    double *dval = malloc(sizeof(double) * (size_t)no * ch->ntheta);
    double lmax = -DBL_MAX;
    for (int i0 = 0; i0 < no; ++i0) {
        walk_tangent_full(ch, i0, &val[i0], &dval[(long)i0 * ch->ntheta], &lg[i0]);
        if (lg[i0] > lmax) lmax = lg[i0];
    }
    double tr = 0.0;
    for (int i0 = 0; i0 < no; ++i0) tr += val[i0] * exp(lg[i0] - lmax);
    for (int t = 0; t < ch->ntheta; ++t) {
        double d = 0.0;
        for (int i0 = 0; i0 < no; ++i0) d += dval[(long)i0 * ch->ntheta + t] * exp(lg[i0] - lmax);
        dlt[t] = d / tr;
    }
    free(val), free(lg), free(dval);
}

/* d ln|det| / d theta = sum_f tr(Delta_f^{-1} dDelta_f) with a fresh LU */
static void det_tangent(ochain_t *ch, double *dld)
{
    for (int t = 0; t < ch->ntheta; ++t) dld[t] = 0.0;
    const double *ftau = ch->ftau;
    for (int f = 0; f < NF(ch); ++f) {
        int k = count_flavour(ch, f) / 2;
        if (k == 0) continue;
        double *D = malloc(sizeof(double) * (size_t)k * k), *inv = malloc(sizeof(double) * (size_t)k * k);
        double *dD = malloc(sizeof(double) * (size_t)k * k);
        build_delta(ch, f, D);
        olog_t d;
        lu_logdet(k, D, &d, inv);
        for (int t = 0; t < ch->ntheta; ++t) {
            ch->ftau = ch->dftau + (long)t * NF(ch) * ch->nftau;   /* dF/dtheta tables */
            build_delta(ch, f, dD);
            ch->ftau = ftau;
            double acc = 0.0;
            for (int i = 0; i < k; ++i)
                for (int j = 0; j < k; ++j) acc += inv[i + (long)j * k] * dD[j + (long)i * k];
            dld[t] += acc;
        }
        free(D), free(inv), free(dD);
    }
}

/* ket and bra of the outer state i0 at beta/2 with tangents (as
 * beta_half_states); returns the superstate at beta/2 */
static int half_tangent(const ochain_t *ch, int i0, double *k, double *dk, double *kl, double *b, double *db,
                        double *bl)
{
    double bh = ch->beta / 2.0;
    int s = ch->outer_sst, ns = ch->sizes[s];
    double lg = 0.0, t0 = 0.0;
    memset(k, 0, sizeof(double) * (size_t)ch->nmax);
    memset(dk, 0, sizeof(double) * (size_t)ch->nmax * ch->ntheta);
    k[i0] = 1.0;
    int j = 0;
    for (; j < ch->nop && ch->op[j].tau <= bh; ++j) {
        const oop_t *o = &ch->op[j];
        int t = conn(ch, s, o->orb, o->spin, o->ca), nt = ch->sizes[t];
        const double *E = ch->evals + ch->offsets[s];
        prop_tangent(ch, s, o->tau - t0, k, dk, ns);
        lg += -(E[0] - ch->egs) * (o->tau - t0);
        apply_tangent(ch, opblk(ch, s, o->orb, o->spin, o->ca), nt, ns, k, dk);
        split_log_tangent(ch, k, dk, nt, &lg);
        s = t, ns = nt, t0 = o->tau;
    }
    {
        const double *E = ch->evals + ch->offsets[s];
        prop_tangent(ch, s, bh - t0, k, dk, ns);
        lg += -(E[0] - ch->egs) * (bh - t0);
        split_log_tangent(ch, k, dk, ns, &lg);
    }
    *kl = lg;
    const int smid = s;
    s = ch->outer_sst, ns = ch->sizes[s], lg = 0.0, t0 = ch->beta;
    memset(b, 0, sizeof(double) * (size_t)ch->nmax);
    memset(db, 0, sizeof(double) * (size_t)ch->nmax * ch->ntheta);
    b[i0] = 1.0;
    for (int q = ch->nop - 1; q >= j; --q) {
        const oop_t *o = &ch->op[q];
        int t = conn(ch, s, o->orb, o->spin, 3 - o->ca), nt = ch->sizes[t];
        const double *E = ch->evals + ch->offsets[s];
        prop_tangent(ch, s, t0 - o->tau, b, db, ns);
        lg += -(E[0] - ch->egs) * (t0 - o->tau);
        apply_tangent(ch, opblk(ch, s, o->orb, o->spin, 3 - o->ca), nt, ns, b, db);
        split_log_tangent(ch, b, db, nt, &lg);
        s = t, ns = nt, t0 = o->tau;
    }
    {
        const double *E = ch->evals + ch->offsets[s];
        prop_tangent(ch, s, t0 - bh, b, db, ns);
        lg += -(E[0] - ch->egs) * (t0 - bh);
        split_log_tangent(ch, b, db, ns, &lg);
    }
    *bl = lg;
    return smid;
}

/* score-function sums of one measurement: s = d ln|w| / d theta;
 * Sum sign s, Sum s x (this measurement's G(tau) bins), Sum s x (this
 * measurement's density-matrix sample) and Sum d(sample)/dtheta at fixed
 * configuration (the density matrix at beta/2 depends on H explicitly). */
static void meas_score(ochain_t *ch)
{
    const int nth = ch->ntheta, nf = NF(ch);
    double sc[8], dlt[8];
    det_tangent(ch, sc);
    trace_tangent(ch, dlt);
    for (int t = 0; t < nth; ++t) {
        sc[t] += dlt[t];
        ch->sc_sum[t] += ch->cfgsign * sc[t];
    }
    /* G(tau) bins of this measurement, as meas_gtau */
    double *gm = calloc((size_t)nf * ch->ntau, sizeof(double));
    int cpos, apos[14];
    memset(apos, 0, sizeof(apos));
    for (int ia = 0; ia < ch->nop; ++ia) {
        const oop_t *a = &ch->op[ia];
        if (a->ca != 2) continue;
        int f = FL(*a);
        apos[f]++;
        cpos = 0;
        for (int ic = 0; ic < ch->nop; ++ic) {
            const oop_t *c = &ch->op[ic];
            if (FL(*c) != f || c->ca != 1) continue;
            cpos++;
            double M = ch->minv[f][(cpos - 1) + (long)(apos[f] - 1) * ch->k[f]];
            int idt;
            if (a->tau - c->tau > 0.0) {
                idt = (int)((a->tau - c->tau) / ch->beta * (double)ch->ntau) + 1;
            } else {
                idt = (int)((ch->beta + a->tau - c->tau) / ch->beta * (double)ch->ntau) + 1;
                M = -M;
            }
            if (idt > ch->ntau) idt = ch->ntau;
            gm[(long)f * ch->ntau + idt - 1] += M * ch->cfgsign;
        }
    }
    for (int t = 0; t < nth; ++t)
        for (long i = 0; i < (long)nf * ch->ntau; ++i) ch->sc_g[(long)t * nf * ch->ntau + i] += sc[t] * gm[i];
    free(gm);
    /* density matrix at beta/2 and its explicit derivative */
    if (ch->meas_dm) {
        int no = ch->sizes[ch->outer_sst], nmax = ch->nmax;
        long ns2 = (long)ch->nstates * ch->nstates, nst = ch->nstates;
        double *k = malloc(sizeof(double) * nmax), *b = malloc(sizeof(double) * nmax);
        double *dk = malloc(sizeof(double) * nmax * nth), *db = malloc(sizeof(double) * nmax * nth);
        for (int i0 = 0; i0 < no; ++i0) {
            double kl, bl;
            int sm = half_tangent(ch, i0, k, dk, &kl, b, db, &bl);
            int n = ch->sizes[sm], off = ch->offsets[sm];
            double fac = ch->trace.sign * (-ch->cfgsign) * exp(kl + bl - ch->trace.log);
            for (int s1 = 0; s1 < n; ++s1)
                for (int s2 = 0; s2 < n; ++s2) {
                    long e = (off + s1) + nst * (off + s2);
                    double smp = fac * b[s1] * k[s2];
                    for (int t = 0; t < nth; ++t) {
                        ch->sc_dm[t * ns2 + e] += sc[t] * smp;
                        ch->sc_dmx[t * ns2 + e] += fac * (db[(long)t * nmax + s1] * k[s2] + b[s1] * dk[(long)t * nmax + s2])
                                                   - smp * dlt[t];
                    }
                }
        }
        free(k), free(b), free(dk), free(db);
    }
}

/* Set the derivative tables: dblk [ntheta][sum n_s^2] (eigenbasis dH blocks,
 * column-major per block in superstate order), dftau [ntheta][nf][nftau];
 * score != 0 adds the score sums to every measurement. */
int oracle_set_derivatives(void *h, int ntheta, const double *dblk, const double *dftau, int score)
{
    ochain_t *ch = h;
    if (ntheta < 1 || ntheta > 8) return 1;
    free(ch->dboff), free(ch->dblk), free(ch->dftau);
    free(ch->sc_sum), free(ch->sc_g), free(ch->sc_dm), free(ch->sc_dmx);
    ch->ntheta = ntheta;
    ch->dboff = malloc(sizeof(long) * (size_t)ch->nsst);
    long o = 0;
    for (int s = 0; s < ch->nsst; ++s) ch->dboff[s] = o, o += (long)ch->sizes[s] * ch->sizes[s];
    ch->dbsize = o;
    ch->dblk = malloc(sizeof(double) * (size_t)(o * ntheta));
    memcpy(ch->dblk, dblk, sizeof(double) * (size_t)(o * ntheta));
    long nft = (long)NF(ch) * ch->nftau * ntheta;
    ch->dftau = malloc(sizeof(double) * (size_t)nft);
    memcpy(ch->dftau, dftau, sizeof(double) * (size_t)nft);
    ch->dket = realloc(ch->dket, sizeof(double) * (size_t)ch->nmax * ntheta);
    ch->score = score;
    ch->sc_sum = calloc((size_t)ntheta, sizeof(double));
    ch->sc_g = calloc((size_t)ntheta * NF(ch) * ch->ntau, sizeof(double));
    ch->sc_dm = calloc((size_t)ntheta * ch->nstates * ch->nstates, sizeof(double));
    ch->sc_dmx = calloc((size_t)ntheta * ch->nstates * ch->nstates, sizeof(double));
    return 0;
}

/* d ln|det| / d theta and d ln|Tr| / d theta of the current configuration:
 * out[2 t] determinant part, out[2 t + 1] trace part */
void oracle_dlogw(void *h, double *out)
{
    ochain_t *ch = h;
    double a[8], b[8];
    det_tangent(ch, a);
    trace_tangent(ch, b);
    for (int t = 0; t < ch->ntheta; ++t) out[2 * t] = a[t], out[2 * t + 1] = b[t];
}

/* score sums: ssum [ntheta], g [ntheta][nf][ntau], dm and dmx [ntheta][nstates^2] */
void oracle_score_sums(void *h, double *ssum, double *g, double *dm, double *dmx)
{
    ochain_t *ch = h;
    long nt = ch->ntheta, ns2 = (long)ch->nstates * ch->nstates;
    memcpy(ssum, ch->sc_sum, sizeof(double) * (size_t)nt);
    memcpy(g, ch->sc_g, sizeof(double) * (size_t)(nt * NF(ch) * ch->ntau));
    memcpy(dm, ch->sc_dm, sizeof(double) * (size_t)(nt * ns2));
    memcpy(dmx, ch->sc_dmx, sizeof(double) * (size_t)(nt * ns2));
}

/* ---------------------------- public API --------------------------------- */
typedef struct {
    int32_t kind, outcome, noper, outer_sst, window;
    double trace_log, trace_sign, det_log, det_sign, cfgsign;
} omove_rec_t;

/* one Metropolis step with the dispatch of ctqmc_warmup / ctqmc_measure
 * (CTQMC.F90:8213-8254, 8432-8533) followed by the window-shift rule */
static void one_move(ochain_t *ch, omove_rec_t *rec)
{
    double r = orng_grnd(&ch->rng);
    int kind, out;
    if (r < (1.0 - ch->p_shift - ch->p_glob) / 2.0) {
        kind = MV_ADD;
        if (orng_grnd(&ch->rng) < ch->p_4op) abort();
        (void)(orng_grnd(&ch->rng) < ch->p_outer);
        out = step_add(ch);
    } else if (r <= 1.0 - ch->p_shift - ch->p_glob) {
        kind = MV_REM;
        if (orng_grnd(&ch->rng) < ch->p_4op) abort();
        (void)(orng_grnd(&ch->rng) < ch->p_outer);
        out = step_rem(ch);
    } else if (r <= 1.0 - ch->p_shift) {
        kind = MV_GLOB;
        out = step_glob(ch);
    } else {
        kind = MV_SHIFT;
        out = step_shift(ch);
    }
    if ((ch->try_add + ch->try_rem) % ch->nslide == 0) shift_window(ch);
    if (rec) {
        rec->kind = kind;
        rec->outcome = out;
        rec->noper = ch->nop;
        rec->outer_sst = ch->outer_sst;
        rec->window = ch->window_position;
        rec->trace_log = ch->trace.log;
        rec->trace_sign = ch->trace.sign;
        rec->det_log = ch->det.log;
        rec->det_sign = ch->det.sign;
        rec->cfgsign = ch->cfgsign;
    }
}

/* Create a chain: init_solver's init_Trace (Trace.F90:340-701) including the
 * initial outer-superstate draw. */
void *oracle_create(int nb, int nsst, int nmax, const int *sizes, const int *offsets,
                    const double *evals, double egs, const int *connect, const long *opoff,
                    const double *ops, const double *evec, const long *evoff,
                    const int *first_state, const int *state_sst, int nftau, const double *ftau,
                    double beta, const double *dparams, const long *iparams, unsigned seed)
{
    ochain_t *ch = calloc(1, sizeof(ochain_t));
    ch->nb = nb, ch->nsst = nsst, ch->nmax = nmax;
    ch->nstates = 1 << (2 * nb);
    ch->sizes = sizes, ch->offsets = offsets, ch->evals = evals, ch->egs = egs;
    ch->connect = connect, ch->opoff = opoff, ch->ops = ops, ch->evec = evec, ch->evoff = evoff;
    ch->first_state = first_state, ch->state_sst = state_sst;
    ch->nftau = nftau, ch->ftau = ftau, ch->beta = beta;
    /* dparams: p_shift, p_glob, p_outer, p_4op, TaudiffMax
     * iparams: NCorr, Ntau, NLegMax, MaxHisto, has_szt, MeasDensityMatrix */
    ch->p_shift = dparams[0], ch->p_glob = dparams[1], ch->p_outer = dparams[2], ch->p_4op = dparams[3];
    ch->ntau = (int)iparams[1], ch->nleg = (int)iparams[2], ch->maxhisto = (int)iparams[3];
    ch->has_szt = (int)iparams[4], ch->meas_dm = (int)iparams[5];
    /* iparams[6] AccumGtau, iparams[7] Nmeas (levels = ilog2(Nmeas), CTQMC.F90:497-500) */
    ch->accum = (int)iparams[6];
    ch->nlev = ilog2_l(iparams[7]);
    if (ch->nlev < 1) ch->nlev = 1;
    /* iparams[8..10] Gtau_mean_step, Gtau_mid_step, sign_step (arrays of Nmeas) */
    ch->step_mean = iparams[8] != 0, ch->step_mid = iparams[9] != 0, ch->step_sign = iparams[10] != 0;
    ch->nmstep = (ch->step_mean || ch->step_mid || ch->step_sign) ? (int)iparams[7] : 0;
    /* iparams[11] Niw when MeasGiw is set, else 0 */
    ch->niw = (int)iparams[11];
    ch->cap = 1024;
    ch->op = malloc(sizeof(oop_t) * (size_t)ch->cap);
    ch->ket = malloc(sizeof(double) * (size_t)nmax);
    ch->tmp = malloc(sizeof(double) * (size_t)nmax);
    ch->kets_b2 = malloc(sizeof(double) * (size_t)nmax * nmax);
    ch->bras_b2 = malloc(sizeof(double) * (size_t)nmax * nmax);
    ch->klog_b2 = malloc(sizeof(double) * (size_t)nmax);
    ch->blog_b2 = malloc(sizeof(double) * (size_t)nmax);
    ch->parttrace = malloc(sizeof(olog_t) * (size_t)nmax);
    ch->tpart = malloc(sizeof(olog_t) * (size_t)nmax);
    int nf = 2 * nb;
    for (int f = 0; f < nf; ++f) ch->minv[f] = malloc(sizeof(double)), ch->k[f] = 0;
    ch->gtau_buf = calloc((size_t)nf * ch->ntau, sizeof(double));
    ch->gleg = calloc((size_t)nf * ch->nleg, sizeof(double));
    ch->giw = calloc((size_t)nf * (ch->niw > 0 ? ch->niw : 1) * 2, sizeof(double));
    ch->gmeas = calloc((size_t)nf * ch->ntau, sizeof(double));
    ch->accpair = calloc((size_t)nf * ACC_CAP, sizeof(double));
    ch->aptau = calloc(APT_CAP, sizeof(double));
    ch->naccpair = ch->ntau < ACC_CAP ? ch->ntau : ACC_CAP;   /* NGtau bins of [0, beta] */
    ch->accpair_max = beta;
    ch->apt_index = 1, ch->apt_n = 0;
    ch->mstep = calloc((size_t)nf * (ch->nmstep + 1), sizeof(double));
    ch->dstep = calloc((size_t)nf * (ch->nmstep + 1), sizeof(double));
    ch->sstep = calloc((size_t)ch->nmstep + 1, sizeof(double));
    ch->acc_spill = calloc((size_t)nf * ch->ntau * ch->nlev, sizeof(double));
    ch->acc_mean = calloc((size_t)nf * ch->ntau * ch->nlev, sizeof(double));
    ch->acc_sq = calloc((size_t)nf * ch->ntau * ch->nlev, sizeof(double));
    ch->dm = calloc((size_t)ch->nstates * ch->nstates, sizeof(double));
    ch->histo = calloc((size_t)nf * (ch->maxhisto + 1), sizeof(double));
    ch->outer_hist = calloc((size_t)nsst, sizeof(double));
    ch->sign_hist_sst = calloc((size_t)nsst, sizeof(double));
    ch->tc_sst = calloc((size_t)nsst, sizeof(double));
    ch->outer_state_hist = calloc((size_t)ch->nstates, sizeof(double));
    ch->sign_hist_states = calloc((size_t)ch->nstates, sizeof(double));
    ch->tc_states = calloc((size_t)ch->nstates, sizeof(double));
    ch->cfgsign = -1.0;
    ch->det.log = 0.0, ch->det.sign = 1.0;
    orng_seed(&ch->rng, seed);
    init_window(ch, dparams[4]);
    /* NSlide = max(2*NBands, floor(NCorr/num_windows)), CTQMC.F90:8069 */
    long ns = (long)floor((double)iparams[0] / (double)ch->num_windows);
    ch->nslide = ns > 2 * nb ? ns : 2 * nb;
    /* initial outer superstate proportional to the empty trace */
    double *vals = malloc(sizeof(double) * (size_t)nsst);
    double sum = 0.0;
    for (int s = 0; s < nsst; ++s) {
        ch->outer_sst = s;
        olog_t t = local_trace(ch, NULL);
        vals[s] = t.sign * exp(t.log);
    }
    for (int s = 0; s < nsst; ++s) sum += vals[s];
    for (int s = 0; s < nsst; ++s) vals[s] /= sum;
    double r = orng_grnd(&ch->rng), acc = 0.0;
    for (int s = 0; s < nsst; ++s) {
        ch->outer_sst = s;
        acc += vals[s];
        if (acc >= r && vals[s] != 0.0) break;
    }
    free(vals);
    accept_trace(ch, local_trace(ch, ch->tpart));
    return ch;
}

void oracle_destroy(void *h)
{
    ochain_t *ch = h;
    for (int f = 0; f < NF(ch); ++f) free(ch->minv[f]);
    free(ch->op), free(ch->ket), free(ch->tmp), free(ch->kets_b2), free(ch->bras_b2);
    free(ch->klog_b2), free(ch->blog_b2), free(ch->parttrace), free(ch->tpart);
    free(ch->giw);
    free(ch->gtau_buf), free(ch->gleg), free(ch->dm), free(ch->histo), free(ch->outer_hist);
    free(ch->gmeas), free(ch->acc_spill), free(ch->acc_mean), free(ch->acc_sq);
    free(ch->accpair), free(ch->aptau), free(ch->mstep), free(ch->dstep), free(ch->sstep);
    free(ch->sign_hist_sst), free(ch->tc_sst), free(ch->outer_state_hist);
    free(ch->sign_hist_states), free(ch->tc_states);
    free(ch->dboff), free(ch->dblk), free(ch->dftau), free(ch->dket), free(ch->symmoves);
    free(ch->sc_sum), free(ch->sc_g), free(ch->sc_dm), free(ch->sc_dmx);
    free(ch);
}

/* User symmetry moves (NSymMove, SymMove01..; Trace.F90:471-478): nsym
 * permutations of the 1-based flavours, each followed in the table by its
 * inverse as the reference stores them (rows nsym + i). */
void oracle_set_symmoves(void *h, int nsym, const int *perms)
{
    ochain_t *ch = h;
    int nf = NF(ch);
    free(ch->symmoves);
    ch->nsym = nsym;
    ch->symmoves = malloc(sizeof(int) * (size_t)(2 * nsym * nf + 1));
    for (int i = 0; i < nsym; ++i)
        for (int j = 0; j < nf; ++j) {
            ch->symmoves[i * nf + j] = perms[i * nf + j];
            ch->symmoves[(nsym + i) * nf + perms[i * nf + j] - 1] = j + 1;
        }
}

/* Re-initialise the window with a new TaudiffMax (set_taudiffmax,
 * CTQMC.F90:8180-8185); NSlide keeps its value as in the reference. */
void oracle_set_taudiffmax(void *h, double tdm)
{
    ochain_t *ch = h;
    if (tdm > 0) init_window(ch, tdm);
}

/* Run n moves; rec may be NULL or an array of n records. */
void oracle_moves(void *h, long n, omove_rec_t *rec)
{
    ochain_t *ch = h;
    for (long i = 0; i < n; ++i) {
        if (ch->nop + 2 >= ch->cap) {
            ch->cap *= 2;
            ch->op = realloc(ch->op, sizeof(oop_t) * (size_t)ch->cap);
        }
        one_move(ch, rec ? rec + i : NULL);
    }
}

/* reset counters (init_counters, CTQMC.F90:847-899): move counters, AccPair,
 * the calibration list index */
void oracle_reset_counters(void *h)
{
    ochain_t *ch = h;
    ch->try_add = ch->try_rem = ch->try_glob = ch->try_shift = 0;
    ch->acc_add = ch->acc_rem = ch->acc_glob = ch->acc_shift = 0;
    ch->accqn_add = ch->accqn_rem = ch->accqn_glob = 0;
    memset(ch->accpair, 0, sizeof(double) * (size_t)NF(ch) * ACC_CAP);
    ch->apt_index = 1;
}

/* ctqmc_warmup's AccPair over NGtau bins of [0, beta] (CTQMC.F90:8199-8204) */
void oracle_begin_warmup(void *h)
{
    ochain_t *ch = h;
    ch->naccpair = ch->ntau < ACC_CAP ? ch->ntau : ACC_CAP;
    ch->accpair_max = ch->beta;
    oracle_reset_counters(h);
}

/* ctqmc_calibrate (CTQMC.F90:8088-8178): moves until, checked every
 * min(1000, nprogress) moves, the distance list holds phase1pairnum entries
 * (list) or every flavour has phase2pairnum accepted pairs in AccPair over
 * phase2taugrid bins of [0, taudiff_max] (not list); returns the moves done */
long oracle_calibrate(void *h, int list, int phase1pairnum, int phase2taugrid, int phase2pairnum, int nprogress)
{
    ochain_t *ch = h;
    if (list) {
        ch->apt_n = phase1pairnum;
    } else {
        ch->apt_n = 0;
        memset(ch->accpair, 0, sizeof(double) * (size_t)NF(ch) * ACC_CAP);
        ch->naccpair = phase2taugrid;
        ch->accpair_max = ch->taudiff_max;
    }
    int every = nprogress < 1000 ? nprogress : 1000;
    long i = 0;
    for (;;) {
        ++i;
        oracle_moves(h, 1, NULL);
        if (i % every != 0) continue;
        if (list) {
            if (ch->apt_index > ch->apt_n) return i;
        } else {
            int ok = 1;
            for (int f = 0; f < NF(ch) && ok; ++f) {
                double n = 0.0;
                for (int b = 0; b < ch->naccpair; ++b) n += ch->accpair[(long)f * ACC_CAP + b];
                if (n < (double)phase2pairnum) ok = 0;
            }
            if (ok) return i;
        }
    }
}

/* accepted-pair statistics: the distance list (returns its length) and AccPair
 * [nf][naccpair] */
int oracle_pair_stats(void *h, double *aptau, double *accpair, int *naccpair)
{
    ochain_t *ch = h;
    int n = ch->apt_index - 1 < ch->apt_n ? ch->apt_index - 1 : ch->apt_n;
    if (aptau) memcpy(aptau, ch->aptau, sizeof(double) * (size_t)(n > 0 ? n : 0));
    if (naccpair) *naccpair = ch->naccpair;
    if (accpair)
        for (int f = 0; f < NF(ch); ++f)
            memcpy(accpair + (long)f * ch->naccpair, ch->accpair + (long)f * ACC_CAP,
                   sizeof(double) * (size_t)ch->naccpair);
    return n > 0 ? n : 0;
}

/* MeasOuterHistograms accumulators (raw) */
void oracle_outer_histos(void *h, double *outer_state, double *sign_sst, double *sign_state, double *contrib_sst,
                         double *contrib_state)
{
    ochain_t *ch = h;
    memcpy(outer_state, ch->outer_state_hist, sizeof(double) * (size_t)ch->nstates);
    memcpy(sign_sst, ch->sign_hist_sst, sizeof(double) * (size_t)ch->nsst);
    memcpy(sign_state, ch->sign_hist_states, sizeof(double) * (size_t)ch->nstates);
    memcpy(contrib_sst, ch->tc_sst, sizeof(double) * (size_t)ch->nsst);
    memcpy(contrib_state, ch->tc_states, sizeof(double) * (size_t)ch->nstates);
}

/* step-resolved arrays (raw): [nmstep][nf], [nmstep][nf], [nmstep]; returns nmstep */
int oracle_steps(void *h, double *mstep, double *dstep, double *sstep)
{
    ochain_t *ch = h;
    size_t n = (size_t)ch->nmstep, nf = (size_t)NF(ch);
    if (mstep) memcpy(mstep, ch->mstep, sizeof(double) * n * nf);
    if (dstep) memcpy(dstep, ch->dstep, sizeof(double) * n * nf);
    if (sstep) memcpy(sstep, ch->sstep, sizeof(double) * n);
    return ch->nmstep;
}

/* ctqmc_measure's sampling loop: nmeas x (ncorr moves + measurements) */
void oracle_measure(void *h, long nmeas, long ncorr)
{
    ochain_t *ch = h;
    for (long m = 0; m < nmeas; ++m) {
        oracle_moves(h, ncorr, NULL);
        ch->cnt_meas++;
        meas_misc(ch);
        if (ch->meas_dm) meas_dm(ch);
        meas_gtau(ch);
        if (ch->score) meas_score(ch);
    }
}

/* raw G(i w_n) sums [nf][niw][2]; returns niw (0: MeasGiw off) */
int oracle_giw(void *h, double *out)
{
    ochain_t *ch = h;
    if (out && ch->niw > 0) memcpy(out, ch->giw, sizeof(double) * (size_t)NF(ch) * ch->niw * 2);
    return ch->niw;
}

/* binning blocks of G(tau) (AccumGtau = 2): [nlev][nf*ntau]; returns nlev */
int oracle_gtau_blocks(void *h, double *out)
{
    ochain_t *ch = h;
    if (out) oacc_blocks((long)NF(ch) * ch->ntau, ch->nlev, ch->acc_n, ch->acc_sq, out);
    return ch->nlev;
}

/* raw accumulators and counters (before normalisation) */
void oracle_accumulators(void *h, double *gtau, double *gleg, double *dm, double *histo,
                         double *outer_hist, double *scalars)
{
    ochain_t *ch = h;
    int nf = NF(ch);
    memcpy(gtau, ch->gtau_buf, sizeof(double) * (size_t)nf * ch->ntau);
    memcpy(gleg, ch->gleg, sizeof(double) * (size_t)nf * ch->nleg);
    memcpy(dm, ch->dm, sizeof(double) * (size_t)ch->nstates * ch->nstates);
    memcpy(histo, ch->histo, sizeof(double) * (size_t)nf * (ch->maxhisto + 1));
    memcpy(outer_hist, ch->outer_hist, sizeof(double) * (size_t)ch->nsst);
    scalars[0] = ch->meas_sign;
    scalars[1] = (double)ch->cnt_meas;
    scalars[2] = (double)ch->acc_add, scalars[3] = (double)ch->try_add;
    scalars[4] = (double)ch->acc_rem, scalars[5] = (double)ch->try_rem;
    scalars[6] = (double)ch->acc_glob, scalars[7] = (double)ch->try_glob;
    scalars[8] = (double)ch->acc_shift, scalars[9] = (double)ch->try_shift;
}

/* current configuration: returns n; fills up to cap entries */
int oracle_get_config(void *h, int cap, double *tau, int *orb, int *spin, int *ca, int *outer,
                      double *weights)
{
    ochain_t *ch = h;
    for (int i = 0; i < ch->nop && i < cap; ++i)
        tau[i] = ch->op[i].tau, orb[i] = ch->op[i].orb, spin[i] = ch->op[i].spin, ca[i] = ch->op[i].ca;
    *outer = ch->outer_sst;
    weights[0] = ch->trace.log, weights[1] = ch->trace.sign;
    weights[2] = ch->det.log, weights[3] = ch->det.sign, weights[4] = ch->cfgsign;
    weights[5] = (double)ch->window_position;
    return ch->nop;
}

/* inverse hybridisation block of flavour f (column-major k x k), returns k */
int oracle_get_minv(void *h, int f, double *out)
{
    ochain_t *ch = h;
    int k = ch->k[f];
    if (out) memcpy(out, ch->minv[f], sizeof(double) * (size_t)k * k);
    return k;
}

/* Evaluate a given configuration like set_mc_config_from_arrays
 * (Trace.F90:6872-7000): full trace, full inverse, sign.  The chain adopts it. */
void oracle_set_config(void *h, int n, const double *tau, const int *orb, const int *spin,
                       const int *ca, int outer, double *weights)
{
    ochain_t *ch = h;
    if (n + 2 >= ch->cap) {
        ch->cap = 2 * n + 16;
        ch->op = realloc(ch->op, sizeof(oop_t) * (size_t)ch->cap);
    }
    ch->nop = n;
    for (int i = 0; i < n; ++i)
        ch->op[i].tau = tau[i], ch->op[i].orb = orb[i], ch->op[i].spin = spin[i], ch->op[i].ca = ca[i];
    ch->outer_sst = outer;
    ch->window_position = 0;
    set_window(ch);
    accept_trace(ch, local_trace(ch, ch->tpart));
    double *ninv[14];
    int nk[14];
    ch->det = full_inverse(ch, ninv, nk);
    for (int f = 0; f < NF(ch); ++f) {
        free(ch->minv[f]);
        ch->minv[f] = ninv[f];
        ch->k[f] = nk[f];
    }
    ch->cfgsign = -config_sign(ch);
    oracle_get_config(h, 0, NULL, NULL, NULL, NULL, &outer, weights);
}

/* one-shot helpers used by the RNG tests */
void oracle_rng_stream(unsigned seed, long n, const int *kind, const int *lo, const int *hi,
                       double *out)
{
    orng_t r;
    orng_seed(&r, seed);
    for (long i = 0; i < n; ++i) {
        if (kind[i] == 0)
            out[i] = orng_grnd(&r);
        else if (kind[i] == 1)
            out[i] = (double)orng_randint(&r, lo[i], hi[i]);
        else
            out[i] = (double)orng_next(&r);
    }
}
