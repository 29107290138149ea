// Batched CT-HYB (matrix/"M-V" Z-space) Monte Carlo engine for MI355X (gfx950):
// the device code lives in ctqmc_kernels.hpp; this translation unit adds the
// host side, i.e. the C-ABI of include/w2dyn_ctqmc.h.
#include "ctqmc_kernels.hpp"

// ===========================================================================
// host side: the C-ABI (include/w2dyn_ctqmc.h)
// ===========================================================================
#include <algorithm>
#include <atomic>
#include <csignal>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <sstream>
#include <vector>

#include "model.hpp"

struct w2d_ctqmc {
    std::string params;
    w2d::LocalModel model;
    w2d::DevModel dm{};
    w2d::ChainLayout L{};
    int nchains = 0, nm = 0, device = 0, kcap = 0, ntau = 0, nleg = 0, maxhisto = 0;
    int lds_max = 64 * 1024;   // LDS bytes a workgroup may allocate on this device
    bool moved = false;        // chains have taken steps since create
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
    // abort word: pinned host memory the sweep kernel polls (request_abort)
    int *abort_host = nullptr;
    int *abort_dev = nullptr;
    int aborted = 0;           // of the last warmup / measure (CTQMC.F90:8648-8658)
    int regrows = 0;           // capacity doublings so far
    void *scratch = nullptr;   // device address of the static scratch on this handle's GPU
    bool grow = true;          // capacity doubles on demand (create's capacity >= 0)
    int3 calib{0, 1, 0};       // calibration mode of the sweeps: phase, check period, pairs per flavour
    // launch machinery: chains ordered longest first before every sweep launch,
    // the launches of a warmup / measure pass captured into one HIP graph,
    // moves per chain and launch (0: whole measurement intervals of ~4000)
    int nm_min = 0;            // smallest trace engine (padded block size) the model allows
    int *order = nullptr;      // device: chain of each workgroup
    bool order_chains = false, graphs = true;   // chain ordering: opt-in
    long long per_launch = 0;
    // MeasGiw: raw G(i w_n) sums [nchains][2N][niw][2] and the per (chain,
    // flavour) visit record of giw_kernel [nchains][2N][2]; niw = 0: off
    int niw = 0;
    double *giw = nullptr;
    long long *giw_seen = nullptr;
    ~w2d_ctqmc();
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

constexpr int KCAP_MAX = 4096;   // pairs per flavour a chain may grow to
constexpr int ACCPAIR_BINS_MAX = 1024;   // accepted-pair histogram bins (NAccPair) at most, beyond Ntau
constexpr int APT_CAP = 65536;           // accepted-pair distances a calibration list holds at most
constexpr int GIW_NIW_MAX = 1 << 20;     // Matsubara frequencies of MeasGiw at most

// Static device scratch of the allocation-free entry point (16 MiB per GPU,
// held until the process ends).  In host code `g_scratch` names a host-side
// stand-in, not device memory: kernels get its device address, looked up once
// per device with hipGetSymbolAddress (scratch_for_device).  Every call on one
// GPU shares the array, so calls that overlap on one GPU race.
__device__ __align__(256) char g_scratch[16u << 20];
constexpr int MAX_DEVICES = 64;
void *g_scratch_dev[MAX_DEVICES];
std::mutex g_scratch_mu;

// device address of g_scratch on the current device `dev` (looked up once)
int scratch_for_device(int dev, void **out)
{
    if (dev < 0 || dev >= MAX_DEVICES) return W2D_ERR_ARG;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(g_scratch));
        if (e != hipSuccess || !p) {
            std::fprintf(stderr, "w2d_ctqmc: hipGetSymbolAddress(g_scratch) failed: %s\n", hipGetErrorString(e));
            return W2D_ERR_HIP;
        }
        g_scratch_dev[dev] = p;
    }
    *out = g_scratch_dev[dev];
    return W2D_OK;
}

// abort words of the live handles, raised together by the signal handler
constexpr int MAX_TRAPPED = 64;
std::atomic<int *> g_abort_words[MAX_TRAPPED];

void register_abort_word(int *w)
{
    for (auto &slot : g_abort_words) {
        int *expect = nullptr;
        if (slot.compare_exchange_strong(expect, w)) return;
    }
}

void unregister_abort_word(int *w)
{
    for (auto &slot : g_abort_words) {
        int *expect = w;
        if (slot.compare_exchange_strong(expect, nullptr)) return;
    }
}

extern "C" void w2d_signal_handler(int)
{
    for (auto &slot : g_abort_words) {
        int *w = slot.load();
        if (w) __atomic_store_n(w, 1, __ATOMIC_SEQ_CST);
    }
}

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
    hd.wvalid = 0;
}

int nbound_for(double tdm, double beta)
{
    if (tdm <= 0.0) return 3;
    int nw = (int)std::floor(beta / tdm) - 1;
    return std::max(nw, 1) + 2;
}

// Byte layout of a chain block and of its LDS for `nbound` window boundaries
// and `kc` pairs per flavour.
int compute_layout(const w2d_ctqmc *h, int nbound, int kc, w2d::ChainLayout &L)
{
    using namespace w2d;
    L = ChainLayout{};
    const int nf = 2 * h->model.nbands, nm = h->nm;
    L.kcap = kc;
    L.cap = 2 * nf * kc;
    L.lcap = L.cap + 2;
    L.nbound = nbound;
    L.kmat = std::max<long>((long)kc * kc, (long)nm * nm);
    long o = 0;
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
    L.fcs = take((long)nf * kc * 4);
    L.fas = take((long)nf * kc * 4);
    L.minv = take((long)nf * kc * kc * 8);
    L.scratch = take((L.kmat + 4L * L.cap) * 8);
    L.kst = take((long)nm * nm * 8);
    L.bnd = take((long)nbound * nm * nm * 8);
    L.bndlog = take((long)nbound * 8);
    L.bndsst = take((long)nbound * 4);
    L.part = take(2L * nm * 8 + 16);
    L.kwork = take(nm > 16 ? 2L * nm * nm * 8 : 8);
    const int nth = h->dm.ntheta;
    L.dwork = take(nth > 0 ? ((4L + 3L * nth) * nm * nm + nm) * 8 : 8);
    L.ltau = take((long)L.lcap * 8);
    L.lcode = take((long)L.lcap * 4);
    L.gtau = take((long)nf * h->ntau * 8);
    L.gleg = take((long)nf * h->nleg * 8);
    L.dm = take(h->dm.dmsize * 8);
    L.histo = take((long)nf * (h->maxhisto + 1) * 8);
    L.outer = take((long)h->model.nsst * 8);
    const bool bins = h->dm.accum == 2;
    L.gbuf = take((bins || h->dm.score) ? (long)nf * h->ntau * 8 : 8);
    L.gacc = take(bins ? 3L * h->dm.nlev * nf * h->ntau * 8 : 8);
    const long nst = h->model.nstates, nss = h->model.nsst;
    L.ohst = take(nst * 8);
    L.shss = take(nss * 8);
    L.shst = take(nst * 8);
    L.tcss = take(nss * 8);
    L.tcst = take(nst * 8);
    L.naccpair_cap = std::max(h->ntau, ACCPAIR_BINS_MAX);
    L.accpair = take((long)nf * L.naccpair_cap * 8);
    const long nms = h->dm.nmstep;
    L.gmstep = take(h->dm.step_mean ? nms * nf * 8 : 8);
    L.gdstep = take(h->dm.step_mid ? nms * nf * 8 : 8);
    L.sstep = take(h->dm.step_sign ? nms * 8 : 8);
    L.apt_cap = APT_CAP;
    L.aptau = take((long)L.apt_cap * 8);
    const bool sc = h->dm.score && nth > 0;
    L.scsum = take(sc ? nth * 8L : 8);
    L.scg = take(sc ? (long)nth * nf * h->ntau * 8 : 8);
    L.scdm = take(sc ? nth * h->dm.dmsize * 8 : 8);
    L.scdmx = take(sc ? nth * h->dm.dmsize * 8 : 8);
    L.stride = o;
    // dynamic LDS
    const w2d::LocalModel &m = h->model;
    long q = 0;
    auto lds = [&](long bytes) {
        long at = q;
        q = (q + bytes + 15) / 16 * 16;
        return at;
    };
    // list capacity: the window's share of the operator capacity with a margin
    const int nwin = std::max(nbound - 2, 1);
    const int lc = (int)std::min<long>(std::min<long>(L.cap + 2, 1024),
                                       std::max<long>(64, 2L * L.cap / (nwin + 1) + 32));
    L.lcap_lds = lc;
    L.l_hdr = lds(sizeof(ChainHdr));
    L.l_mt = lds(624 * 4);
    L.l_sz = lds((long)m.nsst * 4);
    L.l_off = lds((long)m.nsst * 4);
    // Connect and operator-block offsets: LDS for blocks up to 16 states, HBM
    // above (with them a chain of the 5-orbital model fits 20 KB, eight per CU)
    L.l_cn = nm <= 16 ? lds((long)m.nsst * m.nbands * 4 * 4) : -1;
    L.l_oo = nm <= 16 ? lds((long)m.nsst * m.nbands * 4 * 4) : -1;
    // transition records and energies: LDS for blocks up to 16 states, HBM
    // above (there the block products dominate and LDS buys chains per CU)
    L.l_tr = nm <= 16 ? lds((long)m.nsst * m.nbands * 4 * 16) : -1;
    L.l_ev = nm <= 16 ? lds((long)m.nstates * 8) : -1;
    L.l_wtau = lds((long)lc * 8);
    L.l_wcode = lds((long)lc * 4);
    L.l_ttau = lds((long)lc * 8);
    L.l_tcode = lds((long)lc * 4);
    L.l_pc = lds((long)lc * 4);
    L.l_dvec = lds((long)nm * 8);
    const long qn = std::max<long>(kc + 2, m.nsst + 1);
    L.l_q = lds(qn * 8);
    L.l_r = lds(qn * 8);
    L.l_col = lds((long)(kc + 2) * 8);
    L.l_row = lds((long)(kc + 2) * 8);
    L.l_pl = lds((long)(2 * nm + 2) * 8);
    L.l_k = nm <= 16 ? lds(2L * nm * nm * 8) : -1;
    L.l_strip = nm > 16 ? lds(128L * 16 * 8) : -1;
    L.l_legc = lds(2L * std::max(h->nleg, 1) * 8);
    // blocks of up to 4 states: quad-walk chunk buffers (16 operators)
    if (nm <= 4) {
        L.l_qa = lds(16L * 16 * 8);
        L.l_qlc = lds(16L * 8);
        L.l_qrec = lds(16L * 16);
        L.l_qsst = lds(16L * 4);
    } else {
        L.l_qa = L.l_qlc = L.l_qrec = L.l_qsst = -1;
    }
    L.sc = nm <= 4 ? std::min(64, lc - 2) : 0;
    if (L.sc > 0) {
        L.l_rk = lds((long)(L.sc + 1) * nm * nm * 8);
        L.l_lb = lds((long)(L.sc + 1) * nm * nm * 8);
        L.l_rkx = lds((long)(L.sc + 1) * 8);
        L.l_lbx = lds((long)(L.sc + 1) * 8);
        L.l_rks = lds((long)(L.sc + 1) * 4);
        L.l_lbs = lds((long)(L.sc + 1) * 4);
    } else {
        L.l_rk = L.l_lb = L.l_rkx = L.l_lbx = L.l_rks = L.l_lbs = -1;
    }
    const long opsb = (long)m.ops.size() * 8;
    // blocks of up to 4 states: the kernels read operator blocks from LDS only
    // operator blocks join LDS when the total stays within 64 KB (more would
    // cut the number of chains resident per CU); blocks of up to 4 states always
    // (above 16 states they stay in global memory: the MFMA walks read them
    // with global instructions)
    L.l_ops = (nm <= 4 || (nm <= 16 && q + opsb <= 64 * 1024)) ? lds(opsb) : -1;
    L.lds_bytes = q;
    if (L.lds_bytes > h->lds_max) return W2D_ERR_CAPACITY;
    L.hdrs = L.stride * h->nchains;
    return W2D_OK;
}

size_t chains_bytes(const w2d_ctqmc *h, const w2d::ChainLayout &L)
{
    return (size_t)L.hdrs + sizeof(ChainHdr) * (size_t)h->nchains;
}

int layout_and_alloc(w2d_ctqmc *h, int nbound)
{
    w2d::ChainLayout L;
    int st = compute_layout(h, nbound, h->kcap, L);
    if (st) return st;
    if (h->chains) {
        HIPCHK(hipFree(h->chains));
        h->chains = nullptr;
    }
    h->L = L;
    const size_t total = chains_bytes(h, L);
    HIPCHK(hipMalloc(&h->chains, total));
    HIPCHK(hipMemset(h->chains, 0, total));
    h->nwin_alloc = nbound;
    return W2D_OK;
}

// Move every chain into a layout of capacity kc and nbound window boundaries,
// keeping each chain's state (configuration, random stream, inverses,
// accumulators; boundary caches when the window count stays).
int relayout(w2d_ctqmc *h, int kc, int nbound, int nm_copy = 0)
{
    w2d::ChainLayout B;
    int st = compute_layout(h, nbound, kc, B);
    if (st) return st;
    if (nm_copy <= 0) nm_copy = h->nm;
    char *nc = nullptr;
    HIPCHK(hipStreamSynchronize(h->stream));
    HIPCHK(hipMalloc(&nc, chains_bytes(h, B)));
    HIPCHK(hipMemsetAsync(nc, 0, chains_bytes(h, B), h->stream));
    const w2d::ChainLayout &A = h->L;
    hipLaunchKernelGGL(w2d::migrate_kernel, dim3(h->nchains), dim3(w2d::WAVE), 0, h->stream, A, B,
                       (const char *)h->chains, nc, 2 * h->model.nbands, nm_copy, (long)(A.stride - A.gtau));
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(h->stream));
    HIPCHK(hipFree(h->chains));
    h->chains = nc;
    h->L = B;
    h->kcap = kc;
    h->nwin_alloc = nbound;
    return W2D_OK;
}

// Double the pairs-per-flavour capacity of every chain (a chain halted for
// room continues where it stopped).
int regrow(w2d_ctqmc *h)
{
    const int kc = h->kcap * 2;
    if (!h->grow || kc > KCAP_MAX) return W2D_ERR_CAPACITY;
    const int st = relayout(h, kc, h->L.nbound);
    if (!st) h->regrows++;
    return st;
}

// allow the dynamic LDS of this handle (above 64 KB on devices that have it)
template <typename K>
void allow_lds(K kernel, const w2d_ctqmc *h)
{
    if (h->L.lds_bytes > 64 * 1024)
        (void)hipFuncSetAttribute(reinterpret_cast<const void *>(kernel),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)h->L.lds_bytes);
}

template <int NM>
void launch_init(w2d_ctqmc *h)
{
    allow_lds(w2d::init_kernel<NM>, h);
    hipLaunchKernelGGL(w2d::init_kernel<NM>, dim3(h->nchains), dim3(w2d::WAVE), (unsigned)h->L.lds_bytes, h->stream, h->dm, h->L,
                       h->chains);
}

// LDS attribute of the sweep kernels (set outside any stream capture)
template <int NM>
void prepare_sweep(w2d_ctqmc *h)
{
    allow_lds(w2d::sweep_kernel<NM, true>, h);
    allow_lds(w2d::sweep_kernel<NM, false>, h);
}

template <int NM>
void launch_sweep(w2d_ctqmc *h, long long target, long long ncorr, int only, w2d_move_record *rec,
                  long long rec_base)
{
    const int *order = nullptr;
    if (h->order_chains && only < 0) {
        // chains with more operators first (bucket = nop >> shift below 1024)
        int shift = 0;
        while ((h->L.cap >> shift) >= w2d::ORDER_BUCKETS) ++shift;
        hipLaunchKernelGGL(w2d::order_kernel, dim3(1), dim3(1024), 0, h->stream,
                           reinterpret_cast<const ChainHdr *>(h->chains + h->L.hdrs), h->nchains, shift, h->order);
        order = h->order;
    }
    if (h->dm.score && h->dm.ntheta > 0)
        hipLaunchKernelGGL((w2d::sweep_kernel<NM, true>), dim3(h->nchains), dim3(w2d::WAVE), (unsigned)h->L.lds_bytes,
                           h->stream, h->dm, h->L, h->chains, target, ncorr, only, rec, rec_base,
                           (const int *)h->abort_dev, h->calib, order);
    else
        hipLaunchKernelGGL((w2d::sweep_kernel<NM, false>), dim3(h->nchains), dim3(w2d::WAVE), (unsigned)h->L.lds_bytes,
                           h->stream, h->dm, h->L, h->chains, target, ncorr, only, rec, rec_base,
                           (const int *)h->abort_dev, h->calib, order);
    // G(i w) of the configurations this launch measured (launches end on a
    // measurement when MeasGiw is on, see launch_range)
    if (h->niw > 0 && ncorr > 0) {
        const int per_thread = (h->niw + w2d::GIW_R - 1) / w2d::GIW_R;
        const int nt = std::min(w2d::GIW_THREADS, (per_thread + 63) / 64 * 64);
        hipLaunchKernelGGL(w2d::giw_kernel, dim3(h->nchains, h->dm.nf), dim3(nt), 0, h->stream, h->L,
                           (const char *)h->chains, h->dm.nf, h->niw, h->beta, h->giw, h->giw_seen);
    }
}

template <int NM>
void launch_setconfig(w2d_ctqmc *h, int chain)
{
    allow_lds(w2d::setconfig_kernel<NM>, h);
    hipLaunchKernelGGL(w2d::setconfig_kernel<NM>, dim3(h->nchains), dim3(w2d::WAVE), (unsigned)h->L.lds_bytes, h->stream, h->dm,
                       h->L, h->chains, chain);
}

template <int NM>
void launch_dlogw(w2d_ctqmc *h, int chain, double *out)
{
    allow_lds(w2d::dlogw_kernel<NM>, h);
    hipLaunchKernelGGL(w2d::dlogw_kernel<NM>, dim3(1), dim3(w2d::WAVE), (unsigned)h->L.lds_bytes, h->stream, h->dm,
                       h->L, h->chains, chain, out);
}

#define DISPATCH(fn, ...)                    \
    switch (h->nm) {                         \
    case 1: fn<1>(__VA_ARGS__); break;       \
    case 4: fn<4>(__VA_ARGS__); break;       \
    case 16: fn<16>(__VA_ARGS__); break;     \
    default: fn<128>(__VA_ARGS__); break;    \
    }

// device address of chain c's header
char *hdr_of(w2d_ctqmc *h, int c) { return h->chains + h->L.hdrs + (long)c * (long)sizeof(ChainHdr); }

int read_headers(w2d_ctqmc *h, std::vector<ChainHdr> &hd)
{
    hd.resize(h->nchains);
    HIPCHK(hipStreamSynchronize(h->stream));
    HIPCHK(hipMemcpy(hd.data(), h->chains + h->L.hdrs, sizeof(ChainHdr) * (size_t)h->nchains,
                     hipMemcpyDeviceToHost));
    return W2D_OK;
}

int write_headers(w2d_ctqmc *h, const std::vector<ChainHdr> &hd)
{
    HIPCHK(hipMemcpy(h->chains + h->L.hdrs, hd.data(), sizeof(ChainHdr) * (size_t)h->nchains,
                     hipMemcpyHostToDevice));
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
    std::vector<ChainHdr> hds(h->nchains);
    std::vector<uint32_t> mt(624L * h->nchains);
    for (int c = 0; c < h->nchains; ++c) {
        ChainHdr &hd = hds[c];
        std::memset(&hd, 0, sizeof(hd));
        hd.det_sign = 1.0;
        hd.trace_sign = 1.0;
        hd.cfgsign = -1.0;
        hd.mti = 624;
        // AccPair over NGtau bins of [0, beta] (CTQMC.F90:775-777)
        hd.naccpair = h->ntau;
        hd.accpair_max = h->beta;
        hd.apt_index = 1;
        hd.apt_n = 0;
        init_window(hd, h->taudiff_param, h->beta);
        seed_mt(mt.data() + 624L * c, h->seed0 + (uint32_t)c);
    }
    int st0 = write_headers(h, hds);
    if (st0) return st0;
    HIPCHK(hipMemcpy2D(h->chains + h->L.mt, (size_t)h->L.stride, mt.data(), 624 * 4, 624 * 4,
                       (size_t)h->nchains, hipMemcpyHostToDevice));
    DISPATCH(launch_init, h);
    HIPCHK(hipGetLastError());
    std::vector<ChainHdr> hd;
    int st = read_headers(h, hd);
    if (st) return st;
    return first_error(hd) ? W2D_ERR_ZERO_TRACE : W2D_OK;
}

// init_counters (CTQMC.F90:847) and the start of a warmup / measure call
void reset_counters(ChainHdr &x)
{
    x.try_add = x.try_rem = x.try_glob = x.try_shift = 0;
    x.acc_add = x.acc_rem = x.acc_glob = x.acc_shift = 0;
    x.qn_rej = x.moves = 0;
    x.alg_flops = x.alg_bytes = 0.0;
    x.pos = 0;
    x.halt = w2d::HALT_NONE;
    x.apt_index = 1;
}

// AccPair = 0 on every chain (init_counters, CTQMC.F90:851)
int clear_accpair(w2d_ctqmc *h)
{
    const w2d::ChainLayout &L = h->L;
    HIPCHK(hipMemset2DAsync(h->chains + L.accpair, (size_t)L.stride, 0,
                            (size_t)2 * h->model.nbands * L.naccpair_cap * 8, (size_t)h->nchains, h->stream));
    return W2D_OK;
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

// Bring every chain to `target` moves of the current call in launches of at
// most `per` moves (whole measurement intervals), starting from position
// `from`.  Asynchronous: the caller synchronises.
void launch_range(w2d_ctqmc *h, long long from, long long target, long long ncorr, long long per, int only,
                  w2d_move_record *rec, long long rec_base)
{
    // with MeasGiw every launch ends on a measurement: targets on multiples of
    // ncorr, also when a resumed pass starts between two of them
    const bool giw = h->niw > 0 && ncorr > 0;
    for (long long t = from; t < target;) {
        t = std::min(target, giw ? (t / ncorr + 1) * ncorr : t + per);
        DISPATCH(launch_sweep, h, t, ncorr, only, rec, rec_base);
        ++h->last_launches;
    }
}

// After a pass: chains halted for room get a larger capacity and continue;
// returns whether another pass is needed.  Chains that cannot grow further
// are stopped with W2D_ERR_CAPACITY (their results are left out).
int resume_halted(w2d_ctqmc *h, std::vector<ChainHdr> &hd, long long &from, bool &again)
{
    again = false;
    bool cap = false;
    from = -1;
    for (const auto &x : hd) {
        if (x.halt == w2d::HALT_ABORT) h->aborted = 1;
        if (x.halt == w2d::HALT_CAP && x.err == 0) {
            cap = true;
            from = from < 0 ? x.pos : std::min(from, x.pos);
        }
    }
    if (!cap || h->aborted) return W2D_OK;
    const int st = regrow(h);
    if (st && st != W2D_ERR_CAPACITY) return st;
    for (auto &x : hd)
        if (x.halt == w2d::HALT_CAP) {
            x.halt = w2d::HALT_NONE;
            if (st == W2D_ERR_CAPACITY) x.err = W2D_ERR_CAPACITY;
        }
    int wst = write_headers(h, hd);
    if (wst) return wst;
    again = (st == W2D_OK);
    return W2D_OK;
}

// W2D_OK if every chain is fine, W2D_ERR_PARTIAL if some chains stopped on an
// error and others did not, else the (first) error of the chains
int batch_status(const std::vector<ChainHdr> &hd)
{
    int bad = 0, first = 0;
    for (const auto &x : hd)
        if (x.err) {
            ++bad;
            if (!first) first = x.err;
        }
    if (!bad) return W2D_OK;
    return bad < (int)hd.size() ? W2D_ERR_PARTIAL : map_err(first);
}

// run nmoves steps per chain in launches of whole measurement intervals
int run_moves(w2d_ctqmc *h, long long nmoves, long long ncorr)
{
    h->moved = true;
    h->aborted = 0;
    const long long base = h->per_launch > 0 ? h->per_launch : 4000;
    long long per = ncorr > 0 ? ncorr * std::max<long long>(1, base / ncorr) : base;
    if (h->niw > 0 && ncorr > 0) per = ncorr;   // one measurement per launch (MeasGiw)
    h->last_launches = 0;
    h->last_ms = 0.0;
    std::vector<ChainHdr> hd;
    long long from = 0;
    for (bool again = true; again;) {
        DISPATCH(prepare_sweep, h);
        const bool graph = h->graphs && (nmoves - from + per - 1) / per >= 2;
        hipGraph_t g = nullptr;
        hipGraphExec_t ge = nullptr;
        if (graph) {
            // the pass's launches (order + sweep per chunk) as one graph
            HIPCHK(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
            launch_range(h, from, nmoves, ncorr, per, -1, nullptr, 0);
            hipError_t e = hipStreamEndCapture(h->stream, &g);
            if (e == hipSuccess) e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
            if (e != hipSuccess) {
                if (g) (void)hipGraphDestroy(g);
                HIPCHK(e);
            }
        }
        HIPCHK(hipEventRecord(h->ev0, h->stream));
        if (graph) {
            hipError_t e = hipGraphLaunch(ge, h->stream);
            if (e != hipSuccess) {
                (void)hipGraphExecDestroy(ge);
                (void)hipGraphDestroy(g);
                HIPCHK(e);
            }
        } else {
            launch_range(h, from, nmoves, ncorr, per, -1, nullptr, 0);
        }
        HIPCHK(hipGetLastError());
        HIPCHK(hipEventRecord(h->ev1, h->stream));
        hipError_t se = hipEventSynchronize(h->ev1);
        if (graph) {
            (void)hipGraphExecDestroy(ge);
            (void)hipGraphDestroy(g);
        }
        HIPCHK(se);
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, h->ev0, h->ev1));
        h->last_ms += ms;
        int st = read_headers(h, hd);
        if (st) return st;
        if ((st = resume_halted(h, hd, from, again))) return st;
    }
    if (h->aborted) {
        long long done = 0;
        for (const auto &x : hd) done += x.cnt_meas;
        if (ncorr > 0 && done == 0) h->aborted = 2;
    }
    return batch_status(hd);
}

// Device work area of the group sums: [ngroups][nel] partial sums, then
// [ngroups][14] scalars, then [ngroups] chain counts.
size_t group_sums_bytes(const w2d_ctqmc *h, int ngroups)
{
    const long nel = (h->L.stride - h->L.gtau) / 8;
    return (size_t)ngroups * ((size_t)nel * 8 + 14 * 8 + 8);
}

// Group sums into a caller-provided device work area; the results go straight
// to the host arrays (strided copies), so nothing is allocated here.
int group_sums_into(w2d_ctqmc *h, int ngroups, void *work, double *gtau, double *gleg, double *dm_blocks,
                    double *histo, double *outer, double *scalars, int32_t *nused)
{
    const w2d::ChainLayout &L = h->L;
    const int nf = 2 * h->model.nbands, nc = h->nchains;
    const long nel = (L.stride - L.gtau) / 8;
    double *part = static_cast<double *>(work);
    double *scal = part + (long)ngroups * nel;
    int *used = reinterpret_cast<int *>(scal + (long)ngroups * 14);
    const ChainHdr *hdrs = reinterpret_cast<const ChainHdr *>(h->chains + L.hdrs);
    hipLaunchKernelGGL(w2d::group_sum_kernel, dim3((unsigned)((nel + 255) / 256), (unsigned)ngroups), dim3(256), 0,
                       h->stream, (const char *)h->chains, L.stride, L.gtau, nel, nc, ngroups, hdrs, part);
    hipLaunchKernelGGL(w2d::group_scalars_kernel, dim3((unsigned)((ngroups + 63) / 64)), dim3(64), 0, h->stream, hdrs,
                       nc, ngroups, scal, used);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(h->stream));
    auto put = [&](double *dst, long off, long n) -> int {
        if (!dst || n <= 0) return W2D_OK;
        HIPCHK(hipMemcpy2D(dst, (size_t)n * 8, part + (off - L.gtau) / 8, (size_t)nel * 8, (size_t)n * 8,
                           (size_t)ngroups, hipMemcpyDeviceToHost));
        return W2D_OK;
    };
    int st;
    if ((st = put(gtau, L.gtau, (long)nf * h->ntau)) || (st = put(gleg, L.gleg, (long)nf * h->nleg)) ||
        (st = put(dm_blocks, L.dm, h->dm.dmsize)) || (st = put(histo, L.histo, (long)nf * (h->maxhisto + 1))) ||
        (st = put(outer, L.outer, h->model.nsst)))
        return st;
    if (scalars) HIPCHK(hipMemcpy(scalars, scal, (size_t)ngroups * 14 * 8, hipMemcpyDeviceToHost));
    if (nused) HIPCHK(hipMemcpy(nused, used, (size_t)ngroups * 4, hipMemcpyDeviceToHost));
    return W2D_OK;
}

// Zero-padded copies of the operator blocks for the MFMA walks (engine for
// blocks above 16 states): rows and columns padded to multiples of 16,
// column-major with leading dimension 16 ceil(n_t / 16); offsets indexed like
// the transition records.
int build_ops16(w2d_ctqmc *h)
{
    if (h->dm.ops16) return W2D_OK;
    const w2d::LocalModel &m = h->model;
    const int nb = m.nbands;
    std::vector<double> ops16;
    std::vector<int> opoff16(m.opoff.size(), -1);
    for (size_t i = 0; i < m.connect.size(); ++i) {
        const int ts = m.connect[i];
        if (ts < 0 || m.opoff[i] < 0) continue;
        const int src = (int)(i / (4 * (size_t)nb));   // superstate of this record
        const int nt = m.sizes[ts], ns = m.sizes[src];
        const int ntp = (nt + 15) / 16 * 16, nsp = (ns + 15) / 16 * 16;
        opoff16[i] = (int)ops16.size();
        ops16.resize(ops16.size() + (size_t)ntp * nsp, 0.0);
        double *dst = ops16.data() + opoff16[i];
        const double *srcb = m.ops.data() + m.opoff[i];
        for (int k = 0; k < ns; ++k)
            for (int r = 0; r < nt; ++r) dst[r + (size_t)k * ntp] = srcb[r + (size_t)k * nt];
    }
    int st;
    const double *o = nullptr;
    const int *f = nullptr;
    if ((st = upload(h, ops16, &o)) || (st = upload(h, opoff16, &f))) return st;
    h->dm.ops16 = o;
    h->dm.opoff16 = f;
    return W2D_OK;
}

// moves of the next pass count from zero on every chain
int rewind(w2d_ctqmc *h)
{
    std::vector<ChainHdr> hd;
    int st = read_headers(h, hd);
    if (st) return st;
    for (auto &x : hd) {
        x.pos = 0;
        if (x.halt != w2d::HALT_CAP) x.halt = w2d::HALT_NONE;
    }
    return write_headers(h, hd);
}

// Change the trace engine (padded block size 1, 4, 16 or 128, at least what
// the model's blocks need) of every chain: their memory is re-laid out and the
// trace caches are rebuilt by the next moves.
int switch_engine(w2d_ctqmc *h, int nm)
{
    if (nm != 1 && nm != 4 && nm != 16 && nm != 128) return W2D_ERR_ARG;
    if (nm < h->nm_min) return W2D_ERR_ARG;
    if (nm > 16 && h->dm.ntheta > 0) return W2D_ERR_ARG;   // derivative pass: engines up to 16 only
    // the MFMA engine only for models that need it (blocks above 16 states):
    // measurements with it on models of smaller blocks have ended in GPU
    // memory faults whose cause is not found, so that combination is refused
    if (nm > 16 && h->nm_min <= 16) return W2D_ERR_ARG;
    if (nm == h->nm) return W2D_OK;
    int st;
    if (nm > 16 && (st = build_ops16(h))) return st;
    const int old = h->nm;
    h->nm = nm;
    if ((st = relayout(h, h->kcap, h->L.nbound, std::min(old, nm)))) {
        h->nm = old;
        return st;
    }
    std::vector<ChainHdr> hd;
    if ((st = read_headers(h, hd))) return st;
    for (auto &x : hd) {
        x.cache_pos = -1;
        x.wvalid = 0;
    }
    return write_headers(h, hd);
}

}  // namespace

// every resource the handle holds (also on the error paths of create)
w2d_ctqmc::~w2d_ctqmc()
{
    if (stream) (void)hipStreamSynchronize(stream);
    for (void *p : dev_allocs) (void)hipFree(p);
    if (chains) (void)hipFree(chains);
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (stream) (void)hipStreamDestroy(stream);
    if (abort_host) {
        unregister_abort_word(abort_host);
        (void)hipHostFree(abort_host);
    }
}

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
        P.integer("Eigenbasis") == 0)
        return W2D_ERR_ARG;
    if (P.integer("Nd") != nbands) return W2D_ERR_ARG;
    // measurements this library does not produce: refused rather than left
    // unmeasured (the caller would read arrays that were never filled)
    {
        const char *off_keys[] = {"MeasG2iw", "MeasGSigmaiw", "FourPnt", "MeasSusz", "MeasSsquare",
                                  "MeasExpResDensityMatrix", "MeasGtauDetRat", "WormMeasGiw", "WormMeasGtau"};
        for (const char *k : off_keys)
            if (P.has(k) && P.integer(k) != 0) {
                std::fprintf(stderr, "w2d_ctqmc: %s = %lld is not supported (only 0)\n", k, P.integer(k));
                return W2D_ERR_ARG;
            }
        if (P.has("MeasurementTiming") && P.integer("MeasurementTiming") >= 0) {
            std::fprintf(stderr, "w2d_ctqmc: MeasurementTiming >= 0 is not supported\n");
            return W2D_ERR_ARG;
        }
    }
    auto h = std::make_unique<w2d_ctqmc>();
    h->params = params;
    h->nchains = nchains;
    h->seed0 = seed0;
    h->device = device;
    h->kcap = capacity != 0 ? std::abs(capacity) : 256;
    h->grow = capacity >= 0;
    if (h->kcap < 2 || h->kcap > KCAP_MAX) return W2D_ERR_ARG;
    h->beta = P.real("beta");
    h->ntau = (int)P.integer("Ntau");
    h->nleg = (int)P.integer("NLegMax");
    h->maxhisto = (int)P.integer("MaxHisto");
    h->ncorr = P.integer("NCorr");
    h->taudiff_param = P.real("TaudiffMax");
    if (h->beta <= 0 || h->ntau < 1 || h->nleg < 1 || h->maxhisto < 0 || h->ncorr < 1) return W2D_ERR_ARG;
    const double epsblock = P.has("EPSBLOCK") ? P.real("EPSBLOCK") : 1e-13;
    std::string err = w2d::build_local_model(nbands, u_matrix, muimp_full, P.str("QuantumNumbers"), h->model,
                                             epsblock);
    if (!err.empty()) {
        std::fprintf(stderr, "w2d_ctqmc: %s\n", err.c_str());
        return W2D_ERR_ARG;
    }
    const w2d::LocalModel &m = h->model;
    if (m.nsst > w2d::MAXSST) {
        std::fprintf(stderr, "w2d_ctqmc: %d quantum-number blocks, at most %d are supported\n", m.nsst, w2d::MAXSST);
        return W2D_ERR_ARG;
    }
    if (m.nmax <= 1) h->nm = 1;
    else if (m.nmax <= 4) h->nm = 4;
    else if (m.nmax <= 16) h->nm = 16;
    else if (m.nmax <= 128) h->nm = 128;
    else return W2D_ERR_ARG;
    h->nm_min = h->nm;
    HIPCHK(hipSetDevice(device));
    HIPCHK(hipDeviceGetAttribute(&h->lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, device));
    HIPCHK(hipStreamCreate(&h->stream));
    HIPCHK(hipEventCreate(&h->ev0));
    HIPCHK(hipEventCreate(&h->ev1));
    {
        void *o = nullptr;
        HIPCHK(hipMalloc(&o, sizeof(int) * (size_t)nchains));
        h->dev_allocs.push_back(o);
        h->order = static_cast<int *>(o);
    }
    HIPCHK(hipHostMalloc(reinterpret_cast<void **>(&h->abort_host), sizeof(int),
                         hipHostMallocMapped | hipHostMallocCoherent));
    *h->abort_host = 0;
    HIPCHK(hipHostGetDevicePointer(reinterpret_cast<void **>(&h->abort_dev), h->abort_host, 0));
    register_abort_word(h->abort_host);
    {
        const int sst = scratch_for_device(device, &h->scratch);
        if (sst) return sst;
    }
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
    d.nops = (long)m.ops.size();
    // Bonnet recursion P_{l+1} = a_l x P_l - b_l P_{l-1}: a_l = (2l+1)/(l+1), b_l = l/(l+1)
    std::vector<double> legc(2 * (size_t)std::max(h->nleg, 1));
    for (int l = 0; l < h->nleg; ++l) {
        legc[2 * l] = (2.0 * l + 1.0) / (double)(l + 1);
        legc[2 * l + 1] = (double)l / (double)(l + 1);
    }
    std::vector<int> first_state(m.nsst);
    for (int s = 0; s < m.nsst; ++s) first_state[s] = m.states[m.offsets[s]];
    std::vector<int> opoff32(m.opoff.begin(), m.opoff.end());
    std::vector<int4> trans(m.connect.size());
    for (size_t i = 0; i < m.connect.size(); ++i) {
        const int ts = m.connect[i];
        trans[i] = make_int4(ts, ts >= 0 ? m.sizes[ts] : 0, (int)m.opoff[i], ts >= 0 ? m.offsets[ts] : 0);
    }
    int st = 0;
    if ((st = upload(h.get(), m.sizes, &d.sizes)) || (st = upload(h.get(), m.offsets, &d.offsets)) ||
        (st = upload(h.get(), m.connect, &d.connect)) || (st = upload(h.get(), first_state, &d.first_state)) ||
        (st = upload(h.get(), m.state_sst, &d.state_sst)) || (st = upload(h.get(), m.opoff, &d.opoff)) ||
        (st = upload(h.get(), m.ops, &d.ops)) || (st = upload(h.get(), m.evals, &d.evals)) ||
        (st = upload(h.get(), fdiag, &d.ftau)) || (st = upload(h.get(), dmoff, &d.dmoff)) ||
        (st = upload(h.get(), legc, &d.legc)) || (st = upload(h.get(), trans, &d.trans)) ||
        (st = upload(h.get(), opoff32, &d.opoff32)) || (h->nm > 16 && (st = build_ops16(h.get()))))
        return st;
    // user symmetry moves (Trace.F90:471-478): SymMove01.. are 1-based
    // flavour permutations; their inverses follow in the table
    {
        const long long nsym = P.has("NSymMove") ? P.integer("NSymMove") : 0;
        std::vector<int> tab;
        if (nsym > 0) {
            tab.assign((size_t)2 * nsym * nf, 0);
            for (long long i = 0; i < nsym; ++i) {
                char key[16];
                std::snprintf(key, sizeof(key), "SymMove%02lld", i + 1);
                std::istringstream in(P.str(key));
                std::vector<int> perm;
                for (int v; in >> v;) perm.push_back(v);
                std::vector<int> seen(nf + 1, 0);
                bool ok = (int)perm.size() == nf;
                for (int v : perm) ok = ok && v >= 1 && v <= nf && !seen[v]++;
                if (!ok) {
                    std::fprintf(stderr, "w2d_ctqmc: %s is not a permutation of the %d flavours\n", key, nf);
                    return W2D_ERR_ARG;
                }
                for (int j = 0; j < nf; ++j) {
                    tab[(size_t)i * nf + j] = perm[j];
                    tab[(size_t)(nsym + i) * nf + perm[j] - 1] = j + 1;
                }
            }
            if ((st = upload(h.get(), tab, &d.symmoves))) return st;
        }
        d.nsym = (int)std::max<long long>(0, nsym);
    }
    // the MFMA engine (blocks above 16 states) keeps the global-move code of
    // its validated builds: no user symmetry moves
    if (h->nm > 16 && d.nsym > 0) {
        std::fprintf(stderr, "w2d_ctqmc: NSymMove is not supported with blocks above 16 states\n");
        return W2D_ERR_ARG;
    }
    d.p_shift = P.real("PercentageTauShiftMove");
    d.p_glob = P.real("PercentageGlobalMove");
    d.p_outer = 0.0;
    d.p_4op = 0.0;
    d.ntau = h->ntau;
    d.nleg = h->nleg;
    d.maxhisto = h->maxhisto;
    d.has_szt = P.str("QuantumNumbers").find("Szt") != std::string::npos;
    d.meas_dm = P.integer("MeasDensityMatrix") != 0;
    // MeasGiw with Niw frequencies (configspec default 2000)
    if (P.has("MeasGiw") && P.integer("MeasGiw") != 0) {
        const long long niw = P.has("Niw") ? P.integer("Niw") : 2000;
        if (niw < 1 || niw > GIW_NIW_MAX) {
            std::fprintf(stderr, "w2d_ctqmc: Niw = %lld outside [1, %d]\n", niw, GIW_NIW_MAX);
            return W2D_ERR_ARG;
        }
        h->niw = (int)niw;
        const size_t nsum = (size_t)nchains * nf * (size_t)niw * 2, nseen = (size_t)nchains * nf * 2;
        void *a = nullptr, *b = nullptr;
        HIPCHK(hipMalloc(&a, nsum * sizeof(double)));
        h->dev_allocs.push_back(a);
        HIPCHK(hipMalloc(&b, nseen * sizeof(long long)));
        h->dev_allocs.push_back(b);
        h->giw = static_cast<double *>(a);
        h->giw_seen = static_cast<long long *>(b);
        HIPCHK(hipMemset(a, 0, nsum * sizeof(double)));
        HIPCHK(hipMemset(b, 0, nseen * sizeof(long long)));
    }
    // AccumGtau: 1 meanacc, 2 autocorracc with ilog2(Nmeas) levels (CTQMC.F90:491-503)
    d.accum = P.has("AccumGtau") ? (int)P.integer("AccumGtau") : 1;
    if (d.accum != 1 && d.accum != 2) return W2D_ERR_ARG;
    // step-resolved G(tau) and sign over the Nmeas measurements (CTQMC.F90:510-512)
    d.step_mean = P.has("Gtau_mean_step") && P.integer("Gtau_mean_step") != 0;
    d.step_mid = P.has("Gtau_mid_step") && P.integer("Gtau_mid_step") != 0;
    d.step_sign = P.has("sign_step") && P.integer("sign_step") != 0;
    d.nmstep = (d.step_mean || d.step_mid || d.step_sign) ? (int)std::max<long long>(P.integer("Nmeas"), 1) : 0;
    {
        long long nm = P.integer("Nmeas");
        int lv = -1;
        while (nm > 0) ++lv, nm /= 2;
        d.nlev = std::max(lv, 1);
    }
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
    if (nb != h->nwin_alloc) {
        // the window count sizes the boundary caches and the LDS operator
        // lists: before the chains have moved, re-create their memory (their
        // initial state follows from the seeds); after (the calibration of
        // cthyb_solver.py:899-944), move them into the new layout
        int st = h->moved ? relayout(h, h->kcap, nb) : layout_and_alloc(h, nb);
        if (st) return st;
        if (!h->moved && (st = init_chains(h))) return st;
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
    // ctqmc_warmup: AccPair over NGtau bins of [0, beta], then init_counters
    // (CTQMC.F90:8199-8206)
    for (auto &x : hd) {
        reset_counters(x);
        x.naccpair = h->ntau;
        x.accpair_max = h->beta;
    }
    if ((st = write_headers(h, hd)) || (st = clear_accpair(h))) return st;
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
    HIPCHK(hipMemset2DAsync(h->chains + L.gtau, (size_t)L.stride, 0, (size_t)(L.stride - L.gtau),
                            (size_t)h->nchains, h->stream));
    if (h->niw > 0) {
        const size_t nf = (size_t)h->dm.nf, nc = (size_t)h->nchains;
        HIPCHK(hipMemsetAsync(h->giw, 0, nc * nf * (size_t)h->niw * 2 * sizeof(double), h->stream));
        HIPCHK(hipMemsetAsync(h->giw_seen, 0, nc * nf * 2 * sizeof(long long), h->stream));
    }
    st = run_moves(h, nmeas * h->ncorr, h->ncorr);
    h->measured = (st == W2D_OK || st == W2D_ERR_PARTIAL);
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
    // strided gathers: one copy per array over all chains
    auto gather = [&](double *dst, long off, size_t bytes) -> int {
        HIPCHK(hipMemcpy2D(dst, bytes, h->chains + off, (size_t)L.stride, bytes, (size_t)h->nchains,
                           hipMemcpyDeviceToHost));
        return W2D_OK;
    };
    if (gtau && (st = gather(gtau, L.gtau, (size_t)nf * h->ntau * 8))) return st;
    if (gleg && (st = gather(gleg, L.gleg, (size_t)nf * h->nleg * 8))) return st;
    if (histo && (st = gather(histo, L.histo, (size_t)nf * (h->maxhisto + 1) * 8))) return st;
    if (outer && (st = gather(outer, L.outer, (size_t)m.nsst * 8))) return st;
    std::vector<double> dmall;
    if (dm) {
        dmall.resize((size_t)h->dm.dmsize * h->nchains);
        if ((st = gather(dmall.data(), L.dm, (size_t)h->dm.dmsize * 8))) return st;
    }
    for (int c = 0; c < h->nchains; ++c) {
        if (dm) {
            std::copy(dmall.begin() + (long)c * h->dm.dmsize, dmall.begin() + (long)(c + 1) * h->dm.dmsize,
                      blk.begin());
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
    if (h) h->moved = true;
    if (!h || chain < 0 || chain >= h->nchains || n < 0 || (!rec && n > 0)) return W2D_ERR_ARG;
    const long long per = 4000;
    std::vector<ChainHdr> hd;
    int st = read_headers(h, hd);
    if (st) return st;
    hd[chain].pos = 0;
    hd[chain].halt = w2d::HALT_NONE;
    if ((st = write_headers(h, hd))) return st;
    w2d_move_record *d = nullptr;
    HIPCHK(hipMalloc(&d, sizeof(w2d_move_record) * per));
    DISPATCH(prepare_sweep, h);
    for (long long done = 0; done < n && st == W2D_OK;) {
        const long long k = std::min<long long>(per, n - done);
        DISPATCH(launch_sweep, h, done + k, 0, chain, d, done);
        if (hipGetLastError() != hipSuccess || hipStreamSynchronize(h->stream) != hipSuccess) {
            st = W2D_ERR_HIP;
            break;
        }
        if ((st = read_headers(h, hd))) break;
        if (hd[chain].halt == w2d::HALT_CAP && hd[chain].err == 0) {
            // grow and finish this chunk (records continue at pos - done)
            long long from;
            bool again;
            if ((st = resume_halted(h, hd, from, again))) break;
            if (again) continue;
        }
        if (hipMemcpy(rec + done, d, sizeof(w2d_move_record) * (hd[chain].pos - done), hipMemcpyDeviceToHost) !=
            hipSuccess) {
            st = W2D_ERR_HIP;
            break;
        }
        if (hd[chain].err || hd[chain].halt) break;
        done += k;
    }
    (void)hipFree(d);
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
    HIPCHK(hipMemcpy(&x, hdr_of(h, chain), sizeof(x), hipMemcpyDeviceToHost));
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
    h->moved = true;
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
        if (cnt[2 * f] != cnt[2 * f + 1]) return W2D_ERR_ARG;
    // grow until every flavour's order fits
    for (int f = 0; f < 2 * h->model.nbands; ++f)
        while (cnt[2 * f] > h->kcap) {
            const int st = regrow(h);
            if (st) return st;
        }
    char *base = h->chains + (long)chain * h->L.stride;
    HIPCHK(hipStreamSynchronize(h->stream));
    ChainHdr x;
    HIPCHK(hipMemcpy(&x, hdr_of(h, chain), sizeof(x), hipMemcpyDeviceToHost));
    x.nop = n;
    x.outer = outer_sst;
    x.err = 0;
    HIPCHK(hipMemcpy(hdr_of(h, chain), &x, sizeof(x), hipMemcpyHostToDevice));
    if (n > 0) {
        HIPCHK(hipMemcpy(base + h->L.tau, tau, (size_t)n * 8, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(base + h->L.code, code.data(), (size_t)n * 4, hipMemcpyHostToDevice));
    }
    DISPATCH(launch_setconfig, h, chain);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(h->stream));
    HIPCHK(hipMemcpy(&x, hdr_of(h, chain), sizeof(x), hipMemcpyDeviceToHost));
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
    HIPCHK(hipMemcpy(&x, hdr_of(h, chain), sizeof(x), hipMemcpyDeviceToHost));
    *k = x.kf[f];
    if (!out) return W2D_OK;
    // gather the physical block into logical (rank) order
    const int kk = x.kf[f], ld = h->L.kcap;
    std::vector<double> phys((size_t)ld * ld);
    std::vector<int> rs(ld), cs(ld);
    HIPCHK(hipMemcpy(phys.data(), base + h->L.minv + (long)f * ld * ld * 8, phys.size() * 8, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(rs.data(), base + h->L.fcs + (long)f * ld * 4, (size_t)ld * 4, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(cs.data(), base + h->L.fas + (long)f * ld * 4, (size_t)ld * 4, hipMemcpyDeviceToHost));
    for (int j = 0; j < kk; ++j)
        for (int i = 0; i < kk; ++i) out[i + (long)j * kk] = phys[rs[i] + (long)cs[j] * ld];
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
    delete h;
    return W2D_OK;
}

int w2d_ctqmc_set_derivatives(w2d_ctqmc *h, int ntheta, const double *du, const double *dmuimp,
                              const double *dftau, int score)
{
    if (!h || ntheta < 1 || ntheta > w2d::MAXTHETA || !du || !dmuimp || !dftau) return W2D_ERR_ARG;
    // the derivative pass runs with the engines for blocks up to 16 states:
    // with the MFMA engine (128) a measurement run faulted the GPU, cause not
    // found yet, so that combination is refused
    if (h->nm > 16) {
        std::fprintf(stderr, "w2d_ctqmc_set_derivatives: not available with trace engine %d (blocks above 16 "
                             "states)\n", h->nm);
        return W2D_ERR_ARG;
    }
    const int nb = h->model.nbands, n2 = 2 * nb, nf = 2 * nb, nftau = h->dm.nftau;
    const long usz = (long)n2 * n2 * n2 * n2, msz = (long)n2 * n2;
    std::vector<double> blk, all;
    for (int t = 0; t < ntheta; ++t) {
        const std::string err = w2d::derivative_blocks(h->model, du + t * usz, dmuimp + t * msz, blk);
        if (!err.empty()) {
            std::fprintf(stderr, "w2d_ctqmc_set_derivatives: parameter %d: %s\n", t, err.c_str());
            return W2D_ERR_ARG;
        }
        all.insert(all.end(), blk.begin(), blk.end());
    }
    // dF/dtheta(b, s, b, s, :) of each flavour f = b + N s, as the F(tau) tables
    std::vector<double> fd((size_t)ntheta * nf * nftau);
    for (int t = 0; t < ntheta; ++t)
        for (int b = 0; b < nb; ++b)
            for (int sp = 0; sp < 2; ++sp)
                for (int k = 0; k < nftau; ++k) {
                    const long idx = b + (long)nb * (sp + 2L * (b + (long)nb * (sp + 2L * k)));
                    fd[((size_t)t * nf + b + nb * sp) * nftau + k] = dftau[(long)t * msz * nftau + idx];
                }
    int st;
    const double *dblk = nullptr, *dft = nullptr;
    if ((st = upload(h, all, &dblk)) || (st = upload(h, fd, &dft))) return st;
    h->dm.dblk = dblk;
    h->dm.dftau = dft;
    h->dm.ntheta = ntheta;
    h->dm.score = score != 0;
    // room for the derivative work area and the score sums (cleared)
    if ((st = relayout(h, h->kcap, h->L.nbound))) return st;
    const w2d::ChainLayout &L = h->L;
    HIPCHK(hipMemset2DAsync(h->chains + L.scsum, (size_t)L.stride, 0, (size_t)(L.stride - L.scsum),
                            (size_t)h->nchains, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return W2D_OK;
}

int w2d_ctqmc_dlogw(w2d_ctqmc *h, int chain, double *out)
{
    if (!h || !out || chain < 0 || chain >= h->nchains) return W2D_ERR_ARG;
    if (h->dm.ntheta < 1) return W2D_ERR_STATE;
    double *d = nullptr;
    HIPCHK(hipMalloc(&d, 2 * w2d::MAXTHETA * sizeof(double)));
    DISPATCH(launch_dlogw, h, chain, d);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e == hipSuccess) e = hipMemcpy(out, d, 2 * h->dm.ntheta * sizeof(double), hipMemcpyDeviceToHost);
    (void)hipFree(d);
    HIPCHK(e);
    return W2D_OK;
}

int w2d_ctqmc_get_score_sums(w2d_ctqmc *h, int ngroups, double *ssum, double *gtau, double *dm_blocks,
                             double *dm_explicit)
{
    if (!h || ngroups < 1 || ngroups > h->nchains) return W2D_ERR_ARG;
    if (!h->measured || !h->dm.score || h->dm.ntheta < 1) return W2D_ERR_STATE;
    void *work = nullptr;
    HIPCHK(hipMalloc(&work, group_sums_bytes(h, ngroups)));
    int st = group_sums_into(h, ngroups, work, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr);
    const w2d::ChainLayout &L = h->L;
    const long nel = (L.stride - L.gtau) / 8;
    const double *part = static_cast<const double *>(work);
    auto put = [&](double *dst, long off, long n) -> int {
        if (!dst || n <= 0) return W2D_OK;
        HIPCHK(hipMemcpy2D(dst, (size_t)n * 8, part + (off - L.gtau) / 8, (size_t)nel * 8, (size_t)n * 8,
                           (size_t)ngroups, hipMemcpyDeviceToHost));
        return W2D_OK;
    };
    const long nt = h->dm.ntheta, nf = 2 * h->model.nbands;
    if (!st) st = put(ssum, L.scsum, nt);
    if (!st) st = put(gtau, L.scg, nt * nf * h->ntau);
    if (!st) st = put(dm_blocks, L.scdm, nt * h->dm.dmsize);
    if (!st) st = put(dm_explicit, L.scdmx, nt * h->dm.dmsize);
    (void)hipFree(work);
    return st;
}

int w2d_ctqmc_set_launch(w2d_ctqmc *h, int graphs, int order_chains, int64_t moves_per_launch)
{
    if (!h || moves_per_launch < 0) return W2D_ERR_ARG;
    h->graphs = graphs != 0;
    h->order_chains = order_chains != 0;
    h->per_launch = moves_per_launch;
    return W2D_OK;
}

int w2d_ctqmc_set_engine(w2d_ctqmc *h, int nm)
{
    if (!h) return W2D_ERR_ARG;
    return switch_engine(h, nm > 0 ? nm : h->nm_min);
}

int w2d_ctqmc_get_engine(const w2d_ctqmc *h, int *nm, int *nm_min)
{
    if (!h) return W2D_ERR_ARG;
    if (nm) *nm = h->nm;
    if (nm_min) *nm_min = h->nm_min;
    return W2D_OK;
}

int w2d_ctqmc_autotune(w2d_ctqmc *h, int64_t moves, int *nm, double *rates)
{
    if (!h || moves < 1) return W2D_ERR_ARG;
    const int cand[4] = {1, 4, 16, 128};
    int best = h->nm;
    double best_rate = -1.0;
    for (int i = 0; i < 4; ++i) {
        if (rates) rates[i] = 0.0;
        if (cand[i] < h->nm_min || (cand[i] > 16 && (h->dm.ntheta > 0 || h->nm_min <= 16))) continue;
        int st = switch_engine(h, cand[i]);
        if (st == W2D_ERR_CAPACITY) continue;   // this engine's LDS does not fit
        if (st) return st;
        // rebuild the caches untimed, then time the engine (positions count
        // from zero in each pass)
        if ((st = rewind(h)) || ((st = run_moves(h, std::max<int64_t>(1, moves / 4), 0)) && st != W2D_ERR_PARTIAL))
            return st;
        if ((st = rewind(h)) || ((st = run_moves(h, moves, 0)) && st != W2D_ERR_PARTIAL)) return st;
        const double rate = (double)h->nchains * (double)moves / (h->last_ms * 1e-3);
        if (rates) rates[i] = rate;
        if (rate > best_rate) {
            best_rate = rate;
            best = cand[i];
        }
    }
    const int st = switch_engine(h, best);
    if (nm) *nm = h->nm;
    return st;
}

int w2d_ctqmc_get_group_sums(w2d_ctqmc *h, int ngroups, double *gtau, double *gleg, double *dm_blocks,
                             double *histo, double *outer, double *scalars, int32_t *nused)
{
    if (!h || ngroups < 1 || ngroups > h->nchains) return W2D_ERR_ARG;
    if (!h->measured) return W2D_ERR_STATE;
    void *work = nullptr;
    HIPCHK(hipMalloc(&work, group_sums_bytes(h, ngroups)));
    const int st = group_sums_into(h, ngroups, work, gtau, gleg, dm_blocks, histo, outer, scalars, nused);
    (void)hipFree(work);
    return st;
}

int w2d_ctqmc_get_group_sums_noalloc(w2d_ctqmc *h, int ngroups, double *gtau, double *gleg, double *dm_blocks,
                                     double *histo, double *outer, double *scalars, int32_t *nused)
{
    if (!h || ngroups < 1 || ngroups > h->nchains) return W2D_ERR_ARG;
    if (!h->measured) return W2D_ERR_STATE;
    if (!h->scratch) return W2D_ERR_HIP;
    if (group_sums_bytes(h, ngroups) > sizeof(g_scratch)) {
        std::fprintf(stderr,
                     "w2d_ctqmc_get_group_sums_noalloc: %d groups of this handle need %zu bytes of scratch, "
                     "the library's static scratch holds %zu\n",
                     ngroups, group_sums_bytes(h, ngroups), sizeof(g_scratch));
        return W2D_ERR_CAPACITY;
    }
    return group_sums_into(h, ngroups, h->scratch, gtau, gleg, dm_blocks, histo, outer, scalars, nused);
}

int w2d_ctqmc_get_giw(w2d_ctqmc *h, int ngroups, int *niw, double *out)
{
    if (!h || !niw || ngroups < 0 || ngroups > h->nchains) return W2D_ERR_ARG;
    *niw = h->niw;
    if (!out) return W2D_OK;
    if (h->niw == 0 || !h->measured) return W2D_ERR_STATE;
    const int nf = h->dm.nf, nc = h->nchains;
    const size_t per = (size_t)nf * h->niw * 2;
    HIPCHK(hipStreamSynchronize(h->stream));
    std::vector<double> all((size_t)nc * per);
    std::vector<long long> seen((size_t)nc * nf * 2);
    HIPCHK(hipMemcpy(all.data(), h->giw, all.size() * sizeof(double), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(seen.data(), h->giw_seen, seen.size() * sizeof(long long), hipMemcpyDeviceToHost));
    std::vector<ChainHdr> hd;
    int st = read_headers(h, hd);
    if (st) return st;
    // every measurement of a chain that completed has its G(i w) term
    for (int c = 0; c < nc; ++c) {
        if (hd[c].err) continue;
        for (int f = 0; f < nf; ++f)
            if (seen[((size_t)c * nf + f) * 2 + 1] != hd[c].cnt_meas) {
                std::fprintf(stderr, "w2d_ctqmc: chain %d flavour %d has %lld G(iw) terms for %lld measurements\n", c,
                             f, seen[((size_t)c * nf + f) * 2 + 1], hd[c].cnt_meas);
                return W2D_ERR_STATE;
            }
    }
    if (ngroups == 0) {
        std::copy(all.begin(), all.end(), out);
        return W2D_OK;
    }
    // groups of contiguous chains as in w2d_ctqmc_get_group_sums
    std::fill(out, out + (size_t)ngroups * per, 0.0);
    for (int g = 0; g < ngroups; ++g) {
        const int c0 = (int)((long)g * nc / ngroups), c1 = (int)((long)(g + 1) * nc / ngroups);
        double *o = out + (size_t)g * per;
        for (int c = c0; c < c1; ++c) {
            if (hd[c].err) continue;
            const double *a = all.data() + (size_t)c * per;
            for (size_t i = 0; i < per; ++i) o[i] += a[i];
        }
    }
    return W2D_OK;
}

int w2d_ctqmc_get_gtau_blocks(w2d_ctqmc *h, int chain, int *nlevels, double *blocks)
{
    if (!h || chain < 0 || chain >= h->nchains || !nlevels) return W2D_ERR_ARG;
    if (h->dm.accum != 2) return W2D_ERR_STATE;
    *nlevels = h->dm.nlev;
    if (!blocks) return W2D_OK;
    if (!h->measured) return W2D_ERR_STATE;
    const long nc = 2L * h->model.nbands * h->ntau, nl = h->dm.nlev;
    std::vector<double> sq((size_t)nc * nl);
    ChainHdr x;
    HIPCHK(hipStreamSynchronize(h->stream));
    HIPCHK(hipMemcpy(&x, hdr_of(h, chain), sizeof(x), hipMemcpyDeviceToHost));
    const char *base = h->chains + (long)chain * h->L.stride + h->L.gacc + 2 * nc * nl * 8;
    HIPCHK(hipMemcpy(sq.data(), base, sq.size() * 8, hipMemcpyDeviceToHost));
    // autocorracc_blocks (Accumulators.F90:342-362): sqdiff / (n_l - 1), n_l = n >> l
    long long n = x.cnt_meas;
    for (long l = 0; l < nl; ++l, n /= 2)
        for (long i = 0; i < nc; ++i) blocks[l * nc + i] = sq[l * nc + i] / (n > 1 ? (double)(n - 1) : 0.0);
    return W2D_OK;
}

int w2d_ctqmc_init_counters(w2d_ctqmc *h)
{
    if (!h) return W2D_ERR_ARG;
    std::vector<ChainHdr> hd;
    int st = read_headers(h, hd);
    if (st) return st;
    for (auto &x : hd) reset_counters(x);
    if ((st = write_headers(h, hd)) || (st = clear_accpair(h))) return st;
    HIPCHK(hipStreamSynchronize(h->stream));
    return W2D_OK;
}

int w2d_ctqmc_calibrate(w2d_ctqmc *h, int list, int phase1pairnum, int phase2taugrid, int phase2pairnum,
                        int nprogress)
{
    if (!h || nprogress < 1 || phase2pairnum < 0) return W2D_ERR_ARG;
    if (list ? (phase1pairnum < 1 || phase1pairnum > h->L.apt_cap)
             : (phase2taugrid < 1 || phase2taugrid > h->L.naccpair_cap))
        return W2D_ERR_ARG;
    h->moved = true;
    h->aborted = 0;
    std::vector<ChainHdr> hd;
    int st = read_headers(h, hd);
    if (st) return st;
    for (auto &x : hd) {
        x.pos = 0;
        x.halt = w2d::HALT_NONE;
        if (list) {
            x.apt_n = phase1pairnum;   // AccPairTau(phase1pairnum), filled from apt_index
        } else {
            // AccPair(NBands, 2, phase2taugrid) over [0, taudiff_max] (CTQMC.F90:8103-8108)
            x.apt_n = 0;
            x.naccpair = phase2taugrid;
            x.accpair_max = x.taudiff_max;
        }
    }
    if ((st = write_headers(h, hd))) return st;
    if (!list && (st = clear_accpair(h))) return st;
    h->calib = make_int3(list ? 1 : 2, std::min(1000, nprogress), phase2pairnum);
    const long long per = 4000;
    long long from = 0;
    h->last_launches = 0;
    DISPATCH(prepare_sweep, h);
    for (;;) {
        launch_range(h, from, from + per, 0, per, -1, nullptr, 0);
        if (hipGetLastError() != hipSuccess) {
            h->calib = make_int3(0, 1, 0);
            return W2D_ERR_HIP;
        }
        if ((st = read_headers(h, hd))) break;
        bool again = false;
        long long f2;
        if ((st = resume_halted(h, hd, f2, again))) break;
        bool running = again;
        for (const auto &x : hd)
            if (x.err == 0 && x.halt == w2d::HALT_NONE) running = true;
        if (!running || h->aborted) break;
        long long lo = -1;
        for (const auto &x : hd)
            if (x.err == 0 && x.halt == w2d::HALT_NONE) lo = lo < 0 ? x.pos : std::min(lo, x.pos);
        from = lo < 0 ? from + per : lo;
    }
    h->calib = make_int3(0, 1, 0);
    if (st) return st;
    return batch_status(hd);
}

int w2d_ctqmc_get_pair_stats(w2d_ctqmc *h, int chain, double *accpairtau, int *napt, double *accpair,
                             int *naccpair)
{
    if (!h || chain < 0 || chain >= h->nchains) return W2D_ERR_ARG;
    HIPCHK(hipStreamSynchronize(h->stream));
    ChainHdr x;
    HIPCHK(hipMemcpy(&x, hdr_of(h, chain), sizeof(x), hipMemcpyDeviceToHost));
    const char *base = h->chains + (long)chain * h->L.stride;
    const int n = std::min(x.apt_index - 1, x.apt_n);
    if (napt) *napt = std::max(n, 0);
    if (naccpair) *naccpair = x.naccpair;
    if (accpairtau && n > 0) HIPCHK(hipMemcpy(accpairtau, base + h->L.aptau, (size_t)n * 8, hipMemcpyDeviceToHost));
    if (accpair) {
        const int nf = 2 * h->model.nbands;
        HIPCHK(hipMemcpy2D(accpair, (size_t)x.naccpair * 8, base + h->L.accpair, (size_t)h->L.naccpair_cap * 8,
                           (size_t)x.naccpair * 8, (size_t)nf, hipMemcpyDeviceToHost));
    }
    return W2D_OK;
}

int w2d_ctqmc_get_group_histograms(w2d_ctqmc *h, int ngroups, double *outer_state, double *sign_sst,
                                   double *sign_state, double *contrib_sst, double *contrib_state, double *accpair)
{
    if (!h || ngroups < 1 || ngroups > h->nchains) return W2D_ERR_ARG;
    if (!h->measured) return W2D_ERR_STATE;
    void *work = nullptr;
    HIPCHK(hipMalloc(&work, group_sums_bytes(h, ngroups)));
    int st = group_sums_into(h, ngroups, work, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr);
    const w2d::ChainLayout &L = h->L;
    const long nel = (L.stride - L.gtau) / 8;
    const double *part = static_cast<const double *>(work);
    auto put = [&](double *dst, long off, long n) -> int {
        if (!dst || n <= 0) return W2D_OK;
        HIPCHK(hipMemcpy2D(dst, (size_t)n * 8, part + (off - L.gtau) / 8, (size_t)nel * 8, (size_t)n * 8,
                           (size_t)ngroups, hipMemcpyDeviceToHost));
        return W2D_OK;
    };
    const long nst = h->model.nstates, nss = h->model.nsst;
    if (!st) st = put(outer_state, L.ohst, nst);
    if (!st) st = put(sign_sst, L.shss, nss);
    if (!st) st = put(sign_state, L.shst, nst);
    if (!st) st = put(contrib_sst, L.tcss, nss);
    if (!st) st = put(contrib_state, L.tcst, nst);
    if (!st) st = put(accpair, L.accpair, 2L * h->model.nbands * L.naccpair_cap);
    (void)hipFree(work);
    return st;
}

int w2d_ctqmc_get_steps(w2d_ctqmc *h, int chain, int *nmstep, double *gtau_mean_step, double *gtau_mid_step,
                        double *sign_step)
{
    if (!h || chain < 0 || chain >= h->nchains || !nmstep) return W2D_ERR_ARG;
    *nmstep = h->dm.nmstep;
    HIPCHK(hipStreamSynchronize(h->stream));
    const char *base = h->chains + (long)chain * h->L.stride;
    const size_t nf = 2 * (size_t)h->model.nbands, n = (size_t)h->dm.nmstep;
    if (gtau_mean_step && h->dm.step_mean) HIPCHK(hipMemcpy(gtau_mean_step, base + h->L.gmstep, n * nf * 8, hipMemcpyDeviceToHost));
    if (gtau_mid_step && h->dm.step_mid) HIPCHK(hipMemcpy(gtau_mid_step, base + h->L.gdstep, n * nf * 8, hipMemcpyDeviceToHost));
    if (sign_step && h->dm.step_sign) HIPCHK(hipMemcpy(sign_step, base + h->L.sstep, n * 8, hipMemcpyDeviceToHost));
    return W2D_OK;
}

int w2d_ctqmc_chain_status(const w2d_ctqmc *h, int32_t *status, int32_t *kfmax, int *capacity)
{
    if (!h) return W2D_ERR_ARG;
    if (capacity) *capacity = h->kcap;
    if (!status && !kfmax) return W2D_OK;
    std::vector<ChainHdr> hd;
    int st = read_headers(const_cast<w2d_ctqmc *>(h), hd);
    if (st) return st;
    const int nf = 2 * h->model.nbands;
    for (int c = 0; c < h->nchains; ++c) {
        if (status) status[c] = hd[c].err ? map_err(hd[c].err) : W2D_OK;
        if (kfmax)
            for (int f = 0; f < nf; ++f) kfmax[(long)c * nf + f] = hd[c].kfmax[f];
    }
    return W2D_OK;
}

int w2d_ctqmc_op_table_bytes(const w2d_ctqmc *h, int64_t *bytes)
{
    if (!h || !bytes) return W2D_ERR_ARG;
    *bytes = (int64_t)h->dm.nops * 8;
    return W2D_OK;
}

int w2d_ctqmc_get_orders(const w2d_ctqmc *h, int32_t *nop, int32_t *outer_sst)
{
    if (!h || !nop) return W2D_ERR_ARG;
    std::vector<ChainHdr> hd;
    int st = read_headers(const_cast<w2d_ctqmc *>(h), hd);
    if (st) return st;
    for (int c = 0; c < h->nchains; ++c) {
        nop[c] = hd[c].nop;
        if (outer_sst) outer_sst[c] = hd[c].outer;
    }
    return W2D_OK;
}

int w2d_ctqmc_set_abort(w2d_ctqmc *h, int flag)
{
    if (!h) return W2D_ERR_ARG;
    __atomic_store_n(h->abort_host, flag ? 1 : 0, __ATOMIC_SEQ_CST);
    return W2D_OK;
}

int w2d_ctqmc_aborted(const w2d_ctqmc *h, int *aborted)
{
    if (!h || !aborted) return W2D_ERR_ARG;
    *aborted = h->aborted;
    return W2D_OK;
}

int w2d_ctqmc_trap_signals(void)
{
    const int sigs[3] = {SIGTERM, SIGINT, SIGUSR1};
    for (int sg : sigs) {
        struct sigaction sa;
        std::memset(&sa, 0, sizeof(sa));
        sa.sa_handler = w2d_signal_handler;
        sigemptyset(&sa.sa_mask);
        if (sigaction(sg, &sa, nullptr) != 0) return W2D_ERR_ARG;
    }
    return W2D_OK;
}

}  // extern "C"
