// Host-side local Hilbert space (see model.hpp).
#include "model.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <sstream>

namespace w2d {

// --------------------------------------------------------------------------
// parameters
// --------------------------------------------------------------------------
static std::string trim(const std::string &s)
{
    size_t a = s.find_first_not_of(" \t\r");
    if (a == std::string::npos) return "";
    size_t b = s.find_last_not_of(" \t\r");
    return s.substr(a, b - a + 1);
}

// read_ParameterString (Parameters.F90:281-320): one "key = value" per line, a
// line starting with '#' is a comment, '#' ends the value, commas become blanks.
ParamList::ParamList(const std::string &text)
{
    std::istringstream in(text);
    std::string line;
    while (std::getline(in, line)) {
        if (!line.empty() && line[0] == '#') continue;
        for (char &c : line)
            if (c == '\t') c = ' ';
        size_t h = line.find('#');
        if (h != std::string::npos) line = line.substr(0, h);
        size_t e = line.find('=');
        if (e == std::string::npos) continue;
        std::string key = trim(line.substr(0, e));
        std::string val = trim(line.substr(e + 1));
        for (char &c : val)
            if (c == ',') c = ' ';
        entries_.emplace_back(key, val);
    }
}

const std::string *ParamList::find(const std::string &key) const
{
    for (const auto &kv : entries_)
        if (kv.first.compare(0, key.size(), key) == 0) return &kv.second;
    return nullptr;
}

bool ParamList::has(const std::string &key) const { return find(key) != nullptr; }

std::string ParamList::str(const std::string &key) const
{
    const std::string *v = find(key);
    return v ? *v : std::string();
}

double ParamList::real(const std::string &key) const
{
    const std::string *v = find(key);
    if (!v) return -9999999.0;   // RealNotFound
    std::string s = *v;
    for (char &c : s)
        if (c == 'd' || c == 'D') c = 'e';   // Fortran exponent letters
    return std::strtod(s.c_str(), nullptr);
}

long long ParamList::integer(const std::string &key) const
{
    const std::string *v = find(key);
    if (!v) return -9999999;     // IntegerNotFound
    return (long long)std::llround(real(key));
}

// --------------------------------------------------------------------------
// symmetric eigenproblem: cyclic Jacobi rotations, eigenvalues ascending
// --------------------------------------------------------------------------
static void jacobi_eigh(int n, std::vector<double> &a, std::vector<double> &w, std::vector<double> &v)
{
    v.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) v[i + (size_t)i * n] = 1.0;
    auto A = [&](int i, int j) -> double & { return a[i + (size_t)j * n]; };
    double norm = 0.0;
    for (double x : a) norm += x * x;
    norm = std::sqrt(norm);
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0.0;
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < j; ++i) off += A(i, j) * A(i, j);
        if (std::sqrt(off) <= 1e-17 * (norm > 0 ? norm : 1.0)) break;
        for (int p = 0; p < n; ++p)
            for (int q = p + 1; q < n; ++q) {
                double apq = A(p, q);
                if (apq == 0.0) continue;
                double app = A(p, p), aqq = A(q, q);
                double theta = (aqq - app) / (2.0 * apq);
                double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
                double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < n; ++k) {
                    double akp = A(k, p), akq = A(k, q);
                    A(k, p) = c * akp - s * akq;
                    A(k, q) = s * akp + c * akq;
                }
                for (int k = 0; k < n; ++k) {
                    double apk = A(p, k), aqk = A(q, k);
                    A(p, k) = c * apk - s * aqk;
                    A(q, k) = s * apk + c * aqk;
                }
                for (int k = 0; k < n; ++k) {
                    double vkp = v[k + (size_t)p * n], vkq = v[k + (size_t)q * n];
                    v[k + (size_t)p * n] = c * vkp - s * vkq;
                    v[k + (size_t)q * n] = s * vkp + c * vkq;
                }
            }
    }
    std::vector<int> idx(n);
    for (int i = 0; i < n; ++i) idx[i] = i;
    std::stable_sort(idx.begin(), idx.end(), [&](int x, int y) { return A(x, x) < A(y, y); });
    w.resize(n);
    std::vector<double> vs((size_t)n * n);
    for (int j = 0; j < n; ++j) {
        w[j] = A(idx[j], idx[j]);
        for (int i = 0; i < n; ++i) vs[i + (size_t)j * n] = v[i + (size_t)idx[j] * n];
    }
    v.swap(vs);
}

// --------------------------------------------------------------------------
// model
// --------------------------------------------------------------------------
static int popc(unsigned x) { return __builtin_popcount(x); }

// apply c+ (dag = true) or c on bit `bit` of `state`; returns -1 if it vanishes,
// sign = (-1)^(occupied bits below), Operator.F90:160-172
static int apply_op(int state, int bit, bool dag, double &sign)
{
    int occ = (state >> bit) & 1;
    if (dag == (occ != 0)) return -1;
    if (popc((unsigned)state & ((1u << bit) - 1u)) & 1) sign = -sign;
    return state ^ (1 << bit);
}

// Blocks of muimp_12 c+_1 c_2 + 1/2 U_1234 c+_1 c+_2 c_4 c_3 in the occupation
// basis (Operator.F90:481-542, 1041-1109).  Terms that map a block onto
// another are dropped, or with `strict` reported (returns false).
static bool hamiltonian_blocks(int nb, const double *umat, const double *muimp, const std::vector<int> &s2s,
                               const std::vector<std::vector<int>> &blk, const std::vector<int> &index_in_block,
                               const std::vector<int> &sizes, bool strict, std::vector<std::vector<double>> &h)
{
    const int nsst = (int)blk.size();
    bool ok = true;
    h.assign(nsst, {});
    for (int s = 0; s < nsst; ++s) h[s].assign((size_t)sizes[s] * sizes[s], 0.0);
    auto add_term = [&](double val, const int *bits, const bool *dag, int nops) {
        for (int s = 0; s < nsst; ++s)
            for (int st : blk[s]) {
                int cur = st;
                double sg = 1.0;
                for (int o = nops - 1; o >= 0 && cur >= 0; --o) cur = apply_op(cur, bits[o], dag[o], sg);
                if (cur < 0) continue;
                if (s2s[cur] != s) {
                    if (strict) ok = false;
                    continue;
                }
                int n = sizes[s];
                h[s][index_in_block[cur] + (size_t)index_in_block[st] * n] += val * sg;
            }
    };
    const int n2 = 2 * nb;
    for (int b1 = 0; b1 < nb; ++b1)
        for (int s1 = 0; s1 < 2; ++s1)
            for (int b2 = 0; b2 < nb; ++b2)
                for (int s2 = 0; s2 < 2; ++s2) {
                    double v = muimp[b1 + nb * s1 + 2 * nb * b2 + 2 * nb * nb * s2];
                    if (v == 0.0) continue;
                    int bits[2] = {b1 + nb * s1, b2 + nb * s2};
                    bool dag[2] = {true, false};
                    add_term(v, bits, dag, 2);
                }
    for (int i = 0; i < n2; ++i)
        for (int j = 0; j < n2; ++j)
            for (int k = 0; k < n2; ++k)
                for (int l = 0; l < n2; ++l) {
                    double u = umat[i + (size_t)n2 * (j + (size_t)n2 * (k + (size_t)n2 * l))];
                    if (u == 0.0) continue;
                    // flavour index 2*b + s (CTQMC.F90 u_matrix layout)
                    int b[4] = {i / 2, j / 2, k / 2, l / 2}, sp[4] = {i % 2, j % 2, k % 2, l % 2};
                    int bits[4] = {b[0] + nb * sp[0], b[1] + nb * sp[1], b[3] + nb * sp[3], b[2] + nb * sp[2]};
                    bool dag[4] = {true, true, false, false};
                    add_term(0.5 * u, bits, dag, 4);
                }
    return ok;
}

// analyze_hamiltonian (Operator.F90:303-442): within every block of s2s (its
// states blk, H in the occupation basis h), the connected components of
// |H_ij| > thr, numbered in order of discovery (breadth-first from the block's
// lowest unassigned state); then, while some operator c or c+ maps two states
// of one block into two different blocks, those two blocks are merged (the
// larger label into the smaller, labels above it shifted down), rescanning
// from the first operator after every merge.  Returns the number of blocks.
static int refine_blocks(int nb, const std::vector<std::vector<int>> &blk,
                         const std::vector<std::vector<double>> &h, double thr, std::vector<int> &s2s)
{
    const int ns = 1 << (2 * nb);
    s2s.assign(ns, -1);
    int nsub = -1;
    for (size_t sst = 0; sst < blk.size(); ++sst) {
        const std::vector<int> &st = blk[sst];
        const int n = (int)st.size();
        const std::vector<double> &H = h[sst];
        std::vector<std::vector<int>> nets{{0}};
        s2s[st[0]] = ++nsub;
        for (size_t k = 0; k < nets.size(); ++k) {
            for (size_t q = 0; q < nets[k].size(); ++q) {
                const int node = nets[k][q];
                for (int edge = 0; edge < n; ++edge) {
                    const double v = H[node + (size_t)edge * n];
                    if (v == 0.0 || !(std::fabs(v) > thr)) continue;
                    if (std::find(nets[k].begin(), nets[k].end(), edge) == nets[k].end()) {
                        nets[k].push_back(edge);
                        s2s[st[edge]] = nsub;
                    }
                }
            }
            for (int node = 0; node < n; ++node)
                if (s2s[st[node]] == -1) {
                    nets.push_back({node});
                    s2s[st[node]] = ++nsub;
                    break;
                }
        }
    }
    std::vector<int> cs(ns);
    for (bool change = true; change;) {
        change = false;
        for (int ib = 0; ib < nb && !change; ++ib)
            for (int sp = 0; sp < 2 && !change; ++sp)
                for (int ca = 1; ca <= 2 && !change; ++ca) {
                    const int bit = ib + nb * sp;
                    for (int i = 0; i < ns; ++i) {
                        const int occ = (i >> bit) & 1;
                        cs[i] = ((ca == 1 && !occ) || (ca == 2 && occ)) ? s2s[i ^ (1 << bit)] : -1;
                    }
                    for (int i1 = 0; i1 < ns && !change; ++i1) {
                        if (cs[i1] == -1) continue;
                        for (int i2 = i1 + 1; i2 < ns; ++i2) {
                            if (s2s[i2] != s2s[i1] || cs[i2] == -1 || cs[i2] == cs[i1]) continue;
                            const int big = std::max(cs[i1], cs[i2]), small = std::min(cs[i1], cs[i2]);
                            for (int i = 0; i < ns; ++i) {
                                if (s2s[i] == big) s2s[i] = small;
                                if (s2s[i] > big) --s2s[i];
                            }
                            change = true;
                            break;
                        }
                    }
                }
    }
    return *std::max_element(s2s.begin(), s2s.end()) + 1;
}

std::string build_local_model(int nb, const double *umat, const double *muimp,
                              const std::string &qns, LocalModel &m, double epsblock)
{
    if (nb < 1 || nb > 7) return "nbands must be 1..7";
    const bool q_nt = qns.find("Nt") != std::string::npos;
    const bool q_sz = qns.find("Szt") != std::string::npos;
    const bool q_qz = qns.find("Qzt") != std::string::npos;
    const bool q_all = qns.find("All") != std::string::npos;
    for (const char *bad : {"Azt", "Lzt", "Jzt"})
        if (qns.find(bad) != std::string::npos)
            return std::string("unsupported quantum number ") + bad;
    const int ns = 1 << (2 * nb);
    m.nbands = nb;
    m.nstates = ns;
    auto nt = [&](int st) { return popc((unsigned)st); };
    auto sz2 = [&](int st) {   // 2*Szt, bits 0..N-1 count -1/2 (States.F90:224-234)
        return popc((unsigned)st >> nb) - popc((unsigned)st & ((1u << nb) - 1u));
    };
    auto qz = [&](int st) {    // get_Qzt (States.F90:237-247)
        double q = 0;
        for (int b = 0; b < nb; ++b) {
            int u = (st >> b) & 1, d = (st >> (b + nb)) & 1;
            if (u != d) q += std::pow(10.0, b);
        }
        return q;
    };
    // qns2substates (States.F90:151-184)
    std::vector<int> s2s(ns, -1);
    int nsst = 0;
    for (int a = 0; a < ns; ++a) {
        if (s2s[a] != -1) continue;
        s2s[a] = nsst;
        for (int b = a; b < ns; ++b) {
            if (s2s[b] != -1) continue;
            if (q_nt && nt(a) != nt(b)) continue;
            if (q_sz && sz2(a) != sz2(b)) continue;
            if (q_qz && std::fabs(qz(a) - qz(b)) >= 1e-12) continue;
            s2s[b] = nsst;
        }
        ++nsst;
    }
    if (q_all) {
        // "All" (CTQMC.F90:7971-7989): H on the blocks above, its connectivity
        // refines them; threshold EPSBLOCK relative to the largest input element
        std::vector<std::vector<int>> blk0(nsst);
        std::vector<int> iib0(ns), sizes0(nsst);
        for (int st = 0; st < ns; ++st) {
            iib0[st] = (int)blk0[s2s[st]].size();
            blk0[s2s[st]].push_back(st);
        }
        for (int s = 0; s < nsst; ++s) sizes0[s] = (int)blk0[s].size();
        std::vector<std::vector<double>> h0;
        hamiltonian_blocks(nb, umat, muimp, s2s, blk0, iib0, sizes0, false, h0);
        for (int s = 0; s < nsst; ++s) {   // force_hermitian
            const int n = sizes0[s];
            std::vector<double> &a = h0[s];
            for (int i = 0; i < n; ++i)
                for (int j = i + 1; j < n; ++j) {
                    const double x = 0.5 * (a[i + (size_t)j * n] + a[j + (size_t)i * n]);
                    a[i + (size_t)j * n] = a[j + (size_t)i * n] = x;
                }
        }
        double big = 0.0;
        for (int i = 0; i < 4 * nb * nb; ++i) big = std::max(big, std::fabs(muimp[i]));
        for (long i = 0; i < 16L * nb * nb * nb * nb; ++i) big = std::max(big, std::fabs(umat[i]));
        nsst = refine_blocks(nb, blk0, h0, epsblock * big, s2s);
    }
    m.nsst = nsst;
    m.state_sst = s2s;
    m.sizes.assign(nsst, 0);
    m.offsets.assign(nsst, 0);
    std::vector<std::vector<int>> blk(nsst);
    for (int st = 0; st < ns; ++st) blk[s2s[st]].push_back(st);
    std::vector<int> index_in_block(ns);
    m.states.clear();
    int off = 0;
    m.nmax = 0;
    for (int s = 0; s < nsst; ++s) {
        m.sizes[s] = (int)blk[s].size();
        m.offsets[s] = off;
        off += m.sizes[s];
        m.nmax = std::max(m.nmax, m.sizes[s]);
        for (size_t i = 0; i < blk[s].size(); ++i) {
            index_in_block[blk[s][i]] = (int)i;
            m.states.push_back(blk[s][i]);
        }
    }
    // Connect table (Operator.F90:140-190)
    m.connect.assign((size_t)nsst * nb * 4, -1);
    auto cidx = [&](int s, int b, int sp, int ca0) { return ((s * nb + b) * 2 + sp) * 2 + ca0; };
    for (int s = 0; s < nsst; ++s)
        for (int b = 0; b < nb; ++b)
            for (int sp = 0; sp < 2; ++sp)
                for (int ca0 = 0; ca0 < 2; ++ca0) {
                    int target = -1;
                    for (int st : blk[s]) {
                        double sg = 1.0;
                        int t = apply_op(st, b + nb * sp, ca0 == 0, sg);
                        if (t < 0) continue;
                        if (target >= 0 && s2s[t] != target) return "operator maps a block onto several blocks";
                        target = s2s[t];
                    }
                    m.connect[cidx(s, b, sp, ca0)] = target;
                }
    // H_loc per block: one-particle term plus 1/2 U_1234 c+_1 c+_2 c_4 c_3
    std::vector<std::vector<double>> h;
    hamiltonian_blocks(nb, umat, muimp, s2s, blk, index_in_block, m.sizes, false, h);
    // force_hermitian, diagonalise, eigenbasis
    m.evals.assign(ns, 0.0);
    m.evoff.assign(nsst, 0);
    m.evecs.clear();
    std::vector<std::vector<double>> V(nsst);
    for (int s = 0; s < nsst; ++s) {
        int n = m.sizes[s];
        std::vector<double> &a = h[s];
        for (int i = 0; i < n; ++i)
            for (int j = i + 1; j < n; ++j) {
                double x = 0.5 * (a[i + (size_t)j * n] + a[j + (size_t)i * n]);
                a[i + (size_t)j * n] = a[j + (size_t)i * n] = x;
            }
        std::vector<double> w;
        jacobi_eigh(n, a, w, V[s]);
        for (int i = 0; i < n; ++i) m.evals[m.offsets[s] + i] = w[i];
        m.evoff[s] = (long)m.evecs.size();
        m.evecs.insert(m.evecs.end(), V[s].begin(), V[s].end());
    }
    m.egs = m.evals[0];
    for (double e : m.evals) m.egs = std::min(m.egs, e);
    // operator blocks in the eigenbasis: V_t^T Psi V_s (Operator.F90:1130-1168)
    m.opoff.assign(m.connect.size(), -1);
    m.ops.clear();
    for (int s = 0; s < nsst; ++s)
        for (int b = 0; b < nb; ++b)
            for (int sp = 0; sp < 2; ++sp)
                for (int ca0 = 0; ca0 < 2; ++ca0) {
                    int t = m.connect[cidx(s, b, sp, ca0)];
                    if (t < 0) continue;
                    int nsrc = m.sizes[s], ntg = m.sizes[t];
                    std::vector<double> psi((size_t)ntg * nsrc, 0.0);
                    for (int i = 0; i < nsrc; ++i) {
                        double sg = 1.0;
                        int r = apply_op(blk[s][i], b + nb * sp, ca0 == 0, sg);
                        if (r >= 0) psi[index_in_block[r] + (size_t)i * ntg] = sg;
                    }
                    std::vector<double> tmp((size_t)ntg * nsrc, 0.0), out((size_t)ntg * nsrc, 0.0);
                    for (int j = 0; j < nsrc; ++j)       // tmp = psi * V_s
                        for (int k = 0; k < nsrc; ++k) {
                            double v = V[s][k + (size_t)j * nsrc];
                            if (v == 0.0) continue;
                            for (int i = 0; i < ntg; ++i) tmp[i + (size_t)j * ntg] += psi[i + (size_t)k * ntg] * v;
                        }
                    for (int j = 0; j < nsrc; ++j)       // out = V_t^T * tmp
                        for (int i = 0; i < ntg; ++i) {
                            double acc = 0.0;
                            for (int k = 0; k < ntg; ++k) acc += V[t][k + (size_t)i * ntg] * tmp[k + (size_t)j * ntg];
                            out[i + (size_t)j * ntg] = acc;
                        }
                    m.opoff[cidx(s, b, sp, ca0)] = (long)m.ops.size();
                    m.ops.insert(m.ops.end(), out.begin(), out.end());
                }
    return "";
}

std::string derivative_blocks(const LocalModel &m, const double *dumat, const double *dmuimp,
                              std::vector<double> &out)
{
    const int nsst = m.nsst;
    std::vector<std::vector<int>> blk(nsst);
    std::vector<int> index_in_block(m.nstates);
    for (int s = 0; s < nsst; ++s)
        for (int i = 0; i < m.sizes[s]; ++i) {
            const int st = m.states[m.offsets[s] + i];
            blk[s].push_back(st);
            index_in_block[st] = i;
        }
    std::vector<std::vector<double>> h;
    if (!hamiltonian_blocks(m.nbands, dumat, dmuimp, m.state_sst, blk, index_in_block, m.sizes, true, h))
        return "the derivative of H_loc does not conserve the quantum numbers";
    out.clear();
    for (int s = 0; s < nsst; ++s) {
        const int n = m.sizes[s];
        const double *V = m.evecs.data() + m.evoff[s];
        std::vector<double> a = h[s], t((size_t)n * n, 0.0);
        for (int i = 0; i < n; ++i)          // symmetric part, as force_hermitian
            for (int j = i + 1; j < n; ++j) {
                double x = 0.5 * (a[i + (size_t)j * n] + a[j + (size_t)i * n]);
                a[i + (size_t)j * n] = a[j + (size_t)i * n] = x;
            }
        for (int j = 0; j < n; ++j)          // t = a V
            for (int k = 0; k < n; ++k) {
                const double v = V[k + (size_t)j * n];
                if (v == 0.0) continue;
                for (int i = 0; i < n; ++i) t[i + (size_t)j * n] += a[i + (size_t)k * n] * v;
            }
        for (int j = 0; j < n; ++j)          // V^T t
            for (int i = 0; i < n; ++i) {
                double acc = 0.0;
                for (int k = 0; k < n; ++k) acc += V[k + (size_t)i * n] * t[k + (size_t)j * n];
                out.push_back(acc);
            }
    }
    return "";
}

}  // namespace w2d
