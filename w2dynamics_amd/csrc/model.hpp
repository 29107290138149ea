// Host-side local Hilbert space of the impurity: occupation basis, quantum-number
// blocks ("superstates"), the local Hamiltonian and its eigenbasis, and the
// creation/annihilation operator blocks in that eigenbasis.
//
// Follows States.F90:58-184 (basis and blocking), Operator.F90:115-200 (c/c+ blocks
// and the Connect table), Operator.F90:481-542 and 1041-1124 (H_loc), and
// Operator.F90:637-670, 1130-1168 (diagonalisation, eigenbasis operators).
#pragma once

#include <cstdint>
#include <map>
#include <string>
#include <vector>

namespace w2d {

// "key = value #" parameter list with the reference's lookup rule: the first
// entry whose key starts with the requested name wins (Parameters.F90:63-223).
class ParamList {
public:
    explicit ParamList(const std::string &text);
    bool has(const std::string &key) const;
    std::string str(const std::string &key) const;
    double real(const std::string &key) const;
    long long integer(const std::string &key) const;

private:
    std::vector<std::pair<std::string, std::string>> entries_;
    const std::string *find(const std::string &key) const;
};

struct LocalModel {
    int nbands = 0;
    int nsst = 0;       // number of superstates (blocks)
    int nstates = 0;    // 4^nbands
    int nmax = 0;       // largest block
    double egs = 0.0;   // lowest eigenvalue over all blocks (Trace.F90:559-568)
    std::vector<int> sizes, offsets;
    std::vector<int> states;      // occupation-basis state of each block index
    std::vector<int> state_sst;   // block of each occupation-basis state
    std::vector<double> evals;    // eigenvalues, ascending within a block
    std::vector<double> evecs;    // eigenvectors, column-major per block, concatenated
    std::vector<long> evoff;      // offset of each block in evecs
    // connect[((s*nb + b)*2 + sp)*2 + ca0]: target block of c+ (ca0 = 0) or c
    // (ca0 = 1) acting on block s, -1 if none
    std::vector<int> connect;
    // operator blocks in the eigenbasis: opoff (same index as connect) into ops,
    // each a column-major (n_target x n_source) matrix, -1 if none
    std::vector<long> opoff;
    std::vector<double> ops;
};

// Builds the model; returns an empty string on success, else an error message.
// Quantum numbers: any of Nt, Szt, Qzt, plus All (the blocks refined by the
// connectivity of H with threshold epsblock times the largest input element).
std::string build_local_model(int nbands, const double *umatrix_2n4, const double *muimp_n2n2,
                              const std::string &quantum_numbers, LocalModel &m, double epsblock = 1e-13);

// d H_loc / d theta in the eigenbasis for inputs linear in (U matrix, muimp):
// V_s^T (H(dumat, dmuimp))_s V_s per block, column-major, concatenated in block
// order (sum of n_s^2 entries).  Error if the derivative leaves the blocks.
std::string derivative_blocks(const LocalModel &m, const double *dumat, const double *dmuimp,
                              std::vector<double> &out);

}  // namespace w2d
