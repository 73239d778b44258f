// eb_setup.h — host-side model preparation for libendobeams_hip.so (runs once per eb_create).
//
// Device matrices are stored as node-block rows (BSR, 6x6 row-major blocks): block row n holds
// one block per neighbour m of node n (n itself included), sorted by m. The CSC pattern of the
// reference (src/sparse_matrices.jl:6-104) has exactly the same non-zeros; csc2bsr maps between
// the two orders and is applied only when a matrix crosses the ABI.
#pragma once
#include <stdint.h>

#include <string>
#include <vector>

namespace ebh {

struct Pattern {
    int64_t N = 0, E = 0, ndofs = 0, nnz = 0, nb = 0, n_free = 0, nnz_free = 0;
    std::vector<int32_t> n1, n2;                 // 0-based beam nodes
    std::vector<std::vector<int32_t>> adj;       // sorted neighbours incl. self
    std::vector<int64_t> bptr;                   // N+1 block-row pointers
    std::vector<int32_t> bcol;                   // nb block columns
    std::vector<int64_t> colptr, rowval;         // CSC, 0-based
    std::vector<int64_t> csc2bsr;                // nnz
    std::vector<int64_t> smap_csc;               // 144*E, CSC positions (column-major local index)
    std::vector<int64_t> sfree_csc;              // nnz_free
    std::vector<int32_t> sfree_bsr;              // nnz_free
    std::vector<uint8_t> is_fixed;               // ndofs
    std::vector<int32_t> fixed, freed;           // dof lists
    std::vector<int32_t> fixidx;                 // ndofs: index among fixed dofs or -1
    // gather lists: per block, the (beam, row slot, col slot) sub-blocks that add into it,
    // packed as (e << 2) | (slot_row << 1) | slot_col, in increasing beam order
    std::vector<int64_t> blk_ptr;                // nb+1
    std::vector<uint32_t> blk_src;
    // per node, incident beams packed as (e << 1) | slot, increasing beam order
    std::vector<int64_t> inc_ptr;                // N+1
    std::vector<uint32_t> inc_src;
};

// chain (line) decomposition of the node graph for the block-tridiagonal preconditioner
struct Chains {
    std::vector<int64_t> ptr;                    // n_chains+1
    std::vector<int32_t> node;                   // chain-ordered nodes
    std::vector<int32_t> diag_blk, lo_blk, up_blk;  // per chain position; -1 when absent
    std::vector<uint64_t> mD, mL, mU;           // per chain position: entries touching a fixed dof
    std::vector<int32_t> hbase;                  // per chain position: first cyclic-reduction history slot
    int64_t n_hist = 0;
    int64_t max_len = 0;
    // closures: chains of at least three positions whose first and last nodes are coupled (a
    // ring), up to CLOSURE_MAX_LEN positions. Per closure: its chain, the coupling blocks
    // (first, last) and (last, first), their fixed-dof masks, and the offset of its 12 columns of
    // T^-1 U (72 doubles per position)
    std::vector<int32_t> cl_chain, cl_fl, cl_lf;
    std::vector<uint64_t> cl_mfl, cl_mlf;
    std::vector<int64_t> cl_zoff;   // n_closures+1
    int64_t cl_max_len = 0;
};
constexpr int64_t CLOSURE_MAX_LEN = 640;

bool build_pattern(Pattern& p, int64_t N, int64_t E, const int64_t* conn1b, int64_t n_fixed, const int64_t* fixed1b,
                   std::string& err);

// mode 0: path cover preferring straight continuation (exact for chains of beams);
// mode 1: every node alone (6x6 block-Jacobi).
// With `across`, the walk uses only the couplings that none of across's chains holds between
// consecutive positions (a second, crossing cover: the other lines of a net); a node with no such
// coupling left is a chain of its own.
// `cover` (per node, may be null: all): the nodes the chains may hold
void build_chains(Chains& c, const Pattern& p, const double* X0, int mode, const Chains* across = nullptr,
                  const std::vector<uint8_t>* cover = nullptr);
// per node: 1 when it has a coupling that no chain of `c` holds
std::vector<uint8_t> nodes_left_out(const Chains& c, const Pattern& p);
// couplings (node pairs with a matrix block) that no chain of c holds between consecutive
// positions, counted once per pair; closures count as held
int64_t chains_left_out(const Chains& c, const Pattern& p);

// Quadratic B-spline Reflect(OnCell()) coefficients, padded by one cell per side
void prefilter_sdf(const int64_t n[3], const double* samples, std::vector<double>& coef);
// LU factors of the n×n system of prefilter_sdf along one axis: modified super-diagonal cp and
// reciprocal pivots inv (the device prefilter uses the same factors)
void tri_factors(int64_t n, std::vector<double>& cp, std::vector<double>& inv);

}  // namespace ebh
