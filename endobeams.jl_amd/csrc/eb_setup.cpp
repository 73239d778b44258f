// eb_setup.cpp — host-side model preparation (see eb_setup.h).
#include "eb_setup.h"

#include <algorithm>
#include <cmath>

namespace ebh {

bool build_pattern(Pattern& p, int64_t N, int64_t E, const int64_t* conn1b, int64_t n_fixed, const int64_t* fixed1b,
                   std::string& err) {
    p.N = N; p.E = E; p.ndofs = 6 * N;
    p.n1.resize(E); p.n2.resize(E);
    for (int64_t e = 0; e < E; ++e) {
        int64_t a = conn1b[2 * e] - 1, b = conn1b[2 * e + 1] - 1;
        if (a < 0 || a >= N || b < 0 || b >= N || a == b) { err = "invalid connectivity at beam " + std::to_string(e + 1); return false; }
        p.n1[e] = (int32_t)a; p.n2[e] = (int32_t)b;
    }
    if (E >= (int64_t(1) << 30)) { err = "too many beams"; return false; }
    p.adj.assign(N, {});
    for (int64_t n = 0; n < N; ++n) p.adj[n].push_back((int32_t)n);
    for (int64_t e = 0; e < E; ++e) { p.adj[p.n1[e]].push_back(p.n2[e]); p.adj[p.n2[e]].push_back(p.n1[e]); }
    for (auto& a : p.adj) { std::sort(a.begin(), a.end()); a.erase(std::unique(a.begin(), a.end()), a.end()); }
    for (int64_t n = 0; n < N; ++n)
        if (p.adj[n].size() < 2) { err = "node " + std::to_string(n + 1) + " belongs to no beam"; return false; }
    // BSR
    p.bptr.assign(N + 1, 0);
    for (int64_t n = 0; n < N; ++n) p.bptr[n + 1] = p.bptr[n] + (int64_t)p.adj[n].size();
    p.nb = p.bptr[N];
    p.nnz = 36 * p.nb;
    p.bcol.resize(p.nb);
    for (int64_t n = 0; n < N; ++n)
        for (size_t k = 0; k < p.adj[n].size(); ++k) p.bcol[p.bptr[n] + k] = p.adj[n][k];
    auto blk = [&](int32_t n, int32_t m) -> int64_t {
        const auto& a = p.adj[n];
        return p.bptr[n] + (std::lower_bound(a.begin(), a.end(), m) - a.begin());
    };
    // CSC: column j = 6m+c lists rows of neighbours n of m (sorted), 6 each
    p.colptr.assign(p.ndofs + 1, 0);
    for (int64_t j = 0; j < p.ndofs; ++j) p.colptr[j + 1] = p.colptr[j] + 6 * (int64_t)p.adj[j / 6].size();
    p.rowval.resize(p.nnz);
    p.csc2bsr.resize(p.nnz);
    for (int64_t j = 0; j < p.ndofs; ++j) {
        int32_t m = (int32_t)(j / 6);
        int c = (int)(j % 6);
        int64_t q = p.colptr[j];
        for (int32_t n : p.adj[m])
            for (int a = 0; a < 6; ++a) {
                p.rowval[q] = 6 * (int64_t)n + a;
                p.csc2bsr[q] = 36 * blk(n, m) + 6 * a + c;
                ++q;
            }
    }
    auto csc_pos = [&](int64_t row, int64_t col) -> int64_t {
        const auto& a = p.adj[col / 6];
        int64_t pos = std::lower_bound(a.begin(), a.end(), (int32_t)(row / 6)) - a.begin();
        return p.colptr[col] + 6 * pos + row % 6;
    };
    p.smap_csc.resize(144 * E);
    for (int64_t e = 0; e < E; ++e) {
        int64_t dofs[12];
        for (int k = 0; k < 6; ++k) { dofs[k] = 6 * (int64_t)p.n1[e] + k; dofs[6 + k] = 6 * (int64_t)p.n2[e] + k; }
        for (int jl = 0; jl < 12; ++jl)
            for (int il = 0; il < 12; ++il) p.smap_csc[144 * e + 12 * jl + il] = csc_pos(dofs[il], dofs[jl]);
    }
    // fixed / free dofs
    p.is_fixed.assign(p.ndofs, 0);
    p.fixidx.assign(p.ndofs, -1);
    for (int64_t k = 0; k < n_fixed; ++k) {
        int64_t f = fixed1b[k] - 1;
        if (f < 0 || f >= p.ndofs || (k > 0 && fixed1b[k] <= fixed1b[k - 1])) { err = "fixed dofs must be increasing and in range"; return false; }
        p.is_fixed[f] = 1; p.fixidx[f] = (int32_t)k; p.fixed.push_back((int32_t)f);
    }
    for (int64_t j = 0; j < p.ndofs; ++j) if (!p.is_fixed[j]) p.freed.push_back((int32_t)j);
    p.n_free = (int64_t)p.freed.size();
    p.sfree_csc.clear(); p.sfree_bsr.clear();
    for (int64_t j = 0; j < p.ndofs; ++j) {
        if (p.is_fixed[j]) continue;
        for (int64_t q = p.colptr[j]; q < p.colptr[j + 1]; ++q)
            if (!p.is_fixed[p.rowval[q]]) { p.sfree_csc.push_back(q); p.sfree_bsr.push_back((int32_t)p.csc2bsr[q]); }
    }
    p.nnz_free = (int64_t)p.sfree_csc.size();
    // gather lists
    std::vector<int64_t> cnt(p.nb + 1, 0);
    for (int64_t e = 0; e < E; ++e) {
        int32_t nn[2] = {p.n1[e], p.n2[e]};
        for (int sr = 0; sr < 2; ++sr) for (int sc = 0; sc < 2; ++sc) cnt[blk(nn[sr], nn[sc]) + 1]++;
    }
    p.blk_ptr.assign(p.nb + 1, 0);
    for (int64_t b = 0; b < p.nb; ++b) p.blk_ptr[b + 1] = p.blk_ptr[b] + cnt[b + 1];
    p.blk_src.resize(p.blk_ptr[p.nb]);
    std::vector<int64_t> fill(p.blk_ptr.begin(), p.blk_ptr.end() - 1);
    for (int64_t e = 0; e < E; ++e) {
        int32_t nn[2] = {p.n1[e], p.n2[e]};
        for (int sr = 0; sr < 2; ++sr)
            for (int sc = 0; sc < 2; ++sc)
                p.blk_src[fill[blk(nn[sr], nn[sc])]++] = ((uint32_t)e << 2) | (uint32_t)(sr << 1) | (uint32_t)sc;
    }
    p.inc_ptr.assign(N + 1, 0);
    for (int64_t e = 0; e < E; ++e) { p.inc_ptr[p.n1[e] + 1]++; p.inc_ptr[p.n2[e] + 1]++; }
    for (int64_t n = 0; n < N; ++n) p.inc_ptr[n + 1] += p.inc_ptr[n];
    p.inc_src.resize(p.inc_ptr[N]);
    std::vector<int64_t> f2(p.inc_ptr.begin(), p.inc_ptr.end() - 1);
    for (int64_t e = 0; e < E; ++e) {
        p.inc_src[f2[p.n1[e]]++] = ((uint32_t)e << 1) | 0u;
        p.inc_src[f2[p.n2[e]]++] = ((uint32_t)e << 1) | 1u;
    }
    return true;
}

// per node, the sorted neighbours it is joined to by a chain of c (consecutive positions, and the
// two ends of a closure)
static std::vector<std::vector<int32_t>> chain_links(const Chains& c, int64_t N) {
    std::vector<std::vector<int32_t>> l((size_t)N);
    for (size_t ch = 0; ch + 1 < c.ptr.size(); ++ch)
        for (int64_t i = c.ptr[ch]; i + 1 < c.ptr[ch + 1]; ++i) {
            l[c.node[i]].push_back(c.node[i + 1]);
            l[c.node[i + 1]].push_back(c.node[i]);
        }
    for (size_t k = 0; k < c.cl_chain.size(); ++k) {
        const int32_t f = c.node[c.ptr[c.cl_chain[k]]], e = c.node[c.ptr[c.cl_chain[k] + 1] - 1];
        l[f].push_back(e);
        l[e].push_back(f);
    }
    for (auto& v : l) std::sort(v.begin(), v.end());
    return l;
}

int64_t chains_left_out(const Chains& c, const Pattern& p) {
    const auto links = chain_links(c, p.N);
    int64_t n = 0;
    for (int32_t a = 0; a < p.N; ++a)
        for (int32_t m : p.adj[a])
            if (m > a && !std::binary_search(links[a].begin(), links[a].end(), m)) ++n;
    return n;
}

std::vector<uint8_t> nodes_left_out(const Chains& c, const Pattern& p) {
    const auto links = chain_links(c, p.N);
    std::vector<uint8_t> out((size_t)p.N, 0);
    for (int32_t a = 0; a < p.N; ++a)
        for (int32_t m : p.adj[a])
            if (m != a && !std::binary_search(links[a].begin(), links[a].end(), m)) out[a] = 1;
    return out;
}

void build_chains(Chains& c, const Pattern& p, const double* X0, int mode, const Chains* across,
                  const std::vector<uint8_t>* cover) {
    const int64_t N = p.N;
    const std::vector<std::vector<int32_t>> taken = across ? chain_links(*across, N) : std::vector<std::vector<int32_t>>();
    // a coupling the walk may follow: not a self-coupling, and not already held by `across`
    auto free_link = [&](int32_t n, int32_t m) {
        return m != n && (!across || !std::binary_search(taken[n].begin(), taken[n].end(), m));
    };
    c.ptr.assign(1, 0);
    c.node.clear();
    // nodes outside `cover` start as used: no chain holds them
    std::vector<uint8_t> used(N, 0);
    if (cover)
        for (int64_t n = 0; n < N; ++n) used[n] = (*cover)[n] ? 0 : 1;
    auto push = [&](int32_t n) { c.node.push_back(n); used[n] = 1; };
    auto walk = [&](int32_t start) {
        int32_t prev = -1, cur = start;
        push(cur);
        while (true) {
            int32_t best = -1;
            double bestdot = -1e300;
            for (int32_t m : p.adj[cur]) {
                if (!free_link(cur, m) || used[m]) continue;
                double d;
                if (prev < 0) {
                    d = -(double)m;  // first step: smallest index
                } else {
                    double a[3], b[3], na = 0, nb = 0, dd = 0;
                    for (int k = 0; k < 3; ++k) {
                        a[k] = X0[3 * cur + k] - X0[3 * prev + k];
                        b[k] = X0[3 * m + k] - X0[3 * cur + k];
                        na += a[k] * a[k]; nb += b[k] * b[k]; dd += a[k] * b[k];
                    }
                    d = dd / std::sqrt(na * nb + 1e-300);
                }
                if (d > bestdot) { bestdot = d; best = m; }
            }
            if (best < 0) break;
            prev = cur;
            cur = best;
            push(cur);
        }
        c.ptr.push_back((int64_t)c.node.size());
    };
    if (mode == 1) {
        for (int32_t n = 0; n < N; ++n) {
            if (used[n]) continue;
            push(n);
            c.ptr.push_back((int64_t)c.node.size());
        }
    } else {
        // start from path ends (one free neighbour), then from whatever is left (closed loops)
        bool progress = true;
        while (progress) {
            progress = false;
            for (int32_t n = 0; n < N; ++n) {
                if (used[n]) continue;
                int free_nb = 0;
                for (int32_t m : p.adj[n]) if (free_link(n, m) && !used[m]) ++free_nb;
                if (free_nb <= 1) { walk(n); progress = true; }
            }
        }
        for (int32_t n = 0; n < N; ++n) if (!used[n]) walk(n);
    }
    const int64_t L = (int64_t)c.node.size();
    c.diag_blk.assign(L, -1); c.lo_blk.assign(L, -1); c.up_blk.assign(L, -1);
    auto blk = [&](int32_t n, int32_t m) -> int32_t {
        const auto& a = p.adj[n];
        auto it = std::lower_bound(a.begin(), a.end(), m);
        if (it == a.end() || *it != m) return -1;
        return (int32_t)(p.bptr[n] + (it - a.begin()));
    };
    c.max_len = 0;
    for (size_t ch = 0; ch + 1 < c.ptr.size(); ++ch) {
        int64_t s = c.ptr[ch], e = c.ptr[ch + 1];
        c.max_len = std::max(c.max_len, e - s);
        for (int64_t i = s; i < e; ++i) {
            c.diag_blk[i] = blk(c.node[i], c.node[i]);
            if (i > s) c.lo_blk[i] = blk(c.node[i], c.node[i - 1]);
            if (i + 1 < e) c.up_blk[i] = blk(c.node[i], c.node[i + 1]);
        }
    }
    // entry k = 6a+b of block (n, m) is replaced when dof a of n or dof b of m is fixed
    auto mask = [&](int32_t n, int32_t m) -> uint64_t {
        uint64_t bits = 0;
        for (int a = 0; a < 6; ++a)
            for (int b = 0; b < 6; ++b)
                if (p.is_fixed[6 * (int64_t)n + a] || p.is_fixed[6 * (int64_t)m + b]) bits |= uint64_t(1) << (6 * a + b);
        return bits;
    };
    c.mD.assign(L, 0); c.mL.assign(L, 0); c.mU.assign(L, 0);
    for (size_t ch = 0; ch + 1 < c.ptr.size(); ++ch) {
        int64_t s = c.ptr[ch], e = c.ptr[ch + 1];
        for (int64_t i = s; i < e; ++i) {
            c.mD[i] = mask(c.node[i], c.node[i]);
            if (i > s) c.mL[i] = mask(c.node[i], c.node[i - 1]);
            if (i + 1 < e) c.mU[i] = mask(c.node[i], c.node[i + 1]);
        }
    }
    // cyclic reduction keeps position p (chain-relative) through tz(p) levels (p > 0) and position 0
    // through all ceil(log2 n) levels; one history slot per level kept
    c.hbase.assign(L, 0);
    int64_t slot = 0;
    for (size_t ch = 0; ch + 1 < c.ptr.size(); ++ch) {
        int64_t s = c.ptr[ch], e = c.ptr[ch + 1], n = e - s;
        int nl = 0;
        while ((int64_t(1) << nl) < n) ++nl;
        for (int64_t i = s; i < e; ++i) {
            int64_t pos = i - s;
            int kept = 0;
            if (pos == 0) kept = nl;
            else while (((pos >> kept) & 1) == 0) ++kept;
            c.hbase[i] = (int32_t)slot;
            slot += kept;
        }
    }
    c.n_hist = slot;
    // couplings between a chain's two ends, which the block-tridiagonal factorisation leaves out
    c.cl_chain.clear(); c.cl_fl.clear(); c.cl_lf.clear(); c.cl_mfl.clear(); c.cl_mlf.clear();
    c.cl_zoff.assign(1, 0);
    c.cl_max_len = 0;
    for (size_t ch = 0; ch + 1 < c.ptr.size(); ++ch) {
        const int64_t s = c.ptr[ch], e = c.ptr[ch + 1], n = e - s;
        if (n < 3 || n > CLOSURE_MAX_LEN) continue;
        const int32_t f = c.node[s], l = c.node[e - 1];
        const int32_t fl = blk(f, l), lf = blk(l, f);
        if (fl < 0 || lf < 0 || !free_link(f, l)) continue;
        c.cl_chain.push_back((int32_t)ch);
        c.cl_fl.push_back(fl);
        c.cl_lf.push_back(lf);
        c.cl_mfl.push_back(mask(f, l));
        c.cl_mlf.push_back(mask(l, f));
        c.cl_zoff.push_back(c.cl_zoff.back() + 72 * n);
        c.cl_max_len = std::max(c.cl_max_len, n);
    }
}

// Separable solve of  c₀ = c₁, c_{i-1}/8 + 3c_i/4 + c_{i+1}/8 = y_i, c_{n+1} = c_n  along each axis
// (the Reflect(OnCell()) quadratic B-spline system, src/interactions/def_interactions.jl:77).
// The tridiagonal factor is computed once per axis length and reused for every line.
void tri_factors(int64_t n, std::vector<double>& cp, std::vector<double>& inv) {
    cp.assign((size_t)n, 0.0);
    inv.assign((size_t)n, 0.0);
    for (int64_t i = 0; i < n; ++i) {
        double a = i == 0 ? 0.0 : 0.125;
        double b = (i == 0 || i == n - 1) ? 0.875 : 0.75;
        if (n == 1) b = 1.0;
        double c = i == n - 1 ? 0.0 : 0.125;
        double den = b - (i > 0 ? a * cp[i - 1] : 0.0);
        inv[i] = 1.0 / den;
        cp[i] = c * inv[i];
    }
}

namespace {
struct Tri {
    int64_t n;
    std::vector<double> cp, inv;  // modified super-diagonal and reciprocal pivots
    explicit Tri(int64_t n_) : n(n_) { tri_factors(n, cp, inv); }
    // y (stride sy, n values) -> padded c (stride sc, n+2 values)
    void solve(const double* y, int64_t sy, double* c, int64_t sc, double* tmp) const {
        for (int64_t i = 0; i < n; ++i) tmp[i] = (y[i * sy] - (i > 0 ? 0.125 * tmp[i - 1] : 0.0)) * inv[i];
        c[n * sc] = tmp[n - 1];
        for (int64_t i = n - 2; i >= 0; --i) c[(i + 1) * sc] = tmp[i] - cp[i] * c[(i + 2) * sc];
        c[0] = c[sc];
        c[(n + 1) * sc] = c[n * sc];
    }
};
}  // namespace

void prefilter_sdf(const int64_t n[3], const double* samples, std::vector<double>& coef) {
    const int64_t P0 = n[0] + 2, P1 = n[1] + 2, P2 = n[2] + 2;
    coef.assign((size_t)(P0 * P1 * P2), 0.0);
    Tri tx(n[0]), ty(n[1]), tz(n[2]);
    std::vector<double> tmp(std::max({n[0], n[1], n[2]}) + 2), line(std::max({n[0], n[1], n[2]}) + 2);
    for (int64_t k = 0; k < n[2]; ++k)
        for (int64_t j = 0; j < n[1]; ++j)
            tx.solve(samples + (k * n[1] + j) * n[0], 1, &coef[(size_t)(((k + 1) * P1 + (j + 1)) * P0)], 1, tmp.data());
    for (int64_t k = 0; k < n[2]; ++k)
        for (int64_t i = 0; i < P0; ++i) {
            double* base = &coef[(size_t)((k + 1) * P1 * P0 + i)];
            for (int64_t j = 0; j < n[1]; ++j) line[j] = base[(j + 1) * P0];
            ty.solve(line.data(), 1, base, P0, tmp.data());
        }
    for (int64_t j = 0; j < P1; ++j)
        for (int64_t i = 0; i < P0; ++i) {
            double* base = &coef[(size_t)(j * P0 + i)];
            for (int64_t k = 0; k < n[2]; ++k) line[k] = base[(k + 1) * P1 * P0];
            tz.solve(line.data(), 1, base, P1 * P0, tmp.data());
        }
}

}  // namespace ebh
