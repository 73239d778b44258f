// eb_oracle.cpp — TEST INFRASTRUCTURE ONLY.
//
// CPU FP64 restatement of EndoBeams.jl's per-Newton-iteration path, used as the parity oracle
// for libendobeams_hip.so and as the timed CPU baseline. Only tests/, __graft_entry__.smoke()
// and bench.py's cpu_baseline leg load it, and only as the checker; the product never links it.
//
// It exports the product's C ABI (include/endobeams_hip.h) under the prefix `ebo_` with the same
// semantics, except that ebo_solve_free_dofs is a direct solve (banded LU with partial pivoting
// after reverse Cuthill-McKee), standing in for MKL Pardiso (src/linearsolvers.jl:2-28).
//
// Reference files followed (paths relative to the EndoBeams.jl root):
//   src/beams/assemble_beams.jl:1-70             ebo_assemble
//   src/interactions/compute&assemble_interactions_beams.jl:447-500   ebo_assemble_contact
//   src/sparse_matrices.jl:6-104                 build_pattern
//   src/utils_solver.jl:6-304                    HHT combine, extraction, updates, norms
//   src/utils_bcs_and_loads.jl:5-64              BC lift, loads, imposed displacements
//   src/corrector.jl, src/predictor.jl, src/solve_step_dynamics.jl, src/inizialization.jl:37-45
// Element kernels: eb_elements.h. Built with -ffp-contract=off.
#include "../include/endobeams_hip.h"
#include "eb_elements.h"

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <deque>
#include <new>
#include <string>
#include <vector>

using namespace ebo;

namespace {

struct BandLU {
    int64_t n = 0, kl = 0, ku = 0, ldab = 0;
    std::vector<int64_t> perm, iperm;  // perm[new] = old
    std::vector<double> ab;
    std::vector<int64_t> ipiv;
};

struct Oracle {
    int64_t N = 0, E = 0, ndofs = 0, nnz = 0, n_free = 0, n_fixed = 0, nnz_free = 0;
    std::vector<double> X0, l0, Re0, props;
    std::vector<int64_t> n1, n2;  // 0-based
    std::vector<double> node[16];
    std::vector<double> fext, Tint, Tk, Tc, fext_n, Tint_n, Tk_n, Tc_n;
    std::vector<double> D, Dd, Ddd, res, dD, temp;
    std::vector<double> K, C, M, Kn, Cn, Mn, Ktan, r_free, Ktan_free, dD_free, imposed;
    double energy[3] = {0, 0, 0};
    std::vector<int64_t> colptr, rowval, smap, sfree, fixed, freed;  // 0-based
    std::vector<char> is_fixed;
    std::vector<int64_t> fcolptr, frowval;  // Ktan_free pattern
    Gauss g{};
    Surface surf;
    ContactProps cp{};
    std::vector<int64_t> contact_beams;
    BandLU lu;
    bool lu_ready = false;
    int64_t counters[4] = {0, 0, 0, 0};
    int last_k = 0;
    double last_rn = 0, last_dn = 0;
    std::string err;
};

int fail(Oracle* o, int code, const std::string& msg) {
    if (o) o->err = msg;
    return code;
}

std::vector<double>* field_ptr(Oracle* o, int f) {
    if (f >= EB_F_U && f <= EB_F_DR_N) return &o->node[f];
    switch (f) {
        case EB_F_FEXT: return &o->fext;
        case EB_F_TINT: return &o->Tint;
        case EB_F_TK: return &o->Tk;
        case EB_F_TC: return &o->Tc;
        case EB_F_FEXT_N: return &o->fext_n;
        case EB_F_TINT_N: return &o->Tint_n;
        case EB_F_TK_N: return &o->Tk_n;
        case EB_F_TC_N: return &o->Tc_n;
        case EB_F_D: return &o->D;
        case EB_F_DDOT: return &o->Dd;
        case EB_F_DDDOT: return &o->Ddd;
        case EB_F_RES: return &o->res;
        case EB_F_DELTA_D: return &o->dD;
        case EB_F_K: return &o->K;
        case EB_F_C: return &o->C;
        case EB_F_M: return &o->M;
        case EB_F_K_N: return &o->Kn;
        case EB_F_C_N: return &o->Cn;
        case EB_F_M_N: return &o->Mn;
        case EB_F_KTAN: return &o->Ktan;
        case EB_F_RES_FREE: return &o->r_free;
        case EB_F_KTAN_FREE: return &o->Ktan_free;
        case EB_F_DELTA_D_FREE: return &o->dD_free;
        case EB_F_IMPOSED: return &o->imposed;
        default: return nullptr;
    }
}

inline V3 getv(const std::vector<double>& a, int64_t i) { return v3(a[3 * i], a[3 * i + 1], a[3 * i + 2]); }
inline void setv(std::vector<double>& a, int64_t i, const V3& v) { a[3 * i] = v[0]; a[3 * i + 1] = v[1]; a[3 * i + 2] = v[2]; }
inline M3 getm(const std::vector<double>& a, int64_t i) { M3 m; for (int k = 0; k < 9; ++k) m.m[k] = a[9 * i + k]; return m; }
inline void setm(std::vector<double>& a, int64_t i, const M3& m) { for (int k = 0; k < 9; ++k) a[9 * i + k] = m.m[k]; }

// sparsity_beams + sparse_matrices_beams!, src/sparse_matrices.jl:6-104
int build_pattern(Oracle* o) {
    std::vector<std::vector<int64_t>> adj(o->N);
    for (int64_t n = 0; n < o->N; ++n) adj[n].push_back(n);
    for (int64_t e = 0; e < o->E; ++e) { adj[o->n1[e]].push_back(o->n2[e]); adj[o->n2[e]].push_back(o->n1[e]); }
    for (auto& a : adj) { std::sort(a.begin(), a.end()); a.erase(std::unique(a.begin(), a.end()), a.end()); }
    for (int64_t n = 0; n < o->N; ++n)
        if (adj[n].size() < 2) return fail(o, EB_ERR_ARG, "node " + std::to_string(n + 1) + " belongs to no beam");
    o->colptr.assign(o->ndofs + 1, 0);
    for (int64_t j = 0; j < o->ndofs; ++j) o->colptr[j + 1] = o->colptr[j] + 6 * (int64_t)adj[j / 6].size();
    o->nnz = o->colptr[o->ndofs];
    o->rowval.resize(o->nnz);
    for (int64_t j = 0; j < o->ndofs; ++j) {
        int64_t p = o->colptr[j];
        for (int64_t m : adj[j / 6])
            for (int r = 0; r < 6; ++r) o->rowval[p++] = 6 * m + r;
    }
    auto nzidx = [&](int64_t row, int64_t colj) -> int64_t {
        const auto& a = adj[colj / 6];
        int64_t pos = std::lower_bound(a.begin(), a.end(), row / 6) - a.begin();
        return o->colptr[colj] + 6 * pos + row % 6;
    };
    o->smap.resize(144 * o->E);
    for (int64_t e = 0; e < o->E; ++e) {
        int64_t dofs[12];
        for (int k = 0; k < 6; ++k) { dofs[k] = 6 * o->n1[e] + k; dofs[6 + k] = 6 * o->n2[e] + k; }
        for (int jl = 0; jl < 12; ++jl)
            for (int il = 0; il < 12; ++il) o->smap[144 * e + 12 * jl + il] = nzidx(dofs[il], dofs[jl]);
    }
    o->sfree.clear();
    o->fcolptr.assign(1, 0);
    for (int64_t j = 0; j < o->ndofs; ++j) {
        if (!o->is_fixed[j]) {
            for (int64_t p = o->colptr[j]; p < o->colptr[j + 1]; ++p)
                if (!o->is_fixed[o->rowval[p]]) o->sfree.push_back(p);
            o->fcolptr.push_back((int64_t)o->sfree.size());
        }
    }
    o->nnz_free = (int64_t)o->sfree.size();
    std::vector<int64_t> newidx(o->ndofs, -1);
    for (int64_t k = 0; k < o->n_free; ++k) newidx[o->freed[k]] = k;
    o->frowval.resize(o->nnz_free);
    for (int64_t k = 0; k < o->nnz_free; ++k) o->frowval[k] = newidx[o->rowval[o->sfree[k]]];
    return EB_OK;
}

ElemIn gather_elem(const Oracle* o, int64_t e) {
    ElemIn in;
    int64_t a = o->n1[e], b = o->n2[e];
    in.X1 = getv(o->X0, a); in.X2 = getv(o->X0, b);
    in.u1 = getv(o->node[EB_F_U], a); in.u2 = getv(o->node[EB_F_U], b);
    in.ud1 = getv(o->node[EB_F_UDOT], a); in.ud2 = getv(o->node[EB_F_UDOT], b);
    in.udd1 = getv(o->node[EB_F_UDDOT], a); in.udd2 = getv(o->node[EB_F_UDDOT], b);
    in.wd1 = getv(o->node[EB_F_WDOT], a); in.wd2 = getv(o->node[EB_F_WDOT], b);
    in.wdd1 = getv(o->node[EB_F_WDDOT], a); in.wdd2 = getv(o->node[EB_F_WDDOT], b);
    in.R1 = getm(o->node[EB_F_R], a); in.R2 = getm(o->node[EB_F_R], b);
    in.dR1 = getm(o->node[EB_F_DR], a); in.dR2 = getm(o->node[EB_F_DR], b);
    in.l0 = o->l0[e];
    for (int k = 0; k < 9; ++k) in.Re0.m[k] = o->Re0[9 * e + k];
    return in;
}

// ---- banded LU with partial pivoting (LAPACK dgbtf2 algorithm) after RCM -------------------

void rcm(int64_t n, const std::vector<std::vector<int64_t>>& adj, std::vector<int64_t>& perm) {
    perm.clear();
    std::vector<char> seen(n, 0);
    std::vector<int64_t> order(n);
    for (int64_t i = 0; i < n; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int64_t a, int64_t b) { return adj[a].size() < adj[b].size(); });
    for (int64_t s0 : order) {
        if (seen[s0]) continue;
        // pseudo-peripheral start: last node of a BFS from s0
        int64_t start = s0;
        for (int it = 0; it < 2; ++it) {
            std::vector<int64_t> q{start};
            std::vector<char> vis(n, 0);
            vis[start] = 1;
            for (size_t h = 0; h < q.size(); ++h)
                for (int64_t m : adj[q[h]]) if (!vis[m] && !seen[m]) { vis[m] = 1; q.push_back(m); }
            start = q.back();
        }
        size_t base = perm.size();
        perm.push_back(start);
        seen[start] = 1;
        for (size_t h = base; h < perm.size(); ++h) {
            std::vector<int64_t> nb;
            for (int64_t m : adj[perm[h]]) if (!seen[m]) { seen[m] = 1; nb.push_back(m); }
            std::stable_sort(nb.begin(), nb.end(), [&](int64_t a, int64_t b) { return adj[a].size() < adj[b].size(); });
            for (int64_t m : nb) perm.push_back(m);
        }
        std::reverse(perm.begin() + base, perm.end());
    }
}

void band_setup(Oracle* o) {
    BandLU& L = o->lu;
    int64_t n = o->n_free;
    L.n = n;
    std::vector<std::vector<int64_t>> adj(n);
    for (int64_t j = 0; j < n; ++j)
        for (int64_t p = o->fcolptr[j]; p < o->fcolptr[j + 1]; ++p) {
            int64_t i = o->frowval[p];
            if (i != j) { adj[i].push_back(j); adj[j].push_back(i); }
        }
    for (auto& a : adj) { std::sort(a.begin(), a.end()); a.erase(std::unique(a.begin(), a.end()), a.end()); }
    rcm(n, adj, L.perm);
    L.iperm.assign(n, 0);
    for (int64_t k = 0; k < n; ++k) L.iperm[L.perm[k]] = k;
    int64_t kl = 0, ku = 0;
    for (int64_t j = 0; j < n; ++j)
        for (int64_t p = o->fcolptr[j]; p < o->fcolptr[j + 1]; ++p) {
            int64_t i = L.iperm[o->frowval[p]], jj = L.iperm[j];
            if (i > jj) kl = std::max(kl, i - jj); else ku = std::max(ku, jj - i);
        }
    L.kl = kl; L.ku = ku; L.ldab = 2 * kl + ku + 1;
    L.ab.assign((size_t)(L.ldab * std::max<int64_t>(n, 1)), 0.0);
    L.ipiv.assign(n, 0);
    o->lu_ready = true;
}

// A(i,j) lives at ab[kl+ku+i-j + j*ldab]
int band_factor_solve(Oracle* o, const double* vals, const double* b, double* x) {
    BandLU& L = o->lu;
    int64_t n = L.n, kl = L.kl, ku = L.ku, ld = L.ldab, kv = kl + ku;
    std::fill(L.ab.begin(), L.ab.end(), 0.0);
    for (int64_t j = 0; j < n; ++j)
        for (int64_t p = o->fcolptr[j]; p < o->fcolptr[j + 1]; ++p) {
            int64_t i = L.iperm[o->frowval[p]], jj = L.iperm[j];
            L.ab[kv + i - jj + jj * ld] += vals[p];
        }
    auto A = [&](int64_t i, int64_t j) -> double& { return L.ab[kv + i - j + j * ld]; };
    int64_t ju = 0;
    for (int64_t j = 0; j < n; ++j) {
        int64_t km = std::min(kl, n - 1 - j);
        int64_t jp = 0;
        double best = std::fabs(A(j, j));
        for (int64_t i = 1; i <= km; ++i) if (std::fabs(A(j + i, j)) > best) { best = std::fabs(A(j + i, j)); jp = i; }
        L.ipiv[j] = j + jp;
        if (A(j + jp, j) == 0.0) return fail(o, EB_ERR_SOLVE, "singular tangent matrix (zero pivot)");
        ju = std::max(ju, std::min(j + ku + jp, n - 1));
        if (jp != 0)
            for (int64_t c = j; c <= ju; ++c) std::swap(A(j, c), A(j + jp, c));
        double piv = A(j, j);
        for (int64_t i = 1; i <= km; ++i) A(j + i, j) /= piv;
        for (int64_t c = j + 1; c <= ju; ++c) {
            double ajc = A(j, c);
            if (ajc != 0.0)
                for (int64_t i = 1; i <= km; ++i) A(j + i, c) -= A(j + i, j) * ajc;
        }
    }
    std::vector<double> y(n);
    for (int64_t k = 0; k < n; ++k) y[k] = b[L.perm[k]];
    for (int64_t j = 0; j < n; ++j) {
        int64_t km = std::min(kl, n - 1 - j);
        if (L.ipiv[j] != j) std::swap(y[j], y[L.ipiv[j]]);
        double yj = y[j];
        for (int64_t i = 1; i <= km; ++i) y[j + i] -= A(j + i, j) * yj;
    }
    for (int64_t j = n - 1; j >= 0; --j) {
        y[j] /= A(j, j);
        double yj = y[j];
        int64_t i0 = std::max<int64_t>(0, j - kl - ku);
        for (int64_t i = i0; i < j; ++i) y[i] -= A(i, j) * yj;
    }
    for (int64_t k = 0; k < n; ++k) x[L.perm[k]] = y[k];
    return EB_OK;
}

}  // namespace

extern "C" {

int ebo_create(const eb_model_desc* d, void** out) {
    if (!d || !out) return EB_ERR_ARG;
    if (d->abi_version != EB_ABI_VERSION) return EB_ERR_ARG;
    Oracle* o = new (std::nothrow) Oracle();
    if (!o) return EB_ERR_ARG;
    o->N = d->n_nodes; o->E = d->n_beams; o->ndofs = 6 * o->N;
    if (o->N <= 0 || o->E <= 0 || !d->X0 || !d->conn || !d->l0 || !d->Re0 || !d->props) { delete o; return EB_ERR_ARG; }
    o->X0.assign(d->X0, d->X0 + 3 * o->N);
    o->l0.assign(d->l0, d->l0 + o->E);
    o->Re0.assign(d->Re0, d->Re0 + 9 * o->E);
    o->props.assign(d->props, d->props + EB_NPROPS * o->E);
    o->n1.resize(o->E); o->n2.resize(o->E);
    for (int64_t e = 0; e < o->E; ++e) {
        o->n1[e] = d->conn[2 * e] - 1; o->n2[e] = d->conn[2 * e + 1] - 1;
        if (o->n1[e] < 0 || o->n1[e] >= o->N || o->n2[e] < 0 || o->n2[e] >= o->N || o->n1[e] == o->n2[e]) { delete o; return EB_ERR_ARG; }
    }
    o->is_fixed.assign(o->ndofs, 0);
    for (int64_t k = 0; k < d->n_fixed; ++k) {
        int64_t f = d->fixed_dofs[k] - 1;
        if (f < 0 || f >= o->ndofs || (k > 0 && d->fixed_dofs[k] <= d->fixed_dofs[k - 1])) { delete o; return EB_ERR_ARG; }
        o->is_fixed[f] = 1; o->fixed.push_back(f);
    }
    o->n_fixed = d->n_fixed;
    for (int64_t j = 0; j < o->ndofs; ++j) if (!o->is_fixed[j]) o->freed.push_back(j);
    o->n_free = (int64_t)o->freed.size();
    if (d->n_gauss <= 0 || d->n_gauss > 8) { delete o; return EB_ERR_ARG; }
    o->g.n = d->n_gauss;
    for (int k = 0; k < d->n_gauss; ++k) { o->g.w[k] = d->gauss_w[k]; o->g.z[k] = d->gauss_z[k]; }
    o->surf.kind = d->surface;
    o->cp = ContactProps{d->kn, d->mu, d->eps_t, d->eta_n};
    if (d->surface != EB_SURFACE_NONE) {
        for (int64_t k = 0; k < d->n_contact_beams; ++k) o->contact_beams.push_back(d->contact_beams[k] - 1);
        if (d->surface == EB_SURFACE_SPHERE) {
            o->surf.center = v3(d->sphere_center[0], d->sphere_center[1], d->sphere_center[2]);
            o->surf.R = d->sphere_radius;
        } else if (d->surface == EB_SURFACE_SDF) {
            if (!d->sdf_samples || d->sdf_n[0] < 1 || d->sdf_n[1] < 1 || d->sdf_n[2] < 1) { delete o; return EB_ERR_ARG; }
            sdf_build(o->surf.sdf, d->sdf_n, d->sdf_d, d->sdf_x0, d->sdf_samples);
            if (d->sdf_prefiltered)
                o->surf.sdf.coef.assign(d->sdf_samples, d->sdf_samples + o->surf.sdf.coef.size());
        } else { delete o; return EB_ERR_ARG; }
    }
    if (build_pattern(o) != EB_OK) { delete o; return EB_ERR_ARG; }
    for (int f = EB_F_U; f <= EB_F_DR_N; ++f) {
        bool mat = (f == EB_F_R || f == EB_F_DR || f == EB_F_R_N || f == EB_F_DR_N);
        o->node[f].assign((mat ? 9 : 3) * o->N, 0.0);
        if (mat) for (int64_t n = 0; n < o->N; ++n) setm(o->node[f], n, eye());
    }
    for (auto* v : {&o->fext, &o->Tint, &o->Tk, &o->Tc, &o->fext_n, &o->Tint_n, &o->Tk_n, &o->Tc_n, &o->D, &o->Dd,
                    &o->Ddd, &o->res, &o->dD, &o->temp})
        v->assign(o->ndofs, 0.0);
    for (auto* v : {&o->K, &o->C, &o->M, &o->Kn, &o->Cn, &o->Mn, &o->Ktan}) v->assign(o->nnz, 0.0);
    o->r_free.assign(o->n_free, 0.0); o->dD_free.assign(o->n_free, 0.0);
    o->Ktan_free.assign(o->nnz_free, 0.0);
    o->imposed.assign(o->n_fixed, 0.0);
    *out = o;
    return EB_OK;
}

int ebo_destroy(void* h) { delete (Oracle*)h; return EB_OK; }
const char* ebo_last_error(void* h) { return h ? ((Oracle*)h)->err.c_str() : "null context"; }

int ebo_get_sizes(void* h, eb_sizes* s) {
    Oracle* o = (Oracle*)h;
    s->n_nodes = o->N; s->n_beams = o->E; s->ndofs = o->ndofs; s->nnz = o->nnz;
    s->n_free = o->n_free; s->n_fixed = o->n_fixed; s->nnz_free = o->nnz_free;
    return EB_OK;
}

int ebo_get_pattern(void* h, int64_t* colptr, int64_t* rowval, int64_t* smap, int64_t* sfree) {
    Oracle* o = (Oracle*)h;
    if (colptr) for (int64_t j = 0; j <= o->ndofs; ++j) colptr[j] = o->colptr[j] + 1;
    if (rowval) for (int64_t k = 0; k < o->nnz; ++k) rowval[k] = o->rowval[k] + 1;
    if (smap) for (int64_t k = 0; k < 144 * o->E; ++k) smap[k] = o->smap[k] + 1;
    if (sfree) for (int64_t k = 0; k < o->nnz_free; ++k) sfree[k] = o->sfree[k] + 1;
    return EB_OK;
}

int ebo_get_field(void* h, int f, double* out, int64_t count) {
    Oracle* o = (Oracle*)h;
    if (f == EB_F_ENERGY) { if (count != 3) return fail(o, EB_ERR_ARG, "energy has 3 entries"); std::memcpy(out, o->energy, 24); return EB_OK; }
    auto* v = field_ptr(o, f);
    if (!v) return fail(o, EB_ERR_ARG, "unknown field");
    if ((int64_t)v->size() != count) return fail(o, EB_ERR_ARG, "field size mismatch");
    if (count) std::memcpy(out, v->data(), sizeof(double) * count);
    return EB_OK;
}

int ebo_set_field(void* h, int f, const double* in, int64_t count) {
    Oracle* o = (Oracle*)h;
    if (f == EB_F_ENERGY) { if (count != 3) return fail(o, EB_ERR_ARG, "energy has 3 entries"); std::memcpy(o->energy, in, 24); return EB_OK; }
    auto* v = field_ptr(o, f);
    if (!v) return fail(o, EB_ERR_ARG, "unknown field");
    if ((int64_t)v->size() != count) return fail(o, EB_ERR_ARG, "field size mismatch");
    if (count) std::memcpy(v->data(), in, sizeof(double) * count);
    return EB_OK;
}

// assemble!, src/beams/assemble_beams.jl:1-70 (serial loop: the scatter order is beam order)
// The element computations of a chunk of beams run in parallel (OpenMP in the timing build,
// liboracle_omp.so), then their contributions are added in beam order, as the serial loop of
// assemble! adds them: the results do not depend on the thread count.
constexpr int64_t ORACLE_CHUNK = 1024;

int ebo_assemble(void* h) {
    Oracle* o = (Oracle*)h;
    std::fill(o->K.begin(), o->K.end(), 0.0);
    std::fill(o->C.begin(), o->C.end(), 0.0);
    std::fill(o->M.begin(), o->M.end(), 0.0);
    std::fill(o->Tint.begin(), o->Tint.end(), 0.0);
    std::fill(o->Tk.begin(), o->Tk.end(), 0.0);
    std::fill(o->Tc.begin(), o->Tc.end(), 0.0);
    std::vector<BeamOut> outs((size_t)std::min<int64_t>(ORACLE_CHUNK, o->E));
    for (int64_t e0 = 0; e0 < o->E; e0 += ORACLE_CHUNK) {
        const int64_t n = std::min<int64_t>(ORACLE_CHUNK, o->E - e0);
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; ++i) {
            const int64_t e = e0 + i;
            ElemIn in = gather_elem(o, e);
            BeamProps bp = props_from(&o->props[EB_NPROPS * e]);
            compute_beams(in, bp, o->g, outs[i]);
        }
        for (int64_t i = 0; i < n; ++i) {
            const int64_t e = e0 + i;
            const BeamOut& out = outs[i];
            o->energy[0] += out.strain;
            o->energy[1] += out.kinetic;
            int64_t dofs[12];
            for (int k = 0; k < 6; ++k) { dofs[k] = 6 * o->n1[e] + k; dofs[6 + k] = 6 * o->n2[e] + k; }
            for (int k = 0; k < 12; ++k) o->Tk[dofs[k]] += out.Tk[k];
            for (int k = 0; k < 12; ++k) o->Tint[dofs[k]] += out.Tint[k];
            const int64_t* mp = &o->smap[144 * e];
            for (int k = 0; k < 144; ++k) o->K[mp[k]] += out.K.a[k];
            for (int k = 0; k < 144; ++k) o->C[mp[k]] += out.C.a[k];
            for (int k = 0; k < 144; ++k) o->M[mp[k]] += out.M.a[k];
        }
    }
    o->counters[3]++;
    return EB_OK;
}

// assemble_contact!, src/interactions/compute&assemble_interactions_beams.jl:447-500
int ebo_assemble_contact(void* h, double alpha) {
    Oracle* o = (Oracle*)h;
    if (o->surf.kind == EB_SURFACE_NONE) return EB_OK;
    const int64_t nc = (int64_t)o->contact_beams.size();
    std::vector<ContactOut> outs((size_t)std::min<int64_t>(ORACLE_CHUNK, nc));
    for (int64_t c0 = 0; c0 < nc; c0 += ORACLE_CHUNK) {
        const int64_t n = std::min<int64_t>(ORACLE_CHUNK, nc - c0);
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; ++i) {
            const int64_t e = o->contact_beams[c0 + i];
            ElemIn in = gather_elem(o, e);
            double rb = o->props[EB_NPROPS * e + 0];
            compute_contact_beams(in, rb, o->cp, o->surf, o->g, outs[i]);
        }
        for (int64_t i = 0; i < n; ++i) {
            const int64_t e = o->contact_beams[c0 + i];
            const ContactOut& out = outs[i];
            o->energy[2] += out.energy;
            int64_t dofs[12];
            for (int k = 0; k < 6; ++k) { dofs[k] = 6 * o->n1[e] + k; dofs[6 + k] = 6 * o->n2[e] + k; }
            for (int k = 0; k < 12; ++k) o->Tc[dofs[k]] += out.Tc[k];
            const int64_t* mp = &o->smap[144 * e];
            for (int k = 0; k < 144; ++k) o->K[mp[k]] += -out.K.a[k];
            for (int k = 0; k < 144; ++k) o->C[mp[k]] += -((1 + alpha) * out.C.a[k]);
        }
    }
    return EB_OK;
}

// compute_tangent_and_residuals_corrector!, src/utils_solver.jl:80-89
int ebo_tangent_residuals_corrector(void* h, double dt, double a, double b, double g) {
    Oracle* o = (Oracle*)h;
    double c1 = 1 + a, c2 = g / (b * dt), c3 = 1 / (b * (dt * dt));
    for (int64_t k = 0; k < o->nnz; ++k) o->Ktan[k] = (c1 * o->K[k] + c2 * o->C[k]) + c3 * o->M[k];
    for (int64_t i = 0; i < o->ndofs; ++i)
        o->res[i] = ((1 + a) * ((o->fext[i] + o->Tc[i]) - o->Tint[i]) - a * ((o->fext_n[i] + o->Tc_n[i]) - o->Tint_n[i])) - o->Tk[i];
    return EB_OK;
}

// compute_tangent_and_residuals_predictor!, src/utils_solver.jl:54-74
int ebo_tangent_residuals_predictor(void* h, double dt, double a, double b, double g) {
    Oracle* o = (Oracle*)h;
    double c1 = 1 + a, c2 = g / (b * dt), c3 = 1 / (b * (dt * dt));
    for (int64_t k = 0; k < o->nnz; ++k) o->Ktan[k] = (c1 * o->Kn[k] + c2 * o->Cn[k]) + c3 * o->Mn[k];
    for (int64_t i = 0; i < o->ndofs; ++i)
        o->res[i] = (((1 + a) * o->fext[i] - o->Tint_n[i]) - o->Tk_n[i]) - a * o->fext_n[i];
    // r += C·temp, temp = (γ/β)Ḋ − (Δt/2·(2β−γ)/β)D̈   (mul!(r, C, temp, 1, 1): column sweep,
    // r[row] += nzval·(1·temp[j]))
    double t1 = g / b, t2 = ((dt / 2) * (2 * b - g)) / b;
    for (int64_t i = 0; i < o->ndofs; ++i) o->temp[i] = t1 * o->Dd[i] - t2 * o->Ddd[i];
    for (int64_t j = 0; j < o->ndofs; ++j)
        for (int64_t p = o->colptr[j]; p < o->colptr[j + 1]; ++p) o->res[o->rowval[p]] += o->Cn[p] * o->temp[j];
    // r += (1/(βΔt²))·M·temp, temp = ΔtḊ + (Δt²/2)D̈   (mul!(r, M, temp, 1/(βΔt²), 1))
    for (int64_t i = 0; i < o->ndofs; ++i) o->temp[i] = dt * o->Dd[i] + ((dt * dt) / 2) * o->Ddd[i];
    for (int64_t j = 0; j < o->ndofs; ++j) {
        double ax = c3 * o->temp[j];
        for (int64_t p = o->colptr[j]; p < o->colptr[j + 1]; ++p) o->res[o->rowval[p]] += o->Mn[p] * ax;
    }
    return EB_OK;
}

// apply_boundary_conditions!, src/utils_bcs_and_loads.jl:5-29
int ebo_apply_boundary_conditions(void* h) {
    Oracle* o = (Oracle*)h;
    for (int64_t k = 0; k < o->n_fixed; ++k) {
        int64_t dof = o->fixed[k];
        double ch = o->imposed[k] - o->D[dof];
        o->dD[dof] = ch;
        for (int64_t p = o->colptr[dof]; p < o->colptr[dof + 1]; ++p) o->res[o->rowval[p]] -= ch * o->Ktan[p];
    }
    return EB_OK;
}

// extract_free_dofs!, src/utils_solver.jl:94-102
int ebo_extract_free_dofs(void* h) {
    Oracle* o = (Oracle*)h;
    for (int64_t k = 0; k < o->n_free; ++k) o->r_free[k] = o->res[o->freed[k]];
    for (int64_t k = 0; k < o->nnz_free; ++k) o->Ktan_free[k] = o->Ktan[o->sfree[k]];
    return EB_OK;
}

// solve_free_dofs!, src/utils_solver.jl:107-120 (direct solve in place of Pardiso)
int ebo_solve_free_dofs(void* h, const eb_newton_params*, int* its) {
    Oracle* o = (Oracle*)h;
    if (!o->lu_ready) band_setup(o);
    int st = band_factor_solve(o, o->Ktan_free.data(), o->r_free.data(), o->dD_free.data());
    if (st != EB_OK) return st;
    for (int64_t k = 0; k < o->n_free; ++k) o->dD[o->freed[k]] = o->dD_free[k];
    if (its) *its = 1;
    o->counters[2]++;
    return EB_OK;
}

int ebo_band_info(void* h, int64_t* kl, int64_t* ku) {
    Oracle* o = (Oracle*)h;
    if (!o->lu_ready) band_setup(o);
    *kl = o->lu.kl; *ku = o->lu.ku;
    return EB_OK;
}

// initialize_global_corrector!, src/utils_solver.jl:6-24
int ebo_initialize_global_corrector(void* h) {
    Oracle* o = (Oracle*)h;
    for (int64_t n = 0; n < o->N; ++n)
        for (int k = 0; k < 3; ++k) {
            o->dD[6 * n + k] = 0; o->D[6 * n + k] = o->node[EB_F_U][3 * n + k];
            o->Dd[6 * n + k] = o->node[EB_F_UDOT][3 * n + k]; o->Ddd[6 * n + k] = o->node[EB_F_UDDOT][3 * n + k];
            o->dD[6 * n + 3 + k] = 0; o->D[6 * n + 3 + k] = o->node[EB_F_W][3 * n + k];
            o->Dd[6 * n + 3 + k] = o->node[EB_F_WDOT][3 * n + k]; o->Ddd[6 * n + 3 + k] = o->node[EB_F_WDDOT][3 * n + k];
        }
    return EB_OK;
}

// initialize_global_predictor!, src/utils_solver.jl:30-48
int ebo_initialize_global_predictor(void* h) {
    Oracle* o = (Oracle*)h;
    for (int64_t n = 0; n < o->N; ++n)
        for (int k = 0; k < 3; ++k) {
            o->dD[6 * n + k] = 0; o->D[6 * n + k] = o->node[EB_F_U][3 * n + k];
            o->Dd[6 * n + k] = o->node[EB_F_UDOT][3 * n + k]; o->Ddd[6 * n + k] = o->node[EB_F_UDDOT][3 * n + k];
            o->dD[6 * n + 3 + k] = 0; o->D[6 * n + 3 + k] = 0;
            o->Dd[6 * n + 3 + k] = o->node[EB_F_WDOT][3 * n + k]; o->Ddd[6 * n + 3 + k] = o->node[EB_F_WDDOT][3 * n + k];
        }
    return EB_OK;
}

// compute_global_corrector!, src/utils_solver.jl:178-190
int ebo_compute_global_corrector(void* h, double dt, double b, double g) {
    Oracle* o = (Oracle*)h;
    double inv = 1 / (b * (dt * dt)), gb = g / (b * dt);
    for (int64_t i = 0; i < o->ndofs; ++i) {
        o->D[i] += o->dD[i];
        o->Dd[i] += gb * o->dD[i];
        o->Ddd[i] += inv * o->dD[i];
    }
    return EB_OK;
}

// compute_global_predictor!, src/utils_solver.jl:157-172
int ebo_compute_global_predictor(void* h, double dt, double b, double g) {
    Oracle* o = (Oracle*)h;
    double gob = g / b, gbdt = g / (b * dt), inv = 1 / (b * (dt * dt)), dt2b = dt * (2 * b - g) / (2 * b);
    for (int64_t i = 0; i < o->ndofs; ++i) {
        double Ddn = o->Dd[i];
        o->D[i] += o->dD[i];
        o->Dd[i] += (gbdt * o->dD[i] - gob * Ddn) + dt2b * o->Ddd[i];
        o->Ddd[i] += inv * ((o->dD[i] - dt * Ddn) - ((dt * dt) / 2) * o->Ddd[i]);
    }
    return EB_OK;
}

// compute_local_predictor!, src/utils_solver.jl:197-217
int ebo_compute_local_predictor(void* h, double dt, double b, double g) {
    Oracle* o = (Oracle*)h;
    for (int64_t n = 0; n < o->N; ++n) {
        setv(o->node[EB_F_U], n, v3(o->D[6 * n], o->D[6 * n + 1], o->D[6 * n + 2]));
        setv(o->node[EB_F_UDOT], n, v3(o->Dd[6 * n], o->Dd[6 * n + 1], o->Dd[6 * n + 2]));
        setv(o->node[EB_F_UDDOT], n, v3(o->Ddd[6 * n], o->Ddd[6 * n + 1], o->Ddd[6 * n + 2]));
        V3 th = v3(o->D[6 * n + 3], o->D[6 * n + 4], o->D[6 * n + 5]);
        setv(o->node[EB_F_W], n, th);
        M3 dR = rotation_matrix(th);
        setm(o->node[EB_F_DR], n, dR);
        setm(o->node[EB_F_R], n, dR * getm(o->node[EB_F_R_N], n));
        V3 wdn = getv(o->node[EB_F_WDOT_N], n), wddn = getv(o->node[EB_F_WDDOT_N], n);
        V3 a = ((g / (b * dt)) * th + ((b - g) / b) * wdn) + ((dt * (b - g / 2)) / b) * wddn;
        V3 c = ((1 / (b * (dt * dt))) * th - (1 / (b * dt)) * wdn) - (1 / (2 * b) - 1) * wddn;
        setv(o->node[EB_F_WDOT], n, dR * a);
        setv(o->node[EB_F_WDDOT], n, dR * c);
    }
    return EB_OK;
}

// compute_local_corrector!, src/utils_solver.jl:224-245
int ebo_compute_local_corrector(void* h, double dt, double b, double g) {
    Oracle* o = (Oracle*)h;
    for (int64_t n = 0; n < o->N; ++n) {
        setv(o->node[EB_F_U], n, v3(o->D[6 * n], o->D[6 * n + 1], o->D[6 * n + 2]));
        setv(o->node[EB_F_UDOT], n, v3(o->Dd[6 * n], o->Dd[6 * n + 1], o->Dd[6 * n + 2]));
        setv(o->node[EB_F_UDDOT], n, v3(o->Ddd[6 * n], o->Ddd[6 * n + 1], o->Ddd[6 * n + 2]));
        M3 dR = rotation_matrix(v3(o->dD[6 * n + 3], o->dD[6 * n + 4], o->dD[6 * n + 5])) * getm(o->node[EB_F_DR], n);
        setm(o->node[EB_F_DR], n, dR);
        V3 wdn = getv(o->node[EB_F_WDOT_N], n), wddn = getv(o->node[EB_F_WDDOT_N], n);
        V3 w = toangle(dR);
        V3 a = ((g / (b * dt)) * w + ((b - g) / b) * wdn) + ((dt * (b - g / 2)) / b) * wddn;
        V3 c = ((1 / (b * (dt * dt))) * w - (1 / (b * dt)) * wdn) - (1 / (2 * b) - 1) * wddn;
        setv(o->node[EB_F_WDOT], n, dR * a);
        setv(o->node[EB_F_WDDOT], n, dR * c);
        setm(o->node[EB_F_R], n, dR * getm(o->node[EB_F_R_N], n));
    }
    return EB_OK;
}

// compute_norms_corrector, src/utils_solver.jl:126-151
int ebo_compute_norms_corrector(void* h, double* res_norm, double* dD_norm) {
    Oracle* o = (Oracle*)h;
    double sd = 0, sr = 0, sf = 0, se = 0;
    for (int64_t k = 0; k < o->n_free; ++k) { sd += o->dD_free[k] * o->dD_free[k]; sr += o->r_free[k] * o->r_free[k]; }
    for (int64_t k = 0; k < o->n_free; ++k) { int64_t i = o->freed[k]; double f = (o->fext[i] + o->Tc[i]) + o->Tk[i]; sf += f * f; }
    for (int64_t k = 0; k < o->n_fixed; ++k) { int64_t i = o->fixed[k]; double f = (o->fext[i] + o->Tc[i]) + o->Tk[i]; se += f * f; }
    double dn = std::sqrt(sd), rn = std::sqrt(sr), fn = std::sqrt(sf), en = std::sqrt(se);
    if (fn + en > 1e-12) rn = rn / (fn + en);
    *res_norm = rn; *dD_norm = dn;
    return EB_OK;
}

int ebo_update_loads(void* h, int64_t n, const int64_t* dofs, const double* vals) {
    Oracle* o = (Oracle*)h;
    for (int64_t k = 0; k < n; ++k) {
        if (dofs[k] < 1 || dofs[k] > o->ndofs) return fail(o, EB_ERR_ARG, "load dof out of range");
        o->fext[dofs[k] - 1] = vals[k];
    }
    return EB_OK;
}

int ebo_update_boundary_conditions(void* h, const double* vals, int64_t n) {
    Oracle* o = (Oracle*)h;
    if (n != o->n_fixed) return fail(o, EB_ERR_ARG, "imposed displacement count mismatch");
    for (int64_t k = 0; k < n; ++k) o->imposed[k] = vals[k];
    return EB_OK;
}

// initialize_state_simulation!, src/inizialization.jl:37-45
int ebo_initialize_state_simulation(void* h) {
    Oracle* o = (Oracle*)h;
    ebo_assemble(h);
    o->Kn = o->K; o->Cn = o->C; o->Mn = o->M;
    return EB_OK;
}

// predictor!, src/predictor.jl:1-27
int ebo_predictor(void* h, const eb_newton_params* p, double dt) {
    ebo_initialize_global_predictor(h);
    ebo_tangent_residuals_predictor(h, dt, p->alpha, p->beta, p->gamma);
    ebo_apply_boundary_conditions(h);
    ebo_extract_free_dofs(h);
    int st = ebo_solve_free_dofs(h, p, nullptr);
    if (st != EB_OK) return st;
    ebo_compute_global_predictor(h, dt, p->beta, p->gamma);
    ebo_compute_local_predictor(h, dt, p->beta, p->gamma);
    return EB_OK;
}

// corrector!, src/corrector.jl:1-59
int ebo_corrector(void* h, const eb_newton_params* p, double dt, int* kout) {
    Oracle* o = (Oracle*)h;
    ebo_initialize_global_corrector(h);
    int k = 1;
    double dn = INFINITY, rn = INFINITY;
    while ((rn > p->tol_res || dn > p->tol_disp) && k <= p->max_iterations) {
        ebo_assemble(h);
        if (o->surf.kind != EB_SURFACE_NONE) ebo_assemble_contact(h, p->alpha);
        ebo_tangent_residuals_corrector(h, dt, p->alpha, p->beta, p->gamma);
        ebo_apply_boundary_conditions(h);
        ebo_extract_free_dofs(h);
        int st = ebo_solve_free_dofs(h, p, nullptr);
        if (st != EB_OK) return st;
        ebo_compute_global_corrector(h, dt, p->beta, p->gamma);
        ebo_compute_local_corrector(h, dt, p->beta, p->gamma);
        ebo_compute_norms_corrector(h, &rn, &dn);
        o->last_rn = rn;
        o->last_dn = dn;
        o->counters[0]++;
        k += 1;
    }
    o->last_k = k;
    *kout = k;
    return EB_OK;
}

// solve_step_dynamics!, src/solve_step_dynamics.jl:1-14 (loads/BCs updated by the caller)
int ebo_solve_step_dynamics(void* h, const eb_newton_params* p, double dt, int* k) {
    int st = ebo_predictor(h, p, dt);
    if (st != EB_OK) return st;
    return ebo_corrector(h, p, dt, k);
}

// update_converged!, src/utils_solver.jl:252-275
int ebo_update_converged(void* h) {
    Oracle* o = (Oracle*)h;
    for (int f = 0; f < 8; ++f) o->node[EB_F_U_N + f] = o->node[EB_F_U + f];
    o->Tint_n = o->Tint; o->Tk_n = o->Tk; o->Tc_n = o->Tc; o->fext_n = o->fext;
    o->Kn = o->K; o->Cn = o->C; o->Mn = o->M;
    return EB_OK;
}

// update_not_converged!, src/utils_solver.jl:282-304
int ebo_update_not_converged(void* h) {
    Oracle* o = (Oracle*)h;
    for (int f = 0; f < 8; ++f) o->node[EB_F_U + f] = o->node[EB_F_U_N + f];
    o->Tint = o->Tint_n; o->Tk = o->Tk_n; o->Tc = o->Tc_n; o->fext = o->fext_n;
    o->K = o->Kn; o->C = o->Cn; o->M = o->Mn;
    return EB_OK;
}

// Batches: the oracle runs one structure per context; a batch is checked against each of its
// instances run alone, so only a batch of one is accepted here.
// run_simulation!'s time loop (src/run_simulation.jl:38-112) over the oracle's own step functions
int ebo_run_steps(void* h, const eb_newton_params* p, const eb_time_params* tp, int64_t max_steps,
                  int (*callback)(void*, double, void*), void* user, eb_run_state* state) {
    Oracle* o = (Oracle*)h;
    if (!p || !tp || !state || max_steps < 0) return EB_ERR_ARG;
    int64_t done = 0;
    while (done < max_steps && tp->simulation_end_time - state->t > 1.0e-10 && !state->stopped) {
        const double tnp1 = state->t + state->dt;
        if (callback && callback(h, tnp1, user) != 0) return fail(o, EB_ERR_ARG, "step callback failed");
        int k = 0;
        int st = ebo_solve_step_dynamics(h, p, state->dt, &k);
        if (st != EB_OK) return st;
        state->last_k = k;
        if (k > p->max_iterations) {
            state->dt = state->dt / 2;
            state->failed_steps += 1;
            if (state->dt < tp->min_timestep) return fail(o, EB_ERR_STATE, "time step below min_timestep: simulation aborted");
            st = ebo_update_not_converged(h);
            if (st != EB_OK) return st;
            state->successive = 0;
            continue;
        }
        state->successive += 1;
        state->steps += 1;
        state->newton_iterations += k - 1;
        done += 1;
        if (state->successive >= tp->accelerate_after_success_it && k <= 8) {
            state->dt = std::min(state->dt * 1.5, tp->max_timestep);
            state->successive = 0;
        }
        st = ebo_update_converged(h);
        if (st != EB_OK) return st;
        state->t = tnp1;
        if (tp->stop_on_energy_threshold) {
            double en[3];
            st = ebo_get_field(h, EB_F_ENERGY, en, 3);
            if (st != EB_OK) return st;
            if (en[1] < tp->energy_threshold) state->stopped = 1;
        }
    }
    return EB_OK;
}

// the per-instance time loop of a batch (a batch of one here): run_simulation!'s loop with an
// abort recorded as stopped = 2 instead of an error
int ebo_run_steps_batch(void* h, const eb_newton_params* p, const eb_time_params* tp, int64_t max_steps,
                        int (*callback)(void*, int64_t, double, void*), void* user, eb_run_state* state) {
    Oracle* o = (Oracle*)h;
    if (!p || !tp || !state || max_steps < 0) return EB_ERR_ARG;
    if (tp->stop_on_energy_threshold) return fail(o, EB_ERR_ARG, "the energy stop is not available per instance");
    int64_t done = 0;
    while (done < max_steps && tp->simulation_end_time - state->t > 1.0e-10 && !state->stopped) {
        const double tnp1 = state->t + state->dt;
        if (callback && callback(h, 0, tnp1, user) != 0) return fail(o, EB_ERR_ARG, "step callback failed");
        int k = 0;
        int st = ebo_solve_step_dynamics(h, p, state->dt, &k);
        if (st != EB_OK) return st;
        state->last_k = k;
        if (k > p->max_iterations) {
            state->dt = state->dt / 2;
            state->failed_steps += 1;
            state->successive = 0;
            st = ebo_update_not_converged(h);
            if (st != EB_OK) return st;
            if (state->dt < tp->min_timestep) state->stopped = 2;
            continue;
        }
        state->successive += 1;
        state->steps += 1;
        state->newton_iterations += k - 1;
        done += 1;
        if (state->successive >= tp->accelerate_after_success_it && k <= 8) {
            state->dt = std::min(state->dt * 1.5, tp->max_timestep);
            state->successive = 0;
        }
        st = ebo_update_converged(h);
        if (st != EB_OK) return st;
        state->t = tnp1;
    }
    return EB_OK;
}

// the oracle has no contact-tangent launches to count
int ebo_get_contact_stats(void* h, int64_t* out2) {
    out2[0] = out2[1] = 0;
    return EB_OK;
}
// device launch accounting: no counterpart on the CPU
int ebo_get_work_counters(void*, int64_t* out3) {
    out3[0] = out3[1] = out3[2] = 0;
    return EB_OK;
}

int ebo_create_batch(const eb_model_desc* d, int64_t n_instances, void** out) {
    if (n_instances != 1) return EB_ERR_ARG;
    return ebo_create(d, out);
}
int ebo_get_instances(void* h, int64_t* n_instances, int64_t* node_offsets, int64_t* beam_offsets) {
    Oracle* o = (Oracle*)h;
    if (n_instances) *n_instances = 1;
    if (node_offsets) { node_offsets[0] = 0; node_offsets[1] = o->N; }
    if (beam_offsets) { beam_offsets[0] = 0; beam_offsets[1] = o->E; }
    return EB_OK;
}
int ebo_get_instance_iterations(void* h, int32_t* k_out) {
    *k_out = ((Oracle*)h)->last_k;
    return EB_OK;
}
int ebo_get_instance_norms(void* h, double* rn_dn) {
    Oracle* o = (Oracle*)h;
    rn_dn[0] = o->last_rn;
    rn_dn[1] = o->last_dn;
    return EB_OK;
}

int ebo_get_counters(void* h, int64_t* out4) {
    Oracle* o = (Oracle*)h;
    for (int k = 0; k < 4; ++k) out4[k] = o->counters[k];
    return EB_OK;
}

// kernel timing has no CPU counterpart: the oracle reports zero launches
int ebo_profile_enable(void*, uint32_t) { return EB_OK; }
int ebo_profile_read(void*, int, double* ms, int64_t* n) { *ms = 0; *n = 0; return EB_OK; }

// execution mode and launch shapes have no CPU counterpart: the oracle accepts them and computes
// as always
int ebo_set_exec_mode(void*, int mode) { return (mode == EB_EXEC_EAGER || mode == EB_EXEC_GRAPH) ? EB_OK : EB_ERR_ARG; }
int ebo_get_exec_stats(const void*, int32_t* mode, int64_t* captures, int64_t* replays) {
    if (mode) *mode = EB_EXEC_EAGER;
    if (captures) *captures = 0;
    if (replays) *replays = 0;
    return EB_OK;
}
int ebo_autotune(void*, int32_t* choice, double* ms) {
    for (int i = 0; i < EB_TUNE_COUNT; ++i) choice[i] = 0;
    if (ms)
        for (int i = 0; i < EB_TUNE_COUNT * EB_TUNE_CANDIDATES; ++i) ms[i] = -1;
    return EB_OK;
}
int ebo_set_tuning(void*, const int32_t*) { return EB_OK; }
int ebo_get_tuning(const void*, int32_t* choice) {
    for (int i = 0; i < EB_TUNE_COUNT; ++i) choice[i] = 0;
    return EB_OK;
}

// ---- single-element probes for the oracle's own invariant tests ---------------------------

// in: X1,X2,u1,u2,ud1,ud2,udd1,udd2,wd1,wd2,wdd1,wdd2 (12×3), R1,R2,dR1,dR2 (4×9), l0, Re0 (9)
static ElemIn unpack_elem(const double* x) {
    ElemIn in;
    V3* vs[12] = {&in.X1, &in.X2, &in.u1, &in.u2, &in.ud1, &in.ud2, &in.udd1, &in.udd2, &in.wd1, &in.wd2, &in.wdd1, &in.wdd2};
    for (int k = 0; k < 12; ++k) *vs[k] = v3(x[3 * k], x[3 * k + 1], x[3 * k + 2]);
    M3* ms[4] = {&in.R1, &in.R2, &in.dR1, &in.dR2};
    for (int k = 0; k < 4; ++k) for (int q = 0; q < 9; ++q) ms[k]->m[q] = x[36 + 9 * k + q];
    in.l0 = x[72];
    for (int q = 0; q < 9; ++q) in.Re0.m[q] = x[73 + q];
    return in;
}

// out: strain, kinetic, Tint(12), Tk(12), K(144), M(144), C(144)
int ebo_compute_beam(const double* elem82, const double* props14, int n_gauss, const double* gw, const double* gz, double* out) {
    ElemIn in = unpack_elem(elem82);
    BeamProps bp = props_from(props14);
    Gauss g; g.n = n_gauss;
    for (int k = 0; k < n_gauss; ++k) { g.w[k] = gw[k]; g.z[k] = gz[k]; }
    BeamOut o;
    compute_beams(in, bp, g, o);
    out[0] = o.strain; out[1] = o.kinetic;
    std::memcpy(out + 2, o.Tint, 96); std::memcpy(out + 14, o.Tk, 96);
    std::memcpy(out + 26, o.K.a, 1152); std::memcpy(out + 170, o.M.a, 1152); std::memcpy(out + 314, o.C.a, 1152);
    return EB_OK;
}

// out: energy, active, Tc(12), Kc(144), Cc(144); surface given by a model context
int ebo_compute_contact_beam(void* h, const double* elem82, double radius, double* out) {
    Oracle* o = (Oracle*)h;
    if (o->surf.kind == EB_SURFACE_NONE) return fail(o, EB_ERR_STATE, "no contact surface");
    ElemIn in = unpack_elem(elem82);
    ContactOut c;
    compute_contact_beams(in, radius, o->cp, o->surf, o->g, c);
    out[0] = c.energy; out[1] = c.active;
    std::memcpy(out + 2, c.Tc, 96); std::memcpy(out + 14, c.K.a, 1152); std::memcpy(out + 158, c.C.a, 1152);
    return EB_OK;
}

// value, gradient (3), Hessian (9, column-major), raw (not normalised) B-spline SDF evaluation
int ebo_sdf_eval(void* h, const double* p, double* out13) {
    Oracle* o = (Oracle*)h;
    if (o->surf.kind != EB_SURFACE_SDF) return fail(o, EB_ERR_STATE, "no SDF");
    double v; V3 gr; M3 H;
    sdf_eval(o->surf.sdf, v3(p[0], p[1], p[2]), v, gr, H, true);
    out13[0] = v; for (int k = 0; k < 3; ++k) out13[1 + k] = gr[k]; for (int k = 0; k < 9; ++k) out13[4 + k] = H.m[k];
    return EB_OK;
}

// recompute_at_gausspts_beams! (src/IO/visualization_beams.jl:132-181)
int ebo_output_points(void* h, int32_t n_points, double* position, double* displacement, double* velocity,
                      double* strain, double* stress) {
    Oracle* o = (Oracle*)h;
    if (n_points < 1) return fail(o, EB_ERR_ARG, "n_points < 1");
    for (int64_t e = 0; e < o->E; ++e) {
        ElemIn in = gather_elem(o, e);
        V3 x1 = in.X1 + in.u1, x2 = in.X2 + in.u2;
        real ln = norm(x2 - x1);
        LocalFrame f = local_Re_and_aux(x1, x2, in.R1, in.R2, col(in.Re0, 1), ln);
        M3 ReT = tr(f.Re);
        V3 Tb1 = toangle((ReT * in.R1) * in.Re0), Tb2 = toangle((ReT * in.R2) * in.Re0);
        real Emod = o->props[(size_t)(EB_NPROPS * e + 1)];
        for (int i = 0; i < n_points; ++i) {
            int64_t k = e * n_points + i;
            real z = n_points > 1 ? -1.0 + (2.0 * i) / (n_points - 1) : -1.0;
            real xi = in.l0 * (z + 1) / 2;
            real N1 = 1 - xi / in.l0, N2 = 1 - N1;
            real c = 1 - xi / in.l0;
            real N3 = xi * (c * c);
            real N4 = -(1 - xi / in.l0) * ((xi * xi) / in.l0);
            V3 ut = v3(0.0, N3 * Tb1[2] + N4 * Tb2[2], -N3 * Tb1[1] - N4 * Tb2[1]);
            V3 xg = (N1 * x1 + N2 * x2) + f.Re * ut;
            V3 d = N1 * in.u1 + N2 * in.u2, w = N1 * in.ud1 + N2 * in.ud2;
            for (int a = 0; a < 3; ++a) {
                if (position) position[3 * k + a] = xg[a];
                if (displacement) displacement[3 * k + a] = d[a];
                if (velocity) velocity[3 * k + a] = w[a];
            }
            if (strain) strain[k] = 1 - ln / in.l0;
            if (stress) stress[k] = Emod * (1 - ln / in.l0);
        }
    }
    return EB_OK;
}

// the oracle allocates nothing on a device: the same as ebo_output_points
int ebo_output_points_noalloc(void* h, int32_t n_points, double* position, double* displacement, double* velocity,
                              double* strain, double* stress) {
    return ebo_output_points(h, n_points, position, displacement, velocity, strain, stress);
}

int ebo_get_sdf_coefficients(void* h, int64_t instance, double* out, int64_t count) {
    Oracle* o = (Oracle*)h;
    if (instance != 0) return fail(o, EB_ERR_ARG, "one instance");
    if (o->surf.kind != EB_SURFACE_SDF) return fail(o, EB_ERR_STATE, "no SDF");
    if ((int64_t)o->surf.sdf.coef.size() != count) return fail(o, EB_ERR_ARG, "size mismatch");
    std::memcpy(out, o->surf.sdf.coef.data(), sizeof(double) * count);
    return EB_OK;
}

// prefiltered coefficients ((n0+2)(n1+2)(n2+2), x fastest)
int ebo_sdf_coefs(void* h, double* out, int64_t count) {
    Oracle* o = (Oracle*)h;
    if (o->surf.kind != EB_SURFACE_SDF) return fail(o, EB_ERR_STATE, "no SDF");
    if ((int64_t)o->surf.sdf.coef.size() != count) return fail(o, EB_ERR_ARG, "size mismatch");
    std::memcpy(out, o->surf.sdf.coef.data(), sizeof(double) * count);
    return EB_OK;
}

// gap, normal, Hessian after the reference's normalisation (contact_gap_beams) and the contact test
int ebo_contact_gap(void* h, const double* p, double radius, double* out15) {
    Oracle* o = (Oracle*)h;
    if (o->surf.kind == EB_SURFACE_NONE) return fail(o, EB_ERR_STATE, "no contact surface");
    V3 x = v3(p[0], p[1], p[2]);
    out15[0] = incontact(x, o->surf, radius) ? 1.0 : 0.0;
    double g; V3 n; M3 H;
    contact_gap(x, o->surf, radius, g, n, H);
    out15[1] = g; for (int k = 0; k < 3; ++k) out15[2 + k] = n[k]; for (int k = 0; k < 9; ++k) out15[5 + k] = H.m[k];
    return EB_OK;
}

// rotation utilities for closed-form checks: op 0 rotation_matrix(v)→9, 1 toangle(M)→3,
// 2 Tₛ⁻¹(v)→9, 3 Tₛ(v)→9, 4 η,μ(v)→2, 5 get_Rₑ⁰(X1=in[0:3], X2=in[3:6])→9
int ebo_rot_util(int op, const double* in, double* out) {
    switch (op) {
        case 0: { M3 r = rotation_matrix(v3(in[0], in[1], in[2])); std::memcpy(out, r.m, 72); return EB_OK; }
        case 1: { M3 m; std::memcpy(m.m, in, 72); V3 v = toangle(m); out[0] = v[0]; out[1] = v[1]; out[2] = v[2]; return EB_OK; }
        case 2: { M3 r = Ts_inv(v3(in[0], in[1], in[2])); std::memcpy(out, r.m, 72); return EB_OK; }
        case 3: { M3 r = Ts(v3(in[0], in[1], in[2])); std::memcpy(out, r.m, 72); return EB_OK; }
        case 4: { double e, m; eta_mu(v3(in[0], in[1], in[2]), e, m); out[0] = e; out[1] = m; return EB_OK; }
        case 5: { M3 r = get_Re0(v3(in[0], in[1], in[2]), v3(in[3], in[4], in[5])); std::memcpy(out, r.m, 72); return EB_OK; }
        default: return EB_ERR_ARG;
    }
}

// regularize_gap_penalty_beams (p, p', Π) and smoothstep (η, η') for scalar checks
int ebo_regularization(double g, double rb, double eta, double* out5) {
    smoothstep(eta, g, rb, out5[3], out5[4]);
    regularize_gap_penalty(g, rb, out5[0], out5[1], out5[2]);
    return EB_OK;
}

}  // extern "C"
