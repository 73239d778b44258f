// eb_api.hip — libendobeams_hip.so: device state, kernel launches and the C ABI declared in
// include/endobeams_hip.h. Every entry point runs on the context's own HIP stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/endobeams_hip.h"
#include "../../include/endobeams_learned.h"
#include "eb_beam_kernel.h"
#include "eb_contact_kernel.h"
#include "eb_element.h"
#include "eb_fused_launch.h"
#include "eb_learned.h"
#include "eb_linalg.h"
#include "eb_mlp.h"
#include "eb_output.h"
#include "eb_sdf_prefilter.h"
#include "eb_setup.h"

using namespace ebd;

// Device staging of eb_output_points_noalloc: a file-scope array of EB_NOALLOC_SCRATCH_DOUBLES
// doubles, one copy on every GPU the library runs on, held until the process ends. In host code
// `scratch` is a host-side stand-in: its device address on the current GPU comes from
// hipGetSymbolAddress (scratch_setup, once per device at eb_create), and only that address is
// passed to kernels and copies.
static __device__ __align__(256) double scratch[EB_NOALLOC_SCRATCH_DOUBLES];

namespace {

constexpr int SCRATCH_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[SCRATCH_MAX_DEVICES] = {};

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
};

struct eb_ctx_impl;

}  // namespace

struct eb_ctx {
    std::string err;
    int device = 0;
    hipStream_t stream = nullptr;
    int precision = EB_PRECISION_FP64;
    ebh::Pattern pat;
    ebh::Chains chains;
    int64_t N = 0, E = 0, ndofs = 0, nnz = 0, n_free = 0, n_fixed = 0, nnz_free = 0, n_contact = 0;
    GaussPts gauss{};
    int surface = EB_SURFACE_NONE;
    double kn = 0, mu = 0, eps_t = 0, eta_n = 0;
    std::vector<void*> allocs;
    // instances (independent structures stored one after the other)
    int64_t S = 1;
    std::vector<int64_t> node_off, beam_off;          // S+1 each
    int32_t *node_inst = nullptr, *beam_inst = nullptr, *blk_inst = nullptr, *chain_inst = nullptr;
    uint8_t* active = nullptr;                        // [S] Newton activity
    int32_t* kiter = nullptr;                         // [S] k at the end of the corrector loop
    int32_t* kdone = nullptr;                         // [1] Newton iterations finalised since newton_reset
    NewtonAll* nall = nullptr;
    KrylovAll* kall = nullptr;
    int64_t *seg_lo = nullptr, *seg_hi = nullptr;
    int32_t *seg_inst = nullptr, *seg_cptr = nullptr;
    int32_t nchunks = 0;
    NewtonAll* h_nall = nullptr;
    KrylovAll* h_kall = nullptr;
    int32_t* h_kiter = nullptr;
    void* d_sdf = nullptr;          // [S] SDFDev<P>
    SphereDev* d_sphere = nullptr;  // [S]
    // learned warm start
    eb_mlp* mlp = nullptr;
    std::vector<float*> lact;       // the network's layer outputs for this context's nodes (owned here)
    double* ldD = nullptr;          // the increment of the network's start at its last evaluation
    float *lin_scale = nullptr, *lout_scale = nullptr, *ones = nullptr, *lX = nullptr, *lT = nullptr;
    double* hist = nullptr;         // rates of the last three converged steps of each instance (3 slots of 6N)
    int32_t* newest = nullptr;      // [S] slot each instance wrote last
    std::vector<int32_t> hcount;    // [S] converged steps recorded (at most 3)
    uint8_t* eligible = nullptr;    // [S] taking part in the step with three steps recorded
    uint8_t* h_eligible = nullptr;  // pinned copy
    double* pdD = nullptr;          // the reference predictor's increment of the current step
    double* snap = nullptr;         // node fields U..DR, then D, Ḋ, D̈ of the reference predictor's start
    double* snap_l = nullptr;       // the same of the network's start
    double* rn_learned = nullptr;   // [2S]
    uint8_t* use_plain = nullptr;   // [S]
    uint8_t* use_learned = nullptr; // [S]
    int32_t* pred_counts = nullptr; // [3]: starts kept, fallbacks (totals), starts kept by the last choice
    int32_t* h_pred = nullptr;      // pinned copy of pred_counts
    int32_t learn_max_skip = 16;    // backoff of a network whose start keeps losing
    int32_t learn_backoff = 0, learn_wait = 0;
    int64_t learn_skipped = 0;
    // the start of the Newton iteration has been assembled (with its tangent, without energies):
    // the first iteration forms only the residual and the energies
    bool start_assembled = false;
    // per-instance step: coefficients of each instance's Δt, the instances taking part in the
    // step, the predictor's solve failures, the factorisation's singular pivots, final k
    StepCoef* coef = nullptr;       // [S] device
    StepCoef* h_coef = nullptr;     // [S] pinned copy of what coef holds
    std::vector<double> dts;        // [S] Δt the coefficients were computed for
    double coef_b = 0, coef_g = 0;
    bool coef_set = false;
    uint64_t coef_ver = 0;          // incremented by every change of the coefficients
    uint8_t* part = nullptr;        // [S] device
    uint8_t* h_part = nullptr;      // [S] pinned copy
    bool all_part = true;           // every instance takes part
    uint8_t* pfail = nullptr;       // [S] the predictor's linear solve failed
    int32_t* pivot_inst = nullptr;  // [S] singular pivot block in the instance's factorisation
    uint8_t* sel = nullptr;         // [S] instance selection of commits and copy-backs
    uint8_t* h_sel = nullptr;       // pinned copy
    std::vector<int32_t> k_inst;    // [S] k of each instance at the end of the last corrector (0: not taking part)
    // node state
    double* node[16] = {};
    double* X0 = nullptr;
    int32_t *n1 = nullptr, *n2 = nullptr;
    double *l0 = nullptr, *Re0c = nullptr, *propsc = nullptr;
    // staging (precision dependent element type)
    void *sK = nullptr, *sC = nullptr, *sM = nullptr, *sTint = nullptr, *sTk = nullptr, *sTc = nullptr;
    double* senergy = nullptr;
    // pattern
    int32_t *bptr = nullptr, *bcol = nullptr, *blk_ptr = nullptr, *inc_ptr = nullptr;
    uint32_t *blk_src = nullptr, *inc_src = nullptr;
    int32_t *csc2bsr = nullptr, *sfree_bsr = nullptr, *freed = nullptr, *fixidx = nullptr;
    uint8_t* is_fixed = nullptr;
    // matrices (block-row order) and vectors
    double *K = nullptr, *C = nullptr, *M = nullptr, *Kn = nullptr, *Cn = nullptr, *Mn = nullptr, *Ktan = nullptr;
    int64_t chain_lds_max = 0;   // dynamic LDS the chain factorisation may take (bytes)
    int64_t solve_lds_max = 0;   // dynamic LDS the chain solve may take (bytes)
    bool np1_alias = false;  // K,C,M at n+1 currently equal the step-n matrices (after a swap)
    bool kcm_pending = false; // K,C,M at n+1 not yet gathered from the element staging of the last assembly
    // What the tangent, the chain factors and the matrices hold, so that a predictor whose tangent
    // (1+α)Kⁿ + γ/(βΔt) Cⁿ + Mⁿ/(βΔt²) equals the corrector's last one reuses its factorisation:
    // both are gathered from the same element staging with the same coefficients, bit for bit.
    struct TanTag {
        uint64_t ver = 0;           // staging version the tangent was gathered from (0: other)
        double c1 = 0;              // 1 + α
        uint64_t cver = 0;          // version of the per-instance step coefficients (Δt, β, γ)
        bool operator==(const TanTag& o) const { return ver == o.ver && c1 == o.c1 && cver == o.cver; }
    };
    uint64_t stage_ver = 1;         // incremented by every write of the element staging
    TanTag ktan_tag, factor_tag;    // Ktan's content; the tangent the chain factors are of
    uint64_t kcm_ver = 0, kcm_n_ver = 0;  // staging versions of K, C, M at n+1 and at n (0: other)
    double *fext = nullptr, *Tint = nullptr, *Tk = nullptr, *Tc = nullptr;
    double *fext_n = nullptr, *Tint_n = nullptr, *Tk_n = nullptr, *Tc_n = nullptr;
    double *D = nullptr, *Dd = nullptr, *Ddd = nullptr, *res = nullptr, *dD = nullptr;
    double* imposed = nullptr;
    double* energy = nullptr;  // running totals (3)
    // Krylov workspace
    double *kx = nullptr, *kr = nullptr, *krhat = nullptr, *kp = nullptr, *kv = nullptr, *ks = nullptr, *kt = nullptr,
           *ky = nullptr, *kz = nullptr;
    double* partials = nullptr;
    KrylovCtl* kctl = nullptr;
    int* kstatus = nullptr;
    double* norms = nullptr;  // device (2)
    // host-pinned mirrors for per-iteration reads
    KrylovCtl* h_kctl = nullptr;
    double* h_norms = nullptr;
    int* h_status = nullptr;
    // contact
    int32_t* cbeams = nullptr;
    uint8_t* is_contact = nullptr;  // [E]
    uint64_t* clist = nullptr;     // beams found in contact, (beam << 8 | Gauss point mask)
    int32_t* clist_n = nullptr;
    // chains
    int32_t *ch_ptr = nullptr, *ch_node = nullptr, *ch_diag = nullptr, *ch_lo = nullptr, *ch_up = nullptr;
    uint64_t *ch_mD = nullptr, *ch_mL = nullptr, *ch_mU = nullptr;
    uint32_t* ch_mrow = nullptr;   // per position and row: the row's fixed-dof bits of D, L, U
    int32_t* ch_hbase = nullptr;
    // closed chains (rings): their chain, coupling blocks and masks, T⁻¹U columns and H
    int32_t *cl_chain = nullptr, *cl_fl = nullptr, *cl_lf = nullptr;
    uint64_t *cl_mfl = nullptr, *cl_mlf = nullptr;
    int64_t* cl_zoff = nullptr;
    double *cl_Z = nullptr, *cl_H = nullptr;
    int32_t n_closures = 0;
    // the crossing chain cover (the other lines of a net): the couplings the first cover leaves
    // out, corrected after it, x += T₂⁻¹ (b − A x); dev2.n_chains = 0 when the first cover holds all
    ebh::Chains chains2;
    ChainDev dev2{};
    double* cross_r = nullptr;   // [ndofs] b − A x
    double *ch_D = nullptr, *ch_L = nullptr, *ch_U = nullptr, *ch_hist = nullptr, *ch_bw = nullptr, *ch_yw = nullptr;
    int precond_mode = 0;
    double* tmp = nullptr;  // nnz scratch for layout conversion
    int64_t counters[4] = {0, 0, 0, 0};
    // launch timing
    uint32_t prof_mask = 0;
    // work of the launches timed since eb_profile_enable: [0] beams through the element kernel,
    // [1] tangent entries through gather_matrices, [2] chain positions factored; per instance
    // (3 per instance) the same three sizes, summed over the instances active at each launch
    unsigned long long* work = nullptr;
    int64_t* inst_work = nullptr;
    // per instance: 1 where the chain preconditioner holds all couplings (the direct first step);
    // how many instances are of each kind
    uint8_t* inst_exact = nullptr;
    int64_t n_exact = 0, n_inexact = 0;
    std::vector<hipEvent_t> ev_pool;
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> ev_pending;
    double prof_ms[EB_KERNEL_COUNT] = {};
    int64_t prof_n[EB_KERNEL_COUNT] = {};
    hipEvent_t ev_open[EB_KERNEL_COUNT] = {};
    // One Newton iteration (assembly, lift, factorisation, one BiCGStab iteration, update and
    // norms behind the solve's done flag, the two status copies) captured as a HIP graph and
    // replayed while its parameters stay the same. Its launch timings come from event nodes
    // recorded at capture (one pair per timed kernel, read after each replay).
    struct NewtonGraph {
        hipGraphExec_t exec = nullptr;
        double key[9] = {};
        hipEvent_t ev[EB_KERNEL_COUNT][2] = {};
        bool timed[EB_KERNEL_COUNT] = {};
    };
    int exec_mode = EB_EXEC_EAGER;
    bool capturing = false;
    NewtonGraph graph;
    int64_t graph_stats[2] = {0, 0};   // captures, replays
    // launch shapes (eb_autotune)
    int32_t tune[EB_TUNE_COUNT] = {};
    // this device's copy of the file-scope staging array (eb_output_points_noalloc)
    double* scratch = nullptr;
    unsigned long long* contact_stats = nullptr;   // [2] listed beams, Gauss points in contact
    int64_t* fused_out = nullptr;   // [3] what the one-workgroup kernels report (allocated with EB_EXEC_FUSED)
    int fused_batch = 4;            // corrector iterations launched before the first wait
};

namespace {

#define EB_HIP(call)                                                                               \
    do {                                                                                           \
        hipError_t _e = (call);                                                                    \
        if (_e != hipSuccess) {                                                                    \
            ctx->err = std::string(#call) + ": " + hipGetErrorString(_e);                          \
            return EB_ERR_DEVICE;                                                                  \
        }                                                                                          \
    } while (0)

#define EB_TRY(expr)                \
    do {                            \
        int _s = (expr);            \
        if (_s != EB_OK) return _s; \
    } while (0)

hipEvent_t pool_event(eb_ctx* ctx) {
    if (!ctx->ev_pool.empty()) {
        hipEvent_t e = ctx->ev_pool.back();
        ctx->ev_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}
// while a Newton iteration is captured, the timing events become event nodes of the graph (the
// events are created before the capture begins)
bool graph_event(eb_ctx* ctx, int id, int which) {
    hipEvent_t e = ctx->graph.ev[id][which];
    return e && hipEventRecordWithFlags(e, ctx->stream, hipEventRecordExternal) == hipSuccess;
}
inline void prof_begin(eb_ctx* ctx, int id) {
    if (!(ctx->prof_mask & (1u << id))) return;
    if (ctx->capturing) {
        ctx->graph.timed[id] = graph_event(ctx, id, 0);
        return;
    }
    hipEvent_t e = pool_event(ctx);
    if (e && hipEventRecord(e, ctx->stream) == hipSuccess) ctx->ev_open[id] = e;
}
inline void prof_end(eb_ctx* ctx, int id) {
    if (ctx->capturing) {
        if ((ctx->prof_mask & (1u << id)) && ctx->graph.timed[id]) ctx->graph.timed[id] = graph_event(ctx, id, 1);
        return;
    }
    if (!(ctx->prof_mask & (1u << id)) || !ctx->ev_open[id]) return;
    hipEvent_t e = pool_event(ctx);
    if (e && hipEventRecord(e, ctx->stream) == hipSuccess) ctx->ev_pending.push_back({id, {ctx->ev_open[id], e}});
    ctx->ev_open[id] = nullptr;
}
// folds completed event pairs into the totals (the stream has been synchronised by the caller)
void prof_harvest(eb_ctx* ctx) {
    for (auto& p : ctx->ev_pending) {
        float ms = 0;
        if (hipEventElapsedTime(&ms, p.second.first, p.second.second) == hipSuccess) {
            ctx->prof_ms[p.first] += ms;
            ctx->prof_n[p.first] += 1;
        }
        ctx->ev_pool.push_back(p.second.first);
        ctx->ev_pool.push_back(p.second.second);
    }
    ctx->ev_pending.clear();
}
// folds the timings of the last graph replay into the totals (stream synchronised)
void prof_harvest_graph(eb_ctx* ctx) {
    for (int id = 0; id < EB_KERNEL_COUNT; ++id) {
        if (!ctx->graph.timed[id] || !(ctx->prof_mask & (1u << id))) continue;
        float ms = 0;
        if (hipEventElapsedTime(&ms, ctx->graph.ev[id][0], ctx->graph.ev[id][1]) == hipSuccess) {
            ctx->prof_ms[id] += ms;
            ctx->prof_n[id] += 1;
        }
    }
}

inline unsigned nblk(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

// work[which] += Σ over instances active in the launch of inst_work[3 s + which] (one workgroup)
__global__ void __launch_bounds__(256) work_account(int64_t ninst, const uint8_t* __restrict__ active,
                                                     const int64_t* __restrict__ sizes, int which,
                                                     unsigned long long* work) {
    __shared__ unsigned long long part[256];
    unsigned long long acc = 0;
    for (int64_t s = threadIdx.x; s < ninst; s += 256) acc += active[s] ? (unsigned long long)sizes[3 * s + which] : 0ull;
    part[threadIdx.x] = acc;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) part[threadIdx.x] += part[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) work[which] += part[0];
}
// counts the work of a timed launch (only while its kernel group is profiled)
inline void count_work(eb_ctx* ctx, int prof_id, int which) {
    if (!(ctx->prof_mask & (1u << prof_id)) || !ctx->work) return;
    hipLaunchKernelGGL(work_account, dim3(1), dim3(256), 0, ctx->stream, (int64_t)ctx->S, ctx->active, ctx->inst_work,
                       which, ctx->work);
}

// launch shapes chosen by eb_autotune (0: the built-in default)
constexpr int32_t TUNE_DEFAULT[EB_TUNE_COUNT] = {48, 1, 256, EB_CHAIN_BLOCK_LU};
inline int32_t tuned(const eb_ctx* ctx, int id) { return ctx->tune[id] > 0 ? ctx->tune[id] : TUNE_DEFAULT[id]; }
inline int gather_block(const eb_ctx* ctx) { return tuned(ctx, EB_TUNE_GATHER_BLOCK); }
inline int red_grid(int64_t n) { return (int)std::max<int64_t>(1, std::min<int64_t>(RED_MAX_GRID, (n + RED_BLOCK - 1) / RED_BLOCK)); }

template <typename T> int dalloc(eb_ctx* ctx, T*& out, int64_t count) {
    void* p = nullptr;
    size_t bytes = (size_t)std::max<int64_t>(count, 1) * sizeof(T);
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) { ctx->err = std::string("hipMalloc: ") + hipGetErrorString(e); return EB_ERR_DEVICE; }
    e = hipMemsetAsync(p, 0, bytes, ctx->stream);
    if (e != hipSuccess) { ctx->err = std::string("hipMemsetAsync: ") + hipGetErrorString(e); return EB_ERR_DEVICE; }
    ctx->allocs.push_back(p);
    out = (T*)p;
    return EB_OK;
}
template <typename T> int upload(eb_ctx* ctx, T*& out, const T* src, int64_t count) {
    EB_TRY(dalloc(ctx, out, count));
    if (count > 0) EB_HIP(hipMemcpyAsync(out, src, sizeof(T) * count, hipMemcpyHostToDevice, ctx->stream));
    EB_HIP(hipStreamSynchronize(ctx->stream));
    return EB_OK;
}
template <typename T> int upload_vec(eb_ctx* ctx, T*& out, const std::vector<T>& v) {
    return upload(ctx, out, v.data(), (int64_t)v.size());
}
template <typename T, typename S> std::vector<T> narrow(const std::vector<S>& v) {
    return std::vector<T>(v.begin(), v.end());
}

NodeState node_state(eb_ctx* c) {
    return NodeState{c->X0, c->node[EB_F_U], c->node[EB_F_UDOT], c->node[EB_F_UDDOT], c->node[EB_F_WDOT],
                     c->node[EB_F_WDDOT], c->node[EB_F_R], c->node[EB_F_DR]};
}
NodeFields node_fields(eb_ctx* c) {
    NodeFields f;
    double** p = &f.u;
    for (int k = 0; k < 16; ++k) p[k] = c->node[k];
    return f;
}
BeamConsts beam_consts(eb_ctx* c) { return BeamConsts{c->n1, c->n2, c->l0, c->Re0c, c->propsc, c->E}; }
template <typename P> BeamStage<P> beam_stage(eb_ctx* c) {
    return BeamStage<P>{(P*)c->sK, (P*)c->sC, (P*)c->sM, (double*)c->sTint, (double*)c->sTk, c->senergy};
}
ClosureDev closure_dev(eb_ctx* c) {
    return ClosureDev{c->cl_chain, c->cl_fl, c->cl_lf, c->cl_mfl, c->cl_mlf, c->cl_zoff, c->cl_Z, c->cl_H, c->n_closures};
}
ChainDev chain_dev(eb_ctx* c) {
    return ChainDev{c->ch_ptr, c->ch_node, c->ch_diag, c->ch_lo, c->ch_up, c->ch_mD, c->ch_mL, c->ch_mU,
                    c->ch_hbase, c->ch_D, c->ch_L, c->ch_U, c->ch_hist, c->ch_bw, c->ch_yw, c->chain_inst,
                    (int64_t)c->chains.ptr.size() - 1, c->ch_mrow};
}
Active active_dev(eb_ctx* c) { return Active{c->active, c->node_inst, c->beam_inst, c->blk_inst}; }
Segs segs(eb_ctx* c) { return Segs{c->seg_lo, c->seg_hi, c->seg_inst, c->seg_cptr, c->nchunks, (int32_t)c->S}; }
HHTc hht_consts(double dt, double b, double g) {
    HHTc h;
    h.c_w = g / (b * dt);
    h.c_wd = (b - g) / b;
    h.c_wdd = (dt * (b - g / 2)) / b;
    h.d_w = 1 / (b * (dt * dt));
    h.d_wd = 1 / (b * dt);
    h.d_wdd = 1 / (2 * b) - 1;
    return h;
}

// the coefficients of one instance's step, with the expressions the kernels took as arguments
// when every instance shared one Δt
StepCoef step_coef(double dt, double b, double g) {
    StepCoef c;
    c.dt = dt;
    c.gbdt = g / (b * dt);
    c.inv = 1 / (b * (dt * dt));
    c.gob = g / b;
    c.dt2b = dt * (2 * b - g) / (2 * b);
    c.t2c = ((dt / 2) * (2 * b - g)) / b;
    c.h = hht_consts(dt, b, g);
    return c;
}

// Sets the step coefficients of every instance: Δt = dt_inst[s] (one per instance), or `dt` for
// all when dt_inst is null. Uploads and bumps coef_ver only when something changed.
int set_coefs(eb_ctx* ctx, const double* dt_inst, double dt, double b, double g) {
    bool same = ctx->coef_set && b == ctx->coef_b && g == ctx->coef_g;
    for (int64_t s = 0; s < ctx->S && same; ++s) same = ctx->dts[s] == (dt_inst ? dt_inst[s] : dt);
    if (same) return EB_OK;
    // the pinned table may still feed a copy in flight
    EB_HIP(hipStreamSynchronize(ctx->stream));
    for (int64_t s = 0; s < ctx->S; ++s) {
        ctx->dts[s] = dt_inst ? dt_inst[s] : dt;
        ctx->h_coef[s] = step_coef(ctx->dts[s], b, g);
    }
    ctx->coef_b = b;
    ctx->coef_g = g;
    ctx->coef_set = true;
    ctx->coef_ver++;
    EB_HIP(hipMemcpyAsync(ctx->coef, ctx->h_coef, sizeof(StepCoef) * ctx->S, hipMemcpyHostToDevice, ctx->stream));
    return EB_OK;
}

// the device address of `scratch` on the context's device, looked up once per device
int scratch_setup(eb_ctx* ctx) {
    if (ctx->device < 0 || ctx->device >= SCRATCH_MAX_DEVICES) {
        ctx->err = "device index beyond the staging table";
        return EB_ERR_DEVICE;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[ctx->device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            ctx->err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
            return EB_ERR_DEVICE;
        }
        g_scratch_ptr[ctx->device] = (double*)p;
    }
    ctx->scratch = g_scratch_ptr[ctx->device];
    return EB_OK;
}

int check_launch(eb_ctx* ctx, const char* what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) { ctx->err = std::string(what) + ": " + hipGetErrorString(e); return EB_ERR_DEVICE; }
    return EB_OK;
}

// K, C, M at n+1 from the element staging of the last assembly, if that assembly left them pending
template <typename P> int materialize_kcm_t(eb_ctx* ctx) {
    ctx->kcm_ver = ctx->stage_ver;
    const int bs = gather_block(ctx);
    hipLaunchKernelGGL(gather_matrices<P>, dim3(nblk(ctx->nnz / 4, bs)), dim3(bs), 0, ctx->stream, ctx->nnz, ctx->blk_ptr,
                       ctx->blk_src, (const P*)ctx->sK, (const P*)ctx->sC, (const P*)ctx->sM, ctx->K, ctx->C, ctx->M,
                       ctx->Ktan, 0.0, (const StepCoef*)ctx->coef, 0, 1, active_dev(ctx));
    return check_launch(ctx, "gather_matrices");
}
int materialize_kcm(eb_ctx* ctx) {
    if (!ctx->kcm_pending) return EB_OK;
    ctx->kcm_pending = false;
    ctx->np1_alias = false;
    return ctx->precision == EB_PRECISION_FP32 ? materialize_kcm_t<float>(ctx) : materialize_kcm_t<double>(ctx);
}

// materialise K,C,M at n+1 if they are currently aliased to step n
int unalias(eb_ctx* ctx) {
    EB_TRY(materialize_kcm(ctx));
    if (!ctx->np1_alias) return EB_OK;
    size_t b = sizeof(double) * ctx->nnz;
    EB_HIP(hipMemcpyAsync(ctx->K, ctx->Kn, b, hipMemcpyDeviceToDevice, ctx->stream));
    EB_HIP(hipMemcpyAsync(ctx->C, ctx->Cn, b, hipMemcpyDeviceToDevice, ctx->stream));
    EB_HIP(hipMemcpyAsync(ctx->M, ctx->Mn, b, hipMemcpyDeviceToDevice, ctx->stream));
    ctx->np1_alias = false;
    ctx->kcm_ver = ctx->kcm_n_ver;
    return EB_OK;
}

// ---- assembly -------------------------------------------------------------------------------

template <typename P> int launch_beams(eb_ctx* ctx) {
    ctx->start_assembled = false;
    ctx->stage_ver++;
    prof_begin(ctx, EB_KERNEL_BEAM);
    launch_beam_kernel<P>(nblk(ctx->E, BEAM_TILE), ctx->stream, node_state(ctx), beam_consts(ctx), beam_stage<P>(ctx),
                          ctx->gauss, ctx->active, ctx->beam_inst);
    prof_end(ctx, EB_KERNEL_BEAM);
    count_work(ctx, EB_KERNEL_BEAM, 0);
    ctx->counters[3]++;
    return check_launch(ctx, "beam_kernel");
}

template <typename P> ContactArgs<P> contact_args(eb_ctx* ctx, double alpha) {
    ContactArgs<P> ca;
    ca.surface = ctx->surface;
    ca.sphere = ctx->d_sphere;
    ca.sdf = (const SDFDev<P>*)ctx->d_sdf;
    ca.beam_inst = ctx->beam_inst;
    ca.active = ctx->active;
    ca.kn = ctx->kn; ca.mu = ctx->mu; ca.eps_t = ctx->eps_t; ca.eta_n = ctx->eta_n; ca.alpha = alpha;
    ca.beams = ctx->cbeams;
    ca.n = ctx->n_contact;
    ca.Tc = (double*)ctx->sTc;
    ca.energy = ctx->senergy;
    ca.stats = ctx->contact_stats;
    return ca;
}

template <typename P> int launch_contact_tangent(eb_ctx* ctx, const ContactArgs<P>& ca) {
    prof_begin(ctx, EB_KERNEL_CONTACT);
    launch_contact_kernel<P>(nblk(ctx->n_contact, CONTACT_TILE), ctx->stream, node_state(ctx), beam_consts(ctx),
                             beam_stage<P>(ctx), ctx->gauss, ca, ctx->clist, ctx->clist_n);
    prof_end(ctx, EB_KERNEL_CONTACT);
    return check_launch(ctx, "contact_kernel");
}

__global__ void contact_reset(int32_t* count) { *count = 0; }

// element kernel with the contact detection in its static part, then the contact tangent
template <typename P> int launch_beams_contact(eb_ctx* ctx, double alpha) {
    ctx->start_assembled = false;
    const ContactArgs<P> ca = contact_args<P>(ctx, alpha);
    ctx->stage_ver++;
    hipLaunchKernelGGL(contact_reset, dim3(1), dim3(1), 0, ctx->stream, ctx->clist_n);
    prof_begin(ctx, EB_KERNEL_BEAM);
    launch_beam_detect_kernel<P>(nblk(ctx->E, BEAM_TILE), ctx->stream, node_state(ctx), beam_consts(ctx),
                                 beam_stage<P>(ctx), ctx->gauss, ctx->active, ctx->beam_inst, ca, ctx->is_contact,
                                 ctx->clist, ctx->clist_n);
    prof_end(ctx, EB_KERNEL_BEAM);
    count_work(ctx, EB_KERNEL_BEAM, 0);
    ctx->counters[3]++;
    EB_TRY(check_launch(ctx, "beam_detect_kernel"));
    return launch_contact_tangent<P>(ctx, ca);
}

// contact detection and tangent on the staging of the last element assembly
template <typename P> int launch_contact(eb_ctx* ctx, double alpha) {
    ctx->start_assembled = false;
    const ContactArgs<P> ca = contact_args<P>(ctx, alpha);
    ctx->stage_ver++;
    hipLaunchKernelGGL(contact_reset, dim3(1), dim3(1), 0, ctx->stream, ctx->clist_n);
    launch_contact_detect<P>(nblk(ctx->n_contact, 256), ctx->stream, node_state(ctx), beam_consts(ctx), ctx->gauss, ca,
                             ctx->clist, ctx->clist_n);
    EB_TRY(check_launch(ctx, "contact_detect"));
    return launch_contact_tangent<P>(ctx, ca);
}

// matrices false: the vectors (and the residual) only, K, C, M and Ktan left as they are
template <typename P> int launch_gathers(eb_ctx* ctx, bool with_contact, bool tangent, double dt, double a, double b,
                                         double g, bool residual, int energy_mask, bool matrices = true) {
    // the tangent's Δt-dependent factors come from the per-instance step coefficients
    (void)dt; (void)b; (void)g;
    const double c1 = 1 + a;
    prof_begin(ctx, EB_KERNEL_GATHER);
    if (matrices) {
        // inside the Newton iteration only Ktan is needed: K, C, M at n+1 are gathered from the
        // element staging when they are first read or when the step is committed (materialize_kcm)
        // (K, C, M are passed only when written: their buffers are exchanged when a step is committed)
        const int bs = gather_block(ctx);
        if (tangent) prof_begin(ctx, EB_KERNEL_GATHER_MATRICES);
        hipLaunchKernelGGL(gather_matrices<P>, dim3(nblk(ctx->nnz / 4, bs)), dim3(bs), 0, ctx->stream, ctx->nnz,
                           ctx->blk_ptr, ctx->blk_src, (const P*)ctx->sK, (const P*)ctx->sC, (const P*)ctx->sM,
                           tangent ? nullptr : ctx->K, tangent ? nullptr : ctx->C, tangent ? nullptr : ctx->M, ctx->Ktan,
                           c1, (const StepCoef*)ctx->coef, tangent ? 1 : 0, tangent ? 0 : 1, active_dev(ctx));
        if (tangent) {
            prof_end(ctx, EB_KERNEL_GATHER_MATRICES);
            count_work(ctx, EB_KERNEL_GATHER_MATRICES, 1);
        }
        EB_TRY(check_launch(ctx, "gather_matrices"));
        ctx->kcm_pending = tangent;
        if (tangent) ctx->ktan_tag = eb_ctx::TanTag{ctx->stage_ver, c1, ctx->coef_ver};
        else ctx->kcm_ver = ctx->stage_ver;
        ctx->np1_alias = false;
    }
    hipLaunchKernelGGL(gather_vectors<P>, dim3(nblk(ctx->ndofs, 256)), dim3(256), 0, ctx->stream, ctx->ndofs,
                       ctx->inc_ptr, ctx->inc_src, (const double*)ctx->sTint, (const double*)ctx->sTk, (const double*)ctx->sTc,
                       with_contact ? 1 : 0, ctx->Tint, ctx->Tk, ctx->Tc, residual ? 1 : 0, a, ctx->fext, ctx->fext_n,
                       ctx->Tc_n, ctx->Tint_n, ctx->res, active_dev(ctx));
    EB_TRY(check_launch(ctx, "gather_vectors"));
    if (energy_mask) {
        int grid = red_grid(ctx->E);
        hipLaunchKernelGGL(energy_partials, dim3(grid), dim3(RED_BLOCK), 0, ctx->stream, ctx->E, ctx->senergy,
                           ctx->partials, energy_mask, active_dev(ctx));
        hipLaunchKernelGGL(energy_finalize, dim3(1), dim3(RED_BLOCK), 0, ctx->stream, ctx->partials, grid, ctx->energy);
    }
    prof_end(ctx, EB_KERNEL_GATHER);
    return check_launch(ctx, "energy");
}

// assemble! [+ assemble_contact!] [+ Ktan and the corrector residual]
int assemble_all(eb_ctx* ctx, bool contact, double alpha, bool fused, double dt, double b, double g,
                 bool energy = true) {
    bool wc = contact && ctx->surface != EB_SURFACE_NONE;
    const int emask = energy ? (wc ? 3 : 1) : 0;
    if (ctx->precision == EB_PRECISION_FP32) {
        EB_TRY(wc ? launch_beams_contact<float>(ctx, alpha) : launch_beams<float>(ctx));
        return launch_gathers<float>(ctx, wc, fused, dt, alpha, b, g, fused, emask);
    }
    EB_TRY(wc ? launch_beams_contact<double>(ctx, alpha) : launch_beams<double>(ctx));
    return launch_gathers<double>(ctx, wc, fused, dt, alpha, b, g, fused, emask);
}

// the element kernels and the force gathers with the corrector residual, no matrices (a start
// evaluated for its residual only)
int assemble_residual(eb_ctx* ctx, double alpha, double dt, double b, double g) {
    bool wc = ctx->surface != EB_SURFACE_NONE;
    if (ctx->precision == EB_PRECISION_FP32) {
        EB_TRY(wc ? launch_beams_contact<float>(ctx, alpha) : launch_beams<float>(ctx));
        return launch_gathers<float>(ctx, wc, false, dt, alpha, b, g, true, 0, false);
    }
    EB_TRY(wc ? launch_beams_contact<double>(ctx, alpha) : launch_beams<double>(ctx));
    return launch_gathers<double>(ctx, wc, false, dt, alpha, b, g, true, 0, false);
}

// the force gathers, corrector residual and energies of the element staging in place (its
// assembly and tangent already done by start_residual)
int residual_of_staging(eb_ctx* ctx, double alpha, double dt, double b, double g) {
    bool wc = ctx->surface != EB_SURFACE_NONE;
    const int emask = wc ? 3 : 1;
    if (ctx->precision == EB_PRECISION_FP32) return launch_gathers<float>(ctx, wc, false, dt, alpha, b, g, true, emask, false);
    return launch_gathers<double>(ctx, wc, false, dt, alpha, b, g, true, emask, false);
}

int bc_lift_launch(eb_ctx* ctx) {
    if (ctx->n_fixed == 0) return EB_OK;
    hipLaunchKernelGGL(bc_lift, dim3(nblk(ctx->ndofs, 256)), dim3(256), 0, ctx->stream, ctx->ndofs, ctx->bptr, ctx->bcol,
                       ctx->fixidx, ctx->imposed, ctx->D, ctx->Ktan, ctx->dD, ctx->res, active_dev(ctx));
    return check_launch(ctx, "bc_lift");
}

// ---- linear solve ------------------------------------------------------------------------

// the fixed-dof bits of each position's row r in D, L and U (6 bits each, D lowest), the form the
// block-LU kernels read: one word per lane instead of the three 64-bit block masks
std::vector<uint32_t> row_masks(const ebh::Chains& c) {
    std::vector<uint32_t> m(6 * c.mD.size());
    for (size_t g = 0; g < c.mD.size(); ++g)
        for (int r = 0; r < 6; ++r)
            m[6 * g + r] = (uint32_t)((c.mD[g] >> (6 * r)) & 63) | (uint32_t)((c.mL[g] >> (6 * r)) & 63) << 6 |
                           (uint32_t)((c.mU[g] >> (6 * r)) & 63) << 12;
    return m;
}

// the block-LU chain kernels: five chains per 64-lane workgroup
inline unsigned lu_grid(int64_t items) { return (unsigned)((items + LU_CHAINS - 1) / LU_CHAINS); }
inline bool chain_lu(const eb_ctx* ctx) { return tuned(ctx, EB_TUNE_CHAIN_ALGO) == EB_CHAIN_BLOCK_LU; }

int precond_factor(eb_ctx* ctx) {
    int64_t nch = (int64_t)ctx->chains.ptr.size() - 1;
    prof_begin(ctx, EB_KERNEL_PRECOND);
    if (chain_lu(ctx)) {
        hipLaunchKernelGGL(chain_lu_factor, dim3(lu_grid(nch)), dim3(64), 0, ctx->stream, chain_dev(ctx), ctx->Ktan,
                           ctx->pivot_inst, ctx->active);
        if (ctx->n_closures > 0) {
            hipLaunchKernelGGL(closure_columns_lu, dim3(lu_grid(12 * (int64_t)ctx->n_closures)), dim3(64), 0, ctx->stream,
                               chain_dev(ctx), closure_dev(ctx), ctx->active);
            hipLaunchKernelGGL(closure_setup, dim3((unsigned)ctx->n_closures), dim3(64), 0, ctx->stream, chain_dev(ctx),
                               closure_dev(ctx), ctx->Ktan, ctx->active);
        }
        if (ctx->dev2.n_chains > 0)
            hipLaunchKernelGGL(chain_lu_factor, dim3(lu_grid(ctx->dev2.n_chains)), dim3(64), 0, ctx->stream, ctx->dev2,
                               ctx->Ktan, ctx->pivot_inst, ctx->active);
        ctx->factor_tag = ctx->ktan_tag;
        prof_end(ctx, EB_KERNEL_PRECOND);
        count_work(ctx, EB_KERNEL_PRECOND, 2);
        return check_launch(ctx, "chain_lu_factor");
    }
    // the reduced chain stays in LDS once it fits in the budget (48 KB by default: with the
    // kernel's 8 KB of static LDS, two workgroups per CU; 96 KB: one workgroup per CU, the chain
    // in LDS one level earlier). Budgets beyond 64 KB of dynamic LDS were enabled at eb_create.
    const int L0 = chain_lds_level(ctx->chains.max_len, 1024 * (int64_t)tuned(ctx, EB_TUNE_CHAIN_LDS_KB));
    const size_t smem = chain_lds_bytes(ctx->chains.max_len, L0);
    if (smem > (size_t)ctx->chain_lds_max) {
        ctx->err = "chain factorisation: the LDS budget exceeds what the device allows";
        return EB_ERR_ARG;
    }
    hipLaunchKernelGGL(chain_factor, dim3((unsigned)nch), dim3(CHAIN_THREADS), smem, ctx->stream, chain_dev(ctx), ctx->Ktan,
                       ctx->pivot_inst, L0, ctx->active);
    if (ctx->n_closures > 0) {
        // rings: the 12 columns of T⁻¹U (LDS: 12 doubles per position, at most 60 KB), then H
        const size_t cl_smem = (size_t)(12 * ctx->chains.cl_max_len) * sizeof(double);
        hipLaunchKernelGGL(closure_columns<SOLVE_THREADS_MANY>, dim3((unsigned)ctx->n_closures, 12),
                           dim3(SOLVE_THREADS_MANY), cl_smem, ctx->stream, chain_dev(ctx), closure_dev(ctx), ctx->active);
        hipLaunchKernelGGL(closure_setup, dim3((unsigned)ctx->n_closures), dim3(64), 0, ctx->stream, chain_dev(ctx),
                           closure_dev(ctx), ctx->Ktan, ctx->active);
    }
    if (ctx->dev2.n_chains > 0) {
        const int L02 = chain_lds_level(ctx->chains2.max_len, 1024 * (int64_t)tuned(ctx, EB_TUNE_CHAIN_LDS_KB));
        const size_t smem2 = chain_lds_bytes(ctx->chains2.max_len, L02);
        if (smem2 > (size_t)ctx->chain_lds_max) {
            ctx->err = "chain factorisation: the LDS budget exceeds what the device allows";
            return EB_ERR_ARG;
        }
        hipLaunchKernelGGL(chain_factor, dim3((unsigned)ctx->dev2.n_chains), dim3(CHAIN_THREADS), smem2, ctx->stream,
                           ctx->dev2, ctx->Ktan, ctx->pivot_inst, L02, ctx->active);
    }
    ctx->factor_tag = ctx->ktan_tag;
    prof_end(ctx, EB_KERNEL_PRECOND);
    count_work(ctx, EB_KERNEL_PRECOND, 2);
    return check_launch(ctx, "chain_factor");
}

int precond_apply(eb_ctx* ctx, const double* b, double* x, int direct_only = 0) {
    int64_t nch = (int64_t)ctx->chains.ptr.size() - 1;
    if (chain_lu(ctx)) {
        hipLaunchKernelGGL(chain_lu_solve<false>, dim3(lu_grid(nch)), dim3(64), 0, ctx->stream, chain_dev(ctx), b, x,
                           ctx->kctl, direct_only);
        if (ctx->n_closures > 0)
            hipLaunchKernelGGL(closure_apply, dim3((unsigned)ctx->n_closures), dim3(256), 0, ctx->stream, chain_dev(ctx),
                               closure_dev(ctx), x, ctx->kctl, direct_only);
        if (ctx->dev2.n_chains > 0 && !direct_only) {   // no direct instance has a crossing cover
            const Segs sg = segs(ctx);
            hipLaunchKernelGGL(residual_masked, dim3((unsigned)ctx->nchunks), dim3(RED_BLOCK), 0, ctx->stream, sg, ctx->bptr,
                               ctx->bcol, ctx->is_fixed, ctx->Ktan, x, b, ctx->cross_r, ctx->kctl, direct_only);
            hipLaunchKernelGGL(chain_lu_solve<true>, dim3(lu_grid(ctx->dev2.n_chains)), dim3(64), 0, ctx->stream, ctx->dev2,
                               ctx->cross_r, x, ctx->kctl, direct_only);
        }
        return check_launch(ctx, "chain_lu_solve");
    }
    const size_t smem = tuned(ctx, EB_TUNE_CHAIN_SOLVE_LDS) == 1 ? chain_solve_lds_bytes(ctx->chains.max_len, ctx->solve_lds_max) : 0;
    const bool many = nch >= SOLVE_MANY_CHAINS;
    const dim3 grid((unsigned)nch), block(many ? SOLVE_THREADS_MANY : SOLVE_THREADS_FEW);
    if (smem && many)
        hipLaunchKernelGGL((chain_solve<true, SOLVE_THREADS_MANY>), grid, block, smem, ctx->stream, chain_dev(ctx), b, x, ctx->kctl, direct_only);
    else if (smem)
        hipLaunchKernelGGL((chain_solve<true, SOLVE_THREADS_FEW>), grid, block, smem, ctx->stream, chain_dev(ctx), b, x, ctx->kctl, direct_only);
    else if (many)
        hipLaunchKernelGGL((chain_solve<false, SOLVE_THREADS_MANY>), grid, block, 0, ctx->stream, chain_dev(ctx), b, x, ctx->kctl, direct_only);
    else
        hipLaunchKernelGGL((chain_solve<false, SOLVE_THREADS_FEW>), grid, block, 0, ctx->stream, chain_dev(ctx), b, x, ctx->kctl, direct_only);
    if (ctx->n_closures > 0)
        hipLaunchKernelGGL(closure_apply, dim3((unsigned)ctx->n_closures), dim3(256), 0, ctx->stream, chain_dev(ctx),
                           closure_dev(ctx), x, ctx->kctl, direct_only);
    if (ctx->dev2.n_chains > 0 && !direct_only) {   // no direct instance has a crossing cover
        // the crossing cover: x += T₂⁻¹ (b − A x)
        const Segs sg = segs(ctx);
        hipLaunchKernelGGL(residual_masked, dim3((unsigned)ctx->nchunks), dim3(RED_BLOCK), 0, ctx->stream, sg, ctx->bptr,
                           ctx->bcol, ctx->is_fixed, ctx->Ktan, x, b, ctx->cross_r, ctx->kctl, direct_only);
        const int64_t n2 = ctx->dev2.n_chains;
        const size_t smem2 = tuned(ctx, EB_TUNE_CHAIN_SOLVE_LDS) == 1 ? chain_solve_lds_bytes(ctx->chains2.max_len, ctx->solve_lds_max) : 0;
        const bool many2 = n2 >= SOLVE_MANY_CHAINS;
        const dim3 grid2((unsigned)n2), block2(many2 ? SOLVE_THREADS_MANY : SOLVE_THREADS_FEW);
        if (smem2 && many2)
            hipLaunchKernelGGL((chain_solve<true, SOLVE_THREADS_MANY, true>), grid2, block2, smem2, ctx->stream, ctx->dev2,
                               ctx->cross_r, x, ctx->kctl, direct_only);
        else if (smem2)
            hipLaunchKernelGGL((chain_solve<true, SOLVE_THREADS_FEW, true>), grid2, block2, smem2, ctx->stream, ctx->dev2,
                               ctx->cross_r, x, ctx->kctl, direct_only);
        else if (many2)
            hipLaunchKernelGGL((chain_solve<false, SOLVE_THREADS_MANY, true>), grid2, block2, 0, ctx->stream, ctx->dev2,
                               ctx->cross_r, x, ctx->kctl, direct_only);
        else
            hipLaunchKernelGGL((chain_solve<false, SOLVE_THREADS_FEW, true>), grid2, block2, 0, ctx->stream, ctx->dev2,
                               ctx->cross_r, x, ctx->kctl, direct_only);
    }
    return check_launch(ctx, "chain_solve");
}

int krylov_iteration(eb_ctx* ctx) {
    const Segs sg = segs(ctx);
    const unsigned nc = (unsigned)ctx->nchunks;
    prof_begin(ctx, EB_KERNEL_KRYLOV);
    hipLaunchKernelGGL(krylov_p, dim3(nc), dim3(RED_BLOCK), 0, ctx->stream, sg, ctx->kctl, ctx->kr, ctx->kp, ctx->kv);
    EB_TRY(precond_apply(ctx, ctx->kp, ctx->ky));
    hipLaunchKernelGGL(spmv_masked<1>, dim3(nc), dim3(RED_BLOCK), 0, ctx->stream, sg, ctx->bptr, ctx->bcol,
                       ctx->is_fixed, ctx->Ktan, ctx->ky, ctx->kv, ctx->krhat, ctx->partials, ctx->kctl);
    hipLaunchKernelGGL(krylov_alpha_fin, dim3(1), dim3(FIN_BLOCK), 0, ctx->stream, ctx->partials, sg, ctx->kctl);
    hipLaunchKernelGGL(krylov_s, dim3(nc), dim3(RED_BLOCK), 0, ctx->stream, sg, ctx->kctl, ctx->kr, ctx->kv, ctx->ky,
                       ctx->ks, ctx->kx, ctx->partials);
    hipLaunchKernelGGL(krylov_s_fin, dim3(1), dim3(FIN_BLOCK), 0, ctx->stream, ctx->partials, sg, ctx->kctl);
    EB_TRY(precond_apply(ctx, ctx->ks, ctx->kz));
    hipLaunchKernelGGL(spmv_masked<2>, dim3(nc), dim3(RED_BLOCK), 0, ctx->stream, sg, ctx->bptr, ctx->bcol,
                       ctx->is_fixed, ctx->Ktan, ctx->kz, ctx->kt, ctx->ks, ctx->partials, ctx->kctl);
    hipLaunchKernelGGL(krylov_omega_fin, dim3(1), dim3(FIN_BLOCK), 0, ctx->stream, ctx->partials, sg, ctx->kctl);
    hipLaunchKernelGGL(krylov_x, dim3(nc), dim3(RED_BLOCK), 0, ctx->stream, sg, ctx->kctl, ctx->kz, ctx->ks, ctx->kt,
                       ctx->krhat, ctx->kx, ctx->kr, ctx->partials);
    hipLaunchKernelGGL(krylov_rho_fin, dim3(1), dim3(FIN_BLOCK), 0, ctx->stream, ctx->partials, sg, ctx->kctl, ctx->kall);
    prof_end(ctx, EB_KERNEL_KRYLOV);
    return check_launch(ctx, "krylov_iteration");
}

// the first step of a solve: x = M⁻¹ b for the instances whose chain preconditioner holds every
// coupling (open chains, rings), which ends their solve; a BiCGStab iteration for the others (a
// net's crossing lines). Which instance takes which depends on the instance alone, so that its
// arithmetic does not depend on what else its batch holds.
int direct_step(eb_ctx* ctx) {
    const Segs sg = segs(ctx);
    const unsigned nc = (unsigned)ctx->nchunks;
    prof_begin(ctx, EB_KERNEL_KRYLOV);
    EB_TRY(precond_apply(ctx, ctx->krhat, ctx->kx, 1));
    hipLaunchKernelGGL(krylov_direct_residual, dim3(nc), dim3(RED_BLOCK), 0, ctx->stream, sg, ctx->bptr, ctx->bcol,
                       ctx->is_fixed, ctx->Ktan, ctx->kx, ctx->krhat, ctx->kr, ctx->partials, ctx->kctl);
    hipLaunchKernelGGL(krylov_direct_fin, dim3(1), dim3(FIN_BLOCK), 0, ctx->stream, ctx->partials, sg, ctx->kctl,
                       ctx->kall);
    prof_end(ctx, EB_KERNEL_KRYLOV);
    return check_launch(ctx, "krylov_direct");
}
int first_iteration(eb_ctx* ctx) {
    if (ctx->n_exact > 0) EB_TRY(direct_step(ctx));
    return ctx->n_inexact > 0 ? krylov_iteration(ctx) : EB_OK;
}

// The linear solve in pieces: start (factorisation, BiCGStab initialisation), iterations, a poll
// of the all-instances record (one small copy and one synchronisation), the acceptance check and
// the scatter of the solution into ΔD. Ktan (block-row order) and res must be current.
int solve_start(eb_ctx* ctx, const eb_newton_params* p, bool factor = true) {
    const Segs sg = segs(ctx);
    const double tol = (p && p->krylov_tol > 0) ? p->krylov_tol : 1e-11;
    // the pivot flags are those of the factorisation in use: cleared only when it is redone
    if (factor) {
        hipLaunchKernelGGL(solve_begin, dim3(1), dim3(256), 0, ctx->stream, (int)ctx->S, ctx->pivot_inst, ctx->kall);
        EB_TRY(precond_factor(ctx));
    }
    hipLaunchKernelGGL(krylov_init, dim3((unsigned)ctx->nchunks), dim3(RED_BLOCK), 0, ctx->stream, sg, ctx->is_fixed,
                       ctx->res, ctx->kx, ctx->kr, ctx->krhat, ctx->kp, ctx->kv, ctx->partials);
    hipLaunchKernelGGL(krylov_init_fin, dim3(1), dim3(FIN_BLOCK), 0, ctx->stream, ctx->partials, sg, ctx->kctl, tol,
                       ctx->active, ctx->pivot_inst, ctx->inst_exact, ctx->kall);
    return check_launch(ctx, "krylov_init");
}

int solve_poll(eb_ctx* ctx) {
    EB_HIP(hipMemcpyAsync(ctx->h_kall, ctx->kall, sizeof(KrylovAll), hipMemcpyDeviceToHost, ctx->stream));
    EB_HIP(hipStreamSynchronize(ctx->stream));
    return EB_OK;
}

// iterates (in doubling batches, polling after each) until done; `launched` iterations have run
int solve_finish(eb_ctx* ctx, const eb_newton_params* p, int launched) {
    const int maxit = (p && p->krylov_maxit > 0) ? p->krylov_maxit : 200;
    int batch = std::max(1, launched);
    while (!ctx->h_kall->all_done && launched < maxit) {
        const int m = std::min(batch, maxit - launched);
        for (int k = 0; k < m; ++k) EB_TRY(krylov_iteration(ctx));
        launched += m;
        batch = std::min(batch * 2, 16);
        EB_TRY(solve_poll(ctx));
    }
    return EB_OK;
}

// counts the finished solve (Krylov iterations and solves)
void solve_account(eb_ctx* ctx, int* its_out) {
    const KrylovAll& ka = *ctx->h_kall;
    ctx->counters[1] += ka.max_iters;
    ctx->counters[2] += 1;
    if (its_out) *its_out = ka.max_iters;
}

// the stand-alone solve (eb_solve_free_dofs) fails when any instance's solve failed
int solve_check(eb_ctx* ctx, int* its_out) {
    const KrylovAll& ka = *ctx->h_kall;
    solve_account(ctx, its_out);
    if (ka.pivot_status != 0) { ctx->err = "singular 6x6 pivot block in the chain preconditioner"; return EB_ERR_SOLVE; }
    if (!ka.all_done || ka.fail_inst >= 0) {
        char buf[200];
        if (ka.fail_inst >= 0)
            snprintf(buf, sizeof buf, "BiCGStab did not converge (instance %d): relative residual %.3e%s", ka.fail_inst,
                     ka.fail_rel, ka.breakdown ? " (breakdown)" : "");
        else
            snprintf(buf, sizeof buf, "BiCGStab did not converge in %d iterations", ka.max_iters);
        ctx->err = buf;
        return EB_ERR_SOLVE;
    }
    return EB_OK;
}

// Inside the Newton loops a failed solve ends the Newton iteration of its instance only: the
// corrector records k = max_iterations + 1 for it (pfail null), the predictor flags it in pfail.
int screen_launch(eb_ctx* ctx, int32_t kfail, uint8_t* pfail, const KrylovAll* gate) {
    hipLaunchKernelGGL(solve_screen, dim3(1), dim3(256), 0, ctx->stream, (int)ctx->S, ctx->kctl, ctx->active, ctx->kiter,
                       kfail, pfail, gate);
    return check_launch(ctx, "solve_screen");
}

int scatter_launch(eb_ctx* ctx, const KrylovAll* gate) {
    hipLaunchKernelGGL(scatter_solution, dim3(nblk(ctx->ndofs, 256)), dim3(256), 0, ctx->stream, ctx->ndofs,
                       ctx->is_fixed, ctx->kx, ctx->dD, active_dev(ctx), gate);
    return check_launch(ctx, "scatter_solution");
}

// solves the masked free-dof system of every instance active in the Newton iteration and writes
// their ΔD[free]
int solve_free(eb_ctx* ctx, const eb_newton_params* p, int* its_out, bool factor = true) {
    EB_TRY(solve_start(ctx, p, factor));
    EB_TRY(first_iteration(ctx));
    EB_TRY(solve_poll(ctx));
    EB_TRY(solve_finish(ctx, p, 1));
    EB_TRY(solve_check(ctx, its_out));
    return scatter_launch(ctx, nullptr);
}

// the predictor's solve: instances whose solve fails are flagged in pfail (and leave `active`)
int solve_free_screened(eb_ctx* ctx, const eb_newton_params* p, bool factor) {
    EB_TRY(solve_start(ctx, p, factor));
    EB_TRY(first_iteration(ctx));
    EB_TRY(solve_poll(ctx));
    EB_TRY(solve_finish(ctx, p, 1));
    solve_account(ctx, nullptr);
    EB_TRY(screen_launch(ctx, 0, ctx->pfail, nullptr));
    return scatter_launch(ctx, nullptr);
}

// per-instance norms; in a Newton iteration (nc.newton) also the convergence bookkeeping. Kernels
// only (no synchronisation); behind `gate` they wait for the solve on the device.
int norms_kernels(eb_ctx* ctx, const NewtonCtl& nc, const KrylovAll* gate) {
    const Segs sg = segs(ctx);
    prof_begin(ctx, EB_KERNEL_UPDATE);
    hipLaunchKernelGGL(norms_partials, dim3((unsigned)ctx->nchunks), dim3(RED_BLOCK), 0, ctx->stream, sg, ctx->is_fixed,
                       ctx->dD, ctx->res, ctx->fext, ctx->Tc, ctx->Tk, ctx->partials);
    hipLaunchKernelGGL(norms_finalize, dim3(1), dim3(FIN_BLOCK), 0, ctx->stream, ctx->partials, sg, ctx->norms, nc,
                       ctx->active, ctx->kiter, ctx->nall, gate, ctx->kdone);
    prof_end(ctx, EB_KERNEL_UPDATE);
    return check_launch(ctx, "norms");
}

// norms with the per-instance values copied to the host
int norms_launch(eb_ctx* ctx, const NewtonCtl& nc, int* n_active) {
    EB_TRY(norms_kernels(ctx, nc, nullptr));
    EB_HIP(hipMemcpyAsync(ctx->h_nall, ctx->nall, sizeof(NewtonAll), hipMemcpyDeviceToHost, ctx->stream));
    EB_HIP(hipMemcpyAsync(ctx->h_norms, ctx->norms, 2 * sizeof(double) * ctx->S, hipMemcpyDeviceToHost, ctx->stream));
    EB_HIP(hipStreamSynchronize(ctx->stream));
    if (!ctx->ev_pending.empty()) prof_harvest(ctx);
    if (n_active) *n_active = ctx->h_nall->n_active;
    return EB_OK;
}

// node state and force vectors, n+1 → n (commit) or n → n+1 (rollback), in one launch; with
// `sel` (device, per instance) only the selected instances
int copy_state(eb_ctx* ctx, bool commit, const uint8_t* sel = nullptr) {
    ctx->start_assembled = false;
    CopyList cl{};
    cl.sel = sel;
    int k = 0;
    int64_t nmax = 0;
    auto add = [&](double* cur, double* old, int64_t n, int32_t per) {
        cl.src[k] = commit ? cur : old;
        cl.dst[k] = commit ? old : cur;
        cl.n[k] = n;
        cl.map[k] = ctx->node_inst;
        cl.per[k] = per;
        nmax = std::max(nmax, n);
        ++k;
    };
    for (int f = 0; f < 8; ++f) {
        const int32_t per = f >= 6 ? 9 : 3;
        add(ctx->node[EB_F_U + f], ctx->node[EB_F_U_N + f], per * ctx->N, per);
    }
    add(ctx->Tint, ctx->Tint_n, ctx->ndofs, 6);
    add(ctx->Tk, ctx->Tk_n, ctx->ndofs, 6);
    add(ctx->Tc, ctx->Tc_n, ctx->ndofs, 6);
    add(ctx->fext, ctx->fext_n, ctx->ndofs, 6);
    hipLaunchKernelGGL(copy_many, dim3(std::min<unsigned>(nblk(nmax, 256), 4096), (unsigned)k), dim3(256), 0, ctx->stream, cl);
    return check_launch(ctx, "copy_state");
}

// active = the instances taking part in the step; at the start of a corrector loop (`start`)
// an instance whose predictor solve failed stays frozen with k = kfail
int newton_reset_launch(eb_ctx* ctx, bool start = false, int32_t kfail = 0) {
    hipLaunchKernelGGL(newton_reset, dim3(nblk(ctx->S, 256)), dim3(256), 0, ctx->stream, (int)ctx->S, ctx->part,
                       start ? ctx->pfail : nullptr, ctx->active, ctx->kiter, ctx->kdone, kfail);
    return check_launch(ctx, "newton_reset");
}

// Which instances take part in the next step (all when `part` is null); instances that do not
// are frozen by every kernel of the predictor, the corrector and the step's commit.
int set_participation(eb_ctx* ctx, const uint8_t* part) {
    bool all = true, same = true;
    for (int64_t s = 0; s < ctx->S; ++s) {
        const uint8_t v = part ? (part[s] ? 1 : 0) : 1;
        all = all && v;
        same = same && ctx->h_part[s] == v;
    }
    if (same) return EB_OK;
    EB_HIP(hipStreamSynchronize(ctx->stream));
    for (int64_t s = 0; s < ctx->S; ++s) ctx->h_part[s] = part ? (part[s] ? 1 : 0) : 1;
    ctx->all_part = all;
    EB_HIP(hipMemcpyAsync(ctx->part, ctx->h_part, ctx->S, hipMemcpyHostToDevice, ctx->stream));
    return newton_reset_launch(ctx);
}

int global_corrector_async(eb_ctx* ctx, const KrylovAll* gate = nullptr) {
    ctx->start_assembled = false;
    hipLaunchKernelGGL(global_corrector, dim3(nblk(ctx->ndofs, 256)), dim3(256), 0, ctx->stream, ctx->ndofs,
                       (const StepCoef*)ctx->coef, ctx->D, ctx->Dd, ctx->Ddd, ctx->dD, active_dev(ctx), gate);
    return check_launch(ctx, "global_corrector");
}
int local_corrector_async(eb_ctx* ctx, const KrylovAll* gate = nullptr) {
    ctx->start_assembled = false;
    hipLaunchKernelGGL(local_corrector, dim3(nblk(ctx->N, 256)), dim3(256), 0, ctx->stream, ctx->N, node_fields(ctx),
                       ctx->D, ctx->Dd, ctx->Ddd, ctx->dD, (const StepCoef*)ctx->coef, active_dev(ctx), gate);
    return check_launch(ctx, "local_corrector");
}
int global_predictor_async(eb_ctx* ctx) {
    ctx->start_assembled = false;
    hipLaunchKernelGGL(global_predictor, dim3(nblk(ctx->ndofs, 256)), dim3(256), 0, ctx->stream, ctx->ndofs,
                       (const StepCoef*)ctx->coef, ctx->D, ctx->Dd, ctx->Ddd, ctx->dD, active_dev(ctx));
    return check_launch(ctx, "global_predictor");
}
int local_predictor_async(eb_ctx* ctx) {
    ctx->start_assembled = false;
    hipLaunchKernelGGL(local_predictor, dim3(nblk(ctx->N, 256)), dim3(256), 0, ctx->stream, ctx->N, node_fields(ctx),
                       ctx->D, ctx->Dd, ctx->Ddd, (const StepCoef*)ctx->coef, active_dev(ctx));
    return check_launch(ctx, "local_predictor");
}
int init_global_launch(eb_ctx* ctx, int predictor) {
    hipLaunchKernelGGL(init_global, dim3(nblk(ctx->N, 256)), dim3(256), 0, ctx->stream, ctx->N, node_fields(ctx), ctx->D,
                       ctx->Dd, ctx->Ddd, ctx->dD, predictor, active_dev(ctx));
    return check_launch(ctx, "init_global");
}

// ---- one Newton iteration ------------------------------------------------------------------

// The launches of one corrector iteration (src/corrector.jl:18-51): the assembly with contact and
// the tangent (or, on the first iteration of a start already assembled, its residual only), the
// boundary-condition lift, the factorisation and one BiCGStab iteration, then the update and the
// norms behind the solve's all-done flag, and the copies of the two status records. Nothing here
// synchronises: the caller waits on the stream.
int newton_body(eb_ctx* ctx, const eb_newton_params* p, double dt, bool start_done) {
    const NewtonCtl nc{p->tol_res, p->tol_disp, 0, p->max_iterations, 1};
    if (start_done) EB_TRY(residual_of_staging(ctx, p->alpha, dt, p->beta, p->gamma));
    else EB_TRY(assemble_all(ctx, true, p->alpha, true, dt, p->beta, p->gamma));
    ctx->start_assembled = false;
    EB_TRY(bc_lift_launch(ctx));
    EB_TRY(solve_start(ctx, p));
    EB_TRY(first_iteration(ctx));
    EB_TRY(screen_launch(ctx, p->max_iterations + 1, nullptr, ctx->kall));
    EB_TRY(scatter_launch(ctx, ctx->kall));
    EB_TRY(global_corrector_async(ctx, ctx->kall));
    EB_TRY(local_corrector_async(ctx, ctx->kall));
    EB_TRY(norms_kernels(ctx, nc, ctx->kall));
    EB_HIP(hipMemcpyAsync(ctx->h_nall, ctx->nall, sizeof(NewtonAll), hipMemcpyDeviceToHost, ctx->stream));
    EB_HIP(hipMemcpyAsync(ctx->h_kall, ctx->kall, sizeof(KrylovAll), hipMemcpyDeviceToHost, ctx->stream));
    return EB_OK;
}

// What the host side of newton_body(start_done = false) records about the device buffers; a
// graph replay applies it without running the body.
struct BodyNotes {
    bool start_assembled, kcm_pending, np1_alias;
    uint64_t stage_ver;
    eb_ctx::TanTag ktan_tag, factor_tag;
    int64_t beam_launches;
    bool operator==(const BodyNotes& o) const {
        return start_assembled == o.start_assembled && kcm_pending == o.kcm_pending && np1_alias == o.np1_alias &&
               stage_ver == o.stage_ver && ktan_tag == o.ktan_tag && factor_tag == o.factor_tag &&
               beam_launches == o.beam_launches;
    }
};
BodyNotes body_notes(const eb_ctx* ctx) {
    return BodyNotes{ctx->start_assembled, ctx->kcm_pending, ctx->np1_alias, ctx->stage_ver,
                     ctx->ktan_tag,        ctx->factor_tag,  ctx->counters[3]};
}
void set_body_notes(eb_ctx* ctx, const BodyNotes& n) {
    ctx->start_assembled = n.start_assembled;
    ctx->kcm_pending = n.kcm_pending;
    ctx->np1_alias = n.np1_alias;
    ctx->stage_ver = n.stage_ver;
    ctx->ktan_tag = n.ktan_tag;
    ctx->factor_tag = n.factor_tag;
    ctx->counters[3] = n.beam_launches;
}
void apply_body_notes(eb_ctx* ctx, const eb_newton_params* p, double dt) {
    ctx->start_assembled = false;
    ctx->stage_ver++;
    ctx->counters[3]++;
    ctx->kcm_pending = true;
    ctx->ktan_tag = eb_ctx::TanTag{ctx->stage_ver, 1 + p->alpha, ctx->coef_ver};
    ctx->np1_alias = false;
    ctx->factor_tag = ctx->ktan_tag;
}

void graph_release(eb_ctx* ctx) {
    if (ctx->graph.exec) (void)hipGraphExecDestroy(ctx->graph.exec);
    ctx->graph.exec = nullptr;
}

// newton_body(start_done = false) as a replay of the captured graph, captured first when there is
// none for these parameters. Every buffer the body touches is fixed for the context's lifetime
// (K, C, M, whose buffers are exchanged at commits, are not passed to it).
int graph_iteration(eb_ctx* ctx, const eb_newton_params* p, double dt) {
    eb_ctx::NewtonGraph& g = ctx->graph;
    // Δt enters through the per-instance coefficient table in device memory, not through
    // kernel arguments: a captured iteration serves every Δt
    (void)dt;
    const double key[9] = {0.0, p->alpha, p->beta, p->gamma, p->tol_res, p->tol_disp, (double)p->max_iterations,
                           p->krylov_tol, (double)ctx->prof_mask};
    bool captured = false;
    BodyNotes after_capture{};
    const BodyNotes before = body_notes(ctx);
    if (!g.exec || std::memcmp(key, g.key, sizeof key) != 0) {
        graph_release(ctx);
        for (int id = 0; id < EB_KERNEL_COUNT; ++id) {
            g.timed[id] = false;
            if (!(ctx->prof_mask & (1u << id))) continue;
            for (hipEvent_t& e : g.ev[id])
                if (!e) EB_HIP(hipEventCreate(&e));
        }
        EB_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        ctx->capturing = true;
        const int st = newton_body(ctx, p, dt, false);
        ctx->capturing = false;
        hipGraph_t graph = nullptr;
        const hipError_t e = hipStreamEndCapture(ctx->stream, &graph);
        after_capture = body_notes(ctx);
        set_body_notes(ctx, before);
        if (st != EB_OK || e != hipSuccess) {
            if (graph) (void)hipGraphDestroy(graph);
            if (st == EB_OK) ctx->err = std::string("hipStreamEndCapture: ") + hipGetErrorString(e);
            return st != EB_OK ? st : EB_ERR_DEVICE;
        }
        const hipError_t ei = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (ei != hipSuccess) {
            g.exec = nullptr;
            ctx->err = std::string("hipGraphInstantiate: ") + hipGetErrorString(ei);
            return EB_ERR_DEVICE;
        }
        std::memcpy(g.key, key, sizeof key);
        ctx->graph_stats[0]++;
        captured = true;
    }
    EB_HIP(hipGraphLaunch(g.exec, ctx->stream));
    ctx->graph_stats[1]++;
    apply_body_notes(ctx, p, dt);
    if (captured && !(body_notes(ctx) == after_capture)) {
        ctx->err = "captured Newton iteration: host bookkeeping differs from the replay's";
        return EB_ERR_STATE;
    }
    return EB_OK;
}

// ---- launch-shape autotuning ----------------------------------------------------------------

// average milliseconds per launch of `launch` over `reps` launches after an untimed one
template <typename F> int time_launches(eb_ctx* ctx, int reps, F&& launch, double* ms_out) {
    hipEvent_t a = pool_event(ctx), b = pool_event(ctx);
    if (!a || !b) { ctx->err = "autotune: hipEventCreate failed"; return EB_ERR_DEVICE; }
    int st = launch();
    float ms = 0;
    if (st == EB_OK && hipEventRecord(a, ctx->stream) != hipSuccess) st = EB_ERR_DEVICE;
    for (int r = 0; r < reps && st == EB_OK; ++r) st = launch();
    if (st == EB_OK && (hipEventRecord(b, ctx->stream) != hipSuccess || hipEventSynchronize(b) != hipSuccess ||
                        hipEventElapsedTime(&ms, a, b) != hipSuccess))
        st = EB_ERR_DEVICE;
    ctx->ev_pool.push_back(a);
    ctx->ev_pool.push_back(b);
    if (st == EB_ERR_DEVICE && ctx->err.empty()) ctx->err = "autotune: event timing failed";
    *ms_out = ms / reps;
    return st;
}

// Times the candidate launch shapes of the chain factorisation (LDS budget of the reduced
// chain), the chain solve (work vectors in LDS or in global memory) and the matrix gather
// (workgroup size) on the context's own sizes and its current tangent, and keeps the fastest.
// The results of the timed launches are discarded: the gathers write a scratch buffer, the solve
// a Krylov work vector, and the factors are marked stale.
int autotune_impl(eb_ctx* ctx, int32_t* choice, double* ms_out) {
    if (ctx->ktan_tag.ver == 0) { ctx->err = "eb_autotune needs a tangent: call it after a time step"; return EB_ERR_STATE; }
    const int reps = 5;
    static const int32_t cand_lds[EB_TUNE_CANDIDATES] = {16, 32, 48, 96};
    static const int32_t cand_solve[EB_TUNE_CANDIDATES] = {1, 2, 0, 0};
    static const int32_t cand_gather[EB_TUNE_CANDIDATES] = {64, 128, 256, 0};
    static const int32_t cand_algo[EB_TUNE_CANDIDATES] = {EB_CHAIN_CYCLIC_REDUCTION, EB_CHAIN_BLOCK_LU, 0, 0};
    const int32_t* cands[EB_TUNE_COUNT] = {cand_lds, cand_solve, cand_gather, cand_algo};
    for (int i = 0; i < EB_TUNE_COUNT * EB_TUNE_CANDIDATES; ++i) ms_out[i] = -1;
    EB_TRY(materialize_kcm(ctx));   // the gather timing reads the staging; K, C, M stay as they are
    graph_release(ctx);
    int32_t saved[EB_TUNE_COUNT];
    for (int id = 0; id < EB_TUNE_COUNT; ++id) saved[id] = ctx->tune[id];
    // the chain solve skips the chains of finished solves, and after a step every solve is
    // finished: the solve is timed with every instance's control record cleared (not done), and
    // the records are put back afterwards
    std::vector<KrylovCtl> kc_saved((size_t)ctx->S);
    EB_HIP(hipMemcpyAsync(kc_saved.data(), ctx->kctl, sizeof(KrylovCtl) * ctx->S, hipMemcpyDeviceToHost, ctx->stream));
    EB_HIP(hipStreamSynchronize(ctx->stream));
    int st = EB_OK;
    for (int id = 0; id < EB_TUNE_COUNT && st == EB_OK; ++id) {
        double best = 1e30;
        int32_t pick = TUNE_DEFAULT[id];
        for (int c = 0; c < EB_TUNE_CANDIDATES && st == EB_OK; ++c) {
            const int32_t v = cands[id][c];
            if (v == 0) continue;
            if (id == EB_TUNE_CHAIN_SOLVE_LDS && v == 1 && chain_solve_lds_bytes(ctx->chains.max_len, ctx->solve_lds_max) == 0) continue;
            if (id == EB_TUNE_CHAIN_LDS_KB &&
                chain_lds_bytes(ctx->chains.max_len, chain_lds_level(ctx->chains.max_len, 1024 * (int64_t)v)) >
                    (size_t)ctx->chain_lds_max)
                continue;
            ctx->tune[id] = v;
            double ms = 0;
            if (id == EB_TUNE_CHAIN_LDS_KB || id == EB_TUNE_CHAIN_SOLVE_LDS) {
                // shapes of the cyclic-reduction kernels: timed with that algorithm
                ctx->tune[EB_TUNE_CHAIN_ALGO] = EB_CHAIN_CYCLIC_REDUCTION;
            }
            if (id == EB_TUNE_CHAIN_LDS_KB) {
                st = time_launches(ctx, reps, [&] { return precond_factor(ctx); }, &ms);
            } else if (id == EB_TUNE_CHAIN_ALGO) {
                // a factorisation and one application, the pair a Newton iteration's solve needs
                EB_HIP(hipMemsetAsync(ctx->kctl, 0, sizeof(KrylovCtl) * ctx->S, ctx->stream));
                st = time_launches(ctx, reps, [&] {
                    EB_TRY(precond_factor(ctx));
                    return precond_apply(ctx, ctx->res, ctx->ky);
                }, &ms);
            } else if (id == EB_TUNE_CHAIN_SOLVE_LDS) {
                EB_HIP(hipMemsetAsync(ctx->kctl, 0, sizeof(KrylovCtl) * ctx->S, ctx->stream));
                st = time_launches(ctx, reps, [&] { return precond_apply(ctx, ctx->res, ctx->ky); }, &ms);
            } else {
                const int bs = v;
                const eb_ctx::TanTag t = ctx->ktan_tag;
                st = time_launches(ctx, reps, [&] {
                    if (ctx->precision == EB_PRECISION_FP32)
                        hipLaunchKernelGGL(gather_matrices<float>, dim3(nblk(ctx->nnz / 4, bs)), dim3(bs), 0, ctx->stream, ctx->nnz,
                                           ctx->blk_ptr, ctx->blk_src, (const float*)ctx->sK, (const float*)ctx->sC,
                                           (const float*)ctx->sM, nullptr, nullptr, nullptr, ctx->tmp, t.c1,
                                           (const StepCoef*)ctx->coef, 1, 0, active_dev(ctx));
                    else
                        hipLaunchKernelGGL(gather_matrices<double>, dim3(nblk(ctx->nnz / 4, bs)), dim3(bs), 0, ctx->stream, ctx->nnz,
                                           ctx->blk_ptr, ctx->blk_src, (const double*)ctx->sK, (const double*)ctx->sC,
                                           (const double*)ctx->sM, nullptr, nullptr, nullptr, ctx->tmp, t.c1,
                                           (const StepCoef*)ctx->coef, 1, 0, active_dev(ctx));
                    return check_launch(ctx, "gather_matrices");
                }, &ms);
            }
            ms_out[id * EB_TUNE_CANDIDATES + c] = ms;
            if (st == EB_OK && ms < best) { best = ms; pick = v; }
        }
        ctx->tune[id] = pick;
    }
    ctx->factor_tag = eb_ctx::TanTag{};   // the factors were overwritten by the timed launches
    EB_HIP(hipMemcpyAsync(ctx->kctl, kc_saved.data(), sizeof(KrylovCtl) * ctx->S, hipMemcpyHostToDevice, ctx->stream));
    EB_HIP(hipStreamSynchronize(ctx->stream));
    if (st != EB_OK) {
        for (int id = 0; id < EB_TUNE_COUNT; ++id) ctx->tune[id] = saved[id];
        return st;
    }
    for (int id = 0; id < EB_TUNE_COUNT; ++id) choice[id] = ctx->tune[id];
    return EB_OK;
}

// ---- learned warm start ---------------------------------------------------------------------

const int NODE_FIELD_WIDTH[8] = {3, 3, 3, 3, 3, 3, 9, 9};   // U UDOT UDDOT W WDOT WDDOT R DR

// copies the current start (node fields U..DR, D, Ḋ, D̈) into the snapshot, or back for the
// instances flagged in `flag` (all when null)
int snapshot_start(eb_ctx* ctx, bool restore, const uint8_t* flag, double* snap) {
    const int64_t N = ctx->N;
    int64_t off = 0;
    auto cp = [&](double* cur, int per) {
        double* sn = snap + off;
        hipLaunchKernelGGL(copy_node_fields, dim3(nblk(N * per, 256)), dim3(256), 0, ctx->stream, N, per,
                           restore ? (const double*)sn : (const double*)cur, restore ? cur : sn, flag, ctx->node_inst);
        off += per * N;
    };
    for (int f = 0; f < 8; ++f) cp(ctx->node[EB_F_U + f], NODE_FIELD_WIDTH[f]);
    cp(ctx->D, 6);
    cp(ctx->Dd, 6);
    cp(ctx->Ddd, 6);
    return check_launch(ctx, "snapshot_start");
}

// relative residual of every instance at the current start: assembly with contact, the corrector
// residual and the boundary-condition lift, as the first Newton iteration sees them (energy not
// accumulated; with the tangent when `tangent`); per-instance norms into ctx->norms
int start_residual(eb_ctx* ctx, const eb_newton_params* p, double dt, bool tangent) {
    if (tangent) EB_TRY(assemble_all(ctx, true, p->alpha, true, dt, p->beta, p->gamma, false));
    else EB_TRY(assemble_residual(ctx, p->alpha, dt, p->beta, p->gamma));
    EB_TRY(bc_lift_launch(ctx));
    NewtonCtl nc{0, 0, 0, 0, 0};
    return norms_launch(ctx, nc, nullptr);
}

// After the reference predictor: the network's start from step n, and per instance the start with
// the smaller residual. The network's start is evaluated first (residual only), the reference
// predictor's last with its tangent, so that when every instance keeps the reference start the
// first Newton iteration finds its assembly done (start_assembled).
int learned_predictor(eb_ctx* ctx, const eb_newton_params* p) {
    const double dt = 0.0;   // Δt enters through the per-instance step coefficients
    const int64_t N = ctx->N;
    EB_TRY(snapshot_start(ctx, false, nullptr, ctx->snap));
    // back to step n (the predictor starts from it), then the network's increment
    for (int f = 0; f < 8; ++f)
        hipLaunchKernelGGL(copy_node_fields, dim3(nblk(N * NODE_FIELD_WIDTH[f], 256)), dim3(256), 0, ctx->stream, N,
                           NODE_FIELD_WIDTH[f], (const double*)ctx->node[EB_F_U_N + f], ctx->node[EB_F_U + f],
                           (const uint8_t*)nullptr, ctx->node_inst);
    EB_TRY(init_global_launch(ctx, 1));
    hipLaunchKernelGGL(learned_features, dim3(nblk(N, 256)), dim3(256), 0, ctx->stream, N, ctx->pdD, ctx->hist,
                       ctx->newest, ctx->node_inst, ctx->Tc_n, (const StepCoef*)ctx->coef, ctx->lin_scale, ctx->lX);
    EB_TRY(check_launch(ctx, "learned_features"));
    const float* Y = nullptr;
    if (ebm::mlp_infer(ctx->mlp, ctx->lX, N, ctx->stream, ctx->lact.data(), &Y) != EB_OK) {
        ctx->err = "network inference: forward pass launch failed";
        return EB_ERR_DEVICE;
    }
    hipLaunchKernelGGL(learned_apply, dim3(nblk(6 * N, 256)), dim3(256), 0, ctx->stream, N, Y, ctx->pdD,
                       ctx->lout_scale, ctx->is_fixed, ctx->dD);
    EB_TRY(check_launch(ctx, "learned_apply"));
    EB_HIP(hipMemcpyAsync(ctx->ldD, ctx->dD, sizeof(double) * ctx->ndofs, hipMemcpyDeviceToDevice, ctx->stream));
    EB_TRY(global_predictor_async(ctx));
    EB_TRY(local_predictor_async(ctx));
    EB_TRY(start_residual(ctx, p, dt, false));
    EB_HIP(hipMemcpyAsync(ctx->rn_learned, ctx->norms, sizeof(double) * 2 * ctx->S, hipMemcpyDeviceToDevice, ctx->stream));
    EB_TRY(snapshot_start(ctx, false, nullptr, ctx->snap_l));
    // the reference predictor's start, assembled with its tangent
    EB_TRY(snapshot_start(ctx, true, nullptr, ctx->snap));
    EB_TRY(start_residual(ctx, p, dt, true));
    hipLaunchKernelGGL(learned_choose, dim3(1), dim3(256), 0, ctx->stream, (int)ctx->S, ctx->norms, ctx->rn_learned,
                       ctx->eligible, ctx->active, ctx->use_plain, ctx->use_learned, ctx->pred_counts);
    EB_TRY(check_launch(ctx, "learned_choose"));
    EB_HIP(hipMemcpyAsync(ctx->h_pred, ctx->pred_counts, 3 * sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
    EB_HIP(hipStreamSynchronize(ctx->stream));
    if (ctx->h_pred[2] == 0) {
        ctx->start_assembled = true;
        return EB_OK;
    }
    // some instances take the network's start: theirs back, and a full first assembly
    return snapshot_start(ctx, true, ctx->use_learned, ctx->snap_l);
}

int field_info(eb_ctx* ctx, int f, double*& ptr, int64_t& len, int& kind) {
    // kind 0 plain, 1 matrix (block-row storage, CSC at the ABI), 2 gather by free index,
    // 3 gather Ktan by sparsity_free, 4 energy
    kind = 0;
    if (f >= EB_F_U && f <= EB_F_DR_N) {
        bool mat = (f == EB_F_R || f == EB_F_DR || f == EB_F_R_N || f == EB_F_DR_N);
        ptr = ctx->node[f];
        len = (mat ? 9 : 3) * ctx->N;
        return EB_OK;
    }
    len = ctx->ndofs;
    switch (f) {
        case EB_F_FEXT: ptr = ctx->fext; return EB_OK;
        case EB_F_TINT: ptr = ctx->Tint; return EB_OK;
        case EB_F_TK: ptr = ctx->Tk; return EB_OK;
        case EB_F_TC: ptr = ctx->Tc; return EB_OK;
        case EB_F_FEXT_N: ptr = ctx->fext_n; return EB_OK;
        case EB_F_TINT_N: ptr = ctx->Tint_n; return EB_OK;
        case EB_F_TK_N: ptr = ctx->Tk_n; return EB_OK;
        case EB_F_TC_N: ptr = ctx->Tc_n; return EB_OK;
        case EB_F_D: ptr = ctx->D; return EB_OK;
        case EB_F_DDOT: ptr = ctx->Dd; return EB_OK;
        case EB_F_DDDOT: ptr = ctx->Ddd; return EB_OK;
        case EB_F_RES: ptr = ctx->res; return EB_OK;
        case EB_F_DELTA_D: ptr = ctx->dD; return EB_OK;
        default: break;
    }
    kind = 1;
    len = ctx->nnz;
    if (f >= EB_F_K && f <= EB_F_M) EB_TRY(materialize_kcm(ctx));
    switch (f) {
        case EB_F_K: ptr = ctx->np1_alias ? ctx->Kn : ctx->K; return EB_OK;
        case EB_F_C: ptr = ctx->np1_alias ? ctx->Cn : ctx->C; return EB_OK;
        case EB_F_M: ptr = ctx->np1_alias ? ctx->Mn : ctx->M; return EB_OK;
        case EB_F_K_N: ptr = ctx->Kn; return EB_OK;
        case EB_F_C_N: ptr = ctx->Cn; return EB_OK;
        case EB_F_M_N: ptr = ctx->Mn; return EB_OK;
        case EB_F_KTAN: ptr = ctx->Ktan; return EB_OK;
        default: break;
    }
    switch (f) {
        case EB_F_RES_FREE: kind = 2; ptr = ctx->res; len = ctx->n_free; return EB_OK;
        case EB_F_DELTA_D_FREE: kind = 2; ptr = ctx->dD; len = ctx->n_free; return EB_OK;
        case EB_F_KTAN_FREE: kind = 3; ptr = ctx->Ktan; len = ctx->nnz_free; return EB_OK;
        case EB_F_IMPOSED: kind = 0; ptr = ctx->imposed; len = ctx->n_fixed; return EB_OK;
        case EB_F_ENERGY: kind = 4; ptr = ctx->energy; len = 3; return EB_OK;
        default: break;
    }
    ctx->err = "unknown field " + std::to_string(f);
    return EB_ERR_ARG;
}

}  // namespace

// ---- one time step of a small structure in one launch (eb_fused.h) --------------------------

namespace {

bool fused_eligible(const eb_ctx* ctx, std::string* why) {
    if (ctx->S != 1) { *why = "a context of one structure"; return false; }
    if (ctx->E > FUSED_MAX_BEAMS) { *why = "at most " + std::to_string(FUSED_MAX_BEAMS) + " beams"; return false; }
    if (ctx->n_inexact > 0 || ctx->dev2.n_chains > 0) {
        *why = "a chain preconditioner that holds every coupling (open chains, rings)";
        return false;
    }
    if (!chain_lu(ctx)) { *why = "the block-LU chain algorithm"; return false; }
    if (ctx->mlp) { *why = "no learned start installed"; return false; }
    return true;
}

template <typename P> FusedArgs<P> fused_args(eb_ctx* ctx, const eb_newton_params* p, bool reuse) {
    FusedArgs<P> a{};
    a.st = beam_stage<P>(ctx);
    a.sTc = (const double*)ctx->sTc;
    a.E = ctx->E; a.N = ctx->N; a.ndofs = ctx->ndofs; a.nnz = ctx->nnz;
    a.blk_ptr = ctx->blk_ptr; a.blk_src = ctx->blk_src; a.inc_ptr = ctx->inc_ptr; a.inc_src = ctx->inc_src;
    a.Kn = ctx->Kn; a.Cn = ctx->Cn; a.Mn = ctx->Mn; a.Ktan = ctx->Ktan;
    a.coef = (const StepCoef*)ctx->coef;
    a.Tint = ctx->Tint; a.Tk = ctx->Tk; a.Tc = ctx->Tc; a.fext = ctx->fext; a.fext_n = ctx->fext_n; a.Tc_n = ctx->Tc_n;
    a.Tint_n = ctx->Tint_n; a.Tk_n = ctx->Tk_n; a.res = ctx->res; a.energy = ctx->energy; a.partials = ctx->partials;
    a.energy_grid = red_grid(ctx->E);
    a.bptr = ctx->bptr; a.bcol = ctx->bcol; a.fixidx = ctx->fixidx; a.imposed = ctx->imposed; a.is_fixed = ctx->is_fixed;
    a.n_fixed = ctx->n_fixed;
    a.nf = node_fields(ctx);
    a.D = ctx->D; a.Dd = ctx->Dd; a.Ddd = ctx->Ddd; a.dD = ctx->dD; a.pdD = ctx->pdD;
    a.cd = chain_dev(ctx);
    a.cl = closure_dev(ctx);
    a.n_closures = ctx->n_closures;
    a.n_chains = (int64_t)ctx->chains.ptr.size() - 1;
    a.sg = segs(ctx);
    a.kc = ctx->kctl; a.kall = ctx->kall; a.pivot_inst = ctx->pivot_inst; a.inst_exact = ctx->inst_exact;
    a.kx = ctx->kx; a.kr = ctx->kr; a.krhat = ctx->krhat; a.kp = ctx->kp; a.kv = ctx->kv; a.ky = ctx->ky; a.ks = ctx->ks;
    a.kz = ctx->kz; a.kt = ctx->kt;
    a.krylov_tol = p->krylov_tol > 0 ? p->krylov_tol : 1e-11;
    a.krylov_maxit = p->krylov_maxit > 0 ? p->krylov_maxit : 200;
    a.act = active_dev(ctx);
    a.active = ctx->active; a.part = ctx->part; a.pfail = ctx->pfail; a.kiter = ctx->kiter; a.kdone = ctx->kdone;
    a.nall = ctx->nall; a.norms = ctx->norms;
    a.alpha = p->alpha;
    a.nc = NewtonCtl{p->tol_res, p->tol_disp, 0, p->max_iterations, 1};
    a.reuse_factor = reuse ? 1 : 0;
    a.out = ctx->fused_out;
    return a;
}

}  // namespace

static bool predictor_reuses_factor(const eb_ctx* ctx, double a);

// solve_step_dynamics! (src/solve_step_dynamics.jl:1-14) with the predictor as one launch and each
// corrector iteration as the element (and contact) kernels and one launch (eb_fused.h); the host
// bookkeeping is that of eb_predictor + eb_corrector
template <typename P> int step_fused_t(eb_ctx* ctx, const eb_newton_params* p, int* kout) {
    const bool reuse = predictor_reuses_factor(ctx, p->alpha);
    const bool contact = ctx->surface != EB_SURFACE_NONE;
    const FusedArgs<P> a = fused_args<P>(ctx, p, reuse);
    auto launched = [&](hipError_t e, const char* what) {
        if (e == hipSuccess) return EB_OK;
        ctx->err = std::string(what) + ": " + hipGetErrorString(e);
        return EB_ERR_DEVICE;
    };
    EB_TRY(launched(launch_fused_predictor<P>(a, ctx->stream), "fused_predictor"));
    ctx->start_assembled = false;
    if (!reuse) ctx->ktan_tag = eb_ctx::TanTag{};
    ctx->factor_tag = ctx->ktan_tag;
    ctx->counters[2] += 1;
    // corrector iterations launched in batches without waiting (an iteration after the loop's end returns
    // at once, a few µs for its two launches, where a wait and the next batch's launch cost tens of µs):
    // the first batch two more than the last step took, then batches of a quarter of that (at least
    // two); one wait per batch, which also fetches each instance's iteration count
    int64_t launched_its = 0, out[3] = {0, 0, 0};
    int batch = std::max(1, ctx->fused_batch + 2);
    const int more = std::max(2, batch / 4);
    while (launched_its < p->max_iterations) {
        const int m = (int)std::min<int64_t>(batch, p->max_iterations - launched_its);
        for (int j = 0; j < m; ++j) {
            EB_TRY(contact ? launch_beams_contact<P>(ctx, p->alpha) : launch_beams<P>(ctx));
            EB_TRY(launched(launch_fused_iteration<P>(a, contact, ctx->stream), "fused_iteration"));
        }
        launched_its += m;
        EB_HIP(hipMemcpyAsync(ctx->h_nall, ctx->nall, sizeof(NewtonAll), hipMemcpyDeviceToHost, ctx->stream));
        EB_HIP(hipMemcpyAsync(out, ctx->fused_out, sizeof out, hipMemcpyDeviceToHost, ctx->stream));
        EB_HIP(hipMemcpyAsync(ctx->h_kiter, ctx->kiter, sizeof(int32_t) * ctx->S, hipMemcpyDeviceToHost, ctx->stream));
        EB_HIP(hipStreamSynchronize(ctx->stream));
        if (!ctx->ev_pending.empty()) prof_harvest(ctx);
        if (ctx->h_nall->n_active == 0) break;
        batch = more;
    }
    const int64_t iters = out[1];
    ctx->fused_batch = (int)std::max<int64_t>(1, iters);
    // the bookkeeping of the iterations that ran (the element launches past the loop's end wrote nothing)
    if (iters > 0) {
        ctx->kcm_pending = true;
        ctx->np1_alias = false;
        ctx->ktan_tag = eb_ctx::TanTag{ctx->stage_ver, 1 + p->alpha, ctx->coef_ver};
        ctx->factor_tag = ctx->ktan_tag;
    }
    ctx->start_assembled = false;
    ctx->counters[3] -= launched_its - iters;
    ctx->counters[0] += iters;
    ctx->counters[1] += out[0] + out[2];
    ctx->counters[2] += iters;
    const int k = (int)(1 + iters);
    // h_kiter holds the counts of the last wait (nothing has run since)
    if (launched_its == 0) {
        EB_HIP(hipMemcpyAsync(ctx->h_kiter, ctx->kiter, sizeof(int32_t) * ctx->S, hipMemcpyDeviceToHost, ctx->stream));
        EB_HIP(hipStreamSynchronize(ctx->stream));
    }
    // as after corrector_impl: every instance taking part active again, the predictor's failure flags spent
    EB_TRY(newton_reset_launch(ctx));
    EB_HIP(hipMemsetAsync(ctx->pfail, 0, ctx->S, ctx->stream));
    int kmax = 0;
    for (int64_t s = 0; s < ctx->S; ++s) {
        ctx->k_inst[s] = !ctx->h_part[s] ? 0 : (ctx->h_kiter[s] > 0 ? ctx->h_kiter[s] : k);
        kmax = std::max(kmax, ctx->k_inst[s]);
    }
    *kout = kmax;
    return EB_OK;
}

int step_fused(eb_ctx* ctx, const eb_newton_params* p, double dt, int* kout) {
    std::string why;
    if (!fused_eligible(ctx, &why)) {   // e.g. a chain algorithm set after the mode
        ctx->err = "fused execution: " + why;
        return EB_ERR_STATE;
    }
    EB_TRY(set_participation(ctx, nullptr));
    EB_TRY(set_coefs(ctx, nullptr, dt, p->beta, p->gamma));
    return ctx->precision == EB_PRECISION_FP32 ? step_fused_t<float>(ctx, p, kout) : step_fused_t<double>(ctx, p, kout);
}

// =============================================================================================
namespace {

bool desc_ok(const eb_model_desc* d) {
    if (d->abi_version != EB_ABI_VERSION) return false;
    if (d->n_nodes <= 0 || d->n_beams <= 0 || !d->X0 || !d->conn || !d->l0 || !d->Re0 || !d->props) return false;
    // the reference's Gauss rule has at most three points (ωᴳ_beams is a Vec3)
    if (d->n_gauss <= 0 || d->n_gauss > 3 || !d->gauss_w || !d->gauss_z) return false;
    if (d->precision != EB_PRECISION_FP64 && d->precision != EB_PRECISION_FP32) return false;
    if (d->n_fixed < 0 || (d->n_fixed > 0 && !d->fixed_dofs)) return false;
    return true;
}

// settings every instance of a batch shares
bool same_settings(const eb_model_desc* a, const eb_model_desc* b) {
    if (a->precision != b->precision || a->n_gauss != b->n_gauss || a->surface != b->surface) return false;
    for (int k = 0; k < a->n_gauss; ++k)
        if (a->gauss_w[k] != b->gauss_w[k] || a->gauss_z[k] != b->gauss_z[k]) return false;
    if (a->surface != EB_SURFACE_NONE && (a->kn != b->kn || a->mu != b->mu || a->eps_t != b->eps_t || a->eta_n != b->eta_n))
        return false;
    return true;
}

// one model made of the instances one after the other (node, beam and dof numbers shifted)
struct Concat {
    std::vector<double> X0, l0, Re0, props;
    std::vector<int64_t> conn, fixed, cbeams;
    eb_model_desc d{};
};

void concat_descs(const eb_model_desc* ds, int64_t S, Concat& c, std::vector<int64_t>& noff, std::vector<int64_t>& boff) {
    noff.assign(1, 0);
    boff.assign(1, 0);
    for (int64_t s = 0; s < S; ++s) {
        const eb_model_desc& d = ds[s];
        const int64_t n0 = noff.back(), b0 = boff.back();
        c.X0.insert(c.X0.end(), d.X0, d.X0 + 3 * d.n_nodes);
        for (int64_t k = 0; k < 2 * d.n_beams; ++k) c.conn.push_back(d.conn[k] + n0);
        c.l0.insert(c.l0.end(), d.l0, d.l0 + d.n_beams);
        c.Re0.insert(c.Re0.end(), d.Re0, d.Re0 + 9 * d.n_beams);
        c.props.insert(c.props.end(), d.props, d.props + EB_NPROPS * d.n_beams);
        for (int64_t k = 0; k < d.n_fixed; ++k) c.fixed.push_back(d.fixed_dofs[k] + 6 * n0);
        if (d.surface != EB_SURFACE_NONE)
            for (int64_t k = 0; k < d.n_contact_beams; ++k) c.cbeams.push_back(d.contact_beams[k] + b0);
        noff.push_back(n0 + d.n_nodes);
        boff.push_back(b0 + d.n_beams);
    }
    c.d = ds[0];
    c.d.n_nodes = noff.back();
    c.d.n_beams = boff.back();
    c.d.X0 = c.X0.data();
    c.d.conn = c.conn.data();
    c.d.l0 = c.l0.data();
    c.d.Re0 = c.Re0.data();
    c.d.props = c.props.data();
    c.d.n_fixed = (int64_t)c.fixed.size();
    c.d.fixed_dofs = c.fixed.data();
    c.d.n_contact_beams = (int64_t)c.cbeams.size();
    c.d.contact_beams = c.cbeams.data();
}

// The instances' volumes, prefiltered on the device (or taken as given when the descriptor
// holds prefiltered coefficients), stored in the storage type P. One pair of FP64 scratch
// buffers, sized for the largest volume, serves every instance.
template <typename P> int upload_sdfs(eb_ctx* ctx, const eb_model_desc* ds, int64_t S) {
    std::vector<SDFDev<P>> h((size_t)S);
    int64_t max_samples = 0, max_coef = 0;
    for (int64_t s = 0; s < S; ++s) {
        const eb_model_desc* d = ds + s;
        SDFDev<P>& f = h[s];
        if (!d->sdf_samples || d->sdf_n[0] < 1 || d->sdf_n[1] < 1 || d->sdf_n[2] < 1) {
            ctx->err = "instance " + std::to_string(s) + ": signed distance field missing";
            return EB_ERR_ARG;
        }
        for (int a = 0; a < 3; ++a) {
            f.n[a] = (int32_t)d->sdf_n[a]; f.d[a] = d->sdf_d[a]; f.x0[a] = d->sdf_x0[a];
            f.dom[2 * a] = d->sdf_x0[a];
            f.dom[2 * a + 1] = d->sdf_x0[a] + (d->sdf_n[a] - 1) * d->sdf_d[a];
        }
        max_samples = std::max(max_samples, d->sdf_n[0] * d->sdf_n[1] * d->sdf_n[2]);
        max_coef = std::max(max_coef, (d->sdf_n[0] + 2) * (d->sdf_n[1] + 2) * (d->sdf_n[2] + 2));
    }
    double *d_samples = nullptr, *d_coef = nullptr, *d_tri = nullptr;
    auto release = [&] {
        for (double* q : {d_samples, d_coef, d_tri})
            if (q) (void)hipFree(q);
    };
    // LU factors of every axis length in use, concatenated: [cp | inv] per length
    std::vector<int64_t> lens;
    for (int64_t s = 0; s < S; ++s)
        for (int a = 0; a < 3; ++a) lens.push_back(ds[s].sdf_n[a]);
    std::sort(lens.begin(), lens.end());
    lens.erase(std::unique(lens.begin(), lens.end()), lens.end());
    std::vector<double> tri;
    std::vector<int64_t> tri_off;
    for (int64_t n : lens) {
        std::vector<double> cp, inv;
        ebh::tri_factors(n, cp, inv);
        tri_off.push_back((int64_t)tri.size());
        tri.insert(tri.end(), cp.begin(), cp.end());
        tri.insert(tri.end(), inv.begin(), inv.end());
    }
    if (hipMalloc(&d_samples, sizeof(double) * max_samples) != hipSuccess ||
        hipMalloc(&d_coef, sizeof(double) * max_coef) != hipSuccess ||
        hipMalloc(&d_tri, sizeof(double) * tri.size()) != hipSuccess ||
        hipMemcpy(d_tri, tri.data(), sizeof(double) * tri.size(), hipMemcpyHostToDevice) != hipSuccess) {
        release();
        ctx->err = "SDF prefilter: device scratch allocation failed";
        return EB_ERR_DEVICE;
    }
    auto lu = [&](int64_t n) {
        const size_t k = (size_t)(std::lower_bound(lens.begin(), lens.end(), n) - lens.begin());
        return TriLU{d_tri + tri_off[k], d_tri + tri_off[k] + n, n};
    };
    int st = EB_OK;
    for (int64_t s = 0; s < S && st == EB_OK; ++s) {
        const eb_model_desc* d = ds + s;
        const SdfGrid g{d->sdf_n[0], d->sdf_n[1], d->sdf_n[2]};
        const int64_t P0 = g.n0 + 2, P1 = g.n1 + 2, ncoef = P0 * P1 * (g.n2 + 2);
        if (d->sdf_prefiltered) {
            if (hipMemcpyAsync(d_coef, d->sdf_samples, sizeof(double) * ncoef, hipMemcpyHostToDevice, ctx->stream) != hipSuccess)
                st = EB_ERR_DEVICE;
        } else {
            if (hipMemcpyAsync(d_samples, d->sdf_samples, sizeof(double) * g.n0 * g.n1 * g.n2, hipMemcpyHostToDevice,
                               ctx->stream) != hipSuccess)
                st = EB_ERR_DEVICE;
            if (st == EB_OK) {
                hipLaunchKernelGGL(sdf_prefilter_x, dim3(nblk(g.n1 * g.n2, 256)), dim3(256), 0, ctx->stream, g, lu(g.n0),
                                   (const double*)d_samples, d_coef);
                hipLaunchKernelGGL(sdf_prefilter_y, dim3(nblk(P0 * g.n2, 256)), dim3(256), 0, ctx->stream, g, lu(g.n1), d_coef);
                hipLaunchKernelGGL(sdf_prefilter_z, dim3(nblk(P0 * P1, 256)), dim3(256), 0, ctx->stream, g, lu(g.n2), d_coef);
                st = check_launch(ctx, "sdf_prefilter");
            }
        }
        P* dp = nullptr;
        if (st == EB_OK) st = dalloc(ctx, dp, ncoef);
        if (st == EB_OK) {
            hipLaunchKernelGGL(sdf_to_storage<P>, dim3(std::min<unsigned>(nblk(ncoef, 256), 4096)), dim3(256), 0, ctx->stream,
                               ncoef, (const double*)d_coef, dp);
            st = check_launch(ctx, "sdf_to_storage");
            h[s].coef = dp;
        }
        // coarse lower bounds of the stored coefficients (the contact detection's early out)
        P* cm = nullptr;
        for (int a = 0; a < 3; ++a) h[s].cn[a] = (int32_t)((d->sdf_n[a] + SDF_BLOCK - 1) / SDF_BLOCK);
        const int64_t ncm = (int64_t)h[s].cn[0] * h[s].cn[1] * h[s].cn[2];
        if (st == EB_OK) st = dalloc(ctx, cm, ncm);
        if (st == EB_OK) {
            hipLaunchKernelGGL(sdf_block_min<P>, dim3(nblk(ncm, 256)), dim3(256), 0, ctx->stream, g, SDF_BLOCK,
                               (const P*)dp, h[s].cn[0], h[s].cn[1], h[s].cn[2], cm);
            st = check_launch(ctx, "sdf_block_min");
            h[s].cmin = cm;
        }
        // the scratch is reused by the next instance
        if (st == EB_OK && hipStreamSynchronize(ctx->stream) != hipSuccess) st = EB_ERR_DEVICE;
    }
    release();
    if (st != EB_OK) {
        if (ctx->err.empty()) ctx->err = "SDF prefilter on the device failed";
        return st;
    }
    SDFDev<P>* dptr;
    EB_TRY(upload_vec(ctx, dptr, h));
    ctx->d_sdf = dptr;
    return EB_OK;
}

int create_impl(const eb_model_desc* ds, int64_t S, eb_ctx** out) {
    if (!ds || !out || S < 1) return EB_ERR_ARG;
    *out = nullptr;
    for (int64_t s = 0; s < S; ++s) {
        if (!desc_ok(ds + s) || !same_settings(ds, ds + s)) {
            fprintf(stderr, "eb_create: invalid descriptor or settings that differ from instance 0 (instance %lld)\n",
                    (long long)s);
            return EB_ERR_ARG;
        }
    }
    Concat cc;
    std::vector<int64_t> noff, boff;
    concat_descs(ds, S, cc, noff, boff);
    const eb_model_desc* d = &cc.d;
    eb_ctx* ctx = new (std::nothrow) eb_ctx();
    if (!ctx) return EB_ERR_ARG;
    ctx->S = S;
    ctx->node_off = noff;
    ctx->beam_off = boff;
    auto bail = [&](int code) { eb_destroy(ctx); return code; };
    if (hipGetDevice(&ctx->device) != hipSuccess) return bail(EB_ERR_DEVICE);
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) return bail(EB_ERR_DEVICE);
    if (scratch_setup(ctx) != EB_OK) {
        fprintf(stderr, "eb_create: %s\n", ctx->err.c_str());
        return bail(EB_ERR_DEVICE);
    }
    {
        // dynamic LDS for the chain factorisation: what a workgroup may hold beside the kernel's
        // static LDS, asked for explicitly beyond the 64 KB every device grants
        int max_lds = 0;
        hipFuncAttributes fa;
        if (hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, ctx->device) != hipSuccess ||
            hipFuncGetAttributes(&fa, (const void*)chain_factor) != hipSuccess)
            return bail(EB_ERR_DEVICE);
        const int64_t dflt = 65536 - (int64_t)fa.sharedSizeBytes;
        ctx->chain_lds_max = std::max<int64_t>(dflt, (int64_t)max_lds - (int64_t)fa.sharedSizeBytes);
        if (ctx->chain_lds_max > dflt &&
            hipFuncSetAttribute((const void*)chain_factor, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)ctx->chain_lds_max) != hipSuccess) {
            (void)hipGetLastError();
            ctx->chain_lds_max = dflt;
        }
        // the chain solve (no static LDS): every instantiation that takes its work vectors in LDS
        ctx->solve_lds_max = std::max<int64_t>(65536, max_lds);
        const void* solves[4] = {(const void*)chain_solve<true, SOLVE_THREADS_MANY>,
                                 (const void*)chain_solve<true, SOLVE_THREADS_FEW>,
                                 (const void*)chain_solve<true, SOLVE_THREADS_MANY, true>,
                                 (const void*)chain_solve<true, SOLVE_THREADS_FEW, true>};
        for (const void* f : solves)
            if (ctx->solve_lds_max > 65536 &&
                hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)ctx->solve_lds_max) != hipSuccess) {
                (void)hipGetLastError();
                ctx->solve_lds_max = 65536;
            }
    }
    ctx->precision = d->precision;
    std::string err;
    if (!ebh::build_pattern(ctx->pat, d->n_nodes, d->n_beams, d->conn, d->n_fixed, d->fixed_dofs, err)) {
        fprintf(stderr, "eb_create: %s\n", err.c_str());
        return bail(EB_ERR_ARG);
    }
    const ebh::Pattern& P = ctx->pat;
    ctx->N = P.N; ctx->E = P.E; ctx->ndofs = P.ndofs; ctx->nnz = P.nnz; ctx->n_free = P.n_free;
    ctx->n_fixed = d->n_fixed; ctx->nnz_free = P.nnz_free;
    if (ctx->nnz >= (int64_t(1) << 31)) { fprintf(stderr, "eb_create: matrix too large\n"); return bail(EB_ERR_ARG); }
    ctx->gauss.n = d->n_gauss;
    for (int k = 0; k < d->n_gauss; ++k) { ctx->gauss.w[k] = d->gauss_w[k]; ctx->gauss.z[k] = d->gauss_z[k]; }
    int st = EB_OK;
#define CK(x) do { st = (x); if (st != EB_OK) { fprintf(stderr, "eb_create: %s\n", ctx->err.c_str()); return bail(st); } } while (0)
    const int64_t N = ctx->N, E = ctx->E;
    // node state
    for (int f = EB_F_U; f <= EB_F_DR_N; ++f) {
        bool mat = (f == EB_F_R || f == EB_F_DR || f == EB_F_R_N || f == EB_F_DR_N);
        if (mat) {
            std::vector<double> I((size_t)(9 * N), 0.0);
            for (int64_t n = 0; n < N; ++n) { I[9 * n] = 1; I[9 * n + 4] = 1; I[9 * n + 8] = 1; }
            CK(upload_vec(ctx, ctx->node[f], I));
        } else {
            CK(dalloc(ctx, ctx->node[f], 3 * N));
        }
    }
    CK(upload(ctx, ctx->X0, d->X0, 3 * N));
    CK(upload_vec(ctx, ctx->n1, P.n1));
    CK(upload_vec(ctx, ctx->n2, P.n2));
    CK(upload(ctx, ctx->l0, d->l0, E));
    {
        std::vector<double> re((size_t)(9 * E)), pr((size_t)(EB_NPROPS * E));
        for (int64_t e = 0; e < E; ++e) {
            for (int k = 0; k < 9; ++k) re[k * E + e] = d->Re0[9 * e + k];
            for (int k = 0; k < EB_NPROPS; ++k) pr[k * E + e] = d->props[EB_NPROPS * e + k];
        }
        CK(upload_vec(ctx, ctx->Re0c, re));
        CK(upload_vec(ctx, ctx->propsc, pr));
    }
    const size_t esz = ctx->precision == EB_PRECISION_FP32 ? sizeof(float) : sizeof(double);
    {
        double* tmp;
        CK(dalloc(ctx, tmp, (int64_t)((144 * E * esz + 7) / 8))); ctx->sK = tmp;
        CK(dalloc(ctx, tmp, (int64_t)((144 * E * esz + 7) / 8))); ctx->sC = tmp;
        CK(dalloc(ctx, tmp, (int64_t)((144 * E * esz + 7) / 8))); ctx->sM = tmp;
        CK(dalloc(ctx, tmp, 12 * E)); ctx->sTint = tmp;   // forces are FP64 in both forms
        CK(dalloc(ctx, tmp, 12 * E)); ctx->sTk = tmp;
        CK(dalloc(ctx, tmp, 12 * E)); ctx->sTc = tmp;
    }
    CK(dalloc(ctx, ctx->senergy, 3 * E));
    CK(upload_vec(ctx, ctx->bptr, narrow<int32_t>(P.bptr)));
    CK(upload_vec(ctx, ctx->bcol, P.bcol));
    CK(upload_vec(ctx, ctx->blk_ptr, narrow<int32_t>(P.blk_ptr)));
    CK(upload_vec(ctx, ctx->blk_src, P.blk_src));
    CK(upload_vec(ctx, ctx->inc_ptr, narrow<int32_t>(P.inc_ptr)));
    CK(upload_vec(ctx, ctx->inc_src, P.inc_src));
    CK(upload_vec(ctx, ctx->csc2bsr, narrow<int32_t>(P.csc2bsr)));
    CK(upload_vec(ctx, ctx->sfree_bsr, P.sfree_bsr));
    CK(upload_vec(ctx, ctx->freed, P.freed));
    CK(upload_vec(ctx, ctx->fixidx, P.fixidx));
    CK(upload_vec(ctx, ctx->is_fixed, P.is_fixed));
    for (double** m : {&ctx->K, &ctx->C, &ctx->M, &ctx->Kn, &ctx->Cn, &ctx->Mn, &ctx->Ktan, &ctx->tmp}) CK(dalloc(ctx, *m, ctx->nnz));
    for (double** v : {&ctx->fext, &ctx->Tint, &ctx->Tk, &ctx->Tc, &ctx->fext_n, &ctx->Tint_n, &ctx->Tk_n, &ctx->Tc_n,
                       &ctx->D, &ctx->Dd, &ctx->Ddd, &ctx->res, &ctx->dD, &ctx->kx, &ctx->kr, &ctx->krhat, &ctx->kp,
                       &ctx->kv, &ctx->ks, &ctx->kt, &ctx->ky, &ctx->kz})
        CK(dalloc(ctx, *v, ctx->ndofs));
    CK(dalloc(ctx, ctx->imposed, std::max<int64_t>(ctx->n_fixed, 1)));
    CK(dalloc(ctx, ctx->energy, 3));
    // instance tables and per-instance reduction chunks
    {
        std::vector<int32_t> ni((size_t)N), bi((size_t)E), ki;
        for (int64_t s = 0; s < S; ++s) {
            for (int64_t n = noff[s]; n < noff[s + 1]; ++n) ni[n] = (int32_t)s;
            for (int64_t e = boff[s]; e < boff[s + 1]; ++e) bi[e] = (int32_t)s;
        }
        ki.resize((size_t)(ctx->nnz / 36));
        for (int64_t n = 0; n < N; ++n)
            for (int64_t b = P.bptr[n]; b < P.bptr[n + 1]; ++b) ki[b] = ni[n];
        std::vector<int64_t> lo, hi;
        std::vector<int32_t> ci, cp;
        for (int64_t s = 0; s < S; ++s) {
            cp.push_back((int32_t)lo.size());
            for (int64_t c0 = 6 * noff[s]; c0 < 6 * noff[s + 1]; c0 += SEG_CHUNK) {
                lo.push_back(c0);
                hi.push_back(std::min<int64_t>(c0 + SEG_CHUNK, 6 * noff[s + 1]));
                ci.push_back((int32_t)s);
            }
        }
        cp.push_back((int32_t)lo.size());
        ctx->nchunks = (int32_t)lo.size();
        CK(upload_vec(ctx, ctx->node_inst, ni));
        CK(upload_vec(ctx, ctx->beam_inst, bi));
        CK(upload_vec(ctx, ctx->blk_inst, ki));
        CK(upload_vec(ctx, ctx->seg_lo, lo));
        CK(upload_vec(ctx, ctx->seg_hi, hi));
        CK(upload_vec(ctx, ctx->seg_inst, ci));
        CK(upload_vec(ctx, ctx->seg_cptr, cp));
        std::vector<uint8_t> on((size_t)S, 1);
        CK(upload_vec(ctx, ctx->active, on));
        CK(upload_vec(ctx, ctx->part, on));
        CK(dalloc(ctx, ctx->pfail, S));
        CK(dalloc(ctx, ctx->pivot_inst, S));
        CK(dalloc(ctx, ctx->sel, S));
        CK(dalloc(ctx, ctx->coef, S));
        CK(dalloc(ctx, ctx->newest, S));
        CK(dalloc(ctx, ctx->eligible, S));
        ctx->dts.assign((size_t)S, 0.0);
        ctx->k_inst.assign((size_t)S, 0);
        ctx->hcount.assign((size_t)S, 0);
        CK(dalloc(ctx, ctx->kiter, S));
        CK(dalloc(ctx, ctx->kdone, 1));
        CK(dalloc(ctx, ctx->nall, 1));
        CK(dalloc(ctx, ctx->kall, 1));
    }
    CK(dalloc(ctx, ctx->hist, 18 * N));
    CK(dalloc(ctx, ctx->pdD, 6 * N));
    CK(dalloc(ctx, ctx->ldD, 6 * N));
    CK(dalloc(ctx, ctx->snap, 54 * N));
    CK(dalloc(ctx, ctx->snap_l, 54 * N));
    CK(dalloc(ctx, ctx->rn_learned, 2 * S));
    CK(dalloc(ctx, ctx->use_plain, S));
    CK(dalloc(ctx, ctx->use_learned, S));
    CK(dalloc(ctx, ctx->pred_counts, 3));
    CK(dalloc(ctx, ctx->lX, LEARN_IN * N));
    CK(dalloc(ctx, ctx->lT, LEARN_OUT * N));
    {
        std::vector<float> one(LEARN_IN, 1.0f);
        CK(upload_vec(ctx, ctx->ones, one));
        CK(dalloc(ctx, ctx->lin_scale, LEARN_IN));
        CK(dalloc(ctx, ctx->lout_scale, LEARN_OUT));
    }
    CK(dalloc(ctx, ctx->partials, std::max<int64_t>(8 * (int64_t)ctx->nchunks, 8 * RED_MAX_GRID)));
    CK(dalloc(ctx, ctx->kctl, S));
    CK(dalloc(ctx, ctx->kstatus, 1));
    CK(dalloc(ctx, ctx->norms, 2 * S));
    if (hipHostMalloc((void**)&ctx->h_kctl, sizeof(KrylovCtl) * S, hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void**)&ctx->h_norms, 2 * sizeof(double) * S, hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void**)&ctx->h_status, sizeof(int), hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void**)&ctx->h_kall, sizeof(KrylovAll), hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void**)&ctx->h_nall, sizeof(NewtonAll), hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void**)&ctx->h_kiter, sizeof(int32_t) * S, hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void**)&ctx->h_pred, 3 * sizeof(int32_t), hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void**)&ctx->h_coef, sizeof(StepCoef) * S, hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void**)&ctx->h_part, S, hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void**)&ctx->h_sel, S, hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void**)&ctx->h_eligible, S, hipHostMallocDefault) != hipSuccess)
        return bail(EB_ERR_DEVICE);
    for (int k = 0; k < 3; ++k) ctx->h_pred[k] = 0;
    for (int64_t s = 0; s < S; ++s) {
        ctx->h_part[s] = 1;
        ctx->h_sel[s] = 0;
        ctx->h_eligible[s] = 0;
    }
    // contact
    ctx->surface = d->surface;
    if (d->surface != EB_SURFACE_NONE) {
        if (d->surface != EB_SURFACE_SPHERE && d->surface != EB_SURFACE_SDF) return bail(EB_ERR_ARG);
        ctx->kn = d->kn; ctx->mu = d->mu; ctx->eps_t = d->eps_t; ctx->eta_n = d->eta_n;
        std::vector<int32_t> cb((size_t)d->n_contact_beams);
        std::vector<uint8_t> seen((size_t)E, 0);
        for (int64_t k = 0; k < d->n_contact_beams; ++k) {
            int64_t e = d->contact_beams[k] - 1;
            if (e < 0 || e >= E || seen[e]) { fprintf(stderr, "eb_create: invalid or repeated contact beam\n"); return bail(EB_ERR_ARG); }
            seen[e] = 1;
            cb[k] = (int32_t)e;
        }
        ctx->n_contact = d->n_contact_beams;
        CK(upload_vec(ctx, ctx->cbeams, cb));
        CK(upload_vec(ctx, ctx->is_contact, seen));
        CK(dalloc(ctx, ctx->clist, ctx->n_contact));
        CK(dalloc(ctx, ctx->clist_n, 1));
        CK(dalloc(ctx, ctx->contact_stats, 2));
        if (d->surface == EB_SURFACE_SPHERE) {
            std::vector<SphereDev> sp((size_t)S);
            for (int64_t s = 0; s < S; ++s) {
                for (int a = 0; a < 3; ++a) sp[s].c[a] = ds[s].sphere_center[a];
                sp[s].R = ds[s].sphere_radius;
            }
            CK(upload_vec(ctx, ctx->d_sphere, sp));
        } else if (ctx->precision == EB_PRECISION_FP32) {
            CK(upload_sdfs<float>(ctx, ds, S));
        } else {
            CK(upload_sdfs<double>(ctx, ds, S));
        }
    }
    // chains for the preconditioner (a chain never leaves its instance: instances share no node)
    ebh::build_chains(ctx->chains, P, d->X0, 0);
    {
        const ebh::Chains& c = ctx->chains;
        std::vector<int32_t> chi(c.ptr.size() - 1);
        for (size_t k = 0; k + 1 < c.ptr.size(); ++k) {
            const int64_t n0 = c.node[c.ptr[k]];
            chi[k] = (int32_t)(std::upper_bound(noff.begin(), noff.end(), n0) - noff.begin() - 1);
        }
        CK(upload_vec(ctx, ctx->chain_inst, chi));
        CK(upload_vec(ctx, ctx->ch_ptr, narrow<int32_t>(c.ptr)));
        CK(upload_vec(ctx, ctx->ch_node, c.node));
        CK(upload_vec(ctx, ctx->ch_diag, c.diag_blk));
        CK(upload_vec(ctx, ctx->ch_lo, c.lo_blk));
        CK(upload_vec(ctx, ctx->ch_up, c.up_blk));
        CK(upload_vec(ctx, ctx->ch_mD, c.mD));
        CK(upload_vec(ctx, ctx->ch_mL, c.mL));
        CK(upload_vec(ctx, ctx->ch_mU, c.mU));
        CK(upload_vec(ctx, ctx->ch_hbase, c.hbase));
        CK(upload_vec(ctx, ctx->ch_mrow, row_masks(c)));
        CK(dalloc(ctx, ctx->ch_D, 36 * N));
        CK(dalloc(ctx, ctx->ch_L, 36 * N));
        CK(dalloc(ctx, ctx->ch_U, 36 * N));
        CK(dalloc(ctx, ctx->ch_hist, 72 * c.n_hist));
        CK(dalloc(ctx, ctx->ch_bw, 6 * N));
        CK(dalloc(ctx, ctx->ch_yw, 6 * N));
        ctx->n_closures = (int32_t)c.cl_chain.size();
        std::vector<uint8_t> exact((size_t)S, 1);
        if (ebh::chains_left_out(c, P) > 0) {
            // the crossing cover holds the nodes of the instances that have couplings the first
            // cover leaves out (nets), and no node of the others, whose solve is complete without it
            const std::vector<uint8_t> left = ebh::nodes_left_out(c, P);
            std::vector<uint8_t> inst_left((size_t)S, 0), cover((size_t)N, 0);
            for (int64_t s = 0; s < S; ++s)
                for (int64_t n = noff[s]; n < noff[s + 1]; ++n) inst_left[s] |= left[n];
            for (int64_t s = 0; s < S; ++s)
                for (int64_t n = noff[s]; n < noff[s + 1]; ++n) cover[n] = inst_left[s];
            for (int64_t s = 0; s < S; ++s) exact[s] = inst_left[s] ? 0 : 1;
            ebh::build_chains(ctx->chains2, P, d->X0, 0, &c, &cover);
            const ebh::Chains& c2 = ctx->chains2;
            std::vector<int32_t> chi2(c2.ptr.size() - 1);
            for (size_t k = 0; k + 1 < c2.ptr.size(); ++k) {
                const int64_t n0 = c2.node[c2.ptr[k]];
                chi2[k] = (int32_t)(std::upper_bound(noff.begin(), noff.end(), n0) - noff.begin() - 1);
            }
            int32_t *ptr2, *node2, *diag2, *lo2, *up2, *hb2, *inst2;
            uint64_t *mD2, *mL2, *mU2;
            uint32_t* mrow2;
            double *D2, *L2, *U2, *hist2, *bw2, *yw2;
            CK(upload_vec(ctx, inst2, chi2));
            CK(upload_vec(ctx, ptr2, narrow<int32_t>(c2.ptr)));
            CK(upload_vec(ctx, node2, c2.node));
            CK(upload_vec(ctx, diag2, c2.diag_blk));
            CK(upload_vec(ctx, lo2, c2.lo_blk));
            CK(upload_vec(ctx, up2, c2.up_blk));
            CK(upload_vec(ctx, mD2, c2.mD));
            CK(upload_vec(ctx, mL2, c2.mL));
            CK(upload_vec(ctx, mU2, c2.mU));
            CK(upload_vec(ctx, hb2, c2.hbase));
            CK(upload_vec(ctx, mrow2, row_masks(c2)));
            CK(dalloc(ctx, D2, 36 * N));
            CK(dalloc(ctx, L2, 36 * N));
            CK(dalloc(ctx, U2, 36 * N));
            CK(dalloc(ctx, hist2, 72 * c2.n_hist));
            CK(dalloc(ctx, bw2, 6 * N));
            CK(dalloc(ctx, yw2, 6 * N));
            CK(dalloc(ctx, ctx->cross_r, 6 * N));
            ctx->dev2 = ChainDev{ptr2, node2, diag2, lo2, up2, mD2, mL2, mU2, hb2, D2, L2, U2, hist2, bw2, yw2, inst2,
                                 (int64_t)c2.ptr.size() - 1, mrow2};
        }
        if (ctx->n_closures > 0) {
            CK(upload_vec(ctx, ctx->cl_chain, c.cl_chain));
            CK(upload_vec(ctx, ctx->cl_fl, c.cl_fl));
            CK(upload_vec(ctx, ctx->cl_lf, c.cl_lf));
            CK(upload_vec(ctx, ctx->cl_mfl, c.cl_mfl));
            CK(upload_vec(ctx, ctx->cl_mlf, c.cl_mlf));
            CK(upload_vec(ctx, ctx->cl_zoff, c.cl_zoff));
            CK(dalloc(ctx, ctx->cl_Z, c.cl_zoff.back()));
            CK(dalloc(ctx, ctx->cl_H, 144 * (int64_t)ctx->n_closures));
        }
        // per instance: beams, tangent entries and chain positions (both covers), the units of the
        // work counters
        std::vector<int64_t> iw((size_t)(3 * S), 0);
        for (int64_t s = 0; s < S; ++s) {
            iw[3 * s] = boff[s + 1] - boff[s];
            iw[3 * s + 1] = 36 * (int64_t)(P.bptr[noff[s + 1]] - P.bptr[noff[s]]);
        }
        for (const ebh::Chains* cv : {&ctx->chains, &ctx->chains2})
            for (size_t k = 0; k + 1 < cv->ptr.size(); ++k) {
                const int64_t n0 = cv->node[cv->ptr[k]];
                const int64_t s = std::upper_bound(noff.begin(), noff.end(), n0) - noff.begin() - 1;
                iw[3 * s + 2] += cv->ptr[k + 1] - cv->ptr[k];
            }
        CK(upload_vec(ctx, ctx->inst_work, iw));
        CK(upload_vec(ctx, ctx->inst_exact, exact));
        for (int64_t s = 0; s < S; ++s) (exact[s] ? ctx->n_exact : ctx->n_inexact) += 1;
        CK(dalloc(ctx, ctx->work, 4));
        if (hipMemsetAsync(ctx->work, 0, 4 * sizeof(unsigned long long), ctx->stream) != hipSuccess) return bail(EB_ERR_DEVICE);
    }
#undef CK
    if (hipStreamSynchronize(ctx->stream) != hipSuccess) return bail(EB_ERR_DEVICE);
    *out = ctx;
    return EB_OK;
}


}  // namespace

extern "C" {

int eb_create(const eb_model_desc* d, eb_ctx** out) { return create_impl(d, 1, out); }

int eb_create_batch(const eb_model_desc* descs, int64_t n_instances, eb_ctx** out) {
    return create_impl(descs, n_instances, out);
}

int eb_destroy(eb_ctx* ctx) {
    if (!ctx) return EB_OK;
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    for (auto& p : ctx->ev_pending) { ctx->ev_pool.push_back(p.second.first); ctx->ev_pool.push_back(p.second.second); }
    for (hipEvent_t e : ctx->ev_pool) (void)hipEventDestroy(e);
    graph_release(ctx);
    for (auto& pair : ctx->graph.ev)
        for (hipEvent_t e : pair)
            if (e) (void)hipEventDestroy(e);
    for (void* p : ctx->allocs) (void)hipFree(p);
    if (ctx->h_kctl) (void)hipHostFree(ctx->h_kctl);
    if (ctx->h_norms) (void)hipHostFree(ctx->h_norms);
    if (ctx->h_status) (void)hipHostFree(ctx->h_status);
    if (ctx->h_kall) (void)hipHostFree(ctx->h_kall);
    if (ctx->h_nall) (void)hipHostFree(ctx->h_nall);
    if (ctx->h_pred) (void)hipHostFree(ctx->h_pred);
    if (ctx->h_kiter) (void)hipHostFree(ctx->h_kiter);
    for (float* q : ctx->lact) (void)hipFree(q);
    for (void* q : {(void*)ctx->h_coef, (void*)ctx->h_part, (void*)ctx->h_sel, (void*)ctx->h_eligible})
        if (q) (void)hipHostFree(q);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return EB_OK;
}

const char* eb_last_error(const eb_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int eb_get_sizes(const eb_ctx* ctx, eb_sizes* s) {
    if (!ctx || !s) return EB_ERR_ARG;
    s->n_nodes = ctx->N; s->n_beams = ctx->E; s->ndofs = ctx->ndofs; s->nnz = ctx->nnz;
    s->n_free = ctx->n_free; s->n_fixed = ctx->n_fixed; s->nnz_free = ctx->nnz_free;
    return EB_OK;
}

int eb_get_instances(const eb_ctx* ctx, int64_t* n_instances, int64_t* node_offsets, int64_t* beam_offsets) {
    if (!ctx) return EB_ERR_ARG;
    if (n_instances) *n_instances = ctx->S;
    if (node_offsets) for (int64_t s = 0; s <= ctx->S; ++s) node_offsets[s] = ctx->node_off[s];
    if (beam_offsets) for (int64_t s = 0; s <= ctx->S; ++s) beam_offsets[s] = ctx->beam_off[s];
    return EB_OK;
}

int eb_get_instance_iterations(const eb_ctx* ctx, int32_t* k_out) {
    if (!ctx || !k_out) return EB_ERR_ARG;
    for (int64_t s = 0; s < ctx->S; ++s) k_out[s] = ctx->k_inst[s];
    return EB_OK;
}

int eb_get_instance_norms(const eb_ctx* ctx, double* rn_dn) {
    if (!ctx || !rn_dn) return EB_ERR_ARG;
    if (hipMemcpy(rn_dn, ctx->norms, 2 * sizeof(double) * ctx->S, hipMemcpyDeviceToHost) != hipSuccess) return EB_ERR_DEVICE;
    return EB_OK;
}

int eb_get_pattern(const eb_ctx* ctx, int64_t* colptr, int64_t* rowval, int64_t* smap, int64_t* sfree) {
    if (!ctx) return EB_ERR_ARG;
    const ebh::Pattern& P = ctx->pat;
    if (colptr) for (int64_t j = 0; j <= P.ndofs; ++j) colptr[j] = P.colptr[j] + 1;
    if (rowval) for (int64_t k = 0; k < P.nnz; ++k) rowval[k] = P.rowval[k] + 1;
    if (smap) for (int64_t k = 0; k < 144 * P.E; ++k) smap[k] = P.smap_csc[k] + 1;
    if (sfree) for (int64_t k = 0; k < P.nnz_free; ++k) sfree[k] = P.sfree_csc[k] + 1;
    return EB_OK;
}

int eb_get_field(eb_ctx* ctx, int f, double* out, int64_t count) {
    if (!ctx || (!out && count)) return EB_ERR_ARG;
    double* ptr; int64_t len; int kind;
    EB_TRY(field_info(ctx, f, ptr, len, kind));
    if (count != len) { ctx->err = "field size mismatch"; return EB_ERR_ARG; }
    if (len == 0) return EB_OK;
    const double* src = ptr;
    if (kind == 1) {
        hipLaunchKernelGGL(gather_idx32, dim3(nblk(len, 256)), dim3(256), 0, ctx->stream, len, ctx->csc2bsr, ptr, ctx->tmp);
        src = ctx->tmp;
    } else if (kind == 2) {
        hipLaunchKernelGGL(gather_idx32, dim3(nblk(len, 256)), dim3(256), 0, ctx->stream, len, ctx->freed, ptr, ctx->tmp);
        src = ctx->tmp;
    } else if (kind == 3) {
        hipLaunchKernelGGL(gather_idx32, dim3(nblk(len, 256)), dim3(256), 0, ctx->stream, len, ctx->sfree_bsr, ptr, ctx->tmp);
        src = ctx->tmp;
    }
    EB_TRY(check_launch(ctx, "get_field"));
    EB_HIP(hipMemcpyAsync(out, src, sizeof(double) * len, hipMemcpyDeviceToHost, ctx->stream));
    EB_HIP(hipStreamSynchronize(ctx->stream));
    return EB_OK;
}

int eb_set_field(eb_ctx* ctx, int f, const double* in, int64_t count) {
    if (!ctx || (!in && count)) return EB_ERR_ARG;
    ctx->start_assembled = false;
    if (f == EB_F_K || f == EB_F_C || f == EB_F_M) EB_TRY(unalias(ctx));
    ctx->ktan_tag = ctx->factor_tag = eb_ctx::TanTag{};   // matrices set from outside
    ctx->kcm_ver = ctx->kcm_n_ver = 0;
    double* ptr; int64_t len; int kind;
    EB_TRY(field_info(ctx, f, ptr, len, kind));
    if (count != len) { ctx->err = "field size mismatch"; return EB_ERR_ARG; }
    if (len == 0) return EB_OK;
    if (kind == 0 || kind == 4) {
        EB_HIP(hipMemcpyAsync(ptr, in, sizeof(double) * len, hipMemcpyHostToDevice, ctx->stream));
    } else {
        EB_HIP(hipMemcpyAsync(ctx->tmp, in, sizeof(double) * len, hipMemcpyHostToDevice, ctx->stream));
        const int32_t* idx = kind == 1 ? ctx->csc2bsr : (kind == 2 ? ctx->freed : ctx->sfree_bsr);
        hipLaunchKernelGGL(scatter_idx32, dim3(nblk(len, 256)), dim3(256), 0, ctx->stream, len, idx, ctx->tmp, ptr);
        EB_TRY(check_launch(ctx, "set_field"));
    }
    EB_HIP(hipStreamSynchronize(ctx->stream));
    return EB_OK;
}

int eb_assemble(eb_ctx* ctx) {
    if (!ctx) return EB_ERR_ARG;
    EB_TRY(assemble_all(ctx, false, 0.0, false, 1.0, 1.0, 1.0));
    EB_HIP(hipStreamSynchronize(ctx->stream));
    return EB_OK;
}

int eb_assemble_contact(eb_ctx* ctx, double alpha) {
    if (!ctx) return EB_ERR_ARG;
    if (ctx->surface == EB_SURFACE_NONE) return EB_OK;
    // the contact kernel folds into the element staging written by the last eb_assemble
    if (ctx->precision == EB_PRECISION_FP32) {
        EB_TRY(launch_contact<float>(ctx, alpha));
        EB_TRY(launch_gathers<float>(ctx, true, false, 1.0, alpha, 1.0, 1.0, false, 2));
    } else {
        EB_TRY(launch_contact<double>(ctx, alpha));
        EB_TRY(launch_gathers<double>(ctx, true, false, 1.0, alpha, 1.0, 1.0, false, 2));
    }
    EB_HIP(hipStreamSynchronize(ctx->stream));
    return EB_OK;
}

int eb_tangent_residuals_corrector(eb_ctx* ctx, double dt, double a, double b, double g) {
    if (!ctx) return EB_ERR_ARG;
    ctx->start_assembled = false;
    EB_TRY(unalias(ctx));
    EB_TRY(set_coefs(ctx, nullptr, dt, b, g));
    hipLaunchKernelGGL(combine_tangent, dim3(nblk(ctx->nnz, 256)), dim3(256), 0, ctx->stream, ctx->nnz, ctx->K, ctx->C,
                       ctx->M, ctx->Ktan, 1 + a, (const StepCoef*)ctx->coef, ctx->blk_inst);
    ctx->ktan_tag = eb_ctx::TanTag{};
    hipLaunchKernelGGL(residual_corrector, dim3(nblk(ctx->ndofs, 256)), dim3(256), 0, ctx->stream, ctx->ndofs, a,
                       ctx->fext, ctx->Tc, ctx->Tint, ctx->fext_n, ctx->Tc_n, ctx->Tint_n, ctx->Tk, ctx->res);
    EB_TRY(check_launch(ctx, "tangent_residuals_corrector"));
    EB_HIP(hipStreamSynchronize(ctx->stream));
    return EB_OK;
}

// the predictor's tangent and residual with the step coefficients in place; `form_tangent`
// false: Ktan already holds this tangent (see predictor_reuses_factor)
static int tangent_residuals_predictor_async(eb_ctx* ctx, double a, bool form_tangent = true) {
    if (form_tangent) {
        hipLaunchKernelGGL(combine_tangent, dim3(nblk(ctx->nnz, 256)), dim3(256), 0, ctx->stream, ctx->nnz, ctx->Kn,
                           ctx->Cn, ctx->Mn, ctx->Ktan, 1 + a, (const StepCoef*)ctx->coef, ctx->blk_inst);
        ctx->ktan_tag = eb_ctx::TanTag{};
    }
    hipLaunchKernelGGL(residual_predictor, dim3(nblk(ctx->ndofs, 256)), dim3(256), 0, ctx->stream, ctx->ndofs, ctx->bptr,
                       ctx->bcol, a, (const StepCoef*)ctx->coef, ctx->node_inst, ctx->fext, ctx->Tint_n, ctx->Tk_n,
                       ctx->fext_n, ctx->Dd, ctx->Ddd, ctx->Cn, ctx->Mn, ctx->res);
    return check_launch(ctx, "tangent_residuals_predictor");
}

int eb_tangent_residuals_predictor(eb_ctx* ctx, double dt, double a, double b, double g) {
    if (!ctx) return EB_ERR_ARG;
    ctx->start_assembled = false;
    EB_TRY(set_coefs(ctx, nullptr, dt, b, g));
    EB_TRY(tangent_residuals_predictor_async(ctx, a));
    EB_HIP(hipStreamSynchronize(ctx->stream));
    return EB_OK;
}

int eb_apply_boundary_conditions(eb_ctx* ctx) {
    if (!ctx) return EB_ERR_ARG;
    ctx->start_assembled = false;
    EB_TRY(bc_lift_launch(ctx));
    EB_HIP(hipStreamSynchronize(ctx->stream));
    return EB_OK;
}

// r_free and Ktan_free are views of r and Ktan through the free-dof gathers (see eb_get_field)
int eb_extract_free_dofs(eb_ctx* ctx) { return ctx ? EB_OK : EB_ERR_ARG; }

int eb_solve_free_dofs(eb_ctx* ctx, const eb_newton_params* p, int* its) {
    if (!ctx) return EB_ERR_ARG;
    ctx->start_assembled = false;
    EB_TRY(solve_free(ctx, p, its));
    EB_HIP(hipStreamSynchronize(ctx->stream));
    return EB_OK;
}

int eb_initialize_global_corrector(eb_ctx* ctx) {
    if (!ctx) return EB_ERR_ARG;
    ctx->start_assembled = false;
    EB_TRY(init_global_launch(ctx, 0));
    EB_HIP(hipStreamSynchronize(ctx->stream));
    return EB_OK;
}

int eb_initialize_global_predictor(eb_ctx* ctx) {
    if (!ctx) return EB_ERR_ARG;
    ctx->start_assembled = false;
    EB_TRY(init_global_launch(ctx, 1));
    EB_HIP(hipStreamSynchronize(ctx->stream));
    return EB_OK;
}

int eb_compute_global_corrector(eb_ctx* ctx, double dt, double b, double g) {
    if (!ctx) return EB_ERR_ARG;
    EB_TRY(set_coefs(ctx, nullptr, dt, b, g));
    EB_TRY(global_corrector_async(ctx));
    EB_HIP(hipStreamSynchronize(ctx->stream));
    return EB_OK;
}
int eb_compute_global_predictor(eb_ctx* ctx, double dt, double b, double g) {
    if (!ctx) return EB_ERR_ARG;
    EB_TRY(set_coefs(ctx, nullptr, dt, b, g));
    EB_TRY(global_predictor_async(ctx));
    EB_HIP(hipStreamSynchronize(ctx->stream));
    return EB_OK;
}
int eb_compute_local_corrector(eb_ctx* ctx, double dt, double b, double g) {
    if (!ctx) return EB_ERR_ARG;
    EB_TRY(set_coefs(ctx, nullptr, dt, b, g));
    EB_TRY(local_corrector_async(ctx));
    EB_HIP(hipStreamSynchronize(ctx->stream));
    return EB_OK;
}
int eb_compute_local_predictor(eb_ctx* ctx, double dt, double b, double g) {
    if (!ctx) return EB_ERR_ARG;
    EB_TRY(set_coefs(ctx, nullptr, dt, b, g));
    EB_TRY(local_predictor_async(ctx));
    EB_HIP(hipStreamSynchronize(ctx->stream));
    return EB_OK;
}

int eb_compute_norms_corrector(eb_ctx* ctx, double* rn, double* dn) {
    if (!ctx || !rn || !dn) return EB_ERR_ARG;
    NewtonCtl nc{0, 0, 0, 0, 0};
    EB_TRY(norms_launch(ctx, nc, nullptr));
    // the largest over the instances (the one instance's norms for a single structure)
    *rn = 0; *dn = 0;
    for (int64_t s = 0; s < ctx->S; ++s) {
        *rn = std::max(*rn, ctx->h_norms[2 * s]);
        *dn = std::max(*dn, ctx->h_norms[2 * s + 1]);
        if (std::isnan(ctx->h_norms[2 * s])) *rn = ctx->h_norms[2 * s];
        if (std::isnan(ctx->h_norms[2 * s + 1])) *dn = ctx->h_norms[2 * s + 1];
    }
    return EB_OK;
}

int eb_update_loads(eb_ctx* ctx, int64_t n, const int64_t* dofs, const double* values) {
    if (!ctx) return EB_ERR_ARG;
    ctx->start_assembled = false;
    if (n == 0) return EB_OK;
    // loads are few or dense: upload the touched dofs through the scratch buffers
    std::vector<int32_t> idx((size_t)n);
    for (int64_t k = 0; k < n; ++k) {
        if (dofs[k] < 1 || dofs[k] > ctx->ndofs) { ctx->err = "load dof out of range"; return EB_ERR_ARG; }
        idx[k] = (int32_t)(dofs[k] - 1);
    }
    double* vals = ctx->tmp;
    int32_t* didx = (int32_t*)(ctx->kt);  // Krylov scratch is idle between solves
    if (n > ctx->ndofs) { ctx->err = "too many loads"; return EB_ERR_ARG; }
    EB_HIP(hipMemcpyAsync(vals, values, sizeof(double) * n, hipMemcpyHostToDevice, ctx->stream));
    EB_HIP(hipMemcpyAsync(didx, idx.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(scatter_idx32, dim3(nblk(n, 256)), dim3(256), 0, ctx->stream, n, didx, vals, ctx->fext);
    EB_TRY(check_launch(ctx, "update_loads"));
    EB_HIP(hipStreamSynchronize(ctx->stream));
    return EB_OK;
}

int eb_update_boundary_conditions(eb_ctx* ctx, const double* values, int64_t n_fixed) {
    if (!ctx) return EB_ERR_ARG;
    ctx->start_assembled = false;
    if (n_fixed != ctx->n_fixed) { ctx->err = "imposed displacement count mismatch"; return EB_ERR_ARG; }
    if (n_fixed == 0) return EB_OK;
    EB_HIP(hipMemcpyAsync(ctx->imposed, values, sizeof(double) * n_fixed, hipMemcpyHostToDevice, ctx->stream));
    EB_HIP(hipStreamSynchronize(ctx->stream));
    return EB_OK;
}

int eb_initialize_state_simulation(eb_ctx* ctx) {
    if (!ctx) return EB_ERR_ARG;
    ctx->start_assembled = false;
    for (int64_t s = 0; s < ctx->S; ++s) ctx->hcount[s] = 0;
    EB_TRY(assemble_all(ctx, false, 0.0, false, 1.0, 1.0, 1.0));
    size_t b = sizeof(double) * ctx->nnz;
    EB_HIP(hipMemcpyAsync(ctx->Kn, ctx->K, b, hipMemcpyDeviceToDevice, ctx->stream));
    EB_HIP(hipMemcpyAsync(ctx->Cn, ctx->C, b, hipMemcpyDeviceToDevice, ctx->stream));
    EB_HIP(hipMemcpyAsync(ctx->Mn, ctx->M, b, hipMemcpyDeviceToDevice, ctx->stream));
    EB_HIP(hipStreamSynchronize(ctx->stream));
    return EB_OK;
}

// The predictor's tangent from Kⁿ, Cⁿ, Mⁿ equals the tangent of the last corrector iteration when
// both were gathered from the same element staging with the same HHT coefficients (every
// instance's step size unchanged): then Ktan and the chain factors of that iteration are reused
// as they are. Only when every instance took part in that iteration (the factors of the others
// are older) and committed its step (kcm_n_ver is 0 after a step some instances repeated).
static bool predictor_reuses_factor(const eb_ctx* ctx, double a) {
    const eb_ctx::TanTag want{ctx->kcm_n_ver, 1 + a, ctx->coef_ver};
    return ctx->all_part && ctx->kcm_n_ver != 0 && ctx->ktan_tag == want && ctx->factor_tag == want;
}

// uploads which instances may take the network's start (taking part, three steps recorded);
// false when none may
static int learned_eligibility(eb_ctx* ctx, bool* any) {
    *any = false;
    bool same = true;
    for (int64_t s = 0; s < ctx->S; ++s) {
        const uint8_t e = (ctx->h_part[s] && ctx->hcount[s] >= 3) ? 1 : 0;
        *any = *any || e;
        same = same && ctx->h_eligible[s] == e;
    }
    if (same || !*any) return EB_OK;
    EB_HIP(hipStreamSynchronize(ctx->stream));
    for (int64_t s = 0; s < ctx->S; ++s) ctx->h_eligible[s] = (ctx->h_part[s] && ctx->hcount[s] >= 3) ? 1 : 0;
    EB_HIP(hipMemcpyAsync(ctx->eligible, ctx->h_eligible, ctx->S, hipMemcpyHostToDevice, ctx->stream));
    return EB_OK;
}

// predictor! (src/predictor.jl:1-27) for the instances taking part, with their step coefficients
// in place; an instance whose linear solve fails is flagged in pfail (its corrector reports
// k = max_iterations + 1)
static int predictor_impl(eb_ctx* ctx, const eb_newton_params* p) {
    EB_HIP(hipMemsetAsync(ctx->pfail, 0, ctx->S, ctx->stream));
    EB_TRY(newton_reset_launch(ctx));   // active = the instances taking part
    EB_TRY(init_global_launch(ctx, 1));
    const bool reuse = predictor_reuses_factor(ctx, p->alpha);
    EB_TRY(tangent_residuals_predictor_async(ctx, p->alpha, !reuse));
    EB_TRY(bc_lift_launch(ctx));
    EB_TRY(solve_free_screened(ctx, p, !reuse));
    EB_TRY(global_predictor_async(ctx));
    EB_TRY(local_predictor_async(ctx));
    // the reference predictor's increment: an input of the learned start and of its samples
    EB_HIP(hipMemcpyAsync(ctx->pdD, ctx->dD, sizeof(double) * ctx->ndofs, hipMemcpyDeviceToDevice, ctx->stream));
    bool any = false;
    if (ctx->mlp) EB_TRY(learned_eligibility(ctx, &any));
    if (any) {
        if (ctx->learn_wait > 0) {
            ctx->learn_wait--;
            ctx->learn_skipped++;
        } else {
            EB_TRY(learned_predictor(ctx, p));
            if (ctx->h_pred[2] == 0) {
                ctx->learn_backoff = std::min(std::max(1, 2 * ctx->learn_backoff), ctx->learn_max_skip);
            } else {
                ctx->learn_backoff = 0;
            }
            ctx->learn_wait = ctx->learn_backoff;
        }
    }
    EB_HIP(hipStreamSynchronize(ctx->stream));
    return EB_OK;
}

int eb_predictor(eb_ctx* ctx, const eb_newton_params* p, double dt) {
    if (!ctx || !p) return EB_ERR_ARG;
    EB_TRY(set_participation(ctx, nullptr));
    EB_TRY(set_coefs(ctx, nullptr, dt, p->beta, p->gamma));
    return predictor_impl(ctx, p);
}

// corrector! (src/corrector.jl:1-59) for the instances taking part; ctx->k_inst receives each
// instance's k (0 for instances not taking part), *kout the largest
static int corrector_impl(eb_ctx* ctx, const eb_newton_params* p, int* kout) {
    const int32_t kfail = p->max_iterations + 1;
    EB_TRY(init_global_launch(ctx, 0));
    EB_TRY(newton_reset_launch(ctx, true, kfail));
    // corrector.jl:17-36 per instance: an instance leaves the loop when its norms pass or its
    // iterations are used up; the others keep iterating with it frozen
    int k = 1, n_active = 0;
    for (int64_t s = 0; s < ctx->S; ++s) n_active += ctx->h_part[s] ? 1 : 0;
    int st = EB_OK;
    const NewtonCtl nc{p->tol_res, p->tol_disp, 0, p->max_iterations, 1};
    const double dt = ctx->dts.empty() ? 0.0 : ctx->dts[0];
    while (n_active > 0 && k <= p->max_iterations) {
        // assembly, lift, factorisation and one BiCGStab iteration, then the update and norms
        // behind the solve's all-done flag: one synchronisation when one iteration solves (the
        // usual case with the chain preconditioner); otherwise the solve continues and the update
        // and norms are launched again after it. In graph mode the body is one graph replay.
        const bool start_done = k == 1 && ctx->start_assembled;
        const bool graph = ctx->exec_mode == EB_EXEC_GRAPH && !start_done;
        st = graph ? graph_iteration(ctx, p, dt) : newton_body(ctx, p, dt, start_done);
        if (st == EB_OK && hipStreamSynchronize(ctx->stream) != hipSuccess) {
            ctx->err = "Newton iteration: stream synchronisation failed";
            st = EB_ERR_DEVICE;
        }
        if (st == EB_OK && graph) prof_harvest_graph(ctx);
        if (st == EB_OK && !ctx->h_kall->all_done) {
            st = solve_finish(ctx, p, 1);
            if (st == EB_OK) solve_account(ctx, nullptr);
            if (st == EB_OK) st = screen_launch(ctx, kfail, nullptr, nullptr);
            if (st == EB_OK) st = scatter_launch(ctx, nullptr);
            if (st == EB_OK) st = global_corrector_async(ctx);
            if (st == EB_OK) st = local_corrector_async(ctx);
            if (st == EB_OK) st = norms_kernels(ctx, nc, nullptr);
            if (st == EB_OK) {
                if (hipMemcpyAsync(ctx->h_nall, ctx->nall, sizeof(NewtonAll), hipMemcpyDeviceToHost, ctx->stream) !=
                        hipSuccess ||
                    hipStreamSynchronize(ctx->stream) != hipSuccess)
                    st = EB_ERR_DEVICE;
            }
        } else if (st == EB_OK) {
            solve_account(ctx, nullptr);
        }
        if (st != EB_OK) break;
        if (!ctx->ev_pending.empty()) prof_harvest(ctx);
        n_active = ctx->h_nall->n_active;
        ctx->counters[0]++;
        k += 1;
    }
    if (st == EB_OK) {
        EB_HIP(hipMemcpyAsync(ctx->h_kiter, ctx->kiter, sizeof(int32_t) * ctx->S, hipMemcpyDeviceToHost, ctx->stream));
        EB_HIP(hipStreamSynchronize(ctx->stream));
    }
    // every instance taking part in the step takes part in whatever runs outside the corrector
    // loop; the predictor's failure flags are spent
    int st2 = newton_reset_launch(ctx);
    if (st2 == EB_OK && hipMemsetAsync(ctx->pfail, 0, ctx->S, ctx->stream) != hipSuccess) st2 = EB_ERR_DEVICE;
    if (st != EB_OK) return st;
    EB_TRY(st2);
    int kmax = 0;
    for (int64_t s = 0; s < ctx->S; ++s) {
        ctx->k_inst[s] = !ctx->h_part[s] ? 0 : (ctx->h_kiter[s] > 0 ? ctx->h_kiter[s] : k);
        kmax = std::max(kmax, ctx->k_inst[s]);
    }
    *kout = kmax;
    return EB_OK;
}

int eb_corrector(eb_ctx* ctx, const eb_newton_params* p, double dt, int* kout) {
    if (!ctx || !p || !kout) return EB_ERR_ARG;
    EB_TRY(set_participation(ctx, nullptr));
    EB_TRY(set_coefs(ctx, nullptr, dt, p->beta, p->gamma));
    return corrector_impl(ctx, p, kout);
}

int eb_solve_step_dynamics(eb_ctx* ctx, const eb_newton_params* p, double dt, int* k) {
    if (ctx && p && k && ctx->exec_mode == EB_EXEC_FUSED) return step_fused(ctx, p, dt, k);
    EB_TRY(eb_predictor(ctx, p, dt));
    return eb_corrector(ctx, p, dt, k);
}

static int update_converged(eb_ctx* ctx, bool wait);

int eb_run_steps(eb_ctx* ctx, const eb_newton_params* p, const eb_time_params* tp, int64_t max_steps,
                 eb_step_callback callback, void* user, eb_run_state* state) {
    if (!ctx || !p || !tp || !state || max_steps < 0) return EB_ERR_ARG;
    int64_t done = 0;
    while (done < max_steps && tp->simulation_end_time - state->t > 1.0e-10 && !state->stopped) {
        const double tnp1 = state->t + state->dt;
        if (callback && callback(ctx, tnp1, user) != 0) {
            ctx->err = "step callback failed at t = " + std::to_string(tnp1);
            return EB_ERR_ARG;
        }
        int k = 0;
        EB_TRY(eb_solve_step_dynamics(ctx, p, state->dt, &k));
        state->last_k = k;
        if (k > p->max_iterations) {
            state->dt /= 2;
            state->failed_steps += 1;
            if (state->dt < tp->min_timestep) {
                char buf[96];
                snprintf(buf, sizeof buf, "Δt < %.2e: simulation aborted", tp->min_timestep);
                ctx->err = buf;
                return EB_ERR_STATE;
            }
            EB_TRY(eb_update_not_converged(ctx));
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
        // a callback may read the state through its own copies: the commit is then waited for
        EB_TRY(update_converged(ctx, callback != nullptr));
        state->t = tnp1;
        if (tp->stop_on_energy_threshold) {
            double en[3];
            EB_HIP(hipMemcpyAsync(en, ctx->energy, sizeof en, hipMemcpyDeviceToHost, ctx->stream));
            EB_HIP(hipStreamSynchronize(ctx->stream));
            if (en[1] < tp->energy_threshold) state->stopped = 1;
        }
    }
    EB_HIP(hipStreamSynchronize(ctx->stream));
    return EB_OK;
}

// The rates of the step just converged into the learned start's history of the instances in
// `sel_host` (all when null); `sel_dev` is its device copy
static int record_history(eb_ctx* ctx, const uint8_t* sel_host, const uint8_t* sel_dev) {
    if (!ctx->coef_set) return EB_OK;
    for (int64_t s = 0; s < ctx->S; ++s)
        if (!sel_host || sel_host[s]) ctx->hcount[s] = std::min(ctx->hcount[s] + 1, 3);
    // the slot index advances on the device: no host copy, so no wait for the previous one
    hipLaunchKernelGGL(learned_advance, dim3(nblk(ctx->S, 256)), dim3(256), 0, ctx->stream, ctx->S, sel_dev, ctx->newest);
    hipLaunchKernelGGL(learned_record, dim3(nblk(ctx->N, 256)), dim3(256), 0, ctx->stream, ctx->N, ctx->node[EB_F_U],
                       ctx->node[EB_F_U_N], ctx->node[EB_F_W], (const StepCoef*)ctx->coef, ctx->node_inst, sel_dev,
                       ctx->newest, ctx->hist);
    return check_launch(ctx, "learned_record");
}

// wait false: the commit stays queued on the stream (the time loop's next step follows it in order)
static int update_converged(eb_ctx* ctx, bool wait) {
    // rate of the step that converged, newest first in the history of the learned warm start
    EB_TRY(record_history(ctx, nullptr, nullptr));
    EB_TRY(copy_state(ctx, true));
    EB_TRY(materialize_kcm(ctx));
    if (!ctx->np1_alias) {
        // matricesⁿ = matricesⁿ⁺¹ by exchanging buffers; n+1 now aliases n until the next assembly
        std::swap(ctx->K, ctx->Kn);
        std::swap(ctx->C, ctx->Cn);
        std::swap(ctx->M, ctx->Mn);
        ctx->np1_alias = true;
        ctx->kcm_n_ver = ctx->kcm_ver;
    }
    if (wait) EB_HIP(hipStreamSynchronize(ctx->stream));
    return EB_OK;
}

int eb_update_converged(eb_ctx* ctx) {
    if (!ctx) return EB_ERR_ARG;
    return update_converged(ctx, true);
}

int eb_update_not_converged(eb_ctx* ctx) {
    if (!ctx) return EB_ERR_ARG;
    EB_TRY(copy_state(ctx, false));
    ctx->kcm_pending = false;
    ctx->np1_alias = true;
    ctx->kcm_ver = ctx->kcm_n_ver;
    EB_HIP(hipStreamSynchronize(ctx->stream));
    return EB_OK;
}

// End of a batch step in which the instances flagged in `commit` converged and the others taking
// part repeat it: update_converged! for the first, update_not_converged! for the second
// (src/utils_solver.jl:252-304), instance by instance. Instances not taking part are unchanged.
static int step_end_mixed(eb_ctx* ctx, const uint8_t* commit) {
    int64_t n_commit = 0, n_part = 0;
    for (int64_t s = 0; s < ctx->S; ++s) {
        n_commit += commit[s] ? 1 : 0;
        n_part += ctx->h_part[s] ? 1 : 0;
    }
    if (n_commit == ctx->S) return eb_update_converged(ctx);
    if (n_commit == 0) {
        // only instances taking part changed their n+1 state
        EB_TRY(copy_state(ctx, false, ctx->part));
        ctx->kcm_pending = false;
        ctx->np1_alias = true;
        ctx->kcm_ver = ctx->kcm_n_ver;
        EB_HIP(hipStreamSynchronize(ctx->stream));
        return EB_OK;
    }
    (void)n_part;
    // selection: the committing instances
    EB_HIP(hipStreamSynchronize(ctx->stream));
    for (int64_t s = 0; s < ctx->S; ++s) ctx->h_sel[s] = commit[s] ? 1 : 0;
    EB_HIP(hipMemcpyAsync(ctx->sel, ctx->h_sel, ctx->S, hipMemcpyHostToDevice, ctx->stream));
    EB_TRY(record_history(ctx, ctx->h_sel, ctx->sel));
    EB_TRY(copy_state(ctx, true, ctx->sel));
    // K, C, M at n+1 of the committing instances; the others take their step-n matrices back
    // before the exchange of buffers
    EB_TRY(materialize_kcm(ctx));
    if (ctx->np1_alias) EB_TRY(unalias(ctx));
    EB_HIP(hipStreamSynchronize(ctx->stream));
    for (int64_t s = 0; s < ctx->S; ++s) ctx->h_sel[s] = commit[s] ? 0 : 1;
    EB_HIP(hipMemcpyAsync(ctx->sel, ctx->h_sel, ctx->S, hipMemcpyHostToDevice, ctx->stream));
    EB_TRY(copy_state(ctx, false, ctx->sel));   // rollback (a no-op for instances not taking part)
    {
        CopyList cl{};
        cl.sel = ctx->sel;
        double* src[3] = {ctx->Kn, ctx->Cn, ctx->Mn};
        double* dst[3] = {ctx->K, ctx->C, ctx->M};
        for (int k = 0; k < 3; ++k) {
            cl.src[k] = src[k];
            cl.dst[k] = dst[k];
            cl.n[k] = ctx->nnz;
            cl.map[k] = ctx->blk_inst;
            cl.per[k] = 36;
        }
        hipLaunchKernelGGL(copy_many, dim3(std::min<unsigned>(nblk(ctx->nnz, 256), 4096), 3), dim3(256), 0, ctx->stream, cl);
        EB_TRY(check_launch(ctx, "copy_back_matrices"));
    }
    std::swap(ctx->K, ctx->Kn);
    std::swap(ctx->C, ctx->Cn);
    std::swap(ctx->M, ctx->Mn);
    ctx->np1_alias = true;
    ctx->kcm_pending = false;
    // the step-n matrices now come from two element stagings: no predictor reuses a factorisation
    ctx->kcm_n_ver = 0;
    ctx->kcm_ver = 0;
    EB_HIP(hipStreamSynchronize(ctx->stream));
    return EB_OK;
}

int eb_run_steps_batch(eb_ctx* ctx, const eb_newton_params* p, const eb_time_params* tp, int64_t max_steps,
                       eb_instance_callback callback, void* user, eb_run_state* states) {
    if (!ctx || !p || !tp || !states || max_steps < 0) return EB_ERR_ARG;
    if (tp->stop_on_energy_threshold) {
        ctx->err = "eb_run_steps_batch: the energy stop needs per-instance kinetic energies, which the batch does not keep";
        return EB_ERR_ARG;
    }
    const int64_t S = ctx->S;
    std::vector<int64_t> done((size_t)S, 0);
    std::vector<uint8_t> part((size_t)S, 0), commit((size_t)S, 0);
    std::vector<double> dt_inst((size_t)S, 0.0);
    int st = EB_OK;
    for (;;) {
        // every instance that has not stopped or reached the end time steps in every round, until
        // none of them is short of max_steps: an instance that repeats a step with Δt/2 does not
        // hold the others up
        int64_t n_short = 0;
        for (int64_t s = 0; s < S; ++s) {
            const eb_run_state& r = states[s];
            part[s] = (!r.stopped && tp->simulation_end_time - r.t > 1.0e-10) ? 1 : 0;
            n_short += (part[s] && done[s] < max_steps) ? 1 : 0;
            dt_inst[s] = r.dt > 0 ? r.dt : tp->initial_timestep;
        }
        if (n_short == 0) break;
        for (int64_t s = 0; s < S && callback; ++s) {
            if (!part[s]) continue;
            const double tnp1 = states[s].t + states[s].dt;
            if (callback(ctx, s, tnp1, user) != 0) {
                ctx->err = "step callback failed for instance " + std::to_string(s) + " at t = " + std::to_string(tnp1);
                st = EB_ERR_ARG;
                break;
            }
        }
        if (st != EB_OK) break;
        if ((st = set_participation(ctx, part.data())) != EB_OK) break;
        if ((st = set_coefs(ctx, dt_inst.data(), 0.0, p->beta, p->gamma)) != EB_OK) break;
        int kmax = 0;
        if ((st = predictor_impl(ctx, p)) != EB_OK) break;
        if ((st = corrector_impl(ctx, p, &kmax)) != EB_OK) break;
        // run_simulation.jl:53-84 per instance
        for (int64_t s = 0; s < S; ++s) {
            commit[s] = 0;
            if (!part[s]) continue;
            eb_run_state& r = states[s];
            const int k = ctx->k_inst[s];
            r.last_k = k;
            if (k > p->max_iterations) {
                r.dt /= 2;
                r.failed_steps += 1;
                r.successive = 0;
                if (r.dt < tp->min_timestep) r.stopped = 2;   // Δt below min_timestep: this instance aborts
                continue;
            }
            commit[s] = 1;
            const double tnp1 = r.t + r.dt;
            r.successive += 1;
            r.steps += 1;
            r.newton_iterations += k - 1;
            done[s] += 1;
            if (r.successive >= tp->accelerate_after_success_it && k <= 8) {
                r.dt = std::min(r.dt * 1.5, tp->max_timestep);
                r.successive = 0;
            }
            r.t = tnp1;
        }
        if ((st = step_end_mixed(ctx, commit.data())) != EB_OK) break;
    }
    // every instance takes part in what the caller runs next
    int st2 = set_participation(ctx, nullptr);
    return st != EB_OK ? st : st2;
}

int eb_learned_sample(eb_ctx* ctx, float* X, float* T) {
    if (!ctx || !X || !T) return EB_ERR_ARG;
    for (int64_t s = 0; s < ctx->S; ++s)
        if (ctx->hcount[s] < 3 || !ctx->coef_set) {
            ctx->err = "fewer than three converged steps recorded";
            return EB_ERR_STATE;
        }
    const int64_t N = ctx->N;
    hipLaunchKernelGGL(learned_features, dim3(nblk(N, 256)), dim3(256), 0, ctx->stream, N, ctx->pdD, ctx->hist,
                       ctx->newest, ctx->node_inst, ctx->Tc_n, (const StepCoef*)ctx->coef, ctx->ones, ctx->lX);
    hipLaunchKernelGGL(learned_targets, dim3(nblk(N, 256)), dim3(256), 0, ctx->stream, N, ctx->node[EB_F_U],
                       ctx->node[EB_F_U_N], ctx->node[EB_F_W], ctx->pdD, ctx->ones, ctx->lT);
    EB_TRY(check_launch(ctx, "learned_sample"));
    EB_HIP(hipMemcpyAsync(X, ctx->lX, sizeof(float) * LEARN_IN * N, hipMemcpyDeviceToHost, ctx->stream));
    EB_HIP(hipMemcpyAsync(T, ctx->lT, sizeof(float) * LEARN_OUT * N, hipMemcpyDeviceToHost, ctx->stream));
    EB_HIP(hipStreamSynchronize(ctx->stream));
    return EB_OK;
}

int eb_set_learned_predictor(eb_ctx* ctx, eb_mlp* mlp, const float* in_scale, const float* out_scale) {
    if (!ctx) return EB_ERR_ARG;
    if (!mlp) { ctx->mlp = nullptr; return EB_OK; }
    if (!in_scale || !out_scale || ebm::mlp_width(mlp, 0) != LEARN_IN || ebm::mlp_width(mlp, 1) != LEARN_OUT) {
        ctx->err = "the network must map 27 inputs to 6 outputs";
        return EB_ERR_ARG;
    }
    EB_HIP(hipMemcpyAsync(ctx->lin_scale, in_scale, sizeof(float) * LEARN_IN, hipMemcpyHostToDevice, ctx->stream));
    EB_HIP(hipMemcpyAsync(ctx->lout_scale, out_scale, sizeof(float) * LEARN_OUT, hipMemcpyHostToDevice, ctx->stream));
    EB_HIP(hipStreamSynchronize(ctx->stream));
    // this context's layer outputs for its nodes (replacing those of a network installed before)
    EB_HIP(hipStreamSynchronize(ctx->stream));
    for (float* q : ctx->lact) (void)hipFree(q);
    ctx->lact.assign((size_t)ebm::mlp_layers(mlp), nullptr);
    for (int l = 0; l < ebm::mlp_layers(mlp); ++l) {
        void* q = nullptr;
        if (hipMalloc(&q, sizeof(float) * (size_t)std::max<int64_t>(ctx->N, 1) * ebm::mlp_layer_width(mlp, l)) != hipSuccess) {
            ctx->err = "learned predictor: the network's layer outputs could not be allocated";
            return EB_ERR_DEVICE;
        }
        ctx->lact[l] = (float*)q;
    }
    ctx->mlp = mlp;
    ctx->learn_backoff = ctx->learn_wait = 0;
    return EB_OK;
}

int eb_get_learned_increment(eb_ctx* ctx, double* out, int64_t count) {
    if (!ctx || !out) return EB_ERR_ARG;
    if (count != ctx->ndofs) { ctx->err = "increment size mismatch"; return EB_ERR_ARG; }
    EB_HIP(hipMemcpyAsync(out, ctx->ldD, sizeof(double) * count, hipMemcpyDeviceToHost, ctx->stream));
    EB_HIP(hipStreamSynchronize(ctx->stream));
    return EB_OK;
}

int eb_set_learned_backoff(eb_ctx* ctx, int32_t max_skip) {
    if (!ctx || max_skip < 0) return EB_ERR_ARG;
    ctx->learn_max_skip = max_skip;
    ctx->learn_backoff = std::min(ctx->learn_backoff, max_skip);
    ctx->learn_wait = std::min(ctx->learn_wait, max_skip);
    return EB_OK;
}
int eb_get_predictor_stats(const eb_ctx* ctx, int64_t* out3) {
    if (!ctx || !out3) return EB_ERR_ARG;
    int32_t c[2] = {0, 0};
    if (hipMemcpy(c, ctx->pred_counts, sizeof c, hipMemcpyDeviceToHost) != hipSuccess) return EB_ERR_DEVICE;
    out3[0] = c[0];
    out3[1] = c[1];
    out3[2] = ctx->learn_skipped;
    return EB_OK;
}

int eb_get_counters(const eb_ctx* ctx, int64_t* out4) {
    if (!ctx || !out4) return EB_ERR_ARG;
    for (int k = 0; k < 4; ++k) out4[k] = ctx->counters[k];
    return EB_OK;
}

int eb_profile_enable(eb_ctx* ctx, uint32_t mask) {
    if (!ctx) return EB_ERR_ARG;
    EB_HIP(hipStreamSynchronize(ctx->stream));
    prof_harvest(ctx);
    for (int k = 0; k < EB_KERNEL_COUNT; ++k) { ctx->prof_ms[k] = 0; ctx->prof_n[k] = 0; ctx->ev_open[k] = nullptr; }
    ctx->prof_mask = mask;
    if (ctx->work) {
        EB_HIP(hipMemsetAsync(ctx->work, 0, 4 * sizeof(unsigned long long), ctx->stream));
        EB_HIP(hipStreamSynchronize(ctx->stream));
    }
    graph_release(ctx);   // a captured iteration holds the profiling launches of its capture
    return EB_OK;
}

int eb_get_work_counters(eb_ctx* ctx, int64_t* out3) {
    if (!ctx || !out3) return EB_ERR_ARG;
    unsigned long long w[4] = {};
    EB_HIP(hipStreamSynchronize(ctx->stream));
    if (ctx->work) EB_HIP(hipMemcpy(w, ctx->work, sizeof w, hipMemcpyDeviceToHost));
    for (int k = 0; k < 3; ++k) out3[k] = (int64_t)w[k];
    return EB_OK;
}

int eb_output_points(eb_ctx* ctx, int32_t n_points, double* position, double* displacement, double* velocity,
                     double* strain, double* stress) {
    if (!ctx || n_points < 1) return EB_ERR_ARG;
    const int64_t np = (int64_t)n_points * ctx->E;
    double* outs[5] = {position, displacement, velocity, strain, stress};
    const int64_t width[5] = {3, 3, 3, 1, 1};
    double* dev[5] = {};
    auto release = [&] {
        for (double* q : dev)
            if (q) (void)hipFree(q);
    };
    for (int k = 0; k < 5; ++k)
        if (outs[k] && hipMalloc(&dev[k], sizeof(double) * width[k] * std::max<int64_t>(np, 1)) != hipSuccess) {
            release();
            ctx->err = "eb_output_points: device allocation failed";
            return EB_ERR_DEVICE;
        }
    hipLaunchKernelGGL(beam_output_points, dim3(nblk(ctx->E, 256)), dim3(256), 0, ctx->stream, node_state(ctx),
                       beam_consts(ctx), (int)n_points, dev[0], dev[1], dev[2], dev[3], dev[4]);
    int st = check_launch(ctx, "beam_output_points");
    for (int k = 0; k < 5 && st == EB_OK; ++k)
        if (outs[k] && hipMemcpyAsync(outs[k], dev[k], sizeof(double) * width[k] * np, hipMemcpyDeviceToHost, ctx->stream) !=
                           hipSuccess)
            st = EB_ERR_DEVICE;
    if (st == EB_OK && hipStreamSynchronize(ctx->stream) != hipSuccess) st = EB_ERR_DEVICE;
    release();
    if (st != EB_OK && ctx->err.empty()) ctx->err = "eb_output_points failed";
    return st;
}

int eb_get_contact_stats(eb_ctx* ctx, int64_t* out2) {
    if (!ctx || !out2) return EB_ERR_ARG;
    unsigned long long v[2] = {0, 0};
    if (ctx->contact_stats) {
        EB_HIP(hipStreamSynchronize(ctx->stream));
        EB_HIP(hipMemcpy(v, ctx->contact_stats, sizeof v, hipMemcpyDeviceToHost));
    }
    out2[0] = (int64_t)v[0];
    out2[1] = (int64_t)v[1];
    return EB_OK;
}

int eb_output_points_noalloc(eb_ctx* ctx, int32_t n_points, double* position, double* displacement, double* velocity,
                             double* strain, double* stress) {
    if (!ctx || n_points < 1) return EB_ERR_ARG;
    if (!ctx->scratch) { ctx->err = "eb_output_points_noalloc: no staging on this device"; return EB_ERR_STATE; }
    const int64_t np = (int64_t)n_points * ctx->E;
    double* outs[5] = {position, displacement, velocity, strain, stress};
    const int64_t width[5] = {3, 3, 3, 1, 1};
    int64_t need = 0;
    for (int k = 0; k < 5; ++k)
        if (outs[k]) need += width[k] * np;
    if (need > EB_NOALLOC_SCRATCH_DOUBLES) {
        ctx->err = "eb_output_points_noalloc: " + std::to_string(need) + " doubles of output exceed the staging of " +
                   std::to_string((long long)EB_NOALLOC_SCRATCH_DOUBLES) + " (use eb_output_points)";
        return EB_ERR_ARG;
    }
    // every requested output is written in full by the kernel: nothing to clear
    double* dev[5] = {};
    int64_t off = 0;
    for (int k = 0; k < 5; ++k)
        if (outs[k]) { dev[k] = ctx->scratch + off; off += width[k] * np; }
    hipLaunchKernelGGL(beam_output_points, dim3(nblk(ctx->E, 256)), dim3(256), 0, ctx->stream, node_state(ctx),
                       beam_consts(ctx), (int)n_points, dev[0], dev[1], dev[2], dev[3], dev[4]);
    EB_TRY(check_launch(ctx, "beam_output_points"));
    for (int k = 0; k < 5; ++k)
        if (outs[k] && np > 0)
            EB_HIP(hipMemcpyAsync(outs[k], dev[k], sizeof(double) * width[k] * np, hipMemcpyDeviceToHost, ctx->stream));
    EB_HIP(hipStreamSynchronize(ctx->stream));
    return EB_OK;
}

int eb_get_sdf_coefficients(eb_ctx* ctx, int64_t instance, double* out, int64_t count) {
    if (!ctx || !out || instance < 0 || instance >= ctx->S) return EB_ERR_ARG;
    if (ctx->surface != EB_SURFACE_SDF) { ctx->err = "no signed distance field"; return EB_ERR_STATE; }
    const bool f32 = ctx->precision == EB_PRECISION_FP32;
    const size_t rec = f32 ? sizeof(SDFDev<float>) : sizeof(SDFDev<double>);
    std::vector<char> hdr(rec);
    EB_HIP(hipMemcpy(hdr.data(), (const char*)ctx->d_sdf + rec * instance, rec, hipMemcpyDeviceToHost));
    const int32_t* n = f32 ? ((const SDFDev<float>*)hdr.data())->n : ((const SDFDev<double>*)hdr.data())->n;
    const int64_t ncoef = (int64_t)(n[0] + 2) * (n[1] + 2) * (n[2] + 2);
    if (count != ncoef) { ctx->err = "coefficient count mismatch"; return EB_ERR_ARG; }
    if (f32) {
        std::vector<float> tmp((size_t)ncoef);
        EB_HIP(hipMemcpy(tmp.data(), ((const SDFDev<float>*)hdr.data())->coef, sizeof(float) * ncoef, hipMemcpyDeviceToHost));
        for (int64_t q = 0; q < ncoef; ++q) out[q] = tmp[q];
    } else {
        EB_HIP(hipMemcpy(out, ((const SDFDev<double>*)hdr.data())->coef, sizeof(double) * ncoef, hipMemcpyDeviceToHost));
    }
    return EB_OK;
}

int eb_set_exec_mode(eb_ctx* ctx, int mode) {
    if (!ctx || (mode != EB_EXEC_EAGER && mode != EB_EXEC_GRAPH && mode != EB_EXEC_FUSED)) return EB_ERR_ARG;
    if (mode == EB_EXEC_FUSED) {
        std::string why;
        if (!fused_eligible(ctx, &why)) {
            ctx->err = "fused execution: " + why;
            return EB_ERR_ARG;
        }
        if (!ctx->fused_out) EB_TRY(dalloc(ctx, ctx->fused_out, 3));
    }
    EB_HIP(hipStreamSynchronize(ctx->stream));
    ctx->exec_mode = mode;
    if (mode == EB_EXEC_EAGER) graph_release(ctx);
    return EB_OK;
}

int eb_get_exec_stats(const eb_ctx* ctx, int32_t* mode, int64_t* captures, int64_t* replays) {
    if (!ctx) return EB_ERR_ARG;
    if (mode) *mode = ctx->exec_mode;
    if (captures) *captures = ctx->graph_stats[0];
    if (replays) *replays = ctx->graph_stats[1];
    return EB_OK;
}

int eb_autotune(eb_ctx* ctx, int32_t* choice, double* ms_per_candidate) {
    if (!ctx || !choice) return EB_ERR_ARG;
    double scratch[EB_TUNE_COUNT * EB_TUNE_CANDIDATES];
    return autotune_impl(ctx, choice, ms_per_candidate ? ms_per_candidate : scratch);
}

int eb_set_tuning(eb_ctx* ctx, const int32_t* choice) {
    if (!ctx || !choice) return EB_ERR_ARG;
    if (choice[EB_TUNE_CHAIN_LDS_KB] != 0 && (choice[EB_TUNE_CHAIN_LDS_KB] < 1 || choice[EB_TUNE_CHAIN_LDS_KB] > 96))
        return EB_ERR_ARG;
    if (choice[EB_TUNE_CHAIN_SOLVE_LDS] < 0 || choice[EB_TUNE_CHAIN_SOLVE_LDS] > 2) return EB_ERR_ARG;
    if (choice[EB_TUNE_CHAIN_ALGO] < 0 || choice[EB_TUNE_CHAIN_ALGO] > 2) return EB_ERR_ARG;
    const int32_t g = choice[EB_TUNE_GATHER_BLOCK];
    if (g != 0 && g != 64 && g != 128 && g != 256) return EB_ERR_ARG;
    EB_HIP(hipStreamSynchronize(ctx->stream));
    for (int id = 0; id < EB_TUNE_COUNT; ++id) ctx->tune[id] = choice[id];
    graph_release(ctx);
    return EB_OK;
}

int eb_get_tuning(const eb_ctx* ctx, int32_t* choice) {
    if (!ctx || !choice) return EB_ERR_ARG;
    for (int id = 0; id < EB_TUNE_COUNT; ++id) choice[id] = tuned(ctx, id);
    return EB_OK;
}

int eb_profile_read(eb_ctx* ctx, int id, double* total_ms, int64_t* launches) {
    if (!ctx || id < 0 || id >= EB_KERNEL_COUNT || !total_ms || !launches) return EB_ERR_ARG;
    EB_HIP(hipStreamSynchronize(ctx->stream));
    prof_harvest(ctx);
    *total_ms = ctx->prof_ms[id];
    *launches = ctx->prof_n[id];
    return EB_OK;
}

}  // extern "C"
