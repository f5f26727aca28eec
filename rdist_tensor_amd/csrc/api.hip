// C-ABI of libexafac_hip.so (include/exafac_hip.h): object lifetimes, the
// randomized ALS step orchestration and the statistics JSON.
//
// Reference flow: ALS_Optimizer::execute_ALS_rounds (als_optimizer.hpp:40-124)
// over AccumulatorStationary::execute_ALS_step (accumulator_stationary.hpp:139-295).
#include <algorithm>
#include <atomic>
#include <chrono>
#include <dlfcn.h>
#include <fstream>
#include <exception>
#include <thread>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/exafac_hip.h"
#include "common.hpp"

using namespace exafac;

// Scratch of ex_gram_dev_static, the entry point that allocates nothing: the
// per-block Gram partials live in this file-scope device array (one copy per
// GPU, held until the process ends).  Host code never uses the array's name
// as an address: ex_ctx_create looks up each device's copy once
// (hipGetSymbolAddress, with that device current) and the entry point passes
// that address to the kernels.
constexpr size_t kStaticScratchBytes = (size_t)8 << 20;
static __device__ __align__(256) unsigned char scratch[kStaticScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// The address of `device`'s copy of the scratch array (the device is current).
void lookup_static_scratch(int device) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[device]) return;
    void* p = nullptr;
    EX_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    EX_CHECK(p != nullptr, EX_ERR_HIP, "no device address for the static scratch array");
    g_scratch_ptr[device] = p;
}

thread_local std::string g_last_error;

// Launch variants (common.hpp TuneKnob): names, variant counts, current values.
const TuneInfo kTuneTable[kTuneKnobs] = {
    {"sts_draw.fp64.r16", 4}, {"sts_draw.fp64.r32", 6}, {"sts_draw.fp32.r16", 4},
    {"sts_draw.fp32.r32", 7}, {"sts_draw.fp32.r64", 4}, {"row_scan.fp64", 2},
    {"row_scan.fp32", 2}, {"sts_first.fp32", 2},
};
std::atomic<int> g_tune[kTuneKnobs];
std::once_flag g_tune_once;

int tune_index(const std::string& name) {
    for (int q = 0; q < kTuneKnobs; q++)
        if (name == kTuneTable[q].name) return q;
    return -1;
}

// "name value" lines ('#' starts a comment); every name must be a knob and
// every value one of its variants.
void load_tuning_file(const std::string& path) {
    std::ifstream in(path);
    EX_CHECK(in.good(), EX_ERR_ARG, "cannot open tuning file " + path);
    std::string line;
    int ln = 0;
    while (std::getline(in, line)) {
        ln++;
        const size_t h = line.find('#');
        if (h != std::string::npos) line.resize(h);
        std::istringstream ss(line);
        std::string name;
        if (!(ss >> name)) continue;
        int v = -1;
        EX_CHECK((bool)(ss >> v), EX_ERR_ARG, path + ":" + std::to_string(ln) + ": expected 'name value'");
        const int q = tune_index(name);
        EX_CHECK(q >= 0, EX_ERR_ARG, path + ":" + std::to_string(ln) + ": unknown knob " + name);
        EX_CHECK(v >= 0 && v < kTuneTable[q].variants, EX_ERR_ARG,
                 path + ":" + std::to_string(ln) + ": " + name + " has no variant " + std::to_string(v));
        g_tune[q].store(v);
    }
}

// tuning.txt beside the library (or $EXAFAC_TUNING; "off" keeps the
// defaults), read once per process by the first context.
void load_default_tuning() {
    std::call_once(g_tune_once, [] {
        const char* env = std::getenv("EXAFAC_TUNING");
        std::string path;
        if (env && *env) {
            if (std::string(env) == "off") return;
            path = env;
        } else {
            Dl_info info;
            if (!dladdr(reinterpret_cast<void*>(&load_tuning_file), &info) || !info.dli_fname) return;
            path = info.dli_fname;
            const size_t sl = path.rfind('/');
            path = (sl == std::string::npos ? std::string(".") : path.substr(0, sl)) + "/tuning.txt";
            std::ifstream probe(path);
            if (!probe.good()) return;
        }
        load_tuning_file(path);
    });
}

template <class F>
int guarded(F&& f) {
    try {
        f();
        g_last_error.clear();
        return EX_OK;
    } catch (const Error& e) {
        g_last_error = e.what();
        return e.code;
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return EX_ERR_STATE;
    } catch (...) {
        g_last_error = "unknown error";
        return EX_ERR_STATE;
    }
}

void h2d(hipStream_t s, void* d, const void* h, size_t bytes) {
    if (bytes) EX_HIP(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, s));
}
void d2h(hipStream_t s, void* h, const void* d, size_t bytes) {
    if (bytes) EX_HIP(hipMemcpyAsync(h, d, bytes, hipMemcpyDeviceToHost, s));
    EX_HIP(hipStreamSynchronize(s));
}

std::string json_num(double v) {
    if (!std::isfinite(v)) return "null";
    std::ostringstream o;
    o.precision(17);
    o << v;
    return o.str();
}

// compute_dstat (common.h:613-675) at one rank: mean = min = max = x; the
// reference's std divides 0 by (world_size - 1) = 0, giving NaN (JSON null).
std::string dstat(double x) {
    return "{\"max\": " + json_num(x) + ", \"mean\": " + json_num(x) + ", \"min\": " + json_num(x) +
           ", \"std\": null}";
}

template <class T>
std::string json_list(const std::vector<T>& v) {
    std::string s = "[";
    for (size_t i = 0; i < v.size(); i++) {
        if (i) s += ", ";
        if constexpr (std::is_floating_point<T>::value) s += json_num((double)v[i]);
        else s += std::to_string(v[i]);
    }
    return s + "]";
}

}  // namespace

namespace exafac {
const TuneInfo& tune_info(int knob) { return kTuneTable[knob]; }
int tune_value(int knob) { return g_tune[knob].load(std::memory_order_relaxed); }
void tune_select(int knob, int value) { g_tune[knob].store(value); }
}  // namespace exafac

struct ex_ctx {
    Ctx c;
    // second stream for work that overlaps the main stream inside a step
    // (the design Gram and its inverse run beside the sampled MTTKRP)
    hipStream_t aux = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    hipEvent_t ev_zfork = nullptr, ev_zjoin = nullptr;   // Y cleared on aux beside the draw
    hipEvent_t ev_pfork = nullptr, ev_pjoin = nullptr;   // mixed STS-CP: Gram and M beside the row tree
    // opt-in per-launch-group timing (ex_ctx_timing): events are recorded
    // around named launch groups on the context's stream and resolved after
    // the stream is synchronised
    bool timing = false;
    std::vector<std::pair<std::string, std::pair<hipEvent_t, hipEvent_t>>> pending;
    std::vector<hipEvent_t> pool;
    std::vector<std::pair<std::string, std::pair<double, int64_t>>> totals;
    std::vector<std::pair<std::string, double>> counters;
    hipEvent_t get_event() {
        if (!pool.empty()) {
            hipEvent_t e = pool.back();
            pool.pop_back();
            return e;
        }
        hipEvent_t e;
        EX_HIP(hipEventCreate(&e));
        return e;
    }
    void add_counter(const std::string& n, double v) {
        for (auto& c2 : counters)
            if (c2.first == n) { c2.second += v; return; }
        counters.push_back({n, v});
    }
    // device-side running totals reported as counters: value = current - start
    struct Watch {
        const uint64_t* dev;
        std::string name;
        uint64_t start;
    };
    std::vector<Watch> watches;
    void watch(const uint64_t* dev, const std::string& n) {
        for (auto& w : watches)
            if (w.dev == dev) return;
        uint64_t v = 0;
        EX_HIP(hipMemcpyAsync(&v, dev, 8, hipMemcpyDeviceToHost, c.stream));
        EX_HIP(hipStreamSynchronize(c.stream));
        watches.push_back({dev, n, v});
    }
    // the caller has synchronised the stream
    void resolve() {
        for (auto& w : watches) {
            uint64_t v = 0;
            EX_HIP(hipMemcpy(&v, w.dev, 8, hipMemcpyDeviceToHost));
            bool found = false;
            for (auto& c2 : counters)
                if (c2.first == w.name) { c2.second = (double)(v - w.start); found = true; }
            if (!found) counters.push_back({w.name, (double)(v - w.start)});
        }
        for (auto& p : pending) {
            float ms = 0.f;
            EX_HIP(hipEventElapsedTime(&ms, p.second.first, p.second.second));
            bool found = false;
            for (auto& t : totals)
                if (t.first == p.first) { t.second.first += ms; t.second.second += 1; found = true; break; }
            if (!found) totals.push_back({p.first, {ms, 1}});
            pool.push_back(p.second.first);
            pool.push_back(p.second.second);
        }
        pending.clear();
    }
};

// Marketing name when the runtime reports one, else the ISA name.
std::string device_label(const ex_ctx* c) {
    std::string n = c->c.prop.name;
    std::string arch = c->c.prop.gcnArchName;
    if (n.empty()) return arch;
    return n + " (" + arch + ")";
}

// Records a named interval on the context stream when timing is enabled.
struct Span {
    ex_ctx* c;
    std::string name;
    hipStream_t st;
    hipEvent_t a = nullptr;
    Span(ex_ctx* ctx, const char* n, hipStream_t s = nullptr) : c(ctx), name(n), st(s ? s : ctx->c.stream) {
        if (c->timing) {
            a = c->get_event();
            EX_HIP(hipEventRecord(a, st));
        }
    }
    void end() {
        if (c->timing && a) {
            hipEvent_t b = c->get_event();
            EX_HIP(hipEventRecord(b, st));
            c->pending.push_back({name, {a, b}});
            a = nullptr;
        }
    }
};

// Temporaries of the differentiable sampled MTTKRP (grown on demand, kept
// with the tensor so that a call enqueued on a caller's stream allocates
// nothing once they are large enough).
struct AdjScratch {
    DevBuf kd, sd, wd, dg, dm, dD, table, work, cnt;
    DevBuf tmp[4];
};

struct ex_tensor {
    ex_ctx* ctx = nullptr;
    int N = 0;
    int64_t dims[kMaxModes] = {};
    int64_t nnz = 0;
    DevBuf coords;  // u32 nnz x N
    DevBuf vals;    // f64 nnz
    double norm = 0.0;
    bool built = false;
    ModeLookup lk[kMaxModes];
    DevBuf tmp[4];
    AdjScratch adj;
};

struct ex_factors {
    ex_ctx* ctx = nullptr;
    int N = 0, R = 0;
    int64_t dims[kMaxModes] = {};
    DevBuf U[kMaxModes];
    DevBuf sigma;
    DevBuf part, gram;
    const double* ptrs[kMaxModes] = {};
    // FP32 mirrors for the mixed-precision path, allocated when a mixed
    // sampler is created.  Each mode's current values live in the FP64 copy,
    // the FP32 copy or both; a stale copy is refreshed on its next read.
    bool has32 = false;
    DevBuf U32[kMaxModes];
    const float* ptrs32[kMaxModes] = {};
    bool stale32[kMaxModes] = {};
    bool stale64[kMaxModes] = {};

    void enable32() {
        if (has32) return;
        for (int k2 = 0; k2 < N; k2++) {
            U32[k2].ensure((size_t)dims[k2] * R * 4);
            ptrs32[k2] = U32[k2].as<float>();
            stale32[k2] = true;
        }
        has32 = true;
    }
    // the FP64 copy of `mode` is about to be read
    void sync64(int mode) {
        if (!stale64[mode]) return;
        k::to_f64(ctx->c.stream, U32[mode].as<float>(), dims[mode] * R, U[mode].as<double>());
        stale64[mode] = false;
    }
    void sync64_all() {
        for (int k2 = 0; k2 < N; k2++) sync64(k2);
    }
    void sync32(int mode) {
        if (!stale32[mode]) return;
        k::to_f32(ctx->c.stream, U[mode].as<double>(), dims[mode] * R, U32[mode].as<float>());
        stale32[mode] = false;
    }
    void sync32_all() {
        for (int k2 = 0; k2 < N; k2++) sync32(k2);
    }
    // bumped by every write of a factor (a batch checks that nothing else
    // wrote its factors between its launches)
    uint64_t version = 0;
    // the FP64 (wrote64) or FP32 (wrote32) copy of `mode` was written
    void wrote64(int mode) {
        stale64[mode] = false;
        if (has32) stale32[mode] = true;
        version++;
    }
    void wrote32(int mode) {
        stale32[mode] = false;
        stale64[mode] = true;
        version++;
    }
};

struct ex_sampler {
    ex_ctx* ctx = nullptr;
    ex_factors* f = nullptr;
    int kind = 0, precision = 0;
    uint64_t seed = 0;
    // set while a sweep is captured: the draws read the sweep number here
    const uint64_t* sweep_dev = nullptr;
    // CP-ARLS-LEV
    DevBuf scores[kMaxModes], cdf[kMaxModes], sums, gtmp, gpinv, pivflag;
    // STS-CP (row-tree nodes FP64, or FP32 on the mixed path)
    TreeGeom g[kMaxModes];
    TreeGeom ge;
    DevBuf nodes[kMaxModes];
    DevBuf grams, M, lam, eig, eig32, pinvp;
    DevBuf part;
    // draw outputs and injected draws
    DevBuf samples, weights, draws_row, draws_eig;
    // u64: STS-CP draws given weight 0 since creation (mixed path: a
    // rounding residue routed the draw into a leaf of zero mass)
    DevBuf dropped;
    DevBuf first_cdf;   // mixed: prefix sums of the first non-target mode's leaf masses
    // set when a batch advanced the factors: every mode's state is rebuilt
    // before the next draw or score read
    bool refresh = false;
    // mixed STS-CP: the target mode whose M, eigen tree and packed weights
    // were prepared by the last update (-1: none); any update or M write
    // clears it
    int m_ready = -1;
};

struct ex_als {
    ex_ctx* ctx = nullptr;
    ex_tensor* t = nullptr;
    ex_factors* f = nullptr;
    ex_sampler* s = nullptr;
    int dist = 0, precision = 0;
    bool initialized = false;
    // ExactALS (no sampler): Grams of every factor, N x R x R FP64, current
    // once grams_valid is set (each step refreshes its own mode's)
    bool exact = false, grams_valid = false;
    DevBuf grams;
    DevBuf sd, wd, kd, dg, dm, G, Gi, pivflag, GY, Ms, Y, scal;
    DevBuf dht, dhw;   // hashed deduplication: table (zero between steps), per-sample slots
    DevBuf tmp[4];
    DevBuf part;
    // u64[4]: hits (nonzeros iterated) and touched rows summed over steps, the
    // J_u sum, and the current step's J_u (device-side: no host read per step)
    DevBuf cnt;
    // region events of steps not yet read back: 6 per step
    std::vector<hipEvent_t> ev_pending, ev_pool;
    // sweeps replayed from a captured graph (ex_als_use_graph): the sweep
    // number the captured draws read, the graph of one sweep at sample
    // count graph_J, and whether a capture is running (no region events then)
    bool use_graph = false, capturing = false;
    int64_t graph_J = -1;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    DevBuf sweep_ctr;
    std::vector<const void*> graph_fp;   // the buffers the graph was captured on
    double t_region[7] = {};
    std::string stats = "{}";
};

namespace {

// STS-CP draw set-up for target mode j, from the Grams alone
// (sts_cp.hpp:80-188): M's slots, the eigen tree and the packed weights.
void sts_prepare_m(ex_sampler* s, hipStream_t st, int j) {
    const int N = s->f->N, R = s->f->R;
    if (s->precision == EX_MIXED) {
        k::sts_compute_m_mixed_launch(st, N, j, R, s->grams.as<double>(), s->M.as<double>());
        k::eig_tree_build32(st, N, j, R, s->M.as<double>(), s->grams.as<double>(), s->eig32.as<float>());
    } else {
        k::sts_compute_m(st, N, j, R, s->grams.as<double>(), s->M.as<double>(), s->lam.as<double>());
        k::eig_tree_build(st, N, j, R, s->M.as<double>(), s->grams.as<double>(), s->eig.as<double>());
    }
    k::pack_upper(st, s->M.as<double>(), R, s->pinvp.as<double>());
}

// Mixed precision: FP32 factor, BF16 MFMA Grams, FP64 inverse and scores.
// STS-CP: next >= 0 asks for the draw set-up of target mode next.  With
// gram_done the caller has written U_j^T U_j into grams[j] and, for next >= 0,
// enqueued that set-up on the second stream ending with ev_pjoin; otherwise
// both are enqueued here.
void sampler_update_mixed(ex_sampler* s, int j, int next, bool gram_done) {
    ex_factors* f = s->f;
    hipStream_t st = s->ctx->c.stream;
    const int R = f->R;
    const int64_t I = f->dims[j];
    f->sync32(j);
    const float* U = f->U32[j].as<float>();
    Span sp(s->ctx, s->kind == EX_SAMPLER_CP_ARLS_LEV ? "arls_update" : "tree_build");
    if (s->kind == EX_SAMPLER_CP_ARLS_LEV) {
        s->gtmp.ensure((size_t)R * R * 8);
        s->gpinv.ensure((size_t)R * R * 8);
        k::gram_bf16(st, U, I, R, s->gtmp.as<double>(), s->part);
        s->pivflag.ensure(8);
        k::spd_pinv(st, s->gtmp.as<double>(), R, 1e-11, s->gpinv.as<double>(), s->pivflag.as<int>());
        s->scores[j].ensure((size_t)I * 8);
        s->cdf[j].ensure((size_t)I * 8);
        k::dagat32(st, U, I, R, s->gpinv.as<double>(), s->scores[j].as<double>());
        k::arls_cdf(st, s->scores[j].as<double>(), I, s->cdf[j].as<double>(), s->sums.as<double>() + j, s->part);
    } else {
        // the row tree on st; beside it on the second stream U_j^T U_j (the
        // tree's root, from the same BF16 products) and, for the next target
        // mode, the draw set-up, which reads only the Grams
        hipStream_t aux = s->ctx->aux;
        if (!gram_done) {
            EX_HIP(hipEventRecord(s->ctx->ev_pfork, st));
            EX_HIP(hipStreamWaitEvent(aux, s->ctx->ev_pfork, 0));
        }
        k::tree_build_f32(st, U, s->g[j], R, s->nodes[j].as<float>());
        if (!gram_done) {
            k::gram_bf16(aux, U, I, R, s->grams.as<double>() + (int64_t)j * R * R, s->part);
            if (next >= 0) sts_prepare_m(s, aux, next);
            EX_HIP(hipEventRecord(s->ctx->ev_pjoin, aux));
        }
        if (!gram_done || next >= 0) EX_HIP(hipStreamWaitEvent(st, s->ctx->ev_pjoin, 0));
        s->m_ready = next;
    }
    sp.end();
}

void sampler_update_impl(ex_sampler* s, int j, int next = -1, bool gram_done = false) {
    s->m_ready = -1;
    if (s->precision == EX_MIXED) return sampler_update_mixed(s, j, next, gram_done);
    ex_factors* f = s->f;
    hipStream_t st = s->ctx->c.stream;
    const int R = f->R;
    const int64_t I = f->dims[j];
    f->sync64(j);
    const double* U = f->U[j].as<double>();
    Span sp(s->ctx, s->kind == EX_SAMPLER_CP_ARLS_LEV ? "arls_update" : "tree_build");
    if (s->kind == EX_SAMPLER_CP_ARLS_LEV) {
        // cp_arls_lev.hpp:25-34: G = U^T U, pinv (tol 1e-11), l_i = u_i^T G+ u_i
        s->gtmp.ensure((size_t)R * R * 8);
        s->gpinv.ensure((size_t)R * R * 8);
        k::gram(st, U, I, R, s->gtmp.as<double>(), s->part);
        k::pinv(st, s->gtmp.as<double>(), R, 1e-11, s->gpinv.as<double>());
        s->scores[j].ensure((size_t)I * 8);
        s->cdf[j].ensure((size_t)I * 8);
        k::dagat(st, U, I, R, s->gpinv.as<double>(), s->scores[j].as<double>());
        k::arls_cdf(st, s->scores[j].as<double>(), I, s->cdf[j].as<double>(), s->sums.as<double>() + j, s->part);
    } else {
        // sts_cp.hpp:63-65 -> ExactLeverageTree::build_tree: the local tree and
        // ancestor_grams[0] = compute_gram(U_j), which is the tree's root
        k::tree_build(st, U, s->g[j], R, s->nodes[j].as<double>());
        k::unpack_root<double>(st, s->nodes[j].as<double>(), R, s->grams.as<double>() + (int64_t)j * R * R);
    }
    sp.end();
}

// Rebuilds every mode's state after a batch advanced the factors.
void sampler_refresh(ex_sampler* s) {
    if (!s || !s->refresh) return;
    for (int k2 = 0; k2 < s->f->N; k2++) sampler_update_impl(s, k2);
    s->refresh = false;
}

void sampler_draw_impl(ex_sampler* s, int64_t J, int j, uint64_t sweep, const double* draws_row_dev,
                       const double* draws_eig_dev, uint32_t* samples, double* weights) {
    ex_factors* f = s->f;
    hipStream_t st = s->ctx->c.stream;
    const int N = f->N, R = f->R;
    if (s->kind == EX_SAMPLER_CP_ARLS_LEV) {
        const double* cdfs[kMaxModes];
        const double* sc[kMaxModes];
        for (int k2 = 0; k2 < N; k2++) {
            cdfs[k2] = s->cdf[k2].as<double>();
            sc[k2] = s->scores[k2].as<double>();
        }
        Span sp(s->ctx, "arls_draw");
        k::arls_draw(st, N, j, J, cdfs, sc, s->sums.as<double>(), f->dims, s->seed, sweep, draws_row_dev, samples,
                     weights, s->sweep_dev);
        sp.end();
    } else {
        const bool mixed = s->precision == EX_MIXED;
        Span sm(s->ctx, "sts_compute_m");
        if (s->m_ready != j) sts_prepare_m(s, st, j);
        s->m_ready = -1;
        sm.end();
        k::StsModeArgs m[kMaxModes];
        if (mixed) f->sync32_all();
        else f->sync64_all();
        for (int k2 = 0; k2 < N; k2++) {
            m[k2].rows = mixed ? (const void*)f->U32[k2].p : (const void*)f->U[k2].p;
            m[k2].nodes = s->nodes[k2].p;
            if (mixed) m[k2].eig = s->eig32.as<float>() + (int64_t)k2 * s->ge.node_count * tri_stride32(R);
            else m[k2].eig = s->eig.as<double>() + (int64_t)k2 * s->ge.node_count * tri_stride(R);
            int off = (k2 + 1 == j) ? k2 + 2 : k2 + 1;
            m[k2].w = s->M.as<double>() + (int64_t)off * R * R;
            m[k2].g = s->g[k2];
            m[k2].ge = s->ge;
        }
        Span sd(s->ctx, "sts_draw");
        // mixed: the first non-target mode's draws search prefix sums of its
        // leaf masses (its running product is 1, so R queries cover them all)
        const double* first_cdf = nullptr;
        if (mixed && tune_value(kTuneFirstF32) == 0) {
            const int k0 = j == 0 ? 1 : 0;
            const size_t L0 = s->g[k0].leaf_count;
            s->first_cdf.ensure((size_t)R * (L0 + (L0 + 31) / 32) * 8);
            k::sts_first_cdf(st, m[k0], R, s->first_cdf.as<double>());
            first_cdf = s->first_cdf.as<double>();
        }
        k::sts_draw(st, N, j, R, J, m, s->pinvp.as<double>(), s->seed, sweep, draws_row_dev, draws_eig_dev, samples,
                    weights, mixed ? 4 : 8, s->sweep_dev, s->dropped.as<unsigned long long>(), first_cdf);
        sd.end();
    }
}

void tensor_build_lookups(ex_tensor* t) {
    if (t->built) return;
    hipStream_t st = t->ctx->c.stream;
    for (int mode = 0; mode < t->N; mode++)
        k::build_lookup(st, t->coords.as<uint32_t>(), t->vals.as<double>(), t->nnz, t->N, mode, t->dims, t->lk[mode],
                        t->tmp);
    EX_HIP(hipStreamSynchronize(st));
    for (auto& b : t->tmp) b.release();
    t->built = true;
}

hipEvent_t als_event(ex_als* a) {
    if (!a->ev_pool.empty()) {
        hipEvent_t e = a->ev_pool.back();
        a->ev_pool.pop_back();
        return e;
    }
    hipEvent_t e;
    EX_HIP(hipEventCreate(&e));
    return e;
}

// Adds the region times of every finished step to t_region (waits for them).
void als_resolve_regions(ex_als* a) {
    static const int slot[5] = {0, 2, 3, 5, 6};   // R1, R3, R4, R6, R7 (als_optimizer.hpp:21-27)
    for (size_t b = 0; b + 6 <= a->ev_pending.size(); b += 6) {
        EX_HIP(hipEventSynchronize(a->ev_pending[b + 5]));
        for (int r = 0; r < 5; r++) {
            float ms = 0.f;
            EX_HIP(hipEventElapsedTime(&ms, a->ev_pending[b + r], a->ev_pending[b + r + 1]));
            a->t_region[slot[r]] += ms * 1e-3;
        }
    }
    for (hipEvent_t e : a->ev_pending) a->ev_pool.push_back(e);
    a->ev_pending.clear();
}

void als_step_impl(ex_als* a, int j, int64_t J, uint64_t sweep) {
    ex_factors* f = a->f;
    ex_sampler* s = a->s;
    hipStream_t st = a->ctx->c.stream;
    const int N = f->N, R = f->R;
    const int64_t Ij = f->dims[j];
    EX_CHECK(J > 0 && J < (int64_t)1 << 31, EX_ERR_ARG, "sample count J must be in [1, 2^31)");
    hipEvent_t ev[6] = {};
    const bool regions = !a->capturing;   // region events are not captured
    if (regions)
        for (int q = 0; q < 6; q++) ev[q] = als_event(a);
    auto mark = [&](int q) {
        if (regions) EX_HIP(hipEventRecord(ev[q], st));
    };
    mark(0);
    // a large Y is cleared on the second stream while the draw runs (the
    // previous step's factor update, its last reader, is before this point
    // on st); a small one inline, where the fork and join would cost more
    const bool mixed = a->precision == EX_MIXED;
    const size_t vb = mixed ? 4 : 8;   // bytes per design / MTTKRP value
    const size_t ybytes = (size_t)Ij * R * vb;
    a->Y.ensure(ybytes);
    hipStream_t aux = a->ctx->aux;
    const bool early_zero = ybytes >= ((size_t)8 << 20);
    if (early_zero) {
        EX_HIP(hipEventRecord(a->ctx->ev_zfork, st));
        EX_HIP(hipStreamWaitEvent(aux, a->ctx->ev_zfork, 0));
        k::zero_bytes(aux, a->Y.p, ybytes);
        EX_HIP(hipEventRecord(a->ctx->ev_zjoin, aux));
    }
    // Region 1: draw samples (accumulator_stationary.hpp:148-153)
    s->samples.ensure((size_t)J * N * 4);
    s->weights.ensure((size_t)J * 8);
    sampler_draw_impl(s, J, j, sweep, nullptr, nullptr, s->samples.as<uint32_t>(), s->weights.as<double>());
    mark(1);
    // Region 3: dedup, design, Gram, pinv (:176-250)
    a->sd.ensure((size_t)J * N * 4);
    a->wd.ensure((size_t)J * 8);
    a->kd.ensure((size_t)J * 8);
    const KeyLayout kl = a->t->lk[j].kl;
    Span s_dd(a->ctx, "dedup");
    // J_u stays on the device: every later kernel reads it; J bounds the sizes
    uint64_t* cnt = a->cnt.as<uint64_t>();
    const uint64_t* ju = cnt + 3;
    if (kl.bits <= 63)
        k::dedup_hash(st, s->samples.as<uint32_t>(), s->weights.as<double>(), J, N, kl, a->sd.as<uint32_t>(),
                      a->wd.as<double>(), a->kd.as<uint64_t>(), a->dht, a->dhw, cnt + 3, cnt + 2);
    else
        k::dedup(st, s->samples.as<uint32_t>(), s->weights.as<double>(), J, N, kl, a->sd.as<uint32_t>(),
                 a->wd.as<double>(), a->kd.as<uint64_t>(), a->tmp, cnt + 3, cnt + 2);
    const int64_t Ju = J;
    a->dg.ensure((size_t)J * R * vb);
    a->dm.ensure((size_t)J * R * vb);
    a->G.ensure((size_t)R * R * 8);
    a->Gi.ensure((size_t)R * R * 8);
    a->pivflag.ensure(8);
    s_dd.end();
    Span s_de(a->ctx, "design_gather");
    if (mixed) {
        f->sync32_all();
        k::design32(st, a->sd.as<uint32_t>(), a->wd.as<double>(), Ju, N, j, R, f->ptrs32, a->dg.as<float>(),
                    a->dm.as<float>(), ju);
    } else {
        f->sync64_all();
        k::design(st, a->sd.as<uint32_t>(), a->wd.as<double>(), Ju, N, j, R, f->ptrs, a->dg.as<double>(),
                  a->dm.as<double>(), ju);
    }
    s_de.end();
    // the design Gram and its inverse run on the second stream beside the
    // sampled MTTKRP (which needs only the design rows); the factor update
    // waits for them
    EX_HIP(hipEventRecord(a->ctx->ev_fork, st));
    EX_HIP(hipStreamWaitEvent(aux, a->ctx->ev_fork, 0));
    Span s_dp(a->ctx, "design_gram_pinv", aux);
    if (mixed) {
        // FP32 rows, FP64 products and sums: the rank decision of the solve
        // below then sees the design's own rounding (2^-24), not BF16's
        k::gram_f32in(aux, a->dg.as<float>(), Ju, R, a->G.as<double>(), a->part, ju);
        k::spd_pinv(aux, a->G.as<double>(), R, 1e-11, a->Gi.as<double>(), a->pivflag.as<int>());
    } else {
        k::gram(aux, a->dg.as<double>(), Ju, R, a->G.as<double>(), a->part, ju);
        k::pinv(aux, a->G.as<double>(), R, 1e-11, a->Gi.as<double>());
    }
    s_dp.end();
    EX_HIP(hipEventRecord(a->ctx->ev_join, aux));
    mark(2);
    // Region 4: sampled MTTKRP (:252-274)
    ModeLookup& L = a->t->lk[j];
    if (mixed && !L.rval32.p) {
        L.rval32.ensure((size_t)(L.nnz > 0 ? L.nnz : 1) * 4);
        k::to_f32(st, L.rval.as<double>(), L.nnz, L.rval32.as<float>());
    }
    Span s_mt(a->ctx, "sampled_mttkrp");
    hipEvent_t scan_ev[2] = {nullptr, nullptr};
    if (a->ctx->timing) {
        scan_ev[0] = a->ctx->get_event();
        scan_ev[1] = a->ctx->get_event();
    }
    // hits and touched rows accumulate on the device (a->cnt)
    if (early_zero) EX_HIP(hipStreamWaitEvent(st, a->ctx->ev_zjoin, 0));
    if (mixed)
        k::sampled_mttkrp<float>(st, L, L.rval32.as<float>(), a->kd.as<uint64_t>(), Ju, a->dm.as<float>(), R,
                                 a->Y.as<float>(), Ij, a->tmp, cnt, a->ctx->timing ? scan_ev : nullptr, ju,
                                 !early_zero);
    else
        k::sampled_mttkrp<double>(st, L, L.rval.as<double>(), a->kd.as<uint64_t>(), Ju, a->dm.as<double>(), R,
                                  a->Y.as<double>(), Ij, a->tmp, cnt, a->ctx->timing ? scan_ev : nullptr, ju,
                                  !early_zero);
    s_mt.end();
    if (a->ctx->timing) {
        a->ctx->pending.push_back({"row_scan", {scan_ev[0], scan_ev[1]}});
        a->ctx->watch(cnt + 0, "H");
        a->ctx->watch(cnt + 1, "T");
        a->ctx->watch(cnt + 2, "J_u");
        a->ctx->add_counter("mode_steps", 1.0);
        a->ctx->add_counter("nnz", (double)L.nnz);
        a->ctx->add_counter("nfib", (double)L.nfib);
        a->ctx->add_counter("I", (double)Ij);
    }
    mark(3);
    // Region 6: U_j = Y G+, renormalise (:280-289)
    const bool sts_mixed = mixed && s->kind == EX_SAMPLER_STS_CP;
    const int next = sts_mixed && j + 1 < N ? j + 1 : -1;
    EX_HIP(hipStreamWaitEvent(st, a->ctx->ev_join, 0));
    Span s_fu(a->ctx, "factor_update");
    if (mixed) {
        // U_j = Y Gi diag(1 / sigma) in one pass over Y: sigma from the Gram of Y
        float* Uj = f->U32[j].as<float>();
        a->GY.ensure((size_t)R * R * 8);
        a->Ms.ensure((size_t)R * R * 4);
        k::gram_f32in(st, a->Y.as<float>(), Ij, R, a->GY.as<double>(), a->part);
        // STS-CP: U_j^T U_j from the same small matrices, so the next mode's
        // draw set-up (which reads only the Grams) runs on the second stream
        // beside the pass over Y and the row-tree rebuild
        double* gj = sts_mixed ? s->grams.as<double>() + (int64_t)j * R * R : nullptr;
        k::sigma_from_gram(st, a->GY.as<double>(), a->Gi.as<double>(), R, f->sigma.as<double>(), a->Ms.as<float>(),
                           gj);
        if (next >= 0) {
            EX_HIP(hipEventRecord(a->ctx->ev_pfork, st));
            EX_HIP(hipStreamWaitEvent(aux, a->ctx->ev_pfork, 0));
            sts_prepare_m(s, aux, next);
            EX_HIP(hipEventRecord(a->ctx->ev_pjoin, aux));
        }
        k::dsymm32(st, a->Ms.as<float>(), a->Y.as<float>(), Ij, R, Uj);
        f->wrote32(j);
    } else {
        double* Uj = f->U[j].as<double>();
        k::dsymm(st, a->Gi.as<double>(), a->Y.as<double>(), Ij, R, Uj);
        k::col_sumsq(st, Uj, Ij, R, f->sigma.as<double>(), a->part);
        k::scale_cols(st, Uj, Ij, R, f->sigma.as<double>());
        f->wrote64(j);
    }
    s_fu.end();
    mark(4);
    // Region 7: sampler update (:291-294)
    sampler_update_impl(s, j, next, sts_mixed);
    mark(5);
    // region times are read back later (als_resolve_regions): no wait here
    if (!regions) return;
    for (int q = 0; q < 6; q++) a->ev_pending.push_back(ev[q]);
    if (a->ev_pending.size() >= 6 * 64) als_resolve_regions(a);
}

__global__ void set_u64(uint64_t* p, uint64_t v) { *p = v; }

void exact_step_impl(ex_als* a, int j);

// Addresses of every device buffer a sweep can touch.  A captured graph
// holds raw addresses, so a buffer that grew (and moved) since the capture,
// e.g. one the fit enlarges, forces a new capture.
std::vector<const void*> graph_fingerprint(const ex_als* a) {
    std::vector<const void*> v;
    auto add = [&](const DevBuf& b) { v.push_back(b.p); };
    const ex_sampler* s = a->s;
    const ex_factors* f = a->f;
    const ex_tensor* t = a->t;
    for (const DevBuf* b : {&a->sd, &a->wd, &a->kd, &a->dg, &a->dm, &a->G, &a->Gi, &a->pivflag, &a->GY, &a->Ms, &a->Y,
                            &a->scal, &a->part, &a->cnt, &a->grams, &a->sweep_ctr, &a->dht, &a->dhw})
        add(*b);
    for (const DevBuf& b : a->tmp) add(b);
    for (const DevBuf* b : {&s->sums, &s->gtmp, &s->gpinv, &s->grams, &s->M, &s->lam, &s->eig, &s->eig32, &s->pinvp,
                            &s->part, &s->samples, &s->weights, &s->draws_row, &s->draws_eig, &s->pivflag})
        add(*b);
    for (int k2 = 0; k2 < kMaxModes; k2++) {
        add(s->scores[k2]);
        add(s->cdf[k2]);
        add(s->nodes[k2]);
        add(f->U[k2]);
        add(f->U32[k2]);
        const ModeLookup& L = t->lk[k2];
        for (const DevBuf* b : {&L.fkey, &L.rrow, &L.rfib, &L.rval, &L.rval32, &L.cmap, &L.filt, &L.fptr, &L.kpos,
                                &L.cpos, &L.marks, &L.cpart, &L.krow, &L.kval, &L.kval32})
            add(*b);
    }
    for (const DevBuf* b : {&f->sigma, &f->part, &f->gram, &t->coords, &t->vals}) add(*b);
    for (const DevBuf& b : t->tmp) add(b);
    for (const DevBuf& b : a->ctx->c.scratch) add(b);
    return v;
}

void als_drop_graph(ex_als* a) {
    if (a->gexec) (void)hipGraphExecDestroy(a->gexec);
    if (a->graph) (void)hipGraphDestroy(a->graph);
    a->gexec = nullptr;
    a->graph = nullptr;
    a->graph_J = -1;
}

// Captures one sweep (every mode step at sample count J) into a graph whose
// draws read the sweep number from a->sweep_ctr.  The step's buffers must be
// sized already (an eager sweep at the same J).  A captured memset is
// refused: memset nodes have given wrong results on replays here.
void als_capture_sweep(ex_als* a, int64_t J) {
    hipStream_t st = a->ctx->c.stream;
    als_drop_graph(a);
    a->sweep_ctr.ensure(8);
    // bring the factor copy the steps read up to date now: a conversion
    // enqueued inside the capture would be replayed by every later sweep
    if (a->precision == EX_MIXED) a->f->sync32_all();
    else a->f->sync64_all();
    EX_HIP(hipStreamSynchronize(st));
    a->s->sweep_dev = a->sweep_ctr.as<uint64_t>();
    a->capturing = true;
    hipGraph_t g = nullptr;
    EX_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    try {
        for (int j = 0; j < a->f->N; j++) als_step_impl(a, j, J, 0);
    } catch (...) {
        (void)hipStreamEndCapture(st, &g);
        if (g) (void)hipGraphDestroy(g);
        a->capturing = false;
        a->s->sweep_dev = nullptr;
        throw;
    }
    const hipError_t e = hipStreamEndCapture(st, &g);
    a->capturing = false;
    a->s->sweep_dev = nullptr;
    EX_HIP(e);
    a->graph = g;
    size_t n = 0;
    EX_HIP(hipGraphGetNodes(g, nullptr, &n));
    std::vector<hipGraphNode_t> nodes(n);
    if (n) EX_HIP(hipGraphGetNodes(g, nodes.data(), &n));
    for (hipGraphNode_t node : nodes) {
        hipGraphNodeType t;
        EX_HIP(hipGraphNodeGetType(node, &t));
        if (t == hipGraphNodeTypeMemset) {
            als_drop_graph(a);
            EX_CHECK(false, EX_ERR_UNSUPPORTED, "a captured sweep holds a memset node; run without graphs");
        }
    }
    EX_HIP(hipGraphInstantiate(&a->gexec, g, nullptr, nullptr, 0));
    a->graph_J = J;
    a->graph_fp = graph_fingerprint(a);
}

// One sweep at sample count J and sweep number `round`: eager, or with
// graphs on a replay of the captured sweep.  The first sweep at a given J
// runs eagerly (it sizes every buffer) and then captures one.  With the
// context's span timing on, sweeps run eagerly (the spans are host-side
// records of events enqueued step by step).
void als_sweep(ex_als* a, int64_t J, uint64_t round) {
    if (a->exact) {
        for (int j = 0; j < a->f->N; j++) exact_step_impl(a, j);
        return;
    }
    const bool replay = a->use_graph && !a->ctx->timing;
    if (!replay || a->graph_J != J) {
        for (int j = 0; j < a->f->N; j++) als_step_impl(a, j, J, round);
        if (replay) als_capture_sweep(a, J);
        return;
    }
    hipStream_t st = a->ctx->c.stream;
    ex_factors* f = a->f;
    const bool mixed = a->precision == EX_MIXED;
    // the host-side bookkeeping of the factor copies that the captured steps
    // did once: bring the copy the steps read up to date (a host write since
    // the capture) before any new capture, and mark every mode's other copy
    // stale afterwards
    if (mixed) f->sync32_all();
    else f->sync64_all();
    if (graph_fingerprint(a) != a->graph_fp) als_capture_sweep(a, J);
    hipLaunchKernelGGL(set_u64, dim3(1), dim3(1), 0, st, a->sweep_ctr.as<uint64_t>(), round);
    EX_HIP(hipGetLastError());
    EX_HIP(hipGraphLaunch(a->gexec, st));
    for (int j = 0; j < f->N; j++) {
        if (mixed) f->wrote32(j);
        else f->wrote64(j);
    }
}

// Grams of every factor into a->grams (FP64 products and sums in both
// precisions; the mixed path reads the FP32 factors).
void exact_grams(ex_als* a) {
    ex_factors* f = a->f;
    hipStream_t st = a->ctx->c.stream;
    const int N = f->N, R = f->R;
    a->grams.ensure((size_t)N * R * R * 8);
    for (int k2 = 0; k2 < N; k2++) {
        double* Gk = a->grams.as<double>() + (int64_t)k2 * R * R;
        if (a->precision == EX_MIXED) {
            f->sync32(k2);
            k::gram_f32in(st, f->U32[k2].as<float>(), f->dims[k2], R, Gk, a->part);
        } else {
            f->sync64(k2);
            k::gram(st, f->U[k2].as<double>(), f->dims[k2], R, Gk, a->part);
        }
    }
    a->grams_valid = true;
}

// ExactALS::execute_ALS_step (exact_als.hpp:95-141): G = Hadamard product of
// the other factors' Grams, G+ (compute_pinv_square, tolerance 1e-11), the
// exact MTTKRP Y, U_j = Y G+, column renormalisation into sigma, then the new
// Gram of U_j.  Region times: R3 = Hadamard + inverse, R4 = MTTKRP, R6 =
// update, R7 = Gram refresh (the reference leaves them at zero for ExactALS).
void exact_step_impl(ex_als* a, int j) {
    ex_factors* f = a->f;
    hipStream_t st = a->ctx->c.stream;
    const int N = f->N, R = f->R;
    const int64_t Ij = f->dims[j];
    const bool mixed = a->precision == EX_MIXED;
    if (!a->grams_valid) exact_grams(a);
    hipEvent_t ev[6];
    for (int q = 0; q < 6; q++) ev[q] = als_event(a);
    auto mark = [&](int q) { EX_HIP(hipEventRecord(ev[q], st)); };
    mark(0);
    mark(1);
    a->G.ensure((size_t)R * R * 8);
    a->Gi.ensure((size_t)R * R * 8);
    a->pivflag.ensure(8);
    Span s_gp(a->ctx, "exact_gram_pinv");
    k::chain_had(st, a->grams.as<double>(), N, j, R, a->G.as<double>());
    if (mixed) k::spd_pinv(st, a->G.as<double>(), R, 1e-11, a->Gi.as<double>(), a->pivflag.as<int>());
    else k::pinv(st, a->G.as<double>(), R, 1e-11, a->Gi.as<double>());
    s_gp.end();
    mark(2);
    ModeLookup& L = a->t->lk[j];
    const size_t vb = mixed ? 4 : 8;
    a->Y.ensure((size_t)Ij * R * vb);
    Span s_mt(a->ctx, "exact_mttkrp");
    if (mixed) {
        if (!L.rval32.p) {
            L.rval32.ensure((size_t)(L.nnz > 0 ? L.nnz : 1) * 4);
            k::to_f32(st, L.rval.as<double>(), L.nnz, L.rval32.as<float>());
        }
        f->sync32_all();
        k::exact_mttkrp<float>(st, L, L.rval32.as<float>(), f->ptrs32, f->dims, R, a->Y.as<float>(), Ij, a->tmp);
    } else {
        f->sync64_all();
        k::exact_mttkrp<double>(st, L, L.rval.as<double>(), f->ptrs, f->dims, R, a->Y.as<double>(), Ij, a->tmp);
    }
    s_mt.end();
    mark(3);
    Span s_fu(a->ctx, "factor_update");
    double* Gj = a->grams.as<double>() + (int64_t)j * R * R;
    if (mixed) {
        float* Uj = f->U32[j].as<float>();
        a->GY.ensure((size_t)R * R * 8);
        a->Ms.ensure((size_t)R * R * 4);
        k::gram_f32in(st, a->Y.as<float>(), Ij, R, a->GY.as<double>(), a->part);
        k::sigma_from_gram(st, a->GY.as<double>(), a->Gi.as<double>(), R, f->sigma.as<double>(), a->Ms.as<float>());
        k::dsymm32(st, a->Ms.as<float>(), a->Y.as<float>(), Ij, R, Uj);
        f->wrote32(j);
    } else {
        double* Uj = f->U[j].as<double>();
        k::dsymm(st, a->Gi.as<double>(), a->Y.as<double>(), Ij, R, Uj);
        k::col_sumsq(st, Uj, Ij, R, f->sigma.as<double>(), a->part);
        k::scale_cols(st, Uj, Ij, R, f->sigma.as<double>());
        f->wrote64(j);
    }
    s_fu.end();
    mark(4);
    Span s_gr(a->ctx, "exact_gram_update");
    if (mixed) k::gram_f32in(st, f->U32[j].as<float>(), Ij, R, Gj, a->part);
    else k::gram(st, f->U[j].as<double>(), Ij, R, Gj, a->part);
    s_gr.end();
    mark(5);
    for (int q = 0; q < 6; q++) a->ev_pending.push_back(ev[q]);
    if (a->ev_pending.size() >= 6 * 64) als_resolve_regions(a);
}

double als_fit_impl(ex_als* a) {
    ex_factors* f = a->f;
    hipStream_t st = a->ctx->c.stream;
    const int N = f->N, R = f->R;
    a->scal.ensure(64);
    a->G.ensure((size_t)N * R * R * 8 > (size_t)R * R * 8 ? (size_t)N * R * R * 8 : (size_t)R * R * 8);
    DevBuf grams;
    grams.ensure((size_t)N * R * R * 8);
    f->sync64_all();
    for (int k2 = 0; k2 < N; k2++)
        k::gram(st, f->U[k2].as<double>(), f->dims[k2], R, grams.as<double>() + (int64_t)k2 * R * R, a->part);
    double* d = a->scal.as<double>();
    k::chain_had_norm(st, grams.as<double>(), N, R, f->sigma.as<double>(), d);
    DevBuf fpart;
    k::fit_bmb(st, a->t->coords.as<uint32_t>(), a->t->vals.as<double>(), a->t->nnz, N, f->ptrs,
               f->sigma.as<double>(), R, d + 1, fpart);
    double h[2];
    d2h(st, h, d, sizeof(h));
    return 1.0 - std::sqrt(h[0] + h[1]) / a->t->norm;
}

// Keys of device sample tuples for `mode`, checked distinct (the sampled
// MTTKRP sums each fiber once per sample row).
// The samples' fiber keys in their order (t->adj.kd) and, at t->adj.cnt,
// the device count of distinct keys, which equals Ju exactly when the
// samples are deduplicated; enqueued on st.
void smttkrp_keys(ex_tensor* t, hipStream_t st, int mode, int64_t Ju, const uint32_t* samples,
                  const double* weights) {
    AdjScratch& a = t->adj;
    const KeyLayout& kl = t->lk[mode].kl;
    const size_t J1 = (size_t)(Ju > 0 ? Ju : 1);
    a.kd.ensure(J1 * 8);
    a.sd.ensure(J1 * t->N * 4);
    a.wd.ensure(J1 * 8);
    a.cnt.ensure(32);
    uint64_t* cnt = a.cnt.as<uint64_t>();
    if (kl.bits <= 63)
        k::dedup_hash(st, samples, weights, Ju, t->N, kl, a.sd.as<uint32_t>(), a.wd.as<double>(), a.kd.as<uint64_t>(),
                      a.table, a.work, cnt, nullptr);
    else
        k::dedup(st, samples, weights, Ju, t->N, kl, a.sd.as<uint32_t>(), a.wd.as<double>(), a.kd.as<uint64_t>(),
                 a.tmp, cnt, nullptr);
}

// Forward on stream st (nothing waited for); Y is NaN when two samples share
// a fiber key.
void smttkrp_forward_on(ex_tensor* t, hipStream_t st, int mode, int64_t Ju, const uint32_t* samples,
                        const double* weights, const double* const* factors, int R, double* Y) {
    AdjScratch& a = t->adj;
    ModeLookup& L = t->lk[mode];
    const int64_t Ij = t->dims[mode];
    const size_t J1 = (size_t)(Ju > 0 ? Ju : 1);
    if (Ju > 0) smttkrp_keys(t, st, mode, Ju, samples, weights);
    a.dg.ensure(J1 * R * 8);
    a.dm.ensure(J1 * R * 8);
    k::design(st, samples, weights, Ju, t->N, mode, R, factors, a.dg.as<double>(), a.dm.as<double>());
    k::sampled_mttkrp<double>(st, L, L.rval.as<double>(), a.kd.as<uint64_t>(), Ju, a.dm.as<double>(), R, Y, Ij,
                              a.tmp);
    if (Ju > 0) k::poison_unless(st, a.cnt.as<uint64_t>(), Ju, Y, Ij * R);
}

// Backward on stream st (nothing waited for); every requested gradient is NaN
// when two samples share a fiber key.
void smttkrp_backward_on(ex_tensor* t, hipStream_t st, int mode, int64_t Ju, const uint32_t* samples,
                         const double* weights, const double* const* factors, int R, const double* G,
                         double* const* grads) {
    AdjScratch& a = t->adj;
    if (Ju == 0) return;
    smttkrp_keys(t, st, mode, Ju, samples, weights);
    a.dD.ensure((size_t)Ju * R * 8);
    double* gp[kMaxModes] = {};
    for (int k2 = 0; k2 < t->N; k2++) gp[k2] = k2 == mode ? nullptr : grads[k2];
    k::smttkrp_adjoint(st, t->lk[mode], a.kd.as<uint64_t>(), samples, weights, Ju, t->N, mode, R, factors, G, gp,
                       a.dD.as<double>());
    for (int k2 = 0; k2 < t->N; k2++)
        if (gp[k2]) k::poison_unless(st, a.cnt.as<uint64_t>(), Ju, gp[k2], t->dims[k2] * R);
}

// Whether the last smttkrp_keys found the samples deduplicated (waits for st).
bool smttkrp_distinct(ex_tensor* t, hipStream_t st, int64_t Ju) {
    if (Ju == 0) return true;
    uint64_t c = 0;
    d2h(st, &c, t->adj.cnt.p, 8);
    return (int64_t)c == Ju;
}

void smttkrp_check(ex_tensor* t, int mode, int64_t Ju, const void* samples, const void* weights,
                   const double* const* factors, int64_t R) {
    EX_CHECK(t && factors, EX_ERR_ARG, "null argument");
    EX_CHECK(mode >= 0 && mode < t->N, EX_ERR_ARG, "mode out of range");
    EX_CHECK(R >= 1 && R <= EX_MAX_RANK, EX_ERR_UNSUPPORTED, "rank must be in [1, 64]");
    EX_CHECK(Ju >= 0 && Ju < (int64_t)1 << 31, EX_ERR_ARG, "sample count out of range");
    EX_CHECK(Ju == 0 || (samples && weights), EX_ERR_ARG, "null samples or weights");
    for (int k2 = 0; k2 < t->N; k2++)
        EX_CHECK(k2 == mode || factors[k2], EX_ERR_ARG, "null factor pointer");
}
}  // namespace

extern "C" {

const char* ex_last_error(void) { return g_last_error.c_str(); }

int ex_ctx_create(int device, ex_ctx** out) {
    return guarded([&] {
        EX_CHECK(out, EX_ERR_ARG, "null output pointer");
        int n = 0;
        EX_HIP(hipGetDeviceCount(&n));
        EX_CHECK(device >= 0 && device < n, EX_ERR_ARG, "device index out of range");
        EX_CHECK(device < kMaxDevices, EX_ERR_UNSUPPORTED, "device index above 63");
        load_default_tuning();
        auto c = std::make_unique<ex_ctx>();
        c->c.device = device;
        EX_HIP(hipSetDevice(device));
        lookup_static_scratch(device);
        EX_HIP(hipGetDeviceProperties(&c->c.prop, device));
        EX_HIP(hipStreamCreateWithFlags(&c->c.stream, hipStreamNonBlocking));
        EX_HIP(hipStreamCreateWithFlags(&c->aux, hipStreamNonBlocking));
        EX_HIP(hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming));
        EX_HIP(hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming));
        EX_HIP(hipEventCreateWithFlags(&c->ev_zfork, hipEventDisableTiming));
        EX_HIP(hipEventCreateWithFlags(&c->ev_zjoin, hipEventDisableTiming));
        EX_HIP(hipEventCreateWithFlags(&c->ev_pfork, hipEventDisableTiming));
        EX_HIP(hipEventCreateWithFlags(&c->ev_pjoin, hipEventDisableTiming));
        *out = c.release();
    });
}

int ex_ctx_destroy(ex_ctx* ctx) {
    return guarded([&] {
        if (!ctx) return;
        (void)hipStreamSynchronize(ctx->c.stream);
        (void)hipStreamSynchronize(ctx->aux);
        (void)hipEventDestroy(ctx->ev_fork);
        (void)hipEventDestroy(ctx->ev_join);
        (void)hipEventDestroy(ctx->ev_zfork);
        (void)hipEventDestroy(ctx->ev_zjoin);
        (void)hipEventDestroy(ctx->ev_pfork);
        (void)hipEventDestroy(ctx->ev_pjoin);
        (void)hipStreamDestroy(ctx->aux);
        (void)hipStreamDestroy(ctx->c.stream);
        delete ctx;
    });
}

int ex_ctx_sync(ex_ctx* ctx) {
    return guarded([&] {
        EX_CHECK(ctx, EX_ERR_ARG, "null context");
        EX_HIP(hipStreamSynchronize(ctx->c.stream));
    });
}

int ex_ctx_timing(ex_ctx* ctx, int enable) {
    return guarded([&] {
        EX_CHECK(ctx, EX_ERR_ARG, "null context");
        EX_CHECK(enable >= 0 && enable <= 2, EX_ERR_ARG, "enable must be 0, 1 or 2");
        EX_HIP(hipStreamSynchronize(ctx->c.stream));
        ctx->resolve();
        ctx->timing = enable != 0;
        if (enable == 1) {
            ctx->totals.clear();
            ctx->counters.clear();
            ctx->watches.clear();
        }
    });
}

int ex_ctx_timing_json(ex_ctx* ctx, char* buf, size_t* len) {
    return guarded([&] {
        EX_CHECK(ctx && len, EX_ERR_ARG, "null argument");
        EX_HIP(hipStreamSynchronize(ctx->c.stream));
        ctx->resolve();
        std::string js = "{\"spans\": {";
        for (size_t i = 0; i < ctx->totals.size(); i++) {
            if (i) js += ", ";
            js += "\"" + ctx->totals[i].first + "\": {\"ms\": " + json_num(ctx->totals[i].second.first) +
                  ", \"count\": " + std::to_string(ctx->totals[i].second.second) + "}";
        }
        js += "}, \"counters\": {";
        for (size_t i = 0; i < ctx->counters.size(); i++) {
            if (i) js += ", ";
            js += "\"" + ctx->counters[i].first + "\": " + json_num(ctx->counters[i].second);
        }
        js += "}}";
        size_t need = js.size() + 1;
        if (buf && *len > 0) {
            size_t n = *len < need ? *len - 1 : need - 1;
            std::memcpy(buf, js.data(), n);
            buf[n] = 0;
        }
        *len = need;
    });
}

int ex_ctx_device_name(ex_ctx* ctx, char* buf, size_t len) {
    return guarded([&] {
        EX_CHECK(ctx && buf && len > 0, EX_ERR_ARG, "bad arguments");
        std::strncpy(buf, device_label(ctx).c_str(), len - 1);
        buf[len - 1] = 0;
    });
}

// ---------------------------------------------------------------- tensor
int ex_tensor_create(ex_ctx* ctx, int N, const int64_t* dims, int64_t nnz, const uint32_t* idx, const double* vals,
                     int preprocessing, ex_tensor** out) {
    return guarded([&] {
        EX_CHECK(ctx && dims && out, EX_ERR_ARG, "null argument");
        EX_CHECK(N >= 2 && N <= kMaxModes, EX_ERR_UNSUPPORTED, "mode count must be in [2, 8]");
        EX_CHECK(nnz >= 0 && nnz < ((int64_t)1 << 32), EX_ERR_UNSUPPORTED, "nnz must be below 2^32");
        EX_CHECK(nnz == 0 || (idx && vals), EX_ERR_ARG, "null index or value array");
        EX_CHECK(preprocessing >= 0 && preprocessing <= 2, EX_ERR_ARG, "unknown preprocessing");
        for (int k2 = 0; k2 < N; k2++)
            EX_CHECK(dims[k2] >= 1 && dims[k2] <= 0xFFFFFFFFLL, EX_ERR_ARG, "mode size out of range");
        for (int64_t i = 0; i < nnz; i++)
            for (int k2 = 0; k2 < N; k2++)
                EX_CHECK(idx[i * N + k2] < (uint64_t)dims[k2], EX_ERR_ARG,
                         "coordinate out of range at nonzero " + std::to_string(i) + ", mode " + std::to_string(k2));
        auto t = std::make_unique<ex_tensor>();
        t->ctx = ctx;
        t->N = N;
        for (int k2 = 0; k2 < N; k2++) t->dims[k2] = dims[k2];
        t->nnz = nnz;
        hipStream_t st = ctx->c.stream;
        t->coords.ensure((size_t)nnz * N * 4);
        t->vals.ensure((size_t)nnz * 8);
        h2d(st, t->coords.p, idx, (size_t)nnz * N * 4);
        h2d(st, t->vals.p, vals, (size_t)nnz * 8);
        k::preprocess(st, t->vals.as<double>(), nnz, preprocessing);
        DevBuf part, out1;
        out1.ensure(8);
        k::sumsq(st, t->vals.as<double>(), nnz, out1.as<double>(), part);
        double ss = 0;
        d2h(st, &ss, out1.p, 8);
        t->norm = std::sqrt(ss);
        *out = t.release();
    });
}

int ex_tensor_destroy(ex_tensor* t) {
    return guarded([&] { delete t; });
}

int ex_tensor_norm(ex_tensor* t, double* norm) {
    return guarded([&] {
        EX_CHECK(t && norm, EX_ERR_ARG, "null argument");
        *norm = t->norm;
    });
}

int ex_tensor_build_lookups(ex_tensor* t) {
    return guarded([&] {
        EX_CHECK(t, EX_ERR_ARG, "null tensor");
        tensor_build_lookups(t);
    });
}

int ex_tensor_mttkrp_path(ex_tensor* t, int path) {
    return guarded([&] {
        EX_CHECK(t, EX_ERR_ARG, "null tensor");
        EX_CHECK(path == EX_MTTKRP_AUTO || path == EX_MTTKRP_ROW_PASS || path == EX_MTTKRP_HIT_LIST ||
                     path == EX_MTTKRP_BUCKETS,
                 EX_ERR_ARG, "unknown sampled-MTTKRP path");
        for (int k2 = 0; k2 < kMaxModes; k2++) t->lk[k2].mttkrp_path = path;
    });
}

int ex_sampled_mttkrp(ex_tensor* t, int mode, int64_t Ju, const uint32_t* samples, const double* design, int64_t R,
                      double* Y, int64_t* hits) {
    return guarded([&] {
        EX_CHECK(t && Y, EX_ERR_ARG, "null argument");
        EX_CHECK(mode >= 0 && mode < t->N, EX_ERR_ARG, "mode out of range");
        EX_CHECK(R >= 1 && R <= kMaxRank, EX_ERR_UNSUPPORTED, "rank must be in [1, 64]");
        EX_CHECK(Ju >= 0 && (Ju == 0 || (samples && design)), EX_ERR_ARG, "bad sample arrays");
        tensor_build_lookups(t);
        hipStream_t st = t->ctx->c.stream;
        const int N = t->N;
        const KeyLayout& kl = t->lk[mode].kl;
        std::vector<uint64_t> keys((size_t)Ju);
        for (int64_t c = 0; c < Ju; c++) {
            uint64_t key = 0;
            for (int k2 = 0; k2 < N; k2++) {
                if (k2 == mode) continue;
                EX_CHECK(samples[c * N + k2] < (uint64_t)t->dims[k2], EX_ERR_ARG, "sample index out of range");
                key |= (uint64_t)samples[c * N + k2] << kl.shift[k2];
            }
            keys[(size_t)c] = key;
        }
        std::vector<uint64_t> sorted = keys;
        std::sort(sorted.begin(), sorted.end());
        EX_CHECK(std::adjacent_find(sorted.begin(), sorted.end()) == sorted.end(), EX_ERR_ARG,
                 "samples must be deduplicated: two rows share every non-target coordinate");
        DevBuf dd, dk, dy, dc;
        DevBuf tmp[4];
        dd.ensure((size_t)(Ju > 0 ? Ju : 1) * R * 8);
        dk.ensure((size_t)(Ju > 0 ? Ju : 1) * 8);
        dc.ensure(16);
        const int64_t Ij = t->dims[mode];
        dy.ensure((size_t)Ij * R * 8);
        h2d(st, dk.p, keys.data(), (size_t)Ju * 8);
        h2d(st, dd.p, design, (size_t)Ju * R * 8);
        EX_HIP(hipMemsetAsync(dc.p, 0, 16, st));
        k::sampled_mttkrp<double>(st, t->lk[mode], t->lk[mode].rval.as<double>(), dk.as<uint64_t>(), Ju,
                                  dd.as<double>(), (int)R, dy.as<double>(), Ij, tmp, dc.as<uint64_t>());
        d2h(st, Y, dy.p, (size_t)Ij * R * 8);
        uint64_t c[2];
        d2h(st, c, dc.p, sizeof(c));
        if (hits) *hits = (int64_t)c[0];
    });
}

// ---------------------------------------------------------------- differentiable sampled MTTKRP
int ex_smttkrp_forward(ex_tensor* t, int mode, int64_t Ju, const uint32_t* samples, const double* weights,
                       const double* const* factors, int64_t R, double* Y) {
    return guarded([&] {
        smttkrp_check(t, mode, Ju, samples, weights, factors, R);
        EX_CHECK(Y, EX_ERR_ARG, "null output");
        EX_HIP(hipSetDevice(t->ctx->c.device));
        tensor_build_lookups(t);
        hipStream_t st = t->ctx->c.stream;
        smttkrp_forward_on(t, st, mode, Ju, samples, weights, factors, (int)R, Y);
        EX_CHECK(smttkrp_distinct(t, st, Ju), EX_ERR_ARG,
                 "samples must be deduplicated: two rows share every non-target coordinate");
    });
}

int ex_smttkrp_backward(ex_tensor* t, int mode, int64_t Ju, const uint32_t* samples, const double* weights,
                        const double* const* factors, int64_t R, const double* G, double* const* grads) {
    return guarded([&] {
        smttkrp_check(t, mode, Ju, samples, weights, factors, R);
        EX_CHECK(G && grads, EX_ERR_ARG, "null gradient argument");
        EX_HIP(hipSetDevice(t->ctx->c.device));
        tensor_build_lookups(t);
        hipStream_t st = t->ctx->c.stream;
        smttkrp_backward_on(t, st, mode, Ju, samples, weights, factors, (int)R, G, grads);
        EX_CHECK(smttkrp_distinct(t, st, Ju), EX_ERR_ARG,
                 "samples must be deduplicated: two rows share every non-target coordinate");
        EX_HIP(hipStreamSynchronize(st));
    });
}

int ex_smttkrp_forward_async(ex_tensor* t, void* stream, int mode, int64_t Ju, const uint32_t* samples,
                             const double* weights, const double* const* factors, int64_t R, double* Y) {
    return guarded([&] {
        smttkrp_check(t, mode, Ju, samples, weights, factors, R);
        EX_CHECK(Y, EX_ERR_ARG, "null output");
        EX_HIP(hipSetDevice(t->ctx->c.device));
        tensor_build_lookups(t);
        hipStream_t st = static_cast<hipStream_t>(stream);   // NULL: the null stream
        smttkrp_forward_on(t, st, mode, Ju, samples, weights, factors, (int)R, Y);
        EX_HIP(hipGetLastError());
    });
}

int ex_smttkrp_backward_async(ex_tensor* t, void* stream, int mode, int64_t Ju, const uint32_t* samples,
                              const double* weights, const double* const* factors, int64_t R, const double* G,
                              double* const* grads) {
    return guarded([&] {
        smttkrp_check(t, mode, Ju, samples, weights, factors, R);
        EX_CHECK(G && grads, EX_ERR_ARG, "null gradient argument");
        EX_HIP(hipSetDevice(t->ctx->c.device));
        tensor_build_lookups(t);
        hipStream_t st = static_cast<hipStream_t>(stream);   // NULL: the null stream
        smttkrp_backward_on(t, st, mode, Ju, samples, weights, factors, (int)R, G, grads);
        EX_HIP(hipGetLastError());
    });
}


// ---------------------------------------------------------------- factors
int ex_factors_create(ex_ctx* ctx, int N, const int64_t* dims, int64_t R, ex_factors** out) {
    return guarded([&] {
        EX_CHECK(ctx && dims && out, EX_ERR_ARG, "null argument");
        EX_CHECK(N >= 2 && N <= kMaxModes, EX_ERR_UNSUPPORTED, "mode count must be in [2, 8]");
        EX_CHECK(R >= 1 && R <= kMaxRank, EX_ERR_UNSUPPORTED, "rank must be in [1, 64]");
        auto f = std::make_unique<ex_factors>();
        f->ctx = ctx;
        f->N = N;
        f->R = (int)R;
        hipStream_t st = ctx->c.stream;
        for (int k2 = 0; k2 < N; k2++) {
            EX_CHECK(dims[k2] >= 1, EX_ERR_ARG, "mode size must be positive");
            f->dims[k2] = dims[k2];
            f->U[k2].ensure((size_t)dims[k2] * R * 8);
            EX_HIP(hipMemsetAsync(f->U[k2].p, 0, (size_t)dims[k2] * R * 8, st));
            f->ptrs[k2] = f->U[k2].as<double>();
        }
        f->sigma.ensure((size_t)R * 8);
        std::vector<double> ones(R, 1.0);
        h2d(st, f->sigma.p, ones.data(), (size_t)R * 8);
        EX_HIP(hipStreamSynchronize(st));
        *out = f.release();
    });
}

int ex_factors_destroy(ex_factors* f) {
    return guarded([&] { delete f; });
}

int ex_factors_set(ex_factors* f, int mode, const double* host) {
    return guarded([&] {
        EX_CHECK(f && host, EX_ERR_ARG, "null argument");
        EX_CHECK(mode >= 0 && mode < f->N, EX_ERR_ARG, "mode out of range");
        h2d(f->ctx->c.stream, f->U[mode].p, host, (size_t)f->dims[mode] * f->R * 8);
        f->wrote64(mode);
        EX_HIP(hipStreamSynchronize(f->ctx->c.stream));
    });
}

int ex_factors_get(ex_factors* f, int mode, double* host) {
    return guarded([&] {
        EX_CHECK(f && host, EX_ERR_ARG, "null argument");
        EX_CHECK(mode >= 0 && mode < f->N, EX_ERR_ARG, "mode out of range");
        f->sync64(mode);
        d2h(f->ctx->c.stream, host, f->U[mode].p, (size_t)f->dims[mode] * f->R * 8);
    });
}

int ex_factors_get_sigma(ex_factors* f, double* sigma) {
    return guarded([&] {
        EX_CHECK(f && sigma, EX_ERR_ARG, "null argument");
        d2h(f->ctx->c.stream, sigma, f->sigma.p, (size_t)f->R * 8);
    });
}

int ex_factors_set_sigma(ex_factors* f, const double* sigma) {
    return guarded([&] {
        EX_CHECK(f && sigma, EX_ERR_ARG, "null argument");
        h2d(f->ctx->c.stream, f->sigma.p, sigma, (size_t)f->R * 8);
        EX_HIP(hipStreamSynchronize(f->ctx->c.stream));
    });
}

int ex_factors_renormalize(ex_factors* f, int mode, double* sigma) {
    return guarded([&] {
        EX_CHECK(f, EX_ERR_ARG, "null argument");
        EX_CHECK(mode >= 0 && mode < f->N, EX_ERR_ARG, "mode out of range");
        hipStream_t st = f->ctx->c.stream;
        DevBuf sg;
        sg.ensure((size_t)f->R * 8);
        f->sync64(mode);
        k::col_sumsq(st, f->U[mode].as<double>(), f->dims[mode], f->R, sg.as<double>(), f->part);
        k::scale_cols(st, f->U[mode].as<double>(), f->dims[mode], f->R, sg.as<double>());
        f->wrote64(mode);
        if (sigma) d2h(st, sigma, sg.p, (size_t)f->R * 8);
        EX_HIP(hipStreamSynchronize(st));
    });
}

int ex_factors_init_deterministic(ex_factors* f) {
    return guarded([&] {
        EX_CHECK(f, EX_ERR_ARG, "null argument");
        hipStream_t st = f->ctx->c.stream;
        DevBuf sg;
        sg.ensure((size_t)f->R * 8);
        for (int k2 = 0; k2 < f->N; k2++) {
            k::cos_init(st, f->U[k2].as<double>(), f->dims[k2], f->R);
            k::col_sumsq(st, f->U[k2].as<double>(), f->dims[k2], f->R, sg.as<double>(), f->part);
            k::scale_cols(st, f->U[k2].as<double>(), f->dims[k2], f->R, sg.as<double>());
            f->wrote64(k2);
        }
        std::vector<double> ones(f->R, 1.0);
        h2d(st, f->sigma.p, ones.data(), (size_t)f->R * 8);
        EX_HIP(hipStreamSynchronize(st));
    });
}

int ex_factors_gram(ex_factors* f, int mode, double* G) {
    return guarded([&] {
        EX_CHECK(f && G, EX_ERR_ARG, "null argument");
        EX_CHECK(mode >= 0 && mode < f->N, EX_ERR_ARG, "mode out of range");
        hipStream_t st = f->ctx->c.stream;
        f->gram.ensure((size_t)f->R * f->R * 8);
        f->sync64(mode);
        k::gram(st, f->U[mode].as<double>(), f->dims[mode], f->R, f->gram.as<double>(), f->part);
        d2h(st, G, f->gram.p, (size_t)f->R * f->R * 8);
    });
}

// ---------------------------------------------------------------- samplers
int ex_sampler_create(ex_ctx* ctx, ex_factors* f, int kind, int precision, uint64_t seed, ex_sampler** out) {
    return guarded([&] {
        EX_CHECK(ctx && f && out, EX_ERR_ARG, "null argument");
        EX_CHECK(kind == EX_SAMPLER_CP_ARLS_LEV || kind == EX_SAMPLER_STS_CP, EX_ERR_ARG, "unknown sampler kind");
        EX_CHECK(precision == EX_FP64 || precision == EX_MIXED, EX_ERR_ARG, "unknown precision");
        auto s = std::make_unique<ex_sampler>();
        s->ctx = ctx;
        s->f = f;
        s->kind = kind;
        s->precision = precision;
        s->seed = seed;
        const int N = f->N, R = f->R;
        s->sums.ensure((size_t)N * 8);
        s->part.ensure(4096);
        s->dropped.ensure(8);
        EX_HIP(hipMemsetAsync(s->dropped.p, 0, 8, ctx->c.stream));
        if (precision == EX_MIXED) f->enable32();
        if (kind == EX_SAMPLER_STS_CP) {
            const int S = tri_stride(R);
            const size_t node_bytes = precision == EX_MIXED ? (size_t)tri_stride32(R) * 4 : (size_t)S * 8;
            s->ge.init(R, 1);
            // node arrays with zeroed slack after them (the draw's unchecked
            // vector reads past the last node)
            auto with_slack = [&](DevBuf& b, size_t bytes) {
                b.ensure(bytes + kNodeSlackBytes);
                k::zero_bytes(ctx->c.stream, static_cast<char*>(b.p) + bytes, kNodeSlackBytes);
            };
            for (int k2 = 0; k2 < N; k2++) {
                s->g[k2].init(f->dims[k2], R);
                with_slack(s->nodes[k2], (size_t)s->g[k2].node_count * node_bytes);
            }
            s->grams.ensure((size_t)N * R * R * 8);
            s->M.ensure((size_t)(N + 2) * R * R * 8);
            s->lam.ensure((size_t)(N + 1) * R * 8);
            with_slack(s->eig, (size_t)N * s->ge.node_count * S * 8);
            if (precision == EX_MIXED) with_slack(s->eig32, (size_t)N * s->ge.node_count * tri_stride32(R) * 4);
            s->pinvp.ensure((size_t)tri_pad64(R) * 8);
        }
        for (int k2 = 0; k2 < N; k2++) sampler_update_impl(s.get(), k2);
        EX_HIP(hipStreamSynchronize(ctx->c.stream));
        *out = s.release();
    });
}

int ex_sampler_destroy(ex_sampler* s) {
    return guarded([&] { delete s; });
}

int ex_sampler_update(ex_sampler* s, int mode) {
    return guarded([&] {
        EX_CHECK(s, EX_ERR_ARG, "null sampler");
        EX_CHECK(mode >= 0 && mode < s->f->N, EX_ERR_ARG, "mode out of range");
        sampler_update_impl(s, mode);
        EX_HIP(hipStreamSynchronize(s->ctx->c.stream));
    });
}

int ex_sampler_draw(ex_sampler* s, int64_t J, int mode, uint64_t sweep, const double* draws_row,
                    const double* draws_eig, uint32_t* samples, double* weights) {
    return guarded([&] {
        EX_CHECK(s && samples && weights, EX_ERR_ARG, "null argument");
        EX_CHECK(mode >= 0 && mode < s->f->N, EX_ERR_ARG, "mode out of range");
        EX_CHECK(J > 0 && J < (int64_t)1 << 31, EX_ERR_ARG, "sample count J must be in [1, 2^31)");
        sampler_refresh(s);
        const int N = s->f->N;
        hipStream_t st = s->ctx->c.stream;
        const double* dr = nullptr;
        const double* de = nullptr;
        if (draws_row) {
            s->draws_row.ensure((size_t)N * J * 8);
            h2d(st, s->draws_row.p, draws_row, (size_t)N * J * 8);
            dr = s->draws_row.as<double>();
        }
        if (draws_eig) {
            s->draws_eig.ensure((size_t)N * J * 8);
            h2d(st, s->draws_eig.p, draws_eig, (size_t)N * J * 8);
            de = s->draws_eig.as<double>();
        }
        EX_CHECK(s->kind == EX_SAMPLER_CP_ARLS_LEV || ((dr == nullptr) == (de == nullptr)), EX_ERR_ARG,
                 "STS-CP needs both draw arrays or neither");
        s->samples.ensure((size_t)J * N * 4);
        s->weights.ensure((size_t)J * 8);
        sampler_draw_impl(s, J, mode, sweep, dr, de, s->samples.as<uint32_t>(), s->weights.as<double>());
        EX_HIP(hipMemcpyAsync(samples, s->samples.p, (size_t)J * N * 4, hipMemcpyDeviceToHost, st));
        d2h(st, weights, s->weights.p, (size_t)J * 8);
    });
}

int ex_sampler_dropped_draws(ex_sampler* s, int64_t* n) {
    return guarded([&] {
        EX_CHECK(s && n, EX_ERR_ARG, "null argument");
        uint64_t v = 0;
        d2h(s->ctx->c.stream, &v, s->dropped.p, 8);
        *n = (int64_t)v;
    });
}

int ex_sts_compute_m(ex_sampler* s, int mode, double* M) {
    return guarded([&] {
        EX_CHECK(s && M, EX_ERR_ARG, "null argument");
        EX_CHECK(s->kind == EX_SAMPLER_STS_CP, EX_ERR_STATE, "not an STS-CP sampler");
        EX_CHECK(mode >= 0 && mode < s->f->N, EX_ERR_ARG, "mode out of range");
        const int N = s->f->N, R = s->f->R;
        hipStream_t st = s->ctx->c.stream;
        s->m_ready = -1;
        if (s->precision == EX_MIXED) k::sts_compute_m_mixed_launch(st, N, mode, R, s->grams.as<double>(), s->M.as<double>());
        else k::sts_compute_m(st, N, mode, R, s->grams.as<double>(), s->M.as<double>(), s->lam.as<double>());
        d2h(st, M, s->M.p, (size_t)(N + 2) * R * R * 8);
    });
}

int ex_arls_scores(ex_sampler* s, int mode, double* scores) {
    return guarded([&] {
        EX_CHECK(s && scores, EX_ERR_ARG, "null argument");
        EX_CHECK(s->kind == EX_SAMPLER_CP_ARLS_LEV, EX_ERR_STATE, "not a CP-ARLS-LEV sampler");
        EX_CHECK(mode >= 0 && mode < s->f->N, EX_ERR_ARG, "mode out of range");
        sampler_refresh(s);
        d2h(s->ctx->c.stream, scores, s->scores[mode].p, (size_t)s->f->dims[mode] * 8);
    });
}

// ---------------------------------------------------------------- ALS
int ex_als_create(ex_ctx* ctx, ex_tensor* t, ex_factors* f, ex_sampler* s, int distribution, int precision,
                  ex_als** out) {
    return guarded([&] {
        EX_CHECK(ctx && t && f && s && out, EX_ERR_ARG, "null argument");
        EX_CHECK(t->N == f->N, EX_ERR_ARG, "tensor and factor mode counts differ");
        for (int k2 = 0; k2 < t->N; k2++) EX_CHECK(t->dims[k2] == f->dims[k2], EX_ERR_ARG, "mode sizes differ");
        EX_CHECK(s->f == f, EX_ERR_ARG, "sampler was built on another LowRankTensor");
        EX_CHECK(distribution == EX_DIST_ACCUMULATOR_STATIONARY || distribution == EX_DIST_TENSOR_STATIONARY,
                 EX_ERR_ARG, "unknown data distribution");
        EX_CHECK(precision == s->precision, EX_ERR_ARG, "optimizer and sampler precisions differ");
        auto a = std::make_unique<ex_als>();
        a->ctx = ctx;
        a->t = t;
        a->f = f;
        a->s = s;
        a->dist = distribution;
        a->precision = precision;
        a->cnt.ensure(32);
        EX_HIP(hipMemsetAsync(a->cnt.p, 0, 32, ctx->c.stream));
        EX_HIP(hipStreamSynchronize(ctx->c.stream));
        *out = a.release();
    });
}

int ex_exact_als_create(ex_ctx* ctx, ex_tensor* t, ex_factors* f, int precision, ex_als** out) {
    return guarded([&] {
        EX_CHECK(ctx && t && f && out, EX_ERR_ARG, "null argument");
        EX_CHECK(t->N == f->N, EX_ERR_ARG, "tensor and factor mode counts differ");
        for (int k2 = 0; k2 < t->N; k2++) EX_CHECK(t->dims[k2] == f->dims[k2], EX_ERR_ARG, "mode sizes differ");
        EX_CHECK(precision == EX_FP64 || precision == EX_MIXED, EX_ERR_ARG, "unknown precision");
        auto a = std::make_unique<ex_als>();
        a->ctx = ctx;
        a->t = t;
        a->f = f;
        a->exact = true;
        a->dist = EX_DIST_ACCUMULATOR_STATIONARY;
        a->precision = precision;
        a->cnt.ensure(32);
        EX_HIP(hipMemsetAsync(a->cnt.p, 0, 32, ctx->c.stream));
        if (precision == EX_MIXED) f->enable32();
        exact_grams(a.get());
        EX_HIP(hipStreamSynchronize(ctx->c.stream));
        *out = a.release();
    });
}

int ex_als_destroy(ex_als* a) {
    return guarded([&] {
        if (!a) return;
        (void)hipStreamSynchronize(a->ctx->c.stream);
        const uint64_t* c0 = a->cnt.as<uint64_t>();
        auto& w = a->ctx->watches;
        for (size_t i = 0; i < w.size();)
            if (w[i].dev >= c0 && w[i].dev < c0 + 4) w.erase(w.begin() + (long)i);
            else i++;
        for (hipEvent_t e : a->ev_pending) (void)hipEventDestroy(e);
        for (hipEvent_t e : a->ev_pool) (void)hipEventDestroy(e);
        als_drop_graph(a);
        delete a;
    });
}

int ex_als_init_ground_truth(ex_als* a) {
    return guarded([&] {
        EX_CHECK(a, EX_ERR_ARG, "null optimizer");
        tensor_build_lookups(a->t);
        a->initialized = true;
    });
}

int ex_als_step(ex_als* a, int mode, int64_t J, uint64_t sweep) {
    return guarded([&] {
        EX_CHECK(a, EX_ERR_ARG, "null optimizer");
        EX_CHECK(a->initialized, EX_ERR_STATE, "initialize_ground_truth_for_als() has not been called");
        EX_CHECK(mode >= 0 && mode < a->f->N, EX_ERR_ARG, "mode out of range");
        if (a->exact) {
            // the factors may have been changed through the LowRankTensor
            a->grams_valid = false;
            exact_step_impl(a, mode);
        } else {
            sampler_refresh(a->s);
            als_step_impl(a, mode, J, sweep);
        }
    });
}

int ex_als_fit(ex_als* a, double* fit) {
    return guarded([&] {
        EX_CHECK(a && fit, EX_ERR_ARG, "null argument");
        EX_CHECK(a->initialized, EX_ERR_STATE, "initialize_ground_truth_for_als() has not been called");
        *fit = als_fit_impl(a);
    });
}

// Host threads that enqueue the work of ex_als_run_batch.
constexpr int kBatchThreads = 8;

int ex_als_use_graph(ex_als* a, int on) {
    return guarded([&] {
        EX_CHECK(a, EX_ERR_ARG, "null optimizer");
        EX_CHECK(!a->exact || !on, EX_ERR_UNSUPPORTED, "graph replay is for the randomized optimizers");
        EX_HIP(hipSetDevice(a->ctx->c.device));
        EX_HIP(hipStreamSynchronize(a->ctx->c.stream));
        a->use_graph = on != 0;
        if (!a->use_graph) als_drop_graph(a);
    });
}

int ex_als_run_batch(ex_als* const* als, int n, int64_t rounds, int64_t J, double* seconds) {
    return guarded([&] {
        EX_CHECK(als && n >= 1, EX_ERR_ARG, "no optimizers");
        EX_CHECK(rounds >= 0, EX_ERR_ARG, "negative round count");
        int maxN = 0;
        for (int t = 0; t < n; t++) {
            EX_CHECK(als[t], EX_ERR_ARG, "null optimizer");
            EX_CHECK(als[t]->initialized, EX_ERR_STATE, "initialize_ground_truth_for_als() has not been called");
            for (int u = 0; u < t; u++)
                EX_CHECK(als[u]->ctx != als[t]->ctx, EX_ERR_ARG, "each optimizer needs its own context");
            maxN = std::max(maxN, als[t]->f->N);
        }
        bool graphs = true;
        for (int t = 0; t < n; t++) {
            sampler_refresh(als[t]->s);
            als_resolve_regions(als[t]);
            EX_HIP(hipStreamSynchronize(als[t]->ctx->c.stream));
            graphs = graphs && als[t]->use_graph && !als[t]->exact;
        }
        auto t0 = std::chrono::steady_clock::now();
        // with graphs, every optimizer's first sweep runs here, one optimizer
        // after the other (for one without a graph at this J: an eager sweep,
        // then the capture of one), so no capture overlaps another thread's
        // launches; the threads below only replay graphs
        int64_t first = 1;
        if (graphs && rounds >= 1) {
            for (int t = 0; t < n; t++) {
                EX_HIP(hipSetDevice(als[t]->ctx->c.device));
                als_sweep(als[t], J, 1);
            }
            first = 2;
        }
        // host threads enqueue the optimizers' mode steps (optimizer t on
        // thread t mod T), interleaved across each thread's optimizers: every
        // context's stream carries its own chain and the streams run side by
        // side, so the launch work of the host is spread over T threads
        const int T = std::min(n, kBatchThreads);
        std::vector<std::thread> pool;
        std::vector<std::exception_ptr> err((size_t)T);
        for (int th = 0; th < T; th++)
            pool.emplace_back([&, th] {
                try {
                    for (int64_t round = first; round <= rounds; round++) {
                        if (graphs) {
                            // one launch of each optimizer's captured sweep
                            for (int t = th; t < n; t += T) {
                                EX_HIP(hipSetDevice(als[t]->ctx->c.device));
                                als_sweep(als[t], J, (uint64_t)round);
                            }
                            continue;
                        }
                        for (int j = 0; j < maxN; j++)
                            for (int t = th; t < n; t += T) {
                                ex_als* a = als[t];
                                if (j >= a->f->N) continue;
                                EX_HIP(hipSetDevice(a->ctx->c.device));
                                if (a->exact) exact_step_impl(a, j);
                                else als_step_impl(a, j, J, (uint64_t)round);
                            }
                    }
                    for (int t = th; t < n; t += T) EX_HIP(hipStreamSynchronize(als[t]->ctx->c.stream));
                } catch (...) {
                    err[(size_t)th] = std::current_exception();
                }
            });
        for (auto& p : pool) p.join();
        for (auto& e : err)
            if (e) std::rethrow_exception(e);
        if (seconds) *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        for (int t = 0; t < n; t++) als_resolve_regions(als[t]);
    });
}

// ---------------------------------------------------------------- batched sweeps
// Many CP-ARLS-LEV decompositions advanced by one kernel launch per sweep
// (kernels_batch.hip): the descriptors, the per-tensor leverage state and
// scratch live in one device arena; the launches run on the batch's own
// stream, ordered after and before the optimizers' context streams.
struct ex_batch {
    int device = 0;
    hipStream_t stream = nullptr;
    std::vector<ex_als*> als;          // the caller's order
    std::vector<int> slot;             // descriptor slot of each optimizer
    std::vector<uint64_t> version;     // factor versions after the batch's last write
    std::vector<size_t> stats_off;     // arena offset of each optimizer's counters
    int precision = 0;
    int64_t J = 0;
    int grid = 0;
    DevBuf desc, ctl, arena;
    unsigned long long first_sweep = 1;  // host source of the sweep reset
    bool use_graph = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    std::vector<hipEvent_t> events;
};

namespace {

void batch_drop_graph(ex_batch* b) {
    if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->gexec = nullptr;
    b->graph = nullptr;
}

// Orders the batch stream after (to_batch) or each context stream after
// (!to_batch) the other side's queued work.
void batch_order(ex_batch* b, bool to_batch) {
    std::vector<ex_ctx*> ctxs;
    for (ex_als* a : b->als)
        if (std::find(ctxs.begin(), ctxs.end(), a->ctx) == ctxs.end()) ctxs.push_back(a->ctx);
    while (b->events.size() < ctxs.size() + 1) {
        hipEvent_t e;
        EX_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        b->events.push_back(e);
    }
    if (to_batch) {
        for (size_t i = 0; i < ctxs.size(); i++) {
            EX_HIP(hipEventRecord(b->events[i], ctxs[i]->c.stream));
            EX_HIP(hipStreamWaitEvent(b->stream, b->events[i], 0));
        }
    } else {
        hipEvent_t e = b->events[ctxs.size()];
        EX_HIP(hipEventRecord(e, b->stream));
        for (ex_ctx* c : ctxs) EX_HIP(hipStreamWaitEvent(c->c.stream, e, 0));
    }
}

}  // namespace

int ex_batch_create(ex_als* const* als, int n, int64_t J, ex_batch** out) {
    return guarded([&] {
        EX_CHECK(als && n >= 1 && out, EX_ERR_ARG, "no optimizers");
        EX_CHECK(J >= 1 && J <= kBatchMaxJ, EX_ERR_UNSUPPORTED,
                 "batched sweeps take 1 <= J <= " + std::to_string(kBatchMaxJ) + " samples");
        auto b = std::make_unique<ex_batch>();
        b->J = J;
        for (int t = 0; t < n; t++) {
            ex_als* a = als[t];
            EX_CHECK(a, EX_ERR_ARG, "null optimizer");
            EX_CHECK(a->initialized, EX_ERR_STATE, "initialize_ground_truth_for_als() has not been called");
            EX_CHECK(!a->exact && a->s && a->s->kind == EX_SAMPLER_CP_ARLS_LEV, EX_ERR_UNSUPPORTED,
                     "batched sweeps run CP-ARLS-LEV optimizers");
            EX_CHECK(a->f->R <= kBatchMaxRank, EX_ERR_UNSUPPORTED,
                     "batched sweeps take R <= " + std::to_string(kBatchMaxRank));
            if (t == 0) {
                b->precision = a->precision;
                b->device = a->ctx->c.device;
            }
            EX_CHECK(a->precision == b->precision, EX_ERR_ARG, "optimizers of one batch share a precision");
            EX_CHECK(a->ctx->c.device == b->device, EX_ERR_ARG, "optimizers of one batch share a device");
            for (int u = 0; u < t; u++)
                EX_CHECK(als[u]->f != a->f, EX_ERR_ARG, "two optimizers of one batch share a LowRankTensor");
            b->als.push_back(a);
        }
        EX_HIP(hipSetDevice(b->device));
        const bool mixed = b->precision == EX_MIXED;
        const size_t vb = mixed ? 4 : 8;
        // current factors and sampler state on every context stream
        for (ex_als* a : b->als) {
            if (mixed) a->f->sync32_all();
            else a->f->sync64_all();
            for (int k2 = 0; k2 < a->f->N; k2++) {
                ModeLookup& L = a->t->lk[k2];
                if (mixed && !L.rval32.p) {
                    L.rval32.ensure((size_t)(L.nnz > 0 ? L.nnz : 1) * 4);
                    k::to_f32(a->ctx->c.stream, L.rval.as<double>(), L.nnz, L.rval32.as<float>());
                }
            }
            a->s->refresh = false;
            for (int k2 = 0; k2 < a->f->N; k2++) sampler_update_impl(a->s, k2);
        }
        // descriptor slots: the most expensive tensors first (the queue hands
        // them out in slot order)
        std::vector<double> cost((size_t)n);
        for (int t = 0; t < n; t++) {
            const ex_als* a = b->als[(size_t)t];
            double c = (double)a->f->R * a->f->R * (double)J * a->f->N + (double)a->t->nnz;
            for (int k2 = 0; k2 < a->f->N; k2++) c += (double)a->f->dims[k2] * a->f->R;
            cost[(size_t)t] = c;
        }
        std::vector<int> order((size_t)n);
        for (int t = 0; t < n; t++) order[(size_t)t] = t;
        std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return cost[(size_t)x] > cost[(size_t)y]; });
        b->slot.assign((size_t)n, 0);
        for (int q = 0; q < n; q++) b->slot[(size_t)order[(size_t)q]] = q;
        // arena layout
        auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
        struct Plan {
            size_t samples, weights, sd, kd, wd, dg, dm, Yc, hu, hstamp, rstamp, cnt, ssum, stats;
            size_t list[kMaxModes], csc[kMaxModes], ccdf[kMaxModes], rowptr[kMaxModes];
        };
        std::vector<Plan> plan((size_t)n);
        size_t off = 0;
        auto take = [&](size_t bytes) {
            const size_t o = off;
            off += al(bytes > 0 ? bytes : 1);
            return o;
        };
        for (int t = 0; t < n; t++) {
            const ex_als* a = b->als[(size_t)t];
            const int N = a->f->N, R = a->f->R;
            int64_t Imax = 1;
            for (int k2 = 0; k2 < N; k2++) Imax = std::max(Imax, a->f->dims[k2]);
            const size_t nnz = (size_t)std::max<int64_t>(a->t->nnz, 1);
            Plan& P = plan[(size_t)t];
            P.samples = take((size_t)J * N * 4);
            P.weights = take((size_t)J * 8);
            P.sd = take((size_t)J * N * 4);
            P.kd = take((size_t)J * 8);
            P.wd = take((size_t)J * 8);
            P.dg = take((size_t)J * R * vb);
            P.dm = take((size_t)J * R * vb);
            P.Yc = take((size_t)Imax * R * vb);
            P.hu = take(nnz * 4);
            P.hstamp = take(nnz * 4);
            P.rstamp = take((size_t)Imax * 4);
            P.cnt = take(kMaxModes * 4);
            P.ssum = take(kMaxModes * 8);
            P.stats = take(24);
            for (int k2 = 0; k2 < N; k2++) {
                P.list[k2] = take((size_t)a->f->dims[k2] * 4);
                P.csc[k2] = take((size_t)a->f->dims[k2] * 8);
                P.ccdf[k2] = take((size_t)a->f->dims[k2] * 8);
                P.rowptr[k2] = take((size_t)(a->f->dims[k2] + 1) * 4);
            }
        }
        EX_HIP(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
        b->arena.ensure(off);
        b->desc.ensure(sizeof(BatchTensor) * (size_t)n);
        b->ctl.ensure(sizeof(BatchCtl));
        EX_HIP(hipMemsetAsync(b->arena.p, 0, off, b->stream));
        EX_HIP(hipMemsetAsync(b->ctl.p, 0, sizeof(BatchCtl), b->stream));
        unsigned char* base = b->arena.as<unsigned char>();
        std::vector<BatchTensor> host((size_t)n);
        b->stats_off.assign((size_t)n, 0);
        for (int t = 0; t < n; t++) {
            ex_als* a = b->als[(size_t)t];
            const Plan& P = plan[(size_t)t];
            BatchTensor& T = host[(size_t)b->slot[(size_t)t]];
            std::memset(&T, 0, sizeof(T));
            T.N = a->f->N;
            T.R = a->f->R;
            T.nnz = a->t->nnz;
            T.seed = a->s->seed;
            for (int k2 = 0; k2 < T.N; k2++) {
                BatchMode& M = T.m[k2];
                const ModeLookup& L = a->t->lk[k2];
                M.U = mixed ? a->f->U32[k2].p : a->f->U[k2].p;
                M.I = a->f->dims[k2];
                M.fkey = L.fkey.as<uint64_t>();
                M.nfib = L.nfib;
                M.fptr = L.fptr.as<uint32_t>();
                M.kpos = L.kpos.as<uint32_t>();
                M.rrow = L.rrow.as<uint32_t>();
                M.rval = mixed ? L.rval32.p : L.rval.p;
                M.rowptr = reinterpret_cast<uint32_t*>(base + P.rowptr[k2]);
                for (int q = 0; q < kMaxModes; q++) M.shift[q] = q < T.N ? L.kl.shift[q] : -1;
                M.list = reinterpret_cast<uint32_t*>(base + P.list[k2]);
                M.csc = reinterpret_cast<double*>(base + P.csc[k2]);
                M.ccdf = reinterpret_cast<double*>(base + P.ccdf[k2]);
            }
            T.sigma = a->f->sigma.as<double>();
            T.cnt = reinterpret_cast<uint32_t*>(base + P.cnt);
            T.ssum = reinterpret_cast<double*>(base + P.ssum);
            T.stats = reinterpret_cast<unsigned long long*>(base + P.stats);
            T.samples = reinterpret_cast<uint32_t*>(base + P.samples);
            T.weights = reinterpret_cast<double*>(base + P.weights);
            T.sd = reinterpret_cast<uint32_t*>(base + P.sd);
            T.kd = reinterpret_cast<uint64_t*>(base + P.kd);
            T.wd = reinterpret_cast<double*>(base + P.wd);
            T.dg = base + P.dg;
            T.dm = base + P.dm;
            T.Yc = base + P.Yc;
            T.hu = reinterpret_cast<uint32_t*>(base + P.hu);
            T.hstamp = reinterpret_cast<uint32_t*>(base + P.hstamp);
            T.rstamp = reinterpret_cast<uint32_t*>(base + P.rstamp);
            b->stats_off[(size_t)t] = P.stats;
        }
        // the sampler state (every row listed) after the context streams' work
        batch_order(b.get(), true);
        EX_HIP(hipMemcpyAsync(b->desc.p, host.data(), sizeof(BatchTensor) * (size_t)n, hipMemcpyHostToDevice,
                              b->stream));
        k::batch_init(b->stream, b->desc.as<BatchTensor>(), n);
        std::vector<uint32_t> counts((size_t)n * kMaxModes, 0);
        for (int t = 0; t < n; t++) {
            ex_als* a = b->als[(size_t)t];
            const BatchTensor& T = host[(size_t)b->slot[(size_t)t]];
            for (int k2 = 0; k2 < T.N; k2++) {
                const size_t I = (size_t)a->f->dims[k2];
                EX_HIP(hipMemcpyAsync(T.m[k2].csc, a->s->scores[k2].p, I * 8, hipMemcpyDeviceToDevice, b->stream));
                EX_HIP(hipMemcpyAsync(T.m[k2].ccdf, a->s->cdf[k2].p, I * 8, hipMemcpyDeviceToDevice, b->stream));
                counts[(size_t)t * kMaxModes + k2] = (uint32_t)I;
            }
            EX_HIP(hipMemcpyAsync(T.ssum, a->s->sums.p, (size_t)T.N * 8, hipMemcpyDeviceToDevice, b->stream));
            EX_HIP(hipMemcpyAsync(T.cnt, counts.data() + (size_t)t * kMaxModes, kMaxModes * 4, hipMemcpyHostToDevice,
                                  b->stream));
        }
        EX_HIP(hipStreamSynchronize(b->stream));
        // one workgroup per CU at most (1024 threads and about 120 KiB of LDS each)
        const int cus = std::max(1, b->als[0]->ctx->c.prop.multiProcessorCount);
        b->grid = std::min(n, cus);
        b->version.assign((size_t)n, 0);
        for (int t = 0; t < n; t++) b->version[(size_t)t] = b->als[(size_t)t]->f->version;
        *out = b.release();
    });
}

int ex_batch_use_graph(ex_batch* b, int on) {
    return guarded([&] {
        EX_CHECK(b, EX_ERR_ARG, "null batch");
        EX_HIP(hipSetDevice(b->device));
        EX_HIP(hipStreamSynchronize(b->stream));
        b->use_graph = on != 0;
        if (!b->use_graph) batch_drop_graph(b);
    });
}

int ex_batch_run(ex_batch* b, int64_t rounds, double* seconds) {
    return guarded([&] {
        EX_CHECK(b, EX_ERR_ARG, "null batch");
        EX_CHECK(rounds >= 0, EX_ERR_ARG, "negative round count");
        EX_HIP(hipSetDevice(b->device));
        const int n = (int)b->als.size();
        for (int t = 0; t < n; t++)
            EX_CHECK(b->als[(size_t)t]->f->version == b->version[(size_t)t], EX_ERR_STATE,
                     "a factor of the batch was written outside it since its last sweep: create a new batch");
        batch_order(b, true);
        // sweeps are numbered from 1 in every run, as ex_als_run numbers them
        EX_HIP(hipMemcpyAsync(b->ctl.p, &b->first_sweep, sizeof(b->first_sweep), hipMemcpyHostToDevice,
                              b->stream));
        const int vb = b->precision == EX_MIXED ? 4 : 8;
        BatchTensor* desc = b->desc.as<BatchTensor>();
        BatchCtl* ctl = b->ctl.as<BatchCtl>();
        if (b->use_graph && !b->gexec && rounds > 0) {
            EX_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
            k::batch_sweep(b->stream, desc, n, b->J, ctl, b->grid, vb);
            EX_HIP(hipStreamEndCapture(b->stream, &b->graph));
            EX_HIP(hipGraphInstantiate(&b->gexec, b->graph, nullptr, nullptr, 0));
        }
        EX_HIP(hipStreamSynchronize(b->stream));
        auto t0 = std::chrono::steady_clock::now();
        for (int64_t r = 0; r < rounds; r++) {
            if (b->gexec) EX_HIP(hipGraphLaunch(b->gexec, b->stream));
            else k::batch_sweep(b->stream, desc, n, b->J, ctl, b->grid, vb);
        }
        EX_HIP(hipStreamSynchronize(b->stream));
        if (seconds) *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if (rounds > 0)
            for (int t = 0; t < n; t++) {
                ex_als* a = b->als[(size_t)t];
                for (int k2 = 0; k2 < a->f->N; k2++) {
                    if (b->precision == EX_MIXED) a->f->wrote32(k2);
                    else a->f->wrote64(k2);
                }
                a->s->refresh = true;
                a->grams_valid = false;
                b->version[(size_t)t] = a->f->version;
            }
        batch_order(b, false);
    });
}

int ex_batch_counters(ex_batch* b, int64_t* nonzeros, int64_t* unique_samples, int64_t* rows_updated) {
    return guarded([&] {
        EX_CHECK(b, EX_ERR_ARG, "null batch");
        EX_HIP(hipSetDevice(b->device));
        const size_t n = b->als.size();
        std::vector<unsigned long long> v(3);
        for (size_t t = 0; t < n; t++) {
            EX_HIP(hipMemcpyAsync(v.data(), b->arena.as<unsigned char>() + b->stats_off[t], 24,
                                  hipMemcpyDeviceToHost, b->stream));
            EX_HIP(hipStreamSynchronize(b->stream));
            if (nonzeros) nonzeros[t] = (int64_t)v[0];
            if (unique_samples) unique_samples[t] = (int64_t)v[1];
            if (rows_updated) rows_updated[t] = (int64_t)v[2];
        }
    });
}

int ex_batch_destroy(ex_batch* b) {
    return guarded([&] {
        if (!b) return;
        (void)hipSetDevice(b->device);
        if (b->stream) (void)hipStreamSynchronize(b->stream);
        batch_drop_graph(b);
        for (hipEvent_t e : b->events) (void)hipEventDestroy(e);
        if (b->stream) (void)hipStreamDestroy(b->stream);
        delete b;
    });
}

int ex_als_run(ex_als* a, int64_t rounds, int64_t J, int64_t epoch) {
    return guarded([&] {
        EX_CHECK(a, EX_ERR_ARG, "null optimizer");
        EX_CHECK(a->initialized, EX_ERR_STATE, "initialize_ground_truth_for_als() has not been called");
        EX_CHECK(rounds >= 0, EX_ERR_ARG, "negative round count");
        EX_CHECK(epoch >= 1, EX_ERR_ARG, "epoch_interval must be positive");
        hipStream_t st = a->ctx->c.stream;
        sampler_refresh(a->s);
        als_resolve_regions(a);
        for (double& x : a->t_region) x = 0.0;
        a->grams_valid = false;
        uint64_t c0[2], z0 = 0;
        d2h(st, c0, a->cnt.p, sizeof(c0));
        if (a->s) d2h(st, &z0, a->s->dropped.p, 8);
        double als_total = 0.0, fit_total = 0.0;
        std::vector<int64_t> rl{0};
        std::vector<double> fits{0.0}, als_times{0.0}, fit_times{0.0};
        for (int64_t round = 1; round <= rounds; round++) {
            EX_HIP(hipStreamSynchronize(st));
            auto t0 = std::chrono::steady_clock::now();
            als_sweep(a, J, (uint64_t)round);
            EX_HIP(hipStreamSynchronize(st));
            als_total += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            if (round % epoch == 0) {
                auto t1 = std::chrono::steady_clock::now();
                double fit = als_fit_impl(a);
                fit_total += std::chrono::duration<double>(std::chrono::steady_clock::now() - t1).count();
                rl.push_back(round);
                fits.push_back(fit);
                als_times.push_back(als_total);
                fit_times.push_back(fit_total);
            }
        }
        als_resolve_regions(a);
        uint64_t c1[2], z1 = 0;
        d2h(st, c1, a->cnt.p, sizeof(c1));
        if (a->s) d2h(st, &z1, a->s->dropped.p, 8);
        const double nonzeros = (double)(c1[0] - c0[0]);
        std::string js = "{";
        js += "\"als_total_time\": " + json_num(als_total);
        js += ", \"als_times\": " + json_list(als_times);
        js += ", \"avg_nonzeros_iterated\": " + dstat(nonzeros);
        js += ", \"dense_reduce_time\": " + dstat(a->t_region[4]);
        js += ", \"design_matrix_prep_time\": " + dstat(a->t_region[2]);
        js += ", \"fit_computation_time\": " + json_num(fit_total);
        js += ", \"fit_computation_times\": " + json_list(fit_times);
        js += ", \"fits\": " + json_list(fits);
        js += ", \"leverage_sampling_time\": " + dstat(a->t_region[0]);
        js += ", \"num_rounds\": " + std::to_string(rounds);
        js += ", \"postprocessing_time\": " + dstat(a->t_region[5]);
        js += ", \"precision\": \"" + std::string(a->precision == EX_FP64 ? "fp64" : "mixed") + "\"";
        js += ", \"row_gather_time\": " + dstat(a->t_region[1]);
        js += ", \"rounds\": " + json_list(rl);
        js += ", \"sampler\": \"" +
              std::string(a->exact ? "exact" : a->s->kind == EX_SAMPLER_STS_CP ? "sts_cp" : "cp_arls_lev") + "\"";
        js += ", \"sampler_update_time\": " + dstat(a->t_region[6]);
        js += ", \"spmm_time\": " + dstat(a->t_region[3]);
        js += ", \"sum_nonzeros_iterated\": " + json_num(nonzeros);
        js += ", \"zero_mass_draws\": " + std::to_string(z1 - z0);
        js += ", \"device\": \"" + device_label(a->ctx) + "\"";
        js += "}";
        a->stats = js;
    });
}

int ex_als_stats_json(ex_als* a, char* buf, size_t* len) {
    return guarded([&] {
        EX_CHECK(a && len, EX_ERR_ARG, "null argument");
        size_t need = a->stats.size() + 1;
        if (buf && *len > 0) {
            size_t n = *len < need ? *len - 1 : need - 1;
            std::memcpy(buf, a->stats.data(), n);
            buf[n] = 0;
        }
        *len = need;
    });
}

// ---------------------------------------------------------------- dense hooks
int ex_tuning_count(int* n) {
    return guarded([&] {
        EX_CHECK(n, EX_ERR_ARG, "null argument");
        *n = kTuneKnobs;
    });
}

int ex_tuning_info(int knob, char* name, size_t len, int* variants, int* value) {
    return guarded([&] {
        EX_CHECK(knob >= 0 && knob < kTuneKnobs, EX_ERR_ARG, "knob index out of range");
        if (name && len) {
            std::strncpy(name, kTuneTable[knob].name, len - 1);
            name[len - 1] = 0;
        }
        if (variants) *variants = kTuneTable[knob].variants;
        if (value) *value = g_tune[knob].load();
    });
}

int ex_tuning_set(const char* name, int value) {
    return guarded([&] {
        EX_CHECK(name, EX_ERR_ARG, "null argument");
        const int q = tune_index(name);
        EX_CHECK(q >= 0, EX_ERR_ARG, std::string("unknown knob ") + name);
        EX_CHECK(value >= 0 && value < kTuneTable[q].variants, EX_ERR_ARG,
                 std::string(name) + " has no variant " + std::to_string(value));
        g_tune[q].store(value);
    });
}

int ex_tuning_load(const char* path) {
    return guarded([&] {
        EX_CHECK(path, EX_ERR_ARG, "null argument");
        load_tuning_file(path);
    });
}

int ex_gram(ex_ctx* ctx, const double* A, int64_t m, int64_t R, double* G) {
    return guarded([&] {
        EX_CHECK(ctx && G && (m == 0 || A), EX_ERR_ARG, "null argument");
        EX_CHECK(R >= 1 && R <= kMaxRank, EX_ERR_UNSUPPORTED, "rank must be in [1, 64]");
        EX_CHECK(m >= 0, EX_ERR_ARG, "negative row count");
        hipStream_t st = ctx->c.stream;
        DevBuf dA, dG, part;
        dA.ensure((size_t)(m > 0 ? m : 1) * R * 8);
        dG.ensure((size_t)R * R * 8);
        h2d(st, dA.p, A, (size_t)m * R * 8);
        k::gram(st, dA.as<double>(), m, (int)R, dG.as<double>(), part);
        d2h(st, G, dG.p, (size_t)R * R * 8);
    });
}

int ex_gram_dev_static(ex_ctx* ctx, const double* A, int64_t m, int64_t R, double* G) {
    return guarded([&] {
        EX_CHECK(ctx && G && (m == 0 || A), EX_ERR_ARG, "null argument");
        EX_CHECK(R >= 1 && R <= kMaxRank, EX_ERR_UNSUPPORTED, "rank must be in [1, 64]");
        EX_CHECK(m >= 0, EX_ERR_ARG, "negative row count");
        void* part = g_scratch_ptr[ctx->c.device];
        EX_CHECK(part != nullptr, EX_ERR_STATE, "static scratch not set up for this device");
        EX_HIP(hipSetDevice(ctx->c.device));
        k::gram_into(ctx->c.stream, A, m, (int)R, G, static_cast<double*>(part), kStaticScratchBytes);
    });
}

int ex_tns_scan(const char* path, int64_t* nnz, int* nmodes) {
    return guarded([&] {
        EX_CHECK(path && nnz && nmodes, EX_ERR_ARG, "null argument");
        int64_t n = 0;
        int N = 0;
        tns_scan(path, n, N);
        *nnz = n;
        *nmodes = N;
    });
}

int ex_tns_read(const char* path, int64_t nnz, int nmodes, uint64_t* coords, double* vals, uint64_t* min_idx,
                uint64_t* max_idx) {
    return guarded([&] {
        EX_CHECK(path && min_idx && max_idx && (nnz == 0 || (coords && vals)), EX_ERR_ARG, "null argument");
        EX_CHECK(nnz >= 0 && nmodes >= 1 && nmodes <= kMaxModes, EX_ERR_ARG, "bad nonzero or mode count");
        tns_read(path, nnz, nmodes, coords, vals, min_idx, max_idx);
    });
}

int ex_dsymm(ex_ctx* ctx, int precision, const double* G, const double* Y, int64_t m, int64_t R, double* out) {
    return guarded([&] {
        EX_CHECK(ctx && G && out && (m == 0 || Y), EX_ERR_ARG, "null argument");
        EX_CHECK(R >= 1 && R <= kMaxRank, EX_ERR_UNSUPPORTED, "rank must be in [1, 64]");
        EX_CHECK(m >= 0, EX_ERR_ARG, "negative row count");
        EX_CHECK(precision == EX_FP64 || precision == EX_MIXED, EX_ERR_ARG, "precision must be EX_FP64 or EX_MIXED");
        hipStream_t st = ctx->c.stream;
        const size_t n = (size_t)m * R, r2 = (size_t)R * R;
        DevBuf dG, dY, dO;
        if (precision == EX_FP64) {
            dG.ensure(r2 * 8);
            dY.ensure((n > 0 ? n : 1) * 8);
            dO.ensure((n > 0 ? n : 1) * 8);
            h2d(st, dG.p, G, r2 * 8);
            h2d(st, dY.p, Y, n * 8);
            k::dsymm(st, dG.as<double>(), dY.as<double>(), m, (int)R, dO.as<double>());
            d2h(st, out, dO.p, n * 8);
        } else {
            std::vector<float> g32(G, G + r2), y32(Y, Y + n), o32(n);
            dG.ensure(r2 * 4);
            dY.ensure((n > 0 ? n : 1) * 4);
            dO.ensure((n > 0 ? n : 1) * 4);
            h2d(st, dG.p, g32.data(), r2 * 4);
            h2d(st, dY.p, y32.data(), n * 4);
            k::dsymm32(st, dG.as<float>(), dY.as<float>(), m, (int)R, dO.as<float>());
            d2h(st, o32.data(), dO.p, n * 4);
            for (size_t e = 0; e < n; e++) out[e] = o32[e];
        }
    });
}

int ex_eigh(ex_ctx* ctx, const double* A, int64_t R, double* lam, double* V) {
    return guarded([&] {
        EX_CHECK(ctx && A && lam && V, EX_ERR_ARG, "null argument");
        EX_CHECK(R >= 1 && R <= kMaxRank, EX_ERR_UNSUPPORTED, "rank must be in [1, 64]");
        hipStream_t st = ctx->c.stream;
        DevBuf dA, dl, dV;
        dA.ensure((size_t)R * R * 8);
        dl.ensure((size_t)R * 8);
        dV.ensure((size_t)R * R * 8);
        h2d(st, dA.p, A, (size_t)R * R * 8);
        k::eigh(st, dA.as<double>(), (int)R, dl.as<double>(), dV.as<double>());
        EX_HIP(hipMemcpyAsync(lam, dl.p, (size_t)R * 8, hipMemcpyDeviceToHost, st));
        d2h(st, V, dV.p, (size_t)R * R * 8);
    });
}

int ex_pinv(ex_ctx* ctx, const double* M, int64_t R, double tol, double* out) {
    return guarded([&] {
        EX_CHECK(ctx && M && out, EX_ERR_ARG, "null argument");
        EX_CHECK(R >= 1 && R <= kMaxRank, EX_ERR_UNSUPPORTED, "rank must be in [1, 64]");
        hipStream_t st = ctx->c.stream;
        DevBuf dM, dO;
        dM.ensure((size_t)R * R * 8);
        dO.ensure((size_t)R * R * 8);
        h2d(st, dM.p, M, (size_t)R * R * 8);
        k::pinv(st, dM.as<double>(), (int)R, tol, dO.as<double>());
        d2h(st, out, dO.p, (size_t)R * R * 8);
    });
}

int ex_spd_pinv(ex_ctx* ctx, const double* M, int64_t R, double tol_rel, double* out) {
    return guarded([&] {
        EX_CHECK(ctx && M && out, EX_ERR_ARG, "null argument");
        EX_CHECK(R >= 1 && R <= kMaxRank, EX_ERR_UNSUPPORTED, "rank must be in [1, 64]");
        hipStream_t st = ctx->c.stream;
        DevBuf dM, dO, flag;
        dM.ensure((size_t)R * R * 8);
        dO.ensure((size_t)R * R * 8);
        flag.ensure(8);
        h2d(st, dM.p, M, (size_t)R * R * 8);
        k::spd_pinv(st, dM.as<double>(), (int)R, tol_rel, dO.as<double>(), flag.as<int>());
        d2h(st, out, dO.p, (size_t)R * R * 8);
    });
}

int ex_dedup(ex_ctx* ctx, const uint32_t* samples, const double* weights, int64_t J, int N, int mode,
             const int64_t* dims, uint32_t* samples_dedup, double* weights_dedup, int64_t* Ju) {
    return guarded([&] {
        EX_CHECK(ctx && dims && Ju && samples_dedup && weights_dedup, EX_ERR_ARG, "null argument");
        EX_CHECK(N >= 2 && N <= kMaxModes && mode >= 0 && mode < N, EX_ERR_ARG, "bad mode arguments");
        EX_CHECK(J >= 0 && J < (int64_t)1 << 31 && (J == 0 || (samples && weights)), EX_ERR_ARG, "bad sample arrays");
        KeyLayout kl = make_key_layout(N, mode, dims);
        EX_CHECK(kl.bits <= 64, EX_ERR_UNSUPPORTED, "tuple key exceeds 64 bits");
        for (int64_t i = 0; i < J; i++)
            for (int k2 = 0; k2 < N; k2++)
                EX_CHECK(k2 == mode || samples[i * N + k2] < (uint64_t)dims[k2], EX_ERR_ARG, "sample index out of range");
        hipStream_t st = ctx->c.stream;
        DevBuf ds, dw, sd, wd, kd;
        DevBuf tmp[4];
        size_t JJ = (size_t)(J > 0 ? J : 1);
        ds.ensure(JJ * N * 4);
        dw.ensure(JJ * 8);
        sd.ensure(JJ * N * 4);
        wd.ensure(JJ * 8);
        kd.ensure(JJ * 8);
        h2d(st, ds.p, samples, (size_t)J * N * 4);
        h2d(st, dw.p, weights, (size_t)J * 8);
        DevBuf table, work;
        const int64_t nu = kl.bits <= 63
                               ? k::dedup_hash(st, ds.as<uint32_t>(), dw.as<double>(), J, N, kl, sd.as<uint32_t>(),
                                               wd.as<double>(), kd.as<uint64_t>(), table, work)
                               : k::dedup(st, ds.as<uint32_t>(), dw.as<double>(), J, N, kl, sd.as<uint32_t>(),
                                          wd.as<double>(), kd.as<uint64_t>(), tmp);
        EX_HIP(hipMemcpyAsync(samples_dedup, sd.p, (size_t)nu * N * 4, hipMemcpyDeviceToHost, st));
        d2h(st, weights_dedup, wd.p, (size_t)nu * 8);
        *Ju = nu;
    });
}

}  // extern "C"
