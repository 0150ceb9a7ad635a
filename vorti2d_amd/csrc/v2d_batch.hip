// v2d_batch.hip -- lockstep Newton steps of many solver contexts of uneven
// mesh sizes with shared launches.
//
// A batch holds up to kBatchMax contexts of one precision (the caller owns
// them).  One Newton step of the active cases runs
//   * one assembly launch, one residual-norm launch pair and one right-hand-side
//     scaling launch for all of them (k_assemble_bt, k_multidot_bt + k_reduce_bt,
//     k_scale_copy_bt), on the batch stream;
//   * the factorisations and preconditioner applies per case, on the cases' own
//     streams, so that they overlap (each case's block-LU sweeps are chains of
//     small level launches);
//   * mixed precision: the GMRES iterations in lockstep -- per iteration one
//     SpMV launch, two fused multi-dot / reduce / multi-axpy rounds (classical
//     Gram-Schmidt with re-orthogonalisation), one norm round and one
//     normalisation launch for all cases still iterating, and one host
//     synchronisation, after which every case's Givens step and convergence
//     decision are made by the same host code as a single solve (GmresHost);
//   * one Newton-update launch and one update-norm launch pair.
// Every batched launch runs, on each case's blocks, the device code of that
// case's single-context launch (see BMap), so a case's iterates equal those of
// v2d_ctx_newton_step on its context bit for bit.
//
// Block sizes of the pointwise batched launches (assembly, SpMV, multi-axpy,
// scaling, update) do not change any result; v2d_batch_autotune times the
// candidates 128 / 256 / 512 / 1024 on the batch's own cases and keeps the
// fastest per launch kind, and remembers the choice per size class (the
// batch's total node count rounded up to a power of two).  The reductions keep
// 256 threads and each case's own block count.
#include <hip/hip_runtime.h>

#include <cmath>
#include <condition_variable>
#include <cstring>
#include <functional>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/vorti2d.h"
#include "v2d_internal.h"
#include "v2d_ctx.h"

namespace v2d {
namespace {

enum { kAsm = 0, kSpmvK = 1, kVecK = 2, kKinds = 3 };
constexpr int kOut = 130;   // per case: h of pass 1 (64), of pass 2 (64), |w|^2, spare
constexpr int kTables = 12; // BVec tables of one lockstep iteration

std::mutex g_tune_mu;
std::map<int64_t, std::vector<int>> g_tuned;   // size class -> block size per launch kind

int64_t size_class(int64_t nodes)
{
    int64_t c = 1;
    while (c < nodes) c <<= 1;
    return c;
}

// Host threads that enqueue per-case work (factorisations, preconditioner
// applies) concurrently, each on its case's own stream: a factorisation is
// thousands of small launches, whose enqueueing from one thread would
// serialise the cases.  run() returns when every task has returned (their GPU
// work is then queued, not finished); the first failure's message is kept.
class Pool {
  public:
    explicit Pool(int n)
    {
        for (int i = 0; i < n; ++i) th_.emplace_back([this] { loop(); });
    }
    ~Pool()
    {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
        }
        cv_.notify_all();
        for (auto &t : th_) t.join();
    }
    int run(std::vector<std::function<int()>> &tasks)
    {
        if (tasks.size() <= 1 || th_.empty()) {
            for (auto &f : tasks)
                if (f()) return -1;
            return 0;
        }
        {
            std::lock_guard<std::mutex> lk(mu_);
            for (auto &f : tasks) q_.push_back(&f);
            pending_ = (int)tasks.size();
            rc_ = 0;
            err_.clear();
        }
        cv_.notify_all();
        std::unique_lock<std::mutex> lk(mu_);
        done_.wait(lk, [this] { return pending_ == 0; });
        if (rc_) set_error(err_);
        return rc_;
    }

  private:
    void loop()
    {
        for (;;) {
            std::function<int()> *f = nullptr;
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [this] { return stop_ || !q_.empty(); });
                if (stop_ && q_.empty()) return;
                f = q_.back();
                q_.pop_back();
            }
            const int rc = (*f)();
            std::string msg;
            if (rc) msg = v2d_last_error();
            std::lock_guard<std::mutex> lk(mu_);
            if (rc && !rc_) {
                rc_ = -1;
                err_ = msg;
            }
            if (--pending_ == 0) done_.notify_all();
        }
    }
    std::vector<std::thread> th_;
    std::mutex mu_;
    std::condition_variable cv_, done_;
    std::vector<std::function<int()> *> q_;
    int pending_ = 0, rc_ = 0;
    std::string err_;
    bool stop_ = false;
};

struct Batch {
    std::unique_ptr<Pool> pool;
    std::vector<Ctx *> cs;
    int prec = V2D_FP64;
    hipStream_t s = nullptr;
    hipEvent_t eb = nullptr;
    std::vector<hipEvent_t> ec;
    int bs[kKinds] = {256, 256, 256};
    // device tables and buffers
    BAsm *d_asm = nullptr;
    BVec *d_vec = nullptr;       // kTables x kBatchMax
    double *d_out = nullptr;     // kBatchMax x kOut
    double *d_scal = nullptr;    // kBatchMax x 2 (norm scale, hv)
    // pinned host staging
    BAsm *h_asm = nullptr;
    BVec *h_vec = nullptr;
    double *h_out = nullptr;
    double *h_scal = nullptr;
    ~Batch()
    {
        pool.reset();
        if (s) (void)hipStreamSynchronize(s);
        for (void *p : {(void *)d_asm, (void *)d_vec, (void *)d_out, (void *)d_scal})
            if (p) (void)hipFree(p);
        for (void *p : {(void *)h_asm, (void *)h_vec, (void *)h_out, (void *)h_scal})
            if (p) (void)hipHostFree(p);
        for (auto e : ec) if (e) (void)hipEventDestroy(e);
        if (eb) (void)hipEventDestroy(eb);
        if (s) (void)hipStreamDestroy(s);
    }
};

// Block map of one launch: case q of `idx` gets ceil(len_q / threads) blocks.
BMap make_map(const std::vector<int64_t> &len, int threads, unsigned *grid)
{
    BMap m{};
    m.nc = (int)len.size();
    m.off[0] = 0;
    for (int q = 0; q < m.nc; ++q) m.off[q + 1] = m.off[q] + (int)grid_for(len[q], threads);
    *grid = (unsigned)m.off[m.nc];
    return m;
}

int nb_of(int64_t n) { return (int)std::min<unsigned>(1024, grid_for(n)); }

// The batch stream waits for the listed cases' queued work, or they wait for
// the batch stream's.
int join_cases(Batch &B, const std::vector<int> &idx)
{
    for (int q : idx) {
        V2D_HIP(hipEventRecord(B.ec[q], B.cs[q]->s));
        V2D_HIP(hipStreamWaitEvent(B.s, B.ec[q], 0));
    }
    return 0;
}
int fork_cases(Batch &B, const std::vector<int> &idx)
{
    V2D_HIP(hipEventRecord(B.eb, B.s));
    for (int q : idx) V2D_HIP(hipStreamWaitEvent(B.cs[q]->s, B.eb, 0));
    return 0;
}

// ||v_q||_2^2 of 2N-vectors for every listed case, one multi-dot + one reduce
// launch (the blocks of each case's dnorm2), results in d_out[q * kOut].
int batched_norms(Batch &B, const std::vector<int> &idx, double *const *vecs, BVec *hv, BVec *dv)
{
    std::vector<int64_t> len;
    for (size_t t = 0; t < idx.size(); ++t) {
        Ctx *c = B.cs[idx[t]];
        BVec q{};
        q.n = 2 * c->N;
        q.ld = q.n;
        q.m = 1;
        q.nb = nb_of(q.n);
        q.Vb = vecs[t];
        q.w = vecs[t];
        q.partial = c->partial;
        q.out = B.d_out + (int64_t)idx[t] * kOut;
        hv[t] = q;
        len.push_back((int64_t)q.nb * 256);
    }
    V2D_HIP(hipMemcpyAsync(dv, hv, sizeof(BVec) * idx.size(), hipMemcpyHostToDevice, B.s));
    unsigned g = 0;
    const BMap m = make_map(len, 256, &g);
    hipLaunchKernelGGL(k_multidot_bt<double>, dim3(g), dim3(256), 0, B.s, m, dv);
    hipLaunchKernelGGL(k_reduce_bt, dim3((unsigned)idx.size()), dim3(256), 0, B.s, dv);
    V2D_LAUNCH();
    return 0;
}

int fetch_out(Batch &B)
{
    V2D_HIP(hipMemcpyAsync(B.h_out, B.d_out, sizeof(double) * kBatchMax * kOut, hipMemcpyDeviceToHost, B.s));
    V2D_HIP(hipStreamSynchronize(B.s));
    return 0;
}

int batched_assembly(Batch &B, const std::vector<int> &idx, const double *re, const double *invdtau,
                     const double *inv2dt, const double *urot)
{
    std::vector<int64_t> len;
    for (size_t t = 0; t < idx.size(); ++t) {
        const int q = idx[t];
        Ctx *c = B.cs[q];
        const double r = re[q], id = invdtau ? invdtau[q] : 0.0, i2 = inv2dt ? inv2dt[q] : 0.0,
                     ur = urot ? urot[q] : 0.0;
        c->last_re = r; c->last_invdtau = id; c->last_inv2dt = i2; c->last_urot = ur;
        BAsm a{};
        a.a = c->args(r, id, i2, ur);
        a.k = asm_consts(r, c->dksi, c->deta);
        a.flags = c->flags;
        a.val = c->val;
        a.rhs = c->rhs;
        B.h_asm[t] = a;
        len.push_back(c->N);
    }
    V2D_HIP(hipMemcpyAsync(B.d_asm, B.h_asm, sizeof(BAsm) * idx.size(), hipMemcpyHostToDevice, B.s));
    unsigned g = 0;
    const BMap m = make_map(len, B.bs[kAsm], &g);
    if (B.prec == V2D_FP64)
        hipLaunchKernelGGL(k_assemble_bt<double>, dim3(g), dim3(B.bs[kAsm]), 0, B.s, m, B.d_asm);
    else
        hipLaunchKernelGGL(k_assemble_bt<float>, dim3(g), dim3(B.bs[kAsm]), 0, B.s, m, B.d_asm);
    V2D_LAUNCH();
    return 0;
}

// One lockstep GMRES solve of every listed case (mixed precision), with the
// refactor-and-retry rule of v2d_ctx_newton_step.  its[t]: iterations of case
// idx[t] (both attempts).
struct Slot {
    int q = 0;
    bool need = false;
    int phase = 0;        // 0: first attempt, 1: after a refactorisation
    int its = 0;          // iterations so far (earlier attempts)
    bool active = true;
    GmresHost G;
};

int batched_gmres(Batch &B, const std::vector<int> &idx, const std::vector<char> &need, std::vector<int> &its)
{
    std::vector<Slot> sl(idx.size());
    for (size_t t = 0; t < idx.size(); ++t) {
        Ctx *c = B.cs[idx[t]];
        sl[t].q = idx[t];
        sl[t].need = need[t] != 0;
        sl[t].G.begin(sl[t].need ? c->gm_cap : c->gm_max, sl[t].need ? 0 : c->gm_max);
        V2D_HIP(hipMemcpyAsync(c->gV, c->gb, sizeof(float) * 2 * c->N, hipMemcpyDeviceToDevice, B.s));
    }
    for (;;) {
        std::vector<int> act;   // slot numbers still iterating
        for (size_t t = 0; t < sl.size(); ++t)
            if (sl[t].active) act.push_back((int)t);
        if (act.empty()) break;
        std::vector<int> aq;
        for (int t : act) aq.push_back(sl[t].q);
        // preconditioner applies on the cases' streams
        if (fork_cases(B, aq)) return -1;
        {
            std::vector<std::function<int()>> tasks;
            for (int t : act) {
                Ctx *c = B.cs[sl[t].q];
                const int64_t n = 2 * c->N;
                const int j = sl[t].G.j;
                tasks.emplace_back([c, n, j] { return apply_pc(c, c->gV + j * n, c->gZ + j * n); });
            }
            if (B.pool->run(tasks)) return -1;
        }
        if (join_cases(B, aq)) return -1;
        // the operand tables of this iteration: SpMV, 2 x (dot, reduce,
        // axpy), norm dot + reduce, normalisation
        const int na = (int)act.size();
        BVec *H = B.h_vec;
        std::vector<int64_t> len_n, len_2n, len_nb;
        for (int a = 0; a < na; ++a) {
            const Slot &S = sl[act[a]];
            Ctx *c = B.cs[S.q];
            const int64_t n = 2 * c->N;
            const int j = S.G.j;
            float *w = c->gV + (j + 1) * n;
            double *o = B.d_out + (int64_t)S.q * kOut;
            BVec sp{};
            sp.imax = c->imax; sp.jmax = c->jmax; sp.flags = c->flags; sp.val = c->val;
            sp.x = c->gZ + j * n; sp.y = w;
            H[0 * kBatchMax + a] = sp;
            for (int pass = 0; pass < 2; ++pass) {
                BVec d{};
                d.n = n; d.ld = n; d.m = j + 1; d.nb = nb_of(n); d.Vb = c->gV; d.w = w;
                d.partial = c->partial; d.out = o + 64 * pass;
                d.h = d.out; d.y = w;
                H[(1 + pass) * kBatchMax + a] = d;   // dot / reduce / axpy share the entry
            }
            BVec nd{};
            nd.n = n; nd.ld = n; nd.m = 1; nd.nb = nb_of(n); nd.Vb = w; nd.w = w;
            nd.partial = c->partial; nd.out = o + 128;
            H[3 * kBatchMax + a] = nd;
            BVec sc{};
            sc.n = n; sc.x = w; sc.y = w; sc.s = B.d_scal + 2 * S.q + 1; sc.inv = 1;
            H[4 * kBatchMax + a] = sc;
            len_n.push_back(c->N);
            len_2n.push_back(n);
            len_nb.push_back((int64_t)nb_of(n) * 256);
        }
        V2D_HIP(hipMemcpyAsync(B.d_vec, B.h_vec, sizeof(BVec) * 5 * kBatchMax, hipMemcpyHostToDevice, B.s));
        unsigned g_sp = 0, g_ax = 0, g_nb = 0;
        const BMap m_sp = make_map(len_n, B.bs[kSpmvK], &g_sp);
        const BMap m_ax = make_map(len_2n, B.bs[kVecK], &g_ax);
        const BMap m_nb = make_map(len_nb, 256, &g_nb);
        hipLaunchKernelGGL((k_spmv_bt<float, float>), dim3(g_sp), dim3(B.bs[kSpmvK]), 0, B.s, m_sp, B.d_vec);
        for (int pass = 0; pass < 2; ++pass) {
            const BVec *tb = B.d_vec + (1 + pass) * kBatchMax;
            hipLaunchKernelGGL(k_multidot_bt<float>, dim3(g_nb), dim3(256), 0, B.s, m_nb, tb);
            hipLaunchKernelGGL(k_reduce_bt, dim3((unsigned)na), dim3(256), 0, B.s, tb);
            hipLaunchKernelGGL(k_multiaxpy_bt<float>, dim3(g_ax), dim3(B.bs[kVecK]), 0, B.s, m_ax, tb);
        }
        hipLaunchKernelGGL(k_multidot_bt<float>, dim3(g_nb), dim3(256), 0, B.s, m_nb, B.d_vec + 3 * kBatchMax);
        hipLaunchKernelGGL(k_reduce_bt, dim3((unsigned)na), dim3(256), 0, B.s, B.d_vec + 3 * kBatchMax);
        V2D_LAUNCH();
        if (fetch_out(B)) return -1;
        // host: each case's Arnoldi step and decision
        std::vector<int> scale_a;      // table entries to normalise
        std::vector<int> finish;       // slots whose solve ended with a solution
        std::vector<std::function<int()>> refac;   // stale factorisations to redo
        for (int a = 0; a < na; ++a) {
            Slot &S = sl[act[a]];
            Ctx *c = B.cs[S.q];
            if (sweep_failed(c->wf.err_host)) return -1;
            const double *o = B.h_out + (int64_t)S.q * kOut;
            const int j = S.G.j;
            for (int i = 0; i <= j; ++i) {
                S.G.h[i] = o[i];
                S.G.h2[i] = o[64 + i];
            }
            int it = 0;
            double rr = 1.0;
            const int st = S.G.advance(o[128], c->rtol_step, &it, &rr);
            if (st == 3 || st == 0) {
                B.h_scal[2 * S.q + 1] = S.G.hv;
                scale_a.push_back(a);
            }
            if (st == 3) continue;
            // the attempt ended (the rules of v2d_ctx_newton_step)
            if (S.need && S.phase == 0) c->fresh_its = it;
            if (st != 0 || rr > c->rtol_step) {
                if (!S.need && S.phase == 0) {   // stale factorisation: refactor and solve again
                    S.its += it;
                    S.phase = 1;
                    refac.emplace_back([c] { return ctx_factor(c); });   // on the case's stream
                    S.G.begin(c->gm_cap, 0);
                    V2D_HIP(hipMemcpyAsync(c->gV, c->gb, sizeof(float) * 2 * c->N, hipMemcpyDeviceToDevice,
                                           B.s));
                    continue;
                }
                if (st != 0) {
                    set_error("mixed-precision GMRES failed on a fresh FP32 factorisation "
                              "(non-finite Krylov vectors)");
                    return -1;
                }
            }
            if (S.phase == 1) c->fresh_its = it;
            S.its += it;
            S.active = false;
            finish.push_back(act[a]);
        }
        if (B.pool->run(refac)) return -1;
        if (!scale_a.empty()) {
            // normalisations v_{j+1} = w / hv, as in the single solve (also for
            // the iteration that converged)
            BVec *T = B.h_vec + 5 * kBatchMax;
            std::vector<int64_t> len;
            for (size_t u = 0; u < scale_a.size(); ++u) {
                T[u] = B.h_vec[4 * kBatchMax + scale_a[u]];
                len.push_back(T[u].n);
            }
            V2D_HIP(hipMemcpyAsync(B.d_scal, B.h_scal, sizeof(double) * 2 * kBatchMax, hipMemcpyHostToDevice,
                                   B.s));
            V2D_HIP(hipMemcpyAsync(B.d_vec + 5 * kBatchMax, T, sizeof(BVec) * scale_a.size(),
                                   hipMemcpyHostToDevice, B.s));
            unsigned g = 0;
            const BMap m = make_map(len, B.bs[kVecK], &g);
            hipLaunchKernelGGL((k_scale_copy_bt<float, float>), dim3(g), dim3(B.bs[kVecK]), 0, B.s, m,
                               B.d_vec + 5 * kBatchMax);
            V2D_LAUNCH();
        }
        for (int t : finish) {   // x = Z y of the finished solves
            Ctx *c = B.cs[sl[t].q];
            const int64_t n = 2 * c->N;
            const std::vector<double> yv = sl[t].G.solution();
            const int k = (int)yv.size();
            V2D_HIP(hipMemcpyAsync(c->hbuf, yv.data(), sizeof(double) * k, hipMemcpyHostToDevice, B.s));
            hipLaunchKernelGGL((k_fill_zero_d), dim3(grid_for(n)), dim3(256), 0, B.s, n, c->gxd);
            hipLaunchKernelGGL((k_update_sol<float, double>), dim3(grid_for(n)), dim3(256), 0, B.s, n, k,
                               c->gZ, n, c->hbuf, 1.0, c->gxd);
            V2D_LAUNCH();
            V2D_HIP(hipStreamSynchronize(B.s));   // yv is a local
        }
    }
    its.assign(idx.size(), 0);
    for (size_t t = 0; t < sl.size(); ++t) its[t] = sl[t].its;
    return 0;
}

}  // namespace
}  // namespace v2d

using namespace v2d;

struct v2d_batch {
    Batch b;
};

extern "C" {

int v2d_batch_create(v2d_batch **out, int precision, int slots)
{
    if (!out) {
        set_error("batch: null output pointer");
        return -1;
    }
    *out = nullptr;
    if (slots < 1 || slots > kBatchMax) {
        set_error("batch: between 1 and " + std::to_string(kBatchMax) + " slots");
        return -1;
    }
    if (precision != V2D_FP64 && precision != V2D_MIXED) {
        set_error("batch: precision must be V2D_FP64 or V2D_MIXED");
        return -1;
    }
    auto h = std::make_unique<v2d_batch>();
    Batch &B = h->b;
    B.prec = precision;
    B.cs.assign(slots, nullptr);
    B.pool = std::make_unique<Pool>(slots);
    V2D_HIP(hipStreamCreateWithFlags(&B.s, hipStreamNonBlocking));
    V2D_HIP(hipEventCreateWithFlags(&B.eb, hipEventDisableTiming));
    B.ec.assign(slots, nullptr);
    for (auto &e : B.ec) V2D_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    V2D_HIP(hipMalloc((void **)&B.d_asm, sizeof(BAsm) * kBatchMax));
    V2D_HIP(hipMalloc((void **)&B.d_vec, sizeof(BVec) * kTables * kBatchMax));
    V2D_HIP(hipMalloc((void **)&B.d_out, sizeof(double) * kBatchMax * kOut));
    V2D_HIP(hipMalloc((void **)&B.d_scal, sizeof(double) * 2 * kBatchMax));
    V2D_HIP(hipHostMalloc((void **)&B.h_asm, sizeof(BAsm) * kBatchMax, hipHostMallocDefault));
    V2D_HIP(hipHostMalloc((void **)&B.h_vec, sizeof(BVec) * kTables * kBatchMax, hipHostMallocDefault));
    V2D_HIP(hipHostMalloc((void **)&B.h_out, sizeof(double) * kBatchMax * kOut, hipHostMallocDefault));
    V2D_HIP(hipHostMalloc((void **)&B.h_scal, sizeof(double) * 2 * kBatchMax, hipHostMallocDefault));
    std::memset(B.h_scal, 0, sizeof(double) * 2 * kBatchMax);
    *out = h.release();
    return 0;
}

int v2d_batch_set_case(v2d_batch *bh, int slot, v2d_ctx *ctx)
{
    Batch &B = bh->b;
    if (slot < 0 || slot >= (int)B.cs.size()) {
        set_error("batch: slot " + std::to_string(slot) + " out of range");
        return -1;
    }
    Ctx *c = ctx ? ctx->c.get() : nullptr;
    if (c && c->prec != B.prec) {
        set_error("batch: the context's precision differs from the batch's");
        return -1;
    }
    for (int q = 0; q < (int)B.cs.size(); ++q)
        if (c && q != slot && B.cs[q] == c) {
            set_error("batch: the context already sits in slot " + std::to_string(q));
            return -1;
        }
    if (B.cs[slot]) {   // the leaving case's stream sees the batch's work on it
        V2D_HIP(hipEventRecord(B.eb, B.s));
        V2D_HIP(hipStreamWaitEvent(B.cs[slot]->s, B.eb, 0));
    }
    B.cs[slot] = c;
    if (c) {
        // the shared launches read the case through fixed pointers and ordered
        // on the batch stream; block-LU applies on the case's stream must not
        // synchronise the host inside a lockstep iteration
        if (c->wf.sweep_kind == V2D_SWEEP_PERSISTENT) {
            c->wf.sweep_kind = V2D_SWEEP_LEVELS;
            c->drop_graph();
        }
        int64_t nodes = 0;
        for (Ctx *o : B.cs) if (o) nodes += o->N;
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto it = g_tuned.find(size_class(nodes));
        if (it != g_tuned.end())
            for (int k = 0; k < kKinds; ++k) B.bs[k] = it->second[k];
    }
    return 0;
}

int v2d_batch_destroy(v2d_batch *b)
{
    delete b;
    return 0;
}

int v2d_batch_get_blocks(v2d_batch *b, int *threads3)
{
    for (int k = 0; k < kKinds; ++k) threads3[k] = b->b.bs[k];
    return 0;
}

int v2d_batch_set_blocks(v2d_batch *b, const int *threads3)
{
    for (int k = 0; k < kKinds; ++k)
        if (threads3[k] != 128 && threads3[k] != 256 && threads3[k] != 512 && threads3[k] != 1024) {
            set_error("batch: block sizes must be 128, 256, 512 or 1024 threads");
            return -1;
        }
    for (int k = 0; k < kKinds; ++k) b->b.bs[k] = threads3[k];
    return 0;
}

int v2d_batch_newton_step(v2d_batch *bh, const int *active, const double *re, const double *invdtau,
                          const double *inv2dt, const double *urot, double *res, double *rnorm, int *lin_iters)
{
    Batch &B = bh->b;
    std::vector<int> idx;
    for (int q = 0; q < (int)B.cs.size(); ++q)
        if (B.cs[q] && (!active || active[q])) idx.push_back(q);
    if (idx.empty()) return 0;
    for (int q : idx) {
        Ctx *c = B.cs[q];
        if (c->prec == V2D_MIXED && c->refactor_policy != V2D_REFACTOR_FIXED) {
            set_error("batch: mixed-precision cases need the fixed refactor policy");
            return -1;
        }
        c->fac_in_step = false;
    }
    // the cases' queued work (state uploads, earlier solves) before the shared launches
    if (join_cases(B, idx)) return -1;
    if (batched_assembly(B, idx, re, invdtau, inv2dt, urot)) return -1;
    std::vector<double *> vecs;
    for (int q : idx) vecs.push_back(B.cs[q]->rhs);
    if (batched_norms(B, idx, vecs.data(), B.h_vec, B.d_vec)) return -1;
    if (fetch_out(B)) return -1;
    std::vector<double> bn(B.cs.size(), 0.0);
    for (int q : idx) {
        bn[q] = std::sqrt(B.h_out[(int64_t)q * kOut]);
        if (!std::isfinite(bn[q])) {
            set_error("Newton step: non-finite residual norm (degenerate mesh or diverged state)");
            return -1;
        }
        rnorm[q] = bn[q];
        lin_iters[q] = 0;
    }
    if (fork_cases(B, idx)) return -1;
    std::vector<int> gq;      // cases solved by GMRES
    std::vector<char> gneed;
    std::vector<int> zero_q;  // mixed cases with a zero residual: zero update
    std::vector<std::function<int()>> tasks;
    for (int q : idx) {
        Ctx *c = B.cs[q];
        if (c->prec == V2D_FP64) {
            tasks.emplace_back([c] {
                if (ctx_factor(c)) return -1;
                return lin_solve<double>(c, c->rhs, c->delta, false);
            });
            continue;
        }
        const bool need = !c->factored || (c->refactor_every > 0 && c->since_factor >= c->refactor_every);
        if (need) tasks.emplace_back([c] { return ctx_factor(c); });
        if (bn[q] > 0.0) {
            c->rtol_step = forcing_rtol(c, bn[q]);
            gq.push_back(q);
            gneed.push_back(need ? 1 : 0);
        } else {
            zero_q.push_back(q);
        }
    }
    if (B.pool->run(tasks)) return -1;
    if (!gq.empty()) {
        // unit-norm FP32 right-hand sides
        std::vector<int64_t> len;
        BVec *T = B.h_vec + 6 * kBatchMax;
        for (size_t t = 0; t < gq.size(); ++t) {
            Ctx *c = B.cs[gq[t]];
            B.h_scal[2 * gq[t]] = bn[gq[t]];
            BVec sc{};
            sc.n = 2 * c->N; sc.x = c->rhs; sc.y = c->gb; sc.s = B.d_scal + 2 * gq[t]; sc.inv = 1;
            T[t] = sc;
            len.push_back(sc.n);
        }
        V2D_HIP(hipMemcpyAsync(B.d_scal, B.h_scal, sizeof(double) * 2 * kBatchMax, hipMemcpyHostToDevice, B.s));
        V2D_HIP(hipMemcpyAsync(B.d_vec + 6 * kBatchMax, T, sizeof(BVec) * gq.size(), hipMemcpyHostToDevice, B.s));
        unsigned g = 0;
        const BMap m = make_map(len, B.bs[kVecK], &g);
        hipLaunchKernelGGL((k_scale_copy_bt<double, float>), dim3(g), dim3(B.bs[kVecK]), 0, B.s, m,
                           B.d_vec + 6 * kBatchMax);
        V2D_LAUNCH();
        std::vector<int> its;
        if (batched_gmres(B, gq, gneed, its)) return -1;
        // delta = bn * x
        len.clear();
        BVec *U = B.h_vec + 7 * kBatchMax;
        for (size_t t = 0; t < gq.size(); ++t) {
            Ctx *c = B.cs[gq[t]];
            lin_iters[gq[t]] = its[t];
            BVec sc{};
            sc.n = 2 * c->N; sc.x = c->gxd; sc.y = c->delta; sc.s = B.d_scal + 2 * gq[t]; sc.inv = 0;
            U[t] = sc;
            len.push_back(sc.n);
        }
        V2D_HIP(hipMemcpyAsync(B.d_scal, B.h_scal, sizeof(double) * 2 * kBatchMax, hipMemcpyHostToDevice, B.s));
        V2D_HIP(hipMemcpyAsync(B.d_vec + 7 * kBatchMax, U, sizeof(BVec) * gq.size(), hipMemcpyHostToDevice, B.s));
        const BMap m2 = make_map(len, B.bs[kVecK], &g);
        hipLaunchKernelGGL((k_scale_copy_bt<double, double>), dim3(g), dim3(B.bs[kVecK]), 0, B.s, m2,
                           B.d_vec + 7 * kBatchMax);
        V2D_LAUNCH();
    }
    for (int q : zero_q) {
        Ctx *c = B.cs[q];
        hipLaunchKernelGGL(k_fill_zero_d, dim3(grid_for(2 * c->N)), dim3(256), 0, B.s, 2 * c->N, c->delta);
        V2D_LAUNCH();
    }
    for (int q : idx)
        if (B.cs[q]->prec == V2D_MIXED) ++B.cs[q]->since_factor;
    // the FP64 cases' direct solves (and any factorisation) before the updates
    if (join_cases(B, idx)) return -1;
    {
        std::vector<int64_t> len;
        BVec *T = B.h_vec + 8 * kBatchMax;
        for (size_t t = 0; t < idx.size(); ++t) {
            Ctx *c = B.cs[idx[t]];
            BVec u{};
            u.n = c->N; u.x = c->delta; u.psi = c->psi; u.ome = c->ome;
            T[t] = u;
            len.push_back(c->N);
        }
        V2D_HIP(hipMemcpyAsync(B.d_vec + 8 * kBatchMax, T, sizeof(BVec) * idx.size(), hipMemcpyHostToDevice, B.s));
        unsigned g = 0;
        const BMap m = make_map(len, B.bs[kVecK], &g);
        hipLaunchKernelGGL(k_apply_update_bt, dim3(g), dim3(B.bs[kVecK]), 0, B.s, m, B.d_vec + 8 * kBatchMax);
        V2D_LAUNCH();
    }
    vecs.clear();
    for (int q : idx) vecs.push_back(B.cs[q]->delta);
    if (batched_norms(B, idx, vecs.data(), B.h_vec + 9 * kBatchMax, B.d_vec + 9 * kBatchMax)) return -1;
    if (fetch_out(B)) return -1;
    for (int q : idx) {
        Ctx *c = B.cs[q];
        if (sweep_failed(c->prec == V2D_FP64 ? c->wd.err_host : c->wf.err_host)) return -1;
        const double dn = std::sqrt(B.h_out[(int64_t)q * kOut]);
        res[q] = dn / (2.0 * (double)c->N);
        ++c->n_steps;
        c->n_lin += lin_iters[q];
        if (!std::isfinite(res[q])) {
            set_error("Newton step: non-finite update");
            return -1;
        }
    }
    // the cases' next work (their own calls) after the batch's
    return fork_cases(B, idx);
}

int v2d_batch_autotune(v2d_batch *bh, double re, int *threads3)
{
    Batch &B = bh->b;
    std::vector<int> idx;
    for (int q = 0; q < (int)B.cs.size(); ++q)
        if (B.cs[q]) idx.push_back(q);
    if (idx.empty()) {
        set_error("batch: no case to tune on");
        return -1;
    }
    std::vector<double> r(B.cs.size(), re);
    if (join_cases(B, idx)) return -1;
    const int cand[4] = {128, 256, 512, 1024};
    hipEvent_t e0, e1;
    V2D_HIP(hipEventCreate(&e0));
    V2D_HIP(hipEventCreate(&e1));
    int best[kKinds] = {256, 256, 256};
    // assembly (the Jacobian and residual of the current states are rewritten,
    // as any assembly does), SpMV into the spare GMRES basis vector (mixed) or
    // the update buffer (FP64), and the update-vector scaling in place by 1
    float best_ms[kKinds] = {1e30f, 1e30f, 1e30f};
    for (int k = 0; k < kKinds; ++k) {
        for (int cb : cand) {
            const int keep = B.bs[k];
            B.bs[k] = cb;
            float ms = 0.f;
            for (int rep = 0; rep < 6; ++rep) {
                V2D_HIP(hipEventRecord(e0, B.s));
                if (k == kAsm) {
                    if (batched_assembly(B, idx, r.data(), nullptr, nullptr, nullptr)) return -1;
                } else {
                    std::vector<int64_t> len;
                    BVec *T = B.h_vec + 10 * kBatchMax;
                    for (size_t t = 0; t < idx.size(); ++t) {
                        Ctx *c = B.cs[idx[t]];
                        B.h_scal[2 * idx[t] + 1] = 1.0;
                        BVec v{};
                        v.imax = c->imax; v.jmax = c->jmax; v.flags = c->flags; v.val = c->val;
                        if (k == kSpmvK) {
                            v.x = c->prec == V2D_MIXED ? (const void *)c->gZ : (const void *)c->rhs;
                            v.y = c->prec == V2D_MIXED ? (void *)(c->gV + (int64_t)c->gm_cap * 2 * c->N)
                                                       : (void *)c->delta;
                            len.push_back(c->N);
                        } else {
                            v.n = 2 * c->N; v.x = c->delta; v.y = c->delta; v.s = B.d_scal + 2 * idx[t] + 1;
                            v.inv = 0;
                            len.push_back(v.n);
                        }
                        T[t] = v;
                    }
                    V2D_HIP(hipMemcpyAsync(B.d_scal, B.h_scal, sizeof(double) * 2 * kBatchMax,
                                           hipMemcpyHostToDevice, B.s));
                    V2D_HIP(hipMemcpyAsync(B.d_vec + 10 * kBatchMax, T, sizeof(BVec) * idx.size(),
                                           hipMemcpyHostToDevice, B.s));
                    unsigned g = 0;
                    const BMap m = make_map(len, cb, &g);
                    V2D_HIP(hipEventRecord(e0, B.s));
                    if (k == kSpmvK) {
                        if (B.prec == V2D_MIXED)
                            hipLaunchKernelGGL((k_spmv_bt<float, float>), dim3(g), dim3(cb), 0, B.s, m,
                                               B.d_vec + 10 * kBatchMax);
                        else
                            hipLaunchKernelGGL((k_spmv_bt<double, double>), dim3(g), dim3(cb), 0, B.s, m,
                                               B.d_vec + 10 * kBatchMax);
                    } else {
                        hipLaunchKernelGGL((k_scale_copy_bt<double, double>), dim3(g), dim3(cb), 0, B.s, m,
                                           B.d_vec + 10 * kBatchMax);
                    }
                    V2D_LAUNCH();
                }
                V2D_HIP(hipEventRecord(e1, B.s));
                V2D_HIP(hipEventSynchronize(e1));
                float t = 0.f;
                V2D_HIP(hipEventElapsedTime(&t, e0, e1));
                if (rep > 0) ms += t;   // the first run loads the code object
            }
            if (ms < best_ms[k]) {
                best_ms[k] = ms;
                best[k] = cb;
            }
            B.bs[k] = keep;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    for (int k = 0; k < kKinds; ++k) {
        B.bs[k] = best[k];
        threads3[k] = best[k];
    }
    int64_t nodes = 0;
    for (Ctx *c : B.cs) if (c) nodes += c->N;
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tuned[size_class(nodes)] = std::vector<int>(best, best + kKinds);
    return fork_cases(B, idx);
}

}  // extern "C"
