// v2d_ctx.h -- the device-resident solver context shared by the C ABI
// (v2d_capi.hip) and the batched multi-case driver (v2d_batch.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <memory>
#include <string>
#include <vector>

#include "../../include/vorti2d.h"
#include "v2d_internal.h"

namespace v2d {

// RAII device buffer: plain for the stateless host entry points (null
// stream), stream-ordered (hipMallocAsync / hipFreeAsync) on a context's
// stream, so that a context never synchronises the whole device.
struct DBuf {
    void *p = nullptr;
    hipStream_t s = nullptr;
    ~DBuf()
    {
        if (!p) return;
        if (s) (void)hipFreeAsync(p, s);
        else (void)hipFree(p);
    }
    template <typename T> T *as() { return static_cast<T *>(p); }
};
#define V2D_ALLOC(buf, bytes) V2D_HIP(hipMalloc(&(buf).p, std::max<size_t>((bytes), 8)))
#define V2D_ALLOC_S(buf, bytes, st)                                                          \
    do {                                                                                     \
        (buf).s = (st);                                                                      \
        V2D_HIP(hipMallocAsync(&(buf).p, std::max<size_t>((bytes), 8), (st)));               \
    } while (0)

// ----------------------------------------------------------- context -----
struct Ctx {
    int imax = 0, jmax = 0, prec = V2D_FP64, ff_bc = 0;
    int64_t N = 0;
    double dksi = 0, deta = 0, ca = 1, sa = 0;
    hipStream_t s = nullptr;
    int64_t bytes = 0;
    double *x = nullptr, *y = nullptr, *m = nullptr;
    double *psi = nullptr, *ome = nullptr, *omeold = nullptr, *omeoldold = nullptr;
    double *rhs = nullptr, *delta = nullptr, *red = nullptr, *partial = nullptr, *hbuf = nullptr;
    uint8_t *flags = nullptr;
    int *pre_psi = nullptr, *pre_ome = nullptr;
    int npsi_out = 0;
    int64_t nnz = 0;
    std::vector<uint8_t> hflags;
    void *val = nullptr;      // kSlots * N of T: the last assembled Jacobian
    void *val_pc = nullptr;   // the values the stored block LU was built from
    DirectPattern pat;
    DirectWork<double> wd;
    DirectWork<float> wf;
    // linear solver of the Newton systems (V2D_SOLVER_*); the storage of each
    // solver is allocated at its first factorisation
    int solver = V2D_SOLVER_MULTIFRONTAL;
    bool blu_ready = false;
    // multifrontal analysis, device tables uploaded; shared (read only) with
    // the contexts created like this one on the same grid and far field
    std::shared_ptr<MfPattern> mfp;
    MfWork<double> md;
    MfWork<float> mw;
    void drop_mf()
    {
        mf_free(md, s);
        mf_free(mw, s);
        if (mfp) {
            if (mfp.use_count() == 1) {
                mf_free_pattern(*mfp, s);   // stream-ordered after this context's work
            } else {
                (void)hipStreamSynchronize(s);   // other owners keep the tables: finish our reads
            }
        }
        mfp.reset();
    }
    // GMRES (mixed) work, FP32 vectors
    int gm_cap = 30;   // allocated Krylov vectors
    int gm_max = 30;
    double gm_rtol = 1e-6;
    int refactor_every = 1;
    float *gV = nullptr, *gZ = nullptr, *gb = nullptr;
    // captured graphs, built on first use and dropped when the pattern, the
    // solver or the sweep kernel changes: the FP32 preconditioner apply over
    // fixed buffers (ga_in -> ga_out), replayed by every GMRES iteration
    // (V2D_GRAPH_APPLY); or every GMRES iteration (apply, SpMV,
    // Gram-Schmidt, norm and read-backs) as one replay of a graph captured per
    // basis index (V2D_GRAPH_ITERATION, itg[j]); V2D_GRAPH_OFF launches eagerly
    int graph_mode = V2D_GRAPH_APPLY;
    float *ga_in = nullptr, *ga_out = nullptr;
    hipGraph_t ag = nullptr;
    hipGraphExec_t age = nullptr;
    std::vector<hipGraphExec_t> itg;
    void drop_graph()
    {
        if (age) (void)hipGraphExecDestroy(age);
        if (ag) (void)hipGraphDestroy(ag);
        age = nullptr;
        ag = nullptr;
        for (hipGraphExec_t e : itg)
            if (e) (void)hipGraphExecDestroy(e);
        itg.clear();
    }
    double *gxd = nullptr;
    // reverse sweep of the unsteady (BDF2) adjoint: the adjoint vectors of the
    // two later time levels, the current level's, its right-hand side, the
    // accumulated metric-plane gradient and (dJ/dRe, dJ/dalpha [rad]) sums;
    // allocated at the first v2d_ctx_adjoint_bdf2_begin
    double *adj_l1 = nullptr, *adj_l2 = nullptr, *adj_lam = nullptr, *adj_rhs = nullptr, *adj_gm = nullptr;
    double adj_gs[2] = {0.0, 0.0};
    int64_t adj_levels = -1;   // levels swept so far; -1: no sweep begun
    bool factored = false;
    int since_factor = 0;
    // refactorisation decision: V2D_REFACTOR_FIXED allows gm_max iterations on a
    // stored factorisation (a function of the iterates only, so trajectories do
    // not depend on GPU load); V2D_REFACTOR_MEASURED weighs the measured costs
    // below (milliseconds) instead
    int refactor_policy = V2D_REFACTOR_FIXED;
    double fac_ms_last = 0.0, it_ms_avg = 0.0;
    int fresh_its = 0;
    // inexact Newton forcing (Eisenstat-Walker choice 2); eta_max <= 0: off
    double eta_max = 0.0, eta_prev = 0.0, rn_prev = 0.0;
    double r_target = 0.0;     // residual norm the solve aims at (forcing safeguard; 0: off)
    double rtol_step = 1e-6;   // GMRES tolerance of the current Newton step
    double last_re = 0, last_invdtau = 0, last_inv2dt = 0, last_urot = 0;
    // timers: ev[0..3] bracket a Newton step (start, assembled, preconditioner
    // ready, solved); ev[4..5] bracket the factorisation made inside it, if any
    hipEvent_t ev[6];
    bool fac_in_step = false;
    double t_asm = 0, t_fac = 0, t_sol = 0, t_tot = 0;
    int64_t n_steps = 0, n_fac = 0, n_lin = 0;
    std::vector<void *> allocs;

    template <typename T> T *alloc(int64_t count, int &rc)
    {
        void *p = nullptr;
        if (hipMallocAsync(&p, sizeof(T) * std::max<int64_t>(count, 1), s) != hipSuccess) {
            set_error("hipMallocAsync of " + std::to_string(sizeof(T) * count) + " bytes failed");
            rc = -1;
            return nullptr;
        }
        allocs.push_back(p);
        bytes += sizeof(T) * std::max<int64_t>(count, 1);
        return static_cast<T *>(p);
    }
    // One pinned host block (one allocation per context): the solve error
    // words of wd, wf (err_host), then one GMRES iteration's two Gram-Schmidt
    // columns (kGmHost each) and ||w||^2 (gm_host), read back with one
    // synchronisation per iteration
    static constexpr int kGmHost = 128;
    int *err_host = nullptr;
    double *gm_host = nullptr;
    ~Ctx()
    {
        drop_graph();
        for (void *p : allocs) (void)hipFreeAsync(p, s);
        direct_free_pattern(pat, s);
        drop_mf();
        if (s) {
            (void)hipStreamSynchronize(s);
            (void)hipStreamDestroy(s);
        }
        if (err_host) (void)hipHostFree(err_host);
        for (auto &e : ev) if (e) (void)hipEventDestroy(e);
    }
    AsmArgs args(double re, double invdtau, double inv2dt, double urot) const
    {
        AsmArgs a;
        a.imax = imax; a.jmax = jmax; a.ff_bc = ff_bc;
        a.re = re; a.invdtau = invdtau; a.inv2dt = inv2dt; a.urot = urot;
        a.dksi = dksi; a.deta = deta; a.ca = ca; a.sa = sa;
        for (int q = 0; q < 10; ++q) a.m[q] = m + q * N;
        a.psi = psi; a.ome = ome; a.omeold = omeold; a.omeoldold = omeoldold;
        return a;
    }
};

// Shared pieces of the Newton step (v2d_capi.hip).
int ctx_factor(Ctx *c);
int ctx_assemble(Ctx *c, double re, double invdtau, double inv2dt, double urot);
int apply_pc(Ctx *c, const float *v, float *z);
// x = A^-1 b (trans: A^-T b) with the stored factorisation of the context's
// linear solver, on the context's stream; b and x distinct.
template <typename T>
int lin_solve(Ctx *c, const T *b, T *x, bool trans);
int sweep_failed(const int *err_host);
// GMRES tolerance of the Newton step whose residual norm is bn (Eisenstat-Walker
// forcing with the residual-target safeguard); updates the forcing history.
double forcing_rtol(Ctx *c, double bn);

// Host side of one right-preconditioned GMRES solve of v2d_capi.hip's
// gmres_f32 (Arnoldi with classical Gram-Schmidt and one re-orthogonalisation,
// Givens rotations): the single-context solve and the batched driver make
// their decisions through it, so both follow the same iterates.
struct GmresHost {
    int m = 0, budget = 0, j = 0;
    std::vector<double> H, cs, sn, g, h, h2, hist;
    double est = 1.0, hv = 1.0;
    void begin(int cap, int budget_);
    // Column j of the Arnoldi step: h, h2 (the two Gram-Schmidt passes, j + 1
    // values each) and wn2 = |w|^2 are on the host.  Returns 3 to continue
    // (hv: the norm to divide w by), 0 converged or basis full, 1 non-finite
    // or breakdown, 2 gave up (a stored factorisation projected past its
    // budget).  iters, relres: as for gmres_f32.
    int advance(double wn2, double rtol, int *iters, double *relres);
    // Least-squares coefficients of the k = min(j, m) basis vectors.
    std::vector<double> solution() const;
};

}  // namespace v2d

struct v2d_ctx {
    std::unique_ptr<v2d::Ctx> c;
};
