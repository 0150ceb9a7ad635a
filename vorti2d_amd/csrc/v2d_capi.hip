// v2d_capi.hip -- C ABI of libvorti2d: parity entry points, the device-resident
// solver context, the mixed-precision GMRES driver and the Newton step.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <utility>
#include <vector>

#include "../../include/vorti2d.h"
#include "v2d_internal.h"
#include "v2d_ctx.h"

namespace v2d {

static thread_local std::string g_err;
void set_error(const std::string &msg) { g_err = msg; }

static int check_grid(int imax, int jmax)
{
    if (imax < 4 || jmax < 4) {
        set_error("grid must have imax >= 4 and jmax >= 4 (got " + std::to_string(imax) + "x" +
                  std::to_string(jmax) + ")");
        return -1;
    }
    if ((int64_t)imax * jmax * kSlots >= (int64_t)INT32_MAX) {
        set_error("grid too large for 32-bit value indices");
        return -1;
    }
    return 0;
}

// outflow flags of the far-field row (vorti2d_core.f90:255, 308) and their
// exclusive prefix counts (the COO offsets of the reference emission order)
static void outflow_flags(int imax, const double *ex_last, const double *ey_last, int ff_bc,
                          double ca, double sa, std::vector<uint8_t> &fl, std::vector<int> &pp,
                          std::vector<int> &po, int &npsi)
{
    fl.assign(imax, 0);
    pp.assign(imax, 0);
    po.assign(imax, 0);
    int cp = 0, co = 0;
    for (int i = 0; i < imax; ++i) {
        const bool out = (ca * ex_last[i] + sa * ey_last[i]) > 0.0;
        uint8_t f = 0;
        if (ff_bc >= 1 && out) f |= kOutOme;
        if (ff_bc >= 2 && out) f |= kOutPsi;
        fl[i] = f;
        pp[i] = cp;
        po[i] = co;
        cp += (f & kOutPsi) ? 1 : 0;
        co += (f & kOutOme) ? 1 : 0;
    }
    npsi = cp;
}

static int64_t count_rows(int imax, int jmax, int64_t r0, int64_t r1, const std::vector<uint8_t> &fl)
{
    const int64_t N = (int64_t)imax * jmax;
    int64_t c = 0;
    for (int64_t r = r0; r < r1; ++r) {
        const int comp = r >= N;
        const int64_t k = comp ? r - N : r;
        const int j = (int)(k / imax), i = (int)(k % imax);
        c += row_count(jmax, j, comp, j == jmax - 1 ? fl[i] : 0);
    }
    return c;
}

static int ctx_bc(Ctx *c, int ff_bc, double ca, double sa)
{
    c->ff_bc = ff_bc; c->ca = ca; c->sa = sa;
    const int64_t N = c->N;
    std::vector<double> ex(c->imax), ey(c->imax);
    V2D_HIP(hipMemcpyAsync(ex.data(), c->m + 6 * N + (N - c->imax), sizeof(double) * c->imax, hipMemcpyDeviceToHost, c->s));
    V2D_HIP(hipMemcpyAsync(ey.data(), c->m + 7 * N + (N - c->imax), sizeof(double) * c->imax, hipMemcpyDeviceToHost, c->s));
    V2D_HIP(hipStreamSynchronize(c->s));
    std::vector<int> pp, po;
    outflow_flags(c->imax, ex.data(), ey.data(), ff_bc, ca, sa, c->hflags, pp, po, c->npsi_out);
    V2D_HIP(hipMemcpyAsync(c->flags, c->hflags.data(), c->imax, hipMemcpyHostToDevice, c->s));
    V2D_HIP(hipMemcpyAsync(c->pre_psi, pp.data(), sizeof(int) * c->imax, hipMemcpyHostToDevice, c->s));
    V2D_HIP(hipMemcpyAsync(c->pre_ome, po.data(), sizeof(int) * c->imax, hipMemcpyHostToDevice, c->s));
    V2D_HIP(hipStreamSynchronize(c->s));
    c->nnz = count_rows(c->imax, c->jmax, 0, 2 * N, c->hflags);
    c->drop_graph();
    c->drop_mf();
    direct_free_pattern(c->pat, c->s);
    if (direct_build_pattern(c->pat, c->imax, c->jmax, c->hflags.data(), c->s)) return -1;
    c->factored = false;
    return 0;
}

// ------------------------------------------------------- linear solver -----
// Block-LU storage of the context's precision, allocated at its first
// factorisation.
static int ensure_block_lu(Ctx *c)
{
    if (c->blu_ready) return 0;
    int rc = 0;
    const int64_t N = c->N, nm = c->pat.nmax;
    if (c->prec == V2D_FP64) {
        c->wd.sinv = c->alloc<double>(c->pat.sinv_elems, rc);
        c->wd.tmat = c->alloc<double>(2 * nm * c->pat.ldmax, rc);
        c->wd.colk = c->alloc<double>(2 * nm * 64, rc);
        c->wd.rowk = c->alloc<double>(2 * (int64_t)c->pat.ldmax * 64, rc);
        c->wd.colk2 = c->alloc<double>(2 * nm * 64, rc);
        c->wd.rowk2 = c->alloc<double>(2 * (int64_t)c->pat.ldmax * 64, rc);
        c->wd.dinv = c->alloc<double>(2 * 64 * 64, rc);
        c->wd.z = c->alloc<double>(2 * N, rc);
        c->wd.wv = c->alloc<double>(2 * N, rc);
        c->wd.cfL = c->alloc<double>(kCouple * 2 * N, rc);
        c->wd.cfU = c->alloc<double>(kCouple * 2 * N, rc);
        c->wd.cfC = c->alloc<double>(std::max<int64_t>(1, c->pat.ncx), rc);
        c->wd.yv = c->alloc<double>(2 * N, rc);
        c->wd.ctr = c->alloc<int>(std::max(1, c->pat.sweep_grid), rc);
    } else {
        c->wf.sinv = c->alloc<float>(c->pat.sinv_elems, rc);
        c->wf.tmat = c->alloc<float>(2 * nm * c->pat.ldmax, rc);
        c->wf.colk = c->alloc<float>(2 * nm * 64, rc);
        c->wf.rowk = c->alloc<float>(2 * (int64_t)c->pat.ldmax * 64, rc);
        c->wf.colk2 = c->alloc<float>(2 * nm * 64, rc);
        c->wf.rowk2 = c->alloc<float>(2 * (int64_t)c->pat.ldmax * 64, rc);
        c->wf.dinv = c->alloc<float>(2 * 64 * 64, rc);
        c->wf.z = c->alloc<float>(2 * N, rc);
        c->wf.wv = c->alloc<float>(2 * N, rc);
        c->wf.cfL = c->alloc<float>(kCouple * 2 * N, rc);
        c->wf.cfU = c->alloc<float>(kCouple * 2 * N, rc);
        c->wf.cfC = c->alloc<float>(std::max<int64_t>(1, c->pat.ncx), rc);
        c->wf.yv = c->alloc<float>(2 * N, rc);
        c->wf.ctr = c->alloc<int>(std::max(1, c->pat.sweep_grid), rc);
    }
    if (rc) return -1;
    c->blu_ready = true;
    return 0;
}

// The multifrontal analysis of the grid (host, then its tables uploaded).
static int ensure_mf_pattern(Ctx *c)
{
    if (c->mfp) return 0;
    auto p = std::make_unique<MfPattern>();
    if (mf_analyse(*p, c->imax, c->jmax, c->hflags.data())) return -1;
    if (mf_upload(*p, c->s)) return -1;
    c->mfp = std::shared_ptr<MfPattern>(std::move(p));
    return 0;
}

// ... and the multifrontal storage of the context's precision.
static int ensure_mf(Ctx *c)
{
    if (ensure_mf_pattern(c)) return -1;
    const MfPattern &P = *c->mfp;
    const int64_t elems = P.F_elems + P.V_elems + P.Z_elems + P.pool_elems;
    if (c->prec == V2D_FP64 && !c->md.F) {
        if (mf_alloc(P, c->md, c->s)) return -1;
        c->bytes += (int64_t)sizeof(double) * elems;
    }
    if (c->prec != V2D_FP64 && !c->mw.F) {
        if (mf_alloc(P, c->mw, c->s)) return -1;
        c->bytes += (int64_t)sizeof(float) * elems;
    }
    return 0;
}

template <typename T> static DirectWork<T> &blu_work(Ctx *c);
template <> DirectWork<double> &blu_work<double>(Ctx *c) { return c->wd; }
template <> DirectWork<float> &blu_work<float>(Ctx *c) { return c->wf; }
template <typename T> static MfWork<T> &mf_work(Ctx *c);
template <> MfWork<double> &mf_work<double>(Ctx *c) { return c->md; }
template <> MfWork<float> &mf_work<float>(Ctx *c) { return c->mw; }

// x = A^-1 b (trans: A^-T b) with the stored factorisation of the context's
// solver, queued on the context's stream.
template <typename T>
int lin_solve(Ctx *c, const T *b, T *x, bool trans)
{
    if (c->solver == V2D_SOLVER_MULTIFRONTAL) return mf_solve<T>(*c->mfp, mf_work<T>(c), b, x, trans, c->s);
    const T *v = static_cast<const T *>(c->val_pc);
    return trans ? direct_solve_t<T>(c->pat, blu_work<T>(c), v, b, x, c->s)
                 : direct_solve<T>(c->pat, blu_work<T>(c), v, b, x, c->s);
}
template int lin_solve<double>(Ctx *, const double *, double *, bool);
template int lin_solve<float>(Ctx *, const float *, float *, bool);

// A persistent solve sweep that waited past its spin limit leaves 2 in the
// error word; direct_solve copies the word to pinned memory after every solve,
// so it is current after the caller's next synchronisation.
int sweep_failed(const int *err_host)
{
    if (err_host && *err_host) {
        set_error("block-LU solve: a sweep step waited past its spin limit (results invalid)");
        return 1;
    }
    return 0;
}

// z = M^-1 v with the stored FP32 block LU.  With graphs on and the per-level
// sweeps, the apply is one replay of a graph captured over ga_in -> ga_out
// (captured on first use; its kernels read the factorisation through fixed
// pointers, so it stays valid across refactorisations) between two
// device-to-device copies; otherwise the kernels are launched directly.
int apply_pc(Ctx *c, const float *v, float *z)
{
    const int64_t n = 2 * c->N;
    const bool mf = c->solver == V2D_SOLVER_MULTIFRONTAL;
    if (c->graph_mode == V2D_GRAPH_OFF || (!mf && c->wf.sweep_kind == V2D_SWEEP_PERSISTENT) || !c->ga_in || !c->ga_out)
        return lin_solve<float>(c, v, z, false);
    if (!c->age) {
        V2D_HIP(hipStreamBeginCapture(c->s, hipStreamCaptureModeThreadLocal));
        const int rc = lin_solve<float>(c, c->ga_in, c->ga_out, false);
        hipGraph_t g = nullptr;
        const hipError_t ec = hipStreamEndCapture(c->s, &g);
        if (rc || ec != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            if (!rc) set_error(std::string("graph capture of the block-LU apply failed: ") + hipGetErrorString(ec));
            return -1;
        }
        c->ag = g;
        V2D_HIP(hipGraphInstantiate(&c->age, c->ag, nullptr, nullptr, 0));
    }
    V2D_HIP(hipMemcpyAsync(c->ga_in, v, sizeof(float) * n, hipMemcpyDeviceToDevice, c->s));
    V2D_HIP(hipGraphLaunch(c->age, c->s));
    V2D_HIP(hipMemcpyAsync(z, c->ga_out, sizeof(float) * n, hipMemcpyDeviceToDevice, c->s));
    return 0;
}

// ------------------------------------------------------------- GMRES ------
void GmresHost::begin(int cap, int budget_)
{
    m = cap;
    budget = budget_;
    j = 0;
    H.assign((size_t)(m + 1) * m, 0.0);
    cs.assign(m, 0.0);
    sn.assign(m, 0.0);
    g.assign(m + 1, 0.0);
    g[0] = 1.0;
    h.assign(m + 1, 0.0);
    h2.assign(m + 1, 0.0);
    hist.assign(m, 1.0);
    est = 1.0;
    hv = 1.0;
}

int GmresHost::advance(double wn2, double rtol, int *iters, double *relres)
{
    const double hn = std::sqrt(std::max(wn2, 0.0));
    if (!std::isfinite(hn)) {   // overflow in the FP32 vectors: not converged
        *iters = j + 1;
        *relres = INFINITY;
        return 1;
    }
    for (int i = 0; i <= j; ++i) H[i * m + j] = h[i] + h2[i];
    H[(j + 1) * m + j] = hn;
    hv = hn > 0.0 ? hn : 1.0;   // v_{j+1} = w / hv
    for (int i = 0; i < j; ++i) {
        const double t = cs[i] * H[i * m + j] + sn[i] * H[(i + 1) * m + j];
        H[(i + 1) * m + j] = -sn[i] * H[i * m + j] + cs[i] * H[(i + 1) * m + j];
        H[i * m + j] = t;
    }
    const double d = std::hypot(H[j * m + j], H[(j + 1) * m + j]);
    cs[j] = d > 0.0 ? H[j * m + j] / d : 1.0;
    sn[j] = d > 0.0 ? H[(j + 1) * m + j] / d : 0.0;
    H[j * m + j] = d;
    H[(j + 1) * m + j] = 0.0;
    g[j + 1] = -sn[j] * g[j];
    g[j] = cs[j] * g[j];
    est = std::fabs(g[j + 1]);
    hist[j] = est;
    if (!std::isfinite(est) || !(d > 0.0)) {
        *iters = j + 1;
        *relres = INFINITY;
        return 1;
    }
    if (est <= rtol || hn == 0.0) {
        ++j;
        *iters = j;
        *relres = est;
        return 0;
    }
    if (budget > 0 && j >= 2) {
        // geometric projection of the iterations this preconditioner needs,
        // from the average rate so far, and of the iterations still needed
        // from the rate of the last two (a stale factorisation often gives
        // one large first reduction and then stalls, which only the recent
        // rate shows)
        const double rate = std::pow(est, 1.0 / (j + 1));
        const double need = rate < 1.0 ? std::log(rtol) / std::log(rate) : INFINITY;
        double rest = 0.0;
        if (j >= 3) {
            const double rr = std::sqrt(est / hist[j - 2]);
            rest = rr < 1.0 ? std::log(rtol / est) / std::log(rr) : INFINITY;
        }
        if (need > budget || rest > budget) {
            *iters = j + 1;
            *relres = est;
            return 2;
        }
    }
    ++j;
    if (j >= m) {   // basis full
        *iters = m;
        *relres = est;
        return 0;
    }
    return 3;
}

std::vector<double> GmresHost::solution() const
{
    const int k = std::min(j, m);
    std::vector<double> yv(k, 0.0);
    for (int i = k - 1; i >= 0; --i) {
        double acc = g[i];
        for (int q = i + 1; q < k; ++q) acc -= H[i * m + q] * yv[q];
        yv[i] = acc / H[i * m + i];
    }
    return yv;
}

// Enqueues GMRES iteration j's device work on the context stream: z_j = M^-1
// v_j, w = A z_j, two classical Gram-Schmidt rounds of w against v_0 .. v_j,
// ||w||^2 and v_{j+1} = w / ||w||, with both Gram-Schmidt columns and ||w||^2
// copied to pinned host memory.  captured: the enqueue is being recorded into
// a graph, so the apply is issued directly (apply_pc would nest its own
// capture and copy through ga_in / ga_out).
static int gmres_enqueue(Ctx *c, int j, bool captured)
{
    const int64_t n = 2 * c->N;
    const float *val = static_cast<const float *>(c->val);        // operator (the preconditioner: apply_pc)
    const unsigned nb = std::min<unsigned>(1024, grid_for(n));
    float *vj = c->gV + j * n, *zj = c->gZ + j * n, *w = c->gV + (j + 1) * n;
    if (captured ? lin_solve<float>(c, vj, zj, false) : apply_pc(c, vj, zj)) return -1;
    hipLaunchKernelGGL((k_spmv<float, float>), dim3(grid_for(c->N)), dim3(256), 0, c->s, c->imax,
                       c->jmax, c->flags, val, zj, w);
    V2D_LAUNCH();
    // The columns and ||w||^2 go to pinned host memory without a wait, and
    // v_{j+1} = w / ||w|| is formed on the device before the host reads
    // them (unused when the host then stops), so an iteration waits for
    // the host once.
    for (int pass = 0; pass < 2; ++pass) {
        hipLaunchKernelGGL(k_multidot<float>, dim3(nb), dim3(256), 0, c->s, n, j + 1, c->gV, n, w,
                           c->partial);
        hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(256), 0, c->s, (int)nb, j + 1,
                           c->partial, c->hbuf);
        hipLaunchKernelGGL(k_multiaxpy<float>, dim3(grid_for(n)), dim3(256), 0, c->s, n, j + 1,
                           c->gV, n, c->hbuf, w);
        V2D_LAUNCH();
        V2D_HIP(hipMemcpyAsync(c->gm_host + pass * Ctx::kGmHost, c->hbuf, sizeof(double) * (j + 1),
                               hipMemcpyDeviceToHost, c->s));
    }
    hipLaunchKernelGGL(k_multidot<float>, dim3(nb), dim3(256), 0, c->s, n, 1, w, n, w, c->partial);
    hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(256), 0, c->s, (int)nb, 1, c->partial,
                       c->red);
    V2D_LAUNCH();
    V2D_HIP(hipMemcpyAsync(c->gm_host + 2 * Ctx::kGmHost, c->red, sizeof(double), hipMemcpyDeviceToHost, c->s));
    hipLaunchKernelGGL((k_scale_copy<float, float>), dim3(grid_for(n)), dim3(256), 0, c->s, n, w, c->red, 2,
                       w);
    V2D_LAUNCH();
    return 0;
}

// Iteration j's device work as one replay of a graph captured on first use
// (kernel and device-to-pinned-host copy nodes only; every buffer it touches
// lives as long as the context and the factorisation is read through fixed
// pointers, so the graph stays valid across refactorisations).
static int gmres_replay(Ctx *c, int j)
{
    if ((int)c->itg.size() <= j) c->itg.resize(j + 1, nullptr);
    if (!c->itg[j]) {
        V2D_HIP(hipStreamBeginCapture(c->s, hipStreamCaptureModeThreadLocal));
        const int rc = gmres_enqueue(c, j, true);
        hipGraph_t g = nullptr;
        const hipError_t ec = hipStreamEndCapture(c->s, &g);
        hipGraphExec_t e = nullptr;
        const hipError_t ei = (!rc && ec == hipSuccess) ? hipGraphInstantiate(&e, g, nullptr, nullptr, 0) : ec;
        if (g) (void)hipGraphDestroy(g);
        if (rc || ei != hipSuccess) {
            if (!rc) set_error(std::string("graph capture of a GMRES iteration failed: ") + hipGetErrorString(ei));
            return -1;
        }
        c->itg[j] = e;
    }
    V2D_HIP(hipGraphLaunch(c->itg[j], c->s));
    return 0;
}

// Right-preconditioned GMRES with the stored FP32 LU M (multifrontal or block
// LU), classical Gram-Schmidt with
// one re-orthogonalisation (fused multi-dot / multi-axpy), Givens rotations on
// the host (GmresHost).  Solves A x = b for unit-norm b (FP32 vectors, FP64 dot
// products); Z_j = M V_j is kept so the solution is x = Z y (flexible form).
// budget > 0 (a stored, older factorisation): give up (return 2) once the
// convergence rate seen so far projects more than `budget` iterations.
static int gmres_f32(Ctx *c, int cap, int budget, int *iters, double *relres)
{
    const int64_t n = 2 * c->N;
    const bool whole = c->graph_mode == V2D_GRAPH_ITERATION &&
                       (c->solver == V2D_SOLVER_MULTIFRONTAL || c->wf.sweep_kind != V2D_SWEEP_PERSISTENT);
    GmresHost G;
    G.begin(cap, budget);
    // v0 = b (unit norm already) -> gV[0]
    V2D_HIP(hipMemcpyAsync(c->gV, c->gb, sizeof(float) * n, hipMemcpyDeviceToDevice, c->s));
    for (;;) {
        const int j = G.j;
        if (whole ? gmres_replay(c, j) : gmres_enqueue(c, j, false)) return -1;
        V2D_HIP(hipStreamSynchronize(c->s));
        if (sweep_failed(c->wf.err_host)) return -1;
        std::copy(c->gm_host, c->gm_host + j + 1, G.h.begin());
        std::copy(c->gm_host + Ctx::kGmHost, c->gm_host + Ctx::kGmHost + j + 1, G.h2.begin());
        const int st = G.advance(c->gm_host[2 * Ctx::kGmHost], c->rtol_step, iters, relres);
        if (st == 1 || st == 2) return st;
        if (st == 0) break;
    }
    const std::vector<double> yv = G.solution();
    const int k = (int)yv.size();
    V2D_HIP(hipMemcpyAsync(c->hbuf, yv.data(), sizeof(double) * k, hipMemcpyHostToDevice, c->s));
    hipLaunchKernelGGL((k_fill_zero_d), dim3(grid_for(n)), dim3(256), 0, c->s, n, c->gxd);
    hipLaunchKernelGGL((k_update_sol<float, double>), dim3(grid_for(n)), dim3(256), 0, c->s, n, k, c->gZ,
                       n, c->hbuf, 1.0, c->gxd);
    V2D_LAUNCH();
    V2D_HIP(hipStreamSynchronize(c->s));   // yv is a local
    *iters = k;
    return 0;
}

}  // namespace v2d

using namespace v2d;


struct v2d_mf_analysis {
    v2d::MfPattern p;
};

extern "C" {

const char *v2d_last_error(void) { return g_err.c_str(); }
int v2d_version(void) { return 100; }

int v2d_node_neighbors(int i, int j, int imax, int out[9])
{
    if (imax < 4 || i < 1 || i > imax || j < 1) {
        set_error("node_neighbors: index out of range");
        return -1;
    }
    const Nbr b = neighbors(i - 1, j - 1, imax);
    const int v[9] = {b.k, b.e, b.w, b.n, b.s, b.ne, b.nw, b.se, b.sw};
    for (int q = 0; q < 9; ++q) out[q] = v[q] + 1;
    return 0;
}

int v2d_compute_metrics_dev(int imax, int jmax, double dksi, double deta, const double *d_xg,
                            const double *d_yg, double *d_out, void *stream)
{
    if (check_grid(imax, jmax)) return -1;
    const int64_t N = (int64_t)imax * jmax;
    hipLaunchKernelGGL(k_metrics, dim3(grid_for(N)), dim3(256), 0, (hipStream_t)stream, imax, jmax,
                       dksi, deta, d_xg, d_yg, d_out);
    V2D_LAUNCH();
    return 0;
}

int v2d_compute_metrics(int imax, int jmax, double dksi, double deta, const double *xg,
                        const double *yg, double *out)
{
    if (check_grid(imax, jmax)) return -1;
    const int64_t N = (int64_t)imax * jmax;
    DBuf dx, dy, dout;
    V2D_ALLOC(dx, sizeof(double) * N);
    V2D_ALLOC(dy, sizeof(double) * N);
    V2D_ALLOC(dout, sizeof(double) * 10 * N);
    V2D_HIP(hipMemcpy(dx.p, xg, sizeof(double) * N, hipMemcpyHostToDevice));
    V2D_HIP(hipMemcpy(dy.p, yg, sizeof(double) * N, hipMemcpyHostToDevice));
    if (v2d_compute_metrics_dev(imax, jmax, dksi, deta, dx.as<double>(), dy.as<double>(),
                                dout.as<double>(), nullptr))
        return -1;
    V2D_HIP(hipMemcpy(out, dout.p, sizeof(double) * 10 * N, hipMemcpyDeviceToHost));
    return 0;
}

int v2d_compute_metrics_local(int gxm, int jmax, int il0, int il1, double dksi, double deta,
                              const double *xgl, const double *ygl, double *out)
{
    if (gxm < 3 || jmax < 4 || il0 < 2 || il1 > gxm - 1 || il0 > il1) {
        set_error("compute_metrics_local: owned columns must satisfy 2 <= il0 <= il1 <= gxm-1");
        return -1;
    }
    const int64_t N = (int64_t)gxm * jmax;
    DBuf dx, dy, dout;
    V2D_ALLOC(dx, sizeof(double) * N);
    V2D_ALLOC(dy, sizeof(double) * N);
    V2D_ALLOC(dout, sizeof(double) * 10 * N);
    V2D_HIP(hipMemcpy(dx.p, xgl, sizeof(double) * N, hipMemcpyHostToDevice));
    V2D_HIP(hipMemcpy(dy.p, ygl, sizeof(double) * N, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_metrics_local, dim3(grid_for(N)), dim3(256), 0, nullptr, gxm, jmax, il0 - 1,
                       il1 - 1, dksi, deta, dx.as<double>(), dy.as<double>(), dout.as<double>());
    V2D_LAUNCH();
    V2D_HIP(hipMemcpy(out, dout.p, sizeof(double) * 10 * N, hipMemcpyDeviceToHost));
    return 0;
}

// upload 10 metric planes + 4 state planes into one device block (14 N)
static int upload_fields(DBuf &buf, int64_t N, const double *const *metrics, const double *psi,
                         const double *ome, const double *omeold, const double *omeoldold,
                         AsmArgs &a)
{
    V2D_ALLOC(buf, sizeof(double) * 14 * N);
    double *b = buf.as<double>();
    for (int q = 0; q < 10; ++q) {
        V2D_HIP(hipMemcpy(b + q * N, metrics[q], sizeof(double) * N, hipMemcpyHostToDevice));
        a.m[q] = b + q * N;
    }
    const double *st[4] = {psi, ome, omeold, omeoldold};
    for (int q = 0; q < 4; ++q)
        V2D_HIP(hipMemcpy(b + (10 + q) * N, st[q], sizeof(double) * N, hipMemcpyHostToDevice));
    a.psi = b + 10 * N; a.ome = b + 11 * N; a.omeold = b + 12 * N; a.omeoldold = b + 13 * N;
    return 0;
}

int v2d_assemble_coo(int imax, int jmax, double re, double invdtau, double inv2dt, double urot,
                     double dksi, double deta, const double *const *metrics, const double *psi,
                     const double *ome, const double *omeold, const double *omeoldold, int r0,
                     int r1, int maxnnz, int ff_bc, double ca, double sa, int *coo_i, int *coo_j,
                     double *coo_v, int *nnz, double *bvec)
{
    if (check_grid(imax, jmax)) return -1;
    const int64_t N = (int64_t)imax * jmax;
    if (r0 < 0 || r1 > 2 * N || r0 > r1) {
        set_error("assemble_coo: row range outside [0, 2*imax*jmax)");
        return -1;
    }
    std::vector<uint8_t> fl;
    std::vector<int> pp, po;
    int npsi = 0;
    outflow_flags(imax, metrics[6] + (N - imax), metrics[7] + (N - imax), ff_bc, ca, sa, fl, pp, po, npsi);
    const int64_t cnt = count_rows(imax, jmax, r0, r1, fl);
    if (cnt > maxnnz) {
        set_error("assemble_coo: " + std::to_string(cnt) + " entries exceed maxnnz=" +
                  std::to_string(maxnnz));
        return -1;
    }
    AsmArgs a;
    a.imax = imax; a.jmax = jmax; a.ff_bc = ff_bc;
    a.re = re; a.invdtau = invdtau; a.inv2dt = inv2dt; a.urot = urot;
    a.dksi = dksi; a.deta = deta; a.ca = ca; a.sa = sa;
    DBuf f, dfl, dpp, dpo, dci, dcj, dcv, db;
    if (upload_fields(f, N, metrics, psi, ome, omeold, omeoldold, a)) return -1;
    V2D_ALLOC(dfl, imax);
    V2D_ALLOC(dpp, sizeof(int) * imax);
    V2D_ALLOC(dpo, sizeof(int) * imax);
    V2D_ALLOC(dci, sizeof(int) * cnt);
    V2D_ALLOC(dcj, sizeof(int) * cnt);
    V2D_ALLOC(dcv, sizeof(double) * cnt);
    V2D_ALLOC(db, sizeof(double) * (r1 - r0));
    V2D_HIP(hipMemcpy(dfl.p, fl.data(), imax, hipMemcpyHostToDevice));
    V2D_HIP(hipMemcpy(dpp.p, pp.data(), sizeof(int) * imax, hipMemcpyHostToDevice));
    V2D_HIP(hipMemcpy(dpo.p, po.data(), sizeof(int) * imax, hipMemcpyHostToDevice));
    if (r1 > r0) {
        hipLaunchKernelGGL(k_assemble_coo, dim3(grid_for(r1 - r0)), dim3(256), 0, nullptr, a,
                           dfl.as<uint8_t>(), dpp.as<int>(), dpo.as<int>(), npsi, (int64_t)r0,
                           (int64_t)r1, dci.as<int>(), dcj.as<int>(), dcv.as<double>(),
                           db.as<double>());
        V2D_LAUNCH();
    }
    V2D_HIP(hipMemcpy(coo_i, dci.p, sizeof(int) * cnt, hipMemcpyDeviceToHost));
    V2D_HIP(hipMemcpy(coo_j, dcj.p, sizeof(int) * cnt, hipMemcpyDeviceToHost));
    V2D_HIP(hipMemcpy(coo_v, dcv.p, sizeof(double) * cnt, hipMemcpyDeviceToHost));
    V2D_HIP(hipMemcpy(bvec, db.p, sizeof(double) * (r1 - r0), hipMemcpyDeviceToHost));
    *nnz = (int)cnt;
    return 0;
}

int v2d_assemble_coo_local(int gxm, int jmax, int il0, int il1, double re, double invdtau,
                           double inv2dt, double urot, double dksi, double deta,
                           const double *const *metrics, const double *psi, const double *ome,
                           const double *omeold, const double *omeoldold, int r0, int nloc,
                           const int *lg, int maxnnz, int ff_bc, double ca, double sa,
                           int *coo_i, int *coo_j, double *coo_v, int *nnz, double *bvec)
{
    if (gxm < 3 || jmax < 4 || il0 < 2 || il1 > gxm - 1 || il0 > il1) {
        set_error("assemble_coo_local: owned columns must satisfy 2 <= il0 <= il1 <= gxm-1");
        return -1;
    }
    const int64_t N = (int64_t)gxm * jmax;
    const int ow = il1 - il0 + 1;
    const int64_t nown = (int64_t)ow * jmax;
    // entry offsets per owned node (j outer, i inner; psi row then omega row)
    std::vector<int64_t> off(nown + 1, 0);
    const double *ex = metrics[6], *ey = metrics[7];
    for (int64_t t = 0; t < nown; ++t) {
        const int j = (int)(t / ow), i = il0 - 1 + (int)(t % ow);
        uint8_t f = 0;
        if (j == jmax - 1) {
            const int64_t k = (int64_t)gxm * j + i;
            const bool out = (ca * ex[k] + sa * ey[k]) > 0.0;
            if (ff_bc >= 1 && out) f |= kOutOme;
            if (ff_bc >= 2 && out) f |= kOutPsi;
        }
        off[t + 1] = off[t] + row_count(jmax, j, 0, f) + row_count(jmax, j, 1, f);
    }
    const int64_t cnt = off[nown];
    if (cnt > maxnnz) {
        set_error("assemble_coo_local: " + std::to_string(cnt) + " entries exceed maxnnz=" +
                  std::to_string(maxnnz));
        return -1;
    }
    AsmArgs a;
    a.imax = gxm; a.jmax = jmax; a.ff_bc = ff_bc;
    a.re = re; a.invdtau = invdtau; a.inv2dt = inv2dt; a.urot = urot;
    a.dksi = dksi; a.deta = deta; a.ca = ca; a.sa = sa;
    DBuf f, dlg, doff, dci, dcj, dcv, db;
    if (upload_fields(f, N, metrics, psi, ome, omeold, omeoldold, a)) return -1;
    V2D_ALLOC(dlg, sizeof(int) * 2 * N);
    V2D_ALLOC(doff, sizeof(int64_t) * (nown + 1));
    V2D_ALLOC(dci, sizeof(int) * cnt);
    V2D_ALLOC(dcj, sizeof(int) * cnt);
    V2D_ALLOC(dcv, sizeof(double) * cnt);
    V2D_ALLOC(db, sizeof(double) * nloc);
    V2D_HIP(hipMemcpy(dlg.p, lg, sizeof(int) * 2 * N, hipMemcpyHostToDevice));
    V2D_HIP(hipMemcpy(doff.p, off.data(), sizeof(int64_t) * (nown + 1), hipMemcpyHostToDevice));
    V2D_HIP(hipMemset(db.p, 0, sizeof(double) * std::max(nloc, 1)));
    hipLaunchKernelGGL(k_assemble_coo_local, dim3(grid_for(nown)), dim3(256), 0, nullptr, a, il0 - 1,
                       ow, dlg.as<int>(), doff.as<int64_t>(), r0, nloc, dci.as<int>(), dcj.as<int>(),
                       dcv.as<double>(), db.as<double>());
    V2D_LAUNCH();
    V2D_HIP(hipMemcpy(coo_i, dci.p, sizeof(int) * cnt, hipMemcpyDeviceToHost));
    V2D_HIP(hipMemcpy(coo_j, dcj.p, sizeof(int) * cnt, hipMemcpyDeviceToHost));
    V2D_HIP(hipMemcpy(coo_v, dcv.p, sizeof(double) * cnt, hipMemcpyDeviceToHost));
    V2D_HIP(hipMemcpy(bvec, db.p, sizeof(double) * nloc, hipMemcpyDeviceToHost));
    *nnz = (int)cnt;
    return 0;
}

int v2d_force_coeffs(int imax, int jmax, double re, double dksi, double deta,
                     const double *beta, const double *gama, const double *detadx,
                     const double *detady, const double *xphys, const double *yphys,
                     const double *ome, double ref_length, double x0, double y0, double *out)
{
    if (check_grid(imax, jmax)) return -1;
    const int64_t N = (int64_t)imax * jmax;
    DBuf dm, dome, dout;
    V2D_ALLOC(dm, sizeof(double) * 10 * N);
    V2D_ALLOC(dome, sizeof(double) * N);
    V2D_ALLOC(dout, sizeof(double) * 10);
    const double *planes[10] = {nullptr, nullptr, beta, gama, nullptr, nullptr, detadx, detady, xphys, yphys};
    for (int q = 0; q < 10; ++q)
        if (planes[q])
            V2D_HIP(hipMemcpy(dm.as<double>() + q * N, planes[q], sizeof(double) * N, hipMemcpyHostToDevice));
    V2D_HIP(hipMemcpy(dome.p, ome, sizeof(double) * N, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_forces, dim3(1), dim3(1024), 0, nullptr, imax, re, dksi, deta, dm.as<double>(),
                       N, dome.as<double>(), ref_length, x0, y0, dout.as<double>());
    V2D_LAUNCH();
    V2D_HIP(hipMemcpy(out, dout.p, sizeof(double) * 10, hipMemcpyDeviceToHost));
    return 0;
}

// ------------------------------------------------------------ context -----
// A context on the same mesh, far-field setting and state as src: the mesh
// and state are copied on the device, the block-LU pattern is copied and the
// multifrontal analysis (with its device tables) is shared, not rebuilt on the
// host.
static int ctx_bc_copy(Ctx *c, const Ctx *src)
{
    c->ff_bc = src->ff_bc; c->ca = src->ca; c->sa = src->sa;
    c->hflags = src->hflags;
    c->npsi_out = src->npsi_out;
    c->nnz = src->nnz;
    V2D_HIP(hipMemcpyAsync(c->flags, src->flags, c->imax, hipMemcpyDeviceToDevice, c->s));
    V2D_HIP(hipMemcpyAsync(c->pre_psi, src->pre_psi, sizeof(int) * c->imax, hipMemcpyDeviceToDevice, c->s));
    V2D_HIP(hipMemcpyAsync(c->pre_ome, src->pre_ome, sizeof(int) * c->imax, hipMemcpyDeviceToDevice, c->s));
    c->drop_graph();
    c->drop_mf();
    c->mfp = src->mfp;
    c->solver = src->solver;
    if (direct_copy_pattern(c->pat, src->pat, c->s)) return -1;
    c->factored = false;
    return 0;
}

static int ctx_create(v2d_ctx **out, int imax, int jmax, const double *xg, const double *yg, int ff_bc, double ca,
                      double sa, int precision, const Ctx *src)
{
    *out = nullptr;
    if (check_grid(imax, jmax)) return -1;
    if (precision != V2D_FP64 && precision != V2D_MIXED) {
        set_error("precision must be V2D_FP64 (0) or V2D_MIXED (1)");
        return -1;
    }
    auto h = std::make_unique<v2d_ctx>();
    h->c = std::make_unique<Ctx>();
    Ctx *c = h->c.get();
    c->imax = imax; c->jmax = jmax; c->prec = precision;
    c->N = (int64_t)imax * jmax;
    c->dksi = 1.0 / (imax - 1);
    c->deta = 1.0 / (jmax - 1);
    const int64_t N = c->N;
    // non-blocking: independent contexts (a batch of solves) never wait on each
    // other through the null stream
    V2D_HIP(hipStreamCreateWithFlags(&c->s, hipStreamNonBlocking));
    for (auto &e : c->ev) V2D_HIP(hipEventCreate(&e));
    int rc = 0;
    c->x = c->alloc<double>(N, rc); c->y = c->alloc<double>(N, rc);
    c->m = c->alloc<double>(10 * N, rc);
    c->psi = c->alloc<double>(N, rc); c->ome = c->alloc<double>(N, rc);
    c->omeold = c->alloc<double>(N, rc); c->omeoldold = c->alloc<double>(N, rc);
    c->rhs = c->alloc<double>(2 * N, rc); c->delta = c->alloc<double>(2 * N, rc);
    c->red = c->alloc<double>(16, rc);
    c->partial = c->alloc<double>(1024 * 64, rc);
    c->hbuf = c->alloc<double>(128, rc);
    c->flags = c->alloc<uint8_t>(imax, rc);
    c->pre_psi = c->alloc<int>(imax, rc); c->pre_ome = c->alloc<int>(imax, rc);
    if (rc) return -1;
    if (src) {
        V2D_HIP(hipStreamSynchronize(src->s));   // the source's queued work (state updates) first
        const std::pair<double *, const double *> cp[] = {{c->x, src->x}, {c->y, src->y}, {c->psi, src->psi},
                                                          {c->ome, src->ome}, {c->omeold, src->omeold},
                                                          {c->omeoldold, src->omeoldold}};
        for (const auto &q : cp)
            V2D_HIP(hipMemcpyAsync(q.first, q.second, sizeof(double) * N, hipMemcpyDeviceToDevice, c->s));
    } else {
        V2D_HIP(hipMemcpyAsync(c->x, xg, sizeof(double) * N, hipMemcpyHostToDevice, c->s));
        V2D_HIP(hipMemcpyAsync(c->y, yg, sizeof(double) * N, hipMemcpyHostToDevice, c->s));
        V2D_HIP(hipMemsetAsync(c->psi, 0, sizeof(double) * N, c->s));
        V2D_HIP(hipMemsetAsync(c->ome, 0, sizeof(double) * N, c->s));
        V2D_HIP(hipMemsetAsync(c->omeold, 0, sizeof(double) * N, c->s));
        V2D_HIP(hipMemsetAsync(c->omeoldold, 0, sizeof(double) * N, c->s));
    }
    if (v2d_compute_metrics_dev(imax, jmax, c->dksi, c->deta, c->x, c->y, c->m, c->s)) return -1;
    V2D_HIP(hipStreamSynchronize(c->s));
    if (src ? ctx_bc_copy(c, src) : ctx_bc(c, ff_bc, ca, sa)) return -1;
    const size_t tsz = precision == V2D_FP64 ? sizeof(double) : sizeof(float);
    c->val = c->alloc<unsigned char>((int64_t)kSlots * N * tsz, rc);
    c->val_pc = c->alloc<unsigned char>((int64_t)kSlots * N * tsz, rc);
    if (precision != V2D_FP64) {
        c->gV = c->alloc<float>((int64_t)(c->gm_cap + 1) * 2 * N, rc);
        c->gZ = c->alloc<float>((int64_t)c->gm_cap * 2 * N, rc);
        c->gb = c->alloc<float>(2 * N, rc);
        c->ga_in = c->alloc<float>(2 * N, rc);
        c->ga_out = c->alloc<float>(2 * N, rc);
        c->gxd = c->alloc<double>(2 * N, rc);
    }
    // pivot-ready flags (2 per precision) and the error word of the dataflow steps
    int *sync_words = c->alloc<int>(8, rc);
    if (rc) return -1;
    V2D_HIP(hipMemsetAsync(sync_words, 0, sizeof(int) * 8, c->s));
    c->wd.flags = sync_words;
    c->wd.err = sync_words + 2;
    c->wf.flags = sync_words + 4;
    c->wf.err = sync_words + 6;
    // FP32 solve sweeps: the persistent launch when it fills the GPU (two
    // chains of one workgroup per compute unit), else launch pairs.  On a
    // 1025^2 mesh the persistent launch takes 11.0 us per level step against
    // 12.3 us for graph-replayed launch pairs (the fused FP32 step, 11.6 us,
    // gains nothing once the apply is a graph replay).  A smaller persistent
    // launch is not used by default: beside other contexts' persistent
    // launches (a batch of solves) its workgroups may not all be resident,
    // and its bounded waits would expire.  FP64 solves, one per Newton step,
    // use launch pairs.
    c->wd.sweep_kind = V2D_SWEEP_LEVELS;
    c->wf.sweep_kind = (c->pat.sweep_ok && 2 * c->pat.sweep_C >= c->pat.ncu) ? V2D_SWEEP_PERSISTENT
                                                                           : V2D_SWEEP_LEVELS;
    V2D_HIP(hipHostMalloc((void **)&c->err_host, (2 + 2 * Ctx::kGmHost + 1) * sizeof(double), hipHostMallocDefault));
    c->err_host[0] = c->err_host[1] = 0;
    c->gm_host = reinterpret_cast<double *>(c->err_host) + 2;
    c->wd.err_host = c->err_host;
    c->wf.err_host = c->err_host + 1;
    V2D_HIP(hipStreamSynchronize(c->s));
    *out = h.release();
    return 0;
}

int v2d_ctx_create(v2d_ctx **out, int imax, int jmax, const double *xg, const double *yg,
                   int ff_bc, double ca, double sa, int precision)
{
    return ctx_create(out, imax, jmax, xg, yg, ff_bc, ca, sa, precision, nullptr);
}

int v2d_ctx_create_like(v2d_ctx **out, v2d_ctx *src, int precision)
{
    if (out) *out = nullptr;
    if (!out || !src || !src->c) {
        set_error("create_like: null context");
        return -1;
    }
    const Ctx *sc = src->c.get();
    return ctx_create(out, sc->imax, sc->jmax, nullptr, nullptr, sc->ff_bc, sc->ca, sc->sa, precision, sc);
}

int v2d_ctx_destroy(v2d_ctx *ctx)
{
    delete ctx;
    return 0;
}

int v2d_ctx_device_bytes(v2d_ctx *ctx, int64_t *bytes)
{
    *bytes = ctx->c->bytes;
    return 0;
}

int v2d_ctx_metrics(v2d_ctx *ctx, double *out)
{
    Ctx *c = ctx->c.get();
    V2D_HIP(hipStreamSynchronize(c->s));
    V2D_HIP(hipMemcpyAsync(out, c->m, sizeof(double) * 10 * c->N, hipMemcpyDeviceToHost, c->s));
    V2D_HIP(hipStreamSynchronize(c->s));
    return 0;
}

int v2d_ctx_set_state(v2d_ctx *ctx, const double *psi, const double *ome, const double *omeold,
                      const double *omeoldold)
{
    Ctx *c = ctx->c.get();
    const double *src[4] = {psi, ome, omeold, omeoldold};
    double *dst[4] = {c->psi, c->ome, c->omeold, c->omeoldold};
    V2D_HIP(hipStreamSynchronize(c->s));
    for (int q = 0; q < 4; ++q)
        if (src[q]) V2D_HIP(hipMemcpyAsync(dst[q], src[q], sizeof(double) * c->N, hipMemcpyHostToDevice, c->s));
    V2D_HIP(hipStreamSynchronize(c->s));
    // a new state makes the stored block LU (of an earlier Jacobian) stale
    if (psi || ome) {
        c->factored = false;
        c->rn_prev = 0.0;
        c->eta_prev = 0.0;
    }
    return 0;
}

int v2d_ctx_get_state(v2d_ctx *ctx, double *psi, double *ome, double *omeold, double *omeoldold)
{
    Ctx *c = ctx->c.get();
    double *dst[4] = {psi, ome, omeold, omeoldold};
    const double *src[4] = {c->psi, c->ome, c->omeold, c->omeoldold};
    V2D_HIP(hipStreamSynchronize(c->s));
    for (int q = 0; q < 4; ++q)
        if (dst[q]) V2D_HIP(hipMemcpyAsync(dst[q], src[q], sizeof(double) * c->N, hipMemcpyDeviceToHost, c->s));
    V2D_HIP(hipStreamSynchronize(c->s));
    return 0;
}

int v2d_ctx_init_state(v2d_ctx *ctx)
{
    Ctx *c = ctx->c.get();
    const int64_t N = c->N;
    V2D_HIP(hipStreamSynchronize(c->s));
    std::vector<double> psi(N, 0.0), z(N, 0.0);
    std::vector<double> xl(c->imax), yl(c->imax);
    V2D_HIP(hipMemcpyAsync(xl.data(), c->m + 8 * N + (N - c->imax), sizeof(double) * c->imax, hipMemcpyDeviceToHost, c->s));
    V2D_HIP(hipMemcpyAsync(yl.data(), c->m + 9 * N + (N - c->imax), sizeof(double) * c->imax, hipMemcpyDeviceToHost, c->s));
    V2D_HIP(hipStreamSynchronize(c->s));
    for (int i = 0; i < c->imax; ++i) psi[N - c->imax + i] = c->ca * yl[i] - c->sa * xl[i];
    return v2d_ctx_set_state(ctx, psi.data(), z.data(), z.data(), z.data());
}

int v2d_ctx_set_solver(v2d_ctx *ctx, int refactor_every, double gmres_rtol, int gmres_max_iter)
{
    Ctx *c = ctx->c.get();
    if (gmres_max_iter < 1 || gmres_max_iter > c->gm_cap) {
        set_error("gmres_max_iter must be in [1, " + std::to_string(c->gm_cap) + "]");
        return -1;
    }
    if (!(gmres_rtol > 0.0) || refactor_every < 0) {
        set_error("gmres_rtol must be > 0 and refactor_every >= 0");
        return -1;
    }
    c->refactor_every = refactor_every;
    c->gm_rtol = gmres_rtol;
    c->gm_max = gmres_max_iter;
    return 0;
}

int v2d_ctx_set_forcing(v2d_ctx *ctx, double eta_max, double r_target)
{
    Ctx *c = ctx->c.get();
    if (!(eta_max < 1.0)) {
        set_error("forcing eta_max must be < 1 (<= 0 disables it)");
        return -1;
    }
    if (!(r_target >= 0.0)) {
        set_error("forcing r_target must be >= 0 (0 disables it)");
        return -1;
    }
    c->eta_max = eta_max > 0.0 ? eta_max : 0.0;
    c->r_target = r_target;
    c->rn_prev = c->eta_prev = 0.0;
    return 0;
}

}  // extern "C"

namespace v2d {
int ctx_assemble(Ctx *c, double re, double invdtau, double inv2dt, double urot)
{
    const AsmArgs a = c->args(re, invdtau, inv2dt, urot);
    c->last_re = re; c->last_invdtau = invdtau; c->last_inv2dt = inv2dt; c->last_urot = urot;
    const AsmConsts k = asm_consts(re, c->dksi, c->deta);
    if (c->prec == V2D_FP64)
        hipLaunchKernelGGL(k_assemble<double>, dim3(grid_for(c->N)), dim3(256), 0, c->s, a, k, c->flags,
                           static_cast<double *>(c->val), c->rhs);
    else
        hipLaunchKernelGGL(k_assemble<float>, dim3(grid_for(c->N)), dim3(256), 0, c->s, a, k, c->flags,
                           static_cast<float *>(c->val), c->rhs);
    V2D_LAUNCH();
    return 0;
}
}  // namespace v2d

extern "C" {

int v2d_ctx_set_sweep_kernel(v2d_ctx *ctx, int kind)
{
    if (kind != V2D_SWEEP_LEVELS && kind != V2D_SWEEP_PERSISTENT && kind != V2D_SWEEP_FUSED) {
        set_error("sweep kernel must be V2D_SWEEP_LEVELS, V2D_SWEEP_PERSISTENT or V2D_SWEEP_FUSED");
        return -1;
    }
    ctx->c->wd.sweep_kind = kind;
    ctx->c->wf.sweep_kind = kind;
    ctx->c->drop_graph();
    return 0;
}

int v2d_ctx_get_sweep_kernel(v2d_ctx *ctx, int *kind, int64_t *fallbacks)
{
    const Ctx *c = ctx->c.get();
    const auto &w = c->prec == V2D_FP64 ? c->wd.sweep_kind : c->wf.sweep_kind;
    *kind = w;
    *fallbacks = c->prec == V2D_FP64 ? c->wd.sweep_fallbacks : c->wf.sweep_fallbacks;
    return 0;
}

int v2d_ctx_set_refactor_policy(v2d_ctx *ctx, int policy)
{
    if (policy != V2D_REFACTOR_FIXED && policy != V2D_REFACTOR_MEASURED) {
        set_error("refactor policy must be V2D_REFACTOR_FIXED or V2D_REFACTOR_MEASURED");
        return -1;
    }
    ctx->c->refactor_policy = policy;
    return 0;
}

int v2d_ctx_set_sweep_spin_limit(v2d_ctx *ctx, int polls)
{
    if (polls < 0) {
        set_error("sweep spin limit must be >= 0");
        return -1;
    }
    ctx->c->wd.sweep_spins = ctx->c->wf.sweep_spins = polls;
    return 0;
}

int v2d_ctx_set_graph(v2d_ctx *ctx, int mode)
{
    if (mode != V2D_GRAPH_OFF && mode != V2D_GRAPH_APPLY && mode != V2D_GRAPH_ITERATION) {
        set_error("graph mode must be V2D_GRAPH_OFF (0), V2D_GRAPH_APPLY (1) or V2D_GRAPH_ITERATION (2)");
        return -1;
    }
    Ctx *c = ctx->c.get();
    if (mode != c->graph_mode) {
        V2D_HIP(hipStreamSynchronize(c->s));
        c->drop_graph();
        c->graph_mode = mode;
    }
    return 0;
}

int v2d_ctx_set_linear_solver(v2d_ctx *ctx, int kind)
{
    if (kind != V2D_SOLVER_BLOCK_LU && kind != V2D_SOLVER_MULTIFRONTAL) {
        set_error("linear solver must be V2D_SOLVER_BLOCK_LU (0) or V2D_SOLVER_MULTIFRONTAL (1)");
        return -1;
    }
    Ctx *c = ctx->c.get();
    if (c->solver != kind) {
        V2D_HIP(hipStreamSynchronize(c->s));
        c->drop_graph();   // a captured apply belongs to one solver
        c->solver = kind;
        c->factored = false;
    }
    return 0;
}

int v2d_ctx_set_step_pairing(v2d_ctx *ctx, int paired)
{
    Ctx *c = ctx->c.get();
    V2D_HIP(hipStreamSynchronize(c->s));
    c->md.paired = c->mw.paired = paired != 0;
    return 0;
}

int v2d_ctx_set_front_fusion(v2d_ctx *ctx, int fused)
{
    Ctx *c = ctx->c.get();
    V2D_HIP(hipStreamSynchronize(c->s));
    c->md.fused = c->mw.fused = fused != 0;
    return 0;
}

int v2d_ctx_get_linear_solver(v2d_ctx *ctx, int *kind)
{
    *kind = ctx->c->solver;
    return 0;
}

int v2d_ctx_mf_info(v2d_ctx *ctx, int64_t *out, double *flops)
{
    Ctx *c = ctx->c.get();
    if (ensure_mf_pattern(c)) return -1;
    const MfPattern &P = *c->mfp;
    const int64_t v[8] = {P.nfront, P.nheight, P.nbig, P.F_elems, P.V_elems, P.pool_elems, P.nvals, P.npos};
    for (int q = 0; q < 8; ++q) out[q] = v[q];
    if (flops) *flops = P.flops;
    return 0;
}

int v2d_ctx_profile_mf(v2d_ctx *ctx, double *out)
{
    Ctx *c = ctx->c.get();
    if (c->solver != V2D_SOLVER_MULTIFRONTAL) {
        set_error("profile_mf: the context's linear solver is not the multifrontal LU");
        return -1;
    }
    if (ensure_mf(c)) return -1;
    V2D_HIP(hipStreamSynchronize(c->s));
    const size_t tsz = c->prec == V2D_FP64 ? sizeof(double) : sizeof(float);
    V2D_HIP(hipMemcpyAsync(c->val_pc, c->val, (size_t)kSlots * c->N * tsz, hipMemcpyDeviceToDevice, c->s));
    MfProbe pr;
    hipEvent_t e0, e1;
    V2D_HIP(hipEventCreate(&e0));
    V2D_HIP(hipEventCreate(&e1));
    int rc;
    V2D_HIP(hipEventRecord(e0, c->s));
    if (c->prec == V2D_FP64)
        rc = mf_factor<double>(*c->mfp, c->md, static_cast<const double *>(c->val_pc), c->s, &pr);
    else
        rc = mf_factor<float>(*c->mfp, c->mw, static_cast<const float *>(c->val_pc), c->s, &pr);
    V2D_HIP(hipEventRecord(e1, c->s));
    if (!rc) rc = pr.flush();
    float fac_ms = 0.f;
    if (!rc) {
        V2D_HIP(hipEventSynchronize(e1));
        V2D_HIP(hipEventElapsedTime(&fac_ms, e0, e1));
    }
    float sol_ms = 0.f;
    if (!rc) {
        MfProbe ps;
        V2D_HIP(hipEventRecord(e0, c->s));
        if (c->prec == V2D_FP64) {
            rc = mf_solve<double>(*c->mfp, c->md, c->rhs, c->delta, false, c->s, &ps);
        } else {
            V2D_HIP(hipMemsetAsync(c->ga_in, 0, sizeof(float) * 2 * c->N, c->s));
            rc = mf_solve<float>(*c->mfp, c->mw, c->ga_in, c->ga_out, false, c->s, &ps);
        }
        V2D_HIP(hipEventRecord(e1, c->s));
        if (!rc) rc = ps.flush();
        if (!rc) {
            V2D_HIP(hipEventSynchronize(e1));
            V2D_HIP(hipEventElapsedTime(&sol_ms, e0, e1));
            for (int g = MfProbe::kGather; g <= MfProbe::kBwd; ++g) {
                pr.ms[g] = ps.ms[g];
                pr.launches[g] = ps.launches[g];
                pr.flops[g] = ps.flops[g];
                pr.bytes[g] = ps.bytes[g];
            }
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return -1;
    c->factored = true;   // the factorisation of the last assembled Jacobian
    for (int g = 0; g < MfProbe::kGroups; ++g) {
        out[4 * g] = pr.ms[g];
        out[4 * g + 1] = (double)pr.launches[g];
        out[4 * g + 2] = pr.flops[g];
        out[4 * g + 3] = pr.bytes[g];
    }
    out[4 * MfProbe::kGroups] = fac_ms;
    out[4 * MfProbe::kGroups + 1] = sol_ms;
    return 0;
}

int v2d_ctx_mf_fronts(v2d_ctx *ctx, double *F)
{
    Ctx *c = ctx->c.get();
    if (c->solver != V2D_SOLVER_MULTIFRONTAL || !c->factored || !c->mfp) {
        set_error("mf_fronts: no multifrontal factorisation");
        return -1;
    }
    const int64_t n = c->mfp->F_elems;
    V2D_HIP(hipStreamSynchronize(c->s));
    if (c->prec == V2D_FP64) {
        V2D_HIP(hipMemcpyAsync(F, c->md.F, sizeof(double) * n, hipMemcpyDeviceToHost, c->s));
        V2D_HIP(hipStreamSynchronize(c->s));
        return 0;
    }
    std::vector<float> t(n);
    V2D_HIP(hipMemcpyAsync(t.data(), c->mw.F, sizeof(float) * n, hipMemcpyDeviceToHost, c->s));
    V2D_HIP(hipStreamSynchronize(c->s));
    for (int64_t q = 0; q < n; ++q) F[q] = t[q];
    return 0;
}

int v2d_mf_analyse(int imax, int jmax, const uint8_t *ff_flags, v2d_mf_analysis **out)
{
    *out = nullptr;
    if (check_grid(imax, jmax)) return -1;
    auto a = std::make_unique<v2d_mf_analysis>();
    std::vector<uint8_t> fl(imax, 0);
    if (ff_flags) std::copy(ff_flags, ff_flags + imax, fl.begin());
    if (mf_analyse(a->p, imax, jmax, fl.data())) return -1;
    *out = a.release();
    return 0;
}

int v2d_mf_sizes(const v2d_mf_analysis *a, int64_t *out, double *flops)
{
    const MfPattern &P = a->p;
    const int64_t v[16] = {P.nfront, P.nheight, P.nbig, P.F_elems, P.V_elems, P.Z_elems, P.nfu, P.nmap,
                           P.nch, P.npos, P.nvals, P.nrows, 0, 0, 0, 0};
    for (int q = 0; q < 16; ++q) out[q] = v[q];
    if (flops) *flops = P.flops;
    return 0;
}

int v2d_mf_export(const v2d_mf_analysis *a, int64_t *fronts, int *fu, int *ch, int *map, int64_t *aoff,
                  int *aseg, int *avid)
{
    const MfPattern &P = a->p;
    if (fronts) {
        for (int h = 0; h < P.nheight; ++h)
            for (int f = P.lv[h].f0; f < P.lv[h].f1; ++f) {
                const MfFront &F = P.fr[f];
                const int64_t r[14] = {F.fo, F.vo, F.zo, F.uo, F.mo, F.ns, F.nsp, F.nb, F.nf, F.ld, F.nch,
                                       F.co, F.big, h};
                std::copy(r, r + 14, fronts + 14 * (int64_t)f);
            }
    }
    if (fu) std::copy(P.fu.begin(), P.fu.end(), fu);
    if (ch) std::copy(P.ch.begin(), P.ch.end(), ch);
    if (map) std::copy(P.map.begin(), P.map.end(), map);
    if (aoff) std::copy(P.aoff.begin(), P.aoff.end(), aoff);
    if (aseg) std::copy(P.aseg.begin(), P.aseg.end(), aseg);
    if (avid) std::copy(P.avid.begin(), P.avid.end(), avid);
    return 0;
}

int v2d_mf_destroy(v2d_mf_analysis *a)
{
    delete a;
    return 0;
}

int v2d_ctx_set_update_kernel(v2d_ctx *ctx, int kind)
{
    if (kind != V2D_UPDATE_VALU64 && kind != V2D_UPDATE_VALU128 && kind != V2D_UPDATE_MFMA64) {
        set_error("update kernel must be V2D_UPDATE_VALU64, V2D_UPDATE_VALU128 or V2D_UPDATE_MFMA64");
        return -1;
    }
    ctx->c->wd.update_kind = kind;
    ctx->c->wf.update_kind = kind;
    ctx->c->factored = false;   // the panels of a stored factorisation are kernel-specific
    return 0;
}


int v2d_ctx_profile_factor(v2d_ctx *ctx, double *out)
{
    Ctx *c = ctx->c.get();
    LaunchProbe probe;
    c->wd.probe = c->wf.probe = &probe;
    V2D_HIP(hipStreamSynchronize(c->s));
    const auto t0 = std::chrono::steady_clock::now();
    const int rc = ctx_factor(c);
    c->wd.probe = c->wf.probe = nullptr;
    if (rc) return -1;
    V2D_HIP(hipStreamSynchronize(c->s));
    const double wall = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    out[0] = (double)probe.launches;
    out[1] = probe.ms;
    out[2] = probe.flops;
    out[3] = probe.bytes;
    out[4] = wall;
    return 0;
}

int v2d_ctx_assemble(v2d_ctx *ctx, double re, double invdtau, double inv2dt, double urot)
{
    return ctx_assemble(ctx->c.get(), re, invdtau, inv2dt, urot);
}

static int dnorm2(Ctx *c, const double *v, int64_t n, double *out)
{
    const unsigned nb = std::min<unsigned>(1024, grid_for(n));
    hipLaunchKernelGGL(k_multidot<double>, dim3(nb), dim3(256), 0, c->s, n, 1, v, n, v, c->partial);
    hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(256), 0, c->s, (int)nb, 1, c->partial, c->red);
    V2D_LAUNCH();
    double s2 = 0.0;
    V2D_HIP(hipMemcpyAsync(&s2, c->red, sizeof(double), hipMemcpyDeviceToHost, c->s));
    V2D_HIP(hipStreamSynchronize(c->s));
    *out = std::sqrt(s2);
    return 0;
}

int v2d_ctx_rhs_norm(v2d_ctx *ctx, double *out)
{
    Ctx *c = ctx->c.get();
    return dnorm2(c, c->rhs, 2 * c->N, out);
}

int v2d_ctx_nnz(v2d_ctx *ctx, int64_t *nnz)
{
    *nnz = ctx->c->nnz;
    return 0;
}

int v2d_ctx_export_coo(v2d_ctx *ctx, int *coo_i, int *coo_j, double *coo_v, double *bvec)
{
    Ctx *c = ctx->c.get();
    const int64_t N = c->N, nnz = c->nnz;
    DBuf dci, dcj, dcv;
    V2D_ALLOC_S(dci, sizeof(int) * nnz, c->s);
    V2D_ALLOC_S(dcj, sizeof(int) * nnz, c->s);
    V2D_ALLOC_S(dcv, sizeof(double) * nnz, c->s);
    if (c->prec == V2D_FP64)
        hipLaunchKernelGGL(k_slots_to_coo<double>, dim3(grid_for(2 * N)), dim3(256), 0, c->s, c->imax,
                           c->jmax, c->flags, c->pre_psi, c->pre_ome, c->npsi_out,
                           static_cast<const double *>(c->val), dci.as<int>(), dcj.as<int>(), dcv.as<double>());
    else
        hipLaunchKernelGGL(k_slots_to_coo<float>, dim3(grid_for(2 * N)), dim3(256), 0, c->s, c->imax,
                           c->jmax, c->flags, c->pre_psi, c->pre_ome, c->npsi_out,
                           static_cast<const float *>(c->val), dci.as<int>(), dcj.as<int>(), dcv.as<double>());
    V2D_LAUNCH();
    V2D_HIP(hipStreamSynchronize(c->s));
    V2D_HIP(hipMemcpyAsync(coo_i, dci.p, sizeof(int) * nnz, hipMemcpyDeviceToHost, c->s));
    V2D_HIP(hipMemcpyAsync(coo_j, dcj.p, sizeof(int) * nnz, hipMemcpyDeviceToHost, c->s));
    V2D_HIP(hipMemcpyAsync(coo_v, dcv.p, sizeof(double) * nnz, hipMemcpyDeviceToHost, c->s));
    if (bvec) V2D_HIP(hipMemcpyAsync(bvec, c->rhs, sizeof(double) * 2 * N, hipMemcpyDeviceToHost, c->s));
    V2D_HIP(hipStreamSynchronize(c->s));
    return 0;
}

int v2d_ctx_export_csr(v2d_ctx *ctx, int64_t *indptr, int *indices, double *data)
{
    Ctx *c = ctx->c.get();
    const int64_t N = c->N, nnz = c->nnz;
    DBuf dp, di, dd;
    V2D_ALLOC_S(dp, sizeof(int64_t) * (2 * N + 1), c->s);
    V2D_ALLOC_S(di, sizeof(int) * nnz, c->s);
    V2D_ALLOC_S(dd, sizeof(double) * nnz, c->s);
    if (c->prec == V2D_FP64)
        hipLaunchKernelGGL(k_slots_to_csr<double>, dim3(grid_for(2 * N + 1)), dim3(256), 0, c->s, c->imax,
                           c->jmax, c->flags, c->pre_psi, c->pre_ome, c->npsi_out,
                           static_cast<const double *>(c->val), dp.as<int64_t>(), di.as<int>(), dd.as<double>());
    else
        hipLaunchKernelGGL(k_slots_to_csr<float>, dim3(grid_for(2 * N + 1)), dim3(256), 0, c->s, c->imax,
                           c->jmax, c->flags, c->pre_psi, c->pre_ome, c->npsi_out,
                           static_cast<const float *>(c->val), dp.as<int64_t>(), di.as<int>(), dd.as<double>());
    V2D_LAUNCH();
    V2D_HIP(hipStreamSynchronize(c->s));
    V2D_HIP(hipMemcpyAsync(indptr, dp.p, sizeof(int64_t) * (2 * N + 1), hipMemcpyDeviceToHost, c->s));
    V2D_HIP(hipMemcpyAsync(indices, di.p, sizeof(int) * nnz, hipMemcpyDeviceToHost, c->s));
    V2D_HIP(hipMemcpyAsync(data, dd.p, sizeof(double) * nnz, hipMemcpyDeviceToHost, c->s));
    V2D_HIP(hipStreamSynchronize(c->s));
    return 0;
}

int v2d_ctx_spmv(v2d_ctx *ctx, const double *x, double *y)
{
    Ctx *c = ctx->c.get();
    const int64_t n = 2 * c->N;
    DBuf dx, dy;
    V2D_ALLOC_S(dx, sizeof(double) * n, c->s);
    V2D_ALLOC_S(dy, sizeof(double) * n, c->s);
    V2D_HIP(hipStreamSynchronize(c->s));
    V2D_HIP(hipMemcpyAsync(dx.p, x, sizeof(double) * n, hipMemcpyHostToDevice, c->s));
    // FP64 vectors through the FP64-widening residual kernel: y = 0 - A x
    DBuf dz;
    V2D_ALLOC_S(dz, sizeof(double) * n, c->s);
    V2D_HIP(hipMemsetAsync(dz.p, 0, sizeof(double) * n, c->s));
    if (c->prec == V2D_FP64)
        hipLaunchKernelGGL(k_residual<double>, dim3(grid_for(c->N)), dim3(256), 0, c->s, c->imax, c->jmax,
                           c->flags, static_cast<const double *>(c->val), dx.as<double>(), dz.as<double>(),
                           dy.as<double>());
    else
        hipLaunchKernelGGL(k_residual<float>, dim3(grid_for(c->N)), dim3(256), 0, c->s, c->imax, c->jmax,
                           c->flags, static_cast<const float *>(c->val), dx.as<double>(), dz.as<double>(),
                           dy.as<double>());
    V2D_LAUNCH();
    V2D_HIP(hipStreamSynchronize(c->s));
    std::vector<double> t(n);
    V2D_HIP(hipMemcpyAsync(t.data(), dy.p, sizeof(double) * n, hipMemcpyDeviceToHost, c->s));
    V2D_HIP(hipStreamSynchronize(c->s));
    for (int64_t q = 0; q < n; ++q) y[q] = -t[q];
    return 0;
}

}  // extern "C"

namespace v2d {
int ctx_factor(Ctx *c)
{
    int rc;
    V2D_HIP(hipEventRecord(c->ev[4], c->s));
    // The block LU is a self-contained operator: its solve sweeps use the
    // inter-level couplings of the same Jacobian as its dense Schur inverses,
    // so a stored factorisation stays consistent after later assemblies.
    const size_t tsz = c->prec == V2D_FP64 ? sizeof(double) : sizeof(float);
    V2D_HIP(hipMemcpyAsync(c->val_pc, c->val, (size_t)kSlots * c->N * tsz, hipMemcpyDeviceToDevice, c->s));
    if (c->solver == V2D_SOLVER_MULTIFRONTAL) {
        if (ensure_mf(c)) return -1;
        rc = c->prec == V2D_FP64
                 ? mf_factor<double>(*c->mfp, c->md, static_cast<const double *>(c->val_pc), c->s)
                 : mf_factor<float>(*c->mfp, c->mw, static_cast<const float *>(c->val_pc), c->s);
    } else {
        if (ensure_block_lu(c)) return -1;
        rc = c->prec == V2D_FP64
                 ? direct_factor<double>(c->pat, c->wd, static_cast<const double *>(c->val_pc), c->s)
                 : direct_factor<float>(c->pat, c->wf, static_cast<const float *>(c->val_pc), c->s);
    }
    if (rc) return rc;
    V2D_HIP(hipEventRecord(c->ev[5], c->s));
    c->fac_in_step = true;
    c->factored = true;
    c->since_factor = 0;
    ++c->n_fac;
    return 0;
}
}  // namespace v2d

extern "C" {

int v2d_ctx_factor(v2d_ctx *ctx)
{
    Ctx *c = ctx->c.get();
    if (ctx_factor(c)) return -1;
    V2D_HIP(hipStreamSynchronize(c->s));
    return 0;
}

int v2d_ctx_solve(v2d_ctx *ctx, const double *b, double *x)
{
    Ctx *c = ctx->c.get();
    if (!c->factored) {
        set_error("solve: no factorisation (call v2d_ctx_factor first)");
        return -1;
    }
    const int64_t n = 2 * c->N;
    V2D_HIP(hipStreamSynchronize(c->s));
    if (c->prec == V2D_FP64) {
        DBuf db, dx;
        V2D_ALLOC_S(db, sizeof(double) * n, c->s);
        V2D_ALLOC_S(dx, sizeof(double) * n, c->s);
        V2D_HIP(hipMemcpyAsync(db.p, b, sizeof(double) * n, hipMemcpyHostToDevice, c->s));
        if (lin_solve<double>(c, db.as<double>(), dx.as<double>(), false)) return -1;
        V2D_HIP(hipStreamSynchronize(c->s));
        if (sweep_failed(c->wd.err_host)) return -1;
        V2D_HIP(hipMemcpyAsync(x, dx.p, sizeof(double) * n, hipMemcpyDeviceToHost, c->s));
        V2D_HIP(hipStreamSynchronize(c->s));
        return 0;
    }
    std::vector<float> bf(n), xf(n);
    for (int64_t q = 0; q < n; ++q) bf[q] = (float)b[q];
    DBuf db, dx;
    V2D_ALLOC_S(db, sizeof(float) * n, c->s);
    V2D_ALLOC_S(dx, sizeof(float) * n, c->s);
    V2D_HIP(hipMemcpyAsync(db.p, bf.data(), sizeof(float) * n, hipMemcpyHostToDevice, c->s));
    if (lin_solve<float>(c, db.as<float>(), dx.as<float>(), false)) return -1;
    V2D_HIP(hipStreamSynchronize(c->s));
    if (sweep_failed(c->wf.err_host)) return -1;
    V2D_HIP(hipMemcpyAsync(xf.data(), dx.p, sizeof(float) * n, hipMemcpyDeviceToHost, c->s));
    V2D_HIP(hipStreamSynchronize(c->s));
    for (int64_t q = 0; q < n; ++q) x[q] = xf[q];
    return 0;
}

}  // extern "C"

namespace v2d {
// GMRES tolerance: gmres_rtol, or the Eisenstat-Walker forcing term
// eta_k = 0.9 (||R_k|| / ||R_k-1||)^2 (safeguarded, capped at eta_max)
double forcing_rtol(Ctx *c, double bn)
{
    double rtol = c->gm_rtol;
    if (c->eta_max > 0.0) {
        double eta = c->eta_max;
        if (c->rn_prev > 0.0) {
            const double r = bn / c->rn_prev;
            eta = 0.9 * r * r;
            const double sg = 0.9 * c->eta_prev * c->eta_prev;
            if (sg > 0.1) eta = std::max(eta, sg);
        }
        eta = std::min(eta, c->eta_max);
        // no more accuracy than the residual target needs: eta_k >=
        // 0.5 r_target / ||R_k|| (Kelley, Iterative Methods for Linear
        // and Nonlinear Equations, 1995, section 6.3.2)
        if (c->r_target > 0.0) eta = std::min(c->eta_max, std::max(eta, 0.5 * c->r_target / bn));
        rtol = std::max(eta, c->gm_rtol);
        c->eta_prev = eta;
    }
    c->rn_prev = bn;
    return rtol;
}
}  // namespace v2d

extern "C" {

int v2d_ctx_newton_step(v2d_ctx *ctx, double re, double invdtau, double inv2dt, double urot,
                        double *res, double *rnorm, int *lin_iters)
{
    Ctx *c = ctx->c.get();
    const int64_t N = c->N, n = 2 * N;
    c->fac_in_step = false;
    V2D_HIP(hipEventRecord(c->ev[0], c->s));
    if (ctx_assemble(c, re, invdtau, inv2dt, urot)) return -1;
    V2D_HIP(hipEventRecord(c->ev[1], c->s));
    double bn = 0.0;
    if (dnorm2(c, c->rhs, n, &bn)) return -1;
    if (!std::isfinite(bn)) {
        set_error("Newton step: non-finite residual norm (degenerate mesh or diverged state)");
        return -1;
    }
    *rnorm = bn;
    int its = 0;
    float fac_ms = 0.f, sol_ms = 0.f;
    if (c->prec == V2D_FP64) {
        if (ctx_factor(c)) return -1;
        V2D_HIP(hipEventRecord(c->ev[2], c->s));
        if (lin_solve<double>(c, c->rhs, c->delta, false)) return -1;
    } else {
        // Factorisation policy: refactor_every N > 0 refactors every N steps;
        // N = 0 reuses the last FP32 block LU while GMRES reaches its tolerance
        // within gmres_max_iter iterations, and refactors otherwise.
        const bool need = !c->factored || (c->refactor_every > 0 && c->since_factor >= c->refactor_every);
        if (need && ctx_factor(c)) return -1;
        V2D_HIP(hipEventRecord(c->ev[2], c->s));
        if (bn > 0.0) {
            // unit-norm FP32 right-hand side
            V2D_HIP(hipMemcpyAsync(c->red, &bn, sizeof(double), hipMemcpyHostToDevice, c->s));
            hipLaunchKernelGGL((k_scale_copy<double, float>), dim3(grid_for(n)), dim3(256), 0, c->s, n,
                               c->rhs, c->red, 1, c->gb);
            V2D_LAUNCH();
            double rr = 1.0;
            c->rtol_step = forcing_rtol(c, bn);
            // Iterations on the stored factorisation are worth spending while
            // they cost less than a refactorisation plus the iterations a
            // fresh one needs (both measured on this context).
            int budget = c->gm_max;
            if (!need && c->refactor_policy == V2D_REFACTOR_MEASURED && c->fac_ms_last > 0.0 &&
                c->it_ms_avg > 0.0)
                budget = std::max(3, std::min(c->gm_max, (int)(c->fac_ms_last / c->it_ms_avg) + c->fresh_its));
            int st = gmres_f32(c, need ? c->gm_cap : c->gm_max, need ? 0 : budget, &its, &rr);
            if (st < 0) return -1;
            if (need) c->fresh_its = its;
            if (st != 0 || rr > c->rtol_step) {
                if (!need) {   // stale factorisation: refactor and solve again
                    if (ctx_factor(c)) return -1;
                    int its2 = 0;
                    st = gmres_f32(c, c->gm_cap, 0, &its2, &rr);
                    if (st < 0) return -1;
                    c->fresh_its = its2;
                    its += its2;
                }
                if (st != 0) {
                    set_error("mixed-precision GMRES failed on a fresh FP32 factorisation "
                              "(non-finite Krylov vectors)");
                    return -1;
                }
            }
            V2D_HIP(hipMemcpyAsync(c->red, &bn, sizeof(double), hipMemcpyHostToDevice, c->s));
            hipLaunchKernelGGL((k_scale_copy<double, double>), dim3(grid_for(n)), dim3(256), 0, c->s, n,
                               c->gxd, c->red, 0, c->delta);
            V2D_LAUNCH();
        } else {
            V2D_HIP(hipMemsetAsync(c->delta, 0, sizeof(double) * n, c->s));
        }
        ++c->since_factor;
    }
    V2D_HIP(hipEventRecord(c->ev[3], c->s));
    hipLaunchKernelGGL(k_apply_update, dim3(grid_for(N)), dim3(256), 0, c->s, N, c->delta, c->psi, c->ome);
    V2D_LAUNCH();
    double dn = 0.0;
    if (dnorm2(c, c->delta, n, &dn)) return -1;
    if (sweep_failed(c->prec == V2D_FP64 ? c->wd.err_host : c->wf.err_host)) return -1;
    float a_ms = 0.f, t_ms = 0.f;
    V2D_HIP(hipEventElapsedTime(&a_ms, c->ev[0], c->ev[1]));
    if (c->fac_in_step) V2D_HIP(hipEventElapsedTime(&fac_ms, c->ev[4], c->ev[5]));
    V2D_HIP(hipEventElapsedTime(&sol_ms, c->ev[1], c->ev[3]));
    sol_ms -= fac_ms;   // the factorisation may come before or after a GMRES attempt
    V2D_HIP(hipEventElapsedTime(&t_ms, c->ev[0], c->ev[3]));
    c->t_asm += a_ms; c->t_fac += fac_ms; c->t_sol += sol_ms; c->t_tot += t_ms;
    if (c->fac_in_step) c->fac_ms_last = fac_ms;
    if (its > 0) {
        const double per = sol_ms / its;
        c->it_ms_avg = c->it_ms_avg > 0.0 ? 0.5 * (c->it_ms_avg + per) : per;
    }
    ++c->n_steps;
    c->n_lin += its;
    *res = dn / (2.0 * (double)N);
    *lin_iters = its;
    if (!std::isfinite(*res)) {
        set_error("Newton step: non-finite update");
        return -1;
    }
    return 0;
}

int v2d_ctx_bdf2_shift(v2d_ctx *ctx)
{
    Ctx *c = ctx->c.get();
    V2D_HIP(hipMemcpyAsync(c->omeoldold, c->omeold, sizeof(double) * c->N, hipMemcpyDeviceToDevice, c->s));
    V2D_HIP(hipMemcpyAsync(c->omeold, c->ome, sizeof(double) * c->N, hipMemcpyDeviceToDevice, c->s));
    V2D_HIP(hipStreamSynchronize(c->s));
    return 0;
}

int v2d_ctx_forces(v2d_ctx *ctx, double re, double ref_length, double x0, double y0, double *out)
{
    Ctx *c = ctx->c.get();
    hipLaunchKernelGGL(k_forces, dim3(1), dim3(1024), 0, c->s, c->imax, re, c->dksi, c->deta, c->m, c->N,
                       c->ome, ref_length, x0, y0, c->hbuf);
    V2D_LAUNCH();
    V2D_HIP(hipMemcpyAsync(out, c->hbuf, sizeof(double) * 10, hipMemcpyDeviceToHost, c->s));
    V2D_HIP(hipStreamSynchronize(c->s));
    return 0;
}

int v2d_ctx_timers(v2d_ctx *ctx, double *ms4, int64_t *counts3, int reset)
{
    Ctx *c = ctx->c.get();
    ms4[0] = c->t_asm; ms4[1] = c->t_fac; ms4[2] = c->t_sol; ms4[3] = c->t_tot;
    counts3[0] = c->n_steps; counts3[1] = c->n_fac; counts3[2] = c->n_lin;
    if (reset) {
        c->t_asm = c->t_fac = c->t_sol = c->t_tot = 0;
        c->n_steps = c->n_fac = c->n_lin = 0;
    }
    return 0;
}

int v2d_ctx_time_assembly(v2d_ctx *ctx, double re, double invdtau, double inv2dt, double urot,
                          int launches, double *ms_per_launch)
{
    Ctx *c = ctx->c.get();
    if (launches < 1) { set_error("launches must be >= 1"); return -1; }
    if (ctx_assemble(c, re, invdtau, inv2dt, urot)) return -1;   // warm-up
    V2D_HIP(hipEventRecord(c->ev[0], c->s));
    for (int q = 0; q < launches; ++q)
        if (ctx_assemble(c, re, invdtau, inv2dt, urot)) return -1;
    V2D_HIP(hipEventRecord(c->ev[1], c->s));
    V2D_HIP(hipEventSynchronize(c->ev[1]));
    float ms = 0.f;
    V2D_HIP(hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
    *ms_per_launch = ms / launches;
    return 0;
}

int v2d_ctx_time_spmv(v2d_ctx *ctx, int launches, double *ms_per_launch)
{
    Ctx *c = ctx->c.get();
    if (launches < 1) { set_error("launches must be >= 1"); return -1; }
    const int64_t n = 2 * c->N;
    DBuf dx, dy;
    const size_t vs = c->prec == V2D_FP64 ? sizeof(double) : sizeof(float);
    V2D_ALLOC_S(dx, vs * n, c->s);
    V2D_ALLOC_S(dy, vs * n, c->s);
    auto launch = [&]() {
        if (c->prec == V2D_FP64)
            hipLaunchKernelGGL((k_spmv<double, double>), dim3(grid_for(c->N)), dim3(256), 0, c->s, c->imax,
                               c->jmax, c->flags, static_cast<const double *>(c->val), dx.as<double>(),
                               dy.as<double>());
        else
            hipLaunchKernelGGL((k_spmv<float, float>), dim3(grid_for(c->N)), dim3(256), 0, c->s, c->imax,
                               c->jmax, c->flags, static_cast<const float *>(c->val), dx.as<float>(),
                               dy.as<float>());
    };
    if (c->prec == V2D_FP64)
        hipLaunchKernelGGL((k_fill_one<double>), dim3(grid_for(n)), dim3(256), 0, c->s, n, dx.as<double>());
    else
        hipLaunchKernelGGL((k_fill_one<float>), dim3(grid_for(n)), dim3(256), 0, c->s, n, dx.as<float>());
    launch();
    V2D_LAUNCH();
    V2D_HIP(hipEventRecord(c->ev[0], c->s));
    for (int q = 0; q < launches; ++q) launch();
    V2D_HIP(hipEventRecord(c->ev[1], c->s));
    V2D_LAUNCH();
    V2D_HIP(hipEventSynchronize(c->ev[1]));
    float ms = 0.f;
    V2D_HIP(hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
    *ms_per_launch = ms / launches;
    return 0;
}

// ------------------------------------------------------------ adjoint ----
int v2d_ctx_solve_transpose(v2d_ctx *ctx, const double *b, double *x)
{
    Ctx *c = ctx->c.get();
    if (!c->factored) {
        set_error("solve_transpose: no factorisation (call v2d_ctx_factor first)");
        return -1;
    }
    const int64_t n = 2 * c->N;
    V2D_HIP(hipStreamSynchronize(c->s));
    if (c->prec == V2D_FP64) {
        DBuf db, dx;
        V2D_ALLOC_S(db, sizeof(double) * n, c->s);
        V2D_ALLOC_S(dx, sizeof(double) * n, c->s);
        V2D_HIP(hipMemcpyAsync(db.p, b, sizeof(double) * n, hipMemcpyHostToDevice, c->s));
        if (lin_solve<double>(c, db.as<double>(), dx.as<double>(), true)) return -1;
        V2D_HIP(hipStreamSynchronize(c->s));
        V2D_HIP(hipMemcpyAsync(x, dx.p, sizeof(double) * n, hipMemcpyDeviceToHost, c->s));
        V2D_HIP(hipStreamSynchronize(c->s));
        return 0;
    }
    std::vector<float> bf(n), xf(n);
    for (int64_t q = 0; q < n; ++q) bf[q] = (float)b[q];
    DBuf db, dx;
    V2D_ALLOC_S(db, sizeof(float) * n, c->s);
    V2D_ALLOC_S(dx, sizeof(float) * n, c->s);
    V2D_HIP(hipMemcpyAsync(db.p, bf.data(), sizeof(float) * n, hipMemcpyHostToDevice, c->s));
    if (lin_solve<float>(c, db.as<float>(), dx.as<float>(), true)) return -1;
    V2D_HIP(hipStreamSynchronize(c->s));
    V2D_HIP(hipMemcpyAsync(xf.data(), dx.p, sizeof(float) * n, hipMemcpyDeviceToHost, c->s));
    V2D_HIP(hipStreamSynchronize(c->s));
    for (int64_t q = 0; q < n; ++q) x[q] = xf[q];
    return 0;
}

// lambda^T db/d(metrics, Re, alpha) on device buffers: g_m (10N) is added to.
static int residual_vjp_dev(Ctx *c, double re, double inv2dt, double urot, const double *d_lam,
                            double *d_gm, double *gscal)
{
    const AsmArgs a = c->args(re, 0.0, inv2dt, urot);
    const unsigned nb = grid_for(c->N);
    DBuf part, red;
    V2D_ALLOC_S(part, sizeof(double) * 2 * nb, c->s);
    V2D_ALLOC_S(red, sizeof(double) * 2, c->s);
    hipLaunchKernelGGL(k_residual_vjp, dim3(nb), dim3(256), 0, c->s, a, c->flags, d_lam, d_gm, part.as<double>());
    hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(256), 0, c->s, (int)nb, 2, part.as<double>(),
                       red.as<double>());
    V2D_LAUNCH();
    V2D_HIP(hipMemcpyAsync(gscal, red.p, sizeof(double) * 2, hipMemcpyDeviceToHost, c->s));
    V2D_HIP(hipStreamSynchronize(c->s));
    return 0;
}

int v2d_ctx_residual_vjp(v2d_ctx *ctx, double re, double inv2dt, double urot, const double *lam,
                         double *g_m, double *gscal)
{
    Ctx *c = ctx->c.get();
    const int64_t N = c->N;
    DBuf dl, dg;
    V2D_ALLOC_S(dl, sizeof(double) * 2 * N, c->s);
    V2D_ALLOC_S(dg, sizeof(double) * 10 * N, c->s);
    V2D_HIP(hipMemcpyAsync(dl.p, lam, sizeof(double) * 2 * N, hipMemcpyHostToDevice, c->s));
    hipLaunchKernelGGL(k_fill_zero_d, dim3(grid_for(10 * N)), dim3(256), 0, c->s, 10 * N, dg.as<double>());
    V2D_LAUNCH();
    if (residual_vjp_dev(c, re, inv2dt, urot, dl.as<double>(), dg.as<double>(), gscal)) return -1;
    V2D_HIP(hipMemcpyAsync(g_m, dg.p, sizeof(double) * 10 * N, hipMemcpyDeviceToHost, c->s));
    V2D_HIP(hipStreamSynchronize(c->s));
    return 0;
}

static int metrics_vjp_dev(int imax, int jmax, double dksi, double deta, const double *dx, const double *dy,
                           const double *d_gm, double *d_gx, double *d_gy, hipStream_t s)
{
    const int64_t N = (int64_t)imax * jmax;
    DBuf geta;
    V2D_ALLOC(geta, sizeof(double) * 2 * N);
    hipLaunchKernelGGL(k_fill_zero_d, dim3(grid_for(2 * N)), dim3(256), 0, s, 2 * N, geta.as<double>());
    hipLaunchKernelGGL(k_fill_zero_d, dim3(grid_for(N)), dim3(256), 0, s, N, d_gx);
    hipLaunchKernelGGL(k_fill_zero_d, dim3(grid_for(N)), dim3(256), 0, s, N, d_gy);
    hipLaunchKernelGGL(k_metrics_vjp_1, dim3(grid_for(N)), dim3(256), 0, s, imax, jmax, dksi, deta, dx, dy,
                       d_gm, d_gx, d_gy, geta.as<double>());
    hipLaunchKernelGGL(k_metrics_vjp_2, dim3(grid_for(N)), dim3(256), 0, s, imax, jmax, deta,
                       geta.as<double>(), d_gx, d_gy);
    V2D_LAUNCH();
    V2D_HIP(hipStreamSynchronize(s));
    return 0;
}

int v2d_metrics_vjp(int imax, int jmax, double dksi, double deta, const double *xg, const double *yg,
                    const double *g_m, double *g_x, double *g_y)
{
    if (check_grid(imax, jmax)) return -1;
    const int64_t N = (int64_t)imax * jmax;
    DBuf dx, dy, dg, gx, gy;
    V2D_ALLOC(dx, sizeof(double) * N);
    V2D_ALLOC(dy, sizeof(double) * N);
    V2D_ALLOC(dg, sizeof(double) * 10 * N);
    V2D_ALLOC(gx, sizeof(double) * N);
    V2D_ALLOC(gy, sizeof(double) * N);
    V2D_HIP(hipMemcpy(dx.p, xg, sizeof(double) * N, hipMemcpyHostToDevice));
    V2D_HIP(hipMemcpy(dy.p, yg, sizeof(double) * N, hipMemcpyHostToDevice));
    V2D_HIP(hipMemcpy(dg.p, g_m, sizeof(double) * 10 * N, hipMemcpyHostToDevice));
    if (metrics_vjp_dev(imax, jmax, dksi, deta, dx.as<double>(), dy.as<double>(), dg.as<double>(),
                        gx.as<double>(), gy.as<double>(), nullptr))
        return -1;
    V2D_HIP(hipMemcpy(g_x, gx.p, sizeof(double) * N, hipMemcpyDeviceToHost));
    V2D_HIP(hipMemcpy(g_y, gy.p, sizeof(double) * N, hipMemcpyDeviceToHost));
    return 0;
}

static int forces_vjp_dev(Ctx *c, double re, double d, double wd, double wl, double *d_gome, double *d_gm,
                          double *out3)
{
    const int64_t N = c->N;
    if (!(d > 0.0)) {
        set_error("force adjoint: ref_length must be given (> 0)");
        return -1;
    }
    hipLaunchKernelGGL(k_forces_vjp, dim3(1), dim3(1024), 0, c->s, c->imax, re, c->dksi, c->deta, c->m, N,
                       c->ome, d, wd, wl, d_gome, d_gm, c->hbuf);
    V2D_LAUNCH();
    V2D_HIP(hipMemcpyAsync(out3, c->hbuf, sizeof(double) * 3, hipMemcpyDeviceToHost, c->s));
    V2D_HIP(hipStreamSynchronize(c->s));
    return 0;
}

int v2d_ctx_forces_vjp(v2d_ctx *ctx, double re, double ref_length, double wd, double wl, double *g_ome,
                       double *g_m, double *out3)
{
    Ctx *c = ctx->c.get();
    const int64_t N = c->N;
    DBuf go, gm;
    V2D_ALLOC_S(go, sizeof(double) * N, c->s);
    V2D_ALLOC_S(gm, sizeof(double) * 10 * N, c->s);
    hipLaunchKernelGGL(k_fill_zero_d, dim3(grid_for(N)), dim3(256), 0, c->s, N, go.as<double>());
    hipLaunchKernelGGL(k_fill_zero_d, dim3(grid_for(10 * N)), dim3(256), 0, c->s, 10 * N, gm.as<double>());
    V2D_LAUNCH();
    if (forces_vjp_dev(c, re, ref_length, wd, wl, go.as<double>(), gm.as<double>(), out3)) return -1;
    V2D_HIP(hipMemcpyAsync(g_ome, go.p, sizeof(double) * N, hipMemcpyDeviceToHost, c->s));
    V2D_HIP(hipMemcpyAsync(g_m, gm.p, sizeof(double) * 10 * N, hipMemcpyDeviceToHost, c->s));
    V2D_HIP(hipStreamSynchronize(c->s));
    return 0;
}

int v2d_ctx_adjoint_bdf2_begin(v2d_ctx *ctx)
{
    Ctx *c = ctx->c.get();
    if (c->prec != V2D_FP64) {
        set_error("unsteady adjoint: needs an FP64 context (exact transposed solves)");
        return -1;
    }
    const int64_t N = c->N;
    if (!c->adj_l1) {
        int rc = 0;
        c->adj_l1 = c->alloc<double>(2 * N, rc);
        c->adj_l2 = c->alloc<double>(2 * N, rc);
        c->adj_lam = c->alloc<double>(2 * N, rc);
        c->adj_rhs = c->alloc<double>(2 * N, rc);
        c->adj_gm = c->alloc<double>(10 * N, rc);
        if (rc) return -1;
    }
    hipLaunchKernelGGL(k_fill_zero_d, dim3(grid_for(2 * N)), dim3(256), 0, c->s, 2 * N, c->adj_l1);
    hipLaunchKernelGGL(k_fill_zero_d, dim3(grid_for(2 * N)), dim3(256), 0, c->s, 2 * N, c->adj_l2);
    hipLaunchKernelGGL(k_fill_zero_d, dim3(grid_for(10 * N)), dim3(256), 0, c->s, 10 * N, c->adj_gm);
    V2D_LAUNCH();
    c->adj_gs[0] = c->adj_gs[1] = 0.0;
    c->adj_levels = 0;
    return 0;
}

int v2d_ctx_adjoint_bdf2_step(v2d_ctx *ctx, double re, double inv2dt, double urot, double ref_length,
                              double wd, double wl, double *coef)
{
    Ctx *c = ctx->c.get();
    if (c->adj_levels < 0) {
        set_error("unsteady adjoint: call v2d_ctx_adjoint_bdf2_begin first");
        return -1;
    }
    const int64_t N = c->N;
    // Newton matrix of this level's converged residual (no pseudo-time shift)
    if (ctx_assemble(c, re, 0.0, inv2dt, urot)) return -1;
    if (ctx_factor(c)) return -1;
    hipLaunchKernelGGL(k_fill_zero_d, dim3(grid_for(2 * N)), dim3(256), 0, c->s, 2 * N, c->adj_rhs);
    V2D_LAUNCH();
    double f3[3] = {0.0, 0.0, 0.0};
    if (wd != 0.0 || wl != 0.0 || coef) {
        // dJ_n/domega in the omega half; explicit dJ_n/dm and dJ_n/dRe
        if (forces_vjp_dev(c, re, ref_length, wd, wl, c->adj_rhs + N, c->adj_gm, f3)) return -1;
        c->adj_gs[0] += f3[2];
    }
    if (c->adj_levels > 0 && inv2dt != 0.0) {
        hipLaunchKernelGGL(k_bdf2_adjoint_rhs, dim3(grid_for(N)), dim3(256), 0, c->s, c->imax, c->jmax, inv2dt,
                           c->adj_l1, c->adj_l2, c->adj_rhs);
        V2D_LAUNCH();
    }
    if (lin_solve<double>(c, c->adj_rhs, c->adj_lam, true)) return -1;
    double rs[2];
    if (residual_vjp_dev(c, re, inv2dt, urot, c->adj_lam, c->adj_gm, rs)) return -1;
    c->adj_gs[0] += rs[0];
    c->adj_gs[1] += rs[1];
    std::swap(c->adj_l2, c->adj_l1);    // lambda_{n+2} <- lambda_{n+1}
    std::swap(c->adj_l1, c->adj_lam);   // lambda_{n+1} <- lambda_n
    ++c->adj_levels;
    if (coef) {
        coef[0] = f3[0];
        coef[1] = f3[1];
    }
    return 0;
}

int v2d_ctx_adjoint_bdf2_end(v2d_ctx *ctx, double *g_x, double *g_y, double *gscal)
{
    Ctx *c = ctx->c.get();
    if (c->adj_levels < 0) {
        set_error("unsteady adjoint: no sweep begun");
        return -1;
    }
    const int64_t N = c->N;
    DBuf gx, gy;
    V2D_ALLOC_S(gx, sizeof(double) * N, c->s);
    V2D_ALLOC_S(gy, sizeof(double) * N, c->s);
    if (metrics_vjp_dev(c->imax, c->jmax, c->dksi, c->deta, c->x, c->y, c->adj_gm, gx.as<double>(),
                        gy.as<double>(), c->s))
        return -1;
    V2D_HIP(hipMemcpyAsync(g_x, gx.p, sizeof(double) * N, hipMemcpyDeviceToHost, c->s));
    V2D_HIP(hipMemcpyAsync(g_y, gy.p, sizeof(double) * N, hipMemcpyDeviceToHost, c->s));
    V2D_HIP(hipStreamSynchronize(c->s));
    gscal[0] = c->adj_gs[1] * (M_PI / 180.0);   // per degree of alpha
    gscal[1] = c->adj_gs[0];                    // per unit Re
    c->adj_levels = -1;
    return 0;
}

int v2d_ctx_adjoint_forces(v2d_ctx *ctx, double re, double urot, double ref_length, double wd, double wl,
                           double *g_x, double *g_y, double *gscal, double *coef)
{
    Ctx *c = ctx->c.get();
    if (c->prec != V2D_FP64) {
        set_error("force adjoint: needs an FP64 context (exact transposed solve)");
        return -1;
    }
    const int64_t N = c->N;
    // Jacobian of the steady residual at the stored state (no pseudo-time shift)
    if (ctx_assemble(c, re, 0.0, 0.0, urot)) return -1;
    if (ctx_factor(c)) return -1;
    DBuf gu, lam, gm, gx, gy;
    V2D_ALLOC_S(gu, sizeof(double) * 2 * N, c->s);
    V2D_ALLOC_S(lam, sizeof(double) * 2 * N, c->s);
    V2D_ALLOC_S(gm, sizeof(double) * 10 * N, c->s);
    V2D_ALLOC_S(gx, sizeof(double) * N, c->s);
    V2D_ALLOC_S(gy, sizeof(double) * N, c->s);
    hipLaunchKernelGGL(k_fill_zero_d, dim3(grid_for(2 * N)), dim3(256), 0, c->s, 2 * N, gu.as<double>());
    hipLaunchKernelGGL(k_fill_zero_d, dim3(grid_for(10 * N)), dim3(256), 0, c->s, 10 * N, gm.as<double>());
    V2D_LAUNCH();
    double f3[3];
    // dJ/dU lives in the omega half of the field-blocked vector
    if (forces_vjp_dev(c, re, ref_length, wd, wl, gu.as<double>() + N, gm.as<double>(), f3)) return -1;
    if (lin_solve<double>(c, gu.as<double>(), lam.as<double>(), true)) return -1;
    double rs[2];
    if (residual_vjp_dev(c, re, 0.0, urot, lam.as<double>(), gm.as<double>(), rs)) return -1;
    if (metrics_vjp_dev(c->imax, c->jmax, c->dksi, c->deta, c->x, c->y, gm.as<double>(), gx.as<double>(),
                        gy.as<double>(), c->s))
        return -1;
    V2D_HIP(hipMemcpyAsync(g_x, gx.p, sizeof(double) * N, hipMemcpyDeviceToHost, c->s));
    V2D_HIP(hipMemcpyAsync(g_y, gy.p, sizeof(double) * N, hipMemcpyDeviceToHost, c->s));
    V2D_HIP(hipStreamSynchronize(c->s));
    gscal[0] = rs[1] * (M_PI / 180.0);   // per degree of alpha
    gscal[1] = f3[2] + rs[0];            // per unit Re
    coef[0] = f3[0];
    coef[1] = f3[1];
    return 0;
}

}  // extern "C"
