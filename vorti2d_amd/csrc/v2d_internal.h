// v2d_internal.h -- declarations shared by the libvorti2d translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "v2d_common.h"

namespace v2d {

// Arguments of the residual / Jacobian builders: scalars of assemble_coo
// (vorti2d_core.f90:219-234) plus device pointers to the 10 metric planes and
// the four state planes.  m[] order: jac, alfa, beta, gama, pmet, qmet,
// detadx, detady, xphys, yphys.
struct AsmArgs {
    int imax, jmax, ff_bc;
    double re, invdtau, inv2dt, urot, dksi, deta, ca, sa;
    const double *m[10];
    const double *psi, *ome, *omeold, *omeoldold;
};

// Reciprocal launch constants of the production assembly.
struct AsmConsts {
    double ire, idk2, ide2, i2dk, i2de, i4dkde, ide;
    double c_al, c_bt, c_p, c_q, c_g;   // 1/(re dk^2), 1/(re de^2), 1/(2 re dk), 1/(2 re de), 2/(4 re dk de)
};
inline AsmConsts asm_consts(double re, double dk, double de)
{
    AsmConsts c;
    c.ire = 1.0 / re;
    c.idk2 = 1.0 / (dk * dk);
    c.ide2 = 1.0 / (de * de);
    c.i2dk = 1.0 / (2.0 * dk);
    c.i2de = 1.0 / (2.0 * de);
    c.i4dkde = 1.0 / (4.0 * de * dk);
    c.ide = 1.0 / de;
    c.c_al = 1.0 / (re * (dk * dk));
    c.c_bt = 1.0 / (re * (de * de));
    c.c_p = 1.0 / (2.0 * re * dk);
    c.c_q = 1.0 / (2.0 * re * de);
    c.c_g = 2.0 / (4.0 * re * dk * de);
    return c;
}

// ------------------------------------------------------------ kernels ------
__global__ void k_metrics(int imax, int jmax, double dksi, double deta, const double *x,
                          const double *y, double *out);
__global__ void k_metrics_local(int gxm, int jmax, int il0, int il1, double dksi, double deta,
                                const double *x, const double *y, double *out);
template <typename T>
__global__ void k_assemble(AsmArgs a, AsmConsts c, const uint8_t *flags, T *val, double *rhs);
__global__ void k_assemble_coo(AsmArgs a, const uint8_t *flags, const int *pre_psi,
                               const int *pre_ome, int npsi_out, int64_t r0, int64_t r1, int *ci,
                               int *cj, double *cv, double *bvec);
__global__ void k_assemble_coo_local(AsmArgs a, int il0, int ow, const int *lg,
                                     const int64_t *node_off, int r0, int nloc, int *ci, int *cj,
                                     double *cv, double *bvec);
template <typename T>
__global__ void k_slots_to_coo(int imax, int jmax, const uint8_t *flags, const int *pre_psi,
                               const int *pre_ome, int npsi_out, const T *val, int *ci, int *cj,
                               double *cv);
template <typename T>
__global__ void k_slots_to_csr(int imax, int jmax, const uint8_t *flags, const int *pre_psi,
                               const int *pre_ome, int npsi_out, const T *val, int64_t *indptr,
                               int *indices, double *data);
template <typename T, typename V>
__global__ void k_spmv(int imax, int jmax, const uint8_t *flags, const T *val, const V *x, V *y);
template <typename V>
__global__ void k_multidot(int64_t n, int m, const V *Vb, int64_t ld, const V *w, double *partial);
__global__ void k_reduce_partials(int nb, int m, const double *partial, double *out);
template <typename V>
__global__ void k_multiaxpy(int64_t n, int m, const V *Vb, int64_t ld, const double *h, V *w);
template <typename A, typename B>
__global__ void k_scale_copy(int64_t n, const A *x, const double *s, int inv, B *y);
template <typename Z, typename X>
__global__ void k_update_sol(int64_t n, int m, const Z *Zb, int64_t ld, const double *y,
                             double scale, X *x);
template <typename T>
__global__ void k_residual(int imax, int jmax, const uint8_t *flags, const T *val, const double *x,
                           const double *b, double *r);
__global__ void k_fill_zero_d(int64_t n, double *p);
template <typename T>
__global__ void k_fill_one(int64_t n, T *p);
__global__ void k_apply_update(int64_t N, const double *d, double *psi, double *ome);
__global__ void k_forces(int imax, double re, double dksi, double deta, const double *m, int64_t N,
                         const double *ome, double ref_length, double x0, double y0, double *out);

// ------------------------------------------------------ batched launches --
// One launch over up to kBatchMax cases of uneven sizes: case c owns the
// blocks [off[c], off[c + 1]) of the grid and runs exactly what its own
// single-case launch would run on them (the same per-node or per-block device
// code, the same block count for the reductions), so a batched result equals
// the single-case result bit for bit.
constexpr int kBatchMax = 16;
struct BMap {
    int nc;
    int off[kBatchMax + 1];
};
__device__ __forceinline__ int bmap_case(const BMap &bm, int b)
{
    int c = 0;
    while (c + 1 < bm.nc && b >= bm.off[c + 1]) ++c;
    return c;
}
// One case's assembly.
struct BAsm {
    AsmArgs a;
    AsmConsts k;
    const uint8_t *flags;
    void *val;
    double *rhs;
};
// One case's operands of a batched vector operation (the fields an
// operation does not use are ignored).
struct BVec {
    int imax, jmax;       // mesh (SpMV)
    int m;                // basis vectors taking part (multi-dot, multi-axpy)
    int nb;               // blocks of the case's single-context reduction
    int inv;              // scale: divide by s (1) or multiply (0)
    int64_t n, ld;        // vector length, basis leading dimension
    const uint8_t *flags;
    const void *val;      // Jacobian slots (SpMV)
    const void *x;        // SpMV / scale / Newton-update input
    void *y;              // SpMV / scale output; multi-axpy in-out vector
    const void *Vb;       // basis (multi-dot, multi-axpy)
    const void *w;        // multi-dot right operand
    double *partial;      // per-block partial sums
    double *out;          // reduced sums
    const double *h;      // multi-axpy coefficients
    const double *s;      // scale factor
    double *psi, *ome;    // Newton update targets
};
template <typename T>
__global__ void k_assemble_bt(BMap bm, const BAsm *cs);
template <typename T, typename V>
__global__ void k_spmv_bt(BMap bm, const BVec *v);
template <typename V>
__global__ void k_multidot_bt(BMap bm, const BVec *v);
__global__ void k_reduce_bt(const BVec *v);
template <typename V>
__global__ void k_multiaxpy_bt(BMap bm, const BVec *v);
template <typename A, typename B>
__global__ void k_scale_copy_bt(BMap bm, const BVec *v);
__global__ void k_apply_update_bt(BMap bm, const BVec *v);

// ------------------------------------------------------- error state -------
void set_error(const std::string &msg);
#define V2D_HIP(call)                                                                       \
    do {                                                                                    \
        hipError_t _e = (call);                                                             \
        if (_e != hipSuccess) {                                                             \
            ::v2d::set_error(std::string(#call) + ": " + hipGetErrorString(_e));            \
            return -1;                                                                      \
        }                                                                                   \
    } while (0)
#define V2D_LAUNCH()                                                                        \
    do {                                                                                    \
        hipError_t _e = hipGetLastError();                                                  \
        if (_e != hipSuccess) {                                                             \
            ::v2d::set_error(std::string("kernel launch: ") + hipGetErrorString(_e));      \
            return -1;                                                                      \
        }                                                                                   \
    } while (0)

inline unsigned grid_for(int64_t n, int b = 256) { return (unsigned)((n + b - 1) / b); }

// ------------------------------------------------- device reductions ----
// Wave64 sum with cross-lane shuffles, then one LDS slot per wavefront.
__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Block-wide sum returned to every thread (sh: one slot per wavefront + 1).
__device__ __forceinline__ double block_sum(double v, double *sh)
{
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) sh[wid] = v;
    __syncthreads();
    if (wid == 0) {
        double t = (threadIdx.x < (blockDim.x >> 6)) ? sh[threadIdx.x] : 0.0;
        t = wave_sum(t);
        if (lane == 0) sh[16] = t;
    }
    __syncthreads();
    const double total = sh[16];
    __syncthreads();
    return total;
}


// ------------------------------------------- block-tridiagonal direct solver
// Exact LU of the Newton matrix in radial-level blocks: level l holds all
// unknowns of its j rows (psi then omega, i fastest).  The factorisation keeps
// the explicit inverse of every Schur complement S_l = D_l - L_l S_{l-1}^-1 U_{l-1}.
struct DirectPattern {
    Levels L;
    int64_t N = 0;
    std::vector<int> n;            // level sizes
    std::vector<int> ld;           // row stride of the dense inverses (n rounded up to 64)
    std::vector<int64_t> voff;     // level offsets in a level-ordered 2N vector
    std::vector<int64_t> soff;     // offsets of the dense inverses (elements)
    int64_t sinv_elems = 0;        // total dense storage
    int nmax = 0, ldmax = 0;
    // device CSR (rows / columns indexed in level order), entries (local, vidx)
    int *d_Dptr = nullptr, *d_Lptr = nullptr, *d_Uptr = nullptr, *d_UCptr = nullptr, *d_LCptr = nullptr;
    int2 *d_D = nullptr, *d_Lr = nullptr, *d_Ur = nullptr, *d_Uc = nullptr, *d_Lc = nullptr;
    // The couplings of every level row into level l-1 (L) and l+1 (U), padded
    // to kCouple entries, stored [kCouple][n_l] at kCouple * voff[l]: the
    // level-ordered source index and the value index (-1: padding).
    int *d_ixL = nullptr, *d_ixU = nullptr, *d_vxL = nullptr, *d_vxU = nullptr;
    // The couplings into every level column q of level l from the rows of level
    // l-1 (side 0, entries of U_{l-1}) and l+1 (side 1, entries of L_{l+1}),
    // padded to cx_w[side][l] (1 to 8) entries, stored [cx_w][n_l] at
    // cx_off[side][l]: the row in the neighbouring level and the value index
    // (-1: padding).  ncx entries in all.
    int *d_jxC = nullptr, *d_vxC = nullptr;
    std::vector<int64_t> cx_off[2];
    std::vector<int> cx_w[2];
    int64_t ncx = 0;
    bool pad_ok = false;     // padded row and column tables built
    int nsteps = 0;          // sweep steps: forward, middle level, backward
    int sweep_C = 1;         // workgroups per elimination chain
    int sweep_grid = 0;      // persistent workgroups: 2 sweep_C, one per compute unit
    int sweep_rows = 1;      // most rows of one workgroup at one level
    bool sweep_ok = false;   // closed-form level offsets verified, tables built
    int ncu = 0;             // compute units of the device (grid of the fused solve steps)
    int64_t nent[5] = {0, 0, 0, 0, 0};   // entries of the D, L, U, UC, LC lists
};

constexpr int kCouple = 4;   // widest inter-level coupling of a row (host-checked)

// Device lists are allocated and freed in stream order on s.
int direct_build_pattern(DirectPattern &P, int imax, int jmax, const uint8_t *flags_host, hipStream_t s);
void direct_free_pattern(DirectPattern &P, hipStream_t s);
// D becomes a copy of the built pattern S, its device lists copied on stream s
// (S is not changed after its build, so only S's lifetime matters: keep it
// until s has run the copies).
int direct_copy_pattern(DirectPattern &D, const DirectPattern &S, hipStream_t s);

// Gauss-Jordan update kernels: VALU with 64- or 128-wide output tiles, or
// MFMA 16x16x4 (FP32 / FP64 matrix cores) on 64-wide tiles.
enum UpdateKind { kUpdateValu64 = 0, kUpdateValu128 = 1, kUpdateMfma64 = 2 };
// Launch modes of the dataflow Gauss-Jordan step over job tables.  Full: every
// tile takes the step's rank-64 update.  Light: only the tiles of the next
// pivot's row and column blocks (they form the next panels); the others wait
// for the heavy step that follows, which applies both steps' panels in one
// rank-128 update.  Begin: no update; the first pivot block is inverted and
// the first panels and pivot rows formed.
enum StepMode { kStepFull = 0, kStepLight = 1, kStepHeavy = 2, kStepBegin = 3 };

// Optional timing of the Gauss-Jordan update kernel on the factorisation
// stream: an event pair brackets each level's run of back-to-back update
// launches (MFMA steps; each launch with the VALU kernels, whose row-panel
// launches sit in between); pairs are read back in groups of kPairs.
// flops / bytes are the algorithmic counts of the timed launches (2 (n-kb) n kb
// flops per job; the updated rows read and written once, both panels read once).
struct LaunchProbe {
    static constexpr int kPairs = 64;
    hipEvent_t ev[2 * kPairs] = {};
    int used = 0;
    int64_t launches = 0;
    double ms = 0.0, flops = 0.0, bytes = 0.0;
    ~LaunchProbe();
};

template <typename T>
struct DirectWork {
    T *sinv = nullptr;   // all dense inverses
    T *tmat = nullptr;   // Schur scratch, 2 x (nmax x ldmax) (one per elimination chain)
    T *colk = nullptr;   // Gauss-Jordan column panels, 2 x (nmax x 64)
    T *rowk = nullptr;   // Gauss-Jordan pivot-row panels, transposed, 2 x (ldmax x 64)
    T *dinv = nullptr;   // pivot-block inverses, 2 x (64 x 64)
    T *z = nullptr;      // level-ordered 2N work vector
    T *colk2 = nullptr, *rowk2 = nullptr;   // the other buffer of each panel (dataflow steps)
    int *flags = nullptr;   // 2 per-job "pivot inverse ready" words (device)
    int *err = nullptr;     // device error word of the dataflow steps
    int epoch = 0;          // value the current step's pivot block publishes
    int update_kind = kUpdateMfma64;
    LaunchProbe *probe = nullptr;   // set to time the update launches
    // solve sweeps: 0 one launch pair per level step, 1 one persistent launch,
    // 2 one fused launch per FP32 level step (launch pairs elsewhere)
    int sweep_kind = 1;
    T *cfL = nullptr, *cfU = nullptr;   // coupling values of the factored Jacobian (padded rows)
    T *cfC = nullptr;                   // the same, padded columns (ncx values)
    T *wv = nullptr;                    // level-ordered backward-sweep results
    T *yv = nullptr;                    // level-ordered right-hand sides of the sweep steps
    int *ctr = nullptr;                 // progress word of every sweep workgroup
    int *err_host = nullptr;            // pinned copy of err, refreshed after every solve
    int sweep_spins = 1 << 21;          // bound of each persistent-sweep wait (polls)
    int64_t sweep_fallbacks = 0;        // persistent solves redone with launch pairs
};

// One dense Gauss-Jordan job: the matrix S (n x n, row stride ld) has its
// pivots [0, np) eliminated in place (np = n: S becomes its inverse; np < n, a
// multiple of 64: a partial sweep, S = [A B; C D] becoming
// [A^-1, A^-1 B; -C A^-1, D - C A^-1 B]).
template <typename T>
struct Job {
    T *S;                  // dense matrix (n x ld, row-major), swept in place
    int n, ld, l;          // size, row stride, level (block LU only)
    int np;                // pivots
    int64_t voff;          // offset of the level in level-ordered vectors (block LU only)
    T *colk, *dinv, *tmat; // per-job scratch (tmat: block LU only)
    T *rowk;               // transposed pivot-row panel (MFMA update only, else null)
    T *colk_w, *rowk_w;    // panels of the next step (MFMA dataflow step only)
    T *colk_x, *rowk_x;    // job tables: third panel buffers (panels rotate over three)
};

// Sweeps of a device table of jobs with the MFMA dataflow steps (colk, colk_w,
// colk_x: n x 64; rowk, rowk_w, rowk_x: 64-rounded n x 64; dinv 64 x 64 per
// job).
template <typename T>
int gj_table_begin(const Job<T> *jobs, int nj, int nmax, int *flags, int *err, int &epoch, hipStream_t s);
// mode: a StepMode; compact (light steps only): no job of the table ends at
// this step, so the grid covers only the next pivot's row and column blocks.
template <typename T>
int gj_table_step(const Job<T> *jobs, int nj, int nmax, int k0, int mode, bool compact, int *flags, int *err,
                  int &epoch, hipStream_t s);

template <typename T>
int direct_factor(const DirectPattern &P, DirectWork<T> &W, const T *val, hipStream_t s);
template <typename T>
int direct_solve(const DirectPattern &P, DirectWork<T> &W, const T *val, const T *b, T *x, hipStream_t s);
// Solve A^T x = b with the same stored factorisation.
template <typename T>
int direct_solve_t(const DirectPattern &P, const DirectWork<T> &W, const T *val, const T *b, T *x,
                   hipStream_t s);

// adjoint kernels (v2d_adjoint.hip)
__global__ void k_forces_vjp(int imax, double re, double dksi, double deta, const double *m, int64_t N,
                             const double *ome, double d, double wd, double wl, double *g_ome,
                             double *g_m, double *out);
__global__ void k_bdf2_adjoint_rhs(int imax, int jmax, double inv2dt, const double *l1, const double *l2,
                                   double *g);
__global__ void k_residual_vjp(AsmArgs a, const uint8_t *flags, const double *lam, double *g_m,
                               double *partial);
__global__ void k_metrics_vjp_1(int imax, int jmax, double dksi, double deta, const double *x,
                                const double *y, const double *g_m, double *g_x, double *g_y,
                                double *g_eta);
__global__ void k_metrics_vjp_2(int imax, int jmax, double deta, const double *g_eta, double *g_x,
                                double *g_y);

// ------------------------------------------- multifrontal direct solver
// Nested-dissection multifrontal LU of the Newton matrix (the structure of the
// reference's MUMPS factorisation, vorti2d/petsc_solver.py:37-62).  The node
// graph is dissected by grid lines; every separator is one front, a dense
// (nf x nf) matrix over its own unknowns (ns pivots, padded with identity to
// nsp) and the unknowns of its ancestors it couples to (nb).  A partial
// Gauss-Jordan sweep of the nsp pivots leaves [G W; X C] in the front:
// G = F11^-1, W = G F12, X = -F21 G, C = F22 - F21 G F12, the contribution
// block the parent gathers.  Fronts are processed by height in the
// elimination tree, one batch of launches per height.
struct MfFront {
    int64_t fo;    // front matrix (elements), nf rows of stride ld
    int64_t vo;    // front vector, nf entries
    int64_t zo;    // pivot-part vector, nsp entries
    int64_t uo;    // unknowns of the front rows (field-blocked global index, -1: padding)
    int64_t mo;    // child maps: nch arrays of nf entries (child's struct row or -1)
    int64_t ao;    // first of the front's assembly positions
    int na;        // number of them
    int ns, nsp, nb, nf, ld;
    int nch, co;   // children: ch[co .. co + nch)
    int big;       // 1: swept by the dataflow Gauss-Jordan steps (nsp a multiple of 64)
};

// Big fronts of one 64-pivot step with two children and nf <= kFront1MaxNf,
// assembled (children, then the entries they own) and swept in two launches:
// one workgroup per front for the pivot block and pivot rows, one per front
// and 64-row tile for the rest.  Fronts f0 .. f0 + nfront - 1 (nf <= nfmax),
// jobs[q] the Gauss-Jordan job (panels) of front f0 + q; kp = 32 when every
// one of them has ns <= 32 (the products then run over 32 pivots), else 64.
constexpr int kFront1MaxNf = 512;
template <typename T>
int mf_front1(const MfFront *fr, int f0, int nfront, int nfmax, int kp, const Job<T> *jobs, const int *ch,
              const int *map, const int64_t *aoff, const int *aseg, const int *avid, const T *val, T *F, hipStream_t s);

struct MfLevel {
    int f0, f1, fbig;        // fronts [f0, f1); [fbig, f1) are big
    int64_t r0, r1;          // rows table range (every front row)
    int64_t g0, g1;          // forward row blocks (all front rows)
    int64_t b0, b1;          // backward row blocks (pivot rows)
    int ntg, ntb;            // threads per forward / backward row block (64, 128 or 256)
    int64_t c0, c1;          // transposed column blocks (64 columns of all front columns)
    int64_t p0, p1;          // transposed column blocks of the pivot columns
    int64_t a0, a1;          // assembly positions
    int64_t rbig, abig;      // rows and positions of the big fronts: [rbig, r1), [abig, a1)
    int64_t s0, s1;          // assembly blocks of the big fronts
    int nch2;                // 1: every big front of the height has two children
    int nfmax_small, nfmax_big, npmax_big;
    int small_ni[10];        // first small front swept with a register tile of NI 16-row blocks
    int one_step;            // 1: the big fronts fit mf_front1 (one 64-pivot step, two children,
                             //    nf <= kFront1MaxNf, every entry in the pivot rows or columns)
    int one_kp;              // mf_front1's product width: 32 if every big front has ns <= 32, else 64
    int64_t pool;            // Gauss-Jordan panel elements of the big fronts
    int64_t job0;            // first job of the level in the job table
};

struct MfPattern {
    int64_t N = 0;
    int imax = 0, jmax = 0;
    int nfront = 0, nheight = 0, nbig = 0;
    std::vector<MfFront> fr;
    std::vector<MfLevel> lv;
    int64_t F_elems = 0, V_elems = 0, Z_elems = 0, pool_elems = 0;
    int64_t nrows = 0, ngemv = 0, nbwd = 0, ncolt = 0, npcolt = 0, npos = 0, nvals = 0;
    int64_t nfu = 0, nmap = 0, nch = 0;
    double flops = 0.0;      // floating-point operations of one factorisation
    // host copies (kept for export)
    std::vector<int> fu, ch, map;
    std::vector<int64_t> aoff;
    std::vector<int> aseg, avid;
    std::vector<int2> rows, gemv, bwd, colt, pcolt;
    std::vector<int2> asmb;     // assembly blocks of the big fronts: (front, first row)
    std::vector<int64_t> cvo;   // per child-list entry: the child's coupling part in Vf
    // device copies
    MfFront *d_fr = nullptr;
    int *d_fu = nullptr, *d_ch = nullptr, *d_map = nullptr;
    int2 *d_rows = nullptr, *d_gemv = nullptr, *d_bwd = nullptr, *d_colt = nullptr, *d_pcolt = nullptr;
    int2 *d_asmb = nullptr;
    int64_t *d_cvo = nullptr;
    int64_t *d_aoff = nullptr;
    int *d_aseg = nullptr, *d_avid = nullptr;
    bool on_device = false;
};

// Per-kernel timing of a factorisation and a solve (HIP events around every
// launch on the stream) with the algorithmic counts of what was timed.
struct MfProbe {
    enum { kAssemble, kScatter, kSmall, kGjBegin, kGjFull, kGjLight, kGjHeavy, kFront, kGather, kFwd, kBwd, kGroups };
    double ms[kGroups] = {}, flops[kGroups] = {}, bytes[kGroups] = {};
    int64_t launches[kGroups] = {};
    std::vector<hipEvent_t> ev;     // pairs, recorded
    std::vector<int> grp;           // group of each pair
    std::vector<hipEvent_t> pool;   // created events for reuse
    size_t used = 0;
    ~MfProbe();
    int begin(hipStream_t s);
    int end(hipStream_t s, int group, double flops, double bytes);
    int flush();
};

// Host analysis only (no device call).
int mf_analyse(MfPattern &P, int imax, int jmax, const uint8_t *flags);
int mf_upload(MfPattern &P, hipStream_t s);
void mf_free_pattern(MfPattern &P, hipStream_t s);

template <typename T>
struct MfWork {
    T *F = nullptr, *Vf = nullptr, *Zf = nullptr, *pool = nullptr;
    Job<T> *jobs = nullptr;   // Gauss-Jordan jobs of the big fronts, level by level
    int *flags = nullptr, *err = nullptr;
    int epoch = 0;
    bool factored = false;
    bool paired = true;       // Gauss-Jordan steps in light / heavy pairs (else every step full)
    bool fused = true;        // one-step big fronts by mf_front1 (else assembly, scatter and steps)
    // two more streams for the small-front launches of a height (one per tile
    // size, independent), joined back to the factorisation stream by events
    hipStream_t aux[2] = {nullptr, nullptr};
    hipEvent_t fork = nullptr, join[2] = {nullptr, nullptr};
};
template <typename T>
int mf_alloc(const MfPattern &P, MfWork<T> &W, hipStream_t s);
template <typename T>
void mf_free(MfWork<T> &W, hipStream_t s);
// val: the Jacobian in the slot layout (kSlots planes of N).
template <typename T>
int mf_factor(const MfPattern &P, MfWork<T> &W, const T *val, hipStream_t s, MfProbe *probe = nullptr);
// x = A^-1 b (trans: A^-T b); b and x distinct 2N vectors.
template <typename T>
int mf_solve(const MfPattern &P, MfWork<T> &W, const T *b, T *x, bool trans, hipStream_t s,
             MfProbe *probe = nullptr);

}  // namespace v2d
