// v2d_mf.hip -- nested-dissection multifrontal LU of the Newton matrix.
//
// The reference factors the assembled Newton matrix with MUMPS, a multifrontal
// LU (vorti2d/petsc_solver.py:37-62: KSP PREONLY, PC LU, factor solver
// "mumps"; vorti2d/dist_solver.py:133-137; the default linsolve="mumps" of
// vorti2d/config.py:74-77).  This is the device counterpart, designed for the
// structured O-grid:
//
//   * analysis (host, once per grid): the node graph of the matrix (psi and
//     omega of a node share a vertex) is dissected recursively by grid lines:
//     a node set is split at the median of its longer extent (i periodic, the
//     two seam copies sharing i = 0), the separator is the smaller of the two
//     sides' boundary nodes, the rest splits into connected components, down to
//     leaves of kMfLeaf nodes.  Every separator is one front; its rows are its
//     own unknowns (pivots) and the ancestor unknowns it couples to, found by
//     the usual symbolic pass (a front's coupling set is its separator's graph
//     neighbours above it and its children's coupling sets, less itself).
//   * factorisation (device, per Jacobian): height by height in the
//     elimination tree, every front of a height is assembled -- its children's
//     contribution blocks gathered (fixed child order, so the sums are
//     deterministic), then the Jacobian entries it owns added -- and swept by
//     partial Gauss-Jordan on its pivots: small fronts in LDS, one workgroup
//     each; large ones by the MFMA dataflow steps of the block LU, as one job
//     table per height.  The front then holds [G W; X C] (G = F11^-1,
//     W = G F12, X = -F21 G, C = F22 - F21 G F12).
//   * solve: forward by height (front vector = b on the pivots + the
//     children's contribution vectors; z = G v_s; v_b += X v_s), backward from
//     the root (x_s = z - W x_b).  The transposed solve uses the same fronts
//     (the sweep of A^T is the transpose of A's: z = G^T v_s, v_b -= W^T v_s,
//     x_s = z + X^T x_b).
//
// No pivoting, as in the block-tridiagonal LU: the fronts' pivot blocks are
// Schur complements of the same diagonally weighted Newton matrix.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "v2d_internal.h"

namespace v2d {

namespace {

constexpr int kMfLeaf = 32;      // most nodes of a leaf supernode
constexpr int kMfSmall = 144;    // largest front swept in registers by one workgroup (16 x 9)
constexpr int kSmallMinNI = 3, kSmallMaxNI = 9;   // register-tile sizes of the small-front kernel
constexpr int kMfAlign = 4;      // pivot padding of small fronts
constexpr int kMfRow = 32;       // row stride of every front, in elements (128-byte FP32 rows)
constexpr int kNB = 64;          // Gauss-Jordan block of the large fronts
constexpr int kMfParFronts = 16384;   // fronts from which a height's small fronts use extra streams

__host__ __device__ inline int round_up(int x, int m) { return (x + m - 1) / m * m; }

// Lanes that share one row product of ncol columns (a power of two, 4 .. 64:
// about one 16-byte FP32 vector per lane and pass), and the rows of one
// workgroup's block at that width (four passes of 256 / lanes rows).
__host__ __device__ inline int row_lanes(int ncol)
{
    int l = 4;
    while (l < 64 && 4 * l < ncol) l <<= 1;
    return l;
}
// rows of a row-product block of nt threads (four rows per lane group)
__host__ __device__ inline int block_rows(int ncol, int nt) { return 4 * (nt / row_lanes(ncol)); }

// Rows of one assembly block of a big front of row stride ld (a multiple of
// 4): about kAsmItems items of 4 columns.
constexpr int kAsmItems = 4096;
__host__ __device__ inline int asm_rows(int ld)
{
    const int r = kAsmItems / (ld / 4);
    return r > 1 ? r : 1;
}


// Symmetrised node graph of the Jacobian pattern (CSR, no self loops).
struct NodeGraph {
    std::vector<int64_t> ptr;
    std::vector<int> adj;
};

void node_graph(int imax, int jmax, const uint8_t *flags, NodeGraph &G)
{
    const int64_t N = (int64_t)imax * jmax;
    std::vector<int64_t> cnt(N + 1, 0);
    auto visit = [&](auto &&edge) {
        int64_t cols[13];
        for (int comp = 0; comp < 2; ++comp)
            for (int64_t k = 0; k < N; ++k) {
                const int j = (int)(k / imax), i = (int)(k % imax);
                const uint8_t fl = j == jmax - 1 ? flags[i] : 0;
                const int c = row_columns(imax, jmax, i, j, comp, fl, N, cols);
                for (int q = 0; q < c; ++q) {
                    const int64_t kc = cols[q] >= N ? cols[q] - N : cols[q];
                    if (kc != k) edge(k, kc);
                }
            }
    };
    visit([&](int64_t a, int64_t b) { ++cnt[a + 1]; ++cnt[b + 1]; });
    for (int64_t k = 0; k < N; ++k) cnt[k + 1] += cnt[k];
    std::vector<int> raw(cnt[N]);
    std::vector<int64_t> fill(cnt.begin(), cnt.end() - 1);
    visit([&](int64_t a, int64_t b) { raw[fill[a]++] = (int)b; raw[fill[b]++] = (int)a; });
    G.ptr.assign(N + 1, 0);
    G.adj.clear();
    G.adj.reserve(raw.size() / 2);
    for (int64_t k = 0; k < N; ++k) {
        auto b = raw.begin() + cnt[k], e = raw.begin() + cnt[k + 1];
        std::sort(b, e);
        e = std::unique(b, e);
        G.adj.insert(G.adj.end(), b, e);
        G.ptr[k + 1] = (int64_t)G.adj.size();
    }
}

// Recursive dissection; supernodes are created children first (a postorder).
struct Dissect {
    const NodeGraph &G;
    std::vector<int> ic, jc;
    std::vector<int> stamp, tag;
    int cur = 0;
    std::vector<std::vector<int>> sep, kids;

    Dissect(const NodeGraph &g, int imax, int jmax) : G(g)
    {
        const int64_t N = (int64_t)imax * jmax;
        ic.resize(N);
        jc.resize(N);
        for (int64_t k = 0; k < N; ++k) {
            const int i = (int)(k % imax);
            ic[k] = i == imax - 1 ? 0 : i;
            jc[k] = (int)(k / imax);
        }
        stamp.assign(N, 0);
        tag.assign(N, 0);
    }

    int leaf(std::vector<int> &S)
    {
        std::sort(S.begin(), S.end());
        sep.push_back(std::move(S));
        kids.emplace_back();
        return (int)sep.size() - 1;
    }

    std::vector<int> rec(std::vector<int> S)
    {
        if ((int)S.size() <= kMfLeaf) return {leaf(S)};
        int ext[2];
        for (int ax = 0; ax < 2; ++ax) {
            const std::vector<int> &co = ax ? jc : ic;
            int lo = 1 << 30, hi = -1;
            for (int v : S) { lo = std::min(lo, co[v]); hi = std::max(hi, co[v]); }
            ext[ax] = hi - lo + 1;
        }
        const int axes[2] = {ext[0] > ext[1] ? 0 : 1, ext[0] > ext[1] ? 1 : 0};
        std::vector<int> best;
        bool found = false;
        for (int ax : axes) {
            const std::vector<int> &co = ax ? jc : ic;
            std::vector<int> c(S.size());
            for (size_t q = 0; q < S.size(); ++q) c[q] = co[S[q]];
            std::nth_element(c.begin(), c.begin() + c.size() / 2, c.end());
            const int med = c[c.size() / 2];
            ++cur;
            size_t na = 0;
            for (int v : S) {
                stamp[v] = cur;
                tag[v] = co[v] < med ? 1 : 2;
                na += co[v] < med;
            }
            if (na == 0 || na == S.size()) continue;
            std::vector<int> sa, sb;
            for (int v : S) {
                const int mine = tag[v], other = 3 - mine;
                for (int64_t e = G.ptr[v]; e < G.ptr[v + 1]; ++e) {
                    const int u = G.adj[e];
                    if (stamp[u] == cur && tag[u] == other) {
                        (mine == 1 ? sa : sb).push_back(v);
                        break;
                    }
                }
            }
            best = sa.size() <= sb.size() ? std::move(sa) : std::move(sb);
            found = true;
            break;
        }
        if (!found) return {leaf(S)};
        // the rest, split into connected components (discovery order)
        ++cur;
        for (int v : S) { stamp[v] = cur; tag[v] = 0; }
        for (int v : best) tag[v] = 3;
        std::vector<std::vector<int>> comps;
        std::vector<int> queue;
        for (int v0 : S) {
            if (tag[v0] != 0) continue;
            tag[v0] = 4;
            queue.assign(1, v0);
            for (size_t h = 0; h < queue.size(); ++h) {
                const int v = queue[h];
                for (int64_t e = G.ptr[v]; e < G.ptr[v + 1]; ++e) {
                    const int u = G.adj[e];
                    if (stamp[u] == cur && tag[u] == 0) {
                        tag[u] = 4;
                        queue.push_back(u);
                    }
                }
            }
            comps.push_back(queue);
        }
        S.clear();
        S.shrink_to_fit();
        std::vector<int> ch;
        for (auto &cpt : comps) {
            std::vector<int> r = rec(std::move(cpt));
            ch.insert(ch.end(), r.begin(), r.end());
        }
        if (best.empty()) return ch;
        std::sort(best.begin(), best.end());
        sep.push_back(std::move(best));
        kids.push_back(std::move(ch));
        return {(int)sep.size() - 1};
    }
};

}  // namespace

int mf_analyse(MfPattern &P, int imax, int jmax, const uint8_t *flags)
{
    if (imax < 4 || jmax < 4) {
        set_error("multifrontal: grid smaller than 4 x 4");
        return -1;
    }
    const int64_t N = (int64_t)imax * jmax;
    if (2 * N >= (int64_t)1 << 31) {
        set_error("multifrontal: more than 2^31 unknowns");
        return -1;
    }
    P = MfPattern{};
    P.N = N;
    P.imax = imax;
    P.jmax = jmax;
    NodeGraph G;
    node_graph(imax, jmax, flags, G);
    Dissect D(G, imax, jmax);
    {
        std::vector<int> all(N);
        std::iota(all.begin(), all.end(), 0);
        (void)D.rec(std::move(all));
    }
    const int ns_ = (int)D.sep.size();
    std::vector<int> owner(N, -1), pos(N, -1);
    for (int s = 0; s < ns_; ++s)
        for (size_t q = 0; q < D.sep[s].size(); ++q) {
            const int v = D.sep[s][q];
            if (owner[v] >= 0) {
                set_error("multifrontal: a node in two separators");
                return -1;
            }
            owner[v] = s;
            pos[v] = (int)q;
        }
    for (int64_t v = 0; v < N; ++v)
        if (owner[v] < 0) {
            set_error("multifrontal: a node in no separator");
            return -1;
        }
    // heights and coupling sets (children come before parents)
    std::vector<int> height(ns_, 0);
    std::vector<std::vector<int>> st(ns_);
    {
        std::vector<int> mark(N, -1);
        for (int s = 0; s < ns_; ++s) {
            for (int c : D.kids[s]) height[s] = std::max(height[s], height[c] + 1);
            std::vector<int> &L = st[s];
            for (int v : D.sep[s])
                for (int64_t e = G.ptr[v]; e < G.ptr[v + 1]; ++e) {
                    const int u = G.adj[e];
                    if (owner[u] > s && mark[u] != s) { mark[u] = s; L.push_back(u); }
                }
            for (int c : D.kids[s])
                for (int u : st[c])
                    if (owner[u] != s && mark[u] != s) { mark[u] = s; L.push_back(u); }
            std::sort(L.begin(), L.end(), [&](int a, int b) {
                return owner[a] != owner[b] ? owner[a] < owner[b] : pos[a] < pos[b];
            });
        }
    }
    // fronts ordered by height (stable in creation order); small before big
    // within a height
    std::vector<int> order(ns_);
    std::iota(order.begin(), order.end(), 0);
    auto is_big = [&](int s) {
        return round_up(2 * (int)D.sep[s].size(), kMfAlign) + 2 * (int)st[s].size() > kMfSmall;
    };
    // small fronts by size within a height, so that each register-tile size
    // (16-row blocks, see k_mf_small_fused) covers a contiguous range
    auto nf_small = [&](int s) {
        return is_big(s) ? 0 : round_up(2 * (int)D.sep[s].size(), kMfAlign) + 2 * (int)st[s].size();
    };
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
        if (height[a] != height[b]) return height[a] < height[b];
        if (is_big(a) != is_big(b)) return (int)is_big(a) < (int)is_big(b);
        return nf_small(a) < nf_small(b);
    });
    std::vector<int> fid(ns_);
    for (int f = 0; f < ns_; ++f) fid[order[f]] = f;
    P.nfront = ns_;
    P.nheight = ns_ ? height[order[ns_ - 1]] + 1 : 0;
    P.fr.resize(ns_);
    int64_t fo = 0, vo = 0, zo = 0, uo = 0, mo = 0;
    for (int f = 0; f < ns_; ++f) {
        const int s = order[f];
        MfFront &F = P.fr[f];
        F.ns = 2 * (int)D.sep[s].size();
        F.nb = 2 * (int)st[s].size();
        F.big = is_big(s) ? 1 : 0;
        F.nsp = round_up(F.ns, F.big ? kNB : kMfAlign);
        F.nf = F.nsp + F.nb;
        F.ld = round_up(F.nf, kMfRow);
        F.nch = (int)D.kids[s].size();
        F.co = (int)P.ch.size();
        for (int c : D.kids[s]) P.ch.push_back(fid[c]);
        F.fo = fo; fo += (int64_t)F.nf * F.ld;
        F.vo = vo; vo += F.nf;
        F.zo = zo; zo += F.nsp;
        F.uo = uo; uo += F.nf;
        F.mo = mo; mo += (int64_t)F.nch * F.nf;
        P.nbig += F.big;
        P.flops += 2.0 * (F.big ? F.nsp : F.ns) * (double)F.nf * F.nf;
    }
    P.cvo.resize(P.ch.size());
    for (size_t q = 0; q < P.ch.size(); ++q) P.cvo[q] = P.fr[P.ch[q]].vo + P.fr[P.ch[q]].nsp;
    P.F_elems = fo;
    P.V_elems = vo;
    P.Z_elems = zo;
    P.nfu = uo;
    P.nmap = mo;
    P.nch = (int64_t)P.ch.size();
    // unknowns of the front rows
    P.fu.assign(uo, -1);
    for (int f = 0; f < ns_; ++f) {
        const int s = order[f];
        const MfFront &F = P.fr[f];
        int *u = P.fu.data() + F.uo;
        for (size_t q = 0; q < D.sep[s].size(); ++q) {
            u[2 * q] = D.sep[s][q];
            u[2 * q + 1] = (int)(N + D.sep[s][q]);
        }
        for (size_t q = 0; q < st[s].size(); ++q) {
            u[F.nsp + 2 * q] = st[s][q];
            u[F.nsp + 2 * q + 1] = (int)(N + st[s][q]);
        }
    }
    // child maps: parent front row -> child's coupling row
    P.map.assign(mo, -1);
    {
        std::vector<int> idx(N, -1);
        for (int f = 0; f < ns_; ++f) {
            const int s = order[f];
            const MfFront &F = P.fr[f];
            for (size_t q = 0; q < st[s].size(); ++q) idx[st[s][q]] = (int)q;
            for (int k = 0; k < F.nch; ++k) {
                const int c = D.kids[s][k];
                int *m = P.map.data() + F.mo + (int64_t)k * F.nf;
                for (size_t q = 0; q < st[c].size(); ++q) {
                    const int u = st[c][q];
                    const int a = owner[u] == s ? 2 * pos[u] : F.nsp + 2 * idx[u];
                    if (owner[u] != s && (owner[u] < s || idx[u] < 0)) {
                        set_error("multifrontal: a child coupling outside its parent front");
                        return -1;
                    }
                    m[a] = 2 * (int)q;
                    m[a + 1] = 2 * (int)q + 1;
                }
            }
            for (size_t q = 0; q < st[s].size(); ++q) idx[st[s][q]] = -1;
        }
    }
    // Jacobian entries: each belongs to the front of the earlier of its row's
    // and column's supernodes, where both are front rows
    {
        std::vector<int64_t> cnt(ns_ + 1, 0);
        int64_t cols[13];
        auto each = [&](auto &&fn) {
            for (int comp = 0; comp < 2; ++comp)
                for (int64_t k = 0; k < N; ++k) {
                    const int j = (int)(k / imax), i = (int)(k % imax);
                    const uint8_t fl = j == jmax - 1 ? flags[i] : 0;
                    const int c = row_columns(imax, jmax, i, j, comp, fl, N, cols);
                    for (int q = 0; q < c; ++q) {
                        const int64_t kc = cols[q] >= N ? cols[q] - N : cols[q];
                        const int s = std::min(owner[k], owner[kc]);
                        const int vidx = (int)(((comp ? kPsiSlots : 0) + q) * N + k);
                        fn(fid[s], s, comp * N + k, cols[q], vidx);
                    }
                }
        };
        each([&](int f, int, int64_t, int64_t, int) { ++cnt[f + 1]; });
        for (int f = 0; f < ns_; ++f) cnt[f + 1] += cnt[f];
        P.nvals = cnt[ns_];
        std::vector<int> eg(P.nvals), ec(P.nvals), ev(P.nvals);
        {
            std::vector<int64_t> fill(cnt.begin(), cnt.end() - 1);
            each([&](int f, int, int64_t g, int64_t gc, int vidx) {
                const int64_t e = fill[f]++;
                eg[e] = (int)g;
                ec[e] = (int)gc;
                ev[e] = vidx;
            });
        }
        std::vector<int> idx(N, -1);
        std::vector<std::pair<int64_t, int>> pv;
        std::vector<int64_t> pend(ns_, 0);   // end of every front's positions
        P.aseg.push_back(0);
        for (int f = 0; f < ns_; ++f) {
            const int s = order[f];
            const MfFront &F = P.fr[f];
            for (size_t q = 0; q < st[s].size(); ++q) idx[st[s][q]] = (int)q;
            auto loc = [&](int64_t g) -> int {
                const int64_t u = g >= N ? g - N : g;
                const int fld = g >= N;
                if (owner[u] == s) return 2 * pos[u] + fld;
                return idx[u] < 0 ? -1 : F.nsp + 2 * idx[u] + fld;
            };
            pv.clear();
            for (int64_t e = cnt[f]; e < cnt[f + 1]; ++e) {
                const int a = loc(eg[e]), b = loc(ec[e]);
                if (a < 0 || b < 0) {
                    set_error("multifrontal: a Jacobian entry outside its front");
                    return -1;
                }
                pv.emplace_back(F.fo + (int64_t)a * F.ld + b, ev[e]);
            }
            std::sort(pv.begin(), pv.end());
            for (size_t q = 0; q < pv.size(); ++q) {
                if (q == 0 || pv[q].first != pv[q - 1].first) {
                    if (q) P.aseg.push_back((int)P.avid.size());
                    P.aoff.push_back(pv[q].first);
                }
                P.avid.push_back(pv[q].second);
            }
            if (!pv.empty()) P.aseg.push_back((int)P.avid.size());
            for (size_t q = 0; q < st[s].size(); ++q) idx[st[s][q]] = -1;
            P.fr[f].ao = f ? pend[f - 1] : 0;
            pend[f] = (int64_t)P.aoff.size();
            P.fr[f].na = (int)(pend[f] - P.fr[f].ao);
        }
        P.npos = (int64_t)P.aoff.size();
        // level ranges of the positions
        P.lv.assign(P.nheight, MfLevel{});
        int f = 0;
        int64_t a = 0;
        for (int h = 0; h < P.nheight; ++h) {
            MfLevel &L = P.lv[h];
            L.f0 = f;
            L.a0 = a;
            while (f < ns_ && height[order[f]] == h) {
                a = pend[f];
                ++f;
            }
            L.f1 = f;
            L.a1 = a;
        }
    }
    // per-height tables
    int64_t pool_max = 0;
    for (int h = 0; h < P.nheight; ++h) {
        MfLevel &L = P.lv[h];
        L.fbig = L.f1;
        for (int f = L.f0; f < L.f1; ++f)
            if (P.fr[f].big) { L.fbig = f; break; }
        L.r0 = (int64_t)P.rows.size();
        L.g0 = (int64_t)P.gemv.size();
        L.b0 = (int64_t)P.bwd.size();
        L.c0 = (int64_t)P.colt.size();
        L.p0 = (int64_t)P.pcolt.size();
        L.s0 = (int64_t)P.asmb.size();
        L.nch2 = 1;
        L.nfmax_small = L.nfmax_big = L.npmax_big = 0;
        for (int q = 0; q <= kSmallMaxNI; ++q) L.small_ni[q] = L.f0;
        L.pool = 0;
        // The row products of a height take the smallest block (64, 128 or
        // 256 threads) that holds every front's rows at once (rows past a
        // front's end would be read again and dropped); 256 and several
        // blocks per front where none does.
        L.ntg = L.ntb = 64;
        for (int f = L.f0; f < L.f1; ++f) {
            const MfFront &F = P.fr[f];
            while (L.ntg < 256 && block_rows(F.nsp, L.ntg) < F.nf) L.ntg <<= 1;
            while (L.ntb < 256 && block_rows(round_up(F.nb, kMfAlign), L.ntb) < F.ns) L.ntb <<= 1;
        }
        for (int f = L.f0; f < L.f1; ++f) {
            const MfFront &F = P.fr[f];
            for (int r = 0; r < F.nf; ++r) P.rows.push_back(make_int2(f, r));
            for (int r = 0; r < F.nf; r += block_rows(F.nsp, L.ntg)) P.gemv.push_back(make_int2(f, r));
            for (int r = 0; r < F.ns; r += block_rows(round_up(F.nb, kMfAlign), L.ntb))
                P.bwd.push_back(make_int2(f, r));
            for (int c = 0; c < F.nf; c += 64) P.colt.push_back(make_int2(f, c));
            for (int c = 0; c < F.ns; c += 64) P.pcolt.push_back(make_int2(f, c));
            if (F.big) {
                for (int r = 0; r < F.nf; r += asm_rows(F.ld)) P.asmb.push_back(make_int2(f, r));
                if (F.nch != 2) L.nch2 = 0;
                L.nfmax_big = std::max(L.nfmax_big, F.nf);
                L.npmax_big = std::max(L.npmax_big, F.nsp);
                L.pool += 3 * (int64_t)F.nf * kNB + 3 * (int64_t)round_up(F.nf, kNB) * kNB + kNB * kNB;
            } else {
                L.nfmax_small = std::max(L.nfmax_small, F.nf);
            }
        }
        // first small front of each register-tile size: NI = ceil(nf / 16), at
        // least kSmallMinNI (the small fronts are sorted by size)
        for (int q = 0; q <= kSmallMaxNI; ++q) {
            int f = L.f0;
            while (f < L.fbig && std::max(kSmallMinNI, (P.fr[f].nf + 15) / 16) < q) ++f;
            L.small_ni[q] = f;
        }
        L.r1 = (int64_t)P.rows.size();
        L.rbig = L.r1;
        L.abig = L.a1;
        {
            int64_t r = L.r0;
            for (int f = L.f0; f < L.f1; ++f) {
                if (P.fr[f].big) {
                    L.rbig = r;
                    L.abig = P.fr[f].ao;
                    break;
                }
                r += P.fr[f].nf;
            }
        }
        L.one_step = L.fbig < L.f1 && L.nch2 && L.npmax_big == kNB && L.nfmax_big <= kFront1MaxNf;
        L.one_kp = 32;
        for (int f = L.fbig; f < L.f1; ++f)
            if (P.fr[f].ns > 32) L.one_kp = 64;
        for (int f = L.fbig; f < L.f1 && L.one_step; ++f) {
            const MfFront &F = P.fr[f];
            for (int64_t u = F.ao; u < F.ao + F.na; ++u) {
                const int64_t q = P.aoff[u] - F.fo;
                if (q / F.ld >= F.nsp && q % F.ld >= F.nsp) {
                    L.one_step = 0;
                    break;
                }
            }
        }
        L.g1 = (int64_t)P.gemv.size();
        L.b1 = (int64_t)P.bwd.size();
        L.c1 = (int64_t)P.colt.size();
        L.p1 = (int64_t)P.pcolt.size();
        L.s1 = (int64_t)P.asmb.size();
        pool_max = std::max(pool_max, L.pool);
    }
    P.pool_elems = pool_max;
    int64_t job = 0;
    for (int h = 0; h < P.nheight; ++h) {
        P.lv[h].job0 = job;
        job += P.lv[h].f1 - P.lv[h].fbig;
    }
    P.nrows = (int64_t)P.rows.size();
    P.ngemv = (int64_t)P.gemv.size();
    P.nbwd = (int64_t)P.bwd.size();
    P.ncolt = (int64_t)P.colt.size();
    P.npcolt = (int64_t)P.pcolt.size();
    return 0;
}

namespace {
template <typename X>
int upload(X *&d, const std::vector<X> &h, hipStream_t s)
{
    if (h.empty()) return 0;
    V2D_HIP(hipMallocAsync((void **)&d, sizeof(X) * h.size(), s));
    V2D_HIP(hipMemcpyAsync(d, h.data(), sizeof(X) * h.size(), hipMemcpyHostToDevice, s));
    return 0;
}
}  // namespace

int mf_upload(MfPattern &P, hipStream_t s)
{
    if (P.on_device) return 0;
    if (upload(P.d_fr, P.fr, s) || upload(P.d_fu, P.fu, s) || upload(P.d_ch, P.ch, s) ||
        upload(P.d_map, P.map, s) || upload(P.d_rows, P.rows, s) || upload(P.d_cvo, P.cvo, s) || upload(P.d_gemv, P.gemv, s) ||
        upload(P.d_bwd, P.bwd, s) || upload(P.d_colt, P.colt, s) || upload(P.d_pcolt, P.pcolt, s) || upload(P.d_asmb, P.asmb, s) ||
        upload(P.d_aoff, P.aoff, s) || upload(P.d_aseg, P.aseg, s) || upload(P.d_avid, P.avid, s))
        return -1;
    V2D_HIP(hipStreamSynchronize(s));   // the host vectors may change after this returns
    P.on_device = true;
    return 0;
}

void mf_free_pattern(MfPattern &P, hipStream_t s)
{
    void *ptrs[] = {P.d_fr, P.d_fu, P.d_ch, P.d_map, P.d_rows, P.d_cvo, P.d_gemv, P.d_bwd,
                    P.d_colt, P.d_pcolt, P.d_asmb, P.d_aoff, P.d_aseg, P.d_avid};
    for (void *p : ptrs) if (p) (void)hipFreeAsync(p, s);
    P.d_fr = nullptr; P.d_fu = P.d_ch = P.d_map = nullptr;
    P.d_rows = P.d_gemv = P.d_bwd = P.d_colt = P.d_pcolt = P.d_asmb = nullptr;
    P.d_cvo = nullptr;
    P.d_aoff = nullptr; P.d_aseg = P.d_avid = nullptr;
    P.on_device = false;
}

template <typename T>
int mf_alloc(const MfPattern &P, MfWork<T> &W, hipStream_t s)
{
    if (!P.on_device) {
        set_error("multifrontal: pattern not on the device");
        return -1;
    }
    V2D_HIP(hipMallocAsync((void **)&W.F, sizeof(T) * std::max<int64_t>(P.F_elems, 1), s));
    V2D_HIP(hipMallocAsync((void **)&W.Vf, sizeof(T) * std::max<int64_t>(P.V_elems, 1), s));
    V2D_HIP(hipMallocAsync((void **)&W.Zf, sizeof(T) * std::max<int64_t>(P.Z_elems, 1), s));
    V2D_HIP(hipMallocAsync((void **)&W.pool, sizeof(T) * std::max<int64_t>(P.pool_elems, 1), s));
    int nbmax = 1;
    std::vector<Job<T>> jobs;
    for (const MfLevel &L : P.lv) {
        nbmax = std::max(nbmax, L.f1 - L.fbig);
        int64_t off = 0;
        for (int f = L.fbig; f < L.f1; ++f) {
            const MfFront &F = P.fr[f];
            Job<T> J{};
            J.S = W.F + F.fo;
            J.n = F.nf;
            J.ld = F.ld;
            J.l = 0;
            J.np = F.nsp;
            J.voff = 0;
            const int64_t r64 = round_up(F.nf, kNB);
            J.colk = W.pool + off; off += (int64_t)F.nf * kNB;
            J.colk_w = W.pool + off; off += (int64_t)F.nf * kNB;
            J.rowk = W.pool + off; off += r64 * kNB;
            J.rowk_w = W.pool + off; off += r64 * kNB;
            J.colk_x = W.pool + off; off += (int64_t)F.nf * kNB;
            J.rowk_x = W.pool + off; off += r64 * kNB;
            J.dinv = W.pool + off; off += kNB * kNB;
            J.tmat = nullptr;
            jobs.push_back(J);
        }
    }
    if (!jobs.empty()) {
        V2D_HIP(hipMallocAsync((void **)&W.jobs, sizeof(Job<T>) * jobs.size(), s));
        V2D_HIP(hipMemcpyAsync(W.jobs, jobs.data(), sizeof(Job<T>) * jobs.size(), hipMemcpyHostToDevice, s));
    }
    if (P.nfront >= kMfParFronts) {
        for (int k = 0; k < 2; ++k) {
            if (!W.aux[k]) V2D_HIP(hipStreamCreateWithFlags(&W.aux[k], hipStreamNonBlocking));
            if (!W.join[k]) V2D_HIP(hipEventCreateWithFlags(&W.join[k], hipEventDisableTiming));
        }
        if (!W.fork) V2D_HIP(hipEventCreateWithFlags(&W.fork, hipEventDisableTiming));
    }
    V2D_HIP(hipMallocAsync((void **)&W.flags, sizeof(int) * (nbmax + 1), s));
    V2D_HIP(hipMemsetAsync(W.flags, 0, sizeof(int) * (nbmax + 1), s));
    W.err = W.flags + nbmax;
    W.epoch = 0;
    W.factored = false;
    V2D_HIP(hipStreamSynchronize(s));   // jobs is a host temporary
    return 0;
}

template <typename T>
void mf_free(MfWork<T> &W, hipStream_t s)
{
    void *ptrs[] = {W.F, W.Vf, W.Zf, W.pool, W.jobs, W.flags};
    for (void *p : ptrs) if (p) (void)hipFreeAsync(p, s);
    for (int k = 0; k < 2; ++k) {
        if (W.aux[k]) {
            (void)hipStreamSynchronize(W.aux[k]);
            (void)hipStreamDestroy(W.aux[k]);
        }
        if (W.join[k]) (void)hipEventDestroy(W.join[k]);
    }
    if (W.fork) (void)hipEventDestroy(W.fork);
    const bool paired = W.paired, fused = W.fused;
    W = MfWork<T>{};
    W.paired = paired;
    W.fused = fused;
}

// ------------------------------------------------------------ kernels ----
namespace {

// F rows of the big fronts of one height: identity on the padding pivots plus
// the children's contribution blocks, summed in child-list order.  A
// workgroup writes asm_rows(ld) rows of one front as items of 4 consecutive
// columns (one 16-byte store in FP32), U items per thread and round.  Each
// round issues every map load of its items, then every contribution load,
// before the first sum, so a thread waits for two memory round trips per
// round rather than two per child and column.  NC = 2 (every big front of
// the O-grid dissection has two children) keeps both children's loads in one
// batch; NC = 0 takes any child count, one child at a time.
template <typename T, int NC>
__global__ void __launch_bounds__(256) k_mf_assemble(const MfFront *__restrict__ fr, const int2 *__restrict__ blk,
                                                     const int *__restrict__ ch, const int *__restrict__ map,
                                                     T *__restrict__ F)
{
    constexpr int U = 4;
    constexpr int KC = NC > 0 ? NC : 1;
    using P2 = T __attribute__((ext_vector_type(2)));
    const int2 fb = blk[blockIdx.x];
    const MfFront f = fr[fb.x];
    const int nf = f.nf, ld4 = f.ld >> 2;
    const int a0 = fb.y;
    const int items = min(asm_rows(f.ld), nf - a0) * ld4;
    const float inv = 1.0f / (float)ld4;
    T *dst0 = F + f.fo + (int64_t)a0 * f.ld;
    const int nk = NC > 0 ? NC : f.nch;
    for (int e0 = 0; e0 < items; e0 += 256 * U) {
        int r[U], c[U];
        bool ok[U];
        T acc[U][4];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int e = e0 + u * 256 + (int)threadIdx.x;
            ok[u] = e < items;
            int q = (int)((float)e * inv);   // e / ld4, corrected below
            if (q * ld4 > e) --q;
            else if ((q + 1) * ld4 <= e) ++q;
            r[u] = q;
            c[u] = 4 * (e - q * ld4);
            const int a = a0 + q;
            const bool pad = a >= f.ns && a < f.nsp;
#pragma unroll
            for (int v = 0; v < 4; ++v) acc[u][v] = (pad && c[u] + v == a) ? T(1) : T(0);
        }
        for (int k0 = 0; k0 < nk; k0 += KC) {
            int ia[KC][U], ib[KC][U][4];
            const T *cb[KC];
            int cld[KC];
#pragma unroll
            for (int kk = 0; kk < KC; ++kk) {
                const MfFront cf = fr[ch[f.co + k0 + kk]];
                cb[kk] = F + cf.fo + (int64_t)cf.nsp * cf.ld + cf.nsp;
                cld[kk] = cf.ld;
                const int *m = map + f.mo + (int64_t)(k0 + kk) * nf;
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    ia[kk][u] = ok[u] ? m[a0 + r[u]] : -1;
                    // nf is even: columns c, c + 1 exist together, as do c + 2, c + 3
                    const int2 lo = (ok[u] && c[u] < nf) ? *reinterpret_cast<const int2 *>(m + c[u]) : make_int2(-1, -1);
                    const int2 hi = (ok[u] && c[u] + 2 < nf) ? *reinterpret_cast<const int2 *>(m + c[u] + 2)
                                                             : make_int2(-1, -1);
                    ib[kk][u][0] = lo.x;
                    ib[kk][u][1] = lo.y;
                    ib[kk][u][2] = hi.x;
                    ib[kk][u][3] = hi.y;
                }
            }
            // columns c, c + 1 (c even) are the two unknowns of one node, which
            // a child holds at an even column and the next: one load per pair
            // the child holds (the lanes of absent pairs load nothing)
            T x[KC][U][4];
            bool in[KC][U][4];
#pragma unroll
            for (int kk = 0; kk < KC; ++kk)
#pragma unroll
                for (int u = 0; u < U; ++u)
#pragma unroll
                    for (int p = 0; p < 2; ++p) {
                        const bool okp = ia[kk][u] >= 0 && ib[kk][u][2 * p] >= 0;
                        const P2 y = okp ? *reinterpret_cast<const P2 *>(cb[kk] + ia[kk][u] * cld[kk] +
                                                                         ib[kk][u][2 * p])
                                         : P2{T(0), T(0)};
                        in[kk][u][2 * p] = in[kk][u][2 * p + 1] = okp;
                        x[kk][u][2 * p] = y[0];
                        x[kk][u][2 * p + 1] = y[1];
                    }
#pragma unroll
            for (int kk = 0; kk < KC; ++kk)
#pragma unroll
                for (int u = 0; u < U; ++u)
#pragma unroll
                    for (int v = 0; v < 4; ++v)
                        if (in[kk][u][v]) acc[u][v] += x[kk][u][v];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (!ok[u]) continue;
            T *d = dst0 + (int64_t)r[u] * f.ld + c[u];
            if constexpr (sizeof(T) == 4) {
                *reinterpret_cast<float4 *>(d) = make_float4(acc[u][0], acc[u][1], acc[u][2], acc[u][3]);
            } else {
                reinterpret_cast<double2 *>(d)[0] = make_double2(acc[u][0], acc[u][1]);
                reinterpret_cast<double2 *>(d)[1] = make_double2(acc[u][2], acc[u][3]);
            }
        }
    }
}

// The Jacobian entries the fronts own: F[aoff[u]] += sum of their values.
template <typename T>
__global__ void __launch_bounds__(256) k_mf_scatter(int64_t u0, int64_t u1, const int64_t *__restrict__ aoff,
                                                    const int *__restrict__ aseg, const int *__restrict__ avid,
                                                    const T *__restrict__ val, T *__restrict__ F)
{
    const int64_t u = u0 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= u1) return;
    T acc = T(0);
    for (int e = aseg[u]; e < aseg[u + 1]; ++e) acc += val[avid[e]];
    F[aoff[u]] += acc;
}

// A small front assembled and swept in one workgroup: its rows are built in
// LDS in chunks of at most kSmallLds bytes -- identity on the padding pivots,
// the children's contribution blocks gathered in list order, then the
// Jacobian entries it owns (the same sums in the same order as k_mf_assemble
// then k_mf_scatter) -- and loaded into registers: thread (ty, tx) of a
// 16 x 16 grid holds rows ty + 16 i and columns tx + 16 j (nf <= 16 NI).
// Then partial Gauss-Jordan on pivots 0 .. ns-1: per pivot p the owners of
// row p and of column p stage them in LDS (two buffers, alternating, so one
// barrier per pivot), every entry takes the rank-1 update and the pivot row
// and column are set: A[p][c] /= A[p][p], A[r][p] = -A[r][p] / A[p][p],
// A[p][p] = 1 / A[p][p].
constexpr int kSmallLds = 32768;

template <typename T, int NI>
__global__ void __launch_bounds__(256) k_mf_small_fused(const MfFront *__restrict__ fr, int f0,
                                                        const int *__restrict__ ch, const int *__restrict__ map,
                                                        const int64_t *__restrict__ aoff,
                                                        const int *__restrict__ aseg, const int *__restrict__ avid,
                                                        const T *__restrict__ val, T *__restrict__ F)
{
    // the staging buffer holds a whole front where (16 NI)^2 entries fit in
    // kSmallLds (smaller tiles then leave room for more workgroups per CU),
    // else rounds of cr rows
    constexpr int kSF = (int)(kSmallLds / sizeof(T)) < 256 * NI * NI ? (int)(kSmallLds / sizeof(T)) : 256 * NI * NI;
    __shared__ T sF[kSF];
    __shared__ T prow[2][16 * NI], pcol[2][16 * NI];
    const MfFront f = fr[f0 + blockIdx.x];
    const int nf = f.nf, ns = f.ns;
    const int ty = threadIdx.x >> 4, tx = threadIdx.x & 15;
    const int cr = min(nf, kSF / nf);
    T a[NI][NI];
#pragma unroll
    for (int i = 0; i < NI; ++i)
#pragma unroll
        for (int j = 0; j < NI; ++j) a[i][j] = T(0);
    for (int r0 = 0; r0 < nf; r0 += cr) {
        const int r1 = min(nf, r0 + cr), ne = (r1 - r0) * nf;
        // Every element pair (2 q, 2 q + 1) is set and summed by one thread:
        // identity on the padding pivots, then the children's contributions
        // in list order.  Rounds of U pairs per thread issue every map load,
        // then every contribution load (clamped addresses, the absent ones
        // dropped by a select: no load behind a branch), before the first sum.
        constexpr int U = 4;
        using P2 = T __attribute__((ext_vector_type(2)));
        const int np2 = ne >> 1;   // element pairs (2 q, 2 q + 1): nf is even, a pair lies in one row
        for (int e0 = 0; e0 < np2; e0 += 256 * U) {
            int er[U], ec[U];
            T acc[U][2];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int e = 2 * min(e0 + u * 256 + (int)threadIdx.x, np2 - 1);
                er[u] = r0 + e / nf;
                ec[u] = e % nf;
#pragma unroll
                for (int v = 0; v < 2; ++v)
                    acc[u][v] = (er[u] >= ns && er[u] < f.nsp && ec[u] + v == er[u]) ? T(1) : T(0);
            }
            for (int k = 0; k < f.nch; ++k) {
                const int *m = map + f.mo + (int64_t)k * nf;
                const MfFront c = fr[ch[f.co + k]];
                const T *cb = F + c.fo + (int64_t)c.nsp * c.ld + c.nsp;
                int ia[U], ib[U];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    ia[u] = m[er[u]];
                    ib[u] = m[ec[u]];
                }
                // columns ec, ec + 1 are the two unknowns of one node, which the
                // child holds at an even column and the next: one load per pair
                P2 x[U];
#pragma unroll
                for (int u = 0; u < U; ++u)
                    x[u] = *reinterpret_cast<const P2 *>(cb + max(ia[u], 0) * c.ld + max(ib[u], 0));
#pragma unroll
                for (int u = 0; u < U; ++u)
                    if (ia[u] >= 0 && ib[u] >= 0) {
                        acc[u][0] += x[u][0];
                        acc[u][1] += x[u][1];
                    }
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int e = e0 + u * 256 + (int)threadIdx.x;
                if (e < np2) {
                    sF[2 * e] = acc[u][0];
                    sF[2 * e + 1] = acc[u][1];
                }
            }
        }
        __syncthreads();
        // the Jacobian entries the front owns, U positions per thread and
        // round, every load of the round issued before the first sum (every
        // position holds at least one entry)
        for (int64_t u0 = f.ao; u0 < f.ao + f.na; u0 += 256 * U) {
            int64_t q[U];
            int s0[U], s1[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int64_t uu = min(u0 + u * 256 + (int64_t)threadIdx.x, f.ao + f.na - 1);
                q[u] = aoff[uu] - f.fo;
                s0[u] = aseg[uu];
                s1[u] = aseg[uu + 1];
            }
            int v0[U];
#pragma unroll
            for (int u = 0; u < U; ++u) v0[u] = avid[s0[u]];
            T acc[U];
#pragma unroll
            for (int u = 0; u < U; ++u) acc[u] = T(0) + val[v0[u]];
#pragma unroll
            for (int u = 0; u < U; ++u)
                for (int e = s0[u] + 1; e < s1[u]; ++e) acc[u] += val[avid[e]];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int r = (int)(q[u] / f.ld), c = (int)(q[u] % f.ld);
                if (u0 + u * 256 + (int64_t)threadIdx.x < f.ao + f.na && r >= r0 && r < r1)
                    sF[(r - r0) * nf + c] += acc[u];
            }
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < NI; ++i) {
            const int r = ty + 16 * i;
            if (r >= r0 && r < r1)
#pragma unroll
                for (int j = 0; j < NI; ++j) {
                    const int c = tx + 16 * j;
                    if (c < nf) a[i][j] = sF[(r - r0) * nf + c];
                }
        }
        __syncthreads();
    }
    int par = 0;
#pragma unroll
    for (int I = 0; I < NI; ++I) {
        if (16 * I >= ns) break;
        for (int q = 0; q < 16 && 16 * I + q < ns; ++q) {
            const int p = 16 * I + q;
            if (ty == q)
#pragma unroll
                for (int j = 0; j < NI; ++j) prow[par][tx + 16 * j] = a[I][j];
            if (tx == q)
#pragma unroll
                for (int i = 0; i < NI; ++i) pcol[par][ty + 16 * i] = a[i][I];
            __syncthreads();
            const T ip = T(1) / pcol[par][p];
            T pr[NI], pc[NI];
#pragma unroll
            for (int j = 0; j < NI; ++j) pr[j] = prow[par][tx + 16 * j] * ip;
#pragma unroll
            for (int i = 0; i < NI; ++i) pc[i] = pcol[par][ty + 16 * i];
#pragma unroll
            for (int i = 0; i < NI; ++i)
#pragma unroll
                for (int j = 0; j < NI; ++j) a[i][j] = fma(-pc[i], pr[j], a[i][j]);
            if (ty == q)
#pragma unroll
                for (int j = 0; j < NI; ++j) a[I][j] = pr[j];
            if (tx == q)
#pragma unroll
                for (int i = 0; i < NI; ++i) a[i][I] = -pc[i] * ip;
            if (ty == q && tx == q) a[I][I] = ip;
            par ^= 1;
        }
    }
    T *g = F + f.fo;
#pragma unroll
    for (int i = 0; i < NI; ++i)
#pragma unroll
        for (int j = 0; j < NI; ++j) {
            const int r = ty + 16 * i, c = tx + 16 * j;
            if (r < nf && c < nf) g[(int64_t)r * f.ld + c] = a[i][j];
        }
    // the row stride's tail (columns nf .. ld) stays zero, as k_mf_assemble leaves it
    for (int e = threadIdx.x; e < nf * (f.ld - nf); e += blockDim.x) {
        const int r = e / (f.ld - nf), c = nf + e % (f.ld - nf);
        g[(int64_t)r * f.ld + c] = T(0);
    }
}

// Forward gather: v = b on the pivots, 0 on padding and couplings, plus the
// children's contribution vectors (list order).  One thread per front row.
template <typename T>
__global__ void __launch_bounds__(256) k_mf_fgather(const MfFront *__restrict__ fr, const int2 *__restrict__ rows,
                                                    int64_t r0, int64_t r1, const int64_t *__restrict__ cvo,
                                                    const int *__restrict__ map, const int *__restrict__ fu,
                                                    const T *__restrict__ b, T *__restrict__ Vf)
{
    const int64_t t = r0 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= r1) return;
    const int2 fa = rows[t];
    const MfFront f = fr[fa.x];
    const int a = fa.y;
    // clamped loads, absent terms dropped by a select (no load behind a branch)
    const T bv = b[fu[f.uo + min(a, f.ns - 1)]];
    T v = a < f.ns ? bv : T(0);
    int k = 0;
    for (; k + 2 <= f.nch; k += 2) {
        const int m0 = map[f.mo + (int64_t)k * f.nf + a], m1 = map[f.mo + (int64_t)(k + 1) * f.nf + a];
        const T x0 = Vf[cvo[f.co + k] + max(m0, 0)], x1 = Vf[cvo[f.co + k + 1] + max(m1, 0)];
        if (m0 >= 0) v += x0;
        if (m1 >= 0) v += x1;
    }
    for (; k < f.nch; ++k) {
        const int m = map[f.mo + (int64_t)k * f.nf + a];
        const T x = Vf[cvo[f.co + k] + max(m, 0)];
        if (m >= 0) v += x;
    }
    Vf[f.vo + a] = v;
}

template <typename T> struct V16;
template <> struct V16<float> { using type = float4; static constexpr int w = 4; };
template <> struct V16<double> { using type = double2; static constexpr int w = 2; };

template <typename T>
__device__ __forceinline__ double dot16(const typename V16<T>::type &a, const T *x)
{
    const T *e = reinterpret_cast<const T *>(&a);
    double s = 0.0;
#pragma unroll
    for (int q = 0; q < V16<T>::w; ++q) s += (double)e[q] * (double)x[q];
    return s;
}

__device__ __forceinline__ double wave_sum64(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Row products of one front against a vector staged in LDS: out(r, p, s)
// with s = sum_c F[r][c0 + c] xs[c] in FP64, for the rows row0 + q G + g
// (q < 4) of a block of NT threads below rend, G = NT / L rows per pass, L =
// row_lanes(ncol) lanes per row (16-byte vectors, summed across the group by
// shuffles).  The four rows of a lane group are read together: their first
// vectors and pre(r), the per-row operands out() needs, are loaded before
// stage() fills xs (they do not depend on it).  No load sits behind a branch
// (addresses are clamped instead: a load in a branch is waited for before
// the branch joins, which would serialise the rows), so a block waits for one
// memory round trip before its sums.  ncol and c0 are multiples of 4 (the
// vectors stay aligned); rows past rend read row rend - 1 and are dropped.
template <typename T, int NT, typename Stage, typename Pre, typename Out>
__device__ __forceinline__ void mf_rows(const MfFront &f, const T *F, int row0, int rend, int c0, int ncol,
                                        const T *xs, Stage &&stage, Pre &&pre, Out &&out)
{
    using V = typename V16<T>::type;
    constexpr int W = V16<T>::w;
    const int L = row_lanes(ncol), G = NT / L;
    const int l = threadIdx.x & (L - 1), g = threadIdx.x / L;
    const int nv = ncol / W;
    const int cb = nv > 0 ? c0 : 0;              // no columns: read the row start
    const int lv = nv > 0 ? min(l, nv - 1) : 0;
    const V *R[4];
    V x0[4];
    decltype(pre(0)) pv[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int r = min(row0 + q * G + g, rend - 1);
        R[q] = reinterpret_cast<const V *>(F + f.fo + (int64_t)r * f.ld + cb);
        x0[q] = R[q][lv];
        pv[q] = pre(r);
    }
    stage();
    __syncthreads();
    double acc[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) acc[q] = (l < nv) ? dot16<T>(x0[q], xs + l * W) : 0.0;
    for (int c = l + L; c < nv; c += L) {
        V x[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) x[q] = R[q][c];
#pragma unroll
        for (int q = 0; q < 4; ++q) acc[q] += dot16<T>(x[q], xs + c * W);
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        for (int o = L >> 1; o > 0; o >>= 1) acc[q] += __shfl_xor(acc[q], o, 64);
        const int r = row0 + q * G + g;
        if (l == 0 && r < rend) out(r, pv[q], acc[q]);
    }
}

// Forward front products: u = F[:, 0:nsp] v_s (v gathered by k_mf_fgather);
// z = u on the pivots (Zf), v_b + u on the couplings (Vf, read by the parent).
template <typename T, int NT>
__global__ void __launch_bounds__(NT) k_mf_fgemv(const MfFront *__restrict__ fr, const int2 *__restrict__ blk,
                                                 int64_t g0, const T *__restrict__ F, T *__restrict__ Vf,
                                                 T *__restrict__ Zf)
{
    extern __shared__ unsigned char smem[];
    T *xs = reinterpret_cast<T *>(smem);
    const int2 fb = blk[g0 + blockIdx.x];
    const MfFront f = fr[fb.x];
    mf_rows<T, NT>(
        f, F, fb.y, f.nf, 0, f.nsp, xs,
        [&] {
            for (int a = threadIdx.x; a < f.nsp; a += NT) xs[a] = Vf[f.vo + a];
        },
        [&](int r) { return Vf[f.vo + r]; },
        [&](int r, T vb, double acc) {
            if (r < f.nsp) Zf[f.zo + r] = (T)acc;
            else Vf[f.vo + r] = (T)((double)vb + acc);
        });
}

// The per-row operands of a backward row: z and the row's unknown.
template <typename T>
struct ZU {
    T z;
    int u;
};

// Backward row products: x_s = z - W x_b, x_b gathered from the solution.
template <typename T, int NT>
__global__ void __launch_bounds__(NT) k_mf_bwd(const MfFront *__restrict__ fr, const int2 *__restrict__ blk,
                                               int64_t b0, const T *__restrict__ F, const T *__restrict__ Zf,
                                               const int *__restrict__ fu, T *__restrict__ x)
{
    extern __shared__ unsigned char smem[];
    T *xs = reinterpret_cast<T *>(smem);
    const int2 fb = blk[b0 + blockIdx.x];
    const MfFront f = fr[fb.x];
    const int nbp = round_up(f.nb, kMfAlign);
    // the W block is read up to the 4-rounded coupling count: those columns
    // lie inside the row stride and meet zeros in xs
    mf_rows<T, NT>(
        f, F, fb.y, f.ns, f.nsp, nbp, xs,
        [&] {
            for (int q = threadIdx.x; q < nbp; q += NT) xs[q] = q < f.nb ? x[fu[f.uo + f.nsp + q]] : T(0);
        },
        [&](int r) { return ZU<T>{Zf[f.zo + r], fu[f.uo + r]}; },
        [&](int r, const ZU<T> &zu, double acc) { x[zu.u] = (T)((double)zu.z - acc); });
}

// Column products for the transposed solve: out(c) = sum_r F[r0 + r][c] ws[r]
// over rows [r0, r1) for the 64 columns c0 .. of the block; the four
// wavefronts take every fourth row and their sums are added in wavefront
// order.
template <typename T, typename Out>
__device__ __forceinline__ void mf_cols(const MfFront &f, const T *F, int r0, int r1, int c0, int cend,
                                        const T *ws, double *red, Out &&out)
{
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int c = c0 + lane;
    double acc = 0.0;
    if (c < cend)
        for (int r = r0 + wave; r < r1; r += 4) acc += (double)F[f.fo + (int64_t)r * f.ld + c] * (double)ws[r - r0];
    red[wave * 64 + lane] = acc;
    __syncthreads();
    if (wave == 0 && c < cend) out(c, ((red[lane] + red[64 + lane]) + red[128 + lane]) + red[192 + lane]);
}

// Transposed forward: u' = F[0:nsp, :]^T v_s; z = u' on the pivots, v_b -= u'.
template <typename T>
__global__ void __launch_bounds__(256) k_mf_fgemv_t(const MfFront *__restrict__ fr, const int2 *__restrict__ blk,
                                                    int64_t c0, const T *__restrict__ F, T *__restrict__ Vf,
                                                    T *__restrict__ Zf)
{
    extern __shared__ unsigned char smem[];
    double *red = reinterpret_cast<double *>(smem);
    T *ws = reinterpret_cast<T *>(red + 256);
    const int2 fb = blk[c0 + blockIdx.x];
    const MfFront f = fr[fb.x];
    for (int r = threadIdx.x; r < f.nsp; r += blockDim.x) ws[r] = Vf[f.vo + r];
    __syncthreads();
    mf_cols<T>(f, F, 0, f.nsp, fb.y, f.nf, ws, red, [&](int c, double acc) {
        if (c < f.nsp) Zf[f.zo + c] = (T)acc;
        else Vf[f.vo + c] = (T)((double)Vf[f.vo + c] - acc);
    });
}

// Transposed backward: x_s = z + X^T x_b.
template <typename T>
__global__ void __launch_bounds__(256) k_mf_bwd_t(const MfFront *__restrict__ fr, const int2 *__restrict__ blk,
                                                  int64_t p0, const T *__restrict__ F, const T *__restrict__ Zf,
                                                  const int *__restrict__ fu, T *__restrict__ x)
{
    extern __shared__ unsigned char smem[];
    double *red = reinterpret_cast<double *>(smem);
    T *ws = reinterpret_cast<T *>(red + 256);
    const int2 fb = blk[p0 + blockIdx.x];
    const MfFront f = fr[fb.x];
    for (int q = threadIdx.x; q < f.nb; q += blockDim.x) ws[q] = x[fu[f.uo + f.nsp + q]];
    __syncthreads();
    mf_cols<T>(f, F, f.nsp, f.nf, fb.y, f.ns, ws, red, [&](int c, double acc) {
        x[fu[f.uo + c]] = (T)((double)Zf[f.zo + c] + acc);
    });
}

__global__ void k_mf_zero_int(int *p) { *p = 0; }

}  // namespace

// ------------------------------------------------------------ probe -----
MfProbe::~MfProbe()
{
    for (hipEvent_t e : pool) (void)hipEventDestroy(e);
}

int MfProbe::begin(hipStream_t s)
{
    while (pool.size() < used + 2) {
        hipEvent_t e;
        V2D_HIP(hipEventCreate(&e));
        pool.push_back(e);
    }
    V2D_HIP(hipEventRecord(pool[used], s));
    return 0;
}

int MfProbe::end(hipStream_t s, int group, double fl, double by)
{
    V2D_HIP(hipEventRecord(pool[used + 1], s));
    ev.push_back(pool[used]);
    ev.push_back(pool[used + 1]);
    grp.push_back(group);
    used += 2;
    flops[group] += fl;
    bytes[group] += by;
    ++launches[group];
    return 0;
}

int MfProbe::flush()
{
    if (!ev.empty()) V2D_HIP(hipEventSynchronize(ev.back()));
    for (size_t q = 0; q < grp.size(); ++q) {
        float t = 0.f;
        V2D_HIP(hipEventElapsedTime(&t, ev[2 * q], ev[2 * q + 1]));
        ms[grp[q]] += t;
    }
    ev.clear();
    grp.clear();
    used = 0;
    return 0;
}

#define MF_BEGIN()                                                                           \
    do {                                                                                     \
        if (probe && probe->begin(s)) return -1;                                             \
    } while (0)
#define MF_END(g, fl, by)                                                                    \
    do {                                                                                     \
        if (probe && probe->end(s, MfProbe::g, (fl), (by))) return -1;                       \
    } while (0)

template <typename T>
int mf_factor(const MfPattern &P, MfWork<T> &W, const T *val, hipStream_t s, MfProbe *probe)
{
    if (!W.F) {
        set_error("multifrontal: work not allocated");
        return -1;
    }
    W.factored = false;
    const double ts = sizeof(T);
    hipLaunchKernelGGL(k_mf_zero_int, dim3(1), dim3(1), 0, s, W.err);
    // extra streams only for large analyses: small grids are factorised many
    // at a time (the batched sweep), whose streams already fill the queues
    const bool par = !probe && W.aux[0] && W.aux[1] && P.nfront >= kMfParFronts;
    bool pend = false;   // small-front launches of the last height still on the extra streams
    auto join = [&]() -> int {
        if (pend)
            for (int k = 0; k < 2; ++k) {
                V2D_HIP(hipEventRecord(W.join[k], W.aux[k]));
                V2D_HIP(hipStreamWaitEvent(s, W.join[k], 0));
            }
        pend = false;
        return 0;
    };
    for (int h = 0; h < P.nheight; ++h) {
        if (join()) return -1;
        const MfLevel &L = P.lv[h];
        const int64_t nr = L.r1 - L.r0;
        // one-step big fronts: assembled, scattered and swept by mf_front1
        const bool one = L.one_step && W.fused;
        // algorithmic counts of the height (host, only when timed)
        double asm_b = 0, small_fl = 0, small_b = 0, one_fl = 0, one_b = 0;
        if (probe)
            for (int f = L.f0; f < L.f1; ++f) {
                const MfFront &F = P.fr[f];
                double cb = 0;
                for (int k = 0; k < F.nch; ++k) {
                    const MfFront &C = P.fr[P.ch[F.co + k]];
                    cb += ts * (double)C.nb * C.nb;
                }
                if (F.big && one) {
                    one_fl += 2.0 * kNB * (double)(F.nf - kNB) * F.nf;
                    one_b += ts * (double)F.nf * F.ld + cb + (ts + 12.0) * F.na;
                } else if (F.big) {
                    asm_b += ts * (double)F.nf * F.ld + cb;
                } else {
                    small_fl += 2.0 * F.ns * (double)F.nf * F.nf;
                    small_b += ts * (double)F.nf * F.ld + cb + (ts + 12.0) * F.na;
                }
            }
        if (L.fbig > L.f0) {
            MF_BEGIN();
            // the tile sizes' launches are independent of each other and of the
            // height's big fronts: untimed, they run on the two extra streams
            // (and on the factorisation stream when the height has no big
            // front), joined back at the next height
            if (par) {
                V2D_HIP(hipEventRecord(W.fork, s));
                for (int k = 0; k < 2; ++k) V2D_HIP(hipStreamWaitEvent(W.aux[k], W.fork, 0));
                pend = true;
            }
            const int nst = L.fbig < L.f1 ? 2 : 3;
            int nl = 0;
            for (int ni = kSmallMinNI; ni <= kSmallMaxNI; ++ni) {
                const int a0 = L.small_ni[ni], a1 = ni < kSmallMaxNI ? L.small_ni[ni + 1] : L.fbig;
                if (a1 <= a0) continue;
                const dim3 g((unsigned)(a1 - a0));
                const hipStream_t st = !par ? s : nl % nst < 2 ? W.aux[nl % nst] : s;
                ++nl;
#define V2D_SMALL(NI)                                                                                      \
    hipLaunchKernelGGL((k_mf_small_fused<T, NI>), g, dim3(256), 0, st, P.d_fr, a0, P.d_ch, P.d_map, P.d_aoff, \
                       P.d_aseg, P.d_avid, val, W.F)
                switch (ni) {
                case 3: V2D_SMALL(3); break;
                case 4: V2D_SMALL(4); break;
                case 5: V2D_SMALL(5); break;
                case 6: V2D_SMALL(6); break;
                case 7: V2D_SMALL(7); break;
                case 8: V2D_SMALL(8); break;
                default: V2D_SMALL(9); break;
                }
#undef V2D_SMALL
            }
            MF_END(kSmall, small_fl, small_b);
        }
        if (!one && L.s1 > L.s0) {
            const dim3 g((unsigned)(L.s1 - L.s0));
            MF_BEGIN();
            if (L.nch2)
                hipLaunchKernelGGL((k_mf_assemble<T, 2>), g, dim3(256), 0, s, P.d_fr, P.d_asmb + L.s0, P.d_ch,
                                   P.d_map, W.F);
            else
                hipLaunchKernelGGL((k_mf_assemble<T, 0>), g, dim3(256), 0, s, P.d_fr, P.d_asmb + L.s0, P.d_ch,
                                   P.d_map, W.F);
            MF_END(kAssemble, 0.0, asm_b);
        }
        if (!one && L.a1 > L.abig) {
            MF_BEGIN();
            hipLaunchKernelGGL(k_mf_scatter<T>, dim3(grid_for(L.a1 - L.abig)), dim3(256), 0, s, L.abig, L.a1,
                               P.d_aoff, P.d_aseg, P.d_avid, val, W.F);
            MF_END(kScatter, 0.0, (2 * ts + 12.0) * (double)(L.a1 - L.abig));
        }
        V2D_LAUNCH();
        const int nj = L.f1 - L.fbig;
        if (nj <= 0 || L.npmax_big <= 0) continue;
        if (one) {
            MF_BEGIN();
            if (mf_front1<T>(P.d_fr, L.fbig, nj, L.nfmax_big, L.one_kp, W.jobs + L.job0, P.d_ch, P.d_map, P.d_aoff,
                             P.d_aseg, P.d_avid, val, W.F, s))
                return -1;
            MF_END(kFront, one_fl, one_b);
            continue;
        }
        // Steps in pairs: a light step (the next panels only) and a heavy step
        // (every other tile takes both steps' panels), a last unpaired step full.
        const int nsteps = L.npmax_big / kNB;
        std::vector<char> ends(nsteps, 0);   // a job's last step (no compact light grid there)
        for (int f = L.fbig; f < L.f1; ++f) ends[P.fr[f].nsp / kNB - 1] = 1;
        MF_BEGIN();
        if (gj_table_begin<T>(W.jobs + L.job0, nj, L.nfmax_big, W.flags, W.err, W.epoch, s)) return -1;
        MF_END(kGjBegin, 0.0, 0.0);
        for (int st = 0; st < nsteps; ++st) {
            const int k0 = st * kNB;
            const int mode = !W.paired ? kStepFull
                             : (st & 1) ? kStepHeavy
                                        : (st + 1 < nsteps ? kStepLight : kStepFull);
            double fl = 0, by = 0;
            if (probe)
                for (int f = L.fbig; f < L.f1; ++f) {
                    const MfFront &F = P.fr[f];
                    if (k0 >= F.nsp) continue;
                    const double n = F.nf, kb = kNB;
                    const bool last = k0 + kNB >= F.nsp;
                    // elements of a step's update (rows outside the pivots, every
                    // column) and of a light step's share of it
                    const double e_step = (n - kb) * n, e_light = kb * n + kb * (n - 2 * kb);
                    double e = e_step;
                    if (mode == kStepLight && !last) e = e_light;
                    if (mode == kStepHeavy) e = e_step + (e_step - e_light);
                    fl += 2.0 * kb * e;
                    by += ts * (2.0 * e + 2.0 * kb * n * (mode == kStepHeavy ? 2 : 1));
                }
            MF_BEGIN();
            if (gj_table_step<T>(W.jobs + L.job0, nj, L.nfmax_big, k0, mode, !ends[st], W.flags, W.err, W.epoch, s))
                return -1;
            if (probe && probe->end(s, mode == kStepHeavy ? MfProbe::kGjHeavy
                                       : mode == kStepLight ? MfProbe::kGjLight
                                                            : MfProbe::kGjFull,
                                    fl, by))
                return -1;
        }
    }
    if (join()) return -1;
    int e = 0;
    V2D_HIP(hipMemcpyAsync(&e, W.err, sizeof(int), hipMemcpyDeviceToHost, s));
    V2D_HIP(hipStreamSynchronize(s));
    if (e) {
        set_error("multifrontal: a Gauss-Jordan step waited past its spin limit for the pivot inverse "
                  "(results invalid)");
        return -1;
    }
    W.factored = true;
    return 0;
}

template <typename T>
int mf_solve(const MfPattern &P, MfWork<T> &W, const T *b, T *x, bool trans, hipStream_t s, MfProbe *probe)
{
    if (!W.factored) {
        set_error("multifrontal: solve before a factorisation");
        return -1;
    }
    if (b == x) {
        set_error("multifrontal: right-hand side and solution must be distinct");
        return -1;
    }
    const double ts = sizeof(T);
    for (int h = 0; h < P.nheight; ++h) {
        const MfLevel &L = P.lv[h];
        int nspmax = 0;
        double fb = 0, gb = 0;
        for (int f = L.f0; f < L.f1; ++f) {
            const MfFront &F = P.fr[f];
            nspmax = std::max(nspmax, F.nsp);
            if (probe) {
                fb += ts * ((double)F.nf * F.nsp + F.nsp + 2.0 * F.nf);
                gb += ts * (2.0 * F.nf) + 4.0 * F.nf * (1 + F.nch);
            }
        }
        if (L.r1 > L.r0) {
            MF_BEGIN();
            hipLaunchKernelGGL(k_mf_fgather<T>, dim3(grid_for(L.r1 - L.r0)), dim3(256), 0, s, P.d_fr, P.d_rows,
                               L.r0, L.r1, P.d_cvo, P.d_map, P.d_fu, b, W.Vf);
            MF_END(kGather, 0.0, gb);
        }
        if (!trans && L.g1 > L.g0) {
            MF_BEGIN();
            const dim3 g((unsigned)(L.g1 - L.g0));
            const size_t lds = sizeof(T) * nspmax;
            if (L.ntg == 64)
                hipLaunchKernelGGL((k_mf_fgemv<T, 64>), g, dim3(64), lds, s, P.d_fr, P.d_gemv, L.g0, W.F, W.Vf, W.Zf);
            else if (L.ntg == 128)
                hipLaunchKernelGGL((k_mf_fgemv<T, 128>), g, dim3(128), lds, s, P.d_fr, P.d_gemv, L.g0, W.F, W.Vf,
                                   W.Zf);

            else
                hipLaunchKernelGGL((k_mf_fgemv<T, 256>), g, dim3(256), lds, s, P.d_fr, P.d_gemv, L.g0, W.F, W.Vf,
                                   W.Zf);
            MF_END(kFwd, 0.0, fb);
        }
        if (trans && L.c1 > L.c0)
            hipLaunchKernelGGL(k_mf_fgemv_t<T>, dim3((unsigned)(L.c1 - L.c0)), dim3(256),
                               sizeof(double) * 256 + sizeof(T) * nspmax, s, P.d_fr, P.d_colt, L.c0, W.F, W.Vf,
                               W.Zf);
        V2D_LAUNCH();
    }
    for (int h = P.nheight - 1; h >= 0; --h) {
        const MfLevel &L = P.lv[h];
        int nbmax = 0;
        double bb = 0;
        for (int f = L.f0; f < L.f1; ++f) {
            const MfFront &F = P.fr[f];
            nbmax = std::max(nbmax, round_up(F.nb, kMfAlign));
            if (probe) bb += ts * ((double)F.ns * round_up(F.nb, kMfAlign) + F.nb + 2.0 * F.ns);
        }
        if (!trans && L.b1 > L.b0) {
            MF_BEGIN();
            const dim3 g((unsigned)(L.b1 - L.b0));
            const size_t lds = sizeof(T) * std::max(nbmax, 1);
            if (L.ntb == 64)
                hipLaunchKernelGGL((k_mf_bwd<T, 64>), g, dim3(64), lds, s, P.d_fr, P.d_bwd, L.b0, W.F, W.Zf, P.d_fu, x);
            else if (L.ntb == 128)
                hipLaunchKernelGGL((k_mf_bwd<T, 128>), g, dim3(128), lds, s, P.d_fr, P.d_bwd, L.b0, W.F, W.Zf, P.d_fu,
                                   x);

            else
                hipLaunchKernelGGL((k_mf_bwd<T, 256>), g, dim3(256), lds, s, P.d_fr, P.d_bwd, L.b0, W.F, W.Zf, P.d_fu,
                                   x);
            MF_END(kBwd, 0.0, bb);
        }
        if (trans && L.p1 > L.p0)
            hipLaunchKernelGGL(k_mf_bwd_t<T>, dim3((unsigned)(L.p1 - L.p0)), dim3(256),
                               sizeof(double) * 256 + sizeof(T) * std::max(nbmax, 1), s, P.d_fr, P.d_pcolt, L.p0,
                               W.F, W.Zf, P.d_fu, x);
        V2D_LAUNCH();
    }
    return 0;
}

template int mf_alloc<float>(const MfPattern &, MfWork<float> &, hipStream_t);
template int mf_alloc<double>(const MfPattern &, MfWork<double> &, hipStream_t);
template void mf_free<float>(MfWork<float> &, hipStream_t);
template void mf_free<double>(MfWork<double> &, hipStream_t);
template int mf_factor<float>(const MfPattern &, MfWork<float> &, const float *, hipStream_t, MfProbe *);
template int mf_factor<double>(const MfPattern &, MfWork<double> &, const double *, hipStream_t, MfProbe *);
template int mf_solve<float>(const MfPattern &, MfWork<float> &, const float *, float *, bool, hipStream_t,
                             MfProbe *);
template int mf_solve<double>(const MfPattern &, MfWork<double> &, const double *, double *, bool, hipStream_t,
                              MfProbe *);

}  // namespace v2d
