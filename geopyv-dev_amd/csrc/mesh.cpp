// Mesh solve on the engine: reliability-guided propagation of subset solves
// over a triangular mesh, the caller of the subset path (mesh.rs:628-857).
//
// The host keeps the traversal (adjacency, the max-heap of C_ZNCC keys,
// the preconditioning fits, the corrections and the element strains); every
// node solve is a list solve on one device batch that holds all the nodes.
// Serial RG launches one node at a time, as its order demands.  Layer-RG
// launches one list per colour of each frontier round (and a second, zero
// warp list for the nodes whose preconditioned attempt failed), so a round
// of hundreds of nodes costs one or two launches instead of hundreds of
// CPU tasks.  The node solve order, every trusted set and every initial
// warp are the reference's, so the results agree with it to the subset
// solver's own tolerance.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <functional>
#include <mutex>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <utility>
#include <vector>

#include "internal.h"

namespace geodic {

namespace {

// Process-wide host workers for the traversal's larger per-node and
// per-element passes (adjacency, the corrections' statistics, strains, the
// initial warps of long lists).  run(n, grain, f) calls f(begin, end) on
// chunks of [0, n) of `grain` items, taken in turn by the calling thread
// and the workers, and returns when every chunk is done; a job of one
// chunk runs on the caller.  Workers sleep between jobs (spinning ones
// would take cycles from the thread that drives the device).  One job at a
// time: a caller that finds another job running does its work alone.  The
// job's fields are published seqlock-style: gen is odd while the next
// caller rewrites them, which it does only once no worker is inside the
// previous job, and a worker reads them only at an even gen.
class HostPool {
  public:
    static HostPool& get() {
        static HostPool p;
        return p;
    }
    int threads() const { return (int)th_.size() + 1; }
    template <class F>
    void run(int64_t n, int64_t grain, F&& f) {
        if (n <= 0) return;
        grain = std::max<int64_t>(grain, 1);
        std::unique_lock<std::mutex> job(run_mu_, std::try_to_lock);
        if (n <= grain || th_.empty() || !job.owns_lock()) {
            f((int64_t)0, n);
            return;
        }
        gen_.fetch_add(1);  // odd: fields being rewritten
        while (busy_.load() != 0) {
        }
        fn_ = [](void* c, int64_t b, int64_t e) { (*static_cast<F*>(c))(b, e); };
        ctx_ = &f;
        n_ = n;
        grain_ = grain;
        chunks_ = (n + grain - 1) / grain;
        next_.store(0);
        done_.store(0);
        {
            std::lock_guard<std::mutex> lk(mu_);
            gen_.fetch_add(1);  // even: published
        }
        cv_.notify_all();
        work();
        while (done_.load() < chunks_) {
        }
    }
    ~HostPool() {
        stop_.store(true);
        {
            std::lock_guard<std::mutex> lk(mu_);
            cv_.notify_all();
        }
        for (auto& t : th_) t.join();
    }

  private:
    HostPool() {
        const int hc = (int)std::thread::hardware_concurrency();  // 0 when unknown
        const int nt = std::max(0, std::min(16, hc) - 1);
        for (int k = 0; k < nt; ++k) th_.emplace_back([this] { loop(); });
    }
    void work() {
        for (int64_t c; (c = next_.fetch_add(1)) < chunks_;) {
            fn_(ctx_, c * grain_, std::min(n_, (c + 1) * grain_));
            done_.fetch_add(1);
        }
    }
    void loop() {
        uint64_t seen = gen_.load();
        for (;;) {
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [&] {
                    const uint64_t g = gen_.load();
                    return stop_.load() || (g != seen && !(g & 1));
                });
            }
            if (stop_.load()) return;
            busy_.fetch_add(1);
            const uint64_t g = gen_.load();
            if (!(g & 1) && g != seen) {
                work();
                seen = g;
            }
            busy_.fetch_sub(1);
        }
    }
    std::vector<std::thread> th_;
    std::mutex run_mu_, mu_;
    std::condition_variable cv_;
    std::atomic<uint64_t> gen_{0};
    alignas(64) std::atomic<int64_t> next_{0};
    alignas(64) std::atomic<int64_t> done_{0};
    alignas(64) std::atomic<int> busy_{0};
    std::atomic<bool> stop_{false};
    void (*fn_)(void*, int64_t, int64_t) = nullptr;
    void* ctx_ = nullptr;
    int64_t n_ = 0, grain_ = 1, chunks_ = 0;
};

// Ascending neighbour lists of every node in one flat array (node indices
// fit in 32 bits: at most 2^30 nodes).
struct Csr {
    std::vector<int64_t> off;
    std::vector<int32_t> idx;
    const int32_t* begin(int64_t i) const { return idx.data() + off[i]; }
    const int32_t* end(int64_t i) const { return idx.data() + off[i + 1]; }
};

struct Topology {
    int64_t n_nodes = 0, n_el = 0;
    int cols = 3, order = 1;
    const int64_t* el = nullptr;
    Csr full;  // nodes sharing an element (connectivity, full = true; mesh.rs:1705-1764)
    Csr edge;  // edge-adjacent nodes (full = false): corner <-> midside for 6-node elements
    std::vector<int64_t> inc_off, inc;  // elements incident to each node
    int64_t at(int64_t e, int k) const { return el[e * cols + k]; }
    // for 3-node elements every pair of nodes in an element shares an edge
    const Csr& edges() const { return order == 2 ? edge : full; }
};

// Adjacency::build (mesh.rs:1853-1868) as CSR: incident elements per node,
// then each node's neighbours sorted and de-duplicated, the nodes in
// parallel (a count pass, the offsets, a fill pass; a node's list does not
// depend on the thread that builds it).
void build_topology(Topology& t) {
    const int64_t n = t.n_nodes;
    std::vector<int64_t>& cnt = t.inc_off;
    cnt.assign(n + 1, 0);
    for (int64_t e = 0; e < t.n_el; ++e)
        for (int k = 0; k < t.cols; ++k) ++cnt[t.at(e, k) + 1];
    for (int64_t i = 0; i < n; ++i) cnt[i + 1] += cnt[i];
    t.inc.resize(cnt[n]);
    std::vector<int64_t> fill(cnt.begin(), cnt.end() - 1);
    for (int64_t e = 0; e < t.n_el; ++e)
        for (int k = 0; k < t.cols; ++k) t.inc[fill[t.at(e, k)]++] = e;
    // the two edge partners of each position of a 6-node element
    // [c0, c1, c2, m01, m12, m20] (mesh.rs:1745-1752)
    static const int partner[6][2] = {{3, 5}, {3, 4}, {4, 5}, {0, 1}, {1, 2}, {0, 2}};
    HostPool& pool = HostPool::get();
    constexpr int64_t kGrain = 2048;
    for (int pass = 0; pass < (t.order == 2 ? 2 : 1); ++pass) {
        Csr& out = pass == 0 ? t.full : t.edge;
        const auto gather = [&](int64_t i, std::vector<int32_t>& tmp) {
            tmp.clear();
            for (int64_t j = cnt[i]; j < cnt[i + 1]; ++j) {
                const int64_t e = t.inc[j];
                for (int k = 0; k < t.cols; ++k) {
                    if (pass == 0) {
                        if (t.at(e, k) != i) tmp.push_back((int32_t)t.at(e, k));
                    } else if (t.at(e, k) == i) {
                        for (int a = 0; a < 2; ++a)
                            if (t.at(e, partner[k][a]) != i) tmp.push_back((int32_t)t.at(e, partner[k][a]));
                    }
                }
            }
            std::sort(tmp.begin(), tmp.end());
            tmp.erase(std::unique(tmp.begin(), tmp.end()), tmp.end());
        };
        out.off.assign(n + 1, 0);
        pool.run(n, kGrain, [&](int64_t b, int64_t e) {
            std::vector<int32_t> tmp;
            for (int64_t i = b; i < e; ++i) {
                gather(i, tmp);
                out.off[i + 1] = (int64_t)tmp.size();
            }
        });
        for (int64_t i = 0; i < n; ++i) out.off[i + 1] += out.off[i];
        out.idx.resize(out.off[n]);
        pool.run(n, kGrain, [&](int64_t b, int64_t e) {
            std::vector<int32_t> tmp;
            for (int64_t i = b; i < e; ++i) {
                gather(i, tmp);
                std::copy(tmp.begin(), tmp.end(), out.idx.begin() + out.off[i]);
            }
        });
    }
}

// lower and upper quartiles of the finite values: the percentile of
// mesh.rs:1919-1930 (linear interpolation between the order statistics at
// floor and floor + 1 of p/100 (n - 1)), found by selection instead of a
// full sort; v is permuted
void quartiles(std::vector<double>& v, double& lq, double& uq) {
    v.erase(std::remove_if(v.begin(), v.end(), [](double x) { return !std::isfinite(x); }), v.end());
    if (v.empty()) {
        lq = uq = 0.0;
        return;
    }
    const size_t n = v.size();
    auto pct = [&](double p) {
        const double pos = p / 100.0 * ((double)n - 1.0);
        const size_t lo = (size_t)std::floor(pos);
        const size_t hi = std::min(lo + 1, n - 1);
        std::nth_element(v.begin(), v.begin() + lo, v.end());
        const double a = v[lo];
        // after the selection every later element is >= v[lo]: the next
        // order statistic is their minimum
        const double b = hi == lo ? a : *std::min_element(v.begin() + lo + 1, v.end());
        return a + (pos - (double)lo) * (b - a);
    };
    lq = pct(25.0);
    uq = pct(75.0);
}

// f64::total_cmp order: by sign-folded bit pattern
int64_t total_key(double a) {
    int64_t x;
    std::memcpy(&x, &a, 8);
    return x ^ (int64_t)((uint64_t)(x >> 63) >> 1);
}

// Max-heap of (key, node) pairs in std::pair order, four children per
// entry (shallower than a binary heap, so a pop touches fewer cache lines).
// The pairs are distinct, so the pop order is that of any max-heap.
class MaxHeap {
  public:
    using Item = std::pair<uint64_t, int64_t>;
    bool empty() const { return v_.empty(); }
    const Item& top() const { return v_.front(); }
    void push(const Item& x) {
        size_t i = v_.size();
        v_.push_back(x);
        while (i > 0) {
            const size_t p = (i - 1) / 4;
            if (!(v_[p] < x)) break;
            v_[i] = v_[p];
            i = p;
        }
        v_[i] = x;
    }
    void pop() {
        const Item x = v_.back();
        v_.pop_back();
        const size_t n = v_.size();
        if (n == 0) return;
        size_t i = 0;
        for (;;) {
            const size_t c0 = 4 * i + 1;
            if (c0 >= n) break;
            size_t best = c0;
            const size_t c1 = std::min(c0 + 4, n);
            for (size_t c = c0 + 1; c < c1; ++c)
                if (v_[best] < v_[c]) best = c;
            if (!(x < v_[best])) break;
            v_[i] = v_[best];
            i = best;
        }
        v_[i] = x;
    }

  private:
    std::vector<Item> v_;
};

uint64_t heap_key(double c) {
    uint64_t k;
    std::memcpy(&k, &c, 8);
    return k;
}

// 3x3 inverse by cofactors, as nalgebra's try_inverse computes it for
// m = 3; false when the determinant is exactly zero
bool inverse3(const double* A, double* out) {
    const double c00 = A[4] * A[8] - A[7] * A[5];
    const double c01 = A[3] * A[8] - A[6] * A[5];
    const double c02 = A[3] * A[7] - A[6] * A[4];
    const double det = A[0] * c00 - A[1] * c01 + A[2] * c02;
    if (det == 0.0) return false;
    out[0] = c00 / det;
    out[1] = -(A[1] * A[8] - A[7] * A[2]) / det;
    out[2] = (A[1] * A[5] - A[4] * A[2]) / det;
    out[3] = -c01 / det;
    out[4] = (A[0] * A[8] - A[6] * A[2]) / det;
    out[5] = -(A[0] * A[5] - A[3] * A[2]) / det;
    out[6] = c02 / det;
    out[7] = -(A[0] * A[7] - A[6] * A[1]) / det;
    out[8] = (A[0] * A[4] - A[3] * A[1]) / det;
    return true;
}

// m x m inverse as nalgebra's try_inverse computes it for m > 4: LU with
// partial pivoting (the first largest magnitude in the column; a zero
// column is skipped), unit-lower multipliers scaled by the pivot's
// reciprocal, then forward and back substitution of each column of the
// row-permuted identity.  False when a diagonal entry of U is exactly
// zero.  The same operations in the same order keep the preconditioning
// warps, and so the node solve order, identical to the reference's.
bool inverse_lu(int m, const double* A, double* out) {
    double L[36];
    int swaps[6][2];
    int ns = 0;
    std::memcpy(L, A, sizeof(double) * m * m);
    for (int i = 0; i < m; ++i) {
        int piv = i;
        double big = std::fabs(L[i * m + i]);
        for (int r = i + 1; r < m; ++r)
            if (std::fabs(L[r * m + i]) > big) {
                big = std::fabs(L[r * m + i]);
                piv = r;
            }
        const double d = L[piv * m + i];
        if (d == 0.0) continue;
        if (piv != i) {
            swaps[ns][0] = i;
            swaps[ns][1] = piv;
            ++ns;
            for (int c = 0; c < m; ++c) std::swap(L[i * m + c], L[piv * m + c]);
        }
        const double rd = 1.0 / d;
        for (int r = i + 1; r < m; ++r) L[r * m + i] *= rd;
        for (int c = i + 1; c < m; ++c) {
            const double u = L[i * m + c];
            for (int r = i + 1; r < m; ++r) L[r * m + c] = -u * L[r * m + i] + L[r * m + c];
        }
    }
    for (int i = 0; i < m; ++i)
        if (L[i * m + i] == 0.0) return false;
    double B[36] = {0};
    for (int i = 0; i < m; ++i) B[i * m + i] = 1.0;
    for (int k = 0; k < ns; ++k)
        for (int c = 0; c < m; ++c) std::swap(B[swaps[k][0] * m + c], B[swaps[k][1] * m + c]);
    for (int c = 0; c < m; ++c) {
        for (int i = 0; i < m; ++i) {
            const double v = B[i * m + c];
            for (int r = i + 1; r < m; ++r) B[r * m + c] = -v * L[r * m + i] + B[r * m + c];
        }
        for (int i = m - 1; i >= 0; --i) {
            const double v = B[i * m + c] / L[i * m + i];
            B[i * m + c] = v;
            for (int r = 0; r < i; ++r) B[r * m + c] = -v * L[r * m + i] + B[r * m + c];
        }
    }
    std::memcpy(out, B, sizeof(double) * m * m);
    return true;
}

// Taylor shift of a warp by (dx, dy) (project_warp, mesh.rs:2135-2150)
void shift_warp(const double* p, int plen, double dx, double dy, double* out) {
    for (int k = 0; k < plen; ++k) out[k] = p[k];
    if (plen >= 6) {
        out[0] = p[0] + p[2] * dx + p[4] * dy;
        out[1] = p[1] + p[3] * dx + p[5] * dy;
    }
    if (plen >= 12) {
        out[0] += 0.5 * p[6] * dx * dx + p[8] * dx * dy + 0.5 * p[10] * dy * dy;
        out[1] += 0.5 * p[7] * dx * dx + p[9] * dx * dy + 0.5 * p[11] * dy * dy;
        out[2] = p[2] + p[6] * dx + p[8] * dy;
        out[3] = p[3] + p[7] * dx + p[9] * dy;
        out[4] = p[4] + p[8] * dx + p[10] * dy;
        out[5] = p[5] + p[9] * dx + p[11] * dy;
    }
}

// Least-squares field through the trusted neighbours' (u, v), as offsets
// from the target (affine_field_fit, mesh.rs:2181-2227): the basis is
// 1, dx, dy (plus dx², dx·dy, dy² for order 2 with six or more neighbours);
// the inverse of the normal matrix times the right-hand sides (ols_solve,
// mesh.rs:2230-2240), false when the normal matrix is singular.
bool field_fit(const double* nodes, const double* tgt, const int64_t* tr, int64_t nt, const double* P, int plen,
               double* out) {
    const bool quad = plen >= 12 && nt >= 6;
    const int m = quad ? 6 : 3;
    double A[36] = {0}, Y[12] = {0};
    for (int64_t k = 0; k < nt; ++k) {
        const int64_t i = tr[k];
        const double dx = nodes[2 * i] - tgt[0], dy = nodes[2 * i + 1] - tgt[1];
        const double row[6] = {1.0, dx, dy, dx * dx, dx * dy, dy * dy};
        for (int a = 0; a < m; ++a) {
            for (int c = 0; c < m; ++c) A[a * m + c] += row[a] * row[c];
            Y[2 * a] += row[a] * P[i * plen];
            Y[2 * a + 1] += row[a] * P[i * plen + 1];
        }
    }
    double inv[36];
    if (!(m == 3 ? inverse3(A, inv) : inverse_lu(6, A, inv))) return false;
    double cu[6], cv[6];
    for (int a = 0; a < m; ++a) {
        double su = 0.0, sv = 0.0;
        for (int k = 0; k < m; ++k) {
            su += inv[m * a + k] * Y[2 * k];
            sv += inv[m * a + k] * Y[2 * k + 1];
        }
        cu[a] = su;
        cv[a] = sv;
    }
    for (int k = 0; k < plen; ++k) out[k] = 0.0;
    out[0] = cu[0];
    out[1] = cv[0];
    out[2] = cu[1];
    out[3] = cv[1];
    out[4] = cu[2];
    out[5] = cv[2];
    if (quad) {  // u = ... + uxx dx²/2 + uxy dx dy + uyy dy²/2
        out[6] = 2.0 * cu[3];
        out[7] = 2.0 * cv[3];
        out[8] = cu[4];
        out[9] = cv[4];
        out[10] = 2.0 * cu[5];
        out[11] = 2.0 * cv[5];
    }
    return true;
}

// precondition_warp (mesh.rs:2152-2179): the field fit over three or more
// trusted neighbours; else the shifted warp of the best-correlated one (the
// last of equals); else zero.
void initial_warp(const double* nodes, const double* tgt, const std::vector<int64_t>& tr, const double* czn,
                  const double* P, int plen, double* out) {
    const int64_t nt = (int64_t)tr.size();
    if (nt >= 3 && field_fit(nodes, tgt, tr.data(), nt, P, plen, out)) return;
    if (nt > 0) {
        int64_t best = tr[0];
        for (int64_t k = 1; k < nt; ++k)
            if (czn[tr[k]] >= czn[best]) best = tr[k];
        shift_warp(P + best * plen, plen, tgt[0] - nodes[2 * best], tgt[1] - nodes[2 * best + 1], out);
        return;
    }
    for (int k = 0; k < plen; ++k) out[k] = 0.0;
}

// element_strains (mesh.rs:1552-1665): centroid displacement and the
// first (and for 6-node elements second) displacement gradients from the
// shape functions at the centroid; elements on the host workers, the
// first degenerate one (in element order) reported.
int strains(const double* nodes, const Topology& t, const double* disp, double* warps, std::string& err) {
    static const double N1[3] = {1.0 / 3.0, 1.0 / 3.0, 1.0 / 3.0};
    static const double dN1[2][6] = {{1.0, 0.0, -1.0}, {0.0, 1.0, -1.0}};
    static const double N2[6] = {-1.0 / 9.0, -1.0 / 9.0, -1.0 / 9.0, 4.0 / 9.0, 4.0 / 9.0, 4.0 / 9.0};
    static const double dN2[2][6] = {{1.0 / 3.0, 0.0, -1.0 / 3.0, 4.0 / 3.0, -4.0 / 3.0, 0.0},
                                     {0.0, 1.0 / 3.0, -1.0 / 3.0, 4.0 / 3.0, 0.0, -4.0 / 3.0}};
    static const double d2N[3][6] = {
        {4.0, 0.0, 4.0, 0.0, 0.0, -8.0}, {0.0, 0.0, 4.0, 4.0, -4.0, -4.0}, {0.0, 4.0, 4.0, 0.0, -8.0, 0.0}};
    const int ns = t.order == 1 ? 3 : 6;
    const double* N = t.order == 1 ? N1 : N2;
    const double(*dN)[6] = t.order == 1 ? dN1 : dN2;
    std::atomic<int64_t> first_bad{t.n_el};
    HostPool::get().run(t.n_el, 2048, [&](int64_t e0, int64_t e1) {
        for (int64_t e = e0; e < e1; ++e) {
            double* w = warps + 12 * e;
            double X[6][2], U[6][2];
            for (int k = 0; k < ns; ++k) {
                const int64_t i = t.at(e, k);
                X[k][0] = nodes[2 * i];
                X[k][1] = nodes[2 * i + 1];
                U[k][0] = disp[2 * i];
                U[k][1] = disp[2 * i + 1];
            }
            for (int k = 0; k < 12; ++k) w[k] = 0.0;
            for (int k = 0; k < ns; ++k) w[0] += N[k] * U[k][0];
            for (int k = 0; k < ns; ++k) w[1] += N[k] * U[k][1];
            double jx[2][2] = {{0, 0}, {0, 0}}, ju[2][2] = {{0, 0}, {0, 0}};
            for (int r = 0; r < 2; ++r)
                for (int c = 0; c < 2; ++c)
                    for (int k = 0; k < ns; ++k) {
                        jx[r][c] += dN[r][k] * X[k][c];
                        ju[r][c] += dN[r][k] * U[k][c];
                    }
            const double det = jx[0][0] * jx[1][1] - jx[0][1] * jx[1][0];
            if (std::fabs(det) < 2.220446049250313e-16 * 1e6) {
                int64_t cur = first_bad.load();
                while (e < cur && !first_bad.compare_exchange_weak(cur, e)) {
                }
                continue;
            }
            const double ij[2][2] = {{jx[1][1] / det, -jx[0][1] / det}, {-jx[1][0] / det, jx[0][0] / det}};
            // [du/dx, dv/dx, du/dy, dv/dy]
            for (int r = 0; r < 2; ++r)
                for (int c = 0; c < 2; ++c) w[2 + 2 * r + c] = ij[r][0] * ju[0][c] + ij[r][1] * ju[1][c];
            if (t.order == 2) {
                double ku[3][2] = {{0, 0}, {0, 0}, {0, 0}};
                for (int r = 0; r < 3; ++r)
                    for (int c = 0; c < 2; ++c)
                        for (int k = 0; k < 6; ++k) ku[r][c] += d2N[r][k] * U[k][c];
                // area coordinates' gradients of the corner triangle
                const double x0 = X[0][0], y0 = X[0][1], x1 = X[1][0], y1 = X[1][1], x2 = X[2][0], y2 = X[2][1];
                const double a2 = (x1 * y2 - x2 * y1) - (x0 * y2 - x2 * y0) + (x0 * y1 - x1 * y0);
                const double g[2][2] = {{(y1 - y2) / a2, (y2 - y0) / a2}, {(x2 - x1) / a2, (x0 - x2) / a2}};
                const double kx[3][3] = {{g[0][0] * g[0][0], 2.0 * g[0][0] * g[0][1], g[0][1] * g[0][1]},
                                         {g[0][0] * g[1][0], g[0][0] * g[1][1] + g[0][1] * g[1][0], g[0][1] * g[1][1]},
                                         {g[1][0] * g[1][0], 2.0 * g[1][0] * g[1][1], g[1][1] * g[1][1]}};
                // [uxx, vxx, uxy, vxy, uyy, vyy]
                for (int r = 0; r < 3; ++r)
                    for (int c = 0; c < 2; ++c) w[6 + 2 * r + c] = kx[r][0] * ku[0][c] + kx[r][1] * ku[1][c] + kx[r][2] * ku[2][c];
            }
        }
    });
    if (first_bad.load() < t.n_el) {
        err = "degenerate Jacobian at element " + std::to_string(first_bad.load());
        return DIC_ERR_MESH_GENERATION;
    }
    return DIC_OK;
}

// The traversal's per-node state (the arrays of solve_impl, mesh.rs:639-648).
struct Run {
    const Topology* t;
    const double* nodes;
    Batch* b;
    const Image* f;
    const Image* g;
    dic_solve_params sp;
    int plen;
    std::vector<uint8_t> stored, propagated, qok;
    std::vector<double> czn, P, disp, norms;
    std::vector<int32_t> iters;
    MaxHeap heap;
    int64_t n_solves = 0, n_lists = 0;
    double list_seconds = 0.0;  // wall time inside list solves (launches, device work, synchronisation)
    // list-solve buffers
    std::vector<int32_t> ids;
    std::vector<double> p0;
    std::vector<dic_subset_result> ra, rb;
    std::vector<int64_t> retry;
    std::vector<int32_t> cand_pos;  // node -> its index among the current candidates, else -1
    const dic_progress* pg = nullptr;  // observer (may be null)

    bool cancelled() const { return pg && pg->cancelled && pg->cancelled(pg->user) != 0; }
    // progress::node_done (progress.rs:34-37): report a stored node, then
    // honour a cancellation request
    int node_done() const {
        if (pg && pg->node_done) pg->node_done(pg->user);
        if (cancelled()) {
            set_error("solve cancelled");
            return DIC_ERR_CANCELLED;
        }
        return DIC_OK;
    }

    // the node's trusted neighbours, in a buffer reused per host thread
    const std::vector<int64_t>& trusted(int64_t i, const uint8_t* skip) const {
        thread_local std::vector<int64_t> out;
        out.clear();
        for (const int32_t* q = t->full.begin(i); q != t->full.end(i); ++q)
            if (qok[*q] && (skip ? !skip[*q] : stored[*q] != 0)) out.push_back(*q);
        return out;
    }

    // precondition_warp of every listed node (each reads only solved nodes,
    // so they are independent), on the host workers for longer lists
    void warps(const std::vector<int64_t>& members, const uint8_t* skip, std::vector<double>& pa) const {
        const int64_t n = (int64_t)members.size();
        HostPool::get().run(n, n < 256 ? n : 64, [&](int64_t b, int64_t e) {
            for (int64_t k = b; k < e; ++k)
                initial_warp(nodes, &nodes[2 * members[k]], trusted(members[k], skip), czn.data(), P.data(), plen,
                             &pa[k * plen]);
        });
    }

    int list(const std::vector<int64_t>& nodes_in, const double* warps, double tol, std::vector<dic_subset_result>& r) {
        const int64_t n = (int64_t)nodes_in.size();
        ids.resize(n);
        p0.assign(12 * n, 0.0);
        for (int64_t k = 0; k < n; ++k) {
            ids[k] = (int32_t)nodes_in[k];
            for (int j = 0; j < plen; ++j) p0[12 * k + j] = warps[k * plen + j];
        }
        r.resize(n);
        dic_solve_params p = sp;
        p.tolerance = tol;
        const auto t0 = std::chrono::steady_clock::now();
        int st = solve_list(b, f, g, &p, ids.data(), n, p0.data(), r.data());
        list_seconds += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if (st) return st;
        n_solves += n;
        ++n_lists;
        for (int64_t k = 0; k < n; ++k)
            if (r[k].status) {
                set_error("mesh: node " + std::to_string(nodes_in[k]) + ": subset solve failed with status " +
                          std::to_string(r[k].status));
                return r[k].status;
            }
        return DIC_OK;
    }

    // solve_node (mesh.rs:1068-1106) for every listed node at once: the
    // preconditioned attempts in one list, then the zero-warp attempts of
    // those that missed quality_ok from a nonzero warp; kept[k] is the
    // attempt that passes, else the better correlated.
    int solve_nodes(const std::vector<int64_t>& nodes_in, const std::vector<double>& pa, double tol,
                    std::vector<dic_subset_result>& kept) {
        int st = list(nodes_in, pa.data(), tol, ra);
        if (st) return st;
        kept = ra;
        retry.clear();
        std::vector<int64_t> pos;
        for (size_t k = 0; k < nodes_in.size(); ++k) {
            if (ra[k].quality_ok) continue;
            bool zero = true;
            for (int j = 0; j < plen; ++j) zero = zero && pa[k * plen + j] == 0.0;
            if (zero) continue;
            retry.push_back(nodes_in[k]);
            pos.push_back((int64_t)k);
        }
        if (retry.empty()) return DIC_OK;
        const std::vector<double> zeros(retry.size() * plen, 0.0);
        st = list(retry, zeros.data(), tol, rb);
        if (st) return st;
        for (size_t q = 0; q < retry.size(); ++q)
            if (rb[q].quality_ok || rb[q].c_zncc > ra[pos[q]].c_zncc) kept[pos[q]] = rb[q];
        return DIC_OK;
    }

    // store_result (mesh.rs:2385-2404)
    void store(int64_t i, const dic_subset_result& r) {
        qok[i] = r.quality_ok != 0;
        czn[i] = std::fmax(r.c_zncc, 0.0);
        for (int k = 0; k < plen; ++k) P[i * plen + k] = r.p[k];
        disp[2 * i] = r.p[0];
        disp[2 * i + 1] = r.p[1];
        iters[i] = r.iterations;
        norms[i] = r.iterations > 0 ? r.norm : 0.0;
    }

    // stores the nodes in the given (ascending) order, each reported to the
    // observer, which may stop the solve after any of them
    int fold(const std::vector<int64_t>& nodes_in, const std::vector<dic_subset_result>& kept) {
        for (size_t k = 0; k < nodes_in.size(); ++k) {
            store(nodes_in[k], kept[k]);
            const int st = node_done();
            if (st) return st;
            stored[nodes_in[k]] = 1;
            heap.push({heap_key(czn[nodes_in[k]]), nodes_in[k]});
        }
        return DIC_OK;
    }

    // one node at a time (serial RG and the corrections)
    int solve_one(int64_t i, const std::vector<int64_t>& tr, double tol, dic_subset_result& r) {
        std::vector<double> pa(plen);
        initial_warp(nodes, &nodes[2 * i], tr, czn.data(), P.data(), plen, pa.data());
        std::vector<dic_subset_result> kept;
        const int st = solve_nodes({i}, pa, tol, kept);
        if (!st) r = kept[0];
        return st;
    }

    // solve_neighbours_from (mesh.rs:1112-1152)
    int rg_from(int64_t cur) {
        for (const int32_t* q = t->edges().begin(cur); q != t->edges().end(cur); ++q) {
            const int64_t nb = *q;
            if (stored[nb]) continue;
            dic_subset_result r;
            int st = solve_one(nb, trusted(nb, nullptr), sp.tolerance, r);
            if (!st) st = fold({nb}, {r});
            if (st) return st;
        }
        return DIC_OK;
    }

    // expand_parallel (mesh.rs:1169-1252): the unsolved edge neighbours of
    // the roots, greedily coloured over the full adjacency (ascending
    // nodes, lowest free colour); one colour at a time, every node from the
    // trusted set frozen before the colour, folded in ascending order.
    int expand(const std::vector<int64_t>& roots) {
        // each unsolved edge neighbour once (marked in cand_pos while
        // gathering), then in ascending order
        if (cand_pos.size() != (size_t)t->n_nodes) cand_pos.assign(t->n_nodes, -1);
        std::vector<int64_t> cand;
        for (int64_t r : roots)
            for (const int32_t* q = t->edges().begin(r); q != t->edges().end(r); ++q)
                if (!stored[*q] && cand_pos[*q] == -1) {
                    cand_pos[*q] = -2;
                    cand.push_back(*q);
                }
        std::sort(cand.begin(), cand.end());
        if (cand.empty()) return DIC_OK;
        std::vector<int32_t> colour(cand.size(), -1);
        int32_t n_col = 0;
        std::vector<uint8_t> used;
        for (size_t k = 0; k < cand.size(); ++k) cand_pos[cand[k]] = (int32_t)k;
        for (size_t k = 0; k < cand.size(); ++k) {
            // lowest colour no coloured neighbour has: a bit mask while the
            // colours fit in 64, else a flag per colour
            uint64_t mask = 0;
            bool wide = n_col > 63;
            if (wide) used.assign(n_col + 1, 0);
            for (const int32_t* q = t->full.begin(cand[k]); q != t->full.end(cand[k]); ++q) {
                const int32_t j = cand_pos[*q];
                if (j >= 0 && colour[j] >= 0) {
                    if (wide)
                        used[colour[j]] = 1;
                    else
                        mask |= (uint64_t)1 << colour[j];
                }
            }
            int32_t c = 0;
            if (wide) {
                while (used[c]) ++c;
            } else {
                c = __builtin_ctzll(~mask);  // n_col <= 63 keeps a zero bit
            }
            colour[k] = c;
            n_col = std::max(n_col, c + 1);
        }
        for (size_t k = 0; k < cand.size(); ++k) cand_pos[cand[k]] = -1;
        std::vector<int64_t> members;
        std::vector<double> pa;
        std::vector<dic_subset_result> kept;
        for (int32_t c = 0; c < n_col; ++c) {
            members.clear();
            for (size_t k = 0; k < cand.size(); ++k)
                if (colour[k] == c) members.push_back(cand[k]);
            pa.assign(members.size() * plen, 0.0);
            warps(members, nullptr, pa);
            int st = solve_nodes(members, pa, sp.tolerance, kept);
            if (!st) st = fold(members, kept);
            if (st) return st;
        }
        return DIC_OK;
    }

    // flow_calc (mesh.rs:1936-1966): cosine between the node's displacement
    // and the mean of its finite, non-excluded full neighbours'; -1 when
    // undefined
    double flow(int64_t i, const uint8_t* exclude) const {
        const double d0 = disp[2 * i], d1 = disp[2 * i + 1];
        if (!std::isfinite(d0) || !std::isfinite(d1)) return -1.0;
        double sx = 0.0, sy = 0.0;
        int64_t n = 0;
        for (const int32_t* q = t->full.begin(i); q != t->full.end(i); ++q) {
            const int64_t nb = *q;
            if ((exclude && exclude[nb]) || !std::isfinite(disp[2 * nb]) || !std::isfinite(disp[2 * nb + 1])) continue;
            sx += disp[2 * nb];
            sy += disp[2 * nb + 1];
            ++n;
        }
        if (n == 0) return -1.0;
        const double vx = sx / (double)n, vy = sy / (double)n;
        const double vm = std::sqrt(vx * vx + vy * vy), dm = std::sqrt(d0 * d0 + d1 * d1);
        if (vm < 2.220446049250313e-16 || dm < 2.220446049250313e-16) return -1.0;
        return (d0 * vx + d1 * vy) / (dm * vm);
    }

    // The corrections visit their nodes one at a time, each preconditioned
    // from its neighbours outside `skip` (the nodes still to come, and for
    // the first lane those left unimproved) after every earlier node has
    // been settled.  A node reads only its neighbours, so it depends on the
    // earlier nodes of the sequence among its neighbours and on nothing
    // else: wave w holds the nodes whose latest such dependency is in wave
    // w - 1.  The nodes of one wave are pairwise non-adjacent and are solved
    // in one list; settling them in sequence order then gives exactly the
    // one-at-a-time result.  settle(j, result) updates the state and skip.
    template <class Settle>
    int in_waves(const std::vector<int64_t>& seq, const std::vector<uint8_t>& skip, Settle settle) {
        const int64_t n = t->n_nodes;
        std::vector<int64_t> pos(n, -1);
        for (size_t i = 0; i < seq.size(); ++i) pos[seq[i]] = (int64_t)i;
        std::vector<int32_t> wave(seq.size(), 0);
        int32_t n_waves = 0;
        for (size_t i = 0; i < seq.size(); ++i) {
            int32_t w = 0;
            for (const int32_t* q = t->full.begin(seq[i]); q != t->full.end(seq[i]); ++q)
                if (pos[*q] >= 0 && pos[*q] < (int64_t)i) w = std::max(w, wave[pos[*q]] + 1);
            wave[i] = w;
            n_waves = std::max(n_waves, w + 1);
        }
        std::vector<int64_t> members;
        std::vector<double> pa;
        std::vector<dic_subset_result> kept;
        for (int32_t w = 0; w < n_waves; ++w) {
            // the corrections report no nodes (mesh.rs:1265-1414), but a
            // cancellation is honoured at every list boundary
            if (cancelled()) {
                set_error("solve cancelled");
                return DIC_ERR_CANCELLED;
            }
            members.clear();
            for (size_t i = 0; i < seq.size(); ++i)
                if (wave[i] == w) members.push_back(seq[i]);
            pa.assign(members.size() * plen, 0.0);
            warps(members, skip.data(), pa);
            const int st = solve_nodes(members, pa, sp.tolerance, kept);
            if (st) return st;
            for (size_t k = 0; k < members.size(); ++k) settle(members[k], kept[k]);
        }
        return DIC_OK;
    }

    // corrections (mesh.rs:1265-1414)
    int correct() {
        const int64_t n = t->n_nodes;
        // lane 1: C_ZNCC outliers (below LQ - 2.5 IQR), worst first; kept only
        // when better and quality_ok
        {
            std::vector<double> v(czn);
            double lq, uq;
            quartiles(v, lq, uq);
            const double fence = lq - 2.5 * (uq - lq);
            std::vector<int64_t> order;
            for (int64_t i = 0; i < n; ++i)
                if (czn[i] < fence) order.push_back(i);
            std::stable_sort(order.begin(), order.end(),
                             [&](int64_t a, int64_t c) { return total_key(czn[a]) < total_key(czn[c]); });
            std::vector<uint8_t> skip(n, 0);  // later outliers and the unimproved
            for (int64_t j : order) skip[j] = 1;
            const int st = in_waves(order, skip, [&](int64_t j, const dic_subset_result& r) {
                if (r.c_zncc > czn[j] && r.quality_ok) {
                    store(j, r);
                    stored[j] = 1;
                    skip[j] = 0;
                }
            });
            if (st) return st;
        }
        // lane 2: flow outliers (below LQ - 1.5 IQR of the mesh) and
        // magnitude outliers (above UQ + 4 IQR of an element's
        // neighbourhood), mildest magnitude first, ties by node
        HostPool& pool = HostPool::get();
        std::vector<double> fl(n), r(n);
        pool.run(n, 4096, [&](int64_t b, int64_t e) {
            for (int64_t i = b; i < e; ++i) {
                fl[i] = flow(i, nullptr);
                r[i] = std::sqrt(disp[2 * i] * disp[2 * i] + disp[2 * i + 1] * disp[2 * i + 1]);
            }
        });
        double flq, fuq;
        {
            std::vector<double> v(fl);
            quartiles(v, flq, fuq);
        }
        const double fiqr = fuq - flq;
        std::vector<uint8_t> rflag(n, 0), flagged(n, 0);
        for (int64_t i = 0; i < n; ++i)
            if (fl[i] < flq - 1.5 * fiqr) flagged[i] = 1;
        // per element (independent of each other, on the host workers): the
        // nodes of its neighbourhood above its fence are flagged; a node
        // flagged by several elements gets the same mark from each
        pool.run(t->n_el, 1024, [&](int64_t b, int64_t e1) {
            std::vector<int32_t> local;
            std::vector<double> lr;
            for (int64_t e = b; e < e1; ++e) {
                local.clear();
                for (int k = 0; k < 3; ++k)
                    local.insert(local.end(), t->full.begin(t->at(e, k)), t->full.end(t->at(e, k)));
                std::sort(local.begin(), local.end());
                local.erase(std::unique(local.begin(), local.end()), local.end());
                lr.clear();
                for (int32_t q : local) lr.push_back(r[q]);
                double lq, uq;
                quartiles(lr, lq, uq);
                if (lr.empty()) continue;
                const double fence = uq + 4.0 * (uq - lq);
                for (int32_t q : local)
                    if (r[q] > fence) __atomic_store_n(&rflag[q], (uint8_t)1, __ATOMIC_RELAXED);
            }
        });
        std::vector<int64_t> ids;
        for (int64_t i = 0; i < n; ++i)
            if (flagged[i] || rflag[i]) {
                flagged[i] = 1;
                ids.push_back(i);
            }
        std::stable_sort(ids.begin(), ids.end(), [&](int64_t a, int64_t c) { return total_key(r[a]) < total_key(r[c]); });
        // exempt: not a magnitude outlier, and flowing with its unflagged
        // neighbours (above LQ - 2.5 IQR)
        std::vector<int64_t> active;
        for (int64_t j : ids)
            if (rflag[j] || !(flow(j, flagged.data()) > flq - 2.5 * fiqr)) active.push_back(j);
        // always stored: each node preconditioned from neighbours outside
        // itself and the nodes after it
        std::vector<uint8_t> later(n, 0);
        for (int64_t j : active) later[j] = 1;
        return in_waves(active, later, [&](int64_t j, const dic_subset_result& res) {
            store(j, res);
            stored[j] = 1;
            later[j] = 0;
        });
        return DIC_OK;
    }
};

int check_mesh_args(int64_t n_nodes, const double* nodes, int64_t n_el, int mesh_order, const int64_t* el) {
    if (!nodes || !el || n_nodes < 1 || n_el < 0) {
        set_error("invalid input: mesh needs nodes and elements");
        return DIC_ERR_INVALID_INPUT;
    }
    if (mesh_order != 1 && mesh_order != 2) {
        set_error("invalid input: mesh_order must be 1 or 2");
        return DIC_ERR_INVALID_INPUT;
    }
    if (n_nodes > (int64_t)1 << 30) {
        set_error("invalid input: too many nodes");
        return DIC_ERR_INVALID_INPUT;
    }
    const int cols = mesh_order == 1 ? 3 : 6;
    for (int64_t k = 0; k < n_el * cols; ++k)
        if (el[k] < 0 || el[k] >= n_nodes) {
            set_error("invalid input: element node index out of range");
            return DIC_ERR_INVALID_INPUT;
        }
    return DIC_OK;
}

}  // namespace

}  // namespace geodic

using namespace geodic;

extern "C" {

int dic_mesh_element_strains(int64_t n_nodes, const double* nodes, int64_t n_elements, int32_t mesh_order,
                             const int64_t* elements, const double* displacements, double* warps) {
    int st = check_mesh_args(n_nodes, nodes, n_elements, mesh_order, elements);
    if (st) return st;
    if (!displacements || !warps) {
        set_error("dic_mesh_element_strains: NULL array");
        return DIC_ERR_INVALID_INPUT;
    }
    Topology t;
    t.n_nodes = n_nodes;
    t.n_el = n_elements;
    t.order = mesh_order;
    t.cols = mesh_order == 1 ? 3 : 6;
    t.el = elements;
    std::string err;
    st = strains(nodes, t, displacements, warps, err);
    if (st) set_error(err);
    return st;
}

int dic_mesh_solve(dic_engine* d, const dic_image* f, const dic_image* g, int64_t n_nodes, const double* nodes,
                   int64_t n_elements, int32_t mesh_order, const int64_t* elements, const int32_t* tpl_index,
                   int32_t n_templates, const int64_t* tpl_start, const int16_t* tpl_xy, const dic_mesh_params* prm,
                   const dic_progress* progress, dic_mesh_output* out) {
    if (!d || !f || !g || !prm || !out || !out->p || !out->c_zncc || !out->iterations || !out->norms ||
        !out->quality_ok || !out->warps) {
        set_error("dic_mesh_solve: NULL argument");
        return DIC_ERR_INVALID_INPUT;
    }
    int st = check_mesh_args(n_nodes, nodes, n_elements, mesh_order, elements);
    if (st) return st;
    // p_len = 6 * subset_order (mesh.rs:636): order 0 leaves no warp
    // columns to store, so a mesh solves order 1 or 2
    if (prm->subset_order != 1 && prm->subset_order != 2) {
        set_error("invalid input: a mesh solves subset_order 1 or 2");
        return DIC_ERR_INVALID_INPUT;
    }
    if (prm->preconditioning != DIC_PRECOND_RG && prm->preconditioning != DIC_PRECOND_LAYER_RG) {
        set_error("invalid input: preconditioning must be RG or layer-RG");
        return DIC_ERR_INVALID_INPUT;
    }
    if (prm->solve.want_history) {
        set_error("invalid input: a mesh solve records no history");
        return DIC_ERR_INVALID_INPUT;
    }
    const int plen = 6 * prm->subset_order;
    std::vector<int32_t> orders(n_nodes, prm->subset_order);
    Topology t;
    t.n_nodes = n_nodes;
    t.n_el = n_elements;
    t.order = mesh_order;
    t.cols = mesh_order == 1 ? 3 : 6;
    t.el = elements;
    // the adjacency is built on a helper thread while the batch is created
    // and the seed node solved, which do not read it; joined before the
    // propagation (and on every way out)
    std::thread topo([&t] { build_topology(t); });
    struct Join {
        std::thread& th;
        ~Join() {
            if (th.joinable()) th.join();
        }
    } join_topo{topo};
    dic_batch* db = nullptr;
    st = dic_batch_create(d, n_nodes, nodes, orders.data(), tpl_index, n_templates, tpl_start, tpl_xy, &db);
    if (st) return st;
    Batch* b = &db->b;
    Engine* e = &d->e;

    Run R;
    R.t = &t;
    R.nodes = nodes;
    R.b = b;
    R.f = &f->im;
    R.g = &g->im;
    R.sp = prm->solve;
    R.plen = plen;
    R.stored.assign(n_nodes, 0);
    R.propagated.assign(n_nodes, 0);
    R.qok.assign(n_nodes, 0);
    R.czn.assign(n_nodes, 0.0);
    R.P.assign((size_t)n_nodes * plen, 0.0);
    R.disp.assign(2 * n_nodes, 0.0);
    R.norms.assign(n_nodes, 0.0);
    R.iters.assign(n_nodes, 0);
    R.pg = progress;
    if (progress && progress->mesh_start) progress->mesh_start(progress->user, n_nodes);

    // seed node (mesh.rs:659-672): the node nearest the seed coordinate,
    // the user's warp, the seed tolerance
    int64_t seed = 0;
    {
        double bd = 0.0;
        for (int64_t i = 0; i < n_nodes; ++i) {
            const double dx = nodes[2 * i] - prm->seed_coord[0], dy = nodes[2 * i + 1] - prm->seed_coord[1];
            const double dd = dx * dx + dy * dy;
            if (i == 0 || dd < bd) {
                seed = i;
                bd = dd;
            }
        }
    }
    {
        std::lock_guard<std::mutex> lk(e->mu);
        st = hipSetDevice(e->device) == hipSuccess ? DIC_OK : DIC_ERR_DEVICE;
        if (!st) st = validate_solve(b, &f->im, &g->im, &prm->solve);
        std::vector<dic_subset_result> kept;
        if (!st) {
            std::vector<double> pa(prm->seed_warp, prm->seed_warp + plen);
            st = R.solve_nodes({seed}, pa, prm->seed_tolerance, kept);
        }
        if (!st) st = R.fold({seed}, kept);
        topo.join();
        if (!st) {
            R.propagated[seed] = 1;
            if (prm->preconditioning == DIC_PRECOND_RG) {
                // serial cascade, highest C_ZNCC first (mesh.rs:678-700)
                st = R.rg_from(seed);
                while (!st && !R.heap.empty()) {
                    const int64_t cur = R.heap.top().second;
                    R.heap.pop();
                    if (R.propagated[cur]) continue;
                    R.propagated[cur] = 1;
                    st = R.rg_from(cur);
                }
            } else {
                // rounds over an epsilon band of roots (mesh.rs:701-752); a
                // worker is a compute unit unless given
                const int64_t workers = prm->workers > 0 ? prm->workers : e->num_cus;
                const int64_t root_cap = std::max<int64_t>(prm->batch_factor, 1) * std::max<int64_t>(workers, 1);
                st = R.expand({seed});
                std::vector<int64_t> roots;
                while (!st && !R.heap.empty()) {
                    const auto top = R.heap.top();
                    R.heap.pop();
                    if (R.propagated[top.second]) continue;
                    R.propagated[top.second] = 1;
                    roots.assign(1, top.second);
                    double top_val;
                    std::memcpy(&top_val, &top.first, 8);
                    const double threshold = top_val * (1.0 - prm->root_rel_eps);
                    while ((int64_t)roots.size() < root_cap && !R.heap.empty()) {
                        double kv;
                        std::memcpy(&kv, &R.heap.top().first, 8);
                        if (!(kv >= threshold)) break;
                        const int64_t r = R.heap.top().second;
                        R.heap.pop();
                        if (R.propagated[r]) continue;
                        R.propagated[r] = 1;
                        roots.push_back(r);
                    }
                    st = R.expand(roots);
                }
            }
        }
        if (!st) st = R.correct();
    }
    dic_batch_destroy(db);
    if (st) return st;

    int64_t n_bad = 0;
    for (int64_t i = 0; i < n_nodes; ++i) {
        n_bad += !R.qok[i];
        for (int k = 0; k < 12; ++k) out->p[12 * i + k] = k < plen ? R.P[i * plen + k] : 0.0;
        out->c_zncc[i] = R.czn[i];
        out->iterations[i] = R.iters[i];
        out->norms[i] = R.norms[i];
        out->quality_ok[i] = R.qok[i];
    }
    out->seed_node = seed;
    out->n_node_solves = R.n_solves;
    out->n_lists = R.n_lists;
    out->n_unsolvable = n_bad;
    out->list_seconds = R.list_seconds;
    // the quality gate (mesh.rs:754-777), then the element strains and the
    // compatibility check (mesh.rs:798-803, 1417-1433)
    std::string err;
    const int sst = strains(nodes, t, R.disp.data(), out->warps, err);
    if (!prm->override_active && n_bad > 0) {
        set_error("invalid input: mesh unsolvable: " + std::to_string(n_bad) +
                  " subset(s) failed to reach quality_ok after corrections");
        return DIC_ERR_INVALID_INPUT;
    }
    if (sst) {
        set_error(err);
        return sst;
    }
    for (int64_t el = 0; el < n_elements; ++el) {
        const double* w = out->warps + 12 * el;
        const double det = (1.0 + w[2]) * (1.0 + w[5]) - w[3] * w[4];
        if (det <= 0.0) {
            double cx = 0.0, cy = 0.0;
            for (int k = 0; k < 3; ++k) {
                cx += nodes[2 * t.at(el, k)] / 3.0;
                cy += nodes[2 * t.at(el, k) + 1] / 3.0;
            }
            char buf[160];
            std::snprintf(buf, sizeof buf, "invalid input: mesh compatibility violated at element %lld near (%.1f,%.1f): "
                          "det(J)=%.4f", (long long)el, cx, cy, det);
            set_error(buf);
            return DIC_ERR_INVALID_INPUT;
        }
    }
    return DIC_OK;
}

}  // extern "C"
