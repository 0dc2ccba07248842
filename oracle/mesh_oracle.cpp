// mesh_oracle.cpp — FP64 CPU restatement of geopyv-dev's mesh solve
// (src/mesh.rs), the caller of the subset-correlation hot path.
// TEST INFRASTRUCTURE ONLY: the correctness checker for the engine's mesh
// orchestration and the timed CPU baseline of the mesh mode.  Nothing in
// the shipped library links it.
//
// Each function follows the mesh.rs item cited beside it, step by step and
// in the same visit orders, so that the node solve order (and therefore
// every preconditioning warp) is the reference's.  Node solves go through
// orc_subset_solve (dic_oracle.cpp).  Layer-RG colour rounds run their
// nodes in parallel with OpenMP, as the reference runs them with rayon
// (mesh.rs:1224-1233); the fold after each colour is serial and ascending.
//
// Containers: nodes (N, 2) row-major [x, y]; elements (E, cols) int64,
// cols 3 (mesh order 1) or 6 (order 2: c0, c1, c2, m01, m12, m20); warps
// (N, p_len) where p_len = 6 * subset_order (mesh.rs:636).

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <queue>
#include <set>
#include <string>
#include <unordered_set>
#include <utility>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "oracle_types.h"

namespace {

enum { ST_OK = 0, ST_INVALID_INPUT = 3, ST_MESH_GENERATION = 4 };

struct Mesh {
    const double* nodes;
    int64_t n_nodes;
    const int64_t* el;
    int64_t n_el;
    int cols;
    int order;
    int64_t at(int64_t e, int k) const { return el[e * cols + k]; }
};

// connectivity (mesh.rs:1705-1764): brute-force element scan, sorted unique.
std::vector<int64_t> connectivity(const Mesh& m, int64_t idx, bool full) {
    std::set<int64_t> result;
    if (m.order == 1 || full) {
        for (int64_t e = 0; e < m.n_el; ++e) {
            bool found = false;
            for (int k = 0; k < m.cols; ++k)
                if (m.at(e, k) == idx) found = true;
            if (found)
                for (int k = 0; k < m.cols; ++k)
                    if (m.at(e, k) != idx) result.insert(m.at(e, k));
        }
    } else {
        // column position -> the two edge-adjacent positions (mesh.rs:1745-1752)
        static const int adds[6][2] = {{3, 5}, {3, 4}, {4, 5}, {0, 1}, {1, 2}, {0, 2}};
        for (int64_t e = 0; e < m.n_el; ++e)
            for (int col = 0; col < m.cols; ++col) {
                if (m.at(e, col) != idx) continue;
                for (int a = 0; a < 2; ++a) {
                    const int64_t nb = m.at(e, adds[col][a]);
                    if (nb != idx) result.insert(nb);
                }
            }
    }
    return std::vector<int64_t>(result.begin(), result.end());
}

// Adjacency::build (mesh.rs:1853-1868): full and edge lists per node.
struct Adjacency {
    std::vector<std::vector<int64_t>> full, edge;
    explicit Adjacency(const Mesh& m) : full(m.n_nodes), edge(m.n_nodes) {
        // per-node element index first, as build_node_element_index does
        std::vector<std::vector<int64_t>> inc(m.n_nodes);
        for (int64_t e = 0; e < m.n_el; ++e)
            for (int k = 0; k < m.cols; ++k) inc[m.at(e, k)].push_back(e);
        static const int adds[6][2] = {{3, 5}, {3, 4}, {4, 5}, {0, 1}, {1, 2}, {0, 2}};
        for (int64_t i = 0; i < m.n_nodes; ++i) {
            std::set<int64_t> f, g;
            for (int64_t e : inc[i])
                for (int k = 0; k < m.cols; ++k) {
                    if (m.at(e, k) != i) f.insert(m.at(e, k));
                    if (m.order == 2 && m.at(e, k) == i)
                        for (int a = 0; a < 2; ++a)
                            if (m.at(e, adds[k][a]) != i) g.insert(m.at(e, adds[k][a]));
                }
            full[i].assign(f.begin(), f.end());
            if (m.order == 1)
                edge[i] = full[i];
            else
                edge[i].assign(g.begin(), g.end());
        }
    }
    const std::vector<int64_t>& get(int64_t i, bool f) const { return f ? full[i] : edge[i]; }
};

// find_seed_node (mesh.rs:1682-1697): first node of least squared distance.
int64_t find_seed_node(const double* nodes, int64_t n, double x, double y) {
    int64_t best = 0;
    double bd = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        const double dx = nodes[2 * i] - x, dy = nodes[2 * i + 1] - y;
        const double d = dx * dx + dy * dy;
        if (i == 0 || d < bd) {
            best = i;
            bd = d;
        }
    }
    return best;
}

// percentile (mesh.rs:1919-1930): linear interpolation on a sorted slice.
double percentile(const std::vector<double>& s, double p) {
    const size_t n = s.size();
    if (n == 0) return 0.0;
    const double pos = p / 100.0 * ((double)n - 1.0);
    const size_t lo = (size_t)std::floor(pos);
    const size_t hi = std::min(lo + 1, n - 1);
    const double frac = pos - (double)lo;
    return s[lo] + frac * (s[hi] - s[lo]);
}

// corr_1d (mesh.rs:1907-1917): indices below LQ - 2.5 IQR of the finite values.
std::vector<int64_t> corr_1d(const double* v, int64_t n) {
    std::vector<double> s;
    for (int64_t i = 0; i < n; ++i)
        if (std::isfinite(v[i])) s.push_back(v[i]);
    std::sort(s.begin(), s.end());
    const double lq = percentile(s, 25.0), uq = percentile(s, 75.0);
    const double fence = lq - 2.5 * (uq - lq);
    std::vector<int64_t> out;
    for (int64_t i = 0; i < n; ++i)
        if (v[i] < fence) out.push_back(i);
    return out;
}

// f64::total_cmp as a strict weak order (NaN by sign bit at either end).
bool total_less(double a, double b) {
    int64_t x, y;
    std::memcpy(&x, &a, 8);
    std::memcpy(&y, &b, 8);
    x ^= (int64_t)((uint64_t)(x >> 63) >> 1);
    y ^= (int64_t)((uint64_t)(y >> 63) >> 1);
    return x < y;
}

// flow_calc (mesh.rs:1936-1966): cosine between a node's displacement and
// the mean displacement of its finite, non-excluded full neighbours.
double flow_calc(const Adjacency& adj, int64_t idx, const double* disp, const std::vector<char>& exclude,
                 const double* override_disp) {
    const double d0 = override_disp ? override_disp[0] : disp[2 * idx];
    const double d1 = override_disp ? override_disp[1] : disp[2 * idx + 1];
    if (!std::isfinite(d0) || !std::isfinite(d1)) return -1.0;
    std::vector<int64_t> valid;
    for (int64_t nb : adj.get(idx, true))
        if (!exclude[nb] && std::isfinite(disp[2 * nb]) && std::isfinite(disp[2 * nb + 1])) valid.push_back(nb);
    if (valid.empty()) return -1.0;
    const double n = (double)valid.size();
    double sx = 0.0, sy = 0.0;
    for (int64_t nb : valid) sx += disp[2 * nb];
    for (int64_t nb : valid) sy += disp[2 * nb + 1];
    const double vx = sx / n, vy = sy / n;
    const double v_mag = std::sqrt(vx * vx + vy * vy);
    const double d_mag = std::sqrt(d0 * d0 + d1 * d1);
    const double eps = 2.220446049250313e-16;
    if (v_mag < eps || d_mag < eps) return -1.0;
    return (d0 * vx + d1 * vy) / (d_mag * v_mag);
}

// flow_stats (mesh.rs:2059-2076): flow per node, LQ, IQR, ids below LQ - 1.5 IQR.
void flow_stats(const Adjacency& adj, const double* disp, int64_t n, std::vector<double>& flow,
                std::vector<int64_t>& ids, double& lq, double& iqr) {
    std::vector<char> none(n, 0);
    flow.assign(n, 0.0);
    for (int64_t i = 0; i < n; ++i) flow[i] = flow_calc(adj, i, disp, none, nullptr);
    std::vector<double> s;
    for (double f : flow)
        if (std::isfinite(f)) s.push_back(f);
    std::sort(s.begin(), s.end());
    lq = percentile(s, 25.0);
    const double uq = percentile(s, 75.0);
    iqr = uq - lq;
    const double fence = lq - 1.5 * iqr;
    ids.clear();
    for (int64_t i = 0; i < n; ++i)
        if (flow[i] < fence) ids.push_back(i);
}

// r_stats (mesh.rs:2079-2128): per element, the union of its three corners'
// full neighbourhoods; finite magnitudes above UQ + 4 IQR are flagged.
void r_stats(const Adjacency& adj, const Mesh& m, const double* disp, std::vector<double>& r,
             std::vector<char>& flagged) {
    const int64_t n = m.n_nodes;
    r.assign(n, 0.0);
    for (int64_t i = 0; i < n; ++i) r[i] = std::sqrt(disp[2 * i] * disp[2 * i] + disp[2 * i + 1] * disp[2 * i + 1]);
    flagged.assign(n, 0);
    for (int64_t e = 0; e < m.n_el; ++e) {
        std::set<int64_t> local;
        for (int k = 0; k < 3; ++k)
            for (int64_t nb : adj.get(m.at(e, k), true)) local.insert(nb);
        std::vector<double> lr;
        for (int64_t nb : local)
            if (std::isfinite(r[nb])) lr.push_back(r[nb]);
        if (lr.empty()) continue;
        std::sort(lr.begin(), lr.end());
        const double lq = percentile(lr, 25.0), uq = percentile(lr, 75.0);
        const double fence = uq + 4.0 * (uq - lq);
        for (int64_t nb : local)
            if (r[nb] > fence) flagged[nb] = 1;
    }
}

// project_warp (mesh.rs:2135-2150): Taylor shift of a warp by (dx, dy).
void project_warp(const double* p, int plen, double dx, double dy, double* out) {
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

// ols_solve (mesh.rs:2230-2240): normal equations (A^T A) x = A^T y with
// the nalgebra-style inverse; false when A^T A is singular.
bool ols_solve(const std::vector<std::vector<double>>& rows, const std::vector<double>& y, std::vector<double>& x) {
    const size_t n = rows.size(), mm = rows[0].size();
    std::vector<double> ata(mm * mm, 0.0), aty(mm, 0.0), inv(mm * mm);
    for (size_t i = 0; i < mm; ++i)
        for (size_t j = 0; j < mm; ++j) {
            double s = 0.0;
            for (size_t r = 0; r < n; ++r) s += rows[r][i] * rows[r][j];
            ata[i * mm + j] = s;
        }
    for (size_t i = 0; i < mm; ++i) {
        double s = 0.0;
        for (size_t r = 0; r < n; ++r) s += rows[r][i] * y[r];
        aty[i] = s;
    }
    if (!orc_inverse(ata.data(), (int)mm, inv.data())) return false;
    x.assign(mm, 0.0);
    for (size_t i = 0; i < mm; ++i) {
        double s = 0.0;
        for (size_t k = 0; k < mm; ++k) s += inv[i * mm + k] * aty[k];
        x[i] = s;
    }
    return true;
}

// affine_field_fit (mesh.rs:2181-2227): regress the trusted neighbours'
// (u, v) on their offsets from the target (quadratic for order 2 with at
// least six neighbours).
bool affine_field_fit(const double* nodes, const double* target, const int64_t* trusted, int64_t nt,
                      const double* p, int plen, double* out) {
    const bool quad = plen >= 12 && nt >= 6;
    std::vector<std::vector<double>> rows;
    std::vector<double> u, v;
    for (int64_t k = 0; k < nt; ++k) {
        const int64_t i = trusted[k];
        const double dx = nodes[2 * i] - target[0], dy = nodes[2 * i + 1] - target[1];
        if (quad)
            rows.push_back({1.0, dx, dy, dx * dx, dx * dy, dy * dy});
        else
            rows.push_back({1.0, dx, dy});
        u.push_back(p[i * plen + 0]);
        v.push_back(p[i * plen + 1]);
    }
    std::vector<double> cu, cv;
    if (!ols_solve(rows, u, cu)) return false;
    if (!ols_solve(rows, v, cv)) return false;
    for (int k = 0; k < plen; ++k) out[k] = 0.0;
    out[0] = cu[0];
    out[1] = cv[0];
    out[2] = cu[1];
    out[3] = cv[1];
    out[4] = cu[2];
    out[5] = cv[2];
    if (quad) {
        out[6] = 2.0 * cu[3];
        out[7] = 2.0 * cv[3];
        out[8] = cu[4];
        out[9] = cv[4];
        out[10] = 2.0 * cu[5];
        out[11] = 2.0 * cv[5];
    }
    return true;
}

// precondition_warp (mesh.rs:2152-2179): affine fit over >= 3 trusted,
// else the Taylor projection of the best-correlated trusted neighbour (the
// last of equals, as Iterator::max_by keeps), else zero.
void precondition_warp(const double* nodes, const double* target, const int64_t* trusted, int64_t nt,
                       const double* c_zncc, const double* p, int plen, double* out) {
    if (nt >= 3 && affine_field_fit(nodes, target, trusted, nt, p, plen, out)) return;
    if (nt > 0) {
        int64_t best = trusted[0];
        for (int64_t k = 1; k < nt; ++k)
            if (!(c_zncc[trusted[k]] < c_zncc[best])) best = trusted[k];
        project_warp(p + best * plen, plen, target[0] - nodes[2 * best], target[1] - nodes[2 * best + 1], out);
        return;
    }
    for (int k = 0; k < plen; ++k) out[k] = 0.0;
}

// greedy_graph_colour (mesh.rs:1882-1902): ascending candidates, lowest
// colour unused by already-coloured full neighbours among the candidates.
std::vector<int64_t> greedy_colour(const Adjacency& adj, const std::vector<int64_t>& cand) {
    std::vector<int64_t> colour(cand.size(), -1);
    for (size_t i = 0; i < cand.size(); ++i) {
        std::vector<char> used;
        for (int64_t nb : adj.get(cand[i], true)) {
            auto it = std::lower_bound(cand.begin(), cand.end(), nb);
            if (it == cand.end() || *it != nb) continue;
            const int64_t cn = colour[it - cand.begin()];
            if (cn < 0) continue;
            if ((size_t)cn >= used.size()) used.resize(cn + 1, 0);
            used[cn] = 1;
        }
        int64_t k = 0;
        while ((size_t)k < used.size() && used[k]) ++k;
        colour[i] = k;
    }
    return colour;
}

// element_area (mesh.rs:1468-1481)
double element_area(const double* nodes, const Mesh& m, int64_t e) {
    const int64_t i0 = m.at(e, 0), i1 = m.at(e, 1), i2 = m.at(e, 2);
    const double x0 = nodes[2 * i0], y0 = nodes[2 * i0 + 1];
    const double x1 = nodes[2 * i1], y1 = nodes[2 * i1 + 1];
    const double x2 = nodes[2 * i2], y2 = nodes[2 * i2 + 1];
    return 0.5 * ((x1 - x0) * (y2 - y0) - (x2 - x0) * (y1 - y0));
}

// element_strains (mesh.rs:1552-1665) with shape_function (1488-1515) and
// local_coordinates (1519-1544) at the centroid.
int element_strains(const double* nodes, const Mesh& m, const double* disp, double* warps) {
    static const double N1[3] = {1.0 / 3.0, 1.0 / 3.0, 1.0 / 3.0};
    static const double dN1[2][3] = {{1.0, 0.0, -1.0}, {0.0, 1.0, -1.0}};
    static const double N2[6] = {-1.0 / 9.0, -1.0 / 9.0, -1.0 / 9.0, 4.0 / 9.0, 4.0 / 9.0, 4.0 / 9.0};
    static const double dN2[2][6] = {{1.0 / 3.0, 0.0, -1.0 / 3.0, 4.0 / 3.0, -4.0 / 3.0, 0.0},
                                     {0.0, 1.0 / 3.0, -1.0 / 3.0, 4.0 / 3.0, 0.0, -4.0 / 3.0}};
    static const double d2N2[3][6] = {
        {4.0, 0.0, 4.0, 0.0, 0.0, -8.0}, {0.0, 0.0, 4.0, 4.0, -4.0, -4.0}, {0.0, 4.0, 4.0, 0.0, -8.0, 0.0}};
    const int ns = m.order == 1 ? 3 : 6;
    const double* Nf = m.order == 1 ? N1 : N2;
    for (int64_t e = 0; e < m.n_el; ++e) {
        double* w = warps + 12 * e;
        for (int k = 0; k < 12; ++k) w[k] = 0.0;
        double x[6][2], u[6][2];
        for (int k = 0; k < ns; ++k) {
            const int64_t ni = m.at(e, k);
            x[k][0] = nodes[2 * ni];
            x[k][1] = nodes[2 * ni + 1];
            u[k][0] = disp[2 * ni];
            u[k][1] = disp[2 * ni + 1];
        }
        double s0 = 0.0, s1 = 0.0;
        for (int k = 0; k < ns; ++k) s0 += Nf[k] * u[k][0];
        for (int k = 0; k < ns; ++k) s1 += Nf[k] * u[k][1];
        w[0] = s0;
        w[1] = s1;
        double jx[2][2], ju[2][2];
        for (int r = 0; r < 2; ++r)
            for (int c = 0; c < 2; ++c) {
                double a = 0.0, b = 0.0;
                for (int k = 0; k < ns; ++k) a += (m.order == 1 ? dN1[r][k] : dN2[r][k]) * x[k][c];
                for (int k = 0; k < ns; ++k) b += (m.order == 1 ? dN1[r][k] : dN2[r][k]) * u[k][c];
                jx[r][c] = a;
                ju[r][c] = b;
            }
        const double det = jx[0][0] * jx[1][1] - jx[0][1] * jx[1][0];
        if (std::fabs(det) < 2.220446049250313e-16 * 1e6) return ST_MESH_GENERATION;
        const double inv[2][2] = {{jx[1][1] / det, -jx[0][1] / det}, {-jx[1][0] / det, jx[0][0] / det}};
        for (int r = 0; r < 2; ++r)
            for (int c = 0; c < 2; ++c) w[2 + r * 2 + c] = inv[r][0] * ju[0][c] + inv[r][1] * ju[1][c];
        if (m.order == 2) {
            double ku[3][2];
            for (int r = 0; r < 3; ++r)
                for (int c = 0; c < 2; ++c) {
                    double a = 0.0;
                    for (int k = 0; k < 6; ++k) a += d2N2[r][k] * u[k][c];
                    ku[r][c] = a;
                }
            const double x0 = x[0][0], y0 = x[0][1], x1 = x[1][0], y1 = x[1][1], x2 = x[2][0], y2 = x[2][1];
            // det of [[1,1,1],[x0,x1,x2],[y0,y1,y2]] (local_coordinates cols 1-3)
            const double det_a = 1.0 * (x1 * y2 - x2 * y1) - 1.0 * (x0 * y2 - x2 * y0) + 1.0 * (x0 * y1 - x1 * y0);
            const double dz[2][2] = {{(y1 - y2) / det_a, (y2 - y0) / det_a}, {(x2 - x1) / det_a, (x0 - x2) / det_a}};
            const double kx[3][3] = {
                {dz[0][0] * dz[0][0], 2.0 * dz[0][0] * dz[0][1], dz[0][1] * dz[0][1]},
                {dz[0][0] * dz[1][0], dz[0][0] * dz[1][1] + dz[0][1] * dz[1][0], dz[0][1] * dz[1][1]},
                {dz[1][0] * dz[1][0], 2.0 * dz[1][0] * dz[1][1], dz[1][1] * dz[1][1]}};
            for (int r = 0; r < 3; ++r)
                for (int c = 0; c < 2; ++c)
                    w[6 + r * 2 + c] = kx[r][0] * ku[0][c] + kx[r][1] * ku[1][c] + kx[r][2] * ku[2][c];
        }
    }
    return ST_OK;
}

// particle.rs:137-183 local_coordinates: area coordinates (zeta, eta,
// theta) of a point in the corner triangle, and the denominator determinant.
double det3(double a00, double a01, double a02, double a10, double a11, double a12, double a20, double a21,
            double a22) {
    return a00 * (a11 * a22 - a12 * a21) - a01 * (a10 * a22 - a12 * a20) + a02 * (a10 * a21 - a11 * a20);
}

void local_coords(const double* c, const double* en, double* zeta, double* eta, double* theta, double* den) {
    const double px = c[0], py = c[1];
    const double n0x = en[0], n0y = en[1], n1x = en[2], n1y = en[3], n2x = en[4], n2y = en[5];
    *den = det3(1.0, 1.0, 1.0, n0x, n1x, n2x, n0y, n1y, n2y);
    const double dz = det3(1.0, 1.0, 1.0, px, n1x, n2x, py, n1y, n2y);
    const double de = det3(1.0, 1.0, 1.0, px, n2x, n0x, py, n2y, n0y);
    *zeta = dz / *den;
    *eta = de / *den;
    *theta = 1.0 - *zeta - *eta;
}

}  // namespace

extern "C" {

// warp_increment (particle.rs:254-320) with shape_function (186-238): the
// displacement and its gradients (and for 6-node elements the second
// gradients) at a point, interpolated in one element.
void orc_warp_increment(const double* coord, const double* en, const double* ed, int order, double* out) {
    double z, e, t, den;
    local_coords(coord, en, &z, &e, &t, &den);
    const int ns = order == 1 ? 3 : 6;
    double N[6], dN[2][6];
    if (order == 1) {
        N[0] = z; N[1] = e; N[2] = t;
        const double d[2][3] = {{1.0, 0.0, -1.0}, {0.0, 1.0, -1.0}};
        for (int r = 0; r < 2; ++r)
            for (int k = 0; k < 3; ++k) dN[r][k] = d[r][k];
    } else {
        N[0] = z * (2.0 * z - 1.0);
        N[1] = e * (2.0 * e - 1.0);
        N[2] = t * (2.0 * t - 1.0);
        N[3] = 4.0 * z * e;
        N[4] = 4.0 * e * t;
        N[5] = 4.0 * t * z;
        const double d[2][6] = {{4.0 * z - 1.0, 0.0, 1.0 - 4.0 * t, 4.0 * e, -4.0 * e, 4.0 * (t - z)},
                                {0.0, 4.0 * e - 1.0, 1.0 - 4.0 * t, 4.0 * z, 4.0 * (t - e), -4.0 * z}};
        for (int r = 0; r < 2; ++r)
            for (int k = 0; k < 6; ++k) dN[r][k] = d[r][k];
    }
    const int plen = 6 * order;
    for (int k = 0; k < plen; ++k) out[k] = 0.0;
    for (int c = 0; c < 2; ++c) {
        double s = 0.0;
        for (int k = 0; k < ns; ++k) s += N[k] * ed[2 * k + c];
        out[c] = s;
    }
    double jx[2][2], ju[2][2];
    for (int r = 0; r < 2; ++r)
        for (int c = 0; c < 2; ++c) {
            double a = 0.0, b = 0.0;
            for (int k = 0; k < ns; ++k) a += dN[r][k] * en[2 * k + c];
            for (int k = 0; k < ns; ++k) b += dN[r][k] * ed[2 * k + c];
            jx[r][c] = a;
            ju[r][c] = b;
        }
    // inv2x2 (particle.rs:323-333): reciprocal of the determinant
    const double idet = 1.0 / (jx[0][0] * jx[1][1] - jx[0][1] * jx[1][0]);
    const double ij[2][2] = {{jx[1][1] * idet, -jx[0][1] * idet}, {-jx[1][0] * idet, jx[0][0] * idet}};
    for (int r = 0; r < 2; ++r)
        for (int c = 0; c < 2; ++c) out[2 + 2 * r + c] = ij[r][0] * ju[0][c] + ij[r][1] * ju[1][c];
    if (order == 2) {
        static const double d2N[3][6] = {
            {4.0, 0.0, 4.0, 0.0, 0.0, -8.0}, {0.0, 0.0, 4.0, 4.0, -4.0, -4.0}, {0.0, 4.0, 4.0, 0.0, -8.0, 0.0}};
        double ku[3][2];
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 2; ++c) {
                double a = 0.0;
                for (int k = 0; k < 6; ++k) a += d2N[r][k] * ed[2 * k + c];
                ku[r][c] = a;
            }
        const double dz[2][2] = {{(en[3] - en[5]) / den, (en[5] - en[1]) / den},
                                 {(en[4] - en[2]) / den, (en[0] - en[4]) / den}};
        const double kx[3][3] = {{dz[0][0] * dz[0][0], 2.0 * dz[0][0] * dz[0][1], dz[0][1] * dz[0][1]},
                                 {dz[0][0] * dz[1][0], dz[0][0] * dz[1][1] + dz[0][1] * dz[1][0], dz[0][1] * dz[1][1]},
                                 {dz[1][0] * dz[1][0], 2.0 * dz[1][0] * dz[1][1], dz[1][1] * dz[1][1]}};
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 2; ++c) out[6 + 2 * r + c] = kx[r][0] * ku[0][c] + kx[r][1] * ku[1][c] + kx[r][2] * ku[2][c];
    }
}

// element_locator (particle.rs:968-1013): elements by squared centroid
// distance (stable order), the first whose corner triangle holds the point
// (area coordinates all >= 0); else the nearest, flagged adrift.
int64_t orc_element_locator(const double* coord, const double* nodes, const int64_t* el, int64_t n_el, int cols,
                            const double* centroids, int* adrift) {
    std::vector<std::pair<double, int64_t>> d(n_el);
    for (int64_t i = 0; i < n_el; ++i) {
        const double dx = centroids[2 * i] - coord[0], dy = centroids[2 * i + 1] - coord[1];
        d[i] = {dx * dx + dy * dy, i};
    }
    std::stable_sort(d.begin(), d.end(), [](const std::pair<double, int64_t>& a, const std::pair<double, int64_t>& b) {
        return a.first < b.first;
    });
    for (auto& q : d) {
        double en[6];
        for (int k = 0; k < 3; ++k) {
            en[2 * k] = nodes[2 * el[q.second * cols + k]];
            en[2 * k + 1] = nodes[2 * el[q.second * cols + k] + 1];
        }
        double z, e, t, den;
        local_coords(coord, en, &z, &e, &t, &den);
        if (z >= 0.0 && e >= 0.0 && t >= 0.0) {
            *adrift = 0;
            return q.second;
        }
    }
    *adrift = 1;
    return d.empty() ? 0 : d[0].second;
}

// Parameters of one mesh solve: SolveConfig (mesh.rs:222-248), SeedConfig
// (269-278) and LayerRgConfig (339-356).  preconditioning 0 = RG, 1 = layer-RG.
struct OrcMeshParams {
    int32_t subset_order, method, max_iterations, preconditioning;
    int32_t override_active, batch_factor, workers, threads;
    double tolerance, max_norm, root_rel_eps, seed_tolerance;
    double seed_coord[2];
    double seed_warp[12];
};

// Per-node and per-element outputs (MeshSolution, mesh.rs:54-113).
struct OrcMeshOut {
    double* p;            // n_nodes x 12 (p_len columns used)
    double* c_zncc;       // n_nodes, clamped at 0 (mesh.rs:2396)
    int32_t* iterations;  // n_nodes
    double* norms;        // n_nodes: last history norm, 0 without history
    int32_t* quality_ok;  // n_nodes
    double* warps;        // n_el x 12: element_strains
    int64_t seed_node;
    int64_t n_solves;     // subset solves run (retries and corrections included)
    int64_t n_bad;        // nodes without quality_ok after corrections
};

int64_t orc_mesh_connectivity(const int64_t* el, int64_t n_el, int cols, int order, int64_t idx, int full,
                              int64_t* out) {
    Mesh m{nullptr, 0, el, n_el, cols, order};
    std::vector<int64_t> v = connectivity(m, idx, full != 0);
    for (size_t k = 0; k < v.size(); ++k) out[k] = v[k];
    return (int64_t)v.size();
}

int64_t orc_mesh_find_seed(const double* nodes, int64_t n, double x, double y) {
    return find_seed_node(nodes, n, x, y);
}

int64_t orc_mesh_corr_1d(const double* v, int64_t n, int64_t* out) {
    std::vector<int64_t> ids = corr_1d(v, n);
    for (size_t k = 0; k < ids.size(); ++k) out[k] = ids[k];
    return (int64_t)ids.size();
}

double orc_mesh_flow_calc(const int64_t* el, int64_t n_el, int cols, int order, int64_t n_nodes, int64_t idx,
                          const double* disp, const uint8_t* exclude, const double* override_disp) {
    Mesh m{nullptr, n_nodes, el, n_el, cols, order};
    Adjacency adj(m);
    std::vector<char> ex(n_nodes, 0);
    if (exclude)
        for (int64_t i = 0; i < n_nodes; ++i) ex[i] = exclude[i] != 0;
    return flow_calc(adj, idx, disp, ex, override_disp);
}

void orc_mesh_r_stats(const int64_t* el, int64_t n_el, int cols, int order, int64_t n_nodes, const double* disp,
                      double* r_out, uint8_t* flagged_out) {
    Mesh m{nullptr, n_nodes, el, n_el, cols, order};
    Adjacency adj(m);
    std::vector<double> r;
    std::vector<char> f;
    r_stats(adj, m, disp, r, f);
    for (int64_t i = 0; i < n_nodes; ++i) {
        r_out[i] = r[i];
        flagged_out[i] = (uint8_t)f[i];
    }
}

void orc_mesh_project_warp(const double* p, int plen, double dx, double dy, double* out) {
    project_warp(p, plen, dx, dy, out);
}

int orc_mesh_affine_field_fit(const double* nodes, const double* target, const int64_t* trusted, int64_t nt,
                              const double* p, int plen, double* out) {
    return affine_field_fit(nodes, target, trusted, nt, p, plen, out) ? 1 : 0;
}

void orc_mesh_precondition_warp(const double* nodes, const double* target, const int64_t* trusted, int64_t nt,
                                const double* c_zncc, const double* p, int plen, double* out) {
    precondition_warp(nodes, target, trusted, nt, c_zncc, p, plen, out);
}

void orc_mesh_greedy_colour(const int64_t* el, int64_t n_el, int cols, int order, int64_t n_nodes,
                            const int64_t* cand, int64_t nc, int64_t* colour_out) {
    Mesh m{nullptr, n_nodes, el, n_el, cols, order};
    Adjacency adj(m);
    std::vector<int64_t> c(cand, cand + nc);
    std::vector<int64_t> col = greedy_colour(adj, c);
    for (int64_t i = 0; i < nc; ++i) colour_out[i] = col[i];
}

void orc_mesh_element_area(const double* nodes, const int64_t* el, int64_t n_el, int cols, double* out) {
    Mesh m{nodes, 0, el, n_el, cols, cols == 3 ? 1 : 2};
    for (int64_t e = 0; e < n_el; ++e) out[e] = element_area(nodes, m, e);
}

int orc_mesh_element_strains(const double* nodes, const int64_t* el, int64_t n_el, int order, const double* disp,
                             double* warps) {
    Mesh m{nodes, 0, el, n_el, order == 1 ? 3 : 6, order};
    return element_strains(nodes, m, disp, warps);
}

// Mesh::solve_impl (mesh.rs:628-857) for uniform masking and deformable
// texture.  Node i's template is tpl_xy[tpl_start[tpl_index[i]] ..], [x, y]
// double pairs (its local mask clipped by the region of interest, as
// subset_at builds it, mesh.rs:575-593).  Outputs are filled even when the
// status reports the unsolvable gate or a compatibility failure.
int orc_mesh_solve(const double* fcoef, const double* gcoef, int h, int w, int border, const double* nodes,
                   int64_t n_nodes, const int64_t* el, int64_t n_el, int mesh_order, const int32_t* tpl_index,
                   const double* tpl_xy, const int64_t* tpl_start, const OrcMeshParams* prm, OrcMeshOut* out) {
    const int cols = mesh_order == 1 ? 3 : 6;
    Mesh m{nodes, n_nodes, el, n_el, cols, mesh_order};
    const int plen = 6 * prm->subset_order;
    Adjacency adj(m);
    std::vector<char> stored(n_nodes, 0), propagated(n_nodes, 0), qok(n_nodes, 0);
    std::vector<double> czn(n_nodes, 0.0), P((size_t)n_nodes * plen, 0.0), disp(2 * n_nodes, 0.0),
        norms(n_nodes, 0.0);
    std::vector<int32_t> iters(n_nodes, 0);
    std::priority_queue<std::pair<uint64_t, int64_t>> queue;
    int64_t n_solves = 0;
    int threads = prm->threads;
#ifdef _OPENMP
    if (threads <= 0) threads = omp_get_max_threads();
#endif

    auto solve_one = [&](int64_t i, const double* p0, double tol, OrcResult* R) {
        const int t = tpl_index[i];
        return orc_subset_solve(fcoef, gcoef, h, w, border, &nodes[2 * i], &tpl_xy[2 * tpl_start[t]],
                                tpl_start[t + 1] - tpl_start[t], prm->subset_order, p0, plen, prm->method, tol,
                                prm->max_norm, prm->max_iterations, R, nullptr, nullptr, nullptr, nullptr);
    };
    // solve_node (mesh.rs:1068-1106): preconditioned attempt, then a zero
    // warp when it fails quality_ok; the better-correlated is kept.
    auto solve_node = [&](int64_t i, const std::vector<int64_t>& trusted, const double* forced, double tol,
                          OrcResult* R, int64_t* count) -> int {
        double pa[12] = {0}, zero[12] = {0};
        if (forced)
            std::memcpy(pa, forced, sizeof(double) * plen);
        else
            precondition_warp(nodes, &nodes[2 * i], trusted.data(), (int64_t)trusted.size(), czn.data(), P.data(),
                              plen, pa);
        int st = solve_one(i, pa, tol, R);
        ++*count;
        if (st || R->quality_ok) return st;
        bool all_zero = true;
        for (int k = 0; k < plen; ++k) all_zero = all_zero && pa[k] == 0.0;
        if (all_zero) return ST_OK;
        OrcResult B;
        st = solve_one(i, zero, tol, &B);
        ++*count;
        if (st) return st;
        if (B.quality_ok || B.c_zncc > R->c_zncc) *R = B;
        return ST_OK;
    };
    // store_result (mesh.rs:2385-2404)
    auto store = [&](int64_t i, const OrcResult& R) {
        qok[i] = R.quality_ok != 0;
        czn[i] = std::fmax(R.c_zncc, 0.0);
        for (int k = 0; k < plen; ++k) P[i * plen + k] = R.p[k];
        disp[2 * i] = R.p[0];
        disp[2 * i + 1] = R.p[1];
        iters[i] = R.iterations;
        norms[i] = R.iterations > 0 ? R.norm : 0.0;
    };
    auto key = [&](int64_t i) {
        uint64_t b;
        std::memcpy(&b, &czn[i], 8);
        return b;
    };
    auto trusted_of = [&](int64_t i) {
        std::vector<int64_t> t;
        for (int64_t nb : adj.get(i, true))
            if (stored[nb] && qok[nb]) t.push_back(nb);
        return t;
    };

    // seed node: the user's warp, the seed tolerance (mesh.rs:659-672)
    const int64_t seed = find_seed_node(nodes, n_nodes, prm->seed_coord[0], prm->seed_coord[1]);
    {
        OrcResult R;
        int st = solve_node(seed, {}, prm->seed_warp, prm->seed_tolerance, &R, &n_solves);
        if (st) return st;
        store(seed, R);
        stored[seed] = propagated[seed] = 1;
        queue.push({key(seed), seed});
    }

    if (prm->preconditioning == 0) {
        // RG: serial cascade from each popped node (mesh.rs:678-700, 1112-1152)
        auto from = [&](int64_t cur) -> int {
            for (int64_t nb : adj.get(cur, false)) {
                if (stored[nb]) continue;
                OrcResult R;
                int st = solve_node(nb, trusted_of(nb), nullptr, prm->tolerance, &R, &n_solves);
                if (st) return st;
                store(nb, R);
                stored[nb] = 1;
                queue.push({key(nb), nb});
            }
            return ST_OK;
        };
        int st = from(seed);
        if (st) return st;
        while (!queue.empty()) {
            const int64_t cur = queue.top().second;
            queue.pop();
            if (propagated[cur]) continue;
            propagated[cur] = 1;
            st = from(cur);
            if (st) return st;
        }
    } else {
        // layer-RG: colour rounds over the frontier of a root band
        // (mesh.rs:701-752, expand_parallel 1169-1252)
        auto expand = [&](const std::vector<int64_t>& roots) -> int {
            std::vector<int64_t> cand;
            for (int64_t r : roots)
                for (int64_t nb : adj.get(r, false))
                    if (!stored[nb]) cand.push_back(nb);
            std::sort(cand.begin(), cand.end());
            cand.erase(std::unique(cand.begin(), cand.end()), cand.end());
            if (cand.empty()) return ST_OK;
            const std::vector<int64_t> colour = greedy_colour(adj, cand);
            const int64_t nc = *std::max_element(colour.begin(), colour.end()) + 1;
            for (int64_t c = 0; c < nc; ++c) {
                std::vector<int64_t> ids;
                std::vector<std::vector<int64_t>> tr;
                for (size_t k = 0; k < cand.size(); ++k)
                    if (colour[k] == c) {
                        ids.push_back(cand[k]);
                        tr.push_back(trusted_of(cand[k]));
                    }
                std::vector<OrcResult> res(ids.size());
                std::vector<int> sts(ids.size(), 0);
                std::vector<int64_t> cnt(ids.size(), 0);
#pragma omp parallel for num_threads(threads) schedule(dynamic, 1)
                for (int64_t k = 0; k < (int64_t)ids.size(); ++k)
                    sts[k] = solve_node(ids[k], tr[k], nullptr, prm->tolerance, &res[k], &cnt[k]);
                for (size_t k = 0; k < ids.size(); ++k) {
                    n_solves += cnt[k];
                    if (sts[k]) return sts[k];
                }
                for (size_t k = 0; k < ids.size(); ++k) {  // ids ascending: the serial fold
                    store(ids[k], res[k]);
                    stored[ids[k]] = 1;
                    queue.push({key(ids[k]), ids[k]});
                }
            }
            return ST_OK;
        };
        const int64_t workers = prm->workers > 0 ? prm->workers : threads;
        const int64_t root_cap = std::max<int64_t>(prm->batch_factor, 1) * std::max<int64_t>(workers, 1);
        int st = expand({seed});
        if (st) return st;
        while (!queue.empty()) {
            const auto top = queue.top();
            queue.pop();
            if (propagated[top.second]) continue;
            propagated[top.second] = 1;
            std::vector<int64_t> roots{top.second};
            double top_val;
            std::memcpy(&top_val, &top.first, 8);
            const double threshold = top_val * (1.0 - prm->root_rel_eps);
            while ((int64_t)roots.size() < root_cap && !queue.empty()) {
                double kv;
                std::memcpy(&kv, &queue.top().first, 8);
                if (!(kv >= threshold)) break;
                const int64_t r = queue.top().second;
                queue.pop();
                if (propagated[r]) continue;
                propagated[r] = 1;
                roots.push_back(r);
            }
            st = expand(roots);
            if (st) return st;
        }
    }

    // corrections, lane 1: correlation_improvements (mesh.rs:1265-1311)
    {
        std::vector<int64_t> order = corr_1d(czn.data(), n_nodes);
        std::stable_sort(order.begin(), order.end(), [&](int64_t a, int64_t b) { return total_less(czn[a], czn[b]); });
        std::vector<char> unimproved(n_nodes, 0);
        for (size_t i = 0; i < order.size(); ++i) {
            const int64_t j = order[i];
            std::vector<char> excluded(unimproved);
            for (size_t k = i + 1; k < order.size(); ++k) excluded[order[k]] = 1;
            std::vector<int64_t> trusted;
            for (int64_t nb : adj.get(j, true))
                if (!excluded[nb] && qok[nb]) trusted.push_back(nb);
            OrcResult R;
            int st = solve_node(j, trusted, nullptr, prm->tolerance, &R, &n_solves);
            if (st) return st;
            if (R.c_zncc > czn[j] && R.quality_ok) {
                store(j, R);
                stored[j] = 1;
            } else {
                unimproved[j] = 1;
            }
        }
    }
    // corrections, lane 2: flow and displacement-magnitude outliers
    // (mesh.rs:1323-1414); flagged nodes in ascending magnitude, ties by index
    {
        std::vector<double> flow, r;
        std::vector<int64_t> flow_ids;
        std::vector<char> r_flag;
        double flq, fiqr;
        flow_stats(adj, disp.data(), n_nodes, flow, flow_ids, flq, fiqr);
        r_stats(adj, m, disp.data(), r, r_flag);
        std::vector<char> in_full(n_nodes, 0);
        for (int64_t i : flow_ids) in_full[i] = 1;
        for (int64_t i = 0; i < n_nodes; ++i)
            if (r_flag[i]) in_full[i] = 1;
        std::vector<int64_t> full_id;
        for (int64_t i = 0; i < n_nodes; ++i)
            if (in_full[i]) full_id.push_back(i);
        std::stable_sort(full_id.begin(), full_id.end(), [&](int64_t a, int64_t b) { return total_less(r[a], r[b]); });
        std::vector<int64_t> active;
        for (int64_t j : full_id) {
            bool exempt = false;
            if (!r_flag[j]) exempt = flow_calc(adj, j, disp.data(), in_full, nullptr) > flq - 2.5 * fiqr;
            if (!exempt) active.push_back(j);
        }
        for (size_t i = 0; i < active.size(); ++i) {
            const int64_t j = active[i];
            std::vector<char> later(n_nodes, 0);
            for (size_t k = i; k < active.size(); ++k) later[active[k]] = 1;
            std::vector<int64_t> trusted;
            for (int64_t nb : adj.get(j, true))
                if (!later[nb] && qok[nb]) trusted.push_back(nb);
            OrcResult R;
            int st = solve_node(j, trusted, nullptr, prm->tolerance, &R, &n_solves);
            if (st) return st;
            store(j, R);
            stored[j] = 1;
        }
    }

    int64_t n_bad = 0;
    for (int64_t i = 0; i < n_nodes; ++i) n_bad += !qok[i];
    for (int64_t i = 0; i < n_nodes; ++i) {
        for (int k = 0; k < 12; ++k) out->p[12 * i + k] = k < plen ? P[i * plen + k] : 0.0;
        out->c_zncc[i] = czn[i];
        out->iterations[i] = iters[i];
        out->norms[i] = norms[i];
        out->quality_ok[i] = qok[i];
    }
    out->seed_node = seed;
    out->n_solves = n_solves;
    out->n_bad = n_bad;
    // the quality gate (mesh.rs:754-777), then strains and compatibility
    // (mesh.rs:798-803, 1417-1433)
    int st = element_strains(nodes, m, disp.data(), out->warps);
    if (!prm->override_active && n_bad > 0) return ST_INVALID_INPUT;
    if (st) return st;
    for (int64_t e = 0; e < n_el; ++e) {
        const double* wv = out->warps + 12 * e;
        const double det = (1.0 + wv[2]) * (1.0 + wv[5]) - wv[3] * wv[4];
        if (det <= 0.0) return ST_INVALID_INPUT;
    }
    return ST_OK;
}

}  // extern "C"
