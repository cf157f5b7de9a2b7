// Separator tree, nested dissection and symbolic analysis (host side).
#include "tree.hpp"

#include "par.hpp"

#include <algorithm>
#include <climits>
#include <cmath>
#include <functional>
#include <numeric>
#include <stdexcept>

namespace amls {

// ---------------------------------------------------------------------------
// SeparatorTree (semantics of src/amls/separator_tree.cpp:9-107)
// ---------------------------------------------------------------------------
int64_t SeparatorTree::addNode(NodeRole role, std::vector<int64_t> indices) {
    TreeNode nd;
    nd.id = nNodes();
    nd.role = role;
    nd.indices = std::move(indices);
    nodes_.push_back(std::move(nd));
    return nodes_.back().id;
}

void SeparatorTree::setChildren(int64_t id, int64_t left, int64_t right) {
    nodes_.at(static_cast<size_t>(id)).children = {left, right};
}

void SeparatorTree::setParent(int64_t id, int64_t parent) {
    nodes_.at(static_cast<size_t>(id)).parent = parent;
}

int64_t SeparatorTree::root() const {
    for (const auto& nd : nodes_)
        if (nd.parent == -1) return nd.id;
    return -1;
}

std::vector<int64_t> SeparatorTree::postOrder() const {
    std::vector<int64_t> out;
    out.reserve(nodes_.size());
    if (nodes_.empty()) return out;
    // iterative post-order: children left to right, then the node
    std::vector<std::pair<int64_t, size_t>> st;
    st.push_back({root(), 0});
    while (!st.empty()) {
        auto& top = st.back();
        const TreeNode& nd = node(top.first);
        if (top.second < nd.children.size()) {
            const int64_t c = nd.children[top.second++];
            st.push_back({c, 0});
        } else {
            out.push_back(top.first);
            st.pop_back();
        }
    }
    return out;
}

std::vector<int64_t> SeparatorTree::ancestors(int64_t id) const {
    std::vector<int64_t> out;
    for (int64_t p = node(id).parent; p != -1; p = node(p).parent) out.push_back(p);
    return out;
}

std::vector<int64_t> SeparatorTree::descendants(int64_t id) const {
    std::vector<int64_t> out, st(node(id).children.rbegin(), node(id).children.rend());
    while (!st.empty()) {
        const int64_t c = st.back();
        st.pop_back();
        out.push_back(c);
        const auto& ch = node(c).children;
        for (auto it = ch.rbegin(); it != ch.rend(); ++it) st.push_back(*it);
    }
    return out;
}

void SeparatorTree::relabelRoot() {
    const int64_t r = root();
    if (r >= 0 && !nodes_[static_cast<size_t>(r)].children.empty())
        nodes_[static_cast<size_t>(r)].role = NodeRole::Root;
}

bool SeparatorTree::validate(int64_t nTotal, std::string* why) const {
    auto fail = [&](const char* m) {
        if (why) *why = m;
        return false;
    };
    std::vector<char> seen(static_cast<size_t>(nTotal), 0);
    int64_t cnt = 0, roots = 0;
    for (const auto& nd : nodes_) {
        for (int64_t v : nd.indices) {
            if (v < 0 || v >= nTotal) return fail("index out of range");
            if (seen[static_cast<size_t>(v)]) return fail("indices overlap");
            seen[static_cast<size_t>(v)] = 1;
            ++cnt;
        }
        if (nd.parent == -1) ++roots;
        if (!nd.children.empty()) {
            if (nd.children.size() != 2) return fail("separator without two children");
            for (int64_t c : nd.children)
                if (node(c).parent != nd.id) return fail("parent/child asymmetry");
            if (nd.role == NodeRole::Leaf) return fail("leaf with children");
        } else if (nd.role != NodeRole::Leaf) {
            return fail("childless non-leaf");
        }
        if (nd.role == NodeRole::Root && nd.parent != -1) return fail("root with parent");
    }
    if (cnt != nTotal) return fail("indices do not cover 0..n-1");
    if (roots != 1) return fail("not exactly one root");
    return true;
}

// ---------------------------------------------------------------------------
// Graph of |K| + |M|
// ---------------------------------------------------------------------------
namespace {

// general form: both directions of every entry of either pattern, then one
// sorted, duplicate-free list per vertex
Graph symmetrisedGraph(int64_t n, const int64_t* cpK, const int64_t* riK, const int64_t* cpM,
                       const int64_t* riM) {
    Graph g;
    g.n = n;
    g.ptr.assign(static_cast<size_t>(n) + 1, 0);
    std::vector<int64_t> mark(static_cast<size_t>(n), -1);
    auto visit = [&](auto&& fn) {
        for (int64_t c = 0; c < n; ++c) {
            for (int64_t e = cpK[c]; e < cpK[c + 1]; ++e) fn(c, riK[e]);
            if (cpM)
                for (int64_t e = cpM[c]; e < cpM[c + 1]; ++e) fn(c, riM[e]);
        }
    };
    std::vector<int64_t> deg(static_cast<size_t>(n), 0);
    visit([&](int64_t c, int64_t r) {
        if (r == c) return;
        ++deg[static_cast<size_t>(c)];
        ++deg[static_cast<size_t>(r)];
    });
    std::vector<int64_t> ptr(static_cast<size_t>(n) + 1, 0);
    for (int64_t v = 0; v < n; ++v) ptr[v + 1] = ptr[v] + deg[v];
    std::vector<int64_t> adj(static_cast<size_t>(ptr[n]));
    std::vector<int64_t> fill(ptr.begin(), ptr.end() - 1);
    visit([&](int64_t c, int64_t r) {
        if (r == c) return;
        adj[static_cast<size_t>(fill[c]++)] = r;
        adj[static_cast<size_t>(fill[r]++)] = c;
    });
    g.adj.reserve(adj.size() / 2 + 16);
    for (int64_t v = 0; v < n; ++v) {
        for (int64_t e = ptr[v]; e < ptr[v + 1]; ++e) {
            const int64_t w = adj[static_cast<size_t>(e)];
            if (mark[w] != v) {
                mark[w] = v;
                g.adj.push_back(w);
            }
        }
        std::sort(g.adj.begin() + g.ptr[v], g.adj.end());
        g.ptr[v + 1] = static_cast<int64_t>(g.adj.size());
    }
    return g;
}

}  // namespace

Graph graphFromCsc(int64_t n, const int64_t* cpK, const int64_t* riK,
                   const int64_t* cpM, const int64_t* riM, int threads) {
    // Fast path for the usual input, both triangles of a symmetric pattern:
    // vertex c's neighbours are the merged rows of column c of K and M,
    // self loop dropped.  Columns whose rows are not strictly ascending are
    // sorted first.  The result is returned when every neighbour relation
    // holds both ways (checked by binary search); otherwise the pattern is
    // symmetrised the general way.
    const int T = n < 4096 ? 1 : threads > 0 ? threads : hostThreads();
    std::vector<int64_t> deg(static_cast<size_t>(n) + 1, 0);
    auto column = [&](int64_t c, std::vector<int64_t>& buf, int64_t* out) -> int64_t {
        const int64_t* a = riK + cpK[c];
        const int64_t na = cpK[c + 1] - cpK[c];
        const int64_t* b = cpM ? riM + cpM[c] : nullptr;
        const int64_t nb = cpM ? cpM[c + 1] - cpM[c] : 0;
        bool sorted = true;
        for (int64_t e = 1; e < na && sorted; ++e) sorted = a[e] > a[e - 1];
        for (int64_t e = 1; e < nb && sorted; ++e) sorted = b[e] > b[e - 1];
        int64_t m = 0;
        auto emit = [&](int64_t r) {
            if (r == c || (m > 0 && (out ? out[m - 1] : buf[0]) == r)) return;
            if (out) out[m] = r;
            else buf[0] = r;
            ++m;
        };
        if (sorted) {
            int64_t i = 0, j = 0;
            while (i < na || j < nb) {
                const int64_t r = j >= nb || (i < na && a[i] <= b[j]) ? a[i++] : b[j++];
                emit(r);
            }
        } else {
            buf.assign(a, a + na);
            buf.insert(buf.end(), b, b + nb);
            std::sort(buf.begin(), buf.end());
            int64_t last = -1;
            for (int64_t r : buf) {
                if (r == c || r == last) continue;
                last = r;
                if (out) out[m] = r;
                ++m;
            }
        }
        return m;
    };
    parallelRanges(n, T, [&](int64_t lo, int64_t hi, int) {
        std::vector<int64_t> buf(1);
        for (int64_t c = lo; c < hi; ++c) deg[c + 1] = column(c, buf, nullptr);
    });
    Graph g;
    g.n = n;
    g.ptr.assign(static_cast<size_t>(n) + 1, 0);
    for (int64_t v = 0; v < n; ++v) g.ptr[v + 1] = g.ptr[v] + deg[v + 1];
    g.adj.resize(static_cast<size_t>(g.ptr[n]));
    std::vector<char> asym(static_cast<size_t>(T), 0);
    parallelRanges(n, T, [&](int64_t lo, int64_t hi, int) {
        std::vector<int64_t> buf(1);
        for (int64_t c = lo; c < hi; ++c) column(c, buf, g.adj.data() + g.ptr[c]);
    });
    // symmetry: the vertices c of a range [lo, hi), visited in ascending
    // order, must meet each neighbour list adj(r) at its entries in [lo, hi)
    // in order, and use all of them.  One cursor per touched r, started by a
    // binary search for lo.
    parallelRanges(n, T, [&](int64_t lo, int64_t hi, int chunk) {
        std::vector<int64_t> cur(static_cast<size_t>(n), -1), touched;
        for (int64_t c = lo; c < hi; ++c)
            for (int64_t e = g.ptr[c]; e < g.ptr[c + 1]; ++e) {
                const int64_t r = g.adj[e];
                int64_t& k = cur[static_cast<size_t>(r)];
                if (k < 0) {
                    k = std::lower_bound(g.adj.begin() + g.ptr[r], g.adj.begin() + g.ptr[r + 1], lo) -
                        g.adj.begin();
                    touched.push_back(r);
                }
                if (k >= g.ptr[r + 1] || g.adj[k] != c) {
                    asym[static_cast<size_t>(chunk)] = 1;
                    return;
                }
                ++k;
            }
        for (int64_t r : touched) {
            const int64_t k = cur[static_cast<size_t>(r)];
            if (k < g.ptr[r + 1] && g.adj[k] < hi) {
                asym[static_cast<size_t>(chunk)] = 1;
                return;
            }
        }
        // every entry in [lo, hi) of an untouched list would have been met
        // from its column: such lists hold none
    });
    for (char a : asym)
        if (a) return symmetrisedGraph(n, cpK, riK, cpM, riM);
    return g;
}

// ---------------------------------------------------------------------------
// Nested dissection
// ---------------------------------------------------------------------------
namespace {

struct Bisector {
    const Graph& g;
    const double* coords;
    std::vector<int64_t> stamp;   // membership stamp per vertex
    std::vector<int64_t> dist;    // BFS level / side scratch
    int64_t cur = 0;

    Bisector(const Graph& gg, const double* c)
        : g(gg), coords(c), stamp(static_cast<size_t>(gg.n), -1),
          dist(static_cast<size_t>(gg.n), -1) {}

    // side: 0 / 1 / 2 (= separator) for each vertex of verts, in order.
    bool byPlane(const std::vector<int64_t>& verts, std::vector<int>& side) {
        if (!coords) return false;
        const int64_t m = static_cast<int64_t>(verts.size());
        double lo[3], hi[3];
        for (int d = 0; d < 3; ++d) { lo[d] = 1e300; hi[d] = -1e300; }
        for (int64_t v : verts)
            for (int d = 0; d < 3; ++d) {
                lo[d] = std::min(lo[d], coords[3 * v + d]);
                hi[d] = std::max(hi[d], coords[3 * v + d]);
            }
        int axes[3] = {0, 1, 2};
        std::sort(axes, axes + 3, [&](int a, int b) {
            const double ea = hi[a] - lo[a], eb = hi[b] - lo[b];
            return ea != eb ? ea > eb : a < b;
        });
        std::vector<double> vals(static_cast<size_t>(m));
        for (int ai = 0; ai < 3; ++ai) {
            const int d = axes[ai];
            if (!(hi[d] > lo[d])) break;
            for (int64_t t = 0; t < m; ++t) vals[t] = coords[3 * verts[t] + d];
            std::vector<double> u(vals);
            std::sort(u.begin(), u.end());
            std::vector<double> uv;
            std::vector<int64_t> cnt;
            for (double x : u) {
                if (uv.empty() || x != uv.back()) { uv.push_back(x); cnt.push_back(0); }
                ++cnt.back();
            }
            const int64_t K = static_cast<int64_t>(uv.size());
            if (K < 3) continue;
            std::vector<int64_t> pre(static_cast<size_t>(K) + 1, 0);
            for (int64_t j = 0; j < K; ++j) pre[j + 1] = pre[j] + cnt[j];
            std::vector<int64_t> cand;
            for (int64_t j = 1; j + 1 < K; ++j) cand.push_back(j);
            auto bal = [&](int64_t j) {
                const int64_t less = pre[j], greater = m - pre[j + 1];
                return std::llabs(less - greater);
            };
            std::stable_sort(cand.begin(), cand.end(), [&](int64_t a, int64_t b) {
                const int64_t ba = bal(a), bb = bal(b);
                if (ba != bb) return ba < bb;
                return cnt[a] < cnt[b];
            });
            if (cand.size() > 12) cand.resize(12);
            ++cur;
            for (int64_t t = 0; t < m; ++t) stamp[verts[t]] = cur;
            for (int64_t j : cand) {
                const double c = uv[j];
                const int64_t less = pre[j], greater = m - pre[j + 1];
                if (std::min(less, greater) < 2) continue;
                bool ok = true;
                for (int64_t t = 0; t < m && ok; ++t) {
                    const int64_t v = verts[t];
                    if (!(coords[3 * v + d] < c)) continue;
                    for (int64_t e = g.ptr[v]; e < g.ptr[v + 1]; ++e) {
                        const int64_t w = g.adj[e];
                        if (stamp[w] == cur && coords[3 * w + d] > c) { ok = false; break; }
                    }
                }
                if (!ok) continue;
                side.assign(static_cast<size_t>(m), 0);
                for (int64_t t = 0; t < m; ++t) {
                    const double x = vals[t];
                    side[t] = x < c ? 0 : (x > c ? 1 : 2);
                }
                return true;
            }
        }
        return false;
    }

    // Spectral bisection (no coordinates): an approximate Fiedler vector of
    // the induced subgraph's Laplacian L = D - A, by Lanczos with full
    // reorthogonalisation on sigma I - L against the constant vector
    // (sigma = twice the largest degree bounds L's spectrum), at most 160
    // steps from a fixed start; the vertices split at its median, and the
    // smaller of the two sides' cut boundaries becomes the separator.
    // False when the subgraph is too small or too large for it, or the cut
    // is degenerate.  Subsets above 20,000 vertices take the level
    // structure instead (the basis would cost m x 100 doubles per step).
    std::vector<int64_t> local;   // global -> local index of the current subset
    bool bySpectral(const std::vector<int64_t>& verts, std::vector<int>& side) {
        const int64_t m = static_cast<int64_t>(verts.size());
        if (m < 8 || m > 20000) return false;
        if (local.size() != static_cast<size_t>(g.n)) local.assign(static_cast<size_t>(g.n), -1);
        ++cur;
        for (int64_t t = 0; t < m; ++t) {
            stamp[verts[t]] = cur;
            local[verts[t]] = t;
        }
        std::vector<int64_t> lp(static_cast<size_t>(m) + 1, 0), li;
        li.reserve(static_cast<size_t>(std::min<int64_t>(g.adj.size(), 64 * m)));
        int64_t dmax = 0;
        for (int64_t t = 0; t < m; ++t) {
            const int64_t v = verts[t];
            for (int64_t e = g.ptr[v]; e < g.ptr[v + 1]; ++e)
                if (stamp[g.adj[e]] == cur) li.push_back(local[g.adj[e]]);
            lp[t + 1] = static_cast<int64_t>(li.size());
            dmax = std::max(dmax, lp[t + 1] - lp[t]);
        }
        const double sigma = 2.0 * static_cast<double>(std::max<int64_t>(dmax, 1));
        // the vector loops run on host threads for large subsets
        const int NT = m >= 8192 ? hostThreads() : 1;
        auto par = [&](auto&& fn) { parallelRanges(m, NT, [&](int64_t lo, int64_t hi, int) { fn(lo, hi); }); };
        auto dot = [&](const std::vector<double>& x, const std::vector<double>& y) {
            std::vector<double> part(static_cast<size_t>(NT), 0.0);
            parallelRanges(m, NT, [&](int64_t lo, int64_t hi, int c) {
                double acc = 0.0;
                for (int64_t t = lo; t < hi; ++t) acc += x[t] * y[t];
                part[static_cast<size_t>(c)] = acc;
            });
            double acc = 0.0;
            for (double v : part) acc += v;
            return acc;
        };
        auto apply = [&](const std::vector<double>& x, std::vector<double>& y) {
            par([&](int64_t lo, int64_t hi) {
                for (int64_t t = lo; t < hi; ++t) {
                    double acc = (sigma - static_cast<double>(lp[t + 1] - lp[t])) * x[t];
                    for (int64_t e = lp[t]; e < lp[t + 1]; ++e) acc += x[li[e]];
                    y[t] = acc;
                }
            });
        };
        const std::vector<double> ones(static_cast<size_t>(m), 1.0);
        auto deflate = [&](std::vector<double>& x) {   // remove the constant vector
            const double mean = dot(x, ones) / static_cast<double>(m);
            par([&](int64_t lo, int64_t hi) {
                for (int64_t t = lo; t < hi; ++t) x[t] -= mean;
            });
        };
        // Lanczos steps: enough for the lowest modes of a mesh-like graph
        // (the count grows with the diameter, not with m)
        const int kmax = static_cast<int>(std::min<int64_t>(40, m - 2));
        std::vector<std::vector<double>> Q;
        std::vector<double> al, be;
        std::vector<double> q(static_cast<size_t>(m)), w(static_cast<size_t>(m));
        for (int64_t t = 0; t < m; ++t)   // fixed start: a ramp plus a slow oscillation
            q[t] = static_cast<double>(t) / static_cast<double>(m) +
                   0.1 * std::sin(0.37 * static_cast<double>(t));
        deflate(q);
        double nq = std::sqrt(dot(q, q));
        if (!(nq > 0.0)) return false;
        for (double& v : q) v /= nq;
        for (int k = 0; k < kmax; ++k) {
            Q.push_back(q);
            apply(q, w);
            deflate(w);
            al.push_back(dot(w, q));
            // classical Gram-Schmidt against the basis, a second pass when
            // the first removed most of w ("twice is enough")
            for (int pass = 0; pass < 2; ++pass) {
                const double before = std::sqrt(dot(w, w));
                const size_t nb = Q.size();
                std::vector<double> part(static_cast<size_t>(NT) * nb, 0.0), c(nb, 0.0);
                parallelRanges(m, NT, [&](int64_t lo, int64_t hi, int ch) {
                    for (size_t j = 0; j < nb; ++j) {
                        const double* u = Q[j].data();
                        double acc = 0.0;
                        for (int64_t t = lo; t < hi; ++t) acc += w[t] * u[t];
                        part[static_cast<size_t>(ch) * nb + j] = acc;
                    }
                });
                for (int ch = 0; ch < NT; ++ch)
                    for (size_t j = 0; j < nb; ++j) c[j] += part[static_cast<size_t>(ch) * nb + j];
                par([&](int64_t lo, int64_t hi) {
                    for (size_t j = 0; j < nb; ++j) {
                        const double cj = c[j];
                        const double* u = Q[j].data();
                        for (int64_t t = lo; t < hi; ++t) w[t] -= cj * u[t];
                    }
                });
                if (std::sqrt(dot(w, w)) > 0.7 * before) break;
            }
            const double b = std::sqrt(dot(w, w));
            if (!(b > 1e-10 * sigma)) break;
            be.push_back(b);
            par([&](int64_t lo, int64_t hi) {
                for (int64_t t = lo; t < hi; ++t) q[t] = w[t] / b;
            });
        }
        const int kk = static_cast<int>(al.size());
        if (kk < 2) return false;
        // top eigenvector of the kk x kk tridiagonal by cyclic Jacobi
        std::vector<double> T(static_cast<size_t>(kk) * kk, 0.0), V(static_cast<size_t>(kk) * kk, 0.0);
        for (int i = 0; i < kk; ++i) {
            T[i * kk + i] = al[i];
            V[i * kk + i] = 1.0;
            if (i + 1 < kk) T[i * kk + i + 1] = T[(i + 1) * kk + i] = be[i];
        }
        for (int sweep = 0; sweep < 60; ++sweep) {
            double off = 0.0;
            for (int i = 0; i < kk; ++i)
                for (int j = i + 1; j < kk; ++j) off += T[i * kk + j] * T[i * kk + j];
            if (off < 1e-24 * sigma * sigma) break;
            for (int p0 = 0; p0 < kk; ++p0)
                for (int q0 = p0 + 1; q0 < kk; ++q0) {
                    const double apq = T[p0 * kk + q0];
                    if (std::fabs(apq) < 1e-300) continue;
                    const double th = (T[q0 * kk + q0] - T[p0 * kk + p0]) / (2.0 * apq);
                    const double tt = (th >= 0 ? 1.0 : -1.0) / (std::fabs(th) + std::sqrt(th * th + 1.0));
                    const double c = 1.0 / std::sqrt(tt * tt + 1.0), sn = tt * c;
                    for (int r = 0; r < kk; ++r) {   // columns p0, q0
                        const double x = T[r * kk + p0], y = T[r * kk + q0];
                        T[r * kk + p0] = c * x - sn * y;
                        T[r * kk + q0] = sn * x + c * y;
                    }
                    for (int r = 0; r < kk; ++r) {   // rows p0, q0
                        const double x = T[p0 * kk + r], y = T[q0 * kk + r];
                        T[p0 * kk + r] = c * x - sn * y;
                        T[q0 * kk + r] = sn * x + c * y;
                    }
                    for (int r = 0; r < kk; ++r) {
                        const double x = V[r * kk + p0], y = V[r * kk + q0];
                        V[r * kk + p0] = c * x - sn * y;
                        V[r * kk + q0] = sn * x + c * y;
                    }
                }
        }
        int top = 0;
        for (int i = 1; i < kk; ++i)
            if (T[i * kk + i] > T[top * kk + top]) top = i;
        std::vector<double> f(static_cast<size_t>(m), 0.0);
        for (int i = 0; i < kk; ++i) {
            const double c = V[i * kk + top];
            for (int64_t t = 0; t < m; ++t) f[t] += c * Q[i][t];
        }
        // median split (ties by local index), then the cut boundary
        std::vector<int64_t> ord(static_cast<size_t>(m));
        std::iota(ord.begin(), ord.end(), 0);
        std::stable_sort(ord.begin(), ord.end(), [&](int64_t a, int64_t b) { return f[a] < f[b]; });
        side.assign(static_cast<size_t>(m), 1);
        for (int64_t r = 0; r < m / 2; ++r) side[ord[r]] = 0;
        int64_t bnd[2] = {0, 0};
        std::vector<char> onB[2] = {std::vector<char>(static_cast<size_t>(m), 0),
                                    std::vector<char>(static_cast<size_t>(m), 0)};
        for (int64_t t = 0; t < m; ++t)
            for (int64_t e = lp[t]; e < lp[t + 1]; ++e)
                if (side[li[e]] != side[t]) {
                    onB[side[t]][t] = 1;
                    break;
                }
        for (int64_t t = 0; t < m; ++t) bnd[side[t]] += onB[side[t]][t];
        const int pick = bnd[0] <= bnd[1] ? 0 : 1;
        if (bnd[pick] == 0) return false;
        for (int64_t t = 0; t < m; ++t)
            if (side[t] == pick && onB[pick][t]) side[t] = 2;
        return true;
    }

    // Vertex-separator refinement (Fiduccia-Mattheyses moves on a node
    // separator, as METIS refines its separators): moving separator vertex v
    // into part A pulls its part-B neighbours into the separator, a gain of
    // 1 - |N(v) in B|.  Separator vertices with no neighbour in one part
    // always leave it (gain 1), so the result is a minimal separator: every
    // vertex left in it touches both parts.  With fm, each pass then moves
    // the best-gain unlocked separator vertex that keeps the larger part
    // within maxFrac of the subset, up to a bounded number of moves, rolls
    // back to the smallest separator seen, and passes repeat while they
    // improve (a full scan per move: meant for low-degree graphs).
    void refine(const std::vector<int64_t>& verts, std::vector<int>& side, bool fm,
                double maxFrac = 0.6) {
        const int64_t m = static_cast<int64_t>(verts.size());
        if (m < 4) return;
        if (local.size() != static_cast<size_t>(g.n)) local.assign(static_cast<size_t>(g.n), -1);
        ++cur;
        for (int64_t t = 0; t < m; ++t) {
            stamp[verts[t]] = cur;
            local[verts[t]] = t;
        }
        const int64_t mine = cur;
        auto nbrs = [&](int64_t t, auto&& fn) {
            const int64_t v = verts[t];
            for (int64_t e = g.ptr[v]; e < g.ptr[v + 1]; ++e)
                if (stamp[g.adj[e]] == mine) fn(local[g.adj[e]]);
        };
        int64_t cnt[3] = {0, 0, 0};
        for (int64_t t = 0; t < m; ++t) ++cnt[side[t]];
        const int64_t cap = static_cast<int64_t>(maxFrac * static_cast<double>(m));
        auto costTo = [&](int64_t t, int to) {   // vertices of the other part pulled in
            int64_t c = 0;
            nbrs(t, [&](int64_t u) { c += side[u] == 1 - to; });
            return c;
        };
        for (int pass = 0; pass < 8; ++pass) {
            // pure gains first: separator vertices adjacent to one part only
            bool any = false;
            for (int64_t t = 0; t < m; ++t) {
                if (side[t] != 2) continue;
                for (int to = 0; to < 2; ++to)
                    if (costTo(t, to) == 0 && cnt[to] + 1 <= std::max(cap, cnt[to])) {
                        side[t] = to;
                        --cnt[2];
                        ++cnt[to];
                        any = true;
                        break;
                    }
            }
            if (!fm) break;
            // a bounded FM pass with rollback to the best separator
            std::vector<int64_t> seps;
            for (int64_t t = 0; t < m; ++t)
                if (side[t] == 2) seps.push_back(t);
            if (seps.empty() || seps.size() > 4096) break;
            std::vector<char> locked(static_cast<size_t>(m), 0);
            struct Move { int64_t v; int to; std::vector<int64_t> pulled; };
            std::vector<Move> log;
            const int64_t best0 = cnt[2];
            int64_t best = cnt[2], bestAt = 0;
            const int64_t maxMoves = std::min<int64_t>(256, 2 * static_cast<int64_t>(seps.size()));
            for (int64_t mv = 0; mv < maxMoves; ++mv) {
                int64_t bv = -1, bg = INT64_MIN;
                int bto = 0;
                for (int64_t t : seps) {
                    if (side[t] != 2 || locked[t]) continue;
                    for (int to = 0; to < 2; ++to) {
                        const int64_t c = costTo(t, to);
                        if (cnt[to] + 1 > cap || cnt[1 - to] - c < 1) continue;
                        const int64_t gain = 1 - c;
                        if (gain > bg) {
                            bg = gain;
                            bv = t;
                            bto = to;
                        }
                    }
                }
                if (bv < 0 || bg < -8) break;
                Move M{bv, bto, {}};
                side[bv] = bto;
                locked[bv] = 1;
                --cnt[2];
                ++cnt[bto];
                nbrs(bv, [&](int64_t u) {
                    if (side[u] == 1 - bto) {
                        side[u] = 2;
                        --cnt[1 - bto];
                        ++cnt[2];
                        locked[u] = 1;
                        M.pulled.push_back(u);
                        seps.push_back(u);
                    }
                });
                log.push_back(std::move(M));
                if (cnt[2] < best) {
                    best = cnt[2];
                    bestAt = static_cast<int64_t>(log.size());
                }
            }
            for (int64_t i = static_cast<int64_t>(log.size()) - 1; i >= bestAt; --i) {
                const Move& M = log[i];
                for (int64_t u : M.pulled) {
                    side[u] = 1 - M.to;
                    --cnt[2];
                    ++cnt[1 - M.to];
                }
                side[M.v] = 2;
                --cnt[M.to];
                ++cnt[2];
            }
            if (best >= best0 && !any) break;
        }
    }

    // BFS level structure from a pseudo-peripheral vertex; separator = the
    // level where the cumulative count first reaches half.
    void byLevels(const std::vector<int64_t>& verts, std::vector<int>& side) {
        const int64_t m = static_cast<int64_t>(verts.size());
        ++cur;
        const int64_t mine = cur;
        for (int64_t t = 0; t < m; ++t) stamp[verts[t]] = mine;
        auto bfs = [&](int64_t s, std::vector<int64_t>& order, int64_t& ecc) {
            order.clear();
            for (int64_t t = 0; t < m; ++t) dist[verts[t]] = -1;
            dist[s] = 0;
            order.push_back(s);
            for (size_t h = 0; h < order.size(); ++h) {
                const int64_t v = order[h];
                for (int64_t e = g.ptr[v]; e < g.ptr[v + 1]; ++e) {
                    const int64_t w = g.adj[e];
                    if (stamp[w] == mine && dist[w] < 0) {
                        dist[w] = dist[v] + 1;
                        order.push_back(w);
                    }
                }
            }
            ecc = dist[order.back()];
        };
        auto degIn = [&](int64_t v) {
            int64_t d = 0;
            for (int64_t e = g.ptr[v]; e < g.ptr[v + 1]; ++e) d += stamp[g.adj[e]] == mine;
            return d;
        };
        std::vector<int64_t> order;
        int64_t s = verts[0], ecc = 0;
        bfs(s, order, ecc);
        for (int it = 0; it < 8; ++it) {
            int64_t best = -1, bd = 0;
            for (auto rit = order.rbegin(); rit != order.rend() && dist[*rit] == ecc; ++rit) {
                const int64_t d = degIn(*rit);
                if (best < 0 || d < bd || (d == bd && *rit < best)) { best = *rit; bd = d; }
            }
            std::vector<int64_t> o2;
            int64_t e2 = 0;
            bfs(best, o2, e2);
            if (e2 <= ecc) { bfs(s, order, ecc); break; }
            s = best;
            order.swap(o2);
            ecc = e2;
        }
        // cumulative level counts over the reached component
        std::vector<int64_t> lc(static_cast<size_t>(ecc) + 1, 0);
        for (int64_t v : order) ++lc[dist[v]];
        int64_t acc = 0, mlev = 0;
        const int64_t half = (m + 1) / 2;
        for (int64_t l = 0; l <= ecc; ++l) {
            acc += lc[l];
            if (acc >= half) { mlev = l; break; }
        }
        side.assign(static_cast<size_t>(m), 1);
        for (int64_t t = 0; t < m; ++t) {
            const int64_t dv = dist[verts[t]];
            if (dv < 0) side[t] = 1;                  // other component
            else side[t] = dv < mlev ? 0 : (dv == mlev ? 2 : 1);
        }
    }
};

}  // namespace

// FM moves keep the larger part within this share of a subset: looser
// bounds trade separator size for balance, and the fronts above a lopsided
// cut grow (a 300 x 300 grid: 0.6 gave 2.4x the elimination flops of 0.55)
constexpr double kFmMaxFrac = 0.55;

PartitionResult recursiveBisection(const Graph& g, const double* coords,
                                   int nLevels, int64_t leafSize, int method) {
    PartitionResult res;
    const int64_t n = g.n;
    if (n == 0) return res;
    Bisector bis(g, coords);
    std::function<int64_t(std::vector<int64_t>&, int)> recurse =
        [&](std::vector<int64_t>& verts, int depth) -> int64_t {
        const int64_t m = static_cast<int64_t>(verts.size());
        if (m <= leafSize || depth >= nLevels || m < 4)
            return res.tree.addNode(NodeRole::Leaf, verts);
        std::vector<int> side;
        // plane separators of a hexahedral mesh are minimal already; graph
        // separators get the FM moves too where the degree is low
        // Without coordinates, low-degree graphs (grid and path Laplacians,
        // shell and frame models) take the spectral cut and FM refinement;
        // high-degree meshes take BFS level sets, where a Lanczos run per
        // bisection would cost seconds (26,460 vertices of degree 170: 4 s)
        if (method == 1) {
            bis.byLevels(verts, side);
        } else {
            const bool plane = bis.byPlane(verts, side);
            int64_t deg = 0;
            for (int64_t v : verts) deg += g.ptr[v + 1] - g.ptr[v];
            const bool lowDeg = deg <= 32 * m;
            if (!plane && !(lowDeg && bis.bySpectral(verts, side))) bis.byLevels(verts, side);
            bis.refine(verts, side, !plane && lowDeg, kFmMaxFrac);
        }
        std::vector<int64_t> s0, s1, sep;
        for (int64_t t = 0; t < m; ++t) {
            if (side[t] == 2) sep.push_back(verts[t]);
            else if (side[t] == 0) s0.push_back(verts[t]);
            else s1.push_back(verts[t]);
        }
        const int64_t minSide = static_cast<int64_t>(std::min(s0.size(), s1.size()));
        if (sep.empty() || minSide < 2)
            return res.tree.addNode(NodeRole::Leaf, verts);
        const int64_t sepId = res.tree.addNode(NodeRole::Separator, std::move(sep));
        std::vector<int64_t>().swap(verts);
        const int64_t l = recurse(s0, depth + 1);
        const int64_t r = recurse(s1, depth + 1);
        res.tree.setChildren(sepId, l, r);
        res.tree.setParent(l, sepId);
        res.tree.setParent(r, sepId);
        return sepId;
    };
    std::vector<int64_t> all(static_cast<size_t>(n));
    std::iota(all.begin(), all.end(), 0);
    recurse(all, 0);
    res.tree.relabelRoot();
    for (int64_t id : res.tree.postOrder()) {
        const auto& idx = res.tree.node(id).indices;
        res.perm.insert(res.perm.end(), idx.begin(), idx.end());
    }
    std::string why;
    if (!res.tree.validate(n, &why))
        throw std::runtime_error("partition: tree validation failed: " + why);
    return res;
}

// ---------------------------------------------------------------------------
// Symbolic analysis
// ---------------------------------------------------------------------------
Symbolic analyze(const PartitionResult& part, const Graph& g) {
    Symbolic S;
    const SeparatorTree& T = part.tree;
    S.n = g.n;
    S.nNodes = T.nNodes();
    const int64_t N = S.nNodes;
    const std::vector<int64_t> order = T.postOrder();
    std::vector<int64_t> posOf(static_cast<size_t>(N));
    for (int64_t p = 0; p < N; ++p) posOf[order[p]] = p;
    S.treeId = order;
    S.off.resize(N);
    S.size.resize(N);
    S.parent.assign(N, -1);
    S.child0.assign(N, -1);
    S.child1.assign(N, -1);
    S.level.assign(N, 0);
    S.subStart.resize(N);
    int64_t acc = 0;
    for (int64_t p = 0; p < N; ++p) {
        const TreeNode& nd = T.node(order[p]);
        S.off[p] = acc;
        S.size[p] = static_cast<int64_t>(nd.indices.size());
        acc += S.size[p];
        if (nd.parent >= 0) S.parent[p] = posOf[nd.parent];
        if (nd.children.size() == 2) {
            S.child0[p] = posOf[nd.children[0]];
            S.child1[p] = posOf[nd.children[1]];
            S.level[p] = std::max(S.level[S.child0[p]], S.level[S.child1[p]]) + 1;
            S.subStart[p] = std::min(S.subStart[S.child0[p]], S.subStart[S.child1[p]]);
        } else {
            S.subStart[p] = p;
        }
        S.maxLevel = std::max(S.maxLevel, S.level[p]);
    }
    if (acc != S.n) throw std::runtime_error("analyze: tree does not cover n");
    S.perm = part.perm;
    S.iperm.assign(static_cast<size_t>(S.n), 0);
    for (int64_t j = 0; j < S.n; ++j) S.iperm[S.perm[j]] = j;
    std::vector<int64_t> owner(static_cast<size_t>(S.n));
    for (int64_t p = 0; p < N; ++p)
        for (int64_t j = S.off[p]; j < S.off[p] + S.size[p]; ++j) owner[j] = p;

    // separator property: every edge joins a node with itself, an ancestor
    // or a descendant (subtree of q = positions [subStart[q], q])
    auto inSub = [&](int64_t p, int64_t q) { return S.subStart[q] <= p && p <= q; };
    for (int64_t v = 0; v < S.n; ++v) {
        const int64_t pv = owner[S.iperm[v]];
        for (int64_t e = g.ptr[v]; e < g.ptr[v + 1]; ++e) {
            const int64_t pw = owner[S.iperm[g.adj[e]]];
            if (pv != pw && !inSub(pv, pw) && !inSub(pw, pv))
                throw std::runtime_error("analyze: partition is not a vertex-separator tree");
        }
    }

    // boundaries, post-order (children first)
    S.bptr.assign(static_cast<size_t>(N) + 1, 0);
    std::vector<std::vector<int64_t>> B(static_cast<size_t>(N));
    std::vector<int64_t> mark(static_cast<size_t>(S.n), -1);
    for (int64_t p = 0; p < N; ++p) {
        const int64_t hiEnd = S.off[p] + S.size[p];
        auto& out = B[p];
        auto add = [&](int64_t j) {
            if (j >= hiEnd && mark[j] != p) {
                mark[j] = p;
                out.push_back(j);
            }
        };
        for (int64_t j = S.off[p]; j < hiEnd; ++j) {
            const int64_t v = S.perm[j];
            for (int64_t e = g.ptr[v]; e < g.ptr[v + 1]; ++e) add(S.iperm[g.adj[e]]);
        }
        for (int64_t c : {S.child0[p], S.child1[p]})
            if (c >= 0)
                for (int64_t j : B[c]) add(j);
        std::sort(out.begin(), out.end());
    }
    for (int64_t p = 0; p < N; ++p) S.bptr[p + 1] = S.bptr[p] + static_cast<int64_t>(B[p].size());
    S.bidx.resize(static_cast<size_t>(S.bptr[N]));
    for (int64_t p = 0; p < N; ++p) std::copy(B[p].begin(), B[p].end(), S.bidx.begin() + S.bptr[p]);
    // relative indices of each child's boundary inside the parent's front
    S.relptr = S.bptr;
    S.rel.assign(S.bidx.size(), -1);
    for (int64_t p = 0; p < N; ++p) {
        const int64_t q = S.parent[p];
        if (q < 0) {
            if (S.b(p) != 0) throw std::runtime_error("analyze: root has a boundary");
            continue;
        }
        const auto& Bq = B[q];
        for (int64_t t = S.bptr[p]; t < S.bptr[p + 1]; ++t) {
            const int64_t j = S.bidx[t];
            int64_t r;
            if (j < S.off[q]) throw std::runtime_error("analyze: boundary below parent");
            if (j < S.off[q] + S.size[q]) {
                r = j - S.off[q];
            } else {
                auto it = std::lower_bound(Bq.begin(), Bq.end(), j);
                if (it == Bq.end() || *it != j)
                    throw std::runtime_error("analyze: child boundary not in parent front");
                r = S.size[q] + (it - Bq.begin());
            }
            S.rel[t] = r;
        }
    }
    S.byLevel.assign(static_cast<size_t>(S.maxLevel) + 1, {});
    for (int64_t p = 0; p < N; ++p) S.byLevel[S.level[p]].push_back(p);
    return S;
}

Symbolic mergeSymbolic(const std::vector<Symbolic>& parts) {
    if (parts.size() == 1) return parts[0];
    Symbolic S;
    int64_t rowBase = 0, nodeBase = 0, bBase = 0;
    for (const Symbolic& P : parts) {
        S.maxLevel = std::max(S.maxLevel, P.maxLevel);
        for (int64_t p = 0; p < P.nNodes; ++p) {
            S.treeId.push_back(P.treeId[p] + nodeBase);
            S.off.push_back(P.off[p] + rowBase);
            S.size.push_back(P.size[p]);
            S.parent.push_back(P.parent[p] < 0 ? -1 : P.parent[p] + nodeBase);
            S.child0.push_back(P.child0[p] < 0 ? -1 : P.child0[p] + nodeBase);
            S.child1.push_back(P.child1[p] < 0 ? -1 : P.child1[p] + nodeBase);
            S.level.push_back(P.level[p]);
            S.subStart.push_back(P.subStart[p] + nodeBase);
            S.bptr.push_back(P.bptr[p] + bBase);
        }
        for (int64_t j : P.bidx) S.bidx.push_back(j + rowBase);
        S.rel.insert(S.rel.end(), P.rel.begin(), P.rel.end());
        for (int64_t j = 0; j < P.n; ++j) S.perm.push_back(P.perm[j] + rowBase);
        rowBase += P.n;
        nodeBase += P.nNodes;
        bBase += static_cast<int64_t>(P.bidx.size());
    }
    S.bptr.push_back(bBase);
    S.n = rowBase;
    S.nNodes = nodeBase;
    S.relptr = S.bptr;
    S.iperm.assign(static_cast<size_t>(S.n), 0);
    for (int64_t j = 0; j < S.n; ++j) S.iperm[S.perm[j]] = j;
    S.byLevel.assign(static_cast<size_t>(S.maxLevel) + 1, {});
    for (int64_t p = 0; p < S.nNodes; ++p) S.byLevel[S.level[p]].push_back(p);
    return S;
}

}  // namespace amls
