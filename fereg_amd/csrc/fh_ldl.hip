// Exact sparse LDL^T on the host: approximate minimum-degree ordering, elimination tree and column
// counts, up-looking numeric factorization with grounding, and a host solve (for tests and the
// C ABI). The GPU side (assembly of S, level-scheduled triangular solves inside the PCG) is in
// fh_escalate.hip.
//
// What the reference does here (precond/sparse_chol.rs:1-449) is faer's supernodal LDL^T with
// AMD ordering on the unit-diagonal scaled matrix with the dense block's dynamic regularization.
// This is our own simplicial form: the same scaling and grounding (they decide which coordinates
// are grounded, hence the preconditioner), a quotient-graph minimum-degree ordering with
// approximate external degrees (Amestoy, Davis and Duff, SIAM J. Matrix Anal. Appl. 17(4), 1996)
// and the up-looking LDL^T of Davis (ACM TOMS 31(4), 2005, Algorithm 849). The ordering and the
// factor values differ from faer's; the solve they give is the same grounded generalized inverse
// to rounding.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <memory>
#include <numeric>

#include "fh_ldl.hpp"

namespace fh {

namespace {

// Minimum-degree ordering of the graph (xadj, adj) (symmetric, no self loops) over nodes
// [nlast, n); nodes [0, nlast) follow in their order. Variables are eliminated by smallest
// approximate external degree; an eliminated variable becomes an element whose list is the
// union of its variable neighbours and the lists of the elements it absorbs. Rows denser than
// max(16, 10 sqrt(n)) are ordered after the others without taking part (AMD's dense-row rule).
std::vector<uint32_t> md_order(size_t n, const std::vector<uint64_t>& xadj, const std::vector<uint32_t>& adj,
                               size_t nlast) {
    enum : uint8_t { VAR = 0, ELT = 1, DEAD = 2, OUT = 3 };
    std::vector<uint8_t> state(n, VAR);
    std::vector<uint32_t> order;
    order.reserve(n);
    const double dense = std::max(16.0, 10.0 * std::sqrt((double)(n - nlast)));
    std::vector<uint32_t> dense_nodes;
    for (size_t i = 0; i < nlast; i++) state[i] = OUT;
    for (size_t i = nlast; i < n; i++)
        if ((double)(xadj[i + 1] - xadj[i]) > dense) {
            state[i] = OUT;
            dense_nodes.push_back((uint32_t)i);
        }
    std::vector<std::vector<uint32_t>> va(n), ea(n), el(n);
    std::vector<int64_t> deg(n, 0), head(n + 1, -1), nxt(n, -1), prv(n, -1);
    auto insert = [&](uint32_t i) {
        const int64_t d = deg[i];
        nxt[i] = head[d];
        prv[i] = -1;
        if (head[d] >= 0) prv[head[d]] = i;
        head[d] = i;
    };
    auto remove = [&](uint32_t i) {
        if (prv[i] >= 0) nxt[prv[i]] = nxt[i];
        else head[deg[i]] = nxt[i];
        if (nxt[i] >= 0) prv[nxt[i]] = prv[i];
    };
    size_t nvar = 0;
    for (size_t i = nlast; i < n; i++) {
        if (state[i] != VAR) continue;
        for (uint64_t e = xadj[i]; e < xadj[i + 1]; e++) {
            const uint32_t j = adj[e];
            if (j != i && state[j] == VAR) va[i].push_back(j);
        }
        deg[i] = (int64_t)va[i].size();
        nvar++;
    }
    for (size_t i = n; i-- > nlast;)
        if (state[i] == VAR) insert((uint32_t)i);
    std::vector<uint64_t> mark(n, 0), wst(n, 0);
    std::vector<int64_t> w(n, 0);
    uint64_t tag = 0;
    int64_t mindeg = 0;
    std::vector<uint32_t> lp;
    for (size_t step = 0; step < nvar; step++) {
        while (head[mindeg] < 0) mindeg++;
        const uint32_t p = (uint32_t)head[mindeg];
        remove(p);
        order.push_back(p);
        state[p] = ELT;
        mark[p] = ++tag;
        // the new element's list: p's variable neighbours and the lists of its elements, which
        // it absorbs
        lp.clear();
        for (uint32_t v : va[p])
            if (state[v] == VAR && mark[v] != tag) {
                mark[v] = tag;
                lp.push_back(v);
            }
        for (uint32_t e : ea[p]) {
            if (state[e] != ELT) continue;
            for (uint32_t v : el[e])
                if (state[v] == VAR && mark[v] != tag) {
                    mark[v] = tag;
                    lp.push_back(v);
                }
            state[e] = DEAD;
            std::vector<uint32_t>().swap(el[e]);
        }
        std::vector<uint32_t>().swap(va[p]);
        std::vector<uint32_t>().swap(ea[p]);
        for (uint32_t i : lp) remove(i);
        // w[e] = |L_e \ L_p| for every live element next to L_p
        for (uint32_t i : lp)
            for (uint32_t e : ea[i]) {
                if (state[e] != ELT) continue;
                if (wst[e] != tag) {
                    std::vector<uint32_t>& L = el[e];
                    size_t o = 0;
                    for (uint32_t v : L)
                        if (state[v] == VAR) L[o++] = v;
                    L.resize(o);
                    w[e] = (int64_t)o;
                    wst[e] = tag;
                }
                w[e] -= 1;
            }
        const int64_t left = (int64_t)(nvar - step - 1), lsz = (int64_t)lp.size();
        for (uint32_t i : lp) {
            std::vector<uint32_t>& E = ea[i];
            size_t o = 0;
            int64_t ext = 0;
            for (uint32_t e : E) {
                if (state[e] != ELT) continue;
                if (w[e] <= 0) {  // L_e inside L_p: absorbed by p
                    state[e] = DEAD;
                    std::vector<uint32_t>().swap(el[e]);
                    continue;
                }
                E[o++] = e;
                ext += w[e];
            }
            E.resize(o);
            E.push_back(p);
            std::vector<uint32_t>& A = va[i];
            o = 0;
            for (uint32_t v : A)
                if (state[v] == VAR && mark[v] != tag) A[o++] = v;
            A.resize(o);
            int64_t d = (int64_t)A.size() + (lsz - 1) + ext;
            d = std::min(d, deg[i] + lsz - 1);
            d = std::max<int64_t>(std::min(d, left - 1), 0);
            deg[i] = d;
            insert(i);
            mindeg = std::min(mindeg, d);
        }
        el[p] = lp;
    }
    for (uint32_t i : dense_nodes) order.push_back(i);
    for (size_t i = 0; i < nlast; i++) order.push_back((uint32_t)i);
    return order;
}

// Nested dissection by BFS level structures (George 1973, in the form of George and Liu's
// "automatic nested dissection"): a part is cut at the middle BFS level from a pseudo-peripheral
// node; both sides are ordered first (recursively), the separating level last. Parts of at most
// ND_LEAF nodes, and parts a level cut cannot split, are ordered by minimum degree. The
// elimination tree is then about log2(n) separators deep instead of the long chains minimum
// degree leaves on path- and ring-like graphs, which matters for the level-scheduled GPU solves.
constexpr size_t ND_LEAF = 64;

std::vector<uint32_t> nd_order(size_t n, const std::vector<uint64_t>& xadj, const std::vector<uint32_t>& adj,
                               size_t nlast) {
    const double dense = std::max(16.0, 10.0 * std::sqrt((double)(n - nlast)));
    std::vector<uint8_t> out(n, 0);  // 1: not part of the dissection (last group or dense row)
    std::vector<uint32_t> dense_nodes;
    for (size_t i = 0; i < nlast; i++) out[i] = 1;
    for (size_t i = nlast; i < n; i++)
        if ((double)(xadj[i + 1] - xadj[i]) > dense) {
            out[i] = 1;
            dense_nodes.push_back((uint32_t)i);
        }
    std::vector<uint32_t> order;
    order.reserve(n);
    std::vector<int64_t> part(n, -1);  // current part id of each node
    std::vector<int64_t> lev(n, -1);
    std::vector<uint32_t> q;
    // BFS over the nodes of part `pid` from `s`; returns the visit order (levels in lev)
    auto bfs = [&](uint32_t s0, int64_t pid, std::vector<uint32_t>& vis) {
        vis.clear();
        vis.push_back(s0);
        lev[s0] = 0;
        for (size_t h = 0; h < vis.size(); h++) {
            const uint32_t u = vis[h];
            for (uint64_t e = xadj[u]; e < xadj[u + 1]; e++) {
                const uint32_t v = adj[e];
                if (part[v] == pid && lev[v] < 0) {
                    lev[v] = lev[u] + 1;
                    vis.push_back(v);
                }
            }
        }
    };
    // minimum degree on a small induced subgraph, appended to the order
    auto leaf = [&](const std::vector<uint32_t>& nodes, int64_t pid) {
        std::vector<uint32_t> loc(nodes.size());
        std::vector<uint64_t> sx(nodes.size() + 1, 0);
        std::vector<uint32_t> sa;
        std::vector<int64_t> idx(0);
        for (size_t k = 0; k < nodes.size(); k++) lev[nodes[k]] = (int64_t)k;  // local index
        for (size_t k = 0; k < nodes.size(); k++) {
            const uint32_t u = nodes[k];
            for (uint64_t e = xadj[u]; e < xadj[u + 1]; e++)
                if (part[adj[e]] == pid) sa.push_back((uint32_t)lev[adj[e]]);
            sx[k + 1] = sa.size();
        }
        for (uint32_t k : md_order(nodes.size(), sx, sa, 0)) order.push_back(nodes[k]);
        for (uint32_t u : nodes) {
            lev[u] = -1;
            part[u] = -2;  // ordered
        }
    };
    // work stack of parts (node lists); a part is ordered after everything pushed above it, so the
    // separator is emitted once both sides are done: (nodes, is_separator)
    struct Item {
        std::vector<uint32_t> nodes;
        bool sep;
    };
    std::vector<Item> stack;
    int64_t next_pid = 0;
    {
        std::vector<uint32_t> all;
        for (size_t i = 0; i < n; i++)
            if (!out[i]) all.push_back((uint32_t)i);
        stack.push_back({std::move(all), false});
    }
    std::vector<uint32_t> vis;
    while (!stack.empty()) {
        Item it = std::move(stack.back());
        stack.pop_back();
        if (it.sep) {
            for (uint32_t u : it.nodes) {
                order.push_back(u);
                part[u] = -2;
            }
            continue;
        }
        const int64_t pid = next_pid++;
        for (uint32_t u : it.nodes) part[u] = pid;
        if (it.nodes.size() <= ND_LEAF) {
            leaf(it.nodes, pid);
            continue;
        }
        // one connected component from the first node; the rest of the part is its own part
        bfs(it.nodes[0], pid, vis);
        if (vis.size() < it.nodes.size()) {
            std::vector<uint32_t> rest;
            for (uint32_t u : it.nodes)
                if (lev[u] < 0) rest.push_back(u);
            for (uint32_t u : vis) lev[u] = -1;
            for (uint32_t u : rest) part[u] = -1;
            stack.push_back({std::move(rest), false});
            it.nodes = vis;
            if (it.nodes.size() <= ND_LEAF) {
                leaf(it.nodes, pid);
                continue;
            }
        }
        // pseudo-peripheral start: BFS again from the last node reached while the depth grows
        uint32_t s0 = vis.back();
        int64_t depth = lev[vis.back()];
        for (int rep = 0; rep < 4; rep++) {
            for (uint32_t u : vis) lev[u] = -1;
            bfs(s0, pid, vis);
            const int64_t d = lev[vis.back()];
            if (d <= depth && rep > 0) break;
            depth = d;
            s0 = vis.back();
        }
        depth = lev[vis.back()];  // the last BFS's depth (its levels are in lev)
        if (depth < 2) {  // no level separates two non-empty sides
            for (uint32_t u : vis) lev[u] = -1;
            leaf(vis, pid);
            continue;
        }
        // the level where half of the part has been reached
        std::vector<size_t> cnt(depth + 1, 0);
        for (uint32_t u : vis) cnt[lev[u]]++;
        size_t acc = 0;
        int64_t L = 1;
        for (int64_t l = 0; l <= depth; l++) {
            acc += cnt[l];
            if (2 * acc >= vis.size()) {
                L = std::min<int64_t>(std::max<int64_t>(l, 1), depth - 1);
                break;
            }
        }
        std::vector<uint32_t> A, B, Sp;
        for (uint32_t u : vis) (lev[u] < L ? A : lev[u] > L ? B : Sp).push_back(u);
        for (uint32_t u : vis) {
            lev[u] = -1;
            part[u] = -1;
        }
        stack.push_back({std::move(Sp), true});
        stack.push_back({std::move(B), false});
        stack.push_back({std::move(A), false});
    }
    for (uint32_t i : dense_nodes) order.push_back(i);
    for (size_t i = 0; i < nlast; i++) order.push_back((uint32_t)i);
    return order;
}

// Symbolic analysis of the ordering perm: the permuted upper triangle by columns, the elimination
// tree and the column counts (Liu: walking from each entry of column k up the tree until a node
// already visited for k). Returns the elimination tree's height.
uint64_t symbolic(const SymLower& a, const std::vector<uint32_t>& perm, LdlFactor& f) {
    const size_t n = a.n;
    std::vector<uint32_t> inv(n);
    for (size_t k = 0; k < n; k++) inv[perm[k]] = (uint32_t)k;
    f.perm = perm;
    // entry (i, j), j <= i, goes to column max(inv i, inv j) at row min(...)
    f.up.assign(n + 1, 0);
    for (size_t i = 0; i < n; i++)
        for (uint64_t e = a.ptr[i]; e < a.ptr[i + 1]; e++) f.up[std::max(inv[i], inv[a.col[e]]) + 1]++;
    for (size_t k = 0; k < n; k++) f.up[k + 1] += f.up[k];
    f.uri.resize(f.up[n]);
    f.upos.resize(f.up[n]);
    {
        std::vector<uint64_t> fill(f.up.begin(), f.up.end() - 1);
        for (size_t i = 0; i < n; i++)
            for (uint64_t e = a.ptr[i]; e < a.ptr[i + 1]; e++) {
                const uint32_t pi = inv[i], pj = inv[a.col[e]];
                const uint32_t c = std::max(pi, pj);
                f.uri[fill[c]] = std::min(pi, pj);
                f.upos[fill[c]++] = e;
            }
    }
    f.parent.assign(n, -1);
    std::vector<int64_t> flag(n, -1);
    std::vector<uint64_t> lnz(n, 0);
    for (size_t k = 0; k < n; k++) {
        flag[k] = (int64_t)k;
        for (uint64_t e = f.up[k]; e < f.up[k + 1]; e++) {
            int64_t i = f.uri[e];
            while (flag[i] != (int64_t)k) {
                if (f.parent[i] < 0) f.parent[i] = (int64_t)k;
                lnz[i]++;
                flag[i] = (int64_t)k;
                i = f.parent[i];
            }
        }
    }
    f.lp.assign(n + 1, 0);
    double flops = 0.0;
    for (size_t k = 0; k < n; k++) {
        f.lp[k + 1] = f.lp[k] + lnz[k];
        flops += (double)lnz[k] * (double)lnz[k];
    }
    f.flops = flops;
    std::vector<uint64_t> h(n, 1);
    uint64_t height = n ? 1 : 0;
    for (size_t i = 0; i < n; i++) {  // parent[i] > i: a node's height is final when it is reached
        height = std::max(height, h[i]);
        if (f.parent[i] >= 0) h[f.parent[i]] = std::max(h[f.parent[i]], h[i] + 1);
    }
    return height;
}

}  // namespace

bool ldl_analyze(const SymLower& a, size_t nlast, uint64_t max_len, double max_flops, LdlFactor& f) {
    const size_t n = a.n;
    f = LdlFactor();
    f.n = n;
    // full adjacency (both triangles, no diagonal) and the diagonal's position in each row
    std::vector<uint64_t> xadj(n + 1, 0);
    f.diag_pos.assign(n, UINT64_MAX);
    for (size_t i = 0; i < n; i++)
        for (uint64_t e = a.ptr[i]; e < a.ptr[i + 1]; e++) {
            const uint32_t j = a.col[e];
            if (j == i) {
                f.diag_pos[i] = e;
            } else {
                xadj[i + 1]++;
                xadj[j + 1]++;
            }
        }
    for (size_t i = 0; i < n; i++) xadj[i + 1] += xadj[i];
    std::vector<uint32_t> adj(xadj[n]);
    {
        std::vector<uint64_t> fill(xadj.begin(), xadj.end() - 1);
        for (size_t i = 0; i < n; i++)
            for (uint64_t e = a.ptr[i]; e < a.ptr[i + 1]; e++) {
                const uint32_t j = a.col[e];
                if (j == i) continue;
                adj[fill[i]++] = j;
                adj[fill[j]++] = (uint32_t)i;
            }
    }
    // minimum degree (least fill); when its elimination tree is tall, nested dissection too, and
    // the ordering whose level-scheduled solve is estimated cheaper within the budgets is kept:
    // a level costs about one dependent round of loads (~2.5 us per direction), an entry a few
    // bytes of streaming
    auto apply_cost = [](uint64_t height, uint64_t len) { return 5e-6 * (double)height + 2e-10 * (double)len; };
    const uint64_t h_md = symbolic(a, md_order(n, xadj, adj, nlast), f);
    f.height = h_md;
    if (h_md > 2 * ND_LEAF) {
        LdlFactor g;
        g.n = n;
        g.diag_pos = f.diag_pos;
        const uint64_t h_nd = symbolic(a, nd_order(n, xadj, adj, nlast), g);
        g.height = h_nd;
        const bool md_fits = f.lp[n] <= max_len && f.flops <= max_flops;
        const bool nd_fits = g.lp[n] <= max_len && g.flops <= max_flops;
        if (nd_fits && (!md_fits || apply_cost(h_nd, g.lp[n]) < apply_cost(h_md, f.lp[n]))) f = std::move(g);
    }
    const double flops = f.flops;
    if (f.lp[n] > max_len || flops > max_flops) return false;
    return true;
}

bool ldl_numeric(const SymLower& a, LdlFactor& f) {
    const size_t n = f.n;
    std::vector<double> s(n);
    for (size_t i = 0; i < n; i++) {
        const double dii = f.diag_pos[i] == UINT64_MAX ? 0.0 : a.val[f.diag_pos[i]];
        s[i] = dii > 0.0 ? 1.0 / std::sqrt(dii) : 1.0;
    }
    f.scale.resize(n);
    for (size_t k = 0; k < n; k++) f.scale[k] = s[f.perm[k]];
    const uint64_t len = f.lp[n];
    f.li.assign(len, 0);
    f.lx.assign(len, 0.0);
    f.d.assign(n, 0.0);
    f.grounded = 0;
    std::vector<double> y(n, 0.0);
    std::vector<int64_t> flag(n, -1);
    std::vector<uint32_t> pattern(n);
    std::vector<uint64_t> lnz(n, 0);
    for (size_t k = 0; k < n; k++) {
        // row k of L: the sparse solve L[0:k, 0:k] y = A[0:k, k] over the nodes reached from
        // column k's entries in the elimination tree, in topological order
        flag[k] = (int64_t)k;
        size_t top = n;
        const double sk = f.scale[k];
        for (uint64_t e = f.up[k]; e < f.up[k + 1]; e++) {
            uint32_t i = f.uri[e];
            y[i] += a.val[f.upos[e]] * f.scale[i] * sk;
            size_t len_i = 0;
            while (flag[i] != (int64_t)k) {
                pattern[len_i++] = i;
                flag[i] = (int64_t)k;
                i = (uint32_t)f.parent[i];
            }
            while (len_i > 0) pattern[--top] = pattern[--len_i];
        }
        double dk = y[k];
        y[k] = 0.0;
        for (; top < n; top++) {
            const uint32_t i = pattern[top];
            const double yi = y[i];
            y[i] = 0.0;
            const uint64_t p0 = f.lp[i], p1 = f.lp[i] + lnz[i];
            for (uint64_t p = p0; p < p1; p++) y[f.li[p]] -= f.lx[p] * yi;
            const double lki = yi / f.d[i];
            dk -= lki * yi;
            f.li[p1] = (uint32_t)k;
            f.lx[p1] = lki;
            lnz[i]++;
        }
        if (!std::isfinite(dk)) return false;
        if (dk <= LDL_PIVOT_TOL) {  // a null direction: grounded (precond/dense.rs:26-40)
            dk = LDL_GROUND;
            f.grounded++;
        }
        f.d[k] = dk;
    }
    return true;
}

void ldl_solve_host(const LdlFactor& f, double* b, size_t nrhs, size_t ld) {
    const size_t n = f.n;
    std::vector<double> y(n * nrhs);
    for (size_t k = 0; k < n; k++)
        for (size_t c = 0; c < nrhs; c++) y[k * nrhs + c] = f.scale[k] * b[(size_t)f.perm[k] * ld + c];
    for (size_t j = 0; j < n; j++)
        for (uint64_t p = f.lp[j]; p < f.lp[j + 1]; p++)
            for (size_t c = 0; c < nrhs; c++) y[(size_t)f.li[p] * nrhs + c] -= f.lx[p] * y[j * nrhs + c];
    for (size_t k = 0; k < n; k++)
        for (size_t c = 0; c < nrhs; c++) y[k * nrhs + c] /= f.d[k];
    for (size_t j = n; j-- > 0;)
        for (uint64_t p = f.lp[j]; p < f.lp[j + 1]; p++)
            for (size_t c = 0; c < nrhs; c++) y[j * nrhs + c] -= f.lx[p] * y[(size_t)f.li[p] * nrhs + c];
    for (size_t k = 0; k < n; k++)
        for (size_t c = 0; c < nrhs; c++) b[(size_t)f.perm[k] * ld + c] = f.scale[k] * y[k * nrhs + c];
}

double selinv_at(const LdlFactor& f, const std::vector<double>& zd, const std::vector<double>& zx, uint32_t i,
                 uint32_t j) {
    if (i == j) return zd[i];
    if (i < j) std::swap(i, j);  // lower triangle: column j, row i > j
    const uint32_t* b = f.li.data() + f.lp[j];
    const uint32_t* e = f.li.data() + f.lp[j + 1];
    const uint32_t* p = std::lower_bound(b, e, i);
    return (p != e && *p == i) ? zx[f.lp[j] + (p - b)] : 0.0;
}

void ldl_selected_inverse(const LdlFactor& f, std::vector<double>& zd, std::vector<double>& zx) {
    const size_t n = f.n;
    zd.assign(n, 0.0);
    zx.assign(f.len(), 0.0);
    // columns from the last: Z_ij = -sum_k L_kj Z_ik and Z_jj = 1/d_j - sum_k L_kj Z_kj over the
    // rows k of column j; every Z_ik needed lies in columns > j, already done
    std::vector<double> acc;
    for (size_t j = n; j-- > 0;) {
        const uint64_t p0 = f.lp[j], p1 = f.lp[j + 1];
        acc.assign(p1 - p0, 0.0);
        for (uint64_t a = p0; a < p1; a++) {
            const uint32_t i = f.li[a];
            double s = 0.0;
            for (uint64_t b = p0; b < p1; b++) s += f.lx[b] * selinv_at(f, zd, zx, i, f.li[b]);
            acc[a - p0] = -s;
        }
        double dj = 1.0 / f.d[j];
        for (uint64_t a = p0; a < p1; a++) {
            zx[a] = acc[a - p0];
            dj -= f.lx[a] * zx[a];
        }
        zd[j] = dj;
    }
}

}  // namespace fh

// ------------------------------------------------------------------ C ABI (host only)

struct fh_sparse_chol {
    fh::SymLower a;
    fh::LdlFactor f;
};

extern "C" {

int fh_sparse_chol_factor(size_t n, const uint64_t* ptr, const uint32_t* col, const double* val, size_t nlast,
                          fh_sparse_chol** out) {
    try {
        fh::require(out != nullptr && ptr != nullptr && (n == 0 || (col != nullptr && val != nullptr)),
                    "fh_sparse_chol_factor: null argument");
        fh::require(nlast <= n, "fh_sparse_chol_factor: nlast > n");
        *out = nullptr;
        auto h = new fh_sparse_chol();
        std::unique_ptr<fh_sparse_chol> g(h);
        h->a.n = n;
        h->a.ptr.assign(ptr, ptr + n + 1);
        fh::require(h->a.ptr[0] == 0, "fh_sparse_chol_factor: ptr[0] != 0");
        for (size_t i = 0; i < n; i++) {
            fh::require(h->a.ptr[i + 1] >= h->a.ptr[i], "fh_sparse_chol_factor: ptr not ascending");
            for (uint64_t e = h->a.ptr[i]; e < h->a.ptr[i + 1]; e++) {
                fh::require(col[e] <= i, "fh_sparse_chol_factor: an entry above the diagonal");
                fh::require(e == h->a.ptr[i] || col[e] > col[e - 1], "fh_sparse_chol_factor: columns not ascending");
            }
        }
        const uint64_t nnz = h->a.ptr[n];
        h->a.col.assign(col, col + nnz);
        h->a.val.assign(val, val + nnz);
        fh::require(fh::ldl_analyze(h->a, nlast, std::numeric_limits<uint64_t>::max(), INFINITY, h->f),
                    "fh_sparse_chol_factor: analysis failed");
        fh::require(fh::ldl_numeric(h->a, h->f), "fh_sparse_chol_factor: non-finite pivot", FH_INTERNAL);
        *out = g.release();
        return FH_OK;
    } catch (const fh::Error& e) {
        fh::set_last_error(e.what());
        return e.rc;
    } catch (const std::bad_alloc&) {
        fh::set_last_error("out of host memory");
        return FH_OOM;
    } catch (const std::exception& e) {
        fh::set_last_error(e.what());
        return FH_INTERNAL;
    }
}

int fh_sparse_chol_info(const fh_sparse_chol* f, uint64_t* len, double* flops, uint64_t* grounded, uint64_t* height) {
    if (!f) {
        fh::set_last_error("fh_sparse_chol_info: null handle");
        return FH_INVALID;
    }
    if (len) *len = f->f.len();
    if (flops) *flops = f->f.flops;
    if (grounded) *grounded = f->f.grounded;
    if (height) *height = f->f.height;
    return FH_OK;
}

int fh_sparse_chol_solve(const fh_sparse_chol* f, double* b, size_t nrhs) {
    if (!f || (!b && f->f.n * nrhs > 0)) {
        fh::set_last_error("fh_sparse_chol_solve: null argument");
        return FH_INVALID;
    }
    fh::ldl_solve_host(f->f, b, nrhs, nrhs);
    return FH_OK;
}

int fh_sparse_chol_inverse_entries(const fh_sparse_chol* f, size_t npairs, const uint32_t* rows,
                                   const uint32_t* cols, double* out) {
    try {
        fh::require(f != nullptr && (npairs == 0 || (rows && cols && out)), "fh_sparse_chol_inverse_entries: null argument");
        std::vector<double> zd, zx;
        fh::ldl_selected_inverse(f->f, zd, zx);
        std::vector<uint32_t> pos(f->f.n);
        for (size_t k = 0; k < f->f.n; k++) pos[f->f.perm[k]] = (uint32_t)k;
        for (size_t q = 0; q < npairs; q++) {
            fh::require(rows[q] < f->f.n && cols[q] < f->f.n, "fh_sparse_chol_inverse_entries: index out of range");
            const uint32_t i = pos[rows[q]], j = pos[cols[q]];
            out[q] = f->f.scale[i] * f->f.scale[j] * fh::selinv_at(f->f, zd, zx, i, j);
        }
        return FH_OK;
    } catch (const fh::Error& e) {
        fh::set_last_error(e.what());
        return e.rc;
    } catch (const std::exception& e) {
        fh::set_last_error(e.what());
        return FH_INTERNAL;
    }
}

void fh_sparse_chol_destroy(fh_sparse_chol* f) { delete f; }

}  // extern "C"
