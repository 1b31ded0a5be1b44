// tf_api.hip -- C ABI of the MI355X Newton-Raphson step: set-up, state transfer, NR driver.
#include "tf_internal.hpp"
#include <array>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <numeric>

static thread_local std::string g_create_error;

namespace {

template <typename T>
T* dev_alloc(size_t n)
{
    if (n == 0) n = 1;
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, n * sizeof(T));
    if (e != hipSuccess) throw tf_hip_error(e, "hipMalloc", __FILE__, __LINE__);
    return (T*)p;
}

// Uploads complete before returning: callers often pass temporaries.
template <typename T>
void h2d(tf_handle* h, T* dst, const T* src, size_t n)
{
    if (!n) return;
    TF_HIP(hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyHostToDevice, h->stream));
    TF_HIP(hipStreamSynchronize(h->stream));
}
template <typename T>
void d2h(tf_handle* h, T* dst, const T* src, size_t n)
{
    if (n) TF_HIP(hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyDeviceToHost, h->stream));
}

// SoA transposition: out[k*E + e] = in[e*K + k]
template <typename T>
std::vector<T> to_soa(const T* in, int E, int K)
{
    std::vector<T> out((size_t)E * K);
    for (int e = 0; e < E; ++e)
        for (int k = 0; k < K; ++k) out[(size_t)k * E + e] = in[(size_t)e * K + k];
    return out;
}
template <typename T>
std::vector<T> from_soa(const T* in, int E, int K)
{
    std::vector<T> out((size_t)E * K);
    for (int e = 0; e < E; ++e)
        for (int k = 0; k < K; ++k) out[(size_t)e * K + k] = in[(size_t)k * E + e];
    return out;
}

int fail(tf_handle* h, int code, const std::string& msg)
{
    if (h) h->err = msg;
    else g_create_error = msg;
    return code;
}

#define TF_GUARD_BEGIN try {
#define TF_GUARD_END(h)                                                                             \
    }                                                                                               \
    catch (const tf_hip_error& ex) {                                                                \
        return fail(h, ex.code == hipErrorOutOfMemory ? TF_OOM : TF_HIP_ERROR, ex.what);            \
    }                                                                                               \
    catch (const std::exception& ex) { return fail(h, TF_INVALID, ex.what()); }

void free_all(tf_handle* h)
{
    drop_krylov_graphs(h);  // before the buffers the captured chunks use
    void* ptrs[] = {h->conn, h->xref, h->fmw, h->fg, h->invfg, h->lam, h->mu, h->eta, h->rate, h->ginc, h->exposed, h->elem_seg,
                    h->uk, h->un, h->g, h->du, h->eta_ext, h->fixed, h->sa, h->face_area, h->kelem, h->felem,
                    h->status, h->row_ptr, h->col, h->blk_row, h->diag_blk, h->val, h->val32, h->dinv, h->blk_ptr,
                    h->blk_src, h->node_ptr, h->node_src, h->tile_r0, h->tile_r1, h->tile_seg, h->seg_tile_ptr,
                    h->seg_node_ptr, h->seg_counter, h->partial, h->seg, h->seg_norm, h->kr, h->kr0, h->kp, h->kph,
                    h->kv, h->ks, h->ksh, h->kt, h->kx, h->km, h->kq, h->fold_ptr, h->fold_tgt, h->fold_src, h->fold_const,
                    h->val_base, h->bind_nodes, h->bind_m_ptr, h->bind_m_dof, h->bind_s_dof, h->bind_pairs,
                    h->ablated, h->blk_tr, h->rot_tz, h->val_t, h->g_save, h->adj_xbar, h->adj_rhs, h->adj_pass,
                    h->adj_y, h->adj_un_bar, h->adj_invfg_bar, h->adj_fg_bar, h->adj_fg_prev, h->adj_rate_bar,
                    h->adj_lam_bar, h->adj_mu_bar, h->adj_lam_acc, h->adj_mu_acc, h->adj_tz, h->seg_rtol2};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    packing_free(h);
    growth_functions_free(h);
    warm_start_free(h);
    twolevel_free(h);
    if (h->h_seg) (void)hipHostFree(h->h_seg);
    if (h->h_seg_norm) (void)hipHostFree(h->h_seg_norm);
    if (h->tile_info) (void)hipFree(h->tile_info);
    if (h->h_status) (void)hipHostFree(h->h_status);
    if (h->hp_rtol2) (void)hipHostFree(h->hp_rtol2);
    if (h->nr_rep) (void)hipFree(h->nr_rep);
    if (h->hp_nr_rep) (void)hipHostFree(h->hp_nr_rep);
    if (h->hp_pos) (void)hipHostFree(h->hp_pos);
    drop_krylov_graphs(h);
    for (KernelTimer* t : {&h->t_elem, &h->t_gather, &h->t_spmv}) {
        for (hipEvent_t e : t->pending) (void)hipEventDestroy(e);
        for (hipEvent_t e : t->sample)
            if (e) (void)hipEventDestroy(e);
    }
    if (h->stream) (void)hipStreamDestroy(h->stream);
}

__global__ void k_seg_begin(SegScalars* seg, int S)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < S) { seg[s].active = 1; seg[s].converged = 0; seg[s].stalled = 0; }
}
__global__ void k_seg_nr_check(SegScalars* seg, const double* norm, double tol, int S)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < S && seg[s].active && norm[s] <= tol) seg[s].active = 0;
}
__global__ void k_seg_deactivate(SegScalars* seg, int S)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < S) seg[s].active = 0;
}
// a tissue whose stalled solve is unusable (solve_unusable) leaves the NR loop without its update
__global__ void k_seg_drop_stalled(SegScalars* seg, int S)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < S && seg[s].active && seg[s].stalled && solve_unusable(seg[s].rnorm2, seg[s].bnorm2)) seg[s].active = 0;
}
// the assembly's flags: status[0] inverted Gauss points, status[3] non-finite det Fe
__global__ void k_zero_assembly_status(int32_t* status)
{
    if (threadIdx.x == 0) { status[0] = 0; status[3] = 0; }
}
__global__ void k_zero_int(int32_t* p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// The Newton iteration's gate (one workgroup), after the first round of its linear solve: a
// snapshot of every tissue's Krylov scalars and of the assembly's status into the report, and
// ok = the assembly found no inverted or non-finite Gauss point, and every active tissue's solve
// converged, or stalled with a usable residual, on a finite right-hand side.  k_update_uk and
// k_nr_finish apply the Newton update only when ok; otherwise the host finishes the solve.
__global__ __launch_bounds__(256) void k_nr_gate(const SegScalars* __restrict__ seg, int S,
                                                 const int32_t* __restrict__ status, NrReport* __restrict__ rep)
{
    __shared__ int fl[3];
    if (threadIdx.x < 3) fl[threadIdx.x] = 0;
    __syncthreads();
    SegScalars* rs = nr_report_seg(rep);
    int unf = 0, unu = 0, nan = 0;
    for (int s = threadIdx.x; s < S; s += blockDim.x) {
        const SegScalars q = seg[s];
        rs[s] = q;
        if (!q.active) continue;
        unf |= !(q.converged || q.stalled);
        unu |= q.stalled && solve_unusable(q.rnorm2, q.bnorm2);
        nan |= !isfinite(q.bnorm2);
    }
    if (unf) atomicOr(&fl[0], 1);
    if (unu) atomicOr(&fl[1], 1);
    if (nan) atomicOr(&fl[2], 1);
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 0; k < 4; ++k) rep->status[k] = status[k];
        rep->unfinished = fl[0];
        rep->unusable = fl[1];
        rep->nan = fl[2];
        rep->ok = status[0] == 0 && status[3] == 0 && !fl[0] && !fl[1] && !fl[2];
    }
}

// After the gated update: every tissue's ||dU|| into the report and, when the gate was open, the
// Newton convergence test (k_seg_nr_check)
__global__ void k_nr_finish(SegScalars* seg, const double* norm, double tol, int S, NrReport* rep)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= S) return;
    nr_report_norm(rep, S)[s] = norm[s];
    if (rep->ok && seg[s].active && norm[s] <= tol) seg[s].active = 0;
}

}  // namespace

// ---------------------------------------------------------------------------------------------
// timers
void timer_begin(tf_handle* h, KernelTimer& t, hipEvent_t* ev)
{
    if (!t.enabled || t.capturing) return;
    TF_HIP(hipEventCreate(&ev[0]));
    TF_HIP(hipEventCreate(&ev[1]));
    TF_HIP(hipEventRecord(ev[0], h->stream));
}
void timer_end(tf_handle* h, KernelTimer& t, hipEvent_t* ev)
{
    if (!t.enabled || t.capturing) return;
    TF_HIP(hipEventRecord(ev[1], h->stream));
    t.pending.push_back(ev[0]);
    t.pending.push_back(ev[1]);
}
void timer_collect(KernelTimer& t)
{
    if (t.pending.empty()) return;
    TF_HIP(hipEventSynchronize(t.pending.back()));
    for (size_t i = 0; i + 1 < t.pending.size(); i += 2) {
        float ms = 0.f;
        TF_HIP(hipEventElapsedTime(&ms, t.pending[i], t.pending[i + 1]));
        t.total_ms += ms;
        t.launches += 1;
        (void)hipEventDestroy(t.pending[i]);
        (void)hipEventDestroy(t.pending[i + 1]);
    }
    t.pending.clear();
}

// ---------------------------------------------------------------------------------------------
// Row tiles: runs of block rows of one tissue with at most tile_cap blocks and wg/3 rows (so a
// tile's DoFs never outnumber the lanes of a workgroup).  Rebuilt when the tiling is retuned.
void drop_krylov_graphs(tf_handle* h)
{
    for (int k = 0; k < 4; ++k)
        if (h->krylov_graph[k]) {
            (void)hipGraphExecDestroy(h->krylov_graph[k]);
            h->krylov_graph[k] = nullptr;
        }
    for (tf_handle::NrGraph& g : h->nr_graphs) (void)hipGraphExecDestroy(g.exec);
    h->nr_graphs.clear();
}

// LDS of the single-kernel CG for a tiling: products (3 per block), two exchange rows, and the
// reduction scratch (well under 2 KB)
bool tiling_fits(const tf_handle* h, int wg, int tile_cap)
{
    return sizeof(double) * (3 * (size_t)tile_cap + 2 * (size_t)wg) + 2048 <= (size_t)h->max_lds;
}

void build_tiles(tf_handle* h, int wg, int tile_cap)
{
    if (!(wg == 256 || wg == 512 || wg == 1024) || tile_cap < 1 || tile_cap > TF_TILE_CAP_MAX || tile_cap > 4 * wg)
        throw std::runtime_error("unsupported tiling: wg in {256, 512, 1024}, 1 <= tile_cap <= min(4 wg, 2048)");
    if (!tiling_fits(h, wg, tile_cap))
        throw std::runtime_error("tiling needs more LDS than a workgroup may allocate on this device");
    const std::vector<int32_t>& row_ptr = h->h_row_ptr;
    std::vector<int32_t> tr0, tr1, tseg, seg_tile_ptr(h->S + 1, 0);
    const int max_rows = wg / 3;
    for (int s = 0; s < h->S; ++s) {
        int r = h->h_node_seg_ptr[s];
        const int rend = h->h_node_seg_ptr[s + 1];
        while (r < rend) {
            int r1 = r, blocks = 0;
            while (r1 < rend && r1 - r < max_rows && blocks + (row_ptr[r1 + 1] - row_ptr[r1]) <= tile_cap) {
                blocks += row_ptr[r1 + 1] - row_ptr[r1];
                ++r1;
            }
            if (r1 == r) throw std::runtime_error("a block row exceeds the tile capacity");
            tr0.push_back(r); tr1.push_back(r1); tseg.push_back(s);
            r = r1;
        }
        seg_tile_ptr[s + 1] = (int32_t)tr0.size();
    }
    for (int32_t** p : {&h->tile_r0, &h->tile_r1, &h->tile_seg, &h->seg_tile_ptr})
        if (*p) { (void)hipFree(*p); *p = nullptr; }
    if (h->tile_info) { (void)hipFree(h->tile_info); h->tile_info = nullptr; }
    if (h->partial) { (void)hipFree(h->partial); h->partial = nullptr; }
    h->T = (int)tr0.size();
    h->tile_r0 = dev_alloc<int32_t>(h->T);
    h->tile_r1 = dev_alloc<int32_t>(h->T);
    h->tile_seg = dev_alloc<int32_t>(h->T);
    h->seg_tile_ptr = dev_alloc<int32_t>(h->S + 1);
    h->partial = dev_alloc<double>((size_t)h->T * TF_RED_SLOTS);
    h2d(h, h->tile_r0, tr0.data(), h->T);
    h2d(h, h->tile_r1, tr1.data(), h->T);
    h2d(h, h->tile_seg, tseg.data(), h->T);
    h2d(h, h->seg_tile_ptr, seg_tile_ptr.data(), h->S + 1);
    std::vector<TileInfo> info(h->T);
    for (int t = 0; t < h->T; ++t)
        info[t] = TileInfo{tseg[t], row_ptr[tr0[t]], row_ptr[tr1[t]], tr0[t], tr1[t], seg_tile_ptr[tseg[t]],
                           seg_tile_ptr[tseg[t] + 1], 0};
    h->tile_info = dev_alloc<TileInfo>(h->T);
    h2d(h, h->tile_info, info.data(), h->T);
    h->h_seg_tile_ptr = seg_tile_ptr;
    h->wg = wg;
    h->tile_cap = tile_cap;
    drop_krylov_graphs(h);
}

// ---------------------------------------------------------------------------------------------
// BSR pattern, element-block -> BSR-block scatter lists and, when slave/master DoF pairs are
// bound, the fold lists that turn the assembled M into N^T M N + I_slaves.
//
// The element pattern couples node A to every node of every element containing A.  Binding a
// slave DoF s of node S to a master DoF m of node Mn adds the blocks the fold fills in: with
// T(X) = {X} u {nodes of masters of X's DoFs}, every element block (r, c) spreads to
// T(r) x T(c).  Fold targets are the entries in a row or column of a bound DoF: slave rows and
// columns become unit rows/columns (SC/NewtonRaphsonSolver.cpp:205-209), every other target
// entry (a, b) is the sum of M[c][d] over c in {a} u slaves(a), d in {b} u slaves(b) (the rows
// and columns the reference adds into the master's, SC/NewtonRaphsonSolver.cpp:188-206).
void build_pattern(tf_handle* h)
{
    twolevel_invalidate(h);  // the coarse space's block pairs and excluded DoFs follow the pattern
    const int N = h->N, E = h->E;
    const std::vector<int32_t>& conn = h->h_conn;
    const std::vector<int32_t>& node_ptr = h->h_node_ptr;
    const std::vector<int32_t>& node_src = h->h_node_src;
    const int nbound = (int)h->h_bound.size() / 2;
    // masters of each node's DoFs (node level) and DoF-level slave/master maps
    std::vector<int32_t> slave_of((size_t)3 * N, -1);
    std::vector<std::vector<int32_t>> slaves_of_master;
    std::vector<int32_t> master_index((size_t)3 * N, -1), master_list;
    std::vector<std::vector<int32_t>> node_masters(nbound ? N : 0);
    for (int k = 0; k < nbound; ++k) {
        const int sd = h->h_bound[2 * k], md = h->h_bound[2 * k + 1];
        slave_of[sd] = md;
        if (master_index[md] < 0) {
            master_index[md] = (int32_t)master_list.size();
            master_list.push_back(md);
            slaves_of_master.emplace_back();
        }
        slaves_of_master[master_index[md]].push_back(sd);
        node_masters[sd / 3].push_back(md / 3);
    }
    for (auto& v : node_masters) {
        std::sort(v.begin(), v.end());
        v.erase(std::unique(v.begin(), v.end()), v.end());
    }
    // ---- node pattern ----
    std::vector<std::vector<int32_t>> rows(N);
    for (int A = 0; A < N; ++A) {
        std::vector<int32_t>& r = rows[A];
        for (int k = node_ptr[A]; k < node_ptr[A + 1]; ++k) {
            const int e = node_src[k] / 6;
            for (int b = 0; b < 6; ++b) r.push_back(conn[6 * e + b]);
        }
        r.push_back(A);  // every node keeps its diagonal block (isolated and ablated nodes too)
    }
    for (int64_t key : h->pk.h_blocks) {  // blocks of implicit-packing pairs
        const int A = (int)(key / N), B = (int)(key % N);
        rows[A].push_back(B);
        rows[B].push_back(A);
    }
    if (nbound) {
        // spread every existing block (r, c) to T(r) x T(c)
        std::vector<std::vector<int32_t>> extra(N);
        for (int A = 0; A < N; ++A) {
            std::vector<int32_t> ta{A};
            ta.insert(ta.end(), node_masters[A].begin(), node_masters[A].end());
            for (int c : rows[A]) {
                std::vector<int32_t> tc{c};
                tc.insert(tc.end(), node_masters[c].begin(), node_masters[c].end());
                for (int a : ta)
                    for (int b : tc)
                        if (a != A || b != c) extra[a].push_back(b);
            }
        }
        for (int A = 0; A < N; ++A) rows[A].insert(rows[A].end(), extra[A].begin(), extra[A].end());
    }
    std::vector<int32_t> row_ptr(N + 1, 0), col;
    col.reserve((size_t)N * 20);
    for (int A = 0; A < N; ++A) {
        std::vector<int32_t>& r = rows[A];
        std::sort(r.begin(), r.end());
        r.erase(std::unique(r.begin(), r.end()), r.end());
        col.insert(col.end(), r.begin(), r.end());
        row_ptr[A + 1] = (int32_t)col.size();
        std::vector<int32_t>().swap(r);
    }
    if (col.size() > (size_t)INT32_MAX / 9) throw std::runtime_error("Jacobian pattern too large for int32 indexing");
    const int nb = (int)col.size();
    std::vector<int32_t> blk_row(nb), diag_blk(N, -1);
    for (int A = 0; A < N; ++A)
        for (int j = row_ptr[A]; j < row_ptr[A + 1]; ++j) {
            blk_row[j] = A;
            if (col[j] == A) diag_blk[A] = j;
        }
    auto find_blk = [&](int A, int B) -> int {
        const int32_t* c0 = col.data() + row_ptr[A];
        const int32_t* c1 = col.data() + row_ptr[A + 1];
        const int32_t* it = std::lower_bound(c0, c1, B);
        return (it != c1 && *it == B) ? (int)(it - col.data()) : -1;
    };
    // ---- element block -> BSR block, then contributors per block (ascending e*36+ab) ----
    std::vector<int32_t> eblk((size_t)36 * E);
    std::vector<int32_t> blk_ptr(nb + 1, 0);
    for (int e = 0; e < E; ++e)
        for (int a = 0; a < 6; ++a)
            for (int b = 0; b < 6; ++b) {
                const int j = find_blk(conn[6 * e + a], conn[6 * e + b]);
                eblk[(size_t)36 * e + 6 * a + b] = j;
                blk_ptr[j + 1]++;
            }
    for (int j = 0; j < nb; ++j) blk_ptr[j + 1] += blk_ptr[j];
    std::vector<int32_t> blk_src((size_t)36 * E);
    {
        std::vector<int32_t> fill(blk_ptr.begin(), blk_ptr.end() - 1);
        for (size_t k = 0; k < (size_t)36 * E; ++k) blk_src[fill[eblk[k]]++] = (int32_t)k;
    }
    // the pattern is structurally symmetric: block (A, B) has a partner (B, A)
    std::vector<int32_t> blk_tr(nb);
    for (int j = 0; j < nb; ++j) {
        blk_tr[j] = find_blk(col[j], blk_row[j]);
        if (blk_tr[j] < 0) throw std::runtime_error("Jacobian pattern is not structurally symmetric");
    }
    // ---- fold lists ----
    std::vector<int32_t> fold_ptr{0}, bind_nodes, m_ptr{0}, m_dof, s_dof, pairs_dev;
    std::vector<int64_t> fold_tgt, fold_src;
    std::vector<double> fold_const;
    if (nbound) {
        std::vector<uint8_t> is_tgt((size_t)9 * nb, 0);
        // column -> blocks map, for the column targets
        std::vector<int32_t> cptr(N + 1, 0), cblk(nb);
        for (int j = 0; j < nb; ++j) cptr[col[j] + 1]++;
        for (int n = 0; n < N; ++n) cptr[n + 1] += cptr[n];
        {
            std::vector<int32_t> fill(cptr.begin(), cptr.end() - 1);
            for (int j = 0; j < nb; ++j) cblk[fill[col[j]]++] = j;
        }
        std::vector<int32_t> bound_dofs;
        for (int k = 0; k < 2 * nbound; ++k) bound_dofs.push_back(h->h_bound[k]);
        std::sort(bound_dofs.begin(), bound_dofs.end());
        bound_dofs.erase(std::unique(bound_dofs.begin(), bound_dofs.end()), bound_dofs.end());
        std::vector<int64_t> targets;
        for (int x : bound_dofs) {
            const int X = x / 3, xi = x % 3;
            for (int j = row_ptr[X]; j < row_ptr[X + 1]; ++j)
                for (int jj = 0; jj < 3; ++jj) {
                    const int64_t t = (int64_t)(3 * xi + jj) * nb + j;
                    if (!is_tgt[t]) { is_tgt[t] = 1; targets.push_back(t); }
                }
            for (int k = cptr[X]; k < cptr[X + 1]; ++k)
                for (int ii = 0; ii < 3; ++ii) {
                    const int64_t t = (int64_t)(3 * ii + xi) * nb + cblk[k];
                    if (!is_tgt[t]) { is_tgt[t] = 1; targets.push_back(t); }
                }
            bind_nodes.push_back(X);
        }
        std::sort(targets.begin(), targets.end());
        std::sort(bind_nodes.begin(), bind_nodes.end());
        bind_nodes.erase(std::unique(bind_nodes.begin(), bind_nodes.end()), bind_nodes.end());
        auto group = [&](int dof) {
            std::vector<int32_t> g{dof};
            if (master_index[dof] >= 0) {
                const auto& sl = slaves_of_master[master_index[dof]];
                g.insert(g.end(), sl.begin(), sl.end());
            }
            return g;
        };
        for (int64_t t : targets) {
            const int q = (int)(t / nb), j = (int)(t % nb);
            const int a = 3 * blk_row[j] + q / 3, b = 3 * col[j] + q % 3;
            fold_tgt.push_back(t);
            if (slave_of[a] >= 0 || slave_of[b] >= 0) {
                fold_const.push_back(a == b ? 1.0 : 0.0);
            } else {
                fold_const.push_back(0.0);
                for (int c : group(a))
                    for (int d : group(b)) {
                        const int jb = find_blk(c / 3, d / 3);
                        if (jb >= 0) fold_src.push_back((int64_t)(3 * (c % 3) + d % 3) * nb + jb);
                    }
            }
            fold_ptr.push_back((int32_t)fold_src.size());
        }
        for (size_t k = 0; k < master_list.size(); ++k) {
            m_dof.push_back(master_list[k]);
            s_dof.insert(s_dof.end(), slaves_of_master[k].begin(), slaves_of_master[k].end());
            m_ptr.push_back((int32_t)s_dof.size());
        }
        pairs_dev = h->h_bound;
    }
    // ---- device arrays ----
    for (int32_t** p : {&h->row_ptr, &h->col, &h->blk_row, &h->diag_blk, &h->blk_ptr, &h->blk_src, &h->blk_tr, &h->fold_ptr,
                        &h->bind_nodes, &h->bind_m_ptr, &h->bind_m_dof, &h->bind_s_dof, &h->bind_pairs})
        if (*p) { (void)hipFree(*p); *p = nullptr; }
    for (int64_t** p : {&h->fold_tgt, &h->fold_src})
        if (*p) { (void)hipFree(*p); *p = nullptr; }
    for (double** p : {&h->val, &h->val_base, &h->fold_const, &h->val_t})
        if (*p) { (void)hipFree(*p); *p = nullptr; }
    if (h->val32) { (void)hipFree(h->val32); h->val32 = nullptr; }
    h->nb = nb;
    h->row_ptr = dev_alloc<int32_t>(N + 1);
    h->col = dev_alloc<int32_t>(nb);
    h->blk_row = dev_alloc<int32_t>(nb);
    h->diag_blk = dev_alloc<int32_t>(N);
    h->val = dev_alloc<double>((size_t)9 * nb);
    h->val32 = dev_alloc<float>((size_t)9 * nb);
    h->blk_ptr = dev_alloc<int32_t>(nb + 1);
    h->blk_src = dev_alloc<int32_t>((size_t)36 * E);
    h->blk_tr = dev_alloc<int32_t>(nb);
    h2d(h, h->blk_tr, blk_tr.data(), nb);
    h2d(h, h->row_ptr, row_ptr.data(), N + 1);
    h2d(h, h->col, col.data(), nb);
    h2d(h, h->blk_row, blk_row.data(), nb);
    h2d(h, h->diag_blk, diag_blk.data(), N);
    h2d(h, h->blk_ptr, blk_ptr.data(), nb + 1);
    h2d(h, h->blk_src, blk_src.data(), blk_src.size());
    h->n_bound = nbound;
    h->n_fold = (int32_t)fold_tgt.size();
    h->n_bind_nodes = (int32_t)bind_nodes.size();
    h->n_masters = (int32_t)m_dof.size();
    if (nbound) {
        h->val_base = dev_alloc<double>((size_t)9 * nb);
        h->fold_ptr = dev_alloc<int32_t>(fold_ptr.size());
        h->fold_tgt = dev_alloc<int64_t>(fold_tgt.size());
        h->fold_src = dev_alloc<int64_t>(std::max<size_t>(1, fold_src.size()));
        h->fold_const = dev_alloc<double>(fold_const.size());
        h->bind_nodes = dev_alloc<int32_t>(bind_nodes.size());
        h->bind_m_ptr = dev_alloc<int32_t>(m_ptr.size());
        h->bind_m_dof = dev_alloc<int32_t>(m_dof.size());
        h->bind_s_dof = dev_alloc<int32_t>(s_dof.size());
        h->bind_pairs = dev_alloc<int32_t>(pairs_dev.size());
        h2d(h, h->fold_ptr, fold_ptr.data(), fold_ptr.size());
        h2d(h, h->fold_tgt, fold_tgt.data(), fold_tgt.size());
        if (!fold_src.empty()) h2d(h, h->fold_src, fold_src.data(), fold_src.size());
        h2d(h, h->fold_const, fold_const.data(), fold_const.size());
        h2d(h, h->bind_nodes, bind_nodes.data(), bind_nodes.size());
        h2d(h, h->bind_m_ptr, m_ptr.data(), m_ptr.size());
        h2d(h, h->bind_m_dof, m_dof.data(), m_dof.size());
        h2d(h, h->bind_s_dof, s_dof.data(), s_dof.size());
        h2d(h, h->bind_pairs, pairs_dev.data(), pairs_dev.size());
    }
    TF_HIP(hipStreamSynchronize(h->stream));
    h->h_row_ptr = row_ptr;
    h->h_col = col;
}

// ---------------------------------------------------------------------------------------------
extern "C" {

void tf_default_options(tf_solver_opts* o)
{
    std::memset(o, 0, sizeof(*o));
    o->krylov = -1;  // automatic: single-kernel CG when the Jacobian is symmetric (no internal viscosity), else BiCGSTAB
    o->krylov_max_iter = 20000;
    o->krylov_rtol = 1e-12;
    o->krylov_atol = 0.0;
    o->iters_per_sync = 32;
    o->nr_tol = 1e-8;
    o->nr_max_iter = 20;
    o->tangent_precision = TF_PREC_FP64;
    o->matrix_precision = TF_PREC_FP64;
    o->growth_rotation = 1;
    o->use_graph = 1;
    o->krylov_forcing = 1;
    o->xcd_tile_map = 1;
    o->krylov_warm_start = 1;
    o->implicit_packing = 1;
    o->preconditioner = TF_PRECOND_AUTO;
    o->coarse_refresh = 100;
    o->column_blocks = 1;
    o->mixed_fp64_from = 1;
    o->nr_graph = 1;
    o->tl_persistent = 1;
    o->tl_barrier_polls = 0;
}

const char* tf_last_error(const tf_handle* h) { return h ? h->err.c_str() : g_create_error.c_str(); }

int tf_create(const tf_mesh_desc* d, tf_handle** out)
{
    if (!d || !out) return fail(nullptr, TF_INVALID, "null argument");
    *out = nullptr;
    if (d->n_nodes <= 0 || d->n_elements <= 0 || !d->node_pos || !d->conn || !d->lambda || !d->mu)
        return fail(nullptr, TF_INVALID, "mesh needs n_nodes > 0, n_elements > 0, node_pos, conn, lambda, mu");
    const int N = d->n_nodes, E = d->n_elements;
    const int S = d->n_segments > 0 ? d->n_segments : 1;
    if (S > 1 && (!d->node_seg_ptr || !d->elem_seg_ptr))
        return fail(nullptr, TF_INVALID, "n_segments > 1 needs node_seg_ptr and elem_seg_ptr");
    tf_handle* h = new tf_handle();
    try {
        tf_default_options(&h->opts);
        h->N = N; h->E = E; h->S = S;
        TF_HIP(hipGetDevice(&h->device));
        scratch_init(h->device);  // the allocation-free element dump's scratch on this device
        TF_HIP(hipDeviceGetAttribute(&h->max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, h->device));
        TF_HIP(hipDeviceGetAttribute(&h->n_cu, hipDeviceAttributeMultiprocessorCount, h->device));
        if (hipDeviceGetAttribute(&h->n_xcd, hipDeviceAttributeNumberOfXccs, h->device) != hipSuccess || h->n_xcd < 1)
            h->n_xcd = 1;
        TF_HIP(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
        h->h_conn.assign(d->conn, d->conn + (size_t)6 * E);
        h->h_node_seg_ptr.resize(S + 1);
        h->h_elem_seg_ptr.resize(S + 1);
        if (S == 1) {
            h->h_node_seg_ptr = {0, N};
            h->h_elem_seg_ptr = {0, E};
        } else {
            std::copy(d->node_seg_ptr, d->node_seg_ptr + S + 1, h->h_node_seg_ptr.begin());
            std::copy(d->elem_seg_ptr, d->elem_seg_ptr + S + 1, h->h_elem_seg_ptr.begin());
            if (h->h_node_seg_ptr[0] != 0 || h->h_node_seg_ptr[S] != N || h->h_elem_seg_ptr[0] != 0 ||
                h->h_elem_seg_ptr[S] != E)
                throw std::runtime_error("segment pointers must cover [0, n) exactly");
            for (int s = 0; s < S; ++s)
                if (h->h_node_seg_ptr[s + 1] <= h->h_node_seg_ptr[s] || h->h_elem_seg_ptr[s + 1] <= h->h_elem_seg_ptr[s])
                    throw std::runtime_error("every segment needs at least one node and one element");
        }
        // ---- validate connectivity and segment membership ----
        std::vector<int32_t> eseg(E);
        for (int s = 0; s < S; ++s)
            for (int e = h->h_elem_seg_ptr[s]; e < h->h_elem_seg_ptr[s + 1]; ++e) {
                eseg[e] = s;
                for (int a = 0; a < 6; ++a) {
                    const int id = h->h_conn[6 * e + a];
                    if (id < h->h_node_seg_ptr[s] || id >= h->h_node_seg_ptr[s + 1])
                        throw std::runtime_error("element " + std::to_string(e) + " references node " +
                                                 std::to_string(id) + " outside its tissue");
                }
                for (int a = 0; a < 6; ++a)
                    for (int b = a + 1; b < 6; ++b)
                        if (h->h_conn[6 * e + a] == h->h_conn[6 * e + b])
                            throw std::runtime_error("element " + std::to_string(e) + " repeats a node");
            }
        // ---- node -> element corners ----
        std::vector<int32_t> node_ptr(N + 1, 0), node_src((size_t)6 * E);
        for (int k = 0; k < 6 * E; ++k) node_ptr[h->h_conn[k] + 1]++;
        for (int n = 0; n < N; ++n) node_ptr[n + 1] += node_ptr[n];
        {
            std::vector<int32_t> fill(node_ptr.begin(), node_ptr.end() - 1);
            for (int k = 0; k < 6 * E; ++k) node_src[fill[h->h_conn[k]]++] = k;  // k = e*6 + a, ascending
        }
        h->h_node_ptr = node_ptr;
        h->h_node_src = node_src;
        // ---- device allocations ----
        h->conn = dev_alloc<int32_t>((size_t)6 * E);
        h->xref = dev_alloc<double>((size_t)18 * E);
        h->fmw = dev_alloc<double>((size_t)18 * E);
        h->fg = dev_alloc<double>((size_t)9 * E);
        h->invfg = dev_alloc<double>((size_t)9 * E);
        h->lam = dev_alloc<double>(E);
        h->mu = dev_alloc<double>(E);
        h->eta = dev_alloc<double>(E);
        h->rate = dev_alloc<double>((size_t)3 * E);
        h->exposed = dev_alloc<uint8_t>(E);
        h->elem_seg = dev_alloc<int32_t>(E);
        h->uk = dev_alloc<double>((size_t)3 * N);
        h->un = dev_alloc<double>((size_t)3 * N);
        h->g = dev_alloc<double>((size_t)3 * N);
        h->du = dev_alloc<double>((size_t)3 * N);
        h->eta_ext = dev_alloc<double>((size_t)3 * N);
        h->fixed = dev_alloc<uint8_t>((size_t)3 * N);
        h->sa = dev_alloc<double>(N);
        h->face_area = dev_alloc<double>((size_t)2 * E);
        h->rot_tz = dev_alloc<double>(E);
        h->kelem = dev_alloc<double>((size_t)324 * E);
        h->felem = dev_alloc<double>((size_t)18 * E);
        h->status = dev_alloc<int32_t>(4);
        h->dinv = dev_alloc<double>((size_t)9 * N);
        h->node_ptr = dev_alloc<int32_t>(N + 1);
        h->node_src = dev_alloc<int32_t>((size_t)6 * E);
        h->seg_node_ptr = dev_alloc<int32_t>(S + 1);
        h->seg_counter = dev_alloc<int32_t>(S);
        h->seg = dev_alloc<SegScalars>(S);
        h->seg_norm = dev_alloc<double>(S);
        h->seg_rtol2 = dev_alloc<double>(S);
        TF_HIP(hipHostMalloc((void**)&h->h_seg, sizeof(SegScalars) * S, hipHostMallocDefault));
        TF_HIP(hipHostMalloc((void**)&h->h_seg_norm, sizeof(double) * S, hipHostMallocDefault));
        TF_HIP(hipHostMalloc((void**)&h->h_status, 12 * sizeof(int32_t), hipHostMallocDefault));
        TF_HIP(hipHostMalloc((void**)&h->hp_rtol2, sizeof(double) * S, hipHostMallocDefault));
        for (double** v : {&h->kr, &h->kr0, &h->kp, &h->kph, &h->kv, &h->ks, &h->ksh, &h->kt, &h->kx, &h->km, &h->kq})
            *v = dev_alloc<double>((size_t)3 * N);

        // ---- uploads ----
        std::vector<int32_t> conn_soa = to_soa(d->conn, E, 6);
        h2d(h, h->conn, conn_soa.data(), conn_soa.size());
        std::vector<double> xr((size_t)18 * E);
        if (d->ref_pos) std::copy(d->ref_pos, d->ref_pos + xr.size(), xr.begin());
        else
            for (int e = 0; e < E; ++e)
                for (int a = 0; a < 6; ++a)
                    for (int i = 0; i < 3; ++i) xr[(size_t)18 * e + 3 * a + i] = d->node_pos[3 * h->h_conn[6 * e + a] + i];
        std::vector<double> xr_soa = to_soa(xr.data(), E, 18);
        h2d(h, h->xref, xr_soa.data(), xr_soa.size());
        launch_mean_f_weights(h);
        std::vector<double> eye((size_t)9 * E, 0.0);
        for (int e = 0; e < E; ++e) eye[e] = eye[4 * (size_t)E + e] = eye[8 * (size_t)E + e] = 1.0;
        h2d(h, h->fg, eye.data(), eye.size());
        h2d(h, h->invfg, eye.data(), eye.size());
        h2d(h, h->lam, d->lambda, E);
        h2d(h, h->mu, d->mu, E);
        std::vector<double> zeroE(E, 0.0), zero3E((size_t)3 * E, 0.0), zero3N((size_t)3 * N, 0.0);
        h2d(h, h->eta, d->eta_int ? d->eta_int : zeroE.data(), E);
        if (d->eta_int)
            for (int e = 0; e < E; ++e) h->any_viscous |= (d->eta_int[e] != 0.0);
        if (d->growth_rate) {
            std::vector<double> r = to_soa(d->growth_rate, E, 3);
            h2d(h, h->rate, r.data(), r.size());
        } else h2d(h, h->rate, zero3E.data(), zero3E.size());
        std::vector<uint8_t> ex(E, 0), fx((size_t)3 * N, 0);
        h2d(h, h->exposed, d->exposed ? d->exposed : ex.data(), E);
        h2d(h, h->elem_seg, eseg.data(), E);
        h2d(h, h->uk, d->node_pos, (size_t)3 * N);
        h2d(h, h->un, d->node_pos, (size_t)3 * N);
        h2d(h, h->eta_ext, d->eta_ext ? d->eta_ext : zero3N.data(), (size_t)3 * N);
        h2d(h, h->fixed, d->fixed ? d->fixed : fx.data(), (size_t)3 * N);
        h2d(h, h->g, zero3N.data(), zero3N.size());
        h2d(h, h->du, zero3N.data(), zero3N.size());
        std::vector<double> zN(N, 0.0);
        h2d(h, h->sa, zN.data(), N);
        h2d(h, h->rot_tz, zeroE.data(), E);
        h2d(h, h->node_ptr, node_ptr.data(), N + 1);
        h2d(h, h->node_src, node_src.data(), node_src.size());
        build_pattern(h);
        build_tiles(h, 256, 512);
        packing_init(h, d->node_placement, d->exposed, xr.data());
        h2d(h, h->seg_node_ptr, h->h_node_seg_ptr.data(), S + 1);
        std::vector<int32_t> zS(S, 0);
        h2d(h, h->seg_counter, zS.data(), S);
        std::vector<SegScalars> s0(S);
        std::memset(s0.data(), 0, sizeof(SegScalars) * S);
        h2d(h, h->seg, s0.data(), S);
        std::vector<int32_t> z4(4, 0);
        h2d(h, h->status, z4.data(), 4);
        TF_HIP(hipStreamSynchronize(h->stream));
        h->seg_nr_iters.assign(S, 0);
        h->seg_converged.assign(S, 0);
        h->seg_failed.assign(S, 0);
        h->seg_du.assign(S, 0.0);
        h->fc_C.assign(S, -1.0);
        h->fc_du0.assign(S, -1.0);
        h->fc_du_prev.assign(S, -1.0);
        h->h_seg_rtol2.assign(S, 0.0);
        // set-up that depends on the mesh alone, done here rather than inside the first timestep:
        // the packing detection tables and, when the default solver uses it, the coarse space
        packing_prepare(h);
        twolevel_setup(h);
    } catch (const tf_hip_error& ex) {
        free_all(h);
        delete h;
        return fail(nullptr, ex.code == hipErrorOutOfMemory ? TF_OOM : TF_HIP_ERROR, ex.what);
    } catch (const std::exception& ex) {
        free_all(h);
        delete h;
        return fail(nullptr, TF_INVALID, ex.what());
    }
    *out = h;
    return TF_OK;
}

void tf_destroy(tf_handle* h)
{
    if (!h) return;
    (void)hipStreamSynchronize(h->stream);
    free_all(h);
    delete h;
}

int tf_set_options(tf_handle* h, const tf_solver_opts* o)
{
    if (!h || !o) return TF_INVALID;
    if (o->krylov < -1 || o->krylov == 2 || o->krylov > TF_KRYLOV_PIPECG || o->krylov_max_iter <= 0 || o->nr_max_iter < 0 ||
        !(o->krylov_rtol >= 0) || !(o->nr_tol >= 0) ||
        (o->matrix_precision != TF_PREC_FP64 && o->matrix_precision != TF_PREC_FP32) ||
        o->tangent_precision < TF_PREC_FP64 || o->tangent_precision > TF_PREC_BF16 ||
        (o->implicit_packing != 0 && o->implicit_packing != 1) ||
        o->preconditioner < TF_PRECOND_AUTO || o->preconditioner > TF_PRECOND_TWO_LEVEL ||
        o->coarse_aggregates < 0 || o->coarse_refresh < 0 || (o->column_blocks != 0 && o->column_blocks != 1) ||
        (o->precond_fp64 != 0 && o->precond_fp64 != 1) || o->nr_graph < 0 || o->nr_graph > 2 ||
        (o->tl_persistent < 0 || o->tl_persistent > 2) || o->tl_barrier_polls < 0)
        return fail(h, TF_INVALID, "invalid solver options");
    const bool coarse_changed = o->coarse_aggregates != h->opts.coarse_aggregates ||
                                o->column_blocks != h->opts.column_blocks || o->precond_fp64 != h->opts.precond_fp64;
    h->opts = *o;
    drop_krylov_graphs(h);
    TF_GUARD_BEGIN
    if (coarse_changed) twolevel_invalidate(h);  // the aggregates follow the option
    twolevel_setup(h);
    return TF_OK;
    TF_GUARD_END(h)
}

int tf_get_options(const tf_handle* h, tf_solver_opts* o)
{
    if (!h || !o) return TF_INVALID;
    *o = h->opts;
    return TF_OK;
}

int tf_get_sizes(const tf_handle* h, int32_t* n_nodes, int32_t* n_elements, int32_t* n_blocks, int32_t* n_segments)
{
    if (!h) return TF_INVALID;
    if (n_nodes) *n_nodes = h->N;
    if (n_elements) *n_elements = h->E;
    if (n_blocks) *n_blocks = h->nb;
    if (n_segments) *n_segments = h->S;
    return TF_OK;
}

int tf_set_positions(tf_handle* h, const double* x)
{
    if (!h || !x) return TF_INVALID;
    TF_GUARD_BEGIN
    h2d(h, h->uk, x, (size_t)3 * h->N);
    TF_HIP(hipStreamSynchronize(h->stream));
    warm_start_clear(h);  // earlier updates belong to another state
    return TF_OK;
    TF_GUARD_END(h)
}

int tf_get_positions(tf_handle* h, double* x)
{
    if (!h || !x) return TF_INVALID;
    TF_GUARD_BEGIN
    // A kernel writes the positions into pinned, device-mapped host memory: a runtime copy of
    // this size (0.6 MB at 40k prisms) went to a DMA engine that, idle since the last such copy,
    // took 5 ms instead of 0.05 about every other time after a run of timesteps.
    const size_t n = (size_t)3 * h->N;
    if (!h->hp_pos) {
        TF_HIP(hipHostMalloc((void**)&h->hp_pos, sizeof(double) * std::max<size_t>(n, 1), hipHostMallocMapped));
        TF_HIP(hipHostGetDevicePointer((void**)&h->dp_pos, h->hp_pos, 0));
    }
    launch_copy(h, h->dp_pos, h->uk, (int64_t)n);
    TF_HIP(hipStreamSynchronize(h->stream));
    std::copy(h->hp_pos, h->hp_pos + n, x);
    return TF_OK;
    TF_GUARD_END(h)
}

int tf_set_growth_rates(tf_handle* h, const double* rate)
{
    if (!h || !rate) return TF_INVALID;
    TF_GUARD_BEGIN
    std::vector<double> r = to_soa(rate, h->E, 3);
    h2d(h, h->rate, r.data(), r.size());
    TF_HIP(hipStreamSynchronize(h->stream));
    return TF_OK;
    TF_GUARD_END(h)
}

int tf_set_growth_increments(tf_handle* h, const double* dFg)
{
    if (!h) return TF_INVALID;
    TF_GUARD_BEGIN
    growth_functions_free(h);  // host-given increments replace device growth functions
    if (!dFg) {
        if (h->ginc) { (void)hipFree(h->ginc); h->ginc = nullptr; }
        return TF_OK;
    }
    for (size_t k = 0; k < (size_t)9 * h->E; ++k)
        if (!std::isfinite(dFg[k])) return fail(h, TF_INVALID, "growth increments must be finite");
    if (!h->ginc) h->ginc = dev_alloc<double>((size_t)9 * h->E);
    std::vector<double> soa = to_soa(dFg, h->E, 9);
    h2d(h, h->ginc, soa.data(), soa.size());
    return TF_OK;
    TF_GUARD_END(h)
}

int tf_set_growth_functions(tf_handle* h, int32_t n, const tf_growth_function* f, const double* initial_z,
                            const int32_t* rel_src, int32_t pinned, int32_t interpolation)
{
    if (!h || n < 0 || (n > 0 && (!f || !initial_z || !rel_src))) return TF_INVALID;
    TF_GUARD_BEGIN
    if (n == 0) {
        growth_functions_free(h);
        if (h->ginc) { (void)hipFree(h->ginc); h->ginc = nullptr; }
        return TF_OK;
    }
    if (n > 64) return fail(h, TF_INVALID, "at most 64 growth functions");
    if ((pinned != 0 && pinned != 1) || (interpolation != 0 && interpolation != 1))
        return fail(h, TF_INVALID, "pinned and interpolation must be 0 or 1");
    for (int k = 0; k < n; ++k) {
        const tf_growth_function& g = f[k];
        const std::string tag = "growth function " + std::to_string(k) + ": ";
        if (g.kind < TF_GROWTH_UNIFORM || g.kind > TF_GROWTH_GRID) return fail(h, TF_INVALID, tag + "unknown kind");
        if (!(g.t0 <= g.t1)) return fail(h, TF_INVALID, tag + "t0 > t1");
        if (!std::isfinite(g.angle)) return fail(h, TF_INVALID, tag + "angle not finite");
        for (int i = 0; i < 3; ++i)
            if (!std::isfinite(g.rate[i])) return fail(h, TF_INVALID, tag + "rate not finite");
        if (g.kind == TF_GROWTH_RING && !(std::isfinite(g.centre[0]) && std::isfinite(g.centre[1]) &&
                                          std::isfinite(g.inner) && std::isfinite(g.outer) && g.inner < g.outer))
            return fail(h, TF_INVALID, tag + "ring centre and radii must be finite with inner < outer");
        if (g.kind == TF_GROWTH_GRID) {
            if (g.nx < 2 || g.ny < 2 || !g.grid_rates || !g.grid_angles)
                return fail(h, TF_INVALID, tag + "a grid needs nx, ny >= 2 and its rates and angles");
            for (size_t q = 0; q < (size_t)g.nx * g.ny; ++q)
                if (!std::isfinite(g.grid_angles[q]) || !std::isfinite(g.grid_rates[3 * q]) ||
                    !std::isfinite(g.grid_rates[3 * q + 1]) || !std::isfinite(g.grid_rates[3 * q + 2]))
                    return fail(h, TF_INVALID, tag + "grid values must be finite");
        }
    }
    for (int e = 0; e < h->E; ++e) {
        if (rel_src[e] < 0 || rel_src[e] >= h->E) return fail(h, TF_INVALID, "rel_src out of range");
        if (!std::isfinite(initial_z[e])) return fail(h, TF_INVALID, "initial_z must be finite");
    }
    growth_functions_set(h, n, f, initial_z, rel_src, pinned, interpolation);
    TF_HIP(hipStreamSynchronize(h->stream));
    return TF_OK;
    TF_GUARD_END(h)
}

int tf_pin_growth_positions(tf_handle* h)
{
    if (!h) return TF_INVALID;
    TF_GUARD_BEGIN
    growth_functions_pin(h);
    TF_HIP(hipStreamSynchronize(h->stream));
    return TF_OK;
    TF_GUARD_END(h)
}

int tf_set_time(tf_handle* h, double t)
{
    if (!h || !std::isfinite(t)) return TF_INVALID;
    h->time = t;
    return TF_OK;
}

int tf_get_time(const tf_handle* h, double* t)
{
    if (!h || !t) return TF_INVALID;
    *t = h->time;
    return TF_OK;
}

int tf_get_growth_increments(tf_handle* h, double* dFg)
{
    if (!h || !dFg) return TF_INVALID;
    if (!h->ginc) return fail(h, TF_INVALID, "rate-based growth: no increments");
    TF_GUARD_BEGIN
    std::vector<double> soa((size_t)9 * h->E);
    d2h(h, soa.data(), h->ginc, soa.size());
    TF_HIP(hipStreamSynchronize(h->stream));
    for (int e = 0; e < h->E; ++e)
        for (int q = 0; q < 9; ++q) dFg[9 * (size_t)e + q] = soa[(size_t)q * h->E + e];
    return TF_OK;
    TF_GUARD_END(h)
}

int tf_set_fg(tf_handle* h, const double* Fg)
{
    if (!h || !Fg) return TF_INVALID;
    TF_GUARD_BEGIN
    std::vector<double> inv((size_t)9 * h->E);
    for (int e = 0; e < h->E; ++e) {
        const double* M = Fg + 9 * (size_t)e;
        double* I = inv.data() + 9 * (size_t)e;
        double c00 = M[4] * M[8] - M[5] * M[7], c01 = M[5] * M[6] - M[3] * M[8], c02 = M[3] * M[7] - M[4] * M[6];
        double det = M[0] * c00 + M[1] * c01 + M[2] * c02;
        if (!(std::fabs(det) > 0)) throw std::runtime_error("singular Fg");
        double id = 1.0 / det;
        I[0] = c00 * id; I[1] = (M[2] * M[7] - M[1] * M[8]) * id; I[2] = (M[1] * M[5] - M[2] * M[4]) * id;
        I[3] = c01 * id; I[4] = (M[0] * M[8] - M[2] * M[6]) * id; I[5] = (M[2] * M[3] - M[0] * M[5]) * id;
        I[6] = c02 * id; I[7] = (M[1] * M[6] - M[0] * M[7]) * id; I[8] = (M[0] * M[4] - M[1] * M[3]) * id;
    }
    std::vector<double> a = to_soa(Fg, h->E, 9), b = to_soa(inv.data(), h->E, 9);
    h2d(h, h->fg, a.data(), a.size());
    h2d(h, h->invfg, b.data(), b.size());
    TF_HIP(hipStreamSynchronize(h->stream));
    return TF_OK;
    TF_GUARD_END(h)
}

int tf_get_fg(tf_handle* h, double* Fg, double* InvFg)
{
    if (!h) return TF_INVALID;
    TF_GUARD_BEGIN
    std::vector<double> a((size_t)9 * h->E), b((size_t)9 * h->E);
    d2h(h, a.data(), h->fg, a.size());
    d2h(h, b.data(), h->invfg, b.size());
    TF_HIP(hipStreamSynchronize(h->stream));
    if (Fg) { auto r = from_soa(a.data(), h->E, 9); std::copy(r.begin(), r.end(), Fg); }
    if (InvFg) { auto r = from_soa(b.data(), h->E, 9); std::copy(r.begin(), r.end(), InvFg); }
    return TF_OK;
    TF_GUARD_END(h)
}

int tf_set_material(tf_handle* h, const double* lambda, const double* mu, const double* eta_int)
{
    if (!h) return TF_INVALID;
    TF_GUARD_BEGIN
    if (lambda) h2d(h, h->lam, lambda, h->E);
    if (mu) h2d(h, h->mu, mu, h->E);
    if (eta_int) {
        h2d(h, h->eta, eta_int, h->E);
        h->any_viscous = 0;
        for (int e = 0; e < h->E; ++e) h->any_viscous |= (eta_int[e] != 0.0);
    }
    TF_HIP(hipStreamSynchronize(h->stream));
    return TF_OK;
    TF_GUARD_END(h)
}

int tf_grow(tf_handle* h, double dt, int32_t* n_rotated)
{
    if (!h || !(dt > 0)) return TF_INVALID;
    TF_GUARD_BEGIN
    hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(64), 0, h->stream, h->status + 1, 1);
    launch_grow(h, dt);
    int32_t st[4];
    d2h(h, st, h->status, 4);
    TF_HIP(hipStreamSynchronize(h->stream));
    if (n_rotated) *n_rotated = st[1];
    return TF_OK;
    TF_GUARD_END(h)
}

int tf_update_viscosity_surfaces(tf_handle* h)
{
    if (!h) return TF_INVALID;
    TF_GUARD_BEGIN
    launch_surfaces(h);
    TF_HIP(hipStreamSynchronize(h->stream));
    return TF_OK;
    TF_GUARD_END(h)
}

int tf_get_viscosity_surfaces(tf_handle* h, double* Sa)
{
    if (!h || !Sa) return TF_INVALID;
    TF_GUARD_BEGIN
    d2h(h, Sa, h->sa, h->N);
    TF_HIP(hipStreamSynchronize(h->stream));
    return TF_OK;
    TF_GUARD_END(h)
}

int tf_begin_step(tf_handle* h)
{
    if (!h) return TF_INVALID;
    TF_GUARD_BEGIN
    launch_copy(h, h->un, h->uk, 3 * (int64_t)h->N);
    hipLaunchKernelGGL(k_seg_begin, dim3((h->S + 255) / 256), dim3(256), 0, h->stream, h->seg, h->S);
    TF_HIP(hipGetLastError());
    TF_HIP(hipStreamSynchronize(h->stream));
    std::fill(h->seg_nr_iters.begin(), h->seg_nr_iters.end(), 0);
    std::fill(h->seg_converged.begin(), h->seg_converged.end(), 0);
    std::fill(h->seg_failed.begin(), h->seg_failed.end(), 0);
    return TF_OK;
    TF_GUARD_END(h)
}

static int assemble_impl(tf_handle* h, double dt)
{
    hipLaunchKernelGGL(k_zero_assembly_status, dim3(1), dim3(64), 0, h->stream, h->status);
    launch_assemble_elements(h, dt);
    launch_assemble_gather(h, dt);
    launch_residual(h, dt);
    return TF_OK;
}

int tf_assemble(tf_handle* h, double dt)
{
    if (!h || !(dt > 0)) return TF_INVALID;
    TF_GUARD_BEGIN
    assemble_impl(h, dt);
    int32_t st[4];
    d2h(h, st, h->status, 4);
    TF_HIP(hipStreamSynchronize(h->stream));
    timer_collect(h->t_elem);
    timer_collect(h->t_gather);
    if (st[3] > 0) return fail(h, TF_NAN, std::to_string(st[3]) + " Gauss points with a det Fe that is NaN or Inf");
    if (st[0] > 0) return fail(h, TF_NEG_JACOBIAN, std::to_string(st[0]) + " Gauss points with det Fe <= 0");
    return TF_OK;
    TF_GUARD_END(h)
}

int tf_get_residual(tf_handle* h, double* g)
{
    if (!h || !g) return TF_INVALID;
    TF_GUARD_BEGIN
    d2h(h, g, h->g, (size_t)3 * h->N);
    TF_HIP(hipStreamSynchronize(h->stream));
    return TF_OK;
    TF_GUARD_END(h)
}

int tf_get_jacobian_bsr(tf_handle* h, int32_t* row_ptr, int32_t* col, double* val)
{
    if (!h) return TF_INVALID;
    TF_GUARD_BEGIN
    if (row_ptr) std::copy(h->h_row_ptr.begin(), h->h_row_ptr.end(), row_ptr);
    if (col) std::copy(h->h_col.begin(), h->h_col.end(), col);
    if (val) {
        std::vector<double> v((size_t)9 * h->nb);
        d2h(h, v.data(), h->val, v.size());
        TF_HIP(hipStreamSynchronize(h->stream));
        for (int j = 0; j < h->nb; ++j)
            for (int q = 0; q < 9; ++q) val[(size_t)9 * j + q] = v[(size_t)q * h->nb + j];
    }
    return TF_OK;
    TF_GUARD_END(h)
}

int tf_element_dump(tf_handle* h, int32_t e, double dt, double* fe, double* fv, double* Ke, double* Kv)
{
    if (!h || e < 0 || e >= h->E || !(dt > 0)) return TF_INVALID;
    TF_GUARD_BEGIN
    double* d_out = dev_alloc<double>(TF_DUMP_DOUBLES);
    launch_element_dump(h, e, 1, dt, d_out);
    std::vector<double> o(36 + 648);
    d2h(h, o.data(), d_out, o.size());
    TF_HIP(hipStreamSynchronize(h->stream));
    (void)hipFree(d_out);
    if (fe) std::copy(o.begin(), o.begin() + 18, fe);
    if (fv) std::copy(o.begin() + 18, o.begin() + 36, fv);
    if (Ke) std::copy(o.begin() + 36, o.begin() + 360, Ke);
    if (Kv) std::copy(o.begin() + 360, o.end(), Kv);
    return TF_OK;
    TF_GUARD_END(h)
}

int tf_element_dump_range_noalloc(tf_handle* h, int32_t e0, int32_t n, double dt, double* fe, double* fv, double* Ke,
                                  double* Kv)
{
    if (!h) return TF_INVALID;
    if (e0 < 0 || n < 1 || e0 > h->E - n || !(dt > 0))
        return fail(h, TF_INVALID, "element range outside [0, n_elements) or dt <= 0");
    if (n > TF_DUMP_SCRATCH_PRISMS)
        return fail(h, TF_INVALID, "at most " + std::to_string(TF_DUMP_SCRATCH_PRISMS) +
                                       " prisms per call (the static scratch holds that many element dumps)");
    TF_GUARD_BEGIN
    double* d_out = scratch_addr(h->device);  // resolved by tf_create on this device
    if (!d_out) return fail(h, TF_INVALID, "the element-dump scratch of this device is not set up");
    launch_element_dump(h, e0, n, dt, d_out);
    const size_t m = (size_t)n;
    if (fe) d2h(h, fe, d_out, 18 * m);
    if (fv) d2h(h, fv, d_out + 18 * m, 18 * m);
    if (Ke) d2h(h, Ke, d_out + 36 * m, 324 * m);
    if (Kv) d2h(h, Kv, d_out + 360 * m, 324 * m);
    TF_HIP(hipStreamSynchronize(h->stream));
    return TF_OK;
    TF_GUARD_END(h)
}

static int solve_impl(tf_handle* h)
{
    tf_solver_opts saved = h->opts;
    if (h->opts.krylov < 0) h->opts.krylov = h->any_viscous ? TF_KRYLOV_BICGSTAB : TF_KRYLOV_CG1;
    int st = krylov_solve(h);
    h->opts.krylov = saved.krylov;
    launch_bind_du(h);  // slaves follow their masters (SC/NewtonRaphsonSolver.cpp:397-399)
    return st;
}

int tf_solve(tf_handle* h, double* dU, int32_t* krylov_iters, double* rel_residual)
{
    if (!h) return TF_INVALID;
    TF_GUARD_BEGIN
    int st = solve_impl(h);
    if (dU) d2h(h, dU, h->du, (size_t)3 * h->N);
    TF_HIP(hipStreamSynchronize(h->stream));
    if (krylov_iters) *krylov_iters = h->last_krylov_iters;
    if (rel_residual) {
        double worst = 0.0;
        for (int s = 0; s < h->S; ++s)
            if (h->h_seg[s].bnorm2 > 0) worst = std::max(worst, std::sqrt(h->h_seg[s].rnorm2 / h->h_seg[s].bnorm2));
        *rel_residual = worst;
    }
    if (st == TF_NAN) return fail(h, st, "NaN or Inf in the right-hand side");
    if (st != TF_OK) return fail(h, st, "Krylov solver did not reach its tolerance");
    return TF_OK;
    TF_GUARD_END(h)
}

// The host's side of a Newton update: every iterating tissue's ||dU|| (norm), its iteration count
// and its convergence (||dU|| <= nr_tol)
static int update_uk_host(tf_handle* h, int iter_index, const double* norm)
{
    int st = TF_OK;
    for (int s = 0; s < h->S; ++s) {
        if (h->seg_converged[s] || h->seg_failed[s]) continue;
        const double dn = norm[s];
        h->seg_du[s] = dn;
        h->seg_nr_iters[s] = iter_index + 1;
        if (!std::isfinite(dn)) st = TF_NAN;
        else if (dn <= h->opts.nr_tol) h->seg_converged[s] = 1;
    }
    return st;
}

static int update_uk_impl(tf_handle* h, int iter_index)
{
    launch_seg_norm(h, h->du);
    launch_update_uk(h);
    hipLaunchKernelGGL(k_seg_nr_check, dim3((h->S + 255) / 256), dim3(256), 0, h->stream, h->seg, h->seg_norm,
                       h->opts.nr_tol, h->S);
    TF_HIP(hipGetLastError());
    d2h(h, h->h_seg_norm, h->seg_norm, h->S);
    TF_HIP(hipStreamSynchronize(h->stream));
    h->host_waits += 1;
    return update_uk_host(h, iter_index, h->h_seg_norm);
}

int tf_update_uk(tf_handle* h, double* du_norm, int32_t* converged)
{
    if (!h) return TF_INVALID;
    TF_GUARD_BEGIN
    int it = 0;
    for (int s = 0; s < h->S; ++s) it = std::max(it, h->seg_nr_iters[s]);
    int st = update_uk_impl(h, it);
    double worst = 0.0;
    int all = 1;
    for (int s = 0; s < h->S; ++s) { worst = std::max(worst, h->seg_du[s]); all &= h->seg_converged[s]; }
    if (du_norm) *du_norm = worst;
    if (converged) *converged = all;
    if (st != TF_OK) return fail(h, st, "NaN in the Newton update");
    return TF_OK;
    TF_GUARD_END(h)
}

static double ms_since(std::chrono::steady_clock::time_point t0)
{
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

// Inexact-Newton forcing (per tissue).  C is the quadratic-convergence constant of the tissue's
// Newton iteration (||dU_{k+1}|| ~ C ||dU_k||^2, the smallest ratio seen; the first step, solved
// to krylov_rtol, supplies it), p the predicted size of this iteration's update and q = C p^2
// that of the next one:
//  - p <= nr_tol (the last iteration): the solve only decides convergence -> relative 1e-2;
//  - q >= 100 nr_tol (far from convergence): the error left in dU may be 5% of the next update.
//    A block-Jacobi CG iterate's error is up to ~100x its relative residual times ||dU|| (soft,
//    weakly pinned modes keep the slowest error), so rtol = 0.05 q / (100 p);
//  - q < nr_tol (the iteration before the last), symmetric solvers: the error may be 5% of the
//    margin nr_tol - q, so that it cannot move the last ||dU|| across nr_tol:
//    rtol = 0.05 (nr_tol - q) / (100 p), never below krylov_rtol (20 instead of 100 with the
//    two-level preconditioner);
//  - nr_tol <= q < 100 nr_tol, symmetric solvers: the error may be 5% of the margin that keeps the
//    next update above nr_tol (q - nr_tol) and the one after it, C q^2, on its side of nr_tol
//    (|sqrt(nr_tol / C) - q|), over the same amplification;
//  - otherwise krylov_rtol, so that the next ||dU||, and with it the NR iteration count, is that
//    of an exact solve (BiCGSTAB has no such error bound and keeps exact solves there).
static void forcing_tolerances(tf_handle* h, int iter)
{
    // error amplification margin: ||error|| / (||dU|| rtol) of a CG iterate reached ~150 with block
    // Jacobi (smooth error, weakly pinned meshes) but stayed below 10 with the two-level
    // preconditioner, whose coarse space removes the smooth components
    const bool two_level = twolevel_active(h) && !h->tl.failed;
    const double floor_rtol = h->opts.krylov_rtol, tol = h->opts.nr_tol, amplification = two_level ? 20.0 : 100.0;
    const int kind = h->opts.krylov;
    const bool symmetric = !h->any_viscous && (kind < 0 || kind == TF_KRYLOV_CG || kind == TF_KRYLOV_CG1 ||
                                               kind == TF_KRYLOV_PIPECG);
    for (int s = 0; s < h->S; ++s) {
        double eta = floor_rtol;
        const double C = h->fc_C[s];
        double p = -1.0;
        if (iter == 0) p = h->fc_du0[s] > 0 ? 1.5 * h->fc_du0[s] : -1.0;
        else if (h->fc_du_prev[s] > 0) p = C * h->fc_du_prev[s] * h->fc_du_prev[s];
        if (C > 0 && p > 0) {
            const double q = C * p * p;
            if (p <= tol) eta = 1e-2;
            else if (q >= 100.0 * tol) eta = std::min(1e-2, std::max(floor_rtol, 0.05 * q / (amplification * p)));
            else if (q < tol && symmetric)
                eta = std::min(1e-2, std::max(floor_rtol, 0.05 * (tol - q) / (amplification * p)));
            else if (symmetric) {  // tol <= q < 100 tol
                // an error e moves the next update to about q +- e: it stays above tol while
                // e < q - tol, and the update after it (C (q +- e)^2) stays on the side of tol
                // that C q^2 is on while e < |sqrt(tol / C) - q|
                const double margin = std::min(q - tol, std::fabs(std::sqrt(tol / C) - q));
                eta = std::min(1e-2, std::max(floor_rtol, 0.05 * margin / (amplification * p)));
            }
        }
        h->h_seg_rtol2[s] = eta * eta;
    }
    // pinned staging, no host wait (the previous upload ended at the last solve's waits)
    std::copy(h->h_seg_rtol2.begin(), h->h_seg_rtol2.end(), h->hp_rtol2);
    TF_HIP(hipMemcpyAsync(h->seg_rtol2, h->hp_rtol2, sizeof(double) * h->S, hipMemcpyHostToDevice, h->stream));
}

static void forcing_observe(tf_handle* h, int iter)
{
    for (int s = 0; s < h->S; ++s) {
        if (h->seg_nr_iters[s] != iter + 1) continue;  // tissue not iterated this time
        const double d = h->seg_du[s], prev = h->fc_du_prev[s];
        if (iter == 0) h->fc_du0[s] = d;
        else if (prev > 0 && d > 0) {
            const double c = d / (prev * prev);
            h->fc_C[s] = h->fc_C[s] > 0 ? std::min(c, 2.0 * h->fc_C[s]) : c;
        }
        h->fc_du_prev[s] = d;
    }
}

// The device side of a Newton iteration's end, enqueued right behind the first round of its solve:
// the solution into du, slaves to their masters, ||dU|| per tissue, the gate, the gated update and
// convergence test, and the report (nr_epilogue_read copies it to the host).
static void nr_epilogue_launch(tf_handle* h)
{
    launch_solution_to_du(h);
    launch_bind_du(h);  // slaves follow their masters (SC/NewtonRaphsonSolver.cpp:397-399)
    launch_seg_norm(h, h->du);
    hipLaunchKernelGGL(k_nr_gate, dim3(1), dim3(256), 0, h->stream, h->seg, h->S, h->status, h->nr_rep);
    launch_update_uk(h, h->nr_rep);
    hipLaunchKernelGGL(k_nr_finish, dim3((h->S + 255) / 256), dim3(256), 0, h->stream, h->seg, h->seg_norm,
                       h->opts.nr_tol, h->S, h->nr_rep);
    TF_HIP(hipGetLastError());
}

// The one host wait of a Newton iteration: the report, then the host mirrors of the status and of
// the tissues' Krylov scalars
static void nr_epilogue_read(tf_handle* h)
{
    TF_HIP(hipMemcpyAsync(h->hp_nr_rep, h->nr_rep, nr_report_bytes(h->S), hipMemcpyDeviceToHost, h->stream));
    TF_HIP(hipStreamSynchronize(h->stream));
    h->host_waits += 1;
    std::memcpy(h->h_status, h->hp_nr_rep->status, 4 * sizeof(int32_t));
    std::memcpy(h->h_seg, nr_report_seg(h->hp_nr_rep), sizeof(SegScalars) * h->S);
}

static void nr_report_alloc(tf_handle* h)
{
    if (h->nr_rep) return;
    TF_HIP(hipMalloc((void**)&h->nr_rep, nr_report_bytes(h->S)));
    TF_HIP(hipHostMalloc((void**)&h->hp_nr_rep, nr_report_bytes(h->S), hipHostMallocDefault));
}

// Whether Newton iteration it (krylov_call set) can run from a graph: no per-launch timing, no
// packing pairs (their lists change size between timesteps), a first burst predicted from the
// previous timestep, and no coarse build or rebuild due (those read back a flag).
static bool nr_graph_ready(tf_handle* h, int it)
{
    if (h->t_elem.enabled || h->t_gather.enabled || h->t_spmv.enabled || h->pk.n > 0) return false;
    if (it >= (int)h->krylov_pred.size() || h->krylov_pred[it] < 0) return false;
    return !krylov_pre_solve_rebuilds(h);
}

// The device work of Newton iteration it as the one-read path enqueues it, the solve's first round
// as kernels (for the capture); krylov_solve_plan has sized that round.
static void nr_iteration_launch(tf_handle* h, double dt, int it, bool warm, bool tl_prep)
{
    assemble_impl(h, dt);
    if (warm) warm_start_launch(h, it);
    if (tl_prep) krylov_pre_launch(h);
    krylov_solve_launch(h, true);
    nr_epilogue_launch(h);
}

// Replays the graph of Newton iteration it for the shape of its launches, capturing and
// instantiating it on first use (at most 32 kept, the least recently used dropped)
static void nr_graph_run(tf_handle* h, double dt, int it, bool warm, bool tl_prep)
{
    const tf_handle::KrylovRun& R = h->krun;
    // a persistent solve is the same launch whatever its prediction: no chunk counts in its key
    const tf_handle::NrGraphKey key{it, warm ? h->warm_count[it] : 0, R.kind, R.fp32 ? 1 : 0, tangent_precision_now(h),
                                    R.persist ? -1 : R.n_big, R.persist ? -1 : R.n_small, h->seg_rtol_on ? 1 : 0, dt,
                                    h->ablated};
    h->nr_graph_clock += 1;
    hipGraphExec_t exec = nullptr;
    for (tf_handle::NrGraph& g : h->nr_graphs)
        if (g.key == key) { exec = g.exec; g.used = h->nr_graph_clock; break; }
    if (!exec) {
        if (h->nr_graphs.size() >= 32) {
            auto lru = std::min_element(h->nr_graphs.begin(), h->nr_graphs.end(),
                                        [](const tf_handle::NrGraph& a, const tf_handle::NrGraph& b) { return a.used < b.used; });
            (void)hipGraphExecDestroy(lru->exec);
            h->nr_graphs.erase(lru);
        }
        hipGraph_t gr;
        TF_HIP(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
        try {
            nr_iteration_launch(h, dt, it, warm, tl_prep);
        } catch (...) {
            (void)hipStreamEndCapture(h->stream, &gr);
            throw;
        }
        TF_HIP(hipStreamEndCapture(h->stream, &gr));
        const hipError_t e = hipGraphInstantiate(&exec, gr, nullptr, nullptr, 0);
        (void)hipGraphDestroy(gr);
        TF_HIP(e);
        h->nr_graphs.push_back(tf_handle::NrGraph{key, exec, h->nr_graph_clock});
        h->nr_graph_builds += 1;
    }
    TF_HIP(hipGraphLaunch(exec, h->stream));
    if (R.persist) h->persist_solves += 1;
}

// The device side of a timestep's start: un <- uk, the tissues' Newton state reset, and (assemble)
// the assembly of Newton iteration 0 enqueued ahead of the host's set-up of the loop.  Enqueued
// again by a step that finds its input changed (packing pairs), it only repeats the same writes.
static void nr_step_begin(tf_handle* h, double dt, bool assemble)
{
    launch_copy(h, h->un, h->uk, 3 * (int64_t)h->N);
    hipLaunchKernelGGL(k_seg_begin, dim3((h->S + 255) / 256), dim3(256), 0, h->stream, h->seg, h->S);
    TF_HIP(hipGetLastError());
    if (assemble) {
        h->nr_it = 0;  // the precision switch reads the iteration index
        assemble_impl(h, dt);
        h->nr_it = -1;
    }
}

// begun: nr_step_begin(h, dt, true) has been enqueued for this step
static int nr_step_impl(tf_handle* h, double dt, tf_step_stats* stats, bool begun = false)
{
    auto t0 = std::chrono::steady_clock::now();
    h->host_waits = 0;
    h->nr_graph_builds = 0;
    h->persist_solves = 0;
    if (!begun) nr_step_begin(h, dt, false);
    // The assembly of the iteration about to run is already enqueued: the host enqueues it as soon
    // as the previous iteration's report says that one more follows, ahead of its own bookkeeping
    // (warm-start history, forcing, tolerances), which then overlaps the assembly on the device.
    // Not under nr_graph = 2, whose graphs hold the assembly.
    bool asm_ahead = begun;
    double next_asm = 0.0;
    std::fill(h->seg_nr_iters.begin(), h->seg_nr_iters.end(), 0);
    std::fill(h->seg_converged.begin(), h->seg_converged.end(), 0);
    std::fill(h->seg_failed.begin(), h->seg_failed.end(), 0);
    int status = TF_OK;
    int krylov_total = 0;
    bool stalled = false;
    int stalls_accepted = 0;
    h->fallbacks = 0;
    h->tl_used = false;
    double ms_asm = 0, ms_sol = 0;
    const int saved_kind = h->opts.krylov;
    // the per-solve state the Krylov set-up reads; reset on every exit from a solve
    auto solve_state_off = [h, saved_kind]() {
        h->seg_rtol_on = false;
        h->warm_guess = nullptr;
        h->krylov_call = -1;
        h->tl_pre_valid = false;
        h->krun.pending = false;
        h->opts.krylov = saved_kind;
    };
    int it = 0;
    struct NrIterGuard {  // the Newton iteration index the precision switch reads; -1 outside the loop
        tf_handle* h;
        ~NrIterGuard() { h->nr_it = -1; }
    } nr_guard{h};
    while (true) {
        h->nr_it = it;
        // One read per Newton iteration: a solve with a prediction runs its first round, the update
        // and the convergence test behind one gate on the device (graph-replayed rounds, no
        // per-launch timing), from one captured graph under nr_graph = 2.  Otherwise the det Fe
        // status comes back ahead of the solve.
        const bool one_read = h->opts.nr_graph != 0 && h->opts.use_graph != 0 && !h->t_spmv.enabled;
        // a first burst that falls short costs a host wait and the epilogue enqueued behind it; a
        // small chunk of slack costs a few no-op launches once the solve has converged
        h->burst_slack = one_read ? TF_ONE_READ_SLACK : 0;
        auto ts = std::chrono::steady_clock::now();
        double it_asm = 0.0;  // host time enqueueing the assembly (eager launches)
        const bool warm = h->opts.krylov_warm_start != 0 && it < TF_WARM_ITERS;
        int kst;
        bool applied = false;  // the device applied this iteration's update (the gate was open)
        try {
            h->seg_rtol_on = h->opts.krylov_forcing != 0;
            if (h->seg_rtol_on) forcing_tolerances(h, it);
            h->krylov_call = it;
            bool burst = false;
            if (one_read) nr_report_alloc(h);
            if (!asm_ahead && one_read && h->opts.nr_graph == 2 && nr_graph_ready(h, it)) {
                // the host's plan of the iteration (warm start, two-level prepare, the solve's first
                // burst; no launches), then its launches from one graph
                const bool w = warm && warm_start_begin(h, it, false);
                h->warm_guess = w ? h->warm_x0 : nullptr;
                krylov_pre_solve(h, false);
                const bool tl_prep = h->tl_pre_valid && h->tl_pre_ok;
                if (h->opts.krylov < 0) h->opts.krylov = h->any_viscous ? TF_KRYLOV_BICGSTAB : TF_KRYLOV_CG1;
                h->no_persist = h->opts.tl_persistent == 2;  // a cooperative launch cannot be captured
                try {
                    burst = krylov_solve_plan(h);
                } catch (...) {
                    h->no_persist = false;
                    throw;
                }
                h->no_persist = false;
                if (!burst) throw std::logic_error("a Newton-iteration graph needs a predicted first burst");
                nr_graph_run(h, dt, it, w, tl_prep);
                nr_epilogue_read(h);
            } else {
                if (asm_ahead) {
                    it_asm = next_asm;
                } else {
                    auto ta = std::chrono::steady_clock::now();
                    assemble_impl(h, dt);
                    it_asm = ms_since(ta);
                }
                asm_ahead = false;
                h->warm_guess = warm && warm_start_begin(h, it) ? h->warm_x0 : nullptr;
                krylov_pre_solve(h);
                if (h->opts.krylov < 0) h->opts.krylov = h->any_viscous ? TF_KRYLOV_BICGSTAB : TF_KRYLOV_CG1;
                if (one_read) {
                    burst = krylov_solve_begin(h);
                    if (burst) {
                        nr_epilogue_launch(h);
                        nr_epilogue_read(h);
                    } else {
                        TF_HIP(hipMemcpyAsync(h->h_status, h->status, 4 * sizeof(int32_t), hipMemcpyDeviceToHost,
                                              h->stream));
                        TF_HIP(hipStreamSynchronize(h->stream));
                        h->host_waits += 1;
                    }
                } else {
                    TF_HIP(hipMemcpyAsync(h->h_status, h->status, 4 * sizeof(int32_t), hipMemcpyDeviceToHost, h->stream));
                    TF_HIP(hipStreamSynchronize(h->stream));  // the status
                    h->host_waits += 1;
                }
            }
            if (h->h_status[3] > 0) {  // positions or growth state not finite
                solve_state_off();
                status = TF_NAN;
                h->err = std::to_string(h->h_status[3]) + " Gauss points with a det Fe that is NaN or Inf";
                break;
            }
            if (h->h_status[0] > 0) {
                solve_state_off();
                status = TF_NEG_JACOBIAN;
                h->err = std::to_string(h->h_status[0]) + " Gauss points with det Fe <= 0";
                break;
            }
            if (one_read) {
                applied = burst && h->hp_nr_rep->ok;
                kst = krylov_solve_end(h, burst, applied);
            } else {
                kst = krylov_solve(h);
            }
            if (!applied) launch_bind_du(h);  // slaves follow their masters (SC/NewtonRaphsonSolver.cpp:397-399)
        } catch (...) {
            solve_state_off();
            throw;
        }
        solve_state_off();
        h->burst_slack = 0;
        if (kst == TF_NAN) {  // a residual that is not finite: nothing to update
            status = TF_NAN;
            h->err = "NaN or Inf in the residual";
            break;
        }
        if (kst == TF_KRYLOV_STALL) {
            // a tissue whose stalled solve is unusable leaves the loop with the positions of its
            // last update; a usable one goes on as an inexact Newton step, as do the others (an
            // open gate means that every stall was usable)
            for (int s = 0; s < h->S; ++s) {
                const SegScalars& q = h->h_seg[s];
                if (!q.active || !q.stalled || h->seg_converged[s]) continue;
                if (solve_unusable(q.rnorm2, q.bnorm2)) {
                    h->seg_failed[s] = 1;
                    stalled = true;
                } else {
                    stalls_accepted += 1;
                }
            }
            if (!applied) {
                hipLaunchKernelGGL(k_seg_drop_stalled, dim3((h->S + 255) / 256), dim3(256), 0, h->stream, h->seg, h->S);
                TF_HIP(hipGetLastError());
            }
        }
        krylov_total += h->last_krylov_iters;
        if (stats && it < 4) stats->krylov_by_newton[it] = h->last_krylov_iters;
        int ust = applied ? update_uk_host(h, it, nr_report_norm(h->hp_nr_rep, h->S)) : update_uk_impl(h, it);
        if (ust == TF_OK && it + 1 <= h->opts.nr_max_iter && h->opts.nr_graph != 2) {
            bool all = true;
            for (int s = 0; s < h->S; ++s) all &= (bool)(h->seg_converged[s] || h->seg_failed[s]);
            if (!all) {  // another iteration follows: its assembly goes first
                auto ta = std::chrono::steady_clock::now();
                h->nr_it = it + 1;
                assemble_impl(h, dt);
                h->nr_it = it;
                next_asm = ms_since(ta);
                asm_ahead = true;
            }
        }
        if (warm && kst == TF_OK) warm_start_push(h, it);  // a guess for the next timestep's iteration it
        if (h->opts.krylov_forcing) forcing_observe(h, it);
        ms_asm += it_asm;
        ms_sol += ms_since(ts) - it_asm;
        if (ust == TF_NAN) { status = TF_NAN; h->err = "NaN in the Newton update"; break; }
        ++it;
        bool all = true;
        for (int s = 0; s < h->S; ++s) all &= (bool)(h->seg_converged[s] || h->seg_failed[s]);
        if (all) break;
        if (it > h->opts.nr_max_iter) {
            status = TF_NOT_CONVERGED;
            h->err = "Newton-Raphson did not converge";
            hipLaunchKernelGGL(k_seg_deactivate, dim3((h->S + 255) / 256), dim3(256), 0, h->stream, h->seg, h->S);
            break;
        }
    }
    if (stalled && (status == TF_OK || status == TF_NOT_CONVERGED)) {
        int n = 0;
        for (int s = 0; s < h->S; ++s) n += h->seg_failed[s];
        status = TF_KRYLOV_STALL;
        h->err = "the linear solve of " + std::to_string(n) + " tissue(s) did not reach its tolerance (CG and "
                 "BiCGSTAB) and left a residual above 1e-2 of the right-hand side; their positions are those "
                 "of their last applied Newton update";
    }
    timer_collect(h->t_elem);
    timer_collect(h->t_gather);
    if (stats) {
        int maxit = 0, allc = 1;
        double worst = 0.0;
        for (int s = 0; s < h->S; ++s) {
            maxit = std::max(maxit, h->seg_nr_iters[s]);
            allc &= h->seg_converged[s];
            worst = std::max(worst, h->seg_du[s]);
        }
        stats->nr_iterations = maxit;
        stats->krylov_iterations = krylov_total;
        stats->krylov_fallbacks = h->fallbacks;
        stats->krylov_stalls = stalls_accepted;
        stats->packing_pairs = h->pk.n;
        stats->two_level = h->tl_used ? 1 : 0;
        stats->converged = allc;
        stats->last_du_norm = worst;
        stats->ms_assemble = ms_asm;
        stats->ms_solve = ms_sol;
        stats->ms_total = ms_since(t0);
        stats->host_waits = h->host_waits;
        stats->nr_graph_builds = h->nr_graph_builds;
        stats->krylov_persistent = h->persist_solves;
    }
    return status;
}

int tf_nr_step(tf_handle* h, double dt, tf_step_stats* stats)
{
    if (!h || !(dt > 0)) return TF_INVALID;
    TF_GUARD_BEGIN
    if (stats) std::memset(stats, 0, sizeof(*stats));
    return nr_step_impl(h, dt, stats);
    TF_GUARD_END(h)
}

int tf_timestep(tf_handle* h, double dt, tf_step_stats* stats)
{
    if (!h || !(dt > 0)) return TF_INVALID;
    TF_GUARD_BEGIN
    if (stats) std::memset(stats, 0, sizeof(*stats));
    auto t0 = std::chrono::steady_clock::now();
    hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(64), 0, h->stream, h->status + 1, 1);
    launch_grow(h, dt);
    h->time += dt;  // the growth update of this timestep is applied
    launch_surfaces(h);
    // the growth update's counters, read once the step has ended
    TF_HIP(hipMemcpyAsync(h->h_status + 4, h->status, 4 * sizeof(int32_t), hipMemcpyDeviceToHost, h->stream));
    bool begun = false;
    if (h->opts.implicit_packing) {
        packing_detect_launch(h);
        // With no pairs in force, the step's start and the assembly of Newton iteration 0 are
        // enqueued behind the detection before the host waits for its pair count; pairs that
        // formed send the step through its start again, with them.
        if (h->pk.n == 0) {
            nr_step_begin(h, dt, true);
            begun = true;
        }
        if (packing_detect_finish(h) > 0) begun = false;
    } else {
        packing_clear(h);
    }
    double ms_grow = ms_since(t0);
    int st = nr_step_impl(h, dt, stats, begun);
    if (stats) {
        TF_HIP(hipStreamSynchronize(h->stream));  // the counters' copy (the step has read its reports)
        stats->rotated_elements = h->h_status[5];
        stats->ms_grow = ms_grow;
        stats->ms_total = ms_since(t0);
    }
    return st;
    TF_GUARD_END(h)
}

int tf_set_bound_dofs(tf_handle* h, int32_t n_pairs, const int32_t* pairs)
{
    if (!h || n_pairs < 0 || (n_pairs > 0 && !pairs)) return TF_INVALID;
    TF_GUARD_BEGIN
    const int nd = 3 * h->N;
    std::vector<int32_t> role((size_t)nd, 0);  // 1 = slave, 2 = master
    auto seg_of = [&](int node) {
        return (int)(std::upper_bound(h->h_node_seg_ptr.begin(), h->h_node_seg_ptr.end(), node) -
                     h->h_node_seg_ptr.begin()) - 1;
    };
    for (int k = 0; k < n_pairs; ++k) {
        const int sd = pairs[2 * k], md = pairs[2 * k + 1];
        const std::string tag = "bound pair " + std::to_string(k) + " (" + std::to_string(sd) + ", " +
                                std::to_string(md) + "): ";
        if (sd < 0 || sd >= nd || md < 0 || md >= nd) return fail(h, TF_INVALID, tag + "DoF out of range");
        if (sd == md) return fail(h, TF_INVALID, tag + "a DoF cannot be bound to itself");
        if (role[sd] == 1) return fail(h, TF_INVALID, tag + "the slave DoF is already bound");
        if (role[sd] == 2 || role[md] == 1)
            return fail(h, TF_INVALID, tag + "chained binding (a slave that is also a master) is not supported");
        if (seg_of(sd / 3) != seg_of(md / 3)) return fail(h, TF_INVALID, tag + "slave and master in different tissues");
        role[sd] = 1;
        role[md] = 2;
    }
    h->h_bound.assign(pairs, pairs + 2 * (size_t)n_pairs);
    build_pattern(h);
    build_tiles(h, h->wg, h->tile_cap);
    return TF_OK;
    TF_GUARD_END(h)
}

int tf_get_node_masses(tf_handle* h, double* mass)
{
    if (!h || !mass) return TF_INVALID;
    TF_GUARD_BEGIN
    launch_packing_mass(h);
    d2h(h, mass, h->pk.mass, h->N);
    TF_HIP(hipStreamSynchronize(h->stream));
    return TF_OK;
    TF_GUARD_END(h)
}

int tf_detect_packing(tf_handle* h, int32_t* n_pairs)
{
    if (!h) return TF_INVALID;
    TF_GUARD_BEGIN
    const int n = packing_detect(h);
    if (n_pairs) *n_pairs = n;
    return TF_OK;
    TF_GUARD_END(h)
}

int tf_set_packing_pairs(tf_handle* h, int32_t n_pairs, const int32_t* pairs, const double* weights,
                         const double* threshold)
{
    if (!h || n_pairs < 0 || (n_pairs > 0 && (!pairs || !weights || !threshold))) return TF_INVALID;
    TF_GUARD_BEGIN
    auto seg_of = [&](int node) {
        return (int)(std::upper_bound(h->h_node_seg_ptr.begin(), h->h_node_seg_ptr.end(), node) -
                     h->h_node_seg_ptr.begin()) - 1;
    };
    std::vector<int32_t> a(n_pairs), b(n_pairs);
    std::vector<int64_t> keys(n_pairs);
    std::vector<uint8_t> seg_used(h->S, 0);
    for (int k = 0; k < n_pairs; ++k) {
        a[k] = pairs[2 * k];
        b[k] = pairs[2 * k + 1];
        const std::string tag = "packing pair " + std::to_string(k) + ": ";
        if (a[k] < 0 || a[k] >= h->N || b[k] < 0 || b[k] >= h->N) return fail(h, TF_INVALID, tag + "node out of range");
        if (a[k] == b[k]) return fail(h, TF_INVALID, tag + "a node cannot pack against itself");
        const int s = seg_of(a[k]);
        if (s != seg_of(b[k])) return fail(h, TF_INVALID, tag + "nodes in different tissues");
        if (!(threshold[s] > 0)) return fail(h, TF_INVALID, tag + "its tissue needs a threshold > 0");
        for (int c = 0; c < 3; ++c)
            if (!std::isfinite(weights[3 * k + c])) return fail(h, TF_INVALID, tag + "weights must be finite");
        seg_used[s] = 1;
        keys[k] = (int64_t)std::min(a[k], b[k]) * h->N + std::max(a[k], b[k]);
    }
    std::vector<int64_t> sorted = keys;
    std::sort(sorted.begin(), sorted.end());
    if (std::adjacent_find(sorted.begin(), sorted.end()) != sorted.end())
        return fail(h, TF_INVALID, "a node pair appears twice in the packing list");
    packing_prepare(h);
    launch_packing_mass(h);
    if (n_pairs > 0) {
        packing_reserve(h, n_pairs);
        h2d(h, h->pk.a, a.data(), n_pairs);
        h2d(h, h->pk.b, b.data(), n_pairs);
        h2d(h, h->pk.w, weights, (size_t)3 * n_pairs);
        std::vector<double> st((size_t)h->S * PK_NSTAT);
        d2h(h, st.data(), h->pk.seg_stats, st.size());
        TF_HIP(hipStreamSynchronize(h->stream));
        for (int s = 0; s < h->S; ++s)
            if (seg_used[s]) st[(size_t)s * PK_NSTAT + PK_ST_T] = threshold[s];
        h2d(h, h->pk.seg_stats, st.data(), st.size());
    }
    packing_install(h, a, b);
    TF_HIP(hipStreamSynchronize(h->stream));
    return TF_OK;
    TF_GUARD_END(h)
}

int tf_get_packing_pairs(tf_handle* h, int32_t* n_pairs, int32_t* pairs, double* weights, double* threshold,
                         double* detect_grid)
{
    if (!h || !n_pairs) return TF_INVALID;
    TF_GUARD_BEGIN
    const int P = h->pk.n;
    *n_pairs = P;
    if (pairs)
        for (int k = 0; k < P; ++k) {
            pairs[2 * k] = h->pk.h_a[k];
            pairs[2 * k + 1] = h->pk.h_b[k];
        }
    if (weights && P) d2h(h, weights, h->pk.w, (size_t)3 * P);
    if (threshold || detect_grid) {
        std::vector<double> st((size_t)h->S * PK_NSTAT);
        d2h(h, st.data(), h->pk.seg_stats, st.size());
        TF_HIP(hipStreamSynchronize(h->stream));
        for (int s = 0; s < h->S; ++s) {
            if (threshold) threshold[s] = st[(size_t)s * PK_NSTAT + PK_ST_T];
            if (detect_grid)
                for (int q = 0; q < 50; ++q) detect_grid[50 * s + q] = st[(size_t)s * PK_NSTAT + PK_ST_GRID + q];
        }
    }
    TF_HIP(hipStreamSynchronize(h->stream));
    return TF_OK;
    TF_GUARD_END(h)
}

int tf_set_ablated_nodes(tf_handle* h, int32_t n, const int32_t* nodes)
{
    if (!h || n < 0 || (n > 0 && !nodes)) return TF_INVALID;
    TF_GUARD_BEGIN
    std::vector<uint8_t> flag(h->N, 0);
    for (int k = 0; k < n; ++k) {
        if (nodes[k] < 0 || nodes[k] >= h->N) return fail(h, TF_INVALID, "ablated node id out of range");
        flag[nodes[k]] = 1;
    }
    if (!h->ablated) h->ablated = dev_alloc<uint8_t>(h->N);
    h2d(h, h->ablated, flag.data(), h->N);
    TF_HIP(hipStreamSynchronize(h->stream));
    if (n == 0) { (void)hipFree(h->ablated); h->ablated = nullptr; }
    return TF_OK;
    TF_GUARD_END(h)
}

// ---------------------------------------------------------------------------------------------
// discrete adjoint and general linear algebra on the assembled Jacobian
static void ensure_adjoint(tf_handle* h)
{
    const size_t n3 = 3 * (size_t)h->N, E = h->E;
    for (double** p : {&h->g_save, &h->adj_xbar, &h->adj_rhs, &h->adj_pass, &h->adj_y, &h->adj_un_bar})
        if (!*p) *p = dev_alloc<double>(n3);
    for (double** p : {&h->adj_invfg_bar, &h->adj_fg_bar, &h->adj_fg_prev})
        if (!*p) *p = dev_alloc<double>(9 * E);
    if (!h->adj_rate_bar) h->adj_rate_bar = dev_alloc<double>(3 * E);
    for (double** p : {&h->adj_lam_bar, &h->adj_mu_bar, &h->adj_lam_acc, &h->adj_mu_acc, &h->adj_tz})
        if (!*p) *p = dev_alloc<double>(E);
}

static void drop_graph(tf_handle* h)
{
    drop_krylov_graphs(h);
}

// M x = b or M^T x = b for device vectors, with the current Jacobian; FP64 matrix values.
static int solve_system_impl(tf_handle* h, int transpose, const double* d_b, double* d_x)
{
    ensure_adjoint(h);
    const int64_t n3 = 3 * (int64_t)h->N;
    launch_copy(h, h->g_save, h->g, n3);
    launch_copy(h, h->g, d_b, n3);
    hipLaunchKernelGGL(k_seg_begin, dim3((h->S + 255) / 256), dim3(256), 0, h->stream, h->seg, h->S);
    const tf_solver_opts saved = h->opts;
    h->opts.matrix_precision = TF_PREC_FP64;
    if (h->opts.krylov < 0) h->opts.krylov = h->any_viscous ? TF_KRYLOV_BICGSTAB : TF_KRYLOV_CG1;
    if (transpose) {
        if (!h->val_t) h->val_t = dev_alloc<double>((size_t)9 * h->nb);
        launch_transpose_vals(h, h->val_t);
        std::swap(h->val, h->val_t);
        launch_dinv_all(h);
        drop_graph(h);
    }
    int st = TF_OK;
    try {
        st = krylov_solve(h);
    } catch (...) {
        if (transpose) { std::swap(h->val, h->val_t); drop_graph(h); }
        h->opts = saved;
        throw;
    }
    if (transpose) {
        std::swap(h->val, h->val_t);
        launch_dinv_all(h);
        drop_graph(h);
    }
    h->opts = saved;
    launch_copy(h, d_x, h->du, n3);
    launch_copy(h, h->g, h->g_save, n3);
    return st;
}

int tf_set_growth_state(tf_handle* h, const double* Fg, const double* InvFg)
{
    if (!h || !Fg || !InvFg) return TF_INVALID;
    TF_GUARD_BEGIN
    std::vector<double> a = to_soa(Fg, h->E, 9), b = to_soa(InvFg, h->E, 9);
    h2d(h, h->fg, a.data(), a.size());
    h2d(h, h->invfg, b.data(), b.size());
    TF_HIP(hipStreamSynchronize(h->stream));
    return TF_OK;
    TF_GUARD_END(h)
}

int tf_get_growth_rotation(tf_handle* h, double* theta_z)
{
    if (!h || !theta_z) return TF_INVALID;
    TF_GUARD_BEGIN
    d2h(h, theta_z, h->rot_tz, h->E);
    TF_HIP(hipStreamSynchronize(h->stream));
    return TF_OK;
    TF_GUARD_END(h)
}

int tf_apply_jacobian(tf_handle* h, int32_t transpose, const double* x, double* y)
{
    if (!h || !x || !y) return TF_INVALID;
    TF_GUARD_BEGIN
    ensure_adjoint(h);
    const size_t n3 = 3 * (size_t)h->N;
    h2d(h, h->adj_xbar, x, n3);
    if (transpose) launch_spmv_t(h, h->adj_xbar, h->adj_y);
    else {
        const int mp = h->opts.matrix_precision;
        h->opts.matrix_precision = TF_PREC_FP64;
        launch_spmv(h, h->adj_xbar, h->adj_y);
        h->opts.matrix_precision = mp;
    }
    d2h(h, y, h->adj_y, n3);
    TF_HIP(hipStreamSynchronize(h->stream));
    return TF_OK;
    TF_GUARD_END(h)
}

int tf_solve_system(tf_handle* h, int32_t transpose, const double* b, double* x, int32_t* krylov_iters,
                    double* rel_residual)
{
    if (!h || !b || !x) return TF_INVALID;
    TF_GUARD_BEGIN
    ensure_adjoint(h);
    const size_t n3 = 3 * (size_t)h->N;
    h2d(h, h->adj_rhs, b, n3);
    const int st = solve_system_impl(h, transpose, h->adj_rhs, h->adj_y);
    d2h(h, x, h->adj_y, n3);
    TF_HIP(hipStreamSynchronize(h->stream));
    if (krylov_iters) *krylov_iters = h->last_krylov_iters;
    if (rel_residual) {
        double worst = 0.0;
        for (int s = 0; s < h->S; ++s)
            if (h->h_seg[s].bnorm2 > 0) worst = std::max(worst, std::sqrt(h->h_seg[s].rnorm2 / h->h_seg[s].bnorm2));
        *rel_residual = worst;
    }
    if (st != TF_OK) return fail(h, st, "Krylov solver did not reach its tolerance");
    return TF_OK;
    TF_GUARD_END(h)
}

static int adjoint_supported(tf_handle* h)
{
    if (h->any_viscous) return fail(h, TF_INVALID, "the adjoint supports eta_int = 0 only");
    if (h->n_bound) return fail(h, TF_INVALID, "the adjoint does not support bound DoFs");
    if (h->ablated) return fail(h, TF_INVALID, "the adjoint does not support ablated nodes");
    if (h->pk.n) return fail(h, TF_INVALID, "the adjoint does not support implicit-packing pairs");
    if (h->ginc) return fail(h, TF_INVALID, "the adjoint differentiates rate-based growth, not given increments");
    return TF_OK;
}

int tf_adjoint_elements(tf_handle* h, const double* y, double* invfg_bar, double* lambda_bar, double* mu_bar)
{
    if (!h || !y) return TF_INVALID;
    TF_GUARD_BEGIN
    ensure_adjoint(h);
    h2d(h, h->adj_y, y, 3 * (size_t)h->N);
    launch_adj_elements(h, h->adj_y, h->adj_invfg_bar, h->adj_lam_bar, h->adj_mu_bar);
    if (invfg_bar) d2h(h, invfg_bar, h->adj_invfg_bar, 9 * (size_t)h->E);
    if (lambda_bar) d2h(h, lambda_bar, h->adj_lam_bar, h->E);
    if (mu_bar) d2h(h, mu_bar, h->adj_mu_bar, h->E);
    TF_HIP(hipStreamSynchronize(h->stream));
    return TF_OK;
    TF_GUARD_END(h)
}

int tf_adjoint_reset(tf_handle* h)
{
    if (!h) return TF_INVALID;
    TF_GUARD_BEGIN
    ensure_adjoint(h);
    launch_fill(h, h->adj_fg_bar, 0.0, 9 * (int64_t)h->E);
    launch_fill(h, h->adj_rate_bar, 0.0, 3 * (int64_t)h->E);
    launch_fill(h, h->adj_lam_acc, 0.0, h->E);
    launch_fill(h, h->adj_mu_acc, 0.0, h->E);
    TF_HIP(hipStreamSynchronize(h->stream));
    return TF_OK;
    TF_GUARD_END(h)
}

int tf_adjoint_step(tf_handle* h, double dt, const double* xbar_next, const double* fg_prev, const double* theta_z,
                    double* xbar_prev, int32_t* krylov_iters)
{
    if (!h || !(dt > 0) || !xbar_next || !fg_prev || !theta_z || !xbar_prev) return TF_INVALID;
    TF_GUARD_BEGIN
    if (adjoint_supported(h) != TF_OK) return TF_INVALID;
    ensure_adjoint(h);
    const size_t n3 = 3 * (size_t)h->N;
    h2d(h, h->adj_xbar, xbar_next, n3);
    h2d(h, h->adj_fg_prev, fg_prev, 9 * (size_t)h->E);
    h2d(h, h->adj_tz, theta_z, h->E);
    // the Jacobian at the converged state of the step
    assemble_impl(h, dt);
    int32_t st4[4];
    d2h(h, st4, h->status, 4);
    TF_HIP(hipStreamSynchronize(h->stream));
    if (st4[0] > 0) return fail(h, TF_NEG_JACOBIAN, std::to_string(st4[0]) + " Gauss points with det Fe <= 0");
    launch_adj_split(h, h->adj_xbar, h->adj_rhs, h->adj_pass);
    const int st = solve_system_impl(h, 1, h->adj_rhs, h->adj_y);
    launch_adj_elements(h, h->adj_y, h->adj_invfg_bar, h->adj_lam_bar, h->adj_mu_bar);
    launch_accumulate(h, h->adj_lam_bar, h->adj_lam_acc, h->E);
    launch_accumulate(h, h->adj_mu_bar, h->adj_mu_acc, h->E);
    launch_adj_un(h, dt, h->adj_y, h->adj_un_bar);
    launch_accumulate(h, h->adj_pass, h->adj_un_bar, (int64_t)n3);
    launch_adj_growth(h, dt, h->adj_tz, h->adj_fg_prev, h->adj_invfg_bar, h->adj_fg_bar, h->adj_rate_bar);
    d2h(h, xbar_prev, h->adj_un_bar, n3);
    TF_HIP(hipStreamSynchronize(h->stream));
    if (krylov_iters) *krylov_iters = h->last_krylov_iters;
    if (st != TF_OK) return fail(h, st, "adjoint Krylov solve did not reach its tolerance");
    return TF_OK;
    TF_GUARD_END(h)
}

int tf_adjoint_get(tf_handle* h, double* rate_bar, double* lambda_bar, double* mu_bar, double* fg_bar)
{
    if (!h) return TF_INVALID;
    TF_GUARD_BEGIN
    ensure_adjoint(h);
    if (rate_bar) d2h(h, rate_bar, h->adj_rate_bar, 3 * (size_t)h->E);
    if (lambda_bar) d2h(h, lambda_bar, h->adj_lam_acc, h->E);
    if (mu_bar) d2h(h, mu_bar, h->adj_mu_acc, h->E);
    if (fg_bar) d2h(h, fg_bar, h->adj_fg_bar, 9 * (size_t)h->E);
    TF_HIP(hipStreamSynchronize(h->stream));
    return TF_OK;
    TF_GUARD_END(h)
}

int tf_set_tiling(tf_handle* h, int32_t wg, int32_t tile_cap)
{
    if (!h) return TF_INVALID;
    TF_GUARD_BEGIN
    build_tiles(h, wg, tile_cap);
    return TF_OK;
    TF_GUARD_END(h)
}

int tf_get_tiling(const tf_handle* h, int32_t* wg, int32_t* tile_cap, int32_t* n_tiles)
{
    if (!h) return TF_INVALID;
    if (wg) *wg = h->wg;
    if (tile_cap) *tile_cap = h->tile_cap;
    if (n_tiles) *n_tiles = h->T;
    return TF_OK;
}

#define TF_AUTOTUNE_MAX 32
int tf_autotune(tf_handle* h, double dt, int32_t n_iters, int32_t* n_variants, int32_t* wg_out, int32_t* cap_out,
                double* ms_out)
{
    if (!h || !(dt > 0) || n_iters < 2) return TF_INVALID;
    TF_GUARD_BEGIN
    // power-of-two capacities, plus "balanced" capacities that cut the rows into about
    // k x (compute units) tiles, so that every CU holds the same number of tiles
    std::vector<std::array<int, 2>> variants = {{256, 256}, {256, 512}, {256, 1024}, {512, 512},
                                                {512, 1024}, {512, 2048}, {1024, 1024}, {1024, 2048}};
    int max_row = 1;
    for (int r = 0; r < h->N; ++r) max_row = std::max(max_row, h->h_row_ptr[r + 1] - h->h_row_ptr[r]);
    for (int wg : {256, 512})
        for (int k = 1; k <= 4; ++k) {
            const int cap = (int)((h->nb + (int64_t)k * h->n_cu - 1) / ((int64_t)k * h->n_cu)) + max_row / 2;
            if (cap >= max_row && cap <= std::min(4 * wg, TF_TILE_CAP_MAX) && cap % wg)
                variants.push_back({wg, cap});
        }
    const int nv = std::min((int)variants.size(), TF_AUTOTUNE_MAX);
    // a representative system: the Jacobian at the current positions
    launch_copy(h, h->un, h->uk, 3 * (int64_t)h->N);
    hipLaunchKernelGGL(k_seg_begin, dim3((h->S + 255) / 256), dim3(256), 0, h->stream, h->seg, h->S);
    assemble_impl(h, dt);
    TF_HIP(hipStreamSynchronize(h->stream));
    double best = 1e300;
    int bi = 0;
    for (int v = 0; v < nv; ++v) {
        if (!tiling_fits(h, variants[v][0], variants[v][1])) {
            if (wg_out) wg_out[v] = variants[v][0];
            if (cap_out) cap_out[v] = variants[v][1];
            if (ms_out) ms_out[v] = -1.0;  // does not fit this device's LDS
            continue;
        }
        build_tiles(h, variants[v][0], variants[v][1]);
        const double ms = krylov_time_iteration(h, n_iters);
        if (wg_out) wg_out[v] = variants[v][0];
        if (cap_out) cap_out[v] = variants[v][1];
        if (ms_out) ms_out[v] = ms;
        if (ms > 0.0 && ms < best) { best = ms; bi = v; }
    }
    build_tiles(h, variants[bi][0], variants[bi][1]);
    if (n_variants) *n_variants = nv;
    return TF_OK;
    TF_GUARD_END(h)
}

int tf_get_coarse_operator(tf_handle* h, int32_t* nc, double* Ac, double* Ac_inv)
{
    if (!h) return TF_INVALID;
    TF_GUARD_BEGIN
    if (twolevel_export(h, nc, Ac, Ac_inv) != 0)
        return fail(h, TF_INVALID, "no coarse operator: the two-level preconditioner has not been built by a solve");
    return TF_OK;
    TF_GUARD_END(h)
}

int tf_kernel_probe(tf_handle* h, int32_t n_launches, double* avg_ms)
{
    if (!h || !avg_ms || n_launches < 2) return TF_INVALID;
    TF_GUARD_BEGIN
    *avg_ms = krylov_probe(h, n_launches);
    TF_HIP(hipStreamSynchronize(h->stream));
    return TF_OK;
    TF_GUARD_END(h)
}

namespace {
__global__ __launch_bounds__(256) void k_bw_fill(uint4* p, size_t n16)
{
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride)
        p[i] = make_uint4((unsigned)i, 1u, 2u, 3u);
}
__global__ __launch_bounds__(256) void k_bw_copy(const uint4* __restrict__ src, uint4* __restrict__ dst, size_t n16)
{
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride) dst[i] = src[i];
}
}  // namespace

int tf_copy_bandwidth(tf_handle* h, int64_t bytes, int32_t reps, double* best_GBps)
{
    if (!h || !best_GBps || bytes < (1 << 20) || bytes > ((int64_t)16 << 30) || reps < 1) return TF_INVALID;
    TF_GUARD_BEGIN
    const size_t n16 = (size_t)bytes / 16;
    uint4 *a = nullptr, *b = nullptr;
    hipEvent_t ev[2] = {nullptr, nullptr};
    double best = 0.0;
    try {
        TF_HIP(hipMalloc(&a, n16 * 16));
        TF_HIP(hipMalloc(&b, n16 * 16));
        TF_HIP(hipEventCreate(&ev[0]));
        TF_HIP(hipEventCreate(&ev[1]));
        const int grid = 256 * 8 * 4;  // 8 workgroups per CU, grid-stride over the buffers
        hipLaunchKernelGGL(k_bw_fill, dim3(grid), dim3(256), 0, h->stream, a, n16);
        hipLaunchKernelGGL(k_bw_fill, dim3(grid), dim3(256), 0, h->stream, b, n16);
        TF_HIP(hipGetLastError());
        for (int r = 0; r < reps; ++r) {
            TF_HIP(hipEventRecord(ev[0], h->stream));
            hipLaunchKernelGGL(k_bw_copy, dim3(grid), dim3(256), 0, h->stream, a, b, n16);
            TF_HIP(hipGetLastError());
            TF_HIP(hipEventRecord(ev[1], h->stream));
            TF_HIP(hipEventSynchronize(ev[1]));
            float ms = 0.f;
            TF_HIP(hipEventElapsedTime(&ms, ev[0], ev[1]));
            if (ms > 0.f) best = std::max(best, 2.0 * (double)(n16 * 16) / (ms * 1e-3) / 1e9);
        }
    } catch (...) {
        if (a) (void)hipFree(a);
        if (b) (void)hipFree(b);
        for (hipEvent_t e : ev)
            if (e) (void)hipEventDestroy(e);
        throw;
    }
    (void)hipFree(a);
    (void)hipFree(b);
    for (hipEvent_t e : ev) (void)hipEventDestroy(e);
    *best_GBps = best;
    return TF_OK;
    TF_GUARD_END(h)
}

int tf_kernel_probe_named(tf_handle* h, const char* name, int32_t n_launches, double* avg_ms)
{
    if (!h || !name || !avg_ms || n_launches < 2) return TF_INVALID;
    int which = -1;
    if (!std::strcmp(name, "spmv")) which = 0;
    else if (!std::strcmp(name, "coarse")) which = 1;
    else if (!std::strcmp(name, "persistent")) which = 2;
    if (which < 0) return fail(h, TF_INVALID, std::string("unknown kernel ") + name);
    TF_GUARD_BEGIN
    *avg_ms = krylov_probe(h, n_launches, which);
    TF_HIP(hipStreamSynchronize(h->stream));
    return TF_OK;
    TF_GUARD_END(h)
}

int tf_krylov_timing(tf_handle* h, int32_t enable, double* replay_ms, int64_t* launched, int64_t* iterations)
{
    if (!h) return TF_INVALID;
    KernelTimer& t = h->t_spmv;
    if (replay_ms) *replay_ms = t.replay_ms;
    if (launched) *launched = t.replay_launches;
    if (iterations) *iterations = t.replay_iters;
    if (enable >= 0) {
        t.enabled = enable != 0;
        t.replay_ms = 0.0;
        t.replay_launches = 0;
        t.replay_iters = 0;
    }
    return TF_OK;
}

int tf_get_segment_stats(tf_handle* h, int32_t* nr_iters, double* du_norm, int32_t* converged)
{
    if (!h) return TF_INVALID;
    for (int s = 0; s < h->S; ++s) {
        if (nr_iters) nr_iters[s] = h->seg_nr_iters[s];
        if (du_norm) du_norm[s] = h->seg_du[s];
        if (converged) converged[s] = h->seg_converged[s];
    }
    return TF_OK;
}

int tf_kernel_timing(tf_handle* h, const char* name, int32_t enable, double* avg_ms, int64_t* launches)
{
    if (!h || !name) return TF_INVALID;
    KernelTimer* t = nullptr;
    if (!std::strcmp(name, "assemble_elements")) t = &h->t_elem;
    else if (!std::strcmp(name, "assemble_gather")) t = &h->t_gather;
    else if (!std::strcmp(name, "spmv")) t = &h->t_spmv;
    if (!t) return fail(h, TF_INVALID, std::string("unknown kernel ") + name);
    TF_GUARD_BEGIN
    timer_collect(*t);
    if (avg_ms) *avg_ms = t->launches ? t->total_ms / t->launches : 0.0;
    if (launches) *launches = t->launches;
    if (enable >= 0) {
        t->enabled = enable != 0;
        t->total_ms = 0.0;
        t->launches = 0;
    }
    return TF_OK;
    TF_GUARD_END(h)
}

int tf_kernel_bytes(tf_handle* h, const char* name, double* bytes)
{
    if (!h || !name || !bytes) return TF_INVALID;
    const double E = h->E, N = h->N, nb = h->nb, n = 3.0 * h->N;
    const double vb = fp32_values(h) ? 36.0 : 72.0;
    const double mat = nb * (vb + 4.0) + (N + 1) * 4.0;  // values + column indices + row pointers
    int kind = h->opts.krylov;
    if (kind < 0) kind = h->any_viscous ? TF_KRYLOV_BICGSTAB : TF_KRYLOV_CG1;
    if (!std::strcmp(name, "coarse")) {
        // k_tp_z: the FP16 coarse inverse (nc x ldA halves), the FP32 column-block inverses and
        // P's rows once; x, r read and written, p, q read, z written; P^T r of the last iteration
        // and P^T q read, P^T r written
        if (!twolevel_active(h) || !h->tl.built) { *bytes = 0.0; return TF_OK; }
        const tf_handle::TwoLevel& tl = h->tl;
        const double D = 3.0 * tl.L, eb = h->opts.precond_fp64 ? 8.0 : 2.0, db = h->opts.precond_fp64 ? 8.0 : 4.0;
        // (P^T q and P^T r are FP32 unless the preconditioner is stored in FP64: db)
        *bytes = (double)tl.ainv_halves * eb + (double)tl.ncol * D * D * db + N * 18.0 * 8.0 + n * 8.0 * 7.0 +
                 tl.nc * db * 3.0;
    } else if (!std::strcmp(name, "persistent")) {
        // one iteration of k_tp_run: the blocks' values (read again every iteration), P's rows (both
        // halves), the column-block inverses; z and p published and their neighbours' read; P^T q and
        // P^T r written and gathered (the coarse inverse rows stay in LDS for the whole solve)
        if (!twolevel_active(h) || !h->tl.built) { *bytes = 0.0; return TF_OK; }
        const tf_handle::TwoLevel& tl = h->tl;
        const double D = 3.0 * tl.L, db = h->opts.precond_fp64 ? 8.0 : 4.0;
        *bytes = nb * vb + 2.0 * N * 18.0 * 8.0 + (double)tl.ncol * D * D * db + 4.0 * n * 8.0 + 4.0 * tl.nc * db;
    } else if (!std::strcmp(name, "spmv")) {
        // the launch timed under "spmv": the dominant Krylov kernel of the active solver
        if (twolevel_active(h))  // k_tp_q: SpMV on p = z + beta p_old: z, p_old read; p_new, q written; P's rows
            *bytes = mat + n * 8.0 * 4.0 + N * 18.0 * 8.0;
        else if (kind == TF_KRYLOV_CG1)  // SpMV; x,p,r,w,s,u,m,q read and x,p,r,w,s,u,m,q written once; D^-1 rows
            *bytes = mat + n * 8.0 * 16.0 + N * 72.0;
        else if (kind == TF_KRYLOV_PIPECG)  // SpMV on m; z,q,s,p,x,r,u,w,m read, the same and m' written; D^-1 rows
            *bytes = mat + n * 8.0 * 18.0 + N * 72.0;
        else if (kind == TF_KRYLOV_CG)  // SpMV on p = z + beta p_old: z, p_old read; p_new, q written
            *bytes = mat + n * 8.0 * 4.0;
        else  // BiCGSTAB SpMV with one fused dot product: x, w read; y written
            *bytes = mat + n * 8.0 * 3.0;
    } else if (!std::strcmp(name, "assemble_elements"))
        *bytes = E * (24.0 + 144.0 + 72.0 + 16.0 + 144.0 + (h->any_viscous ? 152.0 : 0.0) + 324.0 * 8.0 + 18.0 * 8.0);
    else if (!std::strcmp(name, "assemble_gather"))
        *bytes = E * 324.0 * 8.0 + E * 36.0 * 4.0 + nb * 4.0 * 3 + nb * 72.0 + N * 72.0 + N * (8.0 + 24.0 + 3.0);
    else return fail(h, TF_INVALID, std::string("unknown kernel ") + name);
    return TF_OK;
}

}  // extern "C"
