// tf_krylov.hip -- FP64 preconditioned Krylov solvers for M du = g on the BSR3 Jacobian (gfx950).
//
// Work unit: a row tile (a run of block rows of one tissue holding at most TF_TILE_MAX_BLOCKS
// 3x3 blocks).  One 256-lane workgroup per tile.  The SpMV is CSR-stream: lanes stream the
// tile's blocks contiguously (component-major values, so every load instruction is coalesced),
// park the 3-vector products in LDS, then one lane per (row, component) sums its row.
// Dot products are fused into the kernel that produces their operands; every tile writes its
// partial sums and the last tile of a tissue to arrive (atomic ticket) sums the partials in
// tile order and updates that tissue's scalars -- deterministic, no host round trip.
//
// Preconditioner: 3x3 nodal block Jacobi (diagonal blocks inverted during assembly).
//   BiCGSTAB (van der Vorst), right preconditioned, 5 kernels / iteration, 2 SpMV.
//   CG (symmetric positive definite M, the case of no internal viscosity), 2 kernels /
//   iteration: the direction update p = z + beta p_old is folded into the SpMV, which forms
//   p for neighbouring rows on the fly and writes it for its own rows (ping-pong buffers).
#include "tf_internal.hpp"
#include <algorithm>
#include <math.h>

namespace {

constexpr int WG = 256;

struct TileArgs {
    const int32_t* tile_r0;
    const int32_t* tile_r1;
    const int32_t* tile_seg;
    const int32_t* seg_tile_ptr;
    int32_t* counter;
    double* partial;
    SegScalars* seg;
};

__device__ inline bool seg_done(const SegScalars& s) { return s.converged || s.stalled; }

__device__ inline double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Block-wide sum of two values (all lanes of the block must call). Result valid in lane 0.
__device__ inline void block_sum2(double& a, double& b)
{
    __shared__ double ra[WG / 64], rb[WG / 64];
    a = wave_sum(a);
    b = wave_sum(b);
    const int w = threadIdx.x / 64, l = threadIdx.x % 64;
    if (l == 0) { ra[w] = a; rb[w] = b; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double sa = 0.0, sb = 0.0;
        for (int k = 0; k < WG / 64; ++k) { sa += ra[k]; sb += rb[k]; }
        a = sa; b = sb;
    }
}

// Write this tile's partials; returns true (in lane 0) when this tile is the last of its tissue,
// in which case tot[] holds the tissue totals summed in tile order.
__device__ inline bool tile_reduce(const TileArgs& ta, int t, double a, double b, double* tot)
{
    __shared__ int last;
    block_sum2(a, b);
    const int sg = ta.tile_seg[t];
    if (threadIdx.x == 0) {
        ta.partial[t * TF_RED_SLOTS + 0] = a;
        ta.partial[t * TF_RED_SLOTS + 1] = b;
        __threadfence();
        const int nt = ta.seg_tile_ptr[sg + 1] - ta.seg_tile_ptr[sg];
        last = (atomicAdd(&ta.counter[sg], 1) == nt - 1);
    }
    __syncthreads();
    if (threadIdx.x == 0 && last) {
        __threadfence();
        double s0 = 0.0, s1 = 0.0;
        for (int k = ta.seg_tile_ptr[sg]; k < ta.seg_tile_ptr[sg + 1]; ++k) {
            s0 += __hip_atomic_load(&ta.partial[k * TF_RED_SLOTS + 0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            s1 += __hip_atomic_load(&ta.partial[k * TF_RED_SLOTS + 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        tot[0] = s0;
        tot[1] = s1;
        ta.counter[sg] = 0;
        return true;
    }
    return false;
}

__device__ inline void dinv_apply(const double* __restrict__ dinv, const double* __restrict__ r, int d, double& z)
{
    const int A = d / 3, i = d % 3;
    const double* D = dinv + 9 * A + 3 * i;
    const double* rr = r + 3 * A;
    z = D[0] * rr[0] + D[1] * rr[1] + D[2] * rr[2];
}

// CSR-stream product of one tile: prod[3*(j-b0)+c] = (A_j x_col(j))_c, x given by xin(col, c).
template <typename VT, typename XF>
__device__ inline void tile_products(const int32_t* __restrict__ row_ptr, const int32_t* __restrict__ col,
                                     const VT* __restrict__ val, int nb, int b0, int b1, XF xin, double* prod)
{
    for (int j = b0 + threadIdx.x; j < b1; j += WG) {
        const int c = col[j];
        double x0, x1, x2;
        xin(c, x0, x1, x2);
        const double v0 = val[j], v1 = val[(size_t)nb + j], v2 = val[2 * (size_t)nb + j];
        const double v3 = val[3 * (size_t)nb + j], v4 = val[4 * (size_t)nb + j], v5 = val[5 * (size_t)nb + j];
        const double v6 = val[6 * (size_t)nb + j], v7 = val[7 * (size_t)nb + j], v8 = val[8 * (size_t)nb + j];
        double* p = prod + 3 * (j - b0);
        p[0] = v0 * x0 + v1 * x1 + v2 * x2;
        p[1] = v3 * x0 + v4 * x1 + v5 * x2;
        p[2] = v6 * x0 + v7 * x1 + v8 * x2;
    }
    __syncthreads();
}

__device__ inline double row_sum(const int32_t* __restrict__ row_ptr, int r, int comp, int b0, const double* prod)
{
    double s = 0.0;
    const int j1 = row_ptr[r + 1];
    for (int j = row_ptr[r]; j < j1; ++j) s += prod[3 * (j - b0) + comp];
    return s;
}

// ---------------------------------------------------------------------------------------------
// plain y = M x
template <typename VT>
__global__ __launch_bounds__(WG) void k_spmv(const int32_t* __restrict__ row_ptr, const int32_t* __restrict__ col,
                                             const VT* __restrict__ val, int nb, const int32_t* __restrict__ tile_r0,
                                             const int32_t* __restrict__ tile_r1, const double* __restrict__ x,
                                             double* __restrict__ y)
{
    __shared__ double prod[3 * TF_TILE_MAX_BLOCKS];
    const int t = blockIdx.x;
    const int r0 = tile_r0[t], r1 = tile_r1[t];
    const int b0 = row_ptr[r0], b1 = row_ptr[r1];
    tile_products<VT>(row_ptr, col, val, nb, b0, b1,
                      [&](int c, double& a0, double& a1, double& a2) { a0 = x[3 * c]; a1 = x[3 * c + 1]; a2 = x[3 * c + 2]; },
                      prod);
    for (int task = threadIdx.x; task < 3 * (r1 - r0); task += WG) {
        const int r = r0 + task / 3, comp = task % 3;
        y[3 * r + comp] = row_sum(row_ptr, r, comp, b0, prod);
    }
}

// ---------------------------------------------------------------------------------------------
// BiCGSTAB
__global__ __launch_bounds__(WG) void k_bicg_init(TileArgs ta, const double* __restrict__ b, double* x, double* r, double* r0,
                                                  double* p, double* v, double rtol2, double atol2)
{
    const int t = blockIdx.x;
    const int sg = ta.tile_seg[t];
    const bool active = ta.seg[sg].active;
    const int d0 = 3 * ta.tile_r0[t], d1 = 3 * ta.tile_r1[t];
    double s = 0.0;
    for (int d = d0 + threadIdx.x; d < d1; d += WG) {
        const double bv = active ? b[d] : 0.0;
        x[d] = 0.0; r[d] = bv; r0[d] = bv; p[d] = 0.0; v[d] = 0.0;
        s += bv * bv;
    }
    double tot[2];
    if (tile_reduce(ta, t, s, 0.0, tot)) {
        SegScalars& S = ta.seg[sg];
        S.bnorm2 = tot[0];
        S.rho = tot[0];
        S.rho_prev = 1.0; S.alpha = 1.0; S.omega = 1.0;
        S.beta = tot[0];
        S.rnorm2 = tot[0];
        S.tol2 = fmax(rtol2 * tot[0], atol2);
        S.iters = 0;
        S.stalled = 0;
        S.converged = (!active || tot[0] <= S.tol2) ? 1 : 0;
    }
}

__global__ __launch_bounds__(WG) void k_bicg_p(TileArgs ta, const double* __restrict__ dinv, const double* __restrict__ r,
                                               double* __restrict__ p, const double* __restrict__ v, double* __restrict__ ph)
{
    const int t = blockIdx.x;
    const SegScalars& S = ta.seg[ta.tile_seg[t]];
    if (seg_done(S)) return;
    const double beta = S.beta, omega = S.omega;
    const int d0 = 3 * ta.tile_r0[t], d1 = 3 * ta.tile_r1[t];
    for (int d = d0 + threadIdx.x; d < d1; d += WG) p[d] = r[d] + beta * (p[d] - omega * v[d]);
    __syncthreads();
    for (int d = d0 + threadIdx.x; d < d1; d += WG) {
        double z;
        dinv_apply(dinv, p, d, z);
        ph[d] = z;
    }
}

// y = M x with fused dot products; MODE 0: (r0, y) -> alpha; MODE 1: (y, s), (y, y) -> omega
template <typename VT, int MODE>
__global__ __launch_bounds__(WG) void k_bicg_spmv(TileArgs ta, const int32_t* __restrict__ row_ptr,
                                                  const int32_t* __restrict__ col, const VT* __restrict__ val, int nb,
                                                  const double* __restrict__ x, double* __restrict__ y,
                                                  const double* __restrict__ w)
{
    __shared__ double prod[3 * TF_TILE_MAX_BLOCKS];
    const int t = blockIdx.x;
    const int sg = ta.tile_seg[t];
    if (seg_done(ta.seg[sg])) return;
    const int r0 = ta.tile_r0[t], r1 = ta.tile_r1[t];
    const int b0 = row_ptr[r0], b1 = row_ptr[r1];
    tile_products<VT>(row_ptr, col, val, nb, b0, b1,
                      [&](int c, double& a0, double& a1, double& a2) { a0 = x[3 * c]; a1 = x[3 * c + 1]; a2 = x[3 * c + 2]; },
                      prod);
    double s0 = 0.0, s1 = 0.0;
    for (int task = threadIdx.x; task < 3 * (r1 - r0); task += WG) {
        const int r = r0 + task / 3, comp = task % 3;
        const int d = 3 * r + comp;
        const double yv = row_sum(row_ptr, r, comp, b0, prod);
        y[d] = yv;
        if (MODE == 0) s0 += w[d] * yv;
        else { s0 += yv * w[d]; s1 += yv * yv; }
    }
    double tot[2];
    if (tile_reduce(ta, t, s0, s1, tot)) {
        SegScalars& S = ta.seg[sg];
        if (MODE == 0) {
            if (!(tot[0] != 0.0) || !isfinite(tot[0])) S.stalled = 1;
            else S.alpha = S.rho / tot[0];
        } else {
            S.omega = (tot[1] > 0.0) ? tot[0] / tot[1] : 0.0;
        }
    }
}

__global__ __launch_bounds__(WG) void k_bicg_s(TileArgs ta, const double* __restrict__ dinv, const double* __restrict__ r,
                                               const double* __restrict__ v, double* __restrict__ s, double* __restrict__ sh)
{
    const int t = blockIdx.x;
    const SegScalars& S = ta.seg[ta.tile_seg[t]];
    if (seg_done(S)) return;
    const double alpha = S.alpha;
    const int d0 = 3 * ta.tile_r0[t], d1 = 3 * ta.tile_r1[t];
    for (int d = d0 + threadIdx.x; d < d1; d += WG) s[d] = r[d] - alpha * v[d];
    __syncthreads();
    for (int d = d0 + threadIdx.x; d < d1; d += WG) {
        double z;
        dinv_apply(dinv, s, d, z);
        sh[d] = z;
    }
}

__global__ __launch_bounds__(WG) void k_bicg_x(TileArgs ta, const double* __restrict__ ph, const double* __restrict__ sh,
                                               const double* __restrict__ s, const double* __restrict__ tv,
                                               const double* __restrict__ r0, double* __restrict__ x, double* __restrict__ r,
                                               int maxit)
{
    const int t = blockIdx.x;
    const int sg = ta.tile_seg[t];
    if (seg_done(ta.seg[sg])) return;
    const double alpha = ta.seg[sg].alpha, omega = ta.seg[sg].omega;
    const int d0 = 3 * ta.tile_r0[t], d1 = 3 * ta.tile_r1[t];
    double a = 0.0, b = 0.0;
    for (int d = d0 + threadIdx.x; d < d1; d += WG) {
        x[d] = x[d] + alpha * ph[d] + omega * sh[d];
        const double rn = s[d] - omega * tv[d];
        r[d] = rn;
        a += r0[d] * rn;
        b += rn * rn;
    }
    double tot[2];
    if (tile_reduce(ta, t, a, b, tot)) {
        SegScalars& S = ta.seg[sg];
        S.rho_prev = S.rho;
        S.rho = tot[0];
        S.rnorm2 = tot[1];
        S.iters += 1;
        if (!isfinite(tot[1])) S.stalled = 1;
        else if (tot[1] <= S.tol2) S.converged = 1;
        else if (S.iters >= maxit || tot[0] == 0.0 || S.omega == 0.0) S.stalled = 1;
        else S.beta = (S.rho / S.rho_prev) * (S.alpha / S.omega);
    }
}

// ---------------------------------------------------------------------------------------------
// CG
__global__ __launch_bounds__(WG) void k_cg_init(TileArgs ta, const double* __restrict__ dinv, const double* __restrict__ b,
                                                double* x, double* r, double* z, double* p_old, double rtol2, double atol2)
{
    const int t = blockIdx.x;
    const int sg = ta.tile_seg[t];
    const bool active = ta.seg[sg].active;
    const int d0 = 3 * ta.tile_r0[t], d1 = 3 * ta.tile_r1[t];
    for (int d = d0 + threadIdx.x; d < d1; d += WG) {
        x[d] = 0.0; r[d] = active ? b[d] : 0.0; p_old[d] = 0.0;
    }
    __syncthreads();
    double a = 0.0, bb = 0.0;
    for (int d = d0 + threadIdx.x; d < d1; d += WG) {
        double zz;
        dinv_apply(dinv, r, d, zz);
        z[d] = zz;
        a += r[d] * zz;
        bb += r[d] * r[d];
    }
    double tot[2];
    if (tile_reduce(ta, t, a, bb, tot)) {
        SegScalars& S = ta.seg[sg];
        S.rho = tot[0];
        S.bnorm2 = tot[1];
        S.rnorm2 = tot[1];
        S.beta = 0.0;
        S.tol2 = fmax(rtol2 * tot[1], atol2);
        S.iters = 0;
        S.stalled = 0;
        S.converged = (!active || tot[1] <= S.tol2) ? 1 : 0;
    }
}

// p = z + beta p_old (own rows written to p_new, neighbours formed on the fly), q = M p, (p, q) -> alpha
template <typename VT>
__global__ __launch_bounds__(WG) void k_cg_q(TileArgs ta, const int32_t* __restrict__ row_ptr, const int32_t* __restrict__ col,
                                             const VT* __restrict__ val, int nb, const double* __restrict__ z,
                                             const double* __restrict__ p_old, double* __restrict__ p_new,
                                             double* __restrict__ q)
{
    __shared__ double prod[3 * TF_TILE_MAX_BLOCKS];
    const int t = blockIdx.x;
    const int sg = ta.tile_seg[t];
    if (seg_done(ta.seg[sg])) return;
    const double beta = ta.seg[sg].beta;
    const int r0 = ta.tile_r0[t], r1 = ta.tile_r1[t];
    const int b0 = row_ptr[r0], b1 = row_ptr[r1];
    tile_products<VT>(row_ptr, col, val, nb, b0, b1,
                      [&](int c, double& a0, double& a1, double& a2) {
                          a0 = z[3 * c] + beta * p_old[3 * c];
                          a1 = z[3 * c + 1] + beta * p_old[3 * c + 1];
                          a2 = z[3 * c + 2] + beta * p_old[3 * c + 2];
                      },
                      prod);
    double s0 = 0.0;
    for (int task = threadIdx.x; task < 3 * (r1 - r0); task += WG) {
        const int r = r0 + task / 3, comp = task % 3;
        const int d = 3 * r + comp;
        const double qv = row_sum(row_ptr, r, comp, b0, prod);
        const double pv = z[d] + beta * p_old[d];
        q[d] = qv;
        p_new[d] = pv;
        s0 += pv * qv;
    }
    double tot[2];
    if (tile_reduce(ta, t, s0, 0.0, tot)) {
        SegScalars& S = ta.seg[sg];
        if (!(tot[0] > 0.0) || !isfinite(tot[0])) S.stalled = 1;
        else S.alpha = S.rho / tot[0];
    }
}

__global__ __launch_bounds__(WG) void k_cg_x(TileArgs ta, const double* __restrict__ dinv, const double* __restrict__ p,
                                             const double* __restrict__ q, double* __restrict__ x, double* __restrict__ r,
                                             double* __restrict__ z, int maxit)
{
    const int t = blockIdx.x;
    const int sg = ta.tile_seg[t];
    if (seg_done(ta.seg[sg])) return;
    const double alpha = ta.seg[sg].alpha;
    const int d0 = 3 * ta.tile_r0[t], d1 = 3 * ta.tile_r1[t];
    for (int d = d0 + threadIdx.x; d < d1; d += WG) {
        x[d] = x[d] + alpha * p[d];
        r[d] = r[d] - alpha * q[d];
    }
    __syncthreads();
    double a = 0.0, b = 0.0;
    for (int d = d0 + threadIdx.x; d < d1; d += WG) {
        double zz;
        dinv_apply(dinv, r, d, zz);
        z[d] = zz;
        a += r[d] * zz;
        b += r[d] * r[d];
    }
    double tot[2];
    if (tile_reduce(ta, t, a, b, tot)) {
        SegScalars& S = ta.seg[sg];
        S.iters += 1;
        S.rnorm2 = tot[1];
        if (!isfinite(tot[1])) S.stalled = 1;
        else if (tot[1] <= S.tol2) S.converged = 1;
        else if (S.iters >= maxit) S.stalled = 1;
        else {
            S.beta = tot[0] / S.rho;
            S.rho = tot[0];
        }
    }
}

__global__ void k_mark_active(SegScalars* seg, int S)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < S) { seg[s].converged = 0; seg[s].stalled = 0; }
}

TileArgs tile_args(tf_handle* h)
{
    return TileArgs{h->tile_r0, h->tile_r1, h->tile_seg, h->seg_tile_ptr, h->seg_counter, h->partial, h->seg};
}

template <typename VT>
void enqueue_bicg_iter(tf_handle* h, const VT* val)
{
    TileArgs ta = tile_args(h);
    const dim3 grid(h->T), blk(WG);
    hipLaunchKernelGGL(k_bicg_p, grid, blk, 0, h->stream, ta, h->dinv, h->kr, h->kp, h->kv, h->kph);
    hipEvent_t ev[2];
    timer_begin(h, h->t_spmv, ev);
    hipLaunchKernelGGL((k_bicg_spmv<VT, 0>), grid, blk, 0, h->stream, ta, h->row_ptr, h->col, val, h->nb, h->kph, h->kv,
                       h->kr0);
    timer_end(h, h->t_spmv, ev);
    hipLaunchKernelGGL(k_bicg_s, grid, blk, 0, h->stream, ta, h->dinv, h->kr, h->kv, h->ks, h->ksh);
    timer_begin(h, h->t_spmv, ev);
    hipLaunchKernelGGL((k_bicg_spmv<VT, 1>), grid, blk, 0, h->stream, ta, h->row_ptr, h->col, val, h->nb, h->ksh, h->kt,
                       h->ks);
    timer_end(h, h->t_spmv, ev);
    hipLaunchKernelGGL(k_bicg_x, grid, blk, 0, h->stream, ta, h->kph, h->ksh, h->ks, h->kt, h->kr0, h->kx, h->kr,
                       h->opts.krylov_max_iter);
}

// CG uses kp / kph as the ping-pong direction buffers, kv as q, ks as z.
template <typename VT>
void enqueue_cg_iter(tf_handle* h, const VT* val, int parity)
{
    TileArgs ta = tile_args(h);
    const dim3 grid(h->T), blk(WG);
    double* p_old = parity ? h->kph : h->kp;
    double* p_new = parity ? h->kp : h->kph;
    hipEvent_t ev[2];
    timer_begin(h, h->t_spmv, ev);
    hipLaunchKernelGGL((k_cg_q<VT>), grid, blk, 0, h->stream, ta, h->row_ptr, h->col, val, h->nb, h->ks, p_old, p_new, h->kv);
    timer_end(h, h->t_spmv, ev);
    hipLaunchKernelGGL(k_cg_x, grid, blk, 0, h->stream, ta, h->dinv, p_new, h->kv, h->kx, h->kr, h->ks,
                       h->opts.krylov_max_iter);
}

template <typename VT>
void enqueue_chunk(tf_handle* h, const VT* val, int kind, int iters)
{
    for (int i = 0; i < iters; ++i) {
        if (kind == TF_KRYLOV_CG) enqueue_cg_iter<VT>(h, val, i & 1);
        else enqueue_bicg_iter<VT>(h, val);
    }
}

bool all_done(tf_handle* h)
{
    TF_HIP(hipMemcpyAsync(h->h_seg, h->seg, sizeof(SegScalars) * h->S, hipMemcpyDeviceToHost, h->stream));
    TF_HIP(hipStreamSynchronize(h->stream));
    for (int s = 0; s < h->S; ++s)
        if (!(h->h_seg[s].converged || h->h_seg[s].stalled)) return false;
    return true;
}

template <typename VT>
int solve_kind(tf_handle* h, const VT* val, int kind)
{
    TileArgs ta = tile_args(h);
    const dim3 grid(h->T), blk(WG);
    const double rtol2 = h->opts.krylov_rtol * h->opts.krylov_rtol;
    const double atol2 = h->opts.krylov_atol * h->opts.krylov_atol;
    if (kind == TF_KRYLOV_CG)
        hipLaunchKernelGGL(k_cg_init, grid, blk, 0, h->stream, ta, h->dinv, h->g, h->kx, h->kr, h->ks, h->kp, rtol2, atol2);
    else
        hipLaunchKernelGGL(k_bicg_init, grid, blk, 0, h->stream, ta, h->g, h->kx, h->kr, h->kr0, h->kp, h->kv, rtol2,
                           atol2);
    TF_HIP(hipGetLastError());
    int chunk = std::max(2, h->opts.iters_per_sync);
    chunk += chunk & 1;  // CG ping-pong needs an even chunk
    const bool timing = h->t_spmv.enabled;
    const bool graph = h->opts.use_graph && !timing;
    int done_iters = 0;
    while (true) {
        if (graph) {
            if (!h->krylov_graph || h->krylov_graph_iters != chunk || h->krylov_graph_kind != kind * 4 + (int)sizeof(VT)) {
                if (h->krylov_graph) { hipGraphExecDestroy(h->krylov_graph); h->krylov_graph = nullptr; }
                hipGraph_t gr;
                TF_HIP(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
                enqueue_chunk<VT>(h, val, kind, chunk);
                TF_HIP(hipStreamEndCapture(h->stream, &gr));
                TF_HIP(hipGraphInstantiate(&h->krylov_graph, gr, nullptr, nullptr, 0));
                TF_HIP(hipGraphDestroy(gr));
                h->krylov_graph_iters = chunk;
                h->krylov_graph_kind = kind * 4 + (int)sizeof(VT);
            }
            TF_HIP(hipGraphLaunch(h->krylov_graph, h->stream));
        } else {
            enqueue_chunk<VT>(h, val, kind, chunk);
            TF_HIP(hipGetLastError());
        }
        done_iters += chunk;
        if (all_done(h)) break;
        if (done_iters > h->opts.krylov_max_iter + 2 * chunk) break;  // scalars mark the stall
    }
    timer_collect(h->t_spmv);
    int st = TF_OK;
    int maxit = 0;
    for (int s = 0; s < h->S; ++s) {
        maxit = std::max(maxit, h->h_seg[s].iters);
        if (h->h_seg[s].active && h->h_seg[s].stalled) st = TF_KRYLOV_STALL;
    }
    h->last_krylov_iters = maxit;
    return st;
}

}  // namespace

// Solve M du = g for every active tissue; du is written to h->du.
int krylov_solve(tf_handle* h)
{
    int kind = h->opts.krylov;
    const bool fp32 = h->opts.matrix_precision == TF_PREC_FP32;
    auto run = [&](int k) {
        return fp32 ? solve_kind<float>(h, h->val32, k) : solve_kind<double>(h, h->val, k);
    };
    int st = run(kind);
    if (st == TF_KRYLOV_STALL && kind == TF_KRYLOV_CG) {
        // not positive definite (or too slow): fall back to BiCGSTAB from zero
        int cg_iters = h->last_krylov_iters;
        st = run(TF_KRYLOV_BICGSTAB);
        h->last_krylov_iters += cg_iters;
    }
    launch_copy(h, h->du, h->kx, 3 * (int64_t)h->N);
    return st;
}

void launch_spmv(tf_handle* h, const double* x, double* y)
{
    if (h->opts.matrix_precision == TF_PREC_FP32)
        hipLaunchKernelGGL(k_spmv<float>, dim3(h->T), dim3(WG), 0, h->stream, h->row_ptr, h->col, h->val32, h->nb,
                           h->tile_r0, h->tile_r1, x, y);
    else
        hipLaunchKernelGGL(k_spmv<double>, dim3(h->T), dim3(WG), 0, h->stream, h->row_ptr, h->col, h->val, h->nb,
                           h->tile_r0, h->tile_r1, x, y);
    TF_HIP(hipGetLastError());
}
