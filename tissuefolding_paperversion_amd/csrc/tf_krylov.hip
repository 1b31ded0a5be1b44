// tf_krylov.hip -- FP64 preconditioned Krylov solvers for M du = g on the BSR3 Jacobian (gfx950).
//
// Work unit: a row tile (a run of block rows of one tissue holding at most tile_cap 3x3 blocks
// and at most wg/3 block rows).  One workgroup per tile (256 lanes; the single-kernel CG uses
// the tunable wg).  The SpMV is CSR-stream: lanes stream the
// tile's blocks contiguously (component-major values, so every load instruction is coalesced),
// park the 3-vector products in LDS, then one lane per (row, component) sums its row.
//
// Reductions carry no atomics and no device-wide fences: the kernel producing a dot product's
// operands writes one partial per tile (partial[slot*T + tile]); every tile of the next kernel
// that needs the scalar re-sums its tissue's partials with a fixed lane/tree order.  All tiles
// of a tissue therefore derive bit-identical scalars (alpha, beta, omega, convergence decision)
// and kernel ordering alone makes the partials visible.  The first tile of a tissue records the
// outcome (converged / stalled / iteration count) for the host.
//
// Preconditioner: 3x3 nodal block Jacobi (diagonal blocks inverted during assembly).
//   BiCGSTAB (van der Vorst), right preconditioned, 5 kernels / iteration, 2 SpMV.
//   CG (symmetric positive definite M, the case of no internal viscosity), 2 kernels /
//   iteration: the direction update p = z + beta p_old is folded into the SpMV, which forms
//   p for neighbouring rows on the fly and writes it for its own rows (ping-pong buffers).
#include "tf_internal.hpp"
#include <hip/hip_fp16.h>
#include <algorithm>
#include <cmath>
#include <cstring>
#include <math.h>
#include <numeric>
#include <type_traits>
#include <vector>

namespace {

constexpr int WG = 256;  // workgroup of every tile kernel except the single-kernel CG (tunable)

struct TileArgs {
    const TileInfo* info;  // [T] one record per tile (the arrays below, in one scalar load)
    const int32_t* tile_r0;
    const int32_t* tile_r1;
    const int32_t* tile_seg;
    const int32_t* seg_tile_ptr;
    double* partial;
    SegScalars* seg;
    int T;
    int nxcd;  // XCDs the dispatcher deals workgroups to round-robin (1: identity tile map)
};

// Workgroup b runs on XCD b % nxcd.  Giving XCD x a contiguous run of tiles keeps a tile's
// neighbouring block columns -- mostly in the adjacent tiles -- in the same XCD's L2 instead of
// every XCD fetching nearly all of the gathered vectors.  A bijection of [0, T).
__device__ inline int xcd_tile(int b, int T, int nxcd)
{
    if (nxcd <= 1) return b;
    const int x = b % nxcd, k = b / nxcd, q = T / nxcd, rem = T % nxcd;
    return x * q + (x < rem ? x : rem) + k;
}

// v of the lane CTRL names (a DPP row shift or broadcast), 0 where there is none or where the row is
// masked off (ROW_MASK: one bit per row of 16 lanes)
template <int CTRL, int ROW_MASK>
__device__ inline double dpp_take(double v)
{
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, ROW_MASK, 0xf, true);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, ROW_MASK, 0xf, true);
    return __hiloint2double(hi, lo);
}

// Sum over the 64 lanes of a wave, returned in every lane; every lane must be active.  An inclusive
// scan by DPP row shifts (1, 2, 4, 8) inside each row of 16 lanes, then lane 15 of each row added
// into the next row (rows 1 and 3) and lane 31 into the upper half, so lane 63 holds the total.
// Register moves between neighbouring lanes instead of LDS permutes: the same fixed order in every
// wave.
__device__ inline double wave_sum(double v)
{
    v += dpp_take<0x111, 0xf>(v);  // row_shr:1
    v += dpp_take<0x112, 0xf>(v);  // row_shr:2
    v += dpp_take<0x114, 0xf>(v);  // row_shr:4
    v += dpp_take<0x118, 0xf>(v);  // row_shr:8
    v += dpp_take<0x142, 0xa>(v);  // row_bcast:15 into rows 1 and 3
    v += dpp_take<0x143, 0xc>(v);  // row_bcast:31 into rows 2 and 3
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), 63);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), 63);
    return __hiloint2double(hi, lo);
}

// Tile partials: reduce v[q] over the workgroup and store partial[slot[q]*T + t].
template <int K, int WGT = WG>
__device__ inline void tile_write(const TileArgs& ta, int t, const int (&slot)[K], double (&v)[K])
{
    constexpr int NW = WGT / 64;
    __shared__ double red[K][NW];
    const int w = threadIdx.x / 64, l = threadIdx.x % 64;
#pragma unroll
    for (int q = 0; q < K; ++q) {
        v[q] = wave_sum(v[q]);
        if (l == 0) red[q][w] = v[q];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < K; ++q) {
            double s = 0.0;
#pragma unroll
            for (int k = 0; k < NW; ++k) s += red[q][k];
            ta.partial[(size_t)slot[q] * ta.T + t] = s;
        }
    }
}

// Tissue totals of K partial slots, identical in every lane and every tile of the tissue.
template <int K, int WGT = WG>
__device__ inline void seg_sums_range(const TileArgs& ta, int t0, int t1, const int (&slot)[K], double (&out)[K])
{
    constexpr int NW = WGT / 64;
    __shared__ double red[K][NW];
    double s[K];
#pragma unroll
    for (int q = 0; q < K; ++q) s[q] = 0.0;
    for (int k = t0 + threadIdx.x; k < t1; k += WGT)
#pragma unroll
        for (int q = 0; q < K; ++q) s[q] += ta.partial[(size_t)slot[q] * ta.T + k];
    const int w = threadIdx.x / 64, l = threadIdx.x % 64;
#pragma unroll
    for (int q = 0; q < K; ++q) {
        s[q] = wave_sum(s[q]);
        if (l == 0) red[q][w] = s[q];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < K; ++q) {
        double a = 0.0;
#pragma unroll
        for (int k = 0; k < NW; ++k) a += red[q][k];
        out[q] = a;
    }
    __syncthreads();
}

template <int K, int WGT = WG>
__device__ inline void seg_sums(const TileArgs& ta, int sg, const int (&slot)[K], double (&out)[K])
{
    seg_sums_range<K, WGT>(ta, ta.seg_tile_ptr[sg], ta.seg_tile_ptr[sg + 1], slot, out);
}

__device__ inline bool first_tile(const TileArgs& ta, int t, int sg) { return t == ta.seg_tile_ptr[sg]; }

__device__ inline bool seg_done(const SegScalars& s) { return s.converged || s.stalled; }

__device__ inline double dinv_row(const double* __restrict__ dinv, const double* __restrict__ r, int d)
{
    const int A = d / 3, i = d % 3;
    const double* D = dinv + 9 * A + 3 * i;
    const double* rr = r + 3 * A;
    return D[0] * rr[0] + D[1] * rr[1] + D[2] * rr[2];
}

// CSR-stream products of one tile: prod[3*(j-b0)+c] = (A_j x_col(j))_c, x given by xin(col, c).
template <typename VT, typename XF>
__device__ inline void tile_products(const int32_t* __restrict__ col, const VT* __restrict__ val, int nb, int b0, int b1,
                                     XF xin, double* prod)
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

#define XLOAD(vec) [&](int c, double& a0, double& a1, double& a2) { a0 = vec[3 * c]; a1 = vec[3 * c + 1]; a2 = vec[3 * c + 2]; }

// ---------------------------------------------------------------------------------------------
// plain y = M x
template <typename VT>
__global__ __launch_bounds__(WG) void k_spmv(const int32_t* __restrict__ row_ptr, const int32_t* __restrict__ col,
                                             const VT* __restrict__ val, int nb, const int32_t* __restrict__ tile_r0,
                                             const int32_t* __restrict__ tile_r1, const double* __restrict__ x,
                                             double* __restrict__ y)
{
    extern __shared__ double prod[];  // 3 doubles per block of the tile (tile_cap blocks)
    const int t = blockIdx.x;
    const int r0 = tile_r0[t], r1 = tile_r1[t];
    const int b0 = row_ptr[r0], b1 = row_ptr[r1];
    tile_products<VT>(col, val, nb, b0, b1, XLOAD(x), prod);
    for (int task = threadIdx.x; task < 3 * (r1 - r0); task += WG) {
        const int r = r0 + task / 3, comp = task % 3;
        y[3 * r + comp] = row_sum(row_ptr, r, comp, b0, prod);
    }
}

// one workgroup per tissue: ||b||^2 -> tolerance, initial state of the tissue's solve
// (seg_rtol2, when given, holds a relative tolerance per tissue; else rtol2 applies to all)
__global__ __launch_bounds__(WG) void k_seg_init(TileArgs ta, double rtol2, double atol2, const double* seg_rtol2)
{
    const int sg = blockIdx.x;
    const int sl[1] = {SL_BB};
    double tot[1];
    seg_sums<1>(ta, sg, sl, tot);
    if (threadIdx.x == 0) {
        SegScalars& S = ta.seg[sg];
        S.bnorm2 = tot[0];
        S.rnorm2 = tot[0];
        S.tol2 = fmax((seg_rtol2 ? seg_rtol2[sg] : rtol2) * tot[0], atol2);
        S.iters = 0;
        S.stalled = 0;
        for (int q = 0; q < 2; ++q) {
            S.alpha_hist[q] = 0.0;
            S.gam_hist[q] = 0.0;
            S.best_rr_h[q] = tot[0];
            S.iters_h[q] = 0;
            S.best_it_h[q] = 0;
        }
        S.converged = (!S.active || tot[0] <= S.tol2) ? 1 : 0;
    }
}

// reopens every tissue's solve (timing probes): active, not converged, not stalled
__global__ void k_seg_reopen(SegScalars* seg, int S)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < S) { seg[s].converged = 0; seg[s].stalled = 0; seg[s].active = 1; }
}

// Convergence / cap test shared by the first kernel of every iteration.  Returns true when the
// tissue stops (all tiles of the tissue take the same decision).
__device__ inline bool check_stop(const TileArgs& ta, int t, int sg, int slot_rr, int maxit, double* rr_out = nullptr)
{
    const int sl[1] = {slot_rr};
    double rr[1];
    seg_sums<1>(ta, sg, sl, rr);
    if (rr_out) *rr_out = rr[0];
    SegScalars& S = ta.seg[sg];
    const bool conv = rr[0] <= S.tol2;
    const bool cap = !conv && (S.iters >= maxit || !isfinite(rr[0]));
    if ((conv || cap) && first_tile(ta, t, sg) && threadIdx.x == 0) {
        S.rnorm2 = rr[0];
        if (conv) S.converged = 1;
        else S.stalled = 1;
    }
    return conv || cap;
}

// ---------------------------------------------------------------------------------------------
// BiCGSTAB
__global__ __launch_bounds__(WG) void k_bicg_init(TileArgs ta, const double* __restrict__ b, double* x, double* r, double* r0,
                                                  double* p, double* v)
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
    const int sl[3] = {SL_BB, SL_RR0, SL_RHO0};
    double vv[3] = {s, s, s};
    tile_write<3>(ta, t, sl, vv);
}

template <int P>
__global__ __launch_bounds__(WG) void k_bicg_p(TileArgs ta, const double* __restrict__ dinv, const double* __restrict__ r,
                                               double* __restrict__ p, const double* __restrict__ v, double* __restrict__ ph,
                                               int maxit)
{
    const int t = blockIdx.x;
    const int sg = ta.tile_seg[t];
    SegScalars& S = ta.seg[sg];
    if (seg_done(S)) return;
    double rr = 0.0;
    if (check_stop(ta, t, sg, SL_RR0 + P, maxit, &rr)) return;
    double beta = 0.0, omega = 0.0;
    if (S.iters > 0) {
        const int sl[5] = {SL_RHO0 + P, SL_RHO0 + (1 - P), SL_RV, SL_TS, SL_TT};
        double q[5];
        seg_sums<5>(ta, sg, sl, q);
        const double alpha_prev = q[1] / q[2];
        omega = q[3] / q[4];
        beta = (q[0] / q[1]) * (alpha_prev / omega);
        if (!isfinite(beta) || q[0] == 0.0 || omega == 0.0) {  // breakdown: x keeps residual rr
            if (first_tile(ta, t, sg) && threadIdx.x == 0) { S.rnorm2 = rr; S.stalled = 1; }
            return;
        }
    }
    const int d0 = 3 * ta.tile_r0[t], d1 = 3 * ta.tile_r1[t];
    for (int d = d0 + threadIdx.x; d < d1; d += WG) p[d] = r[d] + beta * (p[d] - omega * v[d]);
    __syncthreads();
    for (int d = d0 + threadIdx.x; d < d1; d += WG) ph[d] = dinv_row(dinv, p, d);
}

// y = M x with fused dot products; MODE 0: (w, y) -> SL_RV; MODE 1: (y, w) -> SL_TS, (y, y) -> SL_TT
template <typename VT, int MODE>
__global__ __launch_bounds__(WG) void k_bicg_spmv(TileArgs ta, const int32_t* __restrict__ row_ptr,
                                                  const int32_t* __restrict__ col, const VT* __restrict__ val, int nb,
                                                  const double* __restrict__ x, double* __restrict__ y,
                                                  const double* __restrict__ w)
{
    extern __shared__ double prod[];  // 3 doubles per block of the tile (tile_cap blocks)
    const int t = blockIdx.x;
    const int sg = ta.tile_seg[t];
    if (seg_done(ta.seg[sg])) return;
    const int r0 = ta.tile_r0[t], r1 = ta.tile_r1[t];
    const int b0 = row_ptr[r0], b1 = row_ptr[r1];
    tile_products<VT>(col, val, nb, b0, b1, XLOAD(x), prod);
    double s0 = 0.0, s1 = 0.0;
    for (int task = threadIdx.x; task < 3 * (r1 - r0); task += WG) {
        const int r = r0 + task / 3, comp = task % 3;
        const int d = 3 * r + comp;
        const double yv = row_sum(row_ptr, r, comp, b0, prod);
        y[d] = yv;
        s0 += w[d] * yv;
        if (MODE == 1) s1 += yv * yv;
    }
    if (MODE == 0) {
        const int sl[1] = {SL_RV};
        double vv[1] = {s0};
        tile_write<1>(ta, t, sl, vv);
    } else {
        const int sl[2] = {SL_TS, SL_TT};
        double vv[2] = {s0, s1};
        tile_write<2>(ta, t, sl, vv);
    }
}

template <int P>
__global__ __launch_bounds__(WG) void k_bicg_s(TileArgs ta, const double* __restrict__ dinv, const double* __restrict__ r,
                                               const double* __restrict__ v, double* __restrict__ s, double* __restrict__ sh)
{
    const int t = blockIdx.x;
    const int sg = ta.tile_seg[t];
    SegScalars& S = ta.seg[sg];
    if (seg_done(S)) return;
    const int sl[3] = {SL_RHO0 + P, SL_RV, SL_RR0 + P};
    double q[3];
    seg_sums<3>(ta, sg, sl, q);
    const double alpha = q[0] / q[1];
    if (!(q[1] != 0.0) || !isfinite(alpha)) {  // breakdown: x keeps residual q[2]
        if (first_tile(ta, t, sg) && threadIdx.x == 0) { S.rnorm2 = q[2]; S.stalled = 1; }
        return;
    }
    const int d0 = 3 * ta.tile_r0[t], d1 = 3 * ta.tile_r1[t];
    for (int d = d0 + threadIdx.x; d < d1; d += WG) s[d] = r[d] - alpha * v[d];
    __syncthreads();
    for (int d = d0 + threadIdx.x; d < d1; d += WG) sh[d] = dinv_row(dinv, s, d);
}

template <int P>
__global__ __launch_bounds__(WG) void k_bicg_x(TileArgs ta, const double* __restrict__ ph, const double* __restrict__ sh,
                                               const double* __restrict__ s, const double* __restrict__ tv,
                                               const double* __restrict__ r0, double* __restrict__ x, double* __restrict__ r)
{
    const int t = blockIdx.x;
    const int sg = ta.tile_seg[t];
    SegScalars& S = ta.seg[sg];
    if (seg_done(S)) return;
    const int sl[4] = {SL_RHO0 + P, SL_RV, SL_TS, SL_TT};
    double q[4];
    seg_sums<4>(ta, sg, sl, q);
    const double alpha = q[0] / q[1];
    const double omega = (q[3] > 0.0) ? q[2] / q[3] : 0.0;
    const int d0 = 3 * ta.tile_r0[t], d1 = 3 * ta.tile_r1[t];
    double a = 0.0, b = 0.0;
    for (int d = d0 + threadIdx.x; d < d1; d += WG) {
        x[d] = x[d] + alpha * ph[d] + omega * sh[d];
        const double rn = s[d] - omega * tv[d];
        r[d] = rn;
        a += r0[d] * rn;
        b += rn * rn;
    }
    const int so[2] = {SL_RHO0 + (1 - P), SL_RR0 + (1 - P)};
    double vv[2] = {a, b};
    tile_write<2>(ta, t, so, vv);
    if (first_tile(ta, t, sg) && threadIdx.x == 0) S.iters += 1;
}

// ---------------------------------------------------------------------------------------------
// CG
__global__ __launch_bounds__(WG) void k_cg_init(TileArgs ta, const double* __restrict__ dinv, const double* __restrict__ b,
                                                double* x, double* r, double* z, double* p_old)
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
        const double zz = dinv_row(dinv, r, d);
        z[d] = zz;
        a += r[d] * zz;
        bb += r[d] * r[d];
    }
    const int sl[3] = {SL_RZ0, SL_RR0, SL_BB};
    double vv[3] = {a, bb, bb};
    tile_write<3>(ta, t, sl, vv);
}

// p = z + beta p_old (own rows written to p_new, neighbours formed on the fly), q = M p, (p, q).
// Each lane loads its first PF matrix blocks, with z and p_old of their columns, before the tile
// re-sums the partials for the convergence test and beta (one reduction for the three slots).
template <typename VT, int P, int WGT>
__global__ __launch_bounds__(WGT) void k_cg_q(TileArgs ta, const int32_t* __restrict__ row_ptr, const int32_t* __restrict__ col,
                                              const VT* __restrict__ val, int nb, const double* __restrict__ z,
                                              const double* __restrict__ p_old, double* __restrict__ p_new,
                                              double* __restrict__ q, int maxit)
{
    constexpr int PF = WGT >= 1024 ? 1 : 2;
    extern __shared__ double prod[];  // 3 doubles per block of the tile (tile_cap blocks)
    const int t = xcd_tile(blockIdx.x, ta.T, ta.nxcd);
    const TileInfo ti = ta.info[t];  // tissue, rows, blocks and the tissue's tiles: one scalar load
    const int sg = ti.seg, r0 = ti.r0, r1 = ti.r1, b0 = ti.b0, b1 = ti.b1;
    SegScalars& S = ta.seg[sg];
    if (seg_done(S)) return;
    double pv[PF][9], pz[PF][3], pp[PF][3];
#pragma unroll
    for (int k = 0; k < PF; ++k) {
        const int j = b0 + (int)threadIdx.x + k * WGT;
        const int c = j < b1 ? col[j] : 0;
        const int jj = j < b1 ? j : b0;
#pragma unroll
        for (int e = 0; e < 9; ++e) pv[k][e] = val[(size_t)e * nb + jj];
#pragma unroll
        for (int e = 0; e < 3; ++e) { pz[k][e] = z[3 * c + e]; pp[k][e] = p_old[3 * c + e]; }
    }
    const int sl[3] = {SL_RR0 + P, SL_RZ0 + P, SL_RZ0 + (1 - P)};
    double sc[3];
    seg_sums_range<3, WGT>(ta, ti.st0, ti.st1, sl, sc);
    const int it = S.iters;
    const double best = S.best_rr_h[P];
    const int best_it = S.best_it_h[P];
    const bool lead = t == ti.st0 && threadIdx.x == 0;
    {
        // the convergence / cap test of check_stop, with the stagnation rule of the single-kernel
        // CG: 256 iterations without a halving of ||r|| stop the solve (as converged when the best
        // residual came within 10x of the tolerance, else as stalled: BiCGSTAB takes over)
        const bool stagnant = it - best_it > 256;
        const bool conv = sc[0] <= S.tol2 || (stagnant && best <= 100.0 * S.tol2);
        const bool cap = !conv && (it >= maxit || !isfinite(sc[0]) || stagnant);
        if (conv || cap) {
            if (lead) {
                S.rnorm2 = sc[0];
                if (conv) S.converged = 1;
                else S.stalled = 1;
            }
            return;
        }
    }
    if (lead) {
        const bool better = sc[0] < 0.5 * best;
        S.best_rr_h[1 - P] = better ? sc[0] : best;
        S.best_it_h[1 - P] = better ? it : best_it;
    }
    const double beta = S.iters > 0 ? sc[1] / sc[2] : 0.0;
#pragma unroll
    for (int k = 0; k < PF; ++k) {
        const int j = b0 + (int)threadIdx.x + k * WGT;
        if (j < b1) {
            const double x0 = pz[k][0] + beta * pp[k][0], x1 = pz[k][1] + beta * pp[k][1], x2 = pz[k][2] + beta * pp[k][2];
            double* pr = prod + 3 * (j - b0);
            pr[0] = pv[k][0] * x0 + pv[k][1] * x1 + pv[k][2] * x2;
            pr[1] = pv[k][3] * x0 + pv[k][4] * x1 + pv[k][5] * x2;
            pr[2] = pv[k][6] * x0 + pv[k][7] * x1 + pv[k][8] * x2;
        }
    }
    for (int j = b0 + (int)threadIdx.x + PF * WGT; j < b1; j += WGT) {
        const int c = col[j];
        const double x0 = z[3 * c] + beta * p_old[3 * c];
        const double x1 = z[3 * c + 1] + beta * p_old[3 * c + 1];
        const double x2 = z[3 * c + 2] + beta * p_old[3 * c + 2];
        const double v0 = val[j], v1 = val[(size_t)nb + j], v2 = val[2 * (size_t)nb + j];
        const double v3 = val[3 * (size_t)nb + j], v4 = val[4 * (size_t)nb + j], v5 = val[5 * (size_t)nb + j];
        const double v6 = val[6 * (size_t)nb + j], v7 = val[7 * (size_t)nb + j], v8 = val[8 * (size_t)nb + j];
        double* pr = prod + 3 * (j - b0);
        pr[0] = v0 * x0 + v1 * x1 + v2 * x2;
        pr[1] = v3 * x0 + v4 * x1 + v5 * x2;
        pr[2] = v6 * x0 + v7 * x1 + v8 * x2;
    }
    __syncthreads();
    double s0 = 0.0;
    for (int task = threadIdx.x; task < 3 * (r1 - r0); task += WGT) {
        const int r = r0 + task / 3, comp = task % 3;
        const int d = 3 * r + comp;
        const double qv = row_sum(row_ptr, r, comp, b0, prod);
        const double pv_ = z[d] + beta * p_old[d];
        q[d] = qv;
        p_new[d] = pv_;
        s0 += pv_ * qv;
    }
    const int so[1] = {SL_PQ};
    double vv[1] = {s0};
    tile_write<1, WGT>(ta, t, so, vv);
}

template <int P>
__global__ __launch_bounds__(WG) void k_cg_x(TileArgs ta, const double* __restrict__ dinv, const double* __restrict__ p,
                                             const double* __restrict__ q, double* __restrict__ x, double* __restrict__ r,
                                             double* __restrict__ z)
{
    const int t = blockIdx.x;
    const int sg = ta.tile_seg[t];
    SegScalars& S = ta.seg[sg];
    if (seg_done(S)) return;
    const int sl[3] = {SL_RZ0 + P, SL_PQ, SL_RR0 + P};
    double sc[3];
    seg_sums<3>(ta, sg, sl, sc);
    if (!(sc[1] > 0.0) || !isfinite(sc[1])) {  // not positive definite: x keeps residual sc[2]
        if (first_tile(ta, t, sg) && threadIdx.x == 0) { S.rnorm2 = sc[2]; S.stalled = 1; }
        return;
    }
    const double alpha = sc[0] / sc[1];
    const int d0 = 3 * ta.tile_r0[t], d1 = 3 * ta.tile_r1[t];
    for (int d = d0 + threadIdx.x; d < d1; d += WG) {
        x[d] = x[d] + alpha * p[d];
        r[d] = r[d] - alpha * q[d];
    }
    __syncthreads();
    double a = 0.0, b = 0.0;
    for (int d = d0 + threadIdx.x; d < d1; d += WG) {
        const double zz = dinv_row(dinv, r, d);
        z[d] = zz;
        a += r[d] * zz;
        b += r[d] * r[d];
    }
    const int so[2] = {SL_RZ0 + (1 - P), SL_RR0 + (1 - P)};
    double vv[2] = {a, b};
    tile_write<2>(ta, t, so, vv);
    if (first_tile(ta, t, sg) && threadIdx.x == 0) S.iters += 1;
}

// ---------------------------------------------------------------------------------------------
// Two-level preconditioner (one tissue): z = Dc^-1 r + P Ac^-1 P^T r, with Dc the column blocks of
// M (the 3L x 3L block of each column of L nodes joined by prisms' vertical edges; single nodes
// when the columns are not uniform), P the rigid-body modes of aggregates of whole columns
// (orthonormalised per aggregate) and Ac = P^T M P.  The PCG that applies it runs in aggregate
// order, two launches per iteration (k_tp_q, k_tp_z below).
// An aggregate lists its nodes column by column, so global column g holds agg_nodes[gL .. gL+L-1].
constexpr int TL_MAX_LEVELS = 4;  // nodes per column at most (12 x 12 column blocks)

// the coarse space as the Galerkin product k_tl_ac reads it
struct TLArgs {
    const double* pcoef;       // [N][3][6] rows of P by node (columns past the aggregate's rank are zero)
    const int32_t* coff;       // [A] first coarse unknown of every aggregate
    const int32_t* krank;      // [A] coarse unknowns of every aggregate
    const int32_t* agg_seg;    // [A] tissue of every aggregate
    const int32_t* seg_c0;     // [S+1] first coarse unknown of every tissue
    const int32_t* seg_ldc;    // [S] row stride of the tissue's block of Ac
    const int64_t* seg_acoff;  // [S] offset of that block
    const uint8_t* refresh;    // [S] 1: the tissue's block is rebuilt (others are left alone)
};

// Inverse of every column block of M, as a symmetric FP32 copy: 64 / 3L columns per wave, one lane
// per block row, Gauss-Jordan elimination without pivoting (the blocks of a positive definite M
// are positive definite; the pivot rows come by lane shuffles).  A column with a pivot that is not
// positive and finite gets its nodes' 3 x 3 inverses (dinv) instead.
template <int L, typename OT>
__global__ __launch_bounds__(WG) void k_tl_dcol(const int32_t* __restrict__ col_blk, const int32_t* __restrict__ agg_nodes,
                                               const double* __restrict__ val, int nb, const double* __restrict__ dinv,
                                               int ncol, OT* __restrict__ dcinv)
{
    constexpr int D = 3 * L, CPW = 64 / D;
    const int lane = threadIdx.x % 64, wave = (int)(blockIdx.x * (WG / 64) + threadIdx.x / 64);
    const int slot = lane / D, li = lane % D, base = slot * D;
    const int g = wave * CPW + slot;
    const bool own = slot < CPW && g < ncol;
    const int gg = own ? g : 0;
    const int i = li / 3, ci = li % 3;
    double a[D], v[D];
#pragma unroll
    for (int j = 0; j < L; ++j) {
        const int blk = col_blk[((size_t)gg * L + i) * L + j];
        const int bb = blk >= 0 ? blk : 0;
#pragma unroll
        for (int cj = 0; cj < 3; ++cj) {
            const double m = val[(size_t)(3 * ci + cj) * nb + bb];
            a[3 * j + cj] = blk >= 0 ? m : 0.0;
        }
    }
#pragma unroll
    for (int j = 0; j < D; ++j) v[j] = j == li ? 1.0 : 0.0;
    bool ok = true;
#pragma unroll
    for (int k = 0; k < D; ++k) {
        double pa[D], pv[D];
#pragma unroll
        for (int j = 0; j < D; ++j) {
            pa[j] = __shfl(a[j], base + k, 64);
            pv[j] = __shfl(v[j], base + k, 64);
        }
        const double piv = pa[k];
        ok = ok && piv > 0.0 && isfinite(piv);
        const double f = li == k ? 0.0 : a[k] / piv;
        const double sc = li == k ? 1.0 / piv : 1.0;
#pragma unroll
        for (int j = 0; j < D; ++j) {
            a[j] = sc * (a[j] - f * pa[j]);
            v[j] = sc * (v[j] - f * pv[j]);
        }
    }
    if (!own) return;
    OT* out = dcinv + (size_t)g * D * D;
    if (ok) {
#pragma unroll
        for (int j = 0; j < D; ++j)
            if (j >= li) {
                const OT e = (OT)v[j];
                out[li * D + j] = e;
                out[j * D + li] = e;
            }
    } else {
        const int node = agg_nodes[(size_t)g * L + i];
#pragma unroll
        for (int j = 0; j < D; ++j) out[li * D + j] = j / 3 == i ? (OT)dinv[9 * (size_t)node + 3 * ci + j % 3] : (OT)0;
    }
}

// ---- the two-level PCG in aggregate order: two launches per iteration --------------------------
// The solve runs in the order of the aggregates: position r holds node agg_nodes[r], aggregate a holds
// positions [agg_ptr[a], agg_ptr[a+1]) (column by column), and the Jacobian is a row-permuted copy
// (val_p, prow_ptr, pcol: the blocks of position r's row, columns as positions).  One workgroup per
// aggregate then owns a contiguous run of rows of every vector, and the restriction P^T of its rows
// is a sum inside the workgroup.  An iteration of parity P is
//   k_tp_q<P>: beta from (r, z); p = z + beta p_old (own rows written, neighbours' formed on the fly);
//              q = M p; the partials of (p, q) and of P^T q (the coarse unknowns of the aggregate);
//   k_tp_z<P>: alpha = (r, z) / (p, q); r_c = rcx[P] - alpha P^T q, with rcx[P] = P^T r of the
//              previous iteration (exact, summed from r by the aggregates), so the coarse right-hand
//              side carries one step of rounding, never an accumulated drift; z_c = Ac^-1 r_c on the
//              aggregate's coarse rows; x += alpha p, r -= alpha q, z = Dc^-1 r + P z_c on its rows;
//              the partials of (r, z) and (r, r) (from the z applied), and rcx[1-P] = P^T r.
// Every sum runs in a fixed order.  Partials are indexed by aggregate (TileArgs with T = n_agg).
constexpr int TP_CAP = 2048;  // blocks per pass of a workgroup (48 KB of LDS products)

// Everything a workgroup needs before its first data load, in one 80-byte record read by scalar
// loads that issue together (the aggregate's tissue, offsets and coarse block would otherwise be
// loads that wait for one another).  Workgroups take the aggregates largest first: the dispatcher
// deals workgroups in order, so the small ones fill in behind the large ones.
struct alignas(16) TPRec {
    int32_t a, sg, st0, st1;  // aggregate, tissue, the tissue's aggregates [st0, st1) (partial slots)
    int32_t r0, r1, b0, b1;   // the aggregate's first pass: rows [r0, r1), permuted blocks [b0, b1)
    int32_t ps0, ps1, n0, n1; // its passes [ps0, ps1) and positions [n0, n1)
    int32_t c0, kr, pad0, pad1;  // its first coarse unknown and rank
    int4 sc;                  // the tissue's first coarse unknown, count, ldA, Ainv block offset (pairs)
};
static_assert(sizeof(TPRec) == 80, "TPRec layout");

struct TPArgs {
    const int32_t* agg_nodes;  // [N] node of every position
    const int32_t* prow_ptr;   // [N+1] permuted blocks of every position's row
    const int32_t* pcol;       // [nb] column position of every permuted block
    const int2* pass_rows;     // [passes] rows [x, y) of every pass (<= WG / 3 rows, <= TP_CAP blocks)
    const TPRec* rec;          // [A] the record of every workgroup
    const double* pcoef;       // [N][3][6] rows of P by position
    const float* dcinv;        // [ncol][3L][3L] inverse column blocks (FP32, symmetric)
    const double* dcinv64;     //   the same in FP64 (FP64 preconditioner storage)
    const __half2* Ainv;       // per tissue [nc_s][ldA_s / 2] Ac_s^-1 / scale_s in FP16 pairs
    const double* Ainv64;      //   per tissue [nc_s][ldA_s] Ac_s^-1 in FP64 (FP64 preconditioner storage)
    const int64_t* seg_aoff64; //   [S] offset of the tissue's FP64 block
    const double* ascale;      // [S][2] scale and its reciprocal
    double* ptq;               // [nc] P^T q of the current iteration
    double* rcx;               // [2][nc] P^T r (ping-pong by parity)
    int nc, n_agg;
    unsigned* bar;             // the persistent solve's barrier words (null without it), cleared by
    int bar_words;             //   the first z of every solve (k_tp_z<P, true>)
    int c32;                   // ptq and rcx hold FP32 values (not under FP64 preconditioner storage):
                               // every workgroup of the persistent solve gathers both, and the
                               // coarse product takes its right-hand side in FP32 anyway
};

// Stores an entry of ptq / rcx in the storage TPArgs::c32 names
__device__ inline void coarse_put(const TPArgs& tp, double* base, size_t i, double v)
{
    if (tp.c32) reinterpret_cast<float*>(base)[i] = (float)v;
    else base[i] = v;
}
template <bool F64>
__device__ inline std::conditional_t<F64, double, float> dc_row(const TPArgs& tp, size_t i)
{
    if constexpr (F64) return tp.dcinv64[i];
    else return tp.dcinv[i];
}

// q = M p over the aggregate's passes; own rows see p_new = z + beta p_old (beta = 0: p = z), and
// the callback gets (position, component, q value, p value) for every own DoF (q = 0 without
// products).
template <typename VT, typename F>
__device__ inline void tp_spmv(const TPArgs& tp, const VT* __restrict__ val, int nb, const double* __restrict__ z,
                               const double* __restrict__ p_old, double beta, int ps0, int ps1, double* prod, F own,
                               bool products = true)
{
    for (int k = ps0; k < ps1; ++k) {
        const int2 rr = tp.pass_rows[k];
        const int b0 = tp.prow_ptr[rr.x], b1 = tp.prow_ptr[rr.y];
        for (int j = b0 + (int)threadIdx.x; products && j < b1; j += WG) {
            const int c = tp.pcol[j];
            const double x0 = z[3 * c] + beta * p_old[3 * c];
            const double x1 = z[3 * c + 1] + beta * p_old[3 * c + 1];
            const double x2 = z[3 * c + 2] + beta * p_old[3 * c + 2];
            const double v0 = val[j], v1 = val[(size_t)nb + j], v2 = val[2 * (size_t)nb + j];
            const double v3 = val[3 * (size_t)nb + j], v4 = val[4 * (size_t)nb + j], v5 = val[5 * (size_t)nb + j];
            const double v6 = val[6 * (size_t)nb + j], v7 = val[7 * (size_t)nb + j], v8 = val[8 * (size_t)nb + j];
            double* pr = prod + 3 * (j - b0);
            pr[0] = v0 * x0 + v1 * x1 + v2 * x2;
            pr[1] = v3 * x0 + v4 * x1 + v5 * x2;
            pr[2] = v6 * x0 + v7 * x1 + v8 * x2;
        }
        __syncthreads();
        const int task = threadIdx.x;  // at most one own DoF per lane: a pass holds <= WG / 3 rows
        if (task < 3 * (rr.y - rr.x)) {
            const int r = rr.x + task / 3, comp = task % 3;
            double qv = 0.0;
            const int j1 = tp.prow_ptr[r + 1];
            for (int j = tp.prow_ptr[r]; products && j < j1; ++j) qv += prod[3 * (j - b0) + comp];
            const int d = 3 * r + comp;
            own(r, comp, qv, products ? z[d] + beta * p_old[d] : 0.0);
        }
        __syncthreads();  // prod is refilled by the next pass
    }
}

// The tissue's partials of K slots, loaded before a kernel's data so that the wait for them does
// not wait for the data (loads return in order): at most TP_SU partials per lane and slot, the
// tissue's aggregates [st0, st1) at most TP_SU * WG (a tissue has at most max_nc / 6 aggregates),
// clamped addresses, masked in tp_partials_sum.
constexpr int TP_SU = 3;

template <int K>
__device__ inline void tp_partials_load(const TileArgs& ta, int st0, int st1, const int (&sl)[K],
                                        double (&v)[K][TP_SU])
{
#pragma unroll
    for (int u = 0; u < TP_SU; ++u) {
        const int k = st0 + (int)threadIdx.x + u * WG;
        const int kk = k < st1 ? k : st0;
#pragma unroll
        for (int q = 0; q < K; ++q) v[q][u] = ta.partial[(size_t)sl[q] * ta.T + kk];
    }
}

// Tissue totals from the loaded partials in the order of seg_sums_range (lane sums in index order,
// then a fixed wave and tree order), identical in every lane and every workgroup of the tissue.
template <int K, int SU = TP_SU>
__device__ inline void tp_partials_sum(const TileArgs& ta, int st0, int st1, const int (&sl)[K],
                                       const double (&v)[K][SU], double (&out)[K])
{
    constexpr int NW = WG / 64;
    __shared__ double red[K][NW];
    double s[K];
#pragma unroll
    for (int q = 0; q < K; ++q) s[q] = 0.0;
#pragma unroll
    for (int u = 0; u < SU; ++u)
        if (st0 + (int)threadIdx.x + u * WG < st1)
#pragma unroll
            for (int q = 0; q < K; ++q) s[q] += v[q][u];
    for (int k = st0 + (int)threadIdx.x + SU * WG; k < st1; k += WG)  // beyond SU * WG: not built
#pragma unroll
        for (int q = 0; q < K; ++q) s[q] += ta.partial[(size_t)sl[q] * ta.T + k];
    const int w = threadIdx.x / 64, l = threadIdx.x % 64;
#pragma unroll
    for (int q = 0; q < K; ++q) {
        s[q] = wave_sum(s[q]);
        if (l == 0) red[q][w] = s[q];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < K; ++q) {
        double a = 0.0;
#pragma unroll
        for (int k = 0; k < NW; ++k) a += red[q][k];
        out[q] = a;
    }
    __syncthreads();
}

// Workgroup sums of K values into partial slots of aggregate a (fixed wave and lane order).
template <int K>
__device__ inline void tp_reduce(double (&v)[K], double* out)
{
    constexpr int NW = WG / 64;
    __shared__ double red[K][NW];
    const int w = threadIdx.x / 64, l = threadIdx.x % 64;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const double t = wave_sum(v[k]);
        if (l == 0) red[k][w] = t;
    }
    __syncthreads();
    if ((int)threadIdx.x < K) {
        double t = 0.0;
#pragma unroll
        for (int k = 0; k < NW; ++k) t += red[threadIdx.x][k];
        out[threadIdx.x] = t;
    }
    __syncthreads();
}

// r = b - M x0 (x0 given, in positions) or b, x = x0 or 0, p_old = 0, (b, b), and rcx[1] = P^T r
template <typename VT>
__global__ __launch_bounds__(WG) void k_tp_init(TileArgs ta, TPArgs tp, const VT* __restrict__ val, int nb,
                                                const double* __restrict__ b_nat, const double* __restrict__ x0p,
                                                double* __restrict__ x, double* __restrict__ r,
                                                double* __restrict__ p_old)
{
    extern __shared__ double prod[];
    const TPRec R = tp.rec[blockIdx.x];
    const int a = R.a;
    const bool active = ta.seg[R.sg].active;
    double bb = 0.0, rc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    auto own = [&](int pos, int comp, double mx, double) {
        const int d = 3 * pos + comp;
        const double bv = active ? b_nat[3 * tp.agg_nodes[pos] + comp] : 0.0;
        const double rv = (x0p && active) ? bv - mx : bv;
        x[d] = (x0p && active) ? x0p[d] : 0.0;
        r[d] = rv;
        p_old[d] = 0.0;
        bb += bv * bv;
        const double* Pr = tp.pcoef + 18 * (size_t)pos + 6 * comp;
#pragma unroll
        for (int j = 0; j < 6; ++j) rc[j] += Pr[j] * rv;
    };
    // the product is M x0 (beta = 0); without x0 there is none
    tp_spmv<VT>(tp, val, nb, x0p, x0p, 0.0, R.ps0, R.ps1, prod, own, x0p != nullptr);
    double v[7] = {bb, rc[0], rc[1], rc[2], rc[3], rc[4], rc[5]};
    __shared__ double out[7];
    tp_reduce<7>(v, out);
    if (threadIdx.x == 0) ta.partial[(size_t)SL_BB * ta.T + a] = out[0];
    if ((int)threadIdx.x < R.kr) coarse_put(tp, tp.rcx, (size_t)tp.nc + R.c0 + threadIdx.x, out[1 + threadIdx.x]);
}

// beta, p = z + beta p_old, q = M p, (p, q) and P^T q; the convergence, cap and stagnation test.
// The tissue's partials are loaded first, then each lane's first TP_PF blocks of the first pass
// (columns, values, then z and p_old of those columns), so all of them are in flight while beta is
// formed.  TP_PF * WG blocks cover an aggregate of up to about 75 nodes of a 3-layer mesh; blocks
// past them are loaded after beta.
constexpr int TP_PF = 5;

template <typename VT, int P>
__global__ __launch_bounds__(WG) void k_tp_q(TileArgs ta, TPArgs tp, const VT* __restrict__ val, int nb,
                                             const double* __restrict__ z, const double* __restrict__ p_old,
                                             double* __restrict__ p_new, double* __restrict__ q, int maxit)
{
    extern __shared__ double prod[];
    const TPRec R = tp.rec[blockIdx.x];
    SegScalars& S = ta.seg[R.sg];
    if (seg_done(S)) return;  // after the solve's end: one load before the exit
    const int tid = threadIdx.x;
    const int sl[3] = {SL_RR0 + P, SL_RZ0 + P, SL_RZ0 + (1 - P)};
    double pl[3][TP_SU];
    tp_partials_load<3>(ta, R.st0, R.st1, sl, pl);
    int2 rr = int2{R.r0, R.r1};
    int b0 = R.b0, b1 = R.b1;
    double pv[TP_PF][9], pz[TP_PF][3], pp[TP_PF][3];
#pragma unroll
    for (int k = 0; k < TP_PF; ++k) {
        const int j = b0 + tid + k * WG;
        const int jj = j < b1 ? j : b0;  // clamped: always a valid address
        const int c = tp.pcol[jj];
#pragma unroll
        for (int e = 0; e < 9; ++e) pv[k][e] = val[(size_t)e * nb + jj];
#pragma unroll
        for (int e = 0; e < 3; ++e) { pz[k][e] = z[3 * c + e]; pp[k][e] = p_old[3 * c + e]; }
    }
    double sc[3];
    tp_partials_sum<3>(ta, R.st0, R.st1, sl, pl, sc);
    const int it = S.iters;
    const double best = S.best_rr_h[P];
    const int best_it = S.best_it_h[P];
    const bool lead = R.a == R.st0 && tid == 0;
    {
        const bool stagnant = it - best_it > 256;
        const bool conv = sc[0] <= S.tol2 || (stagnant && best <= 100.0 * S.tol2);
        const bool cap = !conv && (it >= maxit || !isfinite(sc[0]) || stagnant);
        if (conv || cap) {
            if (lead) {
                S.rnorm2 = sc[0];
                if (conv) S.converged = 1;
                else S.stalled = 1;
            }
            return;
        }
    }
    if (lead) {
        const bool better = sc[0] < 0.5 * best;
        S.best_rr_h[1 - P] = better ? sc[0] : best;
        S.best_it_h[1 - P] = better ? it : best_it;
    }
    const double beta = it > 0 ? sc[1] / sc[2] : 0.0;
    double pq = 0.0, pt[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (int ps = R.ps0; ps < R.ps1; ++ps) {
        if (ps > R.ps0) {
            rr = tp.pass_rows[ps];
            b0 = tp.prow_ptr[rr.x];
            b1 = tp.prow_ptr[rr.y];
        }
        const int jf = ps == R.ps0 ? b0 + TP_PF * WG : b0;  // the first pass's prefetched blocks come first
        if (ps == R.ps0) {
#pragma unroll
            for (int k = 0; k < TP_PF; ++k) {
                const int j = b0 + tid + k * WG;
                if (j < b1) {
                    const double x0 = pz[k][0] + beta * pp[k][0], x1 = pz[k][1] + beta * pp[k][1];
                    const double x2 = pz[k][2] + beta * pp[k][2];
                    double* pr = prod + 3 * (j - b0);
                    pr[0] = pv[k][0] * x0 + pv[k][1] * x1 + pv[k][2] * x2;
                    pr[1] = pv[k][3] * x0 + pv[k][4] * x1 + pv[k][5] * x2;
                    pr[2] = pv[k][6] * x0 + pv[k][7] * x1 + pv[k][8] * x2;
                }
            }
        }
        for (int j = jf + tid; j < b1; j += WG) {
            const int c = tp.pcol[j];
            const double x0 = z[3 * c] + beta * p_old[3 * c];
            const double x1 = z[3 * c + 1] + beta * p_old[3 * c + 1];
            const double x2 = z[3 * c + 2] + beta * p_old[3 * c + 2];
            const double v0 = val[j], v1 = val[(size_t)nb + j], v2 = val[2 * (size_t)nb + j];
            const double v3 = val[3 * (size_t)nb + j], v4 = val[4 * (size_t)nb + j], v5 = val[5 * (size_t)nb + j];
            const double v6 = val[6 * (size_t)nb + j], v7 = val[7 * (size_t)nb + j], v8 = val[8 * (size_t)nb + j];
            double* pr = prod + 3 * (j - b0);
            pr[0] = v0 * x0 + v1 * x1 + v2 * x2;
            pr[1] = v3 * x0 + v4 * x1 + v5 * x2;
            pr[2] = v6 * x0 + v7 * x1 + v8 * x2;
        }
        __syncthreads();
        if (tid < 3 * (rr.y - rr.x)) {  // at most one own DoF per lane: a pass holds <= WG / 3 rows
            const int r = rr.x + tid / 3, comp = tid % 3;
            const int d = 3 * r + comp;
            const double* Pr = tp.pcoef + 18 * (size_t)r + 6 * comp;
            double P6[6];
#pragma unroll
            for (int j = 0; j < 6; ++j) P6[j] = Pr[j];
            const double pvn = z[d] + beta * p_old[d];
            double qv = 0.0;
            const int j1 = tp.prow_ptr[r + 1];
            for (int j = tp.prow_ptr[r]; j < j1; ++j) qv += prod[3 * (j - b0) + comp];
            q[d] = qv;
            p_new[d] = pvn;
            pq += pvn * qv;
#pragma unroll
            for (int j = 0; j < 6; ++j) pt[j] += P6[j] * qv;
        }
        __syncthreads();  // prod is refilled by the next pass
    }
    double v[7] = {pq, pt[0], pt[1], pt[2], pt[3], pt[4], pt[5]};
    __shared__ double out[7];
    tp_reduce<7>(v, out);
    if (tid == 0) ta.partial[(size_t)SL_PQ * ta.T + R.a] = out[0];
    if (tid < R.kr) coarse_put(tp, tp.ptq, (size_t)R.c0 + tid, out[1 + tid]);
}

// alpha, the coarse right-hand side, z_c, then x, r, z on the aggregate's rows, (r, z), (r, r) and
// rcx[1-P] = P^T r.  INIT: r as given (the initial residual), r_c = rcx[1], no update.
// Every global load that does not need alpha is issued first, the tissue's partials leading: the
// aggregate's rows of Ac^-1 (16 bytes per lane and load), this lane's entries of rcx[P] and P^T q,
// and the first pass's r, x, p, q, Dc^-1 rows and P rows; alpha is then formed, r_c = rcx[P] -
// alpha P^T q staged in LDS (FP32: it only meets the FP16 coarse inverse), and z_c = Ac^-1 r_c.
// Eight products of a chunk of a coarse-inverse row (FP16) with the chunk (ca, cb) of r_c (FP32),
// summed by fused multiply-adds in FP32 in a fixed order; the row's sum over its chunks is FP64.
// (The FP16 rounding of the inverse bounds the accuracy of the coarse correction far above the FP32
// sum's.)
__device__ inline double coarse_dot8(const uint4& ch, const float4& ca, const float4& cb)
{
    const __half2* hp = reinterpret_cast<const __half2*>(&ch);
    const float2 f0 = __half22float2(hp[0]), f1 = __half22float2(hp[1]);
    const float2 f2 = __half22float2(hp[2]), f3 = __half22float2(hp[3]);
    float s = f0.x * ca.x;
    s = fmaf(f0.y, ca.y, s);
    s = fmaf(f1.x, ca.z, s);
    s = fmaf(f1.y, ca.w, s);
    s = fmaf(f2.x, cb.x, s);
    s = fmaf(f2.y, cb.y, s);
    s = fmaf(f3.x, cb.z, s);
    s = fmaf(f3.y, cb.w, s);
    return (double)s;
}

constexpr int TP_Q4 = 2;     // 16-byte loads (8 halves) per lane of every coarse row held in registers
constexpr int TP_NV = 16;    // entries of rcx and P^T q per lane (nc_s <= 4096)

template <int P, bool INIT, int L, bool F64 = false>
__global__ __launch_bounds__(WG) void k_tp_z(TileArgs ta, TPArgs tp, const double* __restrict__ p,
                                             const double* __restrict__ q, double* __restrict__ x,
                                             double* __restrict__ r, double* __restrict__ z)
{
    constexpr int D = 3 * L, CPP = WG / D;
    if (INIT && tp.bar && blockIdx.x == 0)  // the words the persistent launch behind it starts from
        for (int i = threadIdx.x; i < tp.bar_words; i += WG) tp.bar[i] = 0u;
    const TPRec R = tp.rec[blockIdx.x];
    const int a = R.a, sg = R.sg;
    SegScalars& S = ta.seg[sg];
    if (seg_done(S)) return;  // after the solve's end: one load before the exit
    const int tid = threadIdx.x;
    const int sl[3] = {SL_RZ0 + P, SL_PQ, SL_RR0 + P};
    double pl[3][TP_SU];
    if (!INIT) tp_partials_load<3>(ta, R.st0, R.st1, sl, pl);
    const int n0 = R.n0, ncols = (R.n1 - n0) / L;
    const int c0 = R.c0, kr = R.kr;
    extern __shared__ float tp_rcs[];  // r_c, ldA entries (zero past nc); FP64 storage: doubles
    double* const rcd = reinterpret_cast<double*>(tp_rcs);
    __shared__ double rl[WG];
    __shared__ double zcs[6];
    const int slot = tid / D, li = tid % D;
    const int nc = tp.nc, C0 = R.sc.x, ncs = R.sc.y, ldA = R.sc.z, ld2 = ldA / 2, lc0 = c0 - C0;
    const int last = max(ncs, 1) - 1;  // a tissue without coarse unknowns has one row of zeros
    uint4 pre[6][F64 ? 1 : TP_Q4];     // 8 halves each: pairs 4 (tid + WG m) .. + 3 of the row
    if constexpr (!F64) {
        const uint4* Ab = reinterpret_cast<const uint4*>(tp.Ainv + R.sc.w);  // ldA is a multiple of 8 halves
        const int lq = ld2 / 4;
#pragma unroll
        for (int k = 0; k < 6; ++k)
#pragma unroll
            for (int m = 0; m < TP_Q4; ++m)
                pre[k][m] = Ab[(size_t)min(lc0 + k, last) * lq + min(tid + WG * m, lq - 1)];
    }
    double vx[TP_NV], vq[TP_NV];
#pragma unroll
    for (int m = 0; m < TP_NV; ++m) {
        const int ic = min(tid + WG * m, last);  // clamped: always a valid address
        if constexpr (F64) {
            vx[m] = tp.rcx[(size_t)P * nc + C0 + ic];
            vq[m] = INIT ? 0.0 : tp.ptq[C0 + ic];
        } else {
            vx[m] = reinterpret_cast<const float*>(tp.rcx)[(size_t)P * nc + C0 + ic];
            vq[m] = INIT ? 0.0 : reinterpret_cast<const float*>(tp.ptq)[C0 + ic];
        }
    }
    bool own = slot < CPP && slot < ncols;
    int g = n0 / L + (own ? slot : 0);  // global column; its positions gL .. gL + L - 1
    int pos = g * L + li / 3;
    int d = 3 * pos + li % 3;
    double rv = r[d], xv = 0.0, pv = 0.0, qv = 0.0;
    if (!INIT) { xv = x[d]; pv = p[d]; qv = q[d]; }
    using DT = std::conditional_t<F64, double, float>;
    DT Dr[D];
#pragma unroll
    for (int j = 0; j < D; ++j) Dr[j] = dc_row<F64>(tp, ((size_t)g * D + li) * D + j);
    double Pr[6];
#pragma unroll
    for (int j = 0; j < 6; ++j) Pr[j] = tp.pcoef[18 * (size_t)pos + 6 * (li % 3) + j];
    double alpha = 0.0;
    if (!INIT) {
        double sc[3];
        tp_partials_sum<3>(ta, R.st0, R.st1, sl, pl, sc);
        // (p, q) > 0 needs M positive definite, (r, z) > 0 the preconditioner: either failing stops
        // the solve as stalled, x keeping residual sc[2]
        if (!(sc[1] > 0.0) || !isfinite(sc[1]) || !(sc[0] > 0.0) || !isfinite(sc[0])) {
            if (a == R.st0 && tid == 0) { S.rnorm2 = sc[2]; S.stalled = 1; }
            return;
        }
        alpha = sc[0] / sc[1];
    }
#pragma unroll
    for (int m = 0; m < TP_NV; ++m) {
        const int i = tid + WG * m;
        if (i < ldA) {
            if constexpr (F64) rcd[i] = i < ncs ? vx[m] - alpha * vq[m] : 0.0;
            else tp_rcs[i] = i < ncs ? (float)(vx[m] - alpha * vq[m]) : 0.f;
        }
    }
    __syncthreads();
    {
        double sv[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        if constexpr (F64) {  // the FP64 rows streamed here, 16 bytes per lane and load
#pragma unroll
            for (int k = 0; k < 6; ++k) {
                const double2* row = reinterpret_cast<const double2*>(tp.Ainv64 + tp.seg_aoff64[sg] +
                                                                     (size_t)min(lc0 + k, last) * ldA);
                for (int j = 2 * tid; j < ldA; j += 2 * WG) {
                    const double2 f = row[j / 2];
                    sv[k] += f.x * rcd[j] + f.y * rcd[j + 1];
                }
            }
        } else {
#pragma unroll
            for (int m = 0; m < TP_Q4; ++m) {
                const int j = 8 * (tid + WG * m);
                if (j < ldA) {  // chunks past ldA hold clamped copies: skipped
                    const float4 ca = *reinterpret_cast<const float4*>(tp_rcs + j);
                    const float4 cb = *reinterpret_cast<const float4*>(tp_rcs + j + 4);
#pragma unroll
                    for (int k = 0; k < 6; ++k) sv[k] += coarse_dot8(pre[k][m], ca, cb);
                }
            }
        }
        __shared__ double zc_red[6][WG / 64];
        const int w = tid / 64, l = tid % 64;
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            const double t = wave_sum(sv[k]);
            if (l == 0) zc_red[k][w] = t;
        }
        __syncthreads();
        if (tid < 6) {
            double t = 0.0;
#pragma unroll
            for (int k = 0; k < WG / 64; ++k) t += zc_red[tid][k];
            // P's columns past the rank are zero; the FP16 copy is scaled
            zcs[tid] = tid < kr ? (F64 ? t : t * tp.ascale[2 * sg]) : 0.0;
        }
        __syncthreads();
    }
    double rz = 0.0, rr = 0.0, rc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    const int sb = min(slot, CPP - 1) * D;  // lanes past the last whole column read a valid one
    for (int cb = 0; cb < ncols; cb += CPP) {
        if (cb > 0) {
            own = slot < CPP && cb + slot < ncols;
            g = n0 / L + (own ? cb + slot : 0);
            pos = g * L + li / 3;
            d = 3 * pos + li % 3;
            rv = r[d];
            if (!INIT) { xv = x[d]; pv = p[d]; qv = q[d]; }
#pragma unroll
            for (int j = 0; j < D; ++j) Dr[j] = dc_row<F64>(tp, ((size_t)g * D + li) * D + j);
#pragma unroll
            for (int j = 0; j < 6; ++j) Pr[j] = tp.pcoef[18 * (size_t)pos + 6 * (li % 3) + j];
        }
        if (!INIT) {
            xv = xv + alpha * pv;
            rv = rv - alpha * qv;
            if (own) { x[d] = xv; r[d] = rv; }
        }
        rl[tid] = rv;
        __syncthreads();
        double zi = 0.0;
#pragma unroll
        for (int j = 0; j < D; ++j) zi += (double)Dr[j] * rl[sb + j];
#pragma unroll
        for (int j = 0; j < 6; ++j) zi += Pr[j] * zcs[j];
        if (own) {
            z[d] = zi;
            rz += rv * zi;
            rr += rv * rv;
#pragma unroll
            for (int j = 0; j < 6; ++j) rc[j] += Pr[j] * rv;
        }
        __syncthreads();  // rl is refilled by the next pass
    }
    double v[8] = {rz, rr, rc[0], rc[1], rc[2], rc[3], rc[4], rc[5]};
    __shared__ double out[8];
    tp_reduce<8>(v, out);
    if (tid == 0) {
        ta.partial[(size_t)(SL_RZ0 + (1 - P)) * ta.T + a] = out[0];
        ta.partial[(size_t)(SL_RR0 + (1 - P)) * ta.T + a] = out[1];
        if (!INIT && a == R.st0) S.iters += 1;
    }
    if (tid < kr) coarse_put(tp, tp.rcx, (size_t)(1 - P) * nc + c0 + tid, out[2 + tid]);
}

static size_t tp_z_lds(int ldA, bool f64) { return (f64 ? sizeof(double) : sizeof(float)) * (size_t)ldA; }  // r_c

// ---- the two-level PCG as one persistent launch --------------------------------------------------
// One workgroup per aggregate, every workgroup resident at once, runs the whole
// solve: the halves of the iteration that k_tp_q and k_tp_z launch, each followed by a grid barrier.
// What a solve does not change and a workgroup can hold -- the aggregate's rows of the coarse inverse
// (LDS), its column-block inverses and rows of P (registers) -- is loaded once; the own rows of x, r,
// z, p and q stay in registers; the Jacobian blocks are read again every iteration (registers and LDS
// cannot hold them beside the rest at two workgroups per compute unit).  What other workgroups read (z and p of the nodes their blocks couple,
// the partial sums, P^T q and P^T r) goes through memory with agent-scope loads and stores: the L2 of
// an XCD is not coherent with the others', and those accesses bypass it.  The arithmetic is that of
// k_tp_q and k_tp_z, in the same order, so the iterates are the same bit for bit.
// Eligible: one tissue; every aggregate one pass (<= WG / 3 rows, <= TP_PF WG blocks, <= WG / 3L
// columns) with at most TPR_MAXU own and neighbouring nodes; TPR_NG to TPR_SU * WG aggregates; a
// grid of n_agg workgroups that the device holds at once.
constexpr int TPR_MAXU = 192;           // own + neighbouring nodes of an aggregate (LDS staging)
constexpr int TPR_NG = 8;               // barrier arrival groups: workgroup b arrives on word b % TPR_NG
constexpr int TPR_LINE = 32;            // the barrier words 128 bytes apart
constexpr unsigned TPR_SPIN = 1u << 22; // polls of one barrier wait before it gives up (abort)
constexpr int TPR_WORDS = (TPR_NG + 2) * TPR_LINE;  // group words, the global word, the abort word
constexpr int TPR_SU = 2;               // partials per lane and slot: at most TPR_SU * WG aggregates
constexpr int TPR_NV = 12;              // coarse values per lane: 6 TPR_SU WG coarse unknowns at most

// doubles of k_tp_run's LDS that the products of q = M p and then r_c share
__host__ __device__ inline int tp_run_share(int ldA, bool f64)
{
    const int rc = f64 ? ldA : (ldA + 1) / 2;
    return 3 * TP_PF * WG > rc ? 3 * TP_PF * WG : rc;
}

struct TPRunArgs {
    const int32_t* pslot;     // [nb] node slot of every permuted block's column: own position - n0,
                              //      or the aggregate's own node count + its index among the neighbours
    const int32_t* halo_ptr;  // [A+1] neighbouring nodes of every aggregate ...
    const int32_t* halo_pos;  //      ... as positions
    unsigned* bar;            // [TPR_WORDS] barrier words, zero at the launch (the solve's first z clears them)
    int maxit;
    unsigned spin;            // polls of one barrier wait before it gives up
};

__device__ inline double ld_ag(const double* p)
{
    return __hip_atomic_load(const_cast<double*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void st_ag(double* p, double v) { __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
__device__ inline float ld_ag32(const float* p)
{
    return __hip_atomic_load(const_cast<float*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void st_ag32(float* p, float v) { __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }

// Grid barrier number k (1-based), in two halves so that loads that do not depend on the barrier can
// be issued between them.  tpr_arrive: every store of every wave of the workgroup completes
// (s_waitcnt in each wave, then the workgroup barrier), the workgroup arrives on its group's word,
// and the group's last arrival adds one to the global word.
// tpr_wait: the workgroup waits for the global word to reach TPR_NG k.  A wait longer than spin
// polls sets the abort word; false when abort is set (the caller leaves its loop; every later wait
// sees it).
__device__ inline void tpr_arrive(unsigned* bar, unsigned k)
{
    // every wave's stores (agent-scope: coherent once complete) are complete before the workgroup
    // arrives; the workgroup barrier itself waits only for LDS
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned* glob = bar + TPR_NG * TPR_LINE;
        const int G = (int)gridDim.x, g = (int)blockIdx.x % TPR_NG;
        const unsigned Gg = (unsigned)((G - g + TPR_NG - 1) / TPR_NG);  // workgroups of group g (G >= TPR_NG)
        const unsigned old = __hip_atomic_fetch_add(bar + g * TPR_LINE, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (old + 1 == Gg * k) __hip_atomic_fetch_add(glob, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

__device__ inline bool tpr_wait(unsigned* bar, unsigned k, unsigned spin)
{
    __shared__ int pass;
    if (threadIdx.x == 0) {
        unsigned* glob = bar + TPR_NG * TPR_LINE;
        unsigned* abort = bar + (TPR_NG + 1) * TPR_LINE;
        int ok = 1;
        unsigned n = 0;
        while (__hip_atomic_load(glob, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < (unsigned)TPR_NG * k) {
            if (__hip_atomic_load(abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) { ok = 0; break; }
            if (++n > spin) {
                __hip_atomic_store(abort, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                ok = 0;
                break;
            }
        }
        pass = ok;
    }
    __syncthreads();
    return pass != 0;
}

// A uniform value the compiler must take as new in every iteration of k_tp_run's loop: addresses
// derived from it are formed where they are used instead of being hoisted out of the loop and held
// in vector registers for the whole launch (the empty asm only pins the value to scalar registers).
template <typename T>
__device__ inline T opq(T v)
{
    asm volatile("" : "+s"(v));
    return v;
}

// tp_partials_load through agent-scope loads (the partials change during the launch).  Lanes past the
// tissue's aggregates load a wrapped index rather than one clamped entry: agent-scope loads are not
// served by the L2, and hundreds of workgroups reading one line at once would queue at its channel.
template <int K>
__device__ inline void tpr_partials_load(const double* part, int T, int st0, int st1, const int (&sl)[K],
                                         double (&v)[K][TPR_SU])
{
#pragma unroll
    for (int u = 0; u < TPR_SU; ++u) {
        const int k = st0 + (int)threadIdx.x + u * WG;
        const int kk = k < st1 ? k : st0 + (k - st0) % max(1, st1 - st0);
#pragma unroll
        for (int q = 0; q < K; ++q) v[q][u] = ld_ag(part + (size_t)sl[q] * T + kk);
    }
}

template <typename VT, int L, bool F64>
__global__ __launch_bounds__(WG, 2) void k_tp_run(TileArgs ta, TPArgs tp, const VT* __restrict__ val, int nb,
                                                  TPRunArgs ra, double* __restrict__ x, double* __restrict__ r,
                                                  double* __restrict__ z, double* __restrict__ pb0,
                                                  double* __restrict__ pb1)
{
    constexpr int D = 3 * L, CPP = WG / D;
    // LDS: the products of q = M p (first half) and r_c (second half) share their room; after it the
    // aggregate's rows of the FP16 coarse inverse (6 rows of ldA halves), loaded once
    extern __shared__ double tpr_lds[];
    double* const prod = tpr_lds;
    float* const tp_rcs = reinterpret_cast<float*>(tpr_lds);
    double* const rcd = tpr_lds;
    __shared__ double zst[3 * TPR_MAXU], pst[3 * TPR_MAXU];  // z and p_old of the own and neighbouring nodes
    __shared__ double rl[WG];
    __shared__ double zcs[6];
    __shared__ double zc_red[6][WG / 64];
    __shared__ double outq[7], outz[8];
    const int tid = threadIdx.x;
    const TPRec R = tp.rec[blockIdx.x];
    const int a = R.a;
    SegScalars& S = ta.seg[R.sg];
    if (seg_done(S)) return;  // nothing to solve: every workgroup leaves (the flags are the same for all)
    const int n0 = R.n0, nown = R.n1 - R.n0, b0 = R.b0, b1 = R.b1;
    const int h0 = ra.halo_ptr[a], nh = ra.halo_ptr[a + 1] - h0;
    const bool mine = tid < 3 * nown;  // the lane's DoF 3 n0 + tid (row tid / 3 of the pass, or
                                       // column slot tid / D, entry tid % D of k_tp_z)
    const int d = 3 * n0 + (mine ? tid : 0);
    const int comp = tid % 3;
    // the blocks' column slots, the own row's block range, P's row, the column-block row (the
    // blocks' values are read every iteration: registers do not hold them beside the rest)
    int psl[TP_PF];
#pragma unroll
    for (int k = 0; k < TP_PF; ++k) {
        const int j = b0 + tid + k * WG;
        psl[k] = ra.pslot[j < b1 ? j : b0];
    }
    const int rrow = n0 + (mine ? tid / 3 : 0);
    const int jlo = tp.prow_ptr[rrow], jhi = tp.prow_ptr[rrow + 1];
    const size_t prw = 18 * (size_t)rrow + 6 * comp;  // the lane's row of P
    const int slot = tid / D, li = tid % D;
    const int g = n0 / L + (mine ? slot : 0);
    const size_t drw = ((size_t)g * D + li) * D;  // the lane's row of its column-block inverse
    const int nc = tp.nc, C0 = R.sc.x, ncs = R.sc.y, ldA = R.sc.z, ld2 = ldA / 2, lc0 = R.c0 - C0, kr = R.kr;
    const int last = max(ncs, 1) - 1;
    const int lq = ld2 / 4;  // 16-byte chunks (8 halves) of a coarse row
    uint4* const arow = reinterpret_cast<uint4*>(tpr_lds + tp_run_share(ldA, F64));  // [6][lq]
    if constexpr (!F64) {
        const uint4* Ab = reinterpret_cast<const uint4*>(tp.Ainv + R.sc.w);
        for (int i = tid; i < 6 * lq; i += WG) arow[i] = Ab[(size_t)min(lc0 + i / lq, last) * lq + i % lq];
    }
    // the neighbouring nodes' DoFs this lane stages (at most two)
    int hd[2];
#pragma unroll
    for (int u = 0; u < 2; ++u) {
        const int t = tid + u * WG;
        hd[u] = t < 3 * nh ? 3 * ra.halo_pos[h0 + t / 3] + t % 3 : -1;
    }
    const int hdum = 3 * n0 + tid % max(1, 3 * nown);
    // the own rows' state
    double xv = mine ? x[d] : 0.0, rv = mine ? r[d] : 0.0, zv = mine ? z[d] : 0.0, pold = mine ? pb0[d] : 0.0;
    double pnew = 0.0, qv = 0.0;
    int it = S.iters, best_it = S.best_it_h[0];
    double best = S.best_rr_h[0];
    const double tol2 = S.tol2;
    const bool lead = a == R.st0 && tid == 0;
    unsigned nbar = 0;
    bool aborted = false;
    // (r, z) of the last iteration (beta's denominator), carried from one iteration to the next;
    // a launch that starts after iteration 0 sums it once
    double rz_prev = 0.0;
    if (it > 0) {
        const int sl0[1] = {SL_RZ0 + (1 - (it & 1))};
        double p0[1][TPR_SU], s0[1];
        tpr_partials_load<1>(ta.partial, ta.T, R.st0, R.st1, sl0, p0);
        tp_partials_sum<1, TPR_SU>(ta, R.st0, R.st1, sl0, p0, s0);
        rz_prev = s0[0];
    }
    // loads that do not depend on other workgroups: the blocks and P's row of q = M p with the
    // partials at the start of that half; the column-block row and P's row of z between the halves
    // of the barrier before it
    VT pv[TP_PF][9];
    double P6[6];
    using DT = std::conditional_t<F64, double, float>;
    DT Dr[D];
    auto load_q = [&]() {
        const VT* const vb = opq(val);
        const int nbv = opq(nb);
#pragma unroll
        for (int k = 0; k < TP_PF; ++k) {
            const int j = b0 + tid + k * WG;
            const int jj = j < b1 ? j : b0;  // clamped: always a valid address
#pragma unroll
            for (int e = 0; e < 9; ++e) pv[k][e] = vb[(size_t)e * nbv + jj];
        }
        const double* pc = opq(tp.pcoef) + prw;
#pragma unroll
        for (int j = 0; j < 6; ++j) P6[j] = pc[j];
    };
    auto load_z = [&]() {
        TPArgs tq = tp;
        if constexpr (F64) tq.dcinv64 = opq(tp.dcinv64);
        else tq.dcinv = opq(tp.dcinv);
#pragma unroll
        for (int j = 0; j < D; ++j) Dr[j] = dc_row<F64>(tq, drw + j);
        const double* pc = opq(tp.pcoef) + prw;
#pragma unroll
        for (int j = 0; j < 6; ++j) P6[j] = pc[j];
    };
    while (true) {
        const int P = it & 1;
        double* const part = opq(ta.partial);
        double* const zb = opq(z);
        const double* pob = opq(P ? pb1 : pb0);  // p of the last iteration
        double* pnb = opq(P ? pb0 : pb1);        // p of this one
        // ---- q = M p (k_tp_q) ----
        const int slq[2] = {SL_RR0 + P, SL_RZ0 + P};
        double pl[2][TPR_SU];
        tpr_partials_load<2>(part, ta.T, R.st0, R.st1, slq, pl);
        __builtin_amdgcn_sched_barrier(0);  // the partials' loads go out before the blocks' (they return in order)
        load_q();
        double hz[2], hp[2];
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            const int q = hd[u] >= 0 ? hd[u] : hdum;  // lanes without a neighbour load an own row (spread)
            hz[u] = ld_ag(zb + q);
            hp[u] = ld_ag(pob + q);
        }
        double sc[2];
        tp_partials_sum<2, TPR_SU>(ta, R.st0, R.st1, slq, pl, sc);
        if (mine) { zst[tid] = zv; pst[tid] = pold; }
#pragma unroll
        for (int u = 0; u < 2; ++u)
            if (hd[u] >= 0) { zst[3 * nown + tid + u * WG] = hz[u]; pst[3 * nown + tid + u * WG] = hp[u]; }
        {
            const bool stagnant = it - best_it > 256;
            const bool conv = sc[0] <= tol2 || (stagnant && best <= 100.0 * tol2);
            const bool cap = !conv && (it >= ra.maxit || !isfinite(sc[0]) || stagnant);
            if (conv || cap) {
                if (lead) {
                    S.rnorm2 = sc[0];
                    if (conv) S.converged = 1;
                    else S.stalled = 1;
                }
                break;
            }
        }
        if (sc[0] < 0.5 * best) { best = sc[0]; best_it = it; }
        const double beta = it > 0 ? sc[1] / rz_prev : 0.0;
        // (r, r) and (r, z) of this iteration: the second half needs them again and the next
        // iteration's beta (r, z) (the same sums of the same partials); only (p, q) is new there
        const double rr_it = sc[0], rz_it = sc[1];
        rz_prev = rz_it;
        __syncthreads();  // zst, pst staged
#pragma unroll
        for (int k = 0; k < TP_PF; ++k) {
            const int j = b0 + tid + k * WG;
            if (j < b1) {
                const int c = 3 * psl[k];
                const double x0 = zst[c] + beta * pst[c], x1 = zst[c + 1] + beta * pst[c + 1];
                const double x2 = zst[c + 2] + beta * pst[c + 2];
                double* pr = prod + 3 * (j - b0);
                pr[0] = pv[k][0] * x0 + pv[k][1] * x1 + pv[k][2] * x2;
                pr[1] = pv[k][3] * x0 + pv[k][4] * x1 + pv[k][5] * x2;
                pr[2] = pv[k][6] * x0 + pv[k][7] * x1 + pv[k][8] * x2;
            }
        }
        __syncthreads();
        double pq = 0.0, pt[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        if (mine) {
            pnew = zv + beta * pold;
            double q = 0.0;
            for (int j = jlo; j < jhi; ++j) q += prod[3 * (j - b0) + comp];
            qv = q;
            pq += pnew * q;
#pragma unroll
            for (int j = 0; j < 6; ++j) pt[j] += P6[j] * q;
            st_ag(pnb + d, pnew);
        }
        {
            double v[7] = {pq, pt[0], pt[1], pt[2], pt[3], pt[4], pt[5]};
            tp_reduce<7>(v, outq);
            if (tid == 0) st_ag(part + (size_t)SL_PQ * ta.T + a, outq[0]);
            if (tid < kr) {
                if constexpr (F64) st_ag(opq(tp.ptq) + R.c0 + tid, outq[1 + tid]);
                else st_ag32(reinterpret_cast<float*>(opq(tp.ptq)) + R.c0 + tid, (float)outq[1 + tid]);
            }
        }
        tpr_arrive(ra.bar, ++nbar);
        load_z();
        // P^T r of the last iteration (written before the last barrier) is loaded while this one
        // completes; P^T q of this iteration after it
        double vx[TPR_NV];
        if constexpr (F64) {
            const double* rcp = opq(tp.rcx) + (size_t)P * nc + C0;
#pragma unroll
            for (int m = 0; m < TPR_NV; ++m) vx[m] = ld_ag(rcp + (tid + WG * m) % (last + 1));  // wrapped, not clamped
        } else {
            const float* rcp = reinterpret_cast<const float*>(opq(tp.rcx)) + (size_t)P * nc + C0;
#pragma unroll
            for (int m = 0; m < TPR_NV; ++m) vx[m] = (double)ld_ag32(rcp + (tid + WG * m) % (last + 1));
        }
        if (!tpr_wait(ra.bar, nbar, ra.spin)) { aborted = true; break; }
        // ---- alpha, the coarse correction, x, r, z (k_tp_z) ----
        const int slz[1] = {SL_PQ};
        double plz[1][TPR_SU];
        tpr_partials_load<1>(opq(ta.partial), ta.T, R.st0, R.st1, slz, plz);
        __builtin_amdgcn_sched_barrier(0);
        double vq[TPR_NV];
        if constexpr (F64) {
            const double* ptq = opq(tp.ptq) + C0;
#pragma unroll
            for (int m = 0; m < TPR_NV; ++m) vq[m] = ld_ag(ptq + (tid + WG * m) % (last + 1));
        } else {
            const float* ptq = reinterpret_cast<const float*>(opq(tp.ptq)) + C0;
#pragma unroll
            for (int m = 0; m < TPR_NV; ++m) vq[m] = (double)ld_ag32(ptq + (tid + WG * m) % (last + 1));
        }
        double pqs[1];
        tp_partials_sum<1, TPR_SU>(ta, R.st0, R.st1, slz, plz, pqs);
        if (!(pqs[0] > 0.0) || !isfinite(pqs[0]) || !(rz_it > 0.0) || !isfinite(rz_it)) {
            if (lead) { S.rnorm2 = rr_it; S.stalled = 1; }
            break;
        }
        const double alpha = rz_it / pqs[0];
#pragma unroll
        for (int m = 0; m < TPR_NV; ++m) {
            const int i = tid + WG * m;
            if (i < ldA) {
                if constexpr (F64) rcd[i] = i < ncs ? vx[m] - alpha * vq[m] : 0.0;
                else tp_rcs[i] = i < ncs ? (float)(vx[m] - alpha * vq[m]) : 0.f;
            }
        }
        __syncthreads();
        {
            double sv[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
            if constexpr (F64) {
#pragma unroll
                for (int k = 0; k < 6; ++k) {
                    const double2* row = reinterpret_cast<const double2*>(tp.Ainv64 + tp.seg_aoff64[R.sg] +
                                                                         (size_t)min(lc0 + k, last) * ldA);
                    for (int j = 2 * tid; j < ldA; j += 2 * WG) {
                        const double2 f = row[j / 2];
                        sv[k] += f.x * rcd[j] + f.y * rcd[j + 1];
                    }
                }
            } else {
#pragma unroll
                for (int m = 0; m < TP_Q4; ++m) {
                    const int j = 8 * (tid + WG * m);
                    if (j < ldA) {
                        const float4 ca = *reinterpret_cast<const float4*>(tp_rcs + j);
                        const float4 cb = *reinterpret_cast<const float4*>(tp_rcs + j + 4);
#pragma unroll
                        for (int k = 0; k < 6; ++k) sv[k] += coarse_dot8(arow[k * lq + j / 8], ca, cb);
                    }
                }
            }
            const int w = tid / 64, l = tid % 64;
#pragma unroll
            for (int k = 0; k < 6; ++k) {
                const double t = wave_sum(sv[k]);
                if (l == 0) zc_red[k][w] = t;
            }
            __syncthreads();
            if (tid < 6) {
                double t = 0.0;
#pragma unroll
                for (int k = 0; k < WG / 64; ++k) t += zc_red[tid][k];
                zcs[tid] = tid < kr ? (F64 ? t : t * tp.ascale[2 * R.sg]) : 0.0;
            }
            __syncthreads();
        }
        double rz = 0.0, rr = 0.0, rc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        if (mine) {
            xv = xv + alpha * pnew;
            rv = rv - alpha * qv;
        }
        rl[tid] = rv;
        __syncthreads();
        {
            const int sb = min(slot, CPP - 1) * D;
            double zi = 0.0;
#pragma unroll
            for (int j = 0; j < D; ++j) zi += (double)Dr[j] * rl[sb + j];
#pragma unroll
            for (int j = 0; j < 6; ++j) zi += P6[j] * zcs[j];
            if (mine) {
                zv = zi;
                rz += rv * zi;
                rr += rv * rv;
#pragma unroll
                for (int j = 0; j < 6; ++j) rc[j] += P6[j] * rv;
                st_ag(opq(z) + d, zi);
            }
        }
        {
            double v[8] = {rz, rr, rc[0], rc[1], rc[2], rc[3], rc[4], rc[5]};
            tp_reduce<8>(v, outz);
            double* const pz = opq(ta.partial);
            if (tid == 0) {
                st_ag(pz + (size_t)(SL_RZ0 + (1 - P)) * ta.T + a, outz[0]);
                st_ag(pz + (size_t)(SL_RR0 + (1 - P)) * ta.T + a, outz[1]);
            }
            if (tid < kr) {
                if constexpr (F64) st_ag(opq(tp.rcx) + (size_t)(1 - P) * nc + R.c0 + tid, outz[2 + tid]);
                else st_ag32(reinterpret_cast<float*>(opq(tp.rcx)) + (size_t)(1 - P) * nc + R.c0 + tid, (float)outz[2 + tid]);
            }
        }
        pold = pnew;
        it += 1;
        tpr_arrive(ra.bar, ++nbar);
        if (!tpr_wait(ra.bar, nbar, ra.spin)) { aborted = true; break; }
    }
    if (mine) { x[d] = xv; r[d] = rv; }
    if (lead) {
        S.iters = it;
        if (aborted) { S.rnorm2 = INFINITY; S.stalled = 1; }  // a barrier that gave up: the host falls back
    }
}

static size_t tp_run_lds(int ldA, bool f64)
{
    return sizeof(double) * tp_run_share(ldA, f64) + (f64 ? 0 : (size_t)6 * ldA * sizeof(__half));
}

// the solve's vectors between node order and aggregate order
__global__ __launch_bounds__(256) void k_tp_gather(const int32_t* __restrict__ agg_nodes, int n, const double* __restrict__ src,
                                                   double* __restrict__ dst)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < 3 * n) dst[k] = src[3 * agg_nodes[k / 3] + k % 3];
}
__global__ __launch_bounds__(256) void k_tp_scatter(const int32_t* __restrict__ agg_nodes, int n,
                                                    const double* __restrict__ src, double* __restrict__ dst)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < 3 * n) dst[3 * agg_nodes[k / 3] + k % 3] = src[k];
}
// the Jacobian's values in aggregate order: block k of the permuted matrix is block pblk[k]
template <typename VT>
__global__ __launch_bounds__(256) void k_tp_permute(const int32_t* __restrict__ pblk, int nb, const VT* __restrict__ val,
                                                    VT* __restrict__ val_p)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nb) return;
    const int b = pblk[k];
#pragma unroll
    for (int e = 0; e < 9; ++e) val_p[(size_t)e * nb + k] = val[(size_t)e * nb + b];
}

// Coarse operator Ac = P^T M P: one workgroup per pair of aggregates (a, b); every thread sums
// the 6x6 products P_A^T M_AB P_B of a strided subset of the pair's blocks, then a fixed tree.
__global__ __launch_bounds__(WG) void k_tl_ac(TLArgs tl, const int32_t* __restrict__ pair_a,
                                              const int32_t* __restrict__ pair_b, const int32_t* __restrict__ pair_ptr,
                                              const int32_t* __restrict__ pair_blk, const int32_t* __restrict__ blk_row,
                                              const int32_t* __restrict__ col, const double* __restrict__ val, int nb,
                                              double* __restrict__ Ac)
{
    const int pr = blockIdx.x;
    const int a = pair_a[pr], b = pair_b[pr];
    const int sg = tl.agg_seg[a];  // a and b are aggregates of one tissue
    if (!tl.refresh[sg]) return;
    double acc[36];
#pragma unroll
    for (int q = 0; q < 36; ++q) acc[q] = 0.0;
    for (int k = pair_ptr[pr] + threadIdx.x; k < pair_ptr[pr + 1]; k += WG) {
        const int blk = pair_blk[k];
        const double* PA = tl.pcoef + 18 * (size_t)blk_row[blk];
        const double* PB = tl.pcoef + 18 * (size_t)col[blk];
        double m[9];
#pragma unroll
        for (int q = 0; q < 9; ++q) m[q] = val[(size_t)q * nb + blk];
        double mp[18];  // M_AB P_B, 3 x 6
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 6; ++j)
                mp[6 * i + j] = m[3 * i] * PB[j] + m[3 * i + 1] * PB[6 + j] + m[3 * i + 2] * PB[12 + j];
#pragma unroll
        for (int pi = 0; pi < 6; ++pi)
#pragma unroll
            for (int j = 0; j < 6; ++j)
                acc[6 * pi + j] += PA[pi] * mp[j] + PA[6 + pi] * mp[6 + j] + PA[12 + pi] * mp[12 + j];
    }
    constexpr int NW = WG / 64;
    __shared__ double red[36][NW];
    const int w = threadIdx.x / 64, l = threadIdx.x % 64;
#pragma unroll
    for (int q = 0; q < 36; ++q) {
        const double t = wave_sum(acc[q]);
        if (l == 0) red[q][w] = t;
    }
    __syncthreads();
    if (threadIdx.x < 36) {
        const int pi = threadIdx.x / 6, ri = threadIdx.x % 6;
        if (pi < tl.krank[a] && ri < tl.krank[b]) {
            double t = 0.0;
#pragma unroll
            for (int k = 0; k < NW; ++k) t += red[threadIdx.x][k];
            const int c0 = tl.seg_c0[sg];
            Ac[tl.seg_acoff[sg] + (size_t)(tl.coff[a] - c0 + pi) * tl.seg_ldc[sg] + tl.coff[b] - c0 + ri] = t;
        }
    }
}

// ---- the coarse inverse: blocked Gauss-Jordan sweep ----------------------------------------------
// Ac (symmetric positive definite, n x n) is held row-major with leading dimension ldc = n rounded
// up to SW_B, the padding rows and columns holding the identity, so the padded matrix stays
// positive definite and its inverse's leading n x n block is Ac^-1.  Sweeping the pivot blocks
// K = 0, 1, ... in turn, each with D = A_KK^-1 (no pivoting: every A_KK met is a Schur complement of
// a positive definite matrix, hence positive definite),
//     A_IJ <- A_IJ - A_IK D A_KJ   (I, J != K),   A_IK <- A_IK D,   A_KJ <- D A_KJ,   A_KK <- -D,
// leaves -Ac^-1 in place.  The matrix stays symmetric, so only the tiles I >= J are kept (diagonal
// tiles whole).  One sweep step is two launches: k_sw_panel forms P_I = A_IK D and saves the old
// column O_I = A_IK for every I != K; k_sw_update then updates every lower tile from P and O (no tile
// reads another's), writes the new column K, and -- first workgroup of the grid -- inverts the
// next pivot tile right after updating it, so the serial 64-pivot inversion overlaps the update.
// FP64 throughout; every sum runs in a fixed order.
constexpr int SW_B = 64;        // tile size of the sweep
constexpr int SW_LD = SW_B + 1;  // LDS row stride of a staged tile (odd: columns hit distinct banks)
constexpr int SW_H = 33;         // LDS row stride of half a tile (32 inner columns + 1)

typedef double sw_v4d __attribute__((ext_vector_type(4)));

// 64 x 64 tile products on the matrix cores, one 32 x 32 quarter per wave (v_mfma_f64_16x16x4f64,
// 2 x 2 blocks of 16 x 16).  Operand layout of the instruction: lane l holds A[l % 16][l / 16] of
// the 16 x 4 left operand and B[l / 16][l % 16] of the 4 x 16 right one; output register r of lane
// l is D[l / 16 + 4 r][l % 16].  acc[bi][bj] is block (2 wr + bi, 2 wc + bj) of the wave's quarter.
struct SwQuarter {
    int l, wr, wc;
    __device__ SwQuarter() : l(threadIdx.x & 63), wr((threadIdx.x >> 6) >> 1), wc((threadIdx.x >> 6) & 1) {}
    __device__ int row(int bi, int r) const { return 16 * (2 * wr + bi) + (l >> 4) + 4 * r; }  // output row of reg r
    __device__ int col(int bj) const { return 16 * (2 * wc + bj) + (l & 15); }
};

// acc += sign * A B over 32 inner indices, A given by rows (a[row * lda + k]) and B by columns
// (b[col * ldb + k]), both staged in LDS
__device__ inline void sw_mfma_rows_cols(const SwQuarter& q, const double* a, int lda, const double* b, int ldb, double sign,
                                         sw_v4d (&acc)[2][2])
{
    const int r = q.l & 15, k0 = q.l >> 4;
#pragma unroll
    for (int k = 0; k < 32; k += 4) {
        const double a0 = sign * a[(16 * (2 * q.wr) + r) * lda + k + k0];
        const double a1 = sign * a[(16 * (2 * q.wr + 1) + r) * lda + k + k0];
        const double b0 = b[(16 * (2 * q.wc) + r) * ldb + k + k0];
        const double b1 = b[(16 * (2 * q.wc + 1) + r) * ldb + k + k0];
        acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
        acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
        acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
        acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
    }
}

// In-place Gauss-Jordan inverse of a 4 x 4 block in registers (no pivoting: a Schur complement of
// a positive definite matrix); ok goes false on a pivot that is not positive and finite
__device__ inline void sw_inv4(const double (&a)[16], double (&d)[16], bool& ok)
{
#pragma unroll
    for (int q = 0; q < 16; ++q) d[q] = a[q];
#pragma unroll
    for (int p = 0; p < 4; ++p) {
        const double piv = d[5 * p];
        ok = ok && piv > 0.0 && isfinite(piv);
        // 1 / piv: the hardware reciprocal refined by two Newton steps (within an ulp or two; the
        // inverse only preconditions)
        double ip = __builtin_amdgcn_rcp(piv);
        ip = ip * (2.0 - piv * ip);
        ip = ip * (2.0 - piv * ip);
        double rp[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) rp[c] = d[4 * p + c];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            if (r == p) continue;
            const double fi = d[4 * r + p] * ip;
#pragma unroll
            for (int c = 0; c < 4; ++c) d[4 * r + c] = c == p ? -fi : d[4 * r + c] - fi * rp[c];
        }
#pragma unroll
        for (int c = 0; c < 4; ++c) d[4 * p + c] = c == p ? ip : rp[c] * ip;
    }
}

// In-place Gauss-Jordan inverse of a 64 x 64 tile held in registers by 256 threads: thread t holds
// v[c] = T[t / 4][16 (t % 4) + c].  The pivots go in blocks of four (16 steps, one barrier each):
// for the pivot block P, D = T_PP^-1 (every thread inverts it in registers), then T_Pc <- D T_Pc,
// T_iP <- -T_iP D and T_ic <- T_ic - T_iP D T_Pc, T_PP <- D.  The block's rows and every row's
// entries in the block's columns go through LDS (SW_INV_LDS doubles, two alternating buffers, so
// one barrier per step suffices: a buffer is written again only after every thread passed the
// barrier behind its last reads).  A pivot that is not positive and finite sets *bad.
constexpr int SW_RS = 68;                            // row stride of the staged pivot rows
constexpr int SW_INV_LDS = 2 * (4 * SW_RS + 64 * 4);  // doubles
__device__ void sw_tile_inverse(double (&v)[16], double* buf, int* bad)
{
    const int t = threadIdx.x, i = t >> 2, j0 = (t & 3) * 16;
    bool ok = true;
    for (int s = 0; s < SW_B / 4; ++s) {
        const int p0 = 4 * s;
        double* rows = buf + (s & 1) * (4 * SW_RS + 64 * 4);  // [4][SW_RS]
        double* cols = rows + 4 * SW_RS;                      // [64][4]
        const int ri = i - p0;                                // the thread's row within the block
        if (ri >= 0 && ri < 4) {
#pragma unroll
            for (int c = 0; c < 16; ++c) rows[ri * SW_RS + j0 + c] = v[c];
        }
        const int pc = p0 - j0;  // the block's columns within the thread's 16, if in [0, 16)
        if (pc >= 0 && pc < 16) {
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                double cv = 0.0;
#pragma unroll
                for (int c = 0; c < 16; ++c) cv = c == pc + k ? v[c] : cv;
                cols[i * 4 + k] = cv;
            }
        }
        __syncthreads();
        double a[16], d[16], aiP[4], rp[4][16];
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
            for (int c = 0; c < 4; ++c) a[4 * r + c] = rows[r * SW_RS + p0 + c];
#pragma unroll
        for (int k = 0; k < 4; ++k) aiP[k] = cols[i * 4 + k];
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
            for (int c = 0; c < 16; ++c) rp[r][c] = rows[r * SW_RS + j0 + c];
        sw_inv4(a, d, ok);
        if (ri >= 0 && ri < 4) {  // a pivot row: D T_Pc, and D in the block's columns
#pragma unroll
            for (int c = 0; c < 16; ++c) {
                double acc = 0.0;
#pragma unroll
                for (int k = 0; k < 4; ++k) acc += d[4 * ri + k] * rp[k][c];
                v[c] = acc;
            }
#pragma unroll
            for (int c = 0; c < 16; ++c)
#pragma unroll
                for (int k = 0; k < 4; ++k) v[c] = j0 + c == p0 + k ? d[4 * ri + k] : v[c];
        } else {  // w = T_iP D; T_ic - w T_Pc, and -w in the block's columns
            double w[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                double acc = 0.0;
#pragma unroll
                for (int m = 0; m < 4; ++m) acc += aiP[m] * d[4 * m + k];
                w[k] = acc;
            }
#pragma unroll
            for (int c = 0; c < 16; ++c) {
                double acc = v[c];
#pragma unroll
                for (int k = 0; k < 4; ++k) acc -= w[k] * rp[k][c];
                v[c] = acc;
            }
#pragma unroll
            for (int c = 0; c < 16; ++c)
#pragma unroll
                for (int k = 0; k < 4; ++k) v[c] = j0 + c == p0 + k ? -w[k] : v[c];
        }
    }
    if (!ok && t == 0) *bad = 1;
}

// Ac's buffer before the assembly: zero, with ones on the padding diagonal (rows n .. ldc - 1)
__global__ __launch_bounds__(256) void k_sw_clear(double* __restrict__ A, int n, int ldc)
{
    const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (size_t)ldc * ldc) return;
    const int i = (int)(k / ldc), j = (int)(k % ldc);
    A[k] = (i == j && i >= n) ? 1.0 : 0.0;
}

// D_0 = A_00^-1, and the failure flag cleared
__global__ __launch_bounds__(256) void k_sw_first(const double* __restrict__ A, int ldc, double* __restrict__ D, int* bad)
{
    __shared__ double buf[SW_INV_LDS];
    const int t = threadIdx.x, i = t >> 2, j0 = (t & 3) * 16;
    if (t == 0) *bad = 0;
    double v[16];
#pragma unroll
    for (int c = 0; c < 16; ++c) v[c] = A[(size_t)i * ldc + j0 + c];
    sw_tile_inverse(v, buf, bad);
#pragma unroll
    for (int c = 0; c < 16; ++c) D[i * SW_B + j0 + c] = v[c];
}

// Step K, workgroup I != K: O_I = A_IK (from the stored lower tile: A_IK itself for I > K, the
// transpose of A_KI for I < K) and P_I = O_I D on the matrix cores (D staged by half tiles)
__global__ __launch_bounds__(256) void k_sw_panel(const double* __restrict__ A, int ldc, int K, const double* __restrict__ D,
                                                  double* __restrict__ Pb, double* __restrict__ Ob)
{
    const int I = blockIdx.x;
    if (I == K) return;
    __shared__ double sO[SW_B * SW_LD];  // O_I, 33.3 KB
    __shared__ double sD[32 * SW_B];     // half of D's rows, 16 KB
    const int t = threadIdx.x, i = t >> 2, j0 = (t & 3) * 16;
#pragma unroll
    for (int c = 0; c < 16; ++c) {
        const int j = j0 + c;
        // thread (i, j) loads stored element (row i, col j) of the tile holding A_IK or A_KI
        const double v = I > K ? A[(size_t)(SW_B * I + i) * ldc + SW_B * K + j] : A[(size_t)(SW_B * K + i) * ldc + SW_B * I + j];
        if (I > K) sO[i * SW_LD + j] = v;
        else sO[j * SW_LD + i] = v;  // O_I = (A_KI)^T
    }
    __syncthreads();
    double* O = Ob + (size_t)I * SW_B * SW_B;
#pragma unroll
    for (int c = 0; c < 16; ++c) O[i * SW_B + j0 + c] = sO[i * SW_LD + j0 + c];
    const SwQuarter q;
    sw_v4d acc[2][2];
#pragma unroll
    for (int bi = 0; bi < 2; ++bi)
#pragma unroll
        for (int bj = 0; bj < 2; ++bj) acc[bi][bj] = sw_v4d{0.0, 0.0, 0.0, 0.0};
    for (int h = 0; h < 2; ++h) {
#pragma unroll
        for (int e = 0; e < 8; ++e) sD[t + 256 * e] = D[SW_B * 32 * h + t + 256 * e];  // rows 32 h .. 32 h + 31
        __syncthreads();
        // B[m][j] = D[m][j]: lane l takes sD[(m0 + l / 16) * 64 + col]
        const int r = q.l & 15, k0 = q.l >> 4;
#pragma unroll
        for (int k = 0; k < 32; k += 4) {
            const double a0 = sO[(16 * (2 * q.wr) + r) * SW_LD + 32 * h + k + k0];
            const double a1 = sO[(16 * (2 * q.wr + 1) + r) * SW_LD + 32 * h + k + k0];
            const double b0 = sD[(k + k0) * SW_B + 16 * (2 * q.wc) + r];
            const double b1 = sD[(k + k0) * SW_B + 16 * (2 * q.wc + 1) + r];
            acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
            acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
            acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
            acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
        }
        __syncthreads();
    }
    double* P = Pb + (size_t)I * SW_B * SW_B;
#pragma unroll
    for (int bi = 0; bi < 2; ++bi)
#pragma unroll
        for (int bj = 0; bj < 2; ++bj)
#pragma unroll
            for (int r = 0; r < 4; ++r) P[q.row(bi, r) * SW_B + q.col(bj)] = acc[bi][bj][r];
}

// lower tile number b (row-major over I >= J) -> (I, J)
__device__ inline void sw_tile_of(int b, int& I, int& J)
{
    I = (int)((sqrt(8.0 * b + 1.0) - 1.0) * 0.5);
    while (I * (I + 1) / 2 > b) --I;
    while ((I + 1) * (I + 2) / 2 <= b) ++I;
    J = b - I * (I + 1) / 2;
}

// Step K over every lower tile (grid nt (nt + 1) / 2).  Workgroup 0 takes tile (K+1, K+1) -- and
// tile (0, 0) the workgroup that tile would have had -- so that the next pivot tile is updated and
// inverted (into Dn) first.
__global__ __launch_bounds__(256) void k_sw_update(double* __restrict__ A, int ldc, int K, int nt,
                                                   const double* __restrict__ Pb, const double* __restrict__ Ob,
                                                   const double* __restrict__ D, double* __restrict__ Dn, int* bad)
{
    int b = blockIdx.x;
    const int look = K + 1 < nt ? (K + 1) * (K + 2) / 2 + (K + 1) : -1;  // index of tile (K+1, K+1)
    if (look >= 0) b = b == 0 ? look : (b == look ? 0 : b);
    int I, J;
    sw_tile_of(b, I, J);
    const int t = threadIdx.x, i = t >> 2, j0 = (t & 3) * 16;
    double* a = A + (size_t)(SW_B * I + i) * ldc + SW_B * J + j0;  // this thread's 16 entries
    if (I == K && J == K) {  // A_KK <- -D
#pragma unroll
        for (int c = 0; c < 16; ++c) a[c] = -D[i * SW_B + j0 + c];
        return;
    }
    if (J == K) {  // I > K: A_IK <- P_I
        const double* P = Pb + (size_t)I * SW_B * SW_B;
#pragma unroll
        for (int c = 0; c < 16; ++c) a[c] = P[i * SW_B + j0 + c];
        return;
    }
    if (I == K) {  // J < K: A_KJ <- D A_KJ = (P_J)^T
        const double* P = Pb + (size_t)J * SW_B * SW_B;
#pragma unroll
        for (int c = 0; c < 16; ++c) a[c] = P[(j0 + c) * SW_B + i];
        return;
    }
    // A_IJ <- A_IJ - P_I O_J^T on the matrix cores, the inner dimension in two halves staged in LDS;
    // the next pivot tile is then inverted through the same buffer (33.8 KB)
    __shared__ double sbuf[2 * SW_B * SW_H];
    static_assert(SW_INV_LDS <= 2 * SW_B * SW_H, "the pivot buffers of the tile inverse fit in the staging");
    double* sP = sbuf;
    double* sO = sbuf + SW_B * SW_H;
    const SwQuarter q;
    double* c0 = A + (size_t)(SW_B * I) * ldc + SW_B * J;  // the tile
    sw_v4d acc[2][2];
#pragma unroll
    for (int bi = 0; bi < 2; ++bi)
#pragma unroll
        for (int bj = 0; bj < 2; ++bj)
#pragma unroll
            for (int r = 0; r < 4; ++r) acc[bi][bj][r] = c0[(size_t)q.row(bi, r) * ldc + q.col(bj)];
    const double* P = Pb + (size_t)I * SW_B * SW_B;
    const double* O = Ob + (size_t)J * SW_B * SW_B;
    for (int h = 0; h < 2; ++h) {
        // rows 0..63, inner columns 32 h .. 32 h + 31 of P_I and O_J: 2048 values each, 8 per thread
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int x = t + 256 * e, r = x >> 5, m = x & 31;
            sP[r * SW_H + m] = P[r * SW_B + 32 * h + m];
            sO[r * SW_H + m] = O[r * SW_B + 32 * h + m];
        }
        __syncthreads();
        sw_mfma_rows_cols(q, sP, SW_H, sO, SW_H, -1.0, acc);  // B[k][j] = O_J[j][k]
        __syncthreads();
    }
#pragma unroll
    for (int bi = 0; bi < 2; ++bi)
#pragma unroll
        for (int bj = 0; bj < 2; ++bj)
#pragma unroll
            for (int r = 0; r < 4; ++r) c0[(size_t)q.row(bi, r) * ldc + q.col(bj)] = acc[bi][bj][r];
    if (I == K + 1 && J == K + 1) {  // the next pivot: invert the updated tile now
        double* T = sbuf;            // SW_B * SW_LD doubles fit; the last barrier above freed the buffer
#pragma unroll
        for (int bi = 0; bi < 2; ++bi)
#pragma unroll
            for (int bj = 0; bj < 2; ++bj)
#pragma unroll
                for (int r = 0; r < 4; ++r) T[q.row(bi, r) * SW_LD + q.col(bj)] = acc[bi][bj][r];
        __syncthreads();
        double v[16];
#pragma unroll
        for (int c = 0; c < 16; ++c) v[c] = T[i * SW_LD + j0 + c];
        __syncthreads();  // T is reused as the pivot buffers
        sw_tile_inverse(v, sbuf, bad);
#pragma unroll
        for (int c = 0; c < 16; ++c) Dn[i * SW_B + j0 + c] = v[c];
    }
}

// The scale of the FP16 copy: the largest diagonal entry of Ac^-1 (the sweep leaves -Ac^-1), which
// bounds every entry of a positive definite matrix (|a_ij| <= sqrt(a_ii a_jj)), so the scaled
// entries lie in [-1, 1].
__global__ __launch_bounds__(256) void k_tl_scale(const double* __restrict__ A, int n, int ldc, double* __restrict__ sc)
{
    __shared__ double sh[256];
    double m = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) m = fmax(m, -A[(size_t)i * ldc + i]);
    sh[threadIdx.x] = m;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double s = sh[0] > 0.0 && isfinite(sh[0]) ? sh[0] : 1.0;
        sc[0] = s;
        sc[1] = 1.0 / s;
    }
}

// Symmetric FP16 copy of Ac^-1 / scale from the lower triangle the sweep leaves (negated), rows of
// ld halves with zeros past column n.
__global__ __launch_bounds__(256) void k_tl_symm(const double* __restrict__ A, int ldc, __half* __restrict__ Ainv, int n,
                                                int ld, const double* __restrict__ sc)
{
    const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (size_t)n * ld) return;
    const int i = (int)(k / ld), j = (int)(k % ld);
    const double v = j < n ? -(i >= j ? A[(size_t)i * ldc + j] : A[(size_t)j * ldc + i]) * sc[1] : 0.0;
    Ainv[k] = __float2half((float)v);
}

// the same, unscaled in FP64 (FP64 preconditioner storage)
__global__ __launch_bounds__(256) void k_tl_symm64(const double* __restrict__ A, int ldc, double* __restrict__ Ainv, int n,
                                                  int ld)
{
    const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (size_t)n * ld) return;
    const int i = (int)(k / ld), j = (int)(k % ld);
    Ainv[k] = j < n ? -(i >= j ? A[(size_t)i * ldc + j] : A[(size_t)j * ldc + i]) : 0.0;
}

// ---------------------------------------------------------------------------------------------
// Single-kernel CG (the single-reduction recurrences of Chronopoulos and Gear):
//   s_{k-1} = w_{k-1} + b_{k-1} s_{k-2},  r_k = r_{k-1} - a_{k-1} s_{k-1},  u_k = D^-1 r_k,
//   p_{k-1} = u_{k-1} + b_{k-1} p_{k-2},  x_k = x_{k-1} + a_{k-1} p_{k-1},
//   w_k = M u_k,  g_k = (r_k, u_k),  d_k = (w_k, u_k),
//   b_k = g_k / g_{k-1},  a_k = g_k / (d_k - b_k g_k / a_{k-1}).
// Launch k forms a_{k-1}, b_{k-1} from the partials of launch k-1 and needs u_k of every
// neighbouring block column for w_k = M u_k.  D^-1 is linear, so
//   u_k = u_{k-1} - a_{k-1} (m_{k-1} + b_{k-1} q_{k-2}),  m = D^-1 w,  q = D^-1 s,
// and each launch writes u, m, q for its own nodes (ping-pong buffers): a neighbour costs three
// 3-vectors, and w stays an explicit product with u (no recurrence for w, as in pipelined CG).
__device__ inline double dinv_apply3(const double* __restrict__ D, double r0, double r1, double r2)
{
    return D[0] * r0 + D[1] * r1 + D[2] * r2;
}

template <typename VT>
__global__ __launch_bounds__(WG) void k_cg1_init(TileArgs ta, const int32_t* __restrict__ row_ptr,
                                                 const int32_t* __restrict__ col, const VT* __restrict__ val, int nb,
                                                 const double* __restrict__ dinv, const double* __restrict__ b,
                                                 double* __restrict__ x, double* __restrict__ r, double* __restrict__ w,
                                                 double* __restrict__ s, double* __restrict__ p, double* __restrict__ u0,
                                                 double* __restrict__ m0, double* __restrict__ q0, int tile_cap)
{
    extern __shared__ double prod[];  // 3 doubles per block of the tile (tile_cap blocks), then the tile's w
    const int t = blockIdx.x;
    const int sg = ta.tile_seg[t];
    const bool active = ta.seg[sg].active;
    const int rr0 = ta.tile_r0[t], rr1 = ta.tile_r1[t];
    const int d0 = 3 * rr0, d1 = 3 * rr1;
    const int b0 = row_ptr[rr0], b1 = row_ptr[rr1];
    double* wloc = prod + 3 * (size_t)tile_cap;
    tile_products<VT>(col, val, nb, b0, b1,
                      [&](int c, double& a0, double& a1, double& a2) {
                          const double q0v = active ? b[3 * c] : 0.0, q1v = active ? b[3 * c + 1] : 0.0,
                                       q2v = active ? b[3 * c + 2] : 0.0;
                          const double* D = dinv + 9 * c;
                          a0 = dinv_apply3(D, q0v, q1v, q2v);
                          a1 = dinv_apply3(D + 3, q0v, q1v, q2v);
                          a2 = dinv_apply3(D + 6, q0v, q1v, q2v);
                      },
                      prod);
    double g = 0.0, dl = 0.0, rr = 0.0;
    for (int task = threadIdx.x; task < d1 - d0; task += WG) {
        const int d = d0 + task, rw = d / 3, comp = d % 3;
        const double bv = active ? b[d] : 0.0;
        const double* D = dinv + 9 * rw + 3 * comp;
        const double* bb = b + 3 * rw;
        const double uv = active ? D[0] * bb[0] + D[1] * bb[1] + D[2] * bb[2] : 0.0;
        const double wv = row_sum(row_ptr, rw, comp, b0, prod);
        x[d] = 0.0; r[d] = bv; s[d] = 0.0; p[d] = 0.0; w[d] = wv; u0[d] = uv; q0[d] = 0.0;
        wloc[task] = wv;
        g += bv * uv;
        dl += wv * uv;
        rr += bv * bv;
    }
    __syncthreads();
    for (int task = threadIdx.x; task < d1 - d0; task += WG) {
        const int d = d0 + task, nd = task - d % 3;
        const double* D = dinv + 9 * (d / 3) + 3 * (d % 3);
        m0[d] = D[0] * wloc[nd] + D[1] * wloc[nd + 1] + D[2] * wloc[nd + 2];
    }
    const int sl[4] = {SL_RHO0, SL_DL0, SL_RR0, SL_BB};
    double vv[4] = {g, dl, rr, rr};
    tile_write<4>(ta, t, sl, vv);
}

// Warm start of the single-kernel CG from a guess x0 (the same Newton iteration's update of the
// previous timestep), in two launches: (a) x = x0, r0 = b - M x0, u0 = D^-1 r0; (b) w0 = M u0,
// m0 = D^-1 w0 and the first partials.  Tissues not being solved start from x = 0, r = 0.
template <typename VT>
__global__ __launch_bounds__(WG) void k_cg1_warm_a(TileArgs ta, const int32_t* __restrict__ row_ptr,
                                                   const int32_t* __restrict__ col, const VT* __restrict__ val, int nb,
                                                   const double* __restrict__ dinv, const double* __restrict__ b,
                                                   const double* __restrict__ x0, double* __restrict__ x,
                                                   double* __restrict__ r, double* __restrict__ s, double* __restrict__ p,
                                                   double* __restrict__ u0, double* __restrict__ q0, int tile_cap)
{
    extern __shared__ double prod[];
    const int t = blockIdx.x;
    const int sg = ta.tile_seg[t];
    const bool active = ta.seg[sg].active;
    const int rr0 = ta.tile_r0[t], rr1 = ta.tile_r1[t];
    const int d0 = 3 * rr0, d1 = 3 * rr1;
    const int b0 = row_ptr[rr0], b1 = row_ptr[rr1];
    double* rloc = prod + 3 * (size_t)tile_cap;
    tile_products<VT>(col, val, nb, b0, b1, XLOAD(x0), prod);
    for (int task = threadIdx.x; task < d1 - d0; task += WG) {
        const int d = d0 + task, rw = d / 3, comp = d % 3;
        const double mx = row_sum(row_ptr, rw, comp, b0, prod);
        const double rv = active ? b[d] - mx : 0.0;
        x[d] = active ? x0[d] : 0.0;
        r[d] = rv; s[d] = 0.0; p[d] = 0.0; q0[d] = 0.0;
        rloc[task] = rv;
    }
    __syncthreads();
    for (int task = threadIdx.x; task < d1 - d0; task += WG) {
        const int d = d0 + task, nd = task - d % 3;
        const double* D = dinv + 9 * (d / 3) + 3 * (d % 3);
        u0[d] = D[0] * rloc[nd] + D[1] * rloc[nd + 1] + D[2] * rloc[nd + 2];
    }
}

template <typename VT>
__global__ __launch_bounds__(WG) void k_cg1_warm_b(TileArgs ta, const int32_t* __restrict__ row_ptr,
                                                   const int32_t* __restrict__ col, const VT* __restrict__ val, int nb,
                                                   const double* __restrict__ dinv, const double* __restrict__ b,
                                                   const double* __restrict__ r, const double* __restrict__ u0,
                                                   double* __restrict__ w, double* __restrict__ m0, int tile_cap)
{
    extern __shared__ double prod[];
    const int t = blockIdx.x;
    const int sg = ta.tile_seg[t];
    const bool active = ta.seg[sg].active;
    const int rr0 = ta.tile_r0[t], rr1 = ta.tile_r1[t];
    const int d0 = 3 * rr0, d1 = 3 * rr1;
    const int b0 = row_ptr[rr0], b1 = row_ptr[rr1];
    double* wloc = prod + 3 * (size_t)tile_cap;
    tile_products<VT>(col, val, nb, b0, b1, XLOAD(u0), prod);
    double g = 0.0, dl = 0.0, rr = 0.0, bb = 0.0;
    for (int task = threadIdx.x; task < d1 - d0; task += WG) {
        const int d = d0 + task, rw = d / 3, comp = d % 3;
        const double wv = row_sum(row_ptr, rw, comp, b0, prod);
        w[d] = wv;
        wloc[task] = wv;
        const double rv = r[d], uv = u0[d], bv = active ? b[d] : 0.0;
        g += rv * uv;
        dl += wv * uv;
        rr += rv * rv;
        bb += bv * bv;
    }
    __syncthreads();
    for (int task = threadIdx.x; task < d1 - d0; task += WG) {
        const int d = d0 + task, nd = task - d % 3;
        const double* D = dinv + 9 * (d / 3) + 3 * (d % 3);
        m0[d] = D[0] * wloc[nd] + D[1] * wloc[nd + 1] + D[2] * wloc[nd + 2];
    }
    const int sl[4] = {SL_RHO0, SL_DL0, SL_RR0, SL_BB};
    double vv[4] = {g, dl, rr, bb};
    tile_write<4>(ta, t, sl, vv);
}

// One iteration per launch.  Everything that does not depend on this iteration's scalars (the
// tile's matrix blocks, the neighbour columns' u, m, q, the own rows' vectors and block-inverse
// rows) is loaded into registers first, so the loads are in flight while the workgroup re-sums
// the partials; then the scalars are formed, own rows updated and the products parked in LDS.
// WGT lanes per workgroup, BPT matrix blocks per lane (tile capacity WGT*BPT), at most one own
// DoF per lane (tiles hold at most WGT/3 block rows).  Launch parity P reads u, m, q of parity
// 1-P and writes parity P; x, p, r, w, s are read and written by their own lane only.
template <typename VT, int P, int WGT, int BPT>
__global__ __launch_bounds__(WGT) void k_cg1(TileArgs ta, const int32_t* __restrict__ row_ptr, const int32_t* __restrict__ col,
                                             const VT* __restrict__ val, int nb, const double* __restrict__ dinv,
                                             double* __restrict__ x, double* __restrict__ p, double* __restrict__ r,
                                             double* __restrict__ w, double* __restrict__ s,
                                             const double* __restrict__ u_old, double* __restrict__ u_new,
                                             const double* __restrict__ m_old, double* __restrict__ m_new,
                                             const double* __restrict__ q_old, double* __restrict__ q_new, int maxit,
                                             int tile_cap)
{
    extern __shared__ double prod[];
    const int t = xcd_tile(blockIdx.x, ta.T, ta.nxcd);
    const TileInfo ti = ta.info[t];  // tissue, blocks, rows and the tissue's tiles: one scalar load
    const int sg = ti.seg, rr0 = ti.r0, rr1 = ti.r1, b0 = ti.b0, b1 = ti.b1;
    SegScalars& S = ta.seg[sg];
    if (seg_done(S)) return;
    const int d0 = 3 * rr0, d1 = 3 * rr1;
    // ---- prefetch ----
    double mv[BPT][9], cu[BPT][3], cm[BPT][3], cq[BPT][3];
#pragma unroll
    for (int k = 0; k < BPT; ++k) {
        const int j = b0 + threadIdx.x + k * WGT;
        if (j < b1) {
            const int c = col[j];
#pragma unroll
            for (int e = 0; e < 9; ++e) mv[k][e] = val[(size_t)e * nb + j];
#pragma unroll
            for (int e = 0; e < 3; ++e) {
                cu[k][e] = u_old[3 * c + e];
                cm[k][e] = m_old[3 * c + e];
                cq[k][e] = q_old[3 * c + e];
            }
        }
    }
    const int d = d0 + threadIdx.x;
    const bool own = d < d1;
    double ow = 0.0, os = 0.0, orr = 0.0, ou = 0.0, op = 0.0, ox = 0.0, od[3] = {0.0, 0.0, 0.0};
    int rb0 = 0, rb1 = 0;
    if (own) {
        ow = w[d]; os = s[d]; orr = r[d]; ou = u_old[d]; op = p[d]; ox = x[d];
        const double* D = dinv + 9 * (d / 3) + 3 * (d % 3);
        od[0] = D[0]; od[1] = D[1]; od[2] = D[2];
        rb0 = row_ptr[d / 3]; rb1 = row_ptr[d / 3 + 1];
    }
    double* rloc = prod + 3 * (size_t)tile_cap;  // the tile's new residual, exchanged in LDS
    double* sloc = rloc + WGT;                    // the tile's new s
    // ---- scalars of iteration k-1: partials of launch k-1 (slots of parity 1-P) and the
    //      history launch k-1 left in the parity-P slots of the tissue's scalars ----
    const int sl[3] = {SL_RHO0 + (1 - P), SL_DL0 + (1 - P), SL_RR0 + (1 - P)};
    double q3[3];
    seg_sums_range<3, WGT>(ta, ti.st0, ti.st1, sl, q3);
    const double gk = q3[0], dk = q3[1], rr = q3[2];
    const int it = S.iters_h[P];
    const double best = S.best_rr_h[P];
    const int best_it = S.best_it_h[P];
    const bool lead = t == ti.st0 && threadIdx.x == 0;
    {
        // The single-reduction recurrences can level off just above a very tight tolerance:
        // without a halving of ||r|| for 256 iterations the solve stops, as converged when the
        // best residual came within 10x of the tolerance, else as stalled (BiCGSTAB takes over).
        const bool stagnant = it - best_it > 256;
        const bool conv = rr <= S.tol2 || (stagnant && best <= 100.0 * S.tol2);
        const bool cap = !conv && (it >= maxit || !isfinite(rr) || stagnant);
        if (conv || cap) {
            if (lead) {
                S.rnorm2 = rr;
                S.iters = it;
                if (conv) S.converged = 1;
                else S.stalled = 1;
            }
            return;
        }
    }
    const bool first = it == 0;
    const double beta = first ? 0.0 : gk / S.gam_hist[P];
    const double den = first ? dk : dk - beta * gk / S.alpha_hist[P];
    const double alpha = gk / den;
    if (!(den > 0.0) || !isfinite(alpha)) {  // M not positive definite along this direction
        if (lead) {  // x is still the iterate whose residual is rr
            S.rnorm2 = rr;
            S.iters = it;
            S.stalled = 1;
        }
        return;
    }
    // ---- own rows ----
    if (own) {
        const double sn = ow + beta * os;
        const double rn = orr - alpha * sn;
        const double pn = ou + beta * op;
        s[d] = sn;
        r[d] = rn;
        rloc[d - d0] = rn;
        sloc[d - d0] = sn;
        p[d] = pn;
        x[d] = ox + alpha * pn;
    }
    // ---- products with the neighbours' u_k = u_{k-1} - a (m_{k-1} + b q_{k-2}) ----
#pragma unroll
    for (int k = 0; k < BPT; ++k) {
        const int j = b0 + threadIdx.x + k * WGT;
        if (j < b1) {
            const double a0 = cu[k][0] - alpha * (cm[k][0] + beta * cq[k][0]);
            const double a1 = cu[k][1] - alpha * (cm[k][1] + beta * cq[k][1]);
            const double a2 = cu[k][2] - alpha * (cm[k][2] + beta * cq[k][2]);
            double* pr = prod + 3 * (j - b0);
            pr[0] = mv[k][0] * a0 + mv[k][1] * a1 + mv[k][2] * a2;
            pr[1] = mv[k][3] * a0 + mv[k][4] * a1 + mv[k][5] * a2;
            pr[2] = mv[k][6] * a0 + mv[k][7] * a1 + mv[k][8] * a2;
        }
    }
    __syncthreads();
    double g = 0.0, dl = 0.0, rrn = 0.0, wv = 0.0;
    const int nd = d - d0 - d % 3;  // first DoF of this lane's node within the tile
    if (own) {
        const double rv = rloc[d - d0];
        const double un = od[0] * rloc[nd] + od[1] * rloc[nd + 1] + od[2] * rloc[nd + 2];
        u_new[d] = un;
        q_new[d] = od[0] * sloc[nd] + od[1] * sloc[nd + 1] + od[2] * sloc[nd + 2];
        const int comp = d % 3;
        for (int j = rb0; j < rb1; ++j) wv += prod[3 * (j - b0) + comp];
        w[d] = wv;
        g = rv * un;
        dl = wv * un;
        rrn = rv * rv;
    }
    __syncthreads();  // every lane has read rloc/sloc: reuse rloc for the tile's new w
    if (own) rloc[d - d0] = wv;
    __syncthreads();
    if (own) m_new[d] = od[0] * rloc[nd] + od[1] * rloc[nd + 1] + od[2] * rloc[nd + 2];
    const int so[3] = {SL_RHO0 + P, SL_DL0 + P, SL_RR0 + P};
    double vv[3] = {g, dl, rrn};
    tile_write<3, WGT>(ta, t, so, vv);
    if (lead) {
        S.alpha_hist[1 - P] = alpha;
        S.gam_hist[1 - P] = gk;
        S.iters_h[1 - P] = it + 1;
        S.iters = it + 1;
        const bool better = rr < 0.5 * best;
        S.best_rr_h[1 - P] = better ? rr : best;
        S.best_it_h[1 - P] = better ? it : best_it;
    }
}

// ---------------------------------------------------------------------------------------------
// Pipelined CG (Ghysels and Vanroose), one kernel per iteration:
//   g_i = (r_i, u_i), d_i = (w_i, u_i), n_i = M m_i with m_i = D^-1 w_i,
//   b_i = g_i / g_{i-1}, a_i = g_i / (d_i - b_i g_i / a_{i-1})   (b_0 = 0, a_0 = g_0 / d_0),
//   z = n + b z, q = m + b q, s = w + b s, p = u + b p,
//   x += a p, r -= a s, u -= a q, w -= a z,  and m_{i+1} = D^-1 w_{i+1}.
// The SpMV operand m_i is complete when launch i starts (launch i-1 wrote it for its own rows), so
// a block needs only its column's m (3 doubles) and the products start before the scalars are
// known.  m is ping-ponged (launch P reads m[1-P], writes m[P]); every other vector is touched by
// its own row's lane only.
template <typename VT>
__global__ __launch_bounds__(WG) void k_pipe_init(TileArgs ta, const int32_t* __restrict__ row_ptr,
                                                  const int32_t* __restrict__ col, const VT* __restrict__ val, int nb,
                                                  const double* __restrict__ dinv, const double* __restrict__ b,
                                                  double* __restrict__ x, double* __restrict__ r, double* __restrict__ u,
                                                  double* __restrict__ w, double* __restrict__ z, double* __restrict__ q,
                                                  double* __restrict__ s, double* __restrict__ p, double* __restrict__ m0,
                                                  int tile_cap)
{
    extern __shared__ double prod[];  // 3 doubles per block of the tile, then the tile's w
    const int t = blockIdx.x;
    const int sg = ta.tile_seg[t];
    const bool active = ta.seg[sg].active;
    const int rr0 = ta.tile_r0[t], rr1 = ta.tile_r1[t];
    const int d0 = 3 * rr0, d1 = 3 * rr1;
    const int b0 = row_ptr[rr0], b1 = row_ptr[rr1];
    double* wloc = prod + 3 * (size_t)tile_cap;
    tile_products<VT>(col, val, nb, b0, b1,
                      [&](int c, double& a0, double& a1, double& a2) {
                          const double q0 = active ? b[3 * c] : 0.0, q1 = active ? b[3 * c + 1] : 0.0,
                                       q2 = active ? b[3 * c + 2] : 0.0;
                          const double* D = dinv + 9 * c;
                          a0 = dinv_apply3(D, q0, q1, q2);
                          a1 = dinv_apply3(D + 3, q0, q1, q2);
                          a2 = dinv_apply3(D + 6, q0, q1, q2);
                      },
                      prod);
    double g = 0.0, dl = 0.0, rr = 0.0;
    for (int task = threadIdx.x; task < d1 - d0; task += WG) {
        const int d = d0 + task, rw = d / 3, comp = d % 3;
        const double bv = active ? b[d] : 0.0;
        const double* D = dinv + 9 * rw + 3 * comp;
        const double* bb = b + 3 * rw;
        const double uv = active ? D[0] * bb[0] + D[1] * bb[1] + D[2] * bb[2] : 0.0;
        const double wv = row_sum(row_ptr, rw, comp, b0, prod);
        x[d] = 0.0; r[d] = bv; u[d] = uv; w[d] = wv;
        z[d] = 0.0; q[d] = 0.0; s[d] = 0.0; p[d] = 0.0;
        wloc[task] = wv;
        g += bv * uv;
        dl += wv * uv;
        rr += bv * bv;
    }
    __syncthreads();
    for (int task = threadIdx.x; task < d1 - d0; task += WG) {
        const int d = d0 + task, nd = task - d % 3;
        const double* D = dinv + 9 * (d / 3) + 3 * (d % 3);
        m0[d] = D[0] * wloc[nd] + D[1] * wloc[nd + 1] + D[2] * wloc[nd + 2];
    }
    const int sl[4] = {SL_RHO0, SL_DL0, SL_RR0, SL_BB};
    double vv[4] = {g, dl, rr, rr};
    tile_write<4>(ta, t, sl, vv);
}

template <typename VT, int P, int WGT, int BPT>
__global__ __launch_bounds__(WGT) void k_pipe(TileArgs ta, const int32_t* __restrict__ row_ptr, const int32_t* __restrict__ col,
                                              const VT* __restrict__ val, int nb, const double* __restrict__ dinv,
                                              double* __restrict__ x, double* __restrict__ r, double* __restrict__ u,
                                              double* __restrict__ w, double* __restrict__ z, double* __restrict__ q,
                                              double* __restrict__ s, double* __restrict__ p,
                                              const double* __restrict__ m_old, double* __restrict__ m_new, int maxit,
                                              int tile_cap)
{
    extern __shared__ double prod[];
    const int t = xcd_tile(blockIdx.x, ta.T, ta.nxcd);
    const int sg = ta.tile_seg[t];
    SegScalars& S = ta.seg[sg];
    if (seg_done(S)) return;
    const int rr0 = ta.tile_r0[t], rr1 = ta.tile_r1[t];
    const int b0 = row_ptr[rr0], b1 = row_ptr[rr1];
    const int d0 = 3 * rr0, d1 = 3 * rr1;
    // ---- n = M m: independent of this iteration's scalars, so it goes first ----
#pragma unroll
    for (int k = 0; k < BPT; ++k) {
        const int j = b0 + threadIdx.x + k * WGT;
        if (j < b1) {
            const int c = col[j];
            const double a0 = m_old[3 * c], a1 = m_old[3 * c + 1], a2 = m_old[3 * c + 2];
            double mv[9];
#pragma unroll
            for (int e = 0; e < 9; ++e) mv[e] = (double)val[(size_t)e * nb + j];
            double* pr = prod + 3 * (j - b0);
            pr[0] = mv[0] * a0 + mv[1] * a1 + mv[2] * a2;
            pr[1] = mv[3] * a0 + mv[4] * a1 + mv[5] * a2;
            pr[2] = mv[6] * a0 + mv[7] * a1 + mv[8] * a2;
        }
    }
    const int d = d0 + threadIdx.x;
    const bool own = d < d1;
    double oz = 0, oq = 0, os = 0, op = 0, ox = 0, orr = 0, ou = 0, ow = 0, om = 0, od[3] = {0, 0, 0};
    int rb0 = 0, rb1 = 0;
    if (own) {
        oz = z[d]; oq = q[d]; os = s[d]; op = p[d]; ox = x[d]; orr = r[d]; ou = u[d]; ow = w[d]; om = m_old[d];
        const double* D = dinv + 9 * (d / 3) + 3 * (d % 3);
        od[0] = D[0]; od[1] = D[1]; od[2] = D[2];
        rb0 = row_ptr[d / 3]; rb1 = row_ptr[d / 3 + 1];
    }
    double* wloc = prod + 3 * (size_t)tile_cap;
    // ---- scalars of iteration i from the partials of launch i-1 ----
    const int sl[3] = {SL_RHO0 + (1 - P), SL_DL0 + (1 - P), SL_RR0 + (1 - P)};
    double q3[3];
    seg_sums<3, WGT>(ta, sg, sl, q3);  // its barriers also publish the products above
    const double gk = q3[0], dk = q3[1], rr = q3[2];
    const int it = S.iters_h[P];
    const double best = S.best_rr_h[P];
    const int best_it = S.best_it_h[P];
    const bool lead = first_tile(ta, t, sg) && threadIdx.x == 0;
    {
        const bool stagnant = it - best_it > 256;
        const bool conv = rr <= S.tol2 || (stagnant && best <= 100.0 * S.tol2);
        const bool cap = !conv && (it >= maxit || !isfinite(rr) || stagnant);
        if (conv || cap) {
            if (lead) {
                S.rnorm2 = rr;
                S.iters = it;
                if (conv) S.converged = 1;
                else S.stalled = 1;
            }
            return;
        }
    }
    const bool first = it == 0;
    const double beta = first ? 0.0 : gk / S.gam_hist[P];
    const double den = first ? dk : dk - beta * gk / S.alpha_hist[P];
    const double alpha = gk / den;
    if (!(den > 0.0) || !isfinite(alpha)) {
        if (lead) {  // x is still the iterate whose residual is rr
            S.rnorm2 = rr;
            S.iters = it;
            S.stalled = 1;
        }
        return;
    }
    double rn = 0.0, un = 0.0, wn = 0.0;
    if (own) {
        const int comp = d % 3;
        double nv = 0.0;
        for (int j = rb0; j < rb1; ++j) nv += prod[3 * (j - b0) + comp];
        const double zn = nv + beta * oz, qn = om + beta * oq, sn = ow + beta * os, pn = ou + beta * op;
        z[d] = zn; q[d] = qn; s[d] = sn; p[d] = pn;
        x[d] = ox + alpha * pn;
        rn = orr - alpha * sn;
        un = ou - alpha * qn;
        wn = ow - alpha * zn;
        r[d] = rn; u[d] = un; w[d] = wn;
        wloc[d - d0] = wn;
    }
    __syncthreads();
    double g = 0.0, dl = 0.0, rrn = 0.0;
    if (own) {
        const int nd = d - d0 - d % 3;
        m_new[d] = od[0] * wloc[nd] + od[1] * wloc[nd + 1] + od[2] * wloc[nd + 2];
        g = rn * un;
        dl = wn * un;
        rrn = rn * rn;
    }
    const int so[3] = {SL_RHO0 + P, SL_DL0 + P, SL_RR0 + P};
    double vv[3] = {g, dl, rrn};
    tile_write<3, WGT>(ta, t, so, vv);
    if (lead) {
        S.alpha_hist[1 - P] = alpha;
        S.gam_hist[1 - P] = gk;
        S.iters_h[1 - P] = it + 1;
        S.iters = it + 1;
        const bool better = rr < 0.5 * best;
        S.best_rr_h[1 - P] = better ? rr : best;
        S.best_it_h[1 - P] = better ? it : best_it;
    }
}

// ---------------------------------------------------------------------------------------------
// per-tissue ||x||_2
__global__ __launch_bounds__(WG) void k_tile_norm(TileArgs ta, const double* __restrict__ x)
{
    const int t = blockIdx.x;
    const int d0 = 3 * ta.tile_r0[t], d1 = 3 * ta.tile_r1[t];
    double s = 0.0;
    for (int d = d0 + threadIdx.x; d < d1; d += WG) s += x[d] * x[d];
    const int sl[1] = {SL_NORM};
    double vv[1] = {s};
    tile_write<1>(ta, t, sl, vv);
}
__global__ __launch_bounds__(WG) void k_seg_norm_fin(TileArgs ta, double* __restrict__ out)
{
    const int sg = blockIdx.x;
    const int sl[1] = {SL_NORM};
    double tot[1];
    seg_sums<1>(ta, sg, sl, tot);
    if (threadIdx.x == 0) out[sg] = sqrt(tot[0]);
}

TileArgs tile_args(tf_handle* h)
{
    return TileArgs{h->tile_info, h->tile_r0, h->tile_r1, h->tile_seg, h->seg_tile_ptr, h->partial, h->seg, h->T,
                    h->opts.xcd_tile_map ? h->n_xcd : 1};
}

// ---- warm start: least-squares combination of earlier updates ---------------------------------
struct WarmArgs {
    const double* const* z;  // [TF_WARM_HIST] the kept updates (a device table, newest first)
    const double* g;
    int m;
    int cap;                 // blocks of a tile (the one-pass Gram kernel's LDS stride)
};

// per tile, in one launch: M z_i on the tile's rows for every kept update z_i (one CSR-stream
// pass per update; the row results stay in registers, at most 4 rows x components per lane),
// then the partials of (Mz_i, Mz_j) for i <= j in row-major upper-triangle order, then (Mz_i, g)
// ONE: all kept updates' products in one pass over the blocks (LDS for TF_WARM_HIST of them)
template <typename VT, bool ONE>
__global__ __launch_bounds__(WG) void k_warm_gram(TileArgs ta, WarmArgs wa, const int32_t* __restrict__ row_ptr,
                                                  const int32_t* __restrict__ col, const VT* __restrict__ val, int nb,
                                                  double* __restrict__ part)
{
    constexpr int MAXT = 4;  // tasks per lane: 3 (rows of a tile) <= 1024
    extern __shared__ double prod[];
    const int t = blockIdx.x;
    const int r0 = ta.tile_r0[t], r1 = ta.tile_r1[t];
    const int b0 = row_ptr[r0], b1 = row_ptr[r1];
    const int ntask = 3 * (r1 - r0);
    double mz[TF_WARM_HIST][MAXT];
    if (ONE) {
        // one pass over the tile's blocks: every block's values read once for all kept updates,
        // products of update i at prod + i * 3 * cap (the same products and row sums as below)
        const int cap = wa.cap;
        const double* z[TF_WARM_HIST];
#pragma unroll
        for (int i = 0; i < TF_WARM_HIST; ++i) z[i] = i < wa.m ? wa.z[i] : nullptr;
        for (int j = b0 + threadIdx.x; j < b1; j += WG) {
            const int c = col[j];
            const double v0 = val[j], v1 = val[(size_t)nb + j], v2 = val[2 * (size_t)nb + j];
            const double v3 = val[3 * (size_t)nb + j], v4 = val[4 * (size_t)nb + j], v5 = val[5 * (size_t)nb + j];
            const double v6 = val[6 * (size_t)nb + j], v7 = val[7 * (size_t)nb + j], v8 = val[8 * (size_t)nb + j];
#pragma unroll
            for (int i = 0; i < TF_WARM_HIST; ++i) {
                if (i >= wa.m) continue;
                const double x0 = z[i][3 * c], x1 = z[i][3 * c + 1], x2 = z[i][3 * c + 2];
                double* p = prod + (size_t)i * 3 * cap + 3 * (j - b0);
                p[0] = v0 * x0 + v1 * x1 + v2 * x2;
                p[1] = v3 * x0 + v4 * x1 + v5 * x2;
                p[2] = v6 * x0 + v7 * x1 + v8 * x2;
            }
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < TF_WARM_HIST; ++i)
#pragma unroll
            for (int k = 0; k < MAXT; ++k) {
                const int task = threadIdx.x + k * WG;
                mz[i][k] = i < wa.m && task < ntask
                               ? row_sum(row_ptr, r0 + task / 3, task % 3, b0, prod + (size_t)i * 3 * cap)
                               : 0.0;
            }
    } else {
#pragma unroll
        for (int i = 0; i < TF_WARM_HIST; ++i) {
#pragma unroll
            for (int k = 0; k < MAXT; ++k) mz[i][k] = 0.0;
            if (i < wa.m) {
                const double* z = wa.z[i];
                tile_products<VT>(col, val, nb, b0, b1, XLOAD(z), prod);
#pragma unroll
                for (int k = 0; k < MAXT; ++k) {
                    const int task = threadIdx.x + k * WG;
                    if (task < ntask) mz[i][k] = row_sum(row_ptr, r0 + task / 3, task % 3, b0, prod);
                }
                __syncthreads();  // prod is reused by the next update
            }
        }
    }
    double acc[TF_WARM_DOTS];
#pragma unroll
    for (int q = 0; q < TF_WARM_DOTS; ++q) acc[q] = 0.0;
#pragma unroll
    for (int k = 0; k < MAXT; ++k) {
        const int task = threadIdx.x + k * WG;
        if (task >= ntask) continue;
        const double gd = wa.g[3 * r0 + task];
        int q = 0;
#pragma unroll
        for (int i = 0; i < TF_WARM_HIST; ++i)
#pragma unroll
            for (int j = i; j < TF_WARM_HIST; ++j) acc[q++] += mz[i][k] * mz[j][k];
#pragma unroll
        for (int i = 0; i < TF_WARM_HIST; ++i) acc[q++] += mz[i][k] * gd;
    }
    constexpr int NW = WG / 64;
    __shared__ double red[TF_WARM_DOTS][NW];
    const int w = threadIdx.x / 64, l = threadIdx.x % 64;
#pragma unroll
    for (int q = 0; q < TF_WARM_DOTS; ++q) {
        const double sv = wave_sum(acc[q]);
        if (l == 0) red[q][w] = sv;
    }
    __syncthreads();
    if (threadIdx.x < TF_WARM_DOTS) {
        double sv = 0.0;
#pragma unroll
        for (int k = 0; k < NW; ++k) sv += red[threadIdx.x][k];
        part[(size_t)threadIdx.x * ta.T + t] = sv;
    }
}

// per tissue and dot product (one wave each): the tile partials, lane l summing tiles l, l + 64,
// ... in order, then a fixed tree over the lanes
__global__ __launch_bounds__(64) void k_warm_seg(TileArgs ta, const double* __restrict__ part, double* __restrict__ out)
{
    const int sg = blockIdx.x, q = blockIdx.y;
    double sv = 0.0;
    for (int t = ta.seg_tile_ptr[sg] + threadIdx.x; t < ta.seg_tile_ptr[sg + 1]; t += 64) sv += part[(size_t)q * ta.T + t];
    sv = wave_sum(sv);
    if (threadIdx.x == 0) out[(size_t)sg * TF_WARM_DOTS + q] = sv;
}

// Coefficients y of the least-squares combination min ||g - sum_j y_j M z_j|| from the tissue's
// Gram entries D (upper triangle of G = (Mz_i, Mz_j) row-major over TF_WARM_HIST updates, then
// r = (Mz_i, g)): G is scaled to unit diagonal and factored by Cholesky in the order of the updates
// (the newest first); a pivot that falls below 1e-12 drops its update (coefficient 0), so updates
// nearly in the span of the ones already taken add nothing.  Every loop has a fixed trip count
// (registers only); all zeros when G has a diagonal that is not positive and finite, or when a
// coefficient is not finite.
__device__ inline void warm_coeffs(int m, const double* __restrict__ D, double (&y)[TF_WARM_HIST])
{
    constexpr int H = TF_WARM_HIST;
    double A[H][H], s[H], b[H], L[H][H], w[H], v[H];
    bool take[H];
    int q = 0;
#pragma unroll
    for (int i = 0; i < H; ++i)
#pragma unroll
        for (int j = i; j < H; ++j, ++q) A[i][j] = A[j][i] = D[q];
    bool ok = true;
#pragma unroll
    for (int i = 0; i < H; ++i) {
        const bool in = i < m;
        ok = ok && (!in || (A[i][i] > 0.0 && isfinite(A[i][i])));
        s[i] = in && A[i][i] > 0.0 ? 1.0 / sqrt(A[i][i]) : 0.0;
        b[i] = s[i] * D[q + i];
    }
#pragma unroll
    for (int i = 0; i < H; ++i)
#pragma unroll
        for (int j = 0; j < H; ++j) { A[i][j] *= s[i] * s[j]; L[i][j] = 0.0; }
#pragma unroll
    for (int k = 0; k < H; ++k) {
        double d = A[k][k];
#pragma unroll
        for (int j = 0; j < k; ++j) d -= L[k][j] * L[k][j];
        take[k] = ok && k < m && d > 1e-12 && isfinite(d);
        const double lkk = take[k] ? sqrt(d) : 1.0;
        L[k][k] = lkk;
#pragma unroll
        for (int i = k + 1; i < H; ++i) {
            double a = A[i][k];
#pragma unroll
            for (int j = 0; j < k; ++j) a -= L[i][j] * L[k][j];
            L[i][k] = take[k] ? a / lkk : 0.0;
        }
    }
#pragma unroll
    for (int k = 0; k < H; ++k) {  // L w = b over the updates taken
        double a = b[k];
#pragma unroll
        for (int j = 0; j < k; ++j) a -= L[k][j] * w[j];
        w[k] = take[k] ? a / L[k][k] : 0.0;
    }
    bool fin = true;
#pragma unroll
    for (int k = H - 1; k >= 0; --k) {  // L^T v = w
        double a = w[k];
#pragma unroll
        for (int i = k + 1; i < H; ++i) a -= L[i][k] * v[i];
        v[k] = take[k] ? a / L[k][k] : 0.0;
        y[k] = s[k] * v[k];
        fin = fin && isfinite(y[k]);
    }
    if (!fin)
#pragma unroll
        for (int k = 0; k < H; ++k) y[k] = 0.0;
}

// x0 = sum_j y_j z_j on every row of the tile, y from the tissue's Gram entries (warm_coeffs, the
// same in every tile of the tissue)
__global__ __launch_bounds__(WG) void k_warm_combine(TileArgs ta, WarmArgs wa, const double* __restrict__ dots,
                                                     double* __restrict__ x0)
{
    const int t = blockIdx.x;
    const int sg = ta.tile_seg[t];
    const int d0 = 3 * ta.tile_r0[t], d1 = 3 * ta.tile_r1[t];
    double c[TF_WARM_HIST];
    warm_coeffs(wa.m, dots + (size_t)sg * TF_WARM_DOTS, c);
    for (int d = d0 + threadIdx.x; d < d1; d += WG) {
        double v = 0.0;
#pragma unroll
        for (int j = 0; j < TF_WARM_HIST; ++j)
            if (j < wa.m) v += c[j] * wa.z[j][d];
        x0[d] = v;
    }
}

}  // namespace

bool warm_start_begin(tf_handle* h, int k, bool launch)
{
    const int m = h->warm_count[k];
    if (m <= 0) return false;
    const size_t n3 = 3 * (size_t)h->N;
    if (!h->warm_x0) {
        TF_HIP(hipMalloc((void**)&h->warm_x0, n3 * sizeof(double)));
        TF_HIP(hipMalloc((void**)&h->warm_seg, sizeof(double) * TF_WARM_DOTS * h->S));
    }
    if (!h->warm_part || h->warm_part_T != h->T) {  // tiles change with the tiling
        if (h->warm_part) (void)hipFree(h->warm_part);
        TF_HIP(hipMalloc((void**)&h->warm_part, sizeof(double) * TF_WARM_DOTS * h->T));
        h->warm_part_T = h->T;
    }
    if (launch) warm_start_launch(h, k);
    return true;
}

void warm_start_launch(tf_handle* h, int k)
{
    TileArgs ta = tile_args(h);
    WarmArgs wa{};
    wa.z = h->warm_tab + (size_t)k * TF_WARM_HIST;
    wa.g = h->g;
    wa.m = h->warm_count[k];
    wa.cap = h->tile_cap;
    // one pass over the blocks when the products of every kept update fit in 64 KB of LDS
    const size_t lds1 = 3 * sizeof(double) * (size_t)h->tile_cap;
    const bool one = TF_WARM_HIST * lds1 <= 65536;
    const size_t lds = one ? TF_WARM_HIST * lds1 : lds1;
    if (fp32_values(h)) {
        if (one) hipLaunchKernelGGL((k_warm_gram<float, true>), dim3(h->T), dim3(WG), lds, h->stream, ta, wa, h->row_ptr,
                                    h->col, h->val32, h->nb, h->warm_part);
        else hipLaunchKernelGGL((k_warm_gram<float, false>), dim3(h->T), dim3(WG), lds, h->stream, ta, wa, h->row_ptr,
                                h->col, h->val32, h->nb, h->warm_part);
    } else {
        if (one) hipLaunchKernelGGL((k_warm_gram<double, true>), dim3(h->T), dim3(WG), lds, h->stream, ta, wa, h->row_ptr,
                                    h->col, h->val, h->nb, h->warm_part);
        else hipLaunchKernelGGL((k_warm_gram<double, false>), dim3(h->T), dim3(WG), lds, h->stream, ta, wa, h->row_ptr,
                                h->col, h->val, h->nb, h->warm_part);
    }
    hipLaunchKernelGGL(k_warm_seg, dim3(h->S, TF_WARM_DOTS), dim3(64), 0, h->stream, ta, h->warm_part, h->warm_seg);
    hipLaunchKernelGGL(k_warm_combine, dim3(h->T), dim3(WG), 0, h->stream, ta, wa, h->warm_seg, h->warm_x0);
    TF_HIP(hipGetLastError());
}

void warm_start_push(tf_handle* h, int k)
{
    const size_t n3 = 3 * (size_t)h->N;
    if (!h->warm_tab) {  // every buffer of the history in one allocation, at the first push
        TF_HIP(hipMalloc((void**)&h->warm_tab, sizeof(double*) * TF_WARM_ITERS * TF_WARM_HIST));
        TF_HIP(hipHostMalloc((void**)&h->hp_warm_tab, sizeof(double*) * TF_WARM_ITERS * TF_WARM_HIST,
                             hipHostMallocDefault));
        TF_HIP(hipMalloc((void**)&h->warm_block, sizeof(double) * n3 * TF_WARM_ITERS * TF_WARM_HIST));
        for (int i = 0; i < TF_WARM_ITERS; ++i)
            for (int j = 0; j < TF_WARM_HIST; ++j) h->warm_hist[i][j] = h->warm_block + n3 * (i * TF_WARM_HIST + j);
    }
    double* last = h->warm_hist[k][TF_WARM_HIST - 1];
    for (int j = TF_WARM_HIST - 1; j > 0; --j) h->warm_hist[k][j] = h->warm_hist[k][j - 1];
    h->warm_hist[k][0] = last;  // the oldest buffer is reused for the newest update
    launch_copy(h, last, h->du, (int64_t)n3);
    h->warm_count[k] = std::min(TF_WARM_HIST, h->warm_count[k] + 1);
    // the device table the warm-start kernels read (pinned staging: row k is rewritten only at the
    // next push of iteration k, a timestep later, after several host waits)
    double** row = h->hp_warm_tab + (size_t)k * TF_WARM_HIST;
    for (int j = 0; j < TF_WARM_HIST; ++j) row[j] = h->warm_hist[k][j];
    TF_HIP(hipMemcpyAsync(h->warm_tab + (size_t)k * TF_WARM_HIST, row, sizeof(double*) * TF_WARM_HIST,
                          hipMemcpyHostToDevice, h->stream));
}

void warm_start_clear(tf_handle* h)
{
    for (int k = 0; k < TF_WARM_ITERS; ++k) h->warm_count[k] = 0;
}

void warm_start_free(tf_handle* h)
{
    for (int k = 0; k < TF_WARM_ITERS; ++k)
        for (int j = 0; j < TF_WARM_HIST; ++j) h->warm_hist[k][j] = nullptr;
    for (double** p : {&h->warm_block, &h->warm_x0, &h->warm_part, &h->warm_seg})
        if (*p) { (void)hipFree(*p); *p = nullptr; }
    if (h->warm_tab) { (void)hipFree(h->warm_tab); h->warm_tab = nullptr; }
    if (h->hp_warm_tab) { (void)hipHostFree(h->hp_warm_tab); h->hp_warm_tab = nullptr; }
    warm_start_clear(h);
}

// ---- two-level preconditioner: coarse space and coarse operator -------------------------------
namespace {

template <typename T>
T* tl_alloc(size_t n)
{
    T* p = nullptr;
    TF_HIP(hipMalloc((void**)&p, std::max<size_t>(1, n) * sizeof(T)));
    return p;
}

template <typename T>
void tl_upload(tf_handle* h, T*& dst, const std::vector<T>& v)
{
    if (dst) { (void)hipFree(dst); dst = nullptr; }
    dst = tl_alloc<T>(v.size());
    if (!v.empty()) TF_HIP(hipMemcpyAsync(dst, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, h->stream));
}

void tl_release(tf_handle* h)
{
    tf_handle::TwoLevel& tl = h->tl;
    for (int32_t** p : {&tl.node_agg, &tl.agg_ptr, &tl.agg_nodes, &tl.coff, &tl.krank, &tl.pair_a,
                        &tl.pair_b, &tl.pair_ptr, &tl.pair_blk, &tl.info, &tl.col_blk, &tl.prow_ptr, &tl.pcol,
                        &tl.pblk, &tl.pass_ptr, &tl.tp_seg_ptr, &tl.agg_seg, &tl.seg_c0, &tl.seg_ldc})
        if (*p) { (void)hipFree(*p); *p = nullptr; }
    if (tl.wg_rec) { (void)hipFree(tl.wg_rec); tl.wg_rec = nullptr; }
    if (tl.seg_acoff) { (void)hipFree(tl.seg_acoff); tl.seg_acoff = nullptr; }
    if (tl.refresh) { (void)hipFree(tl.refresh); tl.refresh = nullptr; }
    for (double** p : {&tl.pcoef, &tl.Ac, &tl.swP, &tl.swO, &tl.swD, &tl.pcoef_p, &tl.val_p,
                       &tl.tpart, &tl.ptq, &tl.rcx})
        if (*p) { (void)hipFree(*p); *p = nullptr; }
    if (tl.pass_rows) { (void)hipFree(tl.pass_rows); tl.pass_rows = nullptr; }
    if (tl.val32_p) { (void)hipFree(tl.val32_p); tl.val32_p = nullptr; }
    if (tl.Ainv) { (void)hipFree(tl.Ainv); tl.Ainv = nullptr; }
    if (tl.ascale) { (void)hipFree(tl.ascale); tl.ascale = nullptr; }
    if (tl.dcinv) { (void)hipFree(tl.dcinv); tl.dcinv = nullptr; }
    if (tl.dcinv64) { (void)hipFree(tl.dcinv64); tl.dcinv64 = nullptr; }
    if (tl.Ainv64) { (void)hipFree(tl.Ainv64); tl.Ainv64 = nullptr; }
    if (tl.seg_aoff64) { (void)hipFree(tl.seg_aoff64); tl.seg_aoff64 = nullptr; }
    for (int32_t** p : {&tl.pslot, &tl.halo_ptr, &tl.halo_pos})
        if (*p) { (void)hipFree(*p); *p = nullptr; }
    if (tl.bar) { (void)hipFree(tl.bar); tl.bar = nullptr; }
    tl.run_ok = false;
    for (int32_t& c : tl.run_cap) c = -1;
    tl.built = tl.fresh = tl.failed = false;
}

// Aggregates: the points' (x, y) (one point per node column) in square cells of side
// sqrt(area / target), so that every aggregate holds whole columns of the tissue's layers; cells
// under min_count points join the nearest (by centroid) cell that has min_count or more.
// Numbered in the order of their cell keys.
std::vector<int32_t> tl_aggregates(const std::vector<double>& x, int N, int target, int min_count, int& n_agg)
{
    double lo[2] = {INFINITY, INFINITY}, hi[2] = {-INFINITY, -INFINITY};
    for (int n = 0; n < N; ++n)
        for (int c = 0; c < 2; ++c) { lo[c] = std::min(lo[c], x[3 * n + c]); hi[c] = std::max(hi[c], x[3 * n + c]); }
    const double area = std::max(hi[0] - lo[0], 1e-12) * std::max(hi[1] - lo[1], 1e-12);
    const double side = std::sqrt(area / std::max(1, target));
    std::vector<int64_t> key(N);
    for (int n = 0; n < N; ++n) {
        const int64_t cx = (int64_t)std::floor((x[3 * n] - lo[0]) / side);
        const int64_t cy = (int64_t)std::floor((x[3 * n + 1] - lo[1]) / side);
        key[n] = cx * 1000003 + cy;
    }
    std::vector<int64_t> keys(key);
    std::sort(keys.begin(), keys.end());
    keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
    const int K = (int)keys.size();
    std::vector<int32_t> lab(N), cnt(K, 0);
    std::vector<double> cen(2 * K, 0.0);
    for (int n = 0; n < N; ++n) {
        lab[n] = (int32_t)(std::lower_bound(keys.begin(), keys.end(), key[n]) - keys.begin());
        cnt[lab[n]] += 1;
        cen[2 * lab[n]] += x[3 * n];
        cen[2 * lab[n] + 1] += x[3 * n + 1];
    }
    for (int k = 0; k < K; ++k) { cen[2 * k] /= cnt[k]; cen[2 * k + 1] /= cnt[k]; }
    std::vector<int32_t> remap(K);
    bool any_big = false;
    for (int k = 0; k < K; ++k) any_big = any_big || cnt[k] >= min_count;
    for (int k = 0; k < K; ++k) {
        remap[k] = k;
        if (!any_big) { remap[k] = 0; continue; }
        if (cnt[k] >= min_count) continue;
        double best = INFINITY;
        for (int j = 0; j < K; ++j) {
            if (cnt[j] < min_count) continue;
            const double dx = cen[2 * j] - cen[2 * k], dy = cen[2 * j + 1] - cen[2 * k + 1];
            const double d = dx * dx + dy * dy;
            if (d < best) { best = d; remap[k] = j; }
        }
    }
    std::vector<int32_t> id(K, -1);
    n_agg = 0;
    for (int k = 0; k < K; ++k)
        if (remap[k] == k) id[k] = n_agg++;
    for (int n = 0; n < N; ++n) lab[n] = id[remap[lab[n]]];
    return lab;
}

// Orthonormal basis of the rigid-body modes (3 translations, 3 rotations about the centroid)
// restricted to the aggregate's nodes: modified Gram-Schmidt run twice, a mode whose remainder
// falls under 1e-10 of its norm is dropped.  DoFs flagged in excl (fixed DoFs and bound slaves,
// whose rows and columns of M are unit ones) get zero rows, so that the correction never moves
// them.  Writes rows 3n..3n+2 of the kept columns into pcoef[18 n ..] and returns the rank.
int tl_modes(const std::vector<double>& x, const std::vector<uint8_t>& excl, const std::vector<int32_t>& nodes,
             std::vector<double>& pcoef)
{
    const size_t m = 3 * nodes.size();
    double c[3] = {0.0, 0.0, 0.0};
    for (int n : nodes)
        for (int d = 0; d < 3; ++d) c[d] += x[3 * n + d];
    for (int d = 0; d < 3; ++d) c[d] /= (double)nodes.size();
    std::vector<std::vector<double>> B(6, std::vector<double>(m, 0.0));
    for (size_t k = 0; k < nodes.size(); ++k) {
        const int n = nodes[k];
        const double px = x[3 * n] - c[0], py = x[3 * n + 1] - c[1], pz = x[3 * n + 2] - c[2];
        for (int d = 0; d < 3; ++d) B[d][3 * k + d] = 1.0;
        B[3][3 * k] = -py; B[3][3 * k + 1] = px;      // rotation about z
        B[4][3 * k + 1] = -pz; B[4][3 * k + 2] = py;  // about x
        B[5][3 * k] = pz; B[5][3 * k + 2] = -px;      // about y
        for (int d = 0; d < 3; ++d)
            if (excl[3 * n + d])
                for (int j = 0; j < 6; ++j) B[j][3 * k + d] = 0.0;
    }
    std::vector<std::vector<double>> Q;
    for (int j = 0; j < 6; ++j) {
        std::vector<double> v = B[j];
        double n0 = 0.0;
        for (double a : v) n0 += a * a;
        n0 = std::sqrt(n0);
        if (!(n0 > 0.0)) continue;
        for (int pass = 0; pass < 2; ++pass)
            for (const auto& q : Q) {
                double d = 0.0;
                for (size_t i = 0; i < m; ++i) d += q[i] * v[i];
                for (size_t i = 0; i < m; ++i) v[i] -= d * q[i];
            }
        double nv = 0.0;
        for (double a : v) nv += a * a;
        nv = std::sqrt(nv);
        if (nv <= 1e-10 * n0) continue;
        for (double& a : v) a /= nv;
        Q.push_back(std::move(v));
    }
    for (size_t k = 0; k < nodes.size(); ++k)
        for (int i = 0; i < 3; ++i)
            for (size_t j = 0; j < Q.size(); ++j) pcoef[18 * (size_t)nodes[k] + 6 * i + j] = Q[j][3 * k + i];
    return (int)Q.size();
}

}  // namespace

// Coarse space of the current positions and pattern: aggregates, their modes and node lists, and
// the aggregate pairs that share Jacobian blocks.
static void twolevel_build(tf_handle* h)
{
    tf_handle::TwoLevel& tl = h->tl;
    drop_krylov_graphs(h);  // captured chunks hold the coarse arrays freed below
    tl_release(h);
    const int N = h->N;
    std::vector<double> x(3 * (size_t)N);
    TF_HIP(hipMemcpyAsync(x.data(), h->uk, x.size() * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    std::vector<uint8_t> excl(3 * (size_t)N);
    TF_HIP(hipMemcpyAsync(excl.data(), h->fixed, excl.size(), hipMemcpyDeviceToHost, h->stream));
    TF_HIP(hipStreamSynchronize(h->stream));
    for (size_t k = 0; k < h->h_bound.size(); k += 2) excl[h->h_bound[k]] = 1;
    // k_tp_z stages a tissue's coarse vector (ldA floats, nc_s rounded up to 8) in LDS beside ~4 KB of
    // its own
    const int max_nc = std::min(4096, (h->max_lds - 4096) / (2 * (int)sizeof(float)) / 8 * 8);
    // default: one aggregate per 44 nodes of the tissue, at most 600; a handle of one tissue also
    // keeps at most two aggregates per compute unit: k_tp_z (one workgroup per aggregate) holds two
    // workgroups per compute unit (its registers), so more would take a second round of workgroups
    const bool by_default = h->opts.coarse_aggregates <= 0;
    const int S = h->S;
    const int max_agg = by_default && S == 1 ? std::min(max_nc / 6, 2 * h->n_cu) : max_nc / 6;
    // Node columns: nodes joined by the vertical edges of prisms (corners k and k + 3), each
    // listed in ascending node order.  Used when column_blocks is on and every column has the same
    // number of nodes L <= TL_MAX_LEVELS; otherwise every node is a column of its own (L = 1,
    // 3 x 3 blocks).
    int L = 1;
    std::vector<int32_t> root(N);
    {
        std::iota(root.begin(), root.end(), 0);
        auto find = [&](int a) {
            while (root[a] != a) { root[a] = root[root[a]]; a = root[a]; }
            return a;
        };
        const std::vector<int32_t>& cn = h->h_conn;
        for (int e = 0; e < h->E; ++e)
            for (int k = 0; k < 3; ++k) {
                const int u = find(cn[6 * e + k]), v = find(cn[6 * e + k + 3]);
                if (u != v) root[std::max(u, v)] = std::min(u, v);  // a set's root is its smallest node
            }
        for (int n = 0; n < N; ++n) root[n] = find(n);
        std::vector<int32_t> cnt(N, 0);
        for (int n = 0; n < N; ++n) cnt[root[n]] += 1;
        L = h->opts.column_blocks ? cnt[root[0]] : 0;
        for (int n = 0; n < N && L > 0; ++n)
            if (cnt[root[n]] != L) L = 0;
        if (L < 1 || L > TL_MAX_LEVELS) {
            L = 1;
            std::iota(root.begin(), root.end(), 0);
        }
    }
    std::vector<int32_t> col_of(N), col_start;  // column of every node; first node of every column
    for (int n = 0; n < N; ++n)
        if (root[n] == n) { col_of[n] = (int32_t)col_start.size(); col_start.push_back(n); }
    for (int n = 0; n < N; ++n) col_of[n] = col_of[root[n]];
    const int ncol = (int)col_start.size();
    std::vector<std::vector<int32_t>> col_nodes(ncol);
    for (int n = 0; n < N; ++n) col_nodes[col_of[n]].push_back(n);
    // aggregates of whole columns, tissue by tissue, placed by each column's first node; cells under
    // 16 nodes merge.  A column lies in one tissue (its nodes share prisms), and the columns of a
    // tissue are consecutive (numbered by first node).
    std::vector<int32_t> aggc(ncol), agg_seg, seg_agg_ptr(S + 1, 0);
    {
        int c = 0;
        for (int sg = 0; sg < S; ++sg) {
            const int n1 = h->h_node_seg_ptr[sg + 1], Ns = n1 - h->h_node_seg_ptr[sg];
            const int cc0 = c;
            while (c < ncol && col_start[c] < n1) ++c;
            const int ncs = c - cc0;
            std::vector<double> xc(3 * (size_t)ncs);
            for (int k = 0; k < ncs; ++k)
                for (int d = 0; d < 3; ++d) xc[3 * k + d] = x[3 * (size_t)col_start[cc0 + k] + d];
            // (the aggregates cover the tissue's part of a grid over its bounding box, so a target
            // above max_agg can still give at most max_agg aggregates; the loop shrinks it when
            // not, down to one aggregate at target 1)
            int target = by_default ? std::max(1, std::min(600, Ns / 44)) : h->opts.coarse_aggregates;
            target = std::min(target, max_nc / 6);
            int As = 0;
            std::vector<int32_t> lab;
            if (ncs > 0) {
                while (true) {
                    lab = tl_aggregates(xc, ncs, target, std::max(1, 16 / L), As);
                    if (As <= max_agg || target <= 1) break;
                    target = std::max(1, std::min(target - 1, (int)(0.9 * target * max_agg / As)));
                }
            }
            const int A0 = seg_agg_ptr[sg];
            for (int k = 0; k < ncs; ++k) aggc[cc0 + k] = A0 + lab[k];
            seg_agg_ptr[sg + 1] = A0 + As;
            agg_seg.insert(agg_seg.end(), As, sg);
        }
        if (c != ncol) throw std::runtime_error("two-level preconditioner: node columns out of tissue order");
    }
    const int A = seg_agg_ptr[S];
    std::vector<int32_t> agg(N);
    for (int n = 0; n < N; ++n) agg[n] = aggc[col_of[n]];
    std::vector<std::vector<int32_t>> members(A);  // column by column
    for (int c = 0; c < ncol; ++c) members[aggc[c]].insert(members[aggc[c]].end(), col_nodes[c].begin(), col_nodes[c].end());
    std::vector<double> pcoef(18 * (size_t)N, 0.0);
    std::vector<int32_t> krank(A), coff(A + 1, 0);
    for (int a = 0; a < A; ++a) {
        krank[a] = tl_modes(x, excl, members[a], pcoef);
        coff[a + 1] = coff[a] + krank[a];
    }
    const int nc = coff[A];
    // every tissue's block: coarse unknowns [C0, C0 + nc_s), Ac block of stride ldc_s (a multiple of
    // the sweep's tile), FP16 inverse of stride ldA_s (a multiple of 8; one row of zeros when the
    // tissue has no coarse unknown)
    tl.h_seg_c0.assign(S + 1, 0);
    tl.h_seg_ldc.assign(S, 0);
    tl.h_seg_lda.assign(S, 0);
    tl.h_seg_acoff.assign(S + 1, 0);
    tl.h_seg_aoff.assign(S + 1, 0);
    std::vector<int4> seg_coarse(S);
    tl.max_ldA = 8;
    int max_ldc = SW_B;
    for (int sg = 0; sg < S; ++sg) {
        const int C0 = coff[seg_agg_ptr[sg]], C1 = coff[seg_agg_ptr[sg + 1]], ncs = C1 - C0;
        if (ncs > max_nc) throw std::runtime_error("two-level preconditioner: coarse space exceeds the LDS of a workgroup");
        tl.h_seg_c0[sg + 1] = C1;
        tl.h_seg_ldc[sg] = (ncs + SW_B - 1) / SW_B * SW_B;
        tl.h_seg_lda[sg] = std::max(8, (ncs + 7) / 8 * 8);
        tl.h_seg_acoff[sg + 1] = tl.h_seg_acoff[sg] + (int64_t)tl.h_seg_ldc[sg] * tl.h_seg_ldc[sg];
        tl.h_seg_aoff[sg + 1] = tl.h_seg_aoff[sg] + (int64_t)std::max(ncs, 1) * tl.h_seg_lda[sg];
        if (tl.h_seg_aoff[sg] / 2 > INT32_MAX) throw std::runtime_error("two-level preconditioner: coarse blocks too large");
        seg_coarse[sg] = int4{C0, ncs, tl.h_seg_lda[sg], (int32_t)(tl.h_seg_aoff[sg] / 2)};
        tl.max_ldA = std::max(tl.max_ldA, tl.h_seg_lda[sg]);
        max_ldc = std::max(max_ldc, tl.h_seg_ldc[sg]);
    }
    tl.ainv_halves = tl.h_seg_aoff[S];
    std::vector<int32_t> agg_ptr(A + 1, 0), agg_nodes;
    agg_nodes.reserve(N);
    for (int a = 0; a < A; ++a) {
        agg_nodes.insert(agg_nodes.end(), members[a].begin(), members[a].end());
        agg_ptr[a + 1] = (int32_t)agg_nodes.size();
    }
    // BSR block of every node pair of every column, in the columns' order in agg_nodes (-1: none)
    std::vector<int32_t> col_blk((size_t)ncol * L * L, -1);
    for (int g = 0; g < ncol; ++g)
        for (int i = 0; i < L; ++i) {
            const int ni = agg_nodes[(size_t)g * L + i];
            for (int j = 0; j < L; ++j) {
                const int nj = agg_nodes[(size_t)g * L + j];
                for (int b = h->h_row_ptr[ni]; b < h->h_row_ptr[ni + 1]; ++b)
                    if (h->h_col[b] == nj) { col_blk[((size_t)g * L + i) * L + j] = b; break; }
            }
        }
    // aggregate pairs with their Jacobian blocks in block order (the pairs of a tissue's aggregates:
    // M couples no two tissues), numbered in increasing (a, b) order through a dense table per tissue
    const std::vector<int32_t>& rp = h->h_row_ptr;
    const std::vector<int32_t>& cl = h->h_col;
    std::vector<int64_t> tab_off(S + 1, 0);
    for (int sg = 0; sg < S; ++sg) {
        const int64_t As = seg_agg_ptr[sg + 1] - seg_agg_ptr[sg];
        tab_off[sg + 1] = tab_off[sg] + As * As;
    }
    auto key_of = [&](int a, int b) {
        const int sg = agg_seg[a], A0 = seg_agg_ptr[sg], As = seg_agg_ptr[sg + 1] - A0;
        return tab_off[sg] + (int64_t)(a - A0) * As + (b - A0);
    };
    std::vector<int64_t> pkey;
    pkey.reserve(h->nb);
    for (int r = 0; r < N; ++r)
        for (int j = rp[r]; j < rp[r + 1]; ++j) pkey.push_back(key_of(agg[r], agg[cl[j]]));
    std::vector<int32_t> pid(tab_off[S], -1);
    for (int64_t k : pkey) pid[k] = 0;
    int P = 0;
    for (int32_t& v : pid)
        if (v >= 0) v = P++;
    std::vector<int32_t> pair_a(P), pair_b(P), pair_ptr(P + 1, 0), pair_blk(pkey.size());
    std::vector<int32_t> bpair(pkey.size());
    for (size_t j = 0; j < pkey.size(); ++j) {
        bpair[j] = pid[pkey[j]];
        pair_ptr[bpair[j] + 1] += 1;
    }
    for (int sg = 0; sg < S; ++sg) {
        const int A0 = seg_agg_ptr[sg], As = seg_agg_ptr[sg + 1] - A0;
        for (int64_t k = 0; k < (int64_t)As * As; ++k) {
            const int32_t q = pid[tab_off[sg] + k];
            if (q >= 0) {
                pair_a[q] = A0 + (int32_t)(k / As);
                pair_b[q] = A0 + (int32_t)(k % As);
            }
        }
    }
    for (int q = 0; q < P; ++q) pair_ptr[q + 1] += pair_ptr[q];
    {
        std::vector<int32_t> fill(pair_ptr.begin(), pair_ptr.end() - 1);
        for (size_t j = 0; j < pkey.size(); ++j) pair_blk[fill[bpair[j]]++] = (int32_t)j;
    }
    tl_upload(h, tl.node_agg, agg);
    tl_upload(h, tl.agg_ptr, agg_ptr);
    tl_upload(h, tl.agg_nodes, agg_nodes);
    tl_upload(h, tl.pcoef, pcoef);
    tl_upload(h, tl.coff, coff);
    tl_upload(h, tl.krank, krank);
    tl_upload(h, tl.pair_a, pair_a);
    tl_upload(h, tl.pair_b, pair_b);
    tl_upload(h, tl.pair_ptr, pair_ptr);
    tl_upload(h, tl.pair_blk, pair_blk);
    tl_upload(h, tl.col_blk, col_blk);
    // the solve in aggregate order: position r holds node agg_nodes[r]; the Jacobian's rows follow
    // the positions (a row keeps its blocks in column order), columns become positions
    {
        std::vector<int32_t> pos_of(N);
        for (int r = 0; r < N; ++r) pos_of[agg_nodes[r]] = r;
        std::vector<int32_t> prow(N + 1, 0), pcol(h->nb), pblk(h->nb);
        for (int r = 0; r < N; ++r) {
            const int n = agg_nodes[r];
            prow[r + 1] = prow[r] + (h->h_row_ptr[n + 1] - h->h_row_ptr[n]);
            for (int j = h->h_row_ptr[n], k = prow[r]; j < h->h_row_ptr[n + 1]; ++j, ++k) {
                pblk[k] = j;
                pcol[k] = pos_of[h->h_col[j]];
            }
        }
        // passes: runs of an aggregate's rows with at most WG / 3 rows and TP_CAP blocks
        std::vector<int32_t> pass_ptr(A + 1, 0);
        std::vector<int2> pass_rows;
        tl.tp_max_blocks = 1;
        for (int a = 0; a < A; ++a) {
            int r = agg_ptr[a];
            while (r < agg_ptr[a + 1]) {
                int r1 = r;
                while (r1 < agg_ptr[a + 1] && r1 - r < WG / 3 && prow[r1 + 1] - prow[r] <= TP_CAP) ++r1;
                if (r1 == r) throw std::runtime_error("two-level preconditioner: a block row exceeds a pass");
                pass_rows.push_back(int2{r, r1});
                tl.tp_max_blocks = std::max(tl.tp_max_blocks, prow[r1] - prow[r]);
                r = r1;
            }
            pass_ptr[a + 1] = (int32_t)pass_rows.size();
        }
        std::vector<double> pcoef_p(18 * (size_t)N);
        for (int r = 0; r < N; ++r)
            std::copy(pcoef.begin() + 18 * (size_t)agg_nodes[r], pcoef.begin() + 18 * (size_t)agg_nodes[r] + 18,
                      pcoef_p.begin() + 18 * (size_t)r);
        tl_upload(h, tl.prow_ptr, prow);
        tl_upload(h, tl.pcol, pcol);
        tl_upload(h, tl.pblk, pblk);
        tl_upload(h, tl.pass_ptr, pass_ptr);
        tl_upload(h, tl.pass_rows, pass_rows);
        // one record per workgroup, the aggregates by descending block count
        std::vector<int32_t> order(A);
        std::iota(order.begin(), order.end(), 0);
        std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
            return prow[agg_ptr[x + 1]] - prow[agg_ptr[x]] > prow[agg_ptr[y + 1]] - prow[agg_ptr[y]];
        });
        std::vector<TPRec> rec(A);
        for (int k = 0; k < A; ++k) {
            const int a = order[k], sg = agg_seg[a];
            const int2 r = pass_rows[pass_ptr[a]];
            if (seg_agg_ptr[sg + 1] - seg_agg_ptr[sg] > TP_SU * WG)
                throw std::runtime_error("two-level preconditioner: more aggregates in a tissue than a workgroup sums");
            rec[k] = TPRec{a, sg, seg_agg_ptr[sg], seg_agg_ptr[sg + 1], r.x, r.y, prow[r.x], prow[r.y],
                           pass_ptr[a], pass_ptr[a + 1], agg_ptr[a], agg_ptr[a + 1], coff[a], coff[a + 1] - coff[a],
                           0, 0, seg_coarse[sg]};
        }
        std::vector<int4> rec4(5 * (size_t)A);
        std::memcpy(rec4.data(), rec.data(), sizeof(TPRec) * (size_t)A);
        tl_upload(h, tl.wg_rec, rec4);
        tl_upload(h, tl.pcoef_p, pcoef_p);
        tl_upload(h, tl.tp_seg_ptr, seg_agg_ptr);
        // the persistent solve (k_tp_run): every aggregate one pass with at most WG / 3L columns and
        // TPR_MAXU own and neighbouring nodes; the column slot of every block; one tissue
        {
            bool ok = S == 1 && A >= TPR_NG && A <= TPR_SU * WG && tl.max_ldA <= WG * TPR_NV;
            std::vector<int32_t> pslot(h->nb, 0), hptr(A + 1, 0), hpos;
            for (int a = 0; a < A; ++a) {
                const int n0 = agg_ptr[a], n1 = agg_ptr[a + 1], nown = n1 - n0;
                if (pass_ptr[a + 1] - pass_ptr[a] != 1 || prow[n1] - prow[n0] > TP_PF * WG || nown / L > WG / (3 * L))
                    ok = false;
                std::vector<int32_t> halo;
                for (int j = prow[n0]; j < prow[n1]; ++j)
                    if (pcol[j] < n0 || pcol[j] >= n1) halo.push_back(pcol[j]);
                std::sort(halo.begin(), halo.end());
                halo.erase(std::unique(halo.begin(), halo.end()), halo.end());
                if (nown + (int)halo.size() > TPR_MAXU) ok = false;
                for (int j = prow[n0]; j < prow[n1]; ++j) {
                    const int c = pcol[j];
                    pslot[j] = c >= n0 && c < n1 ? c - n0
                                                 : nown + (int)(std::lower_bound(halo.begin(), halo.end(), c) - halo.begin());
                }
                hpos.insert(hpos.end(), halo.begin(), halo.end());
                hptr[a + 1] = (int32_t)hpos.size();
            }
            tl.run_ok = ok;
            if (ok) {
                tl_upload(h, tl.pslot, pslot);
                tl_upload(h, tl.halo_ptr, hptr);
                tl_upload(h, tl.halo_pos, hpos);
                tl.bar = tl_alloc<unsigned>(TPR_WORDS);
            }
        }
        tl.tpart = tl_alloc<double>((size_t)TF_RED_SLOTS * A);
        // (a tissue without coarse unknowns reads one clamped entry at its C0, which may be nc)
        tl.ptq = tl_alloc<double>(nc + 8);
        tl.rcx = tl_alloc<double>(2 * (size_t)nc + 8);
        TF_HIP(hipMemsetAsync(tl.ptq, 0, (nc + 8) * sizeof(double), h->stream));
        TF_HIP(hipMemsetAsync(tl.rcx, 0, (2 * (size_t)nc + 8) * sizeof(double), h->stream));
        TF_HIP(hipStreamSynchronize(h->stream));  // the host vectors above are temporaries
    }
    const bool f64 = h->opts.precond_fp64 != 0;  // FP64 preconditioner storage
    if (f64) tl.dcinv64 = tl_alloc<double>((size_t)ncol * 9 * L * L);
    else tl.dcinv = tl_alloc<float>((size_t)ncol * 9 * L * L);
    tl.L = L;
    tl.ncol = ncol;
    tl.ldc = max_ldc;  // the sweep's padded size (the largest block)
    tl.Ac = tl_alloc<double>((size_t)tl.h_seg_acoff[S]);
    tl.swP = tl_alloc<double>((size_t)tl.ldc * SW_B);
    tl.swO = tl_alloc<double>((size_t)tl.ldc * SW_B);
    tl.swD = tl_alloc<double>(2 * (size_t)SW_B * SW_B);
    tl.ldA = tl.h_seg_lda[0];
    if (f64) {
        tl.Ainv64 = tl_alloc<double>((size_t)tl.ainv_halves);
        TF_HIP(hipMemsetAsync(tl.Ainv64, 0, (size_t)tl.ainv_halves * sizeof(double), h->stream));
        tl_upload(h, tl.seg_aoff64, tl.h_seg_aoff);
    } else {
        tl.Ainv = tl_alloc<__half>((size_t)tl.ainv_halves);
        TF_HIP(hipMemsetAsync(tl.Ainv, 0, (size_t)tl.ainv_halves * sizeof(__half), h->stream));
    }
    tl.ascale = tl_alloc<double>(2 * (size_t)S);
    TF_HIP(hipMemsetAsync(tl.ascale, 0, 2 * (size_t)S * sizeof(double), h->stream));
    tl.info = tl_alloc<int32_t>(S);
    tl_upload(h, tl.agg_seg, agg_seg);
    tl_upload(h, tl.seg_c0, tl.h_seg_c0);
    tl_upload(h, tl.seg_ldc, tl.h_seg_ldc);
    tl_upload(h, tl.seg_acoff, tl.h_seg_acoff);
    tl.refresh = tl_alloc<uint8_t>(S);
    tl.base_it1.assign(S, -1);
    tl.stale.assign(S, 0);
    TF_HIP(hipStreamSynchronize(h->stream));  // the host vectors above are temporaries
    tl.n_agg = A;
    tl.nc = nc;
    tl.n_pairs = P;
    tl.h_node_agg = agg;
    tl.built = true;
}

static TLArgs tl_args(tf_handle* h)
{
    tf_handle::TwoLevel& tl = h->tl;
    return TLArgs{tl.pcoef, tl.coff, tl.krank, tl.agg_seg, tl.seg_c0, tl.seg_ldc, tl.seg_acoff, tl.refresh};
}

// the aggregates as the tissues' partial-sum units
static TileArgs tp_tile_args(tf_handle* h)
{
    TileArgs ta = tile_args(h);
    ta.info = nullptr;
    ta.partial = h->tl.tpart;
    ta.seg_tile_ptr = h->tl.tp_seg_ptr;
    ta.T = h->tl.n_agg;
    ta.nxcd = 1;
    return ta;
}

static TPArgs tp_args(tf_handle* h)
{
    tf_handle::TwoLevel& tl = h->tl;
    return TPArgs{tl.agg_nodes, tl.prow_ptr, tl.pcol, tl.pass_rows, reinterpret_cast<const TPRec*>(tl.wg_rec),
                  tl.pcoef_p, tl.dcinv, tl.dcinv64, reinterpret_cast<const __half2*>(tl.Ainv),
                  tl.Ainv64, tl.seg_aoff64, tl.ascale, tl.ptq, tl.rcx, tl.nc, tl.n_agg, tl.bar, TPR_WORDS,
                  h->opts.precond_fp64 ? 0 : 1};
}

template <int P, bool INIT>
static void launch_tp_z(tf_handle* h, const TileArgs& ta, const TPArgs& tp, const double* p, const double* q)
{
    const dim3 grid(tp.n_agg), blk(WG);
    const bool f64 = h->opts.precond_fp64 != 0;
    const size_t lds = tp_z_lds(h->tl.max_ldA, f64);
#define TF_TPZ(LL)                                                                                                        \
    if (f64) hipLaunchKernelGGL((k_tp_z<P, INIT, LL, true>), grid, blk, lds, h->stream, ta, tp, p, q, h->kx, h->kr, h->ks); \
    else hipLaunchKernelGGL((k_tp_z<P, INIT, LL, false>), grid, blk, lds, h->stream, ta, tp, p, q, h->kx, h->kr, h->ks)
    switch (h->tl.L) {
    case 1: TF_TPZ(1); break;
    case 2: TF_TPZ(2); break;
    case 3: TF_TPZ(3); break;
    default: TF_TPZ(4); break;
    }
#undef TF_TPZ
}

// inverse column blocks of the current Jacobian (every solve)
static void launch_tl_dcol(tf_handle* h)
{
    tf_handle::TwoLevel& tl = h->tl;
    const int cpw = 64 / (3 * tl.L), waves = (tl.ncol + cpw - 1) / cpw;
    const dim3 grid((waves + WG / 64 - 1) / (WG / 64)), blk(WG);
#define TF_DCOL(LL)                                                                                                        \
    if (tl.dcinv64) hipLaunchKernelGGL((k_tl_dcol<LL, double>), grid, blk, 0, h->stream, tl.col_blk, tl.agg_nodes, h->val, \
                                       h->nb, h->dinv, tl.ncol, tl.dcinv64);                                           \
    else hipLaunchKernelGGL((k_tl_dcol<LL, float>), grid, blk, 0, h->stream, tl.col_blk, tl.agg_nodes, h->val, h->nb,   \
                            h->dinv, tl.ncol, tl.dcinv)
    switch (tl.L) {
    case 1: TF_DCOL(1); break;
    case 2: TF_DCOL(2); break;
    case 3: TF_DCOL(3); break;
    default: TF_DCOL(4); break;
    }
#undef TF_DCOL
    TF_HIP(hipGetLastError());
}

// Ac = P^T M P from the FP64 Jacobian and its inverse by the blocked Gauss-Jordan sweep (FP64, on
// the handle's stream, 2 launches per 64-wide pivot block), then the scaled FP16 copy; for every
// tissue with which[s] set, block by block (the other tissues keep theirs).  A block that is not
// positive definite (a pivot block with a pivot that is not positive and finite) leaves the
// two-level path off until the next refresh.
static void twolevel_refresh(tf_handle* h, const std::vector<uint8_t>& which)
{
    tf_handle::TwoLevel& tl = h->tl;
    const int S = h->S;
    TF_HIP(hipMemcpyAsync(tl.refresh, which.data(), S, hipMemcpyHostToDevice, h->stream));
    for (int sg = 0; sg < S; ++sg) {
        const int ldc = tl.h_seg_ldc[sg], ncs = tl.h_seg_c0[sg + 1] - tl.h_seg_c0[sg];
        if (!which[sg] || ncs == 0) continue;
        const size_t n2 = (size_t)ldc * ldc;
        hipLaunchKernelGGL(k_sw_clear, dim3((unsigned)((n2 + 255) / 256)), dim3(256), 0, h->stream,
                           tl.Ac + tl.h_seg_acoff[sg], ncs, ldc);
    }
    hipLaunchKernelGGL(k_tl_ac, dim3(tl.n_pairs), dim3(WG), 0, h->stream, tl_args(h), tl.pair_a, tl.pair_b, tl.pair_ptr,
                       tl.pair_blk, h->blk_row, h->col, h->val, h->nb, tl.Ac);
    std::vector<int32_t> bad(S, 0);
    for (int sg = 0; sg < S; ++sg) {
        const int ldc = tl.h_seg_ldc[sg], ncs = tl.h_seg_c0[sg + 1] - tl.h_seg_c0[sg], nt = ldc / SW_B;
        if (!which[sg] || ncs == 0) continue;
        double* As = tl.Ac + tl.h_seg_acoff[sg];
        hipLaunchKernelGGL(k_sw_first, dim3(1), dim3(256), 0, h->stream, As, ldc, tl.swD, tl.info + sg);
        for (int K = 0; K < nt; ++K) {
            const double* D = tl.swD + (size_t)(K & 1) * SW_B * SW_B;
            double* Dn = tl.swD + (size_t)((K + 1) & 1) * SW_B * SW_B;
            hipLaunchKernelGGL(k_sw_panel, dim3(nt), dim3(256), 0, h->stream, As, ldc, K, D, tl.swP, tl.swO);
            hipLaunchKernelGGL(k_sw_update, dim3(nt * (nt + 1) / 2), dim3(256), 0, h->stream, As, ldc, K, nt, tl.swP,
                               tl.swO, D, Dn, tl.info + sg);
        }
        // the FP16 copy is enqueued before the failure flags are read back: the device has work
        // while the host waits, and the copy is simply not used after a failure
        const size_t n2h = (size_t)ncs * tl.h_seg_lda[sg];
        if (tl.Ainv64) {
            hipLaunchKernelGGL(k_tl_symm64, dim3((unsigned)((n2h + 255) / 256)), dim3(256), 0, h->stream, As, ldc,
                               tl.Ainv64 + tl.h_seg_aoff[sg], ncs, tl.h_seg_lda[sg]);
        } else {
            hipLaunchKernelGGL(k_tl_scale, dim3(1), dim3(256), 0, h->stream, As, ncs, ldc, tl.ascale + 2 * sg);
            hipLaunchKernelGGL(k_tl_symm, dim3((unsigned)((n2h + 255) / 256)), dim3(256), 0, h->stream, As, ldc,
                               (__half*)tl.Ainv + tl.h_seg_aoff[sg], ncs, tl.h_seg_lda[sg], tl.ascale + 2 * sg);
        }
        TF_HIP(hipGetLastError());
        TF_HIP(hipMemcpyAsync(bad.data() + sg, tl.info + sg, sizeof(int32_t), hipMemcpyDeviceToHost, h->stream));
    }
    TF_HIP(hipGetLastError());
    TF_HIP(hipStreamSynchronize(h->stream));
    h->host_waits += 1;
    bool failed = false;
    for (int sg = 0; sg < S; ++sg) {
        if (!which[sg]) continue;
        failed = failed || bad[sg] != 0;
        tl.base_it1[sg] = -1;
        tl.stale[sg] = 0;
    }
    tl.fresh = true;
    tl.age = 0;
    tl.failed = failed;
}

// The coarse operator P^T M P of the current Jacobian (assembled again into a temporary) and the FP64
// inverse the last refresh left (-Ac^-1 in the lower tiles, read back symmetric).  Diagnostics.
int twolevel_export(tf_handle* h, int32_t* nc_out, double* Ac_out, double* Ainv_out)
{
    tf_handle::TwoLevel& tl = h->tl;
    if (!tl.built || !tl.fresh || tl.failed || h->S != 1) return -1;
    const int nc = tl.nc, ldc = tl.h_seg_ldc[0];
    if (nc_out) *nc_out = nc;
    const size_t n2 = (size_t)ldc * ldc;
    if (Ainv_out) {
        std::vector<double> A(n2);
        TF_HIP(hipMemcpyAsync(A.data(), tl.Ac, n2 * sizeof(double), hipMemcpyDeviceToHost, h->stream));
        TF_HIP(hipStreamSynchronize(h->stream));
        for (int i = 0; i < nc; ++i)
            for (int j = 0; j < nc; ++j)
                Ainv_out[(size_t)i * nc + j] = -(i >= j ? A[(size_t)i * ldc + j] : A[(size_t)j * ldc + i]);
    }
    if (Ac_out) {
        double* tmp = tl_alloc<double>(n2);
        hipLaunchKernelGGL(k_sw_clear, dim3((unsigned)((n2 + 255) / 256)), dim3(256), 0, h->stream, tmp, nc, ldc);
        const uint8_t one = 1;
        TF_HIP(hipMemcpyAsync(tl.refresh, &one, 1, hipMemcpyHostToDevice, h->stream));
        hipLaunchKernelGGL(k_tl_ac, dim3(tl.n_pairs), dim3(WG), 0, h->stream, tl_args(h), tl.pair_a, tl.pair_b,
                           tl.pair_ptr, tl.pair_blk, h->blk_row, h->col, h->val, h->nb, tmp);
        std::vector<double> A(n2);
        TF_HIP(hipMemcpyAsync(A.data(), tmp, n2 * sizeof(double), hipMemcpyDeviceToHost, h->stream));
        TF_HIP(hipStreamSynchronize(h->stream));
        (void)hipFree(tmp);
        for (int i = 0; i < nc; ++i)
            for (int j = 0; j < nc; ++j) Ac_out[(size_t)i * nc + j] = A[(size_t)i * ldc + j];
    }
    return 0;
}

void twolevel_setup(tf_handle* h)
{
    if (twolevel_active(h) && !h->tl.built) twolevel_build(h);
}

void twolevel_invalidate(tf_handle* h)
{
    h->tl.built = false;
    h->tl.fresh = false;
}

void twolevel_free(tf_handle* h)
{
    tl_release(h);
}

namespace {

template <typename VT, int P>
void enqueue_bicg_iter(tf_handle* h, const VT* val)
{
    TileArgs ta = tile_args(h);
    const dim3 grid(h->T), blk(WG);
    hipLaunchKernelGGL(k_bicg_p<P>, grid, blk, 0, h->stream, ta, h->dinv, h->kr, h->kp, h->kv, h->kph,
                       h->opts.krylov_max_iter);
    hipEvent_t ev[2];
    timer_begin(h, h->t_spmv, ev);
    const size_t lds = 3 * sizeof(double) * (size_t)h->tile_cap;
    hipLaunchKernelGGL((k_bicg_spmv<VT, 0>), grid, blk, lds, h->stream, ta, h->row_ptr, h->col, val, h->nb, h->kph, h->kv,
                       h->kr0);
    timer_end(h, h->t_spmv, ev);
    hipLaunchKernelGGL(k_bicg_s<P>, grid, blk, 0, h->stream, ta, h->dinv, h->kr, h->kv, h->ks, h->ksh);
    timer_begin(h, h->t_spmv, ev);
    hipLaunchKernelGGL((k_bicg_spmv<VT, 1>), grid, blk, lds, h->stream, ta, h->row_ptr, h->col, val, h->nb, h->ksh, h->kt,
                       h->ks);
    timer_end(h, h->t_spmv, ev);
    hipLaunchKernelGGL(k_bicg_x<P>, grid, blk, 0, h->stream, ta, h->kph, h->ksh, h->ks, h->kt, h->kr0, h->kx, h->kr);
}

// CG uses kp / kph as the ping-pong direction buffers, kv as q, ks as z.
template <typename VT, int P>
void enqueue_cg_iter(tf_handle* h, const VT* val)
{
    TileArgs ta = tile_args(h);
    const dim3 grid(h->T), blk(WG);
    double* p_old = P ? h->kph : h->kp;
    double* p_new = P ? h->kp : h->kph;
    hipEvent_t ev[2];
    timer_begin(h, h->t_spmv, ev);
    hipLaunchKernelGGL((k_cg_q<VT, P, WG>), grid, blk, 3 * sizeof(double) * (size_t)h->tile_cap, h->stream, ta, h->row_ptr,
                       h->col, val, h->nb, h->ks, p_old, p_new, h->kv, h->opts.krylov_max_iter);
    timer_end(h, h->t_spmv, ev);
    hipLaunchKernelGGL(k_cg_x<P>, grid, blk, 0, h->stream, ta, h->dinv, p_new, h->kv, h->kx, h->kr, h->ks);
}

constexpr int KIND_TL = 64;  // two-level PCG (internal solver kind)

// k_cg_q at workgroup size wg (the tiling's), CG buffers: z = ks, p ping-pong kp / kph, q = kv
template <typename VT, int P>
void launch_cg_q(tf_handle* h, const VT* val, int wg, int maxit = -1)
{
    TileArgs ta = tile_args(h);
    double* p_old = P ? h->kph : h->kp;
    double* p_new = P ? h->kp : h->kph;
    const size_t lds = 3 * sizeof(double) * (size_t)h->tile_cap;
    if (maxit < 0) maxit = h->opts.krylov_max_iter;
    if (wg == 1024)
        hipLaunchKernelGGL((k_cg_q<VT, P, 1024>), dim3(h->T), dim3(1024), lds, h->stream, ta, h->row_ptr, h->col, val, h->nb,
                           h->ks, p_old, p_new, h->kv, maxit);
    else if (wg == 512)
        hipLaunchKernelGGL((k_cg_q<VT, P, 512>), dim3(h->T), dim3(512), lds, h->stream, ta, h->row_ptr, h->col, val, h->nb,
                           h->ks, p_old, p_new, h->kv, maxit);
    else
        hipLaunchKernelGGL((k_cg_q<VT, P, 256>), dim3(h->T), dim3(256), lds, h->stream, ta, h->row_ptr, h->col, val, h->nb,
                           h->ks, p_old, p_new, h->kv, maxit);
}

// the Jacobian's values in aggregate order, in the solver's value type
template <typename VT>
static const VT* tp_values(tf_handle* h)
{
    if (sizeof(VT) == sizeof(double)) return reinterpret_cast<const VT*>(h->tl.val_p);
    return reinterpret_cast<const VT*>(h->tl.val32_p);
}

static size_t tp_q_lds(tf_handle* h) { return 3 * sizeof(double) * (size_t)h->tl.tp_max_blocks; }

template <typename VT, int P>
static void launch_tp_q(tf_handle* h, const TileArgs& ta, const TPArgs& tp, int maxit)
{
    double* p_old = P ? h->kph : h->kp;
    double* p_new = P ? h->kp : h->kph;
    hipLaunchKernelGGL((k_tp_q<VT, P>), dim3(tp.n_agg), dim3(WG), tp_q_lds(h), h->stream, ta, tp, tp_values<VT>(h), h->nb,
                       h->ks, p_old, p_new, h->kv, maxit);
}

// Two-level PCG in aggregate order (k_tp_q, k_tp_z): x = kx, r = kr, z = ks, q = kv, p ping-pong kp / kph,
// all in positions; the solve gathers its start and scatters du at the end.
template <typename VT, int P>
void enqueue_tl_iter(tf_handle* h, const VT*)
{
    const TileArgs ta = tp_tile_args(h);
    const TPArgs tp = tp_args(h);
    hipEvent_t ev[2];
    timer_begin(h, h->t_spmv, ev);
    launch_tp_q<VT, P>(h, ta, tp, h->opts.krylov_max_iter);
    timer_end(h, h->t_spmv, ev);
    launch_tp_z<P, false>(h, ta, tp, P ? h->kp : h->kph, h->kv);
}

// Single-kernel CG buffers: x = kx, p = kp, r = kr, w = kv, s = ks (own rows only) and the
// ping-pong pairs u = (kph, km), m = (kt, ksh), q = (kr0, kq); launch parity P reads 1-P.
template <typename VT, int P>
void enqueue_cg1_iter(tf_handle* h, const VT* val)
{
    TileArgs ta = tile_args(h);
    double* U[2] = {h->kph, h->km};
    double* Mv[2] = {h->kt, h->ksh};
    double* Q[2] = {h->kr0, h->kq};
    hipEvent_t ev[2];
    timer_begin(h, h->t_spmv, ev);
    const size_t lds = sizeof(double) * (3 * (size_t)h->tile_cap + 2 * (size_t)h->wg);
#define TF_CG1(WGT, BPT)                                                                                          \
    hipLaunchKernelGGL((k_cg1<VT, P, WGT, BPT>), dim3(h->T), dim3(WGT), lds, h->stream, ta, h->row_ptr, h->col, val, \
                       h->nb, h->dinv, h->kx, h->kp, h->kr, h->kv, h->ks, U[1 - P], U[P], Mv[1 - P], Mv[P], Q[1 - P],  \
                       Q[P], h->opts.krylov_max_iter, h->tile_cap)
    const int bpt = h->tile_cap <= h->wg ? 1 : (h->tile_cap <= 2 * h->wg ? 2 : 4);
    if (h->wg == 512) {
        if (bpt == 1) TF_CG1(512, 1); else if (bpt == 2) TF_CG1(512, 2); else TF_CG1(512, 4);
    } else if (h->wg == 1024) {
        if (bpt == 1) TF_CG1(1024, 1); else TF_CG1(1024, 2);
    } else {
        if (bpt == 1) TF_CG1(256, 1); else if (bpt == 2) TF_CG1(256, 2); else TF_CG1(256, 4);
    }
#undef TF_CG1
    timer_end(h, h->t_spmv, ev);
}

// Pipelined CG buffers: x = kx, r = kr, u = kr0, w = kv, z = kt, q = ksh, s = ks, p = kp,
// m ping-pong = (kph, km); launch of parity P reads m[1-P] and writes m[P].
template <typename VT, int P>
void enqueue_pipe_iter(tf_handle* h, const VT* val)
{
    TileArgs ta = tile_args(h);
    double* Mb[2] = {h->kph, h->km};
    hipEvent_t ev[2];
    timer_begin(h, h->t_spmv, ev);
    const size_t lds = sizeof(double) * (3 * (size_t)h->tile_cap + h->wg);
#define TF_PIPE(WGT, BPT)                                                                                           \
    hipLaunchKernelGGL((k_pipe<VT, P, WGT, BPT>), dim3(h->T), dim3(WGT), lds, h->stream, ta, h->row_ptr, h->col, val, \
                       h->nb, h->dinv, h->kx, h->kr, h->kr0, h->kv, h->kt, h->ksh, h->ks, h->kp, Mb[1 - P], Mb[P],       \
                       h->opts.krylov_max_iter, h->tile_cap)
    const int bpt = h->tile_cap <= h->wg ? 1 : (h->tile_cap <= 2 * h->wg ? 2 : 4);
    if (h->wg == 512) {
        if (bpt == 1) TF_PIPE(512, 1); else if (bpt == 2) TF_PIPE(512, 2); else TF_PIPE(512, 4);
    } else if (h->wg == 1024) {
        if (bpt == 1) TF_PIPE(1024, 1); else TF_PIPE(1024, 2);
    } else {
        if (bpt == 1) TF_PIPE(256, 1); else if (bpt == 2) TF_PIPE(256, 2); else TF_PIPE(256, 4);
    }
#undef TF_PIPE
    timer_end(h, h->t_spmv, ev);
}

template <typename VT>
void enqueue_chunk(tf_handle* h, const VT* val, int kind, int iters)
{
    for (int i = 0; i < iters; i += 2) {
        if (kind == TF_KRYLOV_PIPECG) {
            enqueue_pipe_iter<VT, 1>(h, val);  // chunks start at an odd kernel (kernel 0 is the init)
            enqueue_pipe_iter<VT, 0>(h, val);
        } else if (kind == TF_KRYLOV_CG1) {
            enqueue_cg1_iter<VT, 1>(h, val);  // chunks start at an odd kernel (kernel 0 is the init)
            enqueue_cg1_iter<VT, 0>(h, val);
        } else if (kind == KIND_TL) {
            enqueue_tl_iter<VT, 0>(h, val);
            enqueue_tl_iter<VT, 1>(h, val);
        } else if (kind == TF_KRYLOV_CG) {
            enqueue_cg_iter<VT, 0>(h, val);
            enqueue_cg_iter<VT, 1>(h, val);
        } else {
            enqueue_bicg_iter<VT, 0>(h, val);
            enqueue_bicg_iter<VT, 1>(h, val);
        }
    }
}

bool all_done(tf_handle* h)
{
    TF_HIP(hipMemcpyAsync(h->h_seg, h->seg, sizeof(SegScalars) * h->S, hipMemcpyDeviceToHost, h->stream));
    TF_HIP(hipStreamSynchronize(h->stream));
    h->host_waits += 1;
    for (int s = 0; s < h->S; ++s)
        if (!(h->h_seg[s].converged || h->h_seg[s].stalled)) return false;
    return true;
}

// The persistent two-level solve (k_tp_run) for the current options: the kernel instance, its LDS,
// and whether the device holds n_agg of its workgroups at once (queried once per instance).
template <typename VT>
static const void* tp_run_fn(tf_handle* h, int& idx)
{
    const bool f64 = h->opts.precond_fp64 != 0;
    const int L = h->tl.L;
    idx = (std::is_same<VT, float>::value ? 8 : 0) + (f64 ? 4 : 0) + (L - 1);
#define TF_TPR(LL)                                                                           \
    if (L == LL) return f64 ? (const void*)k_tp_run<VT, LL, true> : (const void*)k_tp_run<VT, LL, false>
    TF_TPR(1);
    TF_TPR(2);
    TF_TPR(3);
    TF_TPR(4);
#undef TF_TPR
    return nullptr;
}

template <typename VT>
static bool tp_run_fits(tf_handle* h)
{
    tf_handle::TwoLevel& tl = h->tl;
    if (!tl.run_ok || tl.L < 1 || tl.L > TL_MAX_LEVELS) return false;
    int idx = 0;
    const void* fn = tp_run_fn<VT>(h, idx);
    if (!fn) return false;
    int& cap = tl.run_cap[idx];
    if (cap < 0) {
        int per_cu = 0;
        const size_t lds = tp_run_lds(tl.max_ldA, h->opts.precond_fp64 != 0);
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, WG, lds) != hipSuccess) per_cu = 0;
        cap = per_cu * h->n_cu;
    }
    return cap >= tl.n_agg;
}

template <typename VT>
static void launch_tp_run(tf_handle* h)
{
    tf_handle::TwoLevel& tl = h->tl;
    int idx = 0;
    const void* fn = tp_run_fn<VT>(h, idx);
    const size_t lds = tp_run_lds(tl.max_ldA, h->opts.precond_fp64 != 0);
    // the barrier words are zero: the solve's first z (k_tp_z<P, true>, enqueued before) cleared them
    TileArgs ta = tp_tile_args(h);
    TPArgs tp = tp_args(h);
    const VT* val = tp_values<VT>(h);
    int nb = h->nb;
    TPRunArgs ra{tl.pslot, tl.halo_ptr, tl.halo_pos, tl.bar, h->opts.krylov_max_iter,
                 h->opts.tl_barrier_polls > 0 ? (unsigned)h->opts.tl_barrier_polls : TPR_SPIN};
    double *x = h->kx, *r = h->kr, *z = h->ks, *p0 = h->kp, *p1 = h->kph;
    void* args[] = {&ta, &tp, &val, &nb, &ra, &x, &r, &z, &p0, &p1};
    // The occupancy check of tp_run_fits puts every workgroup on the device at once when nothing
    // else runs on it; the stream's order has the previous kernel finished.  A workgroup that still
    // waits (another process's kernels holding compute units) ends its bounded barrier wait with the
    // abort word and the solve falls back.  The cooperative launch has the runtime guarantee the
    // co-residency instead, at about 30 us of launch overhead on each side of the kernel.
    if (h->opts.tl_persistent == 2)
        TF_HIP(hipLaunchCooperativeKernel(fn, dim3(tl.n_agg), dim3(WG), args, (unsigned)lds, h->stream));
    else
        TF_HIP(hipLaunchKernel(fn, dim3(tl.n_agg), dim3(WG), args, lds, h->stream));
}

// The graph of iters iterations of the given kind (sz 0: a chunk, 1: a small chunk), captured on
// first use and kept per value type and size.  The captured launches carry the iteration cap as an
// argument: a chunk captured under another cap (a timing probe's) is not reused.
template <typename VT>
hipGraphExec_t chunk_graph(tf_handle* h, const VT* val, int kind, int sz, int iters)
{
    KernelTimer& tm = h->t_spmv;
    const int key = kind * 16 + (int)sizeof(VT) * 2;
    const int gi = (sizeof(VT) == sizeof(double) ? 0 : 2) + sz;
    hipGraphExec_t& graph_exec = h->krylov_graph[gi];
    if (!graph_exec || h->krylov_graph_iters[gi] != iters || h->krylov_graph_kind[gi] != key ||
        h->krylov_graph_maxit[gi] != h->opts.krylov_max_iter) {
        if (graph_exec) { (void)hipGraphExecDestroy(graph_exec); graph_exec = nullptr; }
        hipGraph_t gr;
        TF_HIP(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
        tm.capturing = true;
        try {
            enqueue_chunk<VT>(h, val, kind, iters);
        } catch (...) {
            tm.capturing = false;
            (void)hipStreamEndCapture(h->stream, &gr);
            throw;
        }
        tm.capturing = false;
        TF_HIP(hipStreamEndCapture(h->stream, &gr));
        TF_HIP(hipGraphInstantiate(&graph_exec, gr, nullptr, nullptr, 0));
        TF_HIP(hipGraphDestroy(gr));
        h->krylov_graph_iters[gi] = iters;
        h->krylov_graph_kind[gi] = key;
        h->krylov_graph_maxit[gi] = h->opts.krylov_max_iter;
    }
    return graph_exec;
}

// The next round of a running solve (host side): the first (round 0) is the predicted burst, or
// one chunk without a prediction; then small chunks, twice, after a predicted burst, then chunks.
void round_plan(tf_handle* h)
{
    tf_handle::KrylovRun& R = h->krun;
    if (R.persist && R.round == 0) {  // the whole solve in one launch
        R.round = 1;
        R.persist_next = true;
        R.last_launched = -1;  // the launch's iterations are the solve's
        return;
    }
    if (R.graph && R.round > 0) {
        const bool fine = R.predicted && R.round <= 2;
        R.n_big = fine ? 0 : 1;
        R.n_small = fine ? 1 : 0;
    }
    const int launched = R.graph ? R.n_big * R.chunk + R.n_small * R.small : R.chunk;
    R.round += 1;
    R.done_iters += launched;
    R.last_launched = launched;
}

// The launches of the round round_plan sized: graph replays of the captured chunks, or (direct:
// inside the capture of a larger graph, or without use_graph) the iterations' kernels themselves.
template <typename VT>
void round_launch(tf_handle* h, const VT* val, bool direct)
{
    tf_handle::KrylovRun& R = h->krun;
    KernelTimer& tm = h->t_spmv;
    if (R.timing) TF_HIP(hipEventRecord(tm.sample[0], h->stream));
    if (R.persist_next) {
        // a cooperative launch cannot be captured; the ordinary launch is a kernel node like any
        // other (the Newton-iteration graph counts its replays)
        if (direct && h->opts.tl_persistent == 2)
            throw std::logic_error("the cooperative persistent solve cannot be captured");
        R.persist_next = false;
        launch_tp_run<VT>(h);
        if (!direct) h->persist_solves += 1;
    } else if (R.graph && !direct) {
        hipGraphExec_t big = R.n_big ? chunk_graph<VT>(h, val, R.kind, 0, R.chunk) : nullptr;
        hipGraphExec_t sm = R.n_small ? chunk_graph<VT>(h, val, R.kind, 1, R.small) : nullptr;
        for (int k = 0; k < R.n_big; ++k) TF_HIP(hipGraphLaunch(big, h->stream));
        for (int k = 0; k < R.n_small; ++k) TF_HIP(hipGraphLaunch(sm, h->stream));
    } else if (R.graph) {
        for (int k = 0; k < R.n_big; ++k) enqueue_chunk<VT>(h, val, R.kind, R.chunk);
        for (int k = 0; k < R.n_small; ++k) enqueue_chunk<VT>(h, val, R.kind, R.small);
    } else {
        enqueue_chunk<VT>(h, val, R.kind, R.chunk);
    }
    TF_HIP(hipGetLastError());
    if (R.timing) TF_HIP(hipEventRecord(tm.sample[1], h->stream));
}

template <typename VT>
void solve_round(tf_handle* h, const VT* val)
{
    round_plan(h);
    round_launch<VT>(h, val, false);
}

// The set-up kernels of a solve of the given kind: x_0, r_0 and the first direction's ingredients,
// the tissues' tolerances (k_seg_init).
template <typename VT>
void solve_launch_init(tf_handle* h, const VT* val, int kind)
{
    TileArgs ta = tile_args(h);
    const dim3 grid(h->T), blk(WG);
    const double rtol2 = h->opts.krylov_rtol * h->opts.krylov_rtol;
    const double atol2 = h->opts.krylov_atol * h->opts.krylov_atol;
    if (kind == TF_KRYLOV_PIPECG)
        hipLaunchKernelGGL(k_pipe_init<VT>, grid, blk, sizeof(double) * (3 * (size_t)h->tile_cap + h->wg), h->stream, ta,
                           h->row_ptr, h->col, val, h->nb, h->dinv, h->g, h->kx, h->kr, h->kr0, h->kv, h->kt, h->ksh,
                           h->ks, h->kp, h->kph, h->tile_cap);
    else if (kind == TF_KRYLOV_CG1 && h->warm_guess) {
        const size_t lds = sizeof(double) * (3 * (size_t)h->tile_cap + h->wg);
        hipLaunchKernelGGL(k_cg1_warm_a<VT>, grid, blk, lds, h->stream, ta, h->row_ptr, h->col, val, h->nb, h->dinv, h->g,
                           h->warm_guess, h->kx, h->kr, h->ks, h->kp, h->kph, h->kr0, h->tile_cap);
        hipLaunchKernelGGL(k_cg1_warm_b<VT>, grid, blk, lds, h->stream, ta, h->row_ptr, h->col, val, h->nb, h->dinv, h->g,
                           h->kr, h->kph, h->kv, h->kt, h->tile_cap);
    } else if (kind == TF_KRYLOV_CG1)
        hipLaunchKernelGGL(k_cg1_init<VT>, grid, blk, sizeof(double) * (3 * (size_t)h->tile_cap + h->wg), h->stream, ta,
                           h->row_ptr, h->col, val, h->nb, h->dinv, h->g, h->kx, h->kr, h->kv, h->ks, h->kp, h->kph,
                           h->kt, h->kr0, h->tile_cap);
    else if (kind == KIND_TL) {  // in aggregate order: the start gathered, r_0, P^T r_0
        const double* x0p = nullptr;
        if (h->warm_guess) {
            hipLaunchKernelGGL(k_tp_gather, dim3((3 * h->N + 255) / 256), dim3(256), 0, h->stream, h->tl.agg_nodes, h->N,
                               h->warm_guess, h->km);
            x0p = h->km;
        }
        hipLaunchKernelGGL(k_tp_init<VT>, dim3(h->tl.n_agg), blk, tp_q_lds(h), h->stream, tp_tile_args(h), tp_args(h),
                           tp_values<VT>(h), h->nb, h->g, x0p, h->kx, h->kr, h->kp);
    }
    else if (kind == TF_KRYLOV_CG)
        hipLaunchKernelGGL(k_cg_init, grid, blk, 0, h->stream, ta, h->dinv, h->g, h->kx, h->kr, h->ks, h->kp);
    else
        hipLaunchKernelGGL(k_bicg_init, grid, blk, 0, h->stream, ta, h->g, h->kx, h->kr, h->kr0, h->kp, h->kv);
    hipLaunchKernelGGL(k_seg_init, dim3(h->S), blk, 0, h->stream, kind == KIND_TL ? tp_tile_args(h) : ta, rtol2, atol2,
                       h->seg_rtol_on ? h->seg_rtol2 : nullptr);
    if (kind == KIND_TL)  // z_0 = D^-1 r_0 + P Ac^-1 P^T r_0
        launch_tp_z<1, true>(h, tp_tile_args(h), tp_args(h), h->kp, h->kv);
    TF_HIP(hipGetLastError());
}

// The host side of a solve's start: kind, chunk sizes and the first burst (h->krun holds the solve
// until solve_end); no launches.
void solve_plan(tf_handle* h, int kind, bool fp32)
{
    tf_handle::KrylovRun& R = h->krun;
    R = tf_handle::KrylovRun{};
    R.kind = kind;
    R.fp32 = fp32;
    // A captured graph holds at most TF_GRAPH_NODES kernel launches (an even number of
    // iterations); a chunk between convergence checks is one replay of it, at most iters_per_sync
    // iterations.
    const int kpi = kind == KIND_TL ? 2 : kind == TF_KRYLOV_CG ? 2 : kind == TF_KRYLOV_BICGSTAB ? 5 : 1;
    int chunk = std::min(std::max(2, h->opts.iters_per_sync), std::max(2, TF_GRAPH_NODES / kpi));
    chunk -= chunk & 1;  // iterations come in parity pairs
    R.chunk = chunk;
    // a small chunk: a quarter of the chunk (an even count), for the end of a predicted solve
    R.small = std::max(2, (chunk / 4) & ~1);
    R.timing = h->t_spmv.enabled;
    // Graph replay.  While timing, the kinds with more than one kernel per iteration (other than
    // the two-level PCG) launch eagerly, each launch of the dominant kernel bracketed by events;
    // replays are timed whole.
    R.graph = h->opts.use_graph != 0 &&
              (!R.timing || kind == TF_KRYLOV_CG1 || kind == TF_KRYLOV_PIPECG || kind == KIND_TL);
    // First burst, replayed without a convergence check: the iterations this Newton iteration's
    // solve took last timestep plus the launch that sees convergence, rounded up to small chunks
    // (chunks first), plus burst_slack small chunks.  A solve that converges inside a replay
    // leaves the rest of its launches as no-ops.
    R.call = h->krylov_call;
    R.predicted = R.graph && R.call >= 0 && R.call < (int)h->krylov_pred.size() && h->krylov_pred[R.call] >= 0;
    // the persistent launch runs the solve to its end: the round that follows needs no check
    R.persist = kind == KIND_TL && h->opts.tl_persistent != 0 && !h->no_persist &&
                (fp32 ? tp_run_fits<float>(h) : tp_run_fits<double>(h));
    if (R.predicted) {
        const int target = ((h->krylov_pred[R.call] + 1 + R.small - 1) / R.small + h->burst_slack) * R.small;
        R.n_big = target / R.chunk;
        R.n_small = (target - R.n_big * R.chunk + R.small - 1) / R.small;
    }
    if (R.persist) R.predicted = true;
    KernelTimer& tm = h->t_spmv;
    if (R.timing && !tm.sample[0]) {
        TF_HIP(hipEventCreate(&tm.sample[0]));
        TF_HIP(hipEventCreate(&tm.sample[1]));
    }
    R.pending = true;
}

// The set-up of a solve of the given kind and its first round, enqueued without a host wait.
template <typename VT>
void solve_start(tf_handle* h, const VT* val, int kind)
{
    solve_launch_init<VT>(h, val, kind);
    solve_plan(h, kind, sizeof(VT) == sizeof(float));
    solve_round<VT>(h, val);
}

// The end of the solve solve_start began: convergence checks (one host wait each; seg_known: the
// host mirror of the tissues' scalars is already current for the rounds launched so far) and
// further rounds until every tissue converged or stalled, then the solve's bookkeeping.
template <typename VT>
int solve_end(tf_handle* h, const VT* val, bool seg_known)
{
    tf_handle::KrylovRun& R = h->krun;
    KernelTimer& tm = h->t_spmv;
    while (true) {
        bool fin = true;
        if (seg_known) {
            for (int s = 0; s < h->S; ++s) fin = fin && (h->h_seg[s].converged || h->h_seg[s].stalled);
        } else {
            fin = all_done(h);
        }
        seg_known = false;
        if (R.timing && R.last_launched != 0) {
            float ms = 0.f;
            TF_HIP(hipEventElapsedTime(&ms, tm.sample[0], tm.sample[1]));
            tm.replay_ms += ms;
            int launched = R.last_launched;
            if (launched < 0) {  // the persistent launch: the iterations it ran
                launched = 0;
                for (int s = 0; s < h->S; ++s) launched = std::max(launched, (int)h->h_seg[s].iters);
            }
            tm.replay_launches += launched;
            R.last_launched = 0;
        }
        if (fin) break;
        if (R.done_iters > h->opts.krylov_max_iter + 2 * R.chunk) break;  // the cap marks the tissue stalled
        solve_round<VT>(h, val);
    }
    R.pending = false;
    timer_collect(h->t_spmv);
    int st = TF_OK;
    int maxit = 0;
    bool nan = false;
    for (int s = 0; s < h->S; ++s) {
        maxit = std::max(maxit, h->h_seg[s].iters);
        if (h->h_seg[s].active && h->h_seg[s].stalled) st = TF_KRYLOV_STALL;
        nan = nan || (h->h_seg[s].active && !std::isfinite(h->h_seg[s].bnorm2));
    }
    if (nan) st = TF_NAN;  // a right-hand side that is not finite: no solver reaches a tolerance
    h->last_krylov_iters = maxit;
    if (R.timing) tm.replay_iters += maxit;
    if (R.call >= 0) {
        if ((int)h->krylov_pred.size() <= R.call) h->krylov_pred.resize(R.call + 1, -1);
        h->krylov_pred[R.call] = st == TF_OK ? maxit : -1;
    }
    return st;
}

template <typename VT>
int solve_kind(tf_handle* h, const VT* val, int kind)
{
    solve_start<VT>(h, val, kind);
    return solve_end<VT>(h, val, false);
}

}  // namespace

bool twolevel_active(const tf_handle* h)
{
    int kind = h->opts.krylov;
    if (kind < 0) kind = h->any_viscous ? TF_KRYLOV_BICGSTAB : TF_KRYLOV_CG1;
    const bool cg_kind = kind == TF_KRYLOV_CG || kind == TF_KRYLOV_CG1 || kind == TF_KRYLOV_PIPECG;
    const int pc = h->opts.preconditioner;
    const bool want = pc == TF_PRECOND_TWO_LEVEL || (pc == TF_PRECOND_AUTO && h->N >= TF_TWO_LEVEL_MIN_NODES);
    return cg_kind && want;
}

// The per-solve part of the two-level preconditioner: the inverse column blocks of the current
// Jacobian and its values in aggregate order (the arrays are allocated by twolevel_prepare).
static void twolevel_launch_prep(tf_handle* h)
{
    tf_handle::TwoLevel& tl = h->tl;
    launch_tl_dcol(h);
    const dim3 grid((h->nb + 255) / 256), blk(256);
    if (fp32_values(h))
        hipLaunchKernelGGL(k_tp_permute<float>, grid, blk, 0, h->stream, tl.pblk, h->nb, h->val32, tl.val32_p);
    else
        hipLaunchKernelGGL(k_tp_permute<double>, grid, blk, 0, h->stream, tl.pblk, h->nb, h->val, tl.val_p);
    TF_HIP(hipGetLastError());
}

// The coarse space follows the pattern and tiling.  A tissue's block of the coarse operator is
// rebuilt at the first Newton iteration of a timestep once the tissue's Krylov iterations of
// Newton iteration 1 (the largest solve) exceed 1.25x their count in the timestep of its last
// rebuild; every block coarse_refresh timesteps after the last full rebuild, and for every solve
// outside the Newton loop (timing runs keep the operator in place).  A coarse operator from an
// earlier Jacobian is still symmetric positive definite, so it changes the iteration count, not
// the solution.  False when the coarse operator is not positive definite (block Jacobi then).
static bool twolevel_prepare(tf_handle* h, bool reuse = false, bool launch = true)
{
    tf_handle::TwoLevel& tl = h->tl;
    if (!tl.built) twolevel_build(h);
    if (h->krylov_call == 0) tl.age += 1;
    // reuse (timing runs): the coarse operator in place, when there is one
    const bool all = !tl.fresh || (h->krylov_call < 0 && !reuse) ||
                     (h->krylov_call == 0 && tl.age >= std::max(1, h->opts.coarse_refresh));
    std::vector<uint8_t> which(h->S, all ? 1 : 0);
    bool any = all;
    if (!all && h->krylov_call == 0)
        for (int sg = 0; sg < h->S; ++sg)
            if (tl.stale[sg]) { which[sg] = 1; any = true; }
    if (any) twolevel_refresh(h, which);
    if (!tl.failed) {
        if (fp32_values(h)) {
            if (!tl.val32_p) tl.val32_p = tl_alloc<float>(9 * (size_t)h->nb);
        } else if (!tl.val_p) {
            tl.val_p = tl_alloc<double>(9 * (size_t)h->nb);
        }
        if (launch) twolevel_launch_prep(h);
    }
    return !tl.failed;
}

// After a two-level solve of Newton iteration 1, for every tissue: the rebuild's baseline, or the
// staleness test (a stale tissue's block is rebuilt at the next timestep's first solve)
static void twolevel_observe(tf_handle* h)
{
    tf_handle::TwoLevel& tl = h->tl;
    if (h->krylov_call != 1) return;
    for (int sg = 0; sg < h->S; ++sg) {
        const int iters = h->h_seg[sg].iters;
        if (tl.base_it1[sg] < 0) tl.base_it1[sg] = iters;
        else if (iters > 1.25 * tl.base_it1[sg]) tl.stale[sg] = 1;
    }
}

// Average device time (ms) of one Krylov iteration over n_iters iterations on the current
// Jacobian, under the current tiling (tolerance 0; no fallback): whole chunks (graph replays
// under use_graph) divided by the iterations launched.  The two-level PCG when twolevel_active, else the single-kernel CG.
double krylov_time_iteration(tf_handle* h, int n_iters)
{
    const tf_solver_opts saved = h->opts;
    const int kind = twolevel_active(h) && twolevel_prepare(h, true) ? KIND_TL : TF_KRYLOV_CG1;
    h->opts.krylov_rtol = 0.0;
    h->opts.krylov_atol = 0.0;
    h->opts.krylov_max_iter = n_iters - 1;  // launch n_iters sees the cap
    h->opts.iters_per_sync = std::min(n_iters, 32);  // the chunk the solves replay
    KernelTimer& tm = h->t_spmv;
    hipLaunchKernelGGL(k_seg_reopen, dim3((h->S + 255) / 256), dim3(256), 0, h->stream, h->seg, h->S);
    TF_HIP(hipGetLastError());
    const bool was = tm.enabled;
    const double keep_ms = tm.total_ms;
    const int64_t keep_n = tm.launches;
    tm.enabled = true;
    double ms = 0.0;
    const int call = h->krylov_call;
    h->krylov_call = -1;
    for (int rep = 0; rep < 2; ++rep) {  // the first run captures the graph
        tm.replay_ms = 0.0;
        tm.replay_launches = 0;
        if (fp32_values(h)) solve_kind<float>(h, h->val32, kind);
        else solve_kind<double>(h, h->val, kind);
        ms = tm.replay_launches ? tm.replay_ms / tm.replay_launches : 0.0;
    }
    h->krylov_call = call;
    tm.enabled = was;
    tm.total_ms = keep_ms;
    tm.launches = keep_n;
    h->opts = saved;
    return ms;
}

// Back-to-back launches of the dominant Krylov kernel on the current Jacobian, timed whole by
// events around a graph replay (around the launches themselves without use_graph).  A short solve with tolerance 0 first leaves consistent vectors
// and partials; the tissues' flags are then reopened.  The single-kernel and pipelined CG advance
// their iterations (the launches are the solver's own); k_cg_q (two-level and textbook CG) is
// launched with one parity, so every launch repeats the same direction update and SpMV.
double krylov_probe(tf_handle* h, int n_launches, int which)
{
    if (which == 1 && !twolevel_active(h)) return 0.0;  // k_tp_z: the two-level PCG only
    if (which == 2) {  // the persistent solve: ms per iteration of one launch of n_launches iterations
        if (!twolevel_active(h) || !h->opts.tl_persistent) return 0.0;
        if (!h->tl.built) twolevel_build(h);
        const bool fits = fp32_values(h) ? tp_run_fits<float>(h) : tp_run_fits<double>(h);
        return fits ? krylov_time_iteration(h, n_launches) : 0.0;
    }
    int kind = h->opts.krylov;
    if (kind < 0) kind = h->any_viscous ? TF_KRYLOV_BICGSTAB : TF_KRYLOV_CG1;
    const bool tl = twolevel_active(h);
    if (kind == TF_KRYLOV_BICGSTAB && !tl) return 0.0;
    if (kind == TF_KRYLOV_CG1 || kind == TF_KRYLOV_PIPECG) {
        if (!tl) {
            if (which != 0) return 0.0;
            const tf_solver_opts saved = h->opts;
            h->opts.krylov = kind;
            h->opts.preconditioner = TF_PRECOND_BLOCK_JACOBI;
            double ms = 0.0;
            try {
                ms = krylov_time_iteration(h, n_launches);
            } catch (...) {
                h->opts = saved;
                throw;
            }
            h->opts = saved;
            return ms;
        }
    }
    // k_cg_q: set-up solve of 8 iterations (two-level when active, else textbook CG) on every
    // tissue (a finished timestep leaves them inactive)
    hipLaunchKernelGGL(k_seg_reopen, dim3((h->S + 255) / 256), dim3(256), 0, h->stream, h->seg, h->S);
    TF_HIP(hipGetLastError());
    const tf_solver_opts saved = h->opts;
    const int sk = tl && twolevel_prepare(h, true) ? KIND_TL : TF_KRYLOV_CG;
    h->opts.krylov_rtol = 0.0;
    h->opts.krylov_atol = 0.0;
    h->opts.krylov_max_iter = 8;  // the ninth iteration sees the cap: the solve stops there
    h->opts.iters_per_sync = 8;
    const bool was = h->t_spmv.enabled;
    h->t_spmv.enabled = false;
    const int call = h->krylov_call;
    h->krylov_call = -1;
    double ms = 0.0;
    try {
        if (fp32_values(h)) solve_kind<float>(h, h->val32, sk);
        else solve_kind<double>(h, h->val, sk);
    } catch (...) {
        h->opts = saved; h->t_spmv.enabled = was; h->krylov_call = call;
        throw;
    }
    h->opts = saved;
    h->t_spmv.enabled = was;
    h->krylov_call = call;
    hipLaunchKernelGGL(k_seg_reopen, dim3((h->S + 255) / 256), dim3(256), 0, h->stream, h->seg, h->S);
    TF_HIP(hipGetLastError());
    auto launches = [&]() {
        const bool fp32 = fp32_values(h);
        for (int k = 0; k < n_launches; ++k) {
            if (sk == KIND_TL && which == 1)  // k_tp_z: alpha from the same partials every launch
                launch_tp_z<0, false>(h, tp_tile_args(h), tp_args(h), h->kph, h->kv);
            else if (sk == KIND_TL) {  // k_tp_q, as the two-level solve launches it
                if (fp32) launch_tp_q<float, 0>(h, tp_tile_args(h), tp_args(h), 1 << 30);
                else launch_tp_q<double, 0>(h, tp_tile_args(h), tp_args(h), 1 << 30);
            } else if (fp32) launch_cg_q<float, 0>(h, h->val32, WG, 1 << 30);
            else launch_cg_q<double, 0>(h, h->val, WG, 1 << 30);
        }
    };
    // one graph replay (use_graph), else the launches enqueued one by one
    hipGraphExec_t ex = nullptr;
    if (h->opts.use_graph) {
        hipGraph_t gr;
        TF_HIP(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
        launches();
        TF_HIP(hipStreamEndCapture(h->stream, &gr));
        TF_HIP(hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0));
        TF_HIP(hipGraphDestroy(gr));
    }
    hipEvent_t ev[2];
    TF_HIP(hipEventCreate(&ev[0]));
    TF_HIP(hipEventCreate(&ev[1]));
    for (int rep = 0; rep < 2; ++rep) {  // the first pass warms the caches
        TF_HIP(hipEventRecord(ev[0], h->stream));
        if (ex) TF_HIP(hipGraphLaunch(ex, h->stream));
        else launches();
        TF_HIP(hipEventRecord(ev[1], h->stream));
        TF_HIP(hipEventSynchronize(ev[1]));
        float e = 0.f;
        TF_HIP(hipEventElapsedTime(&e, ev[0], ev[1]));
        ms = e / n_launches;
    }
    (void)hipEventDestroy(ev[0]);
    (void)hipEventDestroy(ev[1]);
    if (ex) (void)hipGraphExecDestroy(ex);
    return ms;
}

// The two-level prepare of the next krylov_solve (coarse rebuild when due; launch: also the per-solve
// kernels, else krylov_pre_launch enqueues them)
void krylov_pre_solve(tf_handle* h, bool launch)
{
    h->tl_pre_valid = false;
    int kind = h->opts.krylov;
    if (kind < 0) kind = h->any_viscous ? TF_KRYLOV_BICGSTAB : TF_KRYLOV_CG1;
    const bool cg_kind = kind == TF_KRYLOV_CG || kind == TF_KRYLOV_CG1 || kind == TF_KRYLOV_PIPECG;
    if (!cg_kind || !twolevel_active(h)) return;
    h->tl_pre_ok = twolevel_prepare(h, false, launch);
    h->tl_pre_valid = true;
}

void krylov_pre_launch(tf_handle* h) { twolevel_launch_prep(h); }

// Whether the next krylov_pre_solve of a Newton iteration would build the coarse space or rebuild a
// coarse block (work with host waits of its own); no side effects.
bool krylov_pre_solve_rebuilds(const tf_handle* h)
{
    int kind = h->opts.krylov;
    if (kind < 0) kind = h->any_viscous ? TF_KRYLOV_BICGSTAB : TF_KRYLOV_CG1;
    const bool cg_kind = kind == TF_KRYLOV_CG || kind == TF_KRYLOV_CG1 || kind == TF_KRYLOV_PIPECG;
    if (!cg_kind || !twolevel_active(h)) return false;
    const tf_handle::TwoLevel& tl = h->tl;
    if (!tl.built || !tl.fresh || h->krylov_call < 0) return true;
    if (h->krylov_call != 0) return false;
    if (tl.age + 1 >= std::max(1, h->opts.coarse_refresh)) return true;
    for (int sg = 0; sg < h->S; ++sg)
        if (tl.stale[sg]) return true;
    return false;
}

// Solve M du = g for every active tissue; du is written to h->du.
int krylov_solve(tf_handle* h)
{
    krylov_solve_begin(h);
    return krylov_solve_end(h, false, false);
}

bool krylov_solve_plan(tf_handle* h)
{
    int kind = h->opts.krylov;
    const bool fp32 = fp32_values(h);
    const bool cg_kind = kind == TF_KRYLOV_CG || kind == TF_KRYLOV_CG1 || kind == TF_KRYLOV_PIPECG;
    const bool pre = h->tl_pre_valid;
    h->tl_pre_valid = false;
    if (cg_kind && twolevel_active(h) && (pre ? h->tl_pre_ok : twolevel_prepare(h))) {
        kind = KIND_TL;
        h->tl_used = true;
    }
    solve_plan(h, kind, fp32);
    h->krun.cg = cg_kind;
    round_plan(h);
    return h->krun.predicted;
}

void krylov_solve_launch(tf_handle* h, bool direct)
{
    const tf_handle::KrylovRun& R = h->krun;
    if (R.fp32) {
        solve_launch_init<float>(h, h->val32, R.kind);
        round_launch<float>(h, h->val32, direct);
    } else {
        solve_launch_init<double>(h, h->val, R.kind);
        round_launch<double>(h, h->val, direct);
    }
}

bool krylov_solve_begin(tf_handle* h)
{
    const bool predicted = krylov_solve_plan(h);
    krylov_solve_launch(h, false);
    return predicted;
}

void launch_solution_to_du(tf_handle* h)
{
    if (h->krun.kind == KIND_TL)  // the two-level solve runs in aggregate order
        hipLaunchKernelGGL(k_tp_scatter, dim3((3 * h->N + 255) / 256), dim3(256), 0, h->stream, h->tl.agg_nodes, h->N,
                           h->kx, h->du);
    else
        launch_copy(h, h->du, h->kx, 3 * (int64_t)h->N);
    TF_HIP(hipGetLastError());
}

int krylov_solve_end(tf_handle* h, bool seg_known, bool scattered)
{
    const int kind = h->krun.kind;
    const bool fp32 = h->krun.fp32, cg_kind = h->krun.cg;
    auto run = [&](int k) { return fp32 ? solve_kind<float>(h, h->val32, k) : solve_kind<double>(h, h->val, k); };
    int st = fp32 ? solve_end<float>(h, h->val32, seg_known) : solve_end<double>(h, h->val, seg_known);
    if (kind == KIND_TL && st == TF_OK) twolevel_observe(h);
    if (kind == KIND_TL && st != TF_OK) h->tl.fresh = false;  // the next solve rebuilds the coarse operator
    bool unusable = false;  // a stalled tissue whose residual is above 1e-2 of its right-hand side
    for (int s = 0; s < h->S; ++s)
        unusable = unusable || (h->h_seg[s].active && h->h_seg[s].stalled &&
                                solve_unusable(h->h_seg[s].rnorm2, h->h_seg[s].bnorm2));
    if (st == TF_KRYLOV_STALL && (cg_kind || kind == KIND_TL) && unusable) {
        // not positive definite (or too slow): BiCGSTAB from zero.  A CG solve that stagnated
        // close to its tolerance is kept as it is (an inexact step).
        const int cg_iters = h->last_krylov_iters;
        h->fallbacks += 1;
        const int call = h->krylov_call;
        h->krylov_call = -1;  // the prediction belongs to the first method
        try {
            st = run(TF_KRYLOV_BICGSTAB);
        } catch (...) {
            h->krylov_call = call;
            throw;
        }
        h->krylov_call = call;
        h->last_krylov_iters += cg_iters;
        scattered = false;
    }
    if (!scattered) launch_solution_to_du(h);
    return st;
}

void launch_spmv(tf_handle* h, const double* x, double* y)
{
    if (fp32_values(h))
        hipLaunchKernelGGL(k_spmv<float>, dim3(h->T), dim3(WG), 3 * sizeof(double) * (size_t)h->tile_cap, h->stream, h->row_ptr, h->col, h->val32, h->nb,
                           h->tile_r0, h->tile_r1, x, y);
    else
        hipLaunchKernelGGL(k_spmv<double>, dim3(h->T), dim3(WG), 3 * sizeof(double) * (size_t)h->tile_cap, h->stream, h->row_ptr, h->col, h->val, h->nb,
                           h->tile_r0, h->tile_r1, x, y);
    TF_HIP(hipGetLastError());
}

void launch_seg_norm(tf_handle* h, const double* x)
{
    TileArgs ta = tile_args(h);
    hipLaunchKernelGGL(k_tile_norm, dim3(h->T), dim3(WG), 0, h->stream, ta, x);
    hipLaunchKernelGGL(k_seg_norm_fin, dim3(h->S), dim3(WG), 0, h->stream, ta, h->seg_norm);
    TF_HIP(hipGetLastError());
}
