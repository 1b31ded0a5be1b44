// tf_assemble.hip -- Jacobian / residual assembly, growth and viscosity-surface kernels (gfx950, FP64).
//
// Assembly is two deterministic passes, no atomics on values:
//   k_elements : one workgroup = 32 prisms.  Phase A: one lane per (prism, Gauss point) evaluates
//                the kinematics and stresses and stages them in LDS.  Phase B: one lane per
//                (prism, node pair a-b) contracts the six staged Gauss points into the 3x3 block
//                M_ab and stores it (36 blocks per prism, contiguous).  Phase C: one lane per
//                (prism, node) sums the nodal force.
//   k_gather   : one lane per 3x3 block of the BSR Jacobian sums its contributing element
//                blocks in a fixed (element, pair) order, adds the nodal external-viscosity
//                diagonal, applies fixed DoFs (unit row/column) and inverts diagonal blocks for
//                the block-Jacobi preconditioner.
//   k_residual : one lane per node sums its element corners and the external viscous force.
#include "tf_internal.hpp"
#include "tf_element.hpp"

#include <mutex>

namespace {

constexpr int EPB = 32;   // prisms per workgroup
constexpr int WG = 256;

typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

// Mixed-precision element tangent on the matrix cores.  With the 54-long contraction index
// kk = (g, i', j) (Gauss point g, displacement-gradient component du_i'/dx_j), the prism tangent
// (both KIntegral1 and KIntegral2) is one product  K = X^T Y  where
//   X[kk][(a,i)] = d_{i i'} gN_a,j                                    (gradient operator G_g)
//   Y[kk][(b,k)] = clam d_{i'j} gN_b,k + cmu (d_{i'k} gN_b,j + d_{jk} gN_b,i') + d_{i'k} (sig gN_b)_j
// (Y = A_g G_g with A_g the Gauss point's 9x9 tangent; clam, cmu, sig carry w J D0 as staged).
// The 18x18 result is one 32x32 MFMA tile: M = N = 18 padded to 32, K = 54 padded to 56 (FP32,
// 28 steps of v_mfma_f32_32x32x2_f32) or 64 (BF16, 4 steps of v_mfma_f32_32x32x16_bf16).
// Operands are formed in FP64 from the staged Gauss-point data and rounded to FP32 / BF16;
// accumulation is FP32.
__device__ inline double mx_X(const tf::GPElastic& E, int kk_ij, int m)
{
    if (m >= 18) return 0.0;
    const int ip = kk_ij / 3, j = kk_ij % 3;
    return (m % 3 == ip) ? E.gn[3 * (m / 3) + j] : 0.0;
}
__device__ inline double mx_Y(const tf::GPElastic& E, int kk_ij, int n)
{
    if (n >= 18) return 0.0;
    const int ip = kk_ij / 3, j = kk_ij % 3;
    const int b = n / 3, k = n % 3;
    const double* B = E.gn + 3 * b;
    double sB[3];
    tf::voigt_mv(E.sig, B, sB);
    double y = 0.0;
    if (ip == j) y += E.clam * B[k];
    if (ip == k) y += E.cmu * B[j] + sB[j];
    if (j == k) y += E.cmu * B[ip];
    return y;
}

// one wavefront: tangent of prism el into kelem (FP32 accumulate, stored as FP64)
template <int TP>
__device__ inline void mfma_tangent(const tf::GPElastic* sE, int el, int e, double* __restrict__ kelem)
{
    const int l = threadIdx.x % 64;
    const int row = l % 32;
    floatx16 c;
#pragma unroll
    for (int r = 0; r < 16; ++r) c[r] = 0.f;
    if (TP == TF_PREC_FP32) {
        for (int st = 0; st < 27; ++st) {
            const int kk = 2 * st + l / 32;       // 0..53
            const int g = kk / 9, ij = kk % 9;
            const tf::GPElastic& E = sE[g * EPB + el];
            const float av = (float)mx_X(E, ij, row);
            const float bv = (float)mx_Y(E, ij, row);
            c = __builtin_amdgcn_mfma_f32_32x32x2f32(av, bv, c, 0, 0, 0);
        }
    } else {
        for (int st = 0; st < 4; ++st) {
            bf16x8 av, bv;
#pragma unroll
            for (int t = 0; t < 8; ++t) {
                const int kk = 16 * st + 8 * (l / 32) + t;
                double x = 0.0, y = 0.0;
                if (kk < 54) {
                    const int g = kk / 9, ij = kk % 9;
                    const tf::GPElastic& E = sE[g * EPB + el];
                    x = mx_X(E, ij, row);
                    y = mx_Y(E, ij, row);
                }
                av[t] = (__bf16)(float)x;
                bv[t] = (__bf16)(float)y;
            }
            c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, bv, c, 0, 0, 0);
        }
    }
    // accumulator layout: register r of lane l holds D[8*(r/4) + 4*(l/32) + r%4][l%32]
    const int col = l % 32;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int m = 8 * (r / 4) + 4 * (l / 32) + (r % 4);
        if (m < 18 && col < 18) {
            const int a = m / 3, i = m % 3, b = col / 3, k = col % 3;
            kelem[(size_t)e * 324 + (6 * a + b) * 9 + 3 * i + k] = (double)c[r];
        }
    }
}

template <bool VISC, int TP>
__global__ __launch_bounds__(WG) void k_elements(const int32_t* __restrict__ conn, const double* __restrict__ xref,
                                                 const double* __restrict__ invfg, const double* __restrict__ lam,
                                                 const double* __restrict__ mu, const double* __restrict__ eta,
                                                 const double* __restrict__ uk, const double* __restrict__ un,
                                                 double inv_dt, int E, double* __restrict__ kelem,
                                                 double* __restrict__ felem, int32_t* __restrict__ status)
{
    __shared__ tf::GPElastic sE[6 * EPB];
    __shared__ tf::GPViscous sV[VISC ? 6 * EPB : 1];
    const int tid = threadIdx.x;
    const int e0 = blockIdx.x * EPB;

    // ---- phase A: (prism, Gauss point) ----
    if (tid < 6 * EPB) {
        const int el = tid % EPB, g = tid / EPB;
        const int e = e0 + el;
        if (e < E) {
            double X[18], x[18], Fi[9], v[18];
#pragma unroll
            for (int a = 0; a < 6; ++a) {
                const int id = conn[a * E + e];
#pragma unroll
                for (int i = 0; i < 3; ++i) {
                    X[3 * a + i] = xref[(3 * a + i) * E + e];
                    x[3 * a + i] = uk[3 * id + i];
                    if (VISC) v[3 * a + i] = (uk[3 * id + i] - un[3 * id + i]) * inv_dt;
                }
            }
#pragma unroll
            for (int q = 0; q < 9; ++q) Fi[q] = invfg[q * E + e];
            const double et = VISC ? eta[e] : 0.0;
            tf::GPViscous vv;
            const double Je = tf::gp_eval(g, X, x, Fi, lam[e], mu[e], et, v, sE[tid], VISC ? &vv : nullptr);
            if (VISC) {
                if (et == 0.0) {
                    vv.ceta = 0.0;
#pragma unroll
                    for (int q = 0; q < 6; ++q) vv.sv[q] = 0.0;
                }
                sV[tid] = vv;
            }
            // a det Fe that is not finite (NaN or Inf in the positions or the growth state) is
            // counted apart from an inverted prism
            if (!isfinite(Je)) atomicAdd(&status[3], 1);
            else if (!(Je > 0.0)) atomicAdd(&status[0], 1);
        }
    }
    __syncthreads();

    // ---- phase B: (prism, node pair) ----
    if (TP == TF_PREC_FP64) {
        for (int t = tid; t < 36 * EPB; t += WG) {
            const int el = t / 36, ab = t % 36;
            const int e = e0 + el;
            if (e >= E) break;
            const int a = ab / 6, b = ab % 6;
            double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
            for (int g = 0; g < 6; ++g) {
                tf::block_elastic(sE[g * EPB + el], a, b, acc);
                if (VISC) tf::block_viscous(sE[g * EPB + el], sV[g * EPB + el], a, b, inv_dt, acc);
            }
            double* out = kelem + (size_t)e * 324 + ab * 9;
#pragma unroll
            for (int q = 0; q < 9; ++q) out[q] = acc[q];
        }
    } else {
        // matrix cores: one wavefront per prism for the elastic tangent
        for (int el = tid / 64; el < EPB; el += WG / 64) {
            const int e = e0 + el;
            if (e >= E) break;
            mfma_tangent<TP>(sE, el, e, kelem);
        }
        if (VISC) {  // the viscous blocks stay FP64 and are added afterwards
            __syncthreads();
            for (int t = tid; t < 36 * EPB; t += WG) {
                const int el = t / 36, ab = t % 36;
                const int e = e0 + el;
                if (e >= E) break;
                const int a = ab / 6, b = ab % 6;
                double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
                for (int g = 0; g < 6; ++g) tf::block_viscous(sE[g * EPB + el], sV[g * EPB + el], a, b, inv_dt, acc);
                double* out = kelem + (size_t)e * 324 + ab * 9;
                for (int q = 0; q < 9; ++q) out[q] += acc[q];
            }
        }
    }

    // ---- phase C: (prism, node) forces, negated as the reference stores them ----
    if (tid < 6 * EPB) {
        const int el = tid / 6, a = tid % 6;
        const int e = e0 + el;
        if (e < E) {
            double f[3] = {0, 0, 0};
#pragma unroll
            for (int g = 0; g < 6; ++g) {
                double r[3];
                tf::voigt_mv(sE[g * EPB + el].sig, sE[g * EPB + el].gn + 3 * a, r);
                f[0] += r[0]; f[1] += r[1]; f[2] += r[2];
                if (VISC) {
                    tf::voigt_mv(sV[g * EPB + el].sv, sE[g * EPB + el].gn + 3 * a, r);
                    f[0] += r[0]; f[1] += r[1]; f[2] += r[2];
                }
            }
            double* out = felem + (size_t)e * 18 + 3 * a;
            out[0] = -f[0]; out[1] = -f[1]; out[2] = -f[2];
        }
    }
}

__global__ __launch_bounds__(256) void k_gather(const int32_t* __restrict__ blk_ptr, const int32_t* __restrict__ blk_src,
                                                const int32_t* __restrict__ blk_row, const int32_t* __restrict__ col,
                                                const double* __restrict__ kelem, const double* __restrict__ sa,
                                                const double* __restrict__ eta_ext, const uint8_t* __restrict__ fixed,
                                                const uint8_t* __restrict__ ablated,
                                                const double* __restrict__ pack_d, double inv_dt, int nb,
                                                double* __restrict__ val, float* __restrict__ val32,
                                                double* __restrict__ dinv)
{
    const int blk = blockIdx.x * blockDim.x + threadIdx.x;
    if (blk >= nb) return;
    const int A = blk_row[blk], B = col[blk];
    double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    const int k0 = blk_ptr[blk], k1 = blk_ptr[blk + 1];
    // four contributors per trip: their indices (clamped, always valid), then their 36 values are
    // loaded together, and the sums run in the contributors' order
    for (int k = k0; k < k1; k += 4) {
        int src[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) src[u] = blk_src[min(k + u, k1 - 1)];
        double v[4][9];
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int q = 0; q < 9; ++q) v[u][q] = kelem[(size_t)src[u] * 9 + q];
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (k + u < k1)
#pragma unroll
                for (int q = 0; q < 9; ++q) acc[q] += v[u][q];
    }
    if (A == B) {
        const double s = sa[A] * inv_dt;
#pragma unroll
        for (int i = 0; i < 3; ++i) acc[4 * i] += s * eta_ext[3 * A + i];
        // implicit packing, added before the ablated-node and fixed-DoF rules
        // (SC/Simulation.cpp:6489-6505)
        if (pack_d)
#pragma unroll
            for (int i = 0; i < 3; ++i) acc[4 * i] += pack_d[3 * A + i];
        // nodes without mass keep a solvable system (SC/NewtonRaphsonSolver.cpp:350-361)
        if (ablated && ablated[A])
#pragma unroll
            for (int i = 0; i < 3; ++i)
                if (acc[4 * i] == 0.0) acc[4 * i] = 1.0;
    }
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            if (fixed[3 * A + i] || fixed[3 * B + j]) acc[3 * i + j] = (A == B && i == j) ? 1.0 : 0.0;
#pragma unroll
    for (int q = 0; q < 9; ++q) val[(size_t)q * nb + blk] = acc[q];
    if (val32) {
#pragma unroll
        for (int q = 0; q < 9; ++q) val32[(size_t)q * nb + blk] = (float)acc[q];
    }
    if (A == B) {
        double inv[9];
        tf::inv3(acc, inv);
#pragma unroll
        for (int q = 0; q < 9; ++q) dinv[9 * A + q] = inv[q];
    }
}

// ---- slave/master binding (SC/NewtonRaphsonSolver.cpp:154-235) ----------------------------
// Fold targets: val[t] = const + sum of val_base[src] over the target's sources.  Slaves are
// never masters, so every target reads only unfolded values.
__global__ __launch_bounds__(256) void k_bind_values(const int32_t* __restrict__ fold_ptr,
                                                     const int64_t* __restrict__ fold_tgt,
                                                     const int64_t* __restrict__ fold_src,
                                                     const double* __restrict__ fold_const, int n,
                                                     const double* __restrict__ val_base, double* __restrict__ val,
                                                     float* __restrict__ val32)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    double acc = fold_const[t];
    for (int k = fold_ptr[t]; k < fold_ptr[t + 1]; ++k) acc += val_base[fold_src[k]];
    val[fold_tgt[t]] = acc;
    if (val32) val32[fold_tgt[t]] = (float)acc;
}

__global__ __launch_bounds__(64) void k_bind_dinv(const int32_t* __restrict__ nodes, int n,
                                                  const int32_t* __restrict__ diag_blk, const double* __restrict__ val,
                                                  int nb, double* __restrict__ dinv)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int A = nodes[k], j = diag_blk[A];
    double m[9], inv[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) m[q] = val[(size_t)q * nb + j];
    tf::inv3(m, inv);
#pragma unroll
    for (int q = 0; q < 9; ++q) dinv[9 * A + q] = inv[q];
}

// g_master += g_slave (slaves in binding order), g_slave = 0
__global__ __launch_bounds__(256) void k_bind_g(const int32_t* __restrict__ m_ptr, const int32_t* __restrict__ m_dof,
                                                const int32_t* __restrict__ s_dof, int n, double* __restrict__ g)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    double acc = g[m_dof[k]];
    for (int j = m_ptr[k]; j < m_ptr[k + 1]; ++j) {
        acc += g[s_dof[j]];
        g[s_dof[j]] = 0.0;
    }
    g[m_dof[k]] = acc;
}

// dU_slave = dU_master after the solve
__global__ __launch_bounds__(256) void k_bind_du(const int32_t* __restrict__ pairs, int n, double* __restrict__ du)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) du[pairs[2 * k]] = du[pairs[2 * k + 1]];
}

__global__ __launch_bounds__(256) void k_residual(const int32_t* __restrict__ node_ptr, const int32_t* __restrict__ node_src,
                                                  const double* __restrict__ felem, const double* __restrict__ uk,
                                                  const double* __restrict__ un, const double* __restrict__ sa,
                                                  const double* __restrict__ eta_ext, const uint8_t* __restrict__ fixed,
                                                  const double* __restrict__ pack_f, double inv_dt, int N,
                                                  double* __restrict__ g)
{
    const int A = blockIdx.x * blockDim.x + threadIdx.x;
    if (A >= N) return;
    double acc[3] = {0, 0, 0};
    const int k0 = node_ptr[A], k1 = node_ptr[A + 1];
    // eight contributors per trip (indices clamped, loads issued together), summed in their order
    for (int k = k0; k < k1; k += 8) {
        int src[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) src[u] = node_src[min(k + u, k1 - 1)];
        double v[8][3];
#pragma unroll
        for (int u = 0; u < 8; ++u)
#pragma unroll
            for (int i = 0; i < 3; ++i) v[u][i] = felem[(size_t)src[u] * 3 + i];
#pragma unroll
        for (int u = 0; u < 8; ++u)
            if (k + u < k1) { acc[0] += v[u][0]; acc[1] += v[u][1]; acc[2] += v[u][2]; }
    }
    const double s = sa[A];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const int d = 3 * A + i;
        const double vel = (uk[d] - un[d]) * inv_dt;
        const double gext = (s * eta_ext[d] * vel) * -1.0;
        const double gpack = pack_f ? pack_f[d] : 0.0;  // packing forces join gExt (SC/Simulation.cpp:8869-8906)
        g[d] = fixed[d] ? 0.0 : acc[i] + gext + gpack;
    }
}

// Element-local arrays of one prism in the reference layout (parity dumps).
// Prisms e0 .. e0 + n - 1, one workgroup each (prism e0 + k in workgroup k).  out holds n fe[18],
// then n fv[18], then n Ke[324] (TriPointKe), then n Kv[324] (TriPointKv, stored transposed); every
// entry is written (nothing accumulates), so out needs no clearing.
__global__ void k_element_dump(const int32_t* conn, const double* xref, const double* invfg, const double* lam,
                               const double* mu, const double* eta, const double* uk, const double* un,
                               const uint8_t* fixed, double inv_dt, int E, int e0, int n, double* out_base)
{
    const int k = blockIdx.x, e = e0 + k;
    double* o_fe = out_base + 18 * (size_t)k;
    double* o_fv = out_base + 18 * (size_t)n + 18 * (size_t)k;
    double* o_ke = out_base + 36 * (size_t)n + 324 * (size_t)k;
    double* o_kv = out_base + 360 * (size_t)n + 324 * (size_t)k;
    const int t = threadIdx.x;  // 36 lanes: one node pair each
    __shared__ tf::GPElastic sE[6];
    __shared__ tf::GPViscous sV[6];
    if (t < 6) {
        double X[18], x[18], Fi[9], v[18];
        for (int a = 0; a < 6; ++a) {
            const int id = conn[a * E + e];
            for (int i = 0; i < 3; ++i) {
                X[3 * a + i] = xref[(3 * a + i) * E + e];
                x[3 * a + i] = uk[3 * id + i];
                v[3 * a + i] = (uk[3 * id + i] - un[3 * id + i]) * inv_dt;
            }
        }
        for (int q = 0; q < 9; ++q) Fi[q] = invfg[q * E + e];
        tf::gp_eval(t, X, x, Fi, lam[e], mu[e], eta[e], v, sE[t], &sV[t]);
    }
    __syncthreads();
    if (t < 36) {
        const int a = t / 6, b = t % 6;
        double ke[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, kv[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        for (int g = 0; g < 6; ++g) {
            tf::block_elastic(sE[g], a, b, ke);
            if (eta[e] != 0.0) tf::block_viscous(sE[g], sV[g], a, b, inv_dt, kv);
        }
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                o_ke[(3 * a + i) * 18 + 3 * b + j] = ke[3 * i + j];
                o_kv[(3 * b + j) * 18 + 3 * a + i] = kv[3 * i + j];
            }
    }
    if (t < 6) {
        const int a = t;
        const int id = conn[a * E + e];
        double f[3] = {0, 0, 0}, fv[3] = {0, 0, 0};
        for (int g = 0; g < 6; ++g) {
            double r[3];
            tf::voigt_mv(sE[g].sig, sE[g].gn + 3 * a, r);
            for (int i = 0; i < 3; ++i) f[i] += r[i];
            if (eta[e] != 0.0) {
                tf::voigt_mv(sV[g].sv, sE[g].gn + 3 * a, r);
                for (int i = 0; i < 3; ++i) fv[i] += r[i];
            }
        }
        for (int i = 0; i < 3; ++i) {
            const bool fx = fixed[3 * id + i] != 0;
            o_fe[3 * a + i] = fx ? 0.0 : -f[i];
            o_fv[3 * a + i] = fx ? 0.0 : -fv[i];
        }
    }
}

// Rotation factor of the polar decomposition by scaled Newton iteration (Higham):
// R <- (gamma R + R^-T / gamma) / 2.  For det F > 0 this is U V^T of the SVD F = U S V^T.
__device__ inline void polar_rotation(const double* F, double* R)
{
    for (int q = 0; q < 9; ++q) R[q] = F[q];
    bool scale = true;
    for (int it = 0; it < 60; ++it) {
        double Ri[9];
        tf::inv3(R, Ri);
        double n1 = 0.0, n2 = 0.0;
        for (int q = 0; q < 9; ++q) { n1 += R[q] * R[q]; n2 += Ri[q] * Ri[q]; }
        const double gam = scale ? sqrt(sqrt(n2 / n1)) : 1.0;
        double diff = 0.0;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                const double nv = 0.5 * (gam * R[3 * i + j] + Ri[3 * j + i] / gam);
                const double d = nv - R[3 * i + j];
                diff += d * d;
                R[3 * i + j] = nv;
            }
        if (diff < 1e-4) scale = false;
        if (diff < 1e-34) break;
    }
}

// The mean deformation gradient over the six Gauss points (ShapeBase.cpp:1676-1689),
// Fm = 1/6 sum_g G_g G0_g^-1 with G_g = sum_a x_a dN_{g,a}^T, is linear in the positions:
// Fm = sum_a x_a c_a^T with c_a = 1/6 sum_g G0_g^-T dN_{g,a}, fixed by the reference geometry.
// One thread per prism, once per handle.
__global__ __launch_bounds__(256) void k_mean_f_weights(const double* __restrict__ xref, int E, double* __restrict__ fmw)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= E) return;
    double X[18], c[18];
    for (int q = 0; q < 18; ++q) { X[q] = xref[(size_t)q * E + e]; c[q] = 0.0; }
    for (int g = 0; g < 6; ++g) {
        double G0[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, IJ0[9], dN[6][3];
        for (int a = 0; a < 6; ++a) {
            tf::shape_deriv(g, a, dN[a]);
            for (int i = 0; i < 3; ++i)
                for (int k = 0; k < 3; ++k) G0[3 * i + k] += X[3 * a + i] * dN[a][k];
        }
        tf::inv3(G0, IJ0);
        for (int a = 0; a < 6; ++a)
            for (int j = 0; j < 3; ++j)
                c[3 * a + j] += (dN[a][0] * IJ0[j] + dN[a][1] * IJ0[3 + j] + dN[a][2] * IJ0[6 + j]) * (1.0 / 6.0);
    }
    for (int q = 0; q < 18; ++q) fmw[(size_t)q * E + e] = c[q];
}

__global__ __launch_bounds__(256) void k_grow(const int32_t* __restrict__ conn, const double* __restrict__ fmw,
                                              const double* __restrict__ pos, const double* __restrict__ rate,
                                              const double* __restrict__ ginc, double dt, int allow_rotation, int E,
                                              double* __restrict__ fg,
                                              double* __restrict__ invfg, double* __restrict__ rot_tz,
                                              int32_t* __restrict__ status)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= E) return;
    double applied_tz = 0.0;
    double inc[9] = {exp(rate[e] * dt), 0, 0, 0, exp(rate[E + e] * dt), 0, 0, 0, exp(rate[2 * E + e] * dt)};
    if (ginc)  // the product of the step's growth functions' increments, computed by the host
#pragma unroll
        for (int q = 0; q < 9; ++q) inc[q] = ginc[(size_t)q * E + e];
    if (allow_rotation) {
        // mean F at the current positions (ShapeBase.cpp:1676-1689): sum_a x_a c_a^T
        double Fm[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        for (int a = 0; a < 6; ++a) {
            const int id = conn[a * E + e];
            double x[3], c[3];
            for (int i = 0; i < 3; ++i) {
                x[i] = pos[3 * id + i];
                c[i] = fmw[(3 * a + i) * E + e];
            }
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) Fm[3 * i + j] += x[i] * c[j];
        }
        double R[9];
        polar_rotation(Fm, R);
        bool big = false;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                if (i != j && (R[3 * i + j] > 0.017 || R[3 * i + j] < -0.017)) big = true;
        if (big) {
            const double tz = atan2(R[1], R[0]);
            if (tz > 0.017 || tz < -0.017) {
                // increment <- GRM^T inc GRM with GRM = Rz(tz)^T (ShapeBase.cpp:1216-1222, 1697-1712)
                const double c = cos(tz), s = sin(tz);
                const double Rz[9] = {c, -s, 0, s, c, 0, 0, 0, 1};
                const double RzT[9] = {c, s, 0, -s, c, 0, 0, 0, 1};
                double T[9];
                tf::mm3(Rz, inc, T);
                tf::mm3(T, RzT, inc);
                applied_tz = tz;
                atomicAdd(&status[1], 1);
            }
        }
    }
    double F0[9], F1[9], Fi[9];
    rot_tz[e] = applied_tz;
    for (int q = 0; q < 9; ++q) F0[q] = fg[q * E + e];
    tf::mm3(inc, F0, F1);
    tf::inv3(F1, Fi);
    for (int q = 0; q < 9; ++q) {
        fg[q * E + e] = F1[q];
        invfg[q * E + e] = Fi[q];
    }
}

// exposed-face areas (Prism.cpp:2424-2503)
__device__ inline double face_area(const double* P0, const double* P1, const double* P2)
{
    const double Threshold = 1e-5;
    double s1 = 0.0, s2 = 0.0, ct = 0.0;
    for (int i = 0; i < 3; ++i) {
        const double a = P1[i] - P0[i], b = P2[i] - P0[i];
        ct += a * b; s1 += a * a; s2 += b * b;
    }
    double area = 0.0;
    if (s1 > Threshold && s2 > Threshold) {
        s1 = sqrt(s1); s2 = sqrt(s2);
        ct /= (s1 * s2);
        const double st2 = 1 - ct * ct;
        const double st = (st2 > Threshold) ? sqrt(st2) : 0.0;
        area = s1 * s2 * st / 2.0;
    }
    return area;
}

__global__ __launch_bounds__(256) void k_face_areas(const int32_t* __restrict__ conn, const uint8_t* __restrict__ exposed,
                                                    const double* __restrict__ pos, int E, double* __restrict__ area)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= E) return;
    double P[6][3];
    for (int a = 0; a < 6; ++a) {
        const int id = conn[a * E + e];
        for (int i = 0; i < 3; ++i) P[a][i] = pos[3 * id + i];
    }
    const uint8_t x = exposed[e];
    area[e] = (x & 1) ? face_area(P[0], P[1], P[2]) : 0.0;
    area[E + e] = (x & 2) ? face_area(P[3], P[4], P[5]) : 0.0;
}

__global__ __launch_bounds__(256) void k_node_surfaces(const int32_t* __restrict__ node_ptr, const int32_t* __restrict__ node_src,
                                                       const uint8_t* __restrict__ exposed, const double* __restrict__ area,
                                                       int E, int N, double* __restrict__ sa)
{
    const int A = blockIdx.x * blockDim.x + threadIdx.x;
    if (A >= N) return;
    double s = 0.0;
    for (int k = node_ptr[A]; k < node_ptr[A + 1]; ++k) {
        const int src = node_src[k];
        const int e = src / 6, a = src % 6;
        if (a >= 3) { if (exposed[e] & 2) s += area[E + e] / 3; }
        else if (exposed[e] & 1) s += area[e] / 3;
    }
    sa[A] = s;
}

__global__ void k_copy(double* dst, const double* src, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}
__global__ void k_fill(double* dst, double v, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = v;
}

__global__ __launch_bounds__(256) void k_update_uk(double* __restrict__ uk, const double* __restrict__ du,
                                                   const int32_t* __restrict__ tile_r0, const int32_t* __restrict__ tile_r1,
                                                   const int32_t* __restrict__ tile_seg, const SegScalars* __restrict__ seg,
                                                   const NrReport* __restrict__ gate)
{
    const int t = blockIdx.x;
    if (gate && !gate->ok) return;  // the host finishes the solve first
    if (!seg[tile_seg[t]].active) return;
    const int d0 = 3 * tile_r0[t], d1 = 3 * tile_r1[t];
    for (int d = d0 + threadIdx.x; d < d1; d += blockDim.x) uk[d] = uk[d] + du[d];
}

}  // namespace

// ------------------------------------------------------------------------------------------------
template <bool VISC, int TP>
static void launch_elements_t(tf_handle* h, double dt, int grid)
{
    hipLaunchKernelGGL((k_elements<VISC, TP>), dim3(grid), dim3(WG), 0, h->stream, h->conn, h->xref, h->invfg, h->lam,
                       h->mu, h->eta, h->uk, h->un, 1.0 / dt, h->E, h->kelem, h->felem, h->status);
}

void launch_assemble_elements(tf_handle* h, double dt)
{
    const int grid = (h->E + EPB - 1) / EPB;
    hipEvent_t ev[2];
    timer_begin(h, h->t_elem, ev);
    const int tp = tangent_precision_now(h);
    if (h->any_viscous) {
        if (tp == TF_PREC_FP32) launch_elements_t<true, TF_PREC_FP32>(h, dt, grid);
        else if (tp == TF_PREC_BF16) launch_elements_t<true, TF_PREC_BF16>(h, dt, grid);
        else launch_elements_t<true, TF_PREC_FP64>(h, dt, grid);
    } else {
        if (tp == TF_PREC_FP32) launch_elements_t<false, TF_PREC_FP32>(h, dt, grid);
        else if (tp == TF_PREC_BF16) launch_elements_t<false, TF_PREC_BF16>(h, dt, grid);
        else launch_elements_t<false, TF_PREC_FP64>(h, dt, grid);
    }
    TF_HIP(hipGetLastError());
    timer_end(h, h->t_elem, ev);
}

void launch_assemble_gather(tf_handle* h, double dt)
{
    hipEvent_t ev[2];
    timer_begin(h, h->t_gather, ev);
    launch_packing_nodes(h);
    hipLaunchKernelGGL(k_gather, dim3((h->nb + 255) / 256), dim3(256), 0, h->stream, h->blk_ptr, h->blk_src, h->blk_row,
                       h->col, h->kelem, h->sa, h->eta_ext, h->fixed, h->ablated, h->pk.n ? h->pk.d : nullptr, 1.0 / dt,
                       h->nb, h->val, fp32_values(h) ? h->val32 : nullptr, h->dinv);
    TF_HIP(hipGetLastError());
    launch_packing_offdiag(h);
    if (h->n_bound) launch_bind_matrix(h);
    timer_end(h, h->t_gather, ev);
}

void launch_residual(tf_handle* h, double dt)
{
    hipLaunchKernelGGL(k_residual, dim3((h->N + 255) / 256), dim3(256), 0, h->stream, h->node_ptr, h->node_src, h->felem,
                       h->uk, h->un, h->sa, h->eta_ext, h->fixed, h->pk.n ? h->pk.f : nullptr, 1.0 / dt, h->N, h->g);
    TF_HIP(hipGetLastError());
    if (h->n_bound) launch_bind_residual(h);
}

void launch_bind_matrix(tf_handle* h)
{
    launch_copy(h, h->val_base, h->val, 9 * (int64_t)h->nb);
    hipLaunchKernelGGL(k_bind_values, dim3((h->n_fold + 255) / 256), dim3(256), 0, h->stream, h->fold_ptr,
                       h->fold_tgt, h->fold_src, h->fold_const, h->n_fold, h->val_base, h->val,
                       fp32_values(h) ? h->val32 : nullptr);
    hipLaunchKernelGGL(k_bind_dinv, dim3((h->n_bind_nodes + 63) / 64), dim3(64), 0, h->stream, h->bind_nodes,
                       h->n_bind_nodes, h->diag_blk, h->val, h->nb, h->dinv);
    TF_HIP(hipGetLastError());
}

void launch_bind_residual(tf_handle* h)
{
    hipLaunchKernelGGL(k_bind_g, dim3((h->n_masters + 255) / 256), dim3(256), 0, h->stream, h->bind_m_ptr,
                       h->bind_m_dof, h->bind_s_dof, h->n_masters, h->g);
    TF_HIP(hipGetLastError());
}

void launch_bind_du(tf_handle* h)
{
    if (!h->n_bound) return;
    hipLaunchKernelGGL(k_bind_du, dim3((h->n_bound + 255) / 256), dim3(256), 0, h->stream, h->bind_pairs, h->n_bound,
                       h->du);
    TF_HIP(hipGetLastError());
}

void launch_element_dump(tf_handle* h, int32_t e0, int32_t n, double dt, double* d_out)
{
    hipLaunchKernelGGL(k_element_dump, dim3(n), dim3(64), 0, h->stream, h->conn, h->xref, h->invfg, h->lam, h->mu, h->eta,
                       h->uk, h->un, h->fixed, 1.0 / dt, h->E, e0, n, d_out);
    TF_HIP(hipGetLastError());
}

// ---- the element dump's static scratch (callers that cannot allocate) ---------------------------
// One copy per GPU, held until the process ends.  In host code `scratch` names a host-side stand-in,
// not device memory: its device address comes from hipGetSymbolAddress, once per device with that
// device current (tf_create), and only that address is handed to kernels and copies.
static __device__ __align__(256) double scratch[(size_t)TF_DUMP_SCRATCH_PRISMS * TF_DUMP_DOUBLES];
static std::mutex g_scratch_mu;
static double* g_scratch_addr[TF_MAX_DEVICES] = {};

void scratch_init(int device)
{
    if (device < 0 || device >= TF_MAX_DEVICES) throw std::runtime_error("device index beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_addr[device]) return;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) throw tf_hip_error(e != hipSuccess ? e : hipErrorInvalidSymbol,
                                                  "hipGetSymbolAddress(scratch)", __FILE__, __LINE__);
    g_scratch_addr[device] = static_cast<double*>(p);
}

double* scratch_addr(int device)
{
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    return device >= 0 && device < TF_MAX_DEVICES ? g_scratch_addr[device] : nullptr;
}

void launch_grow(tf_handle* h, double dt)
{
    if (h->gf.n_fn > 0) growth_functions_eval(h, dt);  // the step's increments into ginc
    hipLaunchKernelGGL(k_grow, dim3((h->E + 255) / 256), dim3(256), 0, h->stream, h->conn, h->fmw, h->uk, h->rate,
                       h->ginc, dt, h->opts.growth_rotation, h->E, h->fg, h->invfg, h->rot_tz, h->status);
    TF_HIP(hipGetLastError());
}

void launch_mean_f_weights(tf_handle* h)
{
    if (h->E <= 0) return;
    hipLaunchKernelGGL(k_mean_f_weights, dim3((h->E + 255) / 256), dim3(256), 0, h->stream, h->xref, h->E, h->fmw);
    TF_HIP(hipGetLastError());
}

void launch_surfaces(tf_handle* h)
{
    hipLaunchKernelGGL(k_face_areas, dim3((h->E + 255) / 256), dim3(256), 0, h->stream, h->conn, h->exposed, h->uk, h->E,
                       h->face_area);
    hipLaunchKernelGGL(k_node_surfaces, dim3((h->N + 255) / 256), dim3(256), 0, h->stream, h->node_ptr, h->node_src,
                       h->exposed, h->face_area, h->E, h->N, h->sa);
    TF_HIP(hipGetLastError());
}

void launch_copy(tf_handle* h, double* dst, const double* src, int64_t n)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_copy, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, h->stream, dst, src, n);
    TF_HIP(hipGetLastError());
}

void launch_fill(tf_handle* h, double* dst, double v, int64_t n)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_fill, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, h->stream, dst, v, n);
    TF_HIP(hipGetLastError());
}

void launch_update_uk(tf_handle* h, const NrReport* gate)
{
    hipLaunchKernelGGL(k_update_uk, dim3(h->T), dim3(256), 0, h->stream, h->uk, h->du, h->tile_r0, h->tile_r1,
                       h->tile_seg, h->seg, gate);
    TF_HIP(hipGetLastError());
}
