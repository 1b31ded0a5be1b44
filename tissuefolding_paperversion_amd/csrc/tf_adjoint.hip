// tf_adjoint.hip -- discrete adjoint of the implicit timestep (no counterpart in the reference).
//
// One timestep maps (x_n, Fg_n; rates, lambda, mu) to (x_{n+1}, Fg_{n+1}):
//   Fg_{n+1} = inc(rates, theta_z) Fg_n                 growth (ShapeBase.cpp:1215-1251)
//   S_A      = surfaces(x_n)                             (Simulation.cpp:7563-7581)
//   g(x_{n+1}; x_n, S_A, Fg_{n+1}, lambda, mu) = 0     the converged NR residual
// with g = -dPi/dx - S_A eta_ext (x_{n+1} - x_n)/dt on free DoFs (0 on fixed DoFs), where
// Pi = sum_g w D0 det(Fg) psi(Fe) is the neo-Hookean energy whose gradient is the reference's
// elastic force (Prism.cpp:843-846).  With M = -dg/dx_{n+1} (the assembled Jacobian) and
// y = M^-T xbar_{n+1} (fixed DoFs of xbar set aside), every input p of g receives
// pbar += y^T dg/dp.  For the energy terms y^T dg/dtheta = -D_y[dPi/dtheta]: the derivative of
// dPi/dtheta along y, computed here in tangent-linear form at every Gauss point:
//   dPi/dInvFg = w D0 detFg (F^T P - psi Fg^T),  P = mu (Fe - Fe^-T) + lambda lnJe Fe^-T
//   dPi/dlambda = w D0 detFg lnJe^2 / 2,        dPi/dmu = w D0 detFg ((tr C - 3)/2 - lnJe)
// The growth-rotation angle theta_z of each step is held fixed (it is a thresholded decision).
#include <hip/hip_runtime.h>

#include <cmath>

#include "tf_element.hpp"
#include "tf_internal.hpp"

namespace {

// y^T dg/d(InvFg, lambda, mu) per element at the handle's current state (uk, invfg)
__global__ __launch_bounds__(128) void k_adj_elements(const int32_t* __restrict__ conn, const double* __restrict__ xref,
                                                      const double* __restrict__ invfg, const double* __restrict__ lam,
                                                      const double* __restrict__ mu, const double* __restrict__ uk,
                                                      const double* __restrict__ y, int E, double* __restrict__ invfg_bar,
                                                      double* __restrict__ lam_bar, double* __restrict__ mu_bar)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= E) return;
    double X[18], x[18], yl[18], Fi[9], Fg[9];
    for (int a = 0; a < 6; ++a) {
        const int id = conn[a * E + e];
        for (int i = 0; i < 3; ++i) {
            X[3 * a + i] = xref[(3 * a + i) * E + e];
            x[3 * a + i] = uk[3 * id + i];
            yl[3 * a + i] = y[3 * id + i];
        }
    }
    for (int q = 0; q < 9; ++q) Fi[q] = invfg[q * E + e];
    const double detFi = tf::inv3(Fi, Fg);
    const double detFg = 1.0 / detFi;
    const double L = lam[e], Mu = mu[e];
    double dInv[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, dLam = 0.0, dMu = 0.0;
    for (int g = 0; g < 6; ++g) {
        double G0[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, G[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, dG[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        for (int a = 0; a < 6; ++a) {
            double d[3];
            tf::shape_deriv(g, a, d);
            for (int i = 0; i < 3; ++i)
                for (int k = 0; k < 3; ++k) {
                    G0[3 * i + k] += X[3 * a + i] * d[k];
                    G[3 * i + k] += x[3 * a + i] * d[k];
                    dG[3 * i + k] += yl[3 * a + i] * d[k];
                }
        }
        double IJ0[9], F[9], dF[9], Fe[9], dFe[9], FeI[9];
        const double D0 = tf::inv3(G0, IJ0);
        tf::mm3(G, IJ0, F);
        tf::mm3(dG, IJ0, dF);
        tf::mm3(F, Fi, Fe);
        tf::mm3(dF, Fi, dFe);
        const double Je = tf::inv3(Fe, FeI);
        const double lnJ = log(Je);
        double dlnJ = 0.0, trC = 0.0, dtrC = 0.0;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                dlnJ += FeI[3 * i + j] * dFe[3 * j + i];
                trC += Fe[3 * i + j] * Fe[3 * i + j];
                dtrC += 2.0 * Fe[3 * i + j] * dFe[3 * i + j];
            }
        // Fe^-T and its derivative -Fe^-T dFe^T Fe^-T
        double FeIT[9], dFeIT[9], T1[9];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) FeIT[3 * i + j] = FeI[3 * j + i];
        double dFeT[9];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) dFeT[3 * i + j] = dFe[3 * j + i];
        tf::mm3(FeIT, dFeT, T1);
        tf::mm3(T1, FeIT, dFeIT);
        for (int q = 0; q < 9; ++q) dFeIT[q] = -dFeIT[q];
        double P[9], dP[9];
        for (int q = 0; q < 9; ++q) {
            P[q] = Mu * (Fe[q] - FeIT[q]) + L * lnJ * FeIT[q];
            dP[q] = Mu * (dFe[q] - dFeIT[q]) + L * (dlnJ * FeIT[q] + lnJ * dFeIT[q]);
        }
        const double dpsi = 0.5 * Mu * dtrC - Mu * dlnJ + L * lnJ * dlnJ;
        const double c = D0 * detFg / 6.0;
        // d(F^T P) = dF^T P + F^T dP ; d(psi Fg^T) = dpsi Fg^T
        for (int p = 0; p < 3; ++p)
            for (int q = 0; q < 3; ++q) {
                double s = 0.0;
                for (int i = 0; i < 3; ++i) s += dF[3 * i + p] * P[3 * i + q] + F[3 * i + p] * dP[3 * i + q];
                dInv[3 * p + q] += c * (s - dpsi * Fg[3 * q + p]);
            }
        dLam += c * lnJ * dlnJ;
        dMu += c * (0.5 * dtrC - dlnJ);
        (void)trC;
    }
    for (int q = 0; q < 9; ++q) invfg_bar[(size_t)9 * e + q] = -dInv[q];
    lam_bar[e] = -dLam;
    mu_bar[e] = -dMu;
}

// d(area)/d(corner) of the reference's face area (Prism.cpp:2424-2503): 0 where the area is
// clipped to 0 by its thresholds, else the gradient of |a x b| / 2
__device__ inline void face_area_grad(const double* P0, const double* P1, const double* P2, double g0[3], double g1[3],
                                      double g2[3])
{
    const double Threshold = 1e-5;
    double a[3], b[3], ab = 0.0, aa = 0.0, bb = 0.0;
    for (int i = 0; i < 3; ++i) {
        a[i] = P1[i] - P0[i];
        b[i] = P2[i] - P0[i];
        ab += a[i] * b[i]; aa += a[i] * a[i]; bb += b[i] * b[i];
    }
    for (int i = 0; i < 3; ++i) g0[i] = g1[i] = g2[i] = 0.0;
    if (!(aa > Threshold && bb > Threshold)) return;
    const double ct = ab / (sqrt(aa) * sqrt(bb));
    const double st2 = 1 - ct * ct;
    if (!(st2 > Threshold)) return;
    const double area = sqrt(aa) * sqrt(bb) * sqrt(st2) / 2.0;
    const double k = 1.0 / (4.0 * area);
    for (int i = 0; i < 3; ++i) {
        g1[i] = k * (bb * a[i] - ab * b[i]);
        g2[i] = k * (aa * b[i] - ab * a[i]);
        g0[i] = -(g1[i] + g2[i]);
    }
}

// weight of node B's surface S_A(B) in y^T g: sum over free DoFs of -eta_ext v y
__device__ inline double surface_weight(int B, const double* uk, const double* un, const double* eta_ext,
                                        const uint8_t* fixed, const double* y, double inv_dt)
{
    double c = 0.0;
    for (int i = 0; i < 3; ++i) {
        const int d = 3 * B + i;
        if (!fixed[d]) c -= eta_ext[d] * (uk[d] - un[d]) * inv_dt * y[d];
    }
    return c;
}

// y^T dg/dx_n per DoF: the external-viscosity velocity term and the surfaces S_A(x_n)
__global__ __launch_bounds__(256) void k_adj_un(const int32_t* __restrict__ node_ptr, const int32_t* __restrict__ node_src,
                                                const int32_t* __restrict__ conn, const uint8_t* __restrict__ exposed,
                                                const double* __restrict__ uk, const double* __restrict__ un,
                                                const double* __restrict__ sa, const double* __restrict__ eta_ext,
                                                const uint8_t* __restrict__ fixed, const double* __restrict__ y,
                                                double inv_dt, int E, int N, double* __restrict__ un_bar)
{
    const int A = blockIdx.x * blockDim.x + threadIdx.x;
    if (A >= N) return;
    double acc[3];
    for (int i = 0; i < 3; ++i) {
        const int d = 3 * A + i;
        acc[i] = fixed[d] ? 0.0 : sa[A] * eta_ext[d] * inv_dt * y[d];
    }
    for (int k = node_ptr[A]; k < node_ptr[A + 1]; ++k) {
        const int src = node_src[k];
        const int e = src / 6, a = src % 6;
        const int f = a / 3;  // 0 = basal face (corners 0-2), 1 = apical face (corners 3-5)
        if (!(exposed[e] & (1 << f))) continue;
        int id[3];
        double P[3][3];
        for (int c = 0; c < 3; ++c) {
            id[c] = conn[(3 * f + c) * E + e];
            for (int i = 0; i < 3; ++i) P[c][i] = un[3 * id[c] + i];
        }
        const double s = (surface_weight(id[0], uk, un, eta_ext, fixed, y, inv_dt) +
                          surface_weight(id[1], uk, un, eta_ext, fixed, y, inv_dt) +
                          surface_weight(id[2], uk, un, eta_ext, fixed, y, inv_dt)) / 3.0;
        double g[3][3];
        face_area_grad(P[0], P[1], P[2], g[0], g[1], g[2]);
        const double* gg = g[a % 3];
        for (int i = 0; i < 3; ++i) acc[i] += s * gg[i];
    }
    for (int i = 0; i < 3; ++i) un_bar[3 * A + i] = acc[i];
}

// growth step: Fg_{n+1} = inc Fg_n, InvFg_{n+1} = Fg_{n+1}^-1, inc = Rz diag(exp(r dt)) Rz^T.
// fg_bar: in = dL/dFg_{n+1} from later steps, out = dL/dFg_n; rate_bar accumulates dL/drate.
__global__ __launch_bounds__(256) void k_adj_growth(const double* __restrict__ rate, double dt,
                                                    const double* __restrict__ theta_z, const double* __restrict__ fg_prev,
                                                    const double* __restrict__ invfg, const double* __restrict__ invfg_bar,
                                                    int E, double* __restrict__ fg_bar, double* __restrict__ rate_bar)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= E) return;
    double Fi[9], Fib[9], Fb[9], F0[9];
    for (int q = 0; q < 9; ++q) {
        Fi[q] = invfg[q * E + e];
        Fib[q] = invfg_bar[(size_t)9 * e + q];
        Fb[q] = fg_bar[(size_t)9 * e + q];
        F0[q] = fg_prev[(size_t)9 * e + q];
    }
    // d(A^-1) = -A^-1 dA A^-1  =>  Abar += -A^-T (A^-1)bar A^-T
    double FiT[9], T[9], U[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) FiT[3 * i + j] = Fi[3 * j + i];
    tf::mm3(FiT, Fib, T);
    tf::mm3(T, FiT, U);
    for (int q = 0; q < 9; ++q) Fb[q] -= U[q];
    const double ex[3] = {exp(rate[e] * dt), exp(rate[E + e] * dt), exp(rate[2 * E + e] * dt)};
    const double c = cos(theta_z[e]), s = sin(theta_z[e]);
    const double Rz[9] = {c, -s, 0, s, c, 0, 0, 0, 1};
    const double RzT[9] = {c, s, 0, -s, c, 0, 0, 0, 1};
    const double D[9] = {ex[0], 0, 0, 0, ex[1], 0, 0, 0, ex[2]};
    double inc[9];
    tf::mm3(Rz, D, T);
    tf::mm3(T, RzT, inc);
    // Fg_n bar = inc^T Fg_{n+1} bar ; inc bar = Fg_{n+1} bar Fg_n^T ; D bar = Rz^T inc bar Rz
    double incT[9], F0T[9], incb[9], Db[9], out[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            incT[3 * i + j] = inc[3 * j + i];
            F0T[3 * i + j] = F0[3 * j + i];
        }
    tf::mm3(incT, Fb, out);
    tf::mm3(Fb, F0T, incb);
    tf::mm3(RzT, incb, T);
    tf::mm3(T, Rz, Db);
    for (int q = 0; q < 9; ++q) fg_bar[(size_t)9 * e + q] = out[q];
    for (int i = 0; i < 3; ++i) rate_bar[(size_t)3 * e + i] += Db[4 * i] * dt * ex[i];
}

// y = M^T x through the block transpose map (block (A,B) of M^T is block (B,A) of M, transposed)
__global__ __launch_bounds__(256) void k_spmv_t(const int32_t* __restrict__ row_ptr, const int32_t* __restrict__ col,
                                                const int32_t* __restrict__ blk_tr, const double* __restrict__ val,
                                                int nb, int N, const double* __restrict__ x, double* __restrict__ y)
{
    const int d = blockIdx.x * blockDim.x + threadIdx.x;
    if (d >= 3 * N) return;
    const int A = d / 3, i = d % 3;
    double acc = 0.0;
    for (int j = row_ptr[A]; j < row_ptr[A + 1]; ++j) {
        const int B = col[j], jt = blk_tr[j];
        for (int k = 0; k < 3; ++k) acc += val[(size_t)(3 * k + i) * nb + jt] * x[3 * B + k];
    }
    y[d] = acc;
}

__global__ __launch_bounds__(256) void k_transpose_vals(const int32_t* __restrict__ blk_tr, const double* __restrict__ val,
                                                        int nb, double* __restrict__ val_t)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nb) return;
    const int jt = blk_tr[j];
    for (int q = 0; q < 9; ++q) val_t[(size_t)q * nb + j] = val[(size_t)(3 * (q % 3) + q / 3) * nb + jt];
}

__global__ __launch_bounds__(256) void k_dinv_all(const int32_t* __restrict__ diag_blk, const double* __restrict__ val,
                                                  int nb, int N, double* __restrict__ dinv)
{
    const int A = blockIdx.x * blockDim.x + threadIdx.x;
    if (A >= N) return;
    const int j = diag_blk[A];
    double m[9], inv[9];
    for (int q = 0; q < 9; ++q) m[q] = val[(size_t)q * nb + j];
    tf::inv3(m, inv);
    for (int q = 0; q < 9; ++q) dinv[9 * A + q] = inv[q];
}

// xbar split: fixed DoFs pass straight to x_n; the rest is the right-hand side of the solve
__global__ __launch_bounds__(256) void k_adj_split(const uint8_t* __restrict__ fixed, const double* __restrict__ xbar,
                                                   int n, double* __restrict__ rhs, double* __restrict__ pass)
{
    const int d = blockIdx.x * blockDim.x + threadIdx.x;
    if (d >= n) return;
    const bool f = fixed[d] != 0;
    rhs[d] = f ? 0.0 : xbar[d];
    pass[d] = f ? xbar[d] : 0.0;
}

__global__ __launch_bounds__(256) void k_accum(const double* __restrict__ a, double* __restrict__ b, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] += a[i];
}

inline unsigned blocks(int64_t n, int wg) { return (unsigned)((n + wg - 1) / wg); }

}  // namespace

void launch_adj_elements(tf_handle* h, const double* y, double* invfg_bar, double* lam_bar, double* mu_bar)
{
    hipLaunchKernelGGL(k_adj_elements, dim3(blocks(h->E, 128)), dim3(128), 0, h->stream, h->conn, h->xref, h->invfg,
                       h->lam, h->mu, h->uk, y, h->E, invfg_bar, lam_bar, mu_bar);
    TF_HIP(hipGetLastError());
}

void launch_adj_un(tf_handle* h, double dt, const double* y, double* un_bar)
{
    hipLaunchKernelGGL(k_adj_un, dim3(blocks(h->N, 256)), dim3(256), 0, h->stream, h->node_ptr, h->node_src, h->conn,
                       h->exposed, h->uk, h->un, h->sa, h->eta_ext, h->fixed, y, 1.0 / dt, h->E, h->N, un_bar);
    TF_HIP(hipGetLastError());
}

void launch_adj_growth(tf_handle* h, double dt, const double* theta_z, const double* fg_prev, const double* invfg_bar,
                       double* fg_bar, double* rate_bar)
{
    hipLaunchKernelGGL(k_adj_growth, dim3(blocks(h->E, 256)), dim3(256), 0, h->stream, h->rate, dt, theta_z, fg_prev,
                       h->invfg, invfg_bar, h->E, fg_bar, rate_bar);
    TF_HIP(hipGetLastError());
}

void launch_spmv_t(tf_handle* h, const double* x, double* y)
{
    hipLaunchKernelGGL(k_spmv_t, dim3(blocks(3 * (int64_t)h->N, 256)), dim3(256), 0, h->stream, h->row_ptr, h->col,
                       h->blk_tr, h->val, h->nb, h->N, x, y);
    TF_HIP(hipGetLastError());
}

void launch_transpose_vals(tf_handle* h, double* val_t)
{
    hipLaunchKernelGGL(k_transpose_vals, dim3(blocks(h->nb, 256)), dim3(256), 0, h->stream, h->blk_tr, h->val, h->nb,
                       val_t);
    TF_HIP(hipGetLastError());
}

void launch_dinv_all(tf_handle* h)
{
    hipLaunchKernelGGL(k_dinv_all, dim3(blocks(h->N, 256)), dim3(256), 0, h->stream, h->diag_blk, h->val, h->nb, h->N,
                       h->dinv);
    TF_HIP(hipGetLastError());
}

void launch_adj_split(tf_handle* h, const double* xbar, double* rhs, double* pass)
{
    hipLaunchKernelGGL(k_adj_split, dim3(blocks(3 * (int64_t)h->N, 256)), dim3(256), 0, h->stream, h->fixed, xbar,
                       3 * h->N, rhs, pass);
    TF_HIP(hipGetLastError());
}

void launch_accumulate(tf_handle* h, const double* a, double* b, int64_t n)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_accum, dim3(blocks(n, 256)), dim3(256), 0, h->stream, a, b, n);
    TF_HIP(hipGetLastError());
}
