// tf_element.hpp -- per-prism, per-Gauss-point FP64 mechanics shared by the assembly kernels.
//
// Mathematics (reference functions in TissueFolding/TissueFolding/SourceCode):
//   Gauss rule and dN/dxi                         Prism.cpp:123-148, :655-680
//   reference Jacobian G0 = dX/dxi, D0, G0^-1     Prism.cpp:705-728
//   F = G G0^-1, Fe = F Fg^-1, Je = det Fe        Prism.cpp:764-790
//   sigma = Je^-1 Fe S Fe^T, S neo-Hookean        ShapeBase.cpp:2284-2341
//         = (mu/Je)(Fe Fe^T - I) + (lambda lnJe/Je) I   (Fe C^-1 Fe^T = I)
//   grad N_a = G^-T dN_a/dxi (current config.)    ShapeBase.cpp:2378-2411
//   f_a = sum_g w J D0 sigma grad N_a             Prism.cpp:843-846
//   elastic tangent, push-forward of the 729-term KIntegral1 (ShapeBase.cpp:2786-2856):
//     K1_ab(i,k) = w J D0/Je [lambda gNa_i gNb_k + (mu - lambda lnJe)(d_ik gNa.gNb + gNa_k gNb_i)]
//   KIntegral2 (ShapeBase.cpp:2859-2890): d_ik w J D0 gNa^T sigma gNb
//   internal viscosity (ShapeBase.cpp:2413-2617, :2892-3053), natural orientation:
//     l = B v, d = sym, sigma_v = eta d, f^v_a = w J D0 sigma_v gNa
//     Kv1_ab = eta w J D0 Q_ab (I/dt - L),  Q_ab(i,j) = (B_a^T B~_b)_ij = 1/2 (d_ij gNa.gNb + gNa_j gNb_i)
//     Kv2_ab = w J D0 sigma_v (gNa gNb^T - gNb gNa^T)
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

namespace tf {

// Gauss point g: (eta, nu, zeta), weight 1/6 (Prism.cpp:123-148)
__host__ __device__ inline void gauss_point(int g, double& eta, double& nu, double& zeta)
{
    const double s = 0.57735026918962584;  // 1/sqrt(3)
    int r = g % 3;
    eta = (r == 1) ? (2.0 / 3.0) : (1.0 / 6.0);
    nu = (r == 2) ? (2.0 / 3.0) : (1.0 / 6.0);
    zeta = (g < 3) ? s : -s;
}

// dN_a/d(eta, nu, zeta) for node a at Gauss point g (Prism.cpp:655-680)
__host__ __device__ inline void shape_deriv(int g, int a, double d[3])
{
    double eta, nu, zeta;
    gauss_point(g, eta, nu, zeta);
    const double alpha = 0.5 * (1.0 - zeta), beta = 0.5 * (1.0 + zeta), lam = 1.0 - eta - nu;
    const double sgn = (a < 3) ? -1.0 : 1.0;
    const double ab = (a < 3) ? alpha : beta;
    const int c = a % 3;
    d[0] = (c == 0) ? -ab : (c == 1 ? ab : 0.0);
    d[1] = (c == 0) ? -ab : (c == 2 ? ab : 0.0);
    d[2] = sgn * 0.5 * ((c == 0) ? lam : (c == 1 ? eta : nu));
}

__host__ __device__ inline double det3(const double* M)
{
    return M[0] * (M[4] * M[8] - M[5] * M[7]) - M[1] * (M[3] * M[8] - M[5] * M[6]) + M[2] * (M[3] * M[7] - M[4] * M[6]);
}

// inverse by the adjugate; returns the determinant
__host__ __device__ inline double inv3(const double* M, double* I)
{
    double c00 = M[4] * M[8] - M[5] * M[7];
    double c01 = M[5] * M[6] - M[3] * M[8];
    double c02 = M[3] * M[7] - M[4] * M[6];
    double det = M[0] * c00 + M[1] * c01 + M[2] * c02;
    double id = 1.0 / det;
    I[0] = c00 * id;
    I[1] = (M[2] * M[7] - M[1] * M[8]) * id;
    I[2] = (M[1] * M[5] - M[2] * M[4]) * id;
    I[3] = c01 * id;
    I[4] = (M[0] * M[8] - M[2] * M[6]) * id;
    I[5] = (M[2] * M[3] - M[0] * M[5]) * id;
    I[6] = c02 * id;
    I[7] = (M[1] * M[6] - M[0] * M[7]) * id;
    I[8] = (M[0] * M[4] - M[1] * M[3]) * id;
    return det;
}

__host__ __device__ inline void mm3(const double* A, const double* B, double* C)
{
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) C[3 * i + j] = A[3 * i] * B[j] + A[3 * i + 1] * B[3 + j] + A[3 * i + 2] * B[6 + j];
}

// Gauss-point data staged for the tangent and force phases. All scalars carry w*J*D0.
struct GPElastic {
    double gn[18];   // grad N_a (a-major)
    double clam;     // w J D0 lambda / Je
    double cmu;      // w J D0 (mu - lambda lnJe) / Je
    double sig[6];   // w J D0 sigma (xx, yy, zz, xy, yz, xz)
};
struct GPViscous {
    double ceta;     // w J D0 eta
    double sv[6];    // w J D0 sigma_v (xx, yy, zz, xy, yz, xz)
    double L[9];     // velocity gradient L_ij = sum_c v_c,i gN_c,j
};

// Kinematics and stresses at one Gauss point.  X, x: 18 reference / current coordinates,
// v: 18 nodal velocities (uk - un)/dt or nullptr.  Returns Je.
__host__ __device__ inline double gp_eval(int g, const double* X, const double* x, const double* invfg, double lam,
                                          double mu, double eta, const double* v, GPElastic& E, GPViscous* V)
{
    double G0[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, G[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    double dN[6][3];
#pragma unroll
    for (int a = 0; a < 6; ++a) {
        shape_deriv(g, a, dN[a]);
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                G0[3 * i + k] += X[3 * a + i] * dN[a][k];
                G[3 * i + k] += x[3 * a + i] * dN[a][k];
            }
    }
    double IJ0[9], F[9], Fe[9], Gi[9];
    const double D0 = inv3(G0, IJ0);
    mm3(G, IJ0, F);
    mm3(F, invfg, Fe);
    const double Je = det3(Fe);
    const double lnJ = log(Je);
    const double detF = det3(F);
    inv3(G, Gi);
    const double dvw = (1.0 / 6.0) * detF * D0;
    // b = Fe Fe^T
    double b[6];
    b[0] = Fe[0] * Fe[0] + Fe[1] * Fe[1] + Fe[2] * Fe[2];
    b[1] = Fe[3] * Fe[3] + Fe[4] * Fe[4] + Fe[5] * Fe[5];
    b[2] = Fe[6] * Fe[6] + Fe[7] * Fe[7] + Fe[8] * Fe[8];
    b[3] = Fe[0] * Fe[3] + Fe[1] * Fe[4] + Fe[2] * Fe[5];
    b[4] = Fe[3] * Fe[6] + Fe[4] * Fe[7] + Fe[5] * Fe[8];
    b[5] = Fe[0] * Fe[6] + Fe[1] * Fe[7] + Fe[2] * Fe[8];
    const double iJ = 1.0 / Je;
    const double cs = dvw * mu * iJ, cv = dvw * (lam * lnJ - mu) * iJ;
    E.sig[0] = cs * b[0] + cv;
    E.sig[1] = cs * b[1] + cv;
    E.sig[2] = cs * b[2] + cv;
    E.sig[3] = cs * b[3];
    E.sig[4] = cs * b[4];
    E.sig[5] = cs * b[5];
    E.clam = dvw * lam * iJ;
    E.cmu = dvw * (mu - lam * lnJ) * iJ;
#pragma unroll
    for (int a = 0; a < 6; ++a)
#pragma unroll
        for (int m = 0; m < 3; ++m) E.gn[3 * a + m] = Gi[m] * dN[a][0] + Gi[3 + m] * dN[a][1] + Gi[6 + m] * dN[a][2];
    if (V) {
        double L[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
        for (int c = 0; c < 6; ++c)
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int j = 0; j < 3; ++j) L[3 * i + j] += v[3 * c + i] * E.gn[3 * c + j];
        const double ce = dvw * eta;
        V->ceta = ce;
        V->sv[0] = ce * L[0];
        V->sv[1] = ce * L[4];
        V->sv[2] = ce * L[8];
        V->sv[3] = ce * 0.5 * (L[1] + L[3]);
        V->sv[4] = ce * 0.5 * (L[5] + L[7]);
        V->sv[5] = ce * 0.5 * (L[2] + L[6]);
#pragma unroll
        for (int q = 0; q < 9; ++q) V->L[q] = L[q];
    }
    return Je;
}

// symmetric Voigt (xx, yy, zz, xy, yz, xz) times vector
__host__ __device__ inline void voigt_mv(const double* s, const double* a, double* r)
{
    r[0] = s[0] * a[0] + s[3] * a[1] + s[5] * a[2];
    r[1] = s[3] * a[0] + s[1] * a[1] + s[4] * a[2];
    r[2] = s[5] * a[0] + s[4] * a[1] + s[2] * a[2];
}

// Elastic block M_ab += K1' + K2 at one Gauss point (natural orientation, row i, col k)
__host__ __device__ inline void block_elastic(const GPElastic& E, int a, int b, double* acc)
{
    const double* A = E.gn + 3 * a;
    const double* B = E.gn + 3 * b;
    const double ab = A[0] * B[0] + A[1] * B[1] + A[2] * B[2];
    double sB[3];
    voigt_mv(E.sig, B, sB);
    const double diag = E.cmu * ab + (A[0] * sB[0] + A[1] * sB[1] + A[2] * sB[2]);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int k = 0; k < 3; ++k) acc[3 * i + k] += E.clam * A[i] * B[k] + E.cmu * A[k] * B[i] + (i == k ? diag : 0.0);
}

// Viscous block M_ab += Kv1 + Kv2 at one Gauss point (natural orientation)
__host__ __device__ inline void block_viscous(const GPElastic& E, const GPViscous& V, int a, int b, double inv_dt,
                                              double* acc)
{
    const double* A = E.gn + 3 * a;
    const double* B = E.gn + 3 * b;
    const double ab = A[0] * B[0] + A[1] * B[1] + A[2] * B[2];
    double Q[9], P[9];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            Q[3 * i + j] = 0.5 * ((i == j ? ab : 0.0) + A[j] * B[i]);
            P[3 * i + j] = (i == j ? inv_dt : 0.0) - V.L[3 * i + j];
        }
    double sA[3], sB[3];
    voigt_mv(V.sv, A, sA);
    voigt_mv(V.sv, B, sB);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            double qp = Q[3 * i] * P[j] + Q[3 * i + 1] * P[3 + j] + Q[3 * i + 2] * P[6 + j];
            acc[3 * i + j] += V.ceta * qp + (sA[i] * B[j] - sB[i] * A[j]);
        }
}

}  // namespace tf
