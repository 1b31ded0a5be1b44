/*
 * tf_oracle.c -- CPU restatement of TissueFolding's implicit Newton-Raphson hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the oracle: it is compiled into
 * oracle/_build/libtforacle.so and may only be loaded by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg, always as the
 * checker (or as the timed CPU baseline), never as the product path.
 *
 * It restates the reference algorithm step by step, in the reference's order
 * and orientation, without GSL: every routine below cites the reference
 * function it follows (paths relative to TissueFolding/TissueFolding/SourceCode).
 * The elastic tangent is the literal 729-term sum of the reference
 * (ShapeBase.cpp:2786-2856), NOT the closed form the GPU path uses, so that a
 * comparison of the two pins the push-forward identity.
 *
 * Pinning (what checks this file against something it did not produce):
 *   - element forces   : finite differences of the scalar energy
 *                        Pi = sum_g w_g D0_g det(Fg) psi(Fe)  (tests/test_oracle_element.py)
 *   - elastic tangent  : finite differences of the forces, symmetry,
 *                        translation invariance, closed-form push-forward
 *   - viscous forces   : closed form for an affine velocity field
 *   - viscous tangent  : finite differences of the viscous forces at fixed un
 *   - shape functions  : partition of unity, prism volume identity
 *   - areas            : closed-form triangle area
 *   - growth/rotation  : closed-form exponential growth, polar factor vs SVD
 *   - global assembly  : brute force on one- and two-prism meshes
 *   - NR driver        : exact stress-free solution under uniform growth
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <dlfcn.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define NN 6   /* nodes per prism  */
#define ND 3   /* space dimension  */
#define NG 6   /* Gauss points     */

/* ------------------------------------------------------------------------- */
/* Gauss rule and shape-function derivatives (Prism.cpp:123-148, :655-680)   */
/* ------------------------------------------------------------------------- */
void tfo_gauss_point(int g, double* eta, double* nu, double* zeta, double* w)
{
    /* Prism.cpp:123-148: (eta, nu, zeta), zeta = +-1/sqrt(3), weights 1/6 */
    static const double en[3][2] = {{1.0 / 6.0, 1.0 / 6.0}, {2.0 / 3.0, 1.0 / 6.0}, {1.0 / 6.0, 2.0 / 3.0}};
    double s = 1.0 / pow(3.0, 0.5);
    *eta = en[g % 3][0];
    *nu = en[g % 3][1];
    *zeta = (g < 3) ? s : -s;
    *w = 1.0 / 6.0;
}

/* dN[k][a]: derivative of N_a with respect to (eta, nu, zeta)[k]  (Prism.cpp:655-680) */
void tfo_shape_derivs(int g, double dN[3][NN])
{
    double eta, nu, zeta, w;
    tfo_gauss_point(g, &eta, &nu, &zeta, &w);
    double alpha = (1.0 - zeta) / 2.0, beta = (1.0 + zeta) / 2.0, lam = 1.0 - eta - nu;
    double r0[NN] = {-alpha, alpha, 0.0, -beta, beta, 0.0};
    double r1[NN] = {-alpha, 0.0, alpha, -beta, 0.0, beta};
    double r2[NN] = {-lam / 2.0, -eta / 2.0, -nu / 2.0, lam / 2.0, eta / 2.0, nu / 2.0};
    for (int a = 0; a < NN; ++a) { dN[0][a] = r0[a]; dN[1][a] = r1[a]; dN[2][a] = r2[a]; }
}

/* ------------------------------------------------------------------------- */
/* 3x3 helpers                                                               */
/* ------------------------------------------------------------------------- */
static void mm3(const double A[9], const double B[9], double C[9])
{
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double s = 0.0;
            for (int k = 0; k < 3; ++k) s += A[3 * i + k] * B[3 * k + j];
            C[3 * i + j] = s;
        }
}
static void mtm3(const double A[9], const double B[9], double C[9]) /* C = A^T B */
{
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double s = 0.0;
            for (int k = 0; k < 3; ++k) s += A[3 * k + i] * B[3 * k + j];
            C[3 * i + j] = s;
        }
}
static void mmt3(const double A[9], const double B[9], double C[9]) /* C = A B^T */
{
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double s = 0.0;
            for (int k = 0; k < 3; ++k) s += A[3 * i + k] * B[3 * j + k];
            C[3 * i + j] = s;
        }
}

/* ShapeBase::determinant3by3Matrix (ShapeBase.cpp:2050-2057): cofactor expansion on row 0 */
double tfo_det3(const double M[9])
{
    double det = M[0] * (M[4] * M[8] - M[5] * M[7]);
    det -= M[1] * (M[3] * M[8] - M[5] * M[6]);
    det += M[2] * (M[3] * M[7] - M[4] * M[6]);
    return det;
}

/* ShapeBase::InvertMatrix (ShapeBase.cpp:3767-3782): LU with partial pivoting, then
 * inverse by forward/back substitution on the unit vectors (the textbook LU inverse). */
void tfo_inv3(const double Min[9], double Inv[9])
{
    double A[9];
    int p[3] = {0, 1, 2};
    memcpy(A, Min, sizeof(A));
    for (int k = 0; k < 2; ++k) {
        int piv = k;
        double mx = fabs(A[3 * k + k]);
        for (int i = k + 1; i < 3; ++i)
            if (fabs(A[3 * i + k]) > mx) { mx = fabs(A[3 * i + k]); piv = i; }
        if (piv != k) {
            for (int j = 0; j < 3; ++j) { double t = A[3 * k + j]; A[3 * k + j] = A[3 * piv + j]; A[3 * piv + j] = t; }
            int t = p[k]; p[k] = p[piv]; p[piv] = t;
        }
        for (int i = k + 1; i < 3; ++i) {
            double l = A[3 * i + k] / A[3 * k + k];
            A[3 * i + k] = l;
            for (int j = k + 1; j < 3; ++j) A[3 * i + j] -= l * A[3 * k + j];
        }
    }
    for (int c = 0; c < 3; ++c) {
        double y[3];
        for (int i = 0; i < 3; ++i) {
            double s = (p[i] == c) ? 1.0 : 0.0;
            for (int k = 0; k < i; ++k) s -= A[3 * i + k] * y[k];
            y[i] = s;
        }
        for (int i = 2; i >= 0; --i) {
            double s = y[i];
            for (int k = i + 1; k < 3; ++k) s -= A[3 * i + k] * Inv[3 * k + c];
            Inv[3 * i + c] = s / A[3 * i + i];
        }
    }
}

/* ------------------------------------------------------------------------- */
/* Reference geometry of one prism (Prism.cpp:705-728)                       */
/* ------------------------------------------------------------------------- */
void tfo_reference_geometry(const double X[18], double IJ0[NG][9], double D0[NG])
{
    for (int g = 0; g < NG; ++g) {
        double dN[3][NN], dXde[9];
        tfo_shape_derivs(g, dN);
        /* dXde = (dN . X)^T : dXde[i][k] = sum_a X[a][i] dN[k][a] */
        for (int i = 0; i < 3; ++i)
            for (int k = 0; k < 3; ++k) {
                double s = 0.0;
                for (int a = 0; a < NN; ++a) s += dN[k][a] * X[3 * a + i];
                dXde[3 * i + k] = s;
            }
        D0[g] = tfo_det3(dXde);
        tfo_inv3(dXde, IJ0[g]);
    }
}

/* ------------------------------------------------------------------------- */
/* One prism: forces, TriPointKe, TriPointKv, TriPointF                      */
/*   ShapeBase::calculateForces3D     ShapeBase.cpp:2184-2235                 */
/*   Prism::calculateCurrNodalForces  Prism.cpp:749-872                       */
/*   ShapeBase::calculateImplicitKElastic / KIntegral1 / KIntegral2           */
/*                                    ShapeBase.cpp:2619-2639, 2786-2890      */
/*   ShapeBase::calculateImplicitKViscous / KvIntegral1 / KvIntegral2         */
/*                                    ShapeBase.cpp:2641-2705, 2892-3053      */
/* Outputs are in the reference's element-local layout:                      */
/*   fe, fv [18]  : ElementalElasticSystemForces / ElementalInternalViscous-  */
/*                  SystemForces (= -sum_g w_g f_g on non-fixed DoFs, 0 on    */
/*                  fixed ones), node-major 3a+i                              */
/*   Ke [324]     : TriPointKe, row 3a+i, column 3b+k                         */
/*   Kv [324]     : TriPointKv as stored by the reference (transposed         */
/*                  writes: currK[3b+j][3a+i] += Kvab(i,j))                   */
/*   TriPointF[9] : sum_g w_g F_g                                             */
/* ------------------------------------------------------------------------- */
void tfo_element(const double X[18], const double x[18], const double vdt[18], const double InvFg[9],
                 double lambda, double mu, double eta, double dt, const uint8_t* fixedLocal,
                 double fe[18], double fv[18], double Ke[324], double Kv[324], double TriPointF[9])
{
    double IJ0[NG][9], D0[NG];
    tfo_reference_geometry(X, IJ0, D0);

    double TriPointge[18] = {0}, TriPointgv[18] = {0};
    double Fe_g[NG][9], detF_g[NG], nablaN_g[NG][NN][3], sigma_g[NG][9], sigv_g[NG][9];
    double B_g[NG][6][18];
    double D81[NG][3][3][3][3];
    if (TriPointF) memset(TriPointF, 0, 9 * sizeof(double));

    for (int g = 0; g < NG; ++g) {
        double eta_, nu_, zeta_, w;
        tfo_gauss_point(g, &eta_, &nu_, &zeta_, &w);
        double dN[3][NN];
        tfo_shape_derivs(g, dN);
        /* Jacobian = (dN . x)^T  (Prism.cpp:764-766) */
        double J[9];
        for (int i = 0; i < 3; ++i)
            for (int k = 0; k < 3; ++k) {
                double s = 0.0;
                for (int a = 0; a < NN; ++a) s += dN[k][a] * x[3 * a + i];
                J[3 * i + k] = s;
            }
        double F[9], Fe[9];
        mm3(J, IJ0[g], F);        /* F = J . InvdXde   (:768) */
        mm3(F, InvFg, Fe);        /* Fe = F . InvFg    (:771) */
        memcpy(Fe_g[g], Fe, sizeof(Fe));
        /* C = Fe^T Fe (ShapeBase.cpp:2238-2243) */
        double C[9], InvC[9];
        mtm3(Fe, Fe, C);
        double detFe = tfo_det3(Fe);
        double lnJ = log(detFe);
        tfo_inv3(C, InvC);
        /* S = mu (I - C^-1) + lambda lnJ C^-1  (ShapeBase.cpp:2284-2297) */
        double S[9];
        for (int q = 0; q < 9; ++q) {
            double I = (q % 4 == 0) ? 1.0 : 0.0;
            S[q] = InvC[q] * (lambda * lnJ) + (I - InvC[q]) * mu;
        }
        /* D81 (ShapeBase.cpp:2299-2321) */
        double mult = 2.0 * (mu - lambda * lnJ);
        for (int I = 0; I < 3; ++I)
            for (int Jx = 0; Jx < 3; ++Jx)
                for (int K = 0; K < 3; ++K)
                    for (int L = 0; L < 3; ++L) {
                        double Iijkl = 0.5 * (InvC[3 * I + K] * InvC[3 * Jx + L] + InvC[3 * I + L] * InvC[3 * Jx + K]);
                        D81[g][I][Jx][K][L] = lambda * InvC[3 * I + Jx] * InvC[3 * K + L] + mult * Iijkl;
                    }
        /* sigma = detFe^-1 Fe S Fe^T (ShapeBase.cpp:2323-2341) */
        double T1[9], sig[9];
        mm3(Fe, S, T1);
        mmt3(T1, Fe, sig);
        for (int q = 0; q < 9; ++q) sig[q] *= 1.0 / detFe;
        memcpy(sigma_g[g], sig, sizeof(sig));
        double cs[6] = {sig[0], sig[4], sig[8], sig[1], sig[7], sig[2]};
        /* inverse Jacobian stack and B (ShapeBase.cpp:2378-2411):
         * invJShFuncDerS[3d+m][3a+d] = sum_n InvJ[n][m] dN[n][a] = dN_a/dx_m */
        double InvJ[9];
        tfo_inv3(J, InvJ);
        for (int a = 0; a < NN; ++a)
            for (int m = 0; m < 3; ++m) {
                double s = 0.0;
                for (int n = 0; n < 3; ++n) s += InvJ[3 * n + m] * dN[n][a];
                nablaN_g[g][a][m] = s;
            }
        /* B = CoeffMat . invJShFuncDerS (Prism.cpp:351-362): Voigt rows xx,yy,zz,xy,yz,xz */
        double (*B)[18] = B_g[g];
        memset(B, 0, sizeof(B_g[g]));
        for (int a = 0; a < NN; ++a) {
            const double* dn = nablaN_g[g][a];
            B[0][3 * a + 0] = dn[0];
            B[1][3 * a + 1] = dn[1];
            B[2][3 * a + 2] = dn[2];
            B[3][3 * a + 0] = dn[1]; B[3][3 * a + 1] = dn[0];
            B[4][3 * a + 1] = dn[2]; B[4][3 * a + 2] = dn[1];
            B[5][3 * a + 0] = dn[2]; B[5][3 * a + 2] = dn[0];
        }
        double detF = tfo_det3(F);
        detF_g[g] = detF;
        /* ge = B^T sigma detF detdXde  (Prism.cpp:843-846) */
        double currge[18];
        for (int r = 0; r < 18; ++r) {
            double s = 0.0;
            for (int q = 0; q < 6; ++q) s += (B[q][r] * detF * D0[g]) * cs[q];
            currge[r] = s;
        }
        memset(sigv_g[g], 0, sizeof(sigv_g[g]));
        double currgv[18] = {0};
        if (eta != 0.0) {
            /* l = B v_e, d, sigma_v = eta d, gv = B^T sigma_v detF detdXde (ShapeBase.cpp:2413-2617) */
            double l[6];
            for (int q = 0; q < 6; ++q) {
                double s = 0.0;
                for (int r = 0; r < 18; ++r) s += B[q][r] * vdt[r];
                l[q] = s;
            }
            double d[9] = {l[0], 0.5 * l[3], 0.5 * l[5], 0.5 * l[3], l[1], 0.5 * l[4], 0.5 * l[5], 0.5 * l[4], l[2]};
            for (int q = 0; q < 9; ++q) sigv_g[g][q] = d[q] * eta;
            const double* sv = sigv_g[g];
            double csv[6] = {sv[0], sv[4], sv[8], sv[1], sv[7], sv[2]};
            for (int r = 0; r < 18; ++r) {
                double s = 0.0;
                for (int q = 0; q < 6; ++q) s += (B[q][r] * detF * D0[g]) * csv[q];
                currgv[r] = s;
            }
        }
        for (int r = 0; r < 18; ++r) {
            TriPointge[r] += currge[r] * w;
            TriPointgv[r] += currgv[r] * w;
        }
        if (TriPointF)
            for (int q = 0; q < 9; ++q) TriPointF[q] += F[q] * w;
    }
    /* elemental forces = -TriPoint on non-fixed DoFs (ShapeBase.cpp:2198-2226) */
    for (int r = 0; r < 18; ++r) {
        int fx = fixedLocal ? fixedLocal[r] : 0;
        fe[r] = fx ? 0.0 : -TriPointge[r];
        fv[r] = fx ? 0.0 : -TriPointgv[r];
    }

    /* ---- elastic tangent: literal KIntegral1 + KIntegral2 ---- */
    if (Ke) {
        memset(Ke, 0, 324 * sizeof(double));
        for (int g = 0; g < NG; ++g) {
            double eta_, nu_, zeta_, w;
            tfo_gauss_point(g, &eta_, &nu_, &zeta_, &w);
            double currK[324];
            memset(currK, 0, sizeof(currK));
            const double* Fe = Fe_g[g];
            double detFe = tfo_det3(Fe);
            double detF = detF_g[g], detdXde = D0[g];
            for (int a = 0; a < NN; ++a)
                for (int b = 0; b < NN; ++b) {
                    const double* DNa = nablaN_g[g][a];
                    const double* DNb = nablaN_g[g][b];
                    for (int i = 0; i < 3; ++i)
                        for (int k = 0; k < 3; ++k) {
                            double value = 0.0;
                            for (int j = 0; j < 3; ++j)
                                for (int l = 0; l < 3; ++l)
                                    for (int I = 0; I < 3; ++I)
                                        for (int Jx = 0; Jx < 3; ++Jx)
                                            for (int K = 0; K < 3; ++K)
                                                for (int L = 0; L < 3; ++L)
                                                    value += Fe[3 * i + I] * Fe[3 * j + Jx] * Fe[3 * k + K] * Fe[3 * l + L] *
                                                             D81[g][I][Jx][K][L] * DNb[l] * DNa[j];
                            value *= detF * detdXde;
                            value /= detFe;
                            currK[(3 * a + i) * 18 + 3 * b + k] += value;
                        }
                    /* KIntegral2: (DNa^T sigma DNb) detF detdXde on the diagonal of block (a,b) */
                    double t[3];
                    for (int c = 0; c < 3; ++c) {
                        double s = 0.0;
                        for (int i = 0; i < 3; ++i) s += DNa[i] * sigma_g[g][3 * i + c];
                        t[c] = s;
                    }
                    double v = 0.0;
                    for (int c = 0; c < 3; ++c) v += t[c] * DNb[c];
                    v = v * detF * detdXde;
                    for (int i = 0; i < 3; ++i) currK[(3 * a + i) * 18 + 3 * b + i] += v;
                }
            for (int q = 0; q < 324; ++q) Ke[q] += currK[q] * w;
        }
    }

    /* ---- viscous tangent (stored transposed, as the reference does) ---- */
    if (Kv) {
        memset(Kv, 0, 324 * sizeof(double));
        if (eta != 0.0) {
            for (int g = 0; g < NG; ++g) {
                double eta_, nu_, zeta_, w;
                tfo_gauss_point(g, &eta_, &nu_, &zeta_, &w);
                double currK[324];
                memset(currK, 0, sizeof(currK));
                /* velocity gradient L = sum_c v_c (x) DN_c (ShapeBase.cpp:2892-2919) */
                double Lv[9] = {0};
                for (int c = 0; c < NN; ++c)
                    for (int i = 0; i < 3; ++i)
                        for (int j = 0; j < 3; ++j) Lv[3 * i + j] += vdt[3 * c + i] * nablaN_g[g][c][j];
                double P[9];
                for (int q = 0; q < 9; ++q) P[q] = ((q % 4 == 0) ? 1.0 / dt : 0.0) - Lv[q];
                double dv = detF_g[g], dX = D0[g];
                const double (*B)[18] = (const double (*)[18])B_g[g];
                for (int a = 0; a < NN; ++a)
                    for (int b = 0; b < NN; ++b) {
                        /* KvIntegral1 (ShapeBase.cpp:2921-2984) */
                        double BaTBb[9];
                        for (int i = 0; i < 3; ++i)
                            for (int j = 0; j < 3; ++j) {
                                double s = 0.0;
                                for (int q = 0; q < 6; ++q) {
                                    double bb = B[q][3 * b + j];
                                    if (q >= 3) bb *= 0.5;
                                    s += B[q][3 * a + i] * bb;
                                }
                                BaTBb[3 * i + j] = s;
                            }
                        double K1[9];
                        mm3(BaTBb, P, K1);
                        for (int q = 0; q < 9; ++q) K1[q] = K1[q] * dv * dX * eta;
                        for (int i = 0; i < 3; ++i)
                            for (int j = 0; j < 3; ++j) currK[(3 * b + j) * 18 + 3 * a + i] += K1[3 * i + j];
                        /* KvIntegral2 (ShapeBase.cpp:2986-3053) */
                        const double* DNa = nablaN_g[g][a];
                        const double* DNb = nablaN_g[g][b];
                        double par[9], K2[9];
                        for (int i = 0; i < 3; ++i)
                            for (int j = 0; j < 3; ++j) par[3 * i + j] = DNa[i] * DNb[j] - DNb[i] * DNa[j];
                        mm3(sigv_g[g], par, K2);
                        for (int q = 0; q < 9; ++q) K2[q] = K2[q] * dv * dX;
                        for (int i = 0; i < 3; ++i)
                            for (int j = 0; j < 3; ++j) currK[(3 * b + j) * 18 + 3 * a + i] += K2[3 * i + j];
                    }
                for (int q = 0; q < 324; ++q) Kv[q] += currK[q] * w;
            }
        }
    }
}

/* ------------------------------------------------------------------------- */
/* Exposed-face areas (Prism.cpp:2424-2503) and nodal viscosity surfaces     */
/* (ShapeBase.cpp:4041-4077, Simulation.cpp:7563-7581)                       */
/* ------------------------------------------------------------------------- */
double tfo_face_area(const double* P0, const double* P1, const double* P2)
{
    const double Threshold = 1e-5;
    double v1[3], v2[3], costet = 0.0, s1 = 0.0, s2 = 0.0, area = 0.0;
    for (int i = 0; i < 3; ++i) {
        v1[i] = P1[i] - P0[i];
        v2[i] = P2[i] - P0[i];
        costet += v1[i] * v2[i];
        s1 += v1[i] * v1[i];
        s2 += v2[i] * v2[i];
    }
    if (s1 > Threshold && s2 > Threshold) {
        s1 = pow(s1, 0.5);
        s2 = pow(s2, 0.5);
        costet /= (s1 * s2);
        double sinTetSq = 1 - costet * costet;
        double sintet = 0;
        if (sinTetSq > Threshold) sintet = pow(sinTetSq, 0.5);
        area = s1 * s2 * sintet / 2.0;
    }
    return area;
}

/* exposed[e]: bit0 = basal face (nodes 0,1,2) exposed, bit1 = apical face (3,4,5) exposed
 * (Prism.cpp:2674-2711, columnar tissue).  Sa[n] accumulates area/3 in element order. */
void tfo_viscosity_surfaces(int n_nodes, int n_elem, const int32_t* conn, const uint8_t* exposed, const double* pos,
                            double* Sa)
{
    for (int n = 0; n < n_nodes; ++n) Sa[n] = 0.0;
    for (int e = 0; e < n_elem; ++e) {
        const int32_t* c = conn + 6 * e;
        if (exposed[e] & 2) {
            double A = tfo_face_area(pos + 3 * c[3], pos + 3 * c[4], pos + 3 * c[5]);
            for (int a = 3; a < 6; ++a) Sa[c[a]] += A / 3;
        }
        if (exposed[e] & 1) {
            double A = tfo_face_area(pos + 3 * c[0], pos + 3 * c[1], pos + 3 * c[2]);
            for (int a = 0; a < 3; ++a) Sa[c[a]] += A / 3;
        }
    }
}

/* ------------------------------------------------------------------------- */
/* Model and global assembly                                                 */
/* ------------------------------------------------------------------------- */
typedef struct {
    int n_nodes, n_elem;
    const int32_t* conn;     /* [6E]  element-major                     */
    const double* Xref;      /* [18E] element-local reference positions */
    const double* lambda;    /* [E] */
    const double* mu;        /* [E] */
    const double* eta;       /* [E] internal viscosity                  */
    const double* eta_ext;   /* [3N] external viscosity per DoF         */
    const uint8_t* fixed;    /* [3N] 1 = fixed DoF                      */
    const double* InvFg;     /* [9E] */
    const double* Sa;        /* [N] viscosity surface per node          */
} tfo_model;

/* Per-element work of calculateForcesAndJacobianMatrixNR (NewtonRaphsonSolver.cpp:257-271),
 * OpenMP over elements as the reference does.  Outputs element-local arrays. */
void tfo_elements(const tfo_model* m, const double* uk, const double* un, double dt, double* fe, double* fv, double* Ke,
                  double* Kv, double* TriPointF)
{
#pragma omp parallel for schedule(static)
    for (int e = 0; e < m->n_elem; ++e) {
        const int32_t* c = m->conn + 6 * e;
        double x[18], v[18];
        uint8_t fl[18];
        for (int a = 0; a < NN; ++a)
            for (int i = 0; i < 3; ++i) {
                x[3 * a + i] = uk[3 * c[a] + i];
                v[3 * a + i] = (uk[3 * c[a] + i] - un[3 * c[a] + i]) * (1.0 / dt);
                fl[3 * a + i] = m->fixed ? m->fixed[3 * c[a] + i] : 0;
            }
        tfo_element(m->Xref + 18 * e, x, v, m->InvFg + 9 * e, m->lambda[e], m->mu[e], m->eta ? m->eta[e] : 0.0, dt, fl,
                    fe + 18 * e, fv + 18 * e, Ke ? Ke + 324 * (size_t)e : NULL, Kv ? Kv + 324 * (size_t)e : NULL,
                    TriPointF ? TriPointF + 9 * e : NULL);
    }
}

/* residual: writeForcesTogeAndgvInternal (NewtonRaphsonSolver.cpp:273-279, ShapeBase.cpp:2148-2180),
 * calculateExternalViscousForcesForNR (:297-324), calculateSumOfInternalForces (:363-367),
 * gExt = 0, fixed DoFs zeroed by calcutateFixedK (:237-255). */
void tfo_residual(const tfo_model* m, const double* uk, const double* un, double dt, const double* fe, const double* fv,
                  double* g)
{
    int n = 3 * m->n_nodes;
    double* ge = (double*)calloc(n, sizeof(double));
    double* gv = (double*)calloc(n, sizeof(double));
    for (int e = 0; e < m->n_elem; ++e)
        for (int a = 0; a < NN; ++a)
            for (int j = 0; j < 3; ++j) {
                int I = 3 * m->conn[6 * e + a] + j;
                ge[I] = ge[I] + fe[18 * e + 3 * a + j];
                gv[I] = gv[I] + fv[18 * e + 3 * a + j];
            }
    for (int i = 0; i < n; ++i) {
        double vv = (uk[i] - un[i]) * (1.0 / dt);
        double gext = (m->Sa[i / 3] * m->eta_ext[i] * vv) * -1.0;
        g[i] = ge[i] + gv[i] + gext;
        if (m->fixed && m->fixed[i]) g[i] = 0.0;
    }
    free(ge);
    free(gv);
}

/* Dense K in the reference orientation (writeImplicitElementalKToJacobian, NewtonRaphsonSolver.cpp:281-295:
 * viscous first, then elastic; addImplicitKViscousExternalToJacobian :326-348; calcutateFixedK :237-255). */
void tfo_dense_K(const tfo_model* m, double dt, const double* Ke, const double* Kv, double* K)
{
    int n = 3 * m->n_nodes;
    memset(K, 0, (size_t)n * n * sizeof(double));
    for (int pass = 0; pass < 2; ++pass)
        for (int e = 0; e < m->n_elem; ++e) {
            if (pass == 0 && !(m->eta && m->eta[e] != 0.0)) continue;
            const double* T = (pass == 0 ? Kv : Ke) + 324 * (size_t)e;
            for (int a = 0; a < NN; ++a)
                for (int b = 0; b < NN; ++b)
                    for (int i = 0; i < 3; ++i)
                        for (int j = 0; j < 3; ++j) {
                            size_t r = 3 * (size_t)m->conn[6 * e + a] + i, c = 3 * (size_t)m->conn[6 * e + b] + j;
                            K[r * n + c] += T[(3 * a + i) * 18 + 3 * b + j];
                        }
        }
    for (int i = 0; i < m->n_nodes; ++i)
        for (int j = 0; j < 3; ++j) {
            size_t d = 3 * (size_t)i + j;
            K[d * n + d] += (m->Sa[i] / dt) * m->eta_ext[d];
        }
    if (m->fixed)
        for (int d = 0; d < n; ++d)
            if (m->fixed[d]) {
                for (int k = 0; k < n; ++k) {
                    double v = (d == k) ? 1.0 : 0.0;
                    K[(size_t)d * n + k] = v;
                    K[(size_t)k * n + d] = v;
                }
            }
}

/* CSR K in the reference orientation with a pattern built from connectivity ("reference-fair"
 * variant of the CPU path: same element math and same values as tfo_dense_K, without the 3N x 3N
 * dense intermediate).  row_ptr/col are 0-based, columns sorted; the pattern holds every
 * node-pair coupling of the mesh, so fixed DoFs keep their (zeroed) entries. */
void tfo_csr_K(const tfo_model* m, double dt, const double* Ke, const double* Kv, const int64_t* row_ptr,
               const int32_t* col, double* val)
{
    int n = 3 * m->n_nodes;
    memset(val, 0, (size_t)row_ptr[n] * sizeof(double));
    for (int pass = 0; pass < 2; ++pass)
        for (int e = 0; e < m->n_elem; ++e) {
            if (pass == 0 && !(m->eta && m->eta[e] != 0.0)) continue;
            const double* T = (pass == 0 ? Kv : Ke) + 324 * (size_t)e;
            for (int a = 0; a < NN; ++a)
                for (int i = 0; i < 3; ++i) {
                    int r = 3 * m->conn[6 * e + a] + i;
                    for (int b = 0; b < NN; ++b)
                        for (int j = 0; j < 3; ++j) {
                            int c = 3 * m->conn[6 * e + b] + j;
                            int64_t lo = row_ptr[r], hi = row_ptr[r + 1] - 1;
                            while (lo < hi) {
                                int64_t mid = (lo + hi) / 2;
                                if (col[mid] < c) lo = mid + 1; else hi = mid;
                            }
                            val[lo] += T[(3 * a + i) * 18 + 3 * b + j];
                        }
                }
        }
    for (int r = 0; r < n; ++r) {
        for (int64_t k = row_ptr[r]; k < row_ptr[r + 1]; ++k)
            if (col[k] == r) val[k] += (m->Sa[r / 3] / dt) * m->eta_ext[r];
    }
    if (m->fixed)
        for (int r = 0; r < n; ++r)
            for (int64_t k = row_ptr[r]; k < row_ptr[r + 1]; ++k) {
                int c = col[k];
                if (m->fixed[r] || m->fixed[c]) val[k] = (r == c) ? 1.0 : 0.0;
            }
}

/* constructiaForPardiso (NewtonRaphsonSolver.cpp:686-707): dense -> CSR scan dropping |K| <= 1e-13.
 * Returns nnz; with ia/ja/a NULL only counts. */
int64_t tfo_dense_to_csr(int n, const double* K, int64_t* ia, int32_t* ja, double* a)
{
    int64_t cnt = 0;
    for (int i = 0; i < n; ++i) {
        if (ia) ia[i] = cnt;
        for (int j = 0; j < n; ++j) {
            double v = K[(size_t)i * n + j];
            if (v > 1e-13 || v < -1e-13) {
                if (ja) { ja[cnt] = j; a[cnt] = v; }
                cnt++;
            }
        }
    }
    if (ia) ia[n] = cnt;
    return cnt;
}

/* Dense LU solve of K^T x = b (partial pivoting): the reference solves with the transposed
 * matrix (NewtonRaphsonSolver.cpp:636-643, iparm[11]=1).  K is overwritten. */
int tfo_dense_solve_transposed(int n, double* K, const double* b, double* x)
{
    /* factor A = K^T in place into LU of A; work on A stored as K transposed on the fly */
    double* A = (double*)malloc((size_t)n * n * sizeof(double));
    int* p = (int*)malloc(n * sizeof(int));
    if (!A || !p) { free(A); free(p); return -1; }
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) A[(size_t)i * n + j] = K[(size_t)j * n + i];
    for (int i = 0; i < n; ++i) p[i] = i;
    int rc = 0;
    for (int k = 0; k < n; ++k) {
        int piv = k;
        double mx = fabs(A[(size_t)k * n + k]);
        for (int i = k + 1; i < n; ++i) {
            double v = fabs(A[(size_t)i * n + k]);
            if (v > mx) { mx = v; piv = i; }
        }
        if (mx == 0.0) { rc = 1; break; }
        if (piv != k) {
            for (int j = 0; j < n; ++j) {
                double t = A[(size_t)k * n + j]; A[(size_t)k * n + j] = A[(size_t)piv * n + j]; A[(size_t)piv * n + j] = t;
            }
            int t = p[k]; p[k] = p[piv]; p[piv] = t;
        }
        double dk = A[(size_t)k * n + k];
#pragma omp parallel for schedule(static) if (n > 512)
        for (int i = k + 1; i < n; ++i) {
            double l = A[(size_t)i * n + k] / dk;
            A[(size_t)i * n + k] = l;
            if (l != 0.0)
                for (int j = k + 1; j < n; ++j) A[(size_t)i * n + j] -= l * A[(size_t)k * n + j];
        }
    }
    if (rc == 0) {
        for (int i = 0; i < n; ++i) {
            double s = b[p[i]];
            for (int k = 0; k < i; ++k) s -= A[(size_t)i * n + k] * x[k];
            x[i] = s;
        }
        for (int i = n - 1; i >= 0; --i) {
            double s = x[i];
            for (int k = i + 1; k < n; ++k) s -= A[(size_t)i * n + k] * x[k];
            x[i] = s / A[(size_t)i * n + i];
        }
    }
    free(A);
    free(p);
    return rc;
}

/* ------------------------------------------------------------------------- */
/* Sparse direct solve through Intel MKL PARDISO, loaded at run time.        */
/* The reference calls pardiso-project 6.0 with mtype 11, phases 11/22/33,   */
/* one refinement step, no scaling/matching and a transposed solve          */
/* (NewtonRaphsonSolver.cpp:418-683).  MKL is the nearest available relative */
/* and is only called through its documented C interface.                   */
/* ------------------------------------------------------------------------- */
typedef void (*pardiso_fn)(void*, const int*, const int*, const int*, const int*, const int*, const void*, const int*,
                           const int*, int*, const int*, int*, const int*, void*, void*, int*);
typedef void (*pardisoinit_fn)(void*, const int*, int*);
static void* g_mkl = NULL;
static pardiso_fn g_pardiso = NULL;
static pardisoinit_fn g_pardisoinit = NULL;

int tfo_pardiso_available(const char* path)
{
    if (g_pardiso) return 1;
    const char* cands[] = {path, "libmkl_rt.so", "libmkl_rt.so.2", "libmkl_rt.so.1", "/opt/conda/lib/libmkl_rt.so", NULL};
    for (int i = 0; i < 5 && !g_mkl; ++i) {
        if (!cands[i]) continue;
        g_mkl = dlopen(cands[i], RTLD_NOW | RTLD_GLOBAL);
    }
    if (!g_mkl) return 0;
    g_pardiso = (pardiso_fn)dlsym(g_mkl, "pardiso");
    g_pardisoinit = (pardisoinit_fn)dlsym(g_mkl, "pardisoinit");
    if (!g_pardiso || !g_pardisoinit) { g_pardiso = NULL; return 0; }
    return 1;
}

/* Solve K^T x = b for K in 0-based CSR (int32 row pointers as the reference uses). */
int tfo_pardiso_solve_transposed(int n, const int32_t* ia0, const int32_t* ja0, const double* a, const double* b,
                                 double* x, int nthreads)
{
    if (!g_pardiso) return -100;
    void* pt[64];
    int iparm[64];
    memset(pt, 0, sizeof(pt));
    memset(iparm, 0, sizeof(iparm));
    int mtype = 11, maxfct = 1, mnum = 1, nrhs = 1, msglvl = 0, error = 0, phase, idum = 0;
    g_pardisoinit(pt, &mtype, iparm);
    iparm[0] = 1;   /* user-set iparm */
    iparm[2] = nthreads;
    iparm[7] = 1;   /* one iterative refinement step */
    iparm[9] = 13;  /* default pivot perturbation for unsymmetric matrices */
    iparm[10] = 0;  /* no scaling  */
    iparm[12] = 0;  /* no matching */
    iparm[34] = 1;  /* zero-based indexing */
    double ddum = 0.0;
    phase = 11;
    g_pardiso(pt, &maxfct, &mnum, &mtype, &phase, &n, a, ia0, ja0, &idum, &nrhs, iparm, &msglvl, &ddum, &ddum, &error);
    if (error) goto release;
    phase = 22;
    g_pardiso(pt, &maxfct, &mnum, &mtype, &phase, &n, a, ia0, ja0, &idum, &nrhs, iparm, &msglvl, &ddum, &ddum, &error);
    if (error) goto release;
    phase = 33;
    iparm[11] = 2; /* solve with the transposed matrix */
    g_pardiso(pt, &maxfct, &mnum, &mtype, &phase, &n, a, ia0, ja0, &idum, &nrhs, iparm, &msglvl, (void*)b, x, &error);
release:;
    int err2 = 0;
    phase = -1;
    g_pardiso(pt, &maxfct, &mnum, &mtype, &phase, &n, &ddum, ia0, ja0, &idum, &nrhs, iparm, &msglvl, &ddum, &ddum, &err2);
    return error;
}

int tfo_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ------------------------------------------------------------------------- */
/* Growth update (ShapeBase.cpp:396-421, :567-586, :1215-1251, :1656-1811)  */
/* ------------------------------------------------------------------------- */
/* Jacobi eigen-decomposition of a symmetric 3x3 matrix: A = V diag(w) V^T */
static void jacobi_eig3(const double Ain[9], double w[3], double V[9])
{
    double A[9];
    memcpy(A, Ain, sizeof(A));
    for (int q = 0; q < 9; ++q) V[q] = (q % 4 == 0) ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 60; ++sweep) {
        double off = fabs(A[1]) + fabs(A[2]) + fabs(A[5]);
        if (off < 1e-300) break;
        for (int p = 0; p < 2; ++p)
            for (int q = p + 1; q < 3; ++q) {
                double apq = A[3 * p + q];
                if (fabs(apq) < 1e-300) continue;
                double app = A[3 * p + p], aqq = A[3 * q + q];
                double theta = (aqq - app) / (2.0 * apq);
                double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < 3; ++k) {
                    double akp = A[3 * k + p], akq = A[3 * k + q];
                    A[3 * k + p] = c * akp - s * akq;
                    A[3 * k + q] = s * akp + c * akq;
                }
                for (int k = 0; k < 3; ++k) {
                    double apk = A[3 * p + k], aqk = A[3 * q + k];
                    A[3 * p + k] = c * apk - s * aqk;
                    A[3 * q + k] = s * apk + c * aqk;
                }
                for (int k = 0; k < 3; ++k) {
                    double vkp = V[3 * k + p], vkq = V[3 * k + q];
                    V[3 * k + p] = c * vkp - s * vkq;
                    V[3 * k + q] = s * vkp + c * vkq;
                }
            }
    }
    w[0] = A[0]; w[1] = A[4]; w[2] = A[8];
}

/* Rotation factor of the polar decomposition F = R U, R = F U^-1 with U = (F^T F)^(1/2).
 * For det F > 0 this is the U V^T of the SVD F = U S V^T that calculate3DRotMatFromF forms
 * (ShapeBase.cpp:1775-1781). */
void tfo_polar_rotation(const double F[9], double R[9])
{
    double FtF[9], w[3], V[9], Uinv[9];
    mtm3(F, F, FtF);
    jacobi_eig3(FtF, w, V);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double s = 0.0;
            for (int k = 0; k < 3; ++k) s += V[3 * i + k] * V[3 * j + k] / sqrt(w[k]);
            Uinv[3 * i + j] = s;
        }
    mm3(F, Uinv, R);
}

/* One element's growth update for one step.
 *   rates : growth rate per second in x, y, z (already x tissue weight)
 *   F_mean: TriPointF at the start of the step (positions before NR)
 * Reproduces updateGrowthRotationMatrices -> cleanUpGrowthRates -> calculateGrowthUniform ->
 * growShapeByFg for a columnar element with no plastic deformation or shape change.
 * Returns 1 when the growth increment was rotated. */
int tfo_grow_element_inc(const double inc_in[9], const double F_mean[9], int allow_rotation, double Fg[9],
                         double InvFg[9]);
int tfo_grow_element(const double rates[3], double dt, const double F_mean[9], int allow_rotation, double Fg[9],
                     double InvFg[9])
{
    double inc[9] = {exp(rates[0] * dt), 0, 0, 0, exp(rates[1] * dt), 0, 0, 0, exp(rates[2] * dt)};
    return tfo_grow_element_inc(inc, F_mean, allow_rotation, Fg, InvFg);
}

/* The same update from a given growth increment (the product of the step's growth functions'
 * increments, ShapeBase::updateGrowthIncrement, ShapeBase.cpp:567-586). */
int tfo_grow_element_inc(const double inc_in[9], const double F_mean[9], int allow_rotation, double Fg[9],
                         double InvFg[9])
{
    double inc[9];
    memcpy(inc, inc_in, sizeof(inc));
    int rotated = 0;
    if (allow_rotation) {
        double R[9];
        tfo_polar_rotation(F_mean, R);
        int big = 0;
        const double threshold = 0.017;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                if (i != j && (R[3 * i + j] > threshold || R[3 * i + j] < -threshold)) big = 1;
        if (big) {
            double tz = atan2(R[1], R[0]);
            if (tz > 0.017 || tz < -0.017) {
                double c = cos(tz), s = sin(tz);
                /* rotMat = Rz(tz) then transposed (ShapeBase.cpp:1664-1667, 1697-1712) */
                double GRM[9] = {c, s, 0, -s, c, 0, 0, 0, 1};
                double T[9], T2[9];
                mtm3(GRM, inc, T);   /* R^T * growthIncrement   (ShapeBase.cpp:1219) */
                mm3(T, GRM, T2);     /* ... * R                 (ShapeBase.cpp:1221) */
                memcpy(inc, T2, sizeof(inc));
                rotated = 1;
            }
        }
    }
    double newFg[9];
    mm3(inc, Fg, newFg);
    memcpy(Fg, newFg, sizeof(newFg));
    tfo_inv3(Fg, InvFg);
    return rotated;
}

/* Mean deformation gradient of one element at given positions (calculateTriPointFForRatation,
 * ShapeBase.cpp:1676-1689). */
void tfo_mean_F(const double X[18], const double x[18], double Fm[9])
{
    double IJ0[NG][9], D0[NG];
    tfo_reference_geometry(X, IJ0, D0);
    memset(Fm, 0, 9 * sizeof(double));
    for (int g = 0; g < NG; ++g) {
        double dN[3][NN], J[9], F[9], eta, nu, zeta, w;
        tfo_gauss_point(g, &eta, &nu, &zeta, &w);
        tfo_shape_derivs(g, dN);
        for (int i = 0; i < 3; ++i)
            for (int k = 0; k < 3; ++k) {
                double s = 0.0;
                for (int a = 0; a < NN; ++a) s += dN[k][a] * x[3 * a + i];
                J[3 * i + k] = s;
            }
        mm3(J, IJ0[g], F);
        for (int q = 0; q < 9; ++q) Fm[q] += F[q] * w;
    }
}

/* Growth update for all elements from rates (inc == NULL) or from per-element increments
 * (inc[9e..9e+8], row-major); returns the number of rotated increments. */
int tfo_grow_all(int n_elem, const int32_t* conn, const double* Xref, const double* pos, const double* rates, double dt,
                 int allow_rotation, double* Fg, double* InvFg, const double* inc)
{
    int nrot = 0;
#pragma omp parallel for reduction(+ : nrot) schedule(static)
    for (int e = 0; e < n_elem; ++e) {
        double x[18], Fm[9];
        for (int a = 0; a < NN; ++a)
            for (int i = 0; i < 3; ++i) x[3 * a + i] = pos[3 * conn[6 * e + a] + i];
        tfo_mean_F(Xref + 18 * e, x, Fm);
        nrot += inc ? tfo_grow_element_inc(inc + 9 * e, Fm, allow_rotation, Fg + 9 * e, InvFg + 9 * e)
                    : tfo_grow_element(rates + 3 * e, dt, Fm, allow_rotation, Fg + 9 * e, InvFg + 9 * e);
    }
    return nrot;
}

/* Scalar elastic energy of one element, Pi = sum_g w_g D0_g det(Fg) psi(Fe) with the
 * neo-Hookean psi = mu/2 (tr C - 3) - mu ln Je + lambda/2 (ln Je)^2.  Its gradient with respect
 * to x is the internal elastic force of the reference (used to pin forces by finite differences). */
double tfo_element_energy(const double X[18], const double x[18], const double InvFg[9], double lambda, double mu)
{
    double IJ0[NG][9], D0[NG];
    tfo_reference_geometry(X, IJ0, D0);
    double detFg = 1.0 / tfo_det3(InvFg);
    double Pi = 0.0;
    for (int g = 0; g < NG; ++g) {
        double dN[3][NN], J[9], F[9], Fe[9], C[9], eta, nu, zeta, w;
        tfo_gauss_point(g, &eta, &nu, &zeta, &w);
        tfo_shape_derivs(g, dN);
        for (int i = 0; i < 3; ++i)
            for (int k = 0; k < 3; ++k) {
                double s = 0.0;
                for (int a = 0; a < NN; ++a) s += dN[k][a] * x[3 * a + i];
                J[3 * i + k] = s;
            }
        mm3(J, IJ0[g], F);
        mm3(F, InvFg, Fe);
        mtm3(Fe, Fe, C);
        double Je = tfo_det3(Fe), lnJ = log(Je);
        double psi = 0.5 * mu * (C[0] + C[4] + C[8] - 3.0) - mu * lnJ + 0.5 * lambda * lnJ * lnJ;
        Pi += w * D0[g] * detFg * psi;
    }
    return Pi;
}

/* ------------------------------------------------------------------------- */
/* Implicit packing: node masses, pair detection, forces and Jacobian        */
/* ------------------------------------------------------------------------- */
#define PK_MULTIPLIER 5000.0  /* packingMultiplier (Simulation.cpp:8187)          */
#define PK_SATURATION 5.0     /* sigmoidSaturationForPacking (Simulation.cpp:8188) */

/* Prism::calculateReferenceVolume (Prism.cpp:876-908) with
 * ShapeBase::calculateVolumeForInputShapeStructure (ShapeBase.cpp:3362-3397): eight triangles
 * (the two caps, each lateral quad split in two) each close a tetrahedron with the centroid;
 * volume = sum of |height| * area / 3. */
double tfo_reference_volume(const double X[18])
{
    static const int tri[8][3] = {{0, 1, 2}, {3, 4, 5}, {0, 1, 3}, {1, 3, 4}, {1, 2, 4}, {2, 4, 5}, {0, 2, 3}, {2, 3, 5}};
    double c[3] = {0.0, 0.0, 0.0};
    for (int a = 0; a < NN; ++a)
        for (int j = 0; j < 3; ++j) c[j] += X[3 * a + j];
    for (int j = 0; j < 3; ++j) c[j] /= NN;
    double V = 0.0;
    for (int t = 0; t < 8; ++t) {
        const double* p0 = X + 3 * tri[t][0];
        const double* p1 = X + 3 * tri[t][1];
        const double* p2 = X + 3 * tri[t][2];
        double v1[3], v2[3], vm[3], b[3];
        for (int j = 0; j < 3; ++j) {
            v1[j] = p1[j] - p0[j];
            v2[j] = p2[j] - p0[j];
            vm[j] = c[j] - p0[j];
        }
        b[0] = v1[1] * v2[2] - v1[2] * v2[1];
        b[1] = v1[2] * v2[0] - v1[0] * v2[2];
        b[2] = v1[0] * v2[1] - v1[1] * v2[0];
        double nb = sqrt(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
        double area = nb / 2;
        double height = (vm[0] * b[0] + vm[1] * b[1] + vm[2] * b[2]) / nb;
        if (height < 0) height *= -1.0;
        V += 1.0 / 3.0 * (height * area);
    }
    return V;
}

/* Simulation::updateNodeMasses (Simulation.cpp:7552-7560): every prism adds
 * VolumePerNode = det(Fg) V_ref / 6 (ShapeBase.cpp:1243-1245, :3938-3942) to its six nodes, in
 * element order.  Ablated nodes (node_ablated[n] != 0) get mass 0. */
void tfo_node_masses(int n_nodes, int n_elem, const int32_t* conn, const double* Xref, const double* Fg,
                     const uint8_t* node_ablated, double* mass)
{
    memset(mass, 0, (size_t)n_nodes * sizeof(double));
    for (int e = 0; e < n_elem; ++e) {
        double vpn = tfo_det3(Fg + 9 * e) * tfo_reference_volume(Xref + 18 * e) / NN;
        for (int a = 0; a < NN; ++a) mass[conn[6 * e + a]] += vpn;
    }
    if (node_ablated)
        for (int n = 0; n < n_nodes; ++n)
            if (node_ablated[n]) mass[n] = 0.0;
}

/* Prism::getApicalSideLengthAverage / getBasalSideLengthAverage (Prism.cpp:1443-1490): mean
 * length of the cap's three edges, zero-length edges left out (0 when all are). */
static double side_length_average(const double* pos, const int32_t* c6, int apical)
{
    static const int pr[2][3][2] = {{{0, 1}, {0, 2}, {1, 2}}, {{3, 4}, {3, 5}, {4, 5}}};
    double dsum = 0.0;
    int counter = 0;
    for (int i = 0; i < 3; ++i) {
        const double* p = pos + 3 * c6[pr[apical][i][0]];
        const double* q = pos + 3 * c6[pr[apical][i][1]];
        double dx = p[0] - q[0], dy = p[1] - q[1], dz = p[2] - q[2];
        double d2 = dx * dx + dy * dy + dz * dz;
        if (d2 > 0) {
            dsum += pow(d2, 0.5);
            counter++;
        }
    }
    if (counter > 0) dsum /= counter;
    return dsum;
}

static int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

/* Packing thresholds of one tissue, restating Simulation::calculateBoundingBox
 * (Simulation.cpp:9148-9189), ShapeBase::calculateRelativePosInBoundingBox (ShapeBase.cpp:1842-1848)
 * and Simulation::getAverageSideLength (Simulation.cpp:3205-3265) with the scaling of
 * detectPacingNodes (:8175-8186), on the current positions:
 *   - bounding box: "if (x < min) min = x; else if (x > max) max = x" from (1e5, -1e5), so a
 *     node that lowers the running minimum is never compared with the maximum;
 *   - every (non-ablated, columnar) element adds its apical side average, and its basal one when
 *     the element is basal or spans the tissue, to the columnar mean (zero averages left out);
 *   - grid bin (floor(rx), floor(ry / 2)) of the element centre's position relative to the box,
 *     clamped to 10 x 5, sums the last of those values; the bin counters live in a 10 x 4 array
 *     indexed [i][j] with j < 5, so bins (i, 4) and (i + 1, 0) share one counter (flat index
 *     4 i + j; the reference's [9][4] lies past its array, here it is a counter of its own);
 *   - T = 0.4 * columnar mean (packingThreshold), grid thresholds 0.4 * bin mean.
 * elem_basal_side[e] = 1 when the element is basal or spans the tissue.  Outputs: T, grid[50]
 * (row-major [10][5]), bb[4] = (min x, min y, size x, size y). */
void tfo_packing_thresholds(int n_nodes, int n_elem, const int32_t* conn, const uint8_t* elem_basal_side,
                            const double* pos, double* T, double grid[50], double bb[4])
{
    double bmin[3] = {100000.0, 100000.0, 100000.0}, bmax[3] = {-100000.0, -100000.0, -100000.0};
    for (int n = 0; n < n_nodes; ++n)
        for (int i = 0; i < 3; ++i) {
            double p = pos[3 * n + i];
            if (p < bmin[i]) bmin[i] = p;
            else if (p > bmax[i]) bmax[i] = p;
        }
    double size[2] = {bmax[0] - bmin[0], bmax[1] - bmin[1]};
    double sum[50];
    double counter[41];
    memset(sum, 0, sizeof(sum));
    memset(counter, 0, sizeof(counter));
    double dsumCol = 0.0;
    int colCounter = 0;
    for (int e = 0; e < n_elem; ++e) {
        const int32_t* c6 = conn + 6 * e;
        double cur = side_length_average(pos, c6, 1);
        if (cur > 0) { dsumCol += cur; colCounter++; }
        if (elem_basal_side[e]) {
            cur = side_length_average(pos, c6, 0);
            if (cur > 0) { dsumCol += cur; colCounter++; }
        }
        double cen[3] = {0.0, 0.0, 0.0};
        for (int a = 0; a < NN; ++a)
            for (int j = 0; j < 3; ++j) cen[j] += pos[3 * c6[a] + j];
        for (int j = 0; j < 3; ++j) cen[j] /= NN;
        double rx = (cen[0] - bmin[0]) / size[0];
        double ry = (cen[1] - bmin[1]) / size[1];
        int relX = clampi((int)floor(rx), 0, 9);
        int relY = clampi((int)floor(ry / 2.0), 0, 4);
        sum[5 * relX + relY] += cur;
        counter[4 * relX + relY] += 1.0;
    }
    *T = colCounter > 0 ? 0.4 * (dsumCol / (double)colCounter) : 0.0;
    for (int i = 0; i < 10; ++i)
        for (int j = 0; j < 5; ++j) {
            double v = sum[5 * i + j];
            if (counter[4 * i + j] > 0) v /= counter[4 * i + j];
            grid[5 * i + j] = 0.4 * v;
        }
    bb[0] = bmin[0];
    bb[1] = bmin[1];
    bb[2] = size[0];
    bb[3] = size[1];
}

/* Simulation::detectPacingNodes (Simulation.cpp:8167-8440) for one tissue, literally O(N^2):
 *   - a node can pack when its mass is non-zero and it is apical or basal (Node::checkIfNodeHasPacking,
 *     Node.cpp:133-151; there are no lateral elements here);
 *   - master i (ascending) and slave j > i (ascending) pack when they have the same placement, do
 *     not share an element (Node::immediateNeigs, ShapeBase.cpp:3105-3122; adj_ptr/adj_col hold
 *     each node's sorted element neighbours) and |x_i - x_j|^2 < t^2, t the grid threshold of the
 *     master's bin (floor((x - min x) / size x), floor((y - min y) / size y)), clamped to the grid;
 *   - the node range is cut into nSeg = N / max(1, N / 16) chunks (integer division) and the pairs
 *     found from masters in the last chunk are dropped (the merge loop stops at nSeg - 1);
 *   - weights (dx, dy, dz) / d with (dx, dy, dz) = x_i - x_j (initialWeightPointx/y/z).
 * Returns the number of pairs (written up to max_pairs). */
int tfo_packing_detect(int n_nodes, const double* pos, const int8_t* placement, const double* mass,
                       const int32_t* adj_ptr, const int32_t* adj_col, const double grid[50], const double bb[4],
                       int max_pairs, int32_t* pairs, double* weights)
{
    int seg_size = n_nodes / 16;
    if (seg_size < 1) seg_size = 1;
    const int n_seg = n_nodes / seg_size;
    const int master_end = (n_seg - 1) * seg_size;
    int np = 0;
    for (int i = 0; i < master_end; ++i) {
        if (mass[i] == 0 || (placement[i] != 0 && placement[i] != 1)) continue;
        const double* p = pos + 3 * i;
        int gx = clampi((int)floor((p[0] - bb[0]) / bb[2]), 0, 9);
        int gy = clampi((int)floor((p[1] - bb[1]) / bb[3]), 0, 4);
        double t2 = grid[5 * gx + gy] * grid[5 * gx + gy];
        for (int j = i + 1; j < n_nodes; ++j) {
            if (mass[j] == 0 || placement[j] != placement[i]) continue;
            int neig = 0;
            for (int k = adj_ptr[i]; k < adj_ptr[i + 1]; ++k)
                if (adj_col[k] == j) { neig = 1; break; }
            if (neig) continue;
            const double* q = pos + 3 * j;
            double dx = p[0] - q[0], dy = p[1] - q[1], dz = p[2] - q[2];
            double d2 = dx * dx + dy * dy + dz * dz;
            if (d2 < t2) {
                if (np < max_pairs) {
                    double d = pow(d2, 0.5);
                    pairs[2 * np] = i;
                    pairs[2 * np + 1] = j;
                    weights[3 * np] = dx / d;
                    weights[3 * np + 1] = dy / d;
                    weights[3 * np + 2] = dz / d;
                }
                np++;
            }
        }
    }
    return np;
}

/* Simulation::calculatePackingForcesImplicit3D (Simulation.cpp:7149-7203) and
 * calculatePackingJacobian3D (:7205-7295) at positions pos: per pair and component k,
 *   d = x0_k - x1_k, negated when w_k > 0;  s = 1 / (1 + exp(-(S / T) d));
 *   F_k = M * mbar * s * w_k  (mbar the pair's mean mass) added to node 0, subtracted from node 1;
 *   D_k = dF_k / dx0_k = s (1 - s) M mbar w_k (S / T), negated when w_k > 0.
 * F (3N) is accumulated in pair order; D (3 per pair) feeds K[a][a] -= D, K[a][b] += D,
 * K[b][b] -= D, K[b][a] += D on the component's diagonal. */
void tfo_packing_forces(int n_pairs, const int32_t* pairs, const double* weights, const double* mass,
                        const double* pos, double T, double* F, double* D)
{
    const double c = PK_SATURATION / T;
    for (int p = 0; p < n_pairs; ++p) {
        const int id0 = pairs[2 * p], id1 = pairs[2 * p + 1];
        const double mbar = 0.5 * (mass[id0] + mass[id1]);
        for (int k = 0; k < 3; ++k) {
            const double w = weights[3 * p + k];
            double d = pos[3 * id0 + k] - pos[3 * id1 + k];
            if (w > 0) d *= -1.0;
            double f = PK_MULTIPLIER * mbar / (1 + exp(c * (-1.0 * d)));
            f *= w;
            F[3 * id0 + k] += f;
            F[3 * id1 + k] -= f;
            double s = 1 / (1 + exp(c * (-1.0 * d)));
            double dd = s * (1 - s) * PK_MULTIPLIER * mbar * w * c;
            if (w > 0) dd *= -1.0;
            D[3 * p + k] = dd;
        }
    }
}
