/*
 * onsas_oracle.c -- FP64 CPU restatement of the ONSAS.jl nonlinear static path.
 *
 * TEST INFRASTRUCTURE ONLY (see onsas_oracle.h).  Plain C99, no dependencies.
 * File:line citations are relative to the ONSAS.jl v0.4.6 source tree.
 */
#include "onsas_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* small dense helpers (row-major 3x3)                                       */
/* ------------------------------------------------------------------------ */
static double det3(const double A[9]) {
    return A[0] * (A[4] * A[8] - A[5] * A[7]) - A[1] * (A[3] * A[8] - A[5] * A[6]) +
           A[2] * (A[3] * A[7] - A[4] * A[6]);
}

static void inv3(const double A[9], double Ai[9]) {
    double d = det3(A);
    Ai[0] = (A[4] * A[8] - A[5] * A[7]) / d;
    Ai[1] = (A[2] * A[7] - A[1] * A[8]) / d;
    Ai[2] = (A[1] * A[5] - A[2] * A[4]) / d;
    Ai[3] = (A[5] * A[6] - A[3] * A[8]) / d;
    Ai[4] = (A[0] * A[8] - A[2] * A[6]) / d;
    Ai[5] = (A[2] * A[3] - A[0] * A[5]) / d;
    Ai[6] = (A[3] * A[7] - A[4] * A[6]) / d;
    Ai[7] = (A[1] * A[6] - A[0] * A[7]) / d;
    Ai[8] = (A[0] * A[4] - A[1] * A[3]) / d;
}

/* Voigt index pairs 11 22 33 23 13 12 (src/Utils.jl:42) */
static const int VI[6] = {0, 1, 2, 1, 0, 0};
static const int VJ[6] = {0, 1, 2, 2, 2, 1};

/* voigt(T, alpha) (src/Utils.jl:46-48) */
static void voigt(const double T[9], double alpha, double v[6]) {
    v[0] = T[0];
    v[1] = T[4];
    v[2] = T[8];
    v[3] = alpha * T[5];
    v[4] = alpha * T[2];
    v[5] = alpha * T[1];
}

/* ------------------------------------------------------------------------ */
/* materials                                                                 */
/* ------------------------------------------------------------------------ */
void or_lame_from_E_nu(double Emod, double nu, double* lam, double* G) {
    *lam = Emod * nu / ((1 + nu) * (1 - 2 * nu));
    *G = Emod / (2 * (1 + nu));
}

/* cosserat_stress!(S, dSdE, ::SVK, E) (src/Materials/SVKMaterial.jl:89-100).
 * The reference writes only the [1:3,1:3] and [4:6,4:6] blocks of dSdE into a
 * zero-initialised cache; the off-diagonal blocks are zero here. */
void or_svk_stress(double lam, double G, const double E[9], double S[9], double dSdE[36]) {
    double tr = E[0] + E[4] + E[8];
    for (int i = 0; i < 9; ++i) S[i] = 2 * G * E[i];
    S[0] += lam * tr;
    S[4] += lam * tr;
    S[8] += lam * tr;
    memset(dSdE, 0, 36 * sizeof(double));
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) dSdE[a * 6 + b] = lam + (a == b ? 2 * G : 0.0);
    for (int a = 3; a < 6; ++a) dSdE[a * 6 + a] = G;
}

/* strain_energy(::SVK, E) (src/Materials/SVKMaterial.jl:57-60) */
double or_svk_energy(double lam, double G, const double E[9]) {
    double tr = E[0] + E[4] + E[8];
    double trE2 = 0;
    for (int i = 0; i < 3; ++i)
        for (int k = 0; k < 3; ++k) trE2 += E[i * 3 + k] * E[k * 3 + i];
    return lam / 2 * tr * tr + G * trE2;
}

/* --- forward-mode dual numbers: a restatement of what ForwardDiff.gradient
 * (ForwardDiff 0.10.38) evaluates at src/Materials/NeoHookeanMaterial.jl:122.
 * The seed directions are the six upper-triangle entries of E, in Voigt order;
 * lower-triangle entries of the collected matrix carry a zero gradient because
 * `Symmetric(2E + I)` reads only the upper triangle (NeoHookeanMaterial.jl:99). */
typedef struct {
    double v;
    double d[6];
} dual6;

static dual6 dconst(double v) {
    dual6 r;
    r.v = v;
    memset(r.d, 0, sizeof r.d);
    return r;
}
static dual6 dadd(dual6 a, dual6 b) {
    dual6 r;
    r.v = a.v + b.v;
    for (int k = 0; k < 6; ++k) r.d[k] = a.d[k] + b.d[k];
    return r;
}
static dual6 dsub(dual6 a, dual6 b) {
    dual6 r;
    r.v = a.v - b.v;
    for (int k = 0; k < 6; ++k) r.d[k] = a.d[k] - b.d[k];
    return r;
}
static dual6 dmul(dual6 a, dual6 b) {
    dual6 r;
    r.v = a.v * b.v;
    for (int k = 0; k < 6; ++k) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
    return r;
}
static dual6 ddiv(dual6 a, dual6 b) {
    dual6 r;
    r.v = a.v / b.v;
    for (int k = 0; k < 6; ++k) r.d[k] = (a.d[k] * b.v - a.v * b.d[k]) / (b.v * b.v);
    return r;
}
static dual6 dscale(double s, dual6 a) {
    dual6 r;
    r.v = s * a.v;
    for (int k = 0; k < 6; ++k) r.d[k] = s * a.d[k];
    return r;
}
static dual6 dsqrt(dual6 a) {
    dual6 r;
    r.v = sqrt(a.v);
    for (int k = 0; k < 6; ++k) r.d[k] = a.d[k] / (2 * r.v);
    return r;
}

/* _S_analytic(m, E) (src/Materials/NeoHookeanMaterial.jl:96-104) on duals.
 * Cm: the 3x3 Symmetric(2E + I) as duals; S: full 3x3 result. */
static void nh_S_dual(double K, double G, const dual6 Cm[9], dual6 S[9]) {
    /* inv(C) and det(C) of a 3x3 matrix through cofactors */
    dual6 c00 = dsub(dmul(Cm[4], Cm[8]), dmul(Cm[5], Cm[7]));
    dual6 c01 = dsub(dmul(Cm[2], Cm[7]), dmul(Cm[1], Cm[8]));
    dual6 c02 = dsub(dmul(Cm[1], Cm[5]), dmul(Cm[2], Cm[4]));
    dual6 c10 = dsub(dmul(Cm[5], Cm[6]), dmul(Cm[3], Cm[8]));
    dual6 c11 = dsub(dmul(Cm[0], Cm[8]), dmul(Cm[2], Cm[6]));
    dual6 c12 = dsub(dmul(Cm[2], Cm[3]), dmul(Cm[0], Cm[5]));
    dual6 c20 = dsub(dmul(Cm[3], Cm[7]), dmul(Cm[4], Cm[6]));
    dual6 c21 = dsub(dmul(Cm[1], Cm[6]), dmul(Cm[0], Cm[7]));
    dual6 c22 = dsub(dmul(Cm[0], Cm[4]), dmul(Cm[1], Cm[3]));
    dual6 det = dadd(dadd(dmul(Cm[0], c00), dmul(Cm[1], c10)), dmul(Cm[2], c20));
    dual6 Ci[9] = {ddiv(c00, det), ddiv(c01, det), ddiv(c02, det),
                   ddiv(c10, det), ddiv(c11, det), ddiv(c12, det),
                   ddiv(c20, det), ddiv(c21, det), ddiv(c22, det)};
    dual6 J = dsqrt(det);
    dual6 kjj = dscale(K, dmul(J, dsub(J, dconst(1.0))));
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            dual6 eye_minus = dsub(dconst(i == j ? 1.0 : 0.0), Ci[i * 3 + j]);
            S[i * 3 + j] = dadd(dscale(G, eye_minus), dmul(kjj, Ci[i * 3 + j]));
        }
}

/* cosserat_stress!(S, dSdE, ::NeoHookean, E)
 * (src/Materials/NeoHookeanMaterial.jl:134-138, :115-129). */
void or_nh_stress(double K, double G, const double E[9], double S[9], double dSdE[36]) {
    dual6 Cm[9];
    /* Symmetric(2E + I): upper triangle mirrored, seeded on the upper entries */
    for (int p = 0; p < 6; ++p) {
        int i = VI[p], j = VJ[p];
        dual6 c = dconst(2 * E[i * 3 + j] + (i == j ? 1.0 : 0.0));
        c.d[p] = 2.0;
        Cm[i * 3 + j] = c;
        Cm[j * 3 + i] = c;
    }
    dual6 Sd[9];
    nh_S_dual(K, G, Cm, Sd);
    /* S .= Symmetric(S_analytic) -> upper triangle mirrored */
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            int a = i < j ? i : j, b = i < j ? j : i;
            S[i * 3 + j] = Sd[a * 3 + b].v;
        }
    /* row p of dSdE = voigt(gradient of S[i,j], 0.5) */
    for (int p = 0; p < 6; ++p) {
        const dual6* s = &Sd[VI[p] * 3 + VJ[p]];
        for (int q = 0; q < 6; ++q) dSdE[p * 6 + q] = (q < 3 ? 1.0 : 0.5) * s->d[q];
    }
}

/* strain_energy(::NeoHookean, E) (src/Materials/NeoHookeanMaterial.jl:59-66) */
double or_nh_energy(double K, double G, const double E[9]) {
    double C[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            int a = i < j ? i : j, b = i < j ? j : i;
            C[i * 3 + j] = 2 * E[a * 3 + b] + (i == j ? 1.0 : 0.0);
        }
    double J = sqrt(det3(C));
    double I1 = C[0] + C[4] + C[8];
    return G / 2 * (I1 - 2 * log(J)) + K / 2 * (J - 1) * (J - 1);
}

/* stress!(sigma, dsde, ::IsotropicLinearElastic, eps)
 * (src/Materials/IsotropicLinearElasticMaterial.jl:80-92) */
void or_linear_stress(double Emod, double nu, const double eps[9], double sig[9],
                      double dsde[36]) {
    double lam, G;
    or_lame_from_E_nu(Emod, nu, &lam, &G);
    or_svk_stress(lam, G, eps, sig, dsde);
}

/* ------------------------------------------------------------------------ */
/* tetrahedron                                                               */
/* ------------------------------------------------------------------------ */
/* dXdz_1 (src/Entities/Tetrahedrons.jl:269-271) */
static const double DXDZ[3][4] = {{1.0, -1.0, 0.0, 0.0}, {0.0, -1.0, 0.0, 1.0},
                                  {0.0, -1.0, 1.0, 0.0}};

/* J = X dXdz^T (Tetrahedrons.jl:129-131); X is 3x4 with node columns */
static void tet_jacobian(const double X[12], double Jm[9]) {
    for (int i = 0; i < 3; ++i)
        for (int k = 0; k < 3; ++k) {
            double s = 0;
            for (int n = 0; n < 4; ++n) s += X[n * 3 + i] * DXDZ[k][n];
            Jm[i * 3 + k] = s;
        }
}

int or_tet_volume(const double X[12], double* V) {
    double Jm[9];
    tet_jacobian(X, Jm);
    *V = det3(Jm) / 6.0; /* _volume (Tetrahedrons.jl:134-138) */
    return *V > 0 ? OR_OK : OR_ERR_NEGATIVE_VOLUME;
}

/* funder = inv(J)' * dXdz, stored node-major: dN[n*3 + j] = funder[j, n] */
int or_tet_gradients(const double X[12], double dN[12], double* V) {
    double Jm[9], Ji[9];
    tet_jacobian(X, Jm);
    *V = det3(Jm) / 6.0;
    if (!(*V > 0)) return OR_ERR_NEGATIVE_VOLUME;
    inv3(Jm, Ji);
    for (int j = 0; j < 3; ++j)
        for (int n = 0; n < 4; ++n) {
            double s = 0;
            for (int k = 0; k < 3; ++k) s += Ji[k * 3 + j] * DXDZ[k][n]; /* inv(J)'[j,k] */
            dN[n * 3 + j] = s;
        }
    return OR_OK;
}

/* _B_mat!(B, deriv, F) (Tetrahedrons.jl:140-152), B is 6x12 row-major */
static void tet_B(const double dN[12], const double F[9], double B[72]) {
    for (int k = 0; k < 4; ++k) {
        const double* d = &dN[k * 3];
        for (int j = 0; j < 3; ++j) {
            int c = k * 3 + j;
            /* rows 1:3: diagm(deriv[:,k]) * F'  ->  d_i F[j,i] */
            B[0 * 12 + c] = d[0] * F[j * 3 + 0];
            B[1 * 12 + c] = d[1] * F[j * 3 + 1];
            B[2 * 12 + c] = d[2] * F[j * 3 + 2];
            /* rows 4:6: engineering shears 23, 13, 12 */
            B[3 * 12 + c] = d[1] * F[j * 3 + 2] + d[2] * F[j * 3 + 1];
            B[4 * 12 + c] = d[0] * F[j * 3 + 2] + d[2] * F[j * 3 + 0];
            B[5 * 12 + c] = d[0] * F[j * 3 + 1] + d[1] * F[j * 3 + 0];
        }
    }
}

/* K = B' D B vol (12x12), only the upper triangle is kept then mirrored
 * (`Symmetric(B' * dSdE * B * vol)`, Tetrahedrons.jl:210, 261) */
static void tet_BtDB(const double B[72], const double D[36], double vol, double K[144]) {
    double DB[72];
    for (int p = 0; p < 6; ++p)
        for (int c = 0; c < 12; ++c) {
            double s = 0;
            for (int q = 0; q < 6; ++q) s += D[p * 6 + q] * B[q * 12 + c];
            DB[p * 12 + c] = s;
        }
    for (int r = 0; r < 12; ++r)
        for (int c = 0; c < 12; ++c) {
            double s = 0;
            for (int p = 0; p < 6; ++p) s += B[p * 12 + r] * DB[p * 12 + c];
            K[r * 12 + c] = s * vol;
        }
}

static void sym_upper(double* K, int n) {
    for (int r = 0; r < n; ++r)
        for (int c = 0; c < r; ++c) K[r * n + c] = K[c * n + r];
}

int or_tet_hyperelastic(int mat, double p0, double p1, const double X[12], const double u[12],
                        double f[12], double K[144], double P[9], double C[9]) {
    double dN[12], vol;
    int st = or_tet_gradients(X, dN, &vol);
    if (st != OR_OK) return st;
    /* H = U * funder' with U = reshape(u_e, 3, 4) (Tetrahedrons.jl:192,198) */
    double H[9], F[9], E[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double s = 0;
            for (int n = 0; n < 4; ++n) s += u[n * 3 + i] * dN[n * 3 + j];
            H[i * 3 + j] = s;
        }
    for (int i = 0; i < 9; ++i) F[i] = H[i] + ((i % 4) == 0 ? 1.0 : 0.0);
    /* E = Symmetric(0.5 (H + H' + H'H)) (Tetrahedrons.jl:201) */
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double hth = 0;
            for (int k = 0; k < 3; ++k) hth += H[k * 3 + i] * H[k * 3 + j];
            E[i * 3 + j] = 0.5 * (H[i * 3 + j] + H[j * 3 + i] + hth);
        }
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < i; ++j) E[i * 3 + j] = E[j * 3 + i];
    double B[72], S[9], D[36];
    tet_B(dN, F, B);
    if (mat == OR_MAT_SVK)
        or_svk_stress(p0, p1, E, S, D);
    else if (mat == OR_MAT_NEOHOOKEAN)
        or_nh_stress(p0, p1, E, S, D);
    else
        return OR_ERR_ARG;
    /* fint = B' voigt(S) vol (Tetrahedrons.jl:206-207) */
    double Sv[6];
    voigt(S, 1.0, Sv);
    for (int c = 0; c < 12; ++c) {
        double s = 0;
        for (int p = 0; p < 6; ++p) s += B[p * 12 + c] * Sv[p];
        f[c] = s * vol;
    }
    /* material + geometric stiffness (Tetrahedrons.jl:210-215, :161-182) */
    tet_BtDB(B, D, vol, K);
    double aux[16];
    for (int a = 0; a < 4; ++a)
        for (int b = 0; b < 4; ++b) {
            double s = 0;
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) s += dN[a * 3 + i] * S[i * 3 + j] * dN[b * 3 + j];
            aux[a * 4 + b] = s * vol;
        }
    for (int a = 0; a < 4; ++a)
        for (int b = 0; b < 4; ++b)
            for (int d = 0; d < 3; ++d) K[(a * 3 + d) * 12 + (b * 3 + d)] += aux[a * 4 + b];
    sym_upper(K, 12);
    /* P = F S, C = F' F (Tetrahedrons.jl:217-221) */
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double sp = 0, sc = 0;
            for (int k = 0; k < 3; ++k) {
                sp += F[i * 3 + k] * S[k * 3 + j];
                sc += F[k * 3 + i] * F[k * 3 + j];
            }
            P[i * 3 + j] = sp;
            C[i * 3 + j] = sc;
        }
    return OR_OK;
}

int or_tet_linear(double Emod, double nu, const double X[12], const double u[12],
                  double f[12], double K[144], double sig[9], double eps[9]) {
    double dN[12], vol;
    int st = or_tet_gradients(X, dN, &vol);
    if (st != OR_OK) return st;
    double H[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double s = 0;
            for (int n = 0; n < 4; ++n) s += u[n * 3 + i] * dN[n * 3 + j];
            H[i * 3 + j] = s;
        }
    /* eps = Symmetric(0.5 (H + H')), F = I (Tetrahedrons.jl:252-253) */
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) eps[i * 3 + j] = 0.5 * (H[i * 3 + j] + H[j * 3 + i]);
    const double I3[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    double B[72], D[36];
    tet_B(dN, I3, B);
    or_linear_stress(Emod, nu, eps, sig, D);
    tet_BtDB(B, D, vol, K);
    sym_upper(K, 12);
    /* fint = Ks * u_e (Tetrahedrons.jl:263) */
    for (int r = 0; r < 12; ++r) {
        double s = 0;
        for (int c = 0; c < 12; ++c) s += K[r * 12 + c] * u[c];
        f[r] = s;
    }
    return OR_OK;
}

/* ------------------------------------------------------------------------ */
/* truss (src/Entities/Trusses.jl)                                           */
/* ------------------------------------------------------------------------ */
int or_truss(int strain_model, double Emod, double A, const double X[6], const double u[6],
             double f[6], double K[36], double* sigma11, double* eps11) {
    double Xd[6];
    for (int i = 0; i < 6; ++i) Xd[i] = X[i] + u[i];
    /* _lengths (Trusses.jl:232-237): Bdif * X = X2 - X1 */
    double dref[3], ddef[3];
    for (int i = 0; i < 3; ++i) {
        dref[i] = X[3 + i] - X[i];
        ddef[i] = Xd[3 + i] - Xd[i];
    }
    double l_ref = sqrt(dref[0] * dref[0] + dref[1] * dref[1] + dref[2] * dref[2]);
    double l_def = sqrt(ddef[0] * ddef[0] + ddef[1] * ddef[1] + ddef[2] * ddef[2]);
    if (!(l_ref > 0)) return OR_ERR_ARG;
    if (strain_model == OR_STRAIN_ROTATED) {
        /* Trusses.jl:126-155 */
        double t[6];
        for (int i = 0; i < 3; ++i) {
            double e1 = ddef[i] / l_def;
            t[i] = -e1;
            t[3 + i] = e1;
        }
        double eps = (l_def * l_def - l_ref * l_ref) / (l_ref * (l_ref + l_def)); /* :187-189 */
        double S11 = Emod * eps;
        for (int i = 0; i < 6; ++i) f[i] = A * S11 * t[i];
        for (int r = 0; r < 6; ++r)
            for (int c = 0; c < 6; ++c) {
                /* Bdif' * Bdif = [I -I; -I I] */
                double G = (r % 3 == c % 3) ? ((r < 3) == (c < 3) ? 1.0 : -1.0) : 0.0;
                double km = Emod * A / l_ref * (t[r] * t[c]);
                double kg = S11 * A / l_def * (G - t[r] * t[c]);
                K[r * 6 + c] = km + kg;
            }
        *sigma11 = S11 * l_def / l_ref;
        *eps11 = eps;
    } else if (strain_model == OR_STRAIN_GREEN) {
        /* Trusses.jl:159-184 with _aux_b (:221-229) */
        double bsum[6];
        for (int i = 0; i < 3; ++i) {
            /* X_row' * Ge = [X1 - X2, X2 - X1] */
            double bref = (X[3 + i] - X[i]) / (l_ref * l_ref);
            double bdef = (u[3 + i] - u[i]) / (l_ref * l_ref);
            bsum[i] = -(bref + bdef);
            bsum[3 + i] = bref + bdef;
        }
        double eps = (l_def * l_def - l_ref * l_ref) / (2 * l_ref * l_ref); /* :192-194 */
        double S11 = Emod * eps;
        for (int i = 0; i < 6; ++i) f[i] = A * S11 * l_ref * bsum[i];
        for (int r = 0; r < 6; ++r)
            for (int c = 0; c < 6; ++c) {
                double G = (r % 3 == c % 3) ? ((r < 3) == (c < 3) ? 1.0 : -1.0) : 0.0;
                K[r * 6 + c] = S11 * A / l_ref * G + Emod * A * l_ref * bsum[r] * bsum[c];
            }
        *sigma11 = S11 * l_def / l_ref;
        *eps11 = eps;
    } else {
        return OR_ERR_ARG;
    }
    return OR_OK;
}

/* ------------------------------------------------------------------------ */
/* pressure on a triangular face                                             */
/* ------------------------------------------------------------------------ */
void or_face_pressure(const double X[9], double p, double f[9]) {
    double a[3], b[3], Av[3];
    for (int i = 0; i < 3; ++i) {
        a[i] = X[3 + i] - X[i];
        b[i] = X[6 + i] - X[i];
    }
    /* _area_vec = 1/2 cross(x2 - x1, x3 - x1) (TriangularFaces.jl:44-47) */
    Av[0] = 0.5 * (a[1] * b[2] - a[2] * b[1]);
    Av[1] = 0.5 * (a[2] * b[0] - a[0] * b[2]);
    Av[2] = 0.5 * (a[0] * b[1] - a[1] * b[0]);
    double A = sqrt(Av[0] * Av[0] + Av[1] * Av[1] + Av[2] * Av[2]);
    /* p = bc(t) * (-n) * A ; p_nodal = p / 3 (LocalLoadBoundaryConditions.jl:39-45) */
    for (int k = 0; k < 3; ++k)
        for (int i = 0; i < 3; ++i) f[k * 3 + i] = p * (-(Av[i] / A)) * A / 3;
}

/* ------------------------------------------------------------------------ */
/* assembly                                                                  */
/* ------------------------------------------------------------------------ */
void or_csr_free(or_csr* A) {
    free(A->rowptr);
    free(A->col);
    free(A->val);
    A->rowptr = NULL;
    A->col = NULL;
    A->val = NULL;
    A->n = A->nnz = 0;
}

void or_csr_matvec(const or_csr* A, const double* x, double* y) {
    /* rows are independent: the OpenMP build (CPU baseline only) splits them */
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < A->n; ++i) {
        double s = 0;
        for (int64_t k = A->rowptr[i]; k < A->rowptr[i + 1]; ++k) s += A->val[k] * x[A->col[k]];
        y[i] = s;
    }
}

/* Build a CSR matrix from COO triplets, summing duplicates in triplet order:
 * the result of `K[I[k], J[k]] += V[k]` for k = 1..end (Assemblers.jl:84-88).
 * map renumbers (or drops, when < 0) the rows and columns.  Triplets are
 * bucketed by row (stable), then each row's distinct columns get a slot the
 * first time they appear (a marker per column), every value is added to its
 * slot in triplet order starting from a structural 0.0, and the row's slots are
 * sorted by column (insertion sort of the distinct columns, a few dozen). */
static int coo_to_csr(int64_t nt, const int32_t* I, const int32_t* J, const double* V,
                      int64_t n, const int64_t* map, or_csr* A) {
    int64_t* cnt = (int64_t*)calloc((size_t)n + 1, sizeof(int64_t));
    if (!cnt) return OR_ERR_ALLOC;
    for (int64_t k = 0; k < nt; ++k) {
        int64_t r = map ? map[I[k]] : I[k];
        int64_t c = map ? map[J[k]] : J[k];
        if (r >= 0 && c >= 0) cnt[r + 1]++;
    }
    for (int64_t i = 0; i < n; ++i) cnt[i + 1] += cnt[i];
    int64_t m = cnt[n];
    int32_t* tc = (int32_t*)malloc((size_t)(m ? m : 1) * sizeof(int32_t));
    double* tv = (double*)malloc((size_t)(m ? m : 1) * sizeof(double));
    int64_t* pos = (int64_t*)malloc((size_t)(n ? n : 1) * sizeof(int64_t));
    int64_t* mark = (int64_t*)malloc((size_t)(n ? n : 1) * sizeof(int64_t));
    if (!tc || !tv || !pos || !mark) return OR_ERR_ALLOC;
    memcpy(pos, cnt, (size_t)n * sizeof(int64_t));
    for (int64_t k = 0; k < nt; ++k) { /* stable bucket by row */
        int64_t r = map ? map[I[k]] : I[k];
        int64_t c = map ? map[J[k]] : J[k];
        if (r < 0 || c < 0) continue;
        tc[pos[r]] = (int32_t)c;
        tv[pos[r]] = V[k];
        pos[r]++;
    }
    for (int64_t i = 0; i < n; ++i) mark[i] = -1;
    A->n = n;
    A->rowptr = (int64_t*)malloc(((size_t)n + 1) * sizeof(int64_t));
    A->col = (int32_t*)malloc((size_t)(m ? m : 1) * sizeof(int32_t));
    A->val = (double*)malloc((size_t)(m ? m : 1) * sizeof(double));
    /* (column, slot) pairs of one row, for the sort */
    int64_t cap = 64;
    int32_t* pair = (int32_t*)malloc((size_t)cap * 2 * sizeof(int32_t));
    double* rowv = (double*)malloc((size_t)cap * sizeof(double));
    if (!A->rowptr || !A->col || !A->val || !pair || !rowv) return OR_ERR_ALLOC;
    int64_t out = 0;
    A->rowptr[0] = 0;
    for (int64_t r = 0; r < n; ++r) {
        int64_t b = cnt[r], e = cnt[r + 1], nd = 0;
        for (int64_t k = b; k < e; ++k) {
            int32_t c = tc[k];
            if (mark[c] < 0) { /* first appearance in this row: a new slot */
                if (nd == cap) {
                    cap *= 2;
                    pair = (int32_t*)realloc(pair, (size_t)cap * 2 * sizeof(int32_t));
                    rowv = (double*)realloc(rowv, (size_t)cap * sizeof(double));
                    if (!pair || !rowv) return OR_ERR_ALLOC;
                }
                mark[c] = nd;
                pair[2 * nd] = c;
                pair[2 * nd + 1] = (int32_t)nd;
                rowv[nd] = 0.0; /* structural zero then += in triplet order */
                ++nd;
            }
            rowv[mark[c]] += tv[k];
        }
        for (int64_t q = 1; q < nd; ++q) { /* distinct columns, ascending */
            int32_t c = pair[2 * q], sl = pair[2 * q + 1];
            int64_t w = q - 1;
            while (w >= 0 && pair[2 * w] > c) {
                pair[2 * w + 2] = pair[2 * w];
                pair[2 * w + 3] = pair[2 * w + 1];
                --w;
            }
            pair[2 * w + 2] = c;
            pair[2 * w + 3] = sl;
        }
        for (int64_t q = 0; q < nd; ++q) {
            A->col[out] = pair[2 * q];
            A->val[out] = rowv[pair[2 * q + 1]];
            mark[pair[2 * q]] = -1;
            ++out;
        }
        A->rowptr[r + 1] = out;
    }
    A->nnz = out;
    free(cnt);
    free(tc);
    free(tv);
    free(pos);
    free(mark);
    free(pair);
    free(rowv);
    return OR_OK;
}

int or_coo_to_csr(int64_t nt, const int32_t* I, const int32_t* J, const double* V, int64_t n,
                  or_csr* A) {
    return coo_to_csr(nt, I, J, V, n, NULL, A);
}

int or_assemble(const or_model* m, const double* U, double* fint, or_csr* Kfull, or_csr* Kff,
                double* stress, double* strain) {
    int64_t n = 3 * m->nn;
    int64_t nt = m->ntet * 144 + m->ntruss * 36;
    int32_t* I = (int32_t*)malloc((size_t)(nt ? nt : 1) * sizeof(int32_t));
    int32_t* J = (int32_t*)malloc((size_t)(nt ? nt : 1) * sizeof(int32_t));
    double* V = (double*)malloc((size_t)(nt ? nt : 1) * sizeof(double));
    if (!I || !J || !V) return OR_ERR_ALLOC;
    memset(fint, 0, (size_t)n * sizeof(double));
    int64_t t = 0;
    int st = OR_OK;
    /* tetrahedra (Tetrahedrons.jl via Entities.jl:211-218): element e's 144
     * triplets sit at 144 e, column-major (Assemblers.jl:59-66), and its forces
     * are summed into fint in element order after the loop, so the OpenMP
     * build (CPU baseline only) gives the same arrays as the serial one */
    double* FE = (double*)malloc((size_t)(m->ntet ? m->ntet : 1) * 12 * sizeof(double));
    if (!FE) return OR_ERR_ALLOC;
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < m->ntet; ++e) {
        double X[12], ue[12], Ke[144], se[9], ee[9];
        double* fe = &FE[e * 12];
        int32_t dofs[12];
        for (int a = 0; a < 4; ++a) {
            int32_t nd = m->tet[e * 4 + a];
            for (int d = 0; d < 3; ++d) {
                dofs[a * 3 + d] = 3 * nd + d; /* local_dofs (Entities.jl:156-171) */
                X[a * 3 + d] = m->X[3 * (int64_t)nd + d];
                ue[a * 3 + d] = U[3 * (int64_t)nd + d];
            }
        }
        double G = m->tet_G ? m->tet_G[e] : m->tet_p1;
        int ste;
        if (m->tet_mat == OR_MAT_LINEAR)
            ste = or_tet_linear(m->tet_p0, m->tet_p1, X, ue, fe, Ke, se, ee);
        else
            ste = or_tet_hyperelastic(m->tet_mat, m->tet_p0, G, X, ue, fe, Ke, se, ee);
        if (ste != OR_OK) {
#pragma omp atomic write
            st = ste;
            continue;
        }
        int64_t tt = 144 * e;
        for (int c = 0; c < 12; ++c)
            for (int r = 0; r < 12; ++r) {
                I[tt] = dofs[r];
                J[tt] = dofs[c];
                V[tt] = Ke[r * 12 + c];
                ++tt;
            }
        if (stress) memcpy(&stress[e * 9], se, sizeof se);
        if (strain) memcpy(&strain[e * 9], ee, sizeof ee);
    }
    if (st == OR_OK) {
        for (int64_t e = 0; e < m->ntet; ++e) /* StructuralAnalyses.jl:105-107 */
            for (int a = 0; a < 4; ++a)
                for (int d = 0; d < 3; ++d)
                    fint[3 * (int64_t)m->tet[e * 4 + a] + d] += FE[e * 12 + a * 3 + d];
        t = 144 * m->ntet;
    }
    free(FE);
    for (int64_t e = 0; e < m->ntruss && st == OR_OK; ++e) {
        double X[6], ue[6], fe[6], Ke[36], s11, e11;
        int32_t dofs[6];
        for (int a = 0; a < 2; ++a) {
            int32_t nd = m->truss[e * 2 + a];
            for (int d = 0; d < 3; ++d) {
                dofs[a * 3 + d] = 3 * nd + d;
                X[a * 3 + d] = m->X[3 * (int64_t)nd + d];
                ue[a * 3 + d] = U[3 * (int64_t)nd + d];
            }
        }
        st = or_truss(m->truss_strain, m->truss_E, m->truss_A[e], X, ue, fe, Ke, &s11, &e11);
        if (st != OR_OK) break;
        for (int r = 0; r < 6; ++r) fint[dofs[r]] += fe[r];
        for (int c = 0; c < 6; ++c)
            for (int r = 0; r < 6; ++r) {
                I[t] = dofs[r];
                J[t] = dofs[c];
                V[t] = Ke[r * 6 + c];
                ++t;
            }
        int64_t k = m->ntet + e;
        if (stress) {
            memset(&stress[k * 9], 0, 9 * sizeof(double));
            stress[k * 9] = s11;
        }
        if (strain) {
            memset(&strain[k * 9], 0, 9 * sizeof(double));
            strain[k * 9] = e11;
        }
    }
    if (st == OR_OK && Kfull) st = coo_to_csr(t, I, J, V, n, NULL, Kfull);
    if (st == OR_OK && Kff) {
        int64_t* map = (int64_t*)malloc((size_t)n * sizeof(int64_t));
        if (!map) {
            st = OR_ERR_ALLOC;
        } else {
            for (int64_t i = 0; i < n; ++i) map[i] = -1;
            for (int64_t i = 0; i < m->nfree; ++i) map[m->free_dofs[i]] = i;
            st = coo_to_csr(t, I, J, V, m->nfree, map, Kff);
            free(map);
        }
    }
    free(I);
    free(J);
    free(V);
    return st;
}

/* ------------------------------------------------------------------------ */
/* CG (IterativeSolvers 0.9.4 cg_iterator! / iterate, restated)              */
/* ------------------------------------------------------------------------ */
static double nrm2(const double* x, int64_t n) {
    double s = 0;
#pragma omp parallel for reduction(+ : s) schedule(static)
    for (int64_t i = 0; i < n; ++i) s += x[i] * x[i];
    return sqrt(s);
}

/* The OpenMP pragmas take effect only in the CPU-baseline build of this file
 * (-fopenmp); there the reductions' summation order depends on the thread
 * count, so only the serial build is the checker. */
int64_t or_cg(const or_csr* A, const double* b, double* x, double abstol, double reltol,
              int64_t maxiter) {
    int64_t n = A->n;
    double* r = (double*)malloc((size_t)(n ? n : 1) * sizeof(double));
    double* u = (double*)calloc((size_t)(n ? n : 1), sizeof(double));
    double* c = (double*)malloc((size_t)(n ? n : 1) * sizeof(double));
    if (!r || !u || !c) return -1;
    memset(x, 0, (size_t)n * sizeof(double));
    memcpy(r, b, (size_t)n * sizeof(double)); /* r = b - A*0 */
    double residual = nrm2(r, n);
    double prev = 1.0;
    double tol = reltol * residual > abstol ? reltol * residual : abstol;
    int64_t it = 0;
    while (!(residual <= tol) && it < maxiter) {
        double beta = residual * residual / (prev * prev);
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; ++i) u[i] = r[i] + beta * u[i];
        or_csr_matvec(A, u, c);
        double uc = 0;
#pragma omp parallel for reduction(+ : uc) schedule(static)
        for (int64_t i = 0; i < n; ++i) uc += u[i] * c[i];
        double alpha = residual * residual / uc;
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; ++i) {
            x[i] += alpha * u[i];
            r[i] -= alpha * c[i];
        }
        prev = residual;
        residual = nrm2(r, n);
        ++it;
    }
    free(r);
    free(u);
    free(c);
    return it;
}

/* ------------------------------------------------------------------------ */
/* Newton-Raphson driver                                                     */
/* ------------------------------------------------------------------------ */
int or_newton_static(const or_model* m, int nsteps, const double* fext_steps,
                     double rel_U_tol, double rel_res_tol, int max_iter, double cg_reltol,
                     double* U, double* U_steps, or_step_record* records) {
    int64_t n = 3 * m->nn, nf = m->nfree;
    double* fint = (double*)malloc((size_t)n * sizeof(double));
    double* r = (double*)malloc((size_t)(nf ? nf : 1) * sizeof(double));
    double* dU = (double*)malloc((size_t)(nf ? nf : 1) * sizeof(double));
    if (!fint || !r || !dU) return OR_ERR_ALLOC;
    if (cg_reltol < 0) cg_reltol = sqrt(2.220446049250313e-16); /* sqrt(eps(Float64)) */
    int st = OR_OK;
    for (int step = 0; step < nsteps && st == OR_OK; ++step) {
        const double* fext = &fext_steps[(int64_t)step * n];
        /* reset!(ResidualsIterationStep) (StructuralSolvers.jl:115-121) */
        double dU_norm = 1e12, r_norm = 1e12, dU_rel = 1e12, r_rel = 1e12;
        int iter = 0, crit = 0;
        int64_t cg_total = 0;
        for (;;) {
            /* isconverged! (StructuralSolvers.jl:148-171) */
            if (!(dU_rel > 0) || !(r_rel > 0)) {
                st = OR_ERR_ZERO_NORM;
                break;
            }
            if (r_rel <= rel_res_tol) {
                crit = 1;
                break;
            } else if (dU_rel <= rel_U_tol) {
                crit = 2;
                break;
            } else if (iter > max_iter) {
                crit = 3;
                break;
            }
            /* assemble!(s, sa) */
            or_csr Kff = {0, 0, NULL, NULL, NULL};
            st = or_assemble(m, U, fint, NULL, &Kff, NULL, NULL);
            if (st != OR_OK) break;
            /* step! (NonLinearStaticAnalyses.jl:107-148) */
            for (int64_t i = 0; i < nf; ++i) r[i] = fext[m->free_dofs[i]] - fint[m->free_dofs[i]];
            cg_total += or_cg(&Kff, r, dU, 0.0, cg_reltol, nf);
            or_csr_free(&Kff);
            dU_norm = nrm2(dU, nf);
            dU_rel = dU_norm / nrm2(U, n);
            r_norm = nrm2(r, nf);
            r_rel = r_norm / nrm2(fext, n);
            for (int64_t i = 0; i < nf; ++i) U[m->free_dofs[i]] += dU[i];
            ++iter;
        }
        if (records) {
            records[step].dU_norm = dU_norm;
            records[step].dU_rel = dU_rel;
            records[step].r_norm = r_norm;
            records[step].r_rel = r_rel;
            records[step].iters = iter;
            records[step].criterion = crit;
            records[step].cg_iters = cg_total;
        }
        if (U_steps) memcpy(&U_steps[(int64_t)step * n], U, (size_t)n * sizeof(double));
    }
    free(fint);
    free(r);
    free(dU);
    return st;
}
