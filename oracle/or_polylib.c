/*
 * or_polylib.c -- restatement of the Gauss quadrature routines of
 * library/LibUtilities/Polylib/Polylib.cpp that the MODIFIED hex/tet
 * expansions use: Jacobi polynomials (jacobfd, jacobd), their zeros by
 * Newton iteration with polynomial deflation (Jacobz), Gauss-Lobatto-Jacobi
 * and Gauss-Radau-Jacobi(-1) zeros/weights and differentiation matrices.
 *
 * TEST INFRASTRUCTURE ONLY (parity oracle).  Pinned against oracle/_ref,
 * which is the reference Polylib.cpp compiled unmodified.
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>

/* Polylib.cpp:42-44: STOP=30 Newton iterations, EPS = 100*DBL_EPSILON. */
#define OR_NEWTON_MAX 30
#define OR_NEWTON_EPS (100 * DBL_EPSILON)

/* Gamma(x+alpha)/Gamma(y+beta) for integer arguments: Polylib.cpp:1468
 * (gammaFracGammaF, integer branch).  Only integer alpha/beta occur on this
 * path (GLL alpha=beta=0, Radau alpha in {1,2}). */
static double gamma_ratio_int(int x, double alpha, int y, double beta)
{
    int X = x + (int)alpha, Y = y + (int)beta;
    double g = 1.0;
    if (X > Y)
        for (int t = X - 1; t > Y - 1; --t) g *= t;
    else if (Y > X) {
        for (int t = Y - 1; t > X - 1; --t) g *= t;
        g = 1.0 / g;
    }
    return g;
}

/* Gamma(x) = (x-1)! for positive integer x: Polylib.cpp:1413 (gammaF). */
static double gamma_int(double x)
{
    int n = (int)x;
    double g = 1.0, t = x;
    if (n == 0) return 1.0;
    while (--n) { t -= 1.0; g *= t; }
    return g;
}

/* Polylib.cpp:1248-1376 (jacobfd): P_n^{a,b}(z) by the three-term
 * recurrence and, optionally, its derivative from P_n and P_{n-1}. */
void or_jacobfd(int np, const double *z, double *poly, double *polyd, int n,
                double alpha, double beta)
{
    if (np <= 0) return;
    if (n == 0) {
        for (int i = 0; i < np; ++i) {
            if (poly) poly[i] = 1.0;
            if (polyd) polyd[i] = 0.0;
        }
        return;
    }
    if (n == 1) {
        for (int i = 0; i < np; ++i) {
            if (poly) poly[i] = 0.5 * (alpha - beta + (alpha + beta + 2.0) * z[i]);
            if (polyd) polyd[i] = 0.5 * (alpha + beta + 2.0);
        }
        return;
    }
    const double apb = alpha + beta;
    double *work = (double *)malloc(3 * (size_t)np * sizeof(double));
    double *pm1 = work, *pm2 = work + np;
    double *p = poly ? poly : work + 2 * np;
    for (int i = 0; i < np; ++i) {
        pm2[i] = 1.0;
        pm1[i] = 0.5 * (alpha - beta + (alpha + beta + 2.0) * z[i]);
    }
    for (int k = 2; k <= n; ++k) {
        double a1 = 2.0 * k * (k + apb) * (2.0 * k + apb - 2.0);
        double a2 = (2.0 * k + apb - 1.0) * (alpha * alpha - beta * beta);
        double a3 = (2.0 * k + apb - 2.0) * (2.0 * k + apb - 1.0) * (2.0 * k + apb);
        double a4 = 2.0 * (k + alpha - 1.0) * (k + beta - 1.0) * (2.0 * k + apb);
        a2 /= a1;
        a3 /= a1;
        a4 /= a1;
        for (int i = 0; i < np; ++i) {
            p[i] = (a2 + a3 * z[i]) * pm1[i] - a4 * pm2[i];
            pm2[i] = pm1[i];
            pm1[i] = p[i];
        }
    }
    if (polyd) {
        double a1 = n * (alpha - beta);
        double a2 = n * (2.0 * n + alpha + beta);
        double a3 = 2.0 * (n + alpha) * (n + beta);
        double a4 = 2.0 * n + alpha + beta;
        a1 /= a4;
        a2 /= a4;
        a3 /= a4;
        for (int i = 0; i < np; ++i) {
            polyd[i] = (a1 - a2 * z[i]) * p[i] + a3 * pm2[i];
            polyd[i] /= (1.0 - z[i] * z[i]);
        }
    }
    free(work);
}

/* Polylib.cpp:1378-1399 (jacobd): dP_n^{a,b}/dz = (a+b+n+1)/2 P_{n-1}^{a+1,b+1}. */
void or_jacobd(int np, const double *z, double *polyd, int n, double alpha,
               double beta)
{
    if (n == 0) {
        for (int i = 0; i < np; ++i) polyd[i] = 0.0;
        return;
    }
    or_jacobfd(np, z, polyd, NULL, n - 1, alpha + 1.0, beta + 1.0);
    for (int i = 0; i < np; ++i) polyd[i] *= 0.5 * (alpha + beta + (double)n + 1.0);
}

/* Polylib.cpp:1597-1636 (Jacobz): zeros of P_n^{a,b} by Newton iteration
 * from Chebyshev-Gauss guesses with deflation by the zeros already found.
 * Selected by Polylib.cpp:126-129 (POLYNOMIAL_DEFLATION is defined). */
void or_jacobz(int n, double *z, double alpha, double beta)
{
    if (n <= 0) return;
    const double dth = M_PI / (2.0 * (double)n);
    double rlast = 0.0;
    for (int k = 0; k < n; ++k) {
        double r = -cos((2.0 * (double)k + 1.0) * dth);
        if (k) r = 0.5 * (r + rlast);
        for (int it = 1; it < OR_NEWTON_MAX; ++it) {
            double pv, pd;
            or_jacobfd(1, &r, &pv, &pd, n, alpha, beta);
            double defl = 0.0;
            for (int i = 0; i < k; ++i) defl += 1.0 / (r - z[i]);
            double dr = -pv / (pd - defl * pv);
            r += dr;
            if (fabs(dr) < OR_NEWTON_EPS) break;
        }
        z[k] = r;
        rlast = r;
    }
}

/* Polylib.cpp:266-310 (zwglj): Gauss-Lobatto-Jacobi points and weights. */
void or_zwglj(double *z, double *w, int np, double alpha, double beta)
{
    if (np == 1) { z[0] = 0.0; w[0] = 2.0; return; }
    if (np == 2) { z[0] = -1.0; z[1] = 1.0; w[0] = w[1] = 1.0; return; }
    const double apb = alpha + beta;
    z[0] = -1.0;
    z[np - 1] = 1.0;
    or_jacobz(np - 2, z + 1, alpha + 1.0, beta + 1.0);
    or_jacobfd(np, z, w, NULL, np - 1, alpha, beta);
    double fac = pow(2.0, apb + 1.0) * gamma_ratio_int(np, alpha, np, 0.0) *
                 gamma_ratio_int(np, beta, np + 1, apb);
    fac /= (np - 1);
    for (int i = 0; i < np; ++i) w[i] = fac / (w[i] * w[i]);
    w[0] *= (beta + 1.0);
    w[np - 1] *= (alpha + 1.0);
}

/* Polylib.cpp:182-215 (zwgrjm): Gauss-Radau-Jacobi with z=-1 included. */
void or_zwgrjm(double *z, double *w, int np, double alpha, double beta)
{
    if (np == 1) { z[0] = 0.0; w[0] = 2.0; return; }
    const double apb = alpha + beta;
    z[0] = -1.0;
    or_jacobz(np - 1, z + 1, alpha, beta + 1.0);
    or_jacobfd(np, z, w, NULL, np - 1, alpha, beta);
    double fac = pow(2.0, apb) * gamma_ratio_int(np, alpha, np, 0.0) *
                 gamma_ratio_int(np, beta, np + 1, apb);
    fac /= (beta + np);
    for (int i = 0; i < np; ++i) w[i] = fac * (1.0 - z[i]) / (w[i] * w[i]);
    w[0] *= (beta + 1.0);
}

/* Polylib.cpp:833-912 (Dglj).  Storage D[i*np+j] = l_i'(z_j). */
void or_dglj(double *D, const double *z, int np, double alpha, double beta)
{
    if (np <= 1) { D[0] = 0.0; return; }
    double *pd = (double *)malloc((size_t)np * sizeof(double));
    pd[0] = 2.0 * pow(-1.0, np) * gamma_ratio_int(np, beta, np - 1, 0.0);
    pd[0] /= gamma_int(beta + 2.0);
    or_jacobd(np - 2, z + 1, pd + 1, np - 2, alpha + 1.0, beta + 1.0);
    for (int i = 1; i < np - 1; ++i) pd[i] *= (1.0 - z[i] * z[i]);
    pd[np - 1] = -2.0 * gamma_ratio_int(np, alpha, np - 1, 0.0);
    pd[np - 1] /= gamma_int(alpha + 2.0);
    for (int i = 0; i < np; ++i)
        for (int j = 0; j < np; ++j) {
            double v;
            if (i != j) v = pd[j] / (pd[i] * (z[j] - z[i]));
            else if (j == 0)
                v = (alpha - (np - 1) * (np + alpha + beta)) / (2.0 * (beta + 2.0));
            else if (j == np - 1)
                v = -(beta - (np - 1) * (np + alpha + beta)) / (2.0 * (alpha + 2.0));
            else
                v = (alpha - beta + (alpha + beta) * z[j]) / (2.0 * (1.0 - z[j] * z[j]));
            D[i * np + j] = v;
        }
    free(pd);
}

/* Polylib.cpp:704-765 (Dgrjm).  Storage D[i*np+j] = l_i'(z_j). */
void or_dgrjm(double *D, const double *z, int np, double alpha, double beta)
{
    if (np <= 0) { D[0] = 0.0; return; }
    double *pd = (double *)malloc((size_t)np * sizeof(double));
    pd[0] = pow(-1.0, np - 1) * gamma_ratio_int(np + 1, beta, np, 0.0);
    pd[0] /= gamma_int(beta + 2.0);
    or_jacobd(np - 1, z + 1, pd + 1, np - 1, alpha, beta + 1.0);
    for (int i = 1; i < np; ++i) pd[i] *= (1.0 + z[i]);
    for (int i = 0; i < np; ++i)
        for (int j = 0; j < np; ++j) {
            double v;
            if (i != j) v = pd[j] / (pd[i] * (z[j] - z[i]));
            else if (j == 0)
                v = -(np + alpha + beta + 1.0) * (np - 1.0) / (2.0 * (beta + 2.0));
            else
                v = (alpha - beta + 1.0 + (alpha + beta + 1.0) * z[j]) /
                    (2.0 * (1.0 - z[j] * z[j]));
            D[i * np + j] = v;
        }
    free(pd);
}
