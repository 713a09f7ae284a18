/*
 * dadi_oracle.c -- CPU restatement of dadi's implicit ADI sweeps.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP path: it is
 * built into oracle/libdadi_oracle.so and loaded by tests/, by the smoke()
 * entry point and by bench.py's CPU-baseline leg.  Nothing in the dadi_amd
 * package links, loads or calls it.
 *
 * It restates, in one generic routine per variant, what the reference spreads
 * over fourteen hand-unrolled functions:
 *
 *   order ORC_TV    dadi/integration1D.c:7-66 (implicit_1Dx),
 *                   dadi/integration2D.c:19-160 (implicit_2Dx/2Dy),
 *                   dadi/integration3D.c:13-239 (implicit_3Dx/y/z),
 *                   dadi/integration4D.c:13,91,170,249 (implicit_4Dx/y/z/a),
 *                   dadi/integration5D.c:13,94,176,258,340 (implicit_5Dx..b):
 *                   coefficients rebuilt per line with compute_abc_nobc
 *                   (dadi/integration_shared.c:72-93), b starts at 1/dt,
 *                   r = phi/dt, boundary terms on lines whose other
 *                   coordinates are all 0 (first point) or all 1 (last point).
 *   order ORC_PRE   the constant-parameter path for 2 and 3 populations:
 *                   numpy builds ax,bx,cx once (dadi/Integration.py:1037-1095,
 *                   1120-1221) and implicit_precalc_{2,3}D* adds 1/dt per
 *                   step (dadi/integration2D.c:162-217,
 *                   dadi/integration3D.c:241-345): b = bx + 1/dt,
 *                   r = (1/dt)*phi, boundary terms only in the very first and
 *                   very last line of the axis.
 *   order ORC_PRE1  the constant-parameter 1D path
 *                   (dadi/Integration.py:986-1035): like ORC_PRE but
 *                   r = phi/dt.
 *
 * Every floating-point operation is performed in the reference's order so the
 * results are bit-identical to the reference C built with the same compiler
 * flags (no FMA contraction).  Only delj = 0.5 is supported
 * (use_delj_trick = False, dadi/Integration.py:25), as in the reference GPU
 * path (dadi/cuda/Integration.py:171-172).
 */
#include <stdlib.h>
#include <string.h>
#include <math.h>

#define ORC_TV 0
#define ORC_PRE 1
#define ORC_PRE1 2

/* Numerical-Recipes Thomas solve, dadi/tridiag.c:14-34.  gam is caller
 * scratch of length n (the reference keeps it in a process global). */
void orc_tridiag(const double *a, const double *b, const double *c,
                 const double *r, double *u, int n, double *gam)
{
    double bet = b[0];
    int j;
    u[0] = r[0] / bet;
    for (j = 1; j <= n - 1; j++) {
        gam[j] = c[j - 1] / bet;
        bet = b[j] - a[j] * gam[j];
        u[j] = (r[j] - a[j] * u[j - 1]) / bet;
    }
    for (j = n - 2; j >= 0; j--)
        u[j] -= gam[j + 1] * u[j + 1];
}

/* V(x) = x(1-x)/nu, dadi/integration_shared.c:4-6; the beta form
 * (dadi/integration_shared.c:8-10) is used by the 1D sweep only. */
static double orc_V(double x, double nu) { return 1. / nu * x * (1. - x); }
static double orc_Vbeta(double x, double nu, double beta)
{
    return 1. / nu * x * (1. - x) * pow(beta + 1, 2) / (4 * beta);
}

/* M(x; other coordinates) = sum_j m_j (c_j - x) + 2 gamma (h + (1-2h) x) x(1-x),
 * dadi/integration_shared.c:12-30, summed left to right as written there. */
static double orc_M(double x, int nother, const double *coord, const double *mig,
                    double gamma, double h)
{
    double sel = gamma * 2 * (h + (1. - 2 * h) * x) * x * (1. - x);
    if (nother == 0)
        return sel;
    double acc = mig[0] * (coord[0] - x);
    for (int k = 1; k < nother; k++)
        acc = acc + mig[k] * (coord[k] - x);
    return acc + sel;
}

/*
 * One implicit half-step along `axis` of a D-dimensional phi with n points
 * on every axis (grid xx shared by all axes, as dadi's drivers require).
 *   mig[D-1]  migration rates into this population from the other
 *             populations, in increasing population order
 *             (e.g. axis 1 of 3D: m21, m23).
 *   beta      breeding ratio, used only when D == 1.
 */
void orc_sweep(int order, double *phi, int D, int n, int axis, const double *xx,
               double nu, const double *mig, double gamma, double h, double beta,
               double dt)
{
    int L = n;
    if (L < 3) return;
    long stride = 1, nsys = 1;
    for (int k = axis + 1; k < D; k++) stride *= n;
    for (int k = 0; k < D - 1; k++) nsys *= n;

    double *dx = malloc((L - 1) * sizeof(double));
    double *dfactor = malloc(L * sizeof(double));
    double *xInt = malloc((L - 1) * sizeof(double));
    double *V = malloc(L * sizeof(double));
    double *MInt = malloc((L - 1) * sizeof(double));
    double *a = malloc(L * sizeof(double));
    double *b = malloc(L * sizeof(double));
    double *c = malloc(L * sizeof(double));
    double *r = malloc(L * sizeof(double));
    double *u = malloc(L * sizeof(double));
    double *gam = malloc(L * sizeof(double));
    int ii;

    /* compute_dx / compute_dfactor / compute_xInt,
     * dadi/integration_shared.c:32-50 */
    for (ii = 0; ii < L - 1; ii++) dx[ii] = xx[ii + 1] - xx[ii];
    for (ii = 1; ii < L - 1; ii++) dfactor[ii] = 2. / (dx[ii] + dx[ii - 1]);
    dfactor[0] = 2. / dx[0];
    dfactor[L - 1] = 2. / dx[L - 2];
    for (ii = 0; ii < L - 1; ii++) xInt[ii] = 0.5 * (xx[ii + 1] + xx[ii]);
    for (ii = 0; ii < L; ii++)
        V[ii] = (D == 1) ? orc_Vbeta(xx[ii], nu, beta) : orc_V(xx[ii], nu);

    const double delj = 0.5;
    int nother = D - 1;
    double coord[4];
    int oidx[4];

    for (long sys = 0; sys < nsys; sys++) {
        /* Decompose the system index into the other axes' indices, in
         * increasing axis order (the reference's loop nest order). */
        long rem = sys;
        for (int k = nother - 1; k >= 0; k--) {
            oidx[k] = (int)(rem % n);
            rem /= n;
        }
        int all0 = 1, all1 = 1, allfirst = 1, alllast = 1;
        for (int k = 0; k < nother; k++) {
            coord[k] = xx[oidx[k]];
            if (!(coord[k] == 0)) all0 = 0;
            if (!(coord[k] == 1)) all1 = 0;
            if (oidx[k] != 0) allfirst = 0;
            if (oidx[k] != n - 1) alllast = 0;
        }
        /* base offset of this line: place other indices around `axis` */
        long base = 0;
        {
            int k2 = 0;
            for (int k = 0; k < D; k++) {
                base *= n;
                if (k != axis) base += oidx[k2++];
            }
        }

        for (ii = 0; ii < L - 1; ii++)
            MInt[ii] = orc_M(xInt[ii], nother, coord, mig, gamma, h);

        if (order == ORC_TV) {
            double Mfirst = orc_M(xx[0], nother, coord, mig, gamma, h);
            double Mlast = orc_M(xx[L - 1], nother, coord, mig, gamma, h);
            /* compute_abc_nobc, dadi/integration_shared.c:72-93 */
            a[0] = 0;
            c[L - 1] = 0;
            for (ii = 0; ii < L; ii++) b[ii] = 1. / dt;
            for (ii = 0; ii < L - 1; ii++) {
                double atemp = MInt[ii] * delj + V[ii] / (2 * dx[ii]);
                a[ii + 1] = -dfactor[ii + 1] * atemp;
                b[ii] += dfactor[ii] * atemp;
                double ctemp = -MInt[ii] * (1 - delj) + V[ii + 1] / (2 * dx[ii]);
                b[ii + 1] += dfactor[ii + 1] * ctemp;
                c[ii] = -dfactor[ii] * ctemp;
            }
            for (ii = 0; ii < L; ii++) r[ii] = phi[base + ii * stride] / dt;
            /* boundary terms, e.g. dadi/integration2D.c:67-70 (1D: always) */
            if (all0 && (Mfirst <= 0))
                b[0] += (0.5 / nu - Mfirst) * 2. / dx[0];
            if (all1 && (Mlast >= 0))
                b[L - 1] += -(-0.5 / nu - Mlast) * 2. / dx[L - 2];
        } else {
            /* numpy precalc arrays: ax[1:] += dfact[1:]*(-MInt*delj - V[:-1]/(2dx)),
             * cx[:-1] += dfact[:-1]*(MInt*(1-delj) - V[1:]/(2dx)),
             * bx[:-1] += dfact[:-1]*(MInt*delj + V[:-1]/(2dx)),
             * bx[1:]  += dfact[1:]*(-MInt*(1-delj) + V[1:]/(2dx))
             * (dadi/Integration.py:1084-1087; the 1D form at 1006-1015 gives
             * the same values). */
            for (ii = 0; ii < L; ii++) { a[ii] = 0; b[ii] = 0; c[ii] = 0; }
            for (ii = 0; ii < L - 1; ii++) {
                a[ii + 1] += dfactor[ii + 1] * (-MInt[ii] * delj - V[ii] / (2 * dx[ii]));
                c[ii] += dfactor[ii] * (MInt[ii] * (1 - delj) - V[ii + 1] / (2 * dx[ii]));
            }
            for (ii = 0; ii < L - 1; ii++)
                b[ii] += dfactor[ii] * (MInt[ii] * delj + V[ii] / (2 * dx[ii]));
            for (ii = 0; ii < L - 1; ii++)
                b[ii + 1] += dfactor[ii + 1] * (-MInt[ii] * (1 - delj) + V[ii + 1] / (2 * dx[ii]));
            /* corner boundary terms, dadi/Integration.py:1089-1092:
             * Mx[0,..,0] uses every coordinate at xx[0], Mx[-1,..,-1] at xx[-1] */
            if (allfirst) {
                double cf[4];
                for (int k = 0; k < nother; k++) cf[k] = xx[0];
                double M0 = orc_M(xx[0], nother, cf, mig, gamma, h);
                if (M0 <= 0) b[0] += (0.5 / nu - M0) * 2 / dx[0];
            }
            if (alllast) {
                double cl[4];
                for (int k = 0; k < nother; k++) cl[k] = xx[L - 1];
                double M1 = orc_M(xx[L - 1], nother, cl, mig, gamma, h);
                if (M1 >= 0) b[L - 1] += -(-0.5 / nu - M1) * 2 / dx[L - 2];
            }
            /* per step: b + 1/dt (dadi/integration2D.c:181,
             * dadi/Integration.py:1033) */
            for (ii = 0; ii < L; ii++) b[ii] = b[ii] + 1 / dt;
            if (order == ORC_PRE)
                for (ii = 0; ii < L; ii++) r[ii] = 1 / dt * phi[base + ii * stride];
            else
                for (ii = 0; ii < L; ii++) r[ii] = phi[base + ii * stride] / dt;
        }

        orc_tridiag(a, b, c, r, u, L, gam);
        for (ii = 0; ii < L; ii++) phi[base + ii * stride] = u[ii];
    }

    free(dx); free(dfactor); free(xInt); free(V); free(MInt);
    free(a); free(b); free(c); free(r); free(u); free(gam);
}

/*
 * A whole integration with a precomputed schedule, in C, so that timing the
 * CPU baseline is not dominated by per-step Python.  For step s:
 *   inject[s*D + k] is added to phi at index 1 of axis k (others 0) when
 *   inj_on[k]; then every axis with !frozen[k] is swept with
 *   params[s*stride_p ...]:  per axis k: nu, D-1 migration rates, gamma, h
 *   (stride_p = D*(D+2)), and dt[s].
 */
void orc_run(int order, double *phi, int D, int n, const double *xx,
             long nsteps, const double *dts, const double *params,
             const double *inject, const int *inj_on, const int *frozen,
             double beta)
{
    int per = D + 2;
    long sp = (long)D * per;
    for (long s = 0; s < nsteps; s++) {
        for (int k = 0; k < D; k++) {
            if (!inj_on[k]) continue;
            long off = 0;
            for (int q = 0; q < D; q++) off = off * n + (q == k ? 1 : 0);
            phi[off] += inject[s * D + k];
        }
        for (int k = 0; k < D; k++) {
            if (frozen[k]) continue;
            const double *p = params + s * sp + (long)k * per;
            orc_sweep(order, phi, D, n, k, xx, p[0], p + 1, p[D], p[D + 1], beta, dts[s]);
        }
    }
}
